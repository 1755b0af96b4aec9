"""Launch shapes whose kernels, compiled at -O3, gave wrong transition counts on gfx950 that
changed from run to run, while their source runs bit-exact in the host emulation of
tests/kernel_emu (also under AddressSanitizer and MemorySanitizer).  The runtime compiles its
kernels at -O1, where every case below is bit-exact against the oracle and deterministic:

- a 177-node network (3 state words) in blocks of 512 and 1,024 threads with the tables and
  the window-entropy partial sums in LDS (at -O3: 4.04-4.09 million transitions instead of
  4,106,593, different on every run; exact with blocks of 64-256 threads);
- the 4-word class of a sweep without entropy partials in LDS (every variant off by up to 5 %);
- networks of 60 and 130 nodes whose default shape disagreed with five other shapes;
- 1,250- and 2,000-node networks with the active sets in HBM and register rate sums of two
  words;
- the 256-variant sweep (ten word classes), repeated, with and without entropy partials.
"""
import hashlib
import os

import numpy as np
import pytest

from oracle.oracle import OracleModel
from upmabossg_amd import Model, synth

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _sweep_nets():
    return synth.sweep_variants(256)


def _digest(ctx, v):
    """(transitions, digest of the window table and entropies) of sweep variant v."""
    info, prob, th = ctx.variant_result(v)[:3]
    return info.transitions, hashlib.sha1(prob.tobytes() + th.tobytes()).hexdigest()


def _oracle(net, count, seed, vid=0, threads=16):
    return OracleModel(net.bnd(), net.cfg()).run(count=count, n_trajectories=count, seed=seed, variant_id=vid,
                                                 threads=threads)


@pytest.mark.parametrize("block", [0, 512, 1024])
def test_177_node_network_large_blocks(gpu_ctx, block):
    net = _sweep_nets()[91]
    n = net.sample_count
    _flush(f"case: 177 nodes, {n} trajectories, block {block or 'default'}")
    m = Model(net.bnd(), net.cfg())
    r = _oracle(net, n, seed=1, vid=91)
    for rep in range(2):
        g = gpu_ctx.run(m, n_trajectories=n, seed=1, variant_id=91, block=block)
        assert g.transitions == r.flips == 4106593, rep
        assert np.array_equal(g.prob_raw, r.prob) and np.array_equal(g.th_raw, r.th), rep


def test_four_word_sweep_class_without_entropy_partials(gpu_ctx, monkeypatch):
    nets = _sweep_nets()
    ids = [v for v in range(256) if 3 <= -(-nets[v].n // 64) <= 4]
    models = [Model(nets[v].bnd(), nets[v].cfg()) for v in ids]
    vs = [(m, nets[v].sample_count, 1, v) for m, v in zip(models, ids)]

    def sweep():
        gpu_ctx.run_variants(vs)
        return [_digest(gpu_ctx, k) for k in range(len(ids))]

    _flush(f"case: 4-word sweep class, {len(ids)} variants, entropy partials in LDS")
    ref = sweep()
    for v in (140, 158):  # the two smallest variants of the class against the oracle
        k = ids.index(v)
        info, prob, th = gpu_ctx.variant_result(k)[:3]
        r = _oracle(nets[v], nets[v].sample_count, seed=1, vid=v)
        assert info.transitions == r.flips
        assert np.array_equal(prob, r.prob) and np.array_equal(th, r.th)
    monkeypatch.setenv("MBG_TUNE_TH_LDS", "0")
    for rep in range(2):
        _flush(f"case: 4-word sweep class without entropy partials, repetition {rep}")
        assert sweep() == ref, rep


SHAPES = [{}, {"MBG_TUNE_TH_LDS": "0"}, {"MBG_TUNE_SCAN": "8"}, {"MBG_TUNE_DEP_BATCH": "1"},
          {"MBG_TUNE_TABLES": "0"}, {"MBG_TUNE_BLOCK": "256"}]


@pytest.mark.parametrize("n,kin,seed", [(60, 2, 1004), (130, 4, 1008)])
def test_small_networks_agree_across_shapes(gpu_ctx, monkeypatch, n, kin, seed):
    net = synth.random_network(n, 2, kin, seed=seed, rates="loguniform", istate="uniform", readouts=16,
                               max_time=50.0, time_tick=0.5, sample_count=65536)
    m = Model(net.bnd(), net.cfg())
    r = _oracle(net, 65536, seed=3)
    for sh in SHAPES:
        for k in [k for k in list(os.environ) if k.startswith("MBG_TUNE_")]:
            monkeypatch.delenv(k)
        for k, v in sh.items():
            monkeypatch.setenv(k, v)
        for rep in range(2):
            _flush(f"case: {n} nodes, shape {sh}, repetition {rep}")
            g = gpu_ctx.run(m, n_trajectories=65536, seed=3)
            assert g.transitions == r.flips, (sh, rep)
            assert np.array_equal(g.prob_raw, r.prob) and np.array_equal(g.th_raw, r.th), (sh, rep)


@pytest.mark.parametrize("n", [1250, 2000])
def test_active_sets_in_hbm_with_two_word_sum_groups(gpu_ctx, monkeypatch, n):
    """An even number of state words from 18 to 32 (20 and 32 here): by default the active
    sets go to HBM with one register rate sum per two words; the tables equal those of the
    all-LDS layout, run after run, and the oracle's on a sample."""
    net = synth.random_network(n, 2, 3, seed=77 + n, rates="loguniform", istate="uniform", readouts=16,
                               max_time=20.0, time_tick=0.5, sample_count=16384)
    m = Model(net.bnd(), net.cfg())
    _flush(f"case: {n} nodes, active sets in HBM")
    a = [gpu_ctx.run(m, n_trajectories=16384, seed=5) for _ in range(2)]
    assert gpu_ctx.info().state_hbm == 2
    monkeypatch.setenv("MBG_TUNE_STATE_HBM", "0")
    _flush(f"case: {n} nodes, lanes' state in LDS")
    b = gpu_ctx.run(m, n_trajectories=16384, seed=5)
    assert gpu_ctx.info().state_hbm == 0
    for g in a[1:] + [b]:
        assert g.transitions == a[0].transitions
        assert np.array_equal(g.prob_raw, a[0].prob_raw) and np.array_equal(g.th_raw, a[0].th_raw)
    monkeypatch.delenv("MBG_TUNE_STATE_HBM")
    _flush(f"case: {n} nodes, 512 trajectories against the oracle")
    g = gpu_ctx.run(m, count=512, n_trajectories=512, seed=5)
    r = OracleModel(net.bnd(), net.cfg()).run(count=512, n_trajectories=512, seed=5, threads=16)
    assert g.transitions == r.flips
    assert np.array_equal(g.prob_raw, r.prob) and np.array_equal(g.th_raw, r.th)


def test_full_sweep_repeats_exactly(gpu_ctx, monkeypatch):
    """The 256 variants of the sweep benchmark (50-2,000 nodes, ten word classes), run three
    times (the last without entropy partials in LDS): every variant's tables are the same."""
    nets = _sweep_nets()
    models = [Model(n.bnd(), n.cfg()) for n in nets]
    vs = [(m, n.sample_count, 1, v) for v, (m, n) in enumerate(zip(models, nets))]
    runs = []
    for rep in range(3):
        if rep == 2:
            monkeypatch.setenv("MBG_TUNE_TH_LDS", "0")
        _flush(f"case: 256-variant sweep, repetition {rep}")
        gpu_ctx.run_variants(vs)
        runs.append([_digest(gpu_ctx, v) for v in range(len(vs))])
    for rep in (1, 2):
        bad = [v for v in range(len(vs)) if runs[rep][v] != runs[0][v]]
        assert not bad, (rep, bad[:10])

"""The trajectory kernel's own source, compiled for the host and run in the CPU emulation of
tests/kernel_emu (one coroutine per GPU thread, wave collectives and block barriers emulated),
against the oracle bit for bit: transition counts and flip hashes per trajectory, window tables,
entropies, final states and fixed points.  Each case forces a launch shape (state layout, sum
groups, sweep class, sparse tables, logic programs, ...) so that the kernel logic of every
layout is checked on the CPU, before and independently of any GPU run.

The 1,100-node variant in a 20-word sweep class with the active sets in HBM and two-word rate
sum groups is the layout whose GPU run once counted 226,839 transitions against the oracle's
227,473 with kernels compiled at -O3; emulated, the same source, shape, block, grid and draws
give the oracle's numbers exactly (also under AddressSanitizer), and the GPU gives them too
with the -O1 code the runtime now builds (tests/test_gpu_codegen.py covers the shapes whose
-O3 code miscounted).
"""
import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.conftest import model_text
from tests.kernel_emu import emu
from upmabossg_amd import synth


def _oracle(v, trace=True, compat=False):
    o = OracleModel(v["bnd"], v["cfg"])
    if any(v.get(k) is not None for k in ("mu", "md", "istate")):
        o.set_variant(v.get("mu"), v.get("md"), v.get("istate"))
    return o.run(count=v.get("count", v["n"]), first_id=v.get("first_id", 0), n_trajectories=v["n"], seed=v["seed"],
                 variant_id=v.get("vid", 0), threads=8, record_summaries=trace, compat_batch_resume=compat)


def _check(g, r, trace):
    assert g["flips"] == r.flips, "transition counts differ"
    if trace:
        s = r.summaries
        assert np.array_equal(g["summ"][:, 0], s["flips"]), "per-trajectory transition counts differ"
        assert np.array_equal(g["summ"][:, 1], s["flip_hash"]), "flip sequences differ"
        assert np.array_equal(g["summ"][:, 2], s["final_time"].view(np.uint64)), "final times differ"
        assert np.array_equal(g["summ"][:, 3], s["status"].astype(np.uint64)), "statuses differ"
    assert g["FP"] == r.prob_exponent and g["FT"] == r.th_exponent
    assert np.array_equal(g["th"], r.th), "window entropies differ"
    if r.readouts > 20:
        assert g["cell_miss"] == 0
        taken = g["cells"] != np.uint64(0xFFFFFFFFFFFFFFFF)
        keys, vals = g["cells"][taken], g["cellv"][taken]
        order = np.argsort(keys, kind="stable")
        keys, vals = keys[order], vals[order]
        K, W = r.readouts, r.windows
        win = keys >> np.uint64(K)
        st = keys & np.uint64((1 << K) - 1)
        inner = win < np.uint64(W)
        assert np.array_equal(win[inner].astype(np.uint32), r.cells[0]), "window cells differ"
        assert np.array_equal(st[inner], r.cells[1]) and np.array_equal(vals[inner], r.cells[2])
        assert np.array_equal(st[~inner], r.final_cells[0]) and np.array_equal(vals[~inner], r.final_cells[1])
    else:
        assert np.array_equal(g["prob"], r.prob.ravel()), "window tables differ"
        assert np.array_equal(g["fin"], r.final), "final states differ"
    fs, fc = emu.distinct_fixed_points(g["fp_records"], r.fixed_states.shape[1])
    assert np.array_equal(fs, r.fixed_states) and np.array_equal(fc, r.fixed_counts), "fixed points differ"


def _run(variants, grid=2, compat=0, asan=False, trace=True, **shape):
    sh = emu.shape_for(variants, trace=int(trace), **shape)
    exe = emu.build(sh, asan=asan)
    gs = emu.run(exe, variants, grid=grid, compat=compat)
    for v, g in zip(variants, gs):
        _check(g, _oracle(v, trace=trace and not sh["multi"], compat=bool(compat)), trace and not sh["multi"])
    return gs


def _net(n, seed=3, count=96, **kw):
    net = synth.random_network(n, 2, 4, seed=seed, rates="loguniform", istate="uniform", readouts=16,
                               max_time=kw.pop("max_time", 10.0), time_tick=kw.pop("time_tick", 0.5),
                               sample_count=count, **kw)
    return dict(bnd=net.bnd(), cfg=net.cfg(), n=count, seed=5)


def test_small_network_default_layout():
    net = synth.config_small()
    _run([dict(bnd=net.bnd(), cfg=net.cfg(), n=400, seed=1)], block=64, tables_lds=1, wsum_reg=1)


@pytest.mark.parametrize("layout", ["lds", "act_hbm", "act_hbm_pairs", "state_hbm", "x_lds", "lds_sums"])
def test_thousand_node_network_layouts(layout):
    v = _net(1000, count=64)
    over = dict(lds=dict(block=128, tables_lds=1, wsum_reg=1),
                act_hbm=dict(block=128, tables_lds=1, wsum_reg=1, state_hbm=2),
                # rounds of two dependents: kept in the source (the runtime refuses them)
                act_hbm_pairs=dict(block=128, tables_lds=1, wsum_reg=1, state_hbm=2, dep_batch=2),
                state_hbm=dict(block=64, wsum_reg=2, sumw=4, state_hbm=1),
                x_lds=dict(block=128, wsum_reg=2, sumw=3, state_hbm=3),
                lds_sums=dict(block=64, wsum_reg=0, dep_cap=0, scan=8))[layout]
    _run([v], grid=2, **over)


def test_two_word_sum_groups_with_active_sets_in_hbm_single_run():
    v = _net(1100, count=48)
    _run([v], grid=2, block=128, wsum_reg=1, sumw=2, state_hbm=2, nw=20)


def test_sweep_class_of_the_gpu_miscount():
    """The 20-word sweep class (1,100-node variant with rate scales and a knockout, 200
    trajectories, seed 12, variant id 5) with the active sets in HBM and two-word sum groups,
    one block of 768 threads, as the sweep ran it on the GPU: bit-exact, bounds-checked."""
    from tests.test_gpu_sweep import _variants

    b, c, count, (mu, md, ist) = _variants()[5]
    v = dict(bnd=b, cfg=c, n=count, seed=12, vid=5, mu=mu, md=md, istate=ist)
    _run([v], grid=1, asan=True, trace=False, nw=20, block=768, multi=1, wsum_reg=1, sumw=2, state_hbm=2)


@pytest.mark.parametrize("state_hbm,wsum_reg,sumw", [(2, 1, 1), (1, 1, 2), (1, 2, 3)])
def test_sweep_class_several_variants(state_hbm, wsum_reg, sumw):
    """A sweep class of three variants with the active sets (2) or the whole lane state (1) in
    HBM; with the whole state in HBM also with two-word register sum groups and with two-level
    sums."""
    vs = [_net(n, seed=s, count=c) for n, s, c in [(300, 4, 70), (420, 5, 40), (500, 6, 90)]]
    for k, v in enumerate(vs):
        v.update(seed=20 + k, vid=k)
    _run(vs, grid=3, trace=False, nw=8, block=128, multi=1, wsum_reg=wsum_reg, sumw=sumw, state_hbm=state_hbm)


@pytest.mark.parametrize("state_hbm,wsum_reg,sumw", [(2, 1, 1), (0, 1, 1), (0, 2, 3)])
def test_sweep_class_tables_in_lds(state_hbm, wsum_reg, sumw):
    """A sweep class whose blocks stage their current variant's tables in LDS (every wave of a
    block on one variant, blocks moving on together): four variants of different table sizes
    on three blocks, the blocks starting at different variants, with the active sets in HBM or
    the lanes' state in LDS (register sums, or two-level sums)."""
    vs = [_net(n, seed=s, count=c) for n, s, c in [(300, 4, 70), (420, 5, 40), (500, 6, 90), (260, 7, 30)]]
    for k, v in enumerate(vs):
        v.update(seed=30 + k, vid=k)
    _run(vs, grid=3, trace=False, nw=8, block=128, multi=1, tables_lds=1, wsum_reg=wsum_reg, sumw=sumw,
         state_hbm=state_hbm)


def test_wide_model_state_in_hbm_two_level_sums():
    v = _net(2600, count=24, max_time=4.0)
    _run([v], grid=2, block=64, wsum_reg=2, sumw=3, state_hbm=1)


@pytest.mark.parametrize("discrete", [0, 1])
def test_sparse_readouts_and_logic_programs(discrete):
    net = synth.random_network(60, 2, 4, seed=9, rates="loguniform", istate="uniform", readouts=24,
                               max_time=6.0, time_tick=0.5, sample_count=200)
    cfg = net.cfg() + ("\ndiscrete_time = 1;\n" if discrete else "")
    _run([dict(bnd=net.bnd(), cfg=cfg, n=200, seed=3)], grid=2, block=64, wsum_reg=1)
    b, c = synth.ring_network(100, steps_to_fixed_point=50, sample_count=300, time_tick=5.0, max_time=100.0)
    if discrete:
        c += "\ndiscrete_time = 1;\n"
    _run([dict(bnd=b, cfg=c, n=300, seed=4)], grid=2, block=64, wsum_reg=1)


def _check_errors(g, r):
    assert g["err_miss"] == 0
    assert g["FQ"] == r.sq_exponent and g["FQT"] == r.th_sq_exponent
    assert np.array_equal(g["th2"], r.th_sq), "TH sums of squares differ"
    if r.readouts > 20:
        taken = g["cells"] != np.uint64(0xFFFFFFFFFFFFFFFF)
        keys, sq = g["cells"][taken], g["cellv2"][taken]
        order = np.argsort(keys, kind="stable")
        keys, sq = keys[order], sq[order]
        inner = (keys >> np.uint64(r.readouts)) < np.uint64(r.windows)
        assert np.array_equal(sq[inner], r.cell_squares), "cell sums of squares differ"
    else:
        assert np.array_equal(g["prob2"], r.prob_sq.ravel()), "window sums of squares differ"


@pytest.mark.parametrize("case", ["dense_act_hbm", "discrete", "sparse", "sweep"])
def test_error_column_sums_bit_exact(case):
    """MBG_ERRORS: each lane's window list of (state, time), merged on revisits and flushed with
    its squares when the window ends, gives the oracle's sums of squares bit for bit."""
    if case == "dense_act_hbm":
        v = _net(1000, count=48, time_tick=2.0)
        vs, over = [v], dict(block=128, tables_lds=1, wsum_reg=1, state_hbm=2)
    elif case == "discrete":
        b, c = model_text("cellcycle")
        vs, over = [dict(bnd=b, cfg=c, n=300, seed=7)], dict(block=64, wsum_reg=1, tables_lds=1)
    elif case == "sparse":
        net = synth.random_network(60, 2, 4, seed=9, rates="loguniform", istate="uniform", readouts=24,
                                   max_time=6.0, time_tick=1.5, sample_count=150)
        vs, over = [dict(bnd=net.bnd(), cfg=net.cfg(), n=150, seed=3)], dict(block=64, wsum_reg=1)
    else:
        vs = [_net(n, seed=s, count=c, time_tick=1.0) for n, s, c in [(300, 4, 50), (420, 5, 30)]]
        for k, v in enumerate(vs):
            v.update(seed=20 + k, vid=k)
        over = dict(nw=8, block=128, multi=1, wsum_reg=1, state_hbm=2)
    sh = emu.shape_for(vs, errors=1, **over)
    exe = emu.build(sh)
    gs = emu.run(exe, vs, grid=2, err_cap=32)
    for v, g in zip(vs, gs):
        assert g["err_miss"] == 0
        r = _oracle(v, trace=False)
        _check(g, r, False)
        _check_errors(g, r)
    # a list of one entry overflows on a window with two states; the miss is counted (the
    # runtime then repeats the run with a longer list)
    if case == "dense_act_hbm":
        assert emu.run(exe, vs, grid=2, err_cap=1)[0]["err_miss"] > 0


def test_generic_rates_and_batch_resume():
    from tests.test_gpu_parity import EDGE_MODELS

    b, c = EDGE_MODELS["generic"]
    _run([dict(bnd=b, cfg=c, n=500, seed=2)], grid=2, block=64, wsum_reg=1, tables_lds=1, th_lds=1, cells_lds=1)
    net = synth.config_large(n=200)
    _run([dict(bnd=net.bnd(), cfg=net.cfg(), n=64, seed=6)], grid=1, compat=1, block=64, wsum_reg=1)

"""Variant sweep (mbg_run_variants): every variant of one graph launch matches its own oracle
replay bit-exactly, across all state-word classes, with knockouts, rate scales, a discrete-time
model and generic rate expressions; the device marginals equal the window tables reduced."""
import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.conftest import model_text
from tests.test_gpu_parity import EDGE_MODELS
from upmabossg_amd import Model, synth

pytestmark = pytest.mark.gpu


def _variants():
    rng = np.random.default_rng(11)
    out = []
    for k, (n, count) in enumerate([(50, 800), (70, 600), (130, 500), (300, 400), (700, 300), (1100, 200),
                                    (2000, 128), (64, 700), (128, 500)]):
        net = synth.random_network(n, 2, 4, seed=100 + k, rates="loguniform", istate="uniform", readouts=16,
                                   max_time=20.0, time_tick=0.5, sample_count=count)
        mu = np.exp(rng.normal(0, 0.5, n)).astype(np.float32)
        md = np.exp(rng.normal(0, 0.5, n)).astype(np.float32)
        ist = net.istate.copy()
        for i in rng.choice(n, size=k % 4, replace=False):  # knockouts
            mu[i] = 0.0
            ist[i] = 0.0
        out.append((net.bnd(), net.cfg(), count, (mu, md, ist)))
    b, c = model_text("cellcycle")
    out.append((b, c, 3000, None))
    b, c = EDGE_MODELS["generic"]
    out.append((b, c, 2000, None))
    b, c = model_text("metastasis")
    out.append((b, c, 500, None))
    # nodes reading more than 20 nodes (logic programs) in classes shared with table nodes
    b, c = synth.ring_network(100, steps_to_fixed_point=50, sample_count=1000, time_tick=5.0, max_time=100.0)
    out.append((b, c, 1000, None))
    from tests.wide_models import wide_model

    b, c = wide_model(seed=5, sample_count=800)
    out.append((b, c, 800, None))
    return out


def test_sweep_matches_oracle_per_variant(gpu_ctx):
    specs = _variants()
    models = []
    for b, c, count, var in specs:
        m = Model(b, c)
        if var is not None:
            m.set_variant(*var)
        models.append(m)
    print(f"case: sweep of {len(models)} variants", flush=True)
    info = gpu_ctx.run_variants([(m, count, 7 + v, v) for v, (m, (_, _, count, _)) in enumerate(zip(models, specs))])
    assert info.variants == len(specs) and info.classes >= 6
    total = 0
    for v, (b, c, count, var) in enumerate(specs):
        o = OracleModel(b, c)
        if var is not None:
            o.set_variant(*var)
        r = o.run(count=count, seed=7 + v, variant_id=v, threads=16)
        i, prob, th, fin, fs, fc = gpu_ctx.variant_result(v)
        assert i.transitions == r.flips, v
        assert np.array_equal(prob, r.prob), v
        assert np.array_equal(th, r.th), v
        assert np.array_equal(fin, r.final), v
        assert np.array_equal(fs, r.fixed_states) and np.array_equal(fc, r.fixed_counts), v
        # device marginals = the window table summed over the states with each readout ON
        K = r.readouts
        bits = (np.arange(1 << K)[:, None] >> np.arange(K)[None, :]) & 1
        raw = (r.prob.astype(object) @ bits.astype(object)) if K else np.zeros((r.windows, 0))
        if K:
            denom = count if i.discrete else count * float(np.float32(i.time_tick)) * 2.0 ** r.prob_exponent
            want = np.array(raw, dtype=np.float64) / denom
            assert np.allclose(gpu_ctx.variant_marginals(v), want, rtol=1e-12, atol=0)
        total += r.flips
    assert info.transitions == total


def test_sweep_refuses_more_than_20_readouts(gpu_ctx):
    from upmabossg_amd import MbgError

    net = synth.random_network(40, 2, 3, seed=3, readouts=24, sample_count=100)
    with pytest.raises(MbgError) as e:
        gpu_ctx.run_variants([(Model(net.bnd(), net.cfg()), 100, 1, 0)])
    assert e.value.status == 2 and "at most 20" in str(e.value)

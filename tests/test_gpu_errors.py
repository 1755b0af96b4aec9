"""Error columns on the GPU (ErrorProba, ErrorTH: the columns the reference writes as 0,
src/statistics/window_average_small.cpp:211,235,245): the sums of squared per-trajectory window
totals equal the oracle's bit for bit (dense, discrete, sparse, logic programs, a sweep), a run
that needs longer per-lane window lists than the first allocation repeats itself transparently,
and accumulating them leaves the window tables unchanged."""
import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.conftest import model_text
from upmabossg_amd import Model, synth

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _check(ctx, bnd, cfg, count, seed=1, n=None):
    n = n or count
    m = Model(bnd, cfg)
    _flush(f"case: error columns, N={m.num_nodes} count={count}")
    g = ctx.run(m, n_trajectories=n, count=count, seed=seed, error_columns=True)
    r = OracleModel(bnd, cfg).run(count=count, n_trajectories=n, seed=seed, threads=16)
    assert g.transitions == r.flips
    assert g.sq_exponent == r.sq_exponent and g.th_sq_exponent == r.th_sq_exponent
    assert np.array_equal(g.th_raw, r.th) and np.array_equal(g.th_sq, r.th_sq), "TH sums differ"
    if r.readouts > 20:
        for a, b in zip(g.cells, r.cells):
            assert np.array_equal(a, b), "window cells differ"
        assert np.array_equal(g.cell_squares, r.cell_squares), "cell sums of squares differ"
    else:
        assert np.array_equal(g.prob_raw, r.prob), "window tables differ"
        assert np.array_equal(g.prob_sq, r.prob_sq), "window sums of squares differ"
        # the normalised standard errors follow from the exact sums
        P = g.probtraj
        E2 = r.prob_sq.astype(np.float64) / 2.0 ** r.sq_exponent / n
        want = np.sqrt(np.clip(E2 - P * P, 0, None) / (n - 1))
        assert np.allclose(g.prob_error, want, rtol=1e-12, atol=1e-15)
    return m, g


def test_small_network_error_sums(gpu_ctx):
    net = synth.config_small()
    _check(gpu_ctx, net.bnd(), net.cfg(), count=10000)


@pytest.mark.parametrize("name,count", [("cellcycle", 20000), ("metastasis", 1000), ("sizek", 2000)])
def test_reference_models_error_sums(gpu_ctx, name, count):
    b, c = model_text(name)
    _check(gpu_ctx, b, c, count=count)


def test_large_network_error_sums_and_tables_unchanged(gpu_ctx):
    net = synth.config_large()
    m, g = _check(gpu_ctx, net.bnd(), net.cfg(), count=2048)
    plain = gpu_ctx.run(m, n_trajectories=2048, count=2048, seed=1)
    assert np.array_equal(plain.prob_raw, g.prob_raw) and np.array_equal(plain.th_raw, g.th_raw)
    assert plain.transitions == g.transitions


def test_long_windows_repeat_with_longer_lists(gpu_ctx):
    """Windows of 25 time units on a fast 300-node network: a trajectory visits far more than
    the first list length (8) of projected states per window, so the run repeats with longer
    lists and still gives the oracle's sums."""
    net = synth.random_network(300, 2, 4, seed=5, rates="loguniform", istate="uniform", readouts=16,
                               max_time=50.0, time_tick=25.0, sample_count=1024)
    _check(gpu_ctx, net.bnd(), net.cfg(), count=1024)


@pytest.mark.parametrize("readouts,discrete", [(24, False), (24, True)])
def test_sparse_error_sums(gpu_ctx, readouts, discrete):
    net = synth.random_network(60, 2, 4, seed=9, rates="loguniform", istate="uniform", readouts=readouts,
                               max_time=8.0, time_tick=1.0, sample_count=3000)
    cfg = net.cfg() + ("discrete_time = 1;\n" if discrete else "")
    _check(gpu_ctx, net.bnd(), cfg, count=3000)


@pytest.mark.parametrize("state_hbm", ["1", "3"])
def test_lane_state_in_hbm_error_sums(gpu_ctx, monkeypatch, state_hbm):
    """The error sums with the lanes' state in HBM (1) or x alone in LDS with the active sets and
    word sums in HBM (3), the layouts of wide models (each lane's window list is in HBM too)."""
    monkeypatch.setenv("MBG_TUNE_STATE_HBM", state_hbm)
    monkeypatch.setenv("MBG_TUNE_WSUM_REG", "2")
    net = synth.random_network(300, 1, 6, 7, rates="loguniform", istate="uniform", readouts=12,
                               special_frac=0.2, max_time=20.0, time_tick=0.5, sample_count=3000)
    _check(gpu_ctx, net.bnd(), net.cfg(), count=3000)
    assert gpu_ctx.info().state_hbm == int(state_hbm)


def test_logic_programs_error_sums(gpu_ctx):
    b, c = synth.ring_network(100, steps_to_fixed_point=50, sample_count=2000, time_tick=5.0, max_time=100.0)
    _check(gpu_ctx, b, c, count=2000)


def test_sweep_error_sums(gpu_ctx):
    specs = []
    for k, (n, count) in enumerate([(50, 800), (300, 400), (700, 300), (64, 700)]):
        net = synth.random_network(n, 2, 4, seed=200 + k, rates="loguniform", istate="uniform", readouts=16,
                                   max_time=20.0, time_tick=2.0, sample_count=count)
        specs.append((net.bnd(), net.cfg(), count))
    b, c = model_text("cellcycle")
    specs.append((b, c, 3000))
    models = [Model(b, c) for b, c, _ in specs]
    _flush(f"case: error columns, sweep of {len(models)} variants")
    gpu_ctx.run_variants([(m, cnt, 5 + v, v) for v, (m, (_, _, cnt)) in enumerate(zip(models, specs))],
                         error_columns=True)
    for v, (b, c, cnt) in enumerate(specs):
        r = OracleModel(b, c).run(count=cnt, seed=5 + v, variant_id=v, threads=16)
        i, prob, th, _, _, _ = gpu_ctx.variant_result(v)
        assert np.array_equal(prob, r.prob) and np.array_equal(th, r.th), v
        psq, tsq = gpu_ctx.variant_squares(v)
        assert i.sq_exponent == r.sq_exponent and i.th_sq_exponent == r.th_sq_exponent
        assert np.array_equal(psq, r.prob_sq) and np.array_equal(tsq, r.th_sq), v

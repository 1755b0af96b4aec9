"""GPU trajectory kernel vs the CPU oracle, through the C ABI (libmbg.so).

Bit-exact: given the same Philox draws, each trajectory's transition count, FNV-1a hash of
the flipped node sequence, final time and status match the oracle's replay, and the window
tables, final-state counts and fixed-point multisets are identical integers.  At full sizes
(1,048,576 trajectories of the 1,000-node network) the checks are size-independent properties:
windows sum to one, counts add up, results do not depend on block size or grid.
"""
import math

import numpy as np
import pytest

from oracle.oracle import OracleModel
from tests.conftest import REFERENCE_MODELS, model_text
from upmabossg_amd import Model, synth

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _compare(ctx, bnd, cfg, count, n=None, first_id=0, seed=None, threads=16, summaries=True, block=0):
    n = n or count
    m = Model(bnd, cfg)
    o = OracleModel(bnd, cfg)
    _flush(f"case: N={m.num_nodes} count={count} first_id={first_id} block={block}")
    g = ctx.run(m, n_trajectories=n, count=count, first_id=first_id, seed=seed, trace=summaries, block=block)
    r = o.run(count=count, first_id=first_id, n_trajectories=n, seed=seed, threads=threads,
              record_summaries=summaries)
    if summaries:
        s = r.summaries
        assert np.array_equal(g.summaries[:, 0], s["flips"]), "transition counts differ"
        assert np.array_equal(g.summaries[:, 1], s["flip_hash"]), "flip sequences differ"
        assert np.array_equal(g.summaries[:, 2], s["final_time"].view(np.uint64)), "final times differ"
        assert np.array_equal(g.summaries[:, 3], s["status"].astype(np.uint64)), "statuses differ"
    assert g.transitions == r.flips
    assert g.prob_exponent == r.prob_exponent and g.th_exponent == r.th_exponent
    if r.readouts > 20:  # sparse cells
        for a, b in zip(g.cells, r.cells):
            assert np.array_equal(a, b), "window cells differ"
        for a, b in zip(g.final_cells, r.final_cells):
            assert np.array_equal(a, b), "final states differ"
    else:
        assert np.array_equal(g.prob_raw, r.prob), "window tables differ"
        assert np.array_equal(g.final_states, r.final), "final states differ"
    assert np.array_equal(g.th_raw, r.th), "window entropies differ"
    assert np.array_equal(g.fixed_states, r.fixed_states), "fixed points differ"
    assert np.array_equal(g.fixed_counts, r.fixed_counts), "fixed-point counts differ"
    return g, r


def test_small_network_bit_exact(gpu_ctx):
    net = synth.config_small()
    g, r = _compare(gpu_ctx, net.bnd(), net.cfg(), count=10000)
    assert g.transitions > 100000
    # readout marginals reduced on the device = the window table summed over states with each readout ON
    K = r.readouts
    bits = (np.arange(1 << K)[:, None] >> np.arange(K)[None, :]) & 1
    raw = np.array(r.prob.astype(object) @ bits.astype(object), dtype=np.float64)
    want = raw / (10000 * float(np.float32(net.time_tick)) * 2.0 ** r.prob_exponent)
    assert np.allclose(gpu_ctx.marginals(), want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("name,count", [("cellcycle", 20000), ("metastasis", 1000), ("sizek", 2000),
                                        ("Montagud2022_Prostate_Cancer", 2000)])
def test_reference_models_bit_exact(gpu_ctx, name, count):
    b, c = model_text(name)
    _compare(gpu_ctx, b, c, count=count)


def test_large_network_bit_exact(gpu_ctx):
    net = config = synth.config_large()
    _compare(gpu_ctx, net.bnd(), net.cfg(), count=2048, n=1048576)
    # a slice from the middle of the id range
    _compare(gpu_ctx, config.bnd(), config.cfg(), count=512, n=1048576, first_id=700000)


def test_fixed_point_heavy_bit_exact(gpu_ctx):
    net = synth.config_fixed_heavy()
    g, r = _compare(gpu_ctx, net.bnd(), net.cfg(), count=256, n=262144)


@pytest.mark.parametrize("n", [1100, 1300, 2000])
def test_mid_size_networks_bit_exact(gpu_ctx, n):
    """17-32 state words: register rate sums in groups of two words (even word counts) or LDS sums."""
    net = synth.random_network(n, 2, 4, seed=n, rates="loguniform", istate="uniform", readouts=16, max_time=20.0,
                               time_tick=0.5, sample_count=256)
    _compare(gpu_ctx, net.bnd(), net.cfg(), count=256)


def test_batch_resume_compat_bit_exact(gpu_ctx):
    """The reference's 100-record launches (no flip after every 99th step) on the GPU and in
    the oracle: identical; and different from uninterrupted Gillespie."""
    net = synth.random_network(60, 2, 3, 5, rates="loguniform", istate="uniform", readouts=8, max_time=30.0,
                               time_tick=1.0, sample_count=2000)
    m, o = Model(net.bnd(), net.cfg()), OracleModel(net.bnd(), net.cfg())
    print("case: compat_batch_resume 60 nodes x 2000", flush=True)
    g = gpu_ctx.run(m, trace=True, compat_batch_resume=True)
    r = o.run(count=2000, threads=16, record_summaries=True, compat_batch_resume=True)
    assert np.array_equal(g.summaries[:, 0], r.summaries["flips"])
    assert np.array_equal(g.summaries[:, 1], r.summaries["flip_hash"])
    assert np.array_equal(g.prob_raw, r.prob) and np.array_equal(g.th_raw, r.th)
    assert np.array_equal(g.final_states, r.final)
    plain = gpu_ctx.run(m, trace=True)
    assert not np.array_equal(plain.summaries[:, 1], g.summaries[:, 1])


def test_autotuned_shape_gives_identical_results(gpu_ctx):
    """mbg_autotune measures candidate launch shapes; whichever it keeps, the results are the
    same integers as with the default shape."""
    from upmabossg_amd import Context

    net = synth.config_large()
    m = Model(net.bnd(), net.cfg())
    print("case: autotune 1000 nodes", flush=True)
    with Context(0) as ctx:
        ref = ctx.run(m, count=8192, n_trajectories=8192, trace=True)
        info = ctx.autotune(m, sample=32768, trace=True)
        assert info["trials"] >= 4 and info["best_ms"] <= info["default_ms"]
        tuned = ctx.run(m, count=8192, n_trajectories=8192, trace=True)
        assert tuned.transitions == ref.transitions
        assert np.array_equal(tuned.summaries, ref.summaries)
        assert np.array_equal(tuned.prob_raw, ref.prob_raw) and np.array_equal(tuned.final_states, ref.final_states)


@pytest.mark.parametrize("n", [40, 80])
def test_fixed_points_deduplicated_on_device_bit_exact(gpu_ctx, n):
    """20,000 trajectories that almost all end at a fixed point (285 and 3,330 distinct
    states): more than 4,096 records, so the device reduces them to distinct states; the
    multiset equals the oracle's."""
    net = synth.random_network(n, 1, 3, 21, rates="loguniform", istate="uniform", readouts=8, special_frac=0.3,
                               max_time=100.0, time_tick=2.0, sample_count=20000)
    g, r = _compare(gpu_ctx, net.bnd(), net.cfg(), count=20000, summaries=False)
    assert g.fixed_point_trajectories > 4096 and len(g.fixed_counts) == len(r.fixed_counts)


def test_dependent_batches_refused(gpu_ctx, monkeypatch):
    """More than one dependent per round of table reads gave wrong, varying transition counts
    on the GPU in several launch shapes (the kernel source is exact in the CPU emulation); the
    runtime refuses such a shape with a clear error instead of running it."""
    from upmabossg_amd import MbgError

    monkeypatch.setenv("MBG_TUNE_DEP_BATCH", "2")
    net = synth.config_small()
    with pytest.raises(MbgError) as e:
        gpu_ctx.run(Model(net.bnd(), net.cfg()), count=100)
    assert e.value.status == 2 and "one dependent per round" in str(e.value)


@pytest.mark.parametrize("cap", ["1", "2", "0"])
def test_dependents_cap_bit_exact(gpu_ctx, monkeypatch, cap):
    """A cap on the dependents re-evaluated per loop iteration only reschedules the work (a lane
    with more finishes them in later iterations before its next step): with every dependent
    deferred (1), two per iteration, or no cap, the results are the oracle's.  The cap is read
    when a launch shape is built, so each value compiles its own kernel."""
    monkeypatch.setenv("MBG_TUNE_DEP_CAP", cap)
    net = synth.random_network(300, 1, 6, 5, rates="loguniform", istate="uniform", readouts=12, special_frac=0.2,
                               max_time=20.0, time_tick=0.5, sample_count=3000)
    _compare(gpu_ctx, net.bnd(), net.cfg(), count=3000)
    b, c = model_text("sizek")
    _compare(gpu_ctx, b, c, count=1000)


@pytest.mark.parametrize("case", ["random", "sparse", "ring", "discrete", "wide_state"])
def test_lane_state_in_hbm_bit_exact(gpu_ctx, monkeypatch, case):
    """The lanes' state (x, act, word sums) in an HBM buffer instead of LDS (MBG_STATE_HBM, the
    default above one wave's worth of a CU's LDS) is a layout only: forced on for models that
    would keep it in LDS (register word sums, a sparse window table, logic programs, discrete
    time) and left to the default on a 79-word model, the results are the oracle's."""
    monkeypatch.setenv("MBG_TUNE_STATE_HBM", "0" if case == "wide_state" else "1")
    if case == "random":
        net = synth.random_network(300, 1, 6, 7, rates="loguniform", istate="uniform", readouts=12,
                                   special_frac=0.2, max_time=20.0, time_tick=0.5, sample_count=3000)
        _compare(gpu_ctx, net.bnd(), net.cfg(), count=3000)
    elif case == "sparse":
        net = synth.random_network(120, 2, 3, 9, rates="loguniform", istate="uniform", readouts=30,
                                   max_time=10.0, time_tick=0.5, sample_count=2000)
        _compare(gpu_ctx, net.bnd(), net.cfg(), count=2000)
    elif case == "ring":
        b, c = synth.ring_network(100, steps_to_fixed_point=50, sample_count=2000, time_tick=5.0, max_time=100.0)
        _compare(gpu_ctx, b, c, count=2000)
    elif case == "discrete":
        b, c = model_text("cellcycle")
        _compare(gpu_ctx, b, c, count=5000)
    else:
        monkeypatch.delenv("MBG_TUNE_STATE_HBM")
        net = synth.config_fixed_heavy()
        _compare(gpu_ctx, net.bnd(), net.cfg(), count=512, n=262144)


@pytest.mark.parametrize("case", ["random", "sparse", "ring", "discrete", "grid"])
def test_active_sets_in_hbm_bit_exact(gpu_ctx, monkeypatch, case):
    """Only the active sets in HBM (x in LDS, word sums in VGPRs: MBG_STATE_HBM 2, half the LDS
    per lane) is a layout only: on random, sparse-table, logic-program and discrete models and on
    a 3-block grid, the results are the oracle's."""
    monkeypatch.setenv("MBG_TUNE_STATE_HBM", "2")
    if case in ("random", "grid"):
        net = synth.random_network(300, 1, 6, 7, rates="loguniform", istate="uniform", readouts=12,
                                   special_frac=0.2, max_time=20.0, time_tick=0.5, sample_count=3000)
        if case == "grid":
            m = Model(net.bnd(), net.cfg())
            a = gpu_ctx.run(m, count=3000, trace=True, block=128, grid=3)
            assert gpu_ctx.info().state_hbm == 2
            r = OracleModel(net.bnd(), net.cfg()).run(count=3000, threads=8, record_summaries=True)
            assert np.array_equal(a.summaries[:, 1], r.summaries["flip_hash"])
            assert np.array_equal(a.prob_raw, r.prob)
            return
        _compare(gpu_ctx, net.bnd(), net.cfg(), count=3000)
    elif case == "sparse":
        net = synth.random_network(120, 2, 3, 9, rates="loguniform", istate="uniform", readouts=30,
                                   max_time=10.0, time_tick=0.5, sample_count=2000)
        _compare(gpu_ctx, net.bnd(), net.cfg(), count=2000)
    elif case == "ring":
        b, c = synth.ring_network(100, steps_to_fixed_point=50, sample_count=2000, time_tick=5.0, max_time=100.0)
        _compare(gpu_ctx, b, c, count=2000)
    else:
        b, c = model_text("cellcycle")
        _compare(gpu_ctx, b, c, count=5000)
    assert gpu_ctx.info().state_hbm == 2


@pytest.mark.parametrize("case", ["random", "sparse", "ring", "discrete", "grid", "wide"])
def test_x_in_lds_active_sets_and_sums_in_hbm_bit_exact(gpu_ctx, monkeypatch, case):
    """x alone in LDS, the active sets and word sums in HBM, group sums in VGPRs (MBG_STATE_HBM
    3, two-level sums) is a layout only: on random, sparse-table, logic-program and discrete
    models, on a 3-block grid and on the 79-word fixed-point-heavy network, the results are the
    oracle's."""
    monkeypatch.setenv("MBG_TUNE_STATE_HBM", "3")
    monkeypatch.setenv("MBG_TUNE_WSUM_REG", "2")
    if case in ("random", "grid"):
        net = synth.random_network(300, 1, 6, 7, rates="loguniform", istate="uniform", readouts=12,
                                   special_frac=0.2, max_time=20.0, time_tick=0.5, sample_count=3000)
        if case == "grid":
            m = Model(net.bnd(), net.cfg())
            a = gpu_ctx.run(m, count=3000, trace=True, block=128, grid=3)
            assert gpu_ctx.info().state_hbm == 3
            r = OracleModel(net.bnd(), net.cfg()).run(count=3000, threads=8, record_summaries=True)
            assert np.array_equal(a.summaries[:, 1], r.summaries["flip_hash"])
            assert np.array_equal(a.prob_raw, r.prob)
            return
        _compare(gpu_ctx, net.bnd(), net.cfg(), count=3000)
    elif case == "sparse":
        net = synth.random_network(120, 2, 3, 9, rates="loguniform", istate="uniform", readouts=30,
                                   max_time=10.0, time_tick=0.5, sample_count=2000)
        _compare(gpu_ctx, net.bnd(), net.cfg(), count=2000)
    elif case == "ring":
        b, c = synth.ring_network(100, steps_to_fixed_point=50, sample_count=2000, time_tick=5.0, max_time=100.0)
        _compare(gpu_ctx, b, c, count=2000)
    elif case == "discrete":
        b, c = model_text("cellcycle")
        _compare(gpu_ctx, b, c, count=5000)
    else:
        net = synth.config_fixed_heavy()
        _compare(gpu_ctx, net.bnd(), net.cfg(), count=512, n=262144)
    assert gpu_ctx.info().state_hbm == 3


@pytest.mark.parametrize("case", ["random", "large"])
def test_generated_logic_bit_exact(gpu_ctx, monkeypatch, case):
    """The dependents' logic from code generated per node (MBG_TUNE_GENCODE, a switch on the node
    id with each node's logic expression written out) instead of the truth tables: the same
    results as the oracle."""
    monkeypatch.setenv("MBG_TUNE_GENCODE", "1")
    if case == "random":
        net = synth.random_network(300, 1, 6, 7, rates="loguniform", istate="uniform", readouts=12,
                                   special_frac=0.2, max_time=20.0, time_tick=0.5, sample_count=3000)
        count = 3000
    else:
        net = synth.config_large()
        count = 2048
    assert "#define MBG_GENCODE 1" in Model(net.bnd(), net.cfg()).generate_source()
    _compare(gpu_ctx, net.bnd(), net.cfg(), count=count, n=net.sample_count)


def test_more_than_128_words_bit_exact(gpu_ctx):
    """8,300 nodes (130 words): the initial states are drawn per (trajectory, word), the readout
    part of each staged state by its own pass, and the lanes' state (3,120 bytes per lane, more
    than LDS holds for a wave) lives in HBM.  A short run keeps the O(N)-per-step oracle quick."""
    net = synth.random_network(8300, 1, 3, 31, rates="loguniform", istate="uniform", readouts=10, special_frac=0.2,
                               max_time=0.5, time_tick=0.125, sample_count=64)
    g, r = _compare(gpu_ctx, net.bnd(), net.cfg(), count=64)
    assert gpu_ctx.info().state_words == 130 and g.transitions > 10000


def test_block_size_and_grid_independent(gpu_ctx):
    net = synth.config_small()
    m = Model(net.bnd(), net.cfg())
    a = gpu_ctx.run(m, count=5000, trace=True, block=64)
    b = gpu_ctx.run(m, count=5000, trace=True, block=256, grid=3)
    assert np.array_equal(a.summaries, b.summaries)
    assert np.array_equal(a.prob_raw, b.prob_raw) and np.array_equal(a.th_raw, b.th_raw)


def test_variant_multipliers_and_knockout(gpu_ctx):
    net = synth.config_large()
    rng = np.random.default_rng(3)
    mu = np.exp(rng.normal(0, 0.5, net.n)).astype(np.float32)
    md = np.exp(rng.normal(0, 0.5, net.n)).astype(np.float32)
    ist = net.istate.copy()
    for k in (5, 17, 400):  # knockouts: forced OFF
        mu[k] = 0.0
        ist[k] = 0.0
    m = Model(net.bnd(), net.cfg())
    m.set_variant(mu, md, ist)
    o = OracleModel(net.bnd(), net.cfg())
    o.set_variant(mu, md, ist)
    g = gpu_ctx.run(m, count=1024, n_trajectories=1024, variant_id=7, trace=True)
    r = o.run(count=1024, variant_id=7, threads=16, record_summaries=True)
    assert np.array_equal(g.summaries[:, 1], r.summaries["flip_hash"])
    assert np.array_equal(g.prob_raw, r.prob) and np.array_equal(g.final_states, r.final)


EDGE_MODELS = {
    # immediate fixed point: every node already agrees with its logic
    "frozen": ("node A { logic = A; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n"
               "node B { logic = !A; rate_up = @logic ? 2 : 0; rate_down = @logic ? 0 : 2; }\n",
               "A.istate = 1; B.istate = 0; max_time = 5; time_tick = 1; sample_count = 100;"),
    # generic (non fast-path) rate expressions, XOR, arithmetic, a self-loop
    "generic": ("node A { logic = A ^ B; rate_up = B ? 2.5 * $k : 0.5; rate_down = @logic ? 0 : 1 + C; }\n"
                "node B { logic = A | C; rate_up = @logic ? $k : 0; rate_down = @logic ? 0 : 3; }\n"
                "node C { logic = !C; rate_up = A + B; rate_down = 1.25; }\n",
                "$k = 0.75; C.is_internal = 1; max_time = 7.5; time_tick = 0.35; sample_count = 3000;"),
    # no readouts at all (every node internal), max_time shorter than the tick
    "internal_only": ("node A { logic = !B; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n"
                      "node B { logic = A; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n",
                      "A.is_internal = 1; B.is_internal = 1; max_time = 0.5; time_tick = 2; sample_count = 500;"),
    # discrete time with a fixed point reached mid-run
    "discrete_fixed": ("node A { logic = A | B; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n"
                       "node B { logic = !A; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n",
                       "discrete_time = 1; max_time = 3; time_tick = 0.3; sample_count = 2000;"),
}


def _wide_fanin_model():
    # node Z reads 18 inputs (truth table of 2^19 keys: the wide table layout); the inputs toggle
    names = [f"x{i}" for i in range(18)]
    bnd = "".join(f"node {n} {{ logic = !{n}; rate_up = @logic ? {0.5 + 0.1 * i:.1f} : 0; "
                  f"rate_down = @logic ? 0 : {1.5 - 0.05 * i:.2f}; }}\n" for i, n in enumerate(names))
    expr = "((" + " & ".join(names[:6]) + ") | (" + " & ".join(names[6:12]) + ")) ^ (" + " | ".join(names[12:]) + ")"
    bnd += f"node Z {{ logic = {expr}; rate_up = @logic ? 2 : 0; rate_down = @logic ? 0 : 3; }}\n"
    cfg = "".join(f"{n}.is_internal = 1;\n" for n in names[4:]) + \
        "max_time = 4; time_tick = 0.25; sample_count = 3000; seed_pseudorandom = 5;"
    return bnd, cfg


def _twenty_readouts_model():
    # 20 non-internal nodes: 2^20 projected states per window
    bnd = "".join(f"node n{i} {{ logic = n{(i + 1) % 20} | !n{(i + 7) % 20}; rate_up = @logic ? 1.{i % 10} : 0; "
                  f"rate_down = @logic ? 0 : 0.{i % 9 + 1}; }}\n" for i in range(20))
    return bnd, "max_time = 3; time_tick = 0.5; sample_count = 4000;"


EDGE_MODELS["wide_fanin"] = _wide_fanin_model()
EDGE_MODELS["twenty_readouts"] = _twenty_readouts_model()


@pytest.mark.parametrize("name", sorted(EDGE_MODELS))
def test_edge_models_bit_exact(gpu_ctx, name):
    b, c = EDGE_MODELS[name]
    m = Model(b, c)
    n = int(m.constant("sample_count"))
    _compare(gpu_ctx, b, c, count=n)


def test_single_trajectory(gpu_ctx):
    net = synth.config_large()
    _compare(gpu_ctx, net.bnd(), net.cfg(), count=1, n=1)


def test_large_full_size_properties(gpu_ctx):
    """1,048,576 trajectories of the 1,000-node network: every window's probabilities sum to
    one, final states count every trajectory, and two runs give identical tables."""
    net = synth.config_large()
    m = Model(net.bnd(), net.cfg())
    _flush("case: large full size, run 1")
    a = gpu_ctx.run(m)
    assert a.simulated == 1048576
    S = a.prob_raw.sum(axis=1)
    full = np.float64(2.0) ** a.prob_exponent * a.n_trajectories * np.float64(np.float32(0.5))
    assert np.all(np.abs(S.astype(np.float64) - full) <= 1e-9 * full + a.n_trajectories * 300)
    assert int(a.final_states.sum()) == 1048576
    assert np.all(a.th >= 0)
    _flush("case: large full size, run 2")
    b = gpu_ctx.run(m)
    assert np.array_equal(a.prob_raw, b.prob_raw) and np.array_equal(a.th_raw, b.th_raw)
    assert np.array_equal(a.final_states, b.final_states) and a.transitions == b.transitions


def _z_scores(pg, ng, po, no):
    """Two-sample binomial z-scores of probabilities pg (from ng trajectories) and po (from no),
    with the pooled p and a variance floor of one trajectory's weight in the smaller sample, so
    that a state seen in a single trajectory is not a many-sigma event.  For window-averaged
    probabilities (time fractions) the binomial variance is an upper bound."""
    p = (pg * ng + po * no) / (ng + no)
    var = np.maximum(p * (1 - p), 1.0 / min(ng, no)) * (1.0 / ng + 1.0 / no)
    return (pg - po) / np.sqrt(var)


def _family_threshold(m, alpha=1e-3):
    """The per-entry |z| limit: 4 (the 4-sigma rule), raised to the Bonferroni level when m
    entries are tested together (family-wise false-alarm rate alpha), where that is above 4."""
    from scipy.stats import norm

    return max(4.0, float(norm.isf(alpha / (2 * max(1, m)))))


def _stat_report(name, z):
    m = z.size
    thr = _family_threshold(m)
    beyond4 = int(np.sum(np.abs(z) > 4.0))
    _flush(f"{name}: {m} entries, max |z| {np.max(np.abs(z)):.2f}, {beyond4} beyond 4 sigma "
           f"(expected under the null at most {m * 6.334e-5:.2f}), limit {thr:.2f}")
    return thr


def test_large_full_size_statistics_independent_seed(gpu_ctx):
    """1,048,576 GPU trajectories of the 1,000-node network (seed 1) against an independent
    oracle sample (8,192 trajectories, seed 2): every window x projected-state entry where
    either side has mass, every readout marginal and every final state, within 4 sigma (the
    Bonferroni level for the number of entries where that is higher)."""
    net = synth.config_large()
    m = Model(net.bnd(), net.cfg())
    _flush("case: large full size, seed 1, against an oracle sample with seed 2")
    a = gpu_ctx.run(m, seed=1)
    no = 8192
    r = OracleModel(net.bnd(), net.cfg()).run(count=no, n_trajectories=no, seed=2, threads=16, record_fixed=False)
    ng = a.simulated
    tick = float(np.float32(net.time_tick))
    po = r.prob.astype(np.float64) / 2.0 ** r.prob_exponent / (no * tick)
    pg = a.probtraj
    mass = (pg > 0) | (po > 0)
    z = _z_scores(pg[mass], ng, po[mass], no)
    assert np.all(np.abs(z) <= _stat_report("window x state entries", z))
    K = a.readouts
    bits = (np.arange(1 << K)[:, None] >> np.arange(K)[None, :]) & 1
    zm = _z_scores(pg @ bits, ng, po @ bits, no)
    assert np.all(np.abs(zm) <= _stat_report("readout marginals", zm))
    fg, fo = a.final_states / ng, r.final / no
    fm = (fg > 0) | (fo > 0)
    zf = _z_scores(fg[fm], ng, fo[fm], no)
    assert np.all(np.abs(zf) <= _stat_report("final states", zf))


def test_fixed_point_dedup_with_equal_leading_words(gpu_ctx, monkeypatch):
    """Fixed points that agree on their two most significant words (nodes 128-199 frozen OFF
    in a 200-node network, so those words are 0 in every state): the device deduplication
    orders them by the lower words exactly as sorting every record on the host does."""
    net = synth.random_network(200, 1, 3, 21, rates="loguniform", istate="uniform", readouts=10,
                               special_frac=0.3, max_time=200.0, time_tick=2.0, sample_count=20000)
    net.up[128:] = 0.0
    net.down[128:] = 0.0
    net.istate[128:] = 0.0
    m = Model(net.bnd(), net.cfg())
    _flush("case: 200 nodes, frozen leading words, device deduplication")
    a = gpu_ctx.run(m, seed=3)
    assert a.fixed_point_trajectories > 4096 and len(a.fixed_counts) > 100
    assert not a.fixed_states[:, 2:].any()
    _flush("case: 200 nodes, frozen leading words, host deduplication")
    monkeypatch.setenv("MBG_TUNE_FP_HOST", "1")
    h = gpu_ctx.run(m, seed=3)
    monkeypatch.delenv("MBG_TUNE_FP_HOST")
    assert np.array_equal(a.fixed_states, h.fixed_states) and np.array_equal(a.fixed_counts, h.fixed_counts)


def test_fixed_point_heavy_full_size_statistics_and_dedup(gpu_ctx, monkeypatch):
    """262,144 GPU trajectories of the 5,000-node fixed-point-heavy network (seed 1) against an
    independent oracle sample (384 trajectories, seed 2): the fraction reaching a fixed point,
    the final states and the readout marginals within 4 sigma, and every fixed point of the
    sample at its GPU probability within 4 sigma.  The device deduplication of the fixed-point
    records gives the same distinct states and counts as sorting every record on the host."""
    net = synth.config_fixed_heavy()
    m = Model(net.bnd(), net.cfg())
    _flush("case: fixed-point-heavy full size, device deduplication")
    a = gpu_ctx.run(m, seed=1)
    ng = a.simulated
    assert ng == 262144 and int(a.fixed_counts.sum()) == a.fixed_point_trajectories
    _flush("case: fixed-point-heavy full size, host deduplication")
    monkeypatch.setenv("MBG_TUNE_FP_HOST", "1")
    h = gpu_ctx.run(m, seed=1)
    monkeypatch.delenv("MBG_TUNE_FP_HOST")
    assert np.array_equal(a.fixed_states, h.fixed_states) and np.array_equal(a.fixed_counts, h.fixed_counts)
    _flush(f"  {a.fixed_point_trajectories} fixed-point trajectories, {len(a.fixed_counts)} distinct fixed points")
    no = 384
    r = OracleModel(net.bnd(), net.cfg()).run(count=no, n_trajectories=no, seed=2, threads=16)
    fo_frac = float(r.fixed_counts.sum()) / no
    zfp = _z_scores(np.array([a.fixed_point_trajectories / ng]), ng, np.array([fo_frac]), no)
    assert abs(zfp[0]) <= _stat_report("fixed-point fraction", zfp)
    fg, fo = a.final_states / ng, r.final / no
    fm = (fg > 0) | (fo > 0)
    zf = _z_scores(fg[fm], ng, fo[fm], no)
    assert np.all(np.abs(zf) <= _stat_report("final states", zf))
    tick = float(np.float32(net.time_tick))
    K = a.readouts
    bits = (np.arange(1 << K)[:, None] >> np.arange(K)[None, :]) & 1
    po = r.prob.astype(np.float64) / 2.0 ** r.prob_exponent / (no * tick)
    zm = _z_scores(a.probtraj @ bits, ng, po @ bits, no)
    assert np.all(np.abs(zm) <= _stat_report("readout marginals", zm))
    # the sample's fixed points at their GPU probabilities (0 when the GPU run never reached one)
    gpu_fp = {a.fixed_states[i].tobytes(): int(a.fixed_counts[i]) for i in range(len(a.fixed_counts))}
    pgf = np.array([gpu_fp.get(st.tobytes(), 0) / ng for st in r.fixed_states])
    pof = r.fixed_counts.astype(np.float64) / no
    if len(pof):
        zz = _z_scores(pgf, ng, pof, no)
        assert np.all(np.abs(zz) <= _stat_report("fixed points of the sample", zz))


@pytest.mark.parametrize("n,readouts,discrete,count", [(48, 24, False, 4000), (300, 40, False, 2000),
                                                      (64, 58, True, 3000), (2000, 32, False, 256)])
def test_more_than_20_readouts_bit_exact(gpu_ctx, n, readouts, discrete, count):
    """Sparse window cells (device hash table keyed by window and projected state): the cells,
    final states, entropies and fixed points equal the oracle's, for 24 to 58 readouts (58 + 5
    bits of window index fill the 63-bit key); the first two cases hold more cells than the
    table they start from (2^16 slots), so the run is repeated with a larger one."""
    net = synth.random_network(n, 2, 3, seed=n + readouts, rates="loguniform", istate="uniform", readouts=readouts,
                               max_time=10.0, time_tick=0.5, sample_count=count)
    net.discrete = discrete
    g, r = _compare(gpu_ctx, net.bnd(), net.cfg(), count=count)
    assert g.prob_raw is None and g.readouts == readouts
    w, st, v = g.cells
    if n <= 300 and not discrete:
        assert len(v) > 2 ** 16  # more cells than the starting table: the growth path ran
    # the readout marginals of the sparse cells
    K = readouts
    raw = np.zeros((g.windows, K))
    for j in range(K):
        on = ((st >> np.uint64(j)) & np.uint64(1)).astype(bool)
        np.add.at(raw[:, j], w[on].astype(np.int64), v[on].astype(np.float64))
    want = raw / count if discrete else raw / (count * float(np.float32(net.time_tick)) * 2.0 ** g.prob_exponent)
    assert np.allclose(gpu_ctx.marginals(), want, rtol=1e-9, atol=0)


@pytest.mark.parametrize("n,steps,discrete,count", [(100, 0, False, 2000), (100, 50, False, 2000),
                                                   (100, 20, True, 2000), (1000, 200, False, 256)])
def test_ring_models_of_the_reference_generator_bit_exact(gpu_ctx, n, steps, discrete, count):
    """The ring models of data/generate-synth.py: every ring node reads 2 sqrt(N) + 1 other
    nodes (21 at 100 nodes, 63 at 1,000), so its logic runs as a postfix program."""
    b, c = synth.ring_network(n, steps_to_fixed_point=steps, sample_count=count, discrete=discrete,
                              time_tick=5.0, max_time=100.0)
    g, r = _compare(gpu_ctx, b, c, count=count)
    assert g.transitions > 10 * count
    if steps and not discrete:  # (20 discrete steps rarely pick NFP among the active ring nodes)
        assert g.fixed_point_trajectories > 0


@pytest.mark.parametrize("seed,discrete", [(1, False), (2, False), (3, True), (4, False)])
def test_wide_logic_programs_bit_exact(gpu_ctx, seed, discrete):
    """Nodes reading 21-40 nodes through random trees of & | ^ ! ?:, constants, $variables,
    @aliases and unary minus, some reading themselves (tests/wide_models.py)."""
    from tests.wide_models import wide_model

    b, c = wide_model(seed=seed, discrete=discrete)
    _compare(gpu_ctx, b, c, count=2000)


@pytest.mark.parametrize("discrete,compat", [(False, False), (True, False), (False, True)])
def test_all_extensions_together_bit_exact(gpu_ctx, discrete, compat):
    """One model with every extension at once: 24 readouts (sparse cells), wide nodes (logic
    programs), a generic rate expression, XOR, variant multipliers and an istate override,
    in continuous and discrete time and with the batch-resume option."""
    from tests.wide_models import wide_model

    b, c = wide_model(n=48, wide=8, seed=9, discrete=discrete, sample_count=1500, max_time=12.0, time_tick=0.5)
    b += ("node gx { logic = w0 ^ w1; rate_up = (w2 & !w3) ? 1.5 : 0.25; rate_down = w4 ? 0.5 : 2; }\n")
    for i in range(6, 24):  # nodes 0-23 and gx are readouts
        c = c.replace(f"w{i}.is_internal = 1;", "")
    m, o = Model(b, c), OracleModel(b, c)
    assert m.num_readouts > 20
    N = m.num_nodes
    rng = np.random.default_rng(4)
    mu = np.exp(rng.normal(0, 0.4, N)).astype(np.float32)
    md = np.exp(rng.normal(0, 0.4, N)).astype(np.float32)
    ist = rng.uniform(0, 1, N).astype(np.float32)
    m.set_variant(mu, md, ist)
    o.set_variant(mu, md, ist)
    _flush(f"case: all extensions discrete={discrete} compat={compat}")
    g = gpu_ctx.run(m, n_trajectories=1500, count=1500, trace=True, compat_batch_resume=compat)
    r = o.run(count=1500, n_trajectories=1500, threads=16, record_summaries=True, compat_batch_resume=compat)
    assert np.array_equal(g.summaries[:, 0], r.summaries["flips"])
    assert np.array_equal(g.summaries[:, 1], r.summaries["flip_hash"])
    assert g.transitions == r.flips
    for a, bb in zip(g.cells, r.cells):
        assert np.array_equal(a, bb)
    for a, bb in zip(g.final_cells, r.final_cells):
        assert np.array_equal(a, bb)
    assert np.array_equal(g.th_raw, r.th)
    assert np.array_equal(g.fixed_states, r.fixed_states) and np.array_equal(g.fixed_counts, r.fixed_counts)

"""Pins the CPU oracle before it is trusted as the GPU kernel's checker (CPU only).

* Philox4x32-10: the Random123 known-answer vectors (Salmon et al., SC'11).
* log: math.log over a sweep of magnitudes.
* Model facts of the reference's four models: counts, names, constants, istate, is_internal.
* Node rates: brute force over all 2^N states against the fast-path closed form
  (src/generator.cpp:93-107) evaluated by Python.
* Step loop + statistics: exact CTMC (matrix exponential) on tiny networks at 4 binomial sigma,
  exact jump chain in discrete time, and the reference's own test/cellcycle.res windows at 4 sigma.
* A planted mistake in the oracle (wrong selection rule) is caught by the exact-CTMC check.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import ctmc
from tests.conftest import REFERENCE_MODELS, model_text
from upmabossg_amd import synth


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32_10
    assert O.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert O.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert O.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_log_accuracy():
    L = O.lib()
    rng = np.random.default_rng(0)
    xs = np.concatenate([np.logspace(-300, 300, 3001), rng.uniform(0.5, 2.0, 20000), [1.0, 2.0, 0.5, 5e-324]])
    worst = 0.0
    for x in xs:
        ref = math.log(x)
        got = L.orc_ln(float(x))
        if ref == 0.0:
            assert got == 0.0
            continue
        worst = max(worst, abs(got - ref) / (abs(ref) * 2.220446049250313e-16))
    assert worst <= 5.0
    # -ln((v + 1) 2^-32) endpoints
    assert L.orc_neglog_u32(0xFFFFFFFF) == 0.0
    assert abs(L.orc_neglog_u32(0) - 32 * math.log(2)) < 1e-14


GOLDEN_FACTS = {
    # nodes, non-internal, discrete_time, sample_count, max_time, time_tick, seed
    "cellcycle": (10, 4, 1, 1000000, 5.0, 0.2, 100),
    "metastasis": (32, 5, 0, 1000, 100.0, 0.5, 0),
    "sizek": (87, 4, 0, 100000, 100.0, 5.0, 0),
    "Montagud2022_Prostate_Cancer": (133, 3, 0, 1000000, 50.0, 0.1, 100),
}


@pytest.mark.parametrize("name", REFERENCE_MODELS)
def test_reference_model_facts(name):
    m = O.OracleModel(*model_text(name))
    n, k, disc, sc, mt, tt, seed = GOLDEN_FACTS[name]
    assert m.n == n
    assert int((~m.internal()).sum()) == k
    assert m.constant("discrete_time") == disc
    assert m.constant("sample_count") == sc
    assert m.constant("max_time") == np.float32(mt)
    assert m.constant("time_tick") == np.float32(tt)
    assert m.constant("seed_pseudorandom") == seed
    assert m.fast().all()  # all four use "@logic ? $u : 0" / "@logic ? 0 : $d"


def test_cellcycle_details():
    m = O.OracleModel(*model_text("cellcycle"))
    assert m.names() == ["CycD", "CycE", "CycA", "CycB", "Rb", "E2F", "p27", "Cdc20", "UbcH10", "cdh1"]
    assert list(m.istate()) == [1, 0, 0, 0, 0, 0, 0, 0, 0, 1]
    assert list(m.internal()) == [False] * 4 + [True] * 6
    assert m.variable("$fast") == 10 and m.variable("$slow") == 1


def test_rates_brute_force_small():
    net = synth.config_small()
    m = O.OracleModel(net.bnd(), net.cfg())
    states = list(range(0, 1 << net.n, 97))
    rates = ctmc.state_rates(net, states)
    for k, s in enumerate(states):
        bits = [(s >> i) & 1 for i in range(net.n)]
        for i in range(net.n):
            assert m.rate_f32(i, bits) == rates[k, i]


def test_rates_brute_force_cellcycle():
    b, c = model_text("cellcycle")
    m = O.OracleModel(b, c)
    # cellcycle logic rewritten in Python from data/cellcycle.bnd
    def logic(x):
        CycD, CycE, CycA, CycB, Rb, E2F, p27, Cdc20, UbcH10, cdh1 = [bool(v) for v in x]
        return [CycD, (not Rb) and E2F,
                (not Rb) and (not Cdc20) and not (UbcH10 and cdh1) and (CycA or E2F),
                (not Cdc20) and (not cdh1),
                (not CycD) and (not CycB) and (p27 or not (CycA or CycE)),
                (not Rb) and (not CycB) and ((not CycA) or p27),
                (not CycD) and (not CycB) and ((not (CycA or CycE)) or (p27 and not (CycE and CycA))),
                CycB,
                ((not (cdh1 and not UbcH10)) and (CycA or CycB)) or ((not CycA) and (not CycB) and ((not cdh1) or (Cdc20 and UbcH10))),
                Cdc20 or ((not CycB) and ((not CycA) or p27))]
    fast, slow = 10.0, 1.0
    up = [fast, slow, slow, slow, fast, slow, fast, slow, slow, fast]
    down = [fast] * 10
    for s in range(1 << 10):
        x = [(s >> i) & 1 for i in range(10)]
        lg = logic(x)
        for i in range(10):
            want = 0.0 if bool(x[i]) == lg[i] else (down[i] if x[i] else up[i])
            assert m.rate_f32(i, x) == want


def _tiny_networks():
    nets = []
    a = synth.random_network(4, 1, 2, seed=11, rates="loguniform", istate="uniform", readouts=3,
                             max_time=4.0, time_tick=0.5, sample_count=40000)
    nets.append(a)
    b = synth.random_network(5, 1, 3, seed=5, rates="loguniform", istate="uniform", readouts=5,
                             special_frac=0.4, max_time=3.0, time_tick=0.25, sample_count=40000)
    nets.append(b)
    c = synth.random_network(6, 2, 2, seed=3, rates="unit", istate="half", readouts=2,
                             max_time=2.0, time_tick=0.3, sample_count=40000)
    nets.append(c)
    return nets


@pytest.mark.parametrize("k", range(3))
def test_oracle_matches_exact_ctmc(k):
    net = _tiny_networks()[k]
    m = O.OracleModel(net.bnd(), net.cfg())
    n = net.sample_count
    r = m.run(count=n, threads=8)
    probs, th, final, fixed = ctmc.continuous_solution(net, r.windows)
    got = r.prob.astype(np.float64) / 2.0 ** r.prob_exponent / (n * float(np.float32(net.time_tick)))
    assert np.all(ctmc.within_sigmas(got, probs, n)), np.max(np.abs(got - probs))
    # each full window's probabilities sum to one
    full = [w for w in range(r.windows) if (w + 1) * float(np.float32(net.time_tick)) <= float(np.float32(net.max_time))]
    assert np.allclose(got.sum(axis=1)[full], 1.0, atol=1e-9)
    gth = r.th.astype(np.float64) / 2.0 ** r.th_exponent / n / float(np.float32(net.time_tick))
    kmax = max(1.0, math.log2(max(2, len(net.readouts))))
    assert np.all(np.abs(gth - th) <= 4 * kmax / 2 / math.sqrt(n) + 1e-12)
    assert np.all(ctmc.within_sigmas(r.final / n, final, n))
    assert int(r.final.sum()) == n
    # fixed points: full states and their probabilities
    got_fp = {int(sum(int(w) << (64 * i) for i, w in enumerate(st))): int(c) / n
              for st, c in zip(r.fixed_states, r.fixed_counts)}
    for s, p in fixed.items():
        assert abs(got_fp.get(s, 0.0) - p) <= 4 * ctmc.sigma(p, n) + 1e-12
    for s in got_fp:
        assert s in fixed


def test_planted_selection_error_is_caught():
    """Swapping up/down rates in the 'oracle' (simulated by a variant with the two multiplier
    vectors exchanged for a mis-wired rate table) must fail the exact-CTMC check."""
    net = _tiny_networks()[0]
    m = O.OracleModel(net.bnd(), net.cfg())
    ratio_up = (net.down / net.up).astype(np.float32)
    ratio_down = (net.up / net.down).astype(np.float32)
    m.set_variant(mult_up=ratio_up, mult_down=ratio_down)
    n = net.sample_count
    r = m.run(count=n, threads=8)
    probs, _, _, _ = ctmc.continuous_solution(net, r.windows)
    got = r.prob.astype(np.float64) / 2.0 ** r.prob_exponent / (n * float(np.float32(net.time_tick)))
    assert not np.all(ctmc.within_sigmas(got, probs, n))


@pytest.mark.parametrize("k", range(2))
def test_variant_overrides_match_exact_ctmc_of_the_rewritten_network(k):
    """orc_set_variant (rate_up / rate_down multipliers, a knockout as rate_up x 0 with istate 0,
    an istate override) gives the exact CTMC solution of the network written out with the
    scaled rates and initial probabilities, and not that of the unmodified network; every
    sweep and knockout parity test of the GPU compares against these overrides."""
    import copy

    net = _tiny_networks()[k]
    rng = np.random.default_rng(20 + k)
    mu = np.exp(rng.uniform(np.log(0.25), np.log(4.0), net.n)).astype(np.float32)
    md = np.exp(rng.uniform(np.log(0.25), np.log(4.0), net.n)).astype(np.float32)
    ist = rng.uniform(0, 1, net.n).astype(np.float32)
    ko = 1  # knockout: never switched on, starts off
    mu[ko] = 0.0
    ist[ko] = 0.0
    m = O.OracleModel(net.bnd(), net.cfg())
    m.set_variant(mult_up=mu, mult_down=md, istate=ist)
    n = net.sample_count
    r = m.run(count=n, threads=8)
    rew = copy.deepcopy(net)
    rew.up = (net.up * mu).astype(np.float32)
    rew.down = (net.down * md).astype(np.float32)
    rew.istate = ist
    probs, _, final, _ = ctmc.continuous_solution(rew, r.windows)
    got = r.prob.astype(np.float64) / 2.0 ** r.prob_exponent / (n * float(np.float32(net.time_tick)))
    assert np.all(ctmc.within_sigmas(got, probs, n)), np.max(np.abs(got - probs))
    assert np.all(ctmc.within_sigmas(r.final / n, final, n))
    # the knocked-out node is off in every final state that projects it
    if ko in net.readouts:
        bit = net.readouts.index(ko)
        assert all(r.final[s] == 0 for s in range(len(r.final)) if (s >> bit) & 1)
    base, _, _, _ = ctmc.continuous_solution(net, r.windows)
    assert not np.all(ctmc.within_sigmas(got, base, n)), "the overrides changed nothing"
    # the rewritten network itself, run without overrides, gives the same tables bit for bit
    # when its rates are the same FP32 products
    r2 = O.OracleModel(rew.bnd(), rew.cfg()).run(count=n, threads=8)
    assert np.array_equal(r2.prob, r.prob) and np.array_equal(r2.final, r.final) and r2.flips == r.flips


@pytest.mark.parametrize("discrete", [False, True])
def test_error_column_sums_equal_brute_force_per_trajectory(discrete):
    """The sums of squares behind ErrorProba / ErrorTH equal sum_i f_i^2 computed in NumPy from
    every trajectory's own tables: each trajectory run alone (count 1: its window table is its
    time per (window, state), revisits within a window included) gives f_i[w][s] and its
    window-average TH; a wrong revisit merge, a dropped square or a wrong scale fails this."""
    net = synth.random_network(6, 1, 3, seed=8, rates="loguniform", istate="uniform", readouts=4,
                               max_time=6.0, time_tick=2.0, sample_count=60)
    cfg = net.cfg() + ("discrete_time = 1;\n" if discrete else "")
    m = O.OracleModel(net.bnd(), cfg)
    n = 60
    r = m.run(count=n, n_trajectories=n, seed=3, threads=4)
    tick = float(np.float32(net.time_tick))
    f2 = np.zeros(r.prob.shape)
    t2 = np.zeros(r.windows)
    revisits = 0
    for i in range(n):
        one = m.run(count=1, first_id=i, n_trajectories=1, seed=3, threads=1)
        f = one.prob.astype(np.float64) / 2.0 ** one.prob_exponent
        t = one.th.astype(np.float64) / 2.0 ** one.th_exponent
        if not discrete:
            f /= tick
            t /= tick
        f2 += f * f
        t2 += t * t
        revisits += int(np.sum((f > 0) & (f < 1)))
    got = r.prob_sq.astype(np.float64) / 2.0 ** r.sq_exponent
    gth = r.th_sq.astype(np.float64) / 2.0 ** r.th_sq_exponent
    assert np.allclose(got, f2, rtol=1e-9, atol=n * 2.0 ** -40)
    assert np.allclose(gth, t2, rtol=1e-9, atol=n * 2.0 ** -40)
    if not discrete:
        assert revisits > 0  # windows shared by several states: partial fractions occur
    # the same sums from the cell list (window-major, states ascending)
    w, st, _ = r.cells
    assert np.array_equal(r.cell_squares, r.prob_sq[w, st])


def test_error_columns_closed_form_for_frozen_states():
    """A network whose nodes never switch: each trajectory keeps its initial projected state, so
    every fraction is 0 or 1, E[f^2] = P and the standard error is the binomial one,
    sqrt(P (1 - P) / (n - 1)); every TH is 0."""
    bnd = ("node A { logic = A; rate_up = @logic ? 1 : 0; rate_down = @logic ? 0 : 1; }\n"
           "node B { logic = B; rate_up = @logic ? 2 : 0; rate_down = @logic ? 0 : 2; }\n")
    cfg = "A.istate = 0.3; B.istate = 0.6; max_time = 3; time_tick = 1; sample_count = 5000;"
    r = O.OracleModel(bnd, cfg).run(count=5000, threads=4)
    n = 5000
    P = r.prob.astype(np.float64) / 2.0 ** r.prob_exponent / n
    E2 = r.prob_sq.astype(np.float64) / 2.0 ** r.sq_exponent / n
    assert np.allclose(E2, P, rtol=0, atol=1e-12)
    se = np.sqrt((E2 - P * P).clip(0) / (n - 1))
    assert np.allclose(se, np.sqrt(P * (1 - P) / (n - 1)), rtol=0, atol=1e-9)
    assert not np.any(r.th_sq)


def _discrete_exact(net_rates, p0, proj, K, nsteps):
    """Jump chain with the reference's discrete-time window rule (a fixed point is counted only in
    the window where it is reached, src/jit_kernels/window_average_small.cu:101-136)."""
    S = len(p0)
    tot = net_rates.sum(axis=1)
    P = np.zeros((S, S))
    for s in range(S):
        if tot[s] > 0:
            for i in np.nonzero(net_rates[s])[0]:
                P[s, s ^ (1 << i)] += net_rates[s, i] / tot[s]
    absorbing = tot == 0
    a = p0.copy()
    windows = np.zeros((nsteps, 1 << K))
    final = np.zeros(1 << K)
    for k in range(nsteps):
        np.add.at(windows[k], proj, a)
        if k == nsteps - 1:
            np.add.at(final, proj, a)
        else:
            np.add.at(final, proj, np.where(absorbing, a, 0.0))
            a = np.where(absorbing, 0.0, a) @ P
    return windows, final


def test_discrete_cellcycle_matches_jump_chain():
    b, c = model_text("cellcycle")
    m = O.OracleModel(b, c)
    names = m.names()
    S = 1 << 10
    rates = np.zeros((S, 10))
    for s in range(S):
        x = [(s >> i) & 1 for i in range(10)]
        for i in range(10):
            rates[s, i] = m.rate_f32(i, x)
    p0 = np.zeros(S)
    ist = m.istate()
    p0[sum(1 << i for i in range(10) if ist[i] == 1)] = 1.0
    ro = [i for i in range(10) if not m.internal()[i]]
    proj = np.array([sum(((s >> j) & 1) << k for k, j in enumerate(ro)) for s in range(S)])
    n = 200000
    r = m.run(count=n, n_trajectories=n, threads=8)
    exact, final = _discrete_exact(rates, p0, proj, len(ro), r.windows)
    got = r.prob / n
    assert np.all(ctmc.within_sigmas(got, exact, n)), np.max(np.abs(got - exact))
    assert np.all(ctmc.within_sigmas(r.final / n, final, n))
    assert names[0] == "CycD"


def _parse_res_windows(path, names_order):
    windows = []
    with open(path) as f:
        lines = f.read().splitlines()
    cur = None
    for ln in lines:
        if ln.startswith("window ("):
            cur = {}
            windows.append(cur)
        elif cur is not None and ln and not ln.startswith("entropy") and not ln.startswith("transition"):
            p, st = ln.split(" ", 1)
            cur[st] = float(p)
    return windows


def test_cellcycle_res_windows_at_4_sigma():
    """The reference's test/cellcycle.res (1e6 trajectories, discrete time) against the oracle."""
    from tests.conftest import FIXTURES
    import os

    b, c = model_text("cellcycle")
    m = O.OracleModel(b, c)
    names = m.names()
    ro = [i for i in range(m.n) if not m.internal()[i]]
    n = 200000
    r = m.run(count=n, n_trajectories=n, threads=8)
    res = _parse_res_windows(os.path.join(FIXTURES, "cellcycle.res"), names)
    assert len(res) == r.windows == 25
    n_ref = 1000000
    for w, win in enumerate(res):
        mine = {}
        for s in np.nonzero(r.prob[w])[0]:
            st = "-".join(names[ro[j]] for j in range(len(ro)) if (s >> j) & 1) or "<nil>"
            mine[st] = r.prob[w][s] / n
        for st in set(mine) | set(win):
            a, e = mine.get(st, 0.0), win.get(st, 0.0)
            sig = math.sqrt(max(e * (1 - e), 1e-6) / n + max(e * (1 - e), 1e-6) / n_ref)
            assert abs(a - e) <= 4 * sig, (w, st, a, e)

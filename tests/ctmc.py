"""Exact continuous-time / jump-chain solutions for tiny networks (test helper).

Independent of both the oracle and the product: node rates come from the fast-path
closed form rate_i(x) = (x_i == logic_i(x)) ? 0 : (x_i ? D_i : U_i)
(src/generator.cpp:93-107), with the logic evaluated by Python's own `and/or/not`
on the synthetic formulas.
"""
from __future__ import annotations

import math
import re

import numpy as np
import scipy.linalg


def py_logic(expr: str, names):
    """Compiles a synthetic logic formula (& | ! parentheses, node names) to a Python predicate."""
    idx = {n: i for i, n in enumerate(names)}
    toks = re.findall(r"[A-Za-z_][A-Za-z_0-9]*|[&|!()]", expr)
    out = []
    for t in toks:
        if t == "&":
            out.append(" and ")
        elif t == "|":
            out.append(" or ")
        elif t == "!":
            out.append(" not ")
        elif t in "()":
            out.append(t)
        else:
            out.append(f"bool(x[{idx[t]}])")
    code = compile("".join(out).strip(), "<logic>", "eval")
    return lambda x: bool(eval(code, {}, {"x": x}))


def state_rates(net, states):
    """Node rates (FP32 values as float64) of the given states."""
    n = net.n
    preds = [py_logic(e, net.names) for e in net.logic]
    rates = np.zeros((len(states), n))
    for k, s in enumerate(states):
        x = [(s >> i) & 1 for i in range(n)]
        for i in range(n):
            lg = preds[i](x)
            if bool(x[i]) != lg:
                rates[k, i] = float(net.down[i]) if x[i] else float(net.up[i])
    return rates


def rate_matrix(net):
    """Generator Q (2^N x 2^N, row = from-state) and per-state node rates."""
    n = net.n
    S = 1 << n
    rates = state_rates(net, range(S))
    Q = np.zeros((S, S))
    for s in range(S):
        for i in range(n):
            if rates[s, i] > 0:
                Q[s, s ^ (1 << i)] += rates[s, i]
        Q[s, s] = -rates[s].sum()
    return Q, rates


def initial_distribution(net):
    n = net.n
    p = np.ones(1 << n)
    for s in range(1 << n):
        for i in range(n):
            pi = float(np.float32(net.istate[i]))
            p[s] *= pi if (s >> i) & 1 else 1.0 - pi
    return p


def projection(net):
    ro = list(net.readouts)
    n = net.n
    return np.array([sum((((s >> j) & 1) << k) for k, j in enumerate(ro)) for s in range(1 << n)])


def entropy_per_state(net, rates):
    ro = list(net.readouts)
    out = np.zeros(len(rates))
    for s, r in enumerate(rates):
        rr = r[ro]
        R = rr.sum()
        if r.sum() == 0 or R == 0:
            continue
        p = rr[rr > 0] / R
        out[s] = float(-(p * np.log2(p)).sum())
    return out


def continuous_solution(net, n_windows):
    """Window-averaged projected probabilities, window TH, final projected distribution and
    fixed-point (absorbing-state) probabilities at max_time."""
    Q, rates = rate_matrix(net)
    p = initial_distribution(net)
    proj = projection(net)
    K = len(net.readouts)
    th_state = entropy_per_state(net, rates)
    tick = float(np.float32(net.time_tick))
    maxt = float(np.float32(net.max_time))
    S = len(p)
    A = np.zeros((2 * S, 2 * S))
    A[:S, :S] = Q
    A[:S, S:] = np.eye(S)
    E = scipy.linalg.expm(A * tick)
    step, integ = E[:S, :S], E[:S, S:]
    probs = np.zeros((n_windows, 1 << K))
    th = np.zeros(n_windows)
    cur = p.copy()
    for w in range(n_windows):
        span = min(tick, maxt - w * tick)  # the last window may end at max_time; it is still
        if span < tick:                    # normalised by a full tick (window_average_small.cpp:126-140)
            Ep = scipy.linalg.expm(A * span)
            avg = cur @ Ep[:S, S:] / tick
        else:
            avg = cur @ integ / tick
        np.add.at(probs[w], proj, avg)
        th[w] = avg @ th_state
        cur = cur @ step
    pend = p @ scipy.linalg.expm(Q * maxt)
    final = np.zeros(1 << K)
    np.add.at(final, proj, pend)
    absorbing = rates.sum(axis=1) == 0
    fixed = {s: pend[s] for s in range(S) if absorbing[s] and pend[s] > 0}
    return probs, th, final, fixed


def sigma(p, n):
    return np.sqrt(np.maximum(p * (1 - p), 1.0 / n) / n)


def within_sigmas(measured, exact, n, k=4.0):
    return np.abs(measured - exact) <= k * sigma(exact, n) + 1e-12

"""Sweep surrogate: a BF16-MFMA MLP ensemble that maps a variant's parameters to its windowed
readout marginals, trained on data the simulator generates, gated by ensemble spread with a
plain-simulation fallback, and judged against the simulator.

Workload (the "surrogate" configuration): the 200-node synthetic network (in-degree 2-4,
log-uniform base rates, 16 readouts, max_time 50, time_tick 0.5 -> 100 windows).  A variant
multiplies each node's rate_up and rate_down by factors drawn log-uniformly in [0.2, 5] (400
numbers) and draws each node's initial ON probability uniformly in [0, 1] (200 numbers).
Input features: log(multiplier)/log(5) and 2*istate - 1 (600 values).  Output: P(readout ON)
per window, 16 x 100 = 1,600 values (device marginals of the simulator's window tables).
Network: 600 (padded to 640) -> 512 -> 512 -> 512 -> 1,600, GELU, sigmoid output ("4 layers
of width 512" read as four linear layers), MSE, Adam; 4 ensemble members.

    python -m upmabossg_amd.surrogate [--train 50000] [--test 10000] [--report out.json]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import Context, Model, synth
from .mlp import Mlp

LOG5 = math.log(5.0)


class ParamSpace:
    """Variant parameters of a base network: rate multipliers log-uniform in [0.2, 5], istate U[0, 1]."""

    def __init__(self, net):
        self.net = net
        self.n = net.n

    def sample(self, count, seed):
        rng = np.random.default_rng(seed)
        lu = rng.uniform(-1.0, 1.0, (count, 2 * self.n))
        mult = np.exp(lu * LOG5).astype(np.float32)
        ist = rng.uniform(0.0, 1.0, (count, self.n)).astype(np.float32)
        return mult[:, : self.n], mult[:, self.n:], ist

    @staticmethod
    def features(mu, md, ist):
        return np.concatenate([np.log(mu) / LOG5, np.log(md) / LOG5, ist * 2.0 - 1.0], axis=1).astype(np.float32)


def simulate_marginals(ctx, base: Model, mu, md, ist, trajectories, seed=1, chunk=256, variant_base=0):
    """Marginals (count, W*K) of each variant from the simulator (variant sweeps of `chunk`)."""
    count = mu.shape[0]
    out = None
    graph_s = 0.0
    t0 = time.perf_counter()
    for c0 in range(0, count, chunk):
        models = []
        for i in range(c0, min(count, c0 + chunk)):
            m = base.clone()
            m.set_variant(mu[i], md[i], ist[i])
            models.append(m)
        info = ctx.run_variants([(m, trajectories, seed, variant_base + c0 + k) for k, m in enumerate(models)])
        graph_s += info.graph_ms / 1e3
        for k in range(len(models)):
            marg = ctx.variant_marginals(k).ravel()
            if out is None:
                out = np.zeros((count, marg.size), np.float32)
            out[c0 + k] = marg
    return out, graph_s, time.perf_counter() - t0


class Ensemble:
    def __init__(self, n_in, n_out, members=4, hidden=512, hidden_layers=3, seed=0, max_batch=1024):
        self.members = [Mlp(n_in, hidden, hidden_layers, n_out, max_batch=max_batch, seed=seed + 7919 * k)
                        for k in range(members)]

    def fit(self, X, Y, epochs=40, batch=256, lr=1e-3, seed=0):
        losses = []
        for k, m in enumerate(self.members):
            m.load_dataset(X, Y)
            losses.append(m.train(epochs, batch=batch, lr=lr, seed=seed + k))
        return losses

    def predict(self, X):
        preds, ms = [], 0.0
        for m in self.members:
            p, t = m.predict(X)
            preds.append(p)
            ms += t
        P = np.stack(preds)
        mean = P.mean(axis=0)
        spread = P.std(axis=0).max(axis=1)  # per point: the largest member disagreement over outputs
        return mean, spread, ms / 1e3


def forward_flops(n_in_padded, hidden, hidden_layers, n_out):
    dims = [n_in_padded] + [hidden] * hidden_layers + [n_out]
    return sum(2 * a * b for a, b in zip(dims[:-1], dims[1:]))


def run(n_train=50000, n_test=10000, n_ref=128, t_train=1024, t_ref=65536, t_plain=32768, members=4, epochs=40,
        threshold=0.01, seed=1, log=print):
    net = synth.config_surrogate(seed)
    base = Model(net.bnd(), net.cfg())
    space = ParamSpace(net)
    report = {"workload": "surrogate: 200-node network, 16 readouts x 100 windows, rates x[0.2, 5], istate U[0, 1]"}
    with Context(0) as ctx:
        mu, md, ist = space.sample(n_train, seed * 1000 + 1)
        t0 = time.perf_counter()
        Y, gs, ws = simulate_marginals(ctx, base, mu, md, ist, t_train, seed=11, variant_base=0)
        X = ParamSpace.features(mu, md, ist)
        log(f"training data: {n_train} variants x {t_train} trajectories, graph {gs:.1f} s, wall {ws:.1f} s")
        report["train"] = {"points": n_train, "trajectories_per_point": t_train, "sim_graph_s": gs, "sim_wall_s": ws}

        ens = Ensemble(X.shape[1], Y.shape[1], members=members, seed=seed)
        t1 = time.perf_counter()
        losses = ens.fit(X, Y, epochs=epochs)
        fit_s = time.perf_counter() - t1
        log(f"ensemble of {members} trained in {fit_s:.1f} s, final losses {losses}")
        report["fit"] = {"members": members, "epochs": epochs, "seconds": fit_s, "final_loss": losses}

        # held-out screening set
        tmu, tmd, tist = space.sample(n_test, seed * 1000 + 2)
        TX = ParamSpace.features(tmu, tmd, tist)
        mean, spread, t_inf = ens.predict(TX)
        accept = spread <= threshold
        flops = forward_flops(ens.members[0].padded_inputs, 512, 3, Y.shape[1]) * n_test * members
        log(f"screened {n_test} points: {accept.mean() * 100:.1f}% accepted, inference {t_inf * 1e3:.2f} ms")

        # accuracy of accepted predictions against high-precision simulation on a subset
        ref_idx = np.arange(min(n_ref, n_test))
        R, rg, rw = simulate_marginals(ctx, base, tmu[ref_idx], tmd[ref_idx], tist[ref_idx], t_ref, seed=23,
                                       chunk=128, variant_base=10 ** 6)
        err = np.abs(mean[ref_idx] - R).max(axis=1)
        acc_ref = accept[ref_idx]
        log(f"reference subset {len(ref_idx)} x {t_ref}: max error accepted "
            f"{err[acc_ref].max() if acc_ref.any() else float('nan'):.4f}, all {err.max():.4f}")

        # plain simulation throughput at the trajectory count that reaches ~0.01 max error alone
        pn = 64
        pmu, pmd, pist = space.sample(pn, seed * 1000 + 3)
        _, pg, pw = simulate_marginals(ctx, base, pmu, pmd, pist, t_plain, seed=31, chunk=pn, variant_base=2 * 10 ** 6)
        plain_pps = pn / pw
        # screened throughput: inference for all + plain simulation of the rejected points
        n_fb = int((~accept).sum())
        fallback_s = n_fb / plain_pps
        screened_pps = n_test / (t_inf + fallback_s)
        log(f"plain {plain_pps:.1f} points/s; screened {screened_pps:.1f} points/s "
            f"({screened_pps / plain_pps:.1f}x), fallback {n_fb}")
        report["screen"] = {
            "test_points": n_test,
            "threshold_spread": threshold,
            "accepted_fraction": float(accept.mean()),
            "inference_s": t_inf,
            "inference_tflops": flops / t_inf / 1e12 if t_inf > 0 else None,
            "mfma_peak_tflops": 2500.0,
            "reference_points": int(len(ref_idx)),
            "reference_trajectories": t_ref,
            "max_abs_error_accepted": float(err[acc_ref].max()) if acc_ref.any() else None,
            "max_abs_error_all": float(err.max()),
            "mean_max_abs_error": float(err.mean()),
            "plain_trajectories_per_point": t_plain,
            "plain_points_per_s": plain_pps,
            "plain_sim_graph_s": pg,
            "screened_points_per_s": screened_pps,
            "speedup": screened_pps / plain_pps,
        }
    return report


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=50000)
    ap.add_argument("--test", type=int, default=10000)
    ap.add_argument("--ref", type=int, default=128)
    ap.add_argument("--t-train", type=int, default=1024)
    ap.add_argument("--t-ref", type=int, default=65536)
    ap.add_argument("--t-plain", type=int, default=32768)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    rep = run(a.train, a.test, a.ref, a.t_train, a.t_ref, a.t_plain, a.members, a.epochs,
              log=lambda s: print(s, flush=True))
    txt = json.dumps(rep)
    print(txt, flush=True)
    if a.report:
        with open(a.report, "w") as f:
            f.write(txt + "\n")


if __name__ == "__main__":
    main(sys.argv[1:])

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
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import Context, Model, synth
from .mlp import Mlp, ensemble_predict

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

    def fit(self, X, Y, epochs=40, batch=256, lr=1e-3, seed=0, stages=4, decay=0.3):
        """Adam over `epochs` epochs in `stages` equal stages, the rate multiplied by `decay` per stage."""
        per = [epochs // stages + (1 if s < epochs % stages else 0) for s in range(stages)]

        def one(k):
            m = self.members[k]
            m.load_dataset(X, Y)
            loss = None
            for s, e in enumerate(per):
                if e:
                    loss = m.train(e, batch=batch, lr=lr * decay ** s, seed=seed + 1000 * k + s)
            return loss

        # members are independent (own buffers and stream): train them concurrently so their
        # small per-step kernels overlap on the device (MBG_SURROGATE_SERIAL=1: one after
        # another, for tools that trace kernel launches from a single host thread)
        if os.environ.get("MBG_SURROGATE_SERIAL", "0") == "1":
            return [one(k) for k in range(len(self.members))]
        with ThreadPoolExecutor(max_workers=len(self.members)) as pool:
            return list(pool.map(one, range(len(self.members))))

    def predict(self, X):
        """(mean, spread, kernel seconds, wall seconds); spread = per point, the largest standard
        deviation over members of any output."""
        t0 = time.perf_counter()
        mean, spread, ms = ensemble_predict(self.members, X)
        return mean, spread, ms / 1e3, time.perf_counter() - t0


class SweepSurrogate:
    """Screen variants of a base model: the ensemble's mean where its spread is at most
    `threshold`, the simulator's marginals (plain simulation) everywhere else."""

    def __init__(self, ctx, base: Model, members=4, hidden=512, hidden_layers=3, seed=0, max_batch=10240):
        self.ctx, self.base = ctx, base
        self.n = base.num_nodes
        self.ens = None
        self.kw = dict(members=members, hidden=hidden, hidden_layers=hidden_layers, seed=seed, max_batch=max_batch)

    def fit(self, mu, md, ist, Y, **kw):
        X = ParamSpace.features(mu, md, ist)
        self.ens = Ensemble(X.shape[1], Y.shape[1], **self.kw)
        return self.ens.fit(X, Y, **kw)

    def screen(self, mu, md, ist, trajectories, threshold=0.01, seed=1, variant_base=0, chunk=256):
        """(marginals [n][W*K], accepted [n], info): rejected points are simulated with
        `trajectories` trajectories each, variant ids variant_base + point index."""
        if self.ens is None:
            raise RuntimeError("the surrogate is not trained")
        mean, spread, dev_s, wall_s = self.ens.predict(ParamSpace.features(mu, md, ist))
        accepted = spread <= threshold
        out = mean.copy()
        rej = np.flatnonzero(~accepted)
        sim_s = 0.0
        for c0 in range(0, rej.size, chunk):
            idx = rej[c0:c0 + chunk]
            models = []
            for i in idx:
                m = self.base.clone()
                m.set_variant(mu[i], md[i], ist[i])
                models.append(m)
            t0 = time.perf_counter()
            self.ctx.run_variants([(m, trajectories, seed, variant_base + int(i)) for m, i in zip(models, idx)])
            for k, i in enumerate(idx):
                out[i] = self.ctx.variant_marginals(k).ravel()
            sim_s += time.perf_counter() - t0
        return out, accepted, {"inference_wall_s": wall_s, "inference_device_s": dev_s, "fallback_s": sim_s,
                               "spread": spread}


def forward_flops(n_in_padded, hidden, hidden_layers, n_out):
    dims = [n_in_padded] + [hidden] * hidden_layers + [n_out]
    return sum(2 * a * b for a, b in zip(dims[:-1], dims[1:]))


def _error_stats(err, spread, thresholds=(0.005, 0.01, 0.02)):
    out = {"points": int(err.size), "max": float(err.max()), "mean": float(err.mean()),
           "p50": float(np.percentile(err, 50)), "p95": float(np.percentile(err, 95)),
           "fraction_le_0.01": float((err <= 0.01).mean())}
    for th in thresholds:
        acc = spread <= th
        out[f"gate_{th}"] = {"accepted": float(acc.mean()),
                             "max_err_accepted": float(err[acc].max()) if acc.any() else None}
    return out


def run(n_train=50000, n_test=10000, n_ref=64, t_train=4096, t_ref=262144, t_plain=49152, members=4, epochs=160,
        threshold=0.01, seed=1, lr=1e-3, batch=256, plain_counts=(64, 256, 1024, 4096), log=print):
    net = synth.config_surrogate(seed)
    base = Model(net.bnd(), net.cfg())
    space = ParamSpace(net)
    report = {"workload": "surrogate: 200-node network, 16 readouts x 100 windows, rates x[0.2, 5], istate U[0, 1]"}
    with Context(0) as ctx:
        mu, md, ist = space.sample(n_train, seed * 1000 + 1)
        Y, gs, ws = simulate_marginals(ctx, base, mu, md, ist, t_train, seed=11, variant_base=0)
        X = ParamSpace.features(mu, md, ist)
        log(f"training data: {n_train} variants x {t_train} trajectories, graph {gs:.1f} s, wall {ws:.1f} s")
        report["train"] = {"points": n_train, "trajectories_per_point": t_train, "sim_graph_s": gs, "sim_wall_s": ws}

        ens = Ensemble(X.shape[1], Y.shape[1], members=members, seed=seed, max_batch=10240)
        t1 = time.perf_counter()
        losses = ens.fit(X, Y, epochs=epochs, batch=batch, lr=lr)
        fit_s = time.perf_counter() - t1
        log(f"ensemble of {members} trained in {fit_s:.1f} s ({epochs} epochs), final losses {losses}")
        report["fit"] = {"members": members, "epochs": epochs, "batch": batch, "lr": lr, "seconds": fit_s,
                         "final_loss": losses}
        del X, Y

        # held-out screening set
        tmu, tmd, tist = space.sample(n_test, seed * 1000 + 2)
        TX = ParamSpace.features(tmu, tmd, tist)
        ens.predict(TX[:2048])  # warm-up
        mean, spread, t_kernel, t_inf = ens.predict(TX)
        accept = spread <= threshold
        flops = forward_flops(ens.members[0].padded_inputs, 512, 3, mean.shape[1]) * n_test * members
        log(f"screened {n_test} points: {accept.mean() * 100:.1f}% accepted, inference {t_inf * 1e3:.1f} ms wall, "
            f"{t_kernel * 1e3:.2f} ms in kernels ({flops / t_kernel / 1e12:.1f} TFLOP/s)")

        # judged against high-trajectory simulation on a subset of the held-out points
        ri = np.arange(min(n_ref, n_test))
        R, _, rw = simulate_marginals(ctx, base, tmu[ri], tmd[ri], tist[ri], t_ref, seed=23, chunk=64,
                                      variant_base=10 ** 6)
        err = np.abs(mean[ri] - R).max(axis=1)
        stats = _error_stats(err, spread[ri])
        log(f"reference: {len(ri)} points x {t_ref} trajectories ({rw:.1f} s): {stats}")

        # plain simulation on the same subset: its own error at several trajectory counts and its rate
        plain = {}
        rate_n = min(n_test, 1024)
        for T in sorted(set(plain_counts) | {t_plain}):
            P, _, _ = simulate_marginals(ctx, base, tmu[ri], tmd[ri], tist[ri], T, seed=31, chunk=len(ri),
                                         variant_base=2 * 10 ** 6)
            pe = np.abs(P - R)
            # rate from batched sweeps of 256 variants over rate_n points
            _, pg, pw = simulate_marginals(ctx, base, tmu[:rate_n], tmd[:rate_n], tist[:rate_n], T, seed=37,
                                           chunk=256, variant_base=3 * 10 ** 6)
            plain[T] = {"points_per_s": rate_n / pw, "graph_s": pg, "wall_s": pw, "rms": float(np.sqrt((pe ** 2).mean())),
                        "max": float(pe.max()), "max_mean": float(pe.max(axis=1).mean())}
            log(f"plain at {T} trajectories: {plain[T]}")
        plain_pps = plain[t_plain]["points_per_s"]
        n_fb = int((~accept).sum())
        screened_pps = n_test / (t_inf + n_fb / plain_pps)
        log(f"at the spread gate {threshold}: screened {screened_pps:.1f} points/s = {screened_pps / plain_pps:.2f}x "
            f"plain at {t_plain}; fallback {n_fb} points")
        # matched accuracy: the fewest trajectories at which plain simulation is as accurate (RMS) as the surrogate
        rms_s = float(np.sqrt(((mean[ri] - R) ** 2).mean()))
        eq = [T for T in sorted(plain) if plain[T]["rms"] <= rms_s]
        t_eq = eq[0] if eq else None
        surrogate_pps = n_test / t_inf
        report["screen"] = {
            "test_points": n_test,
            "threshold_spread": threshold,
            "accepted_fraction": float(accept.mean()),
            "fallback_points": n_fb,
            "inference_wall_s": t_inf,
            "inference_device_s": t_kernel,
            "inference_tflops": flops / t_kernel / 1e12 if t_kernel > 0 else None,
            "mfma_peak_tflops": 2500.0,
            "reference_trajectories": t_ref,
            "surrogate_error": stats,
            "surrogate_rms": rms_s,
            "plain": {str(k): v for k, v in plain.items()},
            "plain_trajectories_per_point": t_plain,
            "plain_points_per_s": plain_pps,
            "screened_points_per_s": screened_pps,
            "speedup_at_gate": screened_pps / plain_pps,
            "matched_rms_trajectories": t_eq,
            "surrogate_points_per_s": surrogate_pps,
            "speedup_at_matched_rms": surrogate_pps / plain[t_eq]["points_per_s"] if t_eq else None,
        }
    return report


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=50000)
    ap.add_argument("--test", type=int, default=10000)
    ap.add_argument("--ref", type=int, default=64)
    ap.add_argument("--t-train", type=int, default=4096)
    ap.add_argument("--t-ref", type=int, default=262144)
    ap.add_argument("--t-plain", type=int, default=49152)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=160)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    rep = run(a.train, a.test, a.ref, a.t_train, a.t_ref, a.t_plain, a.members, a.epochs, lr=a.lr, batch=a.batch,
              log=lambda s: print(s, flush=True))
    txt = json.dumps(rep)
    print(txt, flush=True)
    if a.report:
        with open(a.report, "w") as f:
            f.write(txt + "\n")


if __name__ == "__main__":
    main(sys.argv[1:])

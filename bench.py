"""Throughput benchmark of the trajectory kernel + statistics stage on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]

Workload ("large"): a synthetic random Boolean network of 1,000 nodes (in-degree 2-4, random
AND/OR/NOT formulas, rate_up/rate_down log-uniform in [0.1, 10] FP32, istate U[0, 1], 16
readout nodes), 1,048,576 trajectories, max_time 50, time_tick 0.5 (100 windows).  One step is
one full run: every trajectory simulated to max_time and the window, final-state and
fixed-point tables accumulated and returned to the host.

Prints one JSON line.  value = transitions/s over the K timed steps (wall clock around the
synchronous runs, inputs already parsed and the kernel already JIT-compiled).  roofline: the
kernel is bounded by vector ALU issue (VALU), priced with a fixed algorithmic op count per
transition, W(N, d, k) = 72 + 2*ceil(N/64) + 2*min(N, 64) + (1 + d)(2k + 6) + 10 lane-ops
(d = mean out-degree, k = mean in-degree), against 256 CUs x 64 lanes x 2.4 GHz = 39.3 T
lane-ops/s, divided by the kernel's own HIP-event duration.  cpu_baseline: the CPU oracle
(the reference algorithm: full O(N) rate recompute and linear flip scan per step) on this
host's cores over trajectory ids 0-4095 of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12  # non-packed 32-bit lane-ops, MI355X


def algorithmic_ops(n, mean_in, mean_out):
    return 72 + 2 * math.ceil(n / 64) + 2 * min(n, 64) + (1 + mean_out) * (2 * mean_in + 6) + 10


def cpu_baseline(net, count=4096):
    from oracle.oracle import OracleModel

    cores = min(16, os.cpu_count() or 1)
    o = OracleModel(net.bnd(), net.cfg())
    r = o.run(count=count, n_trajectories=net.sample_count, threads=cores, record_fixed=True)
    return {"value": r.flips / r.seconds, "unit": "transitions/s", "cores": cores, "kind": "port",
            "sample": f"trajectory ids 0-{count - 1} of the same workload ({r.flips} transitions, "
                      f"{r.seconds:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nodes", type=int, default=1000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one GPU)",
              file=sys.stderr)
        sys.exit(2)

    from upmabossg_amd import Context, Model, synth

    net = synth.config_large(n=args.nodes)
    model = Model(net.bnd(), net.cfg())
    ins = net.inputs()
    mean_in = sum(len(i) for i in ins) / net.n
    ops = algorithmic_ops(net.n, mean_in, mean_in)

    with Context(0) as ctx:
        ctx.compile(model)
        for _ in range(args.warmup):
            ctx.run(model, fetch=False)
        transitions, kernel_ms = 0, []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.run(model, fetch=False)
            i = ctx.info()
            transitions += i.transitions
            kernel_ms.append(i.kernel_ms)
        wall = time.perf_counter() - t0
        info = ctx.info()

    kern_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    per_launch = transitions / args.steps
    achieved = per_launch * ops / kern_s / 1e12
    out = {
        "metric": "transitions/s",
        "value": transitions / wall,
        "unit": "transitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 fixed-point rates, f64 clock",
        "data": "synthetic",
        "config": {
            "workload": f"large: {net.n}-node random network, in-degree 2-4, "
                        f"{info.simulated} trajectories, max_time 50, time_tick 0.5, 16 readouts",
            "nodes": net.n,
            "trajectories": int(info.simulated),
            "windows": int(info.windows),
            "transitions_per_step": per_launch,
            "trajectories_per_s": info.simulated * args.steps / wall,
            "kernel_ms": kern_s * 1e3,
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": VALU_PEAK_TOPS,
            "unit": "T lane-op/s",
            "frac": achieved / VALU_PEAK_TOPS,
            "ops_per_transition": ops,
            "traffic": None,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(net)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

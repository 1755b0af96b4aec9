"""Throughput benchmark of the trajectory kernel + statistics stage on one GPU.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-sweep] [--no-surrogate]

Workload ("large"): a synthetic random Boolean network of 1,000 nodes (in-degree 2-4, random
AND/OR/NOT formulas, rate_up/rate_down log-uniform in [0.1, 10] FP32, istate U[0, 1], 16
readout nodes), 1,048,576 trajectories, max_time 50, time_tick 0.5 (100 windows).  One step is
one full run (after an untimed autotune of the launch shape, like the hipRTC compile): every
trajectory simulated to max_time and the window, final-state and
fixed-point tables accumulated; the small tables (final states, fixed points, counters) return
to the host, the 52 MB window table stays in device memory (the library copies it on demand).
config.transitions_per_s_with_host_copy repeats the steps with the window table copied to
host memory after each (PCIe-inclusive rate).

Prints one JSON line.  value = transitions/s over the K timed steps (wall clock around the
synchronous runs, inputs already parsed and the kernel already JIT-compiled).  roofline: the
kernel is bounded by vector ALU issue (VALU), priced with a fixed algorithmic op count per
transition, W(N, d, k) = 72 + 2*ceil(N/64) + 2*min(N, 64) + (1 + d)(2k + 6) + 10 lane-ops
(d = mean out-degree, k = mean in-degree), against 256 CUs x 64 lanes x 2.4 GHz = 39.3 T
lane-ops/s, divided by the kernel's own HIP-event duration.  cpu_baseline: the CPU oracle
(the reference algorithm: full O(N) rate recompute and linear flip scan per step) on this
host's cores over trajectory ids 0-4095 of the same workload.  roofline.traffic: HBM bytes per
launch of the kernel from the FETCH_SIZE and WRITE_SIZE counter summary committed in
profiles/mbg_simulate_large_traffic.txt (each counter scaled by what it reads for a copy kernel
of known size), null when that file is absent.

More objects ride on the same line:
  sweep      256 variant networks of 50-2,000 nodes (rates scaled by log-normal(0, 0.5), 0-3
             knockouts, 1k-64k trajectories each) in one HIP-graph launch: transitions/s and
             variants/s of the second (warm) sweep, after an untimed per-class autotune
             (mbg_autotune_variants); the classes' chosen layouts and kernel spans.
  fixed_heavy the 5,000-node fixed-point-heavy network (262,144 trajectories, max_time 200):
             transitions/s per step and in the kernel, with its own VALU roofline figure.
  surrogate  the sweep surrogate on the 200-node network (50,000 simulated training variants,
             4-member BF16 MLP ensemble, 10,000 held-out variants, spread gate 0.01, judged
             against 262,144-trajectory simulations): accepted fraction, errors, points/s of
             screening against plain batched simulation.
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


def traffic_from_profile(path=os.path.join(ROOT, "profiles", "mbg_simulate_large_traffic.txt")):
    """HBM bytes per mbg_simulate launch from the counter summary: lines `key: value` with keys
    fetch_kib, write_kib (per dispatch), fetch_factor, write_factor (copy-kernel calibration)."""
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            parts = line.strip().split(":")
            if len(parts) == 2 and parts[0].strip() in ("fetch_kib", "write_kib", "fetch_factor", "write_factor"):
                vals[parts[0].strip()] = float(parts[1])
    if len(vals) != 4:
        return None
    return (vals["fetch_kib"] * vals["fetch_factor"] + vals["write_kib"] * vals["write_factor"]) * 1024.0


def copy_calibration(nbytes=1 << 30):
    """One device-to-device copy of nbytes (the counters' calibration kernel)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def sweep_bench(ctx, count=256, tune=True):
    from upmabossg_amd import Model, synth

    nets = synth.sweep_variants(count)
    models = [Model(n.bnd(), n.cfg()) for n in nets]
    vs = [(m, n.sample_count, 1, v) for v, (m, n) in enumerate(zip(models, nets))]
    t_tune = time.perf_counter()
    if tune:
        ctx.autotune_variants(vs)  # set-up, untimed: per-class layouts measured on these variants
    t_tune = time.perf_counter() - t_tune
    ctx.run_variants(vs)  # JIT of the word classes, buffers
    t0 = time.perf_counter()
    info = ctx.run_variants(vs)
    wall = time.perf_counter() - t0
    classes = [{k: c[k] for k in ("words", "variants", "transitions", "kernel_ms", "block", "blocks_per_cu",
                                  "tables_lds", "state_hbm", "wsum_reg", "sumw", "dep_batch", "dep_cap", "tuned")}
               for c in ctx.sweep_classes()]
    return {
        "workload": f"sweep: {count} variant networks of {min(n.n for n in nets)}-{max(n.n for n in nets)} nodes, "
                    "rates x log-normal(0, 0.5), 0-3 knockouts, 1k-64k trajectories each, one HIP graph",
        "variants": count,
        "word_classes": info.classes,
        "trajectories": int(info.trajectories),
        "transitions": int(info.transitions),
        "graph_ms": info.graph_ms,
        "setup_ms": info.setup_ms,
        "wall_ms": wall * 1e3,
        "transitions_per_s": info.transitions / wall,
        "transitions_per_s_graph": info.transitions / (info.graph_ms / 1e3),
        "variants_per_s": count / wall,
        "autotune_s": t_tune,
        "classes": classes,
    }


def surrogate_bench():
    from upmabossg_amd import surrogate

    rep = surrogate.run(log=lambda m: print(m, file=sys.stderr, flush=True))
    sc = rep["screen"]
    return {
        "workload": rep["workload"],
        "train_points": rep["train"]["points"],
        "train_trajectories_per_point": rep["train"]["trajectories_per_point"],
        "members": rep["fit"]["members"],
        "fit_s": rep["fit"]["seconds"],
        "test_points": sc["test_points"],
        "spread_gate": sc["threshold_spread"],
        "accepted_fraction": sc["accepted_fraction"],
        "reference_trajectories": sc["reference_trajectories"],
        "surrogate_max_abs_error": sc["surrogate_error"]["max"],
        "surrogate_mean_max_abs_error": sc["surrogate_error"]["mean"],
        "surrogate_rms_error": sc["surrogate_rms"],
        "plain_trajectories_per_point": sc["plain_trajectories_per_point"],
        "plain_max_abs_error": sc["plain"][str(sc["plain_trajectories_per_point"])]["max"],
        "plain_points_per_s": sc["plain_points_per_s"],
        "screened_points_per_s": sc["screened_points_per_s"],
        "speedup_at_0.01": sc["speedup_at_gate"],
        "target": ">= 20x the points/s of plain batched simulation at <= 0.01 max abs probability error",
        "target_met": bool(sc["speedup_at_gate"] >= 20.0),
        # not the target metric: plain simulation at the trajectory count whose RMS error matches
        # the surrogate's (its own max error is far above 0.01)
        "secondary": {"matched_rms_trajectories": sc["matched_rms_trajectories"],
                      "surrogate_points_per_s": sc["surrogate_points_per_s"],
                      "speedup_at_matched_rms": sc["speedup_at_matched_rms"]},
        "roofline": {"bound": "mfma", "achieved": sc["inference_tflops"], "peak": 2500.0, "unit": "TFLOP/s",
                     "frac": sc["inference_tflops"] / 2500.0},
    }


def fixed_heavy_bench(ctx, steps=2, tune=True):
    """The fixed-point-heavy workload: 5,000 nodes (79 state words), in-degree 1-3, 30 % of the
    nodes self-sustaining (x = x | f) or absorbing (rate_down 0), 262,144 trajectories, max_time
    200, time_tick 2.  Most trajectories end at a fixed point; every fixed point is a distinct
    5,000-bit state, deduplicated on the device.  Same step definition and roofline as the
    headline line (W(N, d, k) lane-ops per transition against the VALU peak)."""
    from upmabossg_amd import Model, synth

    net = synth.config_fixed_heavy()
    m = Model(net.bnd(), net.cfg())
    ins = net.inputs()
    mean_in = sum(len(i) for i in ins) / net.n
    ops = algorithmic_ops(net.n, mean_in, mean_in)
    tune_info = ctx.autotune(m) if tune else None
    ctx.run(m, fetch=False)  # compile, buffers
    walls, kms, trans = [], [], 0
    for _ in range(steps):
        t0 = time.perf_counter()
        ctx.run(m, fetch=False)
        walls.append(time.perf_counter() - t0)
        i = ctx.info()
        kms.append(i.kernel_ms)
        trans += i.transitions
    i = ctx.info()
    kern_s = sum(kms) / len(kms) / 1e3
    per_launch = trans / steps
    achieved = per_launch * ops / kern_s / 1e12
    return {
        "workload": f"fixed-point heavy: {net.n}-node network, in-degree 1-3, 30 % self-sustaining or absorbing "
                    f"nodes, {i.simulated} trajectories, max_time 200, time_tick 2",
        "trajectories": int(i.simulated),
        "transitions_per_step": per_launch,
        "fixed_point_trajectories": int(i.fixed_point_trajectories),
        "ms_per_step": sum(walls) / len(walls) * 1e3,
        "kernel_ms": kern_s * 1e3,
        "transitions_per_s": trans / sum(walls),
        "transitions_per_s_kernel": per_launch / kern_s,
        "launch_shape": ({k: tune_info[k] for k in ("block", "tables_in_lds", "wsum_reg", "state_hbm", "dep_batch",
                                                     "dep_cap")} if tune_info else "untuned"),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "T lane-op/s",
                     "frac": achieved / VALU_PEAK_TOPS, "ops_per_transition": ops},
    }


def ring_bench(ctx, sizes=(100, 500), cpu=True):
    """The ring models of the reference's own generator (data/generate-synth.py defaults: 5
    readouts, 1,000,000 trajectories, time_tick 5, max_time 100, signal length int(sqrt(N)),
    so every node reads 2 int(sqrt(N)) + 1 others), at its default size and near its default
    512-node build limit; the CPU oracle runs a bounded sample of the same model."""
    from oracle.oracle import OracleModel
    from upmabossg_amd import Model, synth

    out = []
    for n in sizes:
        b, c = synth.ring_network(n, sample_count=1000000)
        m = Model(b, c)
        ctx.run(m, count=65536, n_trajectories=65536, fetch=False)  # compile, buffers
        t0 = time.perf_counter()
        ctx.run(m, fetch=False)
        wall = time.perf_counter() - t0
        i = ctx.info()
        e = {"nodes": n, "inputs_per_node": 2 * int(n ** 0.5) + 1, "trajectories": int(i.simulated),
             "transitions": int(i.transitions), "kernel_ms": i.kernel_ms, "wall_ms": wall * 1e3,
             "transitions_per_s": i.transitions / wall}
        if cpu:
            cores = host_cpu()[2]
            cnt = 16384 if n <= 100 else 2048  # a few seconds of CPU work
            r = OracleModel(b, c).run(count=cnt, n_trajectories=1000000, threads=cores, record_fixed=False)
            e["cpu_transitions_per_s"] = r.flips / r.seconds
            e["cpu_sample"] = f"trajectory ids 0-{cnt - 1}, {cores} threads, {r.seconds:.1f} s"
        out.append(e)
    return out


def host_cpu():
    """CPU model name (from /proc/cpuinfo), logical CPUs the OS reports and the threads used: the
    CPU share of a one-GPU box is 16 threads, whatever the machine's total."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    logical = os.cpu_count() or 1
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = logical
    return model, logical, max(1, min(16, avail))


def cpu_baseline(net, count=4096):
    from oracle.oracle import OracleModel

    model, logical, cores = host_cpu()
    o = OracleModel(net.bnd(), net.cfg())
    r = o.run(count=count, n_trajectories=net.sample_count, threads=cores, record_fixed=True)
    return {"value": r.flips / r.seconds, "unit": "transitions/s", "cores": cores, "kind": "port",
            "cpu_model": model, "logical_cpus_reported": logical,
            "per_thread": r.flips / r.seconds / cores,
            "algorithm": "the reference algorithm in the CPU oracle: every node's rate recomputed each step by "
                         "interpreting its expression tree, sequential sum, linear flip scan; one trajectory per thread",
            "sample": f"trajectory ids 0-{count - 1} of the same workload ({r.flips} transitions, "
                      f"{r.seconds:.1f} s, {cores} threads: the CPU share of a one-GPU box)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--nodes", type=int, default=1000)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-surrogate", action="store_true")
    ap.add_argument("--no-ring", action="store_true", help="skip the reference generator's ring models")
    ap.add_argument("--no-fixed-heavy", action="store_true", help="skip the 5,000-node fixed-point-heavy workload")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the untuned launch shape (and the MBG_TUNE_* environment knobs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first copy 1 GiB device to device (reference kernel for the HBM counters)")
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

    if args.calibrate_copy:
        copy_calibration()
    with Context(0) as ctx:
        # set-up, untimed: measured launch shape for this model (the hipRTC compiles happen here)
        tune = None if args.no_autotune else ctx.autotune(model)
        print(f"autotune: {tune}", file=sys.stderr, flush=True)
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
        # the same steps with the window table returned to host memory (PCIe-inclusive; not value)
        host = None
        t1 = time.perf_counter()
        for _ in range(args.steps):
            ctx.run(model, fetch=False)
            host = ctx.probtraj_raw(host)
        wall_host = time.perf_counter() - t1
        print(f"large: {transitions / wall:.4g} transitions/s over {args.steps} steps", file=sys.stderr, flush=True)
        sweep = None
        if not args.no_sweep:
            try:
                sweep = sweep_bench(ctx, tune=not args.no_autotune)
            except Exception as e:  # reported, never fatal to the headline line
                sweep = {"error": f"{type(e).__name__}: {e}"}
            print(f"sweep: {sweep}", file=sys.stderr, flush=True)
        fixed_heavy = None
        if not args.no_fixed_heavy:
            try:
                fixed_heavy = fixed_heavy_bench(ctx, tune=not args.no_autotune)
            except Exception as e:  # reported, never fatal to the headline line
                fixed_heavy = {"error": f"{type(e).__name__}: {e}"}
            print(f"fixed_heavy: {fixed_heavy}", file=sys.stderr, flush=True)
        ring = None
        if not args.no_ring:
            try:
                ring = ring_bench(ctx, cpu=not args.no_cpu)
            except Exception as e:  # reported, never fatal to the headline line
                ring = {"error": f"{type(e).__name__}: {e}"}
            print(f"ring: {ring}", file=sys.stderr, flush=True)

    kern_s = sum(kernel_ms) / len(kernel_ms) / 1e3
    traffic = traffic_from_profile()
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
        "dtype": "u64, f64",
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
            "transitions_per_s_with_host_copy": transitions / wall_host,
            "launch_shape": ({k: tune[k] for k in ("block", "tables_in_lds", "wsum_reg", "compact", "dep_batch",
                                                   "dep_cap", "scan", "wsmad", "state_hbm")} if tune else "untuned"),
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": VALU_PEAK_TOPS,
            "unit": "T lane-op/s",
            "frac": achieved / VALU_PEAK_TOPS,
            "ops_per_transition": ops,
            "traffic": traffic,
        },
        # the same kernel against HBM (its secondary bound): measured bytes per launch / duration
        "roofline_hbm": {
            "bound": "hbm",
            "achieved": traffic / kern_s / 1e9 if traffic else None,
            "peak": 8000.0,
            "unit": "GB/s",
            "frac": traffic / kern_s / 1e9 / 8000.0 if traffic else None,
            "traffic": traffic,
        },
    }
    if sweep is not None:
        out["sweep"] = sweep
    if fixed_heavy is not None:
        out["fixed_heavy"] = fixed_heavy
    if ring is not None:
        out["reference_generator_rings"] = ring
    if not args.no_surrogate:
        try:
            out["surrogate"] = surrogate_bench()
        except Exception as e:
            out["surrogate"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(net)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Builds and runs the CPU emulation of the trajectory kernel (test infrastructure).

The kernel source for a launch shape (``upmabossg_amd.build.kernel_source``: the shape's
defines plus ``mbg_math.h`` and ``mbg_kernel.hip``) is compiled with g++ against
``emu_device.h`` into an executable per shape (``emu_run.cpp``), linked with the product's host
front end, model compiler and blob packer.  ``run`` drives it on one model (or a sweep class
of models) and returns the raw statistics tables the GPU would produce, to be compared bit for
bit with the oracle.  ``asan=True`` builds it with AddressSanitizer (host code only), so every
access of the kernel to an emulated device buffer is bounds-checked.
"""
from __future__ import annotations

import hashlib
import math
import os
import subprocess
import tempfile

import numpy as np

from upmabossg_amd import build as B

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
HOST_SOURCES = ["front.cpp", "model.cpp", "pack.cpp"]
CXX = os.environ.get("CXX", "g++")


MSAN_CXX = "/opt/rocm/llvm/bin/clang++"
last_stderr = ""  # the last emulated run's diagnostics (sanitizer reports)


def _cxx(msan: bool = False):
    return MSAN_CXX if msan else CXX


def _flags(asan: bool, msan: bool = False):
    f = ["-O2", "-g", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wno-unknown-pragmas",
         "-I" + B.CSRC, "-I" + HERE, "-I" + BUILD]
    if asan:
        f += ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
    if msan:
        # MemorySanitizer: a branch, address or output that depends on a never-written value
        # (a local, LDS, a device buffer) is reported with the origin of that value
        f += ["-fsanitize=memory", "-fsanitize-memory-track-origins=2", "-fsanitize-recover=memory",
              "-fno-omit-frame-pointer", "-fsanitize-ignorelist=" + os.path.join(HERE, "msan_ignore.txt")]
    # extra flags for the emulated kernel (for instance -ftrivial-auto-var-init=pattern, which
    # gives every uninitialised local a fixed non-zero pattern)
    f += os.environ.get("EMU_EXTRA_FLAGS", "").split()
    return f


def _host_objects(asan: bool, msan: bool = False):
    out = os.path.join(BUILD, "host_msan" if msan else "host_asan" if asan else "host")
    os.makedirs(out, exist_ok=True)
    objs, procs = [], []
    for s in HOST_SOURCES:
        src = os.path.join(B.CSRC, s)
        obj = os.path.join(out, s + ".o")
        objs.append(obj)
        deps = [src] + [os.path.join(B.CSRC, h) for h in ("front.h", "model.h", "pack.h")]
        if os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
            continue
        # written under a name of this process, then renamed: test workers running in parallel
        # may build the same object
        tmp = f"{obj}.{os.getpid()}.tmp"
        procs.append((tmp, obj, subprocess.Popen([_cxx(msan), *_flags(asan, msan), "-c", src, "-o", tmp],
                                                 stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
    for tmp, obj, p in procs:
        o, e = p.communicate()
        if p.returncode:
            raise RuntimeError(f"emulator host build failed:\n{o}\n{e}")
        os.replace(tmp, obj)
    return objs


def build(shape: dict, asan: bool = False, msan: bool = False) -> str:
    """Executable emulating the kernel compiled for `shape` (keyword arguments of
    upmabossg_amd.build.kernel_source); cached by the source text and flags."""
    os.makedirs(BUILD, exist_ok=True)
    src = B.kernel_source(**shape)
    with open(os.path.join(HERE, "emu_device.h")) as f:
        dev = f.read()
    with open(os.path.join(HERE, "emu_run.cpp")) as f:
        drv = f.read()
    key = hashlib.sha1((src + dev + drv + repr(_flags(asan, msan))).encode()).hexdigest()[:16]
    exe = os.path.join(BUILD, f"emu_{key}{'_asan' if asan else ''}{'_msan' if msan else ''}")
    # under MemorySanitizer the model parser and packer stay uninstrumented (their outputs are
    # marked written before the kernels run), so that only the kernel code is checked
    objs = _host_objects(asan, False)
    if os.path.exists(exe) and os.path.getmtime(exe) >= max(os.path.getmtime(o) for o in objs):
        return exe
    kfile = os.path.join(BUILD, f"k_{key}.h")
    tmp = f"{exe}.{os.getpid()}.tmp"  # per process: parallel test workers may build the same shape
    with open(f"{kfile}.{os.getpid()}.tmp", "w") as f:
        f.write(src)
    os.replace(f"{kfile}.{os.getpid()}.tmp", kfile)
    cmd = [_cxx(msan), *_flags(asan, msan), f'-DEMU_KERNEL="k_{key}.h"', os.path.join(HERE, "emu_run.cpp"), *objs,
           "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"emulator build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, exe)
    return exe


def _planner() -> str:
    exe = os.path.join(BUILD, "emu_plan")
    objs = _host_objects(False)
    src = os.path.join(HERE, "emu_plan.cpp")
    if not (os.path.exists(exe) and os.path.getmtime(exe) >= max(os.path.getmtime(x) for x in objs + [src])):
        tmp = f"{exe}.{os.getpid()}.tmp"
        r = subprocess.run([CXX, *_flags(False), src, *objs, "-o", tmp], capture_output=True, text=True)
        if r.returncode:
            raise RuntimeError(f"planner build failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, exe)
    return exe


def _job(td, variants, grid=1, compat=0, cell_cap=1 << 16, stride_extra=0, err_cap=8):
    lines = [f"grid {grid}", f"compat {compat}", f"cell_cap {cell_cap}", f"stride_extra {stride_extra}",
             f"err_cap {err_cap}"]
    for i, v in enumerate(variants):
        bnd, cfg = os.path.join(td, f"m{i}.bnd"), os.path.join(td, f"m{i}.cfg")
        with open(bnd, "w") as f:
            f.write(v["bnd"])
        with open(cfg, "w") as f:
            f.write(v["cfg"])
        mu = _f32(os.path.join(td, f"mu{i}"), v.get("mu"))
        md = _f32(os.path.join(td, f"md{i}"), v.get("md"))
        ist = _f32(os.path.join(td, f"is{i}"), v.get("istate"))
        lines.append(f"variant {bnd} {cfg} {v['n']} {v.get('count', v['n'])} {v.get('first_id', 0)} {v['seed']} "
                     f"{v.get('vid', 0)} {mu} {md} {ist}")
    job = os.path.join(td, "job")
    with open(job, "w") as f:
        f.write("\n".join(lines) + "\n")
    return job


def facts(variants):
    """The product's facts of each variant's model (words, readouts, rate widths, compact
    tables, largest in-degree, mean dependents ...), as dicts."""
    with tempfile.TemporaryDirectory() as td:
        job = _job(td, variants)
        r = subprocess.run([_planner(), job], capture_output=True, text=True, timeout=300)
        if r.returncode:
            raise RuntimeError(r.stderr)
        out = []
        for line in r.stdout.splitlines():
            d = {}
            for kv in line.split():
                k, v = kv.split("=")
                d[k] = float(v) if "." in v else int(v)
            out.append(d)
        return out


def shape_for(variants, **over):
    """A launch shape the runtime could choose for these variants (one model, or one sweep class
    with multi=1), with keyword overrides (nw, block, state_hbm, wsum_reg, sumw, ...)."""
    fs = facts(variants)
    nw = max(f["NW"] for f in fs)
    db = over.get("dep_batch", 1)  # the runtime runs one dependent per round
    mean = max(f["mean_deps"] for f in fs)
    sh = dict(nw=nw, block=64, discrete=fs[0]["discrete"], trace=0, generic=int(any(f["generic"] for f in fs)),
              tables_lds=0, max_in=max(1, max(f["max_in"] for f in fs)), dep_batch=db, scan=4,
              wsum_reg=1 if nw <= 16 else 0, rate32=int(all(f["rate32"] for f in fs)), multi=0, sumw=1,
              compact=int(all(f["compact"] for f in fs)), wsmad=0, sparse=int(fs[0]["K"] > 20),
              prog=int(any(f["prog"] for f in fs)), th_lds=1, cells_lds=0,
              dep_cap=max(4, 2 * db, math.ceil(1.5 * mean)), state_hbm=0)
    sh.update(over)
    if sh["multi"] and "sweep_xcd" not in over:
        sh["sweep_xcd"] = 2  # the runtime's default: every wave starts at the class's first variant
    if sh["wsum_reg"] and "wsmad" not in over:
        sh["wsmad"] = int(all(f["wsmad"] for f in fs))
    if not sh["wsum_reg"]:
        sh["wsmad"] = 0
    return sh


def _f32(path, v):
    if v is None:
        return "-"
    np.asarray(v, dtype=np.float32).tofile(path)
    return path


def run(exe: str, variants, grid: int = 1, compat: int = 0, cell_cap: int = 1 << 16, stride_extra: int = 0,
        err_cap: int = 8, timeout: int = 900):
    """variants: list of dicts with bnd, cfg (text), n, count, first_id, seed, vid and optional
    mu, md, istate arrays.  Returns one dict of raw tables per variant."""
    with tempfile.TemporaryDirectory() as td:
        job = _job(td, variants, grid, compat, cell_cap, stride_extra, err_cap)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:detect_stack_use_after_return=0",
                   MSAN_OPTIONS="halt_on_error=0")
        r = subprocess.run([exe, job, td], capture_output=True, text=True, timeout=timeout, env=env)
        if r.returncode:
            raise RuntimeError(f"emulated run failed ({r.returncode}):\n{r.stdout[-4000:]}\n{r.stderr[-4000:]}")
        global last_stderr
        last_stderr = r.stderr
        out = []
        for i in range(len(variants)):
            def rd(name):
                return np.fromfile(os.path.join(td, f"v{i}_{name}.u64"), dtype=np.uint64)
            info = rd("info")
            nw = int(info[6])
            fp = rd("fp").reshape(-1, nw)
            out.append(dict(flips=int(info[0]), nfp=int(info[1]), W=int(info[2]), K=int(info[3]), FP=int(info[4]),
                            FT=int(info[5]), cell_miss=int(info[7]), prob=rd("prob"), th=rd("th"), fin=rd("fin"),
                            fp_records=fp, summ=rd("summ").reshape(-1, 4), cells=rd("cells"), cellv=rd("cellv"),
                            err_miss=int(info[8]), FQ=int(info[9]), FQT=int(info[10]), prob2=rd("prob2"),
                            th2=rd("th2"), cellv2=rd("cellv2")))
        return out


def distinct_fixed_points(records: np.ndarray, words: int):
    """Distinct states (first `words` words of each record) and counts, ordered most significant
    word first (the oracle's and the reference's order)."""
    if len(records) == 0:
        return np.zeros((0, words), np.uint64), np.zeros(0, np.uint64)
    rows = records[:, :words]
    order = np.lexsort(tuple(rows[:, w] for w in range(words)))
    rows = rows[order]
    keep = np.ones(len(rows), bool)
    keep[1:] = np.any(rows[1:] != rows[:-1], axis=1)
    idx = np.flatnonzero(keep)
    counts = np.diff(np.append(idx, len(rows))).astype(np.uint64)
    return rows[idx], counts

"""Build recipe for libmbg.so, the MaBoSSG CLI and the CPU oracle.

    python -m upmabossg_amd.build      (or __graft_entry__.build())

* Embeds the device sources (csrc/device/*) as C strings for hipRTC.
* Compiles the host library with hipcc (host code only; the trajectory kernel is
  compiled at run time by hipRTC for gfx950) and links the CLI against it.
* Compile-checks the generated kernel for gfx950 with hipcc for nine shapes, so a
  device-code error shows up at build time rather than at the first run.
* Compiles the surrogate MLP library (libmbg_mlp.so, BF16 MFMA kernels) for gfx950.
* Builds the CPU oracle (oracle/liboracle.so, test infrastructure).
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
DEV = os.path.join(CSRC, "device")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
# Host code is plain C++ against the HIP runtime API (no device code in these files).
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-pthread", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
              "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]
HOST_LIBS = ["-pthread", "-L" + os.path.join(ROCM, "lib"), "-lhiprtc", "-lamdhip64", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
KERNEL_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_SOURCES = ["front.cpp", "model.cpp", "output.cpp", "pack.cpp", "runtime.cpp", "variants.cpp"]


def _run(cmd, verbose):
    if verbose:
        print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def _c_string(text: str) -> str:
    out = []
    for line in text.splitlines(True):
        esc = line.replace("\\", "\\\\").replace('"', '\\"').replace("\n", "\\n")
        out.append(f'"{esc}"')
    return "\n".join(out) if out else '""'


def embed_sources() -> str:
    with open(os.path.join(DEV, "mbg_math.h")) as f:
        math_src = f.read()
    with open(os.path.join(DEV, "mbg_kernel.hip")) as f:
        kern_src = f.read()
    body = (
        "// Generated by upmabossg_amd/build.py from csrc/device/*: do not edit.\n"
        f"extern const char* const mbg_embedded_math;\nconst char* const mbg_embedded_math =\n{_c_string(math_src)};\n"
        f"extern const char* const mbg_embedded_kernel;\nconst char* const mbg_embedded_kernel =\n{_c_string(kern_src)};\n"
    )
    path = os.path.join(CSRC, "_embedded.cpp")
    with open(path, "w") as f:
        f.write(body)
    return path


def kernel_source(nw: int, block: int = 64, discrete: int = 0, trace: int = 0, generic: int = 0,
                  tables_lds: int = 1, max_in: int = 4, dep_batch: int = 2, scan: int = 4, wsum_reg: int = 0,
                  rate32: int = 0, multi: int = 0, sumw: int = 1, compact: int = 0, wsmad: int = 0,
                  sparse: int = 0, prog: int = 0, th_lds: int = 0, cells_lds: int = 0, dep_cap: int = 0,
                  state_hbm: int = 0, errors: int = 0, sweep_xcd: int = 0) -> str:
    with open(os.path.join(DEV, "mbg_math.h")) as f:
        math_src = f.read()
    with open(os.path.join(DEV, "mbg_kernel.hip")) as f:
        kern_src = f.read()
    return (f"#define MBG_NW {nw}\n#define MBG_BLOCK {block}\n#define MBG_DISCRETE {discrete}\n"
            f"#define MBG_TRACE {trace}\n#define MBG_GENERIC {generic}\n#define MBG_TABLES_LDS {tables_lds}\n"
            f"#define MBG_MAX_IN {max_in}\n#define MBG_DEP_BATCH {dep_batch}\n#define MBG_SCAN {scan}\n"
            f"#define MBG_WSUM_REG {wsum_reg}\n#define MBG_RATE32 {rate32}\n#define MBG_MULTI {multi}\n"
            f"#define MBG_SUMW {sumw}\n#define MBG_COMPACT {compact}\n#define MBG_WSMAD {wsmad}\n"
            f"#define MBG_SPARSE {sparse}\n#define MBG_PROG {prog}\n#define MBG_TH_LDS {th_lds}\n"
            f"#define MBG_CELLS_LDS {cells_lds}\n#define MBG_DEP_CAP {dep_cap}\n#define MBG_STATE_HBM {state_hbm}\n"
            f"#define MBG_ERRORS {errors}\n#define MBG_SWEEP_XCD {sweep_xcd}\n"
            f"#define MBG_HD __device__ __forceinline__\n{math_src}\n{kern_src}\n")


def check_kernel(verbose=False):
    with tempfile.TemporaryDirectory() as td:
        procs = []
        shapes = [(1, 64, 1, 1, 1, 0, 13, 4, 4, 1, 0, 0, 1, 0, 0, 0, 0),
                  (16, 512, 0, 0, 0, 1, 4, 4, 4, 1, 1, 0, 1, 1, 1, 0, 0),
                  (32, 256, 0, 0, 0, 0, 4, 4, 8, 0, 1, 1, 1, 1, 0, 0, 1),
                  (16, 448, 0, 1, 0, 1, 4, 8, 4, 1, 1, 0, 2, 0, 0, 0, 0),
                  (2, 256, 0, 1, 1, 1, 6, 2, 4, 1, 0, 0, 1, 0, 0, 1, 1),
                  (200, 64, 1, 0, 1, 0, 3, 2, 4, 0, 0, 0, 1, 0, 0, 0, 1),
                  (79, 64, 0, 0, 0, 0, 3, 1, 4, 2, 0, 0, 5, 0, 1, 0, 0),
                  (64, 128, 0, 0, 0, 0, 4, 2, 4, 2, 1, 1, 4, 0, 0, 0, 0),
                  (16, 768, 0, 0, 0, 0, 4, 2, 4, 1, 1, 1, 1, 1, 1, 0, 0)]
        for i, (nw, block, disc, trace, gen, lds, mi, db, sc, wr, r32, multi, sumw, cp, mad, sp, pg) in enumerate(shapes):
            src = os.path.join(td, f"k{nw}_{disc}_{trace}_{multi}_{cp}_{sp}_{pg}.hip")
            th = 1  # window-entropy partials in LDS (per block, or per wave in sweeps)
            cells = 1 if (th and not sp and nw <= 2) else 0  # small models: the whole window table too
            cap = 2 * db if i % 2 else 0  # dependents per iteration capped in every other shape
            # lanes' state in HBM (1), only their active sets (2), or all but x (3)
            hbm = 1 if (not multi and i in (3, 5)) else 2 if i in (1, 4, 8) else 3 if i == 6 else 0
            errors = 1 if i in (1, 2, 4) else 0  # the error columns' sums: dense, sweep and sparse
            with open(src, "w") as f:
                f.write(kernel_source(nw, block, disc, trace, gen, lds, mi, db, sc, wr, r32, multi, sumw, cp, mad, sp,
                                      pg, th, cells, cap, hbm, errors, multi))
            cmd = [HIPCC, *KERNEL_FLAGS, "-include", "hip/hip_runtime.h", "--cuda-device-only", "-c", src, "-o", src + ".o"]
            if verbose:
                print("+", " ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        for cmd, p in procs:
            out, err = p.communicate()
            if p.returncode != 0:
                raise RuntimeError(f"kernel compile check failed: {' '.join(cmd)}\n{out}\n{err}")


def build(verbose=False):
    emb = embed_sources()
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in [os.path.join(CSRC, s) for s in HOST_SOURCES] + [emb]:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        cmd = [CXX, *HOST_FLAGS, "-c", src, "-o", obj]
        if verbose:
            print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)))
        objs.append(obj)
    kproc = None
    try:
        import threading

        kerr = []

        def _k():
            try:
                check_kernel(verbose)
            except Exception as e:  # reported after the host build
                kerr.append(e)

        kproc = threading.Thread(target=_k)
        kproc.start()
        for cmd, p in procs:
            out, err = p.communicate()
            if p.returncode != 0:
                raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{out}\n{err}")
        lib = os.path.join(PKG, "libmbg.so")
        _run([CXX, "-shared", "-fPIC", "-o", lib + ".tmp", *objs, *HOST_LIBS], verbose)
        os.replace(lib + ".tmp", lib)
        mlp = os.path.join(PKG, "libmbg_mlp.so")
        _run([HIPCC, *KERNEL_FLAGS, "-fPIC", "-shared", "-o", mlp + ".tmp", os.path.join(CSRC, "mlp", "mbg_mlp.hip")],
             verbose)
        os.replace(mlp + ".tmp", mlp)
        cli = os.path.join(PKG, "MaBoSSG")
        _run([CXX, *HOST_FLAGS, "-o", cli + ".tmp", os.path.join(CSRC, "cli.cpp"), "-L" + PKG, "-lmbg",
              "-Wl,-rpath,$ORIGIN", *HOST_LIBS], verbose)
        os.replace(cli + ".tmp", cli)
        _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)
    finally:
        if kproc is not None:
            kproc.join()
    if kerr:
        raise kerr[0]


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print("built", os.path.join(PKG, "libmbg.so"))

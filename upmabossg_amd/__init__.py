"""MI355X-native MaBoSS-semantics stochastic Boolean network simulator.

Python binding of libmbg.so (C ABI: include/mbg.h).  The simulation runs only
in the HIP trajectory kernel; there is no CPU fallback: if libmbg.so is not
built, importing the binding raises.

    from upmabossg_amd import Model, Context
    m = Model.from_files("model.bnd", "model.cfg")
    with Context() as ctx:
        r = ctx.run(m)                 # sample_count trajectories, seed_pseudorandom
        r.probtraj, r.final_states, r.fixed_points
        ctx.write_csv(m, "out")         # out_probtraj.csv, out_fp.csv

The same run from the shell: ``upmabossg_amd/MaBoSSG -o out model.bnd model.cfg``
(the reference CLI, src/main.cpp:101-169).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

__all__ = ["Model", "Context", "RunResult", "MbgError", "lib", "LIB_PATH", "CLI_PATH"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmbg.so")
CLI_PATH = os.path.join(_HERE, "MaBoSSG")

MBG_OK = 0
_STATUS = {1: "parse error", 2: "limit exceeded", 3: "invalid argument", 4: "HIP error",
           5: "hipRTC compilation error", 6: "no result yet"}


class MbgError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{_STATUS.get(status, status)}: {msg}")
        self.status = status


class _TuneInfo(C.Structure):
    _fields_ = [("block", C.c_int), ("tables_in_lds", C.c_int), ("wsum_reg", C.c_int), ("compact", C.c_int),
                ("dep_batch", C.c_int), ("scan", C.c_int), ("wsmad", C.c_int), ("trials", C.c_int),
                ("sample", C.c_uint64), ("best_ms", C.c_double), ("default_ms", C.c_double), ("dep_cap", C.c_int),
                ("state_hbm", C.c_int)]


class _RunParams(C.Structure):
    _fields_ = [
        ("n_trajectories", C.c_uint64),
        ("first_id", C.c_uint64),
        ("count", C.c_uint64),
        ("seed", C.c_uint64),
        ("use_model_seed", C.c_int),
        ("variant_id", C.c_uint32),
        ("trace", C.c_int),
        ("block", C.c_int),
        ("grid", C.c_int),
        ("compat_batch_resume", C.c_int),
        ("error_columns", C.c_int),
    ]


class _RunInfo(C.Structure):
    _fields_ = [
        ("windows", C.c_int), ("readouts", C.c_int), ("prob_exponent", C.c_int), ("th_exponent", C.c_int),
        ("rate_exponent", C.c_int), ("state_words", C.c_int), ("discrete", C.c_int),
        ("n_trajectories", C.c_uint64), ("simulated", C.c_uint64), ("transitions", C.c_uint64),
        ("fixed_point_trajectories", C.c_uint64), ("kernel_ms", C.c_double),
        ("time_tick", C.c_float), ("max_time", C.c_float),
        ("block", C.c_int), ("tables_in_lds", C.c_int), ("lds_bytes", C.c_int), ("blocks_per_cu", C.c_int),
        ("state_hbm", C.c_int), ("error_columns", C.c_int), ("sq_exponent", C.c_int), ("th_sq_exponent", C.c_int),
    ]


class _Variant(C.Structure):
    _fields_ = [
        ("model", C.c_void_p),
        ("n_trajectories", C.c_uint64),
        ("seed", C.c_uint64),
        ("use_model_seed", C.c_int),
        ("variant_id", C.c_uint32),
    ]


class _SweepInfo(C.Structure):
    _fields_ = [("variants", C.c_int), ("classes", C.c_int), ("trajectories", C.c_uint64),
                ("transitions", C.c_uint64), ("graph_ms", C.c_double), ("setup_ms", C.c_double)]


class _SweepClassInfo(C.Structure):
    _fields_ = [("words", C.c_int), ("discrete", C.c_int), ("variants", C.c_int), ("trajectories", C.c_uint64),
                ("transitions", C.c_uint64), ("kernel_ms", C.c_double)] + \
               [(k, C.c_int) for k in ("block", "blocks_per_cu", "state_hbm", "wsum_reg", "sumw", "dep_batch",
                                       "dep_cap", "scan", "th_lds", "tuned", "tables_lds")]


_LIB = None


def lib():
    """Loads libmbg.so; raises if it is not built (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built: run `python -m upmabossg_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, cp, sz = C.c_void_p, C.c_char_p, C.c_size_t
    u64p, f64p, f32p = C.POINTER(C.c_uint64), C.POINTER(C.c_double), C.POINTER(C.c_float)
    sig = {
        "mbg_model_parse": (C.c_int, [cp, cp, C.POINTER(vp), cp, sz]),
        "mbg_model_parse_files": (C.c_int, [cp, cp, C.POINTER(vp), cp, sz]),
        "mbg_model_free": (None, [vp]),
        "mbg_model_clone": (C.c_int, [vp, C.POINTER(vp)]),
        "mbg_model_num_nodes": (C.c_int, [vp]),
        "mbg_model_node_name": (cp, [vp, C.c_int]),
        "mbg_model_node_istate": (C.c_float, [vp, C.c_int]),
        "mbg_model_node_is_internal": (C.c_int, [vp, C.c_int]),
        "mbg_model_node_is_generic": (C.c_int, [vp, C.c_int]),
        "mbg_model_num_readouts": (C.c_int, [vp]),
        "mbg_model_constant": (C.c_int, [vp, cp, f64p]),
        "mbg_model_set_variant": (C.c_int, [vp, f32p, f32p, f32p]),
        "mbg_model_get_variant": (C.c_int, [vp, f32p, f32p, f32p]),
        "mbg_model_transition_rates": (C.c_int, [vp, u64p, f32p]),
        "mbg_model_generate_source": (C.c_int, [vp, C.c_int, cp, sz, C.POINTER(sz)]),
        "mbg_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "mbg_destroy": (None, [vp]),
        "mbg_last_error": (cp, [vp]),
        "mbg_compile": (C.c_int, [vp, vp, C.c_int]),
        "mbg_run": (C.c_int, [vp, vp, C.POINTER(_RunParams)]),
        "mbg_autotune": (C.c_int, [vp, vp, C.c_uint64, C.c_int, C.POINTER(_TuneInfo)]),
        "mbg_get_run_info": (C.c_int, [vp, C.POINTER(_RunInfo)]),
        "mbg_get_probtraj_raw": (C.c_int, [vp, u64p, u64p]),
        "mbg_get_probtraj": (C.c_int, [vp, f64p, f64p]),
        "mbg_get_final_states": (C.c_int, [vp, u64p]),
        "mbg_get_probtraj_errors": (C.c_int, [vp, f64p, f64p]),
        "mbg_get_probtraj_squares_raw": (C.c_int, [vp, u64p, u64p]),
        "mbg_get_probtraj_cell_squares": (C.c_int, [vp, u64p, sz, C.POINTER(sz)]),
        "mbg_get_probtraj_cells": (C.c_int, [vp, C.POINTER(C.c_uint32), u64p, u64p, sz, C.POINTER(sz)]),
        "mbg_get_final_cells": (C.c_int, [vp, u64p, u64p, sz, C.POINTER(sz)]),
        "mbg_get_fixed_points": (C.c_int, [vp, u64p, u64p, sz, C.POINTER(sz)]),
        "mbg_get_summaries": (C.c_int, [vp, u64p, sz]),
        "mbg_get_marginals": (C.c_int, [vp, f64p]),
        "mbg_run_variants": (C.c_int, [vp, C.POINTER(_Variant), C.c_int]),
        "mbg_run_variants_ex": (C.c_int, [vp, C.POINTER(_Variant), C.c_int, C.c_int]),
        "mbg_autotune_variants": (C.c_int, [vp, C.POINTER(_Variant), C.c_int, C.c_int]),
        "mbg_get_variant_probtraj_squares_raw": (C.c_int, [vp, C.c_int, u64p, u64p]),
        "mbg_get_sweep_info": (C.c_int, [vp, C.POINTER(_SweepInfo)]),
        "mbg_get_sweep_class_info": (C.c_int, [vp, C.c_int, C.POINTER(_SweepClassInfo)]),
        "mbg_get_variant_info": (C.c_int, [vp, C.c_int, C.POINTER(_RunInfo)]),
        "mbg_get_variant_probtraj_raw": (C.c_int, [vp, C.c_int, u64p, u64p]),
        "mbg_get_variant_marginals": (C.c_int, [vp, C.c_int, f64p]),
        "mbg_get_variant_final_states": (C.c_int, [vp, C.c_int, u64p]),
        "mbg_get_variant_fixed_points": (C.c_int, [vp, C.c_int, u64p, u64p, sz, C.POINTER(sz)]),
        "mbg_write_csv": (C.c_int, [vp, vp, cp]),
        "mbg_write_text": (C.c_int, [vp, vp, cp]),
        "mbg_variants_parse": (C.c_int, [vp, cp, sz, C.POINTER(_Variant), cp, C.c_int, C.POINTER(C.c_int), cp, sz]),
        "mbg_variants_parse_file": (C.c_int, [vp, cp, C.POINTER(_Variant), cp, C.c_int, C.POINTER(C.c_int), cp,
                                              sz]),
        "mbg_write_variant_csv": (C.c_int, [vp, C.c_int, cp]),
        "mbg_write_variant_text": (C.c_int, [vp, C.c_int, cp]),
    }
    for name, (rt, args) in sig.items():
        f = getattr(L, name)
        f.restype = rt
        f.argtypes = args
    _LIB = L
    return L


def _u64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint64))


class Model:
    """A parsed and compiled .bnd/.cfg model (host only; no GPU needed)."""

    def __init__(self, bnd_text: str, cfg_text: str):
        L = lib()
        h = C.c_void_p()
        err = C.create_string_buffer(4096)
        st = L.mbg_model_parse(bnd_text.encode(), cfg_text.encode(), C.byref(h), err, 4096)
        if st != MBG_OK:
            raise MbgError(st, err.value.decode())
        self._h = h

    def clone(self) -> "Model":
        """A copy sharing the parsed network (cheap), with the same variant overrides."""
        h = C.c_void_p()
        st = lib().mbg_model_clone(self._h, C.byref(h))
        if st != MBG_OK:
            raise MbgError(st, "clone failed")
        m = Model.__new__(Model)
        m._h = h
        return m

    @classmethod
    def from_files(cls, bnd_path: str, cfg_path: str) -> "Model":
        with open(bnd_path) as f:
            b = f.read()
        with open(cfg_path) as f:
            c = f.read()
        return cls(b, c)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _LIB is not None:
            _LIB.mbg_model_free(h)
            self._h = None

    @property
    def handle(self):
        return self._h

    @property
    def num_nodes(self) -> int:
        return lib().mbg_model_num_nodes(self._h)

    @property
    def num_readouts(self) -> int:
        return lib().mbg_model_num_readouts(self._h)

    def names(self):
        return [lib().mbg_model_node_name(self._h, i).decode() for i in range(self.num_nodes)]

    def istate(self):
        return np.array([lib().mbg_model_node_istate(self._h, i) for i in range(self.num_nodes)], np.float32)

    def internal(self):
        return np.array([lib().mbg_model_node_is_internal(self._h, i) for i in range(self.num_nodes)], bool)

    def generic(self):
        return np.array([lib().mbg_model_node_is_generic(self._h, i) for i in range(self.num_nodes)], bool)

    def constant(self, name: str) -> float:
        v = C.c_double()
        st = lib().mbg_model_constant(self._h, name.encode(), C.byref(v))
        if st != MBG_OK:
            raise KeyError(name)
        return v.value

    def set_variant(self, mult_up=None, mult_down=None, istate=None):
        keep = []

        def ptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            if a.shape != (self.num_nodes,):
                raise ValueError("per-node arrays must have one entry per node")
            keep.append(a)
            return a.ctypes.data_as(C.POINTER(C.c_float))

        st = lib().mbg_model_set_variant(self._h, ptr(mult_up), ptr(mult_down), ptr(istate))
        if st != MBG_OK:
            raise MbgError(st, "invalid variant parameters")

    def transition_rates(self, state_words) -> np.ndarray:
        """FP32 rate of every node at a state given as ceil(N/64) uint64 words."""
        w = np.ascontiguousarray(state_words, dtype=np.uint64)
        out = np.zeros(self.num_nodes, np.float32)
        st = lib().mbg_model_transition_rates(self._h, _u64p(w), out.ctypes.data_as(C.POINTER(C.c_float)))
        if st != MBG_OK:
            raise MbgError(st, "transition_rates failed")
        return out

    def get_variant(self):
        """(mult_up, mult_down, istate) in force, one value per node."""
        N = self.num_nodes
        a = [np.zeros(N, np.float32) for _ in range(3)]
        st = lib().mbg_model_get_variant(self._h, *[x.ctypes.data_as(C.POINTER(C.c_float)) for x in a])
        if st != MBG_OK:
            raise MbgError(st, "get_variant failed")
        return tuple(a)

    def variants(self, text: str):
        """Parses a variants file (text) against this model: a list of
        (name, Model, n_trajectories or 0, seed or None), the models carrying the overrides."""
        L = lib()
        raw = text.encode()
        n = C.c_int()
        err = C.create_string_buffer(4096)
        st = L.mbg_variants_parse(self._h, raw, len(raw), None, None, 0, C.byref(n), err, 4096)
        if st != MBG_OK:
            raise MbgError(st, err.value.decode())
        arr = (_Variant * max(1, n.value))()
        names = C.create_string_buffer(64 * max(1, n.value))
        st = L.mbg_variants_parse(self._h, raw, len(raw), arr, names, n.value, C.byref(n), err, 4096)
        if st != MBG_OK:
            raise MbgError(st, err.value.decode())
        out = []
        for k in range(n.value):
            m = Model.__new__(Model)
            m._h = C.c_void_p(arr[k].model)
            name = names.raw[64 * k:64 * (k + 1)].split(b"\0", 1)[0].decode()
            out.append((name, m, arr[k].n_trajectories, None if arr[k].use_model_seed else arr[k].seed))
        return out

    def generate_source(self, trace=False) -> str:
        n = C.c_size_t()
        lib().mbg_model_generate_source(self._h, int(trace), None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        lib().mbg_model_generate_source(self._h, int(trace), buf, n.value + 1, C.byref(n))
        return buf.value.decode()


@dataclass
class RunResult:
    windows: int
    readouts: int
    prob_exponent: int
    th_exponent: int
    rate_exponent: int
    discrete: bool
    n_trajectories: int
    simulated: int
    transitions: int
    fixed_point_trajectories: int
    kernel_ms: float
    time_tick: float
    max_time: float
    prob_raw: np.ndarray | None      # (W, 2^K) uint64 fixed point (counts when discrete); None above 20 readouts
    th_raw: np.ndarray               # (W,) uint64 fixed point
    probtraj: np.ndarray | None      # (W, 2^K) float64 normalised; None above 20 readouts
    th: np.ndarray                   # (W,) float64
    final_states: np.ndarray | None  # (2^K,) uint64 counts; None above 20 readouts
    fixed_states: np.ndarray         # (F, words) uint64, reference order
    fixed_counts: np.ndarray         # (F,) uint64
    summaries: np.ndarray | None     # (count, 4) uint64 when traced
    # above 20 readouts: the non-zero cells (window uint32, state uint64, raw uint64), window-major
    # with states ascending, and the final states (state, count), states ascending
    cells: tuple | None = None
    final_cells: tuple | None = None
    # error columns (runs with error_columns=True): sums of squared per-trajectory window totals
    # and the standard errors they give (ErrorProba, ErrorTH)
    prob_sq: np.ndarray | None = None      # (W, 2^K) uint64, units 2^-sq_exponent
    th_sq: np.ndarray | None = None        # (W,) uint64, units 2^-th_sq_exponent
    cell_squares: np.ndarray | None = None  # above 20 readouts, aligned with cells
    prob_error: np.ndarray | None = None   # (W, 2^K) float64
    th_error: np.ndarray | None = None     # (W,) float64
    sq_exponent: int = 0
    th_sq_exponent: int = 0


class Context:
    """A HIP device context (one per host thread)."""

    def __init__(self, device: int = 0):
        L = lib()
        h = C.c_void_p()
        st = L.mbg_create(device, C.byref(h))
        if st != MBG_OK:
            raise MbgError(st, f"cannot create a context on HIP device {device}")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.mbg_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, st):
        if st != MBG_OK:
            raise MbgError(st, lib().mbg_last_error(self._h).decode())

    def compile(self, model: Model, trace=False):
        self._check(lib().mbg_compile(self._h, model.handle, int(trace)))

    def run(self, model: Model, n_trajectories=0, count=0, first_id=0, seed=None, variant_id=0, trace=False,
            block=0, grid=0, fetch=True, compat_batch_resume=False, error_columns=False) -> RunResult | None:
        if count and not n_trajectories:
            n_trajectories = count  # a partial run is normalised by its own count unless told otherwise
        p = _RunParams(n_trajectories, first_id, count, 0 if seed is None else seed, int(seed is None),
                       variant_id, int(trace), block, grid, int(compat_batch_resume), int(error_columns))
        self._check(lib().mbg_run(self._h, model.handle, C.byref(p)))
        return self.result(trace) if fetch else None

    def autotune(self, model: Model, sample=0, trace=False) -> dict:
        """Measure candidate launch shapes on `sample` trajectories and keep the fastest for
        later runs of models of this structure (mbg_autotune)."""
        t = _TuneInfo()
        self._check(lib().mbg_autotune(self._h, model.handle, sample, int(trace), C.byref(t)))
        return {k: getattr(t, k) for k, _ in _TuneInfo._fields_}

    def info(self) -> _RunInfo:
        i = _RunInfo()
        self._check(lib().mbg_get_run_info(self._h, C.byref(i)))
        return i

    def cells(self):
        """The last run's non-zero window cells (window, state, raw value), window-major with
        states ascending, and its final states (state, count); any number of readouts."""
        L = lib()
        n = C.c_size_t()
        self._check(L.mbg_get_probtraj_cells(self._h, None, None, None, 0, C.byref(n)))
        w = np.zeros(n.value, np.uint32)
        st = np.zeros(n.value, np.uint64)
        v = np.zeros(n.value, np.uint64)
        self._check(L.mbg_get_probtraj_cells(self._h, w.ctypes.data_as(C.POINTER(C.c_uint32)), _u64p(st), _u64p(v),
                                             n.value, C.byref(n)))
        self._check(L.mbg_get_final_cells(self._h, None, None, 0, C.byref(n)))
        fs = np.zeros(n.value, np.uint64)
        fc = np.zeros(n.value, np.uint64)
        self._check(L.mbg_get_final_cells(self._h, _u64p(fs), _u64p(fc), n.value, C.byref(n)))
        return (w, st, v), (fs, fc)

    def result(self, with_summaries=False) -> RunResult:
        L = lib()
        i = self.info()
        W = i.windows
        th_raw = np.zeros(W, np.uint64)
        th = np.zeros(W, np.float64)
        cells = final_cells = None
        if i.readouts > 20:
            self._check(L.mbg_get_probtraj_raw(self._h, None, _u64p(th_raw)))
            self._check(L.mbg_get_probtraj(self._h, None, th.ctypes.data_as(C.POINTER(C.c_double))))
            cells, final_cells = self.cells()
            prob_raw = prob = fin = None
        else:
            S = 1 << i.readouts
            prob_raw = np.zeros(W * S, np.uint64)
            self._check(L.mbg_get_probtraj_raw(self._h, _u64p(prob_raw), _u64p(th_raw)))
            prob = np.zeros(W * S, np.float64)
            self._check(L.mbg_get_probtraj(self._h, prob.ctypes.data_as(C.POINTER(C.c_double)),
                                           th.ctypes.data_as(C.POINTER(C.c_double))))
            fin = np.zeros(S, np.uint64)
            self._check(L.mbg_get_final_states(self._h, _u64p(fin)))
            prob_raw, prob = prob_raw.reshape(W, S), prob.reshape(W, S)
        n = C.c_size_t()
        self._check(L.mbg_get_fixed_points(self._h, None, None, 0, C.byref(n)))
        fs = np.zeros((n.value, i.state_words), np.uint64)
        fc = np.zeros(n.value, np.uint64)
        if n.value:
            self._check(L.mbg_get_fixed_points(self._h, _u64p(fs), _u64p(fc), n.value, C.byref(n)))
        summ = None
        if with_summaries:
            summ = np.zeros((i.simulated, 4), np.uint64)
            self._check(L.mbg_get_summaries(self._h, _u64p(summ), i.simulated))
        res = RunResult(W, i.readouts, i.prob_exponent, i.th_exponent, i.rate_exponent, bool(i.discrete),
                        i.n_trajectories, i.simulated, i.transitions, i.fixed_point_trajectories, i.kernel_ms,
                        i.time_tick, i.max_time, prob_raw, th_raw, prob, th, fin, fs, fc, summ, cells, final_cells)
        if i.error_columns:
            res.sq_exponent, res.th_sq_exponent = i.sq_exponent, i.th_sq_exponent
            res.th_sq = np.zeros(W, np.uint64)
            res.th_error = np.zeros(W, np.float64)
            if i.readouts > 20:
                self._check(L.mbg_get_probtraj_squares_raw(self._h, None, _u64p(res.th_sq)))
                self._check(L.mbg_get_probtraj_errors(self._h, None, res.th_error.ctypes.data_as(C.POINTER(C.c_double))))
                n = C.c_size_t()
                self._check(L.mbg_get_probtraj_cell_squares(self._h, None, 0, C.byref(n)))
                res.cell_squares = np.zeros(n.value, np.uint64)
                self._check(L.mbg_get_probtraj_cell_squares(self._h, _u64p(res.cell_squares), n.value, C.byref(n)))
            else:
                S = 1 << i.readouts
                res.prob_sq = np.zeros(W * S, np.uint64)
                res.prob_error = np.zeros(W * S, np.float64)
                self._check(L.mbg_get_probtraj_squares_raw(self._h, _u64p(res.prob_sq), _u64p(res.th_sq)))
                self._check(L.mbg_get_probtraj_errors(self._h, res.prob_error.ctypes.data_as(C.POINTER(C.c_double)),
                                                      res.th_error.ctypes.data_as(C.POINTER(C.c_double))))
                res.prob_sq, res.prob_error = res.prob_sq.reshape(W, S), res.prob_error.reshape(W, S)
        return res

    def probtraj_raw(self, out=None) -> np.ndarray:
        """The last run's window table (W * 2^K exact cells), copied to host memory."""
        i = self.info()
        n = i.windows * (1 << i.readouts)
        if out is None or out.size != n:
            out = np.zeros(n, np.uint64)
        self._check(lib().mbg_get_probtraj_raw(self._h, _u64p(out), None))
        return out

    def marginals(self) -> np.ndarray:
        """(W, K) probability that each readout is ON, per window, of the last run."""
        i = self.info()
        out = np.zeros(i.windows * i.readouts, np.float64)
        self._check(lib().mbg_get_marginals(self._h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(i.windows, i.readouts)

    # ---- variant sweeps ----
    @staticmethod
    def _variant_array(vs):
        arr = (_Variant * len(vs))()
        for k, (model, n, seed, vid) in enumerate(vs):
            arr[k] = _Variant(model.handle.value if hasattr(model.handle, "value") else model.handle, n,
                              0 if seed is None else seed, int(seed is None), vid)
        return arr

    def run_variants(self, variants, error_columns=False):
        """variants: iterable of (Model, n_trajectories, seed or None, variant_id).  All run in one
        HIP graph launch; returns the sweep info (per-variant results through the getters below).
        error_columns: also accumulate the error columns' sums of squares (variant_squares)."""
        vs = list(variants)
        arr = self._variant_array(vs)
        self._sweep_models = [v[0] for v in vs]  # keep the models alive with the device results
        self._check(lib().mbg_run_variants_ex(self._h, arr, len(vs), 1 if error_columns else 0))
        return self.sweep_info()

    def autotune_variants(self, variants, error_columns=False):
        """Measures each word class of these variants under candidate layouts and keeps the fastest
        for later sweeps that form the same classes (mbg_autotune_variants)."""
        vs = list(variants)
        arr = self._variant_array(vs)
        self._sweep_models = [v[0] for v in vs]
        self._check(lib().mbg_autotune_variants(self._h, arr, len(vs), 1 if error_columns else 0))

    def variant_squares(self, v):
        """(prob_sq (W, 2^K), th_sq (W,)) of variant v of a sweep run with error_columns."""
        i = self.variant_info(v)
        W, S = i.windows, 1 << i.readouts
        prob = np.zeros(W * S, np.uint64)
        th = np.zeros(W, np.uint64)
        self._check(lib().mbg_get_variant_probtraj_squares_raw(self._h, v, _u64p(prob), _u64p(th)))
        return prob.reshape(W, S), th

    def sweep_info(self) -> _SweepInfo:
        i = _SweepInfo()
        self._check(lib().mbg_get_sweep_info(self._h, C.byref(i)))
        return i

    def sweep_classes(self):
        """Per word class of the last sweep: its launch shape, counts and kernel span, as dicts."""
        out = []
        for i in range(self.sweep_info().classes):
            ci = _SweepClassInfo()
            self._check(lib().mbg_get_sweep_class_info(self._h, i, C.byref(ci)))
            out.append({k: getattr(ci, k) for k, _ in _SweepClassInfo._fields_})
        return out

    def variant_info(self, v) -> _RunInfo:
        i = _RunInfo()
        self._check(lib().mbg_get_variant_info(self._h, v, C.byref(i)))
        return i

    def variant_result(self, v):
        """(info, prob_raw (W, 2^K), th_raw (W,), final (2^K,), fixed_states, fixed_counts)."""
        L = lib()
        i = self.variant_info(v)
        W, S = i.windows, 1 << i.readouts
        prob = np.zeros(W * S, np.uint64)
        th = np.zeros(W, np.uint64)
        self._check(L.mbg_get_variant_probtraj_raw(self._h, v, _u64p(prob), _u64p(th)))
        fin = np.zeros(S, np.uint64)
        self._check(L.mbg_get_variant_final_states(self._h, v, _u64p(fin)))
        n = C.c_size_t()
        self._check(L.mbg_get_variant_fixed_points(self._h, v, None, None, 0, C.byref(n)))
        fs = np.zeros((n.value, i.state_words), np.uint64)
        fc = np.zeros(n.value, np.uint64)
        if n.value:
            self._check(L.mbg_get_variant_fixed_points(self._h, v, _u64p(fs), _u64p(fc), n.value, C.byref(n)))
        return i, prob.reshape(W, S), th, fin, fs, fc

    def variant_marginals(self, v) -> np.ndarray:
        i = self.variant_info(v)
        out = np.zeros(i.windows * i.readouts, np.float64)
        self._check(lib().mbg_get_variant_marginals(self._h, v, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(i.windows, i.readouts)

    def write_variant_csv(self, v, prefix: str):
        self._check(lib().mbg_write_variant_csv(self._h, v, prefix.encode()))

    def write_variant_text(self, v, path: str = "-"):
        self._check(lib().mbg_write_variant_text(self._h, v, path.encode()))

    def write_csv(self, model: Model, prefix: str):
        self._check(lib().mbg_write_csv(self._h, model.handle, prefix.encode()))

    def write_text(self, model: Model, path: str = "-"):
        self._check(lib().mbg_write_text(self._h, model.handle, path.encode()))

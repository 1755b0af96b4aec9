"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OracleError(RuntimeError):
    pass


class _Params(C.Structure):
    _fields_ = [
        ("n_trajectories", C.c_uint64),
        ("first_id", C.c_uint64),
        ("count", C.c_uint64),
        ("seed", C.c_uint64),
        ("variant_id", C.c_uint32),
        ("threads", C.c_int),
        ("record_fixed", C.c_int),
        ("record_summaries", C.c_int),
        ("compat_batch_resume", C.c_int),
    ]


SUMMARY_DTYPE = np.dtype(
    [("flips", "<u8"), ("flip_hash", "<u8"), ("final_time", "<f8"), ("status", "<u4"), ("pad", "<u4")]
)


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        raise OracleError(f"{path} is not built (run __graft_entry__.build())")
    L = C.CDLL(path)
    vp, u64p, f32p = C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_float)
    L.orc_load_text.restype = vp
    L.orc_load_text.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]
    L.orc_free.argtypes = [vp]
    for n, rt in [("orc_num_nodes", C.c_int), ("orc_rate_exponent", C.c_int)]:
        getattr(L, n).restype = rt
        getattr(L, n).argtypes = [vp]
    L.orc_node_name.restype = C.c_char_p
    L.orc_node_name.argtypes = [vp, C.c_int]
    L.orc_node_istate.restype = C.c_float
    L.orc_node_istate.argtypes = [vp, C.c_int]
    L.orc_node_is_internal.restype = C.c_int
    L.orc_node_is_internal.argtypes = [vp, C.c_int]
    L.orc_node_is_fast.restype = C.c_int
    L.orc_node_is_fast.argtypes = [vp, C.c_int]
    L.orc_constant.restype = C.c_double
    L.orc_constant.argtypes = [vp, C.c_char_p]
    L.orc_variable.restype = C.c_double
    L.orc_variable.argtypes = [vp, C.c_char_p]
    L.orc_set_variant.restype = C.c_int
    L.orc_set_variant.argtypes = [vp, f32p, f32p, f32p]
    L.orc_node_rate_f32.restype = C.c_float
    L.orc_node_rate_f32.argtypes = [vp, C.c_int, u64p]
    L.orc_node_rate_fixed.restype = C.c_uint64
    L.orc_node_rate_fixed.argtypes = [vp, C.c_int, u64p]
    L.orc_run.restype = vp
    L.orc_run.argtypes = [vp, C.POINTER(_Params), C.c_char_p, C.c_size_t]
    L.orc_result_free.argtypes = [vp]
    for n in ["orc_windows", "orc_readouts", "orc_prob_exponent", "orc_th_exponent", "orc_state_words",
              "orc_sq_exponent", "orc_th_sq_exponent"]:
        getattr(L, n).restype = C.c_int
        getattr(L, n).argtypes = [vp]
    for n in ["orc_prob_table", "orc_th_table", "orc_final_table", "orc_fixed_states", "orc_fixed_counts",
              "orc_prob_sq_table", "orc_th_sq_table", "orc_cell_squares"]:
        getattr(L, n).restype = u64p
        getattr(L, n).argtypes = [vp]
    for n in ["orc_cell_states", "orc_cell_values", "orc_final_cell_states", "orc_final_cell_counts"]:
        getattr(L, n).restype = u64p
        getattr(L, n).argtypes = [vp]
    L.orc_cell_windows.restype = C.POINTER(C.c_uint32)
    L.orc_cell_windows.argtypes = [vp]
    for n in ["orc_total_flips", "orc_fixed_count", "orc_cell_count", "orc_final_cell_count"]:
        getattr(L, n).restype = C.c_uint64
        getattr(L, n).argtypes = [vp]
    L.orc_elapsed_seconds.restype = C.c_double
    L.orc_elapsed_seconds.argtypes = [vp]
    L.orc_summaries.restype = vp
    L.orc_summaries.argtypes = [vp]
    L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    L.orc_ln.restype = C.c_double
    L.orc_ln.argtypes = [C.c_double]
    L.orc_log2.restype = C.c_double
    L.orc_log2.argtypes = [C.c_double]
    L.orc_neglog_u32.restype = C.c_double
    L.orc_neglog_u32.argtypes = [C.c_uint32]
    _LIB = L
    return L


def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox4x32_10(c, k, o)
    return list(o)


@dataclass
class OracleResult:
    windows: int
    readouts: int
    prob_exponent: int
    th_exponent: int
    prob: np.ndarray | None  # (W, S) uint64 fixed point; None above 20 readouts
    th: np.ndarray  # (W,) uint64 fixed point
    final: np.ndarray | None  # (S,) uint64 counts; None above 20 readouts
    flips: int
    seconds: float
    fixed_states: np.ndarray  # (F, words) uint64, reference order
    fixed_counts: np.ndarray  # (F,) uint64
    summaries: np.ndarray | None
    cells: tuple = None  # non-zero cells (window uint32, state uint64, value uint64), window-major
    final_cells: tuple = None  # non-zero final states (state, count), ascending
    # error columns: sums of squared per-trajectory window totals (units 2^-sq_exponent of the
    # squared window fraction, 2^-th_sq_exponent of the squared window-average TH)
    prob_sq: np.ndarray | None = None  # (W, S) uint64; None above 20 readouts
    th_sq: np.ndarray | None = None  # (W,) uint64
    cell_squares: np.ndarray | None = None  # aligned with cells
    sq_exponent: int = 0
    th_sq_exponent: int = 0


class OracleModel:
    def __init__(self, bnd_text: str, cfg_text: str):
        L = lib()
        err = C.create_string_buffer(4096)
        h = L.orc_load_text(bnd_text.encode(), cfg_text.encode(), err, 4096)
        if not h:
            raise OracleError(err.value.decode())
        self._h = h

    @classmethod
    def from_files(cls, bnd_path, cfg_path):
        with open(bnd_path) as f:
            b = f.read()
        with open(cfg_path) as f:
            c = f.read()
        return cls(b, c)

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.orc_free(self._h)
            self._h = None

    @property
    def n(self):
        return lib().orc_num_nodes(self._h)

    def names(self):
        return [lib().orc_node_name(self._h, i).decode() for i in range(self.n)]

    def istate(self):
        return np.array([lib().orc_node_istate(self._h, i) for i in range(self.n)], dtype=np.float32)

    def internal(self):
        return np.array([lib().orc_node_is_internal(self._h, i) for i in range(self.n)], dtype=bool)

    def fast(self):
        return np.array([lib().orc_node_is_fast(self._h, i) for i in range(self.n)], dtype=bool)

    def constant(self, name):
        return lib().orc_constant(self._h, name.encode())

    def variable(self, name):
        return lib().orc_variable(self._h, name.encode())

    @property
    def rate_exponent(self):
        return lib().orc_rate_exponent(self._h)

    def set_variant(self, mult_up=None, mult_down=None, istate=None):
        def arr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            return a.ctypes.data_as(C.POINTER(C.c_float)), a

        keep = [arr(mult_up), arr(mult_down), arr(istate)]
        ptrs = [k[0] if k else None for k in keep]
        if lib().orc_set_variant(self._h, *ptrs) != 0:
            raise OracleError("invalid variant")

    def _words(self, state_bits):
        w = np.zeros((self.n + 63) // 64, dtype=np.uint64)
        for i, b in enumerate(state_bits):
            if b:
                w[i >> 6] |= np.uint64(1) << np.uint64(i & 63)
        return w

    def rate_f32(self, i, state_bits):
        w = self._words(state_bits)
        return lib().orc_node_rate_f32(self._h, i, w.ctypes.data_as(C.POINTER(C.c_uint64)))

    def rate_fixed(self, i, state_bits):
        w = self._words(state_bits)
        return lib().orc_node_rate_fixed(self._h, i, w.ctypes.data_as(C.POINTER(C.c_uint64)))

    def run(self, count=None, first_id=0, n_trajectories=None, seed=None, variant_id=0, threads=1,
            record_fixed=True, record_summaries=False, compat_batch_resume=False) -> OracleResult:
        L = lib()
        sc = int(self.constant("sample_count"))
        if count is None:
            count = sc
        if n_trajectories is None:
            n_trajectories = count
        if seed is None:
            seed = int(self.constant("seed_pseudorandom"))
        p = _Params(n_trajectories, first_id, count, seed, variant_id, threads, int(record_fixed),
                    int(record_summaries), int(compat_batch_resume))
        err = C.create_string_buffer(4096)
        r = L.orc_run(self._h, C.byref(p), err, 4096)
        if not r:
            raise OracleError(err.value.decode())
        try:
            W = L.orc_windows(r)
            K = L.orc_readouts(r)
            th = np.ctypeslib.as_array(L.orc_th_table(r), shape=(W,)).copy()
            th_sq = np.ctypeslib.as_array(L.orc_th_sq_table(r), shape=(W,)).copy()
            if K > 20:
                prob = fin = prob_sq = None
            else:
                S = 1 << K
                prob = np.ctypeslib.as_array(L.orc_prob_table(r), shape=(W * S,)).copy().reshape(W, S)
                prob_sq = np.ctypeslib.as_array(L.orc_prob_sq_table(r), shape=(W * S,)).copy().reshape(W, S)
                fin = np.ctypeslib.as_array(L.orc_final_table(r), shape=(S,)).copy()

            def arr(ptr, n, dt):
                return np.ctypeslib.as_array(ptr, shape=(n,)).copy() if n else np.zeros(0, dt)

            nc, nf = L.orc_cell_count(r), L.orc_final_cell_count(r)
            cells = (arr(L.orc_cell_windows(r), nc, np.uint32), arr(L.orc_cell_states(r), nc, np.uint64),
                     arr(L.orc_cell_values(r), nc, np.uint64))
            final_cells = (arr(L.orc_final_cell_states(r), nf, np.uint64),
                           arr(L.orc_final_cell_counts(r), nf, np.uint64))
            F = L.orc_fixed_count(r)
            NW = L.orc_state_words(r)
            if F:
                fs = np.ctypeslib.as_array(L.orc_fixed_states(r), shape=(F * NW,)).copy().reshape(F, NW)
                fc = np.ctypeslib.as_array(L.orc_fixed_counts(r), shape=(F,)).copy()
            else:
                fs = np.zeros((0, NW), dtype=np.uint64)
                fc = np.zeros((0,), dtype=np.uint64)
            summ = None
            sp = L.orc_summaries(r)
            if sp:
                buf = (C.c_char * (count * SUMMARY_DTYPE.itemsize)).from_address(sp)
                summ = np.frombuffer(bytes(buf), dtype=SUMMARY_DTYPE).copy()
            return OracleResult(W, K, L.orc_prob_exponent(r), L.orc_th_exponent(r), prob, th, fin,
                                L.orc_total_flips(r), L.orc_elapsed_seconds(r), fs, fc, summ, cells, final_cells,
                                prob_sq, th_sq, arr(L.orc_cell_squares(r), nc, np.uint64), L.orc_sq_exponent(r),
                                L.orc_th_sq_exponent(r))
        finally:
            L.orc_result_free(r)

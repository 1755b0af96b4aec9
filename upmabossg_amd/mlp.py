"""ctypes binding of libmbg_mlp.so (C ABI: include/mbg_mlp.h), the sweep surrogate's network.

Training and inference run only in the library's BF16 MFMA kernels on the GPU; there is no
CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
MLP_LIB_PATH = os.path.join(_HERE, "libmbg_mlp.so")
_LIB = None


class MlpError(RuntimeError):
    def __init__(self, msg, status=None):
        super().__init__(msg)
        self.status = status


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(MLP_LIB_PATH):
        raise ImportError(f"{MLP_LIB_PATH} is not built: run `python -m upmabossg_amd.build`")
    L = C.CDLL(MLP_LIB_PATH)
    vp, f32p = C.c_void_p, C.POINTER(C.c_float)
    sig = {
        "mbg_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint64, C.POINTER(vp)]),
        "mbg_mlp_destroy": (None, [vp]),
        "mbg_mlp_last_error": (C.c_char_p, [vp]),
        "mbg_mlp_num_params": (C.c_size_t, [vp]),
        "mbg_mlp_padded_inputs": (C.c_int, [vp]),
        "mbg_mlp_get_params": (C.c_int, [vp, f32p]),
        "mbg_mlp_set_params": (C.c_int, [vp, f32p]),
        "mbg_mlp_gradients": (C.c_int, [vp, f32p, f32p, C.c_int, f32p, f32p]),
        "mbg_mlp_load_dataset": (C.c_int, [vp, f32p, f32p, C.c_size_t]),
        "mbg_mlp_train": (C.c_int, [vp, C.c_int, C.c_int, C.c_float, C.c_uint64, f32p]),
        "mbg_mlp_predict": (C.c_int, [vp, f32p, C.c_size_t, f32p, C.POINTER(C.c_double)]),
        "mbg_mlp_ensemble_predict": (C.c_int, [C.POINTER(vp), C.c_int, f32p, C.c_size_t, f32p, f32p,
                                               C.POINTER(C.c_double)]),
        "mbg_mlp_ensemble_predict_noalloc": (C.c_int, [C.POINTER(vp), C.c_int, vp, C.c_size_t, vp, vp]),
    }
    for name, (rt, args) in sig.items():
        f = getattr(L, name)
        f.restype = rt
        f.argtypes = args
    _LIB = L
    return L


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Mlp:
    """n_in -> hidden x hidden_layers -> n_out, GELU, sigmoid output, MSE, Adam (BF16 MFMA)."""

    def __init__(self, n_in, hidden, hidden_layers, n_out, max_batch=1024, seed=0, device=0):
        L = lib()
        h = C.c_void_p()
        st = L.mbg_mlp_create(device, n_in, hidden, hidden_layers, n_out, max_batch, seed, C.byref(h))
        if st != 0:
            raise MlpError(f"mbg_mlp_create failed ({st})")
        self._h = h
        self.n_in, self.hidden, self.hidden_layers, self.n_out = n_in, hidden, hidden_layers, n_out
        self.max_batch = max_batch

    def close(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.mbg_mlp_destroy(self._h)
            self._h = None

    __del__ = close

    def _check(self, st):
        if st != 0:
            raise MlpError(lib().mbg_mlp_last_error(self._h).decode())

    @property
    def num_params(self):
        return lib().mbg_mlp_num_params(self._h)

    @property
    def padded_inputs(self):
        return lib().mbg_mlp_padded_inputs(self._h)

    def layer_dims(self):
        dims_in = [self.padded_inputs] + [self.hidden] * self.hidden_layers
        dims_out = [self.hidden] * self.hidden_layers + [self.n_out]
        return list(zip(dims_in, dims_out))

    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        self._check(lib().mbg_mlp_get_params(self._h, _fp(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self._check(lib().mbg_mlp_set_params(self._h, _fp(p)))

    def unpack(self, p):
        """[(W [out][in_padded], b [out]), ...] views of a parameter vector."""
        out, off = [], 0
        for din, dout in self.layer_dims():
            W = p[off:off + din * dout].reshape(dout, din)
            off += din * dout
            b = p[off:off + dout]
            off += dout
            out.append((W, b))
        return out

    def gradients(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_float()
        self._check(lib().mbg_mlp_gradients(self._h, _fp(x), _fp(y), x.shape[0], C.byref(loss), _fp(g)))
        return loss.value, g

    def load_dataset(self, x, y):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        self._check(lib().mbg_mlp_load_dataset(self._h, _fp(x), _fp(y), x.shape[0]))

    def train(self, epochs, batch=256, lr=1e-3, seed=0):
        loss = C.c_float()
        self._check(lib().mbg_mlp_train(self._h, epochs, batch, lr, seed, C.byref(loss)))
        return loss.value

    def predict(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.n_out), np.float32)
        ms = C.c_double()
        self._check(lib().mbg_mlp_predict(self._h, _fp(x), x.shape[0], _fp(y), C.byref(ms)))
        return y, ms.value


def ensemble_predict(members, x):
    """(mean [n][n_out], spread [n], device ms) of networks `members` on rows x [n][n_in]."""
    x = np.ascontiguousarray(x, np.float32)
    n = x.shape[0]
    out = members[0].n_out
    mean = np.zeros((n, out), np.float32)
    spread = np.zeros(n, np.float32)
    ms = C.c_double()
    handles = (C.c_void_p * len(members))(*[m._h.value for m in members])
    st = lib().mbg_mlp_ensemble_predict(handles, len(members), _fp(x), n, _fp(mean), _fp(spread), C.byref(ms))
    if st != 0:
        raise MlpError(lib().mbg_mlp_last_error(members[0]._h).decode())
    return mean, spread, ms.value


def ensemble_predict_noalloc(members, x_dev, n, mean_dev, spread_dev=None):
    """The ensemble prediction on device buffers (addresses as ints: x [n][n_in] float32, mean
    [n][n_out], spread [n] or None), with no device allocation (mbg_mlp_ensemble_predict_noalloc)."""
    handles = (C.c_void_p * len(members))(*[m._h.value for m in members])
    st = lib().mbg_mlp_ensemble_predict_noalloc(handles, len(members), C.c_void_p(x_dev), n, C.c_void_p(mean_dev),
                                                C.c_void_p(spread_dev) if spread_dev else None)
    if st != 0:
        raise MlpError(lib().mbg_mlp_last_error(members[0]._h).decode(), st)
    return st


def _bench(argv=None):
    """Inference benchmark of the surrogate network: `members` random networks of
    600 -> 512 x 3 -> 1600 on `points` rows (device time and TFLOP/s of the whole pass)."""
    import argparse
    import json
    import time

    ap = argparse.ArgumentParser(prog="python -m upmabossg_amd.mlp")
    ap.add_argument("--points", type=int, default=10000)
    ap.add_argument("--members", type=int, default=4)
    ap.add_argument("--reps", type=int, default=5)
    a = ap.parse_args(argv)
    members = [Mlp(600, 512, 3, 1600, max_batch=10240, seed=s) for s in range(a.members)]
    x = np.random.default_rng(0).uniform(-1, 1, (a.points, 600)).astype(np.float32)
    dims = [members[0].padded_inputs, 512, 512, 512, 1600]
    flops = sum(2 * p * q for p, q in zip(dims[:-1], dims[1:])) * a.points * a.members
    ensemble_predict(members, x)
    for _ in range(a.reps):
        t0 = time.perf_counter()
        _, _, ms = ensemble_predict(members, x)
        wall = time.perf_counter() - t0
        print(json.dumps({"points": a.points, "members": a.members, "device_ms": ms, "wall_ms": wall * 1e3,
                          "tflops": flops / (ms / 1e3) / 1e12}), flush=True)


if __name__ == "__main__":
    _bench()

"""mbg_mlp_ensemble_predict_noalloc (the ensemble prediction on device buffers, members' outputs
staged in the library's fixed device scratch array) against mbg_mlp_ensemble_predict (host
buffers, allocated staging): identical means and spreads on two calls in a row, the second
smaller, on one chunk and on several chunks of rows; no change in free device memory across a
call; and a clear error for an ensemble too wide for the scratch."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_HIP = None


def hip():
    global _HIP
    if _HIP is None:
        _HIP = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    return _HIP


class DevArray:
    """A device copy of a float32 array (hipMalloc / hipMemcpy through the HIP runtime)."""

    def __init__(self, a):
        self.shape, self.nbytes = a.shape, a.nbytes
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(4, a.nbytes))) == 0
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def get(self):
        out = np.empty(self.shape, np.float32)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(self.nbytes), 2) == 0
        return out

    def free(self):
        hip().hipFree(self.p)


def _free_bytes():
    free, total = C.c_size_t(), C.c_size_t()
    assert hip().hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_no_alloc_ensemble_matches_the_allocating_entry_point():
    from upmabossg_amd.mlp import Mlp, ensemble_predict, ensemble_predict_noalloc

    members = [Mlp(100, 128, 2, 192, max_batch=256, seed=s) for s in range(3)]
    rng = np.random.default_rng(4)
    for n in (250, 70, 700):  # one chunk, a smaller one, then three chunks of rows
        print(f"case: no-alloc ensemble, 3 members, n={n}", flush=True)
        x = rng.normal(size=(n, 100)).astype(np.float32)
        want_mean, want_spread, _ = ensemble_predict(members, x)
        dx = DevArray(x)
        dm = DevArray(np.full((n, 192), -1.0, np.float32))
        ds = DevArray(np.full(n, -1.0, np.float32))
        before = _free_bytes()
        ensemble_predict_noalloc(members, dx.p.value, n, dm.p.value, ds.p.value)
        assert _free_bytes() == before, "the call allocated device memory"
        assert np.array_equal(dm.get(), want_mean)
        assert np.array_equal(ds.get(), want_spread)
        ensemble_predict_noalloc(members, dx.p.value, n, dm.p.value, None)  # spread optional
        assert np.array_equal(dm.get(), want_mean)
        for a in (dx, dm, ds):
            a.free()
    for m in members:
        m.close()


def test_no_alloc_ensemble_rejects_members_too_wide_for_the_scratch():
    from upmabossg_amd.mlp import Mlp, MlpError, ensemble_predict_noalloc

    print("case: no-alloc ensemble, n_out 262208 (beyond the scratch)", flush=True)
    m = Mlp(8, 64, 1, 262208, max_batch=64, seed=1)  # 64 rows x 262,208 outputs > 16,777,216 floats
    dx = DevArray(np.zeros((4, 8), np.float32))
    dm = DevArray(np.zeros((4, 262208), np.float32))
    with pytest.raises(MlpError) as e:
        ensemble_predict_noalloc([m], dx.p.value, 4, dm.p.value, None)
    assert e.value.status == 3 and "no-allocation scratch" in str(e.value)
    dx.free()
    dm.free()
    m.close()

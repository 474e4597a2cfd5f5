"""gpum_hip_grad_batch_noalloc (device buffers in, scratch in the library's
static array, no device allocation) against gpum_hip_grad_batch, through the
C ABI.  Device buffers come from the HIP runtime the library itself links.

Two calls in a row, the second smaller, each equal to the allocating entry
point's result bit for bit (both run the interpreted FP64 kernels, and the
compiled kernels reproduce those bits), so stale scratch from the first call
cannot leak into the second.  The first call's partial sums exceed one
scratch chunk, so the points are processed in several chunks.
"""
import ctypes as C
import os

import numpy as np
import pytest

from gpumetropolis_amd import _native, workloads

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Dev:
    """hipMalloc / hipMemcpy of the runtime libgpum_hip.so is linked to."""

    def __init__(self):
        _native.lib()
        self.rt = C.CDLL(os.path.join("/opt/rocm/lib", "libamdhip64.so"))
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.bufs.append(p)
        if a.nbytes:
            assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return p

    def empty(self, n):
        return self.put(np.full(n, np.nan))

    def get(self, p, n):
        out = np.empty(n, np.float64)
        assert self.rt.hipMemcpy(out.ctypes.data, p, out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.rt.hipFree(p)
        self.bufs = []


def _noalloc():
    L = _native.lib()
    f = L.gpum_hip_grad_batch_noalloc
    f.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_double),
                  C.c_size_t, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32,
                  C.c_void_p, C.c_uint32, C.c_void_p]
    f.restype = C.c_int
    return f


def _call(f, m, np_, data, pts, dev):
    code = np.ascontiguousarray(m.loglik.code, np.uint32)
    consts = np.ascontiguousarray(m.loglik.consts, np.float64)
    n_obs, n_cols = data.shape
    d_data = dev.put(data.T.copy())            # column-major
    d_pts = dev.put(pts)
    d_out = dev.empty(pts.shape[0] * np_)
    rc = f(code.ctypes.data_as(C.POINTER(C.c_uint32)), code.size,
           consts.ctypes.data_as(C.POINTER(C.c_double)), consts.size, np_,
           d_data, n_cols, n_obs, d_pts, pts.shape[0], d_out)
    assert rc == 0, _native.last_error()
    return dev.get(d_out, pts.shape[0] * np_).reshape(-1, np_)


def test_noalloc_gradient_equals_allocating_entry(hip):
    f = _noalloc()
    dev = _Dev()
    try:
        rng = np.random.default_rng(3)
        wl = workloads.logistic(n=1_000_000, d=16, seed=9)
        pts = wl.meta["true"] + 0.05 * rng.standard_normal((300, 16))
        print("[noalloc] logistic-16, 1e6 rows, 300 points", flush=True)
        got = _call(f, wl.model, 16, wl.data, pts, dev)
        want = hip.grad_batch(wl.model.loglik.code, wl.model.loglik.consts, 16,
                              wl.data, wl.n_cols, wl.n_obs, pts)
        assert np.array_equal(got, want)
        dev.close()
        wg = workloads.gamma(n=5000)
        pts = np.log([2.5, 0.7]) + 0.01 * rng.standard_normal((7, 2))
        print("[noalloc] gamma, 5000 rows, 7 points", flush=True)
        got = _call(f, wg.model, 2, wg.data, pts, dev)
        want = hip.grad_batch(wg.model.loglik.code, wg.model.loglik.consts, 2,
                              wg.data, 1, wg.n_obs, pts)
        assert np.array_equal(got, want)
        assert np.all(np.isfinite(got))
    finally:
        dev.close()


def test_noalloc_rejects_what_does_not_fit(hip):
    f = _noalloc()
    dev = _Dev()
    try:
        wg = workloads.gamma(n=10)
        m = wg.model
        code = np.ascontiguousarray(m.loglik.code, np.uint32)
        consts = np.ascontiguousarray(m.loglik.consts, np.float64)
        d = dev.put(np.ones(64))
        cp = code.ctypes.data_as(C.POINTER(C.c_uint32))
        kp = consts.ctypes.data_as(C.POINTER(C.c_double))
        # more than 16 parameters
        assert f(cp, code.size, kp, consts.size, 17, d, 1, 10, d, 1, d) != 0
        assert "16 parameters" in _native.last_error()
        # partial sums beyond the scratch (checked before any device work)
        assert f(cp, code.size, kp, consts.size, 2, d, 1, 0xFFFFFFFF, d, 1, d) != 0
        assert "scratch" in _native.last_error()
        # an empty point set is a no-op
        assert f(cp, code.size, kp, consts.size, 2, d, 1, 10, d, 0, d) == 0
    finally:
        dev.close()

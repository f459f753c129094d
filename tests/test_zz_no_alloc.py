"""mh_mlp_gradient_noalloc against mh_mlp_gradient through the C ABI.

The caller owns device inputs and outputs (allocated here with the HIP runtime
the library itself links); the entry point allocates nothing and must return
the same bits as mh_mlp_gradient on the same rows, on two calls in a row with
the second one smaller (stale partials of the first call in the library's
static window must not leak into the second).
"""
import ctypes

import numpy as np
import pytest

import mechanic_amd as mh
from mechanic_amd import _lib
from mechanic_amd.learned import WarmStartNet

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    _lib.load()  # the library's own runtime is loaded first; the same soname resolves to it
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    return hip


class DeviceArray:
    def __init__(self, hip, nbytes):
        self.hip, self.nbytes = hip, nbytes
        self.ptr = ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        assert self.hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.hip.hipFree(self.ptr)


def _rows(rng, n):
    f = rng.normal(size=(n, 16)).astype(np.float32) * np.float32(0.7) + np.float32(0.1)
    t = rng.normal(size=(n, 3)).astype(np.float32)
    return f, t


def test_noalloc_gradient_matches_allocating_gradient_on_consecutive_calls():
    if mh.device_count() < 1:
        pytest.skip("no HIP device")
    hip = _hip()
    rng = np.random.default_rng(17)
    net = WarmStartNet(seed=5)
    p = net.param_count
    cap = 10_000
    feat = DeviceArray(hip, cap * 16 * 4)
    targ = DeviceArray(hip, cap * 3 * 4)
    grad = DeviceArray(hip, p * 4)
    loss = DeviceArray(hip, 4)
    try:
        # 10,000 rows: 313 blocks of 32, three windows; then 1,000 rows: one window.
        for n in (10_000, 1_000):
            f, t = _rows(rng, n)
            net.set_dataset(f, t)  # also sets the normalisation both calls use
            g_ref, l_ref = net.gradient()
            feat.upload(f)
            targ.upload(t)
            net.gradient_noalloc(feat.ptr.value, targ.ptr.value, n, grad.ptr.value, loss.ptr.value)
            g = grad.download(p, np.float32)
            l = loss.download(1, np.float32)[0]
            print(f"[noalloc] rows {n}: |g| {np.linalg.norm(g):.6g} loss {l:.6g}", flush=True)
            assert np.array_equal(g.view(np.uint32), g_ref.view(np.uint32)), n
            assert np.float32(l).view(np.uint32) == np.float32(l_ref).view(np.uint32), n
        with pytest.raises(Exception):
            net.gradient_noalloc(feat.ptr.value, targ.ptr.value, 0, grad.ptr.value, loss.ptr.value)
    finally:
        for a in (feat, targ, grad, loss):
            a.free()

"""pfx_est_sums_noalloc (no device allocation; fixed scratch array) against pfc_est_sums.

The caller owns the particle streams and the output (allocated here with the HIP runtime the
library itself links); the entry point must give pfc_est_sums' bits for the same particles,
on two calls in a row, the second with fewer particles (the scratch keeps the first call's
contents), and reject sizes beyond its maximum without launching anything.
"""
import ctypes as C

import numpy as np
import pytest

from cuda_pf_localization_amd import PfKernels, _lib

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    h = C.CDLL("libamdhip64.so.7")  # the runtime libpf_kernels_hip.so is linked against
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    return h


def _state(n, seed):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) * 20.0).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w = np.exp(-rng.gamma(1.0, 4.0, size=n))
    return pos, q.astype(np.float32), (w / w.sum()).astype(np.float32)


def _noalloc_sums(hip, L, pos, quat, w):
    n = len(w)
    streams = [pos[:, 0], pos[:, 1], pos[:, 2], quat[:, 0], quat[:, 1], quat[:, 2], quat[:, 3], w]
    bufs = []
    try:
        for a in streams:
            p = C.c_void_p()
            assert hip.hipMalloc(C.byref(p), 4 * n) == 0
            bufs.append(p)
            a = np.ascontiguousarray(a, np.float32)
            assert hip.hipMemcpy(p, a.ctypes.data, 4 * n, H2D) == 0
        out = C.c_void_p()
        assert hip.hipMalloc(C.byref(out), 8 * 14) == 0
        bufs.append(out)
        assert L.pfx_est_sums_noalloc(*bufs[:8], n, out, None) == 0
        assert hip.hipDeviceSynchronize() == 0
        res = np.zeros(14)
        assert hip.hipMemcpy(res.ctypes.data, out, 8 * 14, D2H) == 0
        return res
    finally:
        for p in bufs:
            hip.hipFree(p)


def test_noalloc_est_sums_equals_pfc_est_sums(gpu):
    hip = _hip()
    L = _lib.kernels()
    for n, seed in ((300_000, 1), (1_029, 2)):  # second call smaller, ragged last tile
        print(f"pfx_est_sums_noalloc n={n}", flush=True)
        pos, quat, w = _state(n, seed)
        k = PfKernels(256)
        k.init(n, 5)
        k.set_state(pos, quat, w)
        ref = k.est_sums()
        k.close()
        got = _noalloc_sums(hip, L, pos, quat, w)
        assert np.array_equal(got, ref), f"n={n}: {got} vs {ref}"


def test_noalloc_rejects_sizes_beyond_its_scratch(gpu):
    L = _lib.kernels()
    d = C.c_void_p(1)  # never dereferenced: the size checks return first
    assert L.pfx_est_sums_noalloc(*([d] * 8), 0, d, None) == -1
    assert L.pfx_est_sums_noalloc(*([d] * 8), 16_777_217, d, None) == -1
    assert L.pfx_est_sums_noalloc(*([None] * 8), 1000, d, None) == -1

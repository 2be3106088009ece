"""doma_net_forward_noalloc (no device allocation: weights and activations in
the library's static scratch array) against doma_net_forward (which allocates
per call), through the C ABI, on two calls in a row -- the second with a smaller
grid -- plus the size limit.  Device buffers for the test come from the HIP
runtime the library itself loaded."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from doma_amd import capi, unet

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    capi.lib()  # the library loads the HIP runtime first
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    return hip


def _dev(hip, a):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), a.nbytes) == 0
    assert hip.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
    return p


def _noalloc(hip, d_params, count, dims, feats):
    D, H, W = dims
    out = np.zeros(D * H * W, dtype=np.float32)
    d_f = _dev(hip, feats)
    d_o = _dev(hip, out)
    try:
        rc = capi.lib().doma_net_forward_noalloc(d_params, count, D, H, W, d_f, d_o, None)
        assert rc == 0, capi.lib().doma_last_error()
        assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_o, out.nbytes, D2H) == 0
    finally:
        hip.hipFree(d_f)
        hip.hipFree(d_o)
    return out.reshape(D, H, W)


def test_noalloc_forward_matches_allocating_forward():
    hip = _hip()
    params = unet.bf16_round(unet.init_params(seed=11))
    rng = np.random.default_rng(12)
    d_params = _dev(hip, params)
    try:
        for dims in [(16, 20, 24), (8, 12, 16)]:  # second call smaller
            print(f"noalloc forward {dims}", flush=True)
            f = rng.standard_normal(dims + (unet.CIN,)).astype(np.float32)
            want = capi.net_forward(params, f)
            got = _noalloc(hip, d_params, params.size, dims, f)
            assert np.array_equal(got, want), np.abs(got - want).max()
        # beyond the scratch limit: rejected before any GPU work
        D = 64
        rc = capi.lib().doma_net_forward_noalloc(d_params, params.size, D, D, D, d_params,
                                                 d_params, None)
        assert rc == capi.DOMA_ERR_INVALID
        assert b"scratch limit" in capi.lib().doma_last_error()
    finally:
        hip.hipFree(d_params)

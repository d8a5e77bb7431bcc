"""dn_calc_res_noalloc (device arrays, no allocation) against dn_calc_res
(host arrays), called through the C-ABI: two calls in a row, the second
with fewer traces and samples, bit-identical residuals and equal FP64 sums;
the maximum trace count is enforced.  Device memory comes from the HIP
runtime the library itself is linked against."""
import ctypes as C
import sys

import numpy as np
import pytest

from common import product

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2   # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    product().lib()    # loads the library and with it its HIP runtime
    # by soname: the runtime copy already loaded for the library (a second
    # copy opened by path would not share its device state)
    hip = C.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


def _dev(hip, nbytes):
    p = C.c_void_p()
    assert hip.hipMalloc(C.byref(p), nbytes) == 0
    return p


def _noalloc(L, hip, syn, obs, dt):
    ntr, ns = syn.shape
    n = syn.nbytes
    d_syn, d_obs, d_diff, d_out = _dev(hip, n), _dev(hip, n), _dev(hip, n), _dev(hip, 16)
    try:
        assert hip.hipMemcpy(d_syn, syn.ctypes.data, n, H2D) == 0
        assert hip.hipMemcpy(d_obs, obs.ctypes.data, n, H2D) == 0
        junk = np.full_like(syn, 7.0)
        assert hip.hipMemcpy(d_diff, junk.ctypes.data, n, H2D) == 0     # must be overwritten
        rc = L.dn_calc_res_noalloc(ntr, ns, C.c_float(dt), d_syn, d_obs, d_diff, d_out, None)
        assert rc == 0, L.dn_last_error(None)
        assert hip.hipDeviceSynchronize() == 0
        diff = np.empty_like(syn)
        out = np.empty(2, np.float64)
        back = np.empty_like(obs)
        assert hip.hipMemcpy(diff.ctypes.data, d_diff, n, D2H) == 0
        assert hip.hipMemcpy(out.ctypes.data, d_out, 16, D2H) == 0
        assert hip.hipMemcpy(back.ctypes.data, d_obs, n, D2H) == 0
        assert np.array_equal(back, obs)          # obs is read only
        return diff, float(out[0]), float(out[1])
    finally:
        for p in (d_syn, d_obs, d_diff, d_out):
            hip.hipFree(p)


def _bind(L):
    L.dn_calc_res_noalloc.argtypes = [C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_void_p]


def test_noalloc_matches_calc_res_twice():
    dn = product()
    L = dn.lib()
    _bind(L)
    hip = _hip()
    dt = 2e-4
    for ntr, ns in ((300, 1500), (41, 257)):
        print("dn_calc_res_noalloc ntr", ntr, "ns", ns)
        sys.stdout.flush()
        rng = np.random.default_rng(ntr + ns)
        syn = rng.standard_normal((ntr, ns)).astype(np.float32)
        obs = (syn + 0.5 * rng.standard_normal((ntr, ns))).astype(np.float32)
        diff, l2, en = _noalloc(L, hip, syn, obs, dt)
        rdiff, rl2, ren, _ = dn.calc_res(syn, obs, dt)
        assert np.array_equal(diff, rdiff)
        assert l2 == rl2 and en == ren


def test_noalloc_rejects_too_many_traces():
    dn = product()
    L = dn.lib()
    _bind(L)
    # rejected on the host before any device work: the pointers are never used
    rc = L.dn_calc_res_noalloc(65537, 4, C.c_float(1e-3), C.c_void_p(16), C.c_void_p(16), C.c_void_p(16),
                               C.c_void_p(16), None)
    assert rc != 0
    assert b"maximum" in L.dn_last_error(None)

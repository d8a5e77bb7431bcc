"""GPU parity of dn_calc_res (calc_res.c + calc_energy.c on one shot's host
seismograms) with the oracle's orc_residual."""
import ctypes as C
import sys

import numpy as np
import pytest

from common import oracle, product

pytestmark = pytest.mark.gpu


def oracle_residual(syn, obs, dt):
    O = oracle("f32")
    ntr, ns = syn.shape
    ob = obs.copy()
    diff = np.zeros_like(syn)
    l2, en = C.c_double(0.0), C.c_double(0.0)
    rp = O.rp
    O.fn("orc_residual")(ntr, ns, C.c_double(dt), syn.ctypes.data_as(rp), ob.ctypes.data_as(rp),
                         diff.ctypes.data_as(rp), C.byref(l2), C.byref(en))
    return diff, l2.value, en.value, ob


@pytest.mark.parametrize("ntr,ns", [(37, 513), (240, 2000), (1, 1), (3, 2)])
def test_calc_res_matches_oracle(ntr, ns):
    print("dn_calc_res ntr", ntr, "ns", ns)
    sys.stdout.flush()
    rng = np.random.default_rng(ntr * 1000 + ns)
    syn = rng.standard_normal((ntr, ns)).astype(np.float32)
    obs = (syn + 0.3 * rng.standard_normal((ntr, ns))).astype(np.float32)
    dt = 1e-4
    diff, l2, en, ob = product().calc_res(syn, obs, dt)
    rdiff, rl2, ren, rob = oracle_residual(syn, obs, float(np.float32(dt)))
    assert np.array_equal(diff, rdiff)
    assert np.array_equal(ob, rob)
    # per-trace FP64 partials summed in trace order vs one running FP64 sum
    assert abs(l2 - rl2) <= 1e-12 * abs(rl2)
    assert abs(en - ren) <= 1e-12 * abs(ren)

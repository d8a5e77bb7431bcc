"""Pins of the oracle's BF16 snapshot storage (orc_snap_bf16: the product's
reduced-precision option; the reference stores FP32 snapshots only).  CPU
only.

  rounding     with a power-of-two scale, every normal value decodes to
               torch's float32 -> bfloat16 cast of itself (round to nearest
               even, an independent implementation), and ties round to even;
  scale        the exponent of snapshot k comes from the maximum of snapshot
               k - 1 of the same field: a snapshot of 1.5 * 2^-60 after one
               with maximum 2^90 flushes to zero (scale 2^-91), and is kept
               exactly after one with maximum 1; the first snapshot and one
               after an all-zero snapshot are unscaled; the exponent is clamped
               at +-100;
  gradient     the flag reaches the gradient: BF16 storage moves it by a
               relative L2 between 1e-6 and 1e-2 (8 significant bits on
               every stored value), and FP32 storage is unchanged.
"""
import numpy as np
import pytest
import torch

from common import model, oracle, rel_l2, ricker

O32 = oracle("f32")


def bf16_torch(x):
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def test_decoded_values_are_bf16_roundings():
    rng = np.random.default_rng(5)
    nsnap, nxny = 6, 999
    s = np.empty((nsnap, 3, nxny), np.float32)
    for k in range(nsnap):
        for f in range(3):
            # magnitudes change by many octaves between snapshots
            s[k, f] = rng.standard_normal(nxny).astype(np.float32) * np.float32(10.0 ** rng.uniform(-12, 8))
    s[3, 1] = 0.0
    x = s.copy()
    O32.snap_bf16(x)
    assert np.array_equal(x, bf16_torch(s))
    err = np.abs(x - s) / np.maximum(np.abs(s), 1e-30)
    assert err[s != 0].max() <= 2.0 ** -8     # unit roundoff of 8 significant bits
    assert np.any(x != s)


def test_ties_round_to_even():
    # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: even (1) wins;
    # 1 + 3 * 2^-8 lies halfway between 1 + 2^-7 and 1 + 2^-6: even (1 + 2^-6) wins
    s = np.zeros((1, 3, 4), np.float32)
    s[0, 0] = [1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 1.0 + 2.0 ** -8 + 2.0 ** -20]
    O32.snap_bf16(s)
    assert s[0, 0].tolist() == [1.0, 1.0 + 2.0 ** -6, -1.0, 1.0 + 2.0 ** -7]


def test_scale_comes_from_the_previous_snapshot():
    small = np.float32(1.5 * 2.0 ** -60)
    s = np.zeros((4, 3, 8), np.float32)
    s[0, 0] = np.float32(2.0 ** 90)     # snapshot 0: unscaled
    s[1, 0] = small                     # scale 2^-91 (max of snapshot 0 is 2^90 = 0.5 * 2^91)
    s[2, 0] = np.float32(1.0)           # scale 2^59 from snapshot 1: exact
    s[3, 0] = small                     # scale 2^-1 from snapshot 2 (max 1 = 0.5 * 2^1): exact
    s[0, 1] = np.float32(2.0 ** 120)    # field 1: exponent 121 clamped to 100
    s[1, 1] = np.float32(2.0 ** -30)    # 2^-130 after scaling: a float32 subnormal, kept exactly
    s[1, 2] = small                     # field 2: snapshot 0 all zero, so unscaled
    x = s.copy()
    O32.snap_bf16(x)
    assert np.all(x[0, 0] == s[0, 0])
    assert np.all(x[1, 0] == 0.0)       # 1.5 * 2^-151 is below the smallest subnormal
    assert np.all(x[2, 0] == 1.0)
    assert np.all(x[3, 0] == small)
    assert np.all(x[1, 1] == np.float32(2.0 ** -30))
    assert np.all(x[1, 2] == small)
    # with the exponent unclamped (121) the same value would fall to 2^-151 and flush:
    # the clamp keeps it
    assert np.all(x[0, 1] == s[0, 1])


@pytest.mark.parametrize("L", [0, 2])
def test_gradient_with_bf16_snapshots(L):
    from pyoracle import make_cfg
    nx, ny, nt, dt, dh = 80, 36, 500, 1e-4, 0.25
    fl = [6.25, 50.0][:L]
    vs, rho, q = model(nx, ny, dh, seed=4, qs=20.0 if L else None)
    kw = dict(nx=nx, ny=ny, fdorder=8, fw=10, free_surf=1, L=L, nt=nt, dtinv=3, dt=dt, dh=dh, fpml=60.0, fl=fl)
    c32, cbf = make_cfg(**kw), make_cfg(snap_bf16=1, **kw)
    rec_i = np.arange(14, nx - 12, 3, dtype=np.int32)
    rec_j = np.full_like(rec_i, 2)
    src_i = np.array([20, 60], np.int32)
    src_j = np.array([3, 3], np.int32)
    sig = np.stack([ricker(nt, dt, 60.0), ricker(nt, dt, 60.0, tshift=0.001)])
    vt = vs * np.float32(1.04)
    obs = np.stack([O32.forward(c32, vt, rho, q, src_i[s:s + 1], src_j[s:s + 1], sig[s], rec_i, rec_j)[0]
                    for s in range(2)])
    a = O32.gradient(c32, vs, rho, q, src_i, src_j, sig, rec_i, rec_j, obs)
    b = O32.gradient(cbf, vs, rho, q, src_i, src_j, sig, rec_i, rec_j, obs)
    a2 = O32.gradient(c32, vs, rho, q, src_i, src_j, sig, rec_i, rec_j, obs)
    e = rel_l2(b["grad_vs"], a["grad_vs"])
    assert 1e-6 < e < 1e-2
    assert 1e-6 < rel_l2(b["grad_rho"], a["grad_rho"]) < 1e-2
    assert np.array_equal(a["grad_vs"], a2["grad_vs"])
    assert b["l2"] == a["l2"]          # the misfit comes from the forward seismograms

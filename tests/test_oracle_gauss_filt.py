"""Pins of the oracle's gauss_filt (SPATFILTER 3, gauss_filt.c:56-160): one
Gaussian for the whole grid, sigma = hfs / 2 in integer arithmetic, weights
normalised by their sum, the gradient continued with its edge values.  CPU
only.

  delta      a unit impulse far from the edges returns the weights, which
             equal the closed-form normalised Gaussian of the integer sigma
             (float64) to float rounding;
  constant   a constant gradient is kept (the weights sum to one);
  correlate  a random gradient equals scipy.ndimage.correlate with the
             closed-form kernel (mode "nearest", the edge continuation) to
             float rounding, and differs from the kernel of a real-valued
             sigma = hfs / 2 (a plausible slip) by far more;
  sizes      FILT_SIZE_GRAD = round(WD_DAMP C_vs / FC_END / DH), made odd;
             0 leaves the gradient alone; hfs < 2 (sigma 0) and negative
             sizes are rejected.
"""
import numpy as np
import pytest
from scipy import ndimage

from common import oracle

O32 = oracle("f32")


def closed_kernel(hfs, sigma):
    ii = np.arange(-hfs, hfs + 1, dtype=np.float64)
    k = np.exp(-(ii[None, :] ** 2) / (2 * sigma * sigma) - (ii[:, None] ** 2) / (2 * sigma * sigma))
    return k / k.sum()      # [jj][ii]: rows y, columns x


def size_of(dh, wd, fc, cvs):
    fs = int(np.round(np.float32(np.float32(wd) * np.float32(np.float32(cvs) / np.float32(fc))) / np.float32(dh)))
    return fs + 1 if fs % 2 == 0 and fs > 0 else fs


# (dh, WD_DAMP, FC_END, C_vs): FILT_SIZE_GRAD 7 (hfs 3, sigma 1; the real-valued
# slip 1.5) and 11 (hfs 5, sigma 2; slip 2.5)
CASES = [(0.25, 0.35, 40.0, 200.0), (0.2, 0.55, 50.0, 200.0)]


@pytest.mark.parametrize("case", CASES)
def test_delta_returns_the_closed_form_kernel(case):
    dh, wd, fc, cvs = case
    hfs = size_of(dh, wd, fc, cvs) // 2
    ny, nx = 4 * hfs + 9, 4 * hfs + 13
    g = np.zeros((ny, nx), np.float32)
    cy, cx = ny // 2, nx // 2
    g[cy, cx] = 1.0
    out, rc = O32.gauss_filt(g, dh, wd, fc, cvs)
    assert rc == 0
    k = out[cy - hfs:cy + hfs + 1, cx - hfs:cx + hfs + 1]      # correlation with a delta: the flipped kernel
    want = closed_kernel(hfs, hfs // 2)[::-1, ::-1]
    assert np.abs(out).sum() == pytest.approx(1.0, rel=1e-5)
    assert np.max(np.abs(k - want)) <= 1e-6 * want.max()
    assert np.count_nonzero(out) == (2 * hfs + 1) ** 2


@pytest.mark.parametrize("case", CASES)
def test_constant_is_kept(case):
    dh, wd, fc, cvs = case
    g = np.full((30, 41), 3.5, np.float32)
    out, rc = O32.gauss_filt(g, dh, wd, fc, cvs)
    assert rc == 0
    assert np.max(np.abs(out - 3.5)) <= 3.5 * 1e-5


@pytest.mark.parametrize("case", CASES)
def test_matches_scipy_correlate_with_edge_continuation(case):
    dh, wd, fc, cvs = case
    hfs = size_of(dh, wd, fc, cvs) // 2
    g = np.random.default_rng(3).standard_normal((37, 53)).astype(np.float32)
    out, rc = O32.gauss_filt(g, dh, wd, fc, cvs)
    assert rc == 0
    want = ndimage.correlate(g.astype(np.float64), closed_kernel(hfs, hfs // 2), mode="nearest")
    err = np.max(np.abs(out - want)) / np.max(np.abs(want))
    slip = ndimage.correlate(g.astype(np.float64), closed_kernel(hfs, hfs / 2.0), mode="nearest")
    err_slip = np.max(np.abs(out - slip)) / np.max(np.abs(want))
    assert err < 1e-5
    assert err_slip > 100 * err


def test_sizes_and_rejections():
    g = np.random.default_rng(4).standard_normal((20, 30)).astype(np.float32)
    out, rc = O32.gauss_filt(g, 0.25, 0.0, 40.0, 200.0)         # FILT_SIZE_GRAD 0: nothing
    assert rc == 1 and np.array_equal(out, g)
    _, rc = O32.gauss_filt(g, 0.25, 0.05, 40.0, 200.0)          # size 1 (hfs 0)
    assert rc == -1
    _, rc = O32.gauss_filt(g, 0.25, 0.15, 40.0, 200.0)          # size 3 (hfs 1, sigma 0)
    assert rc == -1
    _, rc = O32.gauss_filt(g, 0.25, 0.25, 40.0, 200.0)          # size 5 (hfs 2, sigma 1)
    assert rc == 0
    _, rc = O32.gauss_filt(g, 0.25, 0.2, 40.0, 200.0)           # size 4 made odd: 5
    assert rc == 0
    _, rc = O32.gauss_filt(g, 0.25, -0.25, 40.0, 200.0)         # negative size
    assert rc == -1
    assert [size_of(0.25, w, 40.0, 200.0) for w in (0.05, 0.15, 0.2, 0.25)] == [1, 3, 5, 5]
    assert size_of(0.25, 0.35, 40.0, 200.0) == 7 and size_of(0.2, 0.55, 50.0, 200.0) == 11

"""Pins of the oracle's smooth2 (SPATFILTER 2, smooth2.c:28-165 with tripd,
util.c:638-671).  CPU only.

The reference loads the gradient into v[1..NX][1..NY] of a zeroed matrix
v[0..NX][0..NY] and runs its tridiagonal smoothing (r = WD_DAMP^2 / 4) over
the indices 0..n-1 in x, then 1..n-1 in z, so the zero row and column 0 take
part and the last row and column are never touched.

  kept        the last column (x = NX) and the last row (y = NY) come back
              unchanged, all other cells change;
  constant    a constant gradient is kept, to 1e-6, in cells farther than
              40 smoothing lengths from the top and the left edge (the zero
              row and column pull it down there);
  structure   the result equals a float64 model built with
              scipy.linalg.solve_banded from the stated systems (x: diagonal
              1 + 2r with 1 + r at both ends over 0..NX-1; z: 1 + 2r with
              1 + r at the bottom over 1..NY-1, the zero row as the top
              neighbour); the same model with 1-based, whole-grid systems
              (a plausible reading of the intent) does not;
  linear      smooth2(a x + b y) = a smooth2(x) + b smooth2(y) to rounding.
"""
import numpy as np
from scipy.linalg import solve_banded

from common import oracle

O32 = oracle("f32")


def tri(n, r, top, bottom):
    ab = np.zeros((3, n))
    ab[0, 1:] = -r
    ab[2, :-1] = -r
    ab[1, :] = 1.0 + 2.0 * r
    ab[1, 0] = top
    ab[1, -1] = bottom
    return ab


def model(g, wd, shifted=True):
    ny, nx = g.shape
    r = (wd * wd) * 0.25
    if not shifted:
        out = g.astype(np.float64).T.copy()                                  # [x][z], whole grid
        ax = tri(nx, r, 1 + r, 1 + r)
        for iz in range(ny):
            out[:, iz] = solve_banded((1, 1), ax, out[:, iz])
        az = tri(ny, r, 1 + r, 1 + r)
        for ix in range(nx):
            out[ix, :] = solve_banded((1, 1), az, out[ix, :])
        return out.T
    V = np.zeros((nx + 1, ny + 1))
    V[1:, 1:] = g.T
    ax = tri(nx, r, 1 + r, 1 + r)
    for iz in range(ny):
        V[:nx, iz] = solve_banded((1, 1), ax, V[:nx, iz])
    az = tri(ny - 1, r, 1 + 2 * r, 1 + r)
    for ix in range(nx):
        rhs = V[ix, 1:ny].copy()
        rhs[0] += r * V[ix, 0]
        V[ix, 1:ny] = solve_banded((1, 1), az, rhs)
    return V[1:, 1:].T


def field(ny, nx, seed):
    return np.random.default_rng(seed).standard_normal((ny, nx)).astype(np.float32)


def test_last_row_and_column_are_kept():
    g = field(30, 45, 1)
    out = O32.smooth2(g, 3.0)
    assert np.array_equal(out[-1, :], g[-1, :]) and np.array_equal(out[:, -1], g[:, -1])
    assert np.all(out[:-1, :-1] != g[:-1, :-1])


def test_constant_far_from_top_and_left():
    wd = 4.0                              # r = 4: smoothing length about 2 cells
    g = np.full((160, 170), 2.5, np.float32)
    out = O32.smooth2(g, wd)
    assert np.max(np.abs(out[80:, 80:] - 2.5)) <= 2.5e-6
    assert out[0, 100] < 2.0 and out[100, 0] < 2.0


def test_structure_matches_the_banded_model():
    g = field(40, 57, 2)
    out = O32.smooth2(g, 3.0)
    want = model(g, 3.0)
    err = np.max(np.abs(out - want)) / np.max(np.abs(want))
    err_intent = np.max(np.abs(out - model(g, 3.0, shifted=False))) / np.max(np.abs(want))
    assert err < 1e-5
    assert err_intent > 1e3 * err


def test_linear():
    x, y = field(25, 33, 3), field(25, 33, 4)
    a, b = np.float32(1.5), np.float32(-0.75)
    lhs = O32.smooth2(a * x + b * y, 2.0)
    rhs = a * O32.smooth2(x, 2.0) + b * O32.smooth2(y, 2.0)
    assert np.max(np.abs(lhs - rhs)) <= 1e-5 * np.max(np.abs(rhs))

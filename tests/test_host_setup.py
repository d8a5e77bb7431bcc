"""Host-side setup of the product (CPU, no GPU): the source wavelets
(wavelet.c:12-96) and the Butterworth filter of the signals and data
(timedomain_filt.c:10-54 with libcseife) against the oracle, bit for bit.
The oracle's filter is itself pinned to libcseife built from the reference
sources and to scipy's bilinear Butterworth (tests/test_oracle_pins.py).
Also the grid-size check of dn_create, which runs before any device call."""
import numpy as np
import pytest

from common import oracle, product


@pytest.mark.parametrize("quellart", [1, 2, 4, 5, 6])
def test_wavelet_equals_oracle(quellart):
    dn, O = product(), oracle("f32")
    for nt, dt, tshift, fc, amp in ((900, 1e-4, 0.0, 40.0, 1.0), (3000, 5e-5, 0.0123, 25.0, 2.5)):
        try:
            want = O.wavelet(quellart, nt, dt, tshift, fc, amp)
        except AssertionError:
            pytest.skip("the oracle has no QUELLART %d" % quellart)
        got = dn.wavelet(quellart, nt, dt, tshift, fc, amp)
        assert np.max(np.abs(want)) > 0
        assert np.array_equal(got, want.astype(np.float32))


@pytest.mark.parametrize("order", [1, 2, 3, 4, 6])
@pytest.mark.parametrize("method", [1, 2])
def test_timedomain_filt_equals_oracle(order, method):
    dn, O = product(), oracle("f32")
    rng = np.random.default_rng(10 * order + method)
    data = rng.standard_normal((5, 1700)).astype(np.float32)
    for dt, fc in ((1e-4, 30.0), (5e-5, 60.0)):
        got = dn.timedomain_filt(data, dt, fc, order, method)
        want = O.timedomain_filt(data, dt, fc, order, method)
        assert np.array_equal(got, want)
        assert not np.array_equal(got, data)


@pytest.mark.parametrize("nx,ny,L", [(20000, 20000, 2), (30000, 20000, 0)])
def test_create_rejects_grids_beyond_32_bit_offsets(nx, ny, L):
    """dn_create checks the grid size before it touches the device: one
    padded plane times 16 + 8L bytes must stay below 2 GiB."""
    dn = product()
    pitch = (nx + 80 + 31) // 32 * 32
    assert pitch * (ny + 24) * (16 + 8 * L) >= 2 ** 31
    kw = dict(fdorder=8, maxrelerror=1, fw=10, free_surf=1, L=L, dt=1e-4, dh=0.25)
    if L:
        kw["fl"] = [6.25, 50.0][:L]
    with pytest.raises(dn.DeniseError) as e:
        dn.Context(nx, ny, **kw)
    assert "too large" in str(e.value)

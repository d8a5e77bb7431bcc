"""Error handling of the C-ABI (GPU context, host-side checks only): bad
arguments are rejected with a message before any launch, and the context
stays usable after a rejected call.  The instability check follows
denise.c:1183-1185 (a NaN in the centre cell after a pass)."""
import sys

import numpy as np
import pytest

from common import model, product, ricker

pytestmark = pytest.mark.gpu

DT, DH = 1e-4, 0.25


def say(*a):
    print(*a)
    sys.stdout.flush()


def _ctx(nx=80, ny=40, **kw):
    dn = product()
    p = dict(fdorder=4, maxrelerror=1, fw=8, free_surf=1, L=0, dt=DT, dh=DH, damping=300.0, fpml=60.0)
    p.update(kw)
    return dn, dn.Context(nx, ny, **p)


@pytest.mark.parametrize("kw,text", [
    (dict(nx=30, fw=20), "overlap"),
    (dict(fdorder=5), "FDORDER"),
    (dict(L=5), "L=5"),
    (dict(dtinv=0), "DTINV"),
    (dict(snap_precision=3), "snap_precision"),
])
def test_create_rejects_bad_parameters(kw, text):
    dn = product()
    p = dict(nx=80, ny=40, fdorder=4, maxrelerror=1, fw=8, free_surf=1, L=0, dt=DT, dh=DH)
    p.update(kw)
    nx, ny = p.pop("nx"), p.pop("ny")
    with pytest.raises(dn.DeniseError) as e:
        dn.Context(nx, ny, **p)
    say("  %s -> %s" % (kw, e.value))
    assert text in str(e.value)


def test_calls_reject_bad_input_and_the_context_stays_usable():
    dn, ctx = _ctx()
    vs, rho, _ = model(80, 40, DH, seed=2)
    ctx.set_model(vs, rho)
    bad = [
        (lambda: ctx.set_receivers(np.array([0, 5], np.int32), np.array([2, 2], np.int32)), "outside the grid"),
        (lambda: ctx.set_receivers(np.array([5, 5], np.int32), np.array([2, 2], np.int32)), "share a grid point"),
        (lambda: ctx.set_shots(np.array([81], np.int32), np.array([3], np.int32), ricker(100, DT, 60.0)[None]),
         "outside the grid"),
        (lambda: ctx.set_shots(np.array([20], np.int32), np.array([3], np.int32), ricker(100, DT, 60.0)[None],
                               np.array([101], np.int32)), "outside 1..100"),
        (lambda: ctx.set_wave_shots(3), "1 or 2"),
        (lambda: ctx.set_group_layout(3), "adjacent strips"),
    ]
    for call, text in bad:
        with pytest.raises(dn.DeniseError) as e:
            call()
        say("  rejected: %s" % e.value)
        assert text in str(e.value)
    # the context still works
    rec_i = np.arange(10, 70, 4, dtype=np.int32)
    ctx.set_receivers(rec_i, np.full_like(rec_i, 2))
    ctx.set_shots(np.array([20, 60], np.int32), np.full(2, 3, np.int32), np.stack([ricker(150, DT, 60.0)] * 2))
    with pytest.raises(dn.DeniseError) as e:
        ctx.gradient()
    assert "observed data not set" in str(e.value)
    with pytest.raises(dn.DeniseError):
        ctx.forward(shot0=1, nshots=2)
    seis = ctx.forward()
    assert np.all(np.isfinite(seis)) and np.max(np.abs(seis)) > 0
    ctx.close()


def test_unstable_time_step_is_reported():
    """DT ten times above the stability limit: the step kernel latches the
    first step after which the reference's probe cell (NY/2, NX/2) holds a
    NaN, and the call fails with the reference's message and that step,
    long before the end of the 3000-step record."""
    dn, ctx = _ctx(dt=2e-3)
    vs, rho, _ = model(80, 40, DH, seed=2)
    ctx.set_model(vs, rho)
    rec_i = np.arange(10, 70, 4, dtype=np.int32)
    ctx.set_receivers(rec_i, np.full_like(rec_i, 2))
    ctx.set_shots(np.array([40], np.int32), np.array([20], np.int32), ricker(3000, 2e-3, 5.0)[None])
    with pytest.raises(dn.DeniseError) as e:
        ctx.forward()
    say("  %s" % e.value)
    assert "unstable" in str(e.value)
    import re
    step = int(re.search(r"time step (\d+)", str(e.value)).group(1))
    assert 1 < step < 3000
    ctx.close()

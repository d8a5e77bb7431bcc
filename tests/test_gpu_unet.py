"""GPU checks of the learned preconditioner's network: the MFMA convolution
kernels, pooling/upsampling and the hand-written backward against the CPU
oracle (oracle/unet_oracle.py, itself pinned against torch autograd), and a
short training run.  BF16 activations: the bar is bfloat16 accuracy."""
import sys

import numpy as np
import pytest

from common import product, rel_l2
from unet_oracle import LAYERS, UNetOracle, offsets

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a)
    sys.stdout.flush()


def _inputs(seed, n=2, h=32, w=40):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 3, h, w)).astype(np.float32)
    t = (0.5 * x[:, 0] - 0.3 * x[:, 1] * x[:, 2]).astype(np.float32)
    return x, t


def test_unet_forward_matches_oracle():
    dn = product()
    net = dn.UNet(seed=3)
    p = net.get_params()
    wo, bo, n = offsets()
    assert p.size == n == net.nparams
    # nonzero biases exercise the epilogue
    rng = np.random.default_rng(5)
    for l, (ci, co, _) in enumerate(LAYERS):
        p[bo[l]:bo[l] + co] = (rng.standard_normal(co) * 0.05).astype(np.float32)
    net.set_params(p)
    x, _ = _inputs(1)
    say("unet forward 2x3x32x40")
    y = net.forward(x)
    ref = UNetOracle(p, bf16=True).forward(x)
    ref64 = UNetOracle(p, bf16=False).forward(x)
    e = rel_l2(y, ref)
    say("  rel L2 vs bf16 oracle %.3e, vs float64 oracle %.3e" % (e, rel_l2(y, ref64)))
    assert np.all(np.isfinite(y)) and np.max(np.abs(ref)) > 0
    assert e < 1e-2


def test_unet_gradient_matches_oracle():
    dn = product()
    net = dn.UNet(seed=4)
    p = net.get_params()
    x, t = _inputs(2)
    say("unet loss gradient 2x3x32x40")
    loss, g = net.loss_grad(x, t)
    lr, gr = UNetOracle(p, bf16=True).loss_grad(x, t)
    l64, g64 = UNetOracle(p, bf16=False).loss_grad(x, t)
    say("  loss %.6e oracle %.6e (float64 %.6e)" % (loss, lr, l64))
    assert abs(loss - lr) <= 1e-2 * abs(lr)
    wo, bo, n = offsets()
    worst = 0.0
    for l, (ci, co, _) in enumerate(LAYERS):
        sl = slice(wo[l], bo[l] + co)
        if np.max(np.abs(gr[sl])) == 0:
            continue
        e = rel_l2(g[sl], gr[sl])
        say("  layer %2d grad rel L2 vs bf16 oracle %.3e vs float64 %.3e" % (l, e, rel_l2(g[sl], g64[sl])))
        worst = max(worst, e)
    assert worst < 5e-2


def test_unet_training_reduces_loss():
    dn = product()
    net = dn.UNet(seed=7)
    x, t = _inputs(3, n=4, h=32, w=32)
    say("unet training 60 Adam steps")
    first = None
    for k in range(60):
        loss = net.train_step(x, t, 2e-3)
        first = loss if first is None else first
    final, _ = net.loss_grad(x, t)
    say("  loss %.4e -> %.4e" % (first, final))
    assert final < 0.5 * first


@pytest.mark.parametrize("shape", [(1, 16, 328), (2, 200, 72)], ids=["w328", "h200x2"])
def test_unet_tiles_and_chunks_match_oracle(shape):
    """Shapes that exercise the launch tiling of the kernels: rows wider than
    one 256-pixel convolution workgroup and not a multiple of 64 (w328), and
    more than 384 image rows, so that the weight gradient adds partials of
    several rows each (h200x2)."""
    dn = product()
    net = dn.UNet(seed=11)
    p = net.get_params()
    wo, bo, n = offsets()
    rng = np.random.default_rng(12)
    for l, (ci, co, _) in enumerate(LAYERS):
        p[bo[l]:bo[l] + co] = (rng.standard_normal(co) * 0.05).astype(np.float32)
    net.set_params(p)
    x, t = _inputs(13, *shape)
    say("unet %dx3x%dx%d forward and loss gradient" % shape)
    y = net.forward(x)
    ref = UNetOracle(p, bf16=True)
    e = rel_l2(y, ref.forward(x))
    loss, g = net.loss_grad(x, t)
    lr, gr = ref.loss_grad(x, t)
    say("  forward rel L2 %.3e; loss %.6e oracle %.6e" % (e, loss, lr))
    assert e < 1e-2
    assert abs(loss - lr) <= 1e-2 * abs(lr)
    worst = 0.0
    for l, (ci, co, _) in enumerate(LAYERS):
        sl = slice(wo[l], bo[l] + co)
        if np.max(np.abs(gr[sl])) == 0:
            continue
        worst = max(worst, rel_l2(g[sl], gr[sl]))
    say("  worst layer gradient rel L2 %.3e" % worst)
    assert worst < 5e-2

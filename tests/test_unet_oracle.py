"""Pin of the U-Net oracle (oracle/unet_oracle.py) against torch autograd on
the CPU in float64: the same network built from torch.nn.functional (conv2d,
avg_pool2d, nearest interpolation, concatenation) must give the same output,
loss and parameter gradient as the oracle's hand-written backward without
bfloat16 rounding.  CPU only."""
import numpy as np
import pytest

from unet_oracle import LAYERS, UNetOracle, offsets, to_bf16


def _torch():
    """torch, imported only when a test of this file runs: a GPU test
    process that merely collects this file keeps the HIP runtime the
    product library links against (torch bundles its own)."""
    torch = pytest.importorskip("torch")
    import torch.nn.functional as F
    return torch, F


def torch_loss_grad(params, x, t):
    torch, F = _torch()
    wo, bo, n = offsets()
    p = torch.tensor(np.asarray(params, np.float64), requires_grad=True)
    Ws, Bs = [], []
    for l, (ci, co, _) in enumerate(LAYERS):
        w = p[wo[l]:wo[l] + 9 * co * ci].reshape(3, 3, co, ci).permute(2, 3, 0, 1)
        Ws.append(w)
        Bs.append(p[bo[l]:bo[l] + co])
    xt = torch.tensor(np.asarray(x, np.float64))
    xt = torch.cat([xt, torch.zeros(xt.shape[0], 13, xt.shape[2], xt.shape[3], dtype=torch.float64)], dim=1)

    def conv(l, h):
        h = F.conv2d(h, Ws[l], Bs[l], padding=1)
        return F.relu(h) if LAYERS[l][2] else h

    up = lambda h: F.interpolate(h, scale_factor=2, mode="nearest")
    a0 = conv(0, xt)
    a1 = conv(1, a0)
    a3 = conv(3, conv(2, F.avg_pool2d(a1, 2)))
    a5 = conv(5, conv(4, F.avg_pool2d(a3, 2)))
    a7 = conv(7, conv(6, F.avg_pool2d(a5, 2)))
    a9 = conv(9, conv(8, torch.cat([up(a7), a5], 1)))
    a11 = conv(11, conv(10, torch.cat([up(a9), a3], 1)))
    a13 = conv(13, conv(12, torch.cat([up(a11), a1], 1)))
    y = conv(14, a13)[:, 0]
    loss = torch.mean((y - torch.tensor(np.asarray(t, np.float64))) ** 2)
    loss.backward()
    return y.detach().numpy(), float(loss), p.grad.numpy()


def random_params(seed):
    rng = np.random.default_rng(seed)
    wo, bo, n = offsets()
    p = np.zeros(n)
    for l, (ci, co, _) in enumerate(LAYERS):
        p[wo[l]:wo[l] + 9 * co * ci] = rng.standard_normal(9 * co * ci) * np.sqrt(2.0 / (9 * ci))
        p[bo[l]:bo[l] + co] = rng.standard_normal(co) * 0.05
    return p


def test_oracle_forward_and_backward_equal_torch_autograd():
    rng = np.random.default_rng(0)
    params = random_params(1)
    x = rng.standard_normal((2, 3, 16, 24))
    t = rng.standard_normal((2, 16, 24))
    o = UNetOracle(params, bf16=False)
    y = o.forward(x)
    loss, g = o.loss_grad(x, t)
    yt, lt, gt = torch_loss_grad(params, x, t)
    np.testing.assert_allclose(y, yt, rtol=1e-10, atol=1e-12)
    assert abs(loss - lt) <= 1e-10 * abs(lt)
    # the padded input channels and unused head outputs carry no gradient in
    # either; everything else must agree
    np.testing.assert_allclose(g, gt, rtol=1e-8, atol=1e-12 * np.max(np.abs(gt)))


def test_bf16_rounding_is_round_to_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5, 3.0e38], np.float32)
    r = to_bf16(x)
    assert list(r[:4]) == [1.0, 1.0, 1.0 + 2 ** -6, 1.0]
    assert r[4] == -2.5
    assert np.isfinite(r[5])

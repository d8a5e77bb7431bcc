"""CPU restatement of the learned preconditioner's U-Net (numpy).

TEST INFRASTRUCTURE ONLY: checks the product's HIP kernels
(denise-sh_amd/csrc/dn_unet.hip, dn_unet.cpp); only tests/ may use it.  The
network is not part of the reference.

Same network, parameter layout and data flow as the product:
  channels-last activations [n][H][W][C]; per layer W[3][3][co][ci], b[co];
  enc1 3(pad 16)->16->16, pool, enc2 ->32->32, pool, enc3 ->48->48, pool,
  mid ->64->64, dec3 [up(64)|enc3] 112->48->48, dec2 [up(48)|enc2] 80->32->32,
  dec1 [up(32)|enc1] 48->16->16, head 16->16 linear (channel 0 is the output).

With bf16=True the product's rounding points are reproduced (inputs,
weights, every stored activation and activation gradient rounded to
bfloat16, round-to-nearest-even); sums are formed in float64, the product
forms them in float32 on the matrix cores, so results agree to bfloat16
accuracy.  With bf16=False everything is float64: that form is pinned
against torch autograd in the tests.
"""
import numpy as np

LAYERS = [(16, 16, 1), (16, 16, 1), (16, 32, 1), (32, 32, 1), (32, 48, 1), (48, 48, 1), (48, 64, 1), (64, 64, 1),
          (112, 48, 1), (48, 48, 1), (80, 32, 1), (32, 32, 1), (48, 16, 1), (16, 16, 1), (16, 16, 0)]


def offsets():
    off, wo, bo = 0, [], []
    for ci, co, _ in LAYERS:
        wo.append(off)
        off += 9 * co * ci
        bo.append(off)
        off += co
    return wo, bo, off


def to_bf16(x):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


class UNetOracle:
    def __init__(self, params, bf16=True):
        self.bf16 = bf16
        wo, bo, n = offsets()
        p = np.asarray(params, np.float64)
        assert p.size == n
        self.W, self.b = [], []
        for l, (ci, co, _) in enumerate(LAYERS):
            w = p[wo[l]:wo[l] + 9 * co * ci].reshape(3, 3, co, ci)
            self.W.append(self._r(w))
            self.b.append(p[bo[l]:bo[l] + co].astype(np.float32 if bf16 else np.float64))

    # rounding to the stored precision
    def _r(self, x):
        return to_bf16(x).astype(np.float64) if self.bf16 else np.asarray(x, np.float64)

    @staticmethod
    def _conv_sum(x, w):
        """sum over taps and input channels, float64: x [n,H,W,Ci], w [3,3,Co,Ci]"""
        n, H, W_, _ = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        acc = np.zeros((n, H, W_, w.shape[2]))
        for ky in range(3):
            for kx in range(3):
                acc += np.einsum("nhwc,oc->nhwo", xp[:, ky:ky + H, kx:kx + W_, :], w[ky, kx], optimize=True)
        return acc

    def _conv(self, l, x):
        acc = self._conv_sum(x, self.W[l])
        if self.bf16:
            v = acc.astype(np.float32) + self.b[l]
        else:
            v = acc + self.b[l]
        if LAYERS[l][2]:
            v = np.maximum(v, 0)
        return v

    @staticmethod
    def _pool(x):
        return 0.25 * ((x[:, 0::2, 0::2] + x[:, 0::2, 1::2]) + (x[:, 1::2, 0::2] + x[:, 1::2, 1::2]))

    @staticmethod
    def _up(x):
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)

    def pack(self, x):
        """[n][3][H][W] -> [n][H][W][16] (channels 3..15 zero), rounded"""
        x = np.asarray(x, np.float32 if self.bf16 else np.float64)
        n, _, H, W = x.shape
        out = np.zeros((n, H, W, 16))
        out[..., :3] = np.moveaxis(self._r(x), 1, -1)
        return out

    def forward(self, x, keep=False):
        r = self._r
        a = {}
        a["x"] = self.pack(x)
        a[0] = r(self._conv(0, a["x"]))
        a[1] = r(self._conv(1, a[0]))
        a["p1"] = r(self._pool(a[1]))
        a[2] = r(self._conv(2, a["p1"]))
        a[3] = r(self._conv(3, a[2]))
        a["p2"] = r(self._pool(a[3]))
        a[4] = r(self._conv(4, a["p2"]))
        a[5] = r(self._conv(5, a[4]))
        a["p3"] = r(self._pool(a[5]))
        a[6] = r(self._conv(6, a["p3"]))
        a[7] = r(self._conv(7, a[6]))
        a["c8"] = np.concatenate([self._up(a[7]), a[5]], axis=-1)
        a[8] = r(self._conv(8, a["c8"]))
        a[9] = r(self._conv(9, a[8]))
        a["c10"] = np.concatenate([self._up(a[9]), a[3]], axis=-1)
        a[10] = r(self._conv(10, a["c10"]))
        a[11] = r(self._conv(11, a[10]))
        a["c12"] = np.concatenate([self._up(a[11]), a[1]], axis=-1)
        a[12] = r(self._conv(12, a["c12"]))
        a[13] = r(self._conv(13, a[12]))
        out = self._conv(14, a[13])
        a["out"] = out
        y = np.asarray(out[..., 0], np.float32 if self.bf16 else np.float64)
        return (y, a) if keep else y

    # ---- backward ----
    def _conv_bwd(self, l, dy, mask):
        w = self.W[l]
        wt = np.transpose(w[::-1, ::-1], (0, 1, 3, 2))       # [3,3,Ci,Co], flipped taps
        v = self._conv_sum(dy, wt)
        if self.bf16:
            v = v.astype(np.float32)
        if mask is not None:
            v = np.where(mask > 0, v, 0)
        return self._r(v)

    @staticmethod
    def _wgrad(x, dy):
        n, H, W_, _ = x.shape
        xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
        dw = np.zeros((3, 3, dy.shape[-1], x.shape[-1]))
        for ky in range(3):
            for kx in range(3):
                dw[ky, kx] = np.einsum("nhwo,nhwc->oc", dy, xp[:, ky:ky + H, kx:kx + W_, :], optimize=True)
        return dw, dy.sum(axis=(0, 1, 2))

    def _up_bwd(self, dcat, cu, mask):
        d = dcat[..., :cu]
        s = (d[:, 0::2, 0::2] + d[:, 0::2, 1::2]) + (d[:, 1::2, 0::2] + d[:, 1::2, 1::2])
        if self.bf16:
            s = s.astype(np.float32)
        return self._r(np.where(mask > 0, s, 0))

    def _pool_bwd(self, dpool, dskip, mask):
        v = 0.25 * self._up(dpool) + dskip
        if self.bf16:
            v = v.astype(np.float32)
        return self._r(np.where(mask > 0, v, 0))

    def loss_grad(self, x, t):
        """mean squared error and the gradient of every parameter (flat,
        product layout)"""
        y, a = self.forward(x, keep=True)
        t = np.asarray(t, np.float64)
        P = y.size
        diff = np.asarray(y, np.float64) - t
        if self.bf16:
            diff = (np.asarray(y, np.float32) - np.asarray(t, np.float32)).astype(np.float64)
        loss = float(np.mean(diff ** 2))
        d14 = np.zeros(a["out"].shape)
        g0 = 2.0 * diff / P
        if self.bf16:
            g0 = (np.float32(2.0) * diff.astype(np.float32) / np.float32(P)).astype(np.float64)
        d14[..., 0] = self._r(g0)
        grads = {}
        grads[14] = self._wgrad(a[13], d14)
        d13 = self._conv_bwd(14, d14, a[13])
        grads[13] = self._wgrad(a[12], d13)
        d12 = self._conv_bwd(13, d13, a[12])
        grads[12] = self._wgrad(a["c12"], d12)
        dc12 = self._conv_bwd(12, d12, None)
        d11 = self._up_bwd(dc12, 32, a[11])
        grads[11] = self._wgrad(a[10], d11)
        d10 = self._conv_bwd(11, d11, a[10])
        grads[10] = self._wgrad(a["c10"], d10)
        dc10 = self._conv_bwd(10, d10, None)
        d9 = self._up_bwd(dc10, 48, a[9])
        grads[9] = self._wgrad(a[8], d9)
        d8 = self._conv_bwd(9, d9, a[8])
        grads[8] = self._wgrad(a["c8"], d8)
        dc8 = self._conv_bwd(8, d8, None)
        d7 = self._up_bwd(dc8, 64, a[7])
        grads[7] = self._wgrad(a[6], d7)
        d6 = self._conv_bwd(7, d7, a[6])
        grads[6] = self._wgrad(a["p3"], d6)
        dp3 = self._conv_bwd(6, d6, None)
        d5 = self._pool_bwd(dp3, dc8[..., 64:], a[5])
        grads[5] = self._wgrad(a[4], d5)
        d4 = self._conv_bwd(5, d5, a[4])
        grads[4] = self._wgrad(a["p2"], d4)
        dp2 = self._conv_bwd(4, d4, None)
        d3 = self._pool_bwd(dp2, dc10[..., 48:], a[3])
        grads[3] = self._wgrad(a[2], d3)
        d2 = self._conv_bwd(3, d3, a[2])
        grads[2] = self._wgrad(a["p1"], d2)
        dp1 = self._conv_bwd(2, d2, None)
        d1 = self._pool_bwd(dp1, dc12[..., 32:], a[1])
        grads[1] = self._wgrad(a[0], d1)
        d0 = self._conv_bwd(1, d1, a[0])
        grads[0] = self._wgrad(a["x"], d0)
        wo, bo, n = offsets()
        flat = np.zeros(n)
        for l in range(len(LAYERS)):
            dw, db = grads[l]
            flat[wo[l]:wo[l] + dw.size] = dw.ravel()
            flat[bo[l]:bo[l] + db.size] = db
        return loss, flat

"""CPU oracle of the learned preconditioner's U-Net (test infrastructure only).

The same graph as fwi3d_amd/csrc/unet_net.hip -- 3x3x3 "same" convolutions
with ReLU (16/32/64 channels), 2x2x2 average pooling, nearest upsampling, skip
concatenation, 1x1x1 head, MSE loss -- restated in numpy float64 with a
hand-written backward pass.  With bf16=True the values are rounded to
bfloat16 (round to nearest even) where the GPU stores or feeds BF16:
inputs, weights, post-activation tensors, pooled tensors, and the gradient
operands of the weight- and data-gradient products.  With bf16=False
everything is float64, which is what the finite-difference checks use.

Parameter vector layout: for each conv layer W[Cout][27][Cin] then b[Cout];
then head w[16], b.  tap = (dz+1)*9 + (dy+1)*3 + (dx+1).
Tensors are channels-last [B][D][H][W][C].
"""
import numpy as np

CONVS = [(8, 0, 16), (16, 0, 16), (16, 0, 32), (32, 0, 32), (32, 0, 64), (64, 0, 64), (64, 32, 32), (32, 0, 32),
         (32, 16, 16), (16, 0, 16)]
HEAD_C = 16
IN_C = 3


def layout():
    """[(w_off, b_off, ca, cb, cout)], head offset, total size."""
    off, out = 0, []
    for ca, cb, co in CONVS:
        w = off
        off += co * 27 * (ca + cb)
        b = off
        off += co
        out.append((w, b, ca, cb, co))
    return out, off, off + HEAD_C + 1


def bf16(x):
    """Round float64/float32 values to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _taps():
    return [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def conv3(x, w):
    """x [B,D,H,W,Cin], w [Cout,27,Cin] -> [B,D,H,W,Cout] (no bias), zero padding."""
    B, D, H, W, C = x.shape
    xp = np.zeros((B, D + 2, H + 2, W + 2, C), x.dtype)
    xp[:, 1:-1, 1:-1, 1:-1] = x
    out = np.zeros((B, D, H, W, w.shape[0]))
    for t, (dz, dy, dx) in enumerate(_taps()):
        xs = xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]
        out += xs @ w[:, t, :].T
    return out


def conv3_wgrad(x, d):
    """dW[co,t,ci] = sum_v d[v,co] * x[v+tap,ci]."""
    B, D, H, W, C = x.shape
    xp = np.zeros((B, D + 2, H + 2, W + 2, C), x.dtype)
    xp[:, 1:-1, 1:-1, 1:-1] = x
    dw = np.zeros((d.shape[-1], 27, C))
    d2 = d.reshape(-1, d.shape[-1])
    for t, (dz, dy, dx) in enumerate(_taps()):
        xs = xp[:, 1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W].reshape(-1, C)
        dw[:, t, :] = d2.T @ xs
    return dw


def conv3_dgrad(d, w):
    """dx[v,ci] = sum_{t,co} d[v - tap, co] * w[co,t,ci]."""
    B, D, H, W, Co = d.shape
    dp = np.zeros((B, D + 2, H + 2, W + 2, Co))
    dp[:, 1:-1, 1:-1, 1:-1] = d
    dx_ = np.zeros((B, D, H, W, w.shape[2]))
    for t, (dz, dy, dx) in enumerate(_taps()):
        ds = dp[:, 1 - dz:1 - dz + D, 1 - dy:1 - dy + H, 1 - dx:1 - dx + W]
        dx_ += ds @ w[:, t, :]
    return dx_


def pool(x):
    B, D, H, W, C = x.shape
    return x.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(2, 4, 6))


def pool_bwd(d):
    return np.repeat(np.repeat(np.repeat(d, 2, 1), 2, 2), 2, 3) / 8.0


def up(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 1), 2, 2), 2, 3)


def up_bwd(d):
    B, D, H, W, C = d.shape
    return d.reshape(B, D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4, 6))


class UNetRef:
    def __init__(self, params, use_bf16=True):
        self.p = np.asarray(params, np.float64)
        self.r = bf16 if use_bf16 else (lambda v: np.asarray(v, np.float64))
        self.lay, self.hoff, self.n = layout()

    def W(self, l):
        w, b, ca, cb, co = self.lay[l]
        return self.p[w:w + co * 27 * (ca + cb)].reshape(co, 27, ca + cb), self.p[b:b + co]

    def forward(self, x3):
        """x3: planar [3][B][D][H][W] -> y [B][D][H][W]; keeps activations."""
        r = self.r
        x3 = np.asarray(x3, np.float64)
        B, D, H, W = x3.shape[1:]
        x = np.zeros((B, D, H, W, 8))
        x[..., :3] = np.moveaxis(x3, 0, -1)
        x = r(x)
        a = {"x": x}

        def conv(l, inp):
            w, b = self.W(l)
            return r(np.maximum(conv3(inp, r(w)) + b, 0.0))

        a["e1a"] = conv(0, x)
        a["e1b"] = conv(1, a["e1a"])
        a["p1"] = r(pool(a["e1b"]))
        a["e2a"] = conv(2, a["p1"])
        a["e2b"] = conv(3, a["e2a"])
        a["p2"] = r(pool(a["e2b"]))
        a["e3a"] = conv(4, a["p2"])
        a["e3b"] = conv(5, a["e3a"])
        a["u2"] = up(a["e3b"])
        a["c2"] = np.concatenate([a["u2"], a["e2b"]], -1)
        a["d2a"] = conv(6, a["c2"])
        a["d2b"] = conv(7, a["d2a"])
        a["u1"] = up(a["d2b"])
        a["c1"] = np.concatenate([a["u1"], a["e1b"]], -1)
        a["d1a"] = conv(8, a["c1"])
        a["d1b"] = conv(9, a["d1a"])
        hw = self.p[self.hoff:self.hoff + HEAD_C]
        hb = self.p[self.hoff + HEAD_C]
        y = a["d1b"] @ hw + hb
        self.a = a
        return y

    def loss_and_grad(self, x3, target):
        r = self.r
        y = self.forward(x3)
        a = self.a
        t = np.asarray(target, np.float64)
        n = y.size
        loss = float(np.mean((y - t) ** 2))
        g = np.zeros(self.n)
        dy = 2.0 * (y - t) / n
        hw = self.p[self.hoff:self.hoff + HEAD_C]
        g[self.hoff:self.hoff + HEAD_C] = np.tensordot(dy, a["d1b"], axes=(list(range(dy.ndim)),
                                                                           list(range(dy.ndim))))
        g[self.hoff + HEAD_C] = dy.sum()
        d = dy[..., None] * hw

        def layer(l, dact, act, inp, need_dx=True):
            w, _ = self.W(l)
            wo, bo, ca, cb, co = self.lay[l]
            dpre = dact * (act > 0)
            g[bo:bo + co] = dpre.reshape(-1, co).sum(0)
            g[wo:wo + co * 27 * (ca + cb)] = conv3_wgrad(inp, r(dpre)).ravel()
            return conv3_dgrad(r(dpre), r(w)) if need_dx else None

        d = layer(9, d, a["d1b"], a["d1a"])
        d = layer(8, d, a["d1a"], a["c1"])
        g_u1, g_e1b = d[..., :32], d[..., 32:]
        d = up_bwd(g_u1)
        d = layer(7, d, a["d2b"], a["d2a"])
        d = layer(6, d, a["d2a"], a["c2"])
        g_u2, g_e2b = d[..., :64], d[..., 64:]
        d = up_bwd(g_u2)
        d = layer(5, d, a["e3b"], a["e3a"])
        d = layer(4, d, a["e3a"], a["p2"])
        g_e2b = g_e2b + pool_bwd(d)
        d = layer(3, g_e2b, a["e2b"], a["e2a"])
        d = layer(2, d, a["e2a"], a["p1"])
        g_e1b = g_e1b + pool_bwd(d)
        d = layer(1, g_e1b, a["e1b"], a["e1a"])
        layer(0, d, a["e1a"], a["x"], need_dx=False)
        return loss, g

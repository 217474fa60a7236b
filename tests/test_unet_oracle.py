"""Pins the U-Net oracle (oracle/unet_ref.py) on the CPU.

  conv3 / wgrad / dgrad   brute-force direct convolution on tiny shapes, and
                          the adjoint identity <conv(x), d> = <x, dgrad(d)>
  full backward           central finite differences of the FP64 network
  bf16 rounding           torch's float32 -> bfloat16 conversion
"""
import numpy as np
import pytest

from oracle import unet_ref as U


def _direct_conv(x, w):
    B, D, H, W, C = x.shape
    out = np.zeros((B, D, H, W, w.shape[0]))
    for b in range(B):
        for z in range(D):
            for y in range(H):
                for xx in range(W):
                    for t in range(27):
                        dz, dy, dx = t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1
                        zz, yy, xs = z + dz, y + dy, xx + dx
                        if 0 <= zz < D and 0 <= yy < H and 0 <= xs < W:
                            out[b, z, y, xx] += w[:, t, :] @ x[b, zz, yy, xs]
    return out


def test_conv3_matches_direct():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(2, 3, 4, 5, 3))
    w = rng.normal(size=(4, 27, 3))
    assert np.allclose(U.conv3(x, w), _direct_conv(x, w), atol=1e-12)


def test_wgrad_dgrad_adjoint():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(1, 4, 4, 4, 5))
    w = rng.normal(size=(6, 27, 5))
    d = rng.normal(size=(1, 4, 4, 4, 6))
    # <conv(x, w), d> = <x, dgrad(d, w)> = <w, wgrad(x, d)>
    lhs = np.sum(U.conv3(x, w) * d)
    assert abs(lhs - np.sum(x * U.conv3_dgrad(d, w))) < 1e-9 * abs(lhs)
    assert abs(lhs - np.sum(w * U.conv3_wgrad(x, d))) < 1e-9 * abs(lhs)


def test_pool_up_adjoint():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(1, 4, 4, 4, 3))
    c = rng.normal(size=(1, 2, 2, 2, 3))
    assert abs(np.sum(U.pool(a) * c) - np.sum(a * U.pool_bwd(c))) < 1e-12
    assert abs(np.sum(U.up(c) * a) - np.sum(c * U.up_bwd(a))) < 1e-12


def _params(seed=3, scale=0.3):
    lay, hoff, n = U.layout()
    rng = np.random.default_rng(seed)
    p = rng.normal(0, scale, n)
    for w, b, ca, cb, co in lay:
        cin = ca + cb
        p[w:w + co * 27 * cin] *= np.sqrt(2.0 / (27 * cin)) / scale
    return p


def test_backward_finite_differences():
    rng = np.random.default_rng(4)
    p = _params()
    x3 = rng.normal(size=(3, 1, 4, 4, 8))
    t = rng.normal(size=(1, 4, 4, 8))
    net = U.UNetRef(p, use_bf16=False)
    loss, g = net.loss_and_grad(x3, t)
    lay, hoff, n = U.layout()
    idx = []
    for w, b, ca, cb, co in lay:
        idx += list(rng.integers(w, w + co * 27 * (ca + cb), 3)) + [b]
    idx += [hoff, hoff + 16]
    h = 1e-6
    worst = 0.0
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (U.UNetRef(pp, False).loss_and_grad(x3, t)[0] - U.UNetRef(pm, False).loss_and_grad(x3, t)[0]) / (2 * h)
        worst = max(worst, abs(fd - g[i]) / (abs(g).max() + 1e-30))
    assert worst < 1e-6, worst


def test_bf16_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.normal(size=10000) * 10 ** rng.uniform(-20, 20, 10000), [0.0, -0.0, 1.0, 65504.0]])
    want = torch.tensor(v.astype(np.float32)).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(U.bf16(v), want)

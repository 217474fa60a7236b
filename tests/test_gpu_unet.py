"""The U-Net's BF16 MFMA kernels against the CPU oracle (run with -m gpu).

Each convolution layer alone, forward and backward (identical BF16 inputs,
weights and output gradients, FP32 vs FP64 accumulation), must agree to
relative L2 1e-5, and the Adam step to one unit in the last place.  The
whole network's output, loss and parameter gradients are compared with the
oracle run with the same BF16 rounding points; values that sit at BF16
rounding boundaries can round differently after FP32 vs FP64 accumulation
and the difference propagates through the layers, so those end-to-end
comparisons use relative L2 bounds of 2e-2 (output) and 6e-2 (gradients
per layer); the per-layer tests above pin each kernel tightly."""
import numpy as np
import pytest

from oracle import unet_ref as U
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net():
    from fwi3d_amd.unet import UNet
    n = UNet(seed=11)
    yield n
    n.close()


def test_num_params(net):
    assert net.nparams == U.layout()[2]


@pytest.mark.parametrize("layer", list(range(10)))
def test_conv_layer_matches_oracle(net, layer):
    rng = np.random.default_rng(layer)
    cin, cout = net.CONV_CHANNELS[layer]
    x = rng.normal(size=(2, 4, 8, 20, cin)).astype(np.float32)
    got = net.conv_probe(layer, x)
    lay, _, _ = U.layout()
    wo, bo, ca, cb, co = lay[layer]
    w = net.params[wo:wo + co * 27 * (ca + cb)].reshape(co, 27, ca + cb)
    want = U.conv3(U.bf16(x), U.bf16(w))
    assert rel_l2(got, want) <= 1e-5, rel_l2(got, want)


@pytest.mark.parametrize("layer", list(range(10)))
def test_conv_layer_backward_matches_oracle(net, layer):
    """The data- and weight-gradient kernels of each layer against the
    oracle's direct sums on the same BF16 operands (the kernels round the
    output gradient to BF16, as the oracle's bf16 path does): FP32 against
    FP64 accumulation only, so relative L2 <= 1e-5.  Layers 6 and 8 take
    their input as the network's two skip-concatenated tensors."""
    rng = np.random.default_rng(100 + layer)
    cin, cout = net.CONV_CHANNELS[layer]
    x = rng.normal(size=(2, 4, 8, 20, cin)).astype(np.float32)
    d = rng.normal(size=(2, 4, 8, 20, cout)).astype(np.float32)
    dx, dw = net.conv_grad_probe(layer, x, d)
    lay, _, _ = U.layout()
    wo, bo, ca, cb, co = lay[layer]
    w = net.params[wo:wo + co * 27 * (ca + cb)].reshape(co, 27, ca + cb)
    want_dx = U.conv3_dgrad(U.bf16(d), U.bf16(w))
    want_dw = U.conv3_wgrad(U.bf16(x), U.bf16(d))
    assert np.abs(want_dw).max() > 0 and np.abs(want_dx).max() > 0
    assert rel_l2(dx, want_dx) <= 1e-5, rel_l2(dx, want_dx)
    assert rel_l2(dw, want_dw) <= 1e-5, rel_l2(dw, want_dw)


def _ulps(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def test_adam_step_matches_numpy():
    """Two Adam steps (beta 0.9 / 0.999, eps 1e-8) against the same float32
    arithmetic in numpy on the GPU's own gradients: every parameter within
    one unit in the last place."""
    from fwi3d_amd.unet import UNet
    f32 = np.float32
    b1, b2, eps, lr = f32(0.9), f32(0.999), f32(1e-8), f32(2e-3)
    with UNet(seed=21) as n:
        p = n.params.copy()
        m1 = np.zeros_like(p)
        m2 = np.zeros_like(p)
        for step in (1, 2):
            x3, t = _batch(40 + step)
            _, g = n.gradients(x3, t)
            n.train_step(x3, t, float(lr))
            c1 = f32(1.0) / (f32(1.0) - np.power(b1, f32(step)))
            c2 = f32(1.0) / (f32(1.0) - np.power(b2, f32(step)))
            m1 = (b1 * m1 + (f32(1.0) - b1) * g).astype(f32)
            m2 = (b2 * m2 + (f32(1.0) - b2) * g * g).astype(f32)
            p = (p - (lr * (m1 * c1)) / (np.sqrt(m2 * c2) + eps)).astype(f32)
            got = n.params
            assert _ulps(got, p).max() <= 1, (step, _ulps(got, p).max())


def _batch(seed, B=2, D=8, H=8, W=16):
    rng = np.random.default_rng(seed)
    x3 = rng.normal(size=(3, B, D, H, W)).astype(np.float32)
    t = (0.5 * x3[0] - 0.2 * x3[2]).astype(np.float32)
    return x3, t


def test_forward_matches_oracle(net):
    x3, _ = _batch(1)
    y = net.forward(x3)
    ref = U.UNetRef(net.params, use_bf16=True).forward(x3)
    assert rel_l2(y, ref) <= 2e-2, rel_l2(y, ref)


def test_gradients_match_oracle(net):
    x3, t = _batch(2)
    loss, g = net.gradients(x3, t)
    rl, rg = U.UNetRef(net.params, use_bf16=True).loss_and_grad(x3, t)
    assert abs(loss - rl) <= 2e-2 * rl
    lay, hoff, n = U.layout()
    for l, (wo, bo, ca, cb, co) in enumerate(lay):
        sw = slice(wo, wo + co * 27 * (ca + cb))
        sb = slice(bo, bo + co)
        assert rel_l2(g[sw], rg[sw]) <= 6e-2, (l, rel_l2(g[sw], rg[sw]))
        assert rel_l2(g[sb], rg[sb]) <= 6e-2, (l, rel_l2(g[sb], rg[sb]))
    assert rel_l2(g[hoff:], rg[hoff:]) <= 2e-2


def test_gradients_and_training_are_deterministic():
    """The weight, bias and head gradients are summed in a fixed order (no
    float atomics), so the same inputs give the same bits, and two trainings
    from the same seed end with identical parameters."""
    from fwi3d_amd.unet import UNet
    x3, t = _batch(3, B=4, D=16, H=32, W=32)  # the training patch shape: 64 row chunks
    params = []
    for _ in range(2):
        with UNet(seed=9) as n:
            l1, g1 = n.gradients(x3, t)
            l2, g2 = n.gradients(x3, t)
            assert l1 == l2 and np.array_equal(g1, g2)
            for it in range(10):
                xb, tb = _batch(200 + it % 4)
                n.train_step(xb, tb, 2e-3)
            params.append(n.params.copy())
    assert np.array_equal(params[0], params[1])


def test_training_lowers_loss():
    from fwi3d_amd.unet import UNet
    with UNet(seed=5) as n:
        losses = []
        for it in range(60):
            x3, t = _batch(100 + it % 8)
            losses.append(n.train_step(x3, t, 2e-3))
        assert np.mean(losses[-8:]) < 0.5 * np.mean(losses[:8]), (losses[:8], losses[-8:])

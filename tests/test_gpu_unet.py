"""The U-Net's BF16 MFMA kernels against the CPU oracle (run with -m gpu).

Each convolution layer alone (identical BF16 inputs and weights, FP32 vs
FP64 accumulation) must agree to relative L2 1e-5.  The whole network's
output, loss and parameter gradients are compared with the oracle run with
the same BF16 rounding points; values that sit at BF16 rounding boundaries
can round differently after FP32 vs FP64 accumulation, so those comparisons
use relative L2 bounds of 2e-2 (output) and 6e-2 (gradients per layer)."""
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

"""ctypes binding of the learned preconditioner (fwi_unet_* in include/fwi3d.h).

Inputs are planar FP32 arrays [3][B][D][H][W] (gradient, illumination,
model) with D, H, W divisible by 4 (x fastest); outputs and targets are
[B][D][H][W].  There is no CPU fallback.
"""
import ctypes

import numpy as np

from . import lib as _lib

_bound = False

# the fwi_unet_* symbols of include/fwi3d.h
EXPORTS = ["fwi_unet_create", "fwi_unet_destroy", "fwi_unet_last_error", "fwi_unet_num_params",
           "fwi_unet_get_params", "fwi_unet_set_params", "fwi_unet_forward", "fwi_unet_gradients",
           "fwi_unet_train_step", "fwi_unet_conv_probe", "fwi_unet_conv_grad_probe"]


def _bind():
    global _bound
    L = _lib.load()
    if _bound:
        return L
    P, I, F = ctypes.c_void_p, ctypes.c_int, ctypes.c_float
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.fwi_unet_create.argtypes = [I, ctypes.c_uint64, ctypes.POINTER(P)]
    L.fwi_unet_destroy.argtypes = [P]
    L.fwi_unet_last_error.argtypes = [P]
    L.fwi_unet_last_error.restype = ctypes.c_char_p
    L.fwi_unet_num_params.argtypes = [P]
    L.fwi_unet_num_params.restype = ctypes.c_long
    L.fwi_unet_get_params.argtypes = [P, f32]
    L.fwi_unet_set_params.argtypes = [P, f32]
    L.fwi_unet_forward.argtypes = [P, f32, I, I, I, I, f32]
    L.fwi_unet_gradients.argtypes = [P, f32, f32, I, I, I, I, ctypes.POINTER(ctypes.c_double), f32]
    L.fwi_unet_train_step.argtypes = [P, f32, f32, I, I, I, I, F, ctypes.POINTER(ctypes.c_double)]
    L.fwi_unet_conv_probe.argtypes = [P, I, f32, I, I, I, I, f32]
    L.fwi_unet_conv_grad_probe.argtypes = [P, I, f32, f32, I, I, I, I, f32, f32]
    for name in ("fwi_unet_create", "fwi_unet_destroy", "fwi_unet_get_params", "fwi_unet_set_params",
                 "fwi_unet_forward", "fwi_unet_gradients", "fwi_unet_train_step", "fwi_unet_conv_probe",
                 "fwi_unet_conv_grad_probe"):
        getattr(L, name).restype = I
    _bound = True
    return L


class UNet:
    """3-level 3-D U-Net (16/32/64 channels) on the GPU."""

    CONV_CHANNELS = [(8, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (32, 32), (48, 16),
                     (16, 16)]

    def __init__(self, seed=0, device=0):
        self.L = _bind()
        h = ctypes.c_void_p()
        rc = self.L.fwi_unet_create(device, seed, ctypes.byref(h))
        if rc != 0:
            raise _lib.FwiError(rc, "fwi_unet_create failed")
        self.h = h
        self.nparams = self.L.fwi_unet_num_params(h)

    def _chk(self, rc):
        if rc != 0:
            raise _lib.FwiError(rc, self.L.fwi_unet_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.fwi_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def params(self):
        out = np.zeros(self.nparams, np.float32)
        self._chk(self.L.fwi_unet_get_params(self.h, out))
        return out

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.nparams
        self._chk(self.L.fwi_unet_set_params(self.h, p))

    @staticmethod
    def _x(x3):
        x3 = np.ascontiguousarray(x3, np.float32)
        if x3.ndim != 5 or x3.shape[0] != 3:
            raise ValueError("x3 must be [3][B][D][H][W]")
        return x3

    def forward(self, x3):
        x3 = self._x(x3)
        B, D, H, W = x3.shape[1:]
        y = np.zeros((B, D, H, W), np.float32)
        self._chk(self.L.fwi_unet_forward(self.h, x3, B, D, H, W, y))
        return y

    def gradients(self, x3, target):
        x3 = self._x(x3)
        B, D, H, W = x3.shape[1:]
        t = np.ascontiguousarray(target, np.float32).reshape(B, D, H, W)
        g = np.zeros(self.nparams, np.float32)
        loss = ctypes.c_double()
        self._chk(self.L.fwi_unet_gradients(self.h, x3, t, B, D, H, W, ctypes.byref(loss), g))
        return loss.value, g

    def train_step(self, x3, target, lr):
        x3 = self._x(x3)
        B, D, H, W = x3.shape[1:]
        t = np.ascontiguousarray(target, np.float32).reshape(B, D, H, W)
        loss = ctypes.c_double()
        self._chk(self.L.fwi_unet_train_step(self.h, x3, t, B, D, H, W, lr, ctypes.byref(loss)))
        return loss.value

    def conv_grad_probe(self, layer, x, dout):
        """(dx, dw) of one convolution layer's backward kernels: x
        [B][D][H][W][Cin] (rounded to BF16), dout [B][D][H][W][Cout] FP32;
        dw is [Cout][27][Cin]."""
        x = np.ascontiguousarray(x, np.float32)
        dout = np.ascontiguousarray(dout, np.float32)
        B, D, H, W, C = x.shape
        cout = self.CONV_CHANNELS[layer][1]
        assert dout.shape == (B, D, H, W, cout)
        dx = np.zeros(x.shape, np.float32)
        dw = np.zeros((cout, 27, C), np.float32)
        self._chk(self.L.fwi_unet_conv_grad_probe(self.h, layer, x, dout, B, D, H, W, dx, dw))
        return dx, dw

    def conv_probe(self, layer, x):
        x = np.ascontiguousarray(x, np.float32)
        B, D, H, W, C = x.shape
        cout = self.CONV_CHANNELS[layer][1]
        out = np.zeros((B, D, H, W, cout), np.float32)
        self._chk(self.L.fwi_unet_conv_probe(self.h, layer, x, B, D, H, W, out))
        return out


def main(argv=None):
    """Training steps at a fixed shape on synthetic inputs, for kernel traces:
    python -m fwi3d_amd.unet --steps 6 --batch 4 --size 32"""
    import argparse
    ap = argparse.ArgumentParser(description=main.__doc__.split("\n")[0])
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--size", type=int, default=32, help="patch edge (a multiple of 4)")
    a = ap.parse_args(argv)
    rng = np.random.default_rng(0)
    n = a.size
    x3 = rng.normal(size=(3, a.batch, n, n, n)).astype(np.float32)
    t = (0.5 * x3[0] - 0.2 * x3[2]).astype(np.float32)
    with UNet(seed=1) as net:
        for _ in range(a.steps):
            loss = net.train_step(x3, t, 1e-3)
    print("loss after %d steps: %.6f" % (a.steps, loss))
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())

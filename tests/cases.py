"""Small seeded cases shared by the parity tests (host-side set-up only)."""
import numpy as np

import oracle as O


class Case:
    """A padded grid with a 2-layer model, random boundary, one source and a
    receiver patch -- a shrunken version of the 101^3 two-layer workload."""

    def __init__(self, nx=40, ny=36, nz=32, nb=8, dh=10.0, dt=1e-3, nt=120, fmax=45.0, seed=7, shot=0,
                 amp=0.6, rec_stride=3, abc_factor=None, rng_seed=1):
        self.nx, self.ny, self.nz, self.nb, self.dh, self.dt, self.nt = nx, ny, nz, nb, dh, dt, nt
        self.nxx, self.nyy, self.nzz = nx + 2 * nb, ny + 2 * nb, nz + 2 * nb
        self.wavelet, self.tlag = O.wavelet(nt, dt, fmax)
        z = np.arange(nz) * dh
        interior = np.where(z < nz * dh / 2, 1500.0, 2500.0).astype(np.float32)
        self.interior = np.broadcast_to(interior, (ny, nx, nz)).copy()
        rng = np.random.default_rng(rng_seed)
        self.interior += rng.uniform(-50, 50, self.interior.shape).astype(np.float32)
        self.vclamp = np.float32(0.45 * dh / dt)
        padded = O.expand_boundary(self.interior, nb)
        self.vp = O.random_boundary(padded, nb, seed, shot, amp, self.vclamp) if amp > 0 else padded
        self.src = (nb + nx // 2, nb + ny // 2, nb + 2)
        xs = np.arange(nb + 2, nb + nx - 2, rec_stride)
        ys = np.arange(nb + 2, nb + ny - 2, rec_stride)
        X, Y = np.meshgrid(xs, ys, indexing="ij")
        self.rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, nb + 2)], axis=1).astype(np.int32)
        self.abc = O.dampers(nb, abc_factor) if abc_factor is not None else None

    @property
    def steps(self):
        return self.nt + self.tlag

    def context(self, ctx_cls, abc=False):
        c = ctx_cls(self.nxx, self.nyy, self.nzz, self.nb, self.dh, self.dt, self.nt, self.tlag)
        c.set_model(self.vp)
        c.set_wavelet(self.wavelet)
        c.set_source(*self.src)
        c.set_receivers(self.rec)
        if abc:
            c.set_dampers(self.abc[0])
        return c


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))

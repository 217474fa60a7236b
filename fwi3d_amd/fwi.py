"""Batched full-waveform inversion over the C ABI, in Python.

The same algorithm and arithmetic as the C++ Inversion class
(fwi3d_amd/host/inversion.cpp, after src/inversion/inversion.cu): misfit
from an ABC-padded forward pass, gradient from a random-boundary forward pass
and a time-reversed backward pass, illumination-compensated Laplacian
post-processing, parabolic line search over alpha in {0, 0.02, 0.05} * dm,
m = 1/v^2 update.  All shots of the problem run as one batch.  This module
drives the configuration-3/4 experiments and generates the learned
preconditioner's training triples.
"""
import os
import tempfile
import time

import numpy as np

from . import lib


def b1d_table(nb, factor):
    """1-D Cerjan damper, float32 as modeling.cu:96-99."""
    i = np.arange(nb, dtype=np.float32)
    a = (np.float32(factor) * (np.float32(nb) - i)).astype(np.float32)
    return np.exp(-(a.astype(np.float64) ** 2)).astype(np.float32)


def crop(vol, nb):
    return np.ascontiguousarray(vol[nb:-nb, nb:-nb, nb:-nb]) if nb > 0 else vol


class FwiProblem:
    """Observed data, geometry and the two batch contexts of one inversion.

    tune: autotune the launch shapes of the forward passes (absorbing and
    random boundary) when the contexts are built; the choices are cached in
    `tune_cache` (default: one file per process in the temporary
    directory), keyed by grid and shot count, so that problems of the same
    size reuse them (forward results do not depend on the launch shape).
    The backward pass sums the shots' gradient and illumination terms in an
    order that depends on how shots share a block, so its shape is fixed
    rather than timed: one shot per thread and block (each shot its own
    partial sums, folded after the pass, as with the default shape) on
    64 x 16 tiles, so that a seed fixes every result bit for bit."""

    def __init__(self, w, obs, nb_abc=None, nb_rbc=None, abc_factor=0.004, zmask=40.0, seed=None, device=0,
                 tune=True, tune_cache=None):
        self.w = w
        self.obs = [np.ascontiguousarray(o, np.float32) for o in obs]
        self.S = len(w.shots)
        self.nb_abc = nb_abc or w.nb
        self.nb_rbc = nb_rbc or w.nb
        self.zmask = zmask
        self.seed = w.rbc_seed if seed is None else seed
        self.abc = self._context(self.nb_abc, device)
        self.abc.set_dampers(b1d_table(self.nb_abc, abc_factor))
        self.rbc = self._context(self.nb_rbc, device)
        if tune:
            cache = tune_cache or os.path.join(tempfile.gettempdir(), "fwi3d_autotune_%d.txt" % os.getpid())
            self.abc.autotune(lib.KIND_FWD_ABC, 3, cache)
            self.rbc.autotune(lib.KIND_FWD_RBC, 3, cache)
            self.rbc.set_tuning(lib.KIND_BWD_FWI, 16, 1, 1)
        self.forward_passes = 0
        self.seconds = {"misfit": 0.0, "gradient": 0.0}

    def _context(self, nb, device):
        w = self.w
        nx, ny, nz = w.nx + 2 * nb, w.ny + 2 * nb, w.nz + 2 * nb
        c = lib.Context(nx, ny, nz, nb, w.dh, w.dt, w.nt, w.tlag, device=device, nshots=self.S)
        c.set_wavelet(w.wavelet)
        c.set_model_interior(w.interior)
        for s in range(self.S):
            shift = nb - w.nb
            c.set_source(*[v + shift for v in w.src_index(s)], slot=s)
            c.set_receivers(w.rec_index(s) + shift, slot=s)
        return c

    def close(self):
        self.abc.close()
        self.rbc.close()

    def misfit(self, v):
        """Sum over shots of sum r^2 (FP64 on the GPU) for velocity v; leaves
        the residuals on the ABC context."""
        t0 = time.perf_counter()
        self.abc.set_model_interior(v)
        self.abc.forward(lib.BOUNDARY_ABC)
        self.forward_passes += self.S
        s = sum(self.abc.residual(self.obs[k], slot=k) for k in range(self.S))
        self.seconds["misfit"] += time.perf_counter() - t0
        return s

    def gradient(self, v):
        """(sum r^2, normalised gradient, max|g|, illumination) at velocity v."""
        t0 = time.perf_counter()
        f = self.misfit(v)
        c = self.rbc
        c.set_model_interior(v)
        for k in range(self.S):
            c.shot_random_boundary(k, self.seed, k, self.w.rbc_amp, self.w.vclamp)
            c.set_backward_data(self.abc.backward_data(slot=k), slot=k)
        c.zero_volume(lib.VOL_GRAD)
        c.zero_volume(lib.VOL_ILLUM)
        c.forward(lib.BOUNDARY_RBC)
        c.grad_backward()
        g, gmax = c.export_gradient(self.zmask)
        illum = crop(c.get_volume(lib.VOL_ILLUM), self.nb_rbc)
        self.forward_passes += 2 * self.S
        self.seconds["gradient"] += time.perf_counter() - t0
        return f, g, gmax, illum


def parabola_step(a0, a1, a2, f0, f1, f2):
    """Vertex of the parabola through three points, float32 as inversion.cu:304-312.
    Collinear points (den = 0) give a non-finite step, as in the reference;
    the callers take the better trial step then."""
    a0, a1, a2, f0, f1, f2 = (np.float32(v) for v in (a0, a1, a2, f0, f1, f2))
    num = (a1 * a1 - a2 * a2) * f0 + (a2 * a2 - a0 * a0) * f1 + (a0 * a0 - a1 * a1) * f2
    den = (a1 - a2) * f0 + (a2 - a0) * f1 + (a0 - a1) * f2
    with np.errstate(divide="ignore", invalid="ignore"):
        return float(np.float32(0.5) * (num / den))


def velocity_bounds(w, vlo=500.0):
    """Velocities an update may produce: from vlo up to the 3-D scheme's
    stability limit 0.45 dh/dt (the contexts reject models beyond it)."""
    return vlo, 0.45 * w.dh / w.dt


def gradient_step(prob, v, g, f0, vmin, vmax):
    """Parabolic line search along the normalised gradient in m = 1/v^2
    (inversion.cu:291-362); returns (new velocity, step).  Two guards the
    reference lacks keep every trial model valid: slowness^2 is bounded to
    the velocity range of velocity_bounds, and a parabola without a minimum
    (non-finite or non-positive vertex) takes the better trial step."""
    m = (1.0 / (v.astype(np.float32) * v.astype(np.float32))).astype(np.float32)
    dm = np.float32(1.0) / (np.float32(vmin) ** 2) - np.float32(1.0) / (np.float32(vmax) ** 2)
    a1, a2 = np.float32(0.02) * dm, np.float32(0.05) * dm
    vlo, vhi = velocity_bounds(prob.w)
    mlo, mhi = np.float32(1.0 / vhi ** 2), np.float32(1.0 / vlo ** 2)

    def trial(a):
        return (1.0 / np.sqrt(np.clip(m - np.float32(a) * g, mlo, mhi))).astype(np.float32)

    f1 = np.sqrt(prob.misfit(trial(a1)))
    f2 = np.sqrt(prob.misfit(trial(a2)))
    step = parabola_step(0.0, a1, a2, f0, f1, f2)
    if not np.isfinite(step) or step <= 0.0:
        step = float(a1 if f1 < f2 else a2)
    return trial(step), step


def guarded_fwi(prob, v0, iterations, log=None, max_halvings=4, times=None):
    """Plain FWI with the acceptance rule of the learned loop: a gradient
    step (parabolic line search, as plain_fwi) is kept only when the misfit
    of the new model is below the current one; otherwise the step is halved
    (up to max_halvings times), and the run stops when no halving helps.
    Returns (v, misfits, seconds, steps_taken); times (a list) receives the
    elapsed seconds at each misfit."""
    vmin, vmax = float(v0.min()), float(v0.max())
    vlo, vhi = velocity_bounds(prob.w)
    mlo, mhi = np.float32(1.0 / vhi ** 2), np.float32(1.0 / vlo ** 2)
    v = v0.astype(np.float32)
    t0 = time.perf_counter()
    f, g, gmax, illum = prob.gradient(v)
    hist, taken = [float(np.sqrt(f))], []
    if times is not None:
        times.append(time.perf_counter() - t0)
    for k in range(iterations):
        m = (1.0 / (v * v)).astype(np.float32)
        _, step = gradient_step(prob, v, g, hist[-1], vmin, vmax)
        accepted = False
        for h in range(max_halvings + 1):
            v_try = (1.0 / np.sqrt(np.clip(m - np.float32(step) * g, mlo, mhi))).astype(np.float32)
            f_try, g_try, _, il_try = prob.gradient(v_try)
            if np.sqrt(f_try) < hist[-1]:
                v, g, illum = v_try, g_try, il_try
                hist.append(float(np.sqrt(f_try)))
                taken.append(float(step))
                accepted = True
                break
            step *= 0.5
        if times is not None and accepted:
            times.append(time.perf_counter() - t0)
        if log:
            log("guarded FWI iteration %d: misfit %.6g (%s, %.1f s)" % (k + 1, hist[-1],
                                                                       "accepted" if accepted else "no descent",
                                                                       time.perf_counter() - t0))
        if not accepted:
            break
    return v, hist, time.perf_counter() - t0, taken


def plain_fwi(prob, v0, iterations, record=None, log=None):
    """`iterations` gradient steps from v0; returns (v, misfits, seconds)
    with misfits[k] = sqrt(sum r^2) of the model after k steps.  `record`
    (callable) receives (k, v, g, illum, f) at every gradient evaluation."""
    vmin, vmax = float(v0.min()), float(v0.max())
    v = v0.astype(np.float32)
    hist = []
    t0 = time.perf_counter()
    for k in range(iterations + 1):
        f, g, gmax, illum = prob.gradient(v) if k < iterations else (prob.misfit(v), None, None, None)
        hist.append(float(np.sqrt(f)))
        if log:
            log("plain FWI iteration %d: misfit %.6g (%.1f s)" % (k, hist[-1], time.perf_counter() - t0))
        if k == iterations:
            break
        if record:
            record(k, v, g, illum, f)
        v, _ = gradient_step(prob, v, g, hist[-1], vmin, vmax)
    return v, hist, time.perf_counter() - t0

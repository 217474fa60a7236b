"""Learned FWI preconditioner: training data, training loop and the
misfit-guarded learned update.

Training triples come from plain FWI (fwi.plain_fwi) on seeded random
layered models: network inputs are the normalised gradient, the normalised
illumination and the current model; the target is true model - current
model.  At inversion time the network's proposed velocity update is scaled
by a parabolic line search and accepted only if the plain forward model's
misfit drops; otherwise the iteration falls back to the
illumination-compensated plain gradient step.
"""
import time

import numpy as np

from . import configs, fwi
from .unet import UNet

V_REF, V_SCALE = 3000.0, 1000.0  # model normalisation (m/s)


def random_model(seed, nx, ny, nz, dh=10.0, vmin=1500.0, vmax=4500.0):
    """3-8 dipping layers with increasing velocity trend and 0-3 Gaussian
    anomalies, clipped to [vmin, vmax]; returns (ny, nx, nz) float32."""
    rng = np.random.default_rng(seed)
    x, y, z = np.arange(nx) * dh, np.arange(ny) * dh, np.arange(nz) * dh
    Y, X, Z = np.meshgrid(y, x, z, indexing="ij")
    nl = int(rng.integers(3, 9))
    depths = np.sort(rng.uniform(0.1, 0.95, nl)) * z[-1]
    vels = np.sort(rng.uniform(vmin, vmax, nl + 1))
    v = np.full(X.shape, vels[0])
    for l, z0 in enumerate(depths):
        ax, ay = rng.uniform(-0.15, 0.15, 2)
        surf = z0 + ax * (X - x.mean()) + ay * (Y - y.mean())
        v = np.where(Z >= surf, vels[l + 1], v)
    for _ in range(int(rng.integers(0, 4))):
        c = rng.uniform([0, 0, 0.2], [1, 1, 0.9]) * np.array([x[-1], y[-1], z[-1]])
        sig = rng.uniform(0.05, 0.15) * x[-1]
        amp = rng.uniform(-500, 500)
        v += amp * np.exp(-0.5 * ((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) / sig ** 2)
    return np.ascontiguousarray(np.clip(v, vmin, vmax).astype(np.float32))


def smooth(v, sigma_cells):
    """Initial model: 1 / gaussian_filter(1/v) (run/generate_setup.py:64)."""
    from scipy.ndimage import gaussian_filter
    return (1.0 / gaussian_filter(1.0 / v.astype(np.float64), sigma_cells)).astype(np.float32)


def features(g, illum, v):
    """Network inputs [3][D=ny][H=nx][W=nz]... in the (y, x, z) interior layout."""
    il = illum / (np.abs(illum).max() + 1e-30)
    il = np.log10(il + 1e-4) / 4.0 + 1.0
    return np.stack([g, il, (v - V_REF) / V_SCALE]).astype(np.float32)


def pad4(a):
    """Edge-pad the last three axes to multiples of 4; returns (padded, slices)."""
    pads = [(0, 0)] * (a.ndim - 3) + [(0, (-n) % 4) for n in a.shape[-3:]]
    sl = tuple([slice(None)] * (a.ndim - 3) + [slice(0, n) for n in a.shape[-3:]])
    return np.pad(a, pads, mode="edge"), sl


def survey(nx, ny, nz, dh, nt, dt, fmax, nshots_side=2, nb=20, seed=0):
    """A small surface survey: shots on a grid, receivers on a 2-cell grid."""
    interior = np.zeros((ny, nx, nz), np.float32)
    sx = np.linspace(0.25, 0.75, nshots_side) * (nx - 1) * dh
    sy = np.linspace(0.25, 0.75, nshots_side) * (ny - 1) * dh
    shots = np.array([[a, b, 2 * dh] for b in sy for a in sx])
    gx = np.arange(0, nx, 2) * dh
    gy = np.arange(0, ny, 2) * dh
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 2 * dh)], axis=1)
    return configs.Workload("train", interior, nb, dh, dt, nt, fmax, shots, [rec] * len(shots), rbc_seed=seed)


def observed(w, v_true):
    """Observed data: ABC forward modeling of every shot on the true model."""
    from . import lib
    c = lib.Context(w.nxx, w.nyy, w.nzz, w.nb, w.dh, w.dt, w.nt, w.tlag, nshots=len(w.shots))
    c.set_wavelet(w.wavelet)
    c.set_dampers(fwi.b1d_table(w.nb, 0.004))
    c.set_model_interior(v_true)
    for s in range(len(w.shots)):
        c.set_source(*w.src_index(s), slot=s)
        c.set_receivers(w.rec_index(s), slot=s)
    c.forward(lib.BOUNDARY_ABC)
    out = [c.seismogram(slot=s) for s in range(len(w.shots))]
    c.close()
    return out


def make_triples(n_models, nx, ny, nz, iterations, nt, dt=1e-3, fmax=30.0, seed0=1000, sigma=6.0, nb=20, log=None):
    """Plain FWI on seeded random models; returns a list of (x3, target)."""
    triples = []
    for i in range(n_models):
        v_true = random_model(seed0 + i, nx, ny, nz)
        w = survey(nx, ny, nz, 10.0, nt, dt, fmax, nb=nb, seed=seed0 + i)
        w.interior = v_true
        obs = observed(w, v_true)
        v0 = smooth(v_true, sigma)
        prob = fwi.FwiProblem(w, obs, zmask=20.0)

        def record(k, v, g, illum, f):
            triples.append((features(g, illum, v), ((v_true - v) / V_SCALE).astype(np.float32)))

        fwi.plain_fwi(prob, v0, iterations, record=record)
        prob.close()
        if log:
            log("model %d/%d: %d triples" % (i + 1, n_models, len(triples)))
    return triples


def train(net, triples, steps, patch=(16, 32, 32), batch=4, lr=1e-3, seed=0, log=None):
    """Adam on random patches of the triples; returns the loss history."""
    rng = np.random.default_rng(seed)
    pz, py, px = patch
    losses = []
    for it in range(steps):
        xs, ts = [], []
        for _ in range(batch):
            x3, t = triples[int(rng.integers(len(triples)))]
            ny, nx, nz = t.shape
            # network axes (D, H, W) = (y, x, z) of the interior
            a = int(rng.integers(0, ny - pz + 1))
            b = int(rng.integers(0, nx - py + 1))
            c = int(rng.integers(0, nz - px + 1))
            xs.append(x3[:, a:a + pz, b:b + py, c:c + px])
            ts.append(t[a:a + pz, b:b + py, c:c + px])
        x = np.stack(xs, axis=1)
        tt = np.stack(ts)
        losses.append(net.train_step(x, tt, lr))
        if log and (it % 50 == 0 or it == steps - 1):
            log("train step %d: loss %.5f" % (it, losses[-1]))
    return losses


def predict(net, g, illum, v):
    """Proposed velocity update (m/s) for one model."""
    x3, sl = pad4(features(g, illum, v))
    y = net.forward(x3[:, None])[0]
    return (y[sl[1:]] * V_SCALE).astype(np.float32)


def learned_fwi(prob, net, v0, target_misfit=None, max_iterations=5, log=None, times=None):
    """FWI with the learned update and misfit-guarded fallback.  Stops when
    the misfit reaches target_misfit (if given) or after max_iterations.
    Returns (v, misfits, seconds, used) where used[k] says which update was
    taken ('learned' or 'gradient'); times (a list) receives the elapsed
    seconds at each misfit."""
    vmin, vmax = float(v0.min()), float(v0.max())
    vlo, vhi = fwi.velocity_bounds(prob.w)
    v = v0.astype(np.float32)
    hist, used = [], []
    t0 = time.perf_counter()
    f, g, gmax, illum = prob.gradient(v)
    hist.append(float(np.sqrt(f)))
    if times is not None:
        times.append(time.perf_counter() - t0)
    for k in range(max_iterations):
        if target_misfit is not None and hist[-1] <= target_misfit:
            break
        dv = predict(net, g, illum, v)
        # parabolic line search on the learned direction (s in {0, 0.5, 1})
        f1 = np.sqrt(prob.misfit(np.clip(v + 0.5 * dv, vlo, vhi)))
        f2 = np.sqrt(prob.misfit(np.clip(v + 1.0 * dv, vlo, vhi)))
        s = fwi.parabola_step(0.0, 0.5, 1.0, hist[-1], f1, f2)
        if not np.isfinite(s) or s <= 0:
            s = 0.5 if f1 < f2 else 1.0
        s = float(min(s, 2.0))
        v_try = np.clip(v + s * dv, vlo, vhi).astype(np.float32)
        f_try, g_try, _, il_try = prob.gradient(v_try)
        if np.sqrt(f_try) < hist[-1]:
            v, g, illum = v_try, g_try, il_try
            hist.append(float(np.sqrt(f_try)))
            used.append("learned")
        else:
            v, _ = fwi.gradient_step(prob, v, g, hist[-1], vmin, vmax)
            f, g, gmax, illum = prob.gradient(v)
            hist.append(float(np.sqrt(f)))
            used.append("gradient")
        if times is not None:
            times.append(time.perf_counter() - t0)
        if log:
            log("learned FWI iteration %d (%s): misfit %.6g (%.1f s)" % (k + 1, used[-1], hist[-1],
                                                                        time.perf_counter() - t0))
    return v, hist, time.perf_counter() - t0, used

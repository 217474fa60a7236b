"""Synthetic workloads: seeded velocity models, acquisition geometry and the
parameters.txt files that drive the modeling / migration / inversion mains.

Each builder returns a ``Workload`` with everything in memory; ``write()``
lays it out as the reference expects (parameters.txt keys of
run/parameters.txt, raw float32 model with z fastest, SPS/RPS text files
``x,y,z`` in metres written with %.2f as run/generate_setup.py:50-51 does).

Workloads (sizes are the interior grid; nb boundary cells are added per side):
  two_layer_101     101^3 @ 10 m, nb 50, 2 layers (1500 / 2500 m/s, interface
                    at 500 m), one Ricker source (15 Hz peak) at the centre,
                    20 m deep; 81 x 81 receivers on a 10 m grid at 20 m;
                    dt 1 ms, 1000 samples.
  gradient_401      401 x 401 x 201 @ 10 m, nb 50; 1500 -> 3500 m/s linear
                    gradient, 5 dipping layers of +-300 m/s, Gaussian anomaly
                    (sigma 300 m, +400 m/s); Ricker 20 Hz peak; 201 x 201
                    receivers at 20 m; dt 0.8 ms, 3000 samples.
  gradient_401_16   the same model with 16 shots on a 4 x 4 surface grid, each
                    with its own 201 x 201 receiver patch centred on the shot
                    and clipped at the model edges.
"""
import dataclasses
import os

import numpy as np

RBC_SEED = 20261203
RBC_AMP = 0.6      # +-60 % at the outer edge
CFL_SAFE = 0.45    # boundary velocities are clamped to CFL_SAFE * dh / dt


def ricker(nt, dt, fmax):
    """Modeling::set_wavelet (modeling.cu:36-59) in float32: Ricker with peak
    fmax/3, lag t0 = 2 sqrt(pi)/fmax; returns (wavelet, tlag)."""
    f32 = np.float32
    sqpi = np.sqrt(f32(np.pi), dtype=f32)
    t0 = f32(2.0) * sqpi / f32(fmax)
    fc = f32(fmax) / (f32(3.0) * sqpi)
    tlag = int(t0 / f32(dt)) + 1
    n = np.arange(nt, dtype=np.int64)
    td = (n.astype(f32) * f32(dt) - t0).astype(f32)
    arg = (np.pi * np.pi * np.pi * np.float64(fc) * np.float64(fc) * td.astype(np.float64) *
           td.astype(np.float64)).astype(f32)
    e = np.exp(-arg.astype(np.float64)).astype(f32)
    w = (f32(1e5) * (f32(1.0) - f32(2.0) * arg)).astype(f32) * e
    return w.astype(f32), tlag


def grid_index(x, dh, nb):
    """Metres -> padded index, (int)((x + 0.5 dh)/dh) + nb (modeling.cu:270-272)."""
    x = np.asarray(x, np.float32)
    return ((x + np.float32(0.5) * np.float32(dh)) / np.float32(dh)).astype(np.int32) + nb


@dataclasses.dataclass
class Workload:
    name: str
    interior: np.ndarray      # (ny, nx, nz) float32, z fastest in memory
    nb: int
    dh: float
    dt: float
    nt: int
    fmax: float
    shots: np.ndarray         # (nsrc, 3) metres x, y, z
    receivers: list           # per shot: (nrec_s, 3) metres
    rbc_seed: int = RBC_SEED
    rbc_amp: float = RBC_AMP

    def __post_init__(self):
        self.wavelet, self.tlag = ricker(self.nt, self.dt, self.fmax)

    @property
    def ny(self):
        return self.interior.shape[0]

    @property
    def nx(self):
        return self.interior.shape[1]

    @property
    def nz(self):
        return self.interior.shape[2]

    @property
    def nxx(self):
        return self.nx + 2 * self.nb

    @property
    def nyy(self):
        return self.ny + 2 * self.nb

    @property
    def nzz(self):
        return self.nz + 2 * self.nb

    @property
    def vclamp(self):
        return float(np.float32(CFL_SAFE * self.dh / self.dt))

    def src_index(self, s=0):
        x, y, z = self.shots[s]
        return tuple(int(grid_index(v, self.dh, self.nb)) for v in (x, y, z))

    def rec_index(self, s=0):
        r = self.receivers[s]
        return np.stack([grid_index(r[:, 0], self.dh, self.nb), grid_index(r[:, 1], self.dh, self.nb),
                         grid_index(r[:, 2], self.dh, self.nb)], axis=1).astype(np.int32)

    @property
    def src(self):
        return self.src_index(0)

    @property
    def rec(self):
        return self.rec_index(0)

    @property
    def steps(self):
        return self.nt + self.tlag

    def cell_updates(self, fields=1):
        """Updated cells (nxx-8)(nyy-8)(nzz-8) x steps x fields."""
        return (self.nxx - 8) * (self.nyy - 8) * (self.nzz - 8) * self.steps * fields

    # -- reference file layout ---------------------------------------------
    def write(self, folder, prefix=None, model=None, kind="true", max_iteration=5, batch_shots=1, depth_mask=40.0,
              abc_factor=0.004, rbc_style=None):
        """Write model, SPS/RPS and parameters.txt under `folder`; returns the
        parameters.txt path.  `model` replaces the interior written as the
        model file (e.g. a smoothed starting model for inversion, named with
        `kind`).  The reference reads one RPS file for all shots, so shot 0's
        receivers are written.  `rbc_style` ("blobs" or "ramp") is written
        when given (the drivers default to blobs)."""
        prefix = prefix or self.name
        os.makedirs(folder, exist_ok=True)
        for sub in ("inputs/models", "inputs/geometry", "inputs/data", "outputs/models", "outputs/residuo",
                    "outputs/seismic"):
            os.makedirs(os.path.join(folder, sub), exist_ok=True)
        model_path = os.path.join(folder, "inputs/models",
                                  "%s_%s_%dx%dx%d_%dm.bin" % (prefix, kind, self.nz, self.nx, self.ny, int(self.dh)))
        (self.interior if model is None else np.asarray(model)).astype(np.float32).tofile(model_path)
        model = model_path
        sps = os.path.join(folder, "inputs/geometry", prefix + "_SPS.txt")
        rps = os.path.join(folder, "inputs/geometry", prefix + "_RPS.txt")
        np.savetxt(sps, self.shots, fmt="%.2f", delimiter=",")
        np.savetxt(rps, self.receivers[0], fmt="%.2f", delimiter=",")
        length = (self.nb - 1) * self.dh  # nb = (int)(length/dh) + 1 (modeling.cu:90)
        nrec = len(self.receivers[0])
        dtus = int(np.float32(1e6) * np.float32(self.dt))
        text = PARAMS_TEMPLATE.format(
            nx=self.nx, ny=self.ny, nz=self.nz, dh=self.dh, model=model, sps=sps, rps=rps, nt=self.nt,
            dt=self.dt, fmax=self.fmax, length=length, folder=folder, nrec=nrec, dtus=dtus, seed=self.rbc_seed,
            amp=self.rbc_amp, maxit=max_iteration, batch=batch_shots, zmask=depth_mask, factor=abc_factor)
        if rbc_style is not None:
            text += "rbc_style = %s\n" % rbc_style
        path = os.path.join(folder, "parameters.txt")
        with open(path, "w") as f:
            f.write(text)
        return path


PARAMS_TEMPLATE = """#------------------------------------------------------------
# Model parameters ------------------------------------------
#------------------------------------------------------------

x_samples = {nx}
y_samples = {ny}
z_samples = {nz}

model_spacing = {dh}                # [m]

model_file = {model}

#------------------------------------------------------------
# Geometry --------------------------------------------------
#------------------------------------------------------------

SPS = {sps}
RPS = {rps}

#------------------------------------------------------------
# Modeling parameters ---------------------------------------
#------------------------------------------------------------

time_samples = {nt}
time_spacing = {dt}                 # [s]

max_frequency = {fmax}              # [Hz]

abc_length = {length}               # [m]
abc_factor = {factor}

mod_output_folder = {folder}/inputs/data/

#------------------------------------------------------------
# Inversion parameters --------------------------------------
#------------------------------------------------------------

max_iteration = {maxit}

depth_mask = {zmask}                 # [m]

inv_rbc_ratio = 100                 # [m]
inv_rbc_varVp = 1000                # [m/s]
inv_rbc_length = {length}           # [m]

inv_input_folder = {folder}/inputs/data/
inv_input_prefix = seismogram_nt{nt}_nr{nrec}_{dtus}us_shot_

inv_output_folder = {folder}/outputs/models/
inv_residuo_folder = {folder}/outputs/residuo/

#------------------------------------------------------------
# Migration parameters --------------------------------------
#------------------------------------------------------------

mig_rbc_ratio = 100                 # [m]
mig_rbc_varVp = 1000                # [m/s]
mig_rbc_length = {length}           # [m]

mig_input_folder = {folder}/inputs/data/
mig_input_prefix = seismogram_nt{nt}_nr{nrec}_{dtus}us_shot_
mig_output_folder = {folder}/outputs/seismic/

#------------------------------------------------------------
# MI355X build: random boundary and batching ----------------
#------------------------------------------------------------

rbc_seed = {seed}
rbc_amplitude = {amp}
batch_shots = {batch}
"""


def _grid(lo, hi, n):
    return np.linspace(lo, hi, n, dtype=np.float64)


def two_layer_101():
    nx = ny = nz = 101
    dh = 10.0
    z = np.arange(nz) * dh
    col = np.where(z < 500.0, 1500.0, 2500.0).astype(np.float32)
    interior = np.ascontiguousarray(np.broadcast_to(col, (ny, nx, nz)))
    g = _grid(100.0, 900.0, 81)
    X, Y = np.meshgrid(g, g, indexing="ij")
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 20.0)], axis=1)
    return Workload("two_layer", interior, 50, dh, 1e-3, 1000, 45.0, np.array([[500.0, 500.0, 20.0]]), [rec])


def gradient_model(nx, ny, nz, dh, zmax, depths, centre, sigma, seed=1):
    """Linear gradient 1500 -> 3500 m/s over zmax metres of depth, 5 seeded
    dipping planar layers of +-300 m/s at `depths` (dips up to 0.1 in x and y
    about the model centre) and a Gaussian anomaly of +400 m/s (std `sigma`)
    at `centre` (x, y, z in metres); (ny, nx, nz) float32."""
    rng = np.random.default_rng(seed)
    x = np.arange(nx) * dh
    y = np.arange(ny) * dh
    z = np.arange(nz) * dh
    Y, X, Z = np.meshgrid(y, x, z, indexing="ij")
    v = 1500.0 + 2000.0 * Z / zmax
    xc, yc = x.mean(), y.mean()
    for l, z0 in enumerate(depths):
        ax, ay = rng.uniform(-0.1, 0.1, 2)
        surf = z0 + ax * (X - xc) + ay * (Y - yc)
        v += np.where(Z >= surf, 300.0 if l % 2 == 0 else -300.0, 0.0)
    r2 = (X - centre[0]) ** 2 + (Y - centre[1]) ** 2 + (Z - centre[2]) ** 2
    v += 400.0 * np.exp(-0.5 * r2 / sigma ** 2)
    return np.ascontiguousarray(v.astype(np.float32))


def gradient_model_401(seed=1):
    return gradient_model(401, 401, 201, 10.0, 2000.0, np.linspace(300.0, 1700.0, 5), (2000.0, 2000.0, 1200.0),
                          300.0, seed)


def gradient_401():
    interior = gradient_model_401()
    g = _grid(0.0, 4000.0, 201)
    X, Y = np.meshgrid(g, g, indexing="ij")
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 20.0)], axis=1)
    return Workload("gradient", interior, 50, 10.0, 0.8e-3, 3000, 60.0, np.array([[2000.0, 2000.0, 20.0]]), [rec])


def gradient_401_16(interior=None):
    interior = gradient_model_401() if interior is None else interior
    shots, recs = [], []
    sg = [500.0, 1500.0, 2500.0, 3500.0]
    for sy in sg:
        for sx in sg:
            shots.append([sx, sy, 20.0])
            gx = sx + np.arange(-100, 101) * 20.0
            gy = sy + np.arange(-100, 101) * 20.0
            gx = gx[(gx >= 0.0) & (gx <= 4000.0)]
            gy = gy[(gy >= 0.0) & (gy <= 4000.0)]
            X, Y = np.meshgrid(gx, gy, indexing="ij")
            recs.append(np.stack([X.ravel(), Y.ravel(), np.full(X.size, 20.0)], axis=1))
    return Workload("gradient16", interior, 50, 10.0, 0.8e-3, 3000, 60.0, np.array(shots), recs)


def gradient_survey(nx=101, ny=101, nz=51, nb=20, nt=200, dh=10.0, patch=25, spacing=20.0, full_spread=False):
    """A shrunken version of the 16-shot survey on the gradient / dipping-layer
    / anomaly model: interior nx x ny x nz at dh, the model's depth features
    scaled to the grid, 16 shots on a 4 x 4 surface grid at 20 m depth, each
    with a (2 patch + 1)^2 receiver patch at `spacing` centred on the shot and
    clipped at the model edges (full_spread: every shot records the whole
    surface grid instead, as one RPS file of the reference drivers does);
    dt 0.8 ms, fmax 60 Hz (20 Hz Ricker peak), nt samples."""
    zmax = (nz - 1) * dh
    xmax, ymax = (nx - 1) * dh, (ny - 1) * dh
    interior = gradient_model(nx, ny, nz, dh, zmax, np.linspace(0.15 * zmax, 0.85 * zmax, 5),
                              (xmax / 2, ymax / 2, 0.6 * zmax), 0.15 * min(xmax, zmax))
    shots, recs = [], []
    fx = [xmax * f for f in (0.125, 0.375, 0.625, 0.875)]
    fy = [ymax * f for f in (0.125, 0.375, 0.625, 0.875)]
    for sy in fy:
        for sx in fx:
            shots.append([sx, sy, 20.0])
            if full_spread:
                gx = np.arange(0.0, xmax + 1e-6, spacing)
                gy = np.arange(0.0, ymax + 1e-6, spacing)
            else:
                gx = sx + np.arange(-patch, patch + 1) * spacing
                gy = sy + np.arange(-patch, patch + 1) * spacing
                gx = gx[(gx >= 0.0) & (gx <= xmax)]
                gy = gy[(gy >= 0.0) & (gy <= ymax)]
            X, Y = np.meshgrid(gx, gy, indexing="ij")
            recs.append(np.stack([X.ravel(), Y.ravel(), np.full(X.size, 20.0)], axis=1))
    return Workload("gradient16s", interior, nb, dh, 0.8e-3, nt, 60.0, np.array(shots), recs)


def tiny(nshots=2, nx=40, ny=36, nz=30, nt=160):
    """A small two-layer model for driver tests: nb 8, dt 1 ms, 45 Hz."""
    dh = 10.0
    z = np.arange(nz) * dh
    col = np.where(z < nz * dh / 2, 1600.0, 2200.0).astype(np.float32)
    interior = np.ascontiguousarray(np.broadcast_to(col, (ny, nx, nz)))
    xs = np.linspace(80.0, (nx - 9) * dh, nshots)
    shots = np.stack([xs, np.full(nshots, (ny // 2) * dh), np.full(nshots, 20.0)], axis=1)
    gx = np.arange(20.0, (nx - 2) * dh, 30.0)
    gy = np.arange(20.0, (ny - 2) * dh, 30.0)
    X, Y = np.meshgrid(gx, gy, indexing="ij")
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 20.0)], axis=1)
    return Workload("tiny", interior, 8, dh, 1e-3, nt, 45.0, shots, [rec] * nshots)


WORKLOADS = {"two_layer_101": two_layer_101, "gradient_401": gradient_401, "gradient_401_16": gradient_401_16}

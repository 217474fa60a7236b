"""Parity on the benchmark workloads themselves (run with -m gpu).

  two-layer 101^3 forward modeling (nb 50, 1000 samples): the full
      1000 x 6561 seismogram, GPU FP32 against the FP64 oracle (reported
      error; the FP32 bit-exactness is test_gpu_parity's)
  single-shot forward + RTM at 401 x 401 x 201 (nb 50): forward steps
      [0, 160), which cover the source window (tlag 74), and RTM backward
      steps [0, 64) from that state, fields, seismogram rows, image and
      illumination bit for bit against the oracle
  the 16-shot 4 x 4 survey with clipped receiver patches as one batch
      (autotuned, or shot pairs; graph-captured time loop) on a shrunken
      gradient model: seismograms and the stacked image / illumination bit
      for bit against the oracle run in the batch's summation order
  one inversion_main iteration of the 16-shot FWI problem, shrunken
      (smoothed initial model): misfit, gradient, line-search misfits,
      parabola step and updated model against the oracle
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bin")
FP64_BOUND = 1e-5  # GPU FP32 against the FP64 oracle, relative L2


@pytest.fixture(scope="module")
def fl():
    from fwi3d_amd import lib
    lib.load()
    return lib


def _context(fl, w, vp_pad=None, nshots=1):
    c = fl.Context(w.nxx, w.nyy, w.nzz, w.nb, w.dh, w.dt, w.nt, w.tlag, nshots=nshots)
    c.set_model_interior(w.interior)
    c.set_wavelet(w.wavelet)
    return c


def test_two_layer_101_seismogram_fp32_vs_fp64(fl):
    """The FP32 paths' error against an FP64 solution of the same scheme
    (double fields, the same float velocity model) on the full two-layer
    101^3 run: rel-L2 of the 1000 x 6561 seismogram <= 1e-5 in both
    arithmetic modes."""
    from fwi3d_amd import configs
    w = configs.two_layer_101()
    vp = O.random_boundary(O.expand_boundary(w.interior, w.nb), w.nb, w.rbc_seed, 0, w.rbc_amp, w.vclamp)
    _, _, want = O.forward(vp, w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src, w.rec, fp64=True)
    errs = {}
    for arith in (fl.ARITH_REFERENCE, fl.ARITH_FP32):
        with _context(fl, w) as c:
            c.set_arith(arith)
            c.random_boundary(w.rbc_seed, 0, w.rbc_amp, w.vclamp)
            c.set_source(*w.src)
            c.set_receivers(w.rec)
            c.forward(fl.BOUNDARY_RBC)
            errs[arith] = rel_l2(c.seismogram(), want)
    print("two-layer 101^3 seismogram rel-L2 vs FP64: reference %.3g, fp32 %.3g" % (errs[0], errs[1]))
    assert np.abs(want).max() > 0
    assert errs[fl.ARITH_REFERENCE] <= FP64_BOUND and errs[fl.ARITH_FP32] <= FP64_BOUND, errs


def test_gradient_401_forward_and_rtm_segments_bit_exact(fl):
    """401 x 401 x 201 + 50-cell random boundary, 40401 receivers: forward
    steps [0, 160) and RTM backward steps [0, 64) from the forward state, on
    seeded data, against the oracle bit for bit."""
    from fwi3d_amd import configs
    w = configs.gradient_401()
    vp = O.random_boundary(O.expand_boundary(w.interior, w.nb), w.nb, w.rbc_seed, 0, w.rbc_amp, w.vclamp)
    nf, nbk = 160, 64
    with _context(fl, w) as c:
        c.random_boundary(w.rbc_seed, 0, w.rbc_amp, w.vclamp)
        c.set_source(*w.src)
        c.set_receivers(w.rec)
        assert np.array_equal(c.get_volume(fl.VOL_MODEL), vp)
        c.zero_volume(fl.VOL_FIELD)
        c.zero_volume(fl.VOL_FIELD_OLD)
        c.forward_steps(fl.BOUNDARY_RBC, 0, nf)
        g_P, g_Po = c.get_volume(fl.VOL_FIELD), c.get_volume(fl.VOL_FIELD_OLD)
        g_seis = c.seismogram()
        P, Pold, seis = O.forward(vp, w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src, w.rec, n_end=nf)
        assert np.abs(seis).max() > 0
        assert np.array_equal(g_P, P) and np.array_equal(g_Po, Pold)
        assert np.array_equal(g_seis[:, :nf - w.tlag], seis[:, :nf - w.tlag])
        del g_P, g_Po, g_seis
        rng = np.random.default_rng(1)
        data = rng.standard_normal((len(w.rec), w.nt), dtype=np.float32) * np.float32(1e-3)
        c.set_backward_data(data)
        for v in (fl.VOL_RECV, fl.VOL_RECV_OLD, fl.VOL_IMAGE, fl.VOL_ILLUM):
            c.zero_volume(v)
        c.backward_steps(0, 0, nbk)
        got = [c.get_volume(v) for v in (fl.VOL_FIELD, fl.VOL_RECV, fl.VOL_RECV_OLD, fl.VOL_IMAGE, fl.VOL_ILLUM)]
    S, _, R, Ro, image, illum = O.backward(vp, w.dh, w.dt, w.nt, data, w.rec, P, Pold, b_begin=0, b_end=nbk,
                                           tlag=w.tlag)
    assert np.abs(image).max() > 0 and np.abs(R).max() > 0
    for name, a, b in zip(("S", "R", "Ro", "image", "illum"), got, (S, R, Ro, image, illum)):
        assert np.array_equal(a, b), name


@pytest.mark.parametrize("shape", ["autotuned", "pairs"])
def test_16_shot_survey_batch_bit_exact_vs_oracle(fl, shape):
    """The 16-shot 4 x 4 survey (receiver patches clipped at the edges: 1024
    to 1936 receivers per shot) on a 101 x 101 x 51 gradient model with a
    20-cell boundary, autotuned tile and shots per block (or shot pairs on
    64 x 8 tiles, two z-chunks, for both passes), graph-captured passes:
    seismograms bit for bit against per-shot oracle runs, image and
    illumination bit for bit against the oracle in the batch's order."""
    from fwi3d_amd import configs
    w = configs.gradient_survey()
    S = len(w.shots)
    assert len({len(r) for r in w.receivers}) > 1
    pad = O.expand_boundary(w.interior, w.nb)
    vps = [O.random_boundary(pad, w.nb, w.rbc_seed, s, w.rbc_amp, w.vclamp) for s in range(S)]
    c = _context(fl, w, nshots=S)
    for s in range(S):
        c.shot_random_boundary(s, w.rbc_seed, s, w.rbc_amp, w.vclamp)
        c.set_source(*w.src_index(s), slot=s)
        c.set_receivers(w.rec_index(s), slot=s)
    if shape == "pairs":
        c.set_shots_per_thread(fl.KIND_FWD_RBC, 2, 8, 2)
        c.set_shots_per_thread(fl.KIND_BWD_RTM, 2, 8, 2)
    else:
        c.autotune(fl.KIND_FWD_RBC, 3)
        c.autotune(fl.KIND_BWD_RTM, 3)
    sb = c.shots_per_block(fl.KIND_BWD_RTM)
    assert shape != "pairs" or (sb == 2 and c.shots_per_thread(fl.KIND_BWD_RTM) == 2)
    print("16-shot survey tuning: forward %s sb=%d, rtm %s sb=%d"
          % (c.tuning(0), c.shots_per_block(0), c.tuning(2), sb))
    c.set_graph(True)
    c.forward(fl.BOUNDARY_RBC)
    seis = [c.seismogram(slot=s) for s in range(S)]
    for s in range(S):
        c.set_backward_data(seis[s], slot=s)
    c.zero_volume(fl.VOL_IMAGE)
    c.zero_volume(fl.VOL_ILLUM)
    c.rtm_backward()
    img_b, ill_b = c.get_volume(fl.VOL_IMAGE), c.get_volume(fl.VOL_ILLUM)
    c.close()
    states = []
    for s in range(S):
        P, Pold, want = O.forward(vps[s], w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src_index(s), w.rec_index(s))
        assert np.array_equal(seis[s], want), s
        states.append([P, Pold, np.zeros_like(P), np.zeros_like(P)])
    G = S // sb
    imgs = [np.zeros_like(pad) for _ in range(G)]
    ills = [np.zeros_like(pad) for _ in range(G)]
    O.backward_batched(vps, sb, w.dh, w.dt, w.nt, seis, [w.rec_index(s) for s in range(S)], states, imgs, ills,
                       tlag=w.tlag)
    want_img, want_ill = imgs[0], ills[0]
    for g in range(1, G):  # the batch folds its groups in group order
        want_img = want_img + imgs[g]
        want_ill = want_ill + ills[g]
    assert np.abs(want_img).max() > 0
    assert np.array_equal(ill_b, want_ill)
    assert np.array_equal(img_b, want_img)


def _run(exe, params):
    r = subprocess.run([os.path.join(BIN, exe), params], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:]
    return r.stdout


@pytest.mark.parametrize("case", ["vertex", "guarded"])
def test_16_shot_inversion_iteration_vs_oracle(fl, tmp_path, case):
    """One inversion_main iteration (16 shots, batch_shots = 1, reference-style
    blob boundary) on a shrunken gradient model from a smoothed initial model:
    the first misfit, the exported normalised gradient, the two line-search
    misfits, the parabola step (inversion.cu:291-312) and the written model
    (inversion.cu:349-362) against the oracle.  "vertex": 40 x 40 x 32,
    depth mask 100 m, the parabola's vertex gives a usable model and is
    taken, as in the reference.  "guarded": 40 x 40 x 24, depth mask 30 m, the
    vertex lies 25 trial steps out and would drive the slowness negative
    (NaN velocities in the reference); the driver takes the better trial
    step instead."""
    from scipy.ndimage import gaussian_filter
    from fwi3d_amd import configs
    nz, nt, sigma, zmask = (32, 250, 1.0, 100.0) if case == "vertex" else (24, 200, 2.0, 30.0)
    w = configs.gradient_survey(nx=40, ny=40, nz=nz, nb=10, nt=nt, full_spread=True, spacing=40.0)
    folder = str(tmp_path / "fwi3")
    _run("modeling_main", w.write(folder))  # observed data, absorbing boundary
    init = (1.0 / gaussian_filter(1.0 / w.interior.astype(np.float64), sigma)).astype(np.float32)
    params = w.write(folder, model=init, kind="init", max_iteration=1, batch_shots=1, depth_mask=zmask)
    with open(params, "a") as f:
        f.write("inv_export_gradient = true\n")
    out = _run("inversion_main", params)
    line = re.search(r"Line search: a1 = (\S+), a2 = (\S+), f\(0\) = (\S+), f\(a1\) = (\S+), f\(a2\) = (\S+), "
                     r"step = (\S+)\n", out)
    assert line, out[-2000:]
    a1, a2, f0, f1, f2, vertex = (np.float32(float(v)) for v in line.groups())
    guard = re.search(r"Line search: the vertex leaves the valid model range; step = (\S+) ", out)
    step = np.float32(float(guard.group(1))) if guard else vertex
    assert (guard is not None) == (case == "guarded"), out[-1500:]
    dims = "%dx%dx%d_%dm.bin" % (w.nz, w.nx, w.ny, int(w.dh))
    g_file = np.fromfile(os.path.join(folder, "outputs/models", "gradient_iteration_1_60Hz_" + dims), np.float32)
    m_file = np.fromfile(os.path.join(folder, "outputs/models", "model_FWI_iteration_1_60Hz_" + dims), np.float32)

    f32 = np.float32
    nb, S = w.nb, len(w.shots)
    obs = []
    for s in range(S):
        name = "seismogram_nt%d_nr%d_%dus_shot_%d.bin" % (w.nt, len(w.receivers[0]), int(f32(1e6) * f32(w.dt)),
                                                          s + 1)
        obs.append(np.fromfile(os.path.join(folder, "inputs/data", name), np.float32).reshape(-1, w.nt))
    tables = O.dampers(nb, 0.004)
    rec = w.rec_index(0)

    def misfit(v):
        vp = O.expand_boundary(v, nb)
        total, res = 0.0, []
        for s in range(S):
            _, _, cal = O.forward(vp, nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src_index(s), rec,
                                  abc_tables=tables)
            r, ssum, _ = O.misfit(obs[s], cal)
            res.append(r.reshape(cal.shape))
            total += ssum
        return total, res

    m = (f32(1.0) / (init * init)).astype(f32)
    v0 = (f32(1.0) / np.sqrt(m)).astype(f32)
    sum0, res0 = misfit(v0)
    assert abs(float(np.sqrt(sum0)) - f0) <= 1e-6 * f0, (np.sqrt(sum0), f0)
    # gradient: random-boundary forward + FWI backward per shot, shot after shot
    vclamp = f32(0.45) * f32(w.dh) / f32(w.dt)
    ratio, var_vp = f32(100.0), f32(1000.0)
    vlo, vhi = f32(v0.min()) - var_vp, f32(v0.max()) + var_vp
    pad0 = O.expand_boundary(v0, nb)
    grad = np.zeros_like(pad0)
    illum = np.zeros_like(pad0)
    for s in range(S):
        blobs = fl.rbc_blobs(w.rbc_seed, s, w.nxx, w.nyy, w.nzz, nb, w.dh, ratio, var_vp)
        vp = O.random_boundary_blobs(pad0, nb, w.dh, w.rbc_seed, s, var_vp, blobs, vlo, vhi, vclamp)
        P, Pold, _ = O.forward(vp, nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src_index(s), rec, record=False)
        out_b = O.backward(vp, w.dh, w.dt, w.nt, res0[s], rec, P, Pold, image=grad, illum=illum, fwi=True,
                           tlag=w.tlag)
        grad, illum = out_b[4], out_b[5]
    g, _ = O.gradient_post(grad, illum, nb, w.dh, zmask)
    assert np.abs(g).max() > 0
    assert np.array_equal(g_file, g.ravel())
    # line search (inversion.cu:291-347)
    vmin, vmax = f32(init.min()), f32(init.max())
    dm = f32(1.0) / (vmin * vmin) - f32(1.0) / (vmax * vmax)
    assert a1 == f32(0.02) * dm and a2 == f32(0.05) * dm
    for a, f in ((a1, f1), (a2, f2)):
        _, vt = O.model_step(m, g, a)
        st, _ = misfit(vt.reshape(m.shape))
        assert abs(float(np.sqrt(st)) - f) <= 1e-6 * f, (a, np.sqrt(st), f)
    want_vertex = O.parabola_step(f32(0.0), a1, a2, f0, f1, f2)
    assert np.isfinite(want_vertex) and f32(want_vertex) == vertex, (want_vertex, vertex)
    mo, v1 = O.model_step(m, g, vertex)
    if case == "guarded":  # the vertex model is unusable; the better trial is taken
        assert (mo <= 0).any() or not np.isfinite(v1).all() or v1.max() > vclamp
        assert step == (a1 if f1 < f2 else a2)
        _, v1 = O.model_step(m, g, step)
    assert np.isfinite(v1).all() and v1.max() <= vclamp
    assert np.array_equal(m_file, v1)
    conv = np.loadtxt(os.path.join(folder, "outputs/residuo", "convergence_1_iterations.txt"))
    assert conv.shape == (2,) and conv[1] < conv[0]
    s1, _ = misfit(v1.reshape(m.shape))
    assert abs(float(np.sqrt(s1)) - conv[1]) <= 1e-5 * conv[1]

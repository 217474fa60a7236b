"""End-to-end runs of the modeling / migration / inversion drivers on a tiny
parameters.txt (run with -m gpu)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from tests.cases import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bin")


def _run(exe, params):
    r = subprocess.run([os.path.join(BIN, exe), params], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "Run time" in r.stdout
    return r.stdout


def _seis(folder, w, shot):
    f = os.path.join(folder, "inputs/data", "seismogram_nt%d_nr%d_%dus_shot_%d.bin"
                     % (w.nt, len(w.receivers[0]), int(np.float32(1e6) * np.float32(w.dt)), shot + 1))
    return np.fromfile(f, np.float32).reshape(len(w.receivers[0]), w.nt)


@pytest.fixture(scope="module")
def tiny():
    from fwi3d_amd import configs
    return configs.tiny(nshots=3)


def test_modeling_main_matches_oracle(tiny, tmp_path):
    w = tiny
    folder = str(tmp_path / "run")
    _run("modeling_main", w.write(folder))
    vp = O.expand_boundary(w.interior, w.nb)
    tables = O.dampers(w.nb, 0.004)
    for s in range(len(w.shots)):
        _, _, want = O.forward(vp, w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src_index(s), w.rec_index(s),
                               abc_tables=tables)
        assert np.array_equal(_seis(folder, w, s), want), s


def test_modeling_main_batched_is_identical(tiny, tmp_path):
    w = tiny
    a, b = str(tmp_path / "seq"), str(tmp_path / "bat")
    _run("modeling_main", w.write(a))
    _run("modeling_main", w.write(b, batch_shots=2))  # 3 shots: one full batch + one partial
    for s in range(len(w.shots)):
        assert np.array_equal(_seis(a, w, s), _seis(b, w, s)), s


def test_migration_main_sequential_vs_batched(tiny, tmp_path):
    w = tiny
    imgs = []
    for name, batch in (("seq", 1), ("bat", 3)):
        folder = str(tmp_path / name)
        params = w.write(folder, batch_shots=batch)
        _run("modeling_main", params)
        _run("migration_main", params)
        f = os.path.join(folder, "outputs/seismic", "RTM_section_45Hz_%dx%dx%d_10m.bin" % (w.nz, w.nx, w.ny))
        img = np.fromfile(f, np.float32).reshape(w.ny, w.nx, w.nz)
        ill = np.fromfile(f.replace("RTM_section", "RTM_illumination"), np.float32)
        assert np.isfinite(img).all() and np.abs(img).max() > 0 and ill.min() > 0
        imgs.append(img)
    assert rel_l2(imgs[1], imgs[0]) <= 1e-4


def test_partial_batch_matches_per_shot(tiny, tmp_path):
    """3 shots in batches of 2 leave one slot of the last batch unused: it
    is disabled (no source, no receivers), so it adds nothing to the RTM
    illumination or the FWI gradient.  Image, illumination and gradient
    equal the per-shot runs' within the batched summation's rounding."""
    from scipy.ndimage import gaussian_filter
    w = tiny
    outs = {}
    for batch in (1, 2):
        folder = str(tmp_path / ("b%d" % batch))
        params = w.write(folder, batch_shots=batch)
        _run("modeling_main", params)
        _run("migration_main", params)
        f = os.path.join(folder, "outputs/seismic", "RTM_section_45Hz_%dx%dx%d_10m.bin" % (w.nz, w.nx, w.ny))
        img = np.fromfile(f, np.float32)
        ill = np.fromfile(f.replace("RTM_section", "RTM_illumination"), np.float32)
        init = (1.0 / gaussian_filter(1.0 / w.interior.astype(np.float64), 2.0)).astype(np.float32)
        params = w.write(folder, model=init, kind="init", max_iteration=1, batch_shots=batch, depth_mask=20.0)
        with open(params, "a") as fh:
            fh.write("inv_export_gradient = true\n")
        _run("inversion_main", params)
        g = np.fromfile(os.path.join(folder, "outputs/models", "gradient_iteration_1_45Hz_%dx%dx%d_10m.bin"
                                     % (w.nz, w.nx, w.ny)), np.float32)
        outs[batch] = (img, ill, g)
    for a, b in zip(outs[2], outs[1]):
        assert np.abs(b).max() > 0
        assert rel_l2(a, b) <= 1e-4, rel_l2(a, b)


def test_inversion_main_lowers_misfit(tiny, tmp_path):
    from scipy.ndimage import gaussian_filter
    w = tiny
    folder = str(tmp_path / "fwi")
    _run("modeling_main", w.write(folder))  # observed data from the true model
    init = (1.0 / gaussian_filter(1.0 / w.interior.astype(np.float64), 2.0)).astype(np.float32)
    params = w.write(folder, model=init, kind="init", max_iteration=2, batch_shots=3, depth_mask=20.0)
    out = _run("inversion_main", params)
    conv = np.loadtxt(os.path.join(folder, "outputs/residuo", "convergence_2_iterations.txt"))
    assert conv.shape == (3,), out[-1500:]
    assert conv[-1] < conv[0], conv
    final = np.fromfile(os.path.join(folder, "outputs/models", "final_model_FWI_45Hz_%dx%dx%d_10m.bin"
                                     % (w.nz, w.nx, w.ny)), np.float32).reshape(w.interior.shape)
    assert np.isfinite(final).all()
    assert rel_l2(final, w.interior) < rel_l2(init, w.interior)


def test_inversion_main_model_upload_modes(tiny, tmp_path):
    """fwi_model_upload = last_trial reproduces the reference's model
    handling: the next iteration's misfit is computed with the last
    line-search trial model, so it equals that trial's misfit f(a2); the
    default uploads the updated model instead."""
    import re
    from scipy.ndimage import gaussian_filter
    w = tiny
    init = (1.0 / gaussian_filter(1.0 / w.interior.astype(np.float64), 2.0)).astype(np.float32)
    f2s, convs = [], []
    for mode in ("updated", "last_trial"):
        folder = str(tmp_path / mode)
        _run("modeling_main", w.write(folder))
        params = w.write(folder, model=init, kind="init", max_iteration=2, batch_shots=3, depth_mask=20.0)
        with open(params, "a") as f:
            f.write("fwi_model_upload = %s\n" % mode)
        out = _run("inversion_main", params)
        f2s.append([float(x) for x in re.findall(r"f\(a2\) = ([0-9.eE+-]+),", out)])
        convs.append(np.loadtxt(os.path.join(folder, "outputs/residuo", "convergence_2_iterations.txt")))
    assert len(f2s[0]) == len(f2s[1]) == 2
    assert convs[0][0] == convs[1][0] and f2s[0][0] == f2s[1][0]  # same first iteration
    assert abs(convs[1][1] - f2s[1][0]) <= 1e-5 * f2s[1][0]
    assert abs(convs[0][1] - f2s[0][0]) > 1e-4 * f2s[0][0]


def test_migration_main_backward_corrections(tiny, tmp_path):
    """The backward_corrections key reaches the library: with source
    subtraction and aligned imaging the RTM section changes (and stays
    finite); an unknown item is refused."""
    w = tiny
    imgs = []
    for name, corr in (("ref", None), ("corr", "subtract_source, aligned_imaging")):
        folder = str(tmp_path / name)
        params = w.write(folder, batch_shots=3)
        _run("modeling_main", params)
        if corr:
            with open(params, "a") as f:
                f.write("backward_corrections = %s\n" % corr)
        _run("migration_main", params)
        f = os.path.join(folder, "outputs/seismic", "RTM_section_45Hz_%dx%dx%d_10m.bin" % (w.nz, w.nx, w.ny))
        imgs.append(np.fromfile(f, np.float32))
    assert np.isfinite(imgs[1]).all() and np.abs(imgs[1]).max() > 0
    assert rel_l2(imgs[1], imgs[0]) > 1e-3
    lines = [l for l in open(params) if not l.startswith("backward_corrections")]
    with open(params, "w") as f:
        f.writelines(lines + ["backward_corrections = gradient_dyy, bogus\n"])
    r = subprocess.run([os.path.join(BIN, "migration_main"), params], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=240)
    assert r.returncode != 0 and "bogus" in r.stdout


def test_migration_main_rbc_style(tiny, tmp_path):
    """rbc_style selects the boundary: the default (the reference's blobs)
    and the ramp give different finite sections; an unknown style is refused."""
    w = tiny
    imgs = []
    for name, style in (("default", None), ("ramp", "ramp"), ("blobs", "blobs")):
        folder = str(tmp_path / name)
        params = w.write(folder, batch_shots=3, rbc_style=style)
        _run("modeling_main", params)
        _run("migration_main", params)
        f = os.path.join(folder, "outputs/seismic", "RTM_section_45Hz_%dx%dx%d_10m.bin" % (w.nz, w.nx, w.ny))
        imgs.append(np.fromfile(f, np.float32))
        assert np.isfinite(imgs[-1]).all() and np.abs(imgs[-1]).max() > 0
    assert np.array_equal(imgs[0], imgs[2])
    assert rel_l2(imgs[1], imgs[0]) > 1e-3
    with open(params, "a") as f:
        f.write("rbc_style = bogus\n")
    text = open(params).read().replace("rbc_style = blobs\n", "")
    with open(params, "w") as f:
        f.write(text)
    r = subprocess.run([os.path.join(BIN, "migration_main"), params], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True, timeout=240)
    assert r.returncode != 0 and "bogus" in r.stdout


def test_full_workflow_in_cpp(tiny, tmp_path):
    """The reference workflow end to end with no Python step: setup_main
    (acquisition + smoothed initial model), modeling_main, preprocess_main
    (RTM input data), migration_main on the preprocessed data, and
    inversion_main from the initial model (the reference's model_file edited
    to it by hand; here by rewriting that line)."""
    w = tiny
    folder = str(tmp_path / "flow")
    params = w.write(folder, depth_mask=20.0, max_iteration=2, batch_shots=2)
    text = open(params).read()
    text = text.replace("mig_input_folder = %s/inputs/data/" % folder,
                        "mig_input_folder = %s/inputs/data/input_RTM_" % folder)
    # setup_main writes an 8 x 8 receiver grid: the data file names carry nr64
    text = text.replace("_nr%d_" % len(w.receivers[0]), "_nr64_")
    text += ("setup_shots = 2,2\nsetup_shot_range = 100,300,100,250\nsetup_shot_depth = 20\n"
             "setup_receivers = 8,8\nsetup_receiver_range = 20,370,20,330\nsetup_receiver_depth = 20\n"
             "setup_smoothing = 2\n")
    with open(params, "w") as f:
        f.write(text)
    _run_plain = lambda exe: subprocess.run([os.path.join(BIN, exe), params], stdout=subprocess.PIPE,
                                            stderr=subprocess.STDOUT, text=True, timeout=240)
    r = _run_plain("setup_main")
    assert r.returncode == 0, r.stdout
    _run("modeling_main", params)
    r = _run_plain("preprocess_main")
    assert r.returncode == 0 and "4 shots" in r.stdout, r.stdout
    data = os.path.join(folder, "inputs/data")
    assert len([f for f in os.listdir(data) if f.startswith("input_RTM_")]) == 4
    _run("migration_main", params)
    img = np.fromfile(os.path.join(folder, "outputs/seismic", "RTM_section_45Hz_%dx%dx%d_10m.bin"
                                   % (w.nz, w.nx, w.ny)), np.float32)
    assert img.size == w.nx * w.ny * w.nz and np.isfinite(img).all() and np.abs(img).max() > 0
    model_true = [l for l in text.splitlines() if l.startswith("model_file")][0].split("=")[1].strip()
    with open(params, "w") as f:
        f.write(text.replace("model_file = " + model_true, "model_file = " + model_true.replace("true", "init")))
    out = _run("inversion_main", params)
    conv = np.loadtxt(os.path.join(folder, "outputs/residuo", "convergence_2_iterations.txt"))
    assert conv.shape == (3,) and conv[-1] < conv[0], out[-1500:]


def test_drivers_autotune_key(tiny, tmp_path):
    """The drivers' optional `autotune` key.  "on": the forward passes take
    the autotuner's launch shapes and the backward passes fixed one-shot
    tiles, so the seismograms, the RTM image and the FWI convergence curve
    equal the default run's bit for bit; "all": every pass kind is timed
    (shots per block or per thread included), the image sums the shots in
    the chosen grouping and stays within rounding of the default run."""
    from scipy.ndimage import gaussian_filter
    w = tiny
    out = {}
    for mode in ("off", "on", "all"):
        folder = str(tmp_path / mode)
        params = w.write(folder, batch_shots=2)
        with open(params, "a") as f:
            f.write("autotune = %s\nautotune_cache = %s\n" % (mode, os.path.join(folder, "tune.txt")))
        _run("modeling_main", params)
        _run("migration_main", params)
        img = np.fromfile(os.path.join(folder, "outputs/seismic", "RTM_section_45Hz_%dx%dx%d_10m.bin"
                                       % (w.nz, w.nx, w.ny)), np.float32)
        seis = [_seis(folder, w, s) for s in range(len(w.shots))]
        init = (1.0 / gaussian_filter(1.0 / w.interior.astype(np.float64), 2.0)).astype(np.float32)
        fparams = w.write(folder, model=init, kind="init", max_iteration=1, batch_shots=2, depth_mask=20.0)
        with open(fparams, "a") as f:
            f.write("autotune = %s\n" % mode)
        _run("inversion_main", fparams)
        conv = np.loadtxt(os.path.join(folder, "outputs/residuo", "convergence_1_iterations.txt"))
        out[mode] = (seis, img, conv)
        if mode != "off":
            assert os.path.getsize(os.path.join(folder, "tune.txt")) > 0
    for s in range(len(w.shots)):
        assert np.array_equal(out["on"][0][s], out["off"][0][s]), s
        assert np.array_equal(out["all"][0][s], out["off"][0][s]), s
    assert np.array_equal(out["on"][1], out["off"][1])
    assert np.array_equal(out["on"][2], out["off"][2])
    assert rel_l2(out["all"][1], out["off"][1]) <= 1e-4

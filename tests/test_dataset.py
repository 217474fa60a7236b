"""The C++ data-preparation tools against restatements of the reference's
Python steps (no GPU).

run/generate_setup.py and run/generate_input_data.py are restated here with
numpy/scipy, float32 where the scripts compute in float32: acquisition text
files must match character for character and the Gaussian filter bit for bit
(scipy's weights, reflect extension and symmetric-pair summation order are
reproduced); the initial model and the RTM input data may differ from the
restatement by at most one float32 ulp (a division after the filter; libm's
exp against numpy's)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
from scipy.ndimage import gaussian_filter

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bin")
HOST_LIB = os.path.join(ROOT, "fwi3d_amd", "libfwi3d_host.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST_LIB):
        pytest.fail("libfwi3d_host.so is not built (python -m fwi3d_amd.build)")
    L = ctypes.CDLL(HOST_LIB)
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.fwi3d_host_gaussian_filter3d.argtypes = [f32, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    L.fwi3d_host_rtm_input.argtypes = [f32, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, f32, f32]
    return L


def _ulps(a, b):
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    return int(np.abs(a - b).max())


@pytest.mark.parametrize("sigma,shape", [(3.0, (20, 30, 28)), (1.5, (9, 11, 7)), (4.0, (5, 40, 6))])
def test_gaussian_filter_matches_scipy(host, sigma, shape):
    """Reflect boundaries (also where the 4-sigma radius exceeds the axis),
    one pass per axis in float32, symmetric-pair accumulation in double."""
    rng = np.random.default_rng(int(sigma * 10))
    n0, n1, n2 = shape
    v = rng.uniform(1.0 / 4500, 1.0 / 1500, n0 * n1 * n2).astype(np.float32)
    want = gaussian_filter(v.reshape((n0, n1, n2), order="F"), sigma).flatten("F")
    got = v.copy()
    host.fwi3d_host_gaussian_filter3d(got, n0, n1, n2, sigma)
    assert np.array_equal(got, want), _ulps(got, want)


def _params(folder, nx, ny, nz, dh, zmask, nt, dt, extra=""):
    os.makedirs(os.path.join(folder, "data"), exist_ok=True)
    p = os.path.join(folder, "parameters.txt")
    with open(p, "w") as f:
        f.write("x_samples = %d\ny_samples = %d\nz_samples = %d\nmodel_spacing = %g\n" % (nx, ny, nz, dh))
        f.write("model_file = %s/model_true_%dx%dx%d.bin\n" % (folder, nz, nx, ny))
        f.write("SPS = %s/sps.txt\nRPS = %s/rps.txt\n" % (folder, folder))
        f.write("time_samples = %d\ntime_spacing = %g\nmod_output_folder = %s/data/\n" % (nt, dt, folder))
        f.write("depth_mask = %g\n%s" % (zmask, extra))
    return p


def _run(exe, params):
    r = subprocess.run([os.path.join(BIN, exe), params], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    return r.stdout


def test_setup_main_matches_generate_setup(tmp_path):
    folder = str(tmp_path)
    nx, ny, nz, dh, zmask = 30, 28, 20, 25.0, 100.0
    rng = np.random.default_rng(3)
    m_true = rng.uniform(1500, 4500, nz * nx * ny).astype(np.float32)
    params = _params(folder, nx, ny, nz, dh, zmask, 10, 1e-3)
    m_true.tofile(os.path.join(folder, "model_true_%dx%dx%d.bin" % (nz, nx, ny)))
    _run("setup_main", params)

    # generate_setup.py:24-51
    def grid(a0, a1, na, b0, b1, nb, depth):
        A, B = np.meshgrid(np.linspace(a0, a1, na), np.linspace(b0, b1, nb))
        out = np.zeros((na * nb, 3))
        out[:, 0] = np.reshape(A, [na * nb], order="F")
        out[:, 1] = np.reshape(B, [na * nb], order="F")
        out[:, 2] = depth
        return out

    for name, want in (("sps.txt", grid(1000, 4000, 6, 1000, 4000, 6, 100.0)),
                       ("rps.txt", grid(0, 5000, 51, 0, 5000, 51, 0.0))):
        ref = str(tmp_path / ("want_" + name))
        np.savetxt(ref, want, fmt="%.2f", delimiter=",")
        assert open(os.path.join(folder, name)).read() == open(ref).read(), name

    # generate_setup.py:53-68
    mt = m_true.reshape([nz, nx, ny], order="F")
    m_init = 1.0 / gaussian_filter(1.0 / mt, 3.0)
    zid = int(zmask / dh)
    m_init[:zid] = mt[:zid]
    want = m_init.flatten("F").astype(np.float32)
    got = np.fromfile(os.path.join(folder, "model_init_%dx%dx%d.bin" % (nz, nx, ny)), np.float32)
    assert got.size == want.size
    assert _ulps(got, want) <= 1, _ulps(got, want)
    top = np.zeros((nz, nx, ny), bool)
    top[:zid] = True
    assert np.array_equal(got[top.flatten("F")], m_true[top.flatten("F")])


def test_setup_main_optional_geometry_keys(tmp_path):
    folder = str(tmp_path)
    extra = ("setup_shots = 2,3\nsetup_shot_range = 100,200,300,500\nsetup_shot_depth = 20\n"
             "setup_receivers = 3,2\nsetup_receiver_range = 0,50,10,20\nsetup_receiver_depth = 10\n")
    params = _params(folder, 8, 8, 8, 10.0, 0.0, 10, 1e-3, extra)
    np.full(512, 2000.0, np.float32).tofile(os.path.join(folder, "model_true_8x8x8.bin"))
    _run("setup_main", params)
    sps = np.loadtxt(os.path.join(folder, "sps.txt"), delimiter=",")
    assert sps.tolist() == [[100, 300, 20], [100, 400, 20], [100, 500, 20], [200, 300, 20], [200, 400, 20],
                            [200, 500, 20]]
    rps = np.loadtxt(os.path.join(folder, "rps.txt"), delimiter=",")
    assert rps.tolist() == [[0, 10, 10], [0, 20, 10], [25, 10, 10], [25, 20, 10], [50, 10, 10], [50, 20, 10]]


def _input_rtm(data, nt, dt, sps, rps, s):
    """generate_input_data.py:28-46 for shot s (data is (nt, nr), float32)."""
    data = data.copy()
    data *= 1.0 / np.max(np.abs(data))
    distance = np.sqrt((rps[:, 0] - sps[s, 0]) ** 2 + (rps[:, 1] - sps[s, 1]) ** 2)
    tid = np.array((distance / 3600 + 0.18) / dt, dtype=int)
    t = np.arange(nt)
    for r in range(len(rps)):
        data[:tid[r], r] *= np.exp(-0.5 * ((t - tid[r]) / 30) ** 2)[:tid[r]]
    return data


def test_preprocess_main_matches_generate_input_data(tmp_path):
    folder = str(tmp_path)
    nt, dt = 400, 1e-3
    params = _params(folder, 8, 8, 8, 10.0, 0.0, nt, dt)
    sps = np.array([[100.0, 150.0, 20.0], [320.5, 80.25, 20.0]], np.float32)
    rps = np.array([[x, y, 0.0] for x in (0.0, 150.0, 700.0) for y in (10.0, 400.0)], np.float32)
    np.savetxt(os.path.join(folder, "sps.txt"), sps, fmt="%.2f", delimiter=",")
    np.savetxt(os.path.join(folder, "rps.txt"), rps, fmt="%.2f", delimiter=",")
    rng = np.random.default_rng(5)
    raw = []
    for s in range(len(sps)):
        d = rng.normal(0, 3, (len(rps), nt)).astype(np.float32)  # file layout t + r*nt
        name = "seismogram_nt%d_nr%d_%.0fus_shot_%d.bin" % (nt, len(rps), dt * 1e6, s + 1)
        d.tofile(os.path.join(folder, "data", name))
        raw.append((name, d))
    _run("preprocess_main", params)
    sps_r = np.loadtxt(os.path.join(folder, "sps.txt"), dtype=np.float32, delimiter=",")
    rps_r = np.loadtxt(os.path.join(folder, "rps.txt"), dtype=np.float32, delimiter=",")
    for s, (name, d) in enumerate(raw):
        want = _input_rtm(d.T, nt, dt, sps_r, rps_r, s).T
        got = np.fromfile(os.path.join(folder, "data", "input_RTM_" + name), np.float32).reshape(d.shape)
        assert _ulps(got, want) <= 1, (s, _ulps(got, want))
        assert np.abs(got).max() <= 1.0 + 1e-7

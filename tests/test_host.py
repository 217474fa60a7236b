"""Host-side set-up of the drivers (no GPU): parameter parsing with the
reference's semantics and the wavelet / damper arithmetic fed to the
device library, compared bit for bit with the oracle."""
import ctypes
import os

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_LIB = os.path.join(ROOT, "fwi3d_amd", "libfwi3d_host.so")


@pytest.fixture(scope="module")
def host():
    if not os.path.exists(HOST_LIB):
        pytest.fail("libfwi3d_host.so is not built (python -m fwi3d_amd.build)")
    L = ctypes.CDLL(HOST_LIB)
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.fwi3d_host_wavelet.argtypes = [ctypes.c_int, ctypes.c_float, ctypes.c_float, f32]
    L.fwi3d_host_wavelet.restype = ctypes.c_int
    L.fwi3d_host_b1d.argtypes = [ctypes.c_int, ctypes.c_float, f32]
    L.fwi3d_host_catch_parameter.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
    L.fwi3d_host_catch_parameter.restype = ctypes.c_int
    L.fwi3d_host_parabola_vertex.argtypes = [ctypes.c_float] * 6
    L.fwi3d_host_parabola_vertex.restype = ctypes.c_float
    L.fwi3d_host_usable_step.argtypes = [ctypes.c_float] * 5 + [f32, f32, ctypes.c_long, ctypes.c_float,
                                                                ctypes.POINTER(ctypes.c_int)]
    L.fwi3d_host_usable_step.restype = ctypes.c_float
    return L


def test_wavelet_matches_oracle(host):
    for nt, dt, fmax in ((1000, 1e-3, 45.0), (3000, 0.8e-3, 60.0), (2001, 1e-3, 50.0)):
        w = np.zeros(nt, np.float32)
        tlag = host.fwi3d_host_wavelet(nt, dt, fmax, w)
        wo, to = O.wavelet(nt, dt, fmax)
        assert tlag == to
        assert np.array_equal(w, wo)


def test_dampers_match_oracle(host):
    for nb, f in ((51, 0.004), (8, 0.05), (50, 0.004)):
        b = np.zeros(nb, np.float32)
        host.fwi3d_host_b1d(nb, f, b)
        assert np.array_equal(b, O.dampers(nb, f)[0])


def test_python_wavelet_matches_oracle():
    from fwi3d_amd import configs
    for nt, dt, fmax in ((1000, 1e-3, 45.0), (3000, 0.8e-3, 60.0)):
        w, tlag = configs.ricker(nt, dt, fmax)
        wo, to = O.wavelet(nt, dt, fmax)
        assert tlag == to
        # float32 restatement in numpy: exp is evaluated in double and rounded
        assert np.max(np.abs(w.view(np.int32) - wo.view(np.int32))) <= 1


def test_catch_parameter_semantics(host, tmp_path):
    p = tmp_path / "parameters.txt"
    p.write_text("# comment = 1\n\n   indented = 2\nx_samples = 201      # trailing\n"
                 "x_samples_extra = 7\nmodel_file = ../inputs/a b.bin # c\nempty =\n")

    def get(k):
        buf = ctypes.create_string_buffer(256)
        n = host.fwi3d_host_catch_parameter(k.encode(), str(p).encode(), buf, 256)
        return buf.value.decode() if n >= 0 else None

    assert get("x_samples") == "201"            # first prefix match, comment cut
    assert get("x_samp") == "201"               # prefix semantics (admin.cpp:70)
    assert get("model_file") == "../inputs/ab.bin"  # spaces removed (admin.cpp:85)
    assert get("indented") == ""                # lines starting with ' ' skipped
    assert get("missing") == ""                 # missing key -> empty string
    assert get("empty") == ""


def test_configs_write_roundtrip(tmp_path):
    from fwi3d_amd import configs
    w = configs.tiny(nshots=3)
    path = w.write(str(tmp_path / "run"))
    text = open(path).read()
    assert "x_samples = 40" in text and "batch_shots = 1" in text
    sps = np.loadtxt(str(tmp_path / "run/inputs/geometry/tiny_SPS.txt"), delimiter=",")
    assert np.allclose(sps, w.shots)
    model = np.fromfile(str(tmp_path / "run/inputs/models/tiny_true_30x40x36_10m.bin"), np.float32)
    assert np.array_equal(model, w.interior.ravel())
    # nb = (int)(length/dh) + 1 round-trips
    length = float([l for l in text.splitlines() if l.startswith("abc_length")][0].split("=")[1].split("#")[0])
    assert int(np.float32(length) / np.float32(w.dh)) + 1 == w.nb


def _f32(x):
    return float(np.float32(x))


def test_line_search_vertex_matches_oracle(host):
    """Inversion::optimization's vertex (inversion.cu:291-312) equals the
    oracle's parabola_step bit for bit, on the driver's trial steps
    (0, 0.02 dm, 0.05 dm) and random misfits."""
    rng = np.random.default_rng(17)
    vmin, vmax = np.float32(1500.0), np.float32(3500.0)
    dm = np.float32(1.0) / (vmin * vmin) - np.float32(1.0) / (vmax * vmax)
    a0, a1, a2 = np.float32(0.0) * dm, np.float32(0.02) * dm, np.float32(0.05) * dm
    for _ in range(200):
        f0, f1, f2 = (_f32(v) for v in rng.uniform(1e3, 1e5, 3))
        got = host.fwi3d_host_parabola_vertex(a0, a1, a2, f0, f1, f2)
        want = O.parabola_step(a0, a1, a2, f0, f1, f2)
        assert np.float32(got).view(np.int32) == np.float32(want).view(np.int32)


def test_line_search_guard(host):
    """The step the driver takes: the vertex when its model can run, else the
    better trial step, else none -- three collinear misfits (den = 0, a
    non-finite vertex), a vertex that drives a slowness squared below zero,
    one past the stability limit, and trial steps that fail too."""
    n = 1000
    rng = np.random.default_rng(5)
    v = rng.uniform(1500.0, 3500.0, n).astype(np.float32)
    m = (np.float32(1.0) / (v * v)).astype(np.float32)
    g = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    g[0] = np.float32(1.0)  # the cell the positivity test trips on
    vcfl = np.float32(4500.0)
    a1, a2 = _f32(2e-9), _f32(5e-9)
    fb = ctypes.c_int(-1)

    def step(vertex, f1, f2, vc=vcfl):
        s = host.fwi3d_host_usable_step(vertex, a1, a2, f1, f2, m, g, n, vc, ctypes.byref(fb))
        return s, fb.value

    # a usable vertex stands, and its model is the oracle's model_step
    vertex = _f32(3e-9)
    assert step(vertex, 10.0, 20.0) == (vertex, 0)
    mo, vo = O.model_step(m, g, vertex)
    assert np.all(mo > 0) and np.all(vo <= vcfl)
    # collinear misfits: den = 0, the vertex is not finite
    nan_vertex = host.fwi3d_host_parabola_vertex(0.0, a1, a2, 7.0, 7.0, 7.0)
    assert not np.isfinite(nan_vertex)
    assert step(nan_vertex, 7.0, 7.0) == (a2, 1)   # f1 < f2 is false: a2
    assert step(float("inf"), 5.0, 7.0) == (a1, 1)
    # a vertex past the largest slowness squared in the gradient direction
    huge = _f32(float(m.max()) * 2.0)
    assert step(huge, 5.0, 7.0) == (a1, 1)
    # a vertex that lowers a slowness squared below 1/vcfl^2 (velocity past the limit)
    past = _f32(float(m[0] - np.float32(1.0) / (vcfl * vcfl)) * 1.01)
    assert step(past, 9.0, 3.0) == (a2, 1)
    # neither the vertex nor the trial steps can run: no step
    assert step(huge, 5.0, 7.0, vc=_f32(1000.0)) == (0.0, 1)

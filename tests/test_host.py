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

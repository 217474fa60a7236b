"""The allocation-free residual (fwi_residual_noalloc) against fwi_residual
(run with -m gpu).

Both run the same kernels with the same block count; fed the context's own
element order (the time-major seismogram), they give the same residual and
the same FP64 sum bit for bit.  Two calls in a row, the second on a smaller
problem, check that nothing of the first call's partial sums survives in the
library's fixed scratch array.  Device buffers come from the HIP runtime
through ctypes; the call goes through the C ABI."""
import numpy as np
import pytest

from tests.cases import Case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fl():
    from fwi3d_amd import lib
    lib.load()
    return lib


def _context_residual(fl, case, seed):
    """fwi_residual on a recorded seismogram: (obs, cal, residual, sum)."""
    rng = np.random.default_rng(seed)
    with case.context(fl.Context) as c:
        c.forward(fl.BOUNDARY_RBC)
        cal = c.seismogram()
        obs = (cal + rng.normal(0, 0.05 * np.abs(cal).max(), cal.shape)).astype(np.float32)
        total = c.residual(obs)
        res = c.backward_data()
    return obs, cal, res, total


class _Hip:
    """The few HIP runtime calls the test needs for device buffers (through
    ctypes on the runtime the library links)."""

    def __init__(self):
        import ctypes
        self.c = ctypes
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]

    def _ok(self, rc):
        assert rc == 0, "HIP error %d" % rc

    def upload(self, a):
        p = self.c.c_void_p()
        self._ok(self.L.hipMalloc(self.c.byref(p), a.nbytes))
        self._ok(self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1))  # host to device
        return p.value

    def download(self, p, a):
        self._ok(self.L.hipDeviceSynchronize())
        self._ok(self.L.hipMemcpy(a.ctypes.data, p, a.nbytes, 2))  # device to host
        return a

    def free(self, *ps):
        for p in ps:
            self._ok(self.L.hipFree(p))


def _noalloc(fl, hip, obs, cal):
    """The allocation-free call on time-major copies (the context's order)."""
    d_obs = hip.upload(np.ascontiguousarray(obs.T))
    d_cal = hip.upload(np.ascontiguousarray(cal.T))
    d_sum = hip.upload(np.full(1, np.nan))
    fl.residual_noalloc(d_obs, d_cal, obs.size, d_sum, None)
    res = hip.download(d_cal, np.empty(cal.T.shape, np.float32)).T
    total = float(hip.download(d_sum, np.empty(1))[0])
    hip.free(d_obs, d_cal, d_sum)
    return res, total


def test_noalloc_residual_matches_fwi_residual(fl):
    hip = _Hip()
    for case, seed in ((Case(nt=120), 1), (Case(nx=24, ny=22, nz=20, nt=60, rec_stride=4), 2)):
        obs, cal, res, total = _context_residual(fl, case, seed)
        assert total > 0 and res.shape == cal.shape
        got_res, got_total = _noalloc(fl, hip, obs, cal)
        assert np.array_equal(got_res, res)
        assert got_total == total, (got_total, total)


def test_noalloc_residual_rejects_bad_sizes(fl):
    hip = _Hip()
    a = hip.upload(np.zeros(4, np.float32))
    s = hip.upload(np.zeros(1))
    try:
        for n in (0, -1, (1 << 34) + 1):  # rejected before any launch
            with pytest.raises(fl.FwiError) as e:
                fl.residual_noalloc(a, a, n, s)
            assert e.value.code == -5 and "FWI_NOALLOC_MAX_SAMPLES" in str(e.value)
        with pytest.raises(fl.FwiError) as e:
            fl.residual_noalloc(0, a, 4, s)
        assert e.value.code == -1
    finally:
        hip.free(a, s)

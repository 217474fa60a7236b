"""ctypes binding of libfwi3d_hip.so (C ABI in include/fwi3d.h).

Volumes are numpy float32 arrays in the reference host layout: numpy shape
(nyy, nxx, nzz), C order, so the flat index is i + j*nzz + k*nxx*nzz
(src/modeling/modeling.cu:190).  Seismograms are (nrec, nt) arrays, i.e. the
reference file layout t + r*nt (src/modeling/modeling.cu:392).

There is no fallback: if the HIP library is missing or a call fails, an
exception is raised.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(PKG, "libfwi3d_hip.so")

OK = 0
ERR = {-1: "FWI_ERR_ARG", -2: "FWI_ERR_HIP", -3: "FWI_ERR_CFL", -4: "FWI_ERR_STATE",
       -5: "FWI_ERR_RANGE", -6: "FWI_ERR_NONFINITE", -7: "FWI_ERR_NODEV"}
BOUNDARY_RBC, BOUNDARY_ABC = 0, 1
ARITH_REFERENCE, ARITH_FP32 = 0, 1
PHYS_GRAD_DYY, PHYS_SUBTRACT_SOURCE, PHYS_ALIGNED_IMAGING = 1, 2, 4
VOL_MODEL, VOL_FIELD, VOL_FIELD_OLD, VOL_RECV, VOL_RECV_OLD, VOL_IMAGE, VOL_ILLUM, VOL_GRAD = range(8)
KIND_FWD_RBC, KIND_FWD_ABC, KIND_BWD_RTM, KIND_BWD_FWI = range(4)

# every symbol declared in include/fwi3d.h
EXPORTS = [
    "fwi_create", "fwi_destroy", "fwi_last_error", "fwi_set_arith", "fwi_set_physics", "fwi_set_deferred_imaging", "fwi_set_model", "fwi_set_model_interior",
    "fwi_random_boundary", "fwi_set_wavelet", "fwi_set_dampers", "fwi_set_source", "fwi_set_receivers",
    "fwi_forward", "fwi_forward_steps", "fwi_get_seismogram", "fwi_set_backward_data", "fwi_rtm_backward",
    "fwi_grad_backward", "fwi_backward_steps", "fwi_zero_volume", "fwi_get_volume", "fwi_set_volume",
    "fwi_export_image", "fwi_export_gradient", "fwi_residual", "fwi_autotune", "fwi_get_tuning",
    "fwi_set_tuning", "fwi_set_graph", "fwi_last_pass_timing", "fwi_set_kernel_timing",
    "fwi_last_kernel_timing", "fwi_sync", "fwi_create_batch", "fwi_shot_random_boundary",
    "fwi_random_boundary_blobs", "fwi_shot_random_boundary_blobs", "fwi_rbc_blobs",
    "fwi_shot_set_source", "fwi_shot_set_receivers", "fwi_shot_get_seismogram", "fwi_shot_set_backward_data",
    "fwi_shot_residual", "fwi_shot_get_volume", "fwi_set_shots_per_block", "fwi_get_shots_per_block", "fwi_shot_get_backward_data",
    "fwi_copy_bandwidth", "fwi_shot_disable", "fwi_residual_noalloc",
    "fwi_set_prefetch_depth", "fwi_get_prefetch_depth", "fwi_graph_instantiations",
    "fwi_set_two_step", "fwi_get_two_step", "fwi_set_shots_per_thread", "fwi_get_shots_per_thread",
]


class FwiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERR.get(code, "FWI_ERR"), code, msg))
        self.code = code


def rbc_blobs(seed, shot, nxx, nyy, nzz, nb, dh, ratio, var_vp):
    """The blob list of the reference-style random boundary, (n, 5) float32
    rows (xc, yc, zc, r, amplitude); host-side, no device needed."""
    L = load()
    n = ctypes.c_int()
    # points at least `ratio` apart own disjoint balls of radius ratio/2
    # inside the box grown by ratio/2: at most 0.7405 of that volume (dense
    # packing), so one call with this capacity suffices
    ext = [(k - 1) * dh + ratio for k in (nxx, nyy, nzz)]
    cap = int(ext[0] * ext[1] * ext[2] * 0.7405 / (np.pi / 6.0 * ratio ** 3)) + 16
    out = np.zeros((cap, 5), np.float32)
    rc = L.fwi_rbc_blobs(seed, shot, nxx, nyy, nzz, nb, dh, ratio, var_vp, out.ctypes.data_as(ctypes.c_void_p),
                         cap, ctypes.byref(n))
    if rc != OK:
        raise FwiError(rc, "fwi_rbc_blobs")
    if n.value > cap:  # not reached (packing bound); the count is exact, so ask again
        out = np.zeros((n.value, 5), np.float32)
        rc = L.fwi_rbc_blobs(seed, shot, nxx, nyy, nzz, nb, dh, ratio, var_vp,
                             out.ctypes.data_as(ctypes.c_void_p), n.value, ctypes.byref(n))
        if rc != OK:
            raise FwiError(rc, "fwi_rbc_blobs")
    return out[:n.value].copy()


def residual_noalloc(obs_ptr, cal_ptr, n, sum_sq_ptr, stream=None):
    """fwi_residual_noalloc on device addresses (ints): cal <- obs - cal and
    *sum_sq_ptr <- sum r^2 (FP64), asynchronous on `stream` (a hipStream_t
    handle as an int, or None for the null stream); allocates nothing."""
    L = load()
    rc = L.fwi_residual_noalloc(ctypes.c_void_p(obs_ptr), ctypes.c_void_p(cal_ptr), int(n),
                                ctypes.c_void_p(sum_sq_ptr), ctypes.c_void_p(stream or 0))
    if rc != OK:
        raise FwiError(rc, L.fwi_last_error(None).decode())


class Grid(ctypes.Structure):
    _fields_ = [("nxx", ctypes.c_int), ("nyy", ctypes.c_int), ("nzz", ctypes.c_int), ("nb", ctypes.c_int),
                ("dh", ctypes.c_float), ("dt", ctypes.c_float), ("nt", ctypes.c_int), ("tlag", ctypes.c_int)]


_lib = None


def load():
    """Load libfwi3d_hip.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(HIP_LIB):
        raise FileNotFoundError("%s is missing: run `python -m fwi3d_amd.build` (no CPU fallback exists)" % HIP_LIB)
    L = ctypes.CDLL(HIP_LIB)
    P = ctypes.c_void_p
    I, F = ctypes.c_int, ctypes.c_float
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    i32 = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
    sig = {
        "fwi_create": [ctypes.POINTER(Grid), I, ctypes.POINTER(P)],
        "fwi_destroy": [P],
        "fwi_set_arith": [P, I],
        "fwi_set_physics": [P, I],
        "fwi_set_deferred_imaging": [P, I],
        "fwi_set_model": [P, f32],
        "fwi_set_model_interior": [P, f32],
        "fwi_random_boundary": [P, ctypes.c_uint64, I, F, F],
        "fwi_set_wavelet": [P, f32],
        "fwi_set_dampers": [P, f32],
        "fwi_set_source": [P, I, I, I],
        "fwi_set_receivers": [P, I, i32, i32, i32],
        "fwi_forward": [P, I],
        "fwi_forward_steps": [P, I, I, I],
        "fwi_get_seismogram": [P, f32],
        "fwi_set_backward_data": [P, f32],
        "fwi_rtm_backward": [P],
        "fwi_grad_backward": [P],
        "fwi_backward_steps": [P, I, I, I],
        "fwi_zero_volume": [P, I],
        "fwi_get_volume": [P, I, f32],
        "fwi_set_volume": [P, I, f32],
        "fwi_export_image": [P, f32],
        "fwi_export_gradient": [P, F, f32, ctypes.POINTER(F)],
        "fwi_residual": [P, f32, ctypes.POINTER(ctypes.c_double)],
        "fwi_autotune": [P, I, I, ctypes.c_char_p],
        "fwi_get_tuning": [P, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)],
        "fwi_set_tuning": [P, I, I, I, I],
        "fwi_set_graph": [P, I],
        "fwi_last_pass_timing": [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I)],
        "fwi_set_kernel_timing": [P, I],
        "fwi_last_kernel_timing": [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(I)],
        "fwi_sync": [P],
        "fwi_create_batch": [ctypes.POINTER(Grid), I, I, ctypes.POINTER(P)],
        "fwi_shot_random_boundary": [P, I, ctypes.c_uint64, I, F, F],
        "fwi_random_boundary_blobs": [P, ctypes.c_uint64, I, F, F, F],
        "fwi_shot_random_boundary_blobs": [P, I, ctypes.c_uint64, I, F, F, F],
        "fwi_rbc_blobs": [ctypes.c_uint64, I, I, I, I, I, F, F, F, ctypes.c_void_p, I, ctypes.POINTER(I)],
        "fwi_shot_set_source": [P, I, I, I, I],
        "fwi_shot_set_receivers": [P, I, I, i32, i32, i32],
        "fwi_shot_get_seismogram": [P, I, f32],
        "fwi_shot_set_backward_data": [P, I, f32],
        "fwi_shot_residual": [P, I, f32, ctypes.POINTER(ctypes.c_double)],
        "fwi_shot_get_volume": [P, I, I, f32],
        "fwi_set_shots_per_block": [P, I, I],
        "fwi_get_shots_per_block": [P, I, ctypes.POINTER(I)],
        "fwi_shot_get_backward_data": [P, I, f32],
        "fwi_copy_bandwidth": [P, ctypes.c_long, I, ctypes.POINTER(ctypes.c_double)],
        "fwi_shot_disable": [P, I],
        "fwi_residual_noalloc": [P, P, ctypes.c_long, P, P],
        "fwi_set_prefetch_depth": [P, I, I],
        "fwi_get_prefetch_depth": [P, I, ctypes.POINTER(I)],
        "fwi_set_two_step": [P, I, I, I, I, I],
        "fwi_get_two_step": [P, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)],
        "fwi_graph_instantiations": [P, ctypes.POINTER(I)],
        "fwi_set_shots_per_thread": [P, I, I, I, I],
        "fwi_get_shots_per_thread": [P, I, ctypes.POINTER(I)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = I
    L.fwi_last_error.argtypes = [P]
    L.fwi_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


class Context:
    """One padded grid on one GPU: the state the reference keeps in its
    Modeling / Migration / Inversion objects (modeling.cuh:18-90)."""

    def __init__(self, nxx, nyy, nzz, nb, dh, dt, nt, tlag, device=0, nshots=1):
        self.L = load()
        self.grid = Grid(nxx, nyy, nzz, nb, dh, dt, nt, tlag)
        self.shape = (nyy, nxx, nzz)
        self.nb, self.nt, self.tlag, self.dh, self.dt = nb, nt, tlag, dh, dt
        self.nshots = nshots
        self.nrec_slot = [0] * nshots
        h = ctypes.c_void_p()
        rc = self.L.fwi_create_batch(ctypes.byref(self.grid), nshots, device, ctypes.byref(h))
        if rc != OK:
            raise FwiError(rc, self.L.fwi_last_error(None).decode())
        self.h = h

    # -- plumbing ---------------------------------------------------------
    def _chk(self, rc):
        if rc != OK:
            raise FwiError(rc, self.L.fwi_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.L.fwi_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _vol(self, a):
        a = np.ascontiguousarray(a, np.float32)
        if a.shape != self.shape:
            raise ValueError("volume shape %s != %s" % (a.shape, self.shape))
        return a

    # -- set-up ------------------------------------------------------------
    def set_arith(self, arith):
        self._chk(self.L.fwi_set_arith(self.h, arith))

    def set_physics(self, flags):
        """PHYS_* bits of the backward passes (0 = the reference's behaviour)."""
        self._chk(self.L.fwi_set_physics(self.h, flags))

    def set_deferred_imaging(self, enable):
        """Image updated every other backward step and illumination every
        third (every other step with the corrections; default on,
        bit-identical results) or both every step."""
        self._chk(self.L.fwi_set_deferred_imaging(self.h, 1 if enable else 0))

    def set_model(self, vp_pad):
        self._chk(self.L.fwi_set_model(self.h, self._vol(vp_pad)))

    def set_model_interior(self, vp):
        nb = self.nb
        shp = (self.shape[0] - 2 * nb, self.shape[1] - 2 * nb, self.shape[2] - 2 * nb)
        vp = np.ascontiguousarray(vp, np.float32)
        if vp.shape != shp:
            raise ValueError("interior shape %s != %s" % (vp.shape, shp))
        self._chk(self.L.fwi_set_model_interior(self.h, vp))

    def random_boundary(self, seed, shot, amp, vclamp):
        self._chk(self.L.fwi_random_boundary(self.h, seed, shot, amp, vclamp))

    def set_wavelet(self, w):
        w = np.ascontiguousarray(w, np.float32)
        if w.size != self.nt:
            raise ValueError("wavelet needs nt samples")
        self._chk(self.L.fwi_set_wavelet(self.h, w))

    def set_dampers(self, b1d):
        b1d = np.ascontiguousarray(b1d, np.float32)
        if b1d.size != self.nb:
            raise ValueError("b1d needs nb samples")
        self._chk(self.L.fwi_set_dampers(self.h, b1d))

    @property
    def nrec(self):
        return self.nrec_slot[0]

    def set_source(self, sx, sy, sz, slot=0):
        self._chk(self.L.fwi_shot_set_source(self.h, slot, int(sx), int(sy), int(sz)))

    def disable_slot(self, slot):
        """No shot in this slot: no source, no receivers, nothing added to the
        backward pass's sums."""
        self._chk(self.L.fwi_shot_disable(self.h, slot))
        self.nrec_slot[slot] = 0

    def set_receivers(self, rec, slot=0):
        rec = np.asarray(rec, np.int32).reshape(-1, 3)
        rx, ry, rz = (np.ascontiguousarray(rec[:, q]) for q in range(3))
        self._chk(self.L.fwi_shot_set_receivers(self.h, slot, len(rec), rx, ry, rz))
        self.nrec_slot[slot] = len(rec)

    def shot_random_boundary(self, slot, seed, shot, amp, vclamp):
        self._chk(self.L.fwi_shot_random_boundary(self.h, slot, seed, shot, amp, vclamp))

    def random_boundary_blobs(self, seed, shot, ratio, var_vp, vclamp, slot=0):
        """Reference-style random boundary: background noise plus Gaussian
        blobs at Poisson-disk points (ratio in metres, var_vp in m/s)."""
        self._chk(self.L.fwi_shot_random_boundary_blobs(self.h, slot, seed, shot, ratio, var_vp, vclamp))

    def set_shots_per_block(self, kind, sb):
        self._chk(self.L.fwi_set_shots_per_block(self.h, kind, sb))

    def shots_per_block(self, kind):
        sb = ctypes.c_int()
        self._chk(self.L.fwi_get_shots_per_block(self.h, kind, ctypes.byref(sb)))
        return sb.value

    def set_shots_per_thread(self, kind, spt, by=16, zchunks=1):
        """spt = 2: shot pairs (two consecutive slots per thread, 64 x by tiles)."""
        self._chk(self.L.fwi_set_shots_per_thread(self.h, kind, spt, by, zchunks))

    def shots_per_thread(self, kind):
        spt = ctypes.c_int()
        self._chk(self.L.fwi_get_shots_per_thread(self.h, kind, ctypes.byref(spt)))
        return spt.value

    # -- passes -------------------------------------------------------------
    def forward(self, boundary=BOUNDARY_RBC):
        self._chk(self.L.fwi_forward(self.h, boundary))

    def forward_steps(self, boundary, n0, n1):
        self._chk(self.L.fwi_forward_steps(self.h, boundary, n0, n1))

    def seismogram(self, slot=0):
        out = np.zeros((self.nrec_slot[slot], self.nt), np.float32)
        self._chk(self.L.fwi_shot_get_seismogram(self.h, slot, out))
        return out

    def set_backward_data(self, data, slot=0):
        data = np.ascontiguousarray(data, np.float32)
        if data.shape != (self.nrec_slot[slot], self.nt):
            raise ValueError("data must be (nrec, nt)")
        self._chk(self.L.fwi_shot_set_backward_data(self.h, slot, data))

    def backward_data(self, slot=0):
        out = np.zeros((self.nrec_slot[slot], self.nt), np.float32)
        self._chk(self.L.fwi_shot_get_backward_data(self.h, slot, out))
        return out

    def rtm_backward(self):
        self._chk(self.L.fwi_rtm_backward(self.h))

    def grad_backward(self):
        self._chk(self.L.fwi_grad_backward(self.h))

    def backward_steps(self, kind, b0, b1):
        self._chk(self.L.fwi_backward_steps(self.h, kind, b0, b1))

    # -- volumes -------------------------------------------------------------
    def get_volume(self, which, slot=0):
        out = np.zeros(self.shape, np.float32)
        self._chk(self.L.fwi_shot_get_volume(self.h, slot, which, out))
        return out

    def set_volume(self, which, vol):
        self._chk(self.L.fwi_set_volume(self.h, which, self._vol(vol)))

    def zero_volume(self, which):
        self._chk(self.L.fwi_zero_volume(self.h, which))

    def _interior_shape(self):
        nb = self.nb
        return (self.shape[0] - 2 * nb, self.shape[1] - 2 * nb, self.shape[2] - 2 * nb)

    def export_image(self):
        out = np.zeros(self._interior_shape(), np.float32)
        self._chk(self.L.fwi_export_image(self.h, out))
        return out

    def export_gradient(self, zmask):
        out = np.zeros(self._interior_shape(), np.float32)
        g = ctypes.c_float(0.0)
        self._chk(self.L.fwi_export_gradient(self.h, zmask, out, ctypes.byref(g)))
        return out, g.value

    def residual(self, obs, slot=0):
        obs = np.ascontiguousarray(obs, np.float32)
        if obs.shape != (self.nrec_slot[slot], self.nt):
            raise ValueError("obs must be (nrec, nt)")
        s = ctypes.c_double(0.0)
        self._chk(self.L.fwi_shot_residual(self.h, slot, obs, ctypes.byref(s)))
        return s.value

    # -- tuning / timing ------------------------------------------------------
    def autotune(self, kind, steps=4, cache=None):
        self._chk(self.L.fwi_autotune(self.h, kind, steps, cache.encode() if cache else None))

    def tuning(self, kind):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._chk(self.L.fwi_get_tuning(self.h, kind, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def set_tuning(self, kind, by, rpt, zchunks):
        self._chk(self.L.fwi_set_tuning(self.h, kind, by, rpt, zchunks))

    def set_prefetch_depth(self, kind, pd):
        self._chk(self.L.fwi_set_prefetch_depth(self.h, kind, pd))

    def prefetch_depth(self, kind):
        v = ctypes.c_int()
        self._chk(self.L.fwi_get_prefetch_depth(self.h, kind, ctypes.byref(v)))
        return v.value

    def set_two_step(self, kind, enable=True, by=16, rpt=2, zchunks=1):
        """Two time steps per launch for kind 0 (random-boundary forward) or
        2 (RTM backward)."""
        self._chk(self.L.fwi_set_two_step(self.h, kind, int(enable), by, rpt, zchunks))

    def two_step(self, kind):
        """(active, by, rpt, zchunks); active 1 = used by the next pass,
        -1 = requested but not available, 0 = off."""
        v = [ctypes.c_int() for _ in range(4)]
        self._chk(self.L.fwi_get_two_step(self.h, kind, *[ctypes.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def graph_instantiations(self):
        v = ctypes.c_int()
        self._chk(self.L.fwi_graph_instantiations(self.h, ctypes.byref(v)))
        return v.value

    def set_graph(self, enable):
        self._chk(self.L.fwi_set_graph(self.h, 1 if enable else 0))

    def last_pass_timing(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        self._chk(self.L.fwi_last_pass_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def set_kernel_timing(self, enable):
        self._chk(self.L.fwi_set_kernel_timing(self.h, 1 if enable else 0))

    def last_kernel_timing(self):
        ms, n = ctypes.c_double(), ctypes.c_int()
        self._chk(self.L.fwi_last_kernel_timing(self.h, ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def copy_bandwidth(self, nbytes=1 << 30, reps=10):
        g = ctypes.c_double()
        self._chk(self.L.fwi_copy_bandwidth(self.h, nbytes, reps, ctypes.byref(g)))
        return g.value

    def sync(self):
        self._chk(self.L.fwi_sync(self.h))

"""ctypes front end of the CPU oracle (oracle/fwi3d_oracle.cpp).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker; the product never calls it.
All volumes are numpy arrays in the reference host layout, i.e. numpy shape
(nyy, nxx, nzz) in C order so that the flat index is i + j*nzz + k*nxx*nzz
(src/modeling/modeling.cu:190).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfwi3d_oracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_c_int, _c_float, _c_long = ctypes.c_int, ctypes.c_float, ctypes.c_long


def build(force=False):
    """Compile the oracle with its Makefile (g++ -ffp-contract=off -fopenmp)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE] + (["-B"] if force else []), check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = ctypes.CDLL(_LIB_PATH)
    vp = ctypes.c_void_p
    L.oracle_wavelet.argtypes = [_c_int, _c_float, _c_float, _f32p]
    L.oracle_wavelet.restype = _c_int
    L.oracle_dampers.argtypes = [_c_int, _c_float, _f32p, _f32p, _f32p]
    L.oracle_damper_at.argtypes = [_f32p, _f32p, _f32p] + [_c_int] * 7
    L.oracle_damper_at.restype = _c_float
    L.oracle_expand_boundary.argtypes = [_f32p, _f32p] + [_c_int] * 4
    L.oracle_reduce_boundary.argtypes = [_f32p, _f32p] + [_c_int] * 4
    L.oracle_position_index.argtypes = [_c_float, _c_float, _c_int]
    L.oracle_position_index.restype = _c_int
    L.oracle_rbc_uniform.argtypes = [ctypes.c_uint64, _c_int, _c_long]
    L.oracle_rbc_uniform.restype = _c_float
    L.oracle_random_boundary.argtypes = [_f32p] + [_c_int] * 4 + [ctypes.c_uint64, _c_int, _c_float, _c_float]
    L.oracle_random_boundary_blobs.argtypes = [_f32p] + [_c_int] * 4 + [_c_float, ctypes.c_uint64, _c_int,
                                                                        _c_float, _f32p, _c_int] + [_c_float] * 3
    L.oracle_exp_det.argtypes = [ctypes.c_double]
    L.oracle_exp_det.restype = ctypes.c_double
    L.oracle_irwin_hall.argtypes = [ctypes.c_uint64, _c_int, _c_long, _c_int]
    L.oracle_irwin_hall.restype = _c_float
    fwd = [_f32p, _c_int, _c_int, _c_int, _c_int, _c_float, _c_float, _c_int, _c_int, _f32p,
           _c_int, _c_int, _c_int, _c_int, _i32p, _i32p, _i32p, _c_int, vp, vp, vp, _c_int, _c_int]
    L.oracle_forward_f32.argtypes = fwd + [_f32p, _f32p, vp]
    L.oracle_forward_f64.argtypes = fwd + [_f64p, _f64p, vp]
    bwd = [_f32p, _c_int, _c_int, _c_int, _c_float, _c_float, _c_int, _f32p, _c_int, _i32p, _i32p, _i32p,
           _c_int, _c_int, _c_int]
    L.oracle_backward_f32.argtypes = bwd + [_f32p] * 6
    L.oracle_backward_f64.argtypes = bwd + [_f64p] * 6
    phys = [_c_int, _f32p, _c_int, _c_int, _c_int, _c_int]
    L.oracle_backward_phys_f32.argtypes = bwd + [_f32p] * 6 + phys
    L.oracle_backward_phys_f64.argtypes = bwd + [_f64p] * 6 + phys
    pp = ctypes.POINTER(ctypes.c_void_p)
    L.oracle_backward_batched_f32.argtypes = [pp, _c_int, _c_int, _c_int, _c_int, _c_int, _c_float, _c_float, _c_int,
                                              pp, _i32p, pp, pp, pp, _c_int, _c_int, _c_int, pp, pp, pp, pp, pp, pp,
                                              _c_int, _f32p, _i32p, _c_int]
    L.oracle_export_seismic.argtypes = [_f32p, _f32p] + [_c_int] * 4 + [_c_float, _f32p]
    L.oracle_gradient_post.argtypes = [_f32p, _f32p] + [_c_int] * 4 + [_c_float, _c_float, _f32p]
    L.oracle_gradient_post.restype = _c_float
    L.oracle_misfit.argtypes = [_f32p, _f32p, _c_long, vp, vp]
    L.oracle_misfit.restype = ctypes.c_double
    L.oracle_parabola_step.argtypes = [_c_float] * 6
    L.oracle_parabola_step.restype = _c_float
    L.oracle_model_step.argtypes = [_f32p, _f32p, _c_float, _c_long, vp, _f32p]
    L.oracle_num_threads.restype = _c_int
    L.oracle_laplacian_f32.argtypes = [_f32p, _c_int, _c_int, _c_int, _c_float, _f32p]
    L.oracle_laplacian_f64.argtypes = [_f64p, _c_int, _c_int, _c_int, ctypes.c_double, _f64p]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def num_threads():
    return lib().oracle_num_threads()


def laplacian(p, dh, fp64=False):
    """Dxx + Dyy + Dzz of the 8th-order stencil over the updated interior."""
    dtype = np.float64 if fp64 else np.float32
    p = np.ascontiguousarray(p, dtype)
    nyy, nxx, nzz = p.shape
    out = np.zeros_like(p)
    (lib().oracle_laplacian_f64 if fp64 else lib().oracle_laplacian_f32)(p, nxx, nyy, nzz, dh, out)
    return out


def wavelet(nt, dt, fmax):
    """Ricker wavelet and tlag (modeling.cu:36-59)."""
    w = np.zeros(nt, np.float32)
    tlag = lib().oracle_wavelet(nt, dt, fmax, w)
    return w, tlag


def dampers(nb, factor):
    """Cerjan tables b1d, b2d (nb x nb), b3d (nb^3) (modeling.cu:83-140)."""
    b1 = np.zeros(nb, np.float32)
    b2 = np.zeros(nb * nb, np.float32)
    b3 = np.zeros(nb * nb * nb, np.float32)
    lib().oracle_dampers(nb, factor, b1, b2, b3)
    return b1, b2, b3


def damper_at(tables, i, j, k, nxx, nyy, nzz, nb):
    b1, b2, b3 = tables
    return lib().oracle_damper_at(b1, b2, b3, i, j, k, nxx, nyy, nzz, nb)


def expand_boundary(vol, nb):
    """Edge-replicating padding (modeling.cu:181-228). vol shape (ny, nx, nz)."""
    ny, nx, nz = vol.shape
    out = np.zeros((ny + 2 * nb, nx + 2 * nb, nz + 2 * nb), np.float32)
    lib().oracle_expand_boundary(np.ascontiguousarray(vol, np.float32), out, nx, ny, nz, nb)
    return out


def reduce_boundary(vol, nb):
    """Crop to the interior (modeling.cu:230-241)."""
    nyy, nxx, nzz = vol.shape
    nx, ny, nz = nxx - 2 * nb, nyy - 2 * nb, nzz - 2 * nb
    out = np.zeros((ny, nx, nz), np.float32)
    lib().oracle_reduce_boundary(np.ascontiguousarray(vol, np.float32), out, nx, ny, nz, nb)
    return out


def position_index(x, dh, nb):
    """Metres to padded grid index (modeling.cu:270-272)."""
    return lib().oracle_position_index(x, dh, nb)


def rbc_uniform(seed, shot, cell):
    return lib().oracle_rbc_uniform(seed, shot, cell)


def random_boundary(vp_pad, nb, seed, shot, amp, vclamp):
    """Seeded random boundary over an edge-padded model (returns a copy)."""
    out = np.array(vp_pad, np.float32, copy=True, order="C")
    nyy, nxx, nzz = out.shape
    lib().oracle_random_boundary(out, nxx, nyy, nzz, nb, seed, shot, amp, vclamp)
    return out


def random_boundary_blobs(vp_pad, nb, dh, seed, shot, var_vp, blobs, vlo, vhi, vclamp):
    """Reference-style random boundary (background noise + blobs, see the C
    source) over an edge-padded model; blobs is an (n, 5) float32 array of
    (xc, yc, zc, r, amplitude).  Returns a copy."""
    out = np.array(vp_pad, np.float32, copy=True, order="C")
    nyy, nxx, nzz = out.shape
    b = np.ascontiguousarray(np.asarray(blobs, np.float32).reshape(-1, 5))
    lib().oracle_random_boundary_blobs(out, nxx, nyy, nzz, nb, dh, seed, shot, var_vp, b, len(b), vlo, vhi, vclamp)
    return out


def exp_det(x):
    return lib().oracle_exp_det(float(x))


def irwin_hall(seed, shot, cell, axis):
    return lib().oracle_irwin_hall(seed, shot, cell, axis)


def _rec_arrays(rec):
    rec = np.asarray(rec, np.int32).reshape(-1, 3)
    return (np.ascontiguousarray(rec[:, 0]), np.ascontiguousarray(rec[:, 1]),
            np.ascontiguousarray(rec[:, 2]), len(rec))


def forward(vp, nb, dh, dt, nt, tlag, wavelet_, src, rec, abc_tables=None, n_begin=0, n_end=None,
            P=None, Pold=None, record=True, fp64=False):
    """Steps [n_begin, n_end) of Modeling::forward_solver (modeling.cu:309-322).

    src = (sIdx, sIdy, sIdz); rec = int array (nrec, 3) of (rIdx, rIdy, rIdz).
    Returns (P, Pold, seis) with seis of shape (nrec, nt) (file layout t + r*nt).
    """
    vp = np.ascontiguousarray(vp, np.float32)
    nyy, nxx, nzz = vp.shape
    n_end = nt + tlag if n_end is None else n_end
    dtype = np.float64 if fp64 else np.float32
    P = np.zeros(vp.shape, dtype) if P is None else np.array(P, dtype, copy=True, order="C")
    Pold = np.zeros(vp.shape, dtype) if Pold is None else np.array(Pold, dtype, copy=True, order="C")
    rx, ry, rz, nrec = _rec_arrays(rec)
    seis = np.zeros((nrec, nt), dtype) if record else None
    b1, b2, b3 = abc_tables if abc_tables is not None else (None, None, None)
    fn = lib().oracle_forward_f64 if fp64 else lib().oracle_forward_f32
    fn(vp, nxx, nyy, nzz, nb, dh, dt, nt, tlag, np.ascontiguousarray(wavelet_, np.float32),
       int(src[0]), int(src[1]), int(src[2]), nrec, rx, ry, rz, 1 if abc_tables is not None else 0,
       _ptr(b1), _ptr(b2), _ptr(b3), n_begin, n_end, P, Pold, _ptr(seis))
    return P, Pold, seis


def backward(vp, dh, dt, nt, data, rec, S, So, R=None, Ro=None, image=None, illum=None, fwi=False,
             b_begin=0, b_end=None, tlag=None, fp64=False, phys=0, wavelet_=None, src=None):
    """Steps [b_begin, b_end) of the RTM (migration.cu:61-81) or FWI
    (inversion.cu:235-253) backward pass.  data has shape (nrec, nt).
    phys: correction bits (1 gradient with d2/dy2, 2 source subtraction --
    needs wavelet_ and src -- 4 aligned imaging); 0 is the reference.
    Returns (S, So, R, Ro, image, illum)."""
    vp = np.ascontiguousarray(vp, np.float32)
    nyy, nxx, nzz = vp.shape
    if b_end is None:
        b_end = nt + tlag
    dtype = np.float64 if fp64 else np.float32

    def arr(a):
        return np.zeros(vp.shape, dtype) if a is None else np.array(a, dtype, copy=True, order="C")

    S, So, R, Ro, image, illum = arr(S), arr(So), arr(R), arr(Ro), arr(image), arr(illum)
    rx, ry, rz, nrec = _rec_arrays(rec)
    args = (vp, nxx, nyy, nzz, dh, dt, nt, np.ascontiguousarray(data, np.float32), nrec, rx, ry, rz,
            1 if fwi else 0, b_begin, b_end, S, So, R, Ro, image, illum)
    if phys:
        if (phys & 2) and (wavelet_ is None or src is None or tlag is None):
            raise ValueError("source subtraction needs wavelet_, src and tlag")
        w = np.ascontiguousarray(wavelet_ if wavelet_ is not None else np.zeros(nt), np.float32)
        sx, sy, sz = src if src is not None else (-1, -1, -1)
        fn = lib().oracle_backward_phys_f64 if fp64 else lib().oracle_backward_phys_f32
        fn(*args, int(phys), w, int(sx), int(sy), int(sz), int(nt + (tlag or 0)))
    else:
        fn = lib().oracle_backward_f64 if fp64 else lib().oracle_backward_f32
        fn(*args)
    return S, So, R, Ro, image, illum


def _ptrs(arrays):
    return (ctypes.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])


def backward_batched(vps, sb, dh, dt, nt, datas, recs, states, images, illums, fwi=False, b_begin=0, b_end=None,
                     tlag=0, phys=0, wavelet_=None, srcs=None):
    """Backward steps [b_begin, b_end) of a batch of shots in the summation
    order of a batched pass with sb shots per block: every step advances the
    shots in slot order and shot s adds its terms into group s // sb.
    states[s] = [S, So, R, Ro] float32 arrays, updated in place; images /
    illums: one float32 array per group, updated in place."""
    S = len(vps)
    vps = [np.ascontiguousarray(v, np.float32) for v in vps]
    nyy, nxx, nzz = vps[0].shape
    b_end = nt + tlag if b_end is None else b_end
    datas = [np.ascontiguousarray(d, np.float32) for d in datas]
    recs = [np.asarray(r, np.int32).reshape(-1, 3) for r in recs]
    rx = [np.ascontiguousarray(r[:, 0]) for r in recs]
    ry = [np.ascontiguousarray(r[:, 1]) for r in recs]
    rz = [np.ascontiguousarray(r[:, 2]) for r in recs]
    for st in states:
        for a in st:
            assert a.dtype == np.float32 and a.flags.c_contiguous and a.shape == vps[0].shape
    for a in list(images) + list(illums):
        assert a.dtype == np.float32 and a.flags.c_contiguous and a.shape == vps[0].shape
    assert len(images) == len(illums) == (S + sb - 1) // sb
    nrecs = np.array([len(r) for r in recs], np.int32)
    src = np.full(3 * S, -1, np.int32) if srcs is None else np.ascontiguousarray(np.asarray(srcs, np.int32).ravel())
    w = np.ascontiguousarray(wavelet_ if wavelet_ is not None else np.zeros(nt), np.float32)
    lib().oracle_backward_batched_f32(_ptrs(vps), S, sb, nxx, nyy, nzz, dh, dt, nt, _ptrs(datas), nrecs,
                                      _ptrs(rx), _ptrs(ry), _ptrs(rz), 1 if fwi else 0, b_begin, b_end,
                                      _ptrs([st[0] for st in states]), _ptrs([st[1] for st in states]),
                                      _ptrs([st[2] for st in states]), _ptrs([st[3] for st in states]),
                                      _ptrs(list(images)), _ptrs(list(illums)), int(phys), w, src, int(nt + tlag))


def export_seismic(image_pad, illum_pad, nb, dh):
    """Migration::export_seismic arithmetic (migration.cu:90-117)."""
    nyy, nxx, nzz = image_pad.shape
    nx, ny, nz = nxx - 2 * nb, nyy - 2 * nb, nzz - 2 * nb
    out = np.zeros((ny, nx, nz), np.float32)
    lib().oracle_export_seismic(np.ascontiguousarray(image_pad, np.float32),
                                np.ascontiguousarray(illum_pad, np.float32), nx, ny, nz, nb, dh, out)
    return out


def gradient_post(grad_pad, illum_pad, nb, dh, zmask):
    """Inversion::compute_gradient post-processing (inversion.cu:142-169)."""
    nyy, nxx, nzz = grad_pad.shape
    nx, ny, nz = nxx - 2 * nb, nyy - 2 * nb, nzz - 2 * nb
    out = np.zeros((ny, nx, nz), np.float32)
    gmax = lib().oracle_gradient_post(np.ascontiguousarray(grad_pad, np.float32),
                                      np.ascontiguousarray(illum_pad, np.float32), nx, ny, nz, nb, dh, zmask, out)
    return out, gmax


def misfit(obs, cal):
    """Residual and sum of squares (inversion.cu:190-201): returns
    (residual, sum in long double, sum in the reference's sequential float)."""
    obs = np.ascontiguousarray(obs, np.float32).ravel()
    cal = np.ascontiguousarray(cal, np.float32).ravel()
    res = np.zeros_like(obs)
    f = ctypes.c_float(0.0)
    s = lib().oracle_misfit(obs, cal, obs.size, _ptr(res), ctypes.byref(f))
    return res, s, f.value


def parabola_step(a0, a1, a2, f0, f1, f2):
    """Parabolic line-search vertex (inversion.cu:291-312)."""
    return lib().oracle_parabola_step(a0, a1, a2, f0, f1, f2)


def model_step(m, g, alpha):
    """m - alpha*g and 1/sqrt of it (inversion.cu:317-321, 351-356)."""
    m = np.ascontiguousarray(m, np.float32).ravel()
    g = np.ascontiguousarray(g, np.float32).ravel()
    mo = np.zeros_like(m)
    vo = np.zeros_like(m)
    lib().oracle_model_step(m, g, alpha, m.size, _ptr(mo), vo)
    return mo, vo

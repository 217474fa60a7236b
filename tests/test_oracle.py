"""Pins the CPU oracle against closed forms, invariants and brute force.

The reference ships no tests, golden vectors or fixtures, so each oracle
function is checked against what the mathematics fixes:

  function (oracle)          pinned by
  -------------------------  ---------------------------------------------------
  laplacian / stencil        exact on polynomials of degree <= 9 (FP64);
                             coefficient sum = 0; symbol at Nyquist -> CFL 0.452856
  forward step               homogeneous-medium first arrival at d/v; source-
                             receiver reciprocity; leapfrog reversibility;
                             linearity in the wavelet
  backward (RTM/FWI)         reconstruction recovers the forward field; R equals
                             a forward run driven by the reversed data; zero data
                             gives a zero image; image/illum linear/quadratic in
                             scale; FWI gradient term vanishes for a field
                             without x/z curvature
  backward corrections       source subtraction makes the reconstruction exact
                             through the source window; the gradient with d2/dy2
                             is x<->y symmetric for a symmetric experiment;
                             aligned imaging equals a step-by-step replay of
                             So * R and So^2
  wavelet                    Ricker closed form (FP64) and tlag = 79 / 74
  dampers                    b1 = exp(-(f(nb-i))^2); b2 = b1a+b1b-1; b3 ~ 2*sum-5;
                             lookup symmetric under face mirroring
  expand / reduce boundary   numpy edge padding; reduce o expand = identity
  position_index             (int)((x + dh/2)/dh) + nb brute force
  export_seismic             I quadratic in z -> constant -2a; first/last row 0
  gradient_post              same with the depth mask; max |g| = 1 after scaling
  misfit                     math.fsum
  parabola_step              vertex of an exact parabola
  model_step                 1/sqrt(m - a g) in FP64, rounded
  random boundary            interior untouched; |v/v_face - 1| <= amp*ramp;
                             clamp; seeded determinism; U[-1,1) moments
  backward_batched           one shot = the single-shot backward; one shot per
                             group = per-shot backward passes summed in group
                             order; sb > 1 = the step-by-step interleave of
                             single-step backward calls (the batched order)
  blob random boundary       exp_det = math.exp to 1e-15; Irwin-Hall normal
                             moments and bounds; one blob on a constant model
                             = v0 + A e^(-q/2) cut at 4 radii (closed form);
                             interior untouched
"""
import math

import numpy as np
import pytest

import oracle as O

FDM = [2.847222222222, 1.600000000000, 0.200000000000, 0.025396825396, 0.001785714285]


def test_stencil_exact_on_polynomials():
    n = 24
    i = np.arange(n, dtype=np.float64)
    K, J, I = np.meshgrid(i, i, i, indexing="ij")  # k=y, j=x, i=z
    x, y, z = (J - 11.3) / 7.0, (K - 12.1) / 7.0, (I - 10.7) / 7.0
    dh = 1.0 / 7.0
    inner = (slice(4, -4),) * 3
    for deg in range(2, 10):
        p = x ** deg + 0.5 * y ** deg - 0.25 * z ** (deg - 1)
        want = deg * (deg - 1) * x ** (deg - 2) + 0.5 * deg * (deg - 1) * y ** (deg - 2) \
            - 0.25 * (deg - 1) * (deg - 2) * z ** max(deg - 3, 0)
        got = O.laplacian(p, dh, fp64=True)
        scale = np.abs(want[inner]).max() + 1.0
        assert np.abs(got[inner] - want[inner]).max() <= 1e-8 * scale * 10 ** (deg / 3), deg
    # outside the updated interior the Laplacian is not evaluated
    assert np.all(got[:4] == 0) and np.all(got[:, :, -4:] == 0)


def test_coefficients_sum_and_cfl():
    c5, c4, c3, c2, c1 = FDM
    assert abs(2 * (c4 - c3 + c2 - c1) - c5) < 1e-12
    nyq = -c5 + 2 * (-c4 - c3 - c2 - c1)
    assert abs(nyq + 6.501587) < 1e-5
    assert abs(math.sqrt(4.0 / (3 * abs(nyq))) - 0.452856) < 1e-6


def _homog(n=60, nb=6, v=2000.0):
    return np.full((n, n, n), v, np.float32), nb


def test_first_arrival_homogeneous():
    vp, nb = _homog()
    dh, dt, nt, fmax = 10.0, 1e-3, 260, 45.0
    w, tlag = O.wavelet(nt, dt, fmax)
    src = (30, 30, 30)
    rec = np.array([[30 + d, 30, 30] for d in (10, 20)], np.int32)
    _, _, seis = O.forward(vp, nb, dh, dt, nt, tlag, w, src, rec)
    for r, d in enumerate((10, 20)):
        # the Ricker peak sits at sample n - tlag ~ t0/dt after injection; the
        # trace peak arrives d*dh/v later (positive lobe of the Green's function)
        t_peak = np.argmax(np.abs(seis[r])) * dt
        expect = d * dh / 2000.0
        assert abs(t_peak - expect) < 0.012, (r, t_peak, expect)


def test_reciprocity():
    rng = np.random.default_rng(0)
    vp = (2000 + rng.uniform(-200, 200, (40, 44, 36))).astype(np.float32)
    w, tlag = O.wavelet(150, 1e-3, 45.0)
    a, b = (12, 14, 10), (30, 25, 22)
    _, _, sab = O.forward(vp, 6, 10.0, 1e-3, 150, tlag, w, a, np.array([b], np.int32))
    _, _, sba = O.forward(vp, 6, 10.0, 1e-3, 150, tlag, w, b, np.array([a], np.int32))
    # the constant-density operator v^2 L is not symmetric, v^-2 weighting makes it so
    va = vp[a[1], a[0], a[2]] ** 2
    vb = vp[b[1], b[0], b[2]] ** 2
    assert np.allclose(sab[0] / vb, sba[0] / va, rtol=2e-3, atol=1e-6 * np.abs(sab).max() / vb)


def test_linear_in_wavelet():
    vp = np.full((36, 36, 36), 1800.0, np.float32)
    w, tlag = O.wavelet(80, 1e-3, 45.0)
    rec = np.array([[20, 18, 18]], np.int32)
    _, _, s1 = O.forward(vp, 6, 10.0, 1e-3, 80, tlag, w, (18, 18, 18), rec, fp64=True)
    _, _, s2 = O.forward(vp, 6, 10.0, 1e-3, 80, tlag, 2 * w, (18, 18, 18), rec, fp64=True)
    assert np.allclose(s2, 2 * s1, rtol=1e-12, atol=0)


def test_reconstruction_recovers_forward():
    """Leapfrog is time reversible: the backward pass rebuilds u_{N-2-b}."""
    rng = np.random.default_rng(2)
    vp = (2200 + rng.uniform(-300, 300, (36, 40, 32))).astype(np.float32)
    nt, dt, dh = 60, 1e-3, 10.0
    w, tlag = O.wavelet(nt, dt, 45.0)
    N = nt + tlag
    src, rec = (20, 18, 16), np.array([[10, 10, 8]], np.int32)
    # snapshot u_{N-1-m} by running the forward to step N-m
    m = 30
    P_m, _, _ = O.forward(vp, 6, dh, dt, nt, tlag, w, src, rec, n_end=N - m, fp64=True)
    P, Pold, _ = O.forward(vp, 6, dh, dt, nt, tlag, w, src, rec, fp64=True)
    data = np.zeros((1, nt), np.float32)
    # no source during the last N-nt steps, so the reconstruction is exact in FP64
    S, So, *_ = O.backward(vp, dh, dt, nt, data, rec, P, Pold, b_begin=0, b_end=m, tlag=tlag, fp64=True)
    # (S, So) starts at (u_N, u_{N-1}) and steps to (u_{N-b}, u_{N-1-b}); P_m = u_{N-m}
    assert np.abs(S - P_m).max() <= 1e-9 * np.abs(P_m).max()


def test_source_subtraction_makes_reconstruction_exact():
    """With the source term put back (correction bit 2) the reconstruction
    reverses the forward recurrence through the source window too: So after
    m backward steps is the forward's previous field of step N-1-m.  Without
    it (the reference) the fields near the source drift."""
    rng = np.random.default_rng(2)
    vp = (2200 + rng.uniform(-300, 300, (36, 40, 32))).astype(np.float32)
    nt, dt, dh = 60, 1e-3, 10.0
    w, tlag = O.wavelet(nt, dt, 45.0)
    N = nt + tlag
    src, rec = (20, 18, 16), np.array([[10, 10, 8]], np.int32)
    m = N - 25  # back to step 24, inside the source window
    _, Pold_m, _ = O.forward(vp, 6, dh, dt, nt, tlag, w, src, rec, n_end=N - m, fp64=True)
    P, Pold, _ = O.forward(vp, 6, dh, dt, nt, tlag, w, src, rec, fp64=True)
    data = np.zeros((1, nt), np.float32)
    kw = dict(b_begin=0, b_end=m, tlag=tlag, fp64=True)
    _, So, *_ = O.backward(vp, dh, dt, nt, data, rec, P, Pold, phys=2, wavelet_=w, src=src, **kw)
    scale = np.abs(Pold_m).max()
    assert np.abs(So - Pold_m).max() <= 1e-9 * scale
    _, So_ref, *_ = O.backward(vp, dh, dt, nt, data, rec, P, Pold, **kw)
    assert np.abs(So_ref - Pold_m).max() > 1e-3 * scale


def test_gradient_with_dyy_is_symmetric_in_x_and_y():
    """An experiment symmetric under x <-> y (layered model, source on the
    diagonal, square receiver grid) gives, with d2/dy2 in the gradient
    (correction bit 1), a gradient symmetric under x <-> y bit for bit; the
    reference gradient (Dxx + Dzz only) is not symmetric."""
    n, nz = 34, 30
    z = np.arange(nz)
    vp = np.broadcast_to(np.where(z < 15, 1800.0, 2400.0).astype(np.float32), (n, n, nz)).copy()
    nt, dt, dh = 50, 1e-3, 10.0
    w, tlag = O.wavelet(nt, dt, 45.0)
    g = np.arange(9, 26, 4)
    X, Y = np.meshgrid(g, g, indexing="ij")
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 6)], axis=1).astype(np.int32)
    P, Pold, seis = O.forward(vp, 6, dh, dt, nt, tlag, w, (17, 17, 6), rec)
    g1 = O.backward(vp, dh, dt, nt, seis, rec, P, Pold, fwi=True, tlag=tlag, phys=1)[4]
    g0 = O.backward(vp, dh, dt, nt, seis, rec, P, Pold, fwi=True, tlag=tlag)[4]
    assert np.abs(g1).max() > 0
    assert np.array_equal(g1, g1.transpose(1, 0, 2))
    assert not np.array_equal(g0, g0.transpose(1, 0, 2))


def test_aligned_imaging_pairs_the_current_source_field():
    """With aligned imaging (correction bit 4) every step adds So * R and So^2,
    So being the source field the step reads and R the receiver field after
    the step's injection: replay the steps one at a time and accumulate those
    products in numpy (float32, the same order)."""
    rng = np.random.default_rng(8)
    shape = (30, 32, 28)
    vp = (2000 + rng.uniform(-100, 100, shape)).astype(np.float32)
    nt, dt, dh = 40, 1e-3, 10.0
    w, tlag = O.wavelet(nt, dt, 45.0)
    rec = np.array([[12, 12, 8], [18, 16, 8], [12, 12, 8]], np.int32)
    P, Pold, seis = O.forward(vp, 6, dh, dt, nt, tlag, w, (15, 15, 10), rec)
    data = rng.normal(0, 1, seis.shape).astype(np.float32)
    idh3 = np.float32(1.0) / (np.float32(dh) * np.float32(dh) * np.float32(dh))
    inner = (slice(4, -4),) * 3
    state, img, ill = [P, Pold, np.zeros(shape, np.float32), np.zeros(shape, np.float32)], None, None
    want_img, want_ill = np.zeros(shape, np.float32), np.zeros(shape, np.float32)
    for b in range(30):
        c = state[1].copy()
        R = state[2].copy()
        for r, (x, y, zz) in enumerate(rec):
            R[y, x, zz] = R[y, x, zz] + idh3 * data[r, nt - 1 - b]
        want_img[inner] = want_img[inner] + c[inner] * R[inner]
        want_ill[inner] = want_ill[inner] + c[inner] * c[inner]
        out = O.backward(vp, dh, dt, nt, data, rec, *state, img, ill, b_begin=b, b_end=b + 1, tlag=tlag, phys=4)
        state, img, ill = list(out[:4]), out[4], out[5]
    assert np.abs(want_img).max() > 0
    assert np.array_equal(img, want_img)
    assert np.array_equal(ill, want_ill)


def test_receiver_field_equals_forward_of_reversed_data():
    vp = np.full((34, 34, 34), 2000.0, np.float32)
    nt, dt, dh = 50, 1e-3, 10.0
    rng = np.random.default_rng(4)
    data = rng.normal(0, 1, (1, nt)).astype(np.float32)
    r = (17, 17, 17)
    z = np.zeros((34, 34, 34), np.float64)
    _, _, R, Ro, img, ill = O.backward(vp, dh, dt, nt, data, np.array([r], np.int32), z, z, b_begin=0, b_end=40,
                                       tlag=0, fp64=True)
    P, Pold, _ = O.forward(vp, 6, dh, dt, nt, 0, data[0, ::-1].copy(), r, np.zeros((0, 3), np.int32), n_end=40,
                           record=False, fp64=True)
    # forward leaves (P, Pold) = (u_40 without step-40 injection, u_39); backward likewise
    assert np.allclose(R, P, rtol=0, atol=1e-12 * np.abs(P).max())
    assert np.all(img == 0) and np.all(ill == 0)  # zero source field -> nothing imaged


def test_image_scaling():
    rng = np.random.default_rng(6)
    vp = (2000 + rng.uniform(-100, 100, (30, 32, 28))).astype(np.float32)
    nt, dt, dh = 40, 1e-3, 10.0
    w, tlag = O.wavelet(nt, dt, 45.0)
    rec = np.array([[12, 12, 8], [18, 16, 8]], np.int32)
    P, Pold, seis = O.forward(vp, 6, dh, dt, nt, tlag, w, (15, 15, 10), rec, fp64=True)
    out1 = O.backward(vp, dh, dt, nt, seis.astype(np.float32), rec, P, Pold, tlag=tlag, fp64=True)
    out2 = O.backward(vp, dh, dt, nt, (2 * seis).astype(np.float32), rec, 3 * P, 3 * Pold, tlag=tlag, fp64=True)
    assert np.abs(out2[4] - 6 * out1[4]).max() <= 1e-12 * np.abs(out1[4]).max()  # image bilinear
    assert np.abs(out2[5] - 9 * out1[5]).max() <= 1e-12 * np.abs(out1[5]).max()  # illum quadratic in S
    outg = O.backward(vp, dh, dt, nt, seis.astype(np.float32), rec, P, Pold, tlag=tlag, fwi=True, fp64=True)
    assert np.array_equal(outg[5], out1[5])


def test_fwi_gradient_ignores_y_curvature():
    """build_gradient uses (Dxx + Dzz) of the source field only
    (inversion.cu:466): a source field varying only along y gives zero."""
    n = 30
    k = np.arange(n, dtype=np.float64)
    S = np.broadcast_to(np.sin(0.3 * k)[:, None, None], (n, n, n)).copy()
    vp = np.full((n, n, n), 2000.0, np.float32)
    R = np.ones((n, n, n))
    data = np.zeros((0, 10), np.float32)
    out = O.backward(vp, 10.0, 1e-3, 10, data, np.zeros((0, 3), np.int32), S, S, R=R, Ro=R, fwi=True, b_begin=0,
                     b_end=1, tlag=0, fp64=True)
    assert np.abs(out[4]).max() < 1e-12


def test_wavelet_closed_form():
    for nt, dt, fmax, tl in ((1000, 1e-3, 45.0, 79), (3000, 0.8e-3, 60.0, 74), (2001, 1e-3, 50.0, 71)):
        w, tlag = O.wavelet(nt, dt, fmax)
        assert tlag == tl
        t = np.arange(nt) * dt - 2 * math.sqrt(math.pi) / fmax
        a = (math.pi * fmax / 3 * t) ** 2
        want = 1e5 * (1 - 2 * a) * np.exp(-a)
        assert np.abs(w - want).max() <= 2e-6 * 1e5
        assert np.argmax(w) == int(round(2 * math.sqrt(math.pi) / fmax / dt))


def test_dampers_closed_form():
    nb, f = 51, 0.004
    b1, b2, b3 = O.dampers(nb, f)
    i = np.arange(nb)
    assert np.allclose(b1, np.exp(-(f * (nb - i)) ** 2), rtol=1e-6)
    assert abs(b1[0] - 0.95924) < 1e-5
    B2 = b2.reshape(nb, nb)
    assert np.array_equal(B2, B2.T)
    assert np.allclose(B2, b1[:, None] + b1[None, :] - 1, atol=1e-6)
    B3 = b3.reshape(nb, nb, nb)
    assert np.allclose(B3, 2 * (b1[None, None, :] + b1[None, :, None] + b1[:, None, None]) - 5, atol=1e-5)
    assert abs(B2[0, 0] - 0.91848) < 1e-5 and abs(B3[0, 0, 0] - 0.75543) < 1e-5
    # lookup: 1 deep inside, mirrored faces agree, edge/corner use 2-D/3-D
    nxx, nyy, nzz = 200, 190, 180
    t = (b1, b2, b3)
    assert O.damper_at(t, 60, 60, 60, nxx, nyy, nzz, nb) == 1.0
    assert O.damper_at(t, 3, 60, 60, nxx, nyy, nzz, nb) == b1[3]
    assert O.damper_at(t, nzz - 4, 60, 60, nxx, nyy, nzz, nb) == b1[3]
    assert O.damper_at(t, 3, nxx - 6, 60, nxx, nyy, nzz, nb) == B2[5, 3]
    assert O.damper_at(t, 3, 7, nyy - 2, nxx, nyy, nzz, nb) == B3[1, 7, 3]


def test_expand_reduce():
    rng = np.random.default_rng(1)
    v = rng.normal(size=(7, 9, 5)).astype(np.float32)
    nb = 3
    pad = O.expand_boundary(v, nb)
    assert np.array_equal(pad, np.pad(v, nb, mode="edge"))
    assert np.array_equal(O.reduce_boundary(pad, nb), v)


def test_position_index():
    for x in (0.0, 4.99, 5.0, 20.0, 499.7, 1234.5):
        for dh, nb in ((10.0, 50), (25.0, 51)):
            assert O.position_index(x, dh, nb) == int(np.float32((np.float32(x) + np.float32(0.5) * np.float32(dh)) / np.float32(dh))) + nb


def test_export_seismic_quadratic():
    nb, dh = 3, 10.0
    nz, nx, ny = 12, 5, 4
    zz = (np.arange(nz + 2 * nb) * dh).astype(np.float32)
    a = 0.5
    image = np.broadcast_to((a * zz ** 2).astype(np.float32), (ny + 2 * nb, nx + 2 * nb, nz + 2 * nb)).copy()
    illum = np.ones_like(image)
    out = O.export_seismic(image, illum, nb, dh)
    assert np.allclose(out[:, :, 1:-1], -2 * a, rtol=1e-4)
    assert np.all(out[:, :, 0] == 0) and np.all(out[:, :, -1] == 0)


def test_gradient_post_mask_and_scale():
    nb, dh = 3, 10.0
    nz, nx, ny = 12, 5, 4
    rng = np.random.default_rng(3)
    g = rng.normal(size=(ny + 2 * nb, nx + 2 * nb, nz + 2 * nb)).astype(np.float32)
    ill = np.ones_like(g)
    out, gmax = O.gradient_post(g, ill, nb, dh, 40.0)
    assert np.all(out[:, :, :5] == 0) and np.all(out[:, :, -1] == 0)  # i <= (int)(40/10) masked
    assert abs(np.abs(out).max() - 1.0) < 1e-6 and gmax > 0


def test_misfit_fsum():
    rng = np.random.default_rng(8)
    obs = rng.normal(size=100000).astype(np.float32)
    cal = rng.normal(size=100000).astype(np.float32)
    res, s, s32 = O.misfit(obs, cal)
    assert np.array_equal(res, obs - cal)
    exact = math.fsum(float(r) * float(r) for r in res.astype(np.float64))
    assert abs(s - exact) <= 1e-14 * exact
    assert abs(s32 - exact) <= 1e-3 * exact


def test_parabola_vertex():
    for amin in (0.013, 0.03, -0.01):
        f = [np.float32((a - amin) ** 2 * 4.0 + 1.0) for a in (0.0, 0.02, 0.05)]
        st = O.parabola_step(0.0, 0.02, 0.05, *f)
        assert abs(st - amin) < 1e-5


def test_model_step():
    rng = np.random.default_rng(2)
    v = rng.uniform(1500, 4000, 1000).astype(np.float32)
    m = (1.0 / v.astype(np.float64) ** 2).astype(np.float32)
    g = rng.normal(0, 1, 1000).astype(np.float32)
    mo, vo = O.model_step(m, g, 1e-9)
    want = (1.0 / np.sqrt((m - np.float32(1e-9) * g).astype(np.float32).astype(np.float64))).astype(np.float32)
    assert np.array_equal(mo, (m - np.float32(1e-9) * g).astype(np.float32))
    assert np.max(np.abs(vo.view(np.int32) - want.view(np.int32))) <= 1


def test_random_boundary_invariants():
    rng = np.random.default_rng(5)
    nb = 6
    inner = rng.uniform(1500, 2500, (10, 12, 8)).astype(np.float32)
    pad = O.expand_boundary(inner, nb)
    a = O.random_boundary(pad, nb, 123, 4, 0.6, 0.0)
    b = O.random_boundary(pad, nb, 123, 4, 0.6, 0.0)
    c = O.random_boundary(pad, nb, 123, 5, 0.6, 0.0)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert np.array_equal(a[nb:-nb, nb:-nb, nb:-nb], inner)
    ratio = a / pad - 1.0
    nyy, nxx, nzz = pad.shape
    K, J, I = np.meshgrid(np.arange(nyy), np.arange(nxx), np.arange(nzz), indexing="ij")

    def ramp(q, n):
        d = np.where(q < nb, q, n - 1 - q)
        return np.where((q < nb) | (q >= n - nb), 1.0 - d / nb, 0.0)

    r = np.maximum(np.maximum(ramp(I, nzz), ramp(J, nxx)), ramp(K, nyy))
    assert np.all(np.abs(ratio) <= 0.6 * r + 1e-6)
    edge = r == 1.0
    assert np.abs(ratio[edge]).max() > 0.5  # reaches ~60 % at the outer edge
    clamped = O.random_boundary(pad, nb, 123, 4, 0.6, 2600.0)
    assert clamped.max() <= 2600.0
    u = np.array([O.rbc_uniform(9, 0, q) for q in range(20000)])
    assert u.min() >= -1.0 and u.max() < 1.0
    assert abs(u.mean()) < 0.02 and abs(u.var() - 1 / 3) < 0.02


def test_exp_det_matches_math_exp():
    for x in np.linspace(-9.0, 0.0, 2001):
        e = math.exp(x)
        assert abs(O.exp_det(x) - e) <= 2e-15 * e, x


def test_irwin_hall_normal_moments():
    v = np.array([O.irwin_hall(5, 1, c, a) for c in range(6000) for a in range(3)], np.float64)
    assert abs(v.mean()) < 0.03 and abs(v.std() - 1.0) < 0.03
    assert v.min() >= -6.0 and v.max() <= 6.0
    assert abs((v ** 4).mean() - 2.9) < 0.2  # kurtosis of the 12-term sum: 3 - 6/60


def test_blob_boundary_single_blob_closed_form():
    """No background noise (var_vp = 0), one blob, wide clamp bounds: boundary
    cells get v0 + A exp(-q/2) within 4 radii, everything else is unchanged."""
    n, nb, dh = 24, 6, 10.0
    vp = np.full((n, n, n), 2000.0, np.float32)
    blob = np.array([[30.0, 45.0, 200.0, 35.0, -300.0]], np.float32)  # near the x = 0 face
    out = O.random_boundary_blobs(vp, nb, dh, 3, 0, 0.0, blob, 0.0, 1e9, 0.0)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    X, Y, Z = (j * np.float32(dh)).astype(np.float64), (k * np.float32(dh)).astype(np.float64), \
        (i * np.float32(dh)).astype(np.float64)
    q = ((X - 30.0) / 35.0) ** 2 + ((Y - 45.0) / 35.0) ** 2 + ((Z - 200.0) / 35.0) ** 2
    inner = (j >= nb) & (j < n - nb) & (k >= nb) & (k < n - nb) & (i >= nb) & (i < n - nb)
    want = np.where(~inner & (q <= 16.0), np.float32(2000.0) + (-300.0 * np.exp(-0.5 * q)).astype(np.float32),
                    np.float32(2000.0)).astype(np.float32)
    diff = np.abs(out.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    assert diff.max() <= 1
    assert np.array_equal(out[inner], vp[inner])
    assert (out < 2000.0).sum() > 50
    # the clamp follows every blob, also where none is in reach
    far = np.array([[30.0, 45.0, 200.0, 1.0, 50.0]], np.float32)
    out = O.random_boundary_blobs(vp, nb, dh, 3, 0, 0.0, far, 0.0, 1900.0, 0.0)
    assert np.all(out[~inner] == np.float32(1900.0)) and np.array_equal(out[inner], vp[inner])
    out = O.random_boundary_blobs(vp, nb, dh, 3, 0, 0.0, far[:0], 0.0, 1900.0, 0.0)
    assert np.array_equal(out, vp)  # no blob, no clamp


def test_blob_boundary_restarts_from_the_interior():
    """A second application gives the same volume (the base value is read
    from the nearest interior cell, not from the boundary cell itself)."""
    rng = np.random.default_rng(2)
    n, nb = 20, 5
    vp = O.expand_boundary(rng.uniform(1500, 2500, (n - 2 * nb,) * 3).astype(np.float32), nb)
    blobs = np.array([[0.0, 0.0, 0.0, 30.0, 400.0], [190.0, 100.0, 50.0, 20.0, -300.0]], np.float32)
    a = O.random_boundary_blobs(vp, nb, 10.0, 9, 1, 200.0, blobs, 1000.0, 3000.0, 0.0)
    b = O.random_boundary_blobs(a, nb, 10.0, 9, 1, 200.0, blobs, 1000.0, 3000.0, 0.0)
    assert np.array_equal(a, b)
    assert np.array_equal(a[nb:-nb, nb:-nb, nb:-nb], vp[nb:-nb, nb:-nb, nb:-nb])
    assert not np.array_equal(a, vp)


def _batched_case(S=3, seed=5):
    from tests.cases import Case
    case = Case(nx=20, ny=18, nz=16, nb=4, nt=24, amp=0.0)
    rng = np.random.default_rng(seed)
    pad = O.expand_boundary(case.interior, case.nb)
    vps = [O.random_boundary(pad, case.nb, 9, s, 0.5, case.vclamp) for s in range(S)]
    shots = []
    for s in range(S):
        src = (case.nb + 3 + 3 * s, case.nb + 5, case.nb + 2)
        rec = case.rec[s::S] if s != 1 else case.rec[:0]  # one shot without receivers
        shots.append((src, rec))
    states, datas = [], []
    for s, (src, rec) in enumerate(shots):
        P, Pold, _ = O.forward(vps[s], case.nb, case.dh, case.dt, case.nt, case.tlag, case.wavelet, src, rec,
                               record=False)
        states.append([P, Pold, np.zeros_like(P), np.zeros_like(P)])
        datas.append(rng.normal(0, 1e-2, (len(rec), case.nt)).astype(np.float32))
    return case, vps, shots, states, datas


def _copy_states(states):
    return [[a.copy() for a in st] for st in states]


@pytest.mark.parametrize("fwi,phys", [(False, 0), (True, 0), (True, 7)])
def test_backward_batched_orders(fwi, phys):
    case, vps, shots, states0, datas = _batched_case()
    S = len(vps)
    kw = dict(fwi=fwi, tlag=case.tlag, phys=phys, wavelet_=case.wavelet)
    srcs = [src for src, _ in shots]
    zero = np.zeros_like(vps[0])
    # one shot: the single-shot backward
    st = _copy_states(states0[:1])
    img, ill = [zero.copy()], [zero.copy()]
    O.backward_batched(vps[:1], 1, case.dh, case.dt, case.nt, datas[:1], [shots[0][1]], st, img, ill, srcs=srcs[:1],
                       **kw)
    want = O.backward(vps[0], case.dh, case.dt, case.nt, datas[0], shots[0][1], *states0[0], fwi=fwi, tlag=case.tlag,
                      phys=phys, wavelet_=case.wavelet, src=srcs[0])
    for a, b in zip(st[0] + [img[0], ill[0]], want):
        assert np.array_equal(a, b)
    # one shot per group: independent per-shot passes, summed in group order
    st = _copy_states(states0)
    img, ill = [zero.copy() for _ in range(S)], [zero.copy() for _ in range(S)]
    O.backward_batched(vps, 1, case.dh, case.dt, case.nt, datas, [r for _, r in shots], st, img, ill, srcs=srcs, **kw)
    for s in range(S):
        want = O.backward(vps[s], case.dh, case.dt, case.nt, datas[s], shots[s][1], *states0[s], fwi=fwi,
                          tlag=case.tlag, phys=phys, wavelet_=case.wavelet, src=srcs[s])
        for a, b in zip(st[s] + [img[s], ill[s]], want):
            assert np.array_equal(a, b)
    # sb = 2 over an odd segment: the step-by-step interleave of single steps
    b0, b1 = 3, 3 + 17
    st = _copy_states(states0)
    img, ill = [zero.copy(), zero.copy()], [zero.copy(), zero.copy()]
    O.backward_batched(vps, 2, case.dh, case.dt, case.nt, datas, [r for _, r in shots], st, img, ill,
                       b_begin=b0, b_end=b1, srcs=srcs, **kw)
    ref = _copy_states(states0)
    rimg, rill = [zero.copy(), zero.copy()], [zero.copy(), zero.copy()]
    for b in range(b0, b1):
        for s in range(S):
            out = O.backward(vps[s], case.dh, case.dt, case.nt, datas[s], shots[s][1], *ref[s], rimg[s // 2],
                             rill[s // 2], fwi=fwi, b_begin=b, b_end=b + 1, tlag=case.tlag, phys=phys,
                             wavelet_=case.wavelet, src=srcs[s])
            ref[s] = list(out[:4])
            rimg[s // 2], rill[s // 2] = out[4], out[5]
    assert np.abs(rimg[0]).max() > 0 and np.abs(rill[1]).max() > 0
    for s in range(S):
        for a, b in zip(st[s], ref[s]):
            assert np.array_equal(a, b)
    for g in range(2):
        assert np.array_equal(img[g], rimg[g]) and np.array_equal(ill[g], rill[g])

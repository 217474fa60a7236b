"""Batched multi-shot launches against the per-shot path (run with -m gpu).

Shots of uneven aperture (receiver patches clipped at the model edges) share
one step kernel per time step.  Forward seismograms are per-shot independent
and must match single-shot runs bit for bit.  The stacked RTM image sums the
shots per time step instead of shot after shot: it equals the oracle run in
that order bit for bit, and differs from sequential per-shot accumulation by
rounding only (held to relative L2 <= 1e-4 for the batched image)."""
import numpy as np
import pytest

import oracle as O
from tests.cases import Case, rel_l2

pytestmark = pytest.mark.gpu

TOL_BATCH = 1e-4


@pytest.fixture(scope="module")
def fl():
    from fwi3d_amd import lib
    lib.load()
    return lib


def _shots(case, nshots):
    """Sources on a grid near the surface, each with a receiver patch of
    half-width 10 cells centred on it and clipped to the interior."""
    nb = case.nb
    # one column of shots near the low-x edge (clipped patches), one inside
    xs = np.array([nb + 3, nb + case.nx // 2])[: 2 if nshots > 1 else 1]
    ys = np.linspace(nb + 3, nb + case.ny // 2, max(1, nshots // 2)).astype(int)
    out = []
    for sy in ys:
        for sx in xs:
            gx = np.arange(sx - 10, sx + 11, 2)
            gy = np.arange(sy - 10, sy + 11, 2)
            gx = gx[(gx >= nb) & (gx < nb + case.nx)]
            gy = gy[(gy >= nb) & (gy < nb + case.ny)]
            X, Y = np.meshgrid(gx, gy, indexing="ij")
            rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, nb + 2)], axis=1).astype(np.int32)
            out.append(((int(sx), int(sy), nb + 2), rec))
    return out[:nshots]


def _single(fl, case, interior, shot, src, rec, data=None, image=None, illum=None):
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag)
    c.set_model_interior(interior)
    c.random_boundary(99, shot, 0.6, float(case.vclamp))
    c.set_wavelet(case.wavelet)
    c.set_source(*src)
    c.set_receivers(rec)
    if image is not None:
        c.set_volume(fl.VOL_IMAGE, image)
        c.set_volume(fl.VOL_ILLUM, illum)
    c.forward(fl.BOUNDARY_RBC)
    seis = c.seismogram()
    if data is not None:
        c.set_backward_data(data)
        c.rtm_backward()
        image, illum = c.get_volume(fl.VOL_IMAGE), c.get_volume(fl.VOL_ILLUM)
    c.close()
    return seis, image, illum


@pytest.mark.parametrize("sb", [1, 2, 4])
def test_batched_rtm_matches_per_shot(fl, sb):
    case = Case(nx=48, ny=44, nz=32, nt=70, amp=0.0)
    S = 4
    shots = _shots(case, S)
    # per-shot reference path: sequential shots accumulating into one image
    seis_ref, image, illum = [], None, None
    for s, (src, rec) in enumerate(shots):
        seis, _, _ = _single(fl, case, case.interior, s, src, rec)
        seis_ref.append(seis)
    zero = np.zeros((case.nyy, case.nxx, case.nzz), np.float32)
    image, illum = zero, zero
    for s, (src, rec) in enumerate(shots):
        _, image, illum = _single(fl, case, case.interior, s, src, rec, data=seis_ref[s], image=image, illum=illum)

    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
    c.set_model_interior(case.interior)
    for s, (src, rec) in enumerate(shots):
        c.shot_random_boundary(s, 99, s, 0.6, float(case.vclamp))
        c.set_source(*src, slot=s)
        c.set_receivers(rec, slot=s)
    c.set_wavelet(case.wavelet)
    for kind in (0, 2):
        c.set_shots_per_block(kind, sb)
    c.forward(fl.BOUNDARY_RBC)
    for s in range(S):
        assert np.array_equal(c.seismogram(slot=s), seis_ref[s]), s
        c.set_backward_data(seis_ref[s], slot=s)
    c.rtm_backward()
    img_b = c.get_volume(fl.VOL_IMAGE)
    ill_b = c.get_volume(fl.VOL_ILLUM)
    c.close()
    assert len({len(r) for _, r in shots}) > 1 or S == 1  # uneven apertures exercised
    assert np.abs(image).max() > 0
    assert rel_l2(img_b, image) <= TOL_BATCH, rel_l2(img_b, image)
    assert rel_l2(ill_b, illum) <= TOL_BATCH, rel_l2(ill_b, illum)


def _interleaved_oracle_rtm(case, shots, vps, datas, sb, fwi=False, phys=0):
    """The batched RTM's summation order on the CPU oracle: every time step
    advances the shots in slot order, each group of sb consecutive shots
    adds its terms into its own image/illumination partial, and the partials
    are summed in group order after the pass."""
    S = len(shots)
    st = []
    for s, (src, rec) in enumerate(shots):
        P, Pold, _ = O.forward(vps[s], case.nb, case.dh, case.dt, case.nt, case.tlag, case.wavelet, src, rec,
                               record=False)
        st.append([P, Pold, None, None])
    G = S // sb
    img, ill = [None] * G, [None] * G
    for b in range(case.steps):
        for s in range(S):
            gi = s // sb
            out = O.backward(vps[s], case.dh, case.dt, case.nt, datas[s], shots[s][1], *st[s], img[gi], ill[gi],
                             fwi=fwi, b_begin=b, b_end=b + 1, tlag=case.tlag, phys=phys,
                             wavelet_=case.wavelet, src=shots[s][0])
            st[s] = list(out[:4])
            img[gi], ill[gi] = out[4], out[5]
    fimg, fill = img[0].copy(), ill[0].copy()
    for g in range(1, G):
        fimg = fimg + img[g]
        fill = fill + ill[g]
    return fimg, fill


def _set_batch_shape(c, kinds, sb):
    """sb shots per block as wave groups, or "pair<BY>": shot pairs (two
    slots per thread) on 64 x BY tiles, two z-chunks."""
    for kind in kinds:
        if isinstance(sb, str):
            c.set_shots_per_thread(kind, 2, int(sb[4:]), 2)
        else:
            c.set_shots_per_block(kind, sb)


def _sb(sb):
    return 2 if isinstance(sb, str) else sb


@pytest.mark.parametrize(
    "sb,boundary",
    [(1, "ramp"), (2, "ramp"), (4, "ramp"), ("pair4", "ramp"), ("pair8", "ramp"), ("pair16", "ramp"),
     (2, "blobs"), ("pair8", "blobs")],
)
def test_batched_rtm_bit_exact_vs_interleaved_oracle(fl, sb, boundary):
    """With sb shots per block (or shot pairs per thread) the image and
    illumination equal, bit for bit, the oracle run in the batched summation
    order.  The slots' random boundaries differ, so a pair's shared interior
    velocity and its own boundary velocities are both exercised: seeded
    ramps, or the reference-style velocity blobs of the migration and
    inversion drivers (modeling.cu:558-615)."""
    case = Case(nx=40, ny=36, nz=28, nt=50, amp=0.0)
    S = 4
    shots = _shots(case, S)
    padded = O.expand_boundary(case.interior, case.nb)
    ratio, var_vp = 40.0, 800.0
    if boundary == "ramp":
        vps = [O.random_boundary(padded, case.nb, 99, s, 0.6, case.vclamp) for s in range(S)]
    else:
        vlo = np.float32(case.interior.min()) - np.float32(var_vp)
        vhi = np.float32(case.interior.max()) + np.float32(var_vp)
        vps = []
        for s in range(S):
            blobs = fl.rbc_blobs(7, s, case.nxx, case.nyy, case.nzz, case.nb, case.dh, ratio, var_vp)
            vps.append(O.random_boundary_blobs(padded, case.nb, case.dh, 7, s, var_vp, blobs, vlo, vhi,
                                               case.vclamp))
    # what a pair's shared interior velocity rests on: the boundaries differ,
    # the interiors do not
    nb = case.nb
    for s in range(1, S):
        assert not np.array_equal(vps[s], vps[0])
        assert np.array_equal(vps[s][nb:-nb, nb:-nb, nb:-nb], vps[0][nb:-nb, nb:-nb, nb:-nb])
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
    c.set_model_interior(case.interior)
    for s, (src, rec) in enumerate(shots):
        if boundary == "ramp":
            c.shot_random_boundary(s, 99, s, 0.6, float(case.vclamp))
        else:
            c.random_boundary_blobs(7, s, ratio, var_vp, float(case.vclamp), slot=s)
        assert np.array_equal(c.get_volume(fl.VOL_MODEL, slot=s), vps[s]), s
        c.set_source(*src, slot=s)
        c.set_receivers(rec, slot=s)
    c.set_wavelet(case.wavelet)
    _set_batch_shape(c, (0, 2), sb)
    c.set_graph(sb in (2, "pair8"))  # the hipGraph-captured time loop gives the same bits
    c.forward(fl.BOUNDARY_RBC)
    rng = np.random.default_rng(_sb(sb))
    datas = []
    for s in range(S):
        seis = c.seismogram(slot=s)
        _, _, want = O.forward(vps[s], case.nb, case.dh, case.dt, case.nt, case.tlag, case.wavelet, shots[s][0],
                               shots[s][1])
        assert np.array_equal(seis, want), s
        d = (seis + rng.normal(0, 1e-3, (len(shots[s][1]), case.nt))).astype(np.float32)
        datas.append(d)
        c.set_backward_data(d, slot=s)
    c.rtm_backward()
    img_b = c.get_volume(fl.VOL_IMAGE)
    ill_b = c.get_volume(fl.VOL_ILLUM)
    c.close()
    want_img, want_ill = _interleaved_oracle_rtm(case, shots, vps, datas, _sb(sb))
    assert np.abs(want_img).max() > 0
    assert np.array_equal(ill_b, want_ill)
    assert np.array_equal(img_b, want_img)


@pytest.mark.parametrize("fwi", [False, True])
@pytest.mark.parametrize("sb", [2, 4, "pair16"])
def test_batched_deferred_imaging_identical(fl, sb, fwi):
    """Batched RTM with the image updated every other launch and the
    illumination every third (the default; the other groups hand group 0
    each step's terms), and batched FWI with the illumination deferred
    (blocks of up to 512 threads; larger ones update every step), equal
    per-step updates bit for bit, over a whole pass and over the same
    segments of every length modulo 6 (each segment ends by folding the shot
    groups' partial sums into group 0, so segmented and whole passes round
    differently when there are several groups).  An odd number of planes
    and two z-chunks (of 18 and 17 planes) exercise the last-plane path of
    the parity-unrolled loop."""
    case = Case(nx=40, ny=36, nz=27, nt=51, amp=0.0)
    S = 4
    shots = _shots(case, S)
    outs = []
    cuts_odd = [0, 5, 6, 8, 12, 15, 21, 30, case.steps]
    for defer, cuts in ((False, None), (True, None), (False, cuts_odd), (True, cuts_odd)):
        c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
        c.set_model_interior(case.interior)
        for s, (src, rec) in enumerate(shots):
            c.shot_random_boundary(s, 99, s, 0.6, float(case.vclamp))
            c.set_source(*src, slot=s)
            c.set_receivers(rec, slot=s)
        c.set_wavelet(case.wavelet)
        for kind in (0, 3 if fwi else 2):
            _set_batch_shape(c, (kind,), sb)
            by, rpt, _ = c.tuning(kind)
            c.set_tuning(kind, by, rpt, 2)
        c.set_deferred_imaging(defer)
        c.forward(fl.BOUNDARY_RBC)
        rng = np.random.default_rng(_sb(sb))
        for s in range(S):
            d = (c.seismogram(slot=s) + rng.normal(0, 1e-3, (len(shots[s][1]), case.nt))).astype(np.float32)
            c.set_backward_data(d, slot=s)
        if cuts is None:
            (c.grad_backward if fwi else c.rtm_backward)()
        else:
            for b0, b1 in zip(cuts[:-1], cuts[1:]):
                c.backward_steps(1 if fwi else 0, b0, b1)
        outs.append((c.get_volume(fl.VOL_GRAD if fwi else fl.VOL_IMAGE), c.get_volume(fl.VOL_ILLUM)))
        c.close()
    assert np.abs(outs[0][0]).max() > 0
    for k in (1, 3):
        assert np.array_equal(outs[k][0], outs[k - 1][0]), k
        assert np.array_equal(outs[k][1], outs[k - 1][1]), k


@pytest.mark.parametrize("phys", [0, 7])
def test_shot_pair_fwi_gradient_vs_interleaved_oracle(fl, phys):
    """FWI gradient with shot pairs: the reference physics on the pair
    instances (gradient every step, illumination every third launch), and
    with all corrections, which run the pairs as two wave groups per block;
    bit for bit against the oracle in the batched summation order.  The
    absorbing-boundary forward of the same pairs (the FWI calculated data)
    equals the oracle's seismograms."""
    case = Case(nx=40, ny=36, nz=28, nt=50, amp=0.0, abc_factor=0.05)
    S = 4
    shots = _shots(case, S)
    padded = O.expand_boundary(case.interior, case.nb)
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
    c.set_model_interior(case.interior)
    c.set_wavelet(case.wavelet)
    c.set_dampers(case.abc[0])
    for s, (src, rec) in enumerate(shots):
        c.set_source(*src, slot=s)
        c.set_receivers(rec, slot=s)
    _set_batch_shape(c, (1,), "pair16")
    c.forward(fl.BOUNDARY_ABC)
    for s, (src, rec) in enumerate(shots):
        _, _, want = O.forward(padded, case.nb, case.dh, case.dt, case.nt, case.tlag, case.wavelet, src, rec,
                               abc_tables=case.abc)
        assert np.array_equal(c.seismogram(slot=s), want), s
    vps = [O.random_boundary(padded, case.nb, 99, s, 0.6, case.vclamp) for s in range(S)]
    for s in range(S):
        c.shot_random_boundary(s, 99, s, 0.6, float(case.vclamp))
    _set_batch_shape(c, (0, 3), "pair8")
    c.set_physics(phys)
    c.forward(fl.BOUNDARY_RBC)
    rng = np.random.default_rng(11 + phys)
    datas = []
    for s in range(S):
        d = rng.normal(0, 1e-2, (len(shots[s][1]), case.nt)).astype(np.float32)
        datas.append(d)
        c.set_backward_data(d, slot=s)
    c.zero_volume(fl.VOL_GRAD)
    c.grad_backward()
    g_b = c.get_volume(fl.VOL_GRAD)
    ill_b = c.get_volume(fl.VOL_ILLUM)
    c.close()
    want_g, want_ill = _interleaved_oracle_rtm(case, shots, vps, datas, 2, fwi=True, phys=phys)
    assert np.abs(want_g).max() > 0
    assert np.array_equal(ill_b, want_ill)
    assert np.array_equal(g_b, want_g)


@pytest.mark.parametrize("sb", [2, "pair8"])
def test_batched_sources_and_receivers_on_the_updated_edge(fl, sb):
    """A batch whose shots sit on opposite corners of the updated region
    ((4, 4, 4) and its far corner) with receivers on the corners and faces
    next to the rims: seismograms against per-shot oracle runs, image and
    illumination against the oracle in the batch's order, bit for bit."""
    case = Case(nt=60, amp=0.0)
    hi = (case.nxx - 5, case.nyy - 5, case.nzz - 5)
    rec = np.array([(4, 4, 4), hi, (4, hi[1], 4), (hi[0], 4, hi[2]), (4, 4, hi[2]), (hi[0], hi[1], 4),
                    (case.nxx // 2, 4, case.nzz // 2), (4, case.nyy // 2, case.nzz // 2),
                    (case.nxx // 2, case.nyy // 2, hi[2]), (5, 4, 4), (hi[0] - 1, hi[1], hi[2])], np.int32)
    shots = [((4, 4, 4), rec), (hi, rec[::-1].copy())]
    S = len(shots)
    padded = O.expand_boundary(case.interior, case.nb)
    vps = [O.random_boundary(padded, case.nb, 99, s, 0.6, case.vclamp) for s in range(S)]
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
    c.set_model_interior(case.interior)
    for s, (src, r) in enumerate(shots):
        c.shot_random_boundary(s, 99, s, 0.6, float(case.vclamp))
        c.set_source(*src, slot=s)
        c.set_receivers(r, slot=s)
    c.set_wavelet(case.wavelet)
    _set_batch_shape(c, (0, 2), sb)
    c.forward(fl.BOUNDARY_RBC)
    rng = np.random.default_rng(13)
    datas = []
    for s in range(S):
        seis = c.seismogram(slot=s)
        _, _, want = O.forward(vps[s], case.nb, case.dh, case.dt, case.nt, case.tlag, case.wavelet, shots[s][0],
                               shots[s][1])
        assert np.array_equal(seis, want), s
        d = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
        datas.append(d)
        c.set_backward_data(d, slot=s)
    c.rtm_backward()
    img_b = c.get_volume(fl.VOL_IMAGE)
    ill_b = c.get_volume(fl.VOL_ILLUM)
    c.close()
    want_img, want_ill = _interleaved_oracle_rtm(case, shots, vps, datas, _sb(sb))
    assert np.abs(want_img).max() > 0
    assert np.array_equal(ill_b, want_ill)
    assert np.array_equal(img_b, want_img)


def test_shot_pairs_need_an_even_batch(fl):
    case = Case(nx=24, ny=24, nz=20, nt=20, amp=0.0)
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=3)
    with pytest.raises(fl.FwiError):
        c.set_shots_per_thread(2, 2, 16, 1)
    with pytest.raises(fl.FwiError):
        c.set_shots_per_thread(2, 3, 16, 1)
    assert c.shots_per_thread(2) == 1
    c.close()


def test_batched_fwi_gradient_with_a_receiverless_shot(fl):
    """FWI gradient mode with all backward corrections on, two shots per
    block, and a shot with no receivers (its receiver field stays zero, its
    illumination still counts): bit for bit against the oracle in the batched
    summation order."""
    case = Case(nx=40, ny=36, nz=28, nt=50, amp=0.0)
    S, sb = 4, 2
    shots = _shots(case, S)
    shots[1] = (shots[1][0], np.zeros((0, 3), np.int32))
    padded = O.expand_boundary(case.interior, case.nb)
    vps = [O.random_boundary(padded, case.nb, 99, s, 0.6, case.vclamp) for s in range(S)]
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
    c.set_model_interior(case.interior)
    for s, (src, rec) in enumerate(shots):
        c.shot_random_boundary(s, 99, s, 0.6, float(case.vclamp))
        c.set_source(*src, slot=s)
        c.set_receivers(rec, slot=s)
    c.set_wavelet(case.wavelet)
    for kind in (0, 3):
        c.set_shots_per_block(kind, sb)
    c.set_physics(fl.PHYS_GRAD_DYY | fl.PHYS_SUBTRACT_SOURCE | fl.PHYS_ALIGNED_IMAGING)
    c.forward(fl.BOUNDARY_RBC)
    assert c.seismogram(slot=1).shape == (0, case.nt)
    rng = np.random.default_rng(7)
    datas = []
    for s in range(S):
        d = rng.normal(0, 1e-2, (len(shots[s][1]), case.nt)).astype(np.float32)
        datas.append(d)
        c.set_backward_data(d, slot=s)
    c.zero_volume(fl.VOL_GRAD)
    c.grad_backward()
    g_b = c.get_volume(fl.VOL_GRAD)
    ill_b = c.get_volume(fl.VOL_ILLUM)
    c.close()
    want_g, want_ill = _interleaved_oracle_rtm(case, shots, vps, datas, sb, fwi=True, phys=7)
    assert np.abs(want_g).max() > 0
    assert np.array_equal(ill_b, want_ill)
    assert np.array_equal(g_b, want_g)


def test_batched_models_match_oracle(fl):
    case = Case(amp=0.0)
    S = 2
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
    c.set_model_interior(case.interior)
    for s in range(S):
        c.shot_random_boundary(s, 5, 10 + s, 0.6, float(case.vclamp))
    pad = O.expand_boundary(case.interior, case.nb)
    for s in range(S):
        want = O.random_boundary(pad, case.nb, 5, 10 + s, 0.6, case.vclamp)
        assert np.array_equal(c.get_volume(fl.VOL_MODEL, slot=s), want)
    c.close()


def test_batched_blob_boundaries_match_oracle(fl):
    """Per-slot reference-style boundaries: each slot's own shot index."""
    case = Case(amp=0.0)
    S, ratio, var_vp = 2, 40.0, 800.0
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
    c.set_model_interior(case.interior)
    for s in range(S):
        c.random_boundary_blobs(3, 20 + s, ratio, var_vp, float(case.vclamp), slot=s)
    pad = O.expand_boundary(case.interior, case.nb)
    vlo = np.float32(case.interior.min()) - np.float32(var_vp)
    vhi = np.float32(case.interior.max()) + np.float32(var_vp)
    got = [c.get_volume(fl.VOL_MODEL, slot=s) for s in range(S)]
    c.close()
    for s in range(S):
        blobs = fl.rbc_blobs(3, 20 + s, case.nxx, case.nyy, case.nzz, case.nb, case.dh, ratio, var_vp)
        want = O.random_boundary_blobs(pad, case.nb, case.dh, 3, 20 + s, var_vp, blobs, vlo, vhi, case.vclamp)
        assert np.array_equal(got[s], want), s
    assert not np.array_equal(got[0], got[1])


def test_batched_residual_and_fwi_gradient(fl):
    case = Case(nx=40, ny=40, nz=28, nt=50, amp=0.0, abc_factor=0.05)
    S = 2
    shots = _shots(case, S)
    c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=S)
    c.set_model_interior(case.interior)
    c.set_wavelet(case.wavelet)
    c.set_dampers(case.abc[0])
    for s, (src, rec) in enumerate(shots):
        c.set_source(*src, slot=s)
        c.set_receivers(rec, slot=s)
    c.forward(fl.BOUNDARY_ABC)
    sums = []
    for s, (src, rec) in enumerate(shots):
        cal = c.seismogram(slot=s)
        vp = O.expand_boundary(case.interior, case.nb)
        _, _, want = O.forward(vp, case.nb, case.dh, case.dt, case.nt, case.tlag, case.wavelet, src, rec,
                               abc_tables=case.abc)
        assert np.array_equal(cal, want)
        obs = (cal * 0.9).astype(np.float32)
        got = c.residual(obs, slot=s)
        _, ref, _ = O.misfit(obs, cal)
        assert abs(got - ref) <= 1e-12 * ref
        sums.append(got)
    c.close()

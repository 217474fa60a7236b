"""GPU parity of the HIP path against the CPU oracle (run with -m gpu).

The reference arithmetic mode is compared bit for bit: the kernels evaluate
the same float/double operations in the same order as the oracle (both built
without FP contraction).  The FP32 mode is held to relative L2 <= 1e-5.
"""
import numpy as np
import pytest

import oracle as O
from tests.cases import Case, rel_l2

pytestmark = pytest.mark.gpu

TOL_FP32 = 1e-5  # FP32 parity bound against the oracle (relative L2)


@pytest.fixture(scope="module")
def fl():
    from fwi3d_amd import lib
    lib.load()
    return lib


def _oracle_forward(case, abc=False, n_end=None):
    return O.forward(case.vp, case.nb, case.dh, case.dt, case.nt, case.tlag, case.wavelet, case.src, case.rec,
                     abc_tables=case.abc if abc else None, n_end=n_end)


def test_forward_rbc_bit_exact(fl):
    case = Case()
    P, Pold, seis = _oracle_forward(case)
    with case.context(fl.Context) as c:
        c.forward(fl.BOUNDARY_RBC)
        g_seis = c.seismogram()
        g_P = c.get_volume(fl.VOL_FIELD)
        g_Po = c.get_volume(fl.VOL_FIELD_OLD)
    assert np.abs(seis).max() > 0
    assert np.array_equal(g_seis, seis)
    assert np.array_equal(g_P, P)
    assert np.array_equal(g_Po, Pold)
    assert np.all(g_seis[:, 0] == 0)  # sample 0 is never recorded


def test_forward_abc_bit_exact(fl):
    case = Case(amp=0.0, abc_factor=0.05)
    P, Pold, seis = _oracle_forward(case, abc=True)
    with case.context(fl.Context, abc=True) as c:
        c.forward(fl.BOUNDARY_ABC)
        g_seis = c.seismogram()
        g_P = c.get_volume(fl.VOL_FIELD)
        g_Po = c.get_volume(fl.VOL_FIELD_OLD)
    assert np.array_equal(g_seis, seis)
    assert np.array_equal(g_P, P)
    # the device keeps the previous field undamped and damps it on its next
    # read; damping it here must give the oracle's damped copy
    d = np.ones_like(g_Po)
    for k in range(4, case.nyy - 4):
        for j in range(4, case.nxx - 4):
            for i in range(4, case.nzz - 4):
                d[k, j, i] = O.damper_at(case.abc, i, j, k, case.nxx, case.nyy, case.nzz, case.nb)
    inner = np.zeros_like(g_Po, bool)
    inner[4:-4, 4:-4, 4:-4] = True
    assert np.array_equal(np.where(inner, g_Po * d, g_Po), Pold)


@pytest.mark.parametrize("fwi", [False, True])
def test_backward_bit_exact(fl, fwi):
    case = Case()
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(3)
    data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
    S, So, R, Ro, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold, fwi=fwi,
                                            tlag=case.tlag)
    with case.context(fl.Context) as c:
        c.forward(fl.BOUNDARY_RBC)
        c.set_backward_data(data)
        (c.grad_backward if fwi else c.rtm_backward)()
        g_img = c.get_volume(fl.VOL_GRAD if fwi else fl.VOL_IMAGE)
        g_ill = c.get_volume(fl.VOL_ILLUM)
        g_S = c.get_volume(fl.VOL_FIELD)
        g_R = c.get_volume(fl.VOL_RECV)
        g_Ro = c.get_volume(fl.VOL_RECV_OLD)
    assert np.abs(image).max() > 0
    assert np.array_equal(g_ill, illum)
    assert np.array_equal(g_img, image)
    assert np.array_equal(g_S, S)
    assert np.array_equal(g_R, R)
    assert np.array_equal(g_Ro, Ro)


@pytest.mark.parametrize("fwi,two_step", [(False, False), (True, False), (False, True)])
def test_colocated_receivers_bit_exact(fl, fwi, two_step):
    """Receivers sharing a cell (the reference injects them with atomics,
    migration.cu:119-132, inversion.cu:385-398): each records the cell, and
    the backward pass adds their data in receiver order, as the oracle does
    (also inside the two-step launches, which record and inject through a
    map of the receiver planes).  The duplicates sit far apart in the list,
    one cell holds three receivers, and one receiver sits on the source
    cell."""
    case = Case(nx=84, ny=60, nz=31, nt=61) if two_step else Case()
    rec = case.rec
    extra = np.concatenate([rec[::7], rec[len(rec) // 2:len(rec) // 2 + 1], [case.src]]).astype(np.int32)
    rec = np.concatenate([rec[len(rec) // 3:], extra, rec[:len(rec) // 3], rec[len(rec) // 2:len(rec) // 2 + 1]])
    cells = [tuple(r) for r in rec]
    assert max(cells.count(c) for c in cells) == 3
    case.rec = np.ascontiguousarray(rec, np.int32)
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(5)
    data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
    S, So, R, Ro, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold, fwi=fwi,
                                            tlag=case.tlag)
    with case.context(fl.Context) as c:
        if two_step:
            c.set_two_step(fl.KIND_FWD_RBC, True, 16, 2, 1)
            c.set_two_step(fl.KIND_BWD_RTM, True, 16, 2, 1)
        c.forward(fl.BOUNDARY_RBC)
        g_seis = c.seismogram()
        c.set_backward_data(data)
        (c.grad_backward if fwi else c.rtm_backward)()
        if two_step:
            assert c.two_step(fl.KIND_FWD_RBC)[0] == 1 and c.two_step(fl.KIND_BWD_RTM)[0] == 1
        g_img = c.get_volume(fl.VOL_GRAD if fwi else fl.VOL_IMAGE)
        g_ill = c.get_volume(fl.VOL_ILLUM)
        g_R = c.get_volume(fl.VOL_RECV)
        g_Ro = c.get_volume(fl.VOL_RECV_OLD)
    assert np.array_equal(g_seis, seis)
    for i, ci in enumerate(cells):  # co-located receivers record the same trace
        j = cells.index(ci)
        assert np.array_equal(g_seis[i], g_seis[j])
    assert np.abs(image).max() > 0
    assert np.array_equal(g_R, R)
    assert np.array_equal(g_Ro, Ro)
    assert np.array_equal(g_ill, illum)
    assert np.array_equal(g_img, image)


@pytest.mark.parametrize("tuning", [None, (4, 2, 3), (16, 1, 2)])
def test_source_and_receivers_on_the_updated_edge(fl, tuning):
    """The source on the first updated cell (4, 4, 4) and receivers on the
    corners and faces of the updated region (the outermost cells a step
    writes, next to the 4-cell rims): seismogram, image and illumination
    bit for bit against the oracle, with z-chunk boundaries inside the
    grid."""
    case = Case(nt=60)
    hi = (case.nxx - 5, case.nyy - 5, case.nzz - 5)
    case.src = (4, 4, 4)
    rec = [(4, 4, 4), (hi[0], hi[1], hi[2]), (4, hi[1], 4), (hi[0], 4, hi[2]), (4, 4, hi[2]), (hi[0], hi[1], 4),
           (case.nxx // 2, 4, case.nzz // 2), (4, case.nyy // 2, case.nzz // 2), (case.nxx // 2, case.nyy // 2, 4),
           (case.nxx // 2, case.nyy // 2, hi[2]), (5, 4, 4), (4, 5, 4), (4, 4, 5)]
    case.rec = np.array(rec, np.int32)
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(9)
    data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
    S, So, R, Ro, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold,
                                            tlag=case.tlag)
    with case.context(fl.Context) as c:
        if tuning is not None:
            c.set_tuning(fl.KIND_FWD_RBC, *tuning)
            c.set_tuning(fl.KIND_BWD_RTM, *tuning)
        c.forward(fl.BOUNDARY_RBC)
        g_seis = c.seismogram()
        c.set_backward_data(data)
        c.rtm_backward()
        g_img = c.get_volume(fl.VOL_IMAGE)
        g_ill = c.get_volume(fl.VOL_ILLUM)
    assert np.abs(seis).max() > 0 and np.abs(image).max() > 0
    assert np.array_equal(g_seis, seis)
    assert np.array_equal(g_ill, illum)
    assert np.array_equal(g_img, image)


@pytest.mark.parametrize("phys", [0, 3, 7])
@pytest.mark.parametrize("arith", ["reference", "fp32"])
@pytest.mark.parametrize("fwi", [False, True])
def test_deferred_imaging_identical(fl, fwi, arith, phys):
    """Image/illumination (gradient/illumination) updated every second or
    third backward launch, the default, equal per-step updates bit for bit:
    over a whole pass and over segments of every length modulo 6 (a segment
    never leaves a term behind), without corrections, with source
    subtraction and the Dyy gradient term (same cadence), and with all three
    (aligned imaging: every other launch); with reference arithmetic both
    equal the oracle.
    An odd number of planes in two z-chunks (20 and 19) exercises the
    last-plane path of the parity-unrolled loop."""
    case = Case(nz=31, nt=61)
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(11)
    data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
    vol = fl.VOL_GRAD if fwi else fl.VOL_IMAGE
    kind = 1 if fwi else 0
    outs = []
    for defer, cuts in ((False, None), (True, None), (True, [0, 3, 8, 9, 11, 15, 20, 44, case.steps])):
        with case.context(fl.Context) as c:
            c.set_arith(fl.ARITH_FP32 if arith == "fp32" else fl.ARITH_REFERENCE)
            c.set_tuning(2 + kind, 4, 2, 2)
            c.set_physics(phys)
            c.set_deferred_imaging(defer)
            c.forward(fl.BOUNDARY_RBC)
            c.set_backward_data(data)
            if cuts is None:
                (c.grad_backward if fwi else c.rtm_backward)()
            else:
                for b0, b1 in zip(cuts[:-1], cuts[1:]):
                    c.backward_steps(kind, b0, b1)
            outs.append((c.get_volume(vol), c.get_volume(fl.VOL_ILLUM), c.get_volume(fl.VOL_RECV)))
    assert np.abs(outs[0][0]).max() > 0 and np.abs(outs[0][1]).max() > 0
    for k in (1, 2):
        for a, b in zip(outs[0], outs[k]):
            assert np.array_equal(a, b), k
    if arith == "reference":
        _, _, _, _, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold, fwi=fwi,
                                              tlag=case.tlag, phys=phys, wavelet_=case.wavelet, src=case.src)
        assert np.array_equal(outs[1][0], image) and np.array_equal(outs[1][1], illum)


@pytest.mark.parametrize("phys", [1, 2, 4, 7])
@pytest.mark.parametrize("fwi", [False, True])
def test_backward_corrections_bit_exact(fl, fwi, phys):
    """The optional backward corrections (gradient with d2/dy2, source
    subtraction, aligned imaging, all three) against the oracle's."""
    case = Case(nt=70)
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(phys)
    data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
    S, So, R, Ro, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold, fwi=fwi,
                                            tlag=case.tlag, phys=phys, wavelet_=case.wavelet, src=case.src)
    with case.context(fl.Context) as c:
        c.set_physics(phys)
        c.forward(fl.BOUNDARY_RBC)
        c.set_backward_data(data)
        (c.grad_backward if fwi else c.rtm_backward)()
        assert np.array_equal(c.get_volume(fl.VOL_ILLUM), illum)
        assert np.array_equal(c.get_volume(fl.VOL_GRAD if fwi else fl.VOL_IMAGE), image)
        assert np.array_equal(c.get_volume(fl.VOL_FIELD), S)
        assert np.array_equal(c.get_volume(fl.VOL_RECV), R)


def test_step_ranges_compose(fl):
    """forward_steps(0,k) + forward_steps(k,N) == forward_steps(0,N) == oracle."""
    case = Case(nt=60)
    k = 37
    P1, Po1, _ = _oracle_forward(case, n_end=k)
    with case.context(fl.Context) as c:
        c.forward_steps(fl.BOUNDARY_RBC, 0, k)
        assert np.array_equal(c.get_volume(fl.VOL_FIELD), P1)
        assert np.array_equal(c.get_volume(fl.VOL_FIELD_OLD), Po1)
        c.forward_steps(fl.BOUNDARY_RBC, k, case.steps)
        P, Pold, _ = _oracle_forward(case)
        assert np.array_equal(c.get_volume(fl.VOL_FIELD), P)


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_all_tilings_identical(fl, kind):
    """Every tile shape / z-chunking gives the same bits."""
    case = Case(nx=70, ny=44, nz=40, nt=40, abc_factor=0.05)
    ref = None
    P, Pold, seis = _oracle_forward(case, abc=(kind == 1))
    data = seis.copy()
    for by, rpt in [(4, 1), (4, 2), (8, 1), (8, 2), (4, 4), (16, 1), (16, 2)]:
        for zc in (1, 3):
            with case.context(fl.Context, abc=True) as c:
                c.set_tuning(kind, by, rpt, zc)
                if kind < 2:
                    c.forward(kind)
                    out = c.seismogram()
                else:
                    c.forward(fl.BOUNDARY_RBC)
                    c.set_backward_data(data)
                    (c.rtm_backward if kind == 2 else c.grad_backward)()
                    out = c.get_volume(fl.VOL_IMAGE if kind == 2 else fl.VOL_GRAD)
            if ref is None:
                ref = out
                if kind < 2:
                    assert np.array_equal(out, seis)
            assert np.array_equal(out, ref), (by, rpt, zc)


def test_fp32_mode_within_tolerance(fl):
    case = Case(nt=200)
    P, Pold, seis = _oracle_forward(case)
    with case.context(fl.Context) as c:
        c.set_arith(fl.ARITH_FP32)
        c.forward(fl.BOUNDARY_RBC)
        g = c.seismogram()
    err = rel_l2(g, seis)
    assert err <= TOL_FP32, err


def test_graph_matches_stream(fl):
    case = Case(nt=50)
    P, Pold, seis = _oracle_forward(case)
    with case.context(fl.Context) as c:
        c.set_graph(True)
        for _ in range(3):  # replays with eager work in between
            c.forward(fl.BOUNDARY_RBC)
            assert np.array_equal(c.seismogram(), seis)
            c.set_backward_data(seis)
            c.zero_volume(fl.VOL_IMAGE)
            c.zero_volume(fl.VOL_ILLUM)
            c.rtm_backward()
            img_graph = c.get_volume(fl.VOL_IMAGE)
        c.set_graph(False)
        c.forward(fl.BOUNDARY_RBC)
        c.zero_volume(fl.VOL_IMAGE)
        c.zero_volume(fl.VOL_ILLUM)
        c.rtm_backward()
        assert np.array_equal(c.get_volume(fl.VOL_IMAGE), img_graph)


def test_graph_replayed_across_passes(fl):
    """A pass's hipGraph is instantiated once and replayed: three forward +
    RTM rounds instantiate two graphs; a new wavelet (a kernel argument)
    forces a new capture, whose result follows the new wavelet; every
    replay equals the stream path bit for bit."""
    case = Case(nt=50)
    P, Pold, seis = _oracle_forward(case)
    w2 = (case.wavelet * np.float32(0.5)).astype(np.float32)
    with case.context(fl.Context) as c:
        c.set_graph(True)
        for _ in range(3):
            c.forward(fl.BOUNDARY_RBC)
            assert np.array_equal(c.seismogram(), seis)
            c.set_backward_data(seis)
            c.zero_volume(fl.VOL_IMAGE)
            c.zero_volume(fl.VOL_ILLUM)
            c.rtm_backward()
            img = c.get_volume(fl.VOL_IMAGE)
        assert c.graph_instantiations() == 2
        c.set_wavelet(w2)
        c.forward(fl.BOUNDARY_RBC)
        seis2 = c.seismogram()
        assert c.graph_instantiations() == 3
        c.set_graph(False)
        c.forward(fl.BOUNDARY_RBC)
        assert np.array_equal(c.seismogram(), seis2)
        c.set_wavelet(case.wavelet)
        c.forward(fl.BOUNDARY_RBC)
        c.set_backward_data(seis)
        c.zero_volume(fl.VOL_IMAGE)
        c.zero_volume(fl.VOL_ILLUM)
        c.rtm_backward()
        assert np.array_equal(c.get_volume(fl.VOL_IMAGE), img)
    _, _, want2 = O.forward(case.vp, case.nb, case.dh, case.dt, case.nt, case.tlag, w2, case.src, case.rec)
    assert np.array_equal(seis2, want2)


def test_model_setup_matches_oracle(fl):
    case = Case(amp=0.0)
    with case.context(fl.Context) as c:
        c.set_model_interior(case.interior)
        assert np.array_equal(c.get_volume(fl.VOL_MODEL), O.expand_boundary(case.interior, case.nb))
        c.random_boundary(11, 5, 0.6, float(case.vclamp))
        want = O.random_boundary(O.expand_boundary(case.interior, case.nb), case.nb, 11, 5, 0.6, case.vclamp)
        assert np.array_equal(c.get_volume(fl.VOL_MODEL), want)


@pytest.mark.parametrize("ratio,var_vp,dims", [(30.0, 1000.0, (40, 36, 32)), (8.0, 1500.0, (24, 20, 16))])
def test_blob_random_boundary_bit_exact(fl, ratio, var_vp, dims):
    """Reference-style boundary (noise + Poisson-disk blobs) against the
    oracle with the library's own blob list; the small ratio puts more than
    a thousand blobs near some blocks (the gathered list is flushed).  A
    second call starts afresh from the interior."""
    nx, ny, nz = dims
    case = Case(nx=nx, ny=ny, nz=nz, amp=0.0)
    with case.context(fl.Context) as c:
        c.set_model_interior(case.interior)
        c.random_boundary_blobs(7, 2, ratio, var_vp, float(case.vclamp))
        got = c.get_volume(fl.VOL_MODEL)
        c.random_boundary_blobs(7, 2, ratio, var_vp, float(case.vclamp))
        again = c.get_volume(fl.VOL_MODEL)
    blobs = fl.rbc_blobs(7, 2, case.nxx, case.nyy, case.nzz, case.nb, case.dh, ratio, var_vp)
    vlo = np.float32(case.interior.min()) - np.float32(var_vp)
    vhi = np.float32(case.interior.max()) + np.float32(var_vp)
    pad = O.expand_boundary(case.interior, case.nb)
    want = O.random_boundary_blobs(pad, case.nb, case.dh, 7, 2, var_vp, blobs, vlo, vhi, case.vclamp)
    assert len(blobs) > 100
    assert np.array_equal(got, want)
    assert np.array_equal(again, got)
    nb = case.nb
    assert np.array_equal(got[nb:-nb, nb:-nb, nb:-nb], case.interior)
    assert np.abs(got - pad).max() > 0.5 * var_vp


def test_exports_match_oracle(fl):
    case = Case()
    rng = np.random.default_rng(5)
    img = rng.normal(0, 1, (case.nyy, case.nxx, case.nzz)).astype(np.float32)
    ill = rng.uniform(0.5, 2.0, img.shape).astype(np.float32)
    with case.context(fl.Context) as c:
        c.set_volume(fl.VOL_IMAGE, img)
        c.set_volume(fl.VOL_ILLUM, ill)
        c.set_volume(fl.VOL_GRAD, img * 3)
        out = c.export_image()
        assert np.array_equal(out, O.export_seismic(img, ill, case.nb, case.dh))
        g, gmax = c.export_gradient(60.0)
        want, wmax = O.gradient_post(img * 3, ill, case.nb, case.dh, 60.0)
        assert gmax == wmax
        assert np.array_equal(g, want)


def test_residual_fp64(fl):
    case = Case(nt=80)
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(9)
    obs = (seis * 1.1 + rng.normal(0, 1, seis.shape)).astype(np.float32)
    with case.context(fl.Context) as c:
        c.forward(fl.BOUNDARY_RBC)
        s = c.residual(obs)
    res, want, _ = O.misfit(obs, seis)
    assert abs(s - want) <= 1e-12 * want


def test_errors(fl):
    case = Case()
    with case.context(fl.Context) as c:
        bad = case.vp.copy()
        bad[3, 3, 3] = np.nan
        with pytest.raises(fl.FwiError) as e:
            c.set_model(bad)
        assert e.value.code == -6
        with pytest.raises(fl.FwiError) as e:
            c.set_model(case.vp * 3)
        assert e.value.code == -3
        with pytest.raises(fl.FwiError) as e:
            c.set_source(2, 10, 10)
        assert e.value.code == -5
        c.set_model(case.vp)
        with pytest.raises(fl.FwiError) as e:
            c.rtm_backward()
        assert e.value.code == -4


def test_default_launch_shapes(fl):
    """A one-shot context starts from 64 x 32 forward and 64 x 16 backward
    tiles with the fewest z-chunks that give about one block per CU (the
    autotuner's picks at 401 x 401 x 201: (16, 2, 2) and (16, 1, 1)); a
    small grid takes more z-chunks, up to 16; a batched context keeps the
    64 x 8 tiles its shots-per-block instances take."""
    big = fl.Context(501, 501, 301, 50, 10.0, 8e-4, 10, 0)
    assert [big.tuning(k) for k in range(4)] == [(16, 2, 2), (16, 2, 2), (16, 1, 1), (16, 1, 1)]
    big.close()
    case = Case()
    small = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag)
    assert small.tuning(0) == (16, 2, 16) and small.tuning(2) == (16, 1, 16)
    small.close()
    batch = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=4)
    assert all(batch.tuning(k) == (4, 2, 1) for k in range(4))
    batch.set_shots_per_block(2, 4)  # the batched instances take the default tile
    batch.close()


@pytest.mark.slow
def test_two_layer_101_cube_seismogram(fl):
    """The 101^3 two-layer workload (nb = 50, 1000 steps, 81x81 receivers):
    full-size seismogram, bit for bit against the oracle."""
    from fwi3d_amd import configs
    w = configs.two_layer_101()
    vp = O.random_boundary(O.expand_boundary(w.interior, w.nb), w.nb, w.rbc_seed, 0, w.rbc_amp, w.vclamp)
    P, Pold, seis = O.forward(vp, w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src, w.rec)
    with fl.Context(w.nxx, w.nyy, w.nzz, w.nb, w.dh, w.dt, w.nt, w.tlag) as c:
        c.set_model_interior(w.interior)
        c.random_boundary(w.rbc_seed, 0, w.rbc_amp, w.vclamp)
        assert np.array_equal(c.get_volume(fl.VOL_MODEL), vp)
        c.set_wavelet(w.wavelet)
        c.set_source(*w.src)
        c.set_receivers(w.rec)
        c.forward(fl.BOUNDARY_RBC)
        g = c.seismogram()
    assert g.shape == (81 * 81, 1000)
    assert np.array_equal(g, seis), rel_l2(g, seis)


def test_smallest_grids_forward_and_rtm(fl):
    """Grids whose updated region is a few cells wide (no boundary layer;
    one partial tile per axis): forward and RTM bit for bit against the
    oracle."""
    rng = np.random.default_rng(11)
    for nxx, nyy, nzz in ((12, 11, 13), (9, 9, 9), (70, 10, 10)):
        vp = rng.uniform(1500, 2500, (nyy, nxx, nzz)).astype(np.float32)
        wav, tlag = O.wavelet(40, 1e-3, 45.0)
        src = (nxx // 2, nyy // 2, nzz // 2)
        X, Y = np.meshgrid(np.arange(4, nxx - 4), np.arange(4, nyy - 4), indexing="ij")
        rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, 4)], axis=1).astype(np.int32)
        P, Pold, seis = O.forward(vp, 0, 10.0, 1e-3, 40, tlag, wav, src, rec)
        data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
        want = O.backward(vp, 10.0, 1e-3, 40, data, rec, P, Pold, tlag=tlag)
        with fl.Context(nxx, nyy, nzz, 0, 10.0, 1e-3, 40, tlag) as c:
            c.set_model(vp)
            c.set_wavelet(wav)
            c.set_source(*src)
            c.set_receivers(rec)
            c.forward(fl.BOUNDARY_RBC)
            assert np.array_equal(c.seismogram(), seis), (nxx, nyy, nzz)
            c.set_backward_data(data)
            c.rtm_backward()
            assert np.array_equal(c.get_volume(fl.VOL_IMAGE), want[4]), (nxx, nyy, nzz)
            assert np.array_equal(c.get_volume(fl.VOL_ILLUM), want[5]), (nxx, nyy, nzz)


@pytest.mark.parametrize("by", [16, 8])
@pytest.mark.parametrize("mode", ["forward_rbc", "forward_abc", "rtm", "fwi"])
def test_two_plane_prefetch_bit_exact(fl, mode, by):
    """Step kernels with two planes of loads in flight (six-phase loop that
    leaves after a chunk's last plane) against the oracle bit for bit, for
    z-chunk lengths of every residue modulo 6 (39 planes in 1, 2, 4, 7 or
    8 chunks: 39; 20 + 19; 10 x 3 + 9; 6 x 6 + 3; 5 x 7 + 4)."""
    abc = mode == "forward_abc"
    case = Case(nz=31, nt=61, amp=0.0 if abc else 0.6, abc_factor=0.05 if abc else None)
    assert case.nzz - 8 == 39
    P, Pold, seis = _oracle_forward(case, abc=abc)
    kind = {"forward_rbc": 0, "forward_abc": 1, "rtm": 2, "fwi": 3}[mode]
    if kind >= 2:
        rng = np.random.default_rng(5)
        data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
        want = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold, fwi=kind == 3,
                          tlag=case.tlag)
    for zc in (1, 2, 4, 7, 8):
        with case.context(fl.Context, abc=abc) as c:
            c.set_tuning(kind, by, 1, zc)
            c.set_prefetch_depth(kind, 2)
            assert c.prefetch_depth(kind) == 2 and c.tuning(kind) == (by, 1, zc)
            if kind < 2:
                c.forward(fl.BOUNDARY_ABC if abc else fl.BOUNDARY_RBC)
                assert np.array_equal(c.seismogram(), seis), zc
                assert np.array_equal(c.get_volume(fl.VOL_FIELD), P), zc
                continue
            c.set_tuning(0, by, 1, zc)
            c.set_prefetch_depth(0, 2)
            c.forward(fl.BOUNDARY_RBC)
            c.set_backward_data(data)
            (c.grad_backward if kind == 3 else c.rtm_backward)()
            got = (c.get_volume(fl.VOL_GRAD if kind == 3 else fl.VOL_IMAGE), c.get_volume(fl.VOL_ILLUM),
                   c.get_volume(fl.VOL_FIELD), c.get_volume(fl.VOL_RECV))
        assert np.abs(want[4]).max() > 0
        for a, b in zip(got, (want[4], want[5], want[0], want[2])):
            assert np.array_equal(a, b), zc


def test_two_plane_prefetch_with_corrections_uses_one_plane_instances(fl):
    """A backward pass with corrections at prefetch depth 2 runs the
    one-plane instance of the shape: the same bits as depth 1."""
    case = Case(nt=50)
    outs = []
    for pd in (1, 2):
        with case.context(fl.Context) as c:
            c.set_physics(7)
            c.set_tuning(2, 16, 1, 2)
            c.set_prefetch_depth(2, pd)
            c.forward(fl.BOUNDARY_RBC)
            c.set_backward_data(c.seismogram())
            c.rtm_backward()
            outs.append((c.get_volume(fl.VOL_IMAGE), c.get_volume(fl.VOL_ILLUM)))
    assert np.abs(outs[0][0]).max() > 0
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


def test_volume_of_2_gib_and_more_bit_exact(fl):
    """A 1024 x 1024 x 520 padded grid (2.18 GB per volume) runs the
    per-plane-addressed step instances: forward steps [0, 24) with a seeded
    random boundary and a receiver patch, then RTM backward steps [0, 8),
    fields, seismogram rows, image and illumination against the oracle bit
    for bit (tlag 18: a 200 Hz maximum frequency at 1 ms)."""
    nxx, nyy, nzz, nb, dh, dt, nt = 1024, 1024, 520, 8, 10.0, 1e-3, 40
    assert ((nxx + 63) // 64 * 64) * nyy * nzz * 4 >= 2 ** 31
    wavelet, tlag = O.wavelet(nt, dt, 200.0)
    z = np.arange(nzz - 2 * nb) * dh
    col = np.where(z < 2000.0, 1500.0, 2500.0).astype(np.float32)
    interior = np.ascontiguousarray(np.broadcast_to(col, (nyy - 2 * nb, nxx - 2 * nb, nzz - 2 * nb)))
    vclamp = np.float32(0.45 * dh / dt)
    vp = O.random_boundary(O.expand_boundary(interior, nb), nb, 3, 0, 0.6, vclamp)
    del interior
    src = (nxx // 2, nyy // 2, nb + 2)
    xs = np.arange(nxx // 2 - 20, nxx // 2 + 21, 10)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    rec = np.stack([X.ravel(), Y.ravel(), np.full(X.size, nb + 2)], axis=1).astype(np.int32)
    nf, nbk = 24, 8
    with fl.Context(nxx, nyy, nzz, nb, dh, dt, nt, tlag) as c:
        c.set_model(vp)
        c.set_wavelet(wavelet)
        c.set_source(*src)
        c.set_receivers(rec)
        c.forward_steps(fl.BOUNDARY_RBC, 0, nf)
        g_P, g_Po, g_seis = c.get_volume(fl.VOL_FIELD), c.get_volume(fl.VOL_FIELD_OLD), c.seismogram()
        P, Pold, seis = O.forward(vp, nb, dh, dt, nt, tlag, wavelet, src, rec, n_end=nf)
        assert np.abs(seis[:, :nf - tlag]).max() > 0
        assert np.array_equal(g_P, P) and np.array_equal(g_Po, Pold)
        assert np.array_equal(g_seis[:, :nf - tlag], seis[:, :nf - tlag])
        del g_P, g_Po
        data = np.random.default_rng(2).standard_normal((len(rec), nt)).astype(np.float32)
        c.set_backward_data(data)
        for v in (fl.VOL_RECV, fl.VOL_RECV_OLD, fl.VOL_IMAGE, fl.VOL_ILLUM):
            c.zero_volume(v)
        c.backward_steps(0, 0, nbk)
        got = [c.get_volume(v) for v in (fl.VOL_FIELD, fl.VOL_RECV, fl.VOL_IMAGE, fl.VOL_ILLUM)]
    S, _, R, _, image, illum = O.backward(vp, dh, dt, nt, data, rec, P, Pold, b_begin=0, b_end=nbk, tlag=tlag)
    assert np.abs(image).max() > 0
    for name, a, b in zip(("S", "R", "image", "illum"), got, (S, R, image, illum)):
        assert np.array_equal(a, b), name

"""Two time steps per launch (temporal blocking) against the oracle and the
one-step launches (run with -m gpu).

A two-step launch computes the first step on a region 4 cells wider than
its output on every side and the second step on the output cells, with the
arithmetic of the one-step kernel; the results must equal the oracle bit for
bit in the reference arithmetic mode, and equal the one-step launches bit
for bit in the FP32 mode.  Step ranges of odd length end with a one-step
launch; the buffers rotate through four per field.
"""
import numpy as np
import pytest

import oracle as O
from tests.cases import Case

pytestmark = pytest.mark.gpu

TILES = [(16, 2), (12, 3)]


@pytest.fixture(scope="module")
def fl():
    from fwi3d_amd import lib
    lib.load()
    return lib


def _oracle_forward(case, n_end=None):
    return O.forward(case.vp, case.nb, case.dh, case.dt, case.nt, case.tlag, case.wavelet, case.src, case.rec,
                     n_end=n_end)


def _case():
    # two tiles in x (56 output columns each), three or more in y; the
    # source and the receivers share a plane
    return Case(nx=84, ny=60, nz=31, nt=61)


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("zc", [1, 3])
def test_forward_bit_exact(fl, tile, zc):
    case = _case()
    P, Pold, seis = _oracle_forward(case)
    print("two-step forward", tile, zc, case.nxx, case.nyy, case.nzz, case.steps, flush=True)
    with case.context(fl.Context) as c:
        c.set_two_step(0, True, *tile, zc)
        c.forward(fl.BOUNDARY_RBC)
        assert c.two_step(0)[0] == 1
        g_seis = c.seismogram()
        g_P = c.get_volume(fl.VOL_FIELD)
        g_Po = c.get_volume(fl.VOL_FIELD_OLD)
    assert np.abs(seis).max() > 0
    assert np.array_equal(g_seis, seis)
    assert np.array_equal(g_P, P)
    assert np.array_equal(g_Po, Pold)


def test_forward_segments_of_odd_length(fl):
    """Segments of odd and even length (a leftover one-step launch in the
    middle of a pass) compose to the oracle's pass."""
    case = _case()
    cuts = [0, 7, 8, 20, 33, case.steps]
    P7, Po7, _ = _oracle_forward(case, n_end=7)
    P, Pold, seis = _oracle_forward(case)
    print("two-step forward segments", cuts, flush=True)
    with case.context(fl.Context) as c:
        c.set_two_step(0, True, 16, 2, 2)
        for n0, n1 in zip(cuts[:-1], cuts[1:]):
            c.forward_steps(fl.BOUNDARY_RBC, n0, n1)
            if n1 == 7:
                assert np.array_equal(c.get_volume(fl.VOL_FIELD), P7)
                assert np.array_equal(c.get_volume(fl.VOL_FIELD_OLD), Po7)
        assert np.array_equal(c.seismogram(), seis)
        assert np.array_equal(c.get_volume(fl.VOL_FIELD), P)
        assert np.array_equal(c.get_volume(fl.VOL_FIELD_OLD), Pold)


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("zc", [1, 3])
def test_rtm_bit_exact(fl, tile, zc):
    case = _case()
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(3)
    data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
    S, So, R, Ro, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold,
                                            tlag=case.tlag)
    print("two-step rtm", tile, zc, flush=True)
    with case.context(fl.Context) as c:
        c.set_two_step(0, True, *tile, zc)
        c.set_two_step(2, True, *tile, zc)
        c.forward(fl.BOUNDARY_RBC)
        c.set_backward_data(data)
        c.rtm_backward()
        assert c.two_step(2)[0] == 1
        g = {k: c.get_volume(v) for k, v in (("img", fl.VOL_IMAGE), ("ill", fl.VOL_ILLUM), ("S", fl.VOL_FIELD),
                                             ("So", fl.VOL_FIELD_OLD), ("R", fl.VOL_RECV),
                                             ("Ro", fl.VOL_RECV_OLD))}
    assert np.abs(image).max() > 0
    assert np.array_equal(g["ill"], illum)
    assert np.array_equal(g["img"], image)
    assert np.array_equal(g["S"], S)
    assert np.array_equal(g["So"], So)
    assert np.array_equal(g["R"], R)
    assert np.array_equal(g["Ro"], Ro)


def test_rtm_segments_of_odd_length(fl):
    case = _case()
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(4)
    data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
    _, _, _, _, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold,
                                          tlag=case.tlag)
    cuts = [0, 3, 8, 9, 11, 30, case.steps]
    print("two-step rtm segments", cuts, flush=True)
    with case.context(fl.Context) as c:
        c.set_two_step(2, True, 16, 2, 1)
        c.forward(fl.BOUNDARY_RBC)
        c.set_backward_data(data)
        for b0, b1 in zip(cuts[:-1], cuts[1:]):
            c.backward_steps(0, b0, b1)
        assert np.array_equal(c.get_volume(fl.VOL_ILLUM), illum)
        assert np.array_equal(c.get_volume(fl.VOL_IMAGE), image)


@pytest.mark.parametrize("kind", [0, 2])
def test_fp32_mode_equals_one_step(fl, kind):
    """FP32 arithmetic: the two-step launches repeat the one-step kernel's
    operations, so both give the same bits."""
    case = _case()
    _, _, seis = _oracle_forward(case)
    outs = []
    for two in (False, True):
        print("two-step fp32", kind, two, flush=True)
        with case.context(fl.Context) as c:
            c.set_arith(fl.ARITH_FP32)
            c.set_two_step(kind, two, 16, 2, 2)
            c.forward(fl.BOUNDARY_RBC)
            if kind == 0:
                outs.append((c.seismogram(), c.get_volume(fl.VOL_FIELD)))
            else:
                c.set_backward_data(seis)
                c.rtm_backward()
                outs.append((c.get_volume(fl.VOL_IMAGE), c.get_volume(fl.VOL_ILLUM), c.get_volume(fl.VOL_RECV)))
            assert c.two_step(kind)[0] == (1 if two else 0)
    for a, b in zip(*outs):
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b)


@pytest.mark.parametrize("nplanes", [3, 5])
def test_receiver_planes(fl, nplanes):
    """Receivers on three planes (all looked up inside the launch, one of
    them the source plane) and on five (more than a launch looks up: the
    one-step launches run); both bit for bit against the oracle."""
    case = _case()
    nb = case.nb
    zs = nb + 2 + 3 * np.arange(nplanes)
    rec = np.concatenate([np.stack([case.rec[:, 0], case.rec[:, 1], np.full(len(case.rec), z)], axis=1)[k::nplanes]
                          for k, z in enumerate(zs)]).astype(np.int32)
    case.rec = rec
    P, Pold, seis = _oracle_forward(case)
    rng = np.random.default_rng(8)
    data = (seis + rng.normal(0, 1e-3, seis.shape)).astype(np.float32)
    _, _, _, _, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, data, case.rec, P, Pold,
                                          tlag=case.tlag)
    print("two-step receiver planes", nplanes, flush=True)
    with case.context(fl.Context) as c:
        c.set_two_step(0, True, 16, 2, 1)
        c.set_two_step(2, True, 16, 2, 1)
        c.forward(fl.BOUNDARY_RBC)
        assert c.two_step(0)[0] == (1 if nplanes <= 4 else -1)
        assert np.array_equal(c.seismogram(), seis)
        c.set_backward_data(data)
        c.rtm_backward()
        assert np.array_equal(c.get_volume(fl.VOL_IMAGE), image)
        assert np.array_equal(c.get_volume(fl.VOL_ILLUM), illum)


def test_corrections_fall_back(fl):
    """With a backward correction flag the RTM pass runs one-step launches
    (reported as requested but not active) and still equals the oracle."""
    case = Case(nt=50)
    P, Pold, seis = _oracle_forward(case)
    S, So, R, Ro, image, illum = O.backward(case.vp, case.dh, case.dt, case.nt, seis, case.rec, P, Pold,
                                            tlag=case.tlag, phys=2, wavelet_=case.wavelet, src=case.src)
    with case.context(fl.Context) as c:
        c.set_two_step(2, True, 16, 2, 1)
        c.set_physics(2)
        c.forward(fl.BOUNDARY_RBC)
        c.set_backward_data(seis)
        c.rtm_backward()
        assert c.two_step(2)[0] == -1
        assert np.array_equal(c.get_volume(fl.VOL_IMAGE), image)


def test_multi_shot_context_equals_one_step(fl):
    """Three shots in one context (one shot per block), one of them
    disabled: two-step forward and RTM equal the one-step launches bit for
    bit (images summed over the shots in the same order)."""
    case = Case(nx=60, ny=50, nz=28, nt=50, amp=0.0)
    nb = case.nb
    shots = [((nb + 5, nb + 6, nb + 2), case.rec[::2]), ((nb + 40, nb + 30, nb + 2), case.rec[1::2])]
    outs = []
    for two in (False, True):
        print("two-step multi-shot", two, flush=True)
        c = fl.Context(case.nxx, case.nyy, case.nzz, case.nb, case.dh, case.dt, case.nt, case.tlag, nshots=3)
        try:
            c.set_model_interior(case.interior)
            for s in range(3):
                c.shot_random_boundary(s, 99, s, 0.6, float(case.vclamp))
            for s, (src, rec) in enumerate(shots):
                c.set_source(*src, slot=s)
                c.set_receivers(rec, slot=s)
            c.disable_slot(2)
            c.set_wavelet(case.wavelet)
            for kind in (0, 2):
                c.set_two_step(kind, two, 16, 2, 2)
            c.forward(fl.BOUNDARY_RBC)
            seis = [c.seismogram(slot=s) for s in range(2)]
            for s in range(2):
                c.set_backward_data(seis[s], slot=s)
            c.rtm_backward()
            outs.append(seis + [c.get_volume(fl.VOL_IMAGE), c.get_volume(fl.VOL_ILLUM)])
        finally:
            c.close()
    for a, b in zip(*outs):
        assert np.abs(a).max() > 0
        assert np.array_equal(a, b)


def test_graph_replay(fl):
    """Two-step passes captured in hipGraphs and replayed equal the oracle."""
    case = _case()
    P, Pold, seis = _oracle_forward(case)
    _, _, _, _, image, _ = O.backward(case.vp, case.dh, case.dt, case.nt, seis, case.rec, P, Pold, tlag=case.tlag)
    with case.context(fl.Context) as c:
        c.set_two_step(0, True, 16, 2, 1)
        c.set_two_step(2, True, 16, 2, 1)
        c.set_graph(True)
        for _ in range(2):
            c.forward(fl.BOUNDARY_RBC)
            assert np.array_equal(c.seismogram(), seis)
            c.set_backward_data(seis)
            c.zero_volume(fl.VOL_IMAGE)
            c.zero_volume(fl.VOL_ILLUM)
            c.rtm_backward()
            assert np.array_equal(c.get_volume(fl.VOL_IMAGE), image)
        assert c.graph_instantiations() == 2


def test_nonzero_rims_run_one_step(fl):
    """A field state with non-zero values in its 4-cell rim (which no launch
    updates) switches the pass to one-step launches, whose buffers keep
    those values; a new forward pass clears the fields and the two-step
    launches resume."""
    case = _case()
    with case.context(fl.Context) as c:
        c.forward_steps(fl.BOUNDARY_RBC, 0, 20)
        P = c.get_volume(fl.VOL_FIELD)
        Po = c.get_volume(fl.VOL_FIELD_OLD)
    P[0, :, :] = 1.0
    Po[:, :, -1] = -0.5
    outs = []
    for two in (False, True):
        with case.context(fl.Context) as c:
            c.set_two_step(0, two, 16, 2, 1)
            c.set_volume(fl.VOL_FIELD, P)
            c.set_volume(fl.VOL_FIELD_OLD, Po)
            c.forward_steps(fl.BOUNDARY_RBC, 20, 40)
            if two:
                assert c.two_step(0)[0] == -1
            outs.append((c.get_volume(fl.VOL_FIELD), c.get_volume(fl.VOL_FIELD_OLD)))
            if two:
                c.forward(fl.BOUNDARY_RBC)
                assert c.two_step(0)[0] == 1
                _, _, seis = _oracle_forward(case)
                assert np.array_equal(c.seismogram(), seis)
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert outs[0][0][0, 5, 5] == 1.0

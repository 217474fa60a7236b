"""Host-side blob list of the reference-style random boundary (no GPU).

The list (fwi_rbc_blobs) must be a Poisson-disk sample of the padded box
(no two centres closer than `ratio`; every probe point of the boundary shell
within 2 ratio of a centre, Bridson's covering guarantee), with the inner
box removed, and the per-blob draws in their stated ranges; it is
deterministic for (seed, shot)."""
import numpy as np
import pytest

NXX, NYY, NZZ, NB, DH, RATIO, VAR = 41, 37, 33, 10, 10.0, 40.0, 500.0


@pytest.fixture(scope="module")
def blobs():
    from fwi3d_amd import lib
    return lib.rbc_blobs(11, 3, NXX, NYY, NZZ, NB, DH, RATIO, VAR)


def test_blob_centres_are_a_poisson_disk_sample(blobs):
    c = blobs[:, :3].astype(np.float64)
    assert len(c) > 50
    d2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    np.fill_diagonal(d2, np.inf)
    assert np.sqrt(d2.min()) >= RATIO * (1 - 1e-6)
    top = np.array([(NXX - 1) * DH, (NYY - 1) * DH, (NZZ - 1) * DH])
    assert np.all(c >= 0) and np.all(c < top)
    half = 0.5 * NB * DH
    inner = np.all((c > half) & (c < top - half), axis=1)
    assert not inner.any()
    # covering: probes in the boundary shell (away from the inner box by one
    # ratio) lie within 2 ratio of some centre
    rng = np.random.default_rng(0)
    p = rng.uniform(0, 1, (4000, 3)) * top
    shell = ~np.all((p > half - RATIO) & (p < top - half + RATIO), axis=1)
    p = p[shell]
    near = np.sqrt(((p[:, None, :] - c[None, :, :]) ** 2).sum(-1)).min(1)
    assert len(p) > 500 and np.all(near < 2 * RATIO)


def test_blob_draws_in_range_and_seeded(blobs):
    from fwi3d_amd import lib
    r, a = blobs[:, 3], blobs[:, 4]
    assert np.all(r >= 0.2 * RATIO) and np.all(r < RATIO)
    assert np.all(np.abs(a) >= 0.5 * VAR) and np.all(np.abs(a) < VAR)
    neg = np.mean(a < 0)
    assert 0.2 < neg < 0.47, neg  # sign -1 with probability 1/3
    again = lib.rbc_blobs(11, 3, NXX, NYY, NZZ, NB, DH, RATIO, VAR)
    other = lib.rbc_blobs(11, 4, NXX, NYY, NZZ, NB, DH, RATIO, VAR)
    assert np.array_equal(again, blobs)
    assert not (other.shape == blobs.shape and np.array_equal(other, blobs))

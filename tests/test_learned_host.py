"""Host-side pieces of the learned preconditioner (no GPU)."""
import numpy as np

from fwi3d_amd import fwi, learned


def test_random_model_ranges_and_seeding():
    a = learned.random_model(7, 40, 36, 30)
    b = learned.random_model(7, 40, 36, 30)
    c = learned.random_model(8, 40, 36, 30)
    assert a.shape == (36, 40, 30) and a.dtype == np.float32
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 1500.0 and a.max() <= 4500.0
    # velocity increases with depth on average (layer velocities are sorted)
    assert a[:, :, -5:].mean() > a[:, :, :5].mean()


def test_smooth_keeps_range_and_smooths():
    v = learned.random_model(3, 40, 36, 30)
    s = learned.smooth(v, 3.0)
    assert s.min() >= v.min() - 1e-3 and s.max() <= v.max() + 1e-3
    assert np.abs(np.diff(s, axis=2)).max() < np.abs(np.diff(v, axis=2)).max()


def test_features_and_padding():
    rng = np.random.default_rng(0)
    g = rng.uniform(-1, 1, (6, 7, 9)).astype(np.float32)
    il = rng.uniform(0, 5, g.shape).astype(np.float32)
    v = rng.uniform(1500, 4500, g.shape).astype(np.float32)
    x = learned.features(g, il, v)
    assert x.shape == (3, 6, 7, 9)
    assert np.array_equal(x[0], g)
    assert np.all(x[1] <= 1.0 + 1e-4) and np.all(x[1] >= 0.0 - 1e-6)
    assert np.allclose(x[2], (v - 3000) / 1000, atol=1e-6)
    p, sl = learned.pad4(x)
    assert p.shape == (3, 8, 8, 12) and np.array_equal(p[sl], x)


def test_parabola_step_vertex():
    for amin in (0.3, 0.7, 1.4):
        f = [(a - amin) ** 2 + 2.0 for a in (0.0, 0.5, 1.0)]
        assert abs(fwi.parabola_step(0.0, 0.5, 1.0, *f) - amin) < 1e-5


def test_b1d_table_matches_oracle():
    import oracle as O
    for nb, f in ((51, 0.004), (20, 0.004), (8, 0.05)):
        assert np.max(np.abs(fwi.b1d_table(nb, f).view(np.int32) - O.dampers(nb, f)[0].view(np.int32))) <= 1

"""The learned FWI preconditioner on the GPU (run with -m gpu).

* The misfit guard: a network whose every proposal is useless makes
  learned FWI fall back at every iteration, and the fallback is the plain
  illumination-compensated gradient step, so the run is plain FWI exactly
  (same models, same FP64 misfits, bit for bit).
* A learned update is accepted only when the plain forward model's misfit
  drops.
* What training gains: on held-out random models (never seen in training)
  the trained network's proposed update points along the true model error
  v_true - v0 (cosine > 0), while the same network before training does not.
  Measured with this set-up: trained cosines 0.26-0.64, untrained -0.21 to
  -0.15; the bounds below apply to the mean over three held-out models
  (training is deterministic: the weight gradients are summed in a fixed
  order, so a seed fixes the trained network).
"""
import numpy as np
import pytest

from fwi3d_amd import fwi, learned

pytestmark = pytest.mark.gpu

NX, NY, NZ, NT, NB, SIGMA = 48, 48, 32, 500, 20, 4.0


def _problem(seed):
    v_true = learned.random_model(seed, NX, NY, NZ)
    w = learned.survey(NX, NY, NZ, 10.0, NT, 1e-3, 30.0, nb=NB, seed=seed)
    w.interior = v_true
    obs = learned.observed(w, v_true)
    return v_true, learned.smooth(v_true, SIGMA), fwi.FwiProblem(w, obs, zmask=20.0)


def _cos(a, b):
    return float(np.sum(a * b) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-30))


def test_useless_network_falls_back_to_plain_fwi():
    from fwi3d_amd.unet import UNet
    v_true, v0, prob = _problem(4000)
    try:
        with UNet(seed=3) as net:
            p = np.zeros(net.nparams, np.float32)
            p[-1] = 5.0  # head bias only: every proposal is +5000 m/s everywhere
            net.params = p
            _, plain, _ = fwi.plain_fwi(prob, v0, 2)
            _, hist, _, used = learned.learned_fwi(prob, net, v0, max_iterations=2)
    finally:
        prob.close()
    assert used == ["gradient", "gradient"]
    assert hist == plain


def test_training_aligns_updates_with_the_model_error():
    from fwi3d_amd.unet import UNet
    triples = learned.make_triples(12, NX, NY, NZ, 2, NT, dt=1e-3, fmax=30.0, seed0=1000, sigma=SIGMA, nb=NB)
    cos = {"untrained": [], "trained": []}
    with UNet(seed=7) as untrained, UNet(seed=7) as net:
        learned.train(net, triples, 300, patch=(16, 16, 16), batch=4, lr=1e-3, seed=0)
        for seed in (5000, 5001, 5002):
            v_true, v0, prob = _problem(seed)
            try:
                f, g, gmax, il = prob.gradient(v0)
                for name, n in (("untrained", untrained), ("trained", net)):
                    cos[name].append(_cos(learned.predict(n, g, il, v0), v_true - v0))
                if seed == 5000:
                    _, hist, _, used = learned.learned_fwi(prob, net, v0, max_iterations=3)
                    assert len(hist) == len(used) + 1
                    for k, u in enumerate(used):
                        if u == "learned":
                            assert hist[k + 1] < hist[k]
            finally:
                prob.close()
    assert np.mean(cos["trained"]) > 0.15, cos
    assert np.mean(cos["trained"]) > np.mean(cos["untrained"]) + 0.2, cos


def test_learned_fwi_reaches_plain_misfit_sooner():
    """What the trained network gains over the method without it, on eight
    held-out models (never seen in training): the misfit after one learned
    iteration against the misfits plain FWI and its guarded variant (steps
    kept only when the misfit drops) reach after two gradient iterations.
    Training is deterministic, so the comparison is reproducible; measured
    with this set-up, the learned iteration is lower than both on five of the
    eight models (5000, 5001, 5002, 5004, 5005; on 5003 no method lowers the
    misfit, on 5006 the guarded method wins, on 5007 the first learned
    proposal is rejected and the fallback step is taken)."""
    from fwi3d_amd.unet import UNet
    triples = learned.make_triples(12, NX, NY, NZ, 2, NT, dt=1e-3, fmax=30.0, seed0=1000, sigma=SIGMA, nb=NB)
    wins, ratios = 0, []
    with UNet(seed=7) as net:
        learned.train(net, triples, 300, patch=(16, 16, 16), batch=4, lr=1e-3, seed=0)
        for seed in range(5000, 5008):
            v_true, v0, prob = _problem(seed)
            try:
                _, plain, _ = fwi.plain_fwi(prob, v0, 2)
                _, guarded, _, _ = fwi.guarded_fwi(prob, v0, 2)
                _, hist, _, used = learned.learned_fwi(prob, net, v0, max_iterations=1)
            finally:
                prob.close()
            best = min(plain[-1], guarded[-1])
            print("held-out %d: plain %s, guarded %s, learned %s %s" % (seed, plain, guarded, hist, used))
            if used[0] == "learned":
                assert hist[1] < hist[0]  # an accepted learned update lowers the misfit
            wins += hist[1] < best
            ratios.append(hist[1] / best)
    assert wins >= 4, (wins, ratios)
    assert np.median(ratios) < 1.0, ratios

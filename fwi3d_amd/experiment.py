"""Plain FWI against FWI with the learned preconditioner: time to equal misfit.

The two FWI workloads of the path, at a size one GPU runs in minutes:

* plain FWI: 16 shots on a 4 x 4 surface grid over the dipping-layer /
  Gaussian-anomaly velocity model (the 401 x 401 x 201 configuration's
  model family, here on a smaller grid), initial model = 3-D Gaussian
  smoothing of the true model, `iterations` gradient steps with the
  parabolic line search;
* learned FWI: the U-Net preconditioner, trained on (gradient,
  illumination, model) -> (true - current model) triples that plain FWI
  generates on seeded random layered models, proposes each update; the
  update is accepted only when the plain forward model's misfit drops, and
  the iteration falls back to the plain gradient step otherwise.  It runs
  until its misfit reaches plain FWI's final misfit.

Both use the same kernels and the same batched contexts; wall times count
every forward/backward pass, line search, network inference and fallback.

    python -m fwi3d_amd.experiment --scale small --out results.json

The "full" scale is the 401 x 401 x 201 configuration itself (10 m, 50-cell
boundary, 3000 samples at 0.8 ms, 60 Hz, initial model smoothed with
sigma = 150 m); its network is trained on 201 x 201 x 101 models with the
same dt and frequency.  It runs in two phases so that each fits one GPU
session: `--phase train --save-net net.npy`, then `--phase fwi --load-net
net.npy`.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import fwi, learned

SCALES = {
    # grid (nx, ny, nz), boundary, dt, nt, fmax, training models, FWI
    # iterations per training model, training steps, patch (y, x, z)
    "tiny": dict(grid=(48, 48, 32), nb=20, dt=1e-3, nt=500, fmax=30.0, n_train=6, train_iters=2, steps=150,
                 patch=(16, 16, 16), sigma=4.0),
    "small": dict(grid=(96, 96, 48), nb=20, dt=1e-3, nt=800, fmax=30.0, n_train=32, train_iters=3, steps=600,
                  patch=(32, 32, 32), sigma=8.0),
    "medium": dict(grid=(160, 160, 80), nb=30, dt=1e-3, nt=1100, fmax=30.0, n_train=48, train_iters=3, steps=900,
                   patch=(32, 32, 32), sigma=12.0),
    "full": dict(grid=(401, 401, 201), nb=50, dt=8e-4, nt=3000, fmax=60.0, n_train=32, train_iters=3, steps=900,
                 patch=(32, 32, 32), sigma=15.0, train_grid=(201, 201, 101), train_nb=30, train_nt=1500,
                 max_learned=6),
}


def true_model(nx, ny, nz, dh=10.0, seed=1):
    """The dipping-layer / Gaussian-anomaly model family of the 401 x 401 x
    201 workload, scaled to this grid: linear trend 1500 -> 3500 m/s over the
    depth, 5 seeded dipping layers of +-300 m/s, a +400 m/s Gaussian anomaly
    (sigma = 7.5 % of the width) below the centre.  (ny, nx, nz) float32."""
    rng = np.random.default_rng(seed)
    x, y, z = np.arange(nx) * dh, np.arange(ny) * dh, np.arange(nz) * dh
    Y, X, Z = np.meshgrid(y, x, z, indexing="ij")
    lz = z[-1]
    v = 1500.0 + 2000.0 * Z / lz
    for l, z0 in enumerate(np.linspace(0.15, 0.85, 5) * lz):
        ax, ay = rng.uniform(-0.1, 0.1, 2)
        surf = z0 + ax * (X - x.mean()) + ay * (Y - y.mean())
        v += np.where(Z >= surf, 300.0 if l % 2 == 0 else -300.0, 0.0)
    sig = 0.075 * x[-1]
    r2 = (X - x.mean()) ** 2 + (Y - y.mean()) ** 2 + (Z - 0.6 * lz) ** 2
    v += 400.0 * np.exp(-0.5 * r2 / sig ** 2)
    return np.ascontiguousarray(v.astype(np.float32))


def problem(v_true, nb, dt, nt, fmax, nshots_side=4, seed=0):
    """Workload + observed data (ABC modeling on the true model)."""
    ny, nx, nz = v_true.shape
    w = learned.survey(nx, ny, nz, 10.0, nt, dt, fmax, nshots_side=nshots_side, nb=nb, seed=seed)
    w.interior = v_true
    return w, learned.observed(w, v_true)


def run(scale="small", seed=0, log=None, phase="all", save_net="", load_net="", n_train=None, train_offset=0,
        baselines=None, which=("plain", "guarded_plain"), tune=True):
    """phase "all": train, then compare plain and learned FWI; "train": train
    only (save_net: where to store the weights, a float32 .npy; with
    load_net, training continues from those weights, so a large training set
    can be worked through in parts: n_train models starting at model
    train_offset); "fwi": the comparison with the weights of load_net.  The
    comparison can also run in two parts: "baselines" (plain and guarded
    plain FWI, no network) and "learned" (learned FWI with the weights of
    load_net, against the result dict of a "baselines" run passed as
    `baselines`); `which` names the baselines a "baselines" run computes
    (both by default; one each keeps a large run within one GPU session)."""
    p = SCALES[scale]
    nx, ny, nz = p["grid"]
    tnx, tny, tnz = p.get("train_grid", p["grid"])
    log = log or (lambda m: None)
    out = {"scale": scale, "grid": [nx, ny, nz], "nb": p["nb"], "dt": p["dt"], "nt": p["nt"], "fmax": p["fmax"],
           "shots": 16, "train_grid": [tnx, tny, tnz], "seed": seed, "tuned": tune}

    from .unet import UNet
    net = UNet(seed=seed + 7) if phase != "baselines" else None
    if load_net and net is not None:
        net.params = np.load(load_net).astype(np.float32)
        out["network"] = "loaded"
    if phase in ("all", "train"):
        # 1. training triples from plain FWI on seeded random models
        t0 = time.perf_counter()
        n = p["n_train"] if n_train is None else n_train
        out["training_models"] = [train_offset, train_offset + n]
        triples = learned.make_triples(n, tnx, tny, tnz, p["train_iters"], p.get("train_nt", p["nt"]),
                                       dt=p["dt"], fmax=p["fmax"], seed0=1000 + 1000 * seed + train_offset,
                                       sigma=p["sigma"], nb=p.get("train_nb", p["nb"]), log=log)
        out["training_triples"] = len(triples)
        out["triples_seconds"] = round(time.perf_counter() - t0, 2)

        # 2. train the network on the GPU
        t0 = time.perf_counter()
        losses = learned.train(net, triples, p["steps"], patch=p["patch"], batch=4, lr=1e-3, seed=seed, log=log)
        out["train_seconds"] = round(time.perf_counter() - t0, 2)
        out["train_loss_first_last"] = [float(np.mean(losses[:20])), float(np.mean(losses[-20:]))]
        del triples
        if save_net:
            np.save(save_net, np.asarray(net.params, np.float32))
        if phase == "train":
            net.close()
            return out

    # 3. the test problem
    v_true = true_model(nx, ny, nz, seed=seed + 1)
    v0 = learned.smooth(v_true, p["sigma"])
    w, obs = problem(v_true, p["nb"], p["dt"], p["nt"], p["fmax"], seed=seed)
    prob = fwi.FwiProblem(w, obs, zmask=20.0, tune=tune)
    prob.misfit(v0)  # warm-up: first launches, autotune caches

    def model_err(v):
        return float(np.linalg.norm(v - v_true) / np.linalg.norm(v_true))

    if phase == "learned":
        # the baselines of an earlier run on the same problem
        out["plain"], out["guarded_plain"] = baselines["plain"], baselines["guarded_plain"]
        hist_plain, sec_plain = out["plain"]["misfit"], out["plain"]["seconds"]
        hist_guard, sec_guard = out["guarded_plain"]["misfit"], out["guarded_plain"]["seconds"]
    else:
        if "plain" in which or phase != "baselines":
            v_plain, hist_plain, sec_plain = fwi.plain_fwi(prob, v0, 5, log=log)
            out["plain"] = {"misfit": hist_plain, "seconds": round(sec_plain, 3),
                            "model_rel_l2": model_err(v_plain), "initial_model_rel_l2": model_err(v0)}
        if "guarded_plain" in which or phase != "baselines":
            # plain FWI with the learned loop's acceptance rule (a step is
            # kept only when the misfit drops, else halved): the baseline
            # without the network but with the same guard
            t_guard = []
            v_guard, hist_guard, sec_guard, steps_guard = fwi.guarded_fwi(prob, v0, 5, log=log, times=t_guard)
            out["guarded_plain"] = {"misfit": hist_guard, "seconds": round(sec_guard, 3),
                                    "model_rel_l2": model_err(v_guard), "steps": steps_guard}
    if phase == "baselines":
        prob.close()
        return out
    targets = {"plain": hist_plain[-1], "guarded_plain": hist_guard[-1]}
    t_learned = []
    v_l, hist_l, sec_l, used = learned.learned_fwi(prob, net, v0, target_misfit=min(targets.values()),
                                                   max_iterations=p.get("max_learned", 10), log=log,
                                                   times=t_learned)
    out["learned"] = {"misfit": hist_l, "seconds": round(sec_l, 3), "model_rel_l2": model_err(v_l),
                      "updates": used, "elapsed_at_misfit": [round(t, 3) for t in t_learned]}

    def time_to(hist, times, target):
        for f, t in zip(hist, times):
            if f <= target:
                return t
        return None

    ratios = {}
    for name, target in targets.items():
        t = time_to(hist_l, t_learned, target)
        base = sec_plain if name == "plain" else sec_guard
        ratios[name] = round(t / base, 4) if t is not None else None
        out["learned"]["seconds_to_%s_misfit" % name] = round(t, 3) if t is not None else None
    out["learned"]["reached_target"] = ratios["plain"] is not None
    out["time_to_misfit_ratio"] = ratios["plain"]
    out["time_to_misfit_ratio_vs_guarded_plain"] = ratios["guarded_plain"]
    prob.close()
    net.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--scale", choices=sorted(SCALES), default="small")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="")
    ap.add_argument("--phase", choices=["all", "train", "fwi", "baselines", "learned"], default="all")
    ap.add_argument("--baselines", default="",
                    help="phase learned: the JSON output of a --phase baselines run (or of several, comma-separated, "
                         "whose baselines are merged)")
    ap.add_argument("--untuned", action="store_true",
                    help="run the FWI passes with the default launch shape instead of autotuned forward shapes "
                         "and the fixed 64 x 16 backward tiles")
    ap.add_argument("--which", default="plain,guarded_plain",
                    help="phase baselines: the baselines to run (plain, guarded_plain)")
    ap.add_argument("--save-net", default="", help="phase train/all: store the trained weights (.npy)")
    ap.add_argument("--load-net", default="", help="weights to use (phase fwi) or to continue training from (.npy)")
    ap.add_argument("--n-train", type=int, default=None, help="training models (default: the scale's)")
    ap.add_argument("--train-offset", type=int, default=0, help="index of the first training model")
    a = ap.parse_args(argv)
    if a.phase in ("fwi", "learned") and not a.load_net:
        ap.error("--phase %s needs --load-net" % a.phase)
    if a.phase == "learned" and not a.baselines:
        ap.error("--phase learned needs --baselines")
    base = None
    if a.baselines:
        for path in a.baselines.split(","):
            with open(path) as f:
                part = json.loads(f.read())
            if part.get("scale") != a.scale or part.get("seed", a.seed) != a.seed:
                ap.error("--baselines is a run of another scale or seed")
            base = part if base is None else dict(base, **{k: v for k, v in part.items()
                                                            if k in ("plain", "guarded_plain")})
        if "plain" not in base or "guarded_plain" not in base:
            ap.error("--baselines must hold both the plain and the guarded plain run")
    res = run(a.scale, a.seed, log=lambda m: print(m, file=sys.stderr, flush=True), phase=a.phase,
              save_net=a.save_net, load_net=a.load_net, n_train=a.n_train, train_offset=a.train_offset,
              baselines=base, which=tuple(a.which.split(",")), tune=not a.untuned)
    s = json.dumps(res)
    print(s, flush=True)
    if a.out:
        with open(a.out, "w") as f:
            f.write(s + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""Benchmark of the FWI3D wave-propagation hot path on one MI355X.

Workload (one "step"): a single-shot forward modeling pass plus the RTM
backward pass (source-wavefield reconstruction + receiver back-propagation +
cross-correlation imaging + illumination) on the 401 x 401 x 201 interior
grid at 10 m with a 50-cell random boundary (501 x 501 x 301 padded), 3000
samples at dt = 0.8 ms (3074 steps per pass), 201 x 201 receivers.  Inputs
(model, wavelet, geometry, the data injected by the backward pass) are
resident in HBM before timing starts; data are synthetic (the seeded
gradient/dipping-layer/anomaly model of fwi3d_amd.configs).

Prints one JSON line.  `value` is stencil throughput in Gcell-updates/s over
the whole RTM shot (forward: 1 field, backward: 2 fields per cell and step).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E vendor peak
BYTES_PER_UPDATE = 16   # read cur, old, vp; write new


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(kernel_key):
    """HBM bytes per launch of one step-kernel shape, as measured by the two
    rocprofv3 counter passes (FETCH_SIZE, WRITE_SIZE, scaled by the copy
    kernel's calibration) summarised in profiles/*_hbm_traffic.txt.  A
    backward shape with deferred imaging cycles through several instances:
    every six RTM launches two without any update (_IMG=0), two with the
    image only (_IMG=3), one with the illumination only (_IMG=4) and one
    with both (_IMG=5); shapes that keep the two-launch pattern alternate
    _IMG=0 and _IMG=2.  The weighted mean is the per-launch figure.  None
    when no summary covers the shape."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_hbm_traffic.txt"))):
        found = {}
        for line in open(path):
            parts = line.split()
            if len(parts) == 3 and parts[0] == "traffic":
                found[parts[1]] = int(parts[2])
        for cycle in (((0, 2), (3, 2), (4, 1), (5, 1)), ((0, 1), (2, 1))):
            got = [(found.get(kernel_key + "_IMG=%d" % i), w) for i, w in cycle]
            if all(t for t, _ in got):
                return sum(t * w for t, w in got) // sum(w for _, w in got), os.path.relpath(path, ROOT)
        if kernel_key in found:
            return found[kernel_key], os.path.relpath(path, ROOT)
    return None, None


def cpu_baseline(w, vp_pad, seconds):
    """Oracle (OpenMP, reference arithmetic) on a bounded sample of the same
    workload: a few forward steps and a few backward (RTM) steps on the full
    padded grid, scaled to Gcell-updates/s."""
    import oracle as O
    O.build()
    threads = O.num_threads()
    cells = (w.nxx - 8) * (w.nyy - 8) * (w.nzz - 8)
    # forward steps around the source-injection window
    t0 = time.perf_counter()
    P, Pold, _ = O.forward(vp_pad, w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src, w.rec[:1], n_begin=0,
                           n_end=2, record=False)
    t_fwd2 = time.perf_counter() - t0
    nf = max(2, int(0.5 * seconds / max(t_fwd2 / 2, 1e-3)))
    t0 = time.perf_counter()
    P, Pold, _ = O.forward(vp_pad, w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src, w.rec[:1], n_begin=2,
                           n_end=2 + nf, P=P, Pold=Pold, record=False)
    t_fwd = time.perf_counter() - t0
    data = np.zeros((1, w.nt), np.float32)
    nbk = max(1, nf // 2)
    t0 = time.perf_counter()
    O.backward(vp_pad, w.dh, w.dt, w.nt, data, w.rec[:1], P, Pold, b_begin=0, b_end=nbk, tlag=w.tlag)
    t_bwd = time.perf_counter() - t0
    updates = cells * (nf + 2 * nbk)
    rate = updates / (t_fwd + t_bwd) / 1e9
    # the per-run / per-shot host steps the reference runs in host loops:
    # image export (migration.cu:90-117) and residual + misfit of one shot
    # (inversion.cu:190-201), on this workload's sizes
    rng = np.random.default_rng(0)
    img = rng.normal(0, 1, vp_pad.shape).astype(np.float32)
    ill = rng.uniform(0.5, 2.0, vp_pad.shape).astype(np.float32)
    t0 = time.perf_counter()
    O.export_seismic(img, ill, w.nb, w.dh)
    t_exp = time.perf_counter() - t0
    obs = rng.normal(0, 1, (len(w.rec), w.nt)).astype(np.float32)
    cal = rng.normal(0, 1, obs.shape).astype(np.float32)
    t0 = time.perf_counter()
    O.misfit(obs, cal)
    t_mis = time.perf_counter() - t0
    return {"value": round(rate, 4), "unit": "Gcell-upd/s", "cores": threads, "kind": "port",
            "sample": "%d forward + %d RTM backward steps on the full 501x501x301 padded grid (%.1f s)"
                      % (nf, nbk, t_fwd + t_bwd),
            "export_image_ms": round(t_exp * 1e3, 1), "residual_misfit_ms": round(t_mis * 1e3, 1)}


def cpu_vs_gpu_two_layer(lib, configs):
    """The reference's own CPU-sized case -- the 101^3 two-layer forward
    modeling (nb = 50, 1000 samples, 81 x 81 receivers) -- run in full on the
    GPU and by the oracle on the host: wall times and a bit-for-bit check of
    the 1000 x 6561 seismogram."""
    import oracle as O
    w = configs.two_layer_101()
    vp = O.random_boundary(O.expand_boundary(w.interior, w.nb), w.nb, w.rbc_seed, 0, w.rbc_amp, w.vclamp)
    c = lib.Context(w.nxx, w.nyy, w.nzz, w.nb, w.dh, w.dt, w.nt, w.tlag)
    c.set_model_interior(w.interior)
    c.random_boundary(w.rbc_seed, 0, w.rbc_amp, w.vclamp)
    c.set_wavelet(w.wavelet)
    c.set_source(*w.src)
    c.set_receivers(w.rec)
    c.autotune(lib.KIND_FWD_RBC, 4)
    c.forward(lib.BOUNDARY_RBC)
    c.sync()
    t0 = time.perf_counter()
    c.forward(lib.BOUNDARY_RBC)
    c.sync()
    t_gpu = time.perf_counter() - t0
    g = c.seismogram()
    c.close()
    t0 = time.perf_counter()
    _, _, seis = O.forward(vp, w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src, w.rec)
    t_cpu = time.perf_counter() - t0
    updates = w.cell_updates(1)
    return {"workload": "two-layer 101^3 forward modeling, nb=50, 1000 samples (1079 steps), 81x81 receivers",
            "gpu_s": round(t_gpu, 4), "cpu_s": round(t_cpu, 3), "speedup": round(t_cpu / t_gpu, 1),
            "gpu_gcell_upd_s": round(updates / t_gpu / 1e9, 2), "cpu_gcell_upd_s": round(updates / t_cpu / 1e9, 3),
            "seismogram_bit_exact": bool(np.array_equal(g, seis))}


def _rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def fp64_errors(lib, configs, w):
    """Error of the GPU's FP32 fields against an FP64 solution of the same
    scheme (the oracle's double instantiation: double fields, the same float
    velocity model), both arithmetic modes:
      - the full two-layer 101^3 forward modeling (1000 x 6561 seismogram);
      - the 401x401x201 workload: forward steps [0, 160) then RTM steps
        [0, 64) (fields, image, illumination);
      - a whole 3000-sample pass (3074 steps of forward, then 3074 of
        time-reversed reconstruction and imaging) with this workload's dt,
        wavelet and model family on a 101 x 101 x 51 grid with a 20-cell
        boundary: seismogram, image and illumination after the full pass."""
    import oracle as O
    out = {}
    arith = {"reference": lib.ARITH_REFERENCE, "fp32": lib.ARITH_FP32}
    t0 = time.perf_counter()
    # two-layer 101^3, full run
    w0 = configs.two_layer_101()
    vp0 = O.random_boundary(O.expand_boundary(w0.interior, w0.nb), w0.nb, w0.rbc_seed, 0, w0.rbc_amp, w0.vclamp)
    _, _, s64 = O.forward(vp0, w0.nb, w0.dh, w0.dt, w0.nt, w0.tlag, w0.wavelet, w0.src, w0.rec, fp64=True)
    r = {}
    for name, a in arith.items():
        c = lib.Context(w0.nxx, w0.nyy, w0.nzz, w0.nb, w0.dh, w0.dt, w0.nt, w0.tlag)
        c.set_arith(a)
        c.set_model_interior(w0.interior)
        c.random_boundary(w0.rbc_seed, 0, w0.rbc_amp, w0.vclamp)
        c.set_wavelet(w0.wavelet)
        c.set_source(*w0.src)
        c.set_receivers(w0.rec)
        c.forward(lib.BOUNDARY_RBC)
        r[name] = _rel_l2(c.seismogram(), s64)
        c.close()
    out["two_layer_101_seismogram"] = r
    del s64
    # full-size workload: forward [0, 160) and RTM [0, 64)
    nf, nbk = 160, 64
    vp = O.random_boundary(O.expand_boundary(w.interior, w.nb), w.nb, w.rbc_seed, 0, w.rbc_amp, w.vclamp)
    P, Pold, seis = O.forward(vp, w.nb, w.dh, w.dt, w.nt, w.tlag, w.wavelet, w.src, w.rec, n_end=nf, fp64=True)
    rng = np.random.default_rng(1)
    data = rng.standard_normal((len(w.rec), w.nt), dtype=np.float32) * np.float32(1e-3)
    S, _, R, _, image, illum = O.backward(vp, w.dh, w.dt, w.nt, data, w.rec, P, Pold, b_begin=0, b_end=nbk,
                                          tlag=w.tlag, fp64=True)
    r = {}
    for name, a in arith.items():
        c = lib.Context(w.nxx, w.nyy, w.nzz, w.nb, w.dh, w.dt, w.nt, w.tlag)
        c.set_arith(a)
        c.set_model_interior(w.interior)
        c.random_boundary(w.rbc_seed, 0, w.rbc_amp, w.vclamp)
        c.set_wavelet(w.wavelet)
        c.set_source(*w.src)
        c.set_receivers(w.rec)
        c.forward_steps(lib.BOUNDARY_RBC, 0, nf)
        e = {"forward_field": _rel_l2(c.get_volume(lib.VOL_FIELD), P),
             "seismogram": _rel_l2(c.seismogram()[:, :nf - w.tlag], seis[:, :nf - w.tlag])}
        c.set_backward_data(data)
        for v in (lib.VOL_RECV, lib.VOL_RECV_OLD, lib.VOL_IMAGE, lib.VOL_ILLUM):
            c.zero_volume(v)
        c.backward_steps(0, 0, nbk)
        e.update({"source_field": _rel_l2(c.get_volume(lib.VOL_FIELD), S),
                  "receiver_field": _rel_l2(c.get_volume(lib.VOL_RECV), R),
                  "image": _rel_l2(c.get_volume(lib.VOL_IMAGE), image),
                  "illumination": _rel_l2(c.get_volume(lib.VOL_ILLUM), illum)})
        r[name] = e
        c.close()
    out["gradient_401_forward_160_rtm_64_steps"] = r
    del P, Pold, seis, S, R, image, illum, data
    # whole pass, shrunken grid
    ws = configs.gradient_survey(nx=101, ny=101, nz=51, nb=20, nt=w.nt, full_spread=True)
    vps = O.random_boundary(O.expand_boundary(ws.interior, ws.nb), ws.nb, ws.rbc_seed, 0, ws.rbc_amp, ws.vclamp)
    src = ws.src_index(5)  # an inner shot of the 4 x 4 grid
    rec = ws.rec_index(5)
    P, Pold, s64 = O.forward(vps, ws.nb, ws.dh, ws.dt, ws.nt, ws.tlag, ws.wavelet, src, rec, fp64=True)
    d32 = s64.astype(np.float32)
    _, _, _, _, image, illum = O.backward(vps, ws.dh, ws.dt, ws.nt, d32, rec, P, Pold, tlag=ws.tlag, fp64=True)
    r = {}
    for name, a in arith.items():
        c = lib.Context(ws.nxx, ws.nyy, ws.nzz, ws.nb, ws.dh, ws.dt, ws.nt, ws.tlag)
        c.set_arith(a)
        c.set_model_interior(ws.interior)
        c.random_boundary(ws.rbc_seed, 0, ws.rbc_amp, ws.vclamp)
        c.set_wavelet(ws.wavelet)
        c.set_source(*src)
        c.set_receivers(rec)
        c.forward(lib.BOUNDARY_RBC)
        e = {"seismogram": _rel_l2(c.seismogram(), s64)}
        c.set_backward_data(d32)
        c.rtm_backward()
        e.update({"image": _rel_l2(c.get_volume(lib.VOL_IMAGE), image),
                  "illumination": _rel_l2(c.get_volume(lib.VOL_ILLUM), illum)})
        r[name] = e
        c.close()
    out["gradient_101x101x51_full_pass"] = r
    out["note"] = ("relative L2 of the GPU FP32 result against the FP64 oracle; the backward data of the "
                   "full pass is the FP64 seismogram rounded to float")
    out["seconds"] = round(time.perf_counter() - t0, 1)
    log("FP64 errors: %s" % json.dumps(out))
    return out


def blob_boundary(w, ctx, lib, shots=3, ratio=100.0, var_vp=1000.0):
    """The reference-style random boundary (modeling.cu:558-673) on the
    workload grid with the reference parameters.txt's mig_rbc_ratio = 100 m
    and mig_rbc_varVp = 1000 m/s: the host Poisson-disk list alone, and the
    whole library call (list, upload, one fused launch), per shot.  The
    reference launches one full-volume kernel per blob; its HBM floor (read
    and write every padded cell once per blob at 8 TB/s) is reported beside."""
    t_list, nblobs = [], []
    for s in range(shots):
        t0 = time.perf_counter()
        b = lib.rbc_blobs(w.rbc_seed, s, w.nxx, w.nyy, w.nzz, w.nb, w.dh, ratio, var_vp)
        t_list.append(time.perf_counter() - t0)
        nblobs.append(len(b))
    t_call = []
    for s in range(shots):
        ctx.set_model_interior(w.interior)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.random_boundary_blobs(w.rbc_seed, s, ratio, var_vp, w.vclamp)
        ctx.sync()
        t_call.append(time.perf_counter() - t0)
    floor = float(np.mean(nblobs)) * 8.0 * w.nxx * w.nyy * w.nzz / (HBM_PEAK_GBS * 1e9)
    log("blob boundary: %d blobs, list %.0f ms, call %.0f ms" % (nblobs[0], 1e3 * np.mean(t_list),
                                                                  1e3 * np.mean(t_call)))
    return {"ratio_m": ratio, "var_vp": var_vp, "blobs": nblobs, "host_list_ms": round(1e3 * float(np.mean(t_list)), 1),
            "call_ms": round(1e3 * float(np.mean(t_call)), 1),
            "kernel_ms_approx": round(1e3 * float(np.mean(t_call) - np.mean(t_list)), 1),
            "reference_launch_pattern_hbm_floor_ms": round(1e3 * floor, 1)}


def corrections_cost(ctx, lib, steps=198):
    """RTM backward launch time with the reference's physics, with source
    subtraction and d2/dy2 (same imaging cadence) and with all three
    optional corrections (aligned imaging updates image and illumination
    every other launch), over the same steps of the single-shot context."""
    out = {}
    for name, phys in (("reference", 0), ("subtraction_dyy", 3), ("all_corrections", 7)):
        ctx.set_physics(phys)
        ctx.backward_steps(0, 1000, 1000 + steps)
        ms, n = ctx.last_kernel_timing()
        out["rtm_launch_us_" + name] = round(ms / max(n, 1) * 1e3, 1)
    ctx.set_physics(0)
    log("RTM launch: %.1f us (reference physics), %.1f us (subtraction + d2/dy2), %.1f us (all corrections)"
        % (out["rtm_launch_us_reference"], out["rtm_launch_us_subtraction_dyy"], out["rtm_launch_us_all_corrections"]))
    return out


def abc_forward_cost(ctx, w, lib, steps=200):
    """Forward launch time with the Cerjan absorbing boundary (the FWI
    passes that compute calculated data and line-search misfits,
    inversion.cu:178-188,315-347) against the random-boundary forward, over
    the same steps of the single-shot context (tuned for the latter)."""
    from fwi3d_amd import fwi
    ctx.set_dampers(fwi.b1d_table(w.nb, 0.004))
    ctx.autotune(lib.KIND_FWD_ABC, 4)
    out = {"abc_tuning": list(ctx.tuning(lib.KIND_FWD_ABC))}
    for name, boundary in (("rbc", lib.BOUNDARY_RBC), ("abc", lib.BOUNDARY_ABC)):
        ctx.forward_steps(boundary, 0, steps)
        ms, n = ctx.last_kernel_timing()
        out["forward_launch_us_" + name] = round(ms / max(n, 1) * 1e3, 1)
    log("forward launch: %.1f us (random boundary), %.1f us (absorbing boundary)"
        % (out["forward_launch_us_rbc"], out["forward_launch_us_abc"]))
    return out


def post_processing(ctx, w, reps=3):
    """Whole-call times (PCIe transfers included) of the per-run and
    per-shot host-facing steps the reference does in host loops: image
    export (migration.cu:90-117; the gradient export of inversion.cu:142-169
    runs the same kernel) and the residual with its FP64 misfit
    (inversion.cu:190-201) of one shot's 3000 x 40401 seismogram."""
    obs = ctx.seismogram()
    out = {}
    for name, fn in (("export_image_ms", lambda: ctx.export_image()),
                     ("residual_misfit_ms", lambda: ctx.residual(obs))):
        fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        out[name] = round((time.perf_counter() - t0) / reps * 1e3, 1)
    log("post-processing: %s" % out)
    return out


def dataset_tools(w):
    """The C++ initial-model filter (setup_main) against scipy's
    gaussian_filter, which the reference's run/generate_setup.py:64-66 calls,
    on the workload's interior slowness (z fastest) with sigma = 3 cells (the
    script's) and 15 cells (150 m); both must agree bit for bit."""
    import ctypes
    from scipy.ndimage import gaussian_filter
    L = ctypes.CDLL(os.path.join(ROOT, "fwi3d_amd", "libfwi3d_host.so"))
    f32 = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
    L.fwi3d_host_gaussian_filter3d.argtypes = [f32, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]
    ny, nx, nz = w.interior.shape
    slow = np.ascontiguousarray(1.0 / w.interior).reshape(-1)  # x-y-z array in memory = z fastest
    out = {"grid_zxy": [nz, nx, ny], "threads": os.environ.get("OMP_NUM_THREADS", str(os.cpu_count()))}
    for sigma in (3.0, 15.0):
        t0 = time.perf_counter()
        want = gaussian_filter(slow.reshape((nz, nx, ny), order="F"), sigma).flatten("F")
        t_py = time.perf_counter() - t0
        got = slow.copy()
        t0 = time.perf_counter()
        L.fwi3d_host_gaussian_filter3d(got, nz, nx, ny, sigma)
        t_cpp = time.perf_counter() - t0
        out["sigma_%g" % sigma] = {"cpp_s": round(t_cpp, 3), "scipy_s": round(t_py, 3),
                                   "bit_exact": bool(np.array_equal(got, want))}
        log("gaussian filter sigma %g: C++ %.2f s, scipy %.2f s" % (sigma, t_cpp, t_py))
    return out


def batched_rtm(w, single, nshots, lib, configs, tuning="", pass_steps=0, graph=True, real_pre=0):
    """RTM of the first `nshots` shots of the 4x4 surface survey (each shot its
    own random boundary and its own 201x201 receiver patch clipped at the
    model edges), once as one batch and once shot after shot on the
    single-shot context; returns shots/h of both and the image difference.
    With pass_steps (profiling only) the batch runs that many steps of each
    pass on zero data and nothing else; with real_pre as well, it runs the
    whole forward pass and real_pre backward steps on the recorded data
    instead, then pass_steps backward steps on those wavefields."""
    w16 = configs.gradient_401_16(interior=w.interior)
    S = nshots
    log("batched RTM: %d shots, receivers per shot %s" % (S, [len(w16.receivers[s]) for s in range(S)]))
    c = lib.Context(w.nxx, w.nyy, w.nzz, w.nb, w.dh, w.dt, w.nt, w.tlag, nshots=S)
    c.set_model_interior(w.interior)
    c.set_wavelet(w.wavelet)
    for s in range(S):
        c.shot_random_boundary(s, w.rbc_seed, s, w.rbc_amp, w.vclamp)
        c.set_source(*w16.src_index(s), slot=s)
        c.set_receivers(w16.rec_index(s), slot=s)
    t0 = time.perf_counter()
    if tuning:
        kinds = {"fwd": lib.KIND_FWD_RBC, "rtm": lib.KIND_BWD_RTM}
        for item in tuning.split(";"):
            name, vals = item.split("=")
            vals = [int(v) for v in vals.split(",")]
            by, rpt, zc, sb = vals[:4]
            if len(vals) > 4 and vals[4] == 2:  # shot pairs: two slots per thread
                c.set_shots_per_thread(kinds[name.strip()], 2, by, zc)
            else:
                c.set_shots_per_block(kinds[name.strip()], sb)
                c.set_tuning(kinds[name.strip()], by, rpt, zc)
    else:
        c.autotune(lib.KIND_FWD_RBC, 3)
        c.autotune(lib.KIND_BWD_RTM, 3)
    log("batched tuning %.1f s: forward %s sb=%d spt=%d, rtm %s sb=%d spt=%d" % (
        time.perf_counter() - t0, c.tuning(0), c.shots_per_block(0), c.shots_per_thread(0), c.tuning(2),
        c.shots_per_block(2), c.shots_per_thread(2)))
    if pass_steps and real_pre:
        c.forward(lib.BOUNDARY_RBC)
        for s in range(S):
            c.set_backward_data(c.seismogram(slot=s), slot=s)
        c.backward_steps(0, 0, real_pre)
        c.sync()
        log("batch: forward pass and %d backward steps done" % real_pre)
        c.set_kernel_timing(True)
        c.backward_steps(0, real_pre, real_pre + pass_steps)
        ms, n = c.last_kernel_timing()
        c.close()
        us = ms / n * 1e3 / S
        log("batched RTM kernel on real wavefields, steps %d-%d: %.1f us per shot-step"
            % (real_pre, real_pre + pass_steps, us))
        return {"shots": S, "profiling_pass_steps": pass_steps, "profiling_real_pre": real_pre,
                "rtm_us_per_shot_step": round(us, 2)}
    if pass_steps:
        for s in range(S):
            c.set_backward_data(np.zeros((len(w16.receivers[s]), w.nt), np.float32), slot=s)
        c.forward_steps(lib.BOUNDARY_RBC, 0, pass_steps)
        c.backward_steps(0, 0, pass_steps)
        c.sync()
        c.close()
        return {"shots": S, "profiling_pass_steps": pass_steps}
    sb_rtm = c.shots_per_block(lib.KIND_BWD_RTM)
    shape = {"forward": list(c.tuning(0)) + [c.shots_per_block(0), c.shots_per_thread(0)],
             "rtm": list(c.tuning(2)) + [c.shots_per_block(2), c.shots_per_thread(2)]}
    c.set_graph(graph)  # each pass's time loop as one hipGraph, instantiated once and replayed
    c.forward(lib.BOUNDARY_RBC)
    data = [c.seismogram(slot=s) for s in range(S)]
    for s in range(S):
        c.set_backward_data(data[s], slot=s)
    c.rtm_backward()  # warm-up (with graphs: the backward pass's instantiation)
    c.zero_volume(lib.VOL_IMAGE)
    c.zero_volume(lib.VOL_ILLUM)
    c.sync()
    t0 = time.perf_counter()
    c.forward(lib.BOUNDARY_RBC)
    f_ms = c.last_pass_timing()[0]
    c.rtm_backward()
    b_ms = c.last_pass_timing()[0]
    c.sync()
    t_batch = time.perf_counter() - t0
    img_b = c.get_volume(lib.VOL_IMAGE)
    graphs = c.graph_instantiations()
    # the batch's step kernels alone: HIP events around every launch of a
    # 24-step segment past the end of the data (no injection or recording;
    # a multiple of six launches, so the backward segment has the whole
    # pass's mix of imaging cadences), after the timed passes
    c.set_graph(False)
    c.set_kernel_timing(True)
    kernels = {}
    for name, run in (("forward", lambda: c.forward_steps(lib.BOUNDARY_RBC, w.nt, w.nt + 24)),
                      ("rtm", lambda: c.backward_steps(0, w.nt, w.nt + 24))):
        best = None
        for _ in range(2):
            run()
            ms, n = c.last_kernel_timing()
            best = ms / n if best is None else min(best, ms / n)
        kernels[name] = best
    c.close()
    log("batched: %.2f s (forward %.0f ms, backward %.0f ms); kernels per launch: forward %.0f us, rtm %.0f us"
        % (t_batch, f_ms, b_ms, kernels["forward"] * 1e3, kernels["rtm"] * 1e3))

    # per-shot path on the single-shot context
    single.zero_volume(lib.VOL_IMAGE)
    single.zero_volume(lib.VOL_ILLUM)
    single.sync()
    t_seq = 0.0
    for s in range(S):
        # set-up (model, boundary, geometry, data upload) is outside the timing
        single.set_model_interior(w.interior)
        single.random_boundary(w.rbc_seed, s, w.rbc_amp, w.vclamp)
        single.set_source(*w16.src_index(s))
        single.set_receivers(w16.rec_index(s))
        single.set_backward_data(data[s])
        single.sync()
        t0 = time.perf_counter()
        single.forward(lib.BOUNDARY_RBC)
        single.rtm_backward()
        single.sync()
        t_seq += time.perf_counter() - t0
    img_s = single.get_volume(lib.VOL_IMAGE)
    rel = float(np.linalg.norm((img_b - img_s).astype(np.float64)) / np.linalg.norm(img_s.astype(np.float64)))
    log("per-shot: %.2f s; image rel-L2 batched vs per-shot %.3g" % (t_seq, rel))
    cells = (w.nxx - 8) * (w.nyy - 8) * (w.nzz - 8)
    # the passes' own rates at the algorithmic bytes (16 B per cell-update;
    # the backward pass's time includes the receiver injections and the
    # pending-imaging flush, so these bound the step kernels' rates below)
    fwd_gbs = 16.0 * cells * w.steps * S / (f_ms * 1e-3) / 1e9
    bwd_gbs = 32.0 * cells * w.steps * S / (b_ms * 1e-3) / 1e9
    rtm_k = 32.0 * cells * S / (kernels["rtm"] * 1e-3) / 1e9
    fwd_k = 16.0 * cells * S / (kernels["forward"] * 1e-3) / 1e9
    return {"shots": S, "graph": graph, "graph_instantiations": graphs, "sb_rtm": sb_rtm,
            "launch_shape": {"by_rpt_zchunks_sb_spt": shape},
            "roofline": {"bound": "hbm", "kernel": "batched RTM backward step (%d shots per launch)" % S,
                         "achieved": round(rtm_k, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(rtm_k / HBM_PEAK_GBS, 4),
                         "avg_launch_us": round(kernels["rtm"] * 1e3, 1),
                         "us_per_shot_step": round(kernels["rtm"] * 1e3 / S, 2),
                         "algorithmic_bytes_per_launch": 32 * cells * S},
            "forward_kernel": {"avg_launch_us": round(kernels["forward"] * 1e3, 1),
                               "us_per_shot_step": round(kernels["forward"] * 1e3 / S, 2),
                               "achieved_GBs": round(fwd_k, 1), "frac": round(fwd_k / HBM_PEAK_GBS, 4)},
            "rtm_pass_us_per_shot_step": round(b_ms * 1e3 / w.steps / S, 2),
            "rtm_pass_GBs": round(bwd_gbs, 1), "rtm_pass_frac": round(bwd_gbs / HBM_PEAK_GBS, 4),
            "forward_pass_GBs": round(fwd_gbs, 1), "forward_pass_frac": round(fwd_gbs / HBM_PEAK_GBS, 4),
            "shots_per_hour_batched": round(3600.0 * S / t_batch, 1),
            "shots_per_hour_per_shot": round(3600.0 * S / t_seq, 1),
            "speedup": round(t_seq / t_batch, 3),
            "gcell_upd_per_s_batched": round(cells * w.steps * 3 * S / t_batch / 1e9, 2),
            "image_rel_l2_vs_per_shot": rel,
            "note": "wall time of forward + RTM backward passes only; model, boundary, geometry and data "
                    "set-up are outside the timing on both paths"}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--arith", choices=["reference", "fp32"], default="reference")
    ap.add_argument("--graph", action="store_true", help="capture each pass in a hipGraph")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP32-vs-FP64 error report (CPU baseline leg)")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--pass-steps", type=int, default=0,
                    help="profiling only: cut each pass to this many time steps")
    ap.add_argument("--probe-shapes", default="",
                    help="profiling only (with --pass-steps): also run that many forward and RTM steps with each "
                         "of these single-shot shapes, e.g. '16,1,1 8,1,1' (BY, RPT, z-chunks), so that one "
                         "counter pass covers every shape the autotuner may pick")
    ap.add_argument("--copy-probe", action="store_true",
                    help="profiling only: one 1 GiB copy-kernel launch (counter calibration)")
    ap.add_argument("--tuning", default="",
                    help="fixed launch shapes instead of autotuning, e.g. 'fwd=4,2,8;rtm=16,1,1'")
    ap.add_argument("--two-step", default="",
                    help="two time steps per launch with fixed shapes, e.g. 'fwd=16,2,4;rtm=16,2,4' (BY, RPT, "
                         "z-chunks), after the one-step tuning (the autotuner already tries them)")
    ap.add_argument("--fwi-scale", default="small", choices=["none", "tiny", "small", "medium"],
                    help="also run plain FWI against learned-preconditioned FWI (time to equal misfit) at this "
                         "grid scale (fwi3d_amd.experiment)")
    ap.add_argument("--batch-tuning", default="",
                    help="fixed batched launch shapes instead of autotuning, e.g. 'fwd=8,2,16,1;rtm=16,1,1,2,2' "
                         "(BY, RPT, z-chunks, shots per block[, shots per thread: 2 = shot pairs])")
    ap.add_argument("--batch-graph", type=int, default=1, choices=[0, 1],
                    help="run the batched passes' time loops as hipGraphs instantiated once (1) or launch on the "
                         "stream (0)")
    ap.add_argument("--no-aux", action="store_true",
                    help="skip the blob-boundary and dataset-tool timings")
    ap.add_argument("--batch-real-steps", type=int, default=0,
                    help="profiling only (with --pass-steps): the batch runs its whole forward pass and this many "
                         "backward steps on the recorded data before the profiled backward steps")
    ap.add_argument("--batch-shots", type=int, default=16,
                    help="also time this many shots of the 4x4 surface survey as one batch against the "
                         "per-shot path (0: skip; 16: the full survey)")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one MI355X: multi-GPU shot sharding is not part of this build "
              "(run with --gpus 1)", file=sys.stderr)
        return 2

    from fwi3d_amd import configs, lib

    log("building workload gradient_401 ...")
    w = configs.gradient_401()
    log("grid %dx%dx%d padded, nb=%d, nt=%d, tlag=%d, nrec=%d" % (w.nxx, w.nyy, w.nzz, w.nb, w.nt, w.tlag,
                                                                  len(w.rec)))
    ctx = lib.Context(w.nxx, w.nyy, w.nzz, w.nb, w.dh, w.dt, w.nt, w.tlag)
    ctx.set_arith(lib.ARITH_FP32 if args.arith == "fp32" else lib.ARITH_REFERENCE)
    ctx.set_model_interior(w.interior)
    ctx.random_boundary(w.rbc_seed, 0, w.rbc_amp, w.vclamp)
    ctx.set_wavelet(w.wavelet)
    ctx.set_source(*w.src)
    ctx.set_receivers(w.rec)
    if args.tuning:
        kinds = {"fwd": lib.KIND_FWD_RBC, "rtm": lib.KIND_BWD_RTM}
        for item in args.tuning.split(";"):
            name, vals = item.split("=")
            ctx.set_tuning(kinds[name.strip()], *[int(v) for v in vals.split(",")])
        log("fixed tuning: forward %s, rtm %s" % (ctx.tuning(0), ctx.tuning(2)))
    elif not args.no_autotune:
        t0 = time.perf_counter()
        ctx.autotune(lib.KIND_FWD_RBC, 4)
        ctx.autotune(lib.KIND_BWD_RTM, 4)
        log("autotune %.1f s: forward %s, rtm %s" % (time.perf_counter() - t0, ctx.tuning(0), ctx.tuning(2)))
    if args.two_step:
        kinds = {"fwd": lib.KIND_FWD_RBC, "rtm": lib.KIND_BWD_RTM}
        for item in args.two_step.split(";"):
            name, vals = item.split("=")
            ctx.set_two_step(kinds[name.strip()], True, *[int(v) for v in vals.split(",")])
    two_step = {k: ctx.two_step(kind) for k, kind in (("forward", lib.KIND_FWD_RBC), ("rtm", lib.KIND_BWD_RTM))}
    log("two-step launches (active, BY, RPT, z-chunks): %s" % two_step)
    tuning = {"forward": list(ctx.tuning(lib.KIND_FWD_RBC)), "rtm": list(ctx.tuning(lib.KIND_BWD_RTM))}
    ctx.set_graph(args.graph)
    ctx.set_kernel_timing(not args.graph)

    if args.pass_steps:
        # profiling: the injected data's values do not change the traffic
        ctx.set_backward_data(np.zeros((ctx.nrec, w.nt), np.float32))
    else:
        # synthetic observed data: the forward pass's own recording
        ctx.forward(lib.BOUNDARY_RBC)
        ctx.set_backward_data(ctx.seismogram())
    log("data ready")

    def step():
        if args.pass_steps:
            ctx.forward_steps(lib.BOUNDARY_RBC, 0, args.pass_steps)
        else:
            ctx.forward(lib.BOUNDARY_RBC)
        f = ctx.last_pass_timing() + ctx.last_kernel_timing()
        if args.pass_steps:
            ctx.backward_steps(0, 0, args.pass_steps)
        else:
            ctx.rtm_backward()
        b = ctx.last_pass_timing() + ctx.last_kernel_timing()
        return f, b

    copy_gbs = ctx.copy_bandwidth(1 << 30, 1 if args.copy_probe else 10)
    log("copy kernel: %.1f GB/s (read + write)" % copy_gbs)

    for i in range(args.warmup):
        step()
        log("warmup %d done" % i)
    ctx.zero_volume(lib.VOL_IMAGE)
    ctx.zero_volume(lib.VOL_ILLUM)
    ctx.sync()
    t0 = time.perf_counter()
    fwd_k = bwd_k = fwd_pass = bwd_pass = 0.0
    nlaunch_f = nlaunch_b = 0
    for i in range(args.steps):
        (fp, _, fk, fl), (bp, _, bk, bl) = step()
        fwd_pass += fp
        bwd_pass += bp
        fwd_k += fk
        bwd_k += bk
        nlaunch_f += fl
        nlaunch_b += bl
        log("step %d: forward %.1f ms, rtm backward %.1f ms" % (i, fp, bp))
    ctx.sync()
    wall = time.perf_counter() - t0

    cells = (w.nxx - 8) * (w.nyy - 8) * (w.nzz - 8)
    N = args.pass_steps or w.steps
    updates_per_step = cells * N * 3
    sec_per_step = wall / args.steps
    value = updates_per_step / sec_per_step / 1e9

    roofline = None
    extra = {}
    if nlaunch_b and nlaunch_f:
        bwd_avg_s = bwd_k / nlaunch_b / 1e3
        fwd_avg_s = fwd_k / nlaunch_f / 1e3
        # time steps per launch: 1, or about 2 with two-step launches
        bwd_spl = N * args.steps / nlaunch_b
        fwd_spl = N * args.steps / nlaunch_f
        bwd_bytes = int(round(2 * BYTES_PER_UPDATE * cells * bwd_spl))
        fwd_bytes = int(round(BYTES_PER_UPDATE * cells * fwd_spl))
        ach = bwd_bytes / bwd_avg_s / 1e9
        by, rpt, _ = ctx.tuning(lib.KIND_BWD_RTM)
        key = "step_kernel<BWD_RTM,%s>_BY=%d_RPT=%d_SB=1" % ("REF" if args.arith == "reference" else "FP32", by, rpt)
        # the counter summaries cover the one-step shapes
        traffic, traffic_src = measured_traffic(key) if two_step["rtm"][0] != 1 else (None, None)
        if traffic:
            extra["traffic_source"] = traffic_src
            extra["traffic_GBs_at_measured_bytes"] = round(traffic / bwd_avg_s / 1e9, 1)
        roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "frac_of_copy_kernel": round(ach / copy_gbs, 4),
                    "kernel": "step_kernel<BWD_RTM> (fused reconstruction + receiver propagation + imaging)",
                    "avg_launch_us": round(bwd_avg_s * 1e6, 2), "algorithmic_bytes_per_launch": bwd_bytes,
                    "steps_per_launch": round(bwd_spl, 4)}
        extra["forward_kernel"] = {"avg_launch_us": round(fwd_avg_s * 1e6, 2), "steps_per_launch": round(fwd_spl, 4),
                                   "achieved_GBs": round(fwd_bytes / fwd_avg_s / 1e9, 1),
                                   "frac": round(fwd_bytes / fwd_avg_s / 1e9 / HBM_PEAK_GBS, 4)}
    pct_hbm = value * BYTES_PER_UPDATE / HBM_PEAK_GBS
    # the timed shot's padded model (the per-shot leg below replaces its boundary)
    ctx_vp = ctx.get_volume(lib.VOL_MODEL) if not args.no_cpu_baseline else None

    if args.pass_steps and args.probe_shapes:
        for shape in args.probe_shapes.split():
            by, rpt, zc = [int(v) for v in shape.split(",")]
            for kind in (lib.KIND_FWD_RBC, lib.KIND_BWD_RTM):
                ctx.set_tuning(kind, by, rpt, zc)
            ctx.forward_steps(lib.BOUNDARY_RBC, 0, args.pass_steps)
            ctx.backward_steps(0, 0, args.pass_steps)
            log("probe shape %s done" % shape)

    aux = None
    if not args.no_aux and not args.pass_steps:
        aux = {"blob_random_boundary": blob_boundary(w, ctx, lib), "dataset_tools": dataset_tools(w)}
        if not args.graph:
            aux["backward_corrections"] = corrections_cost(ctx, lib)
            aux["absorbing_boundary"] = abc_forward_cost(ctx, w, lib)
        aux["post_processing"] = post_processing(ctx, w)

    batched = None
    if args.batch_shots > 0:
        batched = batched_rtm(w, ctx, args.batch_shots, lib, configs, args.batch_tuning, args.pass_steps,
                              bool(args.batch_graph), args.batch_real_steps)

    fwi_res = None
    if args.fwi_scale != "none" and not args.pass_steps:
        from fwi3d_amd import experiment
        log("FWI: plain vs learned preconditioner, scale %s ..." % args.fwi_scale)
        r = experiment.run(args.fwi_scale, 0, log=log)
        fwi_res = {"scale": r["scale"], "grid": r["grid"], "shots": r["shots"], "nt": r["nt"],
                   "time_to_misfit_ratio": r["time_to_misfit_ratio"],
                   "time_to_misfit_ratio_vs_guarded_plain": r["time_to_misfit_ratio_vs_guarded_plain"],
                   "stand_in_scale": "the 401x401x201 problem and 200 training models of the configuration are "
                                     "measured by `python -m fwi3d_amd.experiment --scale full` (minutes of GPU "
                                     "time); this leg runs the same procedure on the %s grid" % r["scale"],
                   "plain_5_iterations_s": r["plain"]["seconds"],
                   "guarded_plain_5_iterations_s": r["guarded_plain"]["seconds"],
                   "learned_s": r["learned"]["seconds"],
                   "learned_to_plain_misfit_s": r["learned"]["seconds_to_plain_misfit"],
                   "learned_to_guarded_plain_misfit_s": r["learned"]["seconds_to_guarded_plain_misfit"],
                   "plain_misfit": [round(v, 2) for v in r["plain"]["misfit"]],
                   "guarded_plain_misfit": [round(v, 2) for v in r["guarded_plain"]["misfit"]],
                   "learned_misfit": [round(v, 2) for v in r["learned"]["misfit"]],
                   "learned_updates": r["learned"]["updates"],
                   "model_rel_l2": {"initial": round(r["plain"]["initial_model_rel_l2"], 4),
                                    "plain": round(r["plain"]["model_rel_l2"], 4),
                                    "guarded_plain": round(r["guarded_plain"]["model_rel_l2"], 4),
                                    "learned": round(r["learned"]["model_rel_l2"], 4)},
                   "training": {"triples": r["training_triples"], "triples_s": r["triples_seconds"],
                                "train_s": r["train_seconds"], "loss_first_last": r["train_loss_first_last"]},
                   "note": "wall time of each FWI run (every forward/backward pass, line search, network inference "
                           "and fallback); training data generation and training are reported separately"}

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        import oracle as O
        vp_pad = O.random_boundary(O.expand_boundary(w.interior, w.nb), w.nb, w.rbc_seed, 0, w.rbc_amp, w.vclamp)
        if not np.array_equal(ctx_vp, vp_pad):
            log("warning: device model differs from the oracle's")
        ctx.close()
        cpu = cpu_baseline(w, vp_pad, args.cpu_seconds)
        log("two-layer 101^3 forward modeling: GPU and CPU oracle ...")
        cpu["two_layer_101_full_run"] = cpu_vs_gpu_two_layer(lib, configs)
        if not args.no_fp64:
            log("FP32 error against the FP64 oracle ...")
            cpu["fp64_error"] = fp64_errors(lib, configs, w)
    else:
        ctx.close()

    out = {
        "metric": "RTM shots/h batched vs per-shot; stencil Gcell-upd/s %HBM; FWI time-to-misfit",
        "value": round(value, 3),
        "unit": "Gcell-upd/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_per_step * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (per-axis FP64 derivative)" if args.arith == "reference" else "f32",
        "data": "synthetic",
        "config": {"workload": "single-shot forward + RTM, 401x401x201 @10 m, nb=50 random boundary, "
                               "nt=3000, dt=0.8 ms, 201x201 receivers",
                   "padded_grid": [w.nxx, w.nyy, w.nzz], "steps_per_pass": N, "arith": args.arith,
                   "graph": args.graph,
                   "tuning": tuning,
                   "two_step": two_step},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "rtm_shots_per_hour_per_shot_path": round(3600.0 / sec_per_step, 1),
        "pct_hbm_at_16B_per_update": round(pct_hbm, 4),
        "forward_pass_ms": round(fwd_pass / args.steps, 2),
        "rtm_backward_pass_ms": round(bwd_pass / args.steps, 2),
        "batched_rtm": batched,
        "widened_rows": aux,
        "fwi_time_to_misfit": fwi_res,
        "copy_kernel_GBs": round(copy_gbs, 1),
    }
    out.update(extra)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

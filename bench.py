#!/usr/bin/env python3
"""Benchmark: particle-filter ms/cycle at 1M particles (config 1) plus CBGL time-to-fix.

Headline workload (one GPU): 1,000,000 particles, 256 landmarks in a 50 x 50 x 10 m box,
256 range-bearing observations per cycle (1 % range outliers U[0, 60] m), helix motion,
inject ratio 0.01, gate 3 sigma, resample when ESS < N/2.  A step is one full filter cycle
(predict -> weight -> estimate -> [resample -> estimate], particle_filter.rs:110-124).  `value`
is the fused cycle with the particle state and every cycle's inputs resident in HBM when the
timed region starts (estimates read back at its end); the same cycle with a host round trip
per cycle (PCIe-inclusive) and through the C++ host layer over the pfc_* ABI are reported
beside it.

Prints ONE JSON line.  Extra fields: per-stage device times (HIP events on the library
stream), the dominant kernel's roofline, the CBGL config-2 localisation figures, the batched
filters, the learned proposal and the CPU oracle baseline timed on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6

# Algorithmic work of k_weight per range-bearing particle-observation pair: the model of
# cpu.rs:80-96 counted as written (landmark offset 3, world-to-body rotation 15, norm 5, range
# residual + square + scale 3, direction dot product 5 + normalise 3, bearing square + scale 2,
# two gated accumulates 4 -> 40 FLOP, plus sqrt, reciprocal and acos).  This is the figure the
# roofline is computed on.  The kernel's own formulation issues more arithmetic per pair
# (Newton step after the reciprocal square root, the 3-term acos^2 series, fma = 2): 52 FLOP,
# reported beside it as issued work, not as the algorithm.
WEIGHT_FLOP_PER_PAIR = 40.0
WEIGHT_FLOP_PER_PAIR_ISSUED = 52.0
# Algorithmic HBM bytes per particle (SoA f32 pose 28 B, weight 4 B, log-likelihood 8 B,
# fixed-point cumulative weight 8 B).  ABI driver stages:
BYTES = {"predict": 56, "weight": 36, "normalize": 12, "resample_scan": 12, "resample_gather": 72,
         "est_sums": 32, "est_covs": 32}
# Fused (staged) cycle stages: weight = the likelihood (read pose, write ll: 36) plus the previous
# cycle's covariance pass that runs as extra workgroups of the same launch (read pose + w: 32),
# normalize = normalisation + tile aggregates + injection picks (read ll, write w), resample_scan
# = the scan from the tile aggregates (read w, write cum), resample_gather = gather + injection +
# estimate-sum partials (read cum window, pose, list head; write pose, w, ancestor), finalize =
# the one-workgroup-per-filter reductions (no particle traffic).
FUSED_BYTES = {"predict": 56, "weight": 68, "normalize": 12, "resample_scan": 12, "resample_gather": 76,
               "finalize": 0}
# Cycle with resampling (predict 56, weight 32, softmax 8, estimate 64, resample 64) and the
# likelihood's share of it.
CYCLE_BYTES = 224
BYTES_WEIGHT_ALG = 32


PROFILE_PF = os.path.join(ROOT, "profiles", "pf_cycle_1M_particles_256_obs.txt")


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (copy-calibrated), or None when the summary is absent."""
    try:
        with open(PROFILE_PF) as f:
            for line in f:
                if line.startswith(f"{kernel} traffic_bytes_per_launch:"):
                    return int(line.split(":")[1])
    except OSError:
        return None
    return None


def build_cycle_inputs(sc):
    from cuda_pf_localization_amd import _lib
    steps = []
    for st in sc.steps:
        o = _lib.PfhOdom()
        o.t[:] = list(st.odom_t)
        o.q[:] = list(st.odom_q)
        o.trans_noise[:] = list(st.trans_noise)
        o.rot_noise[:] = list(st.rot_noise)
        arr = (_lib.PfhObs * len(st.obs))()
        for j, (i, m, v, r) in enumerate(st.obs):
            arr[j].id, arr[j].mode = i, m
            arr[j].vx, arr[j].vy, arr[j].vz = v
            arr[j].range = r
        steps.append((o, arr, len(st.obs)))
    return steps


def run_pf(n, sc, steps, warmup, fast, profile):
    from cuda_pf_localization_amd import PfConfig, ParticleFilter, PfKernels, _lib
    cfg = PfConfig(particle_count=n, seed=7, range_noise=sc.sigma_r, bearing_noise=sc.sigma_b, gate_sigma=sc.gate,
                   resample_ess_ratio=0.5, random_inject_ratio=0.01, fast=fast)
    pf = ParticleFilter(cfg)
    pf.set_map(sc.landmarks)
    pf.reinitialize([*sc.prior_pos, *sc.prior_q], [0.5] * 3, [0.1] * 3)
    inputs = build_cycle_inputs(sc)
    est = _lib.PfhEstimate()
    kh = pf.kernels_handle()
    L = _lib.kernels()
    for i in range(warmup):
        o, arr, m = inputs[i % len(inputs)]
        pf.update_raw(o, arr, m, est)
    L.pfx_sync(kh)
    if profile:
        L.pfx_profile_enable(kh, 1)
        L.pfx_profile_read(kh, None, None)
    err = 0.0
    n_rs = 0
    t0 = time.perf_counter()
    for i in range(steps):
        o, arr, m = inputs[(warmup + i) % len(inputs)]
        pf.update_raw(o, arr, m, est)
        n_rs += est.resampled
        tp = sc.steps[(warmup + i) % len(inputs)].truth_pos
        err += math.dist(est.pos[:], tp)
    L.pfx_sync(kh)
    t1 = time.perf_counter()
    stages = None
    if profile:
        import ctypes as C
        from cuda_pf_localization_amd.api import STAGES
        ms = (C.c_double * len(STAGES))()
        cnt = (C.c_int * len(STAGES))()
        L.pfx_profile_read(kh, ms, cnt)
        L.pfx_profile_enable(kh, 0)
        stages = {s: {"ms_total": ms[j], "launches": cnt[j], "ms_avg": (ms[j] / cnt[j]) if cnt[j] else 0.0}
                  for j, s in enumerate(STAGES)}
    pf.close()
    return (t1 - t0) * 1000.0 / steps, err / steps, n_rs, stages


def run_gl(n_hyp, n_poses, seed=3):
    from cuda_pf_localization_amd import GlobalLoc, GlParams
    from cuda_pf_localization_amd.sim import f64_scan, free_poses, obstacle_map
    cells, res, org = obstacle_map(seed)
    truths = free_poses(cells, res, org, n_poses, seed=seed, margin_cells=6)
    rng = np.random.default_rng(seed)
    scans = [(f64_scan(cells, res, org, t, 360, 2 * math.pi, 8.0) + rng.normal(size=360) * 0.02).astype(np.float32)
             for t in truths]
    gl = GlobalLoc(cells, res, org, GlParams(hypotheses=n_hyp, rmax=8.0, topk=64, seed=seed))
    gl.localize(scans[0])  # warm-up
    ok = 0
    t0 = time.perf_counter()
    for i, (t, s) in enumerate(zip(truths, scans)):
        gl.set_seed(seed + 100 + i)
        top = gl.localize(s)
        ok += math.hypot(top[0, 0] - t[0], top[0, 1] - t[1]) < 0.5
    ms = (time.perf_counter() - t0) * 1000.0 / len(truths)
    from cuda_pf_localization_amd import _lib
    L = _lib.kernels()
    kh = gl.kernels_handle()
    L.pfglx_profile_enable(kh, 1)
    L.pfglx_profile_read(kh, None, None)
    gl.localize(scans[0])
    import ctypes as C
    msv = (C.c_double * 3)()
    cnt = (C.c_int * 3)()
    L.pfglx_profile_read(kh, msv, cnt)
    st = (C.c_int * 3)()
    L.pfglx_prune_stats(kh, st)
    gl.close()
    # rays actually cast by the profiled localisation: bounds over the rays of every second
    # 64-ray block, then the candidates and survivors in full (one wave per hypothesis) or only
    # their remaining rays (one lane per ray)
    nr = 360
    sel = sum(1 for r in range(nr) if (r >> 6) % 2 == 0)
    m, surv, route = st[0], st[1], st[2]
    if route == 0:
        cast = n_hyp * nr
    else:
        cast = n_hyp * sel + (m + surv) * (nr if route == 1 else nr - sel)
    return {"ms_per_localize": ms, "success": f"{ok}/{len(truths)}", "hypotheses": n_hyp,
            "hyp_per_s": n_hyp / (ms / 1000.0),
            "hypothesis_rays_per_s": n_hyp * nr / (msv[1] / 1000.0) if msv[1] else None,
            "rays_cast_per_localize": cast, "rays_cast_per_s": cast / (msv[1] / 1000.0) if msv[1] else None,
            "pruning": {"candidates": m, "survivors": surv, "survivor_fraction": round(surv / n_hyp, 4),
                        "route": route},
            "stage_ms": {"generate": msv[0], "score": msv[1], "topk": msv[2]}}


PFH_OBS_DTYPE = np.dtype([("id", "<i4"), ("mode", "<i4"), ("vx", "<f8"), ("vy", "<f8"), ("vz", "<f8"),
                          ("range", "<f8")])


def run_batched(nf=64, cycles=100, warmup=2, seed=11):
    """64 independent filters: N log-uniform in [1k, 500k], L uniform in {4..256}, 20 x 20 x 5 m
    boxes, random helix phase, 0.05 m / 0.02 rad observation noise, 100 cycles.  One batched
    cycle (pfb_cycle: one launch per stage for all filters, hipGraph replay, autotuned weight
    tile) against the same 64 filters run one after another through the plain path."""
    import ctypes as C
    from cuda_pf_localization_amd import OBS_DTYPE, BatchedFilters, ParticleFilter, PfConfig, _lib
    from cuda_pf_localization_amd.sim import config3_filters
    ns, ls, hs = config3_filters(nf, cycles, seed)
    obs32, obs64 = [], []
    for h in hs:
        L = len(h.landmarks)
        a = np.zeros((cycles, L), OBS_DTYPE)
        b = np.zeros((cycles, L), PFH_OBS_DTYPE)
        a["id"] = b["id"] = np.arange(L)
        a["mode"] = b["mode"] = 1
        for j, (k32, k64) in enumerate((("dx", "vx"), ("dy", "vy"), ("dz", "vz"))):
            a[k32] = h.obs_v[:, :, j]
            b[k64] = h.obs_v[:, :, j]
        a["range"] = h.obs_r
        b["range"] = h.obs_r
        obs32.append(a)
        obs64.append(b)
    cfgs = [dict(n=int(ns[f]), seed=seed + f, sigma_r=0.05, sigma_b=0.02, gate=3.0, ratio=0.5, inject=0.0)
            for f in range(nf)]
    bf = BatchedFilters(cfgs, obs_cap=256)
    for f, h in enumerate(hs):
        bf.upload_landmarks(f, h.landmarks.astype(np.float32))
    bf.reinit(np.array([[*h.prior_pos, *h.prior_q] for h in hs], np.float32), np.full((nf, 3), 0.5, np.float32),
              np.full((nf, 3), 0.1, np.float32))
    inputs = [bf.make_inputs([(h.odom_t[k], h.odom_q[k], (h.odo_pos,) * 3, (h.odo_rot,) * 3, obs32[f][k])
                              for f, h in enumerate(hs)]) for k in range(cycles)]
    bf.use_graph(True)
    ppt = bf.autotune(inputs[0][0])
    for k in range(warmup):
        bf.cycle(inputs[k][0])
    t0 = time.perf_counter()
    for k in range(warmup, cycles):
        out = bf.cycle(inputs[k][0])
    t_b = (time.perf_counter() - t0) * 1000.0 / (cycles - warmup)
    err_b = float(np.mean([math.dist(out[f].pos[:], hs[f].truth[-1]) for f in range(nf)]))
    bf.close()
    # plain path, one filter after another
    t_p = 0.0
    err_p = []
    est = _lib.PfhEstimate()
    for f, h in enumerate(hs):
        pf = ParticleFilter(PfConfig(particle_count=int(ns[f]), seed=seed + f, range_noise=0.05, bearing_noise=0.02,
                                     gate_sigma=3.0, resample_ess_ratio=0.5, random_inject_ratio=0.0, fast=True))
        pf.set_map({i: tuple(p) for i, p in enumerate(h.landmarks)})
        pf.reinitialize([*h.prior_pos, *h.prior_q], [0.5] * 3, [0.1] * 3)
        odo = []
        for k in range(cycles):
            o = _lib.PfhOdom()
            o.t[:] = list(h.odom_t[k])
            o.q[:] = list(h.odom_q[k])
            o.trans_noise[:] = [h.odo_pos] * 3
            o.rot_noise[:] = [h.odo_rot] * 3
            odo.append(o)
        ptrs = [obs64[f][k].ctypes.data_as(C.POINTER(_lib.PfhObs)) for k in range(cycles)]
        for k in range(warmup):
            pf.update_raw(odo[k], ptrs[k], len(h.landmarks), est)
        t0 = time.perf_counter()
        for k in range(warmup, cycles):
            pf.update_raw(odo[k], ptrs[k], len(h.landmarks), est)
        t_p += time.perf_counter() - t0
        err_p.append(math.dist(est.pos[:], h.truth[-1]))
        pf.close()
    t_p = t_p * 1000.0 / (cycles - warmup)
    return {"workload": "64 filters, N log-uniform [1k, 500k], L uniform [4, 256], 100 cycles, seed 11",
            "particles_total": int(ns.sum()), "landmarks_total": int(ls.sum()),
            "batched_ms_per_cycle": round(t_b, 4), "sequential_plain_ms_per_cycle": round(t_p, 4),
            "speedup": round(t_p / t_b, 2), "particle_cycles_per_s": round(float(ns.sum()) / (t_b * 1e-3), 1),
            "weight_particles_per_lane": int(ppt), "graph": True,
            "mean_final_err_m": {"batched": round(err_b, 4), "plain": round(float(np.mean(err_p)), 4)}}


def run_learned(maps=tuple(range(100, 110)), n_poses=20):
    """Learned proposal on a 10-map family: per map, train the MLP on 1M GPU-rendered scans
    (Adam 1e-3, batch 4096), then localise n_poses scans with the learned proposal (8k
    mixture + 2k uniform, refinement, uniform fallback) and with 200k uniform hypotheses, both
    with the library's pruned scoring (the default) and with every CAER scored, as the
    reference's GL path does."""
    from cuda_pf_localization_amd import GlobalLoc, GlParams, LearnedParams, LearnedProposal
    from cuda_pf_localization_amd.sim import f64_scan, free_poses, obstacle_map
    from cuda_pf_localization_amd import _lib
    L = _lib.kernels()
    ok_l = ok_u = ok_f = fb = 0
    t_l = t_u = t_f = t_train = 0.0
    losses = []
    params = None
    for ms in maps:
        cells, res, org = obstacle_map(ms)
        gl = GlobalLoc(cells, res, org, GlParams(hypotheses=200_000, rmax=8.0, topk=64, seed=ms))
        net = LearnedProposal(360, 512, 3, 16, max_batch=4096, seed=ms)
        params = net.param_count()
        t0 = time.perf_counter()
        first, last = net.train_on_map(gl.kernels_handle(), n_samples=1_000_000, batch=4096, lr=1e-3, rmax=8.0,
                                       ang_span=2 * math.pi, noise_sigma=0.02, seed=ms)
        t_train += time.perf_counter() - t0
        losses.append((round(first, 3), round(last, 3)))
        truths = free_poses(cells, res, org, n_poses, seed=ms, margin_cells=6)
        rng = np.random.default_rng(ms)
        scans = [(f64_scan(cells, res, org, t, 360, 2 * math.pi, 8.0) + rng.normal(size=360) * 0.02)
                 .astype(np.float32) for t in truths]
        lp = LearnedParams(seed=ms)
        gl.localize_learned(net, scans[0], lp)  # warm-up
        gl.localize(scans[0])
        for i, (t, s) in enumerate(zip(truths, scans)):
            lp.seed = ms * 1000 + i
            a = time.perf_counter()
            top, info = gl.localize_learned(net, s, lp)
            t_l += time.perf_counter() - a
            ok_l += math.hypot(top[0, 0] - t[0], top[0, 1] - t[1]) < 0.5
            fb += info["used_fallback"]
            gl.set_seed(ms * 1000 + i)
            a = time.perf_counter()
            top = gl.localize(s)
            t_u += time.perf_counter() - a
            ok_u += math.hypot(top[0, 0] - t[0], top[0, 1] - t[1]) < 0.5
            L.pfglx_set_prune(gl.kernels_handle(), 0)
            gl.set_seed(ms * 1000 + i)
            a = time.perf_counter()
            top = gl.localize(s)
            t_f += time.perf_counter() - a
            ok_f += math.hypot(top[0, 0] - t[0], top[0, 1] - t[1]) < 0.5
            L.pfglx_set_prune(gl.kernels_handle(), 1)
        net.close()
        gl.close()
    n = len(maps) * n_poses
    train_flop = 6.0 * params * 1_000_000 * len(maps)
    return {"workload": f"{len(maps)} obstacle maps (seeds {maps[0]}..{maps[-1]}), {n_poses} poses each, 360 rays, "
                        "rmax 8 m, noise 0.02 m",
            "learned_ms_per_fix": round(t_l * 1000.0 / n, 4), "uniform_ms_per_fix": round(t_u * 1000.0 / n, 4),
            "speedup": round(t_u / t_l, 2), "learned_success": f"{ok_l}/{n}", "uniform_success": f"{ok_u}/{n}",
            "uniform_full_scoring_ms_per_fix": round(t_f * 1000.0 / n, 4),
            "speedup_vs_full_scoring": round(t_f / t_l, 2), "uniform_full_scoring_success": f"{ok_f}/{n}",
            "fallbacks": int(fb), "train_s_per_map": round(t_train / len(maps), 3),
            "train_tflops_end_to_end": round(train_flop / t_train / 1e12, 2), "params": int(params),
            "loss_first_last": losses}


def run_pf_fused(n, sc, steps, warmup, steps_host, weight_ppt=0):
    """The 1M-particle cycle through the fused entry (one filter): every device stage (predict,
    weight, normalise + ESS predicate, resample, estimate sums, Markley mean, covariances) on the
    device.  Two timings:
      * resident (the headline): the inputs of `steps` cycles (odometry + 256 observations each)
        are staged in HBM first; the cycles are then enqueued back to back (each cycle's kernels
        read its parameters in place) and every cycle's estimate is read back at the end
        (pfb_stage_cycles / pfb_run_staged);
      * host inputs: `steps_host` pfb_cycle calls, each with one H2D copy of the cycle's inputs,
        the cycle graph and one D2H of the estimate that the host waits for (PCIe-inclusive).
    Both continue the same filter; the staged cycles follow the host-input ones."""
    from cuda_pf_localization_amd import OBS_DTYPE, BatchedFilters
    bf = BatchedFilters([dict(n=n, seed=7, sigma_r=sc.sigma_r, sigma_b=sc.sigma_b, gate=sc.gate, ratio=0.5,
                              inject=0.01)], obs_cap=256)
    bf.upload_landmarks(0, np.array([sc.landmarks[i] for i in range(len(sc.landmarks))], np.float32))
    bf.reinit(np.array([[*sc.prior_pos, *sc.prior_q]], np.float32), np.full((1, 3), 0.5, np.float32),
              np.full((1, 3), 0.1, np.float32))
    inputs = []
    for st in sc.steps:
        obs = np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], OBS_DTYPE)
        inputs.append(bf.make_inputs([(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise, obs)]))
    assert len(inputs) >= warmup + steps_host + steps
    bf.use_graph(True)
    # the likelihood launch shape from the library's autotuner (times the shapes on this filter
    # and restores the draw state it staged)
    if weight_ppt:
        bf.set_weight_ppt(weight_ppt)
        ppt = weight_ppt
    else:
        ppt = bf.autotune(inputs[0][0])
    for i in range(warmup):
        bf.cycle(inputs[i][0])
    t0 = time.perf_counter()
    for i in range(steps_host):
        bf.cycle(inputs[warmup + i][0])
    ms_host = (time.perf_counter() - t0) * 1000.0 / steps_host
    k0 = warmup + steps_host
    bf.stage_cycles([inputs[k0 + i][0] for i in range(steps)])
    t0 = time.perf_counter()
    outs = bf.run_staged()
    ms = (time.perf_counter() - t0) * 1000.0 / steps
    ms_dev = bf.staged_device_ms() / steps
    err = sum(math.dist(outs[i][0].pos[:], sc.steps[k0 + i].truth_pos) for i in range(steps)) / steps
    n_rs = sum(int(outs[i][0].resampled) for i in range(steps))
    # the same staged cycles again with HIP events around every stage (on the library stream):
    # the per-stage breakdown of the fused cycle and the likelihood kernel's average duration
    k1 = k0 + steps
    n_prof = min(steps, len(inputs) - k1)
    bf.stage_cycles([inputs[k1 + i][0] for i in range(n_prof)])
    bf.profile(True)
    bf.profile_read()
    t0 = time.perf_counter()
    bf.run_staged()
    ms_prof = (time.perf_counter() - t0) * 1000.0 / n_prof
    prof = bf.profile_read()
    bf.profile(False)
    bf.close()
    stages = {k: {"ms_total": v[0], "launches": v[1], "ms_avg": v[0] / v[1] if v[1] else 0.0} for k, v in prof.items()}
    return ms, ms_dev, ms_host, err, n_rs, stages, ms_prof, n_prof, ppt


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _oracle_pf(sc, n, seed, inject, budget_s, max_cycles):
    """Oracle cycles (predict, weight, estimate, resample when ESS < N/2, estimate) on the
    scenario's steps until budget_s or max_cycles; returns (ms per cycle, cycles)."""
    from oracle import OBS_DTYPE, OracleFilter
    o = OracleFilter(n, seed, False)
    o.reinit(np.array([*sc.prior_pos, *sc.prior_q]), np.array([0.5] * 3), np.array([0.1] * 3))
    lm = np.array([sc.landmarks[i] for i in range(len(sc.landmarks))], np.float32)
    obs_all = [np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], OBS_DTYPE) for st in sc.steps[:max_cycles]]
    cycles = 0
    t0 = time.perf_counter()
    while cycles < min(max_cycles, len(sc.steps)) and (cycles == 0 or time.perf_counter() - t0 < budget_s):
        st = sc.steps[cycles]
        o.predict(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise)
        o.weight(obs_all[cycles], lm, sc.sigma_r, sc.sigma_b, sc.gate)
        e = o.estimate()
        if e["ess"] < 0.5 * n:
            o.resample(np.float32(inject))
            o.estimate()
        cycles += 1
    return (time.perf_counter() - t0) * 1000.0 / cycles, cycles


def _oracle_gl(n_hyp, n_poses, seed=3):
    """The CPU twin of CpuGlobalLoc::localize (global_loc.rs:384-398: xorshift dispersal, f64
    DDA + CAER over the hypotheses in parallel, stable sort, top 64) on the 200k-hypothesis
    obstacle-map workload; returns (ms per localisation, successes)."""
    from oracle import OracleGrid
    from cuda_pf_localization_amd.sim import f64_scan, free_poses, obstacle_map
    cells, res, org = obstacle_map(seed)
    truths = free_poses(cells, res, org, n_poses, seed=seed, margin_cells=6)
    rng = np.random.default_rng(seed)
    g = OracleGrid(cells, res, org)
    ok = 0
    t = 0.0
    for i, tp in enumerate(truths):
        scan = f64_scan(cells, res, org, tp, 360, 2 * math.pi, 8.0) + rng.normal(size=360) * 0.02
        scan = scan.astype(np.float32).astype(np.float64)
        a = time.perf_counter()
        top = g.localize64(scan, n_hyp, 64, seed + 100 + i, span=2 * math.pi, rmax=8.0)
        t += time.perf_counter() - a
        ok += math.hypot(top[0, 0] - tp[0], top[0, 1] - tp[1]) < 0.5
    # algorithmic work of one scoring pass: DDA cell steps of the twin over its own hypotheses
    steps = g.dda_steps64(g.disperse64(seed + 100, n_hyp), 360, 2 * math.pi, 8.0)
    return t * 1000.0 / n_poses, ok, steps / (n_hyp * 360.0)


def _oracle_pf_bench_rows(sizes=(1_000, 10_000, 100_000, 500_000), steps_n=100):
    """The reference pf-bench's CPU rows (crates/pf_bench/src/main.rs:21-93) on the oracle: the
    circle scenario (100 cycles, seed 777, Cartesian observations), seed 123, range noise 0.05,
    no gate, no injection, ESS ratio 0.5, prior (1, 1, 1) m / (0.5, 0.5, 0.5) rad, one warm-up
    update, then the mean wall time per update over the 100 steps and the mean position error."""
    from oracle import OBS_DTYPE, OracleFilter
    from cuda_pf_localization_amd.sim import circle_scenario
    sc = circle_scenario(steps_n, 777)
    lm = np.array([sc.landmarks[i] for i in range(len(sc.landmarks))], np.float32)
    steps = [(st, np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], OBS_DTYPE)) for st in sc.steps]
    rows = {}
    for n in sizes:
        o = OracleFilter(n, 123, False)
        o.reinit(np.array([*sc.prior_pos, *sc.prior_q]), np.ones(3), np.full(3, 0.5))

        def update(st, obs):
            o.predict(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise)
            o.weight(obs, lm, 0.05, 0.02, 0.0)
            e = o.estimate()
            if e["ess"] < 0.5 * n:
                o.resample(np.float32(0.0))
                e = o.estimate()
            return e

        update(*steps[0])
        err = 0.0
        t0 = time.perf_counter()
        for st, obs in steps:
            err += math.dist(update(st, obs)["position"], st.truth_pos)
        rows[str(n)] = {"ms_per_cycle": round((time.perf_counter() - t0) * 1000.0 / steps_n, 3),
                        "mean_track_err_m": round(err / steps_n, 4)}
    return rows


def cpu_baseline(sc, n, budget_s=12.0):
    """The oracle (C restatement of cpu.rs and of the global_loc.rs CPU twin, f64, OpenMP in
    place of rayon, the same parallel/sequential split) on this host.  `value` is the config-1
    workload (1M particles, 256 observations) on every thread OpenMP is given, over as many
    cycles as fit ~budget_s; beside it the same on one thread, config 0 (10k particles, 12
    landmarks, 100 cycles) on all threads and on one, and the 200k-hypothesis CBGL
    localisation."""
    from oracle import num_threads, set_threads
    from cuda_pf_localization_amd.sim import config0_scenario
    threads = num_threads()
    ms1, c1 = _oracle_pf(sc, n, 7, 0.01, budget_s, len(sc.steps))
    set_threads(1)
    try:
        ms1_1t, c1_1t = _oracle_pf(sc, n, 7, 0.01, budget_s * 0.75, len(sc.steps))
        sc0 = config0_scenario(cycles=100, seed=42)
        ms0_1t, c0_1t = _oracle_pf(sc0, 10_000, 42, 0.0, 60.0, 100)
    finally:
        set_threads(threads)
    ms0, c0 = _oracle_pf(sc0, 10_000, 42, 0.0, 60.0, 100)
    ms_gl, ok_gl, steps_per_ray = _oracle_gl(200_000, 3)
    pf_bench_rows = _oracle_pf_bench_rows()
    return {"value": ms1, "unit": "ms/cycle", "cores": threads, "kind": "port",
            "sample": f"{c1} cycles of the 1M-particle, 256-observation workload (oracle: C restatement of "
                      f"cpu.rs, f64, OpenMP over particles like rayon)",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "legs": {
                "config1_all_threads": {"particles": n, "ms_per_cycle": round(ms1, 2), "cycles": c1,
                                        "threads": threads},
                "config1_1_thread": {"particles": n, "ms_per_cycle": round(ms1_1t, 2), "cycles": c1_1t, "threads": 1},
                "config0_10k_all_threads": {"ms_per_cycle": round(ms0, 3), "cycles": c0, "threads": threads},
                "config0_10k_1_thread": {"ms_per_cycle": round(ms0_1t, 3), "cycles": c0_1t, "threads": 1},
                "gl_200k_hypotheses_all_threads": {"ms_per_localize": round(ms_gl, 1), "poses": 3,
                                                   "success": f"{ok_gl}/3", "threads": threads,
                                                   "dda_cell_steps_per_ray": round(steps_per_ray, 3)},
                "pf_bench_circle_cpu_rows": pf_bench_rows,
            }}


def copy_calibration(nbytes: int) -> None:
    """A plain copy of known size (read nbytes + write nbytes) for scaling the FETCH_SIZE /
    WRITE_SIZE counters of the profiled kernels."""
    import torch
    a = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "launches": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gl", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-sweep", action="store_true", help="skip the 4M / 16M-particle staged cycles")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    ap.add_argument("--weight-ppt", type=int, default=0, choices=(0, 1, 2, 4),
                    help="likelihood particles per lane of the headline cycle (0: the autotuner's choice)")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="only run a plain device-to-device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: the particle filter is a single-device path here "
              f"(--gpus {args.gpus} is not supported)", file=sys.stderr)
        sys.exit(2)
    from cuda_pf_localization_amd.sim import config1_scenario
    n = args.particles
    steps_host = min(args.steps, 100)
    steps_prof = min(args.steps, 50)
    sc = config1_scenario(cycles=max(args.steps + args.warmup + steps_host + steps_prof, 50), seed=7)
    (ms_fused, ms_fused_dev, ms_host, err_fused, n_rs_fused, fstages, ms_fprof,
     n_fprof, weight_ppt) = run_pf_fused(n, sc, args.steps, args.warmup, steps_host, args.weight_ppt)
    ms_fast, err, n_rs, stages = run_pf(n, sc, args.steps, args.warmup, fast=True, profile=False)
    ms_exact, _, _, _ = run_pf(n, sc, min(args.steps, 50), args.warmup, fast=False, profile=False)
    _, _, _, stages = run_pf(n, sc, min(args.steps, 50), args.warmup, fast=True, profile=True)
    # the likelihood kernel's duration from the profiled staged (fused, inputs-resident) run
    w = fstages["weight"]
    pairs = n * 256
    achieved_tflops = WEIGHT_FLOP_PER_PAIR * pairs / (w["ms_avg"] * 1e-3) / 1e12 if w["ms_avg"] else None
    issued_tflops = WEIGHT_FLOP_PER_PAIR_ISSUED * pairs / (w["ms_avg"] * 1e-3) / 1e12 if w["ms_avg"] else None

    def stage_table(st_ms, nbytes):
        out = {}
        for s_, b_ in nbytes.items():
            st = st_ms.get(s_)
            if st and st["launches"]:
                gbs = b_ * n / (st["ms_avg"] * 1e-3) / 1e9 if b_ else None
                out[s_] = {"ms_avg": round(st["ms_avg"], 5), "GB/s": round(gbs, 1) if gbs else None,
                           "frac_hbm": round(gbs / HBM_PEAK_GBS, 3) if gbs else None}
        return out
    stage_bw = stage_table(fstages, FUSED_BYTES)
    stage_bw_abi = stage_table(stages, BYTES)
    # the same staged cycle at 4M and 16M particles: the per-stage HBM fractions once the small
    # passes are long enough to amortise their launch-to-retire latency
    size_sweep = {}
    if not args.no_sweep:
        for n_sw in (4 * 1024 * 1024, 16 * 1024 * 1024):
            ms_sw, ms_sw_dev, _, err_sw, _, st_sw, _, _, ppt_sw = run_pf_fused(n_sw, sc, 20, 5, 5, args.weight_ppt)
            tab = {}
            for s_, b_ in FUSED_BYTES.items():
                st = st_sw.get(s_)
                if st and st["launches"]:
                    gbs = b_ * n_sw / (st["ms_avg"] * 1e-3) / 1e9 if b_ else None
                    tab[s_] = {"ms_avg": round(st["ms_avg"], 4), "GB/s": round(gbs, 1) if gbs else None,
                               "frac_hbm": round(gbs / HBM_PEAK_GBS, 3) if gbs else None}
            wl = st_sw["weight"]["ms_avg"]
            size_sweep[str(n_sw)] = {
                "ms_per_cycle": round(ms_sw, 4), "particles_per_s": round(n_sw / (ms_sw * 1e-3), 1),
                "mean_track_err_m": round(err_sw, 4), "particles_per_lane": ppt_sw,
                "likelihood_tflops": round(WEIGHT_FLOP_PER_PAIR * n_sw * 256 / (wl * 1e-3) / 1e12, 2) if wl else None,
                "stages": tab}
    # cycle roofline: every stage at the HBM peak on its algorithmic bytes, the likelihood at the
    # FP64 peak on its FLOP (the larger of its two bounds)
    t_roof_us = ((CYCLE_BYTES - BYTES_WEIGHT_ALG) * n / (HBM_PEAK_GBS * 1e9)
                 + max(BYTES_WEIGHT_ALG * n / (HBM_PEAK_GBS * 1e9),
                       WEIGHT_FLOP_PER_PAIR * pairs / (FP64_VALU_PEAK_TFLOPS * 1e12))) * 1e6
    line = {
        "metric": "PF ms/cycle @1M particles; GL time-to-0.5m-fix, learned vs plain dispersal",
        "value": round(ms_fused, 4),
        "unit": "ms/cycle",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_fused, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "storage_dtype": "f32 particle state (f64 log-likelihoods, u64 fixed-point cumulative weights)",
        "data": "synthetic (helix trajectory, uniform landmarks, generated in-process)",
        "config": {"workload": "config 1: 1M-particle PF cycle, 256 landmarks / RB observations, 1% outliers, "
                               "inject 0.01, gate 3", "particles": n, "landmarks": 256, "observations": 256,
                   "driver": "fused cycle, inputs resident in HBM (pfb_stage_cycles + pfb_run_staged: "
                             "one filter, cycles enqueued back to back, estimates read back at the end)"},
        "particles_per_s": round(n / (ms_fused * 1e-3), 1),
        "ms_per_cycle_device": round(ms_fused_dev, 4),
        "ms_per_cycle_host_inputs": round(ms_host, 4),
        "ms_per_cycle_abi_driver": round(ms_fast, 4),
        "ms_per_cycle_exact_driver": round(ms_exact, 4),
        "mean_track_err_m": round(err_fused, 4),
        "mean_track_err_m_abi_driver": round(err, 4),
        "resample_cycles": int(n_rs_fused),
        "stages": stage_bw,
        "stages_profiled_ms_per_cycle": round(ms_fprof, 4),
        "stages_profiled_cycles": n_fprof,
        "stages_abi_driver": stage_bw_abi,
        "size_sweep": size_sweep,
        "cycle_roofline": {"t_roof_us": round(t_roof_us, 1), "t_measured_us": round(ms_fused * 1e3, 1),
                           "frac": round(t_roof_us / (ms_fused * 1e3), 3),
                           "basis": f"{CYCLE_BYTES} B/particle at 8 TB/s, likelihood at the FP64 peak on "
                                    f"{WEIGHT_FLOP_PER_PAIR:.0f} FLOP/pair"},
        "roofline": {"kernel": "k_weight", "bound": "valu", "pipe": "FP64 VALU (no matrix instructions; the peak is "
                     "MI355X's FP64 vector figure)",
                     "achieved": round(achieved_tflops, 2) if achieved_tflops
                     else None, "peak": FP64_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / FP64_VALU_PEAK_TFLOPS, 3) if achieved_tflops else None,
                     "traffic": profiled_traffic("k_weight"),
                     "traffic_source": "rocprofv3 FETCH_SIZE + WRITE_SIZE of the same workload, copy-calibrated, "
                                       "read from profiles/pf_cycle_1M_particles_256_obs.txt (not measured in this "
                                       "run)",
                     "traffic_algorithmic": 68 * n,  # likelihood 36 B + the deferred covariance pass 32 B
                     "ms_avg": round(w["ms_avg"], 5),
                     "launch_includes": "the previous cycle's covariance pass (one extra workgroup per "
                                        "1024-particle tile, dispatched after the likelihood's): the achieved "
                                        "rate is counted on the likelihood's FLOP alone",
                     "ms_avg_plain_kernel_abi_driver_1_per_lane": round(stages["weight"]["ms_avg"], 5)
                     if stages.get("weight") else None,
                     "flop_per_pair": WEIGHT_FLOP_PER_PAIR, "pairs_per_launch": pairs,
                     "particles_per_lane": int(weight_ppt),
                     "issued": {"flop_per_pair": WEIGHT_FLOP_PER_PAIR_ISSUED,
                                "tflops": round(issued_tflops, 2) if issued_tflops else None,
                                "frac": round(issued_tflops / FP64_VALU_PEAK_TFLOPS, 3) if issued_tflops else None}},
    }
    if not args.no_gl:
        line["gl"] = run_gl(200_000, 20)
    from cuda_pf_localization_amd import pf_bench
    line["pf_bench_circle_hip_rows"] = {}
    for nn in (1_000, 10_000, 100_000, 500_000):
        ms_c, err_c = pf_bench.run_pf(nn)
        line["pf_bench_circle_hip_rows"][str(nn)] = {"ms_per_cycle": round(ms_c, 3), "mean_track_err_m": round(err_c, 4)}
    if not args.no_batched:
        line["batched"] = run_batched()
    if not args.no_learned:
        line["learned"] = run_learned()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(sc, n, args.cpu_budget)
        if "gl" in line and line["gl"].get("rays_cast_per_s"):
            spr = line["cpu_baseline"]["legs"]["gl_200k_hypotheses_all_threads"]["dda_cell_steps_per_ray"]
            line["gl"]["dda_cell_steps_per_ray_algorithmic"] = spr
            # (the mean steps per ray of all hypotheses applied to the rays actually cast)
            line["gl"]["cell_steps_per_s_est"] = round(line["gl"]["rays_cast_per_s"] * spr, 1)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

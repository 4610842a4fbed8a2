#!/usr/bin/env python3
"""Benchmark: particle-filter ms/cycle at 1M particles (config 1) plus CBGL time-to-fix.

Headline workload (one GPU): 1,000,000 particles, 256 landmarks in a 50 x 50 x 10 m box,
256 range-bearing observations per cycle (1 % range outliers U[0, 60] m), helix motion,
inject ratio 0.01, gate 3 sigma, resample when ESS < N/2.  A step is one full filter cycle
(predict -> weight -> estimate -> [resample -> estimate], particle_filter.rs:110-124) driven
by the C++ host layer over the pfc_* ABI; inputs (particle state) stay resident in HBM.

Prints ONE JSON line.  Extra fields: per-stage device times (HIP events on the library
stream), the dominant kernel's roofline, the CBGL config-2 localisation figures and the CPU
oracle baseline timed on this host.
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

# Algorithmic work of k_weight per range-bearing particle-observation pair, counted on the
# kernel's f64 formulation (fma = 2 FLOP): landmark offset 3, body rotation 15, |b|^2 5,
# b.v 5, reciprocal-sqrt refinement 6, range residual 2 + square 1 + gated accumulate 2,
# 1 - dot 2, acos^2 series 11, gated accumulate 2 -> 54 FLOP, plus one f32 rsqrt and two
# conversions that are not counted.
WEIGHT_FLOP_PER_PAIR = 54.0
# Algorithmic HBM bytes per particle (SoA f32 pose 28 B, weight 4 B, log-likelihood 8 B).
BYTES = {"predict": 56, "weight": 36, "normalize": 12, "resample_scan": 12, "resample_gather": 72,
         "est_sums": 32, "est_covs": 32}


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
    gl.close()
    return {"ms_per_localize": ms, "success": f"{ok}/{len(truths)}", "hypotheses": n_hyp,
            "hyp_per_s": n_hyp / (ms / 1000.0), "rays_per_s": n_hyp * 360 / (msv[1] / 1000.0) if msv[1] else None,
            "stage_ms": {"generate": msv[0], "score": msv[1], "topk": msv[2]}}


def cpu_baseline(sc, n, budget_s=12.0):
    """The oracle (C restatement of cpu.rs, f64, OpenMP like rayon) on this host, same config-1
    workload at full particle count, timed over as many cycles as fit ~budget_s."""
    from oracle import OBS_DTYPE, OracleFilter, num_threads
    o = OracleFilter(n, 7, False)
    o.reinit(np.array([*sc.prior_pos, *sc.prior_q]), np.array([0.5] * 3), np.array([0.1] * 3))
    lm = np.array([sc.landmarks[i] for i in range(len(sc.landmarks))], np.float32)
    cycles = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and cycles < len(sc.steps):
        st = sc.steps[cycles]
        o.predict(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise)
        obs = np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], OBS_DTYPE)
        o.weight(obs, lm, sc.sigma_r, sc.sigma_b, sc.gate)
        e = o.estimate()
        if e["ess"] < 0.5 * n:
            o.resample(np.float32(0.01))
            o.estimate()
        cycles += 1
    ms = (time.perf_counter() - t0) * 1000.0 / cycles
    return {"value": ms, "unit": "ms/cycle", "cores": num_threads(), "kind": "port",
            "sample": f"{cycles} cycles of the 1M-particle, 256-observation workload (oracle: C restatement of "
                      f"cpu.rs, f64, OpenMP)"}


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
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--particles", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-gl", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
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
    sc = config1_scenario(cycles=max(args.steps + args.warmup, 50), seed=7)
    ms_fast, err, n_rs, stages = run_pf(n, sc, args.steps, args.warmup, fast=True, profile=False)
    ms_exact, _, _, _ = run_pf(n, sc, min(args.steps, 50), args.warmup, fast=False, profile=False)
    _, _, _, stages = run_pf(n, sc, min(args.steps, 50), args.warmup, fast=True, profile=True)
    w = stages["weight"]
    pairs = n * 256
    achieved_tflops = WEIGHT_FLOP_PER_PAIR * pairs / (w["ms_avg"] * 1e-3) / 1e12 if w["ms_avg"] else None
    stage_bw = {}
    for s, b in BYTES.items():
        st = stages.get(s)
        if st and st["launches"]:
            gbs = b * n / (st["ms_avg"] * 1e-3) / 1e9
            stage_bw[s] = {"ms_avg": round(st["ms_avg"], 5), "GB/s": round(gbs, 1),
                           "frac_hbm": round(gbs / HBM_PEAK_GBS, 3)}
    line = {
        "metric": "PF ms/cycle @1M particles; GL time-to-0.5m-fix, learned vs plain dispersal",
        "value": round(ms_fast, 4),
        "unit": "ms/cycle",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_fast, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 likelihood/reduction",
        "data": "synthetic (helix trajectory, uniform landmarks, generated in-process)",
        "config": {"workload": "config 1: 1M-particle PF cycle, 256 landmarks / RB observations, 1% outliers, "
                               "inject 0.01, gate 3", "particles": n, "landmarks": 256, "observations": 256,
                   "driver": "C++ host layer (ParticleFilter) over pfc_* ABI, fast mode"},
        "particles_per_s": round(n / (ms_fast * 1e-3), 1),
        "ms_per_cycle_exact_driver": round(ms_exact, 4),
        "mean_track_err_m": round(err, 4),
        "resample_cycles": int(n_rs),
        "stages": stage_bw,
        "roofline": {"kernel": "k_weight", "bound": "valu", "achieved": round(achieved_tflops, 2) if achieved_tflops
                     else None, "peak": FP64_VALU_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved_tflops / FP64_VALU_PEAK_TFLOPS, 3) if achieved_tflops else None,
                     "traffic": None, "ms_avg": round(w["ms_avg"], 5),
                     "flop_per_pair": WEIGHT_FLOP_PER_PAIR, "pairs_per_launch": pairs},
    }
    if not args.no_gl:
        line["gl"] = run_gl(200_000, 20)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(sc, n, args.cpu_budget)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Command-line benchmark with the tables and output format of the reference's pf-bench
(crates/pf_bench/src/main.rs:62-176), run on the HIP backend.

    python -m cuda_pf_localization_amd.pf_bench [--csv] [--gl]

PF table (bench:62-94): N in {1k, 10k, 100k, 500k}, the frozen circle scenario of sim.rs
(100 cycles, seed 777, Cartesian landmark observations), PfConfig{particle_count: N, seed: 123,
range_noise: 0.05, ..Default}, reinitialise about the prior with (1, 1, 1) m / (0.5, 0.5, 0.5)
rad, one warm-up update, then the mean wall time of update() over the 100 steps and the mean
position error.  GL table (--gl, bench:96-176): room_grid, truth (6, 7, 0.8), a 360-ray 10 m
scan of the map, 5k / 50k / 200k hypotheses with GlParams defaults, 20 localisations with
seeds 100 + i after one warm-up.

The reference also prints rows for its rayon CPU backend.  That backend exists here only as the
test oracle (a C restatement of cpu.rs), which the product never loads, so this module prints the
"hip" rows; the repository's bench.py prints the CPU rows of the same table from the oracle
(`cpu_baseline.legs.pf_bench_circle_cpu_rows`) beside these (`pf_bench_circle_hip_rows`).
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np

from .api import GlobalLoc, GlParams, Observation, OdomDelta, ParticleFilter, PfConfig
from .sim import circle_scenario, f64_scan, room_grid


def run_pf(n: int, steps_n: int = 100, seed: int = 777):
    """bench:21-60 on the HIP backend: (ms per cycle, mean track error)."""
    sc = circle_scenario(steps_n, seed)
    pf = ParticleFilter(PfConfig(particle_count=n, seed=123, range_noise=0.05))
    pf.set_map(sc.landmarks)
    pf.reinitialize([*sc.prior_pos, *sc.prior_q], [1.0, 1.0, 1.0], [0.5, 0.5, 0.5])
    steps = [(OdomDelta(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise),
              [Observation(i, m, v, r) for (i, m, v, r) in st.obs], st.truth_pos) for st in sc.steps]
    pf.update(steps[0][0], steps[0][1])  # warm-up
    t0 = time.perf_counter()
    err = 0.0
    for odom, obs, truth in steps:
        est = pf.update(odom, obs)
        err += math.dist(est.position, truth)
    ms = (time.perf_counter() - t0) * 1000.0 / steps_n
    pf.close()
    return ms, err / steps_n


def run_gl(nh: int, runs: int = 20):
    """bench:96-176 (GPU half): (ms per localisation, error of a final run, successes)."""
    cells, res, org = room_grid()
    truth = (6.0, 7.0, 0.8)
    scan = f64_scan(cells, res, org, truth, 360, 2 * math.pi, 10.0).astype(np.float32)
    gl = GlobalLoc(cells, res, org, GlParams(hypotheses=nh))
    gl.localize(scan)  # warm-up
    ok = 0
    t0 = time.perf_counter()
    for i in range(runs):
        gl.set_seed(100 + i)
        top = gl.localize(scan)
        ok += math.hypot(top[0, 0] - truth[0], top[0, 1] - truth[1]) < 0.5
    ms = (time.perf_counter() - t0) * 1000.0 / runs
    top = gl.localize(scan)
    err = math.hypot(top[0, 0] - truth[0], top[0, 1] - truth[1])
    gl.close()
    return ms, err, ok


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--csv", action="store_true")
    ap.add_argument("--gl", action="store_true")
    args = ap.parse_args(argv)
    sizes = [1_000, 10_000, 100_000, 500_000]
    if args.csv:
        print("n,backend,ms_per_cycle,mean_track_err_m")
    else:
        print(f"{'N':>9} {'backend':<8} {'ms/cycle':>14} {'mean track err':>16}")
        print("-" * 52)
    for n in sizes:
        ms, err = run_pf(n)
        if args.csv:
            print(f"{n},hip,{ms:.4f},{err:.4f}")
        else:
            print(f"{n:>9} {'hip':<8} {ms:>14.3f} {err:>15.4f} m")
        sys.stdout.flush()
    if args.gl:
        print("\n== Global localization (CBGL-style, 10x10 m grid @0.05 m) ==")
        print(f"{'hyps':>10} {'backend':<8} {'ms/localize':>12} {'pos err m':>12} {'success':>10}")
        for nh in (5_000, 50_000, 200_000):
            ms, err, ok = run_gl(nh)
            print(f"{nh:>10} {'hip':<8} {ms:>12.1f} {err:>12.3f} {ok:>8}/20")
            sys.stdout.flush()
    return 0


if __name__ == "__main__":
    sys.exit(main())

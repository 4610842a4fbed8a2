"""GL -> PF hand-off (README:84-86) and checkpoint / resume through the host layer."""
import math

import numpy as np
import pytest

from cuda_pf_localization_amd import (GlobalLoc, GlParams, Observation, OdomDelta, ParticleFilter, PfConfig,
                                      quat_mean)
from cuda_pf_localization_amd.sim import f64_scan, free_poses, helix_scenario, obstacle_map

pytestmark = pytest.mark.gpu


def test_gl_to_pf_handoff(gpu):
    cells, res, org = obstacle_map(5)
    truth = free_poses(cells, res, org, 1, seed=8, margin_cells=6)[0]
    scan = (f64_scan(cells, res, org, truth, 360, 2 * math.pi, 8.0)
            + np.random.default_rng(2).normal(size=360) * 0.02).astype(np.float32)
    gl = GlobalLoc(cells, res, org, GlParams(hypotheses=200_000, rmax=8.0, topk=8, seed=4))
    top = gl.localize(scan)
    assert math.hypot(top[0, 0] - truth[0], top[0, 1] - truth[1]) < 0.5
    pf = ParticleFilter(PfConfig(particle_count=50_000, seed=3))
    pf.reinit_from_gl(top[0, :3], z=1.25, pos_std=(0.05, 0.05, 0.02), rot_std=(0.01, 0.01, 0.02))
    pos, quat, w = pf.snapshot()
    m = pos.mean(axis=0)
    assert abs(m[0] - top[0, 0]) < 2e-3 and abs(m[1] - top[0, 1]) < 2e-3 and abs(m[2] - 1.25) < 1e-3
    q = quat_mean(quat.T @ quat / len(quat))
    yaw = 2.0 * math.atan2(q[3], q[0])
    d = (yaw - top[0, 2] + math.pi) % (2 * math.pi) - math.pi
    assert abs(d) < 2e-3 and abs(q[1]) < 2e-3 and abs(q[2]) < 2e-3


def test_handoff_equals_oracle_reinit(gpu):
    """The hand-off is reinitialize() about (x, y, z) with the yaw as a rotation about z
    (README:84-86): bit-exact against the oracle's reinit (cpu.rs:180-195) with the same pose
    after the host layer's f32 casts (cuda.rs:187-205)."""
    from oracle import OracleFilter
    n = 100_000
    xyt = (4.2, 6.7, -2.3)
    ps, rs = (0.05, 0.05, 0.02), (0.01, 0.01, 0.02)
    pf = ParticleFilter(PfConfig(particle_count=n, seed=3))
    pf.reinit_from_gl(xyt, z=1.25, pos_std=ps, rot_std=rs)
    from cuda_pf_localization_amd import _lib
    pos = np.zeros((n, 3), np.float32)  # the stored f32 state (snapshot renormalises in f64)
    quat = np.zeros((n, 4), np.float32)
    w = np.zeros(n, np.float32)
    assert _lib.kernels().pfx_get_state(pf.kernels_handle(), *(a.ctypes.data_as(_lib.P(_lib.f32))
                                                               for a in (pos, quat, w))) == 0
    f32 = lambda v: np.asarray(v, np.float32).astype(np.float64)
    pose = f32([xyt[0], xyt[1], 1.25, math.cos(0.5 * xyt[2]), 0.0, 0.0, math.sin(0.5 * xyt[2])])
    o = OracleFilter(n, 3, True)
    o.reinit(pose, f32(ps), f32(rs))
    opos, oquat, ow, _ = o.state()
    assert np.array_equal(pos.astype(np.float64), opos) and np.array_equal(quat.astype(np.float64), oquat)
    assert np.all(w == np.float32(1.0 / n))


def test_particle_count_beyond_int_is_rejected(gpu):
    """The device filter indexes particles with int (pfc_init, cu:534): a larger count is an
    error, not a silently truncated filter."""
    with pytest.raises(RuntimeError):
        ParticleFilter(PfConfig(particle_count=2 ** 32 + 5, seed=1))


def test_particle_count_beyond_merge_path_limit_is_rejected(gpu):
    """The resampling merge path indexes 2n items with int: pfc_init refuses n > 2^30 before
    allocating, and the handle keeps its particles."""
    from cuda_pf_localization_amd import PfKernels
    k = PfKernels(16)
    k.init(1000, 3)
    with pytest.raises(RuntimeError):
        k.init(2 ** 30 + 1, 3)
    with pytest.raises(RuntimeError):
        k.init(0, 3)
    assert k.L.pfx_particle_count(k.h) == 1000
    pos, quat, w = k.get_state()
    assert len(pos) == 1000
    k.close()


def _steps(sc, k0, k1):
    out = []
    for st in sc.steps[k0:k1]:
        f32 = lambda v: [float(np.float32(x)) for x in v]
        out.append((OdomDelta(f32(st.odom_t), f32(st.odom_q), f32(st.trans_noise), f32(st.rot_noise)),
                    [Observation(i, m, v, r) for (i, m, v, r) in st.obs]))
    return out


def test_checkpoint_resume_is_bit_identical(gpu, tmp_path):
    sc = helix_scenario(24, (20.0, 20.0, 5.0), 12, 15)
    cfg = PfConfig(particle_count=30_000, seed=9, range_noise=0.05, bearing_noise=0.02, gate_sigma=3.0,
                   resample_ess_ratio=0.5, random_inject_ratio=0.02, fast=True)
    a = ParticleFilter(cfg)
    a.set_map(sc.landmarks)
    a.reinitialize([*sc.prior_pos, *sc.prior_q], [0.5] * 3, [0.1] * 3)
    for od, obs in _steps(sc, 0, 6):
        a.update(od, obs)
    path = tmp_path / "filter.ckpt"
    a.save(path)
    ra = [a.update(od, obs) for od, obs in _steps(sc, 6, 12)]
    b = ParticleFilter(cfg)
    b.set_map(sc.landmarks)
    b.reinitialize([0, 0, 0, 1, 0, 0, 0], [1.0] * 3, [1.0] * 3)  # overwritten by the checkpoint
    b.load(path)
    rb = [b.update(od, obs) for od, obs in _steps(sc, 6, 12)]
    assert any(e.resampled for e in ra)
    for ea, eb in zip(ra, rb):
        assert np.array_equal(ea.position, eb.position) and np.array_equal(ea.attitude, eb.attitude)
        assert np.array_equal(ea.pos_cov, eb.pos_cov) and np.array_equal(ea.att_cov, eb.att_cov)
        assert ea.resampled == eb.resampled
    small = ParticleFilter(PfConfig(particle_count=1000, seed=9))
    with pytest.raises(RuntimeError):
        small.load(path)


def test_pf_bench_circle_workload_matches_oracle(gpu):
    """The pf-bench PF workload (bench:21-60: circle scenario, seed 123, Cartesian observations,
    no gate, prior (1, 1, 1) m / (0.5, 0.5, 0.5) rad) through the host layer against the oracle
    driver (particle_filter.rs:110-124) at 10k particles for 40 cycles: resample decisions equal,
    means within 1e-5, covariances within 1e-4 relative."""
    from oracle import OBS_DTYPE, OracleFilter
    from cuda_pf_localization_amd.sim import circle_scenario
    f32 = lambda v: np.asarray(v, np.float32).astype(np.float64)
    n = 10_000
    sc = circle_scenario(40, 777)
    pf = ParticleFilter(PfConfig(particle_count=n, seed=123, range_noise=0.05))
    pf.set_map(sc.landmarks)
    prior = [*sc.prior_pos, *sc.prior_q]
    pf.reinitialize(prior, [1.0] * 3, [0.5] * 3)
    o = OracleFilter(n, 123, True)
    o.reinit(f32(prior), f32([1.0] * 3), f32([0.5] * 3))
    lm = np.array([sc.landmarks[i] for i in range(len(sc.landmarks))], np.float32)
    n_rs = 0
    for c, st in enumerate(sc.steps):
        eg = pf.update(OdomDelta(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise),
                       [Observation(i, m, v, r) for (i, m, v, r) in st.obs])
        o.predict(f32(st.odom_t), f32(st.odom_q), f32(st.trans_noise), f32(st.rot_noise))
        o.weight(np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], OBS_DTYPE), lm, float(np.float32(0.05)),
                 float(np.float32(0.02)), 0.0)
        eo = o.estimate()
        rs = eo["ess"] < 0.5 * n
        if rs:
            o.resample(np.float32(0.0))
            eo = o.estimate()
        assert eg.resampled == rs, c
        n_rs += rs
        assert np.all(np.abs(eg.position - eo["position"]) < 1e-5), c
        assert abs(eg.ess - eo["ess"]) <= 1e-4 * eo["ess"], c
        assert min(np.abs(eg.attitude - eo["attitude"]).max(), np.abs(eg.attitude + eo["attitude"]).max()) < 1e-5
        assert np.allclose(eg.pos_cov, eo["pos_cov"], rtol=1e-4, atol=1e-9), c
        assert np.allclose(eg.att_cov, eo["att_cov"], rtol=1e-4, atol=1e-9), c
    assert n_rs > 0


def test_pf_bench_cli_tables(gpu, capsys):
    """The pf-bench restatement prints the reference's table layout; the HIP filter tracks the
    circle scenario (bench:21-60) and the GL table localises the room_grid truth."""
    from cuda_pf_localization_amd import pf_bench
    ms, err = pf_bench.run_pf(10_000, steps_n=40)
    assert ms > 0 and err < 0.5
    ms, err, ok = pf_bench.run_gl(5_000, runs=3)
    assert ms > 0 and ok >= 2
    pf_bench.main(["--csv"])
    out = capsys.readouterr().out.splitlines()
    assert out[0] == "n,backend,ms_per_cycle,mean_track_err_m"
    assert [line.split(",")[0] for line in out[1:]] == ["1000", "10000", "100000", "500000"]

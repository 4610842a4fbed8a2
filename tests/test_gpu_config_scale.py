"""The two large configurations at their full size, against the CPU oracle.

Config 1: 1,000,000 particles, 256 landmarks in a 50 x 50 x 10 m box, 256 range-bearing
observations per cycle with 1 % range outliers U[0, 60] m, gate 3 sigma, inject 0.01, resample
when ESS < N/2 (particle_filter.rs:111-124, cpu.rs:220-305).  Three drivers are checked:
  * the reference ABI stage by stage (pfc_predict / weight / est_sums / covs / resample), with
    the oracle given the GPU's f32 weights before each resample, so that the contract's
    "identical weights and offset" holds exactly: poses, ancestors and injection picks
    bit-exact, log-likelihoods within 1e-9 + 1e-12 |ll| plus the truncation bound of the gated
    range-bearing path's 3-term bearing series (series_bound), weights within 1e-5 relative, estimate
    sums and covariances within 1e-9 of the oracle's f64 sums;
  * the C++ host layer (particle_filter.rs:110-124) in the reference's exact call order and in
    the fast order, against an oracle driver that runs its own predict, likelihood and estimate
    and only takes, for its resampling, the weights the GPU's scan used (recovered exactly from
    the scan's fixed-point prefix, pfx_get_cum, after a 1e-5 check against its own);
  * the fused cycle (pfb_cycle, hipGraph), against the same oracle driver.
Config 3: the 64-filter batch (N log-uniform in [1k, 500k], L uniform in [4, 256]) in one
launch per stage; the smallest, the largest, the fewest-landmark filter and a filter given 4
landmarks each run beside their own oracle.  Each cycle the weights the GPU's resampling scan
used (recovered exactly from its fixed-point prefix) are checked against the oracle's within
1e-5 relative and then given to the oracle, so ancestors are compared on identical weights
(weights within 1e-5 can still move an ancestor pointer across a particle boundary).
"""
import math

import numpy as np
import pytest

from cuda_pf_localization_amd import OBS_DTYPE, PfKernels
from oracle import OracleFilter, quat_mean as oracle_quat_mean
from test_gpu_pf_parity import series_bound

pytestmark = pytest.mark.gpu

N1 = 1_000_000


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def obs_array(st):
    return np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], OBS_DTYPE)


def sums_to_M(s):
    M = np.zeros((4, 4))
    e = 3
    for r in range(4):
        for c in range(r, 4):
            M[r, c] = M[c, r] = s[e]
            e += 1
    return M


def assert_poses_equal(pos, quat, opos, oquat, what):
    dp = np.count_nonzero(pos.astype(np.float64) != opos)
    dq = np.count_nonzero(quat.astype(np.float64) != oquat)
    assert dp == 0 and dq == 0, f"{what}: {dp} position and {dq} quaternion mismatches"


def assert_weights_close(w, ow, what):
    w = w.astype(np.float64)
    dw = np.abs(w - ow)
    bad = dw > 1e-5 * ow + 1e-38
    assert not bad.any(), f"{what}: {bad.sum()} weights off by more than 1e-5 relative, worst {dw.max():.3e}"


def scan_weights(cum):
    """The weights a resampling scan used, recovered exactly from its inclusive fixed-point
    prefix: q = floor(w 2^56) of an f32 weight has at most 24 significant bits."""
    return np.diff(cum, prepend=np.uint64(0)).astype(np.float64) * 2.0 ** -56


def assert_scan_weights_close(w_used, ow, what):
    big = ow >= 2.0 ** -32
    assert np.all(np.abs(w_used[big] - ow[big]) <= 1e-5 * ow[big]), f"{what}: scan weights"
    assert np.all(np.abs(w_used[~big] - ow[~big]) <= 2.0 ** -56 + 1e-5 * ow[~big]), f"{what}: scan weights"


def oracle_cycle_on_gpu_weights(o, st, lm, sr, sb, gate, ratio, inject, n, gpu_cum, what):
    """oracle_cycle with the resampling run on the weights the GPU's scan used (gpu_cum() reads
    its prefix), after checking them against the oracle's own within 1e-5: two independent f64
    likelihoods agree to ~1e-8, and weights that close can still move a systematic pointer
    across a particle boundary, so ancestors are compared on identical weights and offset."""
    o.predict(f32(st.odom_t), f32(st.odom_q), f32(st.trans_noise), f32(st.rot_noise))
    o.weight(obs_array(st), lm, sr, sb, gate)
    e = o.estimate()
    if not (ratio > 0 and e["ess"] < ratio * n):
        return e, False
    opos, oquat, ow, _ = o.state()
    w_used = scan_weights(gpu_cum())
    assert_scan_weights_close(w_used, ow, what)
    o.set_state(opos, oquat, w_used)
    o.resample(np.float32(inject))
    return o.estimate(), True


def oracle_cycle(o, st, lm, sr, sb, gate, ratio, inject, n):
    """particle_filter.rs:110-124 on the oracle: predict, weight, estimate, then resample and
    estimate again when ESS < ratio * N."""
    o.predict(f32(st.odom_t), f32(st.odom_q), f32(st.trans_noise), f32(st.rot_noise))
    o.weight(obs_array(st), lm, sr, sb, gate)
    e = o.estimate()
    rs = False
    if ratio > 0 and e["ess"] < ratio * n:
        o.resample(np.float32(inject))
        e = o.estimate()
        rs = True
    return e, rs


def compare_estimate(pos, quat, pos_cov, att_cov, ess, eo, what):
    assert np.all(np.abs(pos - eo["position"]) <= 1e-5 * np.abs(eo["position"]) + 1e-7), what
    assert abs(ess - eo["ess"]) <= 1e-5 * eo["ess"], f"{what}: ess {ess} vs {eo['ess']}"
    dq = min(np.abs(quat - eo["attitude"]).max(), np.abs(quat + eo["attitude"]).max())
    assert dq < 1e-5, f"{what}: attitude {quat} vs {eo['attitude']}"
    for g, c in ((pos_cov, eo["pos_cov"]), (att_cov, eo["att_cov"])):
        scale = np.abs(np.diag(c)).max()
        assert np.all(np.abs(g - c) <= 1e-6 * scale + 1e-12), f"{what}: covariance {g} vs {c}"


@pytest.fixture(scope="module")
def config1():
    from cuda_pf_localization_amd.sim import config1_scenario
    sc = config1_scenario(cycles=4, seed=7)
    lm = np.array([sc.landmarks[i] for i in range(len(sc.landmarks))], np.float32)
    return sc, lm


def test_config1_stagewise_abi_1M(gpu, config1):
    sc, lm = config1
    sr, sb, gate = float(np.float32(sc.sigma_r)), float(np.float32(sc.sigma_b)), float(np.float32(sc.gate))
    k = PfKernels(256)
    k.init(N1, 7)
    o = OracleFilter(N1, 7, True)
    k.upload_landmarks(lm)
    prior = f32([*sc.prior_pos, *sc.prior_q])
    k.reinit(prior, f32([0.5] * 3), f32([0.1] * 3))
    o.reinit(prior, f32([0.5] * 3), f32([0.1] * 3))
    n_rs = 0
    for c, st in enumerate(sc.steps):
        print(f"config 1 stage-wise cycle {c}", flush=True)
        k.predict(f32(st.odom_t), f32(st.odom_q), f32(st.trans_noise), f32(st.rot_noise))
        o.predict(f32(st.odom_t), f32(st.odom_q), f32(st.trans_noise), f32(st.rot_noise))
        pos, quat, _ = k.get_state()
        opos, oquat, _, _ = o.state()
        assert_poses_equal(pos, quat, opos, oquat, f"cycle {c} predict")
        obs = obs_array(st)
        k.weight(obs, sc.sigma_r, sc.sigma_b, sc.gate)
        o.weight(obs, lm, sr, sb, gate)
        ll = k.get_ll()
        _, _, w = k.get_state()
        _, _, ow, oll = o.state()
        dll = np.abs(ll - oll)
        tol = 1e-9 + 1e-12 * np.abs(oll) + series_bound(obs, sb, gate)
        assert np.all(dll <= tol), f"cycle {c}: max |dll| {dll.max():.3e}"
        assert_weights_close(w, ow, f"cycle {c} weights")
        # identical weights from here on: the oracle takes the GPU's f32 weights
        o.set_state(opos, oquat, w.astype(np.float64))
        pre_pos = opos.copy()
        s, os_ = k.est_sums(), o.est_sums()
        assert np.all(np.abs(s - os_) <= 1e-9 * np.abs(os_).max() + 1e-300), f"cycle {c} sums"
        mq = oracle_quat_mean(sums_to_M(os_))
        mp = f32(os_[:3])
        cv, ocv = k.covs(mp, f32(mq)), o.covs(mp, f32(mq))
        for a, b in ((slice(0, 6), (0, 3, 5)), (slice(6, 12), (6, 9, 11))):
            scale = max(abs(ocv[i]) for i in b)
            assert np.all(np.abs(cv[a] - ocv[a]) <= 1e-9 * scale), f"cycle {c} covariances"
        ess = 1.0 / os_[13]
        print(f"  ESS {ess:.1f}", flush=True)
        if ess < 0.5 * N1:
            lcg0 = k.get_lcg()
            assert lcg0 == o.rng()[2]
            k.resample(0.01)
            o.resample(np.float32(0.01))
            assert k.get_lcg() == o.rng()[2] != lcg0
            anc, oanc = k.ancestors(), o.ancestors()
            assert np.array_equal(anc, oanc), f"cycle {c}: {np.count_nonzero(anc != oanc)} ancestor mismatches"
            pos, quat, w = k.get_state()
            opos, oquat, ow, _ = o.state()
            assert_poses_equal(pos, quat, opos, oquat, f"cycle {c} resample + inject")
            assert np.all(w == np.float32(1.0 / N1))
            # the 10,000 injection picks moved that many poses (fewer with duplicate picks) off
            # their ancestors' positions, identically on both sides (bit-exact above)
            moved = np.count_nonzero(np.any(pos.astype(np.float64) != pre_pos[anc], axis=1))
            assert 9_000 <= moved <= 10_000, moved
            n_rs += 1
    assert n_rs >= 1


@pytest.mark.parametrize("fast", [False, True])
def test_config1_host_layer_vs_oracle_driver(gpu, config1, fast):
    from cuda_pf_localization_amd import Observation, OdomDelta, ParticleFilter, PfConfig
    from cuda_pf_localization_amd import _lib
    sc, lm = config1
    cfg = PfConfig(particle_count=N1, seed=7, range_noise=sc.sigma_r, bearing_noise=sc.sigma_b, gate_sigma=sc.gate,
                   resample_ess_ratio=0.5, random_inject_ratio=0.01, fast=fast)
    pf = ParticleFilter(cfg)
    pf.set_map(sc.landmarks)
    prior = [*sc.prior_pos, *sc.prior_q]
    pf.reinitialize(prior, [0.5] * 3, [0.1] * 3)
    o = OracleFilter(N1, 7, True)
    o.reinit(f32(prior), f32([0.5] * 3), f32([0.1] * 3))
    L = _lib.kernels()
    kh = pf.kernels_handle()
    sr, sb, gate = (float(np.float32(v)) for v in (sc.sigma_r, sc.sigma_b, sc.gate))
    n_rs = 0
    for c, st in enumerate(sc.steps[:3]):
        print(f"config 1 host layer (fast={fast}) cycle {c}", flush=True)
        eg = pf.update(OdomDelta(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise),
                       [Observation(i, m, v, r) for (i, m, v, r) in st.obs])
        def gpu_cum():
            c = np.zeros(N1, np.uint64)
            assert L.pfx_get_cum(kh, c.ctypes.data_as(_lib.P(_lib.u64))) == 0
            return c
        eo, rs = oracle_cycle_on_gpu_weights(o, st, lm, sr, sb, gate, 0.5, 0.01, N1, gpu_cum, f"cycle {c}")
        assert eg.resampled == rs
        if rs:
            anc = np.zeros(N1, np.int32)
            assert L.pfx_get_ancestors(kh, anc.ctypes.data_as(_lib.P(_lib.i32))) == 0
            assert np.array_equal(anc, o.ancestors()), f"cycle {c}: ancestors differ"
            n_rs += 1
        compare_estimate(eg.position, eg.attitude, eg.pos_cov, eg.att_cov, eg.ess, eo, f"cycle {c}")
    # the device state itself (the host snapshot renormalises each quaternion in f64 like
    # cuda.rs:343-356, so its values are not the stored f32 ones)
    pos = np.zeros((N1, 3), np.float32)
    quat = np.zeros((N1, 4), np.float32)
    assert L.pfx_get_state(kh, pos.ctypes.data_as(_lib.P(_lib.f32)), quat.ctypes.data_as(_lib.P(_lib.f32)),
                           None) == 0
    opos, oquat, _, _ = o.state()
    assert_poses_equal(pos, quat, opos, oquat, "final state")
    assert n_rs >= 1


def test_config1_fused_cycle_vs_oracle_driver(gpu, config1):
    from cuda_pf_localization_amd import BatchedFilters
    sc, lm = config1
    bf = BatchedFilters([dict(n=N1, seed=7, sigma_r=sc.sigma_r, sigma_b=sc.sigma_b, gate=sc.gate, ratio=0.5,
                              inject=0.01)], obs_cap=256)
    bf.upload_landmarks(0, lm)
    prior = [*sc.prior_pos, *sc.prior_q]
    bf.reinit(np.array([prior], np.float32), np.full((1, 3), 0.5, np.float32), np.full((1, 3), 0.1, np.float32))
    bf.use_graph(True)
    o = OracleFilter(N1, 7, True)
    o.reinit(f32(prior), f32([0.5] * 3), f32([0.1] * 3))
    sr, sb, gate = (float(np.float32(v)) for v in (sc.sigma_r, sc.sigma_b, sc.gate))
    n_rs = 0
    for c, st in enumerate(sc.steps[:3]):
        print(f"config 1 fused cycle {c}", flush=True)
        inp, _keep = bf.make_inputs([(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise, obs_array(st))])
        e = bf.cycle(inp)[0]
        eo, rs = oracle_cycle_on_gpu_weights(o, st, lm, sr, sb, gate, 0.5, 0.01, N1, lambda: bf.cum(0),
                                             f"fused cycle {c}")
        assert bool(e.resampled) == rs
        if rs:
            assert np.array_equal(bf.ancestors(0), o.ancestors()), f"cycle {c}: ancestors differ"
            n_rs += 1
        compare_estimate(np.array(e.pos[:]), np.array(e.quat[:]), np.array(e.pos_cov[:]).reshape(3, 3),
                         np.array(e.att_cov[:]).reshape(3, 3), e.ess, eo, f"fused cycle {c}")
        pos, quat, w = bf.get_state(0)
        opos, oquat, ow, _ = o.state()
        assert_poses_equal(pos, quat, opos, oquat, f"fused cycle {c} state")
        assert_weights_close(w, ow, f"fused cycle {c} weights")
    assert n_rs >= 1
    bf.close()


def test_config3_64_filters_vs_oracle(gpu):
    """The 64-filter batch (4.3M particles, one launch per stage, hipGraph replay, autotuned
    likelihood shape) for 5 cycles; four filters are each run beside their own oracle."""
    from cuda_pf_localization_amd import BatchedFilters
    from cuda_pf_localization_amd.sim import config3_filters, helix_arrays
    nf, cycles = 64, 5
    ns, ls, hs = config3_filters(nf, cycles, 11)
    # the lower end of the landmark range, which this draw does not reach, on filter 0
    ls = ls.copy()
    ls[0] = 4
    hs[0] = helix_arrays(4, (20.0, 20.0, 5.0), cycles, 11 * 1000 + 0, phase=0.3)
    picks = sorted({int(np.argmin(ns)), int(np.argmax(ns)), int(np.argmin(np.where(np.arange(nf) == 0, 999, ls))), 0})
    cfgs = [dict(n=int(ns[f]), seed=11 + f, sigma_r=0.05, sigma_b=0.02, gate=3.0, ratio=0.5, inject=0.0)
            for f in range(nf)]
    bf = BatchedFilters(cfgs, obs_cap=256)
    for f, h in enumerate(hs):
        bf.upload_landmarks(f, h.landmarks.astype(np.float32))
    pos_std, rot_std = np.full((nf, 3), 0.5, np.float32), np.full((nf, 3), 0.1, np.float32)
    poses = np.array([[*h.prior_pos, *h.prior_q] for h in hs], np.float32)
    bf.reinit(poses, pos_std, rot_std)

    def obs_of(h, kc):
        a = np.zeros(len(h.landmarks), OBS_DTYPE)
        a["id"] = np.arange(len(h.landmarks))
        a["mode"] = 1
        a["dx"], a["dy"], a["dz"] = h.obs_v[kc, :, 0], h.obs_v[kc, :, 1], h.obs_v[kc, :, 2]
        a["range"] = h.obs_r[kc]
        return a

    inputs = [bf.make_inputs([(h.odom_t[kc], h.odom_q[kc], (h.odo_pos,) * 3, (h.odo_rot,) * 3, obs_of(h, kc))
                              for h in hs]) for kc in range(cycles)]
    bf.use_graph(True)
    ppt = bf.autotune(inputs[0][0])
    print(f"config 3: {ns.sum()} particles, filters checked {picks} (N {[int(ns[f]) for f in picks]}, "
          f"L {[int(ls[f]) for f in picks]}), particles per lane {ppt}", flush=True)
    # (the autotuner times the likelihood pass only and restores the draw state it staged)
    oracles = {}
    for f in picks:
        o = OracleFilter(int(ns[f]), 11 + f, True)
        o.reinit(f32(poses[f]), f32(pos_std[f]), f32(rot_std[f]))
        oracles[f] = o
    for kc in range(cycles):
        print(f"config 3 cycle {kc}", flush=True)
        out = bf.cycle(inputs[kc][0])
        for f in picks:
            h = hs[f]
            o = oracles[f]
            o.predict(f32(h.odom_t[kc]), f32(h.odom_q[kc]), f32([h.odo_pos] * 3), f32([h.odo_rot] * 3))
            o.weight(obs_of(h, kc), h.landmarks.astype(np.float32), float(np.float32(0.05)),
                     float(np.float32(0.02)), 3.0)
            opos, oquat, ow, _ = o.state()
            eo = o.estimate()
            rs = eo["ess"] < 0.5 * ns[f]
            e = out[f]
            what = f"filter {f} (N {ns[f]}, L {ls[f]}) cycle {kc}"
            assert bool(e.resampled) == rs, what
            # the weights the GPU's resampling scan used, recovered exactly from its fixed-point
            # prefix (q = floor(w 2^56) of an f32 weight has at most 24 significant bits)
            w_used = scan_weights(bf.cum(f))
            assert_scan_weights_close(w_used, ow, what)
            pos, quat, w = bf.get_state(f)
            if rs:
                # identical weights and offset: the oracle resamples the GPU's weights
                o.set_state(opos, oquat, w_used)
                o.resample(np.float32(0.0))
                eo = o.estimate()
                assert np.array_equal(bf.ancestors(f), o.ancestors()), f"{what}: ancestors differ"
                opos, oquat, ow, _ = o.state()
                assert_poses_equal(pos, quat, opos, oquat, what)
                assert np.all(w == np.float32(1.0 / ns[f]))
            else:
                assert_poses_equal(pos, quat, opos, oquat, what)
                assert_weights_close(w, ow, what)
                o.set_state(opos, oquat, w.astype(np.float64))  # the next cycle starts from identical weights
            compare_estimate(np.array(e.pos[:]), np.array(e.quat[:]), np.array(e.pos_cov[:]).reshape(3, 3),
                             np.array(e.att_cov[:]).reshape(3, 3), e.ess, eo, what)
    bf.close()

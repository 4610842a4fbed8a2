"""GPU parity of the particle-filter path (pfc_* ABI) against the CPU oracle.

Every test calls the HIP library through its C ABI and compares with oracle/pf_oracle.c
(the C restatement of crates/pf_core/src/cpu.rs) on identical inputs:
  * reinit / predict poses          bit-exact (deterministic-draw mode, f32 storage)
  * log-likelihoods                 |dll| <= 1e-9 + 1e-12 |ll|, plus on the gated range-bearing
                                    path the truncation bound of its 3-term bearing series
                                    (series_bound: below 1e-7 for 256 observations)
  * weights                         |dw| <= 1e-6 * w_max  (1e-5)
  * resampling ancestors, poses     bit-exact for identical weights and LCG state
  * injection picks                 bit-exact
  * est sums / covariances          relative 1e-9 (contract: covariances within 1e-6)
"""
import math

import numpy as np
import pytest

from cuda_pf_localization_amd import OBS_DTYPE, PfKernels
from oracle import OracleFilter, quat_mean as oracle_quat_mean

pytestmark = pytest.mark.gpu


def f32(x):
    return np.asarray(x, np.float32).astype(np.float64)


def random_state(n, seed, box=20.0, spread_q=True):
    rng = np.random.default_rng(seed)
    pos = (rng.random((n, 3)) * box).astype(np.float32)
    if spread_q:
        q = rng.normal(size=(n, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        q[q[:, 0] < 0] *= -1
    else:
        q = np.tile([1.0, 0, 0, 0], (n, 1))
    return pos, q.astype(np.float32)


def pair(n, seed=42):
    k = PfKernels(256)
    k.init(n, seed)
    o = OracleFilter(n, seed, True)
    return k, o


def assert_state_equal(k, o, what=""):
    pos, quat, w = k.get_state()
    opos, oquat, ow, _ = o.state()
    dp = np.nonzero(pos.astype(np.float64) != opos)
    assert dp[0].size == 0, f"{what}: {dp[0].size} position mismatches, first {dp[0][:5]}"
    dq = np.nonzero(quat.astype(np.float64) != oquat)
    assert dq[0].size == 0, f"{what}: {dq[0].size} quaternion mismatches"
    return pos, quat, w, opos, oquat, ow


@pytest.mark.parametrize("rot_std", [(0.05, 0.07, 0.02), (0.5, 0.5, 0.5), (2.0, 3.0, 1.5), (0.0, 0.0, 0.0)])
def test_reinit_bitexact(gpu, rot_std):
    n = 10_000
    k, o = pair(n, 42)
    pose = f32([1.5, -2.0, 0.5, math.cos(0.3), 0.0, 0.0, math.sin(0.3)])
    ps = f32([1.0, 2.0, 0.5])
    rs = f32(rot_std)
    k.reinit(pose, ps, rs)
    o.reinit(pose, ps, rs)
    assert k.rng()[0] == o.rng()[0] and k.rng()[1] == o.rng()[1]
    _, _, w, _, _, ow = assert_state_equal(k, o, "reinit")
    assert np.all(w == np.float32(1.0 / n))


def test_predict_bitexact_chain(gpu):
    n = 20_000
    k, o = pair(n, 7)
    pose = f32([3.0, 4.0, 1.0, 1.0, 0.0, 0.0, 0.0])
    k.reinit(pose, f32([0.5, 0.5, 0.5]), f32([0.1, 0.1, 0.1]))
    o.reinit(pose, f32([0.5, 0.5, 0.5]), f32([0.1, 0.1, 0.1]))
    rng = np.random.default_rng(3)
    for step in range(5):
        t = f32(rng.normal(size=3) * 0.5)
        ang = rng.normal() * 0.2
        q = f32([math.cos(ang / 2), 0.0, 0.0, math.sin(ang / 2)])
        ts = f32([0.02, 0.03, 0.01])
        rs = f32([0.01, 0.02, 0.03])
        k.predict(t, q, ts, rs)
        o.predict(t, q, ts, rs)
        assert_state_equal(k, o, f"predict step {step}")


def test_predict_bitexact_random_state(gpu):
    n = 5000
    pos, quat = random_state(n, 11)
    k, o = pair(n, 99)
    k.set_state(pos, quat)
    o.set_state(pos.astype(np.float64), quat.astype(np.float64))
    t, q = f32([0.3, -0.1, 0.05]), f32([0.99, 0.1, -0.05, 0.02])
    k.predict(t, q, f32([0.05, 0.05, 0.05]), f32([0.02, 0.02, 0.03]))
    o.predict(t, q, f32([0.05, 0.05, 0.05]), f32([0.02, 0.02, 0.03]))
    assert_state_equal(k, o, "predict random")


def make_landmarks(nl, seed, box=(20.0, 20.0, 5.0)):
    rng = np.random.default_rng(seed)
    return (rng.random((nl, 3)) * np.array(box)).astype(np.float32)


def rb_obs_for(truth_pos, truth_q, lm, seed, sigma_r=0.05, sigma_b=0.02, outliers=0.0):
    """Range-bearing observations (v = unit landmark->sensor vector in the body frame)."""
    from cuda_pf_localization_amd.sim import qconj, qrot
    rng = np.random.default_rng(seed)
    obs = np.zeros(len(lm), OBS_DTYPE)
    for i, p in enumerate(lm.astype(np.float64)):
        b = np.array(qrot(qconj(truth_q), tuple(p - truth_pos)))
        d = np.linalg.norm(b)
        v = -b / d + rng.normal(size=3) * sigma_b
        v /= np.linalg.norm(v)
        r = d + rng.normal() * sigma_r
        if outliers and rng.random() < outliers:
            r = rng.random() * 60.0
        obs[i] = (i, 1, v[0], v[1], v[2], r)
    return obs


def cart_obs_for(truth_pos, truth_q, lm, seed, sigma=0.05):
    from cuda_pf_localization_amd.sim import qconj, qrot
    rng = np.random.default_rng(seed)
    obs = np.zeros(len(lm), OBS_DTYPE)
    for i, p in enumerate(lm.astype(np.float64)):
        b = np.array(qrot(qconj(truth_q), tuple(p - truth_pos))) + rng.normal(size=3) * sigma
        obs[i] = (i, 0, b[0], b[1], b[2], 0.0)
    return obs


def weight_pair(n, obs, lm, sr, sb, gate, seed=5, spread=1.0):
    k, o = pair(n, seed)
    truth = np.array([10.0, 10.0, 2.5])
    pose = f32([*truth, math.cos(0.4), 0.0, 0.0, math.sin(0.4)])
    k.reinit(pose, f32([spread] * 3), f32([0.1, 0.1, 0.3]))
    o.reinit(pose, f32([spread] * 3), f32([0.1, 0.1, 0.3]))
    k.upload_landmarks(lm)
    k.weight(obs, sr, sb, gate)
    o.weight(obs, lm, float(np.float32(sr)), float(np.float32(sb)), float(np.float32(gate)))
    return k, o


def series_bound(obs, sb, gate):
    """Largest log-likelihood change the gated range-bearing path's 3-term acos^2 series can
    make: its first omitted term, (32/1120) u^4, at u <= u_gate ~ 1 - cos(gate sb), times
    1 / (2 sb^2), per range-bearing observation (10 % margin for the kernel's bisected u_gate);
    zero where that path is off (no gate, or u_gate above 0.005)."""
    sb = float(np.float32(sb))
    if not gate > 0.0:
        return 0.0
    ug = 1.0 - math.cos(float(gate) * sb)
    if ug > 0.005:
        return 0.0
    nrb = int(np.sum(np.asarray(obs)["mode"] == 1))
    return 1.1 * nrb * (32.0 / 1120.0) * ug ** 4 / (2.0 * sb * sb)


def check_weights(k, o, ll_check=True, bound=0.0):
    _, _, w = k.get_state()
    _, _, ow, oll = o.state()
    if ll_check:
        ll = k.get_ll()
        dll = np.abs(ll - oll)
        tol = 1e-9 + 1e-12 * np.abs(oll) + bound
        assert np.all(dll <= tol), f"max |dll| {dll.max():.3e} at {np.argmax(dll - tol)}"
    dw = np.abs(w.astype(np.float64) - ow)
    assert dw.max() <= 1e-6 * ow.max(), f"max |dw| {dw.max():.3e} vs w_max {ow.max():.3e}"
    assert abs(w.astype(np.float64).sum() - 1.0) < 1e-5
    return w, ow


@pytest.mark.parametrize("gate", [0.0, 3.0])
@pytest.mark.parametrize("spread", [0.05, 1.0])
def test_weight_range_bearing_parity(gpu, gate, spread):
    n = 20_000
    lm = make_landmarks(12, 1)
    q_t = (math.cos(0.4), 0.0, 0.0, math.sin(0.4))
    obs = rb_obs_for(np.array([10.0, 10.0, 2.5]), q_t, lm, 2)
    k, o = weight_pair(n, obs, lm, 0.05, 0.02, gate, spread=spread)
    check_weights(k, o, bound=series_bound(obs, 0.02, gate))


@pytest.mark.parametrize("gate", [0.0, 3.0])
def test_weight_cartesian_parity(gpu, gate):
    n = 20_000
    lm = make_landmarks(8, 4)
    q_t = (math.cos(0.4), 0.0, 0.0, math.sin(0.4))
    obs = cart_obs_for(np.array([10.0, 10.0, 2.5]), q_t, lm, 5)
    k, o = weight_pair(n, obs, lm, 0.05, 0.02, gate, spread=0.3)
    check_weights(k, o)


def test_weight_256_obs_outliers_parity(gpu):
    """Config-1 observation model (256 RB observations, 1 % outliers, gate 3) at 50k particles."""
    n = 50_000
    lm = make_landmarks(256, 9, box=(50.0, 50.0, 10.0))
    q_t = (math.cos(1.0), 0.0, 0.0, math.sin(1.0))
    truth = np.array([25.0, 25.0, 5.0])
    obs = rb_obs_for(truth, q_t, lm, 10, outliers=0.01)
    k, o = pair(n, 77)
    pose = f32([*truth, *q_t])
    k.reinit(pose, f32([0.5] * 3), f32([0.1] * 3))
    o.reinit(pose, f32([0.5] * 3), f32([0.1] * 3))
    k.upload_landmarks(lm)
    k.weight(obs, 0.05, 0.02, 3.0)
    o.weight(obs, lm, float(np.float32(0.05)), float(np.float32(0.02)), 3.0)
    check_weights(k, o, bound=series_bound(obs, 0.02, 3.0))


def test_weight_mixed_modes_and_unknown_ids(gpu):
    n = 4096
    lm = make_landmarks(6, 3)
    q_t = (1.0, 0.0, 0.0, 0.0)
    truth = np.array([10.0, 10.0, 2.5])
    rb = rb_obs_for(truth, q_t, lm, 1)
    ca = cart_obs_for(truth, q_t, lm, 2)
    obs = np.concatenate([rb[:3], ca[3:], np.array([(17, 0, 1, 0, 0, 0), (-1, 1, 0, 0, 1, 3)], OBS_DTYPE)])
    k, o = weight_pair(n, obs, lm, 0.05, 0.02, 0.0, spread=0.5)
    check_weights(k, o)


def test_weight_empty_and_unknown_reset_uniform(gpu):
    n = 8192
    lm = make_landmarks(1, 0)
    obs = np.array([(0, 0, 0.0, 0.0, 0.0, 0.0)], OBS_DTYPE)
    k, o = weight_pair(n, obs, lm, 0.05, 0.02, 0.0, spread=3.0)
    _, _, w = k.get_state()
    assert 1.0 / np.sum(w.astype(np.float64) ** 2) < 0.5 * n  # informative obs collapses ESS
    k.weight(None, 0.05, 0.02, 0.0)
    _, _, w = k.get_state()
    assert np.all(w == np.float32(1.0 / n))
    k.weight(np.array([(42, 0, 1.0, 0, 0, 0)], OBS_DTYPE), 0.05, 0.02, 0.0)  # id outside the table
    _, _, w = k.get_state()
    assert np.all(w == np.float32(1.0 / n))


def test_weight_obs_cap_truncates(gpu):
    n = 2048
    lm = make_landmarks(6, 8)
    obs = rb_obs_for(np.array([10.0, 10.0, 2.5]), (1.0, 0, 0, 0), lm, 9)
    k = PfKernels(4)  # cap 4: only the first four observations count (pf_kernels.cu:604)
    k.init(n, 3)
    o = OracleFilter(n, 3, True)
    pose = f32([10.0, 10.0, 2.5, 1.0, 0, 0, 0])
    k.reinit(pose, f32([0.5] * 3), f32([0.05] * 3))
    o.reinit(pose, f32([0.5] * 3), f32([0.05] * 3))
    k.upload_landmarks(lm)
    k.weight(obs, 0.05, 0.02, 0.0)
    o.weight(obs[:4], lm, float(np.float32(0.05)), float(np.float32(0.02)), 0.0)
    check_weights(k, o)


def _weight_case(case, seed):
    q_t = (math.cos(0.4), 0.0, 0.0, math.sin(0.4))
    truth = np.array([10.0, 10.0, 2.5])
    if case == "config1":
        lm = make_landmarks(256, seed, box=(50.0, 50.0, 10.0))
        truth = np.array([25.0, 25.0, 5.0])
        return lm, rb_obs_for(truth, q_t, lm, seed + 1, outliers=0.01), truth, q_t, 3.0, 0.5
    lm = make_landmarks(37, seed)
    rb = rb_obs_for(truth, q_t, lm, seed + 1)
    ca = cart_obs_for(truth, q_t, lm, seed + 2)
    if case == "rb_gated":
        return lm, rb, truth, q_t, 3.0, 1.0
    if case == "rb_wide":  # ungated, wide cloud: bearing residuals beyond the 6-term series
        return lm, rb, truth, q_t, 0.0, 3.0
    if case == "cart":
        return lm, ca, truth, q_t, 3.0, 0.5
    mixed = np.concatenate([ca[:5], rb[5:20], ca[20:], rb[:5], np.array([(99, 1, 0, 0, 1, 3)], OBS_DTYPE)])
    return lm, mixed, truth, q_t, 0.0, 0.5


@pytest.mark.parametrize("variant", [1, 2, 4])
@pytest.mark.parametrize("case", ["rb_gated", "rb_wide", "cart", "mixed", "config1"])
def test_weight_variants_match_oracle(gpu, variant, case):
    """Every launch shape of the likelihood kernel (1, 2, 4 particles per lane) against the
    oracle, for range-bearing (gated and ungated, wide clouds beyond the short bearing
    series), Cartesian and mixed observation lists."""
    n = 30_000 if case == "config1" else 5000
    lm, obs, truth, q_t, gate, spread = _weight_case(case, 11)
    k, o = pair(n, 21)
    k.set_weight_variant(variant)
    pose = f32([*truth, *q_t])
    k.reinit(pose, f32([spread] * 3), f32([0.1, 0.1, 0.3]))
    o.reinit(pose, f32([spread] * 3), f32([0.1, 0.1, 0.3]))
    k.upload_landmarks(lm)
    k.weight(obs, 0.05, 0.02, gate)
    o.weight(obs, lm, float(np.float32(0.05)), float(np.float32(0.02)), float(np.float32(gate)))
    ll = k.get_ll()
    _, _, ow, oll = o.state()
    dll = np.abs(ll - oll)
    print(f"variant {variant} {case}: max |dll| {dll.max():.3e} (|ll| max {np.abs(oll).max():.3e})")
    assert np.all(dll <= 1e-9 + 1e-12 * np.abs(oll) + series_bound(obs, 0.02, gate)), f"max |dll| {dll.max():.3e}"
    _, _, w = k.get_state()
    assert np.abs(w.astype(np.float64) - ow).max() <= 1e-6 * ow.max()


@pytest.mark.parametrize("case", ["rb_gated", "rb_wide", "mixed", "config1"])
def test_weight_shapes_bit_identical(gpu, case):
    """The launch shapes (1, 2, 4 particles per lane) give the same log-likelihoods bit for bit
    (each particle's terms are summed in the same two chains in every shape)."""
    n = 30_001 if case == "config1" else 5003
    lm, obs, truth, q_t, gate, spread = _weight_case(case, 13)
    res = []
    for variant in (1, 2, 4):
        k, _ = pair(n, 21)
        k.set_weight_variant(variant)
        pose = f32([*truth, *q_t])
        k.reinit(pose, f32([spread] * 3), f32([0.1, 0.1, 0.3]))
        k.upload_landmarks(lm)
        k.weight(obs, 0.05, 0.02, gate)
        res.append(k.get_ll())
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])


def weighted_state(n, seed, conc=1.0):
    pos, quat = random_state(n, seed)
    rng = np.random.default_rng(seed + 1)
    ll = -rng.gamma(conc, 4.0, size=n)
    w = np.exp(ll - ll.max())
    w = (w / w.sum()).astype(np.float32)
    return pos, quat, w


@pytest.mark.parametrize("n", [1, 2, 3, 1000, 1023, 1024, 1025, 4097, 65_537])
@pytest.mark.parametrize("ratio", [0.0, 0.1])
def test_resample_bitexact(gpu, n, ratio):
    pos, quat, w = weighted_state(n, n)
    k, o = pair(n, 5)
    k.set_state(pos, quat, w)
    o.set_state(pos.astype(np.float64), quat.astype(np.float64), w.astype(np.float64))
    lcg = 0x1234_5678_9ABC_DEF1 ^ n
    k.set_lcg(lcg)
    o.set_rng(o.rng()[0], 0, lcg)
    k.resample(ratio)
    o.resample(ratio)
    anc = k.ancestors()
    oanc = o.ancestors()
    assert np.array_equal(anc, oanc), f"{np.sum(anc != oanc)} ancestor mismatches"
    assert np.all(np.diff(anc) >= 0)
    assert k.get_lcg() == o.rng()[2]
    _, _, wn, _, _, _ = assert_state_equal(k, o, "resample")
    assert np.all(wn == np.float32(1.0 / n))


@pytest.mark.parametrize("n", [4097, 65_537, 1_000_000])
def test_resample_scan_direct_route_bitexact(gpu, n):
    """The look-back scan with a zero wait budget: every tile that finds a predecessor
    unpublished sums the predecessors itself.  Same ancestors as the oracle, twice in a row
    (the second launch runs under the next epoch of the status words)."""
    pos, quat, w = weighted_state(n, n + 3)
    k, o = pair(n, 5)
    k.set_scan_spin_limit(0)
    for rep in range(2):
        k.set_state(pos, quat, w)
        o.set_state(pos.astype(np.float64), quat.astype(np.float64), w.astype(np.float64))
        lcg = 0x0F1E_2D3C_4B5A_6978 ^ (n + rep)
        k.set_lcg(lcg)
        o.set_rng(o.rng()[0], 0, lcg)
        k.resample(0.0)
        o.resample(0.0)
        assert np.array_equal(k.ancestors(), o.ancestors())
        assert k.get_lcg() == o.rng()[2]


@pytest.mark.parametrize("n", [5000, 300_007])
@pytest.mark.parametrize("ratio", [0.0, 0.01])
def test_resample_after_weight_scans_from_aggregates(gpu, n, ratio):
    """pfc_resample right after pfc_weight starts from the fixed-point tile aggregates the
    normalisation left (reduce, then scan); after pfx_set_state it takes the look-back scan.
    On the same weights, poses and LCG both routes give the same ancestors, LCG advance and
    resampled poses, and the aggregate route matches the oracle."""
    lm = make_landmarks(256, 9, box=(50.0, 50.0, 10.0))
    q_t = (math.cos(1.0), 0.0, 0.0, math.sin(1.0))
    truth = np.array([25.0, 25.0, 5.0])
    obs = rb_obs_for(truth, q_t, lm, 10, outliers=0.01)
    a, o = pair(n, 31)
    pose = f32([*truth, *q_t])
    a.reinit(pose, f32([0.5] * 3), f32([0.1] * 3))
    o.reinit(pose, f32([0.5] * 3), f32([0.1] * 3))
    a.upload_landmarks(lm)
    a.weight(obs, 0.05, 0.02, 3.0)
    pos, quat, w = a.get_state()
    lcg = a.get_lcg()
    rng = a.rng()
    a.resample(ratio)  # aggregate route
    b = PfKernels(256)
    b.init(n, 31)
    b.set_state(pos, quat, w)
    b.set_lcg(lcg)
    b.set_rng(*rng)  # the injection draws' generation
    b.resample(ratio)  # look-back route
    assert np.array_equal(a.ancestors(), b.ancestors())
    assert a.get_lcg() == b.get_lcg()
    pa, qa, wa = a.get_state()
    pb, qb, wb = b.get_state()
    assert np.array_equal(pa, pb) and np.array_equal(qa, qb) and np.array_equal(wa, wb)
    o.set_state(pos.astype(np.float64), quat.astype(np.float64), w.astype(np.float64))
    o.set_rng(o.rng()[0], rng[1], lcg)
    o.resample(ratio)
    assert np.array_equal(a.ancestors(), o.ancestors())
    b.close()


@pytest.mark.parametrize("n", [2_200_001, 16_900_003])
def test_aggregate_route_with_tile_prefix_pass(gpu, n):
    """Above 2048 tiles of 1024 particles the aggregate route first scans the tile sums in a
    prefix pass (runs of 16 tiles, 16384 tiles a chunk: 16.9M particles takes two chunks);
    its ancestors, LCG advance and resampled state equal the look-back route's, and at 2.2M
    the oracle's."""
    lm = make_landmarks(64, 4, box=(50.0, 50.0, 10.0))
    q_t = (math.cos(0.4), 0.0, 0.0, math.sin(0.4))
    truth = np.array([20.0, 30.0, 4.0])
    obs = rb_obs_for(truth, q_t, lm, 6, outliers=0.0)
    a = PfKernels(64)
    a.init(n, 77)
    a.reinit(f32([*truth, *q_t]), f32([0.8] * 3), f32([0.2] * 3))
    a.upload_landmarks(lm)
    a.weight(obs, 0.05, 0.02, 3.0)
    pos, quat, w = a.get_state()
    lcg = a.get_lcg()
    rng = a.rng()
    a.resample(0.01)  # aggregate route, prefix pass
    b = PfKernels(64)
    b.init(n, 77)
    b.set_state(pos, quat, w)
    b.set_lcg(lcg)
    b.set_rng(*rng)
    b.resample(0.01)  # look-back route
    anc = a.ancestors()
    assert np.array_equal(anc, b.ancestors())
    assert a.get_lcg() == b.get_lcg()
    pa, qa, wa = a.get_state()
    pb, qb, wb = b.get_state()
    assert np.array_equal(pa, pb) and np.array_equal(qa, qb) and np.array_equal(wa, wb)
    b.close()
    if n < 4_000_000:
        o = OracleFilter(n, 77, True)
        o.set_state(pos.astype(np.float64), quat.astype(np.float64), w.astype(np.float64))
        o.set_rng(o.rng()[0], rng[1], lcg)
        o.resample(0.01)
        assert np.array_equal(anc, o.ancestors())
    a.close()


def test_resample_after_empty_weight_scans_from_aggregates(gpu):
    """pfc_weight with no observations resets the weights to uniform through the normalisation
    (which leaves the aggregates); the resample that follows starts from them and gives the
    look-back route's ancestors and LCG advance."""
    n = 70_001
    a, _ = pair(n, 8)
    a.reinit(f32([1.0, 2.0, 3.0, 1.0, 0.0, 0.0, 0.0]), f32([0.5] * 3), f32([0.1] * 3))
    a.weight(np.zeros(0, OBS_DTYPE), 0.05, 0.02, 3.0)
    pos, quat, w = a.get_state()
    lcg = a.get_lcg()
    a.resample(0.0)
    b = PfKernels(256)
    b.init(n, 8)
    b.set_state(pos, quat, w)
    b.set_lcg(lcg)
    b.resample(0.0)
    assert np.array_equal(a.ancestors(), b.ancestors())
    assert a.get_lcg() == b.get_lcg()
    b.close()


def test_resample_bitexact_more_tiles_than_resident(gpu):
    """4.2M particles: more scan tiles than the device holds at once (later tiles start while
    earlier ones are still finishing); ancestors still bit-exact, over two launches."""
    n = 4_200_000
    pos, quat, w = weighted_state(n, 77)
    k, o = pair(n, 5)
    for rep in range(2):
        k.set_state(pos, quat, w)
        o.set_state(pos.astype(np.float64), quat.astype(np.float64), w.astype(np.float64))
        lcg = 0x5555_AAAA_1234_8765 + rep
        k.set_lcg(lcg)
        o.set_rng(o.rng()[0], 0, lcg)
        k.resample(0.0)
        o.resample(0.0)
        assert np.array_equal(k.ancestors(), o.ancestors())


def test_resample_copies_property(gpu):
    n = 100_000
    pos, quat, w = weighted_state(n, 21, conc=0.5)
    k, _ = pair(n, 5)
    k.set_state(pos, quat, w)
    k.resample(0.0)
    anc = k.ancestors()
    q = np.floor(w.astype(np.float64) * 2.0 ** 56)
    expect = n * q / q.sum()
    copies = np.bincount(anc, minlength=n)
    assert np.all(copies >= np.floor(expect) - 1e-9) and np.all(copies <= np.ceil(expect) + 1e-9)


def test_resample_single_heavy_particle(gpu):
    n = 5000
    pos, quat = random_state(n, 2)
    w = np.zeros(n, np.float32)
    w[1234] = 1.0
    k, _ = pair(n, 5)
    k.set_state(pos, quat, w)
    k.resample(0.0)
    assert np.all(k.ancestors() == 1234)
    p, _, _ = k.get_state()
    assert np.all(p == pos[1234])


def skewed_weights(n, kind, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros(n, np.float64)
    if kind == "first":
        w[0] = 1.0
    elif kind == "last":
        w[-1] = 1.0
    elif kind == "ends":
        w[0], w[-1] = 0.3, 0.7
    elif kind == "sparse":  # a few hundred heavy particles among zeros (the tracking regime)
        idx = rng.choice(n, size=300, replace=False)
        w[idx] = rng.gamma(0.5, 1.0, size=300)
    else:  # "tail": weight rising steeply towards the end (long runs of single copies)
        w = np.exp(np.linspace(-60.0, 0.0, n))
    return (w / w.sum()).astype(np.float32)


@pytest.mark.parametrize("n", [70_001, 1_048_576, 1_048_577])
@pytest.mark.parametrize("kind", ["first", "last", "ends", "sparse", "tail"])
def test_resample_skewed_weights_bitexact(gpu, n, kind):
    """Weights that put whole blocks of outputs on one particle or leave long runs of particles
    without copies: each gather block's merge-path window and its tile search at their edges
    (a filter of exactly 1024 tiles, one more, a ragged last tile); ancestors bit-exact."""
    pos, quat = random_state(n, 31)
    w = skewed_weights(n, kind, n)
    k, o = pair(n, 5)
    k.set_state(pos, quat, w)
    o.set_state(pos.astype(np.float64), quat.astype(np.float64), w.astype(np.float64))
    lcg = 0x0F1E_2D3C_4B5A_6978 ^ n
    k.set_lcg(lcg)
    o.set_rng(o.rng()[0], 0, lcg)
    k.resample(0.0)
    o.resample(0.0)
    anc = k.ancestors()
    assert np.array_equal(anc, o.ancestors()), f"{kind}: {np.sum(anc != o.ancestors())} ancestor mismatches"
    p, q, _ = k.get_state()
    assert np.array_equal(p, pos[anc]) and np.array_equal(q, quat[anc])


def test_inject_duplicates_deterministic(gpu):
    """ratio 1.0 -> n picks, many duplicate targets: pick-ordered f64 sums, bit-exact."""
    n = 3000
    pos, quat, w = weighted_state(n, 8)
    k, o = pair(n, 5)
    k.set_state(pos, quat, w)
    o.set_state(pos.astype(np.float64), quat.astype(np.float64), w.astype(np.float64))
    k.resample(1.0)
    o.resample(1.0)
    assert_state_equal(k, o, "inject ratio 1")


@pytest.mark.parametrize("n", [1, 1000, 100_000])
def test_est_sums_and_covs_parity(gpu, n):
    pos, quat, w = weighted_state(n, n + 3)
    k, o = pair(n, 5)
    k.set_state(pos, quat, w)
    o.set_state(pos.astype(np.float64), quat.astype(np.float64), w.astype(np.float64))
    s = k.est_sums()
    os_ = o.est_sums()
    assert np.allclose(s, os_, rtol=1e-9, atol=1e-15)
    M = np.zeros((4, 4))
    e = 3
    for r in range(4):
        for c in range(r, 4):
            M[r, c] = M[c, r] = os_[e]
            e += 1
    mq = oracle_quat_mean(M)
    mp = f32(os_[:3])
    cv = k.covs(mp, f32(mq))
    ocv = o.covs(mp, f32(mq))
    scale = np.array([max(abs(ocv[0]), abs(ocv[3]), abs(ocv[5]))] * 6 + [max(abs(ocv[6]), abs(ocv[9]), abs(ocv[11]))] * 6)
    assert np.all(np.abs(cv - ocv) <= 1e-9 * scale + 1e-300), f"{np.abs(cv - ocv) / scale}"


# ---------------------------------------------------------------- reference test ports


def test_cuda_covariances_match_gaussian_prior(gpu):
    """crates/pf_core/tests/cuda_agreement.rs:30-80 on the HIP backend."""
    from cuda_pf_localization_amd import ParticleFilter, PfConfig
    n = 200_000
    pf = ParticleFilter(PfConfig(particle_count=n, seed=7))
    pos_std = np.array([1.0, 2.0, 0.5])
    rot_std = np.array([0.05, 0.07, 0.02])
    pf.reinitialize([0, 0, 0, 1, 0, 0, 0], pos_std, rot_std)
    eg = pf.estimate()
    o = OracleFilter(n, 7, False)
    o.reinit(np.array([0, 0, 0, 1.0, 0, 0, 0]), pos_std, rot_std)
    ec = o.estimate()
    assert eg.valid
    assert np.all(np.abs(ec["position"] - eg.position) < 0.05)
    for i in range(3):
        want = pos_std[i] ** 2
        assert abs(eg.pos_cov[i, i] - want) <= 0.12 * want
        assert abs(ec["pos_cov"][i, i] - want) <= 0.12 * want
        wr = rot_std[i] ** 2
        assert abs(eg.att_cov[i, i] - wr) <= 0.15 * wr
        assert abs(ec["att_cov"][i, i] - wr) <= 0.15 * wr
        for j in range(3):
            if i != j:
                assert abs(eg.pos_cov[i, j]) < 0.1 * want
                assert abs(eg.att_cov[i, j]) < 0.1 * wr


def test_cuda_empty_obs_resets_weights_like_cpu(gpu):
    """cuda_agreement.rs:82-131."""
    from cuda_pf_localization_amd import Observation, OdomDelta, ParticleFilter, PfConfig
    n = 8192
    pf = ParticleFilter(PfConfig(particle_count=n, seed=11, resample_ess_ratio=0.0))
    pf.set_map({0: (5.0, 5.0, 1.0)})
    pf.reinitialize([5, 5, 1, 1, 0, 0, 0], [3, 3, 3], [0.1, 0.1, 0.1])
    e1 = pf.update(OdomDelta(), [Observation(0, 0, (0.0, 0.0, 0.0))])
    assert e1.ess < 0.5 * n
    e2 = pf.update(OdomDelta(), [])
    assert e2.ess > 0.95 * n
    e3 = pf.update(OdomDelta(), [Observation(42, 0, (1.0, 0.0, 0.0))])
    assert e3.ess > 0.95 * n


def test_cuda_random_injection_tracks_trajectory(gpu):
    """cuda_agreement.rs:133-174."""
    from cuda_pf_localization_amd import OdomDelta, ParticleFilter, PfConfig
    pf = ParticleFilter(PfConfig(particle_count=4096, seed=13, resample_ess_ratio=0.5, random_inject_ratio=0.1))
    pf.reinitialize([0, 0, 0, 1, 0, 0, 0], [0.5, 0.5, 0.5], [0.01, 0.01, 0.01])
    odo = OdomDelta(translation=(0.5, 0, 0), trans_noise=(0.02, 0.02, 0.02), rot_noise=(0.005, 0.005, 0.005))
    est = pf.update(odo, [])
    for kk in range(1, 20):
        est = pf.update(odo, [])
        assert abs(est.position[0] - 0.5 * (kk + 1)) < 0.6
        assert est.valid and math.isfinite(est.ess)


def test_cuda_tight_cloud_accuracy(gpu):
    """cuda_agreement.rs:177-219: ESS == N +- 0.05 at 100k; att_cov within 2 % of the f64 CPU."""
    from cuda_pf_localization_amd import ParticleFilter, PfConfig
    n = 100_000
    pf = ParticleFilter(PfConfig(particle_count=n, seed=7))
    rot_std = np.array([1e-3, 2e-3, 3e-3])
    pf.reinitialize([0, 0, 0, 1, 0, 0, 0], [0.01] * 3, rot_std)
    eg = pf.estimate()
    o = OracleFilter(n, 7, False)
    o.reinit(np.array([0, 0, 0, 1.0, 0, 0, 0]), np.array([0.01] * 3), rot_std)
    ec = o.estimate()
    assert abs(eg.ess - n) < 0.05
    for i in range(3):
        want = rot_std[i] ** 2
        assert abs(ec["att_cov"][i, i] - want) <= 0.05 * want
        assert abs(eg.att_cov[i, i] - ec["att_cov"][i, i]) / want < 0.02


def test_coasts_without_observations(gpu):
    """integration.rs:40-68 on the HIP backend (Auto -> GPU in the reference)."""
    from cuda_pf_localization_amd import OdomDelta, ParticleFilter, PfConfig
    pf = ParticleFilter(PfConfig(particle_count=1024, seed=5))
    pf.reinitialize([1, 2, 3, 1, 0, 0, 0], [0.1] * 3, [0.01] * 3)
    odo = OdomDelta(translation=(1.0, 0, 0), trans_noise=(0.01,) * 3, rot_noise=(0.005,) * 3)
    for _ in range(20):
        e = pf.update(odo, [])
    assert abs(e.position[0] - 21.0) < 0.5 and abs(e.position[1] - 2.0) < 0.2 and abs(e.position[2] - 3.0) < 0.2
    assert e.valid


def test_ess_drops_on_informative_observation(gpu):
    """integration.rs:70-101."""
    from cuda_pf_localization_amd import Observation, OdomDelta, ParticleFilter, PfConfig
    pf = ParticleFilter(PfConfig(particle_count=2048, seed=77, resample_ess_ratio=0.0))
    pf.set_map({0: (5.0, 5.0, 1.0)})
    pf.reinitialize([5, 5, 1, 1, 0, 0, 0], [3.0] * 3, [0.1] * 3)
    e = pf.update(OdomDelta(), [Observation(0, 0, (0.0, 0.0, 0.0))])
    assert e.ess < 1024.0
    assert abs(e.position[0] - 5.0) < 0.5


def test_quaternion_mean_handles_sign_flip(gpu):
    """integration.rs:103-120 (n = 2, near-zero spread)."""
    from cuda_pf_localization_amd import ParticleFilter, PfConfig
    pf = ParticleFilter(PfConfig(particle_count=2, seed=42))
    pf.reinitialize([0, 0, 0, 1, 0, 0, 0], [1e-9] * 3, [0.0] * 3)
    e = pf.estimate()
    assert abs(abs(e.attitude[0]) - 1.0) < 1e-6


def test_converges_on_3d_trajectory(gpu):
    """integration.rs:11-38 (4096 particles, circle_scenario(100, 777), mean error < 0.25 m)."""
    from cuda_pf_localization_amd import Observation, OdomDelta, ParticleFilter, PfConfig
    from cuda_pf_localization_amd.sim import circle_scenario
    sc = circle_scenario(100, 777)
    pf = ParticleFilter(PfConfig(particle_count=4096, seed=123, range_noise=0.05))
    pf.set_map(sc.landmarks)
    pf.reinitialize([*sc.prior_pos, *sc.prior_q], [1.0] * 3, [0.5] * 3)
    err = 0.0
    for st in sc.steps:
        assert st.obs
        e = pf.update(OdomDelta(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise),
                      [Observation(i, m, v, r) for (i, m, v, r) in st.obs])
        err += np.linalg.norm(e.position - np.array(st.truth_pos))
    assert err / len(sc.steps) < 0.25


# ---------------------------------------------------------------- full cycle vs oracle driver


def oracle_update(o, st, lm, cfg, n):
    o.predict(f32(st.odom_t), f32(st.odom_q), f32(st.trans_noise), f32(st.rot_noise))
    obs = np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], OBS_DTYPE)
    o.weight(obs, lm, float(np.float32(cfg.range_noise)), float(np.float32(cfg.bearing_noise)),
             float(np.float32(cfg.gate_sigma)))
    e = o.estimate()
    rs = False
    if cfg.resample_ess_ratio > 0 and e["ess"] < cfg.resample_ess_ratio * n:
        o.resample(np.float32(cfg.random_inject_ratio))
        e = o.estimate()
        rs = True
    return e, rs


@pytest.mark.parametrize("fast", [False, True])
def test_full_cycle_matches_oracle_driver(gpu, fast):
    """Config-0 tracking (10k particles, 12 landmarks, RB obs, gate 3): the C++ host layer over
    the HIP ABI vs the oracle driver (particle_filter.rs:110-124) for the config's 100 cycles."""
    from cuda_pf_localization_amd import Observation, OdomDelta, ParticleFilter, PfConfig
    from cuda_pf_localization_amd.sim import config0_scenario
    sc = config0_scenario(cycles=100, seed=42)
    n = 10_000
    cfg = PfConfig(particle_count=n, seed=42, range_noise=0.05, bearing_noise=0.02, gate_sigma=3.0,
                   resample_ess_ratio=0.5, random_inject_ratio=0.0, fast=fast)
    pf = ParticleFilter(cfg)
    pf.set_map(sc.landmarks)
    lm = np.array([sc.landmarks[i] for i in range(len(sc.landmarks))], np.float32)
    prior = [*sc.prior_pos, *sc.prior_q]
    pf.reinitialize(prior, [0.5] * 3, [0.1] * 3)
    o = OracleFilter(n, 42, True)
    o.reinit(f32(prior), f32([0.5] * 3), f32([0.1] * 3))
    n_rs = 0
    for st in sc.steps:
        eg = pf.update(OdomDelta(st.odom_t, st.odom_q, st.trans_noise, st.rot_noise),
                       [Observation(i, m, v, r) for (i, m, v, r) in st.obs])
        eo, rs = oracle_update(o, st, lm, cfg, n)
        assert eg.resampled == rs
        n_rs += rs
        assert np.all(np.abs(eg.position - eo["position"]) < 1e-5)
        assert abs(eg.ess - eo["ess"]) <= 1e-4 * eo["ess"]
        dq = min(np.abs(eg.attitude - eo["attitude"]).max(), np.abs(eg.attitude + eo["attitude"]).max())
        assert dq < 1e-5
        assert np.allclose(eg.pos_cov, eo["pos_cov"], rtol=1e-4, atol=1e-9)
        assert np.allclose(eg.att_cov, eo["att_cov"], rtol=1e-4, atol=1e-9)
        assert np.linalg.norm(eg.position - np.array(st.truth_pos)) < 0.5
    assert n_rs > 0

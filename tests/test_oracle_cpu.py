"""Pins the CPU oracle (oracle/pf_oracle.c) before it is trusted as the checker.

The reference ships no golden vectors (its tests are statistical plus a few known answers),
so each oracle function is pinned by what the mathematics or the reference's own tests fix:
  * deterministic math      vs libm (motion -2 ln, sqrt, sin/cos; CBGL sin/cos) and the
                            literal axis-angle formula
  * normals                 moments of the Box-Muller stream
  * Markley mean            vs numpy.linalg.eigh (independent eigen-solver)
  * integer systematic      == literal f64 cpu.rs:264-289 where that is exact (dyadic weights,
                            power-of-two n), == brute force from the definition (Python ints)
  * injection picks         vs an independent Python restatement of pf_kernels.cu:337-351
  * filter behaviour        ports of integration.rs / cuda_agreement.rs on the oracle
  * raycast / CAER          known answers of global_loc.rs tests, brute-force marching,
                            f32 DDA vs f64 DDA, lane-tree CAER vs plain sum
"""
import math

import numpy as np
import pytest

import oracle as O
from cuda_pf_localization_amd import sim

TAU = 2 * math.pi


# ------------------------------------------------------------------ deterministic math

def test_det_sincos_vs_libm():
    """The full-precision sin/cos of the CBGL ray directions and hypotheses."""
    rng = np.random.default_rng(0)
    for a in np.concatenate([rng.random(20000) * 40 - 20, [0.0, math.pi / 4, math.pi / 2, math.pi]]):
        s, c = O.det_sincos(a)
        assert abs(s - math.sin(a)) < 3e-16 * max(1, abs(a)) and abs(c - math.cos(a)) < 3e-16 * max(1, abs(a))


def test_det_cos_sinc_is_axis_angle():
    """(cos h, a sinc(h)/2) equals from_axis_angle(a/|a|, |a|) of cpu.rs:30-42."""
    rng = np.random.default_rng(1)
    for a in rng.normal(size=(5000, 3)) * np.array([[0.01], [0.5], [3.0], [20.0], [1e-7]]).repeat(1000, 0):
        ang = float(np.linalg.norm(a))
        c, s = O.det_cos_sinc(0.25 * ang * ang)
        lit_w = math.cos(0.5 * ang)
        lit_v = a / ang * math.sin(0.5 * ang)
        assert abs(c - lit_w) < 1e-13 * max(1.0, ang)
        assert np.all(np.abs(a * 0.5 * s - lit_v) < 1e-13 * max(1.0, ang))
    # either side of the short-series threshold h^2 = 2^-10 (|a| = 1/16)
    for ang in (0.0625 * (1 - 1e-12), 0.0625, 0.0625 * (1 + 1e-12), 0.06, 0.07):
        c, s = O.det_cos_sinc(0.25 * ang * ang)
        assert abs(c - math.cos(0.5 * ang)) < 1e-16 and abs(0.5 * ang * s - math.sin(0.5 * ang)) < 1e-16, ang


def _sm_final(z):
    m = (1 << 64) - 1
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def test_normals_are_box_muller_of_the_stated_words():
    """The motion normals are the Box-Muller transform (libm log, sqrt, cos, sin here) of the
    six 32-bit words the draw specification names: the particle key (cpu.rs:153-159) and the next
    two splitmix64 outputs (pf_kernels.cu:323-328), radius from each value's high half, angle
    from its low half; the short f32 fmaf sequences agree with libm to 1e-6."""
    m = (1 << 64) - 1
    phi = 0x9E3779B97F4A7C15
    for gseed, i in [(12345, 0), (12345, 7), (0x9E37797979797979, 999_999), (42, 123_456), (1, 2 ** 31 - 1)]:
        x = _sm_final((gseed * phi + i * 0xBF58476D1CE4E5B9) & m)
        z = [x, _sm_final((x + phi) & m), _sm_final((x + 2 * phi) & m)]
        want = []
        for v in z:
            r = math.sqrt(-2.0 * math.log(((v >> 32) + 1) / 2 ** 32))
            th = 2 * math.pi * (v & 0xFFFFFFFF) / 2 ** 32
            want += [r * math.cos(th), r * math.sin(th)]
        got = O.particle_normals(gseed, i)
        assert np.allclose(got, want, rtol=1e-6, atol=1e-6), (gseed, i, got, want)


def test_motion_math_vs_libm():
    """The motion draws' f32 -2 ln, sqrt and sin/cos against libm over 100k random words and the
    end points (words 0, 2^32 - 1, the f32 rounding edges near 2^24 and 2^32, octant
    boundaries).  -2 ln is held to 2e-7 relative plus 2e-7 absolute: the word is rounded to f32
    (24 bits) before the logarithm, which moves ln u by up to 6e-8 near u = 1."""
    import ctypes as C
    L = O.lib()
    rng = np.random.default_rng(4)
    words = [int(a) for a in rng.integers(0, 2 ** 32, 100_000)] + [0, 1, 2 ** 31, 2 ** 32 - 1, 2 ** 24,
                                                                    2 ** 24 + 1, 2 ** 32 - 256, 2 ** 32 - 129]
    for a in words:
        ref = -2.0 * math.log((a + 1) / 2 ** 32)
        assert abs(L.or_mot_neg2ln(a) - ref) <= 2e-7 * ref + 2e-7, a
    assert L.or_mot_neg2ln(2 ** 32 - 1) == 0.0
    assert L.or_mot_sqrt(0.0) == 0.0
    for v in list(rng.random(100_000) * 45.0) + [1e-12, 44.4]:
        v = float(np.float32(v))
        assert abs(L.or_mot_sqrt(v) - math.sqrt(v)) <= 2e-7 * math.sqrt(v), v
    s, c = C.c_float(), C.c_float()
    for b in words + [k * 2 ** 29 + d for k in range(8) for d in (0, 1, 2 ** 29 - 1)]:
        b &= 0xFFFFFFFF
        L.or_mot_sincos_turn(b, C.byref(s), C.byref(c))
        th = 2 * math.pi * b / 2 ** 32
        assert abs(s.value - math.sin(th)) < 1.5e-7 and abs(c.value - math.cos(th)) < 1.5e-7, b


def test_normals_moments():
    z = np.array([O.particle_normals(12345, i) for i in range(40000)]).ravel()
    assert abs(z.mean()) < 0.01
    assert abs(z.var() - 1.0) < 0.01
    assert abs(np.mean(z ** 4) - 3.0) < 0.06  # Gaussian kurtosis, not sim.rs's triangle
    frac = np.mean(np.abs(z) < 1.0)
    assert abs(frac - 0.6826894921) < 0.005


# ------------------------------------------------------------------ Markley mean

def _sym4(rng):
    q = rng.normal(size=(50, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w = rng.random(50)
    return np.einsum("i,ij,ik->jk", w / w.sum(), q, q)


def test_quat_mean_vs_eigh():
    rng = np.random.default_rng(2)
    for _ in range(200):
        M = _sym4(rng)
        q = O.quat_mean(M)
        val, vec = np.linalg.eigh(M)
        v = vec[:, -1] * (1 if vec[0, -1] >= 0 else -1)
        assert np.allclose(q, v, atol=1e-10) or np.allclose(q, -v, atol=1e-10) and abs(q[0]) < 1e-12
        assert abs(np.linalg.norm(q) - 1) < 1e-12 and q[0] >= 0


# ------------------------------------------------------------------ resampling

def test_systematic_int_equals_literal_when_exact():
    """Dyadic weights + power-of-two n: cpu.rs's f64 pointers are exact, so the literal
    algorithm and the fixed-point form must agree index for index."""
    rng = np.random.default_rng(3)
    for n in (1, 2, 1024, 4096, 65536):
        for _ in range(5):
            raw = rng.integers(0, 2 ** 20, size=n).astype(np.float64)
            if raw.sum() == 0:
                raw[0] = 1
            # weights = k_i / 2^40 with sum <= 1
            scale = 2.0 ** -40
            raw = np.floor(raw / raw.sum() * 2 ** 40)
            w = raw * scale
            u32 = int(rng.integers(0, 2 ** 32))
            a_int = O.systematic_int(w, u32)
            a_lit = O.systematic_f64_literal(w, u32 / 2 ** 32)
            assert np.array_equal(a_int, a_lit)


def test_systematic_int_brute_force():
    rng = np.random.default_rng(4)
    for n in (1, 2, 3, 7, 31, 100):
        w = rng.random(n).astype(np.float32).astype(np.float64)
        w[rng.random(n) < 0.3] = 0.0
        if w.sum() == 0:
            w[0] = 0.5
        u32 = int(rng.integers(0, 2 ** 32))
        q = [int(math.floor(x * 2 ** 56)) if x > 0 else 0 for x in w]
        T = sum(q)
        cum = np.cumsum(q).tolist()
        want = []
        for i in range(n):
            rhs = T * (u32 + (i << 32))
            j = next((j for j in range(n) if cum[j] * (n << 32) >= rhs), n - 1)
            want.append(j)
        assert list(O.systematic_int(w, u32)) == want


def test_systematic_int_close_to_literal_random():
    rng = np.random.default_rng(5)
    n = 100_000
    w = rng.gamma(0.5, size=n)
    w /= w.sum()
    w = w.astype(np.float32).astype(np.float64)
    u32 = 0x9E3779B9
    a = O.systematic_int(w, u32)
    b = O.systematic_f64_literal(w, u32 / 2 ** 32)
    diff = np.nonzero(a != b)[0]
    assert len(diff) <= 5 and np.all(np.abs(a[diff] - b[diff]) <= 1)


def _sm64(x):
    x = (x + 0x9E3779B97F4A7C15) & (2 ** 64 - 1)
    z = x
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
    return x, z ^ (z >> 31)


def test_inject_picks_match_python_restatement():
    """pf_kernels.cu:333-351 restated in Python integers; uniform weights keep ancestors = i."""
    n = 5000
    rng = np.random.default_rng(6)
    pos = rng.random((n, 3)) * 10
    pos = pos.astype(np.float32).astype(np.float64)
    q = np.tile([1.0, 0, 0, 0], (n, 1))
    f = O.OracleFilter(n, 42, True)
    f.set_state(pos, q)
    lcg0 = 0x243F6A8885A308D3
    f.resample(np.float32(0.1))
    assert np.array_equal(f.ancestors(), np.arange(n))
    lcg = (lcg0 * 6364136223846793005 + 1442695040888963407) & (2 ** 64 - 1)
    k = int(np.float32(0.1) * np.float32(n))
    exp = pos.copy()
    for j in range(k):
        z = lcg ^ ((0 * 0x9E3779B97F4A7C15) & (2 ** 64 - 1)) ^ (((j + 1) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1))
        if z == 0:
            z = 0x243F6A8885A308D3
        us = []
        for _ in range(4):
            z, out = _sm64(z)
            us.append((out >> 11) * 2.0 ** -53)
        idx = min(int(us[0] * n), n - 1)
        exp[idx] += np.array(us[1:]) - 0.5
    got = f.state()[0]
    assert np.array_equal(got, exp.astype(np.float32).astype(np.float64))


# ------------------------------------------------------------------ filter behaviour (reference tests)

def test_reinit_moments_match_gaussian_prior():
    """cuda_agreement.rs:30-80 on the oracle (n = 200k, seed 7)."""
    n = 200_000
    f = O.OracleFilter(n, 7, False)
    ps = np.array([1.0, 2.0, 0.5])
    rs = np.array([0.05, 0.07, 0.02])
    f.reinit(np.array([0, 0, 0, 1.0, 0, 0, 0]), ps, rs)
    e = f.estimate()
    for i in range(3):
        assert abs(e["pos_cov"][i, i] - ps[i] ** 2) <= 0.12 * ps[i] ** 2
        assert abs(e["att_cov"][i, i] - rs[i] ** 2) <= 0.15 * rs[i] ** 2
        for j in range(3):
            if i != j:
                assert abs(e["pos_cov"][i, j]) < 0.1 * ps[i] ** 2
    assert np.all(np.abs(e["position"]) < 0.05)


def test_tight_cloud_ess_and_att_cov():
    """cuda_agreement.rs:177-219 on the oracle: ESS exactly N, att_cov within 5 %."""
    n = 100_000
    f = O.OracleFilter(n, 7, False)
    rs = np.array([1e-3, 2e-3, 3e-3])
    f.reinit(np.array([0, 0, 0, 1.0, 0, 0, 0]), np.array([0.01] * 3), rs)
    e = f.estimate()
    assert abs(e["ess"] - n) < 1e-6
    for i in range(3):
        assert abs(e["att_cov"][i, i] - rs[i] ** 2) <= 0.05 * rs[i] ** 2


def _cycle(f, n, t, q, tn, rn, obs, lm, sr, sb, gate, ratio=0.5, inject=0.0):
    f.predict(t, q, tn, rn)
    f.weight(obs, lm, sr, sb, gate)
    e = f.estimate()
    if ratio > 0 and e["ess"] < ratio * n:
        f.resample(np.float32(inject))
        e = f.estimate()
    return e


def test_coasts_without_observations():
    """integration.rs:40-68."""
    n = 1024
    f = O.OracleFilter(n, 5, False)
    f.reinit(np.array([1, 2, 3, 1.0, 0, 0, 0]), np.array([0.1] * 3), np.array([0.01] * 3))
    for _ in range(20):
        e = _cycle(f, n, [1.0, 0, 0], [1.0, 0, 0, 0], [0.01] * 3, [0.005] * 3, None, np.zeros((1, 3)), 0.05, 0.02, 0.0)
    assert abs(e["position"][0] - 21.0) < 0.5 and abs(e["position"][1] - 2.0) < 0.2 and abs(e["position"][2] - 3.0) < 0.2


def test_ess_drops_on_informative_observation():
    """integration.rs:70-101 (observation exactly at the landmark)."""
    n = 2048
    f = O.OracleFilter(n, 77, False)
    f.reinit(np.array([5, 5, 1, 1.0, 0, 0, 0]), np.array([3.0] * 3), np.array([0.1] * 3))
    obs = np.array([(0, 0, 0.0, 0.0, 0.0, 0.0)], O.OBS_DTYPE)
    e = _cycle(f, n, [0, 0, 0], [1.0, 0, 0, 0], [0.05, 0.05, 0.05], [0.02, 0.02, 0.03], obs,
               np.array([[5.0, 5.0, 1.0]], np.float32), 0.05, 0.02, 0.0, ratio=0.0)
    assert e["ess"] < 1024
    assert abs(e["position"][0] - 5.0) < 0.5


def test_weight_degenerate_inputs_uniform():
    n = 1000
    f = O.OracleFilter(n, 1, True)
    f.reinit(np.array([5, 5, 1, 1.0, 0, 0, 0]), np.array([3.0] * 3), np.array([0.1] * 3))
    f.weight(None, np.zeros((1, 3)), 0.05, 0.02, 0.0)
    assert np.all(f.state()[2] == np.float32(1.0 / n))
    f.weight(np.array([(9, 0, 1.0, 0, 0, 0)], O.OBS_DTYPE), np.zeros((1, 3)), 0.05, 0.02, 0.0)
    assert np.all(f.state()[2] == np.float32(1.0 / n))


def test_quaternion_mean_handles_sign_flip():
    """integration.rs:103-120."""
    f = O.OracleFilter(2, 42, False)
    f.reinit(np.array([0, 0, 0, 1.0, 0, 0, 0]), np.array([1e-9] * 3), np.zeros(3))
    e = f.estimate()
    assert abs(abs(e["attitude"][0]) - 1.0) < 1e-6
    M = np.diag([0.0, 0.0, 0.0, 1.0])
    q = O.quat_mean(-np.eye(4) + 2 * M)
    assert abs(abs(q[3]) - 1) < 1e-12


def test_converges_on_3d_trajectory():
    """integration.rs:11-38: 4096 particles, circle_scenario(100, 777), mean error < 0.25 m."""
    sc = sim.circle_scenario(100, 777)
    n = 4096
    f = O.OracleFilter(n, 123, False)
    f.reinit(np.array([*sc.prior_pos, *sc.prior_q]), np.array([1.0] * 3), np.array([0.5] * 3))
    lm = np.array([sc.landmarks[i] for i in range(12)], np.float32)
    err = 0.0
    for st in sc.steps:
        assert st.obs
        obs = np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], O.OBS_DTYPE)
        e = _cycle(f, n, st.odom_t, st.odom_q, st.trans_noise, st.rot_noise, obs, lm, 0.05, 0.02, 0.0)
        err += np.linalg.norm(e["position"] - np.array(st.truth_pos))
    assert err / len(sc.steps) < 0.25


def test_config0_tracking_range_bearing():
    """Config 0 (10k particles, 12 landmarks, RB observations gated at 3 sigma) tracks."""
    sc = sim.config0_scenario(cycles=40, seed=42)
    n = 10_000
    f = O.OracleFilter(n, 42, True)
    f.reinit(np.array([*sc.prior_pos, *sc.prior_q]), np.array([0.5] * 3), np.array([0.1] * 3))
    lm = np.array([sc.landmarks[i] for i in range(12)], np.float32)
    errs = []
    for st in sc.steps:
        obs = np.array([(i, m, *v, r) for (i, m, v, r) in st.obs], O.OBS_DTYPE)
        e = _cycle(f, n, st.odom_t, st.odom_q, st.trans_noise, st.rot_noise, obs, lm, 0.05, 0.02, 3.0)
        errs.append(np.linalg.norm(e["position"] - np.array(st.truth_pos)))
    assert np.mean(errs[10:]) < 0.05


# ------------------------------------------------------------------ CBGL

def test_raycast_known_answers():
    """global_loc.rs tests:11-41 on both oracle DDAs."""
    cells, res, org = sim.room_grid()
    g = O.OracleGrid(cells, res, org)
    assert abs(g.cast_ray64(5.0, 5.0, 1.0, 0.0, 20.0) - 4.95) < 0.15
    assert abs(g.cast_ray64(5.0, 5.0, 0.0, 1.0, 20.0) - 4.95) < 0.15
    assert abs(g.cast_ray64(2.5, 5.0, 1.0, 0.0, 20.0) - 1.25) < 0.12
    assert g.cast_ray64(5.0, 5.0, 1.0, 0.0, 2.0) == 2.0
    assert abs(g.cast_ray32(5.0, 5.0, 1.0, 0.0, 20.0) - 4.95) < 0.15
    assert abs(g.cast_ray32(2.5, 5.0, 1.0, 0.0, 20.0) - 1.25) < 0.12
    assert g.cast_ray32(5.0, 5.0, 1.0, 0.0, 2.0) == np.float32(2.0)
    # python restatement of the f64 DDA agrees exactly
    rng = np.random.default_rng(0)
    for _ in range(300):
        x, y, a = rng.random() * 9 + 0.5, rng.random() * 9 + 0.5, rng.random() * TAU
        assert sim._cast_ray64(cells, 200, 200, res, org, x, y, math.cos(a), math.sin(a), 8.0) == \
            g.cast_ray64(x, y, math.cos(a), math.sin(a), 8.0)


def test_dda_step_count_known_answers():
    """The algorithmic work count of a scoring pass: a ray along +x from cell (5, 5) of a 40 x 40
    bordered grid steps through cells 6..39 (34 steps, the last one entering the wall); along +y
    from cell (5, 30) it takes 9; a diagonal one from cell (5, 5) alternates x and y steps
    (33 of each to cell (38, 38)) and enters the wall with its next x step: 67."""
    cells = np.zeros((40, 40), np.uint8)
    cells[0, :] = cells[-1, :] = cells[:, 0] = cells[:, -1] = 1
    g = O.OracleGrid(cells, 0.05, (0.0, 0.0))
    c = lambda i: (i + 0.5) * 0.05
    assert g.dda_steps64(np.array([[c(5), c(5), 0.0]]), 1, 0.0, 10.0) == 34
    assert g.dda_steps64(np.array([[c(5), c(30), math.pi / 2]]), 1, 0.0, 10.0) == 9
    assert g.dda_steps64(np.array([[c(5), c(5), math.pi / 4]]), 1, 0.0, 10.0) == 67
    both = np.array([[c(5), c(5), 0.0], [c(5), c(30), math.pi / 2]])
    assert g.dda_steps64(both, 1, 0.0, 10.0) == 43


def test_raycast_brute_force_marching():
    cells, res, org = sim.obstacle_map(11)
    g = O.OracleGrid(cells, res, org)
    rng = np.random.default_rng(1)
    n_ok = 0
    for _ in range(300):
        x, y, a = rng.random() * 9 + 0.5, rng.random() * 9 + 0.5, rng.random() * TAU
        if cells[int(y / res), int(x / res)]:
            continue
        dx, dy = math.cos(a), math.sin(a)
        d = g.cast_ray64(x, y, dx, dy, 8.0)
        t = 0.0
        while t < 8.0:  # fine march: first occupied cell along the ray
            cx, cy = int((x + t * dx) / res), int((y + t * dy) / res)
            if not (0 <= cx < 200 and 0 <= cy < 200) or cells[cy, cx]:
                break
            t += 1e-4
        n_ok += abs(min(t, 8.0) - d) < 2e-4
    assert n_ok >= 250


def test_f32_dda_agrees_with_f64():
    cells, res, org = sim.obstacle_map(12)
    g = O.OracleGrid(cells, res, org)
    rng = np.random.default_rng(2)
    d = []
    for _ in range(3000):
        x, y, a = rng.random() * 10, rng.random() * 10, rng.random() * TAU
        r64 = g.cast_ray64(x, y, math.cos(a), math.sin(a), 8.0)
        r32 = g.cast_ray32(np.float32(x), np.float32(y), np.float32(math.cos(a)), np.float32(math.sin(a)), 8.0)
        d.append(abs(r64 - r32))
    d = np.array(d)
    assert np.mean(d < 1e-4) > 0.99


def test_caer32_tree_order_close_to_plain_sum():
    cells, res, org = sim.obstacle_map(13)
    g = O.OracleGrid(cells, res, org)
    hyp = g.gen_hyp32(5, 200)
    scan = g.map_scan64(4.0, 4.0, 0.3, 360, TAU, 8.0).astype(np.float32)
    c = g.caer32(scan, hyp, np.float32(TAU), 8.0)
    ranges = g.render32(hyp, 360, np.float32(TAU), 8.0)
    plain = np.abs(scan[None, :].astype(np.float64) - ranges).sum(axis=1)
    assert np.all(np.abs(c - plain) < 1e-3)


def test_gen_hyp32_python_restatement():
    cells, res, org = sim.obstacle_map(14)
    g = O.OracleGrid(cells, res, org)
    fl = g.free_list()
    hyp = g.gen_hyp32(77, 50)
    for h in range(50):
        x = (77 * 0x9E3779B97F4A7C15 + h * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & (2 ** 64 - 1)
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & (2 ** 64 - 1)
        x = z ^ (z >> 31)
        zs = []
        for _ in range(4):
            x, o = _sm64(x)
            zs.append(o)
        c = min(int(((zs[0] >> 11) * 2.0 ** -53) * len(fl)), len(fl) - 1)
        f32 = np.float32
        ux, uy, ut = (f32(z >> 40) * f32(2.0 ** -24) for z in zs[1:])
        ex = f32(org[0]) + (f32(fl[c, 0]) + ux) * f32(res)
        ey = f32(org[1]) + (f32(fl[c, 1]) + uy) * f32(res)
        et = f32(-3.14159265) + f32(6.28318530) * ut
        assert hyp[h, 0] == ex and hyp[h, 1] == ey and hyp[h, 2] == et


def test_cpu_gl_recovers_pose_in_synthetic_room():
    """global_loc.rs tests:43-67: 20k hypotheses, seed 7, error < 0.35 m."""
    cells, res, org = sim.room_grid()
    g = O.OracleGrid(cells, res, org)
    truth = (6.0, 7.0, 0.8)
    scan = g.map_scan64(*truth, 360, TAU, 10.0)
    top = g.localize64(scan, 20_000, 10, 7)
    assert math.hypot(top[0, 0] - truth[0], top[0, 1] - truth[1]) < 0.35
    assert np.all(np.diff(top[:, 3]) >= 0)


def test_caer_zero_at_truth_noiseless():
    cells, res, org = sim.obstacle_map(15)
    g = O.OracleGrid(cells, res, org)
    scan = g.map_scan64(3.0, 6.0, 0.4, 360, TAU, 8.0)
    assert g.caer64(scan, np.array([[3.0, 6.0, 0.4]]), TAU, 8.0)[0] == 0.0


# ------------------------------------------------------------------ fixtures (sim.rs)

def test_sim_fixtures_invariants():
    lm = sim.landmark_box(12, 99)
    a = np.array(list(lm.values()))
    assert a.shape == (12, 3)
    assert np.all(a[:, :2] >= 1.0) and np.all(a[:, :2] < 9.0)
    assert np.all(a[:, 2] >= 0.3) and np.all(a[:, 2] < 2.7)
    sc = sim.circle_scenario(100, 777)
    assert all(st.obs for st in sc.steps)
    cells, res, org = sim.room_grid()
    assert cells.shape == (200, 200) and cells[0].all() and cells[:, -1].all()
    assert cells[100, 80] == 1 and cells[100, 90] == 0
    g = sim.Gauss(7)
    v = [g.normal() for _ in range(20000)]
    assert abs(np.mean(v)) < 0.02 and abs(np.std(v) - 0.707) < 0.02  # sim.rs:38-41 triangle, std ~0.707


def test_helix_arrays_observation_convention():
    """Noise-free helix_arrays observations are the unit landmark->sensor vectors in the
    body frame and the exact ranges, like helix_scenario's."""
    h = sim.helix_arrays(9, (20.0, 20.0, 5.0), 5, 3, phase=0.4, sigma_r=0.0, sigma_b=0.0, odo_pos=0.0)
    s = sim.helix_scenario(9, (20.0, 20.0, 5.0), 5, 3, phase=0.4, sigma_r=0.0, sigma_b=0.0, odo_pos=0.0)
    for k in range(5):
        pos = np.array(s.steps[k].truth_pos)
        assert np.allclose(h.truth[k], pos)
        att = s.steps[k].truth_q
        for i in range(9):
            body = np.array(sim.qrot(sim.qconj(att), tuple(h.landmarks[i] - pos)))
            assert np.allclose(h.obs_v[k, i], -body / np.linalg.norm(body), atol=1e-12)
            assert abs(h.obs_r[k, i] - np.linalg.norm(body)) < 1e-12
        assert np.allclose(h.odom_t[k], s.steps[k].odom_t, atol=1e-12)
        assert np.allclose(h.odom_q[k], s.steps[k].odom_q, atol=1e-12)

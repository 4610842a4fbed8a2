"""Closed-form pins of the oracle's measurement model, motion composition and estimate.

The GPU likelihood, motion and estimate kernels are checked against oracle/pf_oracle.c; these
tests check the oracle itself against values the mathematics fixes for single particles, so a
mistake shared by both sides (a gate constant, sigma^2 for 2 sigma^2, the bearing sign, the
order of a quaternion product, the direction of the world-to-body map) cannot pass:

  * cpu.rs:69-96 MeasModel  exact hit -> 0; range off by d -> -d^2/(2 sr^2); bearing off by
                            theta -> -theta^2/(2 sb^2); each gate just below / above its
                            threshold -> the Gaussian term / -0.5 g^2 (range-bearing) and
                            -1.5 g^2 (Cartesian); observation vector pointing along the
                            predicted direction -> bearing pi; sigma floors at 1e-6; unknown
                            ids dropped; the softmax of two particles.
  * cpu.rs:180-218          reinit q0*dq and predict q*q_odo*dq with p += R(q)(t + noise),
                            checked as rotation-matrix products (scipy.spatial.transform,
                            an independent quaternion implementation) on non-commuting
                            rotations.
  * cpu.rs:308-356          two-particle means and covariances (pos: d d^T; attitude: the
                            log of mean^-1 * q is +-theta a, so the covariance is theta^2 a a^T).

Each check is a plain function taking the oracle library; test_oracle_mutations_cpu.py plants
mistakes in copies of pf_oracle.c and asserts that these same functions catch every one.
"""
import math

import numpy as np
from scipy.spatial.transform import Rotation

import oracle as O

P0 = np.array([1.0, 2.0, 3.0])
IDQ = np.array([1.0, 0.0, 0.0, 0.0])
SR, SB, G = 0.05, 0.02, 3.0


def _obs(rows):
    return np.array(rows, O.OBS_DTYPE)


def ll_single(L, p, q, lm, obs, sr=SR, sb=SB, gate=G):
    """Log-likelihood of one particle at (p, q) for the given landmarks and observations."""
    f = O.OracleFilter(1, 1, False, library=L)
    f.set_state(np.asarray(p, np.float64).reshape(1, 3), np.asarray(q, np.float64).reshape(1, 4))
    f.weight(_obs(obs), np.asarray(lm, np.float32).reshape(-1, 3), sr, sb, gate)
    return float(f.state()[3][0])


def _close(a, b, rel=1e-9, what=""):
    assert abs(a - b) <= rel * max(1.0, abs(b)), f"{what}: {a!r} != {b!r}"


def _quat_wxyz(r: Rotation):
    x, y, z, w = r.as_quat()
    return np.array([w, x, y, z])


def _rot(q_wxyz):
    w, x, y, z = q_wxyz
    return Rotation.from_quat([x, y, z, w])


# ------------------------------------------------------------------ measurement model


def check_range_bearing_closed_forms(L):
    lm = [[5.0, 2.0, 3.0]]  # body vector (4, 0, 0) at P0 with identity attitude
    # exact hit: range 4, vector from the landmark back to the sensor
    assert ll_single(L, P0, IDQ, lm, [(0, 1, -1.0, 0.0, 0.0, 4.0)]) == 0.0
    # range off by d (d exact in f32), both signs: -d^2 / (2 sr^2)
    for d, want in ((0.03125, -0.1953125), (-0.03125, -0.1953125), (0.140625, -3.955078125)):
        _close(ll_single(L, P0, IDQ, lm, [(0, 1, -1.0, 0.0, 0.0, 4.0 - d)]), want, what=f"range off {d}")
    # range gate g sr = 0.15: 5/32 = 0.15625 lies above it -> -0.5 g^2 = -4.5 (Gaussian would be -4.8828125)
    _close(ll_single(L, P0, IDQ, lm, [(0, 1, -1.0, 0.0, 0.0, 4.0 - 0.15625)]), -4.5, what="range gated")
    _close(ll_single(L, P0, IDQ, lm, [(0, 1, -1.0, 0.0, 0.0, 4.0 + 0.15625)]), -4.5, what="range gated +")
    # no gate: the Gaussian term at any distance
    _close(ll_single(L, P0, IDQ, lm, [(0, 1, -1.0, 0.0, 0.0, 4.0 - 0.15625)], gate=0.0), -4.8828125,
           what="range ungated")
    # bearing off by theta: landmark rotated about z in the body frame, vector stays (-1, 0, 0)
    for theta, gated in ((0.03, False), (0.055, False), (0.065, True), (0.5, True)):
        lmt = np.array([P0 + 4.0 * np.array([math.cos(theta), math.sin(theta), 0.0])], np.float32)
        b = lmt[0].astype(np.float64) - P0
        rn = math.sqrt(b @ b)
        th = math.atan2(b[1], b[0])  # the angle from the atan2 of the (f32) landmark, not acos
        want_b = -4.5 if gated else -th * th / (2 * SB * SB)
        _close(ll_single(L, P0, IDQ, lmt, [(0, 1, -1.0, 0.0, 0.0, rn)]), want_b, rel=1e-7, what=f"bearing {theta}")
    # a vector along the predicted direction: bearing error pi
    _close(ll_single(L, P0, IDQ, lm, [(0, 1, 1.0, 0.0, 0.0, 4.0)]), -4.5, what="bearing pi gated")
    _close(ll_single(L, P0, IDQ, lm, [(0, 1, 1.0, 0.0, 0.0, 4.0)], gate=0.0), -math.pi ** 2 / (2 * SB * SB),
           what="bearing pi ungated")
    # both terms add: range off 0.03125 and bearing gated
    _close(ll_single(L, P0, IDQ, lm, [(0, 1, 1.0, 0.0, 0.0, 4.0 - 0.03125)]), -4.5 - 0.1953125, what="sum")


def check_cartesian_closed_forms(L):
    lm = [[5.0, 2.0, 3.0]]
    # exact hit
    assert ll_single(L, P0, IDQ, lm, [(0, 0, 4.0, 0.0, 0.0, 0.0)]) == 0.0
    # residual e = pred - v, |e|^2 below 3 g^2 sr^2 = 0.0675 -> -|e|^2 / (2 sr^2)
    _close(ll_single(L, P0, IDQ, lm, [(0, 0, 4.0 - 0.125, 0.125, -0.0625, 0.0)]), -7.03125, what="cart")
    _close(ll_single(L, P0, IDQ, lm, [(0, 0, 4.0 - 0.25, -0.0625, 0.0, 0.0)]), -13.28125, what="cart below gate")
    # |e|^2 = 0.078125 above the gate -> -0.5 g^2 * 3 = -13.5 (Gaussian would be -15.625)
    _close(ll_single(L, P0, IDQ, lm, [(0, 0, 4.0 - 0.25, 0.125, 0.0, 0.0)]), -13.5, what="cart gated")
    _close(ll_single(L, P0, IDQ, lm, [(0, 0, 4.0 - 0.25, 0.125, 0.0, 0.0)], gate=0.0), -15.625, what="cart ungated")


def check_attitude_and_bookkeeping(L):
    # yaw +90 deg: the world offset (0, 4, 0) is (4, 0, 0) in the body frame (world-to-body = R^T)
    q = np.array([math.cos(math.pi / 4), 0.0, 0.0, math.sin(math.pi / 4)])
    lm = [[1.0, 6.0, 3.0]]
    _close(ll_single(L, P0, q, lm, [(0, 1, -1.0, 0.0, 0.0, 4.0)]), 0.0, what="yaw 90 RB")
    _close(ll_single(L, P0, q, lm, [(0, 0, 4.0, 0.0, 0.0, 0.0)]), 0.0, what="yaw 90 cart")
    # random attitudes, residuals built from scipy's inverse rotation
    # (the observation vector is an axis, the only unit vectors f32 holds exactly: cpu.rs does
    # not normalise it, and acos near 1 would turn a 1e-8 length error into a 1e-4 angle)
    rng = np.random.default_rng(5)
    for _ in range(20):
        r = Rotation.random(random_state=rng)
        p = rng.normal(size=3) * 3
        v = np.zeros(3)
        v[rng.integers(3)] = rng.choice([-1.0, 1.0])
        perp = np.cross(v, rng.normal(size=3))
        u_body = Rotation.from_rotvec(perp / np.linalg.norm(perp) * rng.uniform(0.0, 0.055)).apply(-v)
        lmk = (p + r.apply(u_body * rng.uniform(2.0, 8.0))).astype(np.float32)
        b = r.inv().apply(lmk.astype(np.float64) - p)
        rn = float(np.linalg.norm(b))
        d = float(rng.uniform(-0.1, 0.1))
        o = (0, 1, *v, np.float32(rn - d))
        th = math.atan2(np.linalg.norm(np.cross(b, -v)), float(b @ -v))
        der = rn - float(np.float32(rn - d))
        want = -der * der / (2 * SR * SR) - th * th / (2 * SB * SB)
        _close(ll_single(L, p, _quat_wxyz(r), [lmk], [o]), want, rel=1e-7, what="random attitude")
        e = np.array([0.03, -0.02, 0.01])
        v = (b - e).astype(np.float32)
        ee = b - v.astype(np.float64)
        _close(ll_single(L, p, _quat_wxyz(r), [lmk], [(0, 0, *v, 0.0)]), -(ee @ ee) / (2 * SR * SR), rel=1e-7,
               what="random attitude cart")
    # several observations add; an unknown id is dropped
    lm2 = [[5.0, 2.0, 3.0], [1.0, 2.0, 7.0]]  # body (4,0,0) and (0,0,4)
    obs = [(0, 1, -1.0, 0.0, 0.0, 4.0 - 0.03125), (1, 1, 0.0, 0.0, -1.0, 4.0 + 0.140625), (7, 0, 9.0, 9.0, 9.0, 0.0)]
    _close(ll_single(L, P0, IDQ, lm2, obs), -0.1953125 - 3.955078125, what="sum of observations")
    # sigma floors at 1e-6: d = 2^-20 with sr = 0 -> -d^2 / (2e-12)
    d = 2.0 ** -20
    _close(ll_single(L, P0, IDQ, [[5.0, 2.0, 3.0]], [(0, 0, 4.0 - d, 0.0, 0.0, 0.0)], sr=0.0, gate=0.0),
           -(d * d) / 2e-12, what="sigma floor")


def check_softmax_two_particles(L):
    f = O.OracleFilter(2, 1, False, library=L)
    # the particles sit 1/32 and 1/16 further from the landmark along x: range residuals 1/32
    # and 1/16, ll = -0.1953125 and -0.78125 (neither is 0, so the max shift matters)
    f.set_state(np.array([P0 - [0.03125, 0.0, 0.0], P0 - [0.0625, 0.0, 0.0]]), np.array([IDQ, IDQ]))
    f.weight(_obs([(0, 1, -1.0, 0.0, 0.0, 4.0)]), np.array([[5.0, 2.0, 3.0]], np.float32), SR, SB, G)
    w = f.state()[2]
    w0 = 1.0 / (1.0 + math.exp(-0.78125 + 0.1953125))
    _close(w[0], w0, rel=1e-14, what="softmax w0")
    _close(w[1], 1.0 - w0, rel=1e-14, what="softmax w1")


# ------------------------------------------------------------------ motion


def check_predict_composition(L):
    """cpu.rs:197-218: p' = p + R(q)(t + tstd*n[0:3]);  q' = q * q_odo * exp(rstd*n[3:6])."""
    n = 64
    rng = np.random.default_rng(9)
    r0 = Rotation.from_euler("x", 90, degrees=True)  # attitude about x ...
    ro = Rotation.from_euler("z", 90, degrees=True)  # ... odometry about z: the products do not commute
    for noisy in (False, True):
        f = O.OracleFilter(n, 3, False, library=L)
        pos = rng.normal(size=(n, 3))
        quat = np.tile(_quat_wxyz(r0), (n, 1))
        f.set_state(pos, quat)
        t = np.array([0.0, 1.0, 0.25])
        tstd = np.array([0.02, 0.03, 0.01]) if noisy else np.zeros(3)
        rstd = np.array([0.05, 0.02, 0.08]) if noisy else np.zeros(3)
        f.predict(t, _quat_wxyz(ro), tstd, rstd)
        gs = f.rng()[0]
        p1, q1, _, _ = f.state()
        for i in range(n):
            z = O.particle_normals(gs, i, library=L)
            want_p = pos[i] + r0.apply(t + tstd * z[:3])
            assert np.allclose(p1[i], want_p, rtol=0, atol=1e-13), ("predict position", i)
            want_r = r0 * ro * Rotation.from_rotvec(rstd * z[3:6])
            got = _rot(q1[i])
            assert (got.inv() * want_r).magnitude() < 1e-12, ("predict attitude", i)
        if not noisy:  # q * q_odo is x then z: a rotation matrix product R_x R_z
            assert np.allclose(_rot(q1[0]).as_matrix(), r0.as_matrix() @ ro.as_matrix(), atol=1e-14)


def check_reinit_composition(L):
    """cpu.rs:180-195: p = pose + pos_std*n[0:3];  q = q0 * exp(rot_std*n[3:6])."""
    n = 64
    r0 = Rotation.from_euler("xyz", [0.4, -0.3, 1.1])
    pose = np.array([1.5, -2.0, 0.5, *_quat_wxyz(r0)])
    ps = np.array([1.0, 2.0, 0.5])
    rs = np.array([0.3, 0.1, 0.6])
    f = O.OracleFilter(n, 17, False, library=L)
    f.reinit(pose, ps, rs)
    gs = f.rng()[0]
    p1, q1, w, _ = f.state()
    assert np.all(w == 1.0 / n)
    for i in range(n):
        z = O.particle_normals(gs, i, library=L)
        assert np.allclose(p1[i], pose[:3] + ps * z[:3], rtol=0, atol=1e-14), ("reinit position", i)
        want_r = r0 * Rotation.from_rotvec(rs * z[3:6])
        assert (_rot(q1[i]).inv() * want_r).magnitude() < 1e-12, ("reinit attitude", i)


# ------------------------------------------------------------------ estimate


def check_estimate_two_particles(L):
    """Two particles, weights (a, 1 - a): mean a p1 + (1 - a) p2, covariance a(1 - a) d d^T;
    attitudes mean * exp(+-theta/2 about x): mean attitude = mean, att_cov = theta^2 e_x e_x^T."""
    f = O.OracleFilter(2, 1, False, library=L)
    d = np.array([0.3, -0.2, 0.1])
    qm = Rotation.from_euler("z", 0.7)
    theta = 0.2
    qa = qm * Rotation.from_rotvec([theta, 0, 0])
    qb = qm * Rotation.from_rotvec([-theta, 0, 0])
    f.set_state(np.array([P0 + d, P0 - d]), np.array([_quat_wxyz(qa), _quat_wxyz(qb)]), np.array([0.5, 0.5]))
    e = f.estimate()
    assert np.allclose(e["position"], P0, atol=1e-15)
    assert np.allclose(e["pos_cov"], np.outer(d, d), atol=1e-15)
    assert (_rot(e["attitude"]).inv() * qm).magnitude() < 1e-12
    want = np.zeros((3, 3))
    want[0, 0] = theta * theta
    assert np.allclose(e["att_cov"], want, atol=1e-13), e["att_cov"]
    _close(e["ess"], 2.0, rel=1e-15, what="ess")
    # unequal weights: mean a p1 + (1 - a) p2, covariance a (1 - a) (p1 - p2)(p1 - p2)^T, ESS 1/(a^2+(1-a)^2)
    a = 0.25
    f.set_state(np.array([P0 + d, P0 - d]), np.array([IDQ, IDQ]), np.array([a, 1 - a]))
    e = f.estimate()
    assert np.allclose(e["position"], P0 + (2 * a - 1) * d, atol=1e-15)
    assert np.allclose(e["pos_cov"], a * (1 - a) * np.outer(2 * d, 2 * d), atol=1e-15)
    _close(e["ess"], 1.0 / (a * a + (1 - a) ** 2), rel=1e-15, what="ess unequal")


ALL_CHECKS = [check_range_bearing_closed_forms, check_cartesian_closed_forms, check_attitude_and_bookkeeping,
              check_softmax_two_particles, check_predict_composition, check_reinit_composition,
              check_estimate_two_particles]


def test_range_bearing_closed_forms():
    check_range_bearing_closed_forms(O.lib())


def test_cartesian_closed_forms():
    check_cartesian_closed_forms(O.lib())


def test_attitude_and_bookkeeping():
    check_attitude_and_bookkeeping(O.lib())


def test_softmax_two_particles():
    check_softmax_two_particles(O.lib())


def test_predict_composition_non_commuting():
    check_predict_composition(O.lib())


def test_reinit_composition():
    check_reinit_composition(O.lib())


def test_estimate_two_particles():
    check_estimate_two_particles(O.lib())

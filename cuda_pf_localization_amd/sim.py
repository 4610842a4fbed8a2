"""Deterministic scenario fixtures.

`Gauss`, `landmark_box`, `circle_scenario` and `room_grid` restate the reference's frozen
generators crates/pf_core/src/sim.rs:14-151 (xorshift64 draws in the same order).  The
other generators build the synthetic workloads of the benchmark configurations (helix
tracking with range-bearing observations, random-obstacle occupancy maps); their draws come
from a splitmix64 stream so they are reproducible from the seed alone.

Range-bearing convention: both reference backends score dot(-pred/|pred|, v)
(cpu.rs:82-84, pf_kernels.cu:258-259), i.e. the likelihood peaks when v is the unit vector
from the landmark to the sensor, expressed in the body frame.  The generators emit that v.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

MASK64 = (1 << 64) - 1


# --------------------------------------------------------------- quaternion helpers (wxyz)

def qmul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return (aw * bw - ax * bx - ay * by - az * bz,
            aw * bx + ax * bw + ay * bz - az * by,
            aw * by - ax * bz + ay * bw + az * bx,
            aw * bz + ax * by - ay * bx + az * bw)


def qconj(q):
    return (q[0], -q[1], -q[2], -q[3])


def qrot(q, v):
    w, x, y, z = q
    tx = 2.0 * (y * v[2] - z * v[1])
    ty = 2.0 * (z * v[0] - x * v[2])
    tz = 2.0 * (x * v[1] - y * v[0])
    cx = y * tz - z * ty
    cy = z * tx - x * tz
    cz = x * ty - y * tx
    return (v[0] + w * tx + cx, v[1] + w * ty + cy, v[2] + w * tz + cz)


def yaw_quat(yaw):
    return (math.cos(0.5 * yaw), 0.0, 0.0, math.sin(0.5 * yaw))


# ------------------------------------------------------------------ sim.rs restatement

class Gauss:
    """sim.rs:14-42: xorshift64 source; normal() is the frozen triangular approximation."""

    def __init__(self, seed: int):
        self.s = seed & MASK64

    def _step(self):
        s = self.s
        s ^= (s << 13) & MASK64
        s ^= s >> 7
        s ^= (s << 17) & MASK64
        self.s = s
        return s

    def next_box(self):
        return (self._step() % 8000) / 1000.0 + 1.0

    def next01(self):
        return (self._step() % 10000) / 10000.0

    def normal(self):
        return (self.next01() + self.next01() - 1.0) * 1.732


def landmark_box(n: int, seed: int) -> dict:
    """sim.rs:44-52: n landmarks in an ~8 x 8 x 2.4 m box; {id: (x, y, z)}."""
    g = Gauss(seed)
    out = {}
    for i in range(n):
        x = g.next_box()
        y = g.next_box()
        z = g.next_box() * 0.3
        out[i] = (x, y, z)
    return out


@dataclass
class Step:
    odom_t: tuple
    odom_q: tuple
    trans_noise: tuple
    rot_noise: tuple
    obs: list  # (id, mode, (vx, vy, vz), range)
    truth_pos: tuple
    truth_q: tuple


@dataclass
class Scenario:
    landmarks: dict
    prior_pos: tuple
    prior_q: tuple
    steps: list = field(default_factory=list)


def circle_scenario(steps_n: int, seed: int) -> Scenario:
    """sim.rs:69-121: circular climb at dt = 0.1 s, Cartesian observations within 5 m."""
    lm = landmark_box(12, 99)
    prior_pos = (5.0, 5.0, 0.5)
    prior_q = (1.0, 0.0, 0.0, 0.0)
    g = Gauss(seed)
    dt = 0.1
    pos, att = prior_pos, prior_q
    steps = []
    for t in range(steps_n):
        a = 0.2 * float(t) * dt
        vel = (-0.4 * math.sin(a), 0.4 * math.cos(a), 0.02)
        yaw = math.atan2(vel[1], vel[0])
        natt = yaw_quat(yaw)
        new_pos = (pos[0] + vel[0] * dt, pos[1] + vel[1] * dt, pos[2] + vel[2] * dt)
        tr = qrot(qconj(att), (new_pos[0] - pos[0], new_pos[1] - pos[1], new_pos[2] - pos[2]))
        rot = qmul(qconj(att), natt)
        n0 = g.normal() * 0.05
        n1 = g.normal() * 0.05
        n2 = g.normal() * 0.05
        tr = (tr[0] + n0, tr[1] + n1, tr[2] + n2)
        pos, att = new_pos, natt
        obs = []
        for i in range(12):
            p = lm[i]
            body = qrot(qconj(att), (p[0] - pos[0], p[1] - pos[1], p[2] - pos[2]))
            d = math.sqrt(body[0] ** 2 + body[1] ** 2 + body[2] ** 2)
            if d > 5.0:
                continue
            s = 1.0 + g.normal() * 0.05 / max(d, 0.1)
            obs.append((i, 0, (body[0] * s, body[1] * s, body[2] * s), 0.0))
        steps.append(Step(tr, rot, (0.05, 0.05, 0.05), (0.02, 0.02, 0.03), obs, pos, att))
    return Scenario(lm, prior_pos, prior_q, steps)


def room_grid() -> tuple:
    """sim.rs:123-151: 10 x 10 m room at 0.05 m, border walls + pillar at (4, 5).
    Returns (cells[H, W] uint8, res, (ox, oy))."""
    w = h = 200
    cells = np.zeros((h, w), np.uint8)
    cells[0, :] = 1
    cells[h - 1, :] = 1
    cells[:, 0] = 1
    cells[:, w - 1] = 1
    x0, x1 = int((4.0 - 0.25) / 0.05), int((4.0 + 0.25) / 0.05)
    y0, y1 = int((5.0 - 0.25) / 0.05), int((5.0 + 0.25) / 0.05)
    cells[y0:y1 + 1, x0:x1 + 1] = 1
    return cells, 0.05, (0.0, 0.0)


# ------------------------------------------------------------------ synthetic workloads

class SplitMix:
    """splitmix64 stream (pf_kernels.cu:323-328 mixer) for fixture generation."""

    def __init__(self, seed: int):
        self.x = seed & MASK64

    def next(self) -> int:
        self.x = (self.x + 0x9E3779B97F4A7C15) & MASK64
        z = self.x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def u01(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def uniform(self, a, b):
        return a + (b - a) * self.u01()

    def normal(self):
        u1 = ((self.next() >> 11) + 1) * (1.0 / 9007199254740992.0)
        u2 = self.u01()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


@dataclass
class HelixScenario:
    landmarks: dict
    prior_pos: tuple
    prior_q: tuple
    steps: list
    sigma_r: float
    sigma_b: float
    gate: float


def helix_scenario(n_landmarks: int, box, cycles: int, seed: int, sigma_r: float = 0.05, sigma_b: float = 0.02,
                   outlier_frac: float = 0.0, outlier_max: float = 60.0, odo_pos: float = 0.02,
                   odo_rot: float = 0.01, radius: float = 5.0, pitch: float = 0.5, cycles_per_rev: int = 100,
                   phase: float = 0.0, center=None, gate: float = 3.0) -> HelixScenario:
    """Helix tracking with range-bearing observations of every landmark.

    Landmarks U(box); helix of `radius` m, `pitch` m per revolution, `cycles_per_rev`
    cycles per revolution, yaw along the tangent.  Odometry = true body-frame motion plus
    N(0, odo_pos^2) translation noise; the filter's process noise is (odo_pos, odo_rot).
    Each observation: range + N(0, sigma_r^2) (or, with probability outlier_frac,
    U[0, outlier_max]); unit vector landmark->sensor in the body frame, rotated by
    N(0, sigma_b^2) about a random perpendicular axis."""
    rng = SplitMix(seed)
    bx, by, bz = box
    lm = {i: (rng.uniform(0, bx), rng.uniform(0, by), rng.uniform(0, bz)) for i in range(n_landmarks)}
    cx, cy = center if center is not None else (bx / 2.0, by / 2.0)
    z0 = bz / 2.0

    def pose_at(k):
        th = phase + 2.0 * math.pi * k / cycles_per_rev
        p = (cx + radius * math.cos(th), cy + radius * math.sin(th), z0 + pitch * k / cycles_per_rev)
        return p, yaw_quat(th + 0.5 * math.pi)

    pos, att = pose_at(0)
    steps = []
    for k in range(1, cycles + 1):
        npos, natt = pose_at(k)
        tr = qrot(qconj(att), (npos[0] - pos[0], npos[1] - pos[1], npos[2] - pos[2]))
        tr = (tr[0] + odo_pos * rng.normal(), tr[1] + odo_pos * rng.normal(), tr[2] + odo_pos * rng.normal())
        rot = qmul(qconj(att), natt)
        pos, att = npos, natt
        obs = []
        for i in range(n_landmarks):
            p = lm[i]
            body = qrot(qconj(att), (p[0] - pos[0], p[1] - pos[1], p[2] - pos[2]))
            d = math.sqrt(body[0] ** 2 + body[1] ** 2 + body[2] ** 2)
            v = np.array([-body[0] / d, -body[1] / d, -body[2] / d])
            # perturb the direction by a small rotation about a random perpendicular axis
            a = np.array([rng.normal(), rng.normal(), rng.normal()])
            a -= a.dot(v) * v
            na = np.linalg.norm(a)
            if na > 0:
                ang = sigma_b * rng.normal()
                a /= na
                v = v * math.cos(ang) + np.cross(a, v) * math.sin(ang)
                v /= np.linalg.norm(v)
            if outlier_frac > 0.0 and rng.u01() < outlier_frac:
                r = rng.uniform(0.0, outlier_max)
            else:
                r = d + sigma_r * rng.normal()
            obs.append((i, 1, (float(v[0]), float(v[1]), float(v[2])), float(r)))
        steps.append(Step(tr, rot, (odo_pos,) * 3, (odo_rot,) * 3, obs, npos, natt))
    p0, q0 = pose_at(0)
    return HelixScenario(lm, p0, q0, steps, sigma_r, sigma_b, gate)


def config0_scenario(cycles: int = 100, seed: int = 42) -> HelixScenario:
    """10k-particle tracking: 12 landmarks in a 20 x 20 x 5 m box, RB obs, sigma 0.05 m."""
    return helix_scenario(12, (20.0, 20.0, 5.0), cycles, seed)


def config1_scenario(cycles: int = 200, seed: int = 7) -> HelixScenario:
    """1M-particle tracking: 256 landmarks in 50 x 50 x 10 m, 1% range outliers U[0, 60]."""
    return helix_scenario(256, (50.0, 50.0, 10.0), cycles, seed, outlier_frac=0.01, outlier_max=60.0)


def obstacle_map(seed: int, size_m: float = 10.0, res: float = 0.05, n_rect: int = 6, n_circ: int = 4):
    """Occupancy map: size_m x size_m at res, 1-cell border walls, n_rect axis-aligned
    rectangles (sides U[0.3, 2] m) and n_circ discs (r U[0.2, 0.6] m).
    Returns (cells[H, W] uint8, res, (0, 0))."""
    rng = SplitMix(seed)
    n = int(round(size_m / res))
    cells = np.zeros((n, n), np.uint8)
    cells[0, :] = cells[-1, :] = 1
    cells[:, 0] = cells[:, -1] = 1
    yy, xx = np.mgrid[0:n, 0:n]
    cxm = (xx + 0.5) * res
    cym = (yy + 0.5) * res
    for _ in range(n_rect):
        w = rng.uniform(0.3, 2.0)
        h = rng.uniform(0.3, 2.0)
        x0 = rng.uniform(0.0, size_m - w)
        y0 = rng.uniform(0.0, size_m - h)
        cells[(cxm >= x0) & (cxm <= x0 + w) & (cym >= y0) & (cym <= y0 + h)] = 1
    for _ in range(n_circ):
        r = rng.uniform(0.2, 0.6)
        x0 = rng.uniform(r, size_m - r)
        y0 = rng.uniform(r, size_m - r)
        cells[(cxm - x0) ** 2 + (cym - y0) ** 2 <= r * r] = 1
    return cells, res, (0.0, 0.0)


def free_poses(cells, res, origin, n: int, seed: int, margin_cells: int = 2) -> np.ndarray:
    """n poses (x, y, yaw) uniform over free cells at least margin_cells from any
    obstacle; yaw U[-pi, pi)."""
    from scipy.ndimage import binary_dilation
    occ = cells.astype(bool)
    if margin_cells > 0:
        occ = binary_dilation(occ, iterations=margin_cells)
    free = np.argwhere(~occ)
    rng = SplitMix(seed)
    out = np.zeros((n, 3))
    for i in range(n):
        cy, cx = free[int(rng.u01() * len(free))]
        out[i] = (origin[0] + (cx + rng.u01()) * res, origin[1] + (cy + rng.u01()) * res,
                  -math.pi + 2 * math.pi * rng.u01())
    return out


def f64_scan(cells, res, origin, pose, nrays: int, span: float, rmax: float) -> np.ndarray:
    """Noiseless map_scan (global_loc.rs:200-214) in f64, vectorised over rays."""
    H, W = cells.shape
    x, y, th = pose
    out = np.empty(nrays)
    for r in range(nrays):
        a = th + (span * r / (nrays - 1) if nrays > 1 else 0.0)
        out[r] = _cast_ray64(cells, W, H, res, origin, x, y, math.cos(a), math.sin(a), rmax)
    return out


def _cast_ray64(cells, W, H, res, origin, px, py, dx, dy, rmax):
    """global_loc.rs:152-196."""
    fx = (px - origin[0]) / res
    fy = (py - origin[1]) / res
    if fx < 0.0 or fy < 0.0:
        return rmax
    cx, cy = int(fx), int(fy)
    if not (cx < W and cy < H):
        return rmax
    sx = 1 if dx > 0.0 else -1
    sy = 1 if dy > 0.0 else -1
    adx, ady = abs(dx), abs(dy)
    tdx = res / adx if adx > 1e-9 else math.inf
    tdy = res / ady if ady > 1e-9 else math.inf
    frx = (px - origin[0]) / res - cx
    fry = (py - origin[1]) / res - cy
    tmx = ((1.0 - frx) if dx > 0.0 else frx) * tdx if adx > 1e-9 else math.inf
    tmy = ((1.0 - fry) if dy > 0.0 else fry) * tdy if ady > 1e-9 else math.inf
    for _ in range(8192):
        if tmx <= tmy:
            dist = tmx
            tmx += tdx
            cx += sx
        else:
            dist = tmy
            tmy += tdy
            cy += sy
        if dist > rmax:
            return rmax
        if cx < 0 or cy < 0 or cx >= W or cy >= H:
            return rmax
        if cells[cy, cx] == 1:
            return min(dist, rmax)
    return rmax


# ------------------------------------------------------------------ vectorised helix (batched workloads)


def _qrot_rows(q, v):
    """Rotates each row of v[n, 3] by the unit quaternion q (wxyz)."""
    w, x, y, z = q
    u = np.array([x, y, z])
    t = 2.0 * np.cross(u[None, :], v)
    return v + w * t + np.cross(u[None, :], t)


@dataclass
class HelixArrays:
    """One helix run as arrays: landmarks[L, 3]; odometry t[C, 3], q[C, 4]; observations
    v[C, L, 3] (unit landmark->sensor, body frame) and r[C, L]; truth pos[C, 3]."""
    landmarks: np.ndarray
    prior_pos: tuple
    prior_q: tuple
    odom_t: np.ndarray
    odom_q: np.ndarray
    obs_v: np.ndarray
    obs_r: np.ndarray
    truth: np.ndarray
    odo_pos: float
    odo_rot: float


def helix_arrays(n_landmarks: int, box, cycles: int, seed: int, phase: float = 0.0, sigma_r: float = 0.05,
                 sigma_b: float = 0.02, odo_pos: float = 0.02, odo_rot: float = 0.01, radius: float = 5.0,
                 pitch: float = 0.5, cycles_per_rev: int = 100) -> HelixArrays:
    """The helix_scenario workload generated with numpy (PCG64 draws, not the splitmix64
    stream): same geometry, noise model and observation convention, fast enough for
    batches of many filters."""
    rng = np.random.default_rng(seed)
    bx, by, bz = box
    lm = rng.uniform(0.0, 1.0, (n_landmarks, 3)) * np.array([bx, by, bz])
    cx, cy, z0 = bx / 2.0, by / 2.0, bz / 2.0

    def pose_at(k):
        th = phase + 2.0 * math.pi * k / cycles_per_rev
        return (np.array([cx + radius * math.cos(th), cy + radius * math.sin(th), z0 + pitch * k / cycles_per_rev]),
                yaw_quat(th + 0.5 * math.pi))

    pos, att = pose_at(0)
    ot = np.zeros((cycles, 3))
    oq = np.zeros((cycles, 4))
    ov = np.zeros((cycles, n_landmarks, 3))
    orr = np.zeros((cycles, n_landmarks))
    truth = np.zeros((cycles, 3))
    for k in range(1, cycles + 1):
        npos, natt = pose_at(k)
        ot[k - 1] = _qrot_rows(qconj(att), (npos - pos)[None, :])[0] + odo_pos * rng.normal(size=3)
        oq[k - 1] = qmul(qconj(att), natt)
        pos, att = npos, natt
        body = _qrot_rows(qconj(att), lm - pos)
        d = np.linalg.norm(body, axis=1)
        v = -body / d[:, None]
        a = rng.normal(size=(n_landmarks, 3))
        a -= np.sum(a * v, axis=1)[:, None] * v
        a /= np.maximum(np.linalg.norm(a, axis=1), 1e-300)[:, None]
        ang = sigma_b * rng.normal(size=n_landmarks)
        v = v * np.cos(ang)[:, None] + np.cross(a, v) * np.sin(ang)[:, None]
        v /= np.linalg.norm(v, axis=1)[:, None]
        ov[k - 1] = v
        orr[k - 1] = d + sigma_r * rng.normal(size=n_landmarks)
        truth[k - 1] = npos
    p0, q0 = pose_at(0)
    return HelixArrays(lm, tuple(p0), q0, ot, oq, ov, orr, truth, odo_pos, odo_rot)


def config3_filters(nf: int = 64, cycles: int = 100, seed: int = 11):
    """The batched-filters workload: nf independent filters, particle counts log-uniform in
    [1k, 500k], landmark counts uniform in {4..256} in per-filter 20 x 20 x 5 m boxes, each on
    its own helix with a random phase (config-0 noise).  Returns (ns, ls, [HelixArrays])."""
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(math.log(1000), math.log(500_000), nf)).astype(int)
    ls = rng.integers(4, 257, nf)
    phases = rng.uniform(0.0, 2 * math.pi, nf)
    hs = [helix_arrays(int(ls[f]), (20.0, 20.0, 5.0), cycles, seed * 1000 + f, phase=float(phases[f]))
          for f in range(nf)]
    return ns, ls, hs

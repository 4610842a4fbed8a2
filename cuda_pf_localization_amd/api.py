"""Python-facing wrappers of the gfx950 particle-filter / CBGL libraries.

`PfKernels` and `GlKernels` call the reference ABI (pfc_* / pfgl_*) one to one, so code
written against pf_kernels.cu reads the same here.  `ParticleFilter`, `GlobalLoc` and
`quat_mean` drive the C++ host layer (libpf_host.so) that restates the reference's Rust
engine (crates/pf_core/src/{particle_filter,cuda,global_loc,types}.rs).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib

OBS_DTYPE = np.dtype([("id", "<i4"), ("mode", "<i4"), ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"),
                      ("range", "<f4")])
assert OBS_DTYPE.itemsize == 24  # ObsF / CObservation, pf_kernels.cu:79-84, cuda.rs:18-27

STAGES = ["reinit", "predict", "weight", "normalize", "resample_scan", "resample_gather", "inject", "est_sums",
          "est_covs", "finalize"]
GL_STAGES = ["generate", "score", "topk"]


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed (rc={rc})")


def _text(x) -> str:
    """x itself when it looks like file content (has a newline or a ':'), else the file it names."""
    if isinstance(x, (bytes, bytearray)):
        return bytes(x).decode()
    if "\n" in x or ":" in x:
        return x
    with open(x) as f:
        return f.read()


def load_map_pgm_yaml(pgm, yaml):
    """ROS map_server pair (global_loc.rs:32-130) -> (cells[H, W] uint8, res, (ox, oy)); row 0 is
    the bottom of the map.  `pgm` is a path or the file bytes, `yaml` a path or the text."""
    H = _lib.host()
    if isinstance(pgm, (bytes, bytearray)):
        pb = bytes(pgm)
    else:
        with open(pgm, "rb") as f:
            pb = f.read()
    yt = _text(yaml)
    buf = np.frombuffer(bytes(pb), np.uint8)
    w, h = C.c_int(), C.c_int()
    res, ox, oy = C.c_double(), C.c_double(), C.c_double()
    err = C.create_string_buffer(256)
    args = (buf.ctypes.data_as(C.POINTER(C.c_uint8)), buf.size, yt.encode(), C.byref(w), C.byref(h), C.byref(res),
            C.byref(ox), C.byref(oy))
    if H.pfh_map_from_pgm_yaml(*args, None, 0, err, 256) != 0:
        raise ValueError(err.value.decode())
    cells = np.zeros((h.value, w.value), np.uint8)
    if H.pfh_map_from_pgm_yaml(*args, cells.ctypes.data_as(C.POINTER(C.c_uint8)), cells.size, err, 256) != 0:
        raise ValueError(err.value.decode())
    return cells, res.value, (ox.value, oy.value)


def load_landmarks(text_or_path) -> dict:
    """Landmark text (landmark_map.rs:59-80): `id: [x, y, z]` or `id x y z` lines -> {id: xyz};
    a repeated id replaces the earlier entry."""
    text = _text(text_or_path)
    H = _lib.host()
    cap = text.count("\n") + 1
    ids = np.zeros(cap, np.int32)
    xyz = np.zeros((cap, 3))
    err = C.create_string_buffer(256)
    n = H.pfh_landmarks_from_text(text.encode(), _ip(ids), _dp(xyz), cap, err, 256)
    if n < 0:
        raise ValueError(err.value.decode())
    return {int(ids[i]): tuple(float(v) for v in xyz[i]) for i in range(n)}


def device_available() -> bool:
    return _lib.kernels().pfc_device_available() == 1


def make_obs(ids, modes, vecs, ranges) -> np.ndarray:
    o = np.zeros(len(ids), dtype=OBS_DTYPE)
    o["id"] = ids
    o["mode"] = modes
    v = np.asarray(vecs, dtype=np.float64).reshape(-1, 3)
    o["dx"], o["dy"], o["dz"] = v[:, 0], v[:, 1], v[:, 2]
    o["range"] = ranges
    return o


class PfKernels:
    """One pfc_* handle (pf_kernels.cu:475-735)."""

    def __init__(self, cap_obs: int = 256):
        self.L = _lib.kernels()
        self.h = self.L.pfc_create(cap_obs)
        if not self.h:
            raise RuntimeError("pfc_create failed")
        self.n = 0

    def close(self):
        if self.h:
            self.L.pfc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weight_variant(self, variant: int):
        """Likelihood kernel launch shape: 1, 2 or 4 particles per lane."""
        _check(self.L.pfx_set_weight_variant(self.h, variant), "pfx_set_weight_variant")

    def set_scan_spin_limit(self, limit: int):
        """Reads of a predecessor tile's status in the resampling scan before the tile sums the
        predecessors' weights itself (same bits either way)."""
        _check(self.L.pfx_set_scan_spin_limit(self.h, limit), "pfx_set_scan_spin_limit")

    def init(self, n: int, seed: int):
        _check(self.L.pfc_init(self.h, n, seed & 0xFFFFFFFF), "pfc_init")
        self.n = n

    def upload_landmarks(self, xyz):
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float32).reshape(-1))
        _check(self.L.pfc_upload_landmarks(self.h, _fp(a), a.size // 3), "pfc_upload_landmarks")

    def reinit(self, pose7, pos_std, rot_std):
        ps = np.asarray(pos_std, np.float32)
        rs = np.asarray(rot_std, np.float32)
        _check(self.L.pfc_reinit(self.h, *[float(v) for v in pose7], _fp(ps), _fp(rs)), "pfc_reinit")

    def predict(self, t, q, tstd, rstd):
        ts = np.asarray(tstd, np.float32)
        rs = np.asarray(rstd, np.float32)
        _check(self.L.pfc_predict(self.h, *[float(v) for v in t], *[float(v) for v in q], _fp(ts), _fp(rs)),
               "pfc_predict")

    def weight(self, obs: np.ndarray | None, sigma_r: float, sigma_b: float, gate: float):
        if obs is None or len(obs) == 0:
            rc = self.L.pfc_weight(self.h, None, 0, sigma_r, sigma_b, gate)
        else:
            o = np.ascontiguousarray(obs, dtype=OBS_DTYPE)
            rc = self.L.pfc_weight(self.h, o.ctypes.data, len(o), sigma_r, sigma_b, gate)
        _check(rc, "pfc_weight")

    def resample(self, ratio: float = 0.0):
        _check(self.L.pfc_resample(self.h, ratio), "pfc_resample")

    def est_sums(self) -> np.ndarray:
        out = np.zeros(14)
        _check(self.L.pfc_est_sums(self.h, _dp(out)), "pfc_est_sums")
        return out

    def covs(self, mean_pos, mean_quat) -> np.ndarray:
        out = np.zeros(12)
        _check(self.L.pfc_covs(self.h, *[float(v) for v in mean_pos], *[float(v) for v in mean_quat], _dp(out)),
               "pfc_covs")
        return out

    def snapshot(self):
        pos = np.zeros((self.n, 3), np.float32)
        quat = np.zeros((self.n, 4), np.float32)
        _check(self.L.pfc_snapshot(self.h, _fp(pos), _fp(quat)), "pfc_snapshot")
        return pos, quat

    # extensions
    def set_state(self, pos, quat, w=None):
        p = np.ascontiguousarray(pos, np.float32)
        q = np.ascontiguousarray(quat, np.float32)
        wp = None
        if w is not None:
            wa = np.ascontiguousarray(w, np.float32)
            wp = _fp(wa)
        _check(self.L.pfx_set_state(self.h, _fp(p), _fp(q), wp), "pfx_set_state")

    def get_state(self):
        pos = np.zeros((self.n, 3), np.float32)
        quat = np.zeros((self.n, 4), np.float32)
        w = np.zeros(self.n, np.float32)
        _check(self.L.pfx_get_state(self.h, _fp(pos), _fp(quat), _fp(w)), "pfx_get_state")
        return pos, quat, w

    def get_ll(self):
        ll = np.zeros(self.n)
        _check(self.L.pfx_get_ll(self.h, _dp(ll)), "pfx_get_ll")
        return ll

    def ancestors(self):
        a = np.zeros(self.n, np.int32)
        _check(self.L.pfx_get_ancestors(self.h, _ip(a)), "pfx_get_ancestors")
        return a

    def cum(self):
        """The last resampling scan's inclusive fixed-point prefix (u64 per particle)."""
        c = np.zeros(self.n, np.uint64)
        _check(self.L.pfx_get_cum(self.h, c.ctypes.data_as(_lib.P(_lib.u64))), "pfx_get_cum")
        return c

    def set_lcg(self, v: int):
        _check(self.L.pfx_set_lcg(self.h, v), "pfx_set_lcg")

    def get_lcg(self) -> int:
        v = C.c_uint64()
        _check(self.L.pfx_get_lcg(self.h, C.byref(v)), "pfx_get_lcg")
        return v.value

    def rng(self):
        g, n = C.c_uint64(), C.c_uint64()
        _check(self.L.pfx_get_rng(self.h, C.byref(g), C.byref(n)), "pfx_get_rng")
        return g.value, n.value

    def set_rng(self, gseed: int, gen: int):
        _check(self.L.pfx_set_rng(self.h, gseed, gen), "pfx_set_rng")

    def weight_scalars(self):
        out = np.zeros(4)
        _check(self.L.pfx_weight_scalars(self.h, _dp(out)), "pfx_weight_scalars")
        return out

    def profile(self, enable: bool = True):
        _check(self.L.pfx_profile_enable(self.h, 1 if enable else 0), "pfx_profile_enable")

    def profile_read(self):
        ms = np.zeros(len(STAGES))
        cnt = np.zeros(len(STAGES), np.int32)
        _check(self.L.pfx_profile_read(self.h, _dp(ms), _ip(cnt)), "pfx_profile_read")
        return {s: (float(ms[i]), int(cnt[i])) for i, s in enumerate(STAGES)}

    def sync(self):
        _check(self.L.pfx_sync(self.h), "pfx_sync")


class GlKernels:
    """One pfgl_* handle (pf_kernels.cu:895-1025)."""

    def __init__(self, grid: np.ndarray, res: float, ox: float, oy: float, cap_hyp: int):
        self.L = _lib.kernels()
        g = np.ascontiguousarray(grid, dtype=np.uint8)
        H, W = g.shape
        self.W, self.H = W, H
        self.h = self.L.pfgl_create(W, H, res, ox, oy, g.ctypes.data_as(C.POINTER(C.c_uint8)), cap_hyp)
        if not self.h:
            raise RuntimeError("pfgl_create failed")

    def close(self):
        if self.h:
            self.L.pfgl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_score_mode(self, mode: int):
        """0: by hypothesis count, 1: one wave per hypothesis, 2: one lane per ray, errors summed
        in a second pass (same CAER bits)."""
        _check(self.L.pfglx_set_score_mode(self.h, mode), "pfglx_set_score_mode")

    def set_prune(self, enable: bool):
        _check(self.L.pfglx_set_prune(self.h, 1 if enable else 0), "pfglx_set_prune")

    def free_cells(self) -> int:
        return self.L.pfgl_free_cells(self.h)

    def generate(self, n: int, seed: int):
        _check(self.L.pfgl_generate(self.h, n, seed), "pfgl_generate")

    def score(self, scan, rmax: float, ang_span: float):
        s = np.ascontiguousarray(scan, np.float32)
        _check(self.L.pfgl_score(self.h, _fp(s), s.size, rmax, ang_span), "pfgl_score")

    def topk(self, k: int):
        out = np.zeros((k, 3), np.float32)
        r = self.L.pfgl_topk(self.h, k, _fp(out))
        if r < 0:
            raise RuntimeError("pfgl_topk failed")
        return out[:r]

    def topk_ex(self, k: int):
        out = np.zeros((k, 3), np.float32)
        caer = np.zeros(k, np.float32)
        idx = np.zeros(k, np.int32)
        r = self.L.pfglx_topk_ex(self.h, k, _fp(out), _fp(caer), _ip(idx))
        if r < 0:
            raise RuntimeError("pfglx_topk_ex failed")
        return out[:r], caer[:r], idx[:r]

    def set_hypotheses(self, hyp):
        a = np.ascontiguousarray(hyp, np.float32).reshape(-1, 3)
        _check(self.L.pfglx_set_hypotheses(self.h, _fp(a), len(a)), "pfglx_set_hypotheses")

    def hypotheses(self):
        n = self.L.pfglx_hypothesis_count(self.h)
        a = np.zeros((max(n, 0), 3), np.float32)
        if n > 0 and self.L.pfglx_get_hypotheses(self.h, _fp(a)) < 0:
            raise RuntimeError("pfglx_get_hypotheses failed")
        return a

    def caer(self):
        n = self.L.pfglx_hypothesis_count(self.h)
        a = np.zeros(max(n, 0), np.float32)
        if n > 0 and self.L.pfglx_get_caer(self.h, _fp(a)) < 0:
            raise RuntimeError("pfglx_get_caer failed")
        return a

    def render(self, poses, nrays: int, rmax: float, ang_span: float):
        p = np.ascontiguousarray(poses, np.float32).reshape(-1, 3)
        out = np.zeros((len(p), nrays), np.float32)
        _check(self.L.pfglx_render(self.h, _fp(p), len(p), nrays, rmax, ang_span, _fp(out)), "pfglx_render")
        return out

    def profile(self, enable: bool = True):
        _check(self.L.pfglx_profile_enable(self.h, 1 if enable else 0), "pfglx_profile_enable")

    def profile_read(self):
        ms = np.zeros(len(GL_STAGES))
        cnt = np.zeros(len(GL_STAGES), np.int32)
        _check(self.L.pfglx_profile_read(self.h, _dp(ms), _ip(cnt)), "pfglx_profile_read")
        return {s: (float(ms[i]), int(cnt[i])) for i, s in enumerate(GL_STAGES)}

    def sync(self):
        _check(self.L.pfglx_sync(self.h), "pfglx_sync")


# ------------------------------------------------------------------ host layer


@dataclass
class PfConfig:
    """config.rs:3-64 (defaults config.rs:49-64); obs_cap = cuda.rs:95 OBS_CAP."""
    particle_count: int = 4096
    range_noise: float = 0.05
    bearing_noise: float = 0.02
    resample_ess_ratio: float = 0.5
    random_inject_ratio: float = 0.0
    gate_sigma: float = 0.0
    seed: int = 42
    obs_cap: int = 256
    fast: bool = False


@dataclass
class Observation:
    """types.rs:58-67.  mode 0 Cartesian (vector = body offset); 1 range-bearing
    (vector = unit direction, range in m)."""
    landmark_id: int
    mode: int
    vector: tuple
    range: float = 0.0


@dataclass
class OdomDelta:
    """types.rs:69-90 with its defaults."""
    translation: tuple = (0.0, 0.0, 0.0)
    rotation: tuple = (1.0, 0.0, 0.0, 0.0)
    trans_noise: tuple = (0.05, 0.05, 0.05)
    rot_noise: tuple = (0.02, 0.02, 0.03)


@dataclass
class Estimate:
    position: np.ndarray = field(default_factory=lambda: np.zeros(3))
    attitude: np.ndarray = field(default_factory=lambda: np.array([1.0, 0, 0, 0]))
    pos_cov: np.ndarray = field(default_factory=lambda: np.ones((3, 3)))
    att_cov: np.ndarray = field(default_factory=lambda: np.ones((3, 3)))
    ess: float = 0.0
    valid: bool = False
    resampled: bool = False


def _est(e: _lib.PfhEstimate) -> Estimate:
    return Estimate(np.array(e.pos[:]), np.array(e.quat[:]), np.array(e.pos_cov[:]).reshape(3, 3),
                    np.array(e.att_cov[:]).reshape(3, 3), float(e.ess), bool(e.valid), bool(e.resampled))


def _odom(o: OdomDelta) -> _lib.PfhOdom:
    c = _lib.PfhOdom()
    c.t[:] = list(o.translation)
    c.q[:] = list(o.rotation)
    c.trans_noise[:] = list(o.trans_noise)
    c.rot_noise[:] = list(o.rot_noise)
    return c


def _obs_array(obs) -> C.Array:
    arr = (_lib.PfhObs * max(1, len(obs)))()
    for i, o in enumerate(obs):
        arr[i].id = o.landmark_id
        arr[i].mode = o.mode
        arr[i].vx, arr[i].vy, arr[i].vz = (float(v) for v in o.vector)
        arr[i].range = float(o.range)
    return arr


class ParticleFilter:
    """particle_filter.rs:72-133 driving the HIP backend through the C++ host layer."""

    def __init__(self, cfg: PfConfig):
        self.H = _lib.host()
        c = _lib.PfhConfig(cfg.particle_count, cfg.range_noise, cfg.bearing_noise, cfg.resample_ess_ratio,
                           cfg.random_inject_ratio, cfg.gate_sigma, cfg.seed, cfg.obs_cap, 1 if cfg.fast else 0)
        self.f = self.H.pfh_filter_new(C.byref(c))
        if not self.f:
            raise RuntimeError("no usable GPU (pfc_device_available() == 0) or allocation failed")
        self.cfg = cfg

    def close(self):
        if self.f:
            self.H.pfh_filter_free(self.f)
            self.f = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_map(self, landmarks: dict):
        ids = np.array(sorted(landmarks), np.int32)
        xyz = np.array([landmarks[i] for i in ids], np.float64).reshape(-1, 3)
        _check(self.H.pfh_filter_set_map(self.f, _ip(ids), _dp(xyz), len(ids)), "pfh_filter_set_map")

    def reinit_from_gl(self, xyt, z: float = 0.0, pos_std=(0.25, 0.25, 0.1), rot_std=(0.02, 0.02, 0.1)):
        """GL -> PF hand-off: reinitialise about a global-localisation result (x, y, yaw)."""
        a = np.ascontiguousarray(np.asarray(xyt, np.float64)[:3])
        ps = np.ascontiguousarray(pos_std, np.float64)
        rs = np.ascontiguousarray(rot_std, np.float64)
        _check(self.H.pfh_filter_reinit_from_gl(self.f, _dp(a), float(z), _dp(ps), _dp(rs)),
               "pfh_filter_reinit_from_gl")

    def save(self, path: str):
        """Checkpoint: particles, stored weights and draw state."""
        _check(self.H.pfh_filter_save(self.f, str(path).encode()), "pfh_filter_save")

    def load(self, path: str):
        """Resume from a checkpoint of a filter with the same particle count."""
        _check(self.H.pfh_filter_load(self.f, str(path).encode()), "pfh_filter_load")

    def reinitialize(self, pose7, pos_std, rot_std):
        p = np.asarray(pose7, np.float64)
        ps = np.asarray(pos_std, np.float64)
        rs = np.asarray(rot_std, np.float64)
        _check(self.H.pfh_filter_reinitialize(self.f, _dp(p), _dp(ps), _dp(rs)), "pfh_filter_reinitialize")

    def update(self, odom: OdomDelta, obs) -> Estimate:
        e = _lib.PfhEstimate()
        arr = _obs_array(obs)
        _check(self.H.pfh_filter_update(self.f, C.byref(_odom(odom)), arr, len(obs), C.byref(e)),
               "pfh_filter_update")
        return _est(e)

    def update_raw(self, odom_c: _lib.PfhOdom, obs_arr, nobs: int, est_c: _lib.PfhEstimate) -> None:
        """update() with pre-built ctypes structures (benchmark loop, no per-call packing)."""
        _check(self.H.pfh_filter_update(self.f, C.byref(odom_c), obs_arr, nobs, C.byref(est_c)),
               "pfh_filter_update")

    def estimate(self) -> Estimate:
        e = _lib.PfhEstimate()
        _check(self.H.pfh_filter_estimate(self.f, C.byref(e)), "pfh_filter_estimate")
        return _est(e)

    def snapshot(self):
        n = self.cfg.particle_count
        pos = np.zeros((n, 3))
        quat = np.zeros((n, 4))
        w = np.zeros(n)
        _check(self.H.pfh_filter_snapshot(self.f, _dp(pos), _dp(quat), _dp(w)), "pfh_filter_snapshot")
        return pos, quat, w

    def kernels_handle(self):
        return self.H.pfh_filter_kernels(self.f)


def quat_mean(M) -> np.ndarray:
    """types.rs:128-144 through the host layer's Jacobi solver."""
    m = np.ascontiguousarray(M, np.float64).reshape(16)
    q = np.zeros(4)
    _check(_lib.host().pfh_quat_mean(_dp(m), _dp(q)), "pfh_quat_mean")
    return q


@dataclass
class GlParams:
    """global_loc.rs:236-257 defaults."""
    hypotheses: int = 50_000
    rays: int = 360
    rmax: float = 10.0
    ang_span: float = 2 * np.pi
    topk: int = 10
    seed: int = 42


class GlobalLoc:
    """CudaGlobalLoc (global_loc.rs:274-332) over the pfgl_* ABI; also returns CAER."""

    def __init__(self, grid: np.ndarray, res: float, origin, params: GlParams):
        self.H = _lib.host()
        g = np.ascontiguousarray(grid, np.uint8)
        p = _lib.PfhGlParams(params.hypotheses, params.rays, params.rmax, params.ang_span, params.topk, params.seed)
        self.g = self.H.pfh_gl_new(g.shape[1], g.shape[0], res, float(origin[0]), float(origin[1]),
                                   g.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(p))
        if not self.g:
            raise RuntimeError("pfgl_create failed")
        self.params = params

    def close(self):
        if self.g:
            self.H.pfh_gl_free(self.g)
            self.g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_seed(self, seed: int):
        self.params.seed = seed
        _check(self.H.pfh_gl_set_seed(self.g, seed), "pfh_gl_set_seed")

    def localize(self, scan) -> np.ndarray:
        s = np.ascontiguousarray(scan, np.float32)
        out = np.zeros((self.params.topk, 4))
        k = self.H.pfh_gl_localize(self.g, _fp(s), s.size, _dp(out))
        if k <= 0:
            raise RuntimeError("localize failed")
        return out[:k]

    def kernels_handle(self):
        return self.H.pfh_gl_kernels(self.g)

    def localize_learned(self, proposal: "LearnedProposal", scan, lp: "LearnedParams"):
        """Learned proposal with uniform fallback; returns (rows x, y, theta, caer; info dict)."""
        s = np.ascontiguousarray(scan, np.float32)
        out = np.zeros((self.params.topk, 4))
        p = _lib.PfhLearnedParams(lp.n_mix, lp.n_uniform, lp.fallback_hypotheses, lp.fallback_caer_per_ray, lp.seed,
                                  lp.refine_rounds, lp.refine_parents, lp.refine_children, lp.refine_sigma_xy,
                                  lp.refine_sigma_yaw, lp.refine_shrink)
        info = _lib.PfhLearnedInfo()
        k = self.H.pfh_gl_localize_learned(self.g, proposal.h, _fp(s), s.size, C.byref(p), _dp(out), C.byref(info))
        if k <= 0:
            raise RuntimeError("localize_learned failed")
        return out[:k], dict(used_fallback=bool(info.used_fallback),
                             proposal_best_caer_per_ray=info.proposal_best_caer_per_ray,
                             best_caer_per_ray=info.best_caer_per_ray)


# ------------------------------------------------------------------ learned proposal


@dataclass
class LearnedParams:
    """Learned-proposal localisation: mixture + uniform hypotheses, uniform fallback."""
    n_mix: int = 8000
    n_uniform: int = 2000
    fallback_hypotheses: int = 200_000
    fallback_caer_per_ray: float = 0.05
    seed: int = 1
    refine_rounds: int = 3
    refine_parents: int = 16
    refine_children: int = 64
    refine_sigma_xy: float = 0.12
    refine_sigma_yaw: float = 0.06
    refine_shrink: float = 0.35


class LearnedProposal:
    """pfl_* handle: per-map MLP scan -> Gaussian mixture over (x, y, sin yaw, cos yaw)."""

    def __init__(self, in_dim: int = 360, hid: int = 512, n_hidden: int = 3, ncomp: int = 16,
                 max_batch: int = 4096, seed: int = 0):
        self.L = _lib.kernels()
        self.dims = (in_dim, hid, n_hidden, ncomp)
        self.nout = 9 * ncomp
        self.max_batch = max_batch
        self.h = self.L.pfl_create(in_dim, hid, n_hidden, ncomp, max_batch, seed)
        if not self.h:
            raise RuntimeError("pfl_create failed")

    def close(self):
        if self.h:
            self.L.pfl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def param_count(self) -> int:
        return int(self.L.pfl_param_count(self.h))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.param_count(), np.float32)
        _check(self.L.pfl_get_params(self.h, _fp(p)), "pfl_get_params")
        return p

    def set_params(self, p):
        a = np.ascontiguousarray(p, np.float32)
        if a.size != self.param_count():
            raise ValueError("parameter vector has the wrong length")
        _check(self.L.pfl_set_params(self.h, _fp(a)), "pfl_set_params")

    def _x(self, x):
        a = np.ascontiguousarray(x, np.float32).reshape(-1, self.dims[0])
        if not 0 < len(a) <= self.max_batch:
            raise ValueError("batch size outside [1, max_batch]")
        return a

    def forward(self, x) -> np.ndarray:
        a = self._x(x)
        out = np.zeros((len(a), self.nout), np.float32)
        _check(self.L.pfl_forward(self.h, _fp(a), len(a), _fp(out)), "pfl_forward")
        return out

    def grad(self, x, y):
        a = self._x(x)
        t = np.ascontiguousarray(y, np.float32).reshape(len(a), 4)
        g = np.zeros(self.param_count(), np.float32)
        loss = C.c_double()
        _check(self.L.pfl_grad(self.h, _fp(a), _fp(t), len(a), _fp(g), C.byref(loss)), "pfl_grad")
        return loss.value, g

    def train_step(self, x, y, lr: float = 1e-3) -> float:
        a = self._x(x)
        t = np.ascontiguousarray(y, np.float32).reshape(len(a), 4)
        loss = C.c_double()
        _check(self.L.pfl_train_step(self.h, _fp(a), _fp(t), len(a), lr, C.byref(loss)), "pfl_train_step")
        return loss.value

    def train_on_map(self, gl_handle, n_samples: int = 1_000_000, batch: int = 4096, lr: float = 1e-3,
                     rmax: float = 8.0, ang_span: float = 2 * np.pi, noise_sigma: float = 0.02, seed: int = 1):
        """Trains on scans rendered on the GPU; returns (mean loss of the first 5 % of steps,
        of the last 5 %)."""
        ml = np.zeros(2)
        _check(self.L.pfl_train_on_map(self.h, gl_handle, n_samples, batch, lr, self.dims[0], rmax, ang_span,
                                       noise_sigma, seed, _dp(ml)), "pfl_train_on_map")
        return float(ml[0]), float(ml[1])

    def propose(self, gl_handle, scan, rmax: float, n_mix: int, n_uniform: int, seed: int):
        s = np.ascontiguousarray(scan, np.float32)
        _check(self.L.pfl_propose(self.h, gl_handle, _fp(s), s.size, rmax, n_mix, n_uniform, seed), "pfl_propose")
        return self.L.pfl_last_propose_ms(self.h)


# ------------------------------------------------------------------ batched filters


class BatchedFilters:
    """pfb_* ABI: nf independent filters advanced by one launch per stage."""

    def __init__(self, cfgs: list, obs_cap: int = 256):
        self.L = _lib.kernels()
        arr = (_lib.PfbFilterCfg * len(cfgs))()
        for i, c in enumerate(cfgs):
            arr[i] = _lib.PfbFilterCfg(c["n"], c.get("seed", 42), c.get("sigma_r", 0.05), c.get("sigma_b", 0.02),
                                       c.get("gate", 0.0), c.get("ratio", 0.5), c.get("inject", 0.0))
        self.h = self.L.pfb_create(len(cfgs), arr, obs_cap)
        if not self.h:
            raise RuntimeError("pfb_create failed")
        self.ns = [c["n"] for c in cfgs]
        self.nf = len(cfgs)
        self._keep = []

    def close(self):
        if self.h:
            self.L.pfb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_landmarks(self, f: int, xyz):
        a = np.ascontiguousarray(np.asarray(xyz, np.float32).reshape(-1))
        _check(self.L.pfb_upload_landmarks(self.h, f, _fp(a), a.size // 3), "pfb_upload_landmarks")

    def reinit(self, poses7, pos_std, rot_std):
        p = np.ascontiguousarray(poses7, np.float32)
        ps = np.ascontiguousarray(pos_std, np.float32)
        rs = np.ascontiguousarray(rot_std, np.float32)
        _check(self.L.pfb_reinit(self.h, _fp(p), _fp(ps), _fp(rs)), "pfb_reinit")

    def make_inputs(self, per_filter):
        """per_filter: list of (t3, q4, tstd3, rstd3, obs ndarray OBS_DTYPE or None).

        Returns (array, keep): keep holds the contiguous observation copies the array points
        at; the array also holds them itself, so either one keeps the pointers valid."""
        arr = (_lib.PfbCycleIn * self.nf)()
        keep = []
        for f, (t, q, ts, rs, obs) in enumerate(per_filter):
            arr[f].t[:] = [float(v) for v in t]
            arr[f].q[:] = [float(v) for v in q]
            arr[f].tstd[:] = [float(v) for v in ts]
            arr[f].rstd[:] = [float(v) for v in rs]
            if obs is not None and len(obs):
                o = np.ascontiguousarray(obs, dtype=OBS_DTYPE)
                keep.append(o)
                arr[f].obs = o.ctypes.data_as(C.POINTER(_lib.PfcObs))
                arr[f].nobs = len(o)
            else:
                arr[f].obs = None
                arr[f].nobs = 0
        arr._keep = keep  # the array's obs pointers stay valid as long as the array lives
        return arr, keep

    def cycle(self, inputs) -> C.Array:
        out = (_lib.PfbEstimate * self.nf)()
        _check(self.L.pfb_cycle(self.h, inputs, out), "pfb_cycle")
        return out

    def stage_cycles(self, cycles):
        """Uploads the inputs of several cycles (each a make_inputs() array) to the device."""
        k = len(cycles)
        arr = (_lib.PfbCycleIn * (k * self.nf))()
        for c, inp in enumerate(cycles):
            for f in range(self.nf):
                arr[c * self.nf + f] = inp[f]
        _check(self.L.pfb_stage_cycles(self.h, k, arr), "pfb_stage_cycles")
        self._staged = k

    def run_staged(self):
        """Runs the staged cycles back to back; returns their estimates [cycle][filter]."""
        k = getattr(self, "_staged", 0)
        out = (_lib.PfbEstimate * (k * self.nf))()
        _check(self.L.pfb_run_staged(self.h, out), "pfb_run_staged")
        self._staged = 0
        return [out[c * self.nf:(c + 1) * self.nf] for c in range(k)]

    def staged_device_ms(self) -> float:
        """Device time of the last run_staged (first to last kernel, HIP events)."""
        ms = C.c_float(0.0)
        _check(self.L.pfb_staged_device_ms(self.h, C.byref(ms)), "pfb_staged_device_ms")
        return float(ms.value)

    def use_graph(self, on: bool = True):
        _check(self.L.pfb_use_graph(self.h, 1 if on else 0), "pfb_use_graph")

    def set_weight_ppt(self, ppt: int):
        _check(self.L.pfb_set_weight_ppt(self.h, ppt), "pfb_set_weight_ppt")

    def autotune(self, inputs) -> int:
        r = self.L.pfb_autotune(self.h, inputs)
        if r < 0:
            raise RuntimeError("pfb_autotune failed")
        return r

    def get_state(self, f: int):
        n = self.ns[f]
        pos = np.zeros((n, 3), np.float32)
        quat = np.zeros((n, 4), np.float32)
        w = np.zeros(n, np.float32)
        _check(self.L.pfb_get_state(self.h, f, _fp(pos), _fp(quat), _fp(w)), "pfb_get_state")
        return pos, quat, w

    def ancestors(self, f: int):
        a = np.zeros(self.ns[f], np.int32)
        _check(self.L.pfb_get_ancestors(self.h, f, _ip(a)), "pfb_get_ancestors")
        return a

    def cum(self, f: int):
        """Fixed-point inclusive prefix of filter f's weights from the last resampling scan."""
        a = np.zeros(self.ns[f], np.uint64)
        _check(self.L.pfb_get_cum(self.h, f, a.ctypes.data_as(C.POINTER(C.c_uint64))), "pfb_get_cum")
        return a

    def profile(self, enable: bool = True):
        _check(self.L.pfb_profile_enable(self.h, 1 if enable else 0), "pfb_profile_enable")

    def profile_read(self):
        ms = np.zeros(len(STAGES))
        cnt = np.zeros(len(STAGES), np.int32)
        _check(self.L.pfb_profile_read(self.h, _dp(ms), _ip(cnt)), "pfb_profile_read")
        return {s: (float(ms[i]), int(cnt[i])) for i, s in enumerate(STAGES)}

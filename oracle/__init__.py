"""ctypes wrapper of the CPU oracle (pf_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, never by the product package.  See pf_oracle.c for what each function
restates from the reference (rus1ru/cuda_pf_localization).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

P = C.POINTER
vp = C.c_void_p
f64 = C.c_double
f32 = C.c_float
u64 = C.c_uint64
i32 = C.c_int32

OBS_DTYPE = np.dtype([("id", "<i4"), ("mode", "<i4"), ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"),
                      ("range", "<f4")])

_SIGS = [
    ("or_det_sincos", None, [f64, P(f64), P(f64)]),
    ("or_det_cos_sinc", None, [f64, P(f64), P(f64)]),
    ("or_particle_normals", None, [u64, u64, P(f64)]),
    ("or_mot_neg2ln", C.c_float, [C.c_uint32]),
    ("or_mot_sqrt", C.c_float, [C.c_float]),
    ("or_mot_sincos_turn", None, [C.c_uint32, P(C.c_float), P(C.c_float)]),
    ("or_create", vp, [C.c_int, u64, C.c_int]),
    ("or_destroy", None, [vp]),
    ("or_set_rng", None, [vp, u64, u64, u64]),
    ("or_get_rng", None, [vp, P(u64)]),
    ("or_set_state", None, [vp, P(f64), P(f64), P(f64)]),
    ("or_get_state", None, [vp, P(f64), P(f64), P(f64), P(f64)]),
    ("or_get_ancestors", None, [vp, P(i32)]),
    ("or_reinit", None, [vp, P(f64), P(f64), P(f64)]),
    ("or_predict", None, [vp, P(f64), P(f64), P(f64), P(f64)]),
    ("or_weight", C.c_int, [vp, vp, C.c_int, P(f32), C.c_int, f64, f64, f64]),
    ("or_resample", C.c_int, [vp, f32]),
    ("or_resample_f64_literal", None, [P(f64), C.c_int, f64, P(i32)]),
    ("or_systematic_int", None, [P(f64), C.c_int, C.c_uint32, P(i32)]),
    ("or_est_sums", None, [vp, P(f64)]),
    ("or_covs", None, [vp, P(f64), P(f64), P(f64)]),
    ("or_quat_mean", None, [P(f64), P(f64)]),
    ("or_estimate", None, [vp, P(f64)]),
    ("or_num_threads", C.c_int, []),
    ("or_set_threads", None, [C.c_int]),
    ("or_cast_ray64", f64, [C.c_int, C.c_int, f64, f64, f64, P(C.c_uint8), f64, f64, f64, f64, f64]),
    ("or_map_scan64", None, [C.c_int, C.c_int, f64, f64, f64, P(C.c_uint8), f64, f64, f64, C.c_int, f64, f64,
                             P(f64)]),
    ("or_disperse64", C.c_int, [C.c_int, C.c_int, f64, f64, f64, P(C.c_uint8), u64, C.c_int, P(f64)]),
    ("or_caer64_batch", None, [C.c_int, C.c_int, f64, f64, f64, P(C.c_uint8), P(f64), C.c_int, f64, f64, P(f64),
                               C.c_int, P(f64)]),
    ("or_dda64_steps", C.c_ulonglong, [C.c_int, C.c_int, f64, f64, f64, P(C.c_uint8), P(f64), C.c_int, C.c_int, f64,
                                       f64]),
    ("or_gen_hyp32", None, [P(i32), C.c_int, f32, f32, f32, u64, C.c_int, P(f32)]),
    ("or_cast_ray32", f32, [C.c_int, C.c_int, f32, f32, f32, P(C.c_uint8), f32, f32, f32, f32, f32]),
    ("or_caer32_batch", None, [C.c_int, C.c_int, f32, f32, f32, P(C.c_uint8), P(f32), C.c_int, f32, f32, P(f32),
                               C.c_int, P(f32)]),
    ("or_render32", None, [C.c_int, C.c_int, f32, f32, f32, P(C.c_uint8), C.c_int, f32, f32, P(f32), C.c_int,
                           P(f32)]),
    # learned proposal network (mlp_oracle.c)
    ("om_param_count", C.c_long, [C.c_int, C.c_int, C.c_int, C.c_int]),
    ("om_forward", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P(f32), P(f32), C.c_int, C.c_int, P(f64)]),
    ("om_head", f64, [P(f64), P(f64), C.c_int, C.c_int, P(f64)]),
    ("om_loss_grad", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, P(f32), P(f32), P(f32), C.c_int, C.c_int,
                               P(f64), P(f64)]),
]

_L = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load(path: str) -> C.CDLL:
    """Loads an oracle library built from pf_oracle.c + mlp_oracle.c and binds its signatures
    (the mutation tests load deliberately broken builds of the same sources this way)."""
    L = C.CDLL(path)
    for name, res, args in _SIGS:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def lib() -> C.CDLL:
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        _L = load(LIB_PATH)
    return _L


def _d(a):
    return a.ctypes.data_as(P(f64))


def _f(a):
    return a.ctypes.data_as(P(f32))


def _u8(a):
    return a.ctypes.data_as(P(C.c_uint8))


def _i(a):
    return a.ctypes.data_as(P(i32))


def set_threads(t: int) -> None:
    lib().or_set_threads(t)


def num_threads() -> int:
    return lib().or_num_threads()


class OracleFilter:
    """cpu.rs restated in C (f64); f32_storage rounds poses/weights to f32 after every stage."""

    def __init__(self, n: int, seed: int, f32_storage: bool = True, library: C.CDLL | None = None):
        self.L = library if library is not None else lib()
        self.n = n
        self.h = self.L.or_create(n, seed, 1 if f32_storage else 0)

    def __del__(self):
        try:
            if self.h:
                self.L.or_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def set_rng(self, gseed: int, gen: int, lcg: int):
        self.L.or_set_rng(self.h, gseed, gen, lcg)

    def rng(self):
        o = (u64 * 3)()
        self.L.or_get_rng(self.h, o)
        return tuple(int(v) for v in o)

    def set_state(self, pos, quat, w=None):
        p = np.ascontiguousarray(pos, np.float64)
        q = np.ascontiguousarray(quat, np.float64)
        wp = None
        if w is not None:
            wa = np.ascontiguousarray(w, np.float64)
            wp = _d(wa)
        self.L.or_set_state(self.h, _d(p), _d(q), wp)

    def state(self):
        pos = np.zeros((self.n, 3))
        quat = np.zeros((self.n, 4))
        w = np.zeros(self.n)
        ll = np.zeros(self.n)
        self.L.or_get_state(self.h, _d(pos), _d(quat), _d(w), _d(ll))
        return pos, quat, w, ll

    def ancestors(self):
        a = np.zeros(self.n, np.int32)
        self.L.or_get_ancestors(self.h, _i(a))
        return a

    def reinit(self, pose7, pos_std, rot_std):
        self.L.or_reinit(self.h, _d(np.asarray(pose7, np.float64)), _d(np.asarray(pos_std, np.float64)),
                         _d(np.asarray(rot_std, np.float64)))

    def predict(self, t, q, tstd, rstd):
        self.L.or_predict(self.h, _d(np.asarray(t, np.float64)), _d(np.asarray(q, np.float64)),
                          _d(np.asarray(tstd, np.float64)), _d(np.asarray(rstd, np.float64)))

    def weight(self, obs, lm_xyz, sigma_r, sigma_b, gate):
        o = np.ascontiguousarray(obs if obs is not None else np.zeros(0, OBS_DTYPE), dtype=OBS_DTYPE)
        lm = np.ascontiguousarray(np.asarray(lm_xyz, np.float32).reshape(-1))
        return self.L.or_weight(self.h, o.ctypes.data if len(o) else None, len(o), _f(lm), lm.size // 3,
                                sigma_r, sigma_b, gate)

    def resample(self, ratio: float):
        return self.L.or_resample(self.h, ratio)

    def est_sums(self):
        o = np.zeros(14)
        self.L.or_est_sums(self.h, _d(o))
        return o

    def covs(self, mp, mq):
        o = np.zeros(12)
        self.L.or_covs(self.h, _d(np.asarray(mp, np.float64)), _d(np.asarray(mq, np.float64)), _d(o))
        return o

    def estimate(self):
        """cpu.rs:308-368 -> dict(position, attitude, pos_cov, att_cov, ess)."""
        e = np.zeros(26)
        self.L.or_estimate(self.h, _d(e))
        return {"position": e[0:3], "attitude": e[3:7], "pos_cov": e[7:16].reshape(3, 3),
                "att_cov": e[16:25].reshape(3, 3), "ess": float(e[25])}


def particle_normals(gseed: int, i: int, library: C.CDLL | None = None) -> np.ndarray:
    o = np.zeros(6)
    (library if library is not None else lib()).or_particle_normals(gseed, i, _d(o))
    return o


def det_sincos(x: float):
    s, c = f64(), f64()
    lib().or_det_sincos(x, C.byref(s), C.byref(c))
    return s.value, c.value


def det_cos_sinc(h2: float):
    c, s = f64(), f64()
    lib().or_det_cos_sinc(h2, C.byref(c), C.byref(s))
    return c.value, s.value


def quat_mean(M) -> np.ndarray:
    m = np.ascontiguousarray(M, np.float64).reshape(16)
    q = np.zeros(4)
    lib().or_quat_mean(_d(m), _d(q))
    return q


def systematic_int(w, u32: int) -> np.ndarray:
    wa = np.ascontiguousarray(w, np.float64)
    a = np.zeros(len(wa), np.int32)
    lib().or_systematic_int(_d(wa), len(wa), u32, _i(a))
    return a


def systematic_f64_literal(w, ufrac: float) -> np.ndarray:
    wa = np.ascontiguousarray(w, np.float64)
    a = np.zeros(len(wa), np.int32)
    lib().or_resample_f64_literal(_d(wa), len(wa), ufrac, _i(a))
    return a


# ---- global localisation

class OracleGrid:
    def __init__(self, cells, res, origin):
        self.cells = np.ascontiguousarray(cells, np.uint8)
        self.H, self.W = self.cells.shape
        self.res = float(res)
        self.ox, self.oy = float(origin[0]), float(origin[1])

    def cast_ray64(self, px, py, dx, dy, rmax):
        return lib().or_cast_ray64(self.W, self.H, self.res, self.ox, self.oy, _u8(self.cells), px, py, dx, dy, rmax)

    def map_scan64(self, x, y, th, nrays, span, rmax):
        out = np.zeros(nrays)
        lib().or_map_scan64(self.W, self.H, self.res, self.ox, self.oy, _u8(self.cells), x, y, th, nrays, span, rmax,
                            _d(out))
        return out

    def disperse64(self, seed, n):
        hyp = np.zeros((n, 3))
        rc = lib().or_disperse64(self.W, self.H, self.res, self.ox, self.oy, _u8(self.cells), seed, n, _d(hyp))
        if rc != 0:
            raise RuntimeError("no free cells")
        return hyp

    def caer64(self, scan, hyp, span, rmax):
        s = np.ascontiguousarray(scan, np.float64)
        h = np.ascontiguousarray(hyp, np.float64).reshape(-1, 3)
        out = np.zeros(len(h))
        lib().or_caer64_batch(self.W, self.H, self.res, self.ox, self.oy, _u8(self.cells), _d(s), s.size, span, rmax,
                              _d(h), len(h), _d(out))
        return out

    def dda_steps64(self, hyp, nrays, span, rmax):
        """Total DDA cell steps of the f64 twin over the hypotheses' rays (algorithmic work)."""
        h = np.ascontiguousarray(hyp, np.float64).reshape(-1, 3)
        return int(lib().or_dda64_steps(self.W, self.H, self.res, self.ox, self.oy, _u8(self.cells), _d(h), len(h),
                                        nrays, span, rmax))

    def localize64(self, scan, hypotheses, topk, seed, span=2 * np.pi, rmax=10.0):
        """CpuGlobalLoc::localize (global_loc.rs:384-398): (x, y, th, caer) rows ascending."""
        hyp = self.disperse64(seed, hypotheses)
        caer = self.caer64(scan, hyp, span, rmax)
        order = np.argsort(caer, kind="stable")[:topk]
        return np.column_stack([hyp[order], caer[order]])

    def free_list(self):
        flat = self.cells.reshape(-1)
        idx = np.nonzero(flat == 0)[0]
        return np.ascontiguousarray(np.column_stack([idx % self.W, idx // self.W]).astype(np.int32))

    def gen_hyp32(self, seed, n):
        fl = self.free_list()
        out = np.zeros((n, 3), np.float32)
        lib().or_gen_hyp32(_i(fl), len(fl), self.res, self.ox, self.oy, seed, n, _f(out))
        return out

    def cast_ray32(self, px, py, dx, dy, rmax):
        return lib().or_cast_ray32(self.W, self.H, self.res, self.ox, self.oy, _u8(self.cells), px, py, dx, dy, rmax)

    def caer32(self, scan, hyp, span, rmax):
        s = np.ascontiguousarray(scan, np.float32)
        h = np.ascontiguousarray(hyp, np.float32).reshape(-1, 3)
        out = np.zeros(len(h), np.float32)
        lib().or_caer32_batch(self.W, self.H, self.res, self.ox, self.oy, _u8(self.cells), _f(s), s.size, span, rmax,
                              _f(h), len(h), _f(out))
        return out

    def render32(self, poses, nrays, span, rmax):
        p = np.ascontiguousarray(poses, np.float32).reshape(-1, 3)
        out = np.zeros((len(p), nrays), np.float32)
        lib().or_render32(self.W, self.H, self.res, self.ox, self.oy, _u8(self.cells), nrays, span, rmax, _f(p),
                          len(p), _f(out))
        return out


def bottom_k(caer, k):
    """Stable ascending order by (caer, index), first k indices."""
    return np.lexsort((np.arange(len(caer)), caer))[:k]


# ------------------------------------------------------------------ learned proposal network


class OracleMlp:
    """f64 restatement of the pfl_* network (mlp_oracle.c); `bf16=True` rounds the operands
    the GPU rounds to bfloat16."""

    def __init__(self, in_dim, hid, n_hidden, ncomp):
        self.dims = (in_dim, hid, n_hidden, ncomp)
        self.nout = 9 * ncomp

    def param_count(self) -> int:
        return lib().om_param_count(*self.dims)

    def forward(self, params, x, bf16=True):
        x = np.ascontiguousarray(x, np.float32)
        B = x.shape[0]
        out = np.zeros((B, self.nout))
        p = np.ascontiguousarray(params, np.float32)
        if lib().om_forward(*self.dims, _f(p), _f(x), B, int(bf16), _d(out)) != 0:
            raise ValueError("om_forward")
        return out

    def loss_grad(self, params, x, y, bf16=True):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        p = np.ascontiguousarray(params, np.float32)
        g = np.zeros(p.size)
        loss = C.c_double()
        if lib().om_loss_grad(*self.dims, _f(p), _f(x), _f(y), x.shape[0], int(bf16), _d(g), C.byref(loss)) != 0:
            raise ValueError("om_loss_grad")
        return loss.value, g


def mixture_nll(out, y, ncomp, with_grad=False):
    """Mean mixture NLL of raw output rows (mlp_oracle.c om_head) and optionally d/d(out)."""
    o = np.ascontiguousarray(out, np.float64)
    yy = np.ascontiguousarray(y, np.float64)
    d = np.zeros_like(o) if with_grad else None
    v = lib().om_head(_d(o), _d(yy), o.shape[0], ncomp, _d(d) if with_grad else None)
    return (v, d) if with_grad else v


def adam_step(p, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """Textbook Adam (Kingma & Ba 2015, Algorithm 1) in f64; returns (p, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v

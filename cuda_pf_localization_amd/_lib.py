"""ctypes bindings of the in-tree gfx950 libraries (libpf_kernels_hip.so, libpf_host.so).

There is no CPU fallback: if the libraries are missing, loading raises, naming the build
command.  Every signature mirrors include/pf_kernels.h, include/pf_gl.h and include/pf_host.h.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")
KERNELS_PATH = os.path.join(LIBDIR, "libpf_kernels_hip.so")
HOST_PATH = os.path.join(LIBDIR, "libpf_host.so")

vp = C.c_void_p
i32 = C.c_int32
u32 = C.c_uint32
u64 = C.c_uint64
f32 = C.c_float
f64 = C.c_double
P = C.POINTER


class PfcObs(C.Structure):
    """pfc_obs: the reference's 24-byte ObsF / CObservation record."""
    _fields_ = [("id", i32), ("mode", i32), ("dx", f32), ("dy", f32), ("dz", f32), ("range", f32)]


class PfhConfig(C.Structure):
    _fields_ = [("particle_count", C.c_int64), ("range_noise", f64), ("bearing_noise", f64),
                ("resample_ess_ratio", f64), ("random_inject_ratio", f64), ("gate_sigma", f64),
                ("seed", u64), ("obs_cap", i32), ("fast", i32)]


class PfhObs(C.Structure):
    _fields_ = [("id", i32), ("mode", i32), ("vx", f64), ("vy", f64), ("vz", f64), ("range", f64)]


class PfhOdom(C.Structure):
    _fields_ = [("t", f64 * 3), ("q", f64 * 4), ("trans_noise", f64 * 3), ("rot_noise", f64 * 3)]


class PfhEstimate(C.Structure):
    _fields_ = [("pos", f64 * 3), ("quat", f64 * 4), ("pos_cov", f64 * 9), ("att_cov", f64 * 9),
                ("ess", f64), ("valid", i32), ("resampled", i32)]


class PfhGlParams(C.Structure):
    _fields_ = [("hypotheses", i32), ("rays", i32), ("rmax", f64), ("ang_span", f64), ("topk", i32),
                ("seed", u64)]


class PfhLearnedParams(C.Structure):
    _fields_ = [("n_mix", C.c_int32), ("n_uniform", C.c_int32), ("fallback_hypotheses", C.c_int32),
                ("fallback_caer_per_ray", C.c_double), ("seed", C.c_uint64), ("refine_rounds", C.c_int32),
                ("refine_parents", C.c_int32), ("refine_children", C.c_int32), ("refine_sigma_xy", C.c_double),
                ("refine_sigma_yaw", C.c_double), ("refine_shrink", C.c_double)]


class PfhLearnedInfo(C.Structure):
    _fields_ = [("used_fallback", C.c_int32), ("proposal_best_caer_per_ray", C.c_double),
                ("best_caer_per_ray", C.c_double)]


class PfbFilterCfg(C.Structure):
    _fields_ = [("n", i32), ("seed", u32), ("sigma_r", f32), ("sigma_b", f32), ("gate", f32),
                ("resample_ess_ratio", f32), ("inject_ratio", f32)]


class PfbCycleIn(C.Structure):
    _fields_ = [("t", f32 * 3), ("q", f32 * 4), ("tstd", f32 * 3), ("rstd", f32 * 3), ("obs", C.POINTER(PfcObs)),
                ("nobs", i32)]


class PfbEstimate(C.Structure):
    _fields_ = [("pos", f64 * 3), ("quat", f64 * 4), ("pos_cov", f64 * 9), ("att_cov", f64 * 9), ("ess", f64),
                ("resampled", i32), ("valid", i32)]


# (name, restype, argtypes)
KERNEL_SYMBOLS = [
    # reference ABI (pf_kernels.cu:475-1025)
    ("pfc_device_available", C.c_int, []),
    ("pfc_create", vp, [C.c_int]),
    ("pfc_destroy", C.c_int, [vp]),
    ("pfc_init", C.c_int, [vp, C.c_int, C.c_uint]),
    ("pfc_upload_landmarks", C.c_int, [vp, P(f32), C.c_int]),
    ("pfc_reinit", C.c_int, [vp, f32, f32, f32, f32, f32, f32, f32, P(f32), P(f32)]),
    ("pfc_predict", C.c_int, [vp, f32, f32, f32, f32, f32, f32, f32, P(f32), P(f32)]),
    ("pfc_weight", C.c_int, [vp, vp, C.c_int, f32, f32, f32]),
    ("pfc_resample", C.c_int, [vp, f32]),
    ("pfc_est_sums", C.c_int, [vp, P(f64)]),
    ("pfc_covs", C.c_int, [vp, f32, f32, f32, f32, f32, f32, f32, P(f64)]),
    ("pfc_snapshot", C.c_int, [vp, P(f32), P(f32)]),
    ("pfgl_create", vp, [C.c_int, C.c_int, f32, f32, f32, P(C.c_uint8), C.c_int]),
    ("pfgl_destroy", C.c_int, [vp]),
    ("pfgl_free_cells", C.c_int, [vp]),
    ("pfgl_generate", C.c_int, [vp, C.c_int, C.c_ulonglong]),
    ("pfgl_score", C.c_int, [vp, P(f32), C.c_int, f32, f32]),
    ("pfgl_topk", C.c_int, [vp, C.c_int, P(f32)]),
    # particle-filter extensions
    ("pfx_set_state", C.c_int, [vp, P(f32), P(f32), P(f32)]),
    ("pfx_get_state", C.c_int, [vp, P(f32), P(f32), P(f32)]),
    ("pfx_get_ll", C.c_int, [vp, P(f64)]),
    ("pfx_get_ancestors", C.c_int, [vp, P(i32)]),
    ("pfx_get_cum", C.c_int, [vp, P(u64)]),
    ("pfx_set_lcg", C.c_int, [vp, u64]),
    ("pfx_get_lcg", C.c_int, [vp, P(u64)]),
    ("pfx_get_rng", C.c_int, [vp, P(u64), P(u64)]),
    ("pfx_set_rng", C.c_int, [vp, u64, u64]),
    ("pfx_weight_scalars", C.c_int, [vp, P(f64)]),
    ("pfx_profile_enable", C.c_int, [vp, C.c_int]),
    ("pfx_profile_read", C.c_int, [vp, P(f64), P(C.c_int)]),
    ("pfx_particle_count", C.c_int, [vp]),
    ("pfx_sync", C.c_int, [vp]),
    ("pfx_set_weight_variant", C.c_int, [vp, C.c_int]),
    ("pfx_set_scan_spin_limit", C.c_int, [vp, C.c_int]),
    ("pfx_quat_mean", C.c_int, [P(f64), P(f64)]),
    # global-localisation extensions
    ("pfglx_topk_ex", C.c_int, [vp, C.c_int, P(f32), P(f32), P(i32)]),
    ("pfglx_set_hypotheses", C.c_int, [vp, P(f32), C.c_int]),
    ("pfglx_get_hypotheses", C.c_int, [vp, P(f32)]),
    ("pfglx_get_caer", C.c_int, [vp, P(f32)]),
    ("pfglx_hypothesis_count", C.c_int, [vp]),
    ("pfglx_render", C.c_int, [vp, P(f32), C.c_int, C.c_int, f32, f32, P(f32)]),
    ("pfglx_render_device", C.c_int, [vp, vp, C.c_int, C.c_int, f32, f32, vp, vp]),
    ("pfglx_set_score_mode", C.c_int, [vp, C.c_int]),
    ("pfglx_set_prune", C.c_int, [vp, C.c_int]),
    ("pfglx_prune_stats", C.c_int, [vp, P(C.c_int)]),
    ("pfglx_profile_enable", C.c_int, [vp, C.c_int]),
    ("pfglx_profile_read", C.c_int, [vp, P(f64), P(C.c_int)]),
    ("pfglx_sync", C.c_int, [vp]),
    # batched filters
    ("pfb_create", vp, [C.c_int, P(PfbFilterCfg), C.c_int]),
    ("pfb_destroy", C.c_int, [vp]),
    ("pfb_filter_count", C.c_int, [vp]),
    ("pfb_upload_landmarks", C.c_int, [vp, C.c_int, P(f32), C.c_int]),
    ("pfb_reinit", C.c_int, [vp, P(f32), P(f32), P(f32)]),
    ("pfb_cycle", C.c_int, [vp, P(PfbCycleIn), P(PfbEstimate)]),
    ("pfb_stage_cycles", C.c_int, [vp, C.c_int, P(PfbCycleIn)]),
    ("pfb_run_staged", C.c_int, [vp, P(PfbEstimate)]),
    ("pfb_staged_device_ms", C.c_int, [vp, P(f32)]),
    ("pfb_use_graph", C.c_int, [vp, C.c_int]),
    ("pfb_autotune", C.c_int, [vp, P(PfbCycleIn)]),
    ("pfb_set_weight_ppt", C.c_int, [vp, C.c_int]),
    ("pfb_get_state", C.c_int, [vp, C.c_int, P(f32), P(f32), P(f32)]),
    ("pfb_get_ancestors", C.c_int, [vp, C.c_int, P(i32)]),
    ("pfb_get_cum", C.c_int, [vp, C.c_int, P(u64)]),
    ("pfx_est_sums_noalloc", C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, vp, vp]),
    ("pfb_profile_enable", C.c_int, [vp, C.c_int]),
    ("pfb_profile_read", C.c_int, [vp, P(f64), P(C.c_int)]),
    # learned proposal
    ("pfl_create", vp, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, u64]),
    ("pfl_destroy", C.c_int, [vp]),
    ("pfl_param_count", C.c_int64, [vp]),
    ("pfl_get_params", C.c_int, [vp, P(f32)]),
    ("pfl_set_params", C.c_int, [vp, P(f32)]),
    ("pfl_forward", C.c_int, [vp, P(f32), C.c_int, P(f32)]),
    ("pfl_grad", C.c_int, [vp, P(f32), P(f32), C.c_int, P(f32), P(f64)]),
    ("pfl_train_step", C.c_int, [vp, P(f32), P(f32), C.c_int, f32, P(f64)]),
    ("pfl_train_on_map", C.c_int, [vp, vp, C.c_int64, C.c_int, f32, C.c_int, f32, f32, f32, u64, P(f64)]),
    ("pfl_propose", C.c_int, [vp, vp, P(f32), C.c_int, f32, C.c_int, C.c_int, u64]),
    ("pfl_last_propose_ms", f64, [vp]),
    ("pfglx_sample_poses_device", C.c_int, [vp, C.c_int, u64, vp, vp]),
    ("pfglx_generate_range", C.c_int, [vp, C.c_int, C.c_int, u64]),
    ("pfglx_hypotheses_device", vp, [vp]),
    ("pfglx_map_info", C.c_int, [vp, P(f64)]),
    ("pfglx_set_count", C.c_int, [vp, C.c_int]),
    ("pfglx_stream", vp, [vp]),
    ("pfglx_topk_device", C.c_int, [vp, C.c_int]),
    ("pfglx_set_scan", C.c_int, [vp, P(f32), C.c_int, f32]),
    ("pfglx_score_async", C.c_int, [vp, C.c_int, f32]),
    ("pfglx_refine", C.c_int, [vp, C.c_int, C.c_int, f32, f32, u64]),
]

HOST_SYMBOLS = [
    ("pfh_filter_new", vp, [P(PfhConfig)]),
    ("pfh_filter_free", None, [vp]),
    ("pfh_filter_set_map", C.c_int, [vp, P(i32), P(f64), C.c_int]),
    ("pfh_filter_reinitialize", C.c_int, [vp, P(f64), P(f64), P(f64)]),
    ("pfh_filter_update", C.c_int, [vp, P(PfhOdom), P(PfhObs), C.c_int, P(PfhEstimate)]),
    ("pfh_filter_estimate", C.c_int, [vp, P(PfhEstimate)]),
    ("pfh_filter_snapshot", C.c_int, [vp, P(f64), P(f64), P(f64)]),
    ("pfh_filter_kernels", vp, [vp]),
    ("pfh_quat_mean", C.c_int, [P(f64), P(f64)]),
    ("pfh_gl_new", vp, [C.c_int, C.c_int, f64, f64, f64, P(C.c_uint8), P(PfhGlParams)]),
    ("pfh_gl_free", None, [vp]),
    ("pfh_gl_localize", C.c_int, [vp, P(f32), C.c_int, P(f64)]),
    ("pfh_gl_set_seed", C.c_int, [vp, u64]),
    ("pfh_gl_kernels", vp, [vp]),
    ("pfh_gl_localize_learned", C.c_int, [vp, vp, P(f32), C.c_int, vp, P(f64), vp]),
    ("pfh_map_from_pgm_yaml", C.c_int, [P(C.c_uint8), C.c_size_t, C.c_char_p, P(C.c_int), P(C.c_int), P(f64),
                                        P(f64), P(f64), P(C.c_uint8), C.c_size_t, C.c_char_p, C.c_int]),
    ("pfh_landmarks_from_text", C.c_int, [C.c_char_p, P(i32), P(f64), C.c_int, C.c_char_p, C.c_int]),
    ("pfh_filter_reinit_from_gl", C.c_int, [vp, P(f64), f64, P(f64), P(f64)]),
    ("pfh_filter_save", C.c_int, [vp, C.c_char_p]),
    ("pfh_filter_load", C.c_int, [vp, C.c_char_p]),
]

OPTIONAL_KERNEL_SYMBOLS: list = []  # filled by modules that add entry points (batched, learned)

_kern = None
_host = None


def _bind(lib, table):
    for name, res, args in table:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def kernels() -> C.CDLL:
    """The drop-in library; raises if it has not been built."""
    global _kern
    if _kern is None:
        if not os.path.exists(KERNELS_PATH):
            raise RuntimeError(f"{KERNELS_PATH} is missing: run `python -m cuda_pf_localization_amd.build` "
                               "(there is no CPU fallback)")
        _kern = _bind(C.CDLL(KERNELS_PATH, mode=C.RTLD_GLOBAL), KERNEL_SYMBOLS + OPTIONAL_KERNEL_SYMBOLS)
    return _kern


def host() -> C.CDLL:
    global _host
    if _host is None:
        kernels()
        if not os.path.exists(HOST_PATH):
            raise RuntimeError(f"{HOST_PATH} is missing: run `python -m cuda_pf_localization_amd.build`")
        _host = _bind(C.CDLL(HOST_PATH), HOST_SYMBOLS)
    return _host


def declared_symbols() -> list[str]:
    """Every function name declared in include/*.h (parsed from the headers)."""
    import re
    root = os.path.dirname(HERE)
    names = []
    for h in ("pf_kernels.h", "pf_gl.h", "pf_host.h", "pf_batch_api.h", "pf_learned.h"):
        path = os.path.join(root, "include", h)
        if not os.path.exists(path):
            continue
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(pf[a-z]*_[A-Za-z0-9_]+)\s*\(", text):
            names.append(m.group(1))
    return sorted(set(names))

"""C-ABI checks that need no GPU: the libraries load, export every entry point the headers
declare, keep the reference's record layout and reject bad arguments like the reference
(pf_kernels.cu: -1 / NULL on a NULL handle or invalid sizes) without touching the device."""
import ctypes as C
import os

import numpy as np

from cuda_pf_localization_amd import OBS_DTYPE, _lib

REFERENCE_SYMBOLS = [
    "pfc_device_available", "pfc_create", "pfc_destroy", "pfc_init", "pfc_upload_landmarks", "pfc_reinit",
    "pfc_predict", "pfc_weight", "pfc_resample", "pfc_est_sums", "pfc_covs", "pfc_snapshot",
    "pfgl_create", "pfgl_destroy", "pfgl_free_cells", "pfgl_generate", "pfgl_score", "pfgl_topk",
]


def _exported(path):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


def test_libraries_load_and_export_declared_symbols():
    L = _lib.kernels()
    H = _lib.host()
    assert L is not None and H is not None
    declared = _lib.declared_symbols()
    exported = _exported(_lib.KERNELS_PATH) | _exported(_lib.HOST_PATH)
    missing = [s for s in declared if s not in exported]
    assert not missing, f"declared but not exported: {missing}"
    for s in REFERENCE_SYMBOLS:
        assert s in declared and s in _exported(_lib.KERNELS_PATH)


def test_reference_link_name_alias():
    """crates/pf_core/build.rs:15 links dylib=pf_kernels: libpf_kernels.so must resolve to the HIP library."""
    alias = os.path.join(_lib.LIBDIR, "libpf_kernels.so")
    assert os.path.realpath(alias) == os.path.realpath(_lib.KERNELS_PATH)
    C.CDLL(alias)


def test_observation_record_layout():
    """ObsF (pf_kernels.cu:79-84) / CObservation (cuda.rs:18-27): 24 bytes, id,mode,dx,dy,dz,range."""
    assert OBS_DTYPE.itemsize == 24
    assert [OBS_DTYPE.fields[f][1] for f in ("id", "mode", "dx", "dy", "dz", "range")] == [0, 4, 8, 12, 16, 20]
    assert C.sizeof(_lib.PfcObs) == 24


def test_null_handle_and_bad_argument_errors():
    L = _lib.kernels()
    assert L.pfc_device_available() in (0, 1)
    assert L.pfc_destroy(None) == 0
    assert L.pfc_init(None, 10, 1) == -1
    assert L.pfc_upload_landmarks(None, None, 0) == -1
    z3 = (C.c_float * 3)()
    assert L.pfc_reinit(None, 0, 0, 0, 1, 0, 0, 0, z3, z3) == -1
    assert L.pfc_predict(None, 0, 0, 0, 1, 0, 0, 0, z3, z3) == -1
    assert L.pfc_weight(None, None, 0, 0.05, 0.02, 0.0) == -1
    assert L.pfc_resample(None, 0.0) == -1
    out = (C.c_double * 14)()
    assert L.pfc_est_sums(None, out) == -1
    assert L.pfc_covs(None, 0, 0, 0, 1, 0, 0, 0, out) == -1
    assert L.pfc_snapshot(None, z3, z3) == -1
    assert L.pfgl_destroy(None) == 0
    assert L.pfgl_free_cells(None) == -1
    assert L.pfgl_generate(None, 10, 1) == -1
    assert L.pfgl_score(None, z3, 3, 1.0, 1.0) == -1
    assert L.pfgl_topk(None, 3, z3) == -1
    grid = (C.c_uint8 * 4)()
    assert not L.pfgl_create(0, 2, 0.05, 0, 0, grid, 10)
    assert not L.pfgl_create(2, 2, 0.05, 0, 0, grid, 0)
    assert not L.pfgl_create(2, 2, 0.05, 0, 0, None, 10)
    assert L.pfx_set_state(None, None, None, None) == -1
    assert L.pfx_get_ll(None, None) == -1
    assert L.pfx_particle_count(None) == 0
    assert L.pfglx_topk_ex(None, 1, None, None, None) == -1
    assert L.pfglx_set_hypotheses(None, None, 1) == -1


def test_host_layer_quat_mean_vs_eigh():
    """types.rs:128-144 restated in the C++ host layer (Jacobi) against numpy's eigh."""
    from cuda_pf_localization_amd import quat_mean
    rng = np.random.default_rng(0)
    for _ in range(200):
        q = rng.normal(size=(30, 4))
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        w = rng.random(30)
        M = np.einsum("i,ij,ik->jk", w / w.sum(), q, q)
        m = quat_mean(M)
        val, vec = np.linalg.eigh(M)
        v = vec[:, -1] * (1 if vec[0, -1] >= 0 else -1)
        assert np.allclose(m, v, atol=1e-9)
        assert m[0] >= 0
    # identity moment of a single quaternion and of its negation (sign flip)
    qq = np.array([0.2, -0.4, 0.1, 0.8])
    qq /= np.linalg.norm(qq)
    for s in (1, -1):
        m = quat_mean(np.outer(s * qq, s * qq))
        assert np.allclose(m, qq * np.sign(qq[0]), atol=1e-12)


def test_no_gpu_means_loud_failure():
    """Without a device the host layer refuses to build a filter (no CPU fallback)."""
    L = _lib.kernels()
    if L.pfc_device_available() == 1:
        return
    from cuda_pf_localization_amd import ParticleFilter, PfConfig
    try:
        ParticleFilter(PfConfig(particle_count=16))
    except RuntimeError as e:
        assert "GPU" in str(e)
    else:
        raise AssertionError("ParticleFilter constructed without a GPU")

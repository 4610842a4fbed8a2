"""MI355X (gfx950) backend for the particle-filter localisation engine of
rus1ru/cuda_pf_localization.

The product is the C-ABI library lib/libpf_kernels_hip.so (pfc_* / pfgl_* drop-in plus
extensions) and the C++ host layer lib/libpf_host.so.  This package only binds them.
"""
from .api import (GL_STAGES, BatchedFilters, OBS_DTYPE, STAGES, Estimate, GlKernels, GlobalLoc, GlParams, Observation,  # noqa: F401
                  OdomDelta, PfConfig, PfKernels, ParticleFilter, device_available, make_obs, quat_mean,
                  LearnedParams, LearnedProposal, load_landmarks, load_map_pgm_yaml)

__all__ = ["PfKernels", "BatchedFilters", "GlKernels", "ParticleFilter", "GlobalLoc", "PfConfig", "GlParams", "Observation",
           "OdomDelta", "Estimate", "device_available", "make_obs", "quat_mean", "OBS_DTYPE", "STAGES", "GL_STAGES",
           "LearnedParams", "LearnedProposal", "load_landmarks", "load_map_pgm_yaml"]

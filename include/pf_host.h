/*
 * pf_host.h -- C API of libpf_host.so, the C++ host layer above libpf_kernels_hip.so.
 *
 * The reference's host is Rust (crates/pf_core); its toolchain is not available here, so
 * the host side that sits above the pfc_* / pfgl_* ABI is restated in C++:
 *   - HipBackend   = CudaBackend glue of crates/pf_core/src/cuda.rs:97-357
 *   - ParticleFilter::update = crates/pf_core/src/particle_filter.rs:110-124
 *   - quat_mean    = crates/pf_core/src/types.rs:128-144 (own 4x4 Jacobi)
 *   - GlobalLoc::localize = crates/pf_core/src/global_loc.rs:305-327
 * These pfh_* entry points have no counterpart in the reference; they let Python tests
 * and the benchmark drive that host layer.
 */
#ifndef PF_HOST_H
#define PF_HOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* PfConfig (config.rs:3-64). */
typedef struct pfh_config {
    int64_t particle_count;
    double range_noise, bearing_noise;
    double resample_ess_ratio, random_inject_ratio, gate_sigma;
    uint64_t seed;
    int32_t obs_cap;   /* observations per weight() call (cuda.rs:95 OBS_CAP = 256) */
    int32_t fast;      /* 1: skip the covariance pass of the pre-resample estimate */
} pfh_config;

/* Observation (types.rs:58-67): mode 0 Cartesian, 1 range-bearing. */
typedef struct pfh_obs {
    int32_t id, mode;
    double vx, vy, vz, range;
} pfh_obs;

/* OdomDelta (types.rs:69-90): translation, rotation wxyz, noise stds. */
typedef struct pfh_odom {
    double t[3], q[4], trans_noise[3], rot_noise[3];
} pfh_odom;

/* Estimate (types.rs:92-103). */
typedef struct pfh_estimate {
    double pos[3], quat[4], pos_cov[9], att_cov[9], ess;
    int32_t valid, resampled;
} pfh_estimate;

void *pfh_filter_new(const pfh_config *cfg);
void pfh_filter_free(void *f);
/* Landmark map (landmark_map.rs): n (id, xyz) entries; ids >= 0. */
int pfh_filter_set_map(void *f, const int32_t *ids, const double *xyz, int n);
int pfh_filter_reinitialize(void *f, const double pose7[7], const double pos_std[3], const double rot_std[3]);
int pfh_filter_update(void *f, const pfh_odom *odom, const pfh_obs *obs, int nobs, pfh_estimate *out);
int pfh_filter_estimate(void *f, pfh_estimate *out);
/* Particles as the Rust snapshot returns them (weights 1/n, cuda.rs:354). */
int pfh_filter_snapshot(void *f, double *pos3n, double *quat4n, double *w);
/* The underlying pfc_* handle. */
void *pfh_filter_kernels(void *f);

/* Markley mean of a symmetric 4x4 moment matrix (row-major): q = wxyz, w >= 0. */
int pfh_quat_mean(const double M[16], double q[4]);

/* Global localisation (GlParams global_loc.rs:236-257). */
typedef struct pfh_gl_params {
    int32_t hypotheses, rays;
    double rmax, ang_span;
    int32_t topk;
    uint64_t seed;
} pfh_gl_params;

void *pfh_gl_new(int W, int H, double res, double ox, double oy, const uint8_t *cells, const pfh_gl_params *p);
void pfh_gl_free(void *g);
/* Bottom-k hypotheses (x, y, theta, caer) ascending by CAER into out[4*topk]; returns k or -1. */
int pfh_gl_localize(void *g, const float *scan, int nrays, double *out4k);
int pfh_gl_set_seed(void *g, uint64_t seed);
void *pfh_gl_kernels(void *g);

#ifdef __cplusplus
}
#endif
#endif /* PF_HOST_H */

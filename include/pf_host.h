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

/* Learned-proposal localisation (no reference counterpart; pfl_* in pf_learned.h):
 * n_mix mixture + n_uniform uniform hypotheses are scored by CAER, then refine_rounds
 * rounds of local refinement (refine_children per each of the refine_parents best, jitter
 * sigma_xy / sigma_yaw shrinking by refine_shrink per round) are scored; if the best CAER
 * per ray still exceeds fallback_caer_per_ray, fallback_hypotheses uniform ones are scored
 * instead (the reference's dispersal, global_loc.rs:306-327).  The pfh_gl handle's
 * hypothesis capacity must cover every count. */
typedef struct pfh_learned_params {
    int32_t n_mix, n_uniform, fallback_hypotheses;
    double fallback_caer_per_ray;
    uint64_t seed;
    int32_t refine_rounds, refine_parents, refine_children;
    double refine_sigma_xy, refine_sigma_yaw, refine_shrink;
} pfh_learned_params;
typedef struct pfh_learned_info {
    int32_t used_fallback;
    double proposal_best_caer_per_ray; /* best CAER / nrays of the learned path (after refinement) */
    double best_caer_per_ray;          /* of the returned winner */
} pfh_learned_info;
int pfh_gl_localize_learned(void *g, void *mlp, const float *scan, int nrays, const pfh_learned_params *lp,
                            double *out4k, pfh_learned_info *info);

/* ---- on-disk formats ----
 * ROS map_server pair (global_loc.rs:32-130): P5 PGM bytes + YAML text (resolution, origin
 * [x, y, ...], occupied_thresh default 0.65).  A cell is occupied (1) when
 * (255 - pixel) / 255 > occupied_thresh; rows are flipped so that row 0 is the bottom.
 * cells may be NULL to query W and H.  Returns 0, or -1 with a message in err. */
int pfh_map_from_pgm_yaml(const uint8_t *pgm, size_t pgm_len, const char *yaml, int *W, int *H, double *res,
                          double *ox, double *oy, uint8_t *cells, size_t cells_cap, char *err, int err_len);
/* Landmark text (landmark_map.rs:59-80): lines `id: [x, y, z]` or `id x y z`, '#' starts a
 * comment.  Returns the number of entries written to ids[] / xyz[3 i ..] (at most cap), or
 * -1 with "line N: ..." in err. */
int pfh_landmarks_from_text(const char *text, int32_t *ids, double *xyz, int cap, char *err, int err_len);

/* GL -> PF hand-off (README:84-86): reinitialise the filter about (x, y, z) with the yaw of
 * the localisation result as a rotation about z. */
int pfh_filter_reinit_from_gl(void *f, const double *xyt3, double z, const double *pos_std3,
                              const double *rot_std3);
/* Checkpoint / resume: particles with their stored weights and the draw state (generation
 * seed, generation, resampling LCG); a resumed filter continues bit-identically.  Load needs
 * a filter of the same particle count. */
int pfh_filter_save(void *f, const char *path);
int pfh_filter_load(void *f, const char *path);

#ifdef __cplusplus
}
#endif
#endif /* PF_HOST_H */

/*
 * pf_kernels.h -- C ABI of libpf_kernels_hip.so, the MI355X (gfx950) particle-filter
 * and CBGL global-localization library.
 *
 * The first 18 entry points keep the exact names, argument order, argument meaning
 * and return conventions of the reference library (rus1ru/cuda_pf_localization,
 * src/cuda/pf_kernels.cu), so the reference's Rust host (crates/pf_core/src/cuda.rs
 * and crates/pf_core/src/global_loc.rs) links against this library unchanged.
 * Each declaration cites the reference definition it replaces and the Rust
 * declaration that binds it.
 *
 * Conventions (same as the reference):
 *   - handles are opaque; the library owns all device memory;
 *   - host pointers are borrowed for the duration of the call only;
 *   - int-returning calls return 0 on success and -1 on failure (message on stderr),
 *     create calls return NULL on failure;
 *   - one GPU: device 0.
 *
 * Entry points prefixed pfx_ (particle filter), pfgl_x (global localization),
 * pfb_ (batched filters) and pfl_ (learned proposal) are extensions with no
 * counterpart in the reference; the reference's symbols never depend on them.
 */
#ifndef PF_KERNELS_H
#define PF_KERNELS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 24-byte observation record, identical to the reference's ObsF (pf_kernels.cu:79-84)
 * and CObservation (cuda.rs:18-27). mode 0 = Cartesian offset (dx,dy,dz) in the body
 * frame; mode 1 = range-bearing: `range` plus a unit direction (dx,dy,dz). */
typedef struct pfc_obs {
    int32_t id;
    int32_t mode;
    float dx, dy, dz;
    float range;
} pfc_obs;

/* ------------------------------------------------------------------------------
 * Particle filter (replaces pf_kernels.cu:475-735; bound by cuda.rs:33-88)
 * ---------------------------------------------------------------------------- */

/* 1 if a usable GPU is present, else 0.  Replaces pf_kernels.cu:477-483. */
int pfc_device_available(void);

/* New filter handle; cap_obs outside [1,4096] becomes 64.  Replaces pf_kernels.cu:485-503. */
void *pfc_create(int cap_obs);

/* Frees the handle (NULL is a no-op).  Replaces pf_kernels.cu:505-517 (cuda.rs:35
 * declares it without a return value, which is ABI-compatible). */
int pfc_destroy(void *h);

/* (Re)allocates for n particles; seed 0 becomes 42; resets the generation counter.
 * n must lie in [1, 2^30] (the resampling merge path indexes 2n items with int); a
 * rejected n returns -1 and leaves the handle as it was.  Replaces pf_kernels.cu:519-551. */
int pfc_init(void *h, int n, unsigned int seed);

/* Dense landmark table, 3 floats per id; count = max id + 1 (0 = no landmarks).
 * Replaces pf_kernels.cu:553-565. */
int pfc_upload_landmarks(void *h, const float *xyz, int count);

/* Gaussian prior about (x,y,z | qw,qx,qy,qz); weights become 1/n.
 * Replaces pf_kernels.cu:567-582. */
int pfc_reinit(void *h, float x, float y, float z, float qw, float qx, float qy, float qz,
               const float *pos_std3, const float *rot_std3);

/* Odometry step: p += q*(t + tstd.*n); q = (q (x) q_odo) (x) exp(rstd.*n').
 * Replaces pf_kernels.cu:584-598. */
int pfc_predict(void *h, float tx, float ty, float tz, float qw, float qx, float qy, float qz,
                const float *tstd3, const float *rstd3);

/* Gated log-likelihood weighting + softmax normalisation.  nobs <= 0 resets the
 * weights to 1/n; nobs is clamped to cap_obs; ids outside the landmark table are
 * skipped.  Replaces pf_kernels.cu:600-636. */
int pfc_weight(void *h, const void *obs, int nobs, float sigma_r, float sigma_b, float gate);

/* Systematic resampling + optional uniform re-injection; weights become 1/n.
 * Replaces pf_kernels.cu:638-676. */
int pfc_resample(void *h, float random_inject_ratio);

/* out[14]: sum w*p (3), upper triangle of sum w q q^T (10, order 00,01,02,03,11,12,13,
 * 22,23,33), sum w^2 (1).  Replaces pf_kernels.cu:695-705. */
int pfc_est_sums(void *h, double *out14);

/* out[12]: position covariance (Pxx,Pxy,Pxz,Pyy,Pyz,Pzz) and attitude-tangent covariance
 * (Axx..Azz) about the given means.  Replaces pf_kernels.cu:709-722. */
int pfc_covs(void *h, float mx, float my, float mz, float qw, float qx, float qy, float qz,
             double *out12);

/* Copies positions (3n floats, xyz interleaved) and attitudes (4n floats, wxyz).
 * Replaces pf_kernels.cu:724-732. */
int pfc_snapshot(void *h, float *pos3n, float *quat4n);

/* ------------------------------------------------------------------------------
 * CBGL global localization (replaces pf_kernels.cu:895-1025; bound by global_loc.rs:262-271)
 * ---------------------------------------------------------------------------- */

/* occ: row-major W*H bytes, nonzero = occupied; (ox,oy) = world corner of cell (0,0).
 * Replaces pf_kernels.cu:898-952. */
void *pfgl_create(int W, int H, float res, float ox, float oy, const unsigned char *occ,
                  int cap_hyp);
/* Replaces pf_kernels.cu:954-963. */
int pfgl_destroy(void *h);
/* Number of free cells, -1 on a NULL handle.  Replaces pf_kernels.cu:965-968. */
int pfgl_free_cells(void *h);
/* n uniform hypotheses over free cells x yaw in [-pi, pi).  Replaces pf_kernels.cu:971-979. */
int pfgl_generate(void *h, int n, unsigned long long seed);
/* CAER of every hypothesis against real_scan[nrays]; ray r is at
 * theta + ang_span*r/(nrays-1).  Replaces pf_kernels.cu:982-993. */
int pfgl_score(void *h, const float *real_scan, int nrays, float rmax, float ang_span);
/* Bottom-k hypotheses by CAER, ascending (ties by hypothesis index) into out[3k];
 * returns k (clamped to the hypothesis count) or -1.  Replaces pf_kernels.cu:996-1023. */
int pfgl_topk(void *h, int k, float *out3k);

/* ------------------------------------------------------------------------------
 * Extensions: particle filter (no reference counterpart)
 * ---------------------------------------------------------------------------- */

/* Stage identifiers for pfx_profile_read. */
enum {
    PFX_STAGE_REINIT = 0,
    PFX_STAGE_PREDICT = 1,
    PFX_STAGE_WEIGHT = 2,   /* log-likelihood kernel incl. fused softmax partials */
    PFX_STAGE_NORMALIZE = 3,
    PFX_STAGE_RESAMPLE_SCAN = 4,
    PFX_STAGE_RESAMPLE_GATHER = 5,
    PFX_STAGE_INJECT = 6,
    PFX_STAGE_EST_SUMS = 7,
    PFX_STAGE_EST_COVS = 8,
    PFX_STAGE_FINALIZE = 9, /* one-block-per-filter reductions (weight, ESS, scan top, estimate) */
    PFX_STAGE_COUNT = 10
};

/* Overwrites the particle set (pos 3n xyz, quat 4n wxyz, w n; w may be NULL = 1/n).
 * Checkpoint restore and test injection of identical state. */
int pfx_set_state(void *h, const float *pos3n, const float *quat4n, const float *w);
/* Snapshot including the stored weights (w may be NULL). */
int pfx_get_state(void *h, float *pos3n, float *quat4n, float *w);
/* Per-particle log-likelihoods (f64) of the last pfc_weight call. */
int pfx_get_ll(void *h, double *ll);
/* Ancestor index of every output particle of the last resample (n ints). */
int pfx_get_ancestors(void *h, int32_t *anc);
/* Inclusive fixed-point prefix of the last resampling scan (n u64: cum_i = sum_{j<=i}
 * floor(w_j 2^56)), from which the weights the scan used are recovered exactly. */
int pfx_get_cum(void *h, uint64_t *cum);
/* Sets / reads the resampling LCG state (the u0 and injection stream). */
int pfx_set_lcg(void *h, uint64_t lcg);
int pfx_get_lcg(void *h, uint64_t *lcg);
/* Generation counter and per-generation seed (deterministic-draw mode state); set restores
 * a checkpoint. */
int pfx_get_rng(void *h, uint64_t *gseed, uint64_t *gen);
int pfx_set_rng(void *h, uint64_t gseed, uint64_t gen);
/* Device scalars of the last weight call: [max ll, sum exp(ll-max), sum exp^2, ESS]. */
int pfx_weight_scalars(void *h, double *out4);
/* Per-stage device time: enable=1 starts recording HIP events on the library stream. */
int pfx_profile_enable(void *h, int enable);
/* Accumulated per-stage milliseconds and launch counts since enable (arrays of
 * PFX_STAGE_COUNT); resets the accumulators. */
int pfx_profile_read(void *h, double *ms, int *count);
/* Number of particles of the handle (0 before pfc_init). */
int pfx_particle_count(void *h);
/* Host utility: Markley mean of a 4x4 moment matrix M (row-major; symmetrised), i.e. the
 * unit eigenvector of the largest eigenvalue with w >= 0 (types.rs:128-144), by cyclic
 * Jacobi rotations.  Shared by the host layer and the batched path. */
int pfx_quat_mean(const double *M16, double *q4);
/* Likelihood kernel launch shape: 1, 2 or 4 particles per lane (default 1).  All shapes give
 * bit-identical log-likelihoods; the softmax partials are per 256, 512 or 1024 particles. */
int pfx_set_weight_variant(void *h, int variant);
/* Resampling scan: reads of a predecessor tile's status before the waiting tile sums that
 * predecessor's weights itself (default 64; 0 sums every predecessor not yet published at the
 * first read).  Both routes give the same bits. */
int pfx_set_scan_spin_limit(void *h, int limit);
/* Blocks until every queued operation on the handle's stream has finished. */
int pfx_sync(void *h);

/* Estimate sums without allocation.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Computes pfc_est_sums' out[14] (sum w p, the 10 upper-triangle moments
 * sum w q_r q_c, sum w^2; same kernels, same bits) for n particles in caller-owned device
 * streams px..qz, w (f32, n each) into out14 (device memory, 14 doubles), enqueued on `stream`
 * (a hipStream_t; NULL = the null stream) and not waited for.  The kernels' scratch (layout
 * tables, the filter's scalars, per-1024-particle f64 partials) is a fixed device array of
 * about 2 MB per GPU held until the process ends; no hipMalloc/hipFree is called.  Calls that
 * overlap on one GPU share that array and race.  Returns -1 (message on stderr) for a NULL
 * pointer or n outside [1, PFX_NOALLOC_MAX_PARTICLES]. */
#define PFX_NOALLOC_MAX_PARTICLES 16777216
int pfx_est_sums_noalloc(const float *px, const float *py, const float *pz, const float *qw, const float *qx,
                         const float *qy, const float *qz, const float *w, int n, double *out14, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* PF_KERNELS_H */

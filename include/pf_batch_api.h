/*
 * pf_batch_api.h -- batched filters: many independent particle filters of uneven particle
 * and landmark counts advanced by one launch per stage (libpf_kernels_hip.so).
 *
 * No counterpart in the reference (rus1ru/cuda_pf_localization keeps one PfState per
 * handle, pf_kernels.cu:86-117).  Every filter follows the reference semantics of
 * ParticleFilter::update (particle_filter.rs:110-124) with the deterministic-draw mode of
 * the pfc_* path, so filter f of a batch produces the same results as a single pfc_*
 * handle driven with the same inputs.  The resample decision is taken on the device
 * (ESS of the stored weights < ratio * n), so a cycle needs no host round trip before the
 * estimate; the device stages of a cycle can be replayed from a captured hipGraph.
 */
#ifndef PF_BATCH_API_H
#define PF_BATCH_API_H

#include <stdint.h>

#include "pf_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pfb_filter_cfg {
    int32_t n;                 /* particles */
    uint32_t seed;             /* 0 -> 42 (pf_kernels.cu:523) */
    float sigma_r, sigma_b;    /* measurement noise (config.rs:13-16) */
    float gate;                /* gate_sigma, 0 = off */
    float resample_ess_ratio;  /* resample when ESS < ratio * n; 0 = never */
    float inject_ratio;        /* random_inject_ratio */
} pfb_filter_cfg;

/* Per-filter input of one cycle (odometry + observations, host memory). */
typedef struct pfb_cycle_in {
    float t[3], q[4], tstd[3], rstd[3];
    const pfc_obs *obs;
    int32_t nobs;
} pfb_cycle_in;

/* Per-filter result of one cycle (Estimate, types.rs:92-103). */
typedef struct pfb_estimate {
    double pos[3], quat[4], pos_cov[9], att_cov[9], ess;
    int32_t resampled, valid;
} pfb_estimate;

/* nf filters; obs_cap observations per filter per cycle (clamped like pfc_create).  Every
 * filter's n lies in [1, 2^30], and so does the batch's total rounded up to 1024 per filter. */
void *pfb_create(int nf, const pfb_filter_cfg *cfg, int obs_cap);
int pfb_destroy(void *h);
int pfb_filter_count(void *h);
/* Dense landmark table of filter f (3 floats per id). */
int pfb_upload_landmarks(void *h, int f, const float *xyz, int count);
/* Gaussian prior for every filter: pose7[7 nf], pos_std[3 nf], rot_std[3 nf]. */
int pfb_reinit(void *h, const float *pose7, const float *pos_std, const float *rot_std);
/* One full cycle of every filter; out[nf]. */
int pfb_cycle(void *h, const pfb_cycle_in *in, pfb_estimate *out);
/* Device-resident run of ncycles cycles.  pfb_stage_cycles packs in[k * nf + f] (cycle k,
 * filter f) exactly as pfb_cycle would, advancing the filters' draw state, and uploads every
 * cycle's inputs to device memory; pfb_run_staged then enqueues them back to back (kernels
 * read each cycle's parameters in place) with no host round trip between cycles and
 * returns out[k * nf + f], identical to what ncycles pfb_cycle calls return.  A staging is run
 * once. */
int pfb_stage_cycles(void *h, int ncycles, const pfb_cycle_in *in);
int pfb_run_staged(void *h, pfb_estimate *out);
/* Device time (HIP events on the library stream) from the first to the last kernel of the last
 * pfb_run_staged, in ms. */
int pfb_staged_device_ms(void *h, float *ms);
/* 1: replay the device stages from captured hipGraphs (captured on the next cycle). */
int pfb_use_graph(void *h, int enable);
/* Times the tunable launch shapes on the current state and keeps the fastest;
 * returns the chosen particles-per-lane of the likelihood kernel (1, 2 or 4). */
int pfb_autotune(void *h, const pfb_cycle_in *in);
/* Sets the likelihood kernel's particles per lane (1, 2 or 4) instead of tuning it: a fixed
 * shape for profiling runs.  Every shape gives bit-identical results. */
int pfb_set_weight_ppt(void *h, int ppt);
/* Particles of filter f (pos 3n, quat 4n, weights n; any may be NULL). */
int pfb_get_state(void *h, int f, float *pos3n, float *quat4n, float *w);
int pfb_get_ancestors(void *h, int f, int32_t *anc);
/* Test hook: filter f's fixed-point inclusive prefix cum_j = sum_{i<=j} floor(w_i 2^56) from the
 * last cycle's resampling scan (every filter is scanned, whether or not it resampled); the
 * weights the resampling used are recovered exactly as (cum_j - cum_{j-1}) 2^-56. */
int pfb_get_cum(void *h, int f, uint64_t *cum);
/* Per-stage HIP events (PFX_STAGE_*) around the kernels of eager cycles and of staged runs
 * (not inside captured graphs); pfb_profile_read returns and resets the accumulated times. */
int pfb_profile_enable(void *h, int enable);
int pfb_profile_read(void *h, double *ms, int *count);

#ifdef __cplusplus
}
#endif
#endif /* PF_BATCH_API_H */

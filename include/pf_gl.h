/*
 * pf_gl.h -- extensions of the CBGL global-localization ABI of libpf_kernels_hip.so.
 *
 * The six reference entry points (pfgl_create .. pfgl_topk) are declared in
 * pf_kernels.h.  Everything here is an extension with no counterpart in the reference
 * (rus1ru/cuda_pf_localization); it serves tests, the learned proposal and profiling.
 */
#ifndef PF_GL_H
#define PF_GL_H

#include "pf_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PFGLX_STAGE_GENERATE = 0,
    PFGLX_STAGE_SCORE = 1,
    PFGLX_STAGE_TOPK = 2,
    PFGLX_STAGE_COUNT = 3
};

/* Bottom-k like pfgl_topk, also returning the CAER values and hypothesis indices
 * (either may be NULL).  Returns k or -1. */
int pfglx_topk_ex(void *h, int k, float *out3k, float *caer_k, int32_t *idx_k);
/* Replaces the current hypotheses with n caller poses (x, y, yaw); n <= cap_hyp. */
int pfglx_set_hypotheses(void *h, const float *hyp3n, int n);
/* Copies the current hypotheses; returns their count or -1. */
int pfglx_get_hypotheses(void *h, float *hyp3n);
/* Copies the CAER of every current hypothesis (after pfgl_score); returns the count. */
int pfglx_get_caer(void *h, float *caer);
/* Current hypothesis count. */
int pfglx_hypothesis_count(void *h);
/* Virtual scans: out[n*nrays] for poses[3n], ray r at yaw + ang_span*r/(nrays-1). */
int pfglx_render(void *h, const float *poses3n, int n, int nrays, float rmax, float ang_span, float *out);
/* Same with device pointers, enqueued on `stream` (NULL = the handle's stream). */
int pfglx_render_device(void *h, const float *d_poses3n, int n, int nrays, float rmax, float ang_span,
                        float *d_out, void *stream);
/* Scoring kernel: 0 chooses by hypothesis count (default), 1 one wave per hypothesis (lanes
 * stride the rays), 2 one lane per ray with the per-ray errors summed in a second pass.  All
 * give the same CAER bits. */
int pfglx_set_score_mode(void *h, int mode);
/* Pruned scoring (default on, automatic kernel choice, map in LDS, at least 32,768 hypotheses
 * and 128 rays): pfgl_score and pfglx_score_async compute a lower bound of every CAER (the sum
 * over the rays of every second block of 64) and the next bottom-k scores in full only the
 * max(4k, 64) hypotheses with the smallest bounds and those whose bound does not exceed the
 * k-th smallest CAER among them.  The bottom-k, ties included, is the one full scoring gives;
 * pfglx_get_caer, or a bottom-k with a larger k, completes the scoring of every hypothesis
 * first.  0 disables it. */
int pfglx_set_prune(void *h, int enable);
/* Work of the last pruned bottom-k: out3 = {candidates scored in full, survivors scored after
 * the threshold, route (1 one wave per hypothesis: candidates and survivors cast every ray
 * again; 2 one lane per ray: they cast the rays the bounds left out)}; {0, 0, 0} after a full
 * scoring. */
int pfglx_prune_stats(void *h, int *out3);
int pfglx_profile_enable(void *h, int enable);
int pfglx_profile_read(void *h, double *ms, int *count);
int pfglx_sync(void *h);

#ifdef __cplusplus
}
#endif
#endif /* PF_GL_H */

/*
 * pf_learned.h -- learned proposal for global localization (libpf_kernels_hip.so).
 *
 * No counterpart in the reference (rus1ru/cuda_pf_localization disperses hypotheses
 * uniformly, global_loc.rs:356-382).  A per-map MLP maps a scan (ranges / rmax) to a
 * Gaussian mixture over (x, y, sin yaw, cos yaw); hypotheses drawn from it plus a uniform
 * fraction are scored by CAER (pfgl_score) exactly like dispersed ones, and a caller falls
 * back to uniform dispersal when the best CAER is poor.  Training runs on the GPU on scans
 * rendered by the CBGL raycaster; every GEMM is a hand-written BF16 MFMA kernel with FP32
 * accumulation; master weights and Adam state are FP32.
 *
 * Network: in -> hid -> ... -> hid (n_hidden layers, exact GELU) -> 9*ncomp outputs:
 *   [0, ncomp)            mixture logits
 *   [ncomp, 5 ncomp)      means, component-major (4 per component)
 *   [5 ncomp, 9 ncomp)    log standard deviations (clamped to [-7, 3])
 * Targets: ((x - cx) / span, (y - cy) / span, sin yaw, cos yaw), span = max map side.
 */
#ifndef PF_LEARNED_H
#define PF_LEARNED_H

#include <stdint.h>

#include "pf_gl.h"

#ifdef __cplusplus
extern "C" {
#endif

void *pfl_create(int in_dim, int hid, int n_hidden, int ncomp, int max_batch, uint64_t seed);
int pfl_destroy(void *h);
/* Parameter count and flat FP32 parameter I/O, layer by layer: W (out x in, row-major)
 * then b (out). */
int64_t pfl_param_count(void *h);
int pfl_get_params(void *h, float *out);
int pfl_set_params(void *h, const float *in);
/* Forward on host inputs x[B*in] (already normalised) -> out[B*9*ncomp]. */
int pfl_forward(void *h, const float *x, int B, float *out);
/* Loss and FP32 gradients (same layout as the parameters) of a host batch, no update. */
int pfl_grad(void *h, const float *x, const float *y, int B, float *grads, double *loss);
/* One Adam step on a host batch; returns the batch loss in *loss. */
int pfl_train_step(void *h, const float *x, const float *y, int B, float lr, double *loss);
/* Trains on n_samples scans rendered on the GPU from uniform free poses of the map of the
 * pfgl handle (nrays over ang_span, rmax, Gaussian range noise noise_sigma), in batches of
 * `batch`, Adam at lr.  mean_loss[0] = mean loss of the first 5 % of steps, [1] = last 5 %. */
int pfl_train_on_map(void *h, void *gl, int64_t n_samples, int batch, float lr, int nrays, float rmax,
                     float ang_span, float noise_sigma, uint64_t seed, double *mean_loss2);
/* Proposal for one scan: n_mix hypotheses from the mixture followed by n_uniform uniform
 * ones become the current hypotheses of the pfgl handle (n_mix + n_uniform <= cap_hyp).
 * Enqueued on the pfgl handle's stream; returns without waiting for the device (the scan
 * is copied before returning). */
int pfl_propose(void *h, void *gl, const float *scan, int nrays, float rmax, int n_mix, int n_uniform,
                uint64_t seed);
/* Device time of the last pfl_propose (input copy + forward + sampling), milliseconds;
 * waits for that proposal to finish. */
double pfl_last_propose_ms(void *h);

/* Extension of the pfgl handle used by the learned proposal and its training. */
/* Uniform free-space poses (x, y, yaw) into a caller device buffer, on `stream` (NULL =
 * the handle's stream); same generator as pfgl_generate. */
int pfglx_sample_poses_device(void *gl, int n, uint64_t seed, float *d_out3n, void *stream);
/* Generate n uniform hypotheses into slots [offset, offset + n); count becomes offset + n. */
int pfglx_generate_range(void *gl, int offset, int n, uint64_t seed);
/* Device pointer of the hypothesis buffer (3 floats per hypothesis, cap_hyp entries) and
 * the map geometry: out6 = {W, H, res, ox, oy, nfree}. */
float *pfglx_hypotheses_device(void *gl);
int pfglx_map_info(void *gl, double *out6);
/* Sets the hypothesis count after the caller filled the device buffer. */
int pfglx_set_count(void *gl, int n);
/* The handle's HIP stream (work enqueued there is ordered with the pfgl_* calls). */
void *pfglx_stream(void *gl);
/* Uploads a scan (through pinned staging; returns without waiting for the copy) and
 * prepares the ray table; pfglx_score_async then scores the current hypotheses against it
 * without waiting.  pfgl_score = pfglx_set_scan + pfglx_score_async + wait. */
int pfglx_set_scan(void *gl, const float *scan, int nrays, float ang_span);
int pfglx_score_async(void *gl, int nrays, float rmax);
/* Bottom-k of the current CAERs kept on the device only (no copy, no wait); returns k. */
int pfglx_topk_device(void *gl, int k);
/* Local refinement: the first k_parents rows of the last top-k each spawn per_parent
 * hypotheses (the parent itself plus Gaussian jitter sigma_xy in x, y and sigma_yaw in yaw,
 * clamped to the map); they replace the current hypotheses. */
int pfglx_refine(void *gl, int k_parents, int per_parent, float sigma_xy, float sigma_yaw, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* PF_LEARNED_H */

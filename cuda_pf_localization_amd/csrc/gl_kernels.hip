// gl_kernels.hip -- CBGL global localisation on MI355X (gfx950): pfgl_* / pfglx_* C ABI.
//
// Drop-in for src/cuda/pf_kernels.cu:737-1025 of the reference (bound by
// crates/pf_core/src/global_loc.rs:262-271).  Pipeline per localize():
//   generate  hypotheses uniform over the free cells (counter-based splitmix64 draws;
//             free cells compacted in row-major order, like global_loc.rs:365-368)
//   score     2-D DDA over the occupancy grid framed by set cells, staged in LDS (one byte
//             per cell, or bit-packed); large batches one wave64 per hypothesis with lane l
//             casting rays l, l+64, ...; small batches one lane per ray and a second pass
//             that sums in that same order; CAER = sum |real - virtual| with a fixed lane
//             order and wave tree (deterministic, identical bits on both routes)
//   topk      bottom-k by radix *select* over unique 64-bit keys (ordered CAER bits << 32 |
//             hypothesis index), then one workgroup bitonic-sorts the k winners; ties keep
//             index order as the CPU twin's stable sort does (global_loc.rs:395-397).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/pf_learned.h"
#include "pf_device.h"

using namespace pfd;

#define GL_CHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "pf_hip %s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

#define GL_LAUNCH_CHECK()                                                                        \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "pf_hip %s:%d launch failed: %s\n", __FILE__, __LINE__,              \
                    hipGetErrorString(e_));                                                      \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

namespace glk {

constexpr int kTB = 256;
constexpr int kLdsBitsMaxWords = 6144;   // 24 KB of grid bits in LDS (196,608 cells)
constexpr int kLdsRaysMax = 1024;        // real scan + ray table staged in LDS up to this many rays
constexpr size_t kLdsScoreBytesMax = 64 * 1024;  // byte grid + ray tables per scoring block (two per CU)
constexpr int kScoreBytesMinHyp = 16384;         // below this, 256-thread bit-grid blocks fill the chip better
constexpr int kScoreBlockMaxHyp = 65536;         // below this, one lane per ray in two passes (mode 2)
constexpr size_t kRayDiffsMax = (size_t)32 << 20;  // per-ray |diff| scratch of the two-pass route (128 MB)
constexpr int kSortMax = 4096;           // winners sorted in one workgroup up to this k

struct Grid {
    // Occupancy bits of the map framed by one extra cell on every side, row-major over
    // (W + 2) x (H + 2): bit (cy + 1) * (W + 2) + cx + 1 is map cell (cx, cy), 1 = occupied.
    // The frame cells are set, so a ray stops on the frame where it would leave the map.
    const uint32_t* bits;
    const uint8_t* bytes;  // the same framed grid, one byte per cell: 1 obstacle, 2 frame, 0 free
                           // (nbytes, a multiple of 16)
    int W, H, nwords, nbytes;
    float res, ox, oy;
};

struct GlState {
    hipStream_t stream = nullptr;
    Grid g{};
    uint32_t* d_bits = nullptr;
    uint8_t* d_bytes = nullptr;
    int32_t* d_free = nullptr;  // (cx, cy) pairs, row-major order
    int nfree = 0;
    int cap_hyp = 0;
    int nh = 0;
    int last_k = 0;  // rows of d_out written by the last top-k
    float* d_hyp = nullptr;
    float* d_caer = nullptr;
    // scan + ray table
    float* d_scan = nullptr;
    double* d_cphi = nullptr;
    double* d_sphi = nullptr;
    int ray_cap = 0;
    int tab_nrays = -1;
    float tab_span = 0.f;
    int scan_nrays = 0;  // rays of the scan resident in d_scan
    int score_mode = 0;  // 0 by hypothesis count, 1 one wave per hypothesis, 2 one lane per ray (two passes)
    float* d_diff = nullptr;  // per-ray |scan - cast| of the two-pass scoring
    size_t diff_cap = 0;
    float* h_scan = nullptr;  // pinned staging of the scan upload
    int h_scan_cap = 0;
    hipEvent_t scan_ev = nullptr;  // recorded after each upload from h_scan
    bool scan_ev_live = false;
    // top-k
    uint64_t* d_tk = nullptr;  // [0] prefix, [1] threshold, [2] remaining|done, [3] count
    unsigned* d_hist = nullptr;
    uint64_t* d_win = nullptr;
    uint64_t* d_sortbuf = nullptr;
    int sort_cap = 0;
    float* d_out = nullptr;
    float* d_outcaer = nullptr;
    int32_t* d_outidx = nullptr;
    // render scratch
    float* d_poses = nullptr;
    float* d_ranges = nullptr;
    size_t render_poses_cap = 0, render_ranges_cap = 0;
    // Pruned scoring (large batches): pfgl_score leaves a lower bound of every CAER (the sum
    // over every second ray) and the next bottom-k scores only what can still be among the k
    // smallest.  caer_state: 0 d_caer holds every hypothesis' CAER; 1 only the bounds (d_lb) of
    // the scoring of pend_nrays rays at pend_rmax; 2 exact for every hypothesis that can be among
    // the bottom pruned_k, +inf for the others.
    bool prune = true;
    int caer_state = 0;
    int pend_route = 1;  // of the pending bounds: 1 the wave kernel, 2 one lane per ray
    int pend_nrays = 0, pruned_k = 0;
    float pend_rmax = 0.f;
    float* d_lb = nullptr;
    int32_t* d_surv = nullptr;
    int* d_nsurv = nullptr;  // [0] survivor count
    float* d_tau = nullptr;
    int last_m = 0, last_route = 0;  // candidates and route of the last pruned bottom-k (0: none)
    float* d_cand = nullptr;   // packed poses of the candidates
    float* d_cdiff = nullptr;  // their per-ray errors (kSortMax x kLdsRaysMax)
    // profiling
    bool prof_on = false;
    struct Ev {
        int stage;
        hipEvent_t a, b;
    };
    std::vector<Ev> prof;
};

// Cell lookups for the raycaster: packed bits (small footprint) or one byte per cell (one LDS
// read and one compare per step).  get() is non-zero for a set cell; frame() tells, for the
// set cell a walk ended on, whether it is a frame cell (the ray left the map): the byte grid
// stores frame cells as 2 (obstacles 1), the bit grid is asked by position.
struct OccBits {
    const uint32_t* bits;
    __device__ uint32_t get(int idx) const { return (bits[idx >> 5] >> (idx & 31)) & 1u; }
    __device__ bool frame(int idx, uint32_t, int Wp, int H) const {
        const int qy = idx / Wp, qx = idx - qy * Wp;
        return qx == 0 || qy == 0 || qx == Wp - 1 || qy == H + 1;
    }
};
struct OccBytes {
    const uint8_t* bytes;
    __device__ uint32_t get(int idx) const { return bytes[idx]; }
    __device__ bool frame(int, uint32_t v, int, int) const { return v == 2u; }
};

// f32 DDA with the operation sequence of pf_kernels.cu:784-808 (distance to the first
// occupied cell along a unit direction; rmax on exit, miss or beyond rmax).  The walk tracks
// only the framed linear index: leaving the map means entering the set frame, so the loop
// tests one cell per step and the frame is told apart from an obstacle once, at the end.
template <class Occ>
__device__ inline float cast_ray(Occ occ_at, int W, int H, float res, float ox, float oy, float px, float py,
                                 float dirx, float diry, float rmax) {
    const float fx = (px - ox) / res, fy = (py - oy) / res;
    const int cx = (int)floorf(fx), cy = (int)floorf(fy);
    if (!(cx >= 0 && cy >= 0 && cx < W && cy < H)) return rmax;
    const int sx = dirx > 0.f ? 1 : -1, sy = diry > 0.f ? 1 : -1;
    const float adx = fabsf(dirx), ady = fabsf(diry);
    const float tdx = adx > 1e-9f ? res / adx : INFINITY;
    const float tdy = ady > 1e-9f ? res / ady : INFINITY;
    const float frx = fx - (float)cx, fry = fy - (float)cy;
    float tmx = adx > 1e-9f ? ((dirx > 0.f ? 1.f - frx : frx) * tdx) : INFINITY;
    float tmy = ady > 1e-9f ? ((diry > 0.f ? 1.f - fry : fry) * tdy) : INFINITY;
    const int Wp = W + 2;
    int idx = (cy + 1) * Wp + cx + 1;
    const int dxi = sx, dyi = sy * Wp;
    // Branch-free step: both candidate boundary times are formed and one is selected, which
    // rounds exactly as the reference's if/else.  The loop carries only the walk and the last
    // step's distance and cell; which test ended it is decided once afterwards (beyond rmax
    // first, as the reference tests it first).  The frame ends every walk within W + H steps
    // whatever the direction or rmax; the step bound is a guard only.
    float dist = 0.f;
    uint32_t v = 0;
    for (int step = 0; step < 8192; ++step) {
        const bool xs = tmx <= tmy;
        dist = xs ? tmx : tmy;
        const float tmx1 = tmx + tdx, tmy1 = tmy + tdy;
        tmx = xs ? tmx1 : tmx;
        tmy = xs ? tmy : tmy1;
        idx += xs ? dxi : dyi;
        v = occ_at.get(idx);
        if (v != 0u || dist > rmax) break;
    }
    // A set cell found past rmax (on the step that crossed it) yields rmax: the distance never
    // decreases along the walk, so clamping equals stopping at the first step beyond rmax.
    const float r = (v == 0u || occ_at.frame(idx, v, Wp, H)) ? rmax : dist;
    return fminf(r, rmax);
}

// Hypothesis h (deterministic-draw mode): cell = floor(u53 * nfree); x/y offsets and yaw
// from 24-bit uniforms; f32 arithmetic in the order of pf_kernels.cu:817-822.
__global__ __launch_bounds__(kTB) void k_gl_gen(const int32_t* __restrict__ freexy, int nfree, float res, float ox,
                                                float oy, uint64_t seed, int n, float* __restrict__ hyp) {
    const int h = blockIdx.x * kTB + threadIdx.x;
    if (h >= n) return;
    uint64_t x = sm_final(seed * kPhi + (uint64_t)h * kC1);
    const uint64_t z0 = sm_next(x), z1 = sm_next(x), z2 = sm_next(x), z3 = sm_next(x);
    long c = (long)(u01_co(z0) * (double)nfree);
    if (c > nfree - 1) c = nfree - 1;
    const float ux = (float)(z1 >> 40) * 0x1.0p-24f;
    const float uy = (float)(z2 >> 40) * 0x1.0p-24f;
    const float ut = (float)(z3 >> 40) * 0x1.0p-24f;
    const float fxc = (float)freexy[2 * c] + ux;
    const float fyc = (float)freexy[2 * c + 1] + uy;
    const float fxr = fxc * res;
    const float fyr = fyc * res;
    hyp[3 * h] = ox + fxr;
    hyp[3 * h + 1] = oy + fyr;
    const float tt = 6.28318530f * ut;
    hyp[3 * h + 2] = -3.14159265f + tt;
}

// CAER scoring.  LDS layout: [grid][real scan f32][cos phi f64][sin phi f64]; the grid is
// staged as bytes (kBytes, 1024-thread blocks, two per CU) or as bits (256-thread blocks).
template <int kBT, bool kBytes, bool kLdsGrid, bool kLdsRays>
__global__ __launch_bounds__(kBT) void k_gl_score(Grid g, const float* __restrict__ hyp, int nh,
                                                  const float* __restrict__ scan, const double* __restrict__ cphi,
                                                  const double* __restrict__ sphi, int nrays, float rmax,
                                                  float* __restrict__ caer) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int gbytes = kLdsGrid ? (kBytes ? g.nbytes : 4 * ((g.nwords + 3) & ~3)) : 0;
    double* scphi = reinterpret_cast<double*>(smem + gbytes);
    double* ssphi = scphi + (kLdsRays ? nrays : 0);
    float* sscan = reinterpret_cast<float*>(ssphi + (kLdsRays ? nrays : 0));
    if (kLdsGrid) {
        if (kBytes) {
            const uint4* src = reinterpret_cast<const uint4*>(g.bytes);
            uint4* dst = reinterpret_cast<uint4*>(smem);
            for (int i = threadIdx.x; i < g.nbytes / 16; i += kBT) dst[i] = src[i];
        } else {
            uint32_t* sbits = reinterpret_cast<uint32_t*>(smem);
            for (int i = threadIdx.x; i < g.nwords; i += kBT) sbits[i] = g.bits[i];
        }
    }
    if (kLdsRays)
        for (int i = threadIdx.x; i < nrays; i += kBT) {
            scphi[i] = cphi[i];
            ssphi[i] = sphi[i];
            sscan[i] = scan[i];
        }
    __syncthreads();
    const double* tc = kLdsRays ? scphi : cphi;
    const double* ts = kLdsRays ? ssphi : sphi;
    const float* sc = kLdsRays ? sscan : scan;
    const int lane = threadIdx.x & 63;
    const int waves_total = gridDim.x * (kBT / 64);
    for (int h = blockIdx.x * (kBT / 64) + (threadIdx.x >> 6); h < nh; h += waves_total) {
        const float x = hyp[3 * h], y = hyp[3 * h + 1];
        double st, ct;
        det_sincos((double)hyp[3 * h + 2], st, ct);
        float acc = 0.f;
        for (int r = lane; r < nrays; r += 64) {
            const double cp = tc[r], sp = ts[r];
            const float dx = (float)(ct * cp - st * sp);
            const float dy = (float)(st * cp + ct * sp);
            float vm;
            if (kBytes)
                vm = cast_ray(OccBytes{kLdsGrid ? smem : g.bytes}, g.W, g.H, g.res, g.ox, g.oy, x, y, dx, dy, rmax);
            else
                vm = cast_ray(OccBits{kLdsGrid ? reinterpret_cast<const uint32_t*>(smem) : g.bits}, g.W, g.H, g.res,
                              g.ox, g.oy, x, y, dx, dy, rmax);
            acc += fabsf(sc[r] - vm);
        }
        acc = wave_sum_down(acc);
        if (lane == 0) caer[h] = acc;
    }
}

// Scoring for pruned scoring (the large-batch byte-grid kernel above, with a hypothesis list and
// a ray subset; a kernel of its own so that these options leave the plain scoring kernel's code
// as it is: folded into it, they made that kernel 17 % slower with the same inner loop).
// kSub = 2: every second round of rays only (lane l: rays l, l + 128, l + 256, ...: blocks of 64
// consecutive rays, so a wave's lanes still walk neighbouring directions), a lower bound of the
// CAER up to rounding, for pruned scoring.  list (non-null): score the hypotheses list[i] for i below
// *nlist (nlist non-null) or nh; mark (non-null): set mark[h] = -1 for every hypothesis scored.
template <int kBT, bool kBytes, bool kLdsGrid, bool kLdsRays, int kSub = 1>
__global__ __launch_bounds__(kBT) void k_gl_score_sel(Grid g, const float* __restrict__ hyp, int nh,
                                                  const float* __restrict__ scan, const double* __restrict__ cphi,
                                                  const double* __restrict__ sphi, int nrays, float rmax,
                                                  float* __restrict__ caer, const int32_t* __restrict__ list = nullptr,
                                                  const int* __restrict__ nlist = nullptr,
                                                  float* __restrict__ mark = nullptr) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int gbytes = kLdsGrid ? (kBytes ? g.nbytes : 4 * ((g.nwords + 3) & ~3)) : 0;
    double* scphi = reinterpret_cast<double*>(smem + gbytes);
    double* ssphi = scphi + (kLdsRays ? nrays : 0);
    float* sscan = reinterpret_cast<float*>(ssphi + (kLdsRays ? nrays : 0));
    if (kLdsGrid) {
        if (kBytes) {
            const uint4* src = reinterpret_cast<const uint4*>(g.bytes);
            uint4* dst = reinterpret_cast<uint4*>(smem);
            for (int i = threadIdx.x; i < g.nbytes / 16; i += kBT) dst[i] = src[i];
        } else {
            uint32_t* sbits = reinterpret_cast<uint32_t*>(smem);
            for (int i = threadIdx.x; i < g.nwords; i += kBT) sbits[i] = g.bits[i];
        }
    }
    if (kLdsRays)
        for (int i = threadIdx.x; i < nrays; i += kBT) {
            scphi[i] = cphi[i];
            ssphi[i] = sphi[i];
            sscan[i] = scan[i];
        }
    __syncthreads();
    const double* tc = kLdsRays ? scphi : cphi;
    const double* ts = kLdsRays ? ssphi : sphi;
    const float* sc = kLdsRays ? sscan : scan;
    const int lane = threadIdx.x & 63;
    const int waves_total = gridDim.x * (kBT / 64);
    const int cnt = nlist ? *nlist : nh;
    for (int i = blockIdx.x * (kBT / 64) + (threadIdx.x >> 6); i < cnt; i += waves_total) {
        const int h = list ? list[i] : i;
        const float x = hyp[3 * h], y = hyp[3 * h + 1];
        double st, ct;
        det_sincos((double)hyp[3 * h + 2], st, ct);
        float acc = 0.f;
        for (int r = lane; r < nrays; r += 64 * kSub) {
            const double cp = tc[r], sp = ts[r];
            const float dx = (float)(ct * cp - st * sp);
            const float dy = (float)(st * cp + ct * sp);
            float vm;
            if (kBytes)
                vm = cast_ray(OccBytes{kLdsGrid ? smem : g.bytes}, g.W, g.H, g.res, g.ox, g.oy, x, y, dx, dy, rmax);
            else
                vm = cast_ray(OccBits{kLdsGrid ? reinterpret_cast<const uint32_t*>(smem) : g.bits}, g.W, g.H, g.res,
                              g.ox, g.oy, x, y, dx, dy, rmax);
            acc += fabsf(sc[r] - vm);
        }
        acc = wave_sum_down(acc);
        if (lane == 0) {
            caer[h] = acc;
            if (mark) mark[h] = -1.f;
        }
    }
}


// Pruned scoring, survivors: a hypothesis whose bound, less a margin that covers the rounding of
// both sums, exceeds the threshold cannot be among the bottom k (k candidates have CAER <= tau);
// its CAER is set to +inf.  Every other one, except the candidates (bound marked -1, CAER exact
// already), is appended to the survivor list, one atomic per wave.
__global__ __launch_bounds__(kTB) void k_gl_prune(const float* __restrict__ lb, int n, const float* __restrict__ tau,
                                                  float* __restrict__ caer, int32_t* __restrict__ surv,
                                                  int* __restrict__ nsurv) {
    const int i = blockIdx.x * kTB + threadIdx.x;
    bool keep = false;
    if (i < n) {
        const float b = lb[i];
        if (!(b < 0.f)) {  // (a NaN bound is kept: the comparison below is false for it)
            if ((double)b * (1.0 - 1e-4) > (double)*tau)
                caer[i] = INFINITY;
            else
                keep = true;
        }
    }
    const uint64_t bal = __ballot(keep);
    if (bal == 0) return;
    const int lane = threadIdx.x & 63;
    int base = 0;
    if (lane == __ffsll((long long)bal) - 1) base = atomicAdd(nsurv, __popcll(bal));
    base = __shfl(base, __ffsll((long long)bal) - 1, 64);
    if (keep) surv[base + __popcll(bal & ((1ull << lane) - 1ull))] = i;
}

// One lane per (pose, ray) over a whole batch, consecutive lanes on consecutive rays of one
// pose, so a batch of a few thousand poses still fills every CU.  kDiff = false writes the
// cast range (virtual scans: training data, tests); kDiff = true writes |scan - range|, the
// first pass of the two-pass CAER for batches too small to fill the device one wave per
// hypothesis (2k hypotheses are 2k waves; one lane per ray gives 11k), which k_gl_caer_sum
// then adds in the order of k_gl_score (lane l: rays l, l + 64, ...; then the same
// shuffle-down tree), so CAER keeps its bits.  Directions are computed exactly as in
// k_gl_score (per-pose det_sincos, rotated f64 table entry rounded to f32).  The grid: kGrid 0
// bytes staged in LDS (1024-thread workgroups, two per CU), 1 bits staged in LDS, 2 bits read
// through L2.
template <int kBT, int kGrid, bool kDiff>
__global__ __launch_bounds__(kBT) void k_gl_rays(Grid g, const float* __restrict__ poses, int n,
                                                 const float* __restrict__ scan, const double* __restrict__ cphi,
                                                 const double* __restrict__ sphi, int nrays, float rmax,
                                                 float* __restrict__ out) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (kGrid == 0) {
        const uint4* src = reinterpret_cast<const uint4*>(g.bytes);
        uint4* dst = reinterpret_cast<uint4*>(smem);
        for (int i = threadIdx.x; i < g.nbytes / 16; i += kBT) dst[i] = src[i];
        __syncthreads();
    } else if (kGrid == 1) {
        uint32_t* sbits = reinterpret_cast<uint32_t*>(smem);
        for (int i = threadIdx.x; i < g.nwords; i += kBT) sbits[i] = g.bits[i];
        __syncthreads();
    }
    const size_t total = (size_t)n * (size_t)nrays;
    for (size_t t = (size_t)blockIdx.x * kBT + threadIdx.x; t < total; t += (size_t)gridDim.x * kBT) {
        const int h = (int)(t / (size_t)nrays), r = (int)(t - (size_t)h * nrays);
        const float x = poses[3 * h], y = poses[3 * h + 1];
        double st, ct;
        det_sincos((double)poses[3 * h + 2], st, ct);
        const double cp = cphi[r], sp = sphi[r];
        const float dx = (float)(ct * cp - st * sp);
        const float dy = (float)(st * cp + ct * sp);
        float vm;
        if (kGrid == 0)
            vm = cast_ray(OccBytes{smem}, g.W, g.H, g.res, g.ox, g.oy, x, y, dx, dy, rmax);
        else
            vm = cast_ray(OccBits{kGrid == 1 ? reinterpret_cast<const uint32_t*>(smem) : g.bits}, g.W, g.H, g.res,
                          g.ox, g.oy, x, y, dx, dy, rmax);
        out[t] = kDiff ? fabsf(scan[r] - vm) : vm;
    }
}

// idx (non-null): the sum of row h is the CAER of hypothesis idx[h], and mark[idx[h]] = -1.
__global__ __launch_bounds__(kTB) void k_gl_caer_sum(const float* __restrict__ diff, int nh, int nrays,
                                                     float* __restrict__ caer, const int32_t* __restrict__ idx = nullptr,
                                                     float* __restrict__ mark = nullptr) {
    const int h = blockIdx.x * (kTB / 64) + (threadIdx.x >> 6);
    if (h >= nh) return;
    const int lane = threadIdx.x & 63;
    const float* d = diff + (size_t)h * nrays;
    float acc = 0.f;
    for (int r = lane; r < nrays; r += 64) acc += d[r];
    acc = wave_sum_down(acc);
    if (lane == 0) {
        const int o = idx ? idx[h] : h;
        caer[o] = acc;
        if (mark) mark[o] = -1.f;
    }
}

// Pruned scoring on the one-lane-per-ray route.  Ray r belongs to block (r >> 6); kSel = 1 casts
// the rays of the even blocks (0-63, 128-191, ...: the rays lane l of k_gl_score casts in its
// rounds 0, 2, 4, ...), kSel = 2 those of the odd blocks, each |scan - cast| stored at diff[h
// nrays + r] exactly as k_gl_rays stores it (same direction arithmetic), for the hypotheses
// h = list[i] (i < *nlist, or i < n when nlist is null; h = i when list is null).  The byte grid
// is staged in LDS (1024-thread workgroups).
template <int kSel>
__global__ __launch_bounds__(1024) void k_gl_rays_sel(Grid g, const float* __restrict__ hyp, int n,
                                                      const int32_t* __restrict__ list, const int* __restrict__ nlist,
                                                      const float* __restrict__ scan, const double* __restrict__ cphi,
                                                      const double* __restrict__ sphi, int nrays, float rmax,
                                                      float* __restrict__ diff) {
    extern __shared__ __align__(16) unsigned char smem[];
    {
        const uint4* src = reinterpret_cast<const uint4*>(g.bytes);
        uint4* dst = reinterpret_cast<uint4*>(smem);
        for (int i = threadIdx.x; i < g.nbytes / 16; i += 1024) dst[i] = src[i];
        __syncthreads();
    }
    const int cnt = nlist ? *nlist : n;
    const int per = ((nrays + 127) >> 7) << 6;  // selected-ray slots per hypothesis
    const size_t total = (size_t)cnt * (size_t)per;
    for (size_t t = (size_t)blockIdx.x * 1024 + threadIdx.x; t < total; t += (size_t)gridDim.x * 1024) {
        const int i = (int)(t / (size_t)per), j = (int)(t - (size_t)i * per);
        const int r = ((j >> 6) << 7) + (kSel == 2 ? 64 : 0) + (j & 63);
        if (r >= nrays) continue;
        const int h = list ? list[i] : i;
        const float x = hyp[3 * h], y = hyp[3 * h + 1];
        double st, ct;
        det_sincos((double)hyp[3 * h + 2], st, ct);
        const double cp = cphi[r], sp = sphi[r];
        const float dx = (float)(ct * cp - st * sp);
        const float dy = (float)(st * cp + ct * sp);
        const float vm = cast_ray(OccBytes{smem}, g.W, g.H, g.res, g.ox, g.oy, x, y, dx, dy, rmax);
        diff[(size_t)h * nrays + r] = fabsf(scan[r] - vm);
    }
}

// Bound of every hypothesis from the even blocks' errors: their sum, in any order (the margin of
// k_gl_prune covers its rounding).
__global__ __launch_bounds__(kTB) void k_gl_lb(const float* __restrict__ diff, int nh, int nrays,
                                               float* __restrict__ lb) {
    const int h = blockIdx.x * (kTB / 64) + (threadIdx.x >> 6);
    if (h >= nh) return;
    const int lane = threadIdx.x & 63;
    const float* d = diff + (size_t)h * nrays;
    float acc = 0.f;
    for (int r = lane; r < nrays; r += 128) acc += d[r];
    acc = wave_sum_down(acc);
    if (lane == 0) lb[h] = acc;
}

// CAER of the listed hypotheses from their complete rows of errors, in k_gl_score's order (lane
// l: rays l, l + 64, ...; then the shuffle-down tree): caer[h] for h = list[i], i < *nlist (or
// < n); mark (non-null): mark[h] = -1.
__global__ __launch_bounds__(kTB) void k_gl_caer_rows(const float* __restrict__ diff, const int32_t* __restrict__ list,
                                                      const int* __restrict__ nlist, int n, int nrays,
                                                      float* __restrict__ caer, float* __restrict__ mark) {
    const int cnt = nlist ? *nlist : n;
    const int lane = threadIdx.x & 63;
    for (int i = blockIdx.x * (kTB / 64) + (threadIdx.x >> 6); i < cnt; i += gridDim.x * (kTB / 64)) {
        const int h = list[i];
        const float* d = diff + (size_t)h * nrays;
        float acc = 0.f;
        for (int r = lane; r < nrays; r += 64) acc += d[r];
        acc = wave_sum_down(acc);
        if (lane == 0) {
            caer[h] = acc;
            if (mark) mark[h] = -1.f;
        }
    }
}

// Poses of the listed hypotheses, packed: out[3 i ..] = hyp[3 idx[i] ..] for i < n.
__global__ __launch_bounds__(kTB) void k_gl_pick(const float* __restrict__ hyp, const int32_t* __restrict__ idx, int n,
                                                 float* __restrict__ out) {
    const int i = blockIdx.x * kTB + threadIdx.x;
    if (i >= n) return;
    const int h = idx[i];
    out[3 * i] = hyp[3 * h];
    out[3 * i + 1] = hyp[3 * h + 1];
    out[3 * i + 2] = hyp[3 * h + 2];
}

// ---------------------------------------------------------------- bottom-k selection

__device__ inline uint64_t comp_key(float c, int i) {
    uint32_t b = __float_as_uint(c);
    b = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
    return ((uint64_t)b << 32) | (uint32_t)i;
}

// Pruned scoring, threshold: the CAER of the k-th smallest of the m fully scored candidates
// cand[0 .. m) (k <= m <= kSortMax), in the (CAER bits, index) order of the bottom-k (LDS bitonic
// sort); also resets the survivor count for k_gl_prune.
__global__ __launch_bounds__(1024) void k_gl_tau(const float* __restrict__ caer, const int32_t* __restrict__ cand,
                                                 int m, int k, float* __restrict__ tau, int* __restrict__ nsurv) {
    __shared__ uint64_t key[4096];
    const int t = threadIdx.x;
    int p2 = 1;
    while (p2 < m) p2 <<= 1;
    for (int j = t; j < p2; j += 1024) key[j] = j < m ? comp_key(caer[cand[j]], cand[j]) : ~0ull;
    __syncthreads();
    for (int size = 2; size <= p2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int j = t; j < p2; j += 1024) {
                const int q = j ^ stride;
                if (q > j) {
                    const uint64_t a = key[j], b = key[q];
                    const bool up = (j & size) == 0;
                    if ((a > b) == up) {
                        key[j] = b;
                        key[q] = a;
                    }
                }
            }
            __syncthreads();
        }
    if (t == 0) {
        *tau = caer[(uint32_t)key[k - 1]];
        *nsurv = 0;
    }
}

// tk[0] prefix, tk[1] threshold, tk[2] remaining, tk[3] done, tk[4] collect counter
__global__ void k_tk_init(uint64_t* tk, unsigned* hist, int k) {
    const int t = threadIdx.x;
    if (t < 256) hist[t] = 0;
    if (t == 0) {
        tk[0] = 0;
        tk[1] = ~0ull;
        tk[2] = (uint64_t)k;
        tk[3] = 0;
        tk[4] = 0;
    }
}

__global__ __launch_bounds__(kTB) void k_tk_hist(const float* __restrict__ caer, int n, const uint64_t* tk,
                                                 unsigned* hist, int shift) {
    __shared__ unsigned sh[256];
    if (tk[3]) return;
    sh[threadIdx.x] = 0;
    __syncthreads();
    const uint64_t prefix = tk[0];
    const int hs = shift + 8;
    const uint64_t hiref = hs >= 64 ? 0 : (prefix >> hs);
    for (int i = blockIdx.x * kTB + threadIdx.x; i < n; i += gridDim.x * kTB) {
        const uint64_t c = comp_key(caer[i], i);
        const uint64_t hi = hs >= 64 ? 0 : (c >> hs);
        if (hi == hiref) atomicAdd(&sh[(c >> shift) & 255u], 1u);
    }
    __syncthreads();
    const unsigned v = sh[threadIdx.x];
    if (v) atomicAdd(&hist[threadIdx.x], v);
}

__global__ void k_tk_pick(uint64_t* tk, unsigned* hist, int shift) {
    __shared__ unsigned sh[256];
    const int t = threadIdx.x;
    const bool done = tk[3] != 0;
    sh[t] = hist[t];
    hist[t] = 0;
    __syncthreads();
    if (done || t != 0) return;
    uint64_t rem = tk[2];
    uint64_t before = 0;
    int b = 0;
    for (; b < 256; ++b) {
        if (before + sh[b] >= rem) break;
        before += sh[b];
    }
    if (b > 255) b = 255;  // unreachable when k <= n
    rem -= before;
    const uint64_t prefix = tk[0] | ((uint64_t)b << shift);
    tk[0] = prefix;
    tk[2] = rem;
    const uint64_t low = shift == 0 ? 0 : ((1ull << shift) - 1);
    if (sh[b] == rem || shift == 0) {
        tk[1] = prefix | low;
        tk[3] = 1;
    }
}

__global__ __launch_bounds__(kTB) void k_tk_collect(const float* __restrict__ caer, int n, uint64_t* tk,
                                                    uint64_t* __restrict__ win, int k) {
    const uint64_t thr = tk[1];
    for (int i = blockIdx.x * kTB + threadIdx.x; i < n; i += gridDim.x * kTB) {
        const uint64_t c = comp_key(caer[i], i);
        if (c <= thr) {
            const unsigned long long slot = atomicAdd(reinterpret_cast<unsigned long long*>(&tk[4]), 1ull);
            if (slot < (unsigned long long)k) win[slot] = c;
        }
    }
}

// Sorts k <= kSortMax keys in one workgroup (bitonic in LDS) and gathers the poses.
__global__ __launch_bounds__(1024) void k_tk_sort(const uint64_t* __restrict__ win, const uint64_t* tk, int k, int n,
                                                  const float* __restrict__ hyp, const float* __restrict__ caer,
                                                  float* __restrict__ out, float* __restrict__ outcaer,
                                                  int32_t* __restrict__ outidx) {
    __shared__ uint64_t sk[kSortMax];
    int m = 1;
    while (m < k) m <<= 1;
    const uint64_t got = tk[4];  // == k when the selection is exact
    const int cnt = got < (uint64_t)k ? (int)got : k;
    for (int i = threadIdx.x; i < m; i += 1024) sk[i] = i < cnt ? win[i] : ~0ull;
    __syncthreads();
    for (int size = 2; size <= m; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int i = threadIdx.x; i < m; i += 1024) {
                const int j = i ^ stride;
                if (j > i) {
                    const bool up = (i & size) == 0;
                    const uint64_t a = sk[i], b = sk[j];
                    if ((a > b) == up) {
                        sk[i] = b;
                        sk[j] = a;
                    }
                }
            }
            __syncthreads();
        }
    for (int i = threadIdx.x; i < k; i += 1024) {
        const uint32_t su = (uint32_t)(sk[i] & 0xffffffffu);
        if (su >= (uint32_t)n) {
            outidx[i] = -1;
            continue;
        }
        const int src = (int)su;
        out[3 * i] = hyp[3 * src];
        out[3 * i + 1] = hyp[3 * src + 1];
        out[3 * i + 2] = hyp[3 * src + 2];
        outcaer[i] = caer[src];
        outidx[i] = src;
    }
}

// Global bitonic sort step over m (power of two) composite keys (large-k fallback).
__global__ __launch_bounds__(kTB) void k_bitonic_step(uint64_t* __restrict__ a, int m, int size, int stride) {
    const int i = blockIdx.x * kTB + threadIdx.x;
    if (i >= m) return;
    const int j = i ^ stride;
    if (j > i) {
        const bool up = (i & size) == 0;
        const uint64_t x = a[i], y = a[j];
        if ((x > y) == up) {
            a[i] = y;
            a[j] = x;
        }
    }
}

__global__ __launch_bounds__(kTB) void k_fill_keys(const float* __restrict__ caer, int n, int m,
                                                   uint64_t* __restrict__ a) {
    const int i = blockIdx.x * kTB + threadIdx.x;
    if (i < m) a[i] = i < n ? comp_key(caer[i], i) : ~0ull;
}

__global__ __launch_bounds__(kTB) void k_gather_sorted(const uint64_t* __restrict__ a, int k, int n,
                                                       const float* __restrict__ hyp, const float* __restrict__ caer,
                                                       float* __restrict__ out, float* __restrict__ outcaer,
                                                       int32_t* __restrict__ outidx) {
    const int i = blockIdx.x * kTB + threadIdx.x;
    if (i >= k) return;
    const uint32_t su = (uint32_t)(a[i] & 0xffffffffu);
    if (su >= (uint32_t)n) {
        outidx[i] = -1;
        return;
    }
    const int src = (int)su;
    out[3 * i] = hyp[3 * src];
    out[3 * i + 1] = hyp[3 * src + 1];
    out[3 * i + 2] = hyp[3 * src + 2];
    outcaer[i] = caer[src];
    outidx[i] = src;
}

// ---------------------------------------------------------------- host helpers

void prof_b(GlState* s, int stage) {
    if (!s->prof_on) return;
    GlState::Ev e{stage, nullptr, nullptr};
    (void)hipEventCreate(&e.a);
    (void)hipEventCreate(&e.b);
    (void)hipEventRecord(e.a, s->stream);
    s->prof.push_back(e);
}
void prof_e(GlState* s) {
    if (s->prof_on && !s->prof.empty()) (void)hipEventRecord(s->prof.back().b, s->stream);
}

int ensure_rays(GlState* s, int nrays, float span) {
    if (nrays > s->ray_cap) {
        (void)hipFree(s->d_scan);
        (void)hipFree(s->d_cphi);
        (void)hipFree(s->d_sphi);
        s->d_scan = nullptr;
        s->d_cphi = s->d_sphi = nullptr;
        GL_CHECK(hipMalloc((void**)&s->d_scan, sizeof(float) * (size_t)nrays));
        GL_CHECK(hipMalloc((void**)&s->d_cphi, sizeof(double) * (size_t)nrays));
        GL_CHECK(hipMalloc((void**)&s->d_sphi, sizeof(double) * (size_t)nrays));
        s->ray_cap = nrays;
        s->tab_nrays = -1;
    }
    if (s->tab_nrays != nrays || s->tab_span != span) {
        // ray r at phi = span * r / (nrays - 1) relative to the hypothesis yaw (global_loc.rs:204-209)
        std::vector<double> c((size_t)nrays), sn((size_t)nrays);
        for (int r = 0; r < nrays; ++r) {
            const double phi = nrays > 1 ? (double)span * (double)r / (double)(nrays - 1) : 0.0;
            c[r] = cos(phi);
            sn[r] = sin(phi);
        }
        GL_CHECK(hipMemcpyAsync(s->d_cphi, c.data(), sizeof(double) * (size_t)nrays, hipMemcpyHostToDevice, s->stream));
        GL_CHECK(hipMemcpyAsync(s->d_sphi, sn.data(), sizeof(double) * (size_t)nrays, hipMemcpyHostToDevice, s->stream));
        GL_CHECK(hipStreamSynchronize(s->stream));
        s->tab_nrays = nrays;
        s->tab_span = span;
    }
    return 0;
}

int rays_launch(GlState* s, Grid g, const float* poses, int n, const float* scan, int nrays, float rmax, float* out,
                hipStream_t st);

int ray_diffs_reserve(GlState* s, size_t n) {
    if (n <= s->diff_cap) return 0;
    (void)hipFree(s->d_diff);
    s->d_diff = nullptr;
    s->diff_cap = 0;
    if (hipMalloc((void**)&s->d_diff, sizeof(float) * n) != hipSuccess) return -1;
    s->diff_cap = n;
    return 0;
}

int score_launch(GlState* s, int nrays, float rmax) {
    const bool lds_rays = nrays <= kLdsRaysMax;
    const size_t ray_lds = lds_rays ? (size_t)nrays * (8 + 8 + 4) : 0;
    Grid g = s->g;
    g.bits = s->d_bits;
    g.bytes = s->d_bytes;
    prof_b(s, PFGLX_STAGE_SCORE);
    int mode = s->score_mode != 0 ? s->score_mode : (s->nh < kScoreBlockMaxHyp ? 2 : 1);
    const size_t total = (size_t)s->nh * (size_t)nrays;
    if (mode == 2 && (total > kRayDiffsMax || ray_diffs_reserve(s, total) != 0))
        mode = 1;  // two-pass route unavailable: the wave kernel
    if (mode == 2) {
        if (rays_launch(s, g, s->d_hyp, s->nh, s->d_scan, nrays, rmax, s->d_diff, s->stream) != 0) return -1;
        hipLaunchKernelGGL(k_gl_caer_sum, dim3((s->nh + 3) / 4), dim3(kTB), 0, s->stream, s->d_diff, s->nh, nrays,
                           s->d_caer);
    } else if ((size_t)g.nbytes + ray_lds <= kLdsScoreBytesMax && s->nh >= kScoreBytesMinHyp) {
        // byte grid: 1024-thread blocks, two per CU (enough hypotheses to give every CU two)
        const size_t shm = (size_t)g.nbytes + ray_lds;
        int blocks = (s->nh + 15) / 16;
        blocks = std::max(1, std::min(blocks, 1024));
        if (lds_rays)
            hipLaunchKernelGGL((k_gl_score<1024, true, true, true>), dim3(blocks), dim3(1024), shm, s->stream, g,
                               s->d_hyp, s->nh, s->d_scan, s->d_cphi, s->d_sphi, nrays, rmax, s->d_caer);
        else
            hipLaunchKernelGGL((k_gl_score<1024, true, true, false>), dim3(blocks), dim3(1024), shm, s->stream, g,
                               s->d_hyp, s->nh, s->d_scan, s->d_cphi, s->d_sphi, nrays, rmax, s->d_caer);
    } else {
        const bool lds_grid = s->g.nwords <= kLdsBitsMaxWords;
        const size_t shm = (lds_grid ? 4 * (size_t)((s->g.nwords + 3) & ~3) : 0) + ray_lds;
        int blocks = (s->nh + 3) / 4;
        blocks = std::max(1, std::min(blocks, 4096));
        if (lds_grid && lds_rays)
            hipLaunchKernelGGL((k_gl_score<kTB, false, true, true>), dim3(blocks), dim3(kTB), shm, s->stream, g,
                               s->d_hyp, s->nh, s->d_scan, s->d_cphi, s->d_sphi, nrays, rmax, s->d_caer);
        else if (lds_grid)
            hipLaunchKernelGGL((k_gl_score<kTB, false, true, false>), dim3(blocks), dim3(kTB), shm, s->stream, g,
                               s->d_hyp, s->nh, s->d_scan, s->d_cphi, s->d_sphi, nrays, rmax, s->d_caer);
        else if (lds_rays)
            hipLaunchKernelGGL((k_gl_score<kTB, false, false, true>), dim3(blocks), dim3(kTB), shm, s->stream, g,
                               s->d_hyp, s->nh, s->d_scan, s->d_cphi, s->d_sphi, nrays, rmax, s->d_caer);
        else
            hipLaunchKernelGGL((k_gl_score<kTB, false, false, false>), dim3(blocks), dim3(kTB), shm, s->stream, g,
                               s->d_hyp, s->nh, s->d_scan, s->d_cphi, s->d_sphi, nrays, rmax, s->d_caer);
    }
    prof_e(s);
    GL_LAUNCH_CHECK();
    s->caer_state = 0;
    s->last_m = 0;
    return 0;
}

// Bottom-k for n <= kTopkSmallN, k <= kTopkSmallK in one workgroup: radix select of the
// k-th smallest unique key (CAER bits, index) starting below the bits shared by every key,
// stopping early once the selected digit bucket holds a single key; then the k winners are
// bitonic-sorted in LDS.
constexpr int kTopkSmallN = 1 << 16;
constexpr int kTopkSmallK = 1024;

__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
    const unsigned lo = (unsigned)__shfl_xor((int)(uint32_t)v, m, 64);
    const unsigned hi = (unsigned)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}

__global__ __launch_bounds__(1024) void k_topk_small(const float* __restrict__ caer, int n, int k,
                                                     const float* __restrict__ hyp, float* __restrict__ out,
                                                     float* __restrict__ outcaer, int32_t* __restrict__ outidx) {
    __shared__ unsigned hist[256];
    __shared__ uint64_t red[2][16];
    __shared__ uint64_t win[kTopkSmallK];
    __shared__ uint64_t s_prefix;
    __shared__ int s_d, s_rem, s_done, s_cnt;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint64_t mn = ~0ull, mx = 0;
    for (int i = t; i < n; i += 1024) {
        const uint64_t key = comp_key(caer[i], i);
        mn = key < mn ? key : mn;
        mx = key > mx ? key : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint64_t a = shfl_xor_u64(mn, o), b = shfl_xor_u64(mx, o);
        mn = a < mn ? a : mn;
        mx = b > mx ? b : mx;
    }
    if (lane == 0) {
        red[0][wv] = mn;
        red[1][wv] = mx;
    }
    __syncthreads();
    if (t == 0) {
        uint64_t a = red[0][0], b = red[1][0];
        for (int w = 1; w < 16; ++w) {
            a = red[0][w] < a ? red[0][w] : a;
            b = red[1][w] > b ? red[1][w] : b;
        }
        const int L = a == b ? 64 : __clzll((long long)(a ^ b));
        s_d = L;
        s_prefix = L == 64 ? a : (L == 0 ? 0ull : (a & ~(~0ull >> L)));
        s_rem = k;
        s_done = L == 64 ? 1 : 0;
        s_cnt = 0;
    }
    __syncthreads();
    while (s_done == 0) {
        const int d = s_d, rem = s_rem;
        const uint64_t prefix = s_prefix;
        const int w = 64 - d < 8 ? 64 - d : 8;
        const int shift = 64 - d - w;
        const unsigned dmask = (1u << w) - 1u;
        if (t < 256) hist[t] = 0;
        __syncthreads();
        for (int i = t; i < n; i += 1024) {
            const uint64_t key = comp_key(caer[i], i);
            if (d == 0 || (key >> (64 - d)) == (prefix >> (64 - d))) atomicAdd(&hist[(key >> shift) & dmask], 1u);
        }
        __syncthreads();
        if (wv == 0) {
            unsigned c[4], sum = 0;
            for (int j = 0; j < 4; ++j) {
                c[j] = hist[4 * lane + j];
                sum += c[j];
            }
            unsigned incl = sum;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned y = (unsigned)__shfl_up((int)incl, o, 64);
                if (lane >= o) incl += y;
            }
            const unsigned excl = incl - sum;
            if (excl < (unsigned)rem && (unsigned)rem <= incl) {
                unsigned run = excl;
                int b = 4 * lane + 3;
                unsigned cb = c[3];
                for (int j = 0; j < 4; ++j) {
                    if (run + c[j] >= (unsigned)rem) {
                        b = 4 * lane + j;
                        cb = c[j];
                        break;
                    }
                    run += c[j];
                }
                s_prefix = prefix | ((uint64_t)b << shift);
                s_rem = rem - (int)run;
                s_d = d + w;
                s_done = d + w == 64 ? 1 : (cb == 1 ? 2 : 0);
            }
        }
        __syncthreads();
    }
    if (s_done == 2) {  // the chosen bucket holds exactly one key: find it
        const int d = s_d;
        const uint64_t prefix = s_prefix;
        __syncthreads();
        for (int i = t; i < n; i += 1024) {
            const uint64_t key = comp_key(caer[i], i);
            if ((key >> (64 - d)) == (prefix >> (64 - d))) s_prefix = key;
        }
        __syncthreads();
    }
    const uint64_t kth = s_prefix;
    for (int i = t; i < n; i += 1024) {
        const uint64_t key = comp_key(caer[i], i);
        if (key <= kth) {
            const int slot = atomicAdd(&s_cnt, 1);
            if (slot < kTopkSmallK) win[slot] = key;
        }
    }
    __syncthreads();
    int m = 1;
    while (m < k) m <<= 1;
    const int filled = s_cnt < k ? s_cnt : k;
    for (int i = t; i < m; i += 1024)
        if (i >= filled) win[i] = ~0ull;
    __syncthreads();
    for (int size = 2; size <= m; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            if (t < m) {
                const int p = t ^ stride;
                if (p > t) {
                    const uint64_t a = win[t], b = win[p];
                    const bool up = (t & size) == 0;
                    if ((a > b) == up) {
                        win[t] = b;
                        win[p] = a;
                    }
                }
            }
            __syncthreads();
        }
    if (t < k) {
        const uint64_t key = win[t];
        const uint32_t idx = (uint32_t)key;
        if (idx < (uint32_t)n) {
            out[3 * t] = hyp[3 * idx];
            out[3 * t + 1] = hyp[3 * idx + 1];
            out[3 * t + 2] = hyp[3 * idx + 2];
            outcaer[t] = caer[idx];
            outidx[t] = (int32_t)idx;
        }
    }
}

// Bottom-k of keys[0 .. n) (CAER values, or pruning bounds) by (value bits, index): the winners'
// poses, values and indices into d_out, d_outcaer, d_outidx.
int topk_select(GlState* s, const float* keys, int n, int k) {
    int blocks = (n + kTB - 1) / kTB;
    if (blocks > 1024) blocks = 1024;
    if (n <= kTopkSmallN && k <= kTopkSmallK) {
        hipLaunchKernelGGL(k_topk_small, dim3(1), dim3(1024), 0, s->stream, keys, n, k, s->d_hyp, s->d_out,
                           s->d_outcaer, s->d_outidx);
    } else if (k <= kSortMax) {
        hipLaunchKernelGGL(k_tk_init, dim3(1), dim3(256), 0, s->stream, s->d_tk, s->d_hist, k);
        for (int pass = 0; pass < 8; ++pass) {
            const int shift = 56 - 8 * pass;
            hipLaunchKernelGGL(k_tk_hist, dim3(blocks), dim3(kTB), 0, s->stream, keys, n, s->d_tk, s->d_hist,
                               shift);
            hipLaunchKernelGGL(k_tk_pick, dim3(1), dim3(256), 0, s->stream, s->d_tk, s->d_hist, shift);
        }
        hipLaunchKernelGGL(k_tk_collect, dim3(blocks), dim3(kTB), 0, s->stream, keys, n, s->d_tk, s->d_win, k);
        hipLaunchKernelGGL(k_tk_sort, dim3(1), dim3(1024), 0, s->stream, s->d_win, s->d_tk, k, n, s->d_hyp,
                           keys, s->d_out, s->d_outcaer, s->d_outidx);
    } else {
        int m = 1;
        while (m < n) m <<= 1;
        if (m > s->sort_cap) {
            (void)hipFree(s->d_sortbuf);
            s->d_sortbuf = nullptr;
            GL_CHECK(hipMalloc((void**)&s->d_sortbuf, sizeof(uint64_t) * (size_t)m));
            s->sort_cap = m;
        }
        const int gb = (m + kTB - 1) / kTB;
        hipLaunchKernelGGL(k_fill_keys, dim3(gb), dim3(kTB), 0, s->stream, keys, n, m, s->d_sortbuf);
        for (int size = 2; size <= m; size <<= 1)
            for (int stride = size >> 1; stride > 0; stride >>= 1)
                hipLaunchKernelGGL(k_bitonic_step, dim3(gb), dim3(kTB), 0, s->stream, s->d_sortbuf, m, size, stride);
        hipLaunchKernelGGL(k_gather_sorted, dim3((k + kTB - 1) / kTB), dim3(kTB), 0, s->stream, s->d_sortbuf, k, n,
                           s->d_hyp, keys, s->d_out, s->d_outcaer, s->d_outidx);
    }
    GL_LAUNCH_CHECK();
    return 0;
}

int topk_launch(GlState* s, int k) {
    prof_b(s, PFGLX_STAGE_TOPK);
    if (topk_select(s, s->d_caer, s->nh, k) != 0) return -1;
    prof_e(s);
    s->last_k = k;
    return 0;
}

// ---------------------------------------------------------------- pruned scoring

// Pruning applies, in the automatic kernel choice, to the scoring the large-batch byte-grid
// kernel would run (route 1: the scan and ray tables in LDS) and to the one-lane-per-ray route of
// smaller batches with the byte grid in LDS (route 2, from kPruneRaysMinHyp hypotheses; its
// per-ray errors are kept, so the survivors cast only their remaining rays; below that its
// extra launches cost more than the rays it saves: 10k hypotheses took 0.43 against 0.35 ms).
// 0: not eligible.
constexpr int kPruneRaysMinHyp = 32768;
int prune_route(GlState* s, int nrays) {
    if (!s->prune || s->score_mode != 0 || nrays > kLdsRaysMax || nrays < 128) return 0;
    const size_t ray_lds = (size_t)nrays * (8 + 8 + 4);
    if (s->nh >= kScoreBlockMaxHyp) return (size_t)s->g.nbytes + ray_lds <= kLdsScoreBytesMax ? 1 : 0;
    const size_t total = (size_t)s->nh * (size_t)nrays;
    if (s->nh < kPruneRaysMinHyp || (size_t)s->g.nbytes > kLdsScoreBytesMax || total > kRayDiffsMax ||
        ray_diffs_reserve(s, total) != 0)
        return 0;
    return 2;
}

template <int kSel>
int rays_sel_launch(GlState* s, int nrays, float rmax, const int32_t* list, const int* nlist, int count) {
    Grid g = s->g;
    g.bits = s->d_bits;
    g.bytes = s->d_bytes;
    const size_t per = (size_t)(((nrays + 127) >> 7) << 6);
    const int blocks = (int)std::max<size_t>(1, std::min<size_t>(((size_t)count * per + 1023) / 1024, 2048));
    hipLaunchKernelGGL((k_gl_rays_sel<kSel>), dim3(blocks), dim3(1024), (size_t)g.nbytes, s->stream, g, s->d_hyp,
                       count, list, nlist, s->d_scan, s->d_cphi, s->d_sphi, nrays, rmax, s->d_diff);
    GL_LAUNCH_CHECK();
    return 0;
}

int prune_buffers(GlState* s) {
    if (s->d_lb) return 0;
    GL_CHECK(hipMalloc((void**)&s->d_lb, sizeof(float) * (size_t)s->cap_hyp));
    GL_CHECK(hipMalloc((void**)&s->d_surv, sizeof(int32_t) * (size_t)s->cap_hyp));
    GL_CHECK(hipMalloc((void**)&s->d_nsurv, sizeof(int)));
    GL_CHECK(hipMalloc((void**)&s->d_tau, sizeof(float) * 2));
    GL_CHECK(hipMalloc((void**)&s->d_cand, sizeof(float) * 3 * (size_t)kSortMax));
    GL_CHECK(hipMalloc((void**)&s->d_cdiff, sizeof(float) * (size_t)kSortMax * (size_t)kLdsRaysMax));
    return 0;
}

// The byte-grid scoring kernel over all hypotheses (kSub = 2: bounds into out) or over a list.
// count: the hypotheses (or list entries) to score when nlist is null; else the grid's upper
// bound of the device-side count *nlist.
template <int kSub>
int score_bytes_launch(GlState* s, int nrays, float rmax, float* out, const int32_t* list, const int* nlist, int count,
                       float* mark) {
    Grid g = s->g;
    g.bits = s->d_bits;
    g.bytes = s->d_bytes;
    const size_t shm = (size_t)g.nbytes + (size_t)nrays * (8 + 8 + 4);
    int blocks = (count + 15) / 16;
    blocks = std::max(1, std::min(blocks, 1024));
    hipLaunchKernelGGL((k_gl_score_sel<1024, true, true, true, kSub>), dim3(blocks), dim3(1024), shm, s->stream, g,
                       s->d_hyp, count, s->d_scan, s->d_cphi, s->d_sphi, nrays, rmax, out, list, nlist, mark);
    GL_LAUNCH_CHECK();
    return 0;
}

// First half of a pruned scoring: the bound of every hypothesis.
int bounds_launch(GlState* s, int nrays, float rmax, int route) {
    if (prune_buffers(s) != 0) return -1;
    prof_b(s, PFGLX_STAGE_SCORE);
    if (route == 1) {
        if (score_bytes_launch<2>(s, nrays, rmax, s->d_lb, nullptr, nullptr, s->nh, nullptr) != 0) return -1;
    } else {
        if (rays_sel_launch<1>(s, nrays, rmax, nullptr, nullptr, s->nh) != 0) return -1;
        hipLaunchKernelGGL(k_gl_lb, dim3((s->nh + 3) / 4), dim3(kTB), 0, s->stream, s->d_diff, s->nh, nrays, s->d_lb);
        GL_LAUNCH_CHECK();
    }
    prof_e(s);
    s->caer_state = 1;
    s->pend_route = route;
    s->pend_nrays = nrays;
    s->pend_rmax = rmax;
    return 0;
}

// Second half, for a bottom-k: the m = 4k hypotheses with the smallest bounds are scored in
// full, the k-th smallest of their CAERs is the threshold tau, and of the rest only those whose
// bound does not exceed tau are scored; the others cannot be among the k smallest and get +inf.
// Every hypothesis that can be in the bottom k then holds its exact CAER, so the bottom k (ties
// by index included) is the one full scoring gives.
int finish_pruned(GlState* s, int k) {
    const int n = s->nh, nrays = s->pend_nrays;
    const float rmax = s->pend_rmax;
    const int m = std::min(std::max(4 * k, 64), kSortMax);
    if (m * 2 > n || k > m) return score_launch(s, nrays, rmax);  // nothing to gain: score everything
    prof_b(s, PFGLX_STAGE_SCORE);
    if (topk_select(s, s->d_lb, n, m) != 0) return -1;  // candidates -> d_outidx[0 .. m)
    if (s->pend_route == 2) {
        // the per-ray errors of the even blocks are kept: the candidates, then the survivors,
        // cast only the odd blocks' rays, and each complete row is summed in the standard order
        if (rays_sel_launch<2>(s, nrays, rmax, s->d_outidx, nullptr, m) != 0) return -1;
        hipLaunchKernelGGL(k_gl_caer_rows, dim3((m + 3) / 4), dim3(kTB), 0, s->stream, s->d_diff, s->d_outidx,
                           nullptr, m, nrays, s->d_caer, s->d_lb);
        hipLaunchKernelGGL(k_gl_tau, dim3(1), dim3(1024), 0, s->stream, s->d_caer, s->d_outidx, m, k, s->d_tau,
                           s->d_nsurv);
        hipLaunchKernelGGL(k_gl_prune, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, s->stream, s->d_lb, n, s->d_tau,
                           s->d_caer, s->d_surv, s->d_nsurv);
        GL_LAUNCH_CHECK();
        if (rays_sel_launch<2>(s, nrays, rmax, s->d_surv, s->d_nsurv, n) != 0) return -1;
        hipLaunchKernelGGL(k_gl_caer_rows, dim3(std::max(1, std::min((n + 3) / 4, 4096))), dim3(kTB), 0, s->stream,
                           s->d_diff, s->d_surv, s->d_nsurv, n, nrays, s->d_caer, nullptr);
        GL_LAUNCH_CHECK();
        prof_e(s);
        s->caer_state = 2;
        s->pruned_k = k;
        s->last_m = m;
        s->last_route = 2;
        return 0;
    }
    // the candidates one lane per ray (few hypotheses: one wave each would leave most of the
    // chip idle), their errors summed in k_gl_score's order (the same CAER bits)
    hipLaunchKernelGGL(k_gl_pick, dim3((m + kTB - 1) / kTB), dim3(kTB), 0, s->stream, s->d_hyp, s->d_outidx, m,
                       s->d_cand);
    GL_LAUNCH_CHECK();
    Grid g = s->g;
    g.bits = s->d_bits;
    g.bytes = s->d_bytes;
    if (rays_launch(s, g, s->d_cand, m, s->d_scan, nrays, rmax, s->d_cdiff, s->stream) != 0) return -1;
    hipLaunchKernelGGL(k_gl_caer_sum, dim3((m + 3) / 4), dim3(kTB), 0, s->stream, s->d_cdiff, m, nrays, s->d_caer,
                       s->d_outidx, s->d_lb);
    hipLaunchKernelGGL(k_gl_tau, dim3(1), dim3(1024), 0, s->stream, s->d_caer, s->d_outidx, m, k, s->d_tau,
                       s->d_nsurv);
    hipLaunchKernelGGL(k_gl_prune, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, s->stream, s->d_lb, n, s->d_tau,
                       s->d_caer, s->d_surv, s->d_nsurv);
    GL_LAUNCH_CHECK();
    if (score_bytes_launch<1>(s, nrays, rmax, s->d_caer, s->d_surv, s->d_nsurv, n, nullptr) != 0) return -1;
    prof_e(s);
    s->caer_state = 2;
    s->pruned_k = k;
    s->last_m = m;
    s->last_route = 1;
    return 0;
}

// Before a bottom-k of k: finish a pending pruned scoring, or score everything again when the
// pruned one was for a smaller k.
int caer_for_topk(GlState* s, int k) {
    if (s->caer_state == 1) return finish_pruned(s, k);
    if (s->caer_state == 2 && k > s->pruned_k) return score_launch(s, s->pend_nrays, s->pend_rmax);
    return 0;
}

// Before anything reads every CAER: a pruned scoring is completed in full.
int caer_all(GlState* s) {
    if (s->caer_state == 0) return 0;
    return score_launch(s, s->pend_nrays, s->pend_rmax);
}

// Local refinement: `per` children per parent (the last top-k rows); child 0 is the parent
// itself, the others add N(0, sxy^2) to x, y and N(0, syaw^2) to yaw (wrapped to [-pi, pi)),
// clamped to the map rectangle.
__global__ __launch_bounds__(kTB) void k_gl_refine(const float* __restrict__ parents, int k, int per, float sxy,
                                                   float syaw, uint64_t seed, float x0, float y0, float x1, float y1,
                                                   float* __restrict__ hyp) {
    const int h = blockIdx.x * kTB + threadIdx.x;
    if (h >= k * per) return;
    const int p = h / per, j = h - p * per;
    float x = parents[3 * p], y = parents[3 * p + 1], t = parents[3 * p + 2];
    if (j > 0) {
        uint64_t st = sm_final(seed * kPhi + (uint64_t)h * kC1);
        const uint64_t z0 = sm_next(st), z1 = sm_next(st), z2 = sm_next(st), z3 = sm_next(st);
        const float u1 = ((float)(z0 >> 40) + 1.0f) * 0x1.0p-24f, u2 = (float)(z1 >> 40) * 0x1.0p-24f;
        const float u3 = ((float)(z2 >> 40) + 1.0f) * 0x1.0p-24f, u4 = (float)(z3 >> 40) * 0x1.0p-24f;
        const float r1 = sqrtf(-2.0f * logf(u1)), r2 = sqrtf(-2.0f * logf(u3));
        x = fminf(fmaxf(x + sxy * r1 * cosf(6.2831853f * u2), x0), x1);
        y = fminf(fmaxf(y + sxy * r1 * sinf(6.2831853f * u2), y0), y1);
        t += syaw * r2 * cosf(6.2831853f * u4);
        if (t >= 3.14159265f) t -= 6.28318531f;
        if (t < -3.14159265f) t += 6.28318531f;
    }
    hyp[3 * h] = x;
    hyp[3 * h + 1] = y;
    hyp[3 * h + 2] = t;
}

// One-lane-per-ray pass over n poses (k_gl_rays): scan == nullptr renders ranges, otherwise
// writes |scan - range|.  The byte grid goes to LDS when it fits, else the bits.
int rays_launch(GlState* s, Grid g, const float* poses, int n, const float* scan, int nrays, float rmax, float* out,
                hipStream_t st) {
    g.bits = s->d_bits;
    g.bytes = s->d_bytes;
    const size_t total = (size_t)n * (size_t)nrays;
    const double* cp = s->d_cphi;
    const double* sp = s->d_sphi;
    if ((size_t)g.nbytes <= kLdsScoreBytesMax) {
        const int blocks = (int)std::max<size_t>(1, std::min<size_t>((total + 1023) / 1024, 2048));
        if (scan)
            hipLaunchKernelGGL((k_gl_rays<1024, 0, true>), dim3(blocks), dim3(1024), (size_t)g.nbytes, st, g, poses, n,
                               scan, cp, sp, nrays, rmax, out);
        else
            hipLaunchKernelGGL((k_gl_rays<1024, 0, false>), dim3(blocks), dim3(1024), (size_t)g.nbytes, st, g, poses,
                               n, scan, cp, sp, nrays, rmax, out);
    } else {
        const int blocks = (int)std::max<size_t>(1, std::min<size_t>((total + kTB - 1) / kTB, 16384));
        const bool lds = g.nwords <= kLdsBitsMaxWords;
        const size_t shm = lds ? 4 * (size_t)((g.nwords + 3) & ~3) : 0;
        if (lds && scan)
            hipLaunchKernelGGL((k_gl_rays<kTB, 1, true>), dim3(blocks), dim3(kTB), shm, st, g, poses, n, scan, cp, sp,
                               nrays, rmax, out);
        else if (lds)
            hipLaunchKernelGGL((k_gl_rays<kTB, 1, false>), dim3(blocks), dim3(kTB), shm, st, g, poses, n, scan, cp,
                               sp, nrays, rmax, out);
        else if (scan)
            hipLaunchKernelGGL((k_gl_rays<kTB, 2, true>), dim3(blocks), dim3(kTB), 0, st, g, poses, n, scan, cp, sp,
                               nrays, rmax, out);
        else
            hipLaunchKernelGGL((k_gl_rays<kTB, 2, false>), dim3(blocks), dim3(kTB), 0, st, g, poses, n, scan, cp, sp,
                               nrays, rmax, out);
    }
    GL_LAUNCH_CHECK();
    return 0;
}

int render_launch(GlState* s, const Grid& g, const float* poses, int n, int nrays, float rmax, float* out,
                  hipStream_t st, int /*unused*/) {
    return rays_launch(s, g, poses, n, nullptr, nrays, rmax, out, st);
}

int upload_scan(GlState* s, const float* scan, int nrays, float span) {
    if (ensure_rays(s, nrays, span) != 0) return -1;
    if (s->scan_ev_live) GL_CHECK(hipEventSynchronize(s->scan_ev));  // staging buffer free again
    if (nrays > s->h_scan_cap) {
        if (s->h_scan) GL_CHECK(hipHostFree(s->h_scan));
        s->h_scan = nullptr;
        GL_CHECK(hipHostMalloc((void**)&s->h_scan, sizeof(float) * (size_t)nrays, hipHostMallocDefault));
        s->h_scan_cap = nrays;
    }
    memcpy(s->h_scan, scan, sizeof(float) * (size_t)nrays);
    GL_CHECK(hipMemcpyAsync(s->d_scan, s->h_scan, sizeof(float) * (size_t)nrays, hipMemcpyHostToDevice, s->stream));
    GL_CHECK(hipEventRecord(s->scan_ev, s->stream));
    s->scan_ev_live = true;
    s->scan_nrays = nrays;
    return 0;
}

}  // namespace glk

using namespace glk;

extern "C" {

void* pfgl_create(int W, int H, float res, float ox, float oy, const unsigned char* occ, int cap_hyp) {
    if (W <= 0 || H <= 0 || cap_hyp <= 0 || !occ || (long long)(W + 2) * (H + 2) > (1ll << 30)) return nullptr;
    GlState* s = new (std::nothrow) GlState();
    if (!s) return nullptr;
    const long long ncell = (long long)W * H;
    const long long Wp = W + 2, Hp = H + 2;  // framed grid (frame cells set)
    const int nwords = (int)((Wp * Hp + 31) / 32);
    const int nbytes = (int)((Wp * Hp + 15) / 16 * 16);
    std::vector<uint32_t> bits((size_t)nwords, 0u);
    std::vector<uint8_t> bytes((size_t)nbytes, 0u);
    std::vector<int32_t> freexy;
    freexy.reserve((size_t)ncell * 2);
    for (long long y = 0; y < Hp; ++y)
        for (long long x = 0; x < Wp; ++x) {
            const bool frame = x == 0 || y == 0 || x == Wp - 1 || y == Hp - 1;
            if (frame || occ[(y - 1) * W + (x - 1)]) {
                const long long i = y * Wp + x;
                bits[(size_t)(i >> 5)] |= 1u << (i & 31);
                bytes[(size_t)i] = frame ? 2 : 1;
            }
        }
    for (long long i = 0; i < ncell; ++i)
        if (!occ[i]) {
            freexy.push_back((int32_t)(i % W));
            freexy.push_back((int32_t)(i / W));
        }
    s->nfree = (int)(freexy.size() / 2);
    s->cap_hyp = cap_hyp;
    s->g.W = W;
    s->g.H = H;
    s->g.nwords = nwords;
    s->g.nbytes = nbytes;
    s->g.res = res;
    s->g.ox = ox;
    s->g.oy = oy;
    bool ok = hipSetDevice(0) == hipSuccess && hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&s->scan_ev, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_bits, sizeof(uint32_t) * (size_t)nwords) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_bytes, (size_t)nbytes) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_free, sizeof(int32_t) * (freexy.size() + 2)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_hyp, sizeof(float) * 3 * (size_t)cap_hyp) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_caer, sizeof(float) * (size_t)cap_hyp) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_tk, sizeof(uint64_t) * 8) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_hist, sizeof(unsigned) * 256) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_win, sizeof(uint64_t) * (size_t)std::min(cap_hyp, kSortMax)) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_out, sizeof(float) * 3 * (size_t)cap_hyp) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_outcaer, sizeof(float) * (size_t)cap_hyp) == hipSuccess;
    ok = ok && hipMalloc((void**)&s->d_outidx, sizeof(int32_t) * (size_t)cap_hyp) == hipSuccess;
    ok = ok && hipMemcpy(s->d_bits, bits.data(), sizeof(uint32_t) * (size_t)nwords, hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemcpy(s->d_bytes, bytes.data(), (size_t)nbytes, hipMemcpyHostToDevice) == hipSuccess;
    if (ok && !freexy.empty())
        ok = hipMemcpy(s->d_free, freexy.data(), sizeof(int32_t) * freexy.size(), hipMemcpyHostToDevice) == hipSuccess;
    ok = ok && hipMemset(s->d_hist, 0, sizeof(unsigned) * 256) == hipSuccess;
    if (!ok) {
        fprintf(stderr, "pf_hip pfgl_create: device setup failed\n");
        pfgl_destroy(s);
        return nullptr;
    }
    return s;
}

int pfgl_destroy(void* h) {
    GlState* s = static_cast<GlState*>(h);
    if (!s) return 0;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    void* ptrs[] = {s->d_bits, s->d_bytes, s->d_free,    s->d_hyp,      s->d_caer,   s->d_scan, s->d_cphi,
                    s->d_sphi, s->d_tk,      s->d_hist,     s->d_win,    s->d_sortbuf, s->d_out,
                    s->d_outcaer, s->d_outidx, s->d_poses, s->d_ranges, s->d_diff, s->d_lb, s->d_surv,
                    s->d_nsurv, s->d_tau, s->d_cand, s->d_cdiff};
    for (void* p : ptrs) (void)hipFree(p);
    if (s->h_scan) (void)hipHostFree(s->h_scan);
    if (s->scan_ev) (void)hipEventDestroy(s->scan_ev);
    for (auto& e : s->prof) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return 0;
}

int pfgl_free_cells(void* h) {
    GlState* s = static_cast<GlState*>(h);
    return s ? s->nfree : -1;
}

int pfgl_generate(void* h, int n, unsigned long long seed) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || s->nfree <= 0) return -1;
    if (n <= 0 || n > s->cap_hyp) return -1;
    s->nh = n;
    s->caer_state = 0;
    prof_b(s, PFGLX_STAGE_GENERATE);
    hipLaunchKernelGGL(k_gl_gen, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, s->stream, s->d_free, s->nfree, s->g.res,
                       s->g.ox, s->g.oy, (uint64_t)seed, n, s->d_hyp);
    prof_e(s);
    GL_LAUNCH_CHECK();
    return 0;
}

int pfgl_score(void* h, const float* real_scan, int nrays, float rmax, float ang_span) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || s->nh <= 0 || !real_scan || nrays <= 0) return -1;
    if (upload_scan(s, real_scan, nrays, ang_span) != 0) return -1;
    if (const int route = prune_route(s, nrays)) {  // bounds now, the rest with the bottom-k
        if (bounds_launch(s, nrays, rmax, route) != 0) return -1;
    } else if (score_launch(s, nrays, rmax) != 0) {
        return -1;
    }
    GL_CHECK(hipStreamSynchronize(s->stream));
    return 0;
}

int pfglx_set_scan(void* h, const float* scan, int nrays, float ang_span) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || !scan || nrays <= 0) return -1;
    s->caer_state = 0;  // (a pending pruned scoring was for the previous scan)
    return upload_scan(s, scan, nrays, ang_span);
}

int pfglx_score_async(void* h, int nrays, float rmax) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || s->nh <= 0 || nrays <= 0 || nrays != s->scan_nrays) return -1;
    if (const int route = prune_route(s, nrays)) return bounds_launch(s, nrays, rmax, route);
    return score_launch(s, nrays, rmax);
}

int pfglx_topk_ex(void* h, int k, float* out3k, float* caer_k, int32_t* idx_k) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || k <= 0 || s->nh <= 0) return -1;
    if (k > s->nh) k = s->nh;
    if (caer_for_topk(s, k) != 0 || topk_launch(s, k) != 0) return -1;
    if (out3k) GL_CHECK(hipMemcpyAsync(out3k, s->d_out, sizeof(float) * 3 * (size_t)k, hipMemcpyDeviceToHost, s->stream));
    if (caer_k)
        GL_CHECK(hipMemcpyAsync(caer_k, s->d_outcaer, sizeof(float) * (size_t)k, hipMemcpyDeviceToHost, s->stream));
    if (idx_k)
        GL_CHECK(hipMemcpyAsync(idx_k, s->d_outidx, sizeof(int32_t) * (size_t)k, hipMemcpyDeviceToHost, s->stream));
    GL_CHECK(hipStreamSynchronize(s->stream));
    return k;
}

int pfgl_topk(void* h, int k, float* out) {
    if (!out) return -1;
    return pfglx_topk_ex(h, k, out, nullptr, nullptr);
}

int pfglx_set_hypotheses(void* h, const float* hyp3n, int n) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || !hyp3n || n <= 0 || n > s->cap_hyp) return -1;
    GL_CHECK(hipMemcpyAsync(s->d_hyp, hyp3n, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, s->stream));
    GL_CHECK(hipStreamSynchronize(s->stream));
    s->nh = n;
    s->caer_state = 0;
    return 0;
}

int pfglx_get_hypotheses(void* h, float* hyp3n) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || !hyp3n || s->nh <= 0) return -1;
    GL_CHECK(hipMemcpyAsync(hyp3n, s->d_hyp, sizeof(float) * 3 * (size_t)s->nh, hipMemcpyDeviceToHost, s->stream));
    GL_CHECK(hipStreamSynchronize(s->stream));
    return s->nh;
}

int pfglx_get_caer(void* h, float* caer) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || !caer || s->nh <= 0) return -1;
    if (caer_all(s) != 0) return -1;
    GL_CHECK(hipMemcpyAsync(caer, s->d_caer, sizeof(float) * (size_t)s->nh, hipMemcpyDeviceToHost, s->stream));
    GL_CHECK(hipStreamSynchronize(s->stream));
    return s->nh;
}

int pfglx_hypothesis_count(void* h) {
    GlState* s = static_cast<GlState*>(h);
    return s ? s->nh : -1;
}

int pfglx_render(void* h, const float* poses3n, int n, int nrays, float rmax, float ang_span, float* out) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || !poses3n || !out || n <= 0 || nrays <= 0) return -1;
    if (ensure_rays(s, nrays, ang_span) != 0) return -1;
    const size_t need = (size_t)n;
    if (need > s->render_poses_cap) {
        (void)hipFree(s->d_poses);
        s->d_poses = nullptr;
        GL_CHECK(hipMalloc((void**)&s->d_poses, sizeof(float) * 3 * need));
        s->render_poses_cap = need;
    }
    if (need * (size_t)nrays > s->render_ranges_cap) {
        (void)hipFree(s->d_ranges);
        s->d_ranges = nullptr;
        GL_CHECK(hipMalloc((void**)&s->d_ranges, sizeof(float) * need * (size_t)nrays));
        s->render_ranges_cap = need * (size_t)nrays;
    }
    GL_CHECK(hipMemcpyAsync(s->d_poses, poses3n, sizeof(float) * 3 * need, hipMemcpyHostToDevice, s->stream));
    Grid g = s->g;
    g.bits = s->d_bits;
    int blocks = (n + 3) / 4;
    if (blocks > 8192) blocks = 8192;
    if (render_launch(s, g, s->d_poses, n, nrays, rmax, s->d_ranges, s->stream, blocks) != 0) return -1;
    GL_CHECK(hipMemcpyAsync(out, s->d_ranges, sizeof(float) * need * (size_t)nrays, hipMemcpyDeviceToHost, s->stream));
    GL_CHECK(hipStreamSynchronize(s->stream));
    return 0;
}

// Device-resident render for callers that keep data on the GPU (training-set generation).
int pfglx_render_device(void* h, const float* d_poses3n, int n, int nrays, float rmax, float ang_span, float* d_out,
                        void* stream) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || !d_poses3n || !d_out || n <= 0 || nrays <= 0) return -1;
    if (ensure_rays(s, nrays, ang_span) != 0) return -1;
    Grid g = s->g;
    g.bits = s->d_bits;
    int blocks = (n + 3) / 4;
    if (blocks > 8192) blocks = 8192;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
    if (render_launch(s, g, d_poses3n, n, nrays, rmax, d_out, st, blocks) != 0) return -1;
    return 0;
}

int pfglx_sample_poses_device(void* h, int n, uint64_t seed, float* d_out, void* stream) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || s->nfree <= 0 || n <= 0 || !d_out) return -1;
    hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
    hipLaunchKernelGGL(k_gl_gen, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, st, s->d_free, s->nfree, s->g.res, s->g.ox,
                       s->g.oy, seed, n, d_out);
    GL_LAUNCH_CHECK();
    return 0;
}

int pfglx_generate_range(void* h, int offset, int n, uint64_t seed) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || s->nfree <= 0 || offset < 0 || n < 0 || offset + n > s->cap_hyp) return -1;
    if (n > 0) {
        hipLaunchKernelGGL(k_gl_gen, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, s->stream, s->d_free, s->nfree,
                           s->g.res, s->g.ox, s->g.oy, seed, n, s->d_hyp + 3 * (size_t)offset);
        GL_LAUNCH_CHECK();
    }
    s->nh = offset + n;
    s->caer_state = 0;
    return 0;
}

int pfglx_topk_device(void* h, int k) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || k <= 0 || s->nh <= 0) return -1;
    if (k > s->nh) k = s->nh;
    if (caer_for_topk(s, k) != 0 || topk_launch(s, k) != 0) return -1;
    return k;
}

int pfglx_refine(void* h, int k_parents, int per_parent, float sigma_xy, float sigma_yaw, uint64_t seed) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || k_parents <= 0 || per_parent <= 0 || k_parents > s->last_k) return -1;
    if ((int64_t)k_parents * per_parent > s->cap_hyp) return -1;
    const int n = k_parents * per_parent;
    const float x0 = s->g.ox, y0 = s->g.oy;
    const float x1 = s->g.ox + s->g.W * s->g.res - 1e-4f, y1 = s->g.oy + s->g.H * s->g.res - 1e-4f;
    hipLaunchKernelGGL(k_gl_refine, dim3((n + kTB - 1) / kTB), dim3(kTB), 0, s->stream, s->d_out, k_parents,
                       per_parent, sigma_xy, sigma_yaw, seed, x0, y0, x1, y1, s->d_hyp);
    GL_LAUNCH_CHECK();
    s->nh = n;
    s->caer_state = 0;
    return 0;
}

float* pfglx_hypotheses_device(void* h) {
    GlState* s = static_cast<GlState*>(h);
    return s ? s->d_hyp : nullptr;
}

int pfglx_map_info(void* h, double* out6) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || !out6) return -1;
    out6[0] = s->g.W;
    out6[1] = s->g.H;
    out6[2] = s->g.res;
    out6[3] = s->g.ox;
    out6[4] = s->g.oy;
    out6[5] = s->nfree;
    return 0;
}

int pfglx_set_count(void* h, int n) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || n < 0 || n > s->cap_hyp) return -1;
    s->nh = n;
    s->caer_state = 0;
    return 0;
}

void* pfglx_stream(void* h) {
    GlState* s = static_cast<GlState*>(h);
    return s ? (void*)s->stream : nullptr;
}

int pfglx_prune_stats(void* h, int* out3) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || !out3) return -1;
    out3[0] = s->last_m;
    out3[1] = 0;
    out3[2] = s->last_route;
    if (s->last_m > 0) {
        GL_CHECK(hipMemcpyAsync(out3 + 1, s->d_nsurv, sizeof(int), hipMemcpyDeviceToHost, s->stream));
        GL_CHECK(hipStreamSynchronize(s->stream));
    }
    return 0;
}

int pfglx_set_prune(void* h, int enable) {
    GlState* s = static_cast<GlState*>(h);
    if (!s) return -1;
    s->prune = enable != 0;
    return 0;
}

int pfglx_set_score_mode(void* h, int mode) {
    GlState* s = static_cast<GlState*>(h);
    if (!s || mode < 0 || mode > 2) return -1;
    s->score_mode = mode;
    return 0;
}

int pfglx_profile_enable(void* h, int enable) {
    GlState* s = static_cast<GlState*>(h);
    if (!s) return -1;
    s->prof_on = enable != 0;
    return 0;
}

int pfglx_profile_read(void* h, double* ms, int* count) {
    GlState* s = static_cast<GlState*>(h);
    if (!s) return -1;
    GL_CHECK(hipStreamSynchronize(s->stream));
    for (int k = 0; k < PFGLX_STAGE_COUNT; ++k) {
        if (ms) ms[k] = 0.0;
        if (count) count[k] = 0;
    }
    for (auto& e : s->prof) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, e.a, e.b) == hipSuccess) {
            if (ms) ms[e.stage] += t;
            if (count) count[e.stage] += 1;
        }
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    s->prof.clear();
    return 0;
}

int pfglx_sync(void* h) {
    GlState* s = static_cast<GlState*>(h);
    if (!s) return -1;
    GL_CHECK(hipStreamSynchronize(s->stream));
    return 0;
}

}  // extern "C"

// pf_kernels.hip -- MI355X (gfx950) particle-filter kernels and the pfc_* / pfx_* C ABI.
//
// Drop-in for the reference library src/cuda/pf_kernels.cu (rus1ru/cuda_pf_localization):
// the 12 pfc_* entry points keep its signatures and semantics (bound by
// crates/pf_core/src/cuda.rs:33-88); the kernels are a new design for CDNA4:
//
//   * SoA particle store: seven f32 pose streams (px py pz qw qx qy qz) in two ping-pong
//     sets, f32 weights, f64 log-likelihoods.  Filters are laid out back to back at
//     1024-particle alignment, so one launch serves one filter or a batch of filters.
//   * k_motion: reinit/predict with counter-based splitmix64 draws and an fma-only
//     Box-Muller (no RNG state in memory), pose algebra in f64, f32 storage.
//   * k_weight: one particle per lane (or 2 / 4), observation+landmark records read as
//     wave-uniform scalar loads, residuals in f64 (30 FP64 instructions per pair), online-
//     softmax block partials reduced in a fixed order by a one-workgroup-per-filter finalize.
//   * resampling: fixed-point (2^-56) weights, a single-pass integer scan with decoupled
//     look-back (bounded waits; an unpublished predecessor is summed by the waiting
//     workgroup), exact 128-bit pointer comparisons, ping-pong gather (no copies),
//     deterministic injection (per-target pick lists summed in pick order).
//   * estimate: wave64 butterfly reductions into f64 block partials, a fixed-order
//     one-workgroup-per-filter finalize, and (fused cycle) the Markley mean on the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/pf_kernels.h"
#include "pf_batch.h"
#include "pf_device.h"

using namespace pfd;

#define PF_CHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "pf_hip %s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

#define PF_LAUNCH_CHECK()                                                                        \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "pf_hip %s:%d launch failed: %s\n", __FILE__, __LINE__,              \
                    hipGetErrorString(e_));                                                      \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

namespace pfk {

// Read-only tables and parameter blocks indexed by a wave-uniform value are read through the
// constant address space, so the loads become scalar (s_load) reads through the scalar cache,
// off the vector memory path and out of the vector registers.
#if defined(__HIP_DEVICE_COMPILE__)
template <class T>
using ConstPtr = const T __attribute__((address_space(4)))*;
#else
template <class T>
using ConstPtr = const T*;  // host pass of the single-source compile: never executed
#endif

// Filter of layout block b (256 particles).  A single-filter layout needs no lookup, so a
// kernel's first loads (its filter's record, sizes and parameters) do not wait for this one.
__device__ inline int layout_filter(const Layout& L, int b) {
    if (L.nf == 1) return 0;
    return ((ConstPtr<int>)L.blk_filter)[b];
}

// Parameter block of filter f (wave-uniform): from the per-filter device array or the launch
// argument.
template <class T>
__device__ inline T uniform_params(const ParamRef<T>& A, int f) {
    if (A.dev) return ((ConstPtr<T>)A.dev)[f];
    return A.val;
}

// ============================================================================ kernels

// Reinit (cpu.rs:180-195) or predict (cpu.rs:197-218) of local particle li stored at gi: the
// new pose, rounded to f32, in v (x y z w qx qy qz).  Nothing is stored here.
__device__ inline void motion_pose(const MotionParams& p, const PoseSet& S, int gi, int li, int reinit,
                                   float (&v)[7]) {
    double nn[6];
    particle_normals(p.gseed, (uint64_t)li, nn);
    dquat q, qn;
    double x, y, z;
    if (reinit) {
        x = __fma_rn(p.s_t[0], nn[0], p.t[0]);
        y = __fma_rn(p.s_t[1], nn[1], p.t[1]);
        z = __fma_rn(p.s_t[2], nn[2], p.t[2]);
        q = dquat{p.q[0], p.q[1], p.q[2], p.q[3]};
        qn = qmul(q, noise_dq(p.s_r[0], p.s_r[1], p.s_r[2], nn + 3));
    } else {
        q = dquat{(double)S.c[3][gi], (double)S.c[4][gi], (double)S.c[5][gi], (double)S.c[6][gi]};
        const double bx = __fma_rn(p.s_t[0], nn[0], p.t[0]);
        const double by = __fma_rn(p.s_t[1], nn[1], p.t[1]);
        const double bz = __fma_rn(p.s_t[2], nn[2], p.t[2]);
        double wx, wy, wz;
        qrot(q, bx, by, bz, wx, wy, wz);
        x = (double)S.c[0][gi] + wx;
        y = (double)S.c[1][gi] + wy;
        z = (double)S.c[2][gi] + wz;
        const dquat qo{p.q[0], p.q[1], p.q[2], p.q[3]};
        qn = qmul(qmul(q, qo), noise_dq(p.s_r[0], p.s_r[1], p.s_r[2], nn + 3));
    }
    v[0] = (float)x;
    v[1] = (float)y;
    v[2] = (float)z;
    v[3] = (float)qn.w;
    v[4] = (float)qn.x;
    v[5] = (float)qn.y;
    v[6] = (float)qn.z;
}

// Reinit (reinit = 1) and predict, one particle per lane; the filter, its layout entries and its
// parameter block are wave-uniform scalar loads.  dst_flip = 1: the new poses are stored in the
// idle pose set (the current set is read and left as it is).
template <int reinit>
__global__ __launch_bounds__(kTB) void k_motion(Layout L, PoseBufs B, FilterDev* F, float* __restrict__ w,
                                                ParamRef<MotionParams> A, int dst_flip) {
    const int blk = blockIdx.x;
    const int f = layout_filter(L, blk);
    const int off = ((ConstPtr<int>)L.f_off)[f];
    const int li = blk * kTB + (int)threadIdx.x - off;
    if (li >= ((ConstPtr<int>)L.f_n)[f]) return;
    const int gi = off + li;
    const MotionParams p = uniform_params(A, f);
    const int par = ((ConstPtr<FilterDev>)F)[f].parity;
    const PoseSet S = B.s[par];
    const PoseSet D = B.s[par ^ dst_flip];
    float v[7];
    motion_pose(p, S, gi, li, reinit, v);
    if (li == 0) F[f].est_src = 0;  // poses change: partials of an earlier gather are stale
#pragma unroll
    for (int c = 0; c < 7; ++c) D.c[c][gi] = v[c];
    if (reinit) w[gi] = p.winv;
}

// Block-wide f64 reductions (4 waves): deterministic tree order.
__device__ inline double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wv] = v;
    __syncthreads();
    return (sh[0] + sh[1]) + (sh[2] + sh[3]);
}
__device__ inline double block_max(double v, double* sh) {
    v = wave_max(v);
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[wv] = v;
    __syncthreads();
    return fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// The same reductions over a 1024-thread block (16 waves), in a fixed sequential wave order.
constexpr int kFinTB = 1024;
__device__ inline double block16_sum(double v, double* sh) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int k = 1; k < kFinTB / 64; ++k) r += sh[k];
    return r;
}
__device__ inline double block16_max(double v, double* sh) {
    v = wave_max(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = sh[0];
#pragma unroll
    for (int k = 1; k < kFinTB / 64; ++k) r = fmax(r, sh[k]);
    return r;
}

// Observation records are read through the constant address space: the addresses are
// wave-uniform, so they become scalar (s_load) reads that stay off the vector memory path.
#if defined(__HIP_DEVICE_COMPILE__)
typedef const ObsRec __attribute__((address_space(4)))* ConstObsPtr;
#else
typedef const ObsRec* ConstObsPtr;  // host pass of the single-source compile: never executed
#endif

// Per-particle body map: world -> body rotation qrot(conj(q), .) as a 3x3 matrix (exact
// for non-unit q as well), plus the position.
struct BodyMap {
    double px, py, pz;
    double m00, m01, m02, m10, m11, m12, m20, m21, m22;
    double mpx, mpy, mpz;  // M p, so that b = M l - M p takes three FMAs per component
};

__device__ inline BodyMap body_map(const float (&v)[7]) {
    BodyMap b;
    b.px = (double)v[0];
    b.py = (double)v[1];
    b.pz = (double)v[2];
    const double qw = (double)v[3], qx = (double)v[4], qy = (double)v[5], qz = (double)v[6];
    const double a = 1.0 - 2.0 * (qx * qx + qy * qy + qz * qz);
    b.m00 = a + 2.0 * qx * qx;
    b.m11 = a + 2.0 * qy * qy;
    b.m22 = a + 2.0 * qz * qz;
    b.m01 = 2.0 * (qx * qy + qw * qz);
    b.m10 = 2.0 * (qx * qy - qw * qz);
    b.m02 = 2.0 * (qx * qz - qw * qy);
    b.m20 = 2.0 * (qx * qz + qw * qy);
    b.m12 = 2.0 * (qy * qz + qw * qx);
    b.m21 = 2.0 * (qy * qz - qw * qx);
    b.mpx = __fma_rn(b.m00, b.px, __fma_rn(b.m01, b.py, b.m02 * b.pz));
    b.mpy = __fma_rn(b.m10, b.px, __fma_rn(b.m11, b.py, b.m12 * b.pz));
    b.mpz = __fma_rn(b.m20, b.px, __fma_rn(b.m21, b.py, b.m22 * b.pz));
    return b;
}
__device__ inline BodyMap body_map(const PoseSet& S, int gi) {
    float v[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) v[c] = S.c[c][gi];
    return body_map(v);
}

// acos(1 - u)^2 as a power series in u = 1 - dot: acos(1-u)^2 = sum_n 2^(n+1) u^n / (n^2 C(2n,n))
// (from asin(x)^2 = 1/2 sum (2x)^(2n) / (n^2 C(2n,n)) with x^2 = u/2).  Six terms hold to
// ~1e-18 relative for u <= 0.005, twelve for u <= 0.125.
// Three terms for the gated fast path (u <= u_gate <= 0.005): the first omitted term,
// (32/1120) u^4, is below 1.8e-11 there (< 2e-9 relative), i.e. at most 0.0286 u_gate^4 / (2 sb^2)
// in a log-likelihood term -- 3.8e-10 for sb = 0.02 and a 3-sigma gate -- against the
// 1e-5 relative the weights are held to.
__host__ __device__ inline double acos_sq_p3(double u) {  // the series divided by u
    const double p = fma(16.0 / 180.0, u, 8.0 / 24.0);
    return fma(p, u, 2.0);
}
__host__ __device__ inline double acos_sq_u3(double u) { return u * acos_sq_p3(u); }
__host__ __device__ inline double acos_sq_u6(double u) {
    double p = 128.0 / 33264.0;
    p = fma(p, u, 64.0 / 6300.0);
    p = fma(p, u, 32.0 / 1120.0);
    p = fma(p, u, 16.0 / 180.0);
    p = fma(p, u, 8.0 / 24.0);
    p = fma(p, u, 2.0);
    return u * p;
}
__device__ inline double acos_sq_u12(double u) {
    double p = 8192.0 / 389398464.0;
    p = __fma_rn(p, u, 4096.0 / 85357272.0);
    p = __fma_rn(p, u, 2048.0 / 18475600.0);
    p = __fma_rn(p, u, 1024.0 / 3938220.0);
    p = __fma_rn(p, u, 512.0 / 823680.0);
    p = __fma_rn(p, u, 256.0 / 168168.0);
    p = __fma_rn(p, u, 128.0 / 33264.0);
    p = __fma_rn(p, u, 64.0 / 6300.0);
    p = __fma_rn(p, u, 32.0 / 1120.0);
    p = __fma_rn(p, u, 16.0 / 180.0);
    p = __fma_rn(p, u, 8.0 / 24.0);
    p = __fma_rn(p, u, 2.0);
    return u * p;
}

// One gated log-likelihood term (cpu.rs:69-96).  The kernel keeps two sums per particle, of
// gated squared range (and Cartesian) residuals and of gated squared bearing errors, and forms
// ll = -(sum_r inv2rr + sum_b inv2bb) once at the end.
//   The reference's gate is continuous: the gated constant -g^2/2 equals -G inv2s at the
//   threshold G = g^2 s^2, so a gated term is -min(e^2, G) inv2s (G = +inf when gate = 0); for
//   the range it is taken as min(|e|, g s)^2, one min and one FMA.
//   b = M (l - p) is formed as M l - M p (three FMAs per component).
//   Range-bearing: |b|^2 + 1e-18 (rn >= 1e-9), rn from the v_rsq_f64 seed refined once by
//   Newton (relative ~1e-16); u = 1 - dot = 1 + pv/rn in f64; eb^2 from the series in u.  With
//   a gate narrower than u = 0.005 the series is evaluated at min(u, u_gate), u_gate being the
//   largest u whose series value does not exceed the gate's g^2 sb^2: that min() alone does
//   the gating, and the series' last factor u folds into the accumulating FMA.  u is clamped at
//   0 (an f32 direction of norm 1 + 1e-7 can put it 1e-7 below), as the reference clamps the
//   cosine.
__device__ inline void body_offset(const BodyMap& b, const ObsRec& o, double& bx, double& by, double& bz) {
    bx = __fma_rn(b.m00, o.lx, __fma_rn(b.m01, o.ly, __fma_rn(b.m02, o.lz, -b.mpx)));
    by = __fma_rn(b.m10, o.lx, __fma_rn(b.m11, o.ly, __fma_rn(b.m12, o.lz, -b.mpy)));
    bz = __fma_rn(b.m20, o.lx, __fma_rn(b.m21, o.ly, __fma_rn(b.m22, o.lz, -b.mpz)));
}

// Gated range-bearing pair on the scaled path (26 FP64 instructions with the body offset).
// In: b2 = 2 b (the particle's body map doubled, exact), the record's vs = v / (2 u_gate) and
// rng4 = 4 range.  X = |b2|^2 = 4 |b|^2; y0h = rsq(X) ~ 1 / (2 |b|); one Newton step
// written as q = 3 - X y0h^2 gives 1/|b| = y0h q and |b| = X y0h q / 4, so
//   ec4 = X y0h q - rng4 = 4 (|b| - range)            (gated as min(|ec4|, 4 g sr))
//   us  = clamp((1 + (b.v) / |b|) / u_gate, 0, 1)      (the clamp bit of the FMA is the gate)
// and eb^2 = u acos-series(u) = us (c0 + c1 us + c2 us^2) with u = u_gate us.  The range sum
// comes out 16x, which the caller's final scale removes.
__device__ inline void rb_scaled(double bx, double by, double bz, double vx, double vy, double vz, double rng4,
                                 const WeightParams& p, double& accr, double& accb) {
    const double X = __fma_rn(bx, bx, __fma_rn(by, by, __fma_rn(bz, bz, 4e-18)));
    const double pvs = __fma_rn(bx, vx, __fma_rn(by, vy, bz * vz));
    const double y0h = __builtin_amdgcn_rsq(X);
    const double P = X * y0h;
    const double q = __fma_rn(-P, y0h, 3.0);
    const double ec = fmin(fabs(__fma_rn(P, q, -rng4)), p.rng_gate4);
    accr = __fma_rn(ec, ec, accr);
    const double y = y0h * q;
    const double us = fmin(fmax(__fma_rn(pvs, y, p.cg), 0.0), 1.0);
    double s = __fma_rn(p.c2, us, p.c1);
    s = __fma_rn(s, us, p.c0);
    accb = __fma_rn(us, s, accb);
}

// Gated range-bearing pair for a bearing gate wider than u = 0.005 (or none): series up to
// u = 0.125, acos beyond.
__device__ inline void rb_from_b(double bx, double by, double bz, double vx, double vy, double vz, double range,
                                 const WeightParams& p, double& accr, double& accb) {
    const double r2 = __fma_rn(bx, bx, __fma_rn(by, by, __fma_rn(bz, bz, 1e-18)));
    const double pv = __fma_rn(bx, vx, __fma_rn(by, vy, bz * vz));
    const double y0 = __builtin_amdgcn_rsq(r2);  // v_rsq_f64 seed, refined once below
    const double hy = 0.5 * r2 * y0;
    const double y = __fma_rn(y0, __fma_rn(-hy, y0, 0.5), y0);
    const double ec = fmin(fabs(__fma_rn(r2, y, -range)), p.rng_gate);
    accr = __fma_rn(ec, ec, accr);
    const double u = fmin(fmax(__fma_rn(pv, y, 1.0), 0.0), 2.0);
    double eb2;
    if (u <= 0.005) {
        eb2 = acos_sq_u6(u);
    } else if (u <= 0.125) {
        eb2 = acos_sq_u12(u);
    } else {
        const double eb = acos(1.0 - u);
        eb2 = eb * eb;
    }
    accb += fmin(eb2, p.brg_gate2);
}

// Cartesian pair; hs = 0.5 when b comes from the doubled body map (hs b is then exact).
__device__ inline void cart_from_b(double bx, double by, double bz, double vx, double vy, double vz, double hs,
                                   const WeightParams& p, double& accr) {
    const double ex = __fma_rn(hs, bx, -vx), ey = __fma_rn(hs, by, -vy), ez = __fma_rn(hs, bz, -vz);
    const double r2 = __fma_rn(ex, ex, __fma_rn(ey, ey, ez * ez));
    accr += fmin(r2, p.cart_gate2);
}

// Covariance partials of one 1024-particle tile, and the covariance finalize of one filter by a
// 1024-thread block (both defined with the estimate kernels below).  kWave: each wave stores its
// own partial and the finalize adds a tile's four in the block tree's order (no barrier in the
// tile's code: a barrier there makes the compiler keep a prefetched observation record of the
// likelihood loop, which shares the kernel, in vector registers).
template <bool kWave>
__device__ void covs_tile(const Layout& L, FilterDev* F, const PoseBufs& B, int swapped, const float* __restrict__ w,
                          double* __restrict__ part, int stride, const ParamRef<CovParams>& A, int blk);
__device__ void cov_fin_wave(const Layout& L, FilterDev* F, const double* __restrict__ part, int stride,
                             FilterDev* out, double* __restrict__ fold, int r);
__device__ void cov_fold_sum(FilterDev* F, const double* __restrict__ fold, FilterDev* out, int f);

// Gated log-likelihood (cpu.rs:44-97, 220-248), kPPT particles per lane, plus the block's
// online-softmax partials.  Within each category of observations (range-bearing, then
// Cartesian) the even-numbered and the odd-numbered records are summed as two independent
// chains, each in order, and the two sums added.
//   kDefer (staged cycles, where predict stored the new poses in the idle set and B comes with
// its two sets swapped, so that B.s[parity] is that idle set): workgroups past the likelihood
// grid (blk >= nblk) compute the covariance partials of the previous cycle, one 1024-particle
// tile each, from the current set; they are dispatched last and run on CUs the likelihood's
// final waves leave idle.  (A separate instantiation, so that the deferred work's registers do
// not constrain the plain kernel.)
template <int kPPT, bool kDefer>
__global__ __launch_bounds__(kTB) __attribute__((amdgpu_waves_per_eu(kPPT == 1 ? 5 : 1, 8)))
void k_weight(Layout L, PoseBufs B, FilterDev* F, double* __restrict__ ll, double* __restrict__ part,
              ParamRef<WeightParams> A, const float* __restrict__ cw, double* __restrict__ cpart, int cstride,
              ParamRef<CovParams> cA, int nblk) {
    __shared__ double sh[4];
    constexpr int kSpan = kTB * kPPT;
    const int tl = (int)threadIdx.x;
    const int blk = blockIdx.x;
    if constexpr (kDefer) {
        if (blk >= nblk) {
            covs_tile<true>(L, F, B, 1, cw, cpart, cstride, cA, blk - nblk);
            return;
        }
    }
    const int f = layout_filter(L, blk * kPPT);
    const int off = ((ConstPtr<int>)L.f_off)[f];
    const int nf = ((ConstPtr<int>)L.f_n)[f];
    const WeightParams p = uniform_params(A, f);
    const PoseSet S = B.s[((ConstPtr<FilterDev>)F)[f].parity];
    int li[kPPT];
    bool valid[kPPT];
    BodyMap bm[kPPT];
    double accr[kPPT], accb[kPPT], accc[kPPT], acc[kPPT];
    const bool fast = p.fast_rb != 0;
    const double hs = fast ? 0.5 : 1.0;  // Cartesian records see b = hs * (mapped offset)
#pragma unroll
    for (int e = 0; e < kPPT; ++e) {
        li[e] = blk * kSpan + e * kTB + tl - off;
        valid[e] = li[e] < nf;
        bm[e] = body_map(S, off + (valid[e] ? li[e] : 0));
        if (fast) {  // doubled body map: every offset it forms is exactly 2 b
            BodyMap& m = bm[e];
            m.m00 *= 2.0; m.m01 *= 2.0; m.m02 *= 2.0;
            m.m10 *= 2.0; m.m11 *= 2.0; m.m12 *= 2.0;
            m.m20 *= 2.0; m.m21 *= 2.0; m.m22 *= 2.0;
            m.mpx *= 2.0; m.mpy *= 2.0; m.mpz *= 2.0;
        }
    }
    ConstObsPtr obs = (ConstObsPtr)p.obs;
    const int nobs = p.nobs, nrb = p.nrb;  // range-bearing records first, then Cartesian
    // Each category is summed as (terms of its even-numbered records, in order) + (terms of its
    // odd-numbered records, in order): chain s of a lane sums the records k = c0 + s, c0 + s + 2,
    // ... of a category that starts at record c0.
    double ar[2][kPPT], ab[2][kPPT], ac[2][kPPT];
#pragma unroll
    for (int e = 0; e < kPPT; ++e)
        ar[0][e] = ab[0][e] = ac[0][e] = ar[1][e] = ab[1][e] = ac[1][e] = 0.0;
    {
        if (fast) {
            int k = 0;
            if (nrb >= 2) {
                // Software-pipelined record loads: the next pair of records is requested before
                // the current pair is evaluated, so its scalar-load latency overlaps the
                // arithmetic.  Two register sets (a, b) alternate so that no loaded record is
                // copied.  Scalar loads may return out of order, so each wait for a pair (issued
                // one half-iteration earlier) is explicit, and the scheduler is kept from hoisting
                // the next prefetch above it.
                ObsRec a0 = obs[0], a1 = obs[1];
                for (; k + 3 < nrb; k += 4) {
                    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0): pair a has arrived
                    __builtin_amdgcn_sched_barrier(0);
                    const ObsRec b0 = obs[k + 2], b1 = obs[k + 3];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int e = 0; e < kPPT; ++e) {
                        double bx, by, bz;
                        body_offset(bm[e], a0, bx, by, bz);
                        rb_scaled(bx, by, bz, a0.vx, a0.vy, a0.vz, a0.range, p, ar[0][e], ab[0][e]);
                        body_offset(bm[e], a1, bx, by, bz);
                        rb_scaled(bx, by, bz, a1.vx, a1.vy, a1.vz, a1.range, p, ar[1][e], ab[1][e]);
                    }
                    __builtin_amdgcn_s_waitcnt(0xC07F);  // pair b has arrived
                    __builtin_amdgcn_sched_barrier(0);
                    const int kn = k + 5 < nrb ? k + 4 : k + 2;  // in bounds; unused past the end
                    a0 = obs[kn];
                    a1 = obs[kn + 1];
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int e = 0; e < kPPT; ++e) {
                        double bx, by, bz;
                        body_offset(bm[e], b0, bx, by, bz);
                        rb_scaled(bx, by, bz, b0.vx, b0.vy, b0.vz, b0.range, p, ar[0][e], ab[0][e]);
                        body_offset(bm[e], b1, bx, by, bz);
                        rb_scaled(bx, by, bz, b1.vx, b1.vy, b1.vz, b1.range, p, ar[1][e], ab[1][e]);
                    }
                }
                if (k + 1 < nrb) {  // pair a = records k, k + 1 is loaded
                    __builtin_amdgcn_s_waitcnt(0xC07F);
#pragma unroll
                    for (int e = 0; e < kPPT; ++e) {
                        double bx, by, bz;
                        body_offset(bm[e], a0, bx, by, bz);
                        rb_scaled(bx, by, bz, a0.vx, a0.vy, a0.vz, a0.range, p, ar[0][e], ab[0][e]);
                        body_offset(bm[e], a1, bx, by, bz);
                        rb_scaled(bx, by, bz, a1.vx, a1.vy, a1.vz, a1.range, p, ar[1][e], ab[1][e]);
                    }
                    k += 2;
                }
            }
            if (k < nrb) {
                const ObsRec o = obs[k];
#pragma unroll
                for (int e = 0; e < kPPT; ++e) {
                    double bx, by, bz;
                    body_offset(bm[e], o, bx, by, bz);
                    rb_scaled(bx, by, bz, o.vx, o.vy, o.vz, o.range, p, ar[0][e], ab[0][e]);
                }
            }
        } else {
            for (int k = 0; k < nrb; ++k) {
                const ObsRec o = obs[k];
                const int sk = k & 1;
#pragma unroll
                for (int e = 0; e < kPPT; ++e) {
                    double bx, by, bz;
                    body_offset(bm[e], o, bx, by, bz);
                    double r = sk ? ar[1][e] : ar[0][e], b2 = sk ? ab[1][e] : ab[0][e];
                    rb_from_b(bx, by, bz, o.vx, o.vy, o.vz, o.range, p, r, b2);
                    if (sk) {
                        ar[1][e] = r;
                        ab[1][e] = b2;
                    } else {
                        ar[0][e] = r;
                        ab[0][e] = b2;
                    }
                }
            }
        }
        for (int k = nrb; k < nobs; ++k) {
            const ObsRec o = obs[k];
            const int sk = (k - nrb) & 1;
#pragma unroll
            for (int e = 0; e < kPPT; ++e) {
                double bx, by, bz;
                body_offset(bm[e], o, bx, by, bz);
                double c = sk ? ac[1][e] : ac[0][e];
                cart_from_b(bx, by, bz, o.vx, o.vy, o.vz, hs, p, c);
                if (sk)
                    ac[1][e] = c;
                else
                    ac[0][e] = c;
            }
        }
    }
#pragma unroll
    for (int e = 0; e < kPPT; ++e) {
        accr[e] = ar[0][e] + ar[1][e];
        accb[e] = ab[0][e] + ab[1][e];
        accc[e] = ac[0][e] + ac[1][e];
    }
    const double sr = fast ? p.inv2rr * 0.0625 : p.inv2rr;  // the scaled range sum is 16x
#pragma unroll
    for (int e = 0; e < kPPT; ++e)
        acc[e] = -__fma_rn(accr[e], sr, __fma_rn(accc[e], p.inv2rr, accb[e] * p.inv2bb));
    double mloc = -INFINITY;
#pragma unroll
    for (int e = 0; e < kPPT; ++e)
        if (valid[e]) {
            ll[off + li[e]] = acc[e];
            mloc = fmax(mloc, acc[e]);
        }
    // online softmax partials of this block: m_b, sum exp(ll-m_b), sum exp(2(ll-m_b))
    const double mb = block_max(mloc, sh);
    double es = 0.0, es2 = 0.0;
#pragma unroll
    for (int e = 0; e < kPPT; ++e)
        if (valid[e] && mb > -INFINITY) {
            const double x = exp(acc[e] - mb);
            es += x;
            es2 += x * x;
        }
    const double sb = block_sum(es, sh);
    const double s2b = block_sum(es2, sh);
    if (threadIdx.x == 0) {
        part[3 * blk] = mb;
        part[3 * blk + 1] = sb;
        part[3 * blk + 2] = s2b;
    }
}

// Markley mean on the device (types.rs:128-144): the moment matrix from the estimate sums
// (upper triangle, cuda.rs:277-286), its dominant unit eigenvector with quat_mean4 (the host's
// function, same IEEE operation sequence), and the covariance pass parameters with the mean
// handed over at f32 precision (cuda.rs:291-294).  A degenerate moment matrix yields the
// identity and valid = 0.
__device__ __forceinline__ void markley_mean(FilterDev& fd, const double* sums, CovParams& cov) {
    double M[16];
    int e = 3;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = r; c < 4; ++c) {
            M[4 * r + c] = M[4 * c + r] = sums[e];
            ++e;
        }
    double q[4] = {1.0, 0.0, 0.0, 0.0};
    const int rc = quat_mean4(M, q);
    if (rc != 0) {
        q[0] = 1.0;
        q[1] = q[2] = q[3] = 0.0;
    }
    for (int k = 0; k < 4; ++k) fd.qmean[k] = q[k];
    fd.qvalid = rc == 0 ? 1 : 0;
    for (int k = 0; k < 3; ++k) cov.mp[k] = (double)(float)sums[k];
    for (int k = 0; k < 4; ++k) cov.mq[k] = (double)(float)q[k];
}

// Per filter: M = max m_b, S = sum s_b e^(m_b - M), S2 = sum s2_b e^(2(m_b - M)) over the
// filter's likelihood blocks, in a fixed order (one 1024-thread block per filter, so that each
// lane has few partials to load; the kernel boundary orders it after every partial).  Lane t
// takes blocks t, t + 1024, ... in that order.  Up to four blocks per lane (filters up to 4M
// particles at one particle per lane) the partials stay in registers from a single load; larger
// filters read the maxima once more.  A term whose e^(m_b - M) underflows (m_b - M <= -746, the
// exponential is exactly 0) is skipped without evaluating it: with peaked likelihoods that is
// most blocks.  S and S2 are summed per lane, then over the 16 waves in order (block16_sum's
// tree, two values at once).
__global__ __launch_bounds__(kFinTB) void k_weight_fin(Layout L, FilterDev* F, const double* __restrict__ part,
                                                       int span, int nfilt, int flip, const double* __restrict__ cpart,
                                                       int cstride, FilterDev* cout, double* __restrict__ cfold) {
    __shared__ double sh[kFinTB / 64], sh1[kFinTB / 64], sh2[kFinTB / 64];
    if ((int)blockIdx.x >= nfilt) {  // the previous staged cycle's covariance finalize (its lane groups)
        cov_fin_wave(L, F, cpart, cstride, cout, cfold, blockIdx.x - nfilt);
        return;
    }
    const int f = blockIdx.x;
    const int b0 = L.f_off[f] / span, nb = L.f_padn[f] / span;
    const int t = (int)threadIdx.x;
    const bool resident = nb <= 4 * kFinTB;
    double mb[4], e1[4], e2[4];
    double m = -INFINITY;
    for (int c = 0; c < nb; c += 4 * kFinTB) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int b = c + t + u * kFinTB;
            const bool in = b < nb;
            mb[u] = in ? part[3 * (b0 + b)] : -INFINITY;
            e1[u] = in && resident ? part[3 * (b0 + b) + 1] : 0.0;
            e2[u] = in && resident ? part[3 * (b0 + b) + 2] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) m = fmax(m, mb[u]);
    }
    const double M = block16_max(m, sh);
    double sl = 0.0, s2l = 0.0;
    for (int c = 0; c < nb; c += 4 * kFinTB) {
        if (!resident) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int b = c + t + u * kFinTB;
                const bool in = b < nb;
                mb[u] = in ? part[3 * (b0 + b)] : -INFINITY;
                e1[u] = in ? part[3 * (b0 + b) + 1] : 0.0;
                e2[u] = in ? part[3 * (b0 + b) + 2] : 0.0;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (mb[u] - M > -746.0) {
                const double sc = exp(mb[u] - M);
                sl += e1[u] * sc;
                s2l += e2[u] * (sc * sc);
            }
    }
    sl = wave_sum(sl);
    s2l = wave_sum(s2l);
    if ((t & 63) == 0) {  // (sh may still be read by block16_max's last loop)
        sh1[t >> 6] = sl;
        sh2[t >> 6] = s2l;
    }
    __syncthreads();
    double Ssum = sh1[0], S2sum = sh2[0];
#pragma unroll
    for (int k = 1; k < kFinTB / 64; ++k) {
        Ssum += sh1[k];
        S2sum += sh2[k];
    }
    if (threadIdx.x == 0) {
        FilterDev& fd = F[f];
        const bool ok = (Ssum > 0.0) && isfinite(Ssum) && isfinite(M);
        fd.max_ll = M;
        fd.sum_e = Ssum;
        fd.sum_e2 = S2sum;
        fd.uniform = ok ? 0 : 1;
        if (flip) fd.parity ^= 1;  // predict's new pose set becomes current
    }
}

// Stored weights of the 4 consecutive particles of this lane in 1024-particle tile `tile`:
// w = exp(ll - max) / sum, or 1/n when uni (the weight pass was degenerate, had no
// observations, or uniform weights are forced); 0 past the filter's end.  Below ll - max = -104
// the f32 weight is 0 without evaluating exp: exp(-104) < 2^-150 and sum >= 1 (the maximum's own
// term), so the quotient rounds to 0.  w2 receives the lane's sum of squares of the stored f32
// weights, in particle order.
__device__ inline float4 tile_norm(const double* __restrict__ ll, int tile, const Layout& L, int f,
                                   const FilterDev& fd, bool uni, double& w2) {
    const int base = tile * kTile + (int)threadIdx.x * 4;
    const double2 a = *reinterpret_cast<const double2*>(ll + base);
    const double2 b = *reinterpret_cast<const double2*>(ll + base + 2);
    const double v[4] = {a.x, a.y, b.x, b.y};
    const int li0 = base - L.f_off[f], nf = L.f_n[f];
    const float winv = L.f_winv[f];
    float o[4];
    w2 = 0.0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        o[e] = 0.f;
        if (li0 + e < nf) {
            const double x = v[e] - fd.max_ll;
            o[e] = uni ? winv : (x < -104.0 ? 0.f : (float)(exp(x) / fd.sum_e));
            w2 += (double)o[e] * (double)o[e];
        }
    }
    return make_float4(o[0], o[1], o[2], o[3]);
}

// Normalised weights of every 1024-particle tile plus its partial of sum w^2 (the same tile
// function and block tree as the fused cycle's k_norm_agg, so both give the same bits).
// (defined with the resampling scan below)
__device__ inline void tile_q(const float4 wv, int tile, int off, int nf, uint64_t (&q)[4]);
__device__ inline uint64_t block_sum_u64(uint64_t v, uint64_t* s_w);

// The tile's fixed-point aggregate sum q as well (part_q), which a later resampling scan starts
// from while the weights stay as they are (k_scan_agg).
__global__ __launch_bounds__(kTB) void k_normalize(Layout L, FilterDev* F, const double* __restrict__ ll,
                                                   float* __restrict__ w, double* __restrict__ part,
                                                   uint64_t* __restrict__ part_q, int force_uniform) {
    __shared__ double sh[4];
    __shared__ uint64_t s_w[4];
    const int tile = blockIdx.x;
    const int f = layout_filter(L, tile * (kTile / kTB));
    const FilterDev& fd = F[f];
    double w2;
    const float4 wv = tile_norm(ll, tile, L, f, fd, force_uniform || fd.uniform, w2);
    *reinterpret_cast<float4*>(w + tile * kTile + (int)threadIdx.x * 4) = wv;
    uint64_t q[4];
    tile_q(wv, tile, L.f_off[f], L.f_n[f], q);
    const double bs = block_sum(w2, sh);
    const uint64_t agg = block_sum_u64(q[0] + q[1] + q[2] + q[3], s_w);
    if (threadIdx.x == 0) {
        part[tile] = bs;
        part_q[tile] = agg;
    }
}

// Per filter: ESS = 1 / sum w^2 of the stored weights -- the quantity the reference driver
// tests (particle_filter.rs:115-117 via pfc_est_sums out[13]) -- and, when ratio is given,
// the resample predicate ESS < ratio * n.  (The fused cycle forms them in k_scan_agg, in the
// same order.)
__global__ __launch_bounds__(kFinTB) void k_norm_fin(Layout L, FilterDev* F, const double* __restrict__ part,
                                                     const float* __restrict__ ratio) {
    __shared__ double sh[kFinTB / 64];
    const int f = blockIdx.x;
    const int b0 = L.f_off[f] / kTile, nb = L.f_padn[f] / kTile;
    double v = 0.0;
    for (int b = (int)threadIdx.x; b < nb; b += kFinTB) v += part[b0 + b];
    const double tot = block16_sum(v, sh);
    if (threadIdx.x == 0) {
        FilterDev& fd = F[f];
        fd.est_src = 0;  // weights changed
        fd.sum_w2 = tot;
        fd.ess = tot > 0.0 ? 1.0 / tot : 0.0;
        if (ratio) {
            const double r = (double)ratio[f];
            fd.do_resample = (r > 0.0 && fd.ess < r * (double)L.f_n[f]) ? 1 : 0;
        }
    }
}

// --------------------------------------------------------------------------- resample

// Block-wide inclusive scan of one u64 per thread (4 waves); returns the block total.
__device__ inline uint64_t block_scan_u64(uint64_t x, uint64_t& incl, uint64_t* s_w) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    __syncthreads();
    if (lane == 63) s_w[wv] = x;
    __syncthreads();
    uint64_t pre = 0;
    for (int k = 0; k < wv; ++k) pre += s_w[k];
    incl = pre + x;
    return s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

// Fixed-point weights of the 4 particles of this lane in a 1024-particle tile.
__device__ inline void tile_q(const float4 wv, int tile, int off, int nf, uint64_t (&q)[4]) {
    const int li0 = tile * kTile + (int)threadIdx.x * 4 - off;
    q[0] = (li0 + 0 < nf) ? qweight(wv.x) : 0;
    q[1] = (li0 + 1 < nf) ? qweight(wv.y) : 0;
    q[2] = (li0 + 2 < nf) ? qweight(wv.z) : 0;
    q[3] = (li0 + 3 < nf) ? qweight(wv.w) : 0;
}
__device__ inline float4 tile_w(const float* __restrict__ w, int tile) {
    return *reinterpret_cast<const float4*>(w + tile * kTile + (int)threadIdx.x * 4);
}

// Injection pick j of a filter (cu:333-352 formula) drawn from the LCG state lcg of this resample:
// its target and f64 jitter; the pick pushes itself onto its target's list (head[] is -1
// everywhere at rest: allocation sets it, k_gather resets every head of its filter).
// inj_push returns the previous head, for next[off + j], so that a caller can store it later and
// let the exchange's round trip overlap other work (the lists' order does not matter: the
// gather applies a list's picks in ascending pick order).
__device__ inline int32_t inj_push(int off, int nf, uint64_t lcg, uint64_t gen, int j, double* __restrict__ jit,
                                   int32_t* __restrict__ head) {
    uint64_t z = lcg ^ (gen * kPhi) ^ (((uint64_t)j + 1) * kC1);
    if (z == 0) z = 0x243F6A8885A308D3ull;
    int t = (int)(u01_co(sm_next(z)) * (double)nf);
    if (t >= nf) t = nf - 1;
    const double jx = u01_co(sm_next(z)) - 0.5;
    const double jy = u01_co(sm_next(z)) - 0.5;
    const double jz = u01_co(sm_next(z)) - 0.5;
    jit[3 * (off + j)] = jx;
    jit[3 * (off + j) + 1] = jy;
    jit[3 * (off + j) + 2] = jz;
    return atomicExch(&head[off + t], j);
}
__device__ inline void inj_pick(int off, int nf, uint64_t lcg, uint64_t gen, int j, double* __restrict__ jit,
                                int32_t* __restrict__ head, int32_t* __restrict__ next) {
    next[off + j] = inj_push(off, nf, lcg, gen, j, jit, head);
}

// Single-pass scan of the fixed-point weights with decoupled look-back (integer sums: any
// association gives the same bits): the reference-ABI path's pfc_resample, which cannot rely on
// tile aggregates left by pfc_weight (the state may be set in between).  The fused cycle scans
// in two passes instead (k_norm_agg, k_scan_agg below).
//
// Tiles are 1024 particles, one block each.  Each tile publishes its aggregate, looks back over
// its filter's predecessors until it meets an inclusive prefix, and publishes its own inclusive
// prefix; the first tile of a filter publishes its prefix directly.  The whole block looks
// back, 1024 predecessors per window (four per thread, read together, so one memory round trip
// resolves up to 1024 tiles).  A published 64-bit value occupies two status words, each
// (status << 32) | one 32-bit half, written and read as device-scope atomics; status =
// epoch << 2 | kind, so words from an earlier launch never match (k_flip, which ends every
// resample, advances the epoch).  Tiles of filters that do not resample refresh their words
// with kind 0, so every word is at most one launch old.
//
// Waits are bounded: a predecessor whose status has not appeared within spin_limit reads has
// its aggregate summed by the waiting workgroup itself (the same integer it would publish), so
// no tile depends on another workgroup being scheduled and the result does not depend on the
// order in which workgroups run.  The filter's last tile stores the total T and advances the
// LCG (pf_kernels.cu:647-657: only when T > 0).
constexpr uint32_t kLbAgg = 1, kLbPre = 2;
constexpr int kLbPer = 4, kLbWin = kTB * kLbPer;  // predecessors per thread and per look-back window

__device__ inline void lb_publish(uint64_t* st, int tile, uint32_t status, uint64_t v) {
    __hip_atomic_store(&st[2 * tile], ((uint64_t)status << 32) | (v & 0xffffffffull), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st[2 * tile + 1], ((uint64_t)status << 32) | (v >> 32), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// Kind of the value tile has published in this epoch (0: none yet, or the halves of two
// different publications were read).
__device__ inline uint32_t lb_read(uint64_t* st, int tile, uint32_t epoch, uint64_t& v) {
    const uint64_t lo = __hip_atomic_load(&st[2 * tile], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t hi = __hip_atomic_load(&st[2 * tile + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t sl = (uint32_t)(lo >> 32), sh = (uint32_t)(hi >> 32);
    if (sl != sh || (sl >> 2) != epoch) return 0;
    v = (hi << 32) | (lo & 0xffffffffull);
    return sl & 3u;
}

__device__ inline uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Block sum of one u64 per thread (4 waves); any association gives the same bits.
__device__ inline uint64_t block_sum_u64(uint64_t v, uint64_t* s_w) {
    v = wave_sum_u64(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    return s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

__global__ __launch_bounds__(kTB) void k_scan(Layout L, FilterDev* F, const float* __restrict__ w,
                                              uint64_t* __restrict__ cum, uint64_t* __restrict__ tcum, uint64_t* st,
                                              const unsigned long long* ctl,
                                              int spin_limit) {
    __shared__ uint64_t s_w[4];
    __shared__ int s_dmin, s_nun;
    __shared__ uint64_t s_excl;
    __shared__ int s_un_tile[kLbWin];     // predecessors of a window found unpublished
    __shared__ uint64_t s_un_val[kLbWin];  // their aggregates, summed here
    const int tile = blockIdx.x;
    const uint32_t epoch = (uint32_t)ctl[0] & 0x3fffffffu;
    const int f = layout_filter(L, tile * (kTile / kTB));
    FilterDev& fd = F[f];
    const float4 wv = tile_w(w, tile);
    if (!fd.do_resample) {
        if (threadIdx.x == 0) lb_publish(st, tile, epoch << 2, 0);
        return;
    }
    const int off = L.f_off[f], nf = L.f_n[f];
    const int t0 = off / kTile, nt = L.f_padn[f] / kTile;
    uint64_t q[4];
    tile_q(wv, tile, off, nf, q);
    q[1] += q[0];
    q[2] += q[1];
    q[3] += q[2];
    uint64_t incl;
    const uint64_t agg = block_scan_u64(q[3], incl, s_w);
    if (threadIdx.x == 0) {
        lb_publish(st, tile, (epoch << 2) | (tile == t0 ? kLbPre : kLbAgg), agg);
        s_excl = 0;
    }
    __syncthreads();
    if (tile != t0) {
        // windows of kLbWin predecessors: thread i holds the ones at distance i + 256 j, j < kLbPer
        const int tid = threadIdx.x;
        uint64_t ex = 0;
        for (int base = tile - 1; base >= t0; base -= kLbWin) {
            uint64_t v[kLbPer];
            uint32_t kind[kLbPer];
            bool in[kLbPer];
#pragma unroll
            for (int j = 0; j < kLbPer; ++j) {
                in[j] = base - tid - kTB * j >= t0;
                kind[j] = 0;
                v[j] = 0;
            }
            for (int it = 0;; ++it) {
                bool all = true;
#pragma unroll
                for (int j = 0; j < kLbPer; ++j)
                    if (in[j] && kind[j] == 0) {
                        kind[j] = lb_read(st, base - tid - kTB * j, epoch, v[j]);
                        all = all && kind[j] != 0;
                    }
                if (all || it >= spin_limit) break;
                __builtin_amdgcn_s_sleep(1);
            }
            int bad = 0;
#pragma unroll
            for (int j = 0; j < kLbPer; ++j) bad |= (in[j] && kind[j] == 0) ? 1 : 0;
            if (tid == 0) s_nun = 0;
            if (__syncthreads_or(bad)) {
                // predecessors still unpublished: the workgroup sums each of their tiles itself
                // (the same integer aggregate they would publish), so nothing waits longer
                int slot[kLbPer];
#pragma unroll
                for (int j = 0; j < kLbPer; ++j) {
                    slot[j] = -1;
                    if (in[j] && kind[j] == 0) {
                        slot[j] = atomicAdd(&s_nun, 1);
                        s_un_tile[slot[j]] = base - tid - kTB * j;
                    }
                }
                __syncthreads();
                const int nun = s_nun;
                for (int u = 0; u < nun; ++u) {
                    const int ut = s_un_tile[u];
                    uint64_t qq[4], inc2;
                    tile_q(tile_w(w, ut), ut, off, nf, qq);
                    const uint64_t a = block_scan_u64(qq[0] + qq[1] + qq[2] + qq[3], inc2, s_w);
                    if (tid == 0) s_un_val[u] = a;
                }
                __syncthreads();
#pragma unroll
                for (int j = 0; j < kLbPer; ++j)
                    if (slot[j] >= 0) {
                        v[j] = s_un_val[slot[j]];
                        kind[j] = kLbAgg;
                    }
            }
            int dmin = INT_MAX;
#pragma unroll
            for (int j = 0; j < kLbPer; ++j)
                if (in[j] && kind[j] == kLbPre) dmin = min(dmin, tid + kTB * j);
            if (tid == 0) s_dmin = INT_MAX;
            __syncthreads();
            if (dmin != INT_MAX) atomicMin(&s_dmin, dmin);
            __syncthreads();
            const int dstar = s_dmin;  // nearest predecessor holding an inclusive prefix
            uint64_t part = 0;
#pragma unroll
            for (int j = 0; j < kLbPer; ++j)
                if (in[j] && tid + kTB * j <= dstar) part += v[j];
            ex += block_sum_u64(part, s_w);
            if (dstar != INT_MAX) break;
        }
        if (tid == 0) s_excl = ex;
    }
    __syncthreads();
    const uint64_t excl = s_excl;
    if (threadIdx.x == 0) {
        if (tile != t0) lb_publish(st, tile, (epoch << 2) | kLbPre, excl + agg);
        if (tile == t0 + nt - 1) {
            const uint64_t T = excl + agg;
            fd.total = T;
            if (T > 0) {
                const uint64_t lcg = fd.lcg * 6364136223846793005ull + 1442695040888963407ull;  // cu:653
                fd.lcg = lcg;
                fd.u32 = (uint32_t)(lcg >> 32);
            }
        }
    }
    const uint64_t pre = excl + incl - q[3];
    uint64_t* cp = cum + tile * kTile + (int)threadIdx.x * 4;
    cp[0] = pre + q[0];
    cp[1] = pre + q[1];
    cp[2] = pre + q[2];
    cp[3] = pre + q[3];
    if (threadIdx.x == kTB - 1) tcum[tile] = pre + q[3];
}

// Fused cycle, first pass of the resampling scan: the normalised weights of every 1024-particle
// tile and its sum-w^2 partial (tile_norm and the block tree of k_normalize, so both give the same
// bits), the tile's fixed-point aggregate sum q (the value a look-back tile publishes), and the
// tile's share of the filter's injection picks: drawn before the resample predicate is known,
// with the LCG value its advance will give (k_scan_agg advances fd.lcg only afterwards).
// cout (staged cycles): workgroups past the tiles (blockIdx.x >= ntile) add the previous
// cycle's covariance folds (cov_fold_sum) of filter blockIdx.x - ntile.
__global__ __launch_bounds__(kTB) void k_norm_agg(Layout L, FilterDev* F, const double* __restrict__ ll,
                                                  float* __restrict__ w, double* __restrict__ part_n,
                                                  uint64_t* __restrict__ part_q, ParamRef<InjectParams> inj,
                                                  double* __restrict__ jit, int32_t* __restrict__ head,
                                                  int32_t* __restrict__ next, const double* __restrict__ cfold,
                                                  FilterDev* cout, int ntile) {
    __shared__ double sh[4];
    __shared__ uint64_t s_w[4];
    if (cout && (int)blockIdx.x >= ntile) {
        cov_fold_sum(F, cfold, cout, blockIdx.x - ntile);
        return;
    }
    const int tile = blockIdx.x;
    const int f = layout_filter(L, tile * (kTile / kTB));
    const FilterDev& fd = F[f];
    const int off = L.f_off[f], nf = L.f_n[f];
    // the first pick of each lane is pushed before the tile's own work, its link stored at the
    // end (the exchange's round trip overlaps the log-likelihood loads); further ones after
    const InjectParams ip = inj.get(f);
    int j0 = 0, j1 = 0, pj = -1;
    int32_t plink = -1;
    uint64_t lcg1 = 0;
    if (ip.k > 0) {
        const int t0 = off / kTile, nt = L.f_padn[f] / kTile;
        const int per = (ip.k + nt - 1) / nt;
        j0 = (tile - t0) * per;
        j1 = min(ip.k, j0 + per);
        lcg1 = fd.lcg * 6364136223846793005ull + 1442695040888963407ull;  // cu:653
        if (j0 + (int)threadIdx.x < j1) {
            pj = j0 + (int)threadIdx.x;
            plink = inj_push(off, nf, lcg1, ip.gen, pj, jit, head);
        }
    }
    double w2;
    const float4 wv = tile_norm(ll, tile, L, f, fd, fd.uniform != 0, w2);
    *reinterpret_cast<float4*>(w + tile * kTile + (int)threadIdx.x * 4) = wv;
    uint64_t q[4];
    tile_q(wv, tile, off, nf, q);
    const double bs = block_sum(w2, sh);
    const uint64_t agg = block_sum_u64(q[0] + q[1] + q[2] + q[3], s_w);
    if (threadIdx.x == 0) {
        part_n[tile] = bs;
        part_q[tile] = agg;
    }
    if (pj >= 0) next[off + pj] = plink;
    for (int j = j0 + (int)threadIdx.x + kTB; j < j1; j += kTB) inj_pick(off, nf, lcg1, ip.gen, j, jit, head, next);
}

// Fused cycle, second pass of the resampling scan (reduce, then scan): the first pass
// (k_norm_agg) left every tile's fixed-point aggregate, so a tile's exclusive
// prefix is the sum of its filter's earlier aggregates, read once after that kernel ended (no
// look-back, nothing to wait for); the tile then scans its own weights into cum.  Every filter
// is scanned whatever its predicate.  The filter's last tile stores the total and, as k_norm_fin
// does, forms ESS = 1 / sum w^2 from the tiles' w^2 partials, the resample predicate and the LCG
// advance (pf_kernels.cu:647-657) -- the sum in k_norm_fin's order (1024 lanes, lane v summing
// tiles v, v + 1024, ..., its 16 waves folded by wave_sum and added in sequence), which this
// block's four waves reproduce as virtual waves w, w + 4, w + 8, w + 12, so both paths give the
// same bits.  A null ratio forces the predicate (the reference ABI's pfc_resample, scanning from
// the aggregates k_normalize left).
// Exclusive prefixes of the tile aggregates of every filter (one 1024-thread workgroup per
// filter), for filters too large for k_scan_agg's per-tile sum of the earlier aggregates, which
// grows with the square of the tile count (at 16M particles, 16,384 tiles: 99 us).  Chunks of
// kPrefixRun * 1024 tiles: thread t takes the run of kPrefixRun consecutive tiles from
// c + kPrefixRun t (one 128-byte span, loaded at once), the run totals are scanned across the
// block once, and each thread writes its run's prefixes.  Integer sums: the same prefixes in any
// order.
constexpr int kPrefixTiles = 2048;  // filters above this many tiles take the prefix pass
constexpr int kPrefixRun = 16;
__global__ __launch_bounds__(kFinTB) void k_tile_prefix(Layout L, const uint64_t* __restrict__ part_q,
                                                        uint64_t* __restrict__ tpre) {
    __shared__ uint64_t s_w[kFinTB / 64];
    const int f = blockIdx.x;
    const int t0 = L.f_off[f] / kTile, nt = L.f_padn[f] / kTile;
    if (nt <= kPrefixTiles) return;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint64_t carry = 0;
    for (int c = 0; c < nt; c += kPrefixRun * kFinTB) {
        const int b0 = c + kPrefixRun * (int)threadIdx.x;
        uint64_t a[kPrefixRun];
#pragma unroll
        for (int k = 0; k < kPrefixRun; ++k) a[k] = b0 + k < nt ? part_q[t0 + b0 + k] : 0;
        uint64_t run = 0;
#pragma unroll
        for (int k = 0; k < kPrefixRun; ++k) run += a[k];
        uint64_t inc = run;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t y = __shfl_up(inc, d, 64);
            if (lane >= d) inc += y;
        }
        __syncthreads();  // (s_w of the previous chunk has been read)
        if (lane == 63) s_w[wv] = inc;
        __syncthreads();
        uint64_t pre = carry, tot = 0;
#pragma unroll
        for (int w2 = 0; w2 < kFinTB / 64; ++w2) {
            if (w2 < wv) pre += s_w[w2];
            tot += s_w[w2];
        }
        pre += inc - run;
#pragma unroll
        for (int k = 0; k < kPrefixRun; ++k) {
            if (b0 + k < nt) tpre[t0 + b0 + k] = pre;
            pre += a[k];
        }
        carry += tot;
    }
}

// tpre (non-null): the exclusive prefix of each tile of a filter above kPrefixTiles tiles, from
// k_tile_prefix; smaller filters sum their earlier aggregates here.
__global__ __launch_bounds__(kTB) void k_scan_agg(Layout L, FilterDev* F, const float* __restrict__ w,
                                                  const double* __restrict__ part_n,
                                                  const uint64_t* __restrict__ part_q, uint64_t* __restrict__ cum,
                                                  uint64_t* __restrict__ tcum, const float* __restrict__ ratio,
                                                  const uint64_t* __restrict__ tpre) {
    __shared__ uint64_t s_w[4];
    __shared__ double s_vw[kFinTB / 64];
    const int tile = blockIdx.x;
    const int f = layout_filter(L, tile * (kTile / kTB));
    const int off = L.f_off[f], nf = L.f_n[f];
    const int t0 = off / kTile, nt = L.f_padn[f] / kTile;
    const bool last = tile == t0 + nt - 1;
    // the tile's weights, and for the last tile of a filter of up to 1024 tiles its virtual
    // lanes' single w^2 partials (0.0 + x = x: the same bits as the loop below), are loaded
    // together with the aggregates, so the tile waits for one memory round trip
    const float4 wv = tile_w(w, tile);
    double pn[kFinTB / kTB];
#pragma unroll
    for (int j = 0; j < kFinTB / kTB; ++j) {
        const int vl = (int)threadIdx.x + kTB * j;
        pn[j] = last && nt <= kFinTB && vl < nt ? part_n[t0 + vl] : 0.0;
    }
    uint64_t excl;
    if (tpre && nt > kPrefixTiles) {
        excl = tpre[tile];
    } else {
        uint64_t ex = 0;
        for (int c = t0; c < tile; c += 4 * kTB) {  // (four aggregates per lane in flight)
            uint64_t a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int b = c + (int)threadIdx.x + u * kTB;
                a[u] = b < tile ? part_q[b] : 0;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) ex += a[u];
        }
        excl = block_sum_u64(ex, s_w);
    }
    uint64_t q[4];
    tile_q(wv, tile, off, nf, q);
    q[1] += q[0];
    q[2] += q[1];
    q[3] += q[2];
    uint64_t incl;
    const uint64_t agg = block_scan_u64(q[3], incl, s_w);
    const uint64_t pre = excl + incl - q[3];
    uint64_t* cp = cum + tile * kTile + (int)threadIdx.x * 4;
    cp[0] = pre + q[0];
    cp[1] = pre + q[1];
    cp[2] = pre + q[2];
    cp[3] = pre + q[3];
    if (threadIdx.x == kTB - 1) tcum[tile] = pre + q[3];
    if (!last) return;
#pragma unroll
    for (int j = 0; j < kFinTB / kTB; ++j) {
        const int vl = (int)threadIdx.x + kTB * j;
        double v = pn[j];
        if (nt > kFinTB) {
            v = 0.0;
            for (int b = vl; b < nt; b += kFinTB) v += part_n[t0 + b];
        }
        v = wave_sum(v);
        if ((threadIdx.x & 63) == 0) s_vw[vl >> 6] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tot = s_vw[0];
#pragma unroll
        for (int k = 1; k < kFinTB / 64; ++k) tot += s_vw[k];
        FilterDev& fd = F[f];
        const uint64_t T = excl + agg;
        fd.total = T;
        fd.est_src = 0;  // weights changed
        fd.sum_w2 = tot;
        fd.ess = tot > 0.0 ? 1.0 / tot : 0.0;
        if (ratio) {
            const double r = (double)ratio[f];
            fd.do_resample = (r > 0.0 && fd.ess < r * (double)nf) ? 1 : 0;
        } else {
            fd.do_resample = 1;  // reference ABI: pfc_resample always resamples (cu:638-676)
        }
        if (fd.do_resample && T > 0) {
            const uint64_t lcg = fd.lcg * 6364136223846793005ull + 1442695040888963407ull;  // cu:653
            fd.lcg = lcg;
            fd.u32 = (uint32_t)(lcg >> 32);
        }
    }
}

// Systematic pick: ancestor(i) = first j with cum_j * n 2^32 >= T (U32 + i 2^32) -- the
// pointer t_i = (U + i) T / n of cpu.rs:276-285 compared exactly.  With K_i = ceil(T (U32 +
// i 2^32) / 2^32) (computed once per output) the test cum_j n 2^32 < T (U32 + i 2^32) is
// cum_j n < K_i: one 64 x 32-bit product per probe (at most 96 bits: cum < 2^62, n < 2^31), or
// none when the window's cum_j n are staged.
struct U128 {
    uint64_t hi, lo;
};
__device__ inline bool lt128(const U128& a, const U128& b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
__device__ inline U128 mul_n(uint64_t c, uint32_t n) {
    const uint64_t a = (uint64_t)(uint32_t)c * n;  // low half of c times n
    const uint64_t b = (c >> 32) * n;              // high half of c times n
    U128 r;
    r.lo = a + (b << 32);
    r.hi = (b >> 32) + (r.lo < a ? 1u : 0u);
    return r;
}
__device__ inline U128 pointer_key(uint64_t T, uint64_t Bi) {
    const uint64_t plo = T * Bi, phi = __umul64hi(T, Bi);
    U128 k;
    k.lo = (phi << 32) | (plo >> 32);
    k.hi = phi >> 32;
    if (plo & 0xffffffffull) {
        k.lo += 1;
        if (k.lo == 0) k.hi += 1;
    }
    return k;
}

// Cross-lane folds of 64-bit values on CDNA4 (one 32-bit permutation per half).  Rows are 16
// lanes.  fold32(a, b): v_permlane32_swap exchanges rows 2-3 of a with rows 0-1 of b, so the
// sum leaves a(l) + a(l+32) in lanes 0-31 and b(l-32) + b(l) in lanes 32-63.  fold16(a, b):
// v_permlane16_swap exchanges the odd rows of a with the even rows of b, leaving a summed over
// row pairs (0,1), (2,3) in rows 0 and 2 and b likewise in rows 1 and 3.
__device__ inline double join64(unsigned lo, unsigned hi) {
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | (uint64_t)lo);
}
__device__ inline double fold32(double a, double b) {
    const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    const auto l = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    return join64(l[0], h[0]) + join64(l[1], h[1]);
}
__device__ inline double fold16(double a, double b) {
    const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    const auto l = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32), false, false);
    return join64(l[0], h[0]) + join64(l[1], h[1]);
}
// x of the lane n places further along its row (DPP row_ror).
template <int N>
__device__ inline double row_ror(double x) {
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, 0x120 + N, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), 0x120 + N, 0xf, 0xf, false);
    return join64((unsigned)lo, (unsigned)hi);
}

// Wave sums of NV values per lane (every lane active): the values are folded across the halves
// and then across row pairs (value i with value i + H, so no lane holds more than
// H2 = ceil(NV / 4) of them), then summed within each row by DPP rotations.  wave_fold_store then
// has lane 16 r store the sums of the values base_r .. base_r + H2 - 1 of its row into row[].
template <int NV>
struct Fold {
    static constexpr int H1 = (NV + 1) / 2, H2 = (H1 + 1) / 2;
};
template <int NV>
__device__ inline void wave_fold_store(const double (&acc)[NV], double* __restrict__ row, int rs = 1) {
    constexpr int H1 = Fold<NV>::H1, H2 = Fold<NV>::H2;
    const int lane = threadIdx.x & 63;
    double s1[H1], s2[H2];
#pragma unroll
    for (int i = 0; i < H1; ++i) s1[i] = fold32(acc[i], i + H1 < NV ? acc[i + H1] : 0.0);
#pragma unroll
    for (int i = 0; i < H2; ++i) s2[i] = fold16(s1[i], i + H2 < H1 ? s1[i + H2] : 0.0);
#pragma unroll
    for (int i = 0; i < H2; ++i) {
        s2[i] += row_ror<8>(s2[i]);
        s2[i] += row_ror<4>(s2[i]);
        s2[i] += row_ror<2>(s2[i]);
        s2[i] += row_ror<1>(s2[i]);
    }
    if ((lane & 15) == 0) {
        const int r = lane >> 4;  // rows 0, 1: values [0, H1) split at H2; rows 2, 3: [H1, NV) likewise
        const int base = r == 0 ? 0 : r == 1 ? H2 : r == 2 ? H1 : H1 + H2;
        const int end = r == 0 ? H2 : r == 1 ? H1 : r == 2 ? (H1 + H2 < NV ? H1 + H2 : NV) : NV;
#pragma unroll
        for (int i = 0; i < H2; ++i)
            if (base + i < end) row[(size_t)(base + i) * rs] = s2[i];
    }
}

// Block partial of NV values (4 waves, every lane active): wave folds, then the four waves in a
// fixed order.  Value k of block blk goes to part[k * stride + blk], so that a finalize reading
// one value over consecutive blocks loads consecutive doubles.
template <int NV>
__device__ inline void block_partials(double (&acc)[NV], double* __restrict__ part, int blk, int stride) {
    __shared__ double shv[4][NV];
    wave_fold_store<NV>(acc, shv[threadIdx.x >> 6]);
    __syncthreads();
    if ((int)threadIdx.x < NV) {
        const int k = threadIdx.x;
        part[(size_t)k * stride + blk] = (shv[0][k] + shv[1][k]) + (shv[2][k] + shv[3][k]);
    }
}

// One particle's terms of the estimate sums (cu:354-404 layout): [0..2] w p, [3..12] w q_r q_c
// (r <= c), [13] w^2.
__device__ inline void est_terms(double wi, double px, double py, double pz, const double (&q)[4],
                                 double (&acc)[kEstN]) {
    acc[0] = __fma_rn(wi, px, acc[0]);
    acc[1] = __fma_rn(wi, py, acc[1]);
    acc[2] = __fma_rn(wi, pz, acc[2]);
    int k = 3;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const double wq = wi * q[r];
#pragma unroll
        for (int c = r; c < 4; ++c) {
            acc[k] = __fma_rn(wq, q[c], acc[k]);
            ++k;
        }
    }
    acc[13] = __fma_rn(wi, wi, acc[13]);
}

constexpr int kGatherBlk = kTile;              // a block per 1024-particle tile of the filter
constexpr int kMergeItems = 2 * kGatherBlk;    // merge-path items (outputs + particles) per block

// Jitter of the injection picks linked at head hd (an output's list), added in ascending pick
// order (cu:333-352 adds them one after another) in f64 to the gathered position v[0..2], which
// is then rounded once to f32.
__device__ inline void apply_picks(int off, int hd, const int32_t* __restrict__ next, const double* __restrict__ jit,
                                   float (&v)[7]) {
    double ax = (double)v[0], ay = (double)v[1], az = (double)v[2];
    // a single pick (almost every list): its link and its jitter read together, one round trip
    const int nx = next[off + hd];
    const double jx = jit[3 * (off + hd)], jy = jit[3 * (off + hd) + 1], jz = jit[3 * (off + hd) + 2];
    if (nx < 0) {
        v[0] = (float)(ax + jx);
        v[1] = (float)(ay + jy);
        v[2] = (float)(az + jz);
        return;
    }
    // ascending traversal of the (unordered) list: repeatedly take the smallest pick above the
    // previous one
    int prev = -1;
    while (true) {
        int best = 0x7fffffff;
        for (int q = hd; q >= 0; q = next[off + q])
            if (q > prev && q < best) best = q;
        if (best == 0x7fffffff) break;
        ax += jit[3 * (off + best)];
        ay += jit[3 * (off + best) + 1];
        az += jit[3 * (off + best) + 2];
        prev = best;
    }
    v[0] = (float)ax;
    v[1] = (float)ay;
    v[2] = (float)az;
}

// Merge-path gather.  The outputs (pointer keys K_i = K_0 + i T) and the particles (cum_j n)
// form one merged sequence of 2n items, output i ahead of particle j when K_i <= cum_j n, so that
// ancestor(i) is the first particle behind output i; block b of a filter takes items [2048 b,
// 2048 b + 2048) of it (every block has 2048 items whatever the weights: with a few hundred
// heavy particles the ancestors of 1024 consecutive outputs can span thousands of particles), and
// the grid is one block per tile.  J(d), the number of particles among the first d items, is
// found for the block's two ends in two memory round trips: every tile's last cumulative weight
// (tcum, left by the scan) shows which tile holds J(d) (a particle is among the first d items
// when the output d - 1 - e is not ahead of it); those tiles -- at most three, J(d1) - J(d0) <=
// 2048 -- are staged in LDS, where J(d0), J(d1) and then every output's ancestor are found, the
// latter by a merge of eight items per lane.  Outputs are then moved one per lane (consecutive
// lanes, consecutive outputs: the pose loads and stores coalesce), a lane's loads in flight
// together.  Poses move from the current set to the other set (no copies); weights -> 1/n.  An
// output that injection picks target (linked into a list at head[output]) adds their jitters
// (apply_picks).  Every block then leaves the estimate-sum partial of its outputs' final poses
// and weights (est_src = 1); for a filter that does not resample (its current poses and weights)
// or whose total is zero (weights set to 1/n) the block takes the tile's particles instead, so
// the estimate needs no separate pass.  The other set becomes current afterwards (k_flip, or the
// sums finalize of the fused cycle).
constexpr int kGatherWin = 3 * kTile;  // staged cumulative weights: the tiles from J(d0)'s on
constexpr int kGatherBatch = 4;       // outputs per lane whose loads are in flight together

__device__ inline U128 add128(const U128& a, const U128& b) {
    U128 r;
    r.lo = a.lo + b.lo;
    r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
    return r;
}
// Particle e (filter-local, key cne = cum_e n) lies among the first d merged items: fewer than
// d - e outputs are ahead of it, i.e. output d - 1 - e (if there is one) is not.
__device__ inline bool among_first(int e, int d, const U128& cne, int n, uint64_t T, const U128& K0) {
    const int i = d - 1 - e;
    if (i < 0) return false;
    if (i >= n) return true;
    return lt128(cne, add128(K0, mul_n(T, (uint32_t)i)));
}

__global__ __launch_bounds__(kTB) void k_gather(Layout L, FilterDev* F, PoseBufs B, const uint64_t* __restrict__ cum,
                                                const uint64_t* __restrict__ tcum, float* __restrict__ w,
                                                int32_t* __restrict__ anc, int32_t* __restrict__ head,
                                                const int32_t* __restrict__ next, const double* __restrict__ jit,
                                                double* __restrict__ part_est, int stride) {
    __shared__ uint64_t s_c[kGatherWin];      // cum of the staged tiles
    // ancestor of each of the block's outputs, less J(d0) (ancestors span at most 2049 particles):
    // 16-bit entries keep the block's LDS at 29 KB, five blocks per CU
    __shared__ uint16_t s_anc[kMergeItems];
    __shared__ int s_first[3][2];              // tile search: first probe past J(d), per round and end
    __shared__ int s_j[2];                     // J(d0), J(d1)
    const int blk = blockIdx.x;
    const int f = layout_filter(L, blk * (kGatherBlk / kTB));
    FilterDev& fd = F[f];
    const int off = L.f_off[f];
    const int nf = L.f_n[f];
    const int i0 = blk * kGatherBlk - off;
    const int t = (int)threadIdx.x;
    const uint64_t T = fd.total;
    const int par = fd.parity;
    const float winv = L.f_winv[f];
    const PoseSet S = B.s[par];
    double acc[kEstN];
#pragma unroll
    for (int k = 0; k < kEstN; ++k) acc[k] = 0.0;
    if (!fd.do_resample || T == 0) {
        // no poses move; cu:647-650: a degenerate total only makes the weights uniform.  Picks
        // the fused scan drew in advance are dropped.
#pragma unroll
        for (int e = 0; e < kTile / kTB; ++e) {
            const int li = i0 + t + e * kTB;
            if (li < nf) {
                const int gi = off + li;
                if (head[gi] >= 0) head[gi] = -1;
                float wi;
                if (fd.do_resample) {
                    wi = winv;
                    w[gi] = wi;
                } else {
                    wi = w[gi];
                }
                const double q[4] = {(double)S.c[3][gi], (double)S.c[4][gi], (double)S.c[5][gi], (double)S.c[6][gi]};
                est_terms((double)wi, (double)S.c[0][gi], (double)S.c[1][gi], (double)S.c[2][gi], q, acc);
            }
        }
        block_partials<kEstN>(acc, part_est, blk, stride);
        if (i0 == 0 && t == 0) fd.est_src = 1;
        return;
    }
    const uint32_t n32 = (uint32_t)nf;
    const uint64_t U = (uint64_t)fd.u32;
    const uint64_t* c = cum + off;
    const int t0 = off / kTile, nt = L.f_padn[f] / kTile;
    const int d0 = min((blk - t0) * kMergeItems, 2 * nf), d1 = min(d0 + kMergeItems, 2 * nf);
    if (d0 < d1) {
        // K_i = K_0 + i T with K_0 = ceil(T U / 2^32): no 64 x 64-bit product per output
        const U128 TU = mul_n(T, (uint32_t)U);
        U128 K0;
        K0.lo = (TU.hi << 32) | (TU.lo >> 32);
        K0.hi = TU.hi >> 32;
        if (TU.lo & 0xffffffffull) {
            K0.lo += 1;
            if (K0.lo == 0) K0.hi += 1;
        }
        // 1. The tile holding J(d) for d = d0 (threads 0-127) and d = d1 (128-255): the first tile
        //    whose last particle is not among the first d items (nt: every particle is).  Each
        //    round probes 1024 evenly spaced tiles, eight loads per lane in flight; one round
        //    settles up to 1024 tiles (1M particles), two up to 2^20 tiles.
        const int half = t >> 7, ht = t & 127;
        if (t < 6) s_first[t >> 1][t & 1] = INT_MAX;
        __syncthreads();
        const int rounds = nt <= 1024 ? 1 : nt <= (1 << 20) ? 2 : 3;
        {
            const int d = half ? d1 : d0;
            int lo = 0, hi = nt;  // the answer lies in [lo, hi]; tiles below lo are among the first d
            for (int r = 0; r < rounds; ++r) {
                const int span = hi - lo;
                const int step = (span + 1023) >> 10;
                uint64_t tv[8];
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    const int k = ht + 128 * p;
                    tv[p] = span > 0 && k * step < span ? tcum[t0 + lo + k * step] : 0;
                }
                int kmin = INT_MAX;
#pragma unroll
                for (int p = 0; p < 8; ++p) {
                    const int k = ht + 128 * p;
                    if (span > 0 && k * step < span && k < kmin) {
                        const int e = min((lo + k * step) * kTile + kTile - 1, nf - 1);
                        if (!among_first(e, d, mul_n(tv[p], n32), nf, T, K0)) kmin = k;
                    }
                }
                if (kmin != INT_MAX) atomicMin(&s_first[r][half], kmin);
                __syncthreads();
                const int kf = s_first[r][half];
                if (span > 0) {
                    if (kf == INT_MAX) {
                        lo += ((span - 1) / step) * step + 1;
                    } else {
                        hi = lo + kf * step;
                        lo = kf == 0 ? lo : lo + (kf - 1) * step + 1;
                    }
                }
            }
            if (ht == 0) s_j[half] = lo;  // the tile (lo == hi after the last round)
        }
        __syncthreads();
        const int ta = s_j[0], tb = s_j[1];
        // 2. Stage the tiles from J(d0)'s on (J(d1) <= J(d0) + 2048 lies at most two tiles further).
        const int wb = ta * kTile;  // filter-local index of s_c[0]
        const int wn = min(kGatherWin, nt * kTile - wb);
        {
            uint64_t cv[kGatherWin / kTB];
#pragma unroll
            for (int u = 0; u < kGatherWin / kTB; ++u) {
                const int x = t + kTB * u;
                cv[u] = x < wn ? c[wb + x] : 0;
            }
#pragma unroll
            for (int u = 0; u < kGatherWin / kTB; ++u) s_c[t + kTB * u] = cv[u];
        }
        __syncthreads();
        // J(d0) and J(d1) inside their tiles (wave 0 and wave 1): a 64-ary search, then the exact
        // particle among the 16 the first step leaves.
        const int wv = t >> 6, lane = t & 63;
        if (wv < 2) {
            const int d = wv ? d1 : d0;
            const int tile = wv ? tb : ta;
            int J = nf;
            if (tile < nt) {
                const int A = tile * kTile, E = min(A + kTile - 1, nf - 1);
                auto beyond = [&](int e) {  // e is not among the first d items
                    return !among_first(e, d, mul_n(s_c[min(e - wb, kGatherWin - 1)], n32), nf, T, K0);
                };
                const int e1 = min(A + 16 * lane + 15, E);
                const uint64_t m1 = __ballot(A + 16 * lane <= E && beyond(e1));
                const int l1 = m1 ? __ffsll((long long)m1) - 1 : 0;
                const int e2 = A + 16 * l1 + lane;
                const uint64_t m2 = __ballot(lane < 16 && e2 <= E && beyond(e2));
                J = m2 ? A + 16 * l1 + __ffsll((long long)m2) - 1 : E;
            }
            if (lane == 0) s_j[wv] = J;
        }
        __syncthreads();
        const int ja = s_j[0], jb = s_j[1];
        const int ia = d0 - ja, ib = d1 - jb;  // outputs [ia, ib), ancestors in [ja, jb]
        // 3. Every output's ancestor: lane t merges items [8 t, 8 t + 8) of the block's 2048.
        {
            const int m = d1 - d0;
            int d = d0 + 8 * t;
            if (d < d1) {
                int lo = ja, hi = jb;  // first particle not among the first d items
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (among_first(mid, d, mul_n(s_c[mid - wb], n32), nf, T, K0))
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                int e = lo, i = d - lo;
                U128 K = add128(K0, mul_n(T, (uint32_t)min(i, nf - 1)));
                const int dend = min(d0 + 8 * t + 8, d0 + m);
                for (; d < dend; ++d) {
                    const bool out = i < nf && (e >= nf || !lt128(mul_n(s_c[min(e - wb, kGatherWin - 1)], n32), K));
                    if (out) {
                        s_anc[i - ia] = (uint16_t)(min(e, nf - 1) - ja);
                        ++i;
                        K = add128(K, U128{0, T});
                    } else {
                        ++e;
                    }
                }
            }
        }
        __syncthreads();
        // 4. Move the outputs, one per lane, kGatherBatch per lane with their loads in flight.
        const PoseSet D = B.s[par ^ 1];
        const int no = ib - ia;
        for (int u0 = 0; u0 < no; u0 += kTB * kGatherBatch) {
            int jj[kGatherBatch], hd[kGatherBatch];
            float v[kGatherBatch][7];
#pragma unroll
            for (int r = 0; r < kGatherBatch; ++r) {
                const int u = u0 + t + kTB * r;
                jj[r] = -1;
                hd[r] = -1;
                if (u < no) {
                    const int j = min(ja + (int)s_anc[u], nf - 1);
                    jj[r] = j;
                    hd[r] = head[off + ia + u];
#pragma unroll
                    for (int k = 0; k < 7; ++k) v[r][k] = S.c[k][off + j];
                }
            }
#pragma unroll
            for (int r = 0; r < kGatherBatch; ++r) {
                if (jj[r] < 0) continue;
                const int o = off + ia + u0 + t + kTB * r;
                if (hd[r] >= 0) {
                    apply_picks(off, hd[r], next, jit, v[r]);
                    head[o] = -1;
                }
#pragma unroll
                for (int k = 0; k < 7; ++k) D.c[k][o] = v[r][k];
                anc[o] = jj[r];
                w[o] = winv;
                const double q[4] = {(double)v[r][3], (double)v[r][4], (double)v[r][5], (double)v[r][6]};
                est_terms((double)winv, (double)v[r][0], (double)v[r][1], (double)v[r][2], q, acc);
            }
        }
    }
    block_partials<kEstN>(acc, part_est, blk, stride);
    if (i0 == 0 && t == 0) fd.est_src = 1;
}

// Injection picks of a resample that the scan did not draw (reference ABI path), after the scan
// advanced the LCG.  Runs over the particle-layout blocks that can hold picks: tab lists them
// (blocks whose first local index is below the filter's largest pick count), or, when tab is
// null, the grid is the first ceil(k / 256) blocks of a single filter.
__global__ __launch_bounds__(kTB) void k_inj_pick(Layout L, FilterDev* F, ParamRef<InjectParams> A, const int32_t* __restrict__ tab,
                                                  double* __restrict__ jit, int32_t* __restrict__ head,
                                                  int32_t* __restrict__ next) {
    const int blk = tab ? tab[blockIdx.x] : (int)blockIdx.x;
    const int f = layout_filter(L, blk);
    const FilterDev& fd = F[f];
    const InjectParams& p = A.get(f);
    const int off = L.f_off[f];
    const int j = blk * kTB + (int)threadIdx.x - off;
    if (!fd.do_resample || fd.total == 0 || j >= p.k) return;
    inj_pick(off, L.f_n[f], fd.lcg, p.gen, j, jit, head, next);
}

// The gathered set becomes current for filters that resampled with a positive total; the
// scan's status epoch advances for the next resample.
__global__ void k_flip(FilterDev* F, int nfilt, unsigned long long* scan_ctl) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f < nfilt && F[f].do_resample && F[f].total > 0) F[f].parity ^= 1;
    if (f == 0) scan_ctl[0] = (scan_ctl[0] + 1ull) & 0x3fffffffull;
}

// ---------------------------------------------------------------------------- estimate

// Per filter (one 512-thread block each, so that the register budget holds the loads in flight
// and the Markley mean without spills): lane t adds, for every value, the filter's 1024-particle
// tile partials t, t + 512, ... in that order (value k of tile b at part[k * stride + b]; four
// tiles' worth of loads in flight), then the 8 waves are folded and summed in a fixed order.
// which = 0 -> sums, 1 -> covariances.  With cov
// non-null (sums only) thread 0 then takes the Markley mean of the finished sums and writes the
// covariance pass parameters.  With flip (sums of the fused cycle) the resampled set becomes
// current and the scan's status epoch advances, as k_flip.  With out non-null (covariances of a
// staged cycle) the filter's scalars are copied to out[f]: the other members from F[f] (final
// since earlier kernels), the new values by the lanes that hold them.
constexpr int kEstFinTB = 512;

// Lane vt of the estimate finalize: the sum, in order, of the partials of tiles c + vt + u
// kEstFinTB (u = 0..3, c = 0, 4 kEstFinTB, ...) of one filter (nb tiles; a missing tile reads as
// 0, which leaves the sums unchanged).  kU tiles' loads are in flight at a time; the arithmetic
// is the same for any kU.
template <int NV, int kU>
__device__ inline void fin_lane_sums(const double* __restrict__ pf, int stride, int nb, int vt, double (&acc)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (int c = 0; c < nb; c += 4 * kEstFinTB) {
#pragma unroll 1
        for (int u0 = 0; u0 < 4; u0 += kU) {
            double x[kU][NV];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int b = c + vt + (u0 + u) * kEstFinTB;
#pragma unroll
                for (int k = 0; k < NV; ++k) x[u][k] = b < nb ? pf[(size_t)k * stride + b] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
#pragma unroll
                for (int k = 0; k < NV; ++k) acc[k] += x[u][k];
        }
    }
}

// The same sums over per-wave partials (4 per tile, value k of tile b's wave v at
// k * 4 stride + 4 b + v): each tile's value is formed as (v0 + v1) + (v2 + v3), block_partials'
// order, before it joins the lane's sum, so both layouts give the same bits.
template <int NV, int kU>
__device__ inline void fin_lane_sums_w4(const double* __restrict__ pf, int stride, int nb, int vt,
                                        double (&acc)[NV]) {
#pragma unroll
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    for (int c = 0; c < nb; c += 4 * kEstFinTB) {
#pragma unroll 1
        for (int u0 = 0; u0 < 4; u0 += kU) {
            double x[kU][NV];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int b = c + vt + (u0 + u) * kEstFinTB;
#pragma unroll
                for (int k = 0; k < NV; ++k) {
                    if (b < nb) {
                        const double2 lo = *reinterpret_cast<const double2*>(pf + (size_t)k * 4 * stride + 4 * b);
                        const double2 hi = *reinterpret_cast<const double2*>(pf + (size_t)k * 4 * stride + 4 * b + 2);
                        x[u][k] = (lo.x + lo.y) + (hi.x + hi.y);
                    } else {
                        x[u][k] = 0.0;
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kU; ++u)
#pragma unroll
                for (int k = 0; k < NV; ++k) acc[k] += x[u][k];
        }
    }
}

template <int NV>
__global__ __launch_bounds__(kEstFinTB) void k_est_fin(Layout L, FilterDev* F, const double* __restrict__ part,
                                                       int stride, int which, CovParams* cov, int flip,
                                                       unsigned long long* scan_ctl, FilterDev* out) {
    __shared__ double shv[kEstFinTB / 64][NV];
    __shared__ double sv[NV];
    const int f = blockIdx.x;
    const int b0 = L.f_off[f] / kTile, nb = L.f_padn[f] / kTile;
    const int t = (int)threadIdx.x;
    double acc[NV];
    fin_lane_sums<NV, 4>(part + b0, stride, nb, t, acc);
    wave_fold_store<NV>(acc, shv[t >> 6]);
    __syncthreads();
    if (t < NV) {
        double v = shv[0][t];
#pragma unroll
        for (int wv = 1; wv < kEstFinTB / 64; ++wv) v += shv[wv][t];
        (which == 0 ? F[f].sums : F[f].covs)[t] = v;
        if (out) (which == 0 ? out[f].sums : out[f].covs)[t] = v;
        sv[t] = v;
    }
    if (out) {
        static_assert(sizeof(FilterDev) % sizeof(uint64_t) == 0, "FilterDev copied as 64-bit words");
        constexpr int kWords = (int)(sizeof(FilterDev) / sizeof(uint64_t));
        const int v0 = (int)((which == 0 ? offsetof(FilterDev, sums) : offsetof(FilterDev, covs)) / sizeof(uint64_t));
        if (t < kWords && (t < v0 || t >= v0 + NV))
            reinterpret_cast<uint64_t*>(out + f)[t] = reinterpret_cast<const uint64_t*>(F + f)[t];
    }
    if (flip && threadIdx.x == 0) {
        FilterDev& fd = F[f];
        if (fd.do_resample && fd.total > 0) fd.parity ^= 1;
        if (f == 0) scan_ctl[0] = (scan_ctl[0] + 1ull) & 0x3fffffffull;
    }
    if constexpr (NV == kEstN) {  // (the covariance instantiation carries no Markley code)
        if (!cov) return;
        __syncthreads();
        if (threadIdx.x == 0) markley_mean(F[f], sv, cov[f]);
    }
}

// The covariance finalize of a staged cycle, split over kCovSplit workgroups per filter: workgroup
// r of filter f computes lane group r of k_est_fin<kCovN>'s lanes (lanes vt = 64 r + l), whose
// lane sums run over the tiles vt + 512 u, u = 0, 1, ... (4 per 2048 tiles, tiles past the filter
// counted as 0.0: fin_lane_sums' additions exactly) of the per-wave partials the deferred tiles
// left.  All 1024 threads load and pair up the four per-wave partials of kCovStage rounds of
// those tiles into LDS, then the group's wave adds them in order; the wave's fold goes to
// fold[f][r], and the next kernel adds the kCovSplit folds in k_est_fin's wave order
// (cov_fold_sum).  (One wave loading its own tiles took 116 us at 16M particles.)  Group 0 also
// copies every scalar of the filter to out[f] except the likelihood scalars, the word holding the
// pose parity (which the softmax finalize running beside it rewrites for the next cycle) and the
// covariances.
constexpr int kCovSplit = kEstFinTB / 64;
constexpr int kCovStage = 8;  // rounds of tiles staged at a time: 8 x 12 x 64 doubles (48 KB)
__device__ void cov_fin_wave(const Layout& L, FilterDev* F, const double* __restrict__ part, int stride,
                             FilterDev* out, double* __restrict__ fold, int r) {
    __shared__ double s_t[kCovStage][kCovN][64];
    const int f = r / kCovSplit, j = r % kCovSplit;
    const int b0 = L.f_off[f] / kTile, nb = L.f_padn[f] / kTile;
    const int t = (int)threadIdx.x, lane = t & 63;
    const int nround = 4 * ((nb + 4 * kEstFinTB - 1) / (4 * kEstFinTB));
    const double* pf = part + 4 * (size_t)b0;
    double acc[kCovN];
#pragma unroll
    for (int k = 0; k < kCovN; ++k) acc[k] = 0.0;
    for (int u0 = 0; u0 < nround; u0 += kCovStage) {
        // item i = t + kFinTB m: lane i % 64, value (i / 64) % kCovN, round u0 + i / (64 kCovN)
        constexpr int kItems = kCovStage * kCovN * 64 / kFinTB;
        double2 lo[kItems], hi[kItems];
#pragma unroll
        for (int m = 0; m < kItems; ++m) {
            const int i = t + kFinTB * m, l = i & 63, k = (i >> 6) % kCovN, u = u0 + (i >> 6) / kCovN;
            const int b = 64 * j + l + kEstFinTB * u;
            if (u < nround && b < nb) {
                lo[m] = *reinterpret_cast<const double2*>(pf + (size_t)k * 4 * stride + 4 * b);
                hi[m] = *reinterpret_cast<const double2*>(pf + (size_t)k * 4 * stride + 4 * b + 2);
            } else {
                lo[m] = hi[m] = make_double2(0.0, 0.0);
            }
        }
        __syncthreads();  // (the previous stage has been added)
#pragma unroll
        for (int m = 0; m < kItems; ++m) {
            const int i = t + kFinTB * m, l = i & 63, k = (i >> 6) % kCovN, uu = (i >> 6) / kCovN;
            s_t[uu][k][l] = (lo[m].x + lo[m].y) + (hi[m].x + hi[m].y);
        }
        __syncthreads();
        if (t < 64)
            for (int uu = 0; uu < kCovStage && u0 + uu < nround; ++uu)
#pragma unroll
                for (int k = 0; k < kCovN; ++k) acc[k] += s_t[uu][k][lane];
    }
    if (t >= 64) return;
    wave_fold_store<kCovN>(acc, fold + ((size_t)f * kCovSplit + j) * kCovN);
    if (j != 0) return;
    constexpr int kWords = (int)(sizeof(FilterDev) / sizeof(uint64_t));
    constexpr int c0 = (int)(offsetof(FilterDev, covs) / sizeof(uint64_t));
    constexpr int l1 = (int)(offsetof(FilterDev, sum_w2) / sizeof(uint64_t));  // words [0, l1): likelihood
    constexpr int pw = (int)(offsetof(FilterDev, parity) / sizeof(uint64_t));
    static_assert(offsetof(FilterDev, max_ll) == 0 && offsetof(FilterDev, uniform) / 8 == pw, "FilterDev words");
    for (int t2 = lane; t2 < kWords; t2 += 64)
        if (t2 >= l1 && t2 != pw && (t2 < c0 || t2 >= c0 + kCovN))
            reinterpret_cast<uint64_t*>(out + f)[t2] = reinterpret_cast<const uint64_t*>(F + f)[t2];
}
// The covariances of filter f from the kCovSplit folds cov_fin_wave stored (lanes t < kCovN).
__device__ void cov_fold_sum(FilterDev* F, const double* __restrict__ fold, FilterDev* out, int f) {
    const int t = (int)threadIdx.x;
    if (t >= kCovN) return;
    const double* fr = fold + (size_t)f * kCovSplit * kCovN;
    double v = fr[t];
#pragma unroll
    for (int wv = 1; wv < kCovSplit; ++wv) v += fr[wv * kCovN + t];
    F[f].covs[t] = v;
    out[f].covs[t] = v;
}

// est layout (cu:354-404): [0..2] sum w p, [3..12] sum w q_r q_c (r<=c), [13] sum w^2.
// Filters whose partials the resampling gather already left (est_src = 1) are skipped.
// 1024-particle tiles, lane t taking particles t, t + 256, t + 512, t + 768 of the tile: the
// terms, order and block tree of k_gather's partials, so that either pass gives the same bits
// for the same poses and weights.
__global__ __launch_bounds__(kTB) void k_est_sums(Layout L, FilterDev* F, PoseBufs B, const float* __restrict__ w,
                                                  double* __restrict__ part, int stride) {
    const int blk = blockIdx.x;
    const int f = layout_filter(L, blk * (kTile / kTB));
    if (F[f].est_src) return;
    const int off = L.f_off[f];
    const int nf = L.f_n[f];
    const int i0 = blk * kTile - off;
    const PoseSet S = B.s[F[f].parity];
    double acc[kEstN];
#pragma unroll
    for (int k = 0; k < kEstN; ++k) acc[k] = 0.0;
#pragma unroll
    for (int e = 0; e < kTile / kTB; ++e) {
        const int li = i0 + (int)threadIdx.x + e * kTB;
        if (li < nf) {
            const int gi = off + li;
            const double q[4] = {(double)S.c[3][gi], (double)S.c[4][gi], (double)S.c[5][gi], (double)S.c[6][gi]};
            est_terms((double)w[gi], (double)S.c[0][gi], (double)S.c[1][gi], (double)S.c[2][gi], q, acc);
        }
    }
    block_partials<kEstN>(acc, part, blk, stride);
}

// cov layout (cu:406-471): [0..5] position, [6..11] attitude tangent (axis-angle of qm^-1 q,
// canonical sign, angle = 2 atan2(|v|, w)).
template <bool kWave>
__device__ void covs_tile(const Layout& L, FilterDev* F, const PoseBufs& B, int swapped, const float* __restrict__ w,
                          double* __restrict__ part, int stride, const ParamRef<CovParams>& A, int blk) {
    const int f = layout_filter(L, blk * (kTile / kTB));
    const int off = L.f_off[f];
    const int nf = L.f_n[f];
    const CovParams& p = A.get(f);
    const PoseSet S = B.s[F[f].parity ^ swapped];  // the current set
    double acc[kCovN];
#pragma unroll
    for (int k = 0; k < kCovN; ++k) acc[k] = 0.0;
    const int base = blk * kTile + (int)threadIdx.x * 4;
    const int li0 = base - off;
    if (li0 < nf) {
        float4 v[8];
        v[7] = *reinterpret_cast<const float4*>(w + base);
#pragma unroll
        for (int c = 0; c < 7; ++c) v[c] = *reinterpret_cast<const float4*>(S.c[c] + base);
        const float* vv = reinterpret_cast<const float*>(v);
        const dquat mc{p.mq[0], -p.mq[1], -p.mq[2], -p.mq[3]};
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            if (li0 + e < nf) {
                const double wi = (double)vv[28 + e];
                const double dp[3] = {(double)vv[0 + e] - p.mp[0], (double)vv[4 + e] - p.mp[1],
                                      (double)vv[8 + e] - p.mp[2]};
                int k = 0;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const double wd = wi * dp[r];
#pragma unroll
                    for (int c = r; c < 3; ++c) {
                        acc[k] = __fma_rn(wd, dp[c], acc[k]);
                        ++k;
                    }
                }
                const dquat qi{(double)vv[12 + e], (double)vv[16 + e], (double)vv[20 + e], (double)vv[24 + e]};
                dquat d = qmul(mc, qi);
                if (d.w < 0.0) {
                    d.w = -d.w;
                    d.x = -d.x;
                    d.y = -d.y;
                    d.z = -d.z;
                }
                const double hv2 = d.x * d.x + d.y * d.y + d.z * d.z;
                double aav[3] = {0.0, 0.0, 0.0};
                if (hv2 > 1e-24) {
                    // angle / |v| = 2 atan2(|v|, w) / |v|; for a half-angle tangent z = |v| / w <= 1/8
                    // that is (2 / w) atan(z) / z = (2 / w) sum_k (-1)^k z^2k / (2k + 1), nine terms
                    // (the first omitted one is below 3e-18), with no square root or arctangent
                    double ratio;
                    if (d.w > 0.0 && hv2 <= 0.015625 * (d.w * d.w)) {
                        const double rw = 1.0 / d.w;
                        const double z2 = hv2 * (rw * rw);
                        double sr = 1.0 / 17.0;
                        sr = __fma_rn(sr, -z2, sconst(1.0 / 15.0));
                        sr = __fma_rn(sr, -z2, sconst(1.0 / 13.0));
                        sr = __fma_rn(sr, -z2, sconst(1.0 / 11.0));
                        sr = __fma_rn(sr, -z2, sconst(1.0 / 9.0));
                        sr = __fma_rn(sr, -z2, sconst(1.0 / 7.0));
                        sr = __fma_rn(sr, -z2, sconst(1.0 / 5.0));
                        sr = __fma_rn(sr, -z2, sconst(1.0 / 3.0));
                        sr = __fma_rn(sr, -z2, 1.0);
                        ratio = 2.0 * rw * sr;
                    } else {
                        const double hv = sqrt(hv2);
                        ratio = 2.0 * atan2(hv, d.w) / hv;
                    }
                    aav[0] = d.x * ratio;
                    aav[1] = d.y * ratio;
                    aav[2] = d.z * ratio;
                }
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const double wa = wi * aav[r];
#pragma unroll
                    for (int c = r; c < 3; ++c) {
                        acc[k] = __fma_rn(wa, aav[c], acc[k]);
                        ++k;
                    }
                }
            }
        }
    }
    if constexpr (kWave)
        wave_fold_store<kCovN>(acc, part + 4 * blk + (threadIdx.x >> 6), 4 * stride);
    else
        block_partials<kCovN>(acc, part, blk, stride);
}

__global__ __launch_bounds__(kTB) void k_est_covs(Layout L, FilterDev* F, PoseBufs B, const float* __restrict__ w,
                                                  double* __restrict__ part, int stride, ParamRef<CovParams> A) {
    covs_tile<false>(L, F, B, 0, w, part, stride, A, blockIdx.x);
}

// Sets the per-filter resample predicate: force (ABI call) or ESS < ratio * n (fused cycle).
__global__ void k_resample_predicate(FilterDev* F, int nfilt, const float* ratio_ess, const int* fn, int force) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nfilt) return;
    FilterDev& fd = F[f];
    if (force) {
        fd.do_resample = 1;
    } else {
        const double r = (double)ratio_ess[f];
        fd.do_resample = (r > 0.0 && fd.ess < r * (double)fn[f]) ? 1 : 0;
    }
}

}  // namespace pfk

// ============================================================================ host side

namespace pfk {

constexpr int kMaxFilterN = 1 << 30;

int state_alloc(PfState* s, const std::vector<int>& ns) {
    // Particles are indexed with int: every filter's 2n merge-path items, and the padded total,
    // must fit (at most 2^30 particles per filter and in all); a rejected size leaves the state
    // as it was.
    long long np_total = 0;
    for (int n : ns) {
        if (n <= 0 || n > kMaxFilterN) return -1;
        np_total += (n + kTile - 1) / kTile * kTile;
    }
    if (ns.empty() || np_total > kMaxFilterN) return -1;
    state_free_particles(s);
    s->agg_ok = false;
    const int nf = (int)ns.size();
    s->nf = nf;
    s->n.assign(ns.begin(), ns.end());
    s->off.resize(nf);
    s->padn.resize(nf);
    int NP = 0;
    for (int f = 0; f < nf; ++f) {
        s->off[f] = NP;
        s->padn[f] = (ns[f] + kTile - 1) / kTile * kTile;
        NP += s->padn[f];
    }
    s->NP = NP;
    const size_t fb = sizeof(float) * (size_t)NP;
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 7; ++c) {
            PF_CHECK(hipMalloc((void**)&s->pose.s[b].c[c], fb));
            PF_CHECK(hipMemsetAsync(s->pose.s[b].c[c], 0, fb, s->stream));
        }
    PF_CHECK(hipMalloc((void**)&s->w, fb));
    PF_CHECK(hipMemsetAsync(s->w, 0, fb, s->stream));
    PF_CHECK(hipMalloc((void**)&s->ll, sizeof(double) * (size_t)NP));
    PF_CHECK(hipMemsetAsync(s->ll, 0, sizeof(double) * (size_t)NP, s->stream));
    PF_CHECK(hipMalloc((void**)&s->cum, sizeof(uint64_t) * (size_t)NP));
    PF_CHECK(hipMalloc((void**)&s->anc, sizeof(int32_t) * (size_t)NP));
    PF_CHECK(hipMemsetAsync(s->anc, 0, sizeof(int32_t) * (size_t)NP, s->stream));
    PF_CHECK(hipMalloc((void**)&s->inj_next, sizeof(int32_t) * (size_t)NP));
    PF_CHECK(hipMalloc((void**)&s->inj_head, sizeof(int32_t) * (size_t)NP));
    PF_CHECK(hipMemsetAsync(s->inj_head, 0xff, sizeof(int32_t) * (size_t)NP, s->stream));  // empty lists (-1)
    PF_CHECK(hipMalloc((void**)&s->inj_jit, sizeof(double) * 3 * (size_t)NP));
    const int nb256 = NP / kTB, ntile = NP / kTile;
    PF_CHECK(hipMalloc((void**)&s->part_w, sizeof(double) * 3 * (size_t)nb256));
    PF_CHECK(hipMalloc((void**)&s->part_n, sizeof(double) * (size_t)nb256));
    PF_CHECK(hipMalloc((void**)&s->part_q, sizeof(uint64_t) * (size_t)ntile));
    PF_CHECK(hipMalloc((void**)&s->tcum, sizeof(uint64_t) * (size_t)ntile));
    PF_CHECK(hipMalloc((void**)&s->tpre, sizeof(uint64_t) * (size_t)ntile));
    PF_CHECK(hipMalloc((void**)&s->part_est, sizeof(double) * kEstN * (size_t)ntile));
    PF_CHECK(hipMalloc((void**)&s->part_cov, sizeof(double) * kCovN * 4 * (size_t)ntile));  // per-wave layout
    PF_CHECK(hipMalloc((void**)&s->cov_fold, sizeof(double) * kCovN * kCovSplit * (size_t)nf));
    PF_CHECK(hipMalloc((void**)&s->scan_st, sizeof(uint64_t) * 2 * (size_t)ntile));
    PF_CHECK(hipMemsetAsync(s->scan_st, 0, sizeof(uint64_t) * 2 * (size_t)ntile, s->stream));
    PF_CHECK(hipMalloc((void**)&s->scan_ctl, sizeof(unsigned long long)));
    PF_CHECK(hipMemsetAsync(s->scan_ctl, 0, sizeof(unsigned long long), s->stream));
    // layout tables
    std::vector<int> blk_filter(nb256);
    for (int f = 0; f < nf; ++f)
        for (int b = s->off[f] / kTB; b < (s->off[f] + s->padn[f]) / kTB; ++b) blk_filter[b] = f;
    std::vector<float> winv(nf);
    for (int f = 0; f < nf; ++f) winv[f] = (float)(1.0 / (double)ns[f]);
    PF_CHECK(hipMalloc((void**)&s->d_blk_filter, sizeof(int) * (size_t)nb256));
    PF_CHECK(hipMalloc((void**)&s->d_fn, sizeof(int) * (size_t)nf));
    PF_CHECK(hipMalloc((void**)&s->d_foff, sizeof(int) * (size_t)nf));
    PF_CHECK(hipMalloc((void**)&s->d_fpadn, sizeof(int) * (size_t)nf));
    PF_CHECK(hipMalloc((void**)&s->d_fwinv, sizeof(float) * (size_t)nf));
    PF_CHECK(hipMalloc((void**)&s->fdev, sizeof(FilterDev) * (size_t)nf));
    PF_CHECK(hipMemcpyAsync(s->d_blk_filter, blk_filter.data(), sizeof(int) * (size_t)nb256, hipMemcpyHostToDevice,
                            s->stream));
    PF_CHECK(hipMemcpyAsync(s->d_fn, s->n.data(), sizeof(int) * (size_t)nf, hipMemcpyHostToDevice, s->stream));
    PF_CHECK(hipMemcpyAsync(s->d_foff, s->off.data(), sizeof(int) * (size_t)nf, hipMemcpyHostToDevice, s->stream));
    PF_CHECK(hipMemcpyAsync(s->d_fpadn, s->padn.data(), sizeof(int) * (size_t)nf, hipMemcpyHostToDevice, s->stream));
    PF_CHECK(hipMemcpyAsync(s->d_fwinv, winv.data(), sizeof(float) * (size_t)nf, hipMemcpyHostToDevice, s->stream));
    std::vector<FilterDev> fd(nf);
    for (int f = 0; f < nf; ++f) {
        memset(&fd[f], 0, sizeof(FilterDev));
        fd[f].lcg = 0x243F6A8885A308D3ull;  // cu:116
        fd[f].ess = (double)ns[f];
    }
    PF_CHECK(hipMemcpyAsync(s->fdev, fd.data(), sizeof(FilterDev) * (size_t)nf, hipMemcpyHostToDevice, s->stream));
    PF_CHECK(hipStreamSynchronize(s->stream));
    s->L.blk_filter = s->d_blk_filter;
    s->L.f_n = s->d_fn;
    s->L.f_off = s->d_foff;
    s->L.f_padn = s->d_fpadn;
    s->L.f_winv = s->d_fwinv;
    s->L.nf = nf;
    return 0;
}

void state_free_particles(PfState* s) {
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 7; ++c) {
            (void)hipFree(s->pose.s[b].c[c]);
            s->pose.s[b].c[c] = nullptr;
        }
    void** ptrs[] = {(void**)&s->w,        (void**)&s->ll,          (void**)&s->cum,       (void**)&s->anc,
                     (void**)&s->inj_next,    (void**)&s->inj_head,  (void**)&s->inj_jit,
                     (void**)&s->part_w,   (void**)&s->part_n,  (void**)&s->part_q,  (void**)&s->tcum,  (void**)&s->tpre,  (void**)&s->part_est,    (void**)&s->part_cov,  (void**)&s->cov_fold, (void**)&s->scan_st,
                     (void**)&s->scan_ctl, (void**)&s->d_blk_filter, (void**)&s->d_fn,     (void**)&s->d_foff,
                     (void**)&s->d_fpadn,  (void**)&s->d_fwinv,     (void**)&s->fdev,      (void**)&s->d_inj_blk};
    for (void** p : ptrs) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    s->nf = 0;
    s->NP = 0;
    s->inj_nblk = 0;
}

int set_inject_blocks(PfState* s, const std::vector<int>& kmax) {
    std::vector<int32_t> tab;
    for (int f = 0; f < s->nf; ++f) {
        const int k = std::min(std::max(kmax[f], 0), s->n[f]);
        const int b0 = s->off[f] / kTB;
        for (int b = 0; b < (k + kTB - 1) / kTB; ++b) tab.push_back(b0 + b);
    }
    (void)hipFree(s->d_inj_blk);
    s->d_inj_blk = nullptr;
    s->inj_nblk = (int)tab.size();
    if (tab.empty()) return 0;
    PF_CHECK(hipMalloc((void**)&s->d_inj_blk, sizeof(int32_t) * tab.size()));
    PF_CHECK(hipMemcpy(s->d_inj_blk, tab.data(), sizeof(int32_t) * tab.size(), hipMemcpyHostToDevice));
    return 0;
}

// ---- profiling helpers
void prof_begin(PfState* s, int stage) {
    if (!s->prof_on) return;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    (void)hipEventRecord(a, s->stream);
    s->prof_pending.push_back({stage, a, b});
}
void prof_end(PfState* s) {
    if (!s->prof_on || s->prof_pending.empty()) return;
    (void)hipEventRecord(s->prof_pending.back().b, s->stream);
}

int launch_motion(PfState* s, ParamRef<MotionParams> A, int reinit, int dst_flip) {
    if (reinit) s->agg_ok = false;  // reinit writes the weights
    prof_begin(s, reinit ? PFX_STAGE_REINIT : PFX_STAGE_PREDICT);
    hipLaunchKernelGGL(reinit ? k_motion<1> : k_motion<0>, dim3(s->NP / kTB), dim3(kTB), 0, s->stream, s->L, s->pose,
                       s->fdev, s->w, A, dst_flip);
    prof_end(s);
    PF_LAUNCH_CHECK();
    return 0;
}

int launch_weight(PfState* s, ParamRef<WeightParams> A, const WeightDefer& d) {
    const bool dc = d.cov != nullptr;
    if (dc && (!d.out || !d.flip)) return -1;
    prof_begin(s, PFX_STAGE_WEIGHT);
    const int ppt = s->weight_ppt;
    const int nblk = s->NP / (kTB * ppt);
    const int ntile = s->NP / kTile;
    const int grid = nblk + (dc ? ntile : 0);
    const ParamRef<CovParams> cA = ParamRef<CovParams>::array(d.cov);
    auto kw = ppt == 1 ? (dc ? k_weight<1, true> : k_weight<1, false>)
              : ppt == 4 ? (dc ? k_weight<4, true> : k_weight<4, false>)
                         : (dc ? k_weight<2, true> : k_weight<2, false>);
    PoseBufs B = s->pose;
    if (d.flip) std::swap(B.s[0], B.s[1]);  // B.s[parity]: the set predict wrote
    hipLaunchKernelGGL(kw, dim3(grid), dim3(kTB), 0, s->stream, s->L, B, s->fdev, s->ll, s->part_w, A, s->w,
                       s->part_cov, ntile, cA, nblk);
    prof_end(s);
    PF_LAUNCH_CHECK();
    prof_begin(s, PFX_STAGE_FINALIZE);
    hipLaunchKernelGGL(k_weight_fin, dim3(s->nf + (dc ? s->nf * kCovSplit : 0)), dim3(kFinTB), 0, s->stream, s->L,
                       s->fdev, s->part_w, kTB * ppt, s->nf, d.flip, dc ? s->part_cov : nullptr, ntile, d.out,
                       s->cov_fold);
    prof_end(s);
    PF_LAUNCH_CHECK();
    return 0;
}

// The scan from the tile aggregates (k_scan_agg), after the prefix pass when a filter has more
// than kPrefixTiles tiles.
static int launch_scan_agg(PfState* s, const float* d_ratio) {
    const int ntile = s->NP / kTile;
    int max_nt = 0;
    for (int p : s->padn) max_nt = std::max(max_nt, p / kTile);
    prof_begin(s, PFX_STAGE_RESAMPLE_SCAN);
    if (max_nt > kPrefixTiles) {
        hipLaunchKernelGGL(k_tile_prefix, dim3(s->nf), dim3(kFinTB), 0, s->stream, s->L, s->part_q, s->tpre);
        PF_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_scan_agg, dim3(ntile), dim3(kTB), 0, s->stream, s->L, s->fdev, s->w, s->part_n, s->part_q,
                       s->cum, s->tcum, d_ratio, max_nt > kPrefixTiles ? s->tpre : nullptr);
    prof_end(s);
    PF_LAUNCH_CHECK();
    return 0;
}

int launch_normalize(PfState* s, int force_uniform, const float* d_ratio, bool scan, ParamRef<InjectParams> inj,
                     FilterDev* cov_out) {
    if (cov_out && !scan) return -1;
    const int ntile = s->NP / kTile;
    if (scan) {
        if (!d_ratio) return -1;  // the fused scan forms the resample predicate
        s->agg_ok = false;        // (the gather that follows rewrites the weights)
        prof_begin(s, PFX_STAGE_NORMALIZE);
        hipLaunchKernelGGL(k_norm_agg, dim3(ntile + (cov_out ? s->nf : 0)), dim3(kTB), 0, s->stream, s->L, s->fdev,
                           s->ll, s->w, s->part_n, s->part_q, inj, s->inj_jit, s->inj_head, s->inj_next, s->cov_fold,
                           cov_out, ntile);
        prof_end(s);
        PF_LAUNCH_CHECK();
        return launch_scan_agg(s, d_ratio);
    }
    prof_begin(s, PFX_STAGE_NORMALIZE);
    hipLaunchKernelGGL(k_normalize, dim3(ntile), dim3(kTB), 0, s->stream, s->L, s->fdev, s->ll, s->w, s->part_n,
                       s->part_q, force_uniform);
    prof_end(s);
    s->agg_ok = true;
    PF_LAUNCH_CHECK();
    prof_begin(s, PFX_STAGE_FINALIZE);
    hipLaunchKernelGGL(k_norm_fin, dim3(s->nf), dim3(kFinTB), 0, s->stream, s->L, s->fdev, s->part_n, d_ratio);
    prof_end(s);
    PF_LAUNCH_CHECK();
    return 0;
}

int launch_resample(PfState* s, ParamRef<InjectParams> inj, const int32_t* inj_tab, int inj_nblk, bool flip,
                    bool scan) {
    const int ntile = s->NP / kTile;
    if (scan && s->agg_ok) {
        // the weights are those the last normalisation left aggregates for: reduce-then-scan
        // from them (the same integer prefixes, total and LCG advance as the look-back scan;
        // the predicate forced, as k_resample_predicate does for this path)
        if (launch_scan_agg(s, nullptr) != 0) return -1;
    } else if (scan) {
        prof_begin(s, PFX_STAGE_RESAMPLE_SCAN);
        hipLaunchKernelGGL(k_scan, dim3(ntile), dim3(kTB), 0, s->stream, s->L, s->fdev, s->w, s->cum, s->tcum, s->scan_st,
                           s->scan_ctl, s->scan_spin_limit);
        prof_end(s);
        PF_LAUNCH_CHECK();
    }
    s->agg_ok = false;  // the gather rewrites the weights
    if (scan && inj_nblk > 0) {  // (the fused normalise-and-scan drew them otherwise)
        prof_begin(s, PFX_STAGE_INJECT);
        hipLaunchKernelGGL(k_inj_pick, dim3(inj_nblk), dim3(kTB), 0, s->stream, s->L, s->fdev, inj, inj_tab,
                           s->inj_jit, s->inj_head, s->inj_next);
        prof_end(s);
        PF_LAUNCH_CHECK();
    }
    prof_begin(s, PFX_STAGE_RESAMPLE_GATHER);
    hipLaunchKernelGGL(k_gather, dim3(s->NP / kGatherBlk), dim3(kTB), 0, s->stream, s->L, s->fdev, s->pose, s->cum,
                       s->tcum, s->w, s->anc, s->inj_head, s->inj_next, s->inj_jit, s->part_est, s->NP / kTile);
    prof_end(s);
    PF_LAUNCH_CHECK();
    if (flip) {
        prof_begin(s, PFX_STAGE_FINALIZE);
        hipLaunchKernelGGL(k_flip, dim3((s->nf + 63) / 64), dim3(64), 0, s->stream, s->fdev, s->nf, s->scan_ctl);
        prof_end(s);
        PF_LAUNCH_CHECK();
    }
    return 0;
}

int launch_est_sums(PfState* s, CovParams* d_cov, bool flip, bool pass) {
    if (pass) {
        prof_begin(s, PFX_STAGE_EST_SUMS);
        hipLaunchKernelGGL(k_est_sums, dim3(s->NP / kTile), dim3(kTB), 0, s->stream, s->L, s->fdev, s->pose, s->w,
                           s->part_est, s->NP / kTile);
        prof_end(s);
        PF_LAUNCH_CHECK();
    }
    prof_begin(s, PFX_STAGE_FINALIZE);
    hipLaunchKernelGGL(k_est_fin<kEstN>, dim3(s->nf), dim3(kEstFinTB), 0, s->stream, s->L, s->fdev, s->part_est,
                       s->NP / kTile, 0, d_cov, flip ? 1 : 0, s->scan_ctl, nullptr);
    prof_end(s);
    PF_LAUNCH_CHECK();
    return 0;
}

int launch_est_covs(PfState* s, ParamRef<CovParams> A, FilterDev* results) {
    prof_begin(s, PFX_STAGE_EST_COVS);
    hipLaunchKernelGGL(k_est_covs, dim3(s->NP / kTile), dim3(kTB), 0, s->stream, s->L, s->fdev, s->pose, s->w,
                       s->part_cov, s->NP / kTile, A);
    prof_end(s);
    PF_LAUNCH_CHECK();
    prof_begin(s, PFX_STAGE_FINALIZE);
    hipLaunchKernelGGL(k_est_fin<kCovN>, dim3(s->nf), dim3(kEstFinTB), 0, s->stream, s->L, s->fdev, s->part_cov,
                       s->NP / kTile, 1, nullptr, 0, s->scan_ctl, results);
    prof_end(s);
    PF_LAUNCH_CHECK();
    return 0;
}

int launch_predicate(PfState* s, const float* d_ratio, int force) {
    hipLaunchKernelGGL(k_resample_predicate, dim3((s->nf + 63) / 64), dim3(64), 0, s->stream, s->fdev, s->nf, d_ratio,
                       s->d_fn, force);
    PF_LAUNCH_CHECK();
    return 0;
}

WeightParams make_weight_params(double sigma_r, double sigma_b, double gate) {
    WeightParams p{};
    const double sr = sigma_r > 1e-6 ? sigma_r : 1e-6;  // cpu.rs:56-57
    const double sb = sigma_b > 1e-6 ? sigma_b : 1e-6;
    const double g2 = gate * gate;
    p.inv2rr = 1.0 / (2.0 * sr * sr);
    p.inv2bb = 1.0 / (2.0 * sb * sb);
    p.gate = gate;
    // squared-residual thresholds (cpu.rs:63-65); +inf without a gate (cpu.rs:72, 85, 90)
    const bool gated = gate > 0.0;
    p.cart_gate2 = gated ? g2 * 3.0 * sr * sr : INFINITY;
    p.rng_gate = gated ? fabs(gate) * sr : INFINITY;
    p.brg_gate2 = gated ? g2 * sb * sb : INFINITY;
    // Bearing gate in u = 1 - cos(eb): the largest u whose 3-term series value is <= g^2 sb^2
    // (bisection over doubles with the kernel's own series), so that series(min(u, u_gate))
    // is the gated term; +inf when ungated.
    const double gb = gate * sb;
    p.u_gate = INFINITY;
    if (gated && gb < 0.5) {
        const double target = p.brg_gate2;
        double lo = 0.0, hi = 2.0 * (1.0 - cos(gb)) + 1e-300;
        while (acos_sq_u3(hi) <= target) hi *= 2.0;
        for (int it = 0; it < 2000; ++it) {
            const double mid = 0.5 * (lo + hi);
            if (!(mid > lo && mid < hi)) break;
            if (acos_sq_u3(mid) <= target)
                lo = mid;
            else
                hi = mid;
        }
        p.u_gate = lo;
    }
    // Scaled range-bearing path when the bearing gate bounds u below the 3-term series' range:
    // the series coefficients in us = u / u_gate (u acos-series(u) = us (c0 + c1 us +
    // c2 us^2)), so that us = 1 yields the gated value.
    p.fast_rb = (p.u_gate > 0.0 && p.u_gate <= 0.005) ? 1 : 0;
    p.pad_fast = 0;
    if (p.fast_rb) {
        const double ug = p.u_gate;
        p.cg = 1.0 / ug;
        p.c0 = 2.0 * ug;
        p.c1 = (8.0 / 24.0) * ug * ug;
        p.c2 = (16.0 / 180.0) * ug * ug * ug;
        p.rng_gate4 = 4.0 * p.rng_gate;
    }
    return p;
}

// Joins observations with the landmark table (skips ids outside [0, lm_cap), cu:242).  On the
// scaled range-bearing path (p.fast_rb) the range-bearing records carry v / (2 u_gate) and
// 4 range (see rb_scaled).
int join_obs(const pfc_obs* obs, int nobs, const float* lm, int lm_cap, const WeightParams& p, ObsRec* out,
             int* nrb) {
    const double vscale = p.fast_rb ? 0.5 / p.u_gate : 1.0;
    const double rscale = p.fast_rb ? 4.0 : 1.0;
    // range-bearing records first (in input order), then Cartesian ones (in input order)
    int m = 0, rb = 0;
    for (int pass = 0; pass < 2; ++pass) {
        for (int k = 0; k < nobs; ++k) {
            const pfc_obs& o = obs[k];
            if (o.id < 0 || o.id >= lm_cap) continue;
            if ((o.mode == 0) != (pass == 1)) continue;
            ObsRec& r = out[m++];
            r.lx = lm[3 * o.id];
            r.ly = lm[3 * o.id + 1];
            r.lz = lm[3 * o.id + 2];
            const double vs = pass == 0 ? vscale : 1.0;
            r.vx = (double)o.dx * vs;
            r.vy = (double)o.dy * vs;
            r.vz = (double)o.dz * vs;
            r.range = (double)o.range * (pass == 0 ? rscale : 1.0);
            r.mode = o.mode == 0 ? 0 : 1;
            r.pad = 0;
        }
        if (pass == 0) rb = m;
    }
    if (nrb) *nrb = rb;
    return m;
}

}  // namespace pfk

using namespace pfk;

// ============================================================================ C ABI

extern "C" {

int pfc_device_available(void) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    return prop.major >= 9 ? 1 : 0;
}

constexpr int kSmallRead = 16;  // doubles: the largest synchronous result (pfc_est_sums' 14)

void* pfc_create(int cap_obs) {
    PfState* s = new (std::nothrow) PfState();
    if (!s) return nullptr;
    if (cap_obs < 1 || cap_obs > 4096) cap_obs = 64;  // cu:488
    s->obs_cap = cap_obs;
    if (hipSetDevice(0) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&s->d_obs, sizeof(ObsRec) * (size_t)cap_obs) != hipSuccess ||
        hipHostMalloc((void**)&s->h_obs, sizeof(ObsRec) * (size_t)cap_obs, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s->obs_ev, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc((void**)&s->h_small, sizeof(double) * kSmallRead, hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&s->small_ev, hipEventDisableTiming) != hipSuccess) {
        fprintf(stderr, "pf_hip pfc_create: device setup failed\n");
        pfc_destroy(s);
        return nullptr;
    }
    return s;
}

int pfc_destroy(void* h) {
    PfState* s = static_cast<PfState*>(h);
    if (!s) return 0;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    state_free_particles(s);
    (void)hipFree(s->d_obs);
    (void)hipHostFree(s->h_obs);
    (void)hipHostFree(s->h_small);
    if (s->small_ev) (void)hipEventDestroy(s->small_ev);
    for (auto& p : s->prof_pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    if (s->obs_ev) (void)hipEventDestroy(s->obs_ev);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return 0;
}

int pfc_init(void* h, int n, unsigned int seed) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || n <= 0) return -1;
    if (state_alloc(s, std::vector<int>{n}) != 0) return -1;
    s->gseed = seed ? (uint64_t)seed : 42ull;  // cu:523
    s->gen = 0;
    return 0;
}

int pfc_upload_landmarks(void* h, const float* data, int count) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || count < 0 || (count > 0 && !data)) return -1;
    s->lm.assign(data, data + 3 * (size_t)count);  // host mirror; joined with obs per weight call
    s->lm_cap = count;
    return 0;
}

int pfc_reinit(void* h, float x, float y, float z, float qw, float qx, float qy, float qz, const float* pos_std,
               const float* rot_std) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1 || !pos_std || !rot_std) return -1;
    s->gseed += kGenBump;  // cpu.rs:181
    s->gen += 1;           // cu:572
    MotionParams p{};
    p.t[0] = x;
    p.t[1] = y;
    p.t[2] = z;
    p.q[0] = qw;
    p.q[1] = qx;
    p.q[2] = qy;
    p.q[3] = qz;
    for (int k = 0; k < 3; ++k) {
        p.s_t[k] = pos_std[k];
        p.s_r[k] = rot_std[k];
    }
    p.gseed = s->gseed;
    p.winv = (float)(1.0 / (double)s->n[0]);
    return launch_motion(s, ParamRef<MotionParams>::value(p), 1);
}

int pfc_predict(void* h, float tx, float ty, float tz, float qw, float qx, float qy, float qz, const float* tstd,
                const float* rstd) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1 || !tstd || !rstd) return -1;
    s->gseed += kGenBump;  // cpu.rs:198
    s->gen += 1;           // cu:588
    MotionParams p{};
    p.t[0] = tx;
    p.t[1] = ty;
    p.t[2] = tz;
    p.q[0] = qw;
    p.q[1] = qx;
    p.q[2] = qy;
    p.q[3] = qz;
    for (int k = 0; k < 3; ++k) {
        p.s_t[k] = tstd[k];
        p.s_r[k] = rstd[k];
    }
    p.gseed = s->gseed;
    return launch_motion(s, ParamRef<MotionParams>::value(p), 0);
}

int pfc_weight(void* h, const void* obs, int nobs, float sigma_r, float sigma_b, float gate) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1) return -1;
    nobs = nobs > s->obs_cap ? s->obs_cap : nobs;  // cu:604
    if (nobs <= 0 || !obs) return launch_normalize(s, 1, nullptr);
    // the staging buffer may still feed the previous call's copy
    PF_CHECK(hipEventSynchronize(s->obs_ev));
    int nrb = 0;
    WeightParams p = make_weight_params(sigma_r, sigma_b, gate);
    const int m = join_obs(static_cast<const pfc_obs*>(obs), nobs, s->lm.data(), s->lm_cap, p, s->h_obs, &nrb);
    if (m == 0) return launch_normalize(s, 1, nullptr);  // every id unknown: ll = 0 for all -> uniform
    PF_CHECK(hipMemcpyAsync(s->d_obs, s->h_obs, sizeof(ObsRec) * (size_t)m, hipMemcpyHostToDevice, s->stream));
    PF_CHECK(hipEventRecord(s->obs_ev, s->stream));
    p.obs = s->d_obs;
    p.nobs = m;
    p.nrb = nrb;
    if (launch_weight(s, ParamRef<WeightParams>::value(p)) != 0) return -1;
    return launch_normalize(s, 0, nullptr);
}

int pfc_resample(void* h, float random_inject_ratio) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1) return -1;
    if (launch_predicate(s, nullptr, 1) != 0) return -1;
    InjectParams ip{};
    ip.gen = s->gen;
    int k = 0;
    if (random_inject_ratio > 0.f) {
        k = (int)(random_inject_ratio * (float)s->n[0]);  // cu:667-671
        if (k > s->n[0]) k = s->n[0];
    }
    ip.k = k;
    return launch_resample(s, ParamRef<InjectParams>::value(ip), nullptr, (k + kTB - 1) / kTB, true);
}

// A synchronous read of n doubles after the work queued on the stream: copied into the handle's
// pinned buffer, completion polled on an event (a short cycle's result is ready within tens of
// microseconds; a sleeping wait would add the wake-up latency to every call), then handed to the
// caller's memory.
static int read_small(PfState* s, double* out, const double* dsrc, int n) {
    PF_CHECK(hipMemcpyAsync(s->h_small, dsrc, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, s->stream));
    PF_CHECK(hipEventRecord(s->small_ev, s->stream));
    hipError_t q;
    while ((q = hipEventQuery(s->small_ev)) == hipErrorNotReady) {
    }
    PF_CHECK(q);
    memcpy(out, s->h_small, sizeof(double) * (size_t)n);
    return 0;
}

int pfc_est_sums(void* h, double* out) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1 || !out) return -1;
    if (launch_est_sums(s, nullptr, false) != 0) return -1;
    return read_small(s, out, fdev_field<double>(s, 0, offsetof(FilterDev, sums)), kEstN);
}

int pfc_covs(void* h, float mx, float my, float mz, float qw, float qx, float qy, float qz, double* out) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1 || !out) return -1;
    CovParams p{};
    p.mp[0] = mx;
    p.mp[1] = my;
    p.mp[2] = mz;
    p.mq[0] = qw;
    p.mq[1] = qx;
    p.mq[2] = qy;
    p.mq[3] = qz;
    if (launch_est_covs(s, ParamRef<CovParams>::value(p)) != 0) return -1;
    return read_small(s, out, fdev_field<double>(s, 0, offsetof(FilterDev, covs)), kCovN);
}

static int read_parity(PfState* s, int f, int* par) {
    PF_CHECK(hipMemcpyAsync(par, fdev_field<int>(s, f, offsetof(FilterDev, parity)), sizeof(int), hipMemcpyDeviceToHost,
                            s->stream));
    PF_CHECK(hipStreamSynchronize(s->stream));
    return 0;
}

int pfx_get_state(void* h, float* pos, float* quat, float* w) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1) return -1;
    int par = 0;
    if (read_parity(s, 0, &par) != 0) return -1;
    const int n = s->n[0];
    std::vector<float> tmp((size_t)n);
    for (int c = 0; c < 7; ++c) {
        if ((c < 3 && !pos) || (c >= 3 && !quat)) continue;
        PF_CHECK(hipMemcpy(tmp.data(), s->pose.s[par].c[c], sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
        if (c < 3)
            for (int i = 0; i < n; ++i) pos[3 * i + c] = tmp[i];
        else
            for (int i = 0; i < n; ++i) quat[4 * i + (c - 3)] = tmp[i];
    }
    if (w) PF_CHECK(hipMemcpy(w, s->w, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    return 0;
}

int pfc_snapshot(void* h, float* pos, float* quat) {
    if (!pos || !quat) return -1;
    return pfx_get_state(h, pos, quat, nullptr);
}

int pfx_set_state(void* h, const float* pos, const float* quat, const float* w) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1 || !pos || !quat) return -1;
    const int n = s->n[0];
    if (w) {
        double tot = 0.0;
        for (int i = 0; i < n; ++i) {
            if (!(w[i] >= 0.f) || !std::isfinite(w[i])) return -1;
            tot += w[i];
        }
        if (tot >= 16.0) return -1;  // fixed-point resampling headroom
    }
    PF_CHECK(hipStreamSynchronize(s->stream));
    s->agg_ok = false;  // the weights are replaced
    int par = 0;
    if (read_parity(s, 0, &par) != 0) return -1;
    std::vector<float> tmp((size_t)n);
    for (int c = 0; c < 7; ++c) {
        if (c < 3)
            for (int i = 0; i < n; ++i) tmp[i] = pos[3 * i + c];
        else
            for (int i = 0; i < n; ++i) tmp[i] = quat[4 * i + (c - 3)];
        PF_CHECK(hipMemcpy(s->pose.s[par].c[c], tmp.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    }
    if (w) {
        PF_CHECK(hipMemcpy(s->w, w, sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    } else {
        const float u = (float)(1.0 / (double)n);
        for (int i = 0; i < n; ++i) tmp[i] = u;
        PF_CHECK(hipMemcpy(s->w, tmp.data(), sizeof(float) * (size_t)n, hipMemcpyHostToDevice));
    }
    const int zero = 0;  // the estimate partials of the last gather no longer describe the state
    PF_CHECK(hipMemcpy(fdev_field<int>(s, 0, offsetof(FilterDev, est_src)), &zero, sizeof(int), hipMemcpyHostToDevice));
    return 0;
}

int pfx_get_ll(void* h, double* ll) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1 || !ll) return -1;
    PF_CHECK(hipStreamSynchronize(s->stream));
    PF_CHECK(hipMemcpy(ll, s->ll, sizeof(double) * (size_t)s->n[0], hipMemcpyDeviceToHost));
    return 0;
}

int pfx_get_ancestors(void* h, int32_t* anc) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1 || !anc) return -1;
    PF_CHECK(hipStreamSynchronize(s->stream));
    PF_CHECK(hipMemcpy(anc, s->anc, sizeof(int32_t) * (size_t)s->n[0], hipMemcpyDeviceToHost));
    return 0;
}

int pfx_get_cum(void* h, uint64_t* cum) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf != 1 || !cum) return -1;
    PF_CHECK(hipStreamSynchronize(s->stream));
    PF_CHECK(hipMemcpy(cum, s->cum, sizeof(uint64_t) * (size_t)s->n[0], hipMemcpyDeviceToHost));
    return 0;
}

int pfx_set_lcg(void* h, uint64_t lcg) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf < 1) return -1;
    PF_CHECK(hipStreamSynchronize(s->stream));
    PF_CHECK(hipMemcpy(fdev_field<uint64_t>(s, 0, offsetof(FilterDev, lcg)), &lcg, sizeof(uint64_t), hipMemcpyHostToDevice));
    return 0;
}

int pfx_get_lcg(void* h, uint64_t* lcg) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf < 1 || !lcg) return -1;
    PF_CHECK(hipStreamSynchronize(s->stream));
    PF_CHECK(hipMemcpy(lcg, fdev_field<uint64_t>(s, 0, offsetof(FilterDev, lcg)), sizeof(uint64_t), hipMemcpyDeviceToHost));
    return 0;
}

int pfx_get_rng(void* h, uint64_t* gseed, uint64_t* gen) {
    PfState* s = static_cast<PfState*>(h);
    if (!s) return -1;
    if (gseed) *gseed = s->gseed;
    if (gen) *gen = s->gen;
    return 0;
}

int pfx_set_rng(void* h, uint64_t gseed, uint64_t gen) {
    PfState* s = static_cast<PfState*>(h);
    if (!s) return -1;
    s->gseed = gseed;
    s->gen = gen;
    return 0;
}

int pfx_weight_scalars(void* h, double* out4) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || s->nf < 1 || !out4) return -1;
    FilterDev fd;
    PF_CHECK(hipStreamSynchronize(s->stream));
    PF_CHECK(hipMemcpy(&fd, s->fdev, sizeof(FilterDev), hipMemcpyDeviceToHost));
    out4[0] = fd.max_ll;
    out4[1] = fd.sum_e;
    out4[2] = fd.sum_e2;
    out4[3] = fd.ess;
    return 0;
}

int pfx_profile_enable(void* h, int enable) {
    PfState* s = static_cast<PfState*>(h);
    if (!s) return -1;
    s->prof_on = enable != 0;
    return 0;
}

int pfx_profile_read(void* h, double* ms, int* count) {
    PfState* s = static_cast<PfState*>(h);
    if (!s) return -1;
    PF_CHECK(hipStreamSynchronize(s->stream));
    for (int k = 0; k < PFX_STAGE_COUNT; ++k) {
        if (ms) ms[k] = 0.0;
        if (count) count[k] = 0;
    }
    for (auto& p : s->prof_pending) {
        float t = 0.f;
        if (hipEventElapsedTime(&t, p.a, p.b) == hipSuccess) {
            if (ms) ms[p.stage] += t;
            if (count) count[p.stage] += 1;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    s->prof_pending.clear();
    return 0;
}

int pfx_particle_count(void* h) {
    PfState* s = static_cast<PfState*>(h);
    return (s && s->nf == 1) ? s->n[0] : 0;
}

int pfx_quat_mean(const double* M, double* q) {
    if (!M || !q) return -1;
    return quat_mean4(M, q);
}

int pfx_set_weight_variant(void* h, int variant) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || (variant != 1 && variant != 2 && variant != 4)) return -1;
    s->weight_ppt = variant;
    return 0;
}

int pfx_set_scan_spin_limit(void* h, int limit) {
    PfState* s = static_cast<PfState*>(h);
    if (!s || limit < 0) return -1;
    s->scan_spin_limit = limit;
    return 0;
}

int pfx_sync(void* h) {
    PfState* s = static_cast<PfState*>(h);
    if (!s) return -1;
    PF_CHECK(hipStreamSynchronize(s->stream));
    return 0;
}

}  // extern "C"

// ============================================================ estimate sums without allocation
//
// pfx_est_sums_noalloc runs the estimate-sum kernels of pfc_est_sums (k_est_sums, then
// k_est_fin<14>) over particle streams the caller owns, with every buffer the kernels need
// beyond the caller's inputs and output -- the one-filter layout tables, the filter's scalar
// block and the per-tile f64 partials -- in the file-scope array `scratch` below.  It calls no
// allocator, so it can run where hipMalloc cannot: under HIP graph capture, in serving loops.
// Each GPU holds its own copy of `scratch` (about 2 MB) until the process ends; its device
// address is looked up once per device with hipGetSymbolAddress and only that address is
// handed to kernels and copies.  Every call on a GPU shares that GPU's copy: calls that
// overlap on one GPU (two streams, two threads) race.
namespace pfk {

constexpr int kNoAllocMaxN = 1 << 24;  // particles per call
constexpr size_t kNaAlign = 256;
constexpr size_t naup(size_t b) { return (b + kNaAlign - 1) / kNaAlign * kNaAlign; }
constexpr size_t kNaFilterOff = 0;                                     // FilterDev
constexpr size_t kNaScalarOff = naup(sizeof(FilterDev));               // f_n, f_off, f_padn (int), f_winv (float)
constexpr size_t kNaBlkOff = kNaScalarOff + kNaAlign;                  // blk_filter: one int per 256 particles
constexpr size_t kNaPartOff = kNaBlkOff + naup(sizeof(int) * (kNoAllocMaxN / kTB));  // kEstN x tiles doubles
constexpr size_t kNoAllocBytes = kNaPartOff + naup(sizeof(double) * kEstN * (kNoAllocMaxN / kTile));

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// One-filter layout for n particles at offset 0 (every block belongs to filter 0) and the filter
// fields k_est_sums reads (current pose set 0, no partials left by a gather).  Written on every
// call: the array keeps the previous call's contents.
__global__ void k_noalloc_setup(unsigned char* base, int n, int padn, int nblk) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    int* blk = reinterpret_cast<int*>(base + kNaBlkOff);
    if (i < nblk) blk[i] = 0;
    if (i == 0) {
        int* sc = reinterpret_cast<int*>(base + kNaScalarOff);
        sc[0] = n;
        sc[1] = 0;
        sc[2] = padn;
        reinterpret_cast<float*>(sc)[3] = (float)(1.0 / (double)n);
        FilterDev* F = reinterpret_cast<FilterDev*>(base + kNaFilterOff);
        F->parity = 0;
        F->est_src = 0;
    }
}

}  // namespace pfk

namespace {
constexpr int kNaMaxDevices = 64;
void* g_na_scratch[kNaMaxDevices] = {};
std::mutex g_na_mutex;

// Device address of `scratch` on the current device, looked up on the first call there.
unsigned char* noalloc_scratch() {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices) {
        fprintf(stderr, "pf_hip pfx_est_sums_noalloc: no current device\n");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(pfk::scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "pf_hip pfx_est_sums_noalloc: hipGetSymbolAddress(scratch): %s\n", hipGetErrorString(e));
            return nullptr;
        }
        g_na_scratch[dev] = p;
    }
    return static_cast<unsigned char*>(g_na_scratch[dev]);
}
}  // namespace

extern "C" int pfx_est_sums_noalloc(const float* px, const float* py, const float* pz, const float* qw,
                                    const float* qx, const float* qy, const float* qz, const float* w, int n,
                                    double* out14, void* stream) {
    using namespace pfk;
    if (!px || !py || !pz || !qw || !qx || !qy || !qz || !w || !out14) return -1;
    if (n <= 0 || n > kNoAllocMaxN) {
        fprintf(stderr, "pf_hip pfx_est_sums_noalloc: n = %d outside [1, %d]\n", n, kNoAllocMaxN);
        return -1;
    }
    unsigned char* base = noalloc_scratch();
    if (!base) return -1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int padn = (n + kTile - 1) / kTile * kTile;
    const int nblk = padn / kTB, ntile = padn / kTile;
    hipLaunchKernelGGL(k_noalloc_setup, dim3((nblk + kTB - 1) / kTB), dim3(kTB), 0, st, base, n, padn, nblk);
    PF_LAUNCH_CHECK();
    int* sc = reinterpret_cast<int*>(base + kNaScalarOff);
    Layout L{};
    L.blk_filter = reinterpret_cast<const int*>(base + kNaBlkOff);
    L.f_n = sc;
    L.f_off = sc + 1;
    L.f_padn = sc + 2;
    L.f_winv = reinterpret_cast<const float*>(sc + 3);
    L.nf = 1;
    FilterDev* F = reinterpret_cast<FilterDev*>(base + kNaFilterOff);
    double* part = reinterpret_cast<double*>(base + kNaPartOff);
    PoseBufs B{};
    const float* c[7] = {px, py, pz, qw, qx, qy, qz};
    for (int k = 0; k < 7; ++k) B.s[0].c[k] = B.s[1].c[k] = const_cast<float*>(c[k]);
    hipLaunchKernelGGL(k_est_sums, dim3(ntile), dim3(kTB), 0, st, L, F, B, w, part, ntile);
    PF_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_est_fin<kEstN>, dim3(1), dim3(kEstFinTB), 0, st, L, F, part, ntile, 0,
                       static_cast<CovParams*>(nullptr), 0, static_cast<unsigned long long*>(nullptr),
                       static_cast<FilterDev*>(nullptr));
    PF_LAUNCH_CHECK();
    PF_CHECK(hipMemcpyAsync(out14, F->sums, sizeof(double) * kEstN, hipMemcpyDefault, st));
    return 0;
}

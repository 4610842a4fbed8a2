// mlp_kernels.hip -- learned proposal for CBGL global localisation (pfl_* C ABI).
//
// A per-map MLP (scan -> Gaussian mixture over (x, y, sin yaw, cos yaw)) trained on the GPU
// on scans rendered by the CBGL raycaster.  All matrix products are one hand-written GEMM
// on the gfx950 BF16 MFMA (v_mfma_f32_32x32x16_bf16, FP32 accumulate) with fused epilogues:
//   forward hidden   Z = A W^T + b, H = bf16(GELU(Z))
//   forward output   O = A W^T + b
//   data gradient    dZ_prev = (dZ W) * GELU'(Z_prev)
//   weight gradient  dW = dZ^T A
// Mixture loss/gradients, bias gradients (column sums) and Adam (FP32 master weights with a
// BF16 copy for the GEMMs) are small elementwise kernels.  Everything is deterministic
// (no atomics; fixed reduction orders).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/pf_learned.h"
#include "pf_device.h"

#define ML_CHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "pf_hip %s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)
#define ML_LAUNCH()                                                                              \
    do {                                                                                         \
        hipError_t e_ = hipGetLastError();                                                       \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "pf_hip %s:%d launch failed: %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)

namespace mlk {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short s16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// float -> bf16, round to nearest even (NaN stays NaN).
__host__ __device__ inline uint16_t f2bf(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ inline uint16_t as_bf(float x) { return f2bf(x); }
__device__ inline uint16_t as_bf(uint16_t x) { return x; }

enum { EPI_HIDDEN = 0, EPI_OUT = 1, EPI_DGRAD = 2, EPI_WGRAD = 3 };

struct GemmArgs {
    int M, N, K;
    const void* A;
    long sam, sak;  // A(m, k) = A[m*sam + k*sak]
    const void* B;
    long sbk, sbn;  // B(k, n) = B[k*sbk + n*sbn]
    const float* bias;
    float* C;
    long ldc;
    uint16_t* H;  // EPI_HIDDEN: bf16 GELU(C)
    long ldh;
    const float* Zp;  // EPI_DGRAD: pre-activation of the previous layer
    long ldz;
    int kchunk;  // K range per workgroup along gridDim.z (set by the launcher)
    float* ws;   // split-K partials [gridDim.z][M][N]
};

constexpr int kBM = 64, kBN = 64, kBK = 32, kPad = 8;

__device__ inline float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

template <int EPI>
__device__ inline void epilogue(const GemmArgs& g, int m, int n, float v) {
    if (EPI == EPI_HIDDEN) {
        const float z = v + g.bias[n];
        g.C[(long)m * g.ldc + n] = z;
        g.H[(long)m * g.ldh + n] = f2bf(gelu(z));
    } else if (EPI == EPI_OUT) {
        g.C[(long)m * g.ldc + n] = v + g.bias[n];
    } else if (EPI == EPI_DGRAD) {
        g.C[(long)m * g.ldc + n] = v * gelu_grad(g.Zp[(long)m * g.ldz + n]);
    } else {
        g.C[(long)m * g.ldc + n] = v;
    }
}

// 8 consecutive elements of one operand row -> bf16 (round to nearest even); `cnt` of them
// are in bounds, the rest become 0.  16-byte vector loads when the run is whole and aligned.
__device__ inline void load8(const uint16_t* p, int cnt, uint16_t out[8]) {
    if (cnt == 8 && ((uintptr_t)p & 15) == 0) {
        const uint4 v = *reinterpret_cast<const uint4*>(p);
        out[0] = (uint16_t)v.x; out[1] = (uint16_t)(v.x >> 16);
        out[2] = (uint16_t)v.y; out[3] = (uint16_t)(v.y >> 16);
        out[4] = (uint16_t)v.z; out[5] = (uint16_t)(v.z >> 16);
        out[6] = (uint16_t)v.w; out[7] = (uint16_t)(v.w >> 16);
        return;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) out[j] = j < cnt ? p[j] : (uint16_t)0;
}
__device__ inline void load8(const float* p, int cnt, uint16_t out[8]) {
    if (cnt == 8 && ((uintptr_t)p & 15) == 0) {
        const float4 a = *reinterpret_cast<const float4*>(p);
        const float4 b = *reinterpret_cast<const float4*>(p + 4);
        out[0] = f2bf(a.x); out[1] = f2bf(a.y); out[2] = f2bf(a.z); out[3] = f2bf(a.w);
        out[4] = f2bf(b.x); out[5] = f2bf(b.y); out[6] = f2bf(b.z); out[7] = f2bf(b.w);
        return;
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) out[j] = j < cnt ? f2bf(p[j]) : (uint16_t)0;
}

// Stages a 64 (rows) x 32 (k) tile of an operand into LDS as [row][k] bf16.
//   kKContig: element (r, k) at X[r * ld + k]  (thread: one row, 8 consecutive k)
//   else:     element (r, k) at X[k * ld + r]  (thread: one k, 8 consecutive rows)
template <bool kKContig, class T>
__device__ inline void stage_tile(const T* __restrict__ X, long ld, int R, int K, int r0, int k0, int kend, int tid,
                                  uint16_t (*Xs)[kBK + kPad]) {
    uint16_t v[8];
    if (kKContig) {
        const int r = tid >> 2, kc = (tid & 3) * 8;
        int cnt = kend - (k0 + kc);
        cnt = r0 + r < R ? (cnt < 0 ? 0 : (cnt > 8 ? 8 : cnt)) : 0;
        load8(X + (long)(r0 + r) * ld + (k0 + kc), cnt, v);
        uint4 pk;
        pk.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
        pk.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
        pk.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
        pk.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
        *reinterpret_cast<uint4*>(&Xs[r][kc]) = pk;
    } else {
        const int k = tid >> 3, rc = (tid & 7) * 8;
        int cnt = R - (r0 + rc);
        cnt = k0 + k < kend ? (cnt < 0 ? 0 : (cnt > 8 ? 8 : cnt)) : 0;
        load8(X + (long)(k0 + k) * ld + (r0 + rc), cnt, v);
#pragma unroll
        for (int j = 0; j < 8; ++j) Xs[rc + j][k] = v[j];
    }
}

// C[M x N] = A[M x K] B[K x N] on BF16 MFMA (v_mfma_f32_32x32x16_bf16, f32 accumulate):
// 64 x 64 tile per 4-wave workgroup, 32 x 32 per wave, K in steps of 32 through LDS.
// kAKc: A is k-contiguous (A[m * sam + k]); else m-contiguous (A[k * sak + m]).
// kBKc: B is k-contiguous (B[n * sbn + k]); else n-contiguous (B[k * sbk + n]).
// gridDim.z > 1 splits K: workgroup z covers [z * kchunk, (z + 1) * kchunk) and writes its
// raw partial to ws[z][M][N]; k_splitk_sum then adds the partials in z order.
template <class TA, class TB, int EPI, bool kAKc, bool kBKc>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    __shared__ __align__(16) uint16_t As[kBM][kBK + kPad];
    __shared__ __align__(16) uint16_t Bs[kBN][kBK + kPad];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
    const int kbeg = blockIdx.z * g.kchunk;
    const int kend = kbeg + g.kchunk < g.K ? kbeg + g.kchunk : g.K;
    const TA* A = static_cast<const TA*>(g.A);
    const TB* Bm = static_cast<const TB*>(g.B);
    const long lda = kAKc ? g.sam : g.sak;
    const long ldb = kBKc ? g.sbn : g.sbk;
    f32x16 acc;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc[i] = 0.f;
    for (int k0 = kbeg; k0 < kend; k0 += kBK) {
        stage_tile<kAKc>(A, lda, g.M, g.K, m0, k0, kend, tid, As);
        stage_tile<kBKc>(Bm, ldb, g.N, g.K, n0, k0, kend, tid, Bs);
        __syncthreads();
#pragma unroll
        for (int kh = 0; kh < 2; ++kh) {
            const int kk = kh * 16 + (lane >> 5) * 8;
            const s16x8 av = *reinterpret_cast<const s16x8*>(&As[wm + (lane & 31)][kk]);
            const s16x8 bv = *reinterpret_cast<const s16x8*>(&Bs[wn + (lane & 31)][kk]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                          __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int m = m0 + wm + (i >> 2) * 8 + (lane >> 5) * 4 + (i & 3);
        const int n = n0 + wn + (lane & 31);
        if (m >= g.M || n >= g.N) continue;
        if (gridDim.z > 1)
            g.ws[((long)blockIdx.z * g.M + m) * g.N + n] = acc[i];
        else
            epilogue<EPI>(g, m, n, acc[i]);
    }
}

// Split-K reduction: C = sum over z of ws[z] in z order (weight-gradient epilogue).
__global__ __launch_bounds__(256) void k_splitk_sum(const float* __restrict__ ws, int S, int M, int N,
                                                    float* __restrict__ C, long ldc) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)M * N) return;
    float v = 0.f;
    for (int z = 0; z < S; ++z) v += ws[(long)z * M * N + i];
    C[(i / N) * ldc + (i % N)] = v;
}

// Small-batch forward (M <= kGemvMaxB, bf16 operands): one wave per output column, lanes
// split K, f32 accumulation; used for single-scan inference where a 64 x 64 tile would
// leave the chip idle.
constexpr int kGemvMaxB = 4;

template <int EPI>
__global__ __launch_bounds__(256) void k_gemv(GemmArgs g) {
    const int lane = threadIdx.x & 63;
    const int n = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (n >= g.N) return;
    const uint16_t* A = static_cast<const uint16_t*>(g.A);
    const uint16_t* W = static_cast<const uint16_t*>(g.B);
    float acc[kGemvMaxB];
#pragma unroll
    for (int b = 0; b < kGemvMaxB; ++b) acc[b] = 0.f;
    for (int k = lane; k < g.K; k += 64) {
        const float w = bf2f(W[(long)k * g.sbk + (long)n * g.sbn]);
#pragma unroll
        for (int b = 0; b < kGemvMaxB; ++b)
            if (b < g.M) acc[b] = fmaf(bf2f(A[(long)b * g.sam + (long)k * g.sak]), w, acc[b]);
    }
#pragma unroll
    for (int b = 0; b < kGemvMaxB; ++b)
        for (int o = 32; o > 0; o >>= 1) acc[b] += __shfl_xor(acc[b], o, 64);
    if (lane == 0)
        for (int b = 0; b < g.M && b < kGemvMaxB; ++b) epilogue<EPI>(g, b, n, acc[b]);
}

// Mixture negative log-likelihood and its gradient, one sample per lane; per-block loss
// partials (fixed order).  Layout: logits [0,nc), means [nc,5nc), log-sigmas [5nc,9nc).
__global__ __launch_bounds__(256) void k_gmm_head(const float* __restrict__ O, int ldo, const float* __restrict__ Y,
                                                  int B, int nc, float invB, float* __restrict__ dO,
                                                  double* __restrict__ loss_part) {
    __shared__ double sh[4];
    const int m = blockIdx.x * 256 + threadIdx.x;
    double lm = 0.0;
    if (m < B) {
        const float* o = O + (long)m * ldo;
        float* d = dO + (long)m * ldo;
        const float y[4] = {Y[4 * m], Y[4 * m + 1], Y[4 * m + 2], Y[4 * m + 3]};
        float amax = -INFINITY;
        for (int k = 0; k < nc; ++k) amax = fmaxf(amax, o[k]);
        float asum = 0.f;
        for (int k = 0; k < nc; ++k) asum += expf(o[k] - amax);
        const float lse_a = amax + logf(asum);
        auto comp_lp = [&](int k) {
            float lp = o[k] - lse_a;
            for (int j = 0; j < 4; ++j) {
                const float s = fminf(fmaxf(o[5 * nc + 4 * k + j], -7.f), 3.f);
                const float zz = (y[j] - o[nc + 4 * k + j]) * expf(-s);
                lp += -0.5f * zz * zz - s - 0.91893853320467274f;
            }
            return lp;
        };
        float lmax = -INFINITY;
        for (int k = 0; k < nc; ++k) lmax = fmaxf(lmax, comp_lp(k));
        float lsum = 0.f;
        for (int k = 0; k < nc; ++k) lsum += expf(comp_lp(k) - lmax);
        const float ll = lmax + logf(lsum);
        lm = -(double)ll;
        for (int k = 0; k < nc; ++k) {
            const float r = expf(comp_lp(k) - ll);
            const float pi = expf(o[k] - lse_a);
            d[k] = (pi - r) * invB;
            for (int j = 0; j < 4; ++j) {
                const float raw = o[5 * nc + 4 * k + j];
                const float s = fminf(fmaxf(raw, -7.f), 3.f);
                const float is = expf(-s);
                const float zz = (y[j] - o[nc + 4 * k + j]) * is;
                d[nc + 4 * k + j] = -r * zz * is * invB;
                d[5 * nc + 4 * k + j] = (raw > -7.f && raw < 3.f) ? r * (1.f - zz * zz) * invB : 0.f;
            }
        }
        for (int c = 9 * nc; c < ldo; ++c) d[c] = 0.f;
    }
    lm = pfd::wave_sum(lm);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = lm;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Sum of the loss partials of one step into slot `step` (fixed order).
__global__ void k_loss_reduce(const double* __restrict__ part, int nparts, double* __restrict__ out, int step) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < nparts; ++i) s += part[i];
        out[step] = s;
    }
}

// gb[n] = sum_m dZ[m, n] in two passes: part[s][n] sums rows [s * rows, (s + 1) * rows),
// then the kColsumSplit partials are added in s order.
constexpr int kColsumSplit = 128;
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ dZ, int M, int N, long ld, int rows,
                                                     float* __restrict__ part) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= N) return;
    const int m0 = blockIdx.y * rows;
    const int m1 = m0 + rows < M ? m0 + rows : M;
    float s = 0.f;
    for (int m = m0; m < m1; ++m) s += dZ[(long)m * ld + n];
    part[(long)blockIdx.y * N + n] = s;
}
// One wave per column: lane l adds partials l, l + 64, ...; fixed shuffle tree.
__global__ __launch_bounds__(256) void k_colsum_fin(const float* __restrict__ part, int S, int N,
                                                    float* __restrict__ gb) {
    const int n = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (n >= N) return;
    float s = 0.f;
    for (int z = lane; z < S; z += 64) s += part[(long)z * N + n];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) gb[n] = s;
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, float* __restrict__ mo, float* __restrict__ ve,
                                              const float* __restrict__ g, uint16_t* __restrict__ pb, long n, float lr,
                                              float b1, float b2, float eps, float bc1, float bc2) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float m = b1 * mo[i] + (1.f - b1) * gi;
    const float v = b2 * ve[i] + (1.f - b2) * gi * gi;
    mo[i] = m;
    ve[i] = v;
    const float pn = p[i] - lr * (m / bc1) / (sqrtf(v / bc2) + eps);
    p[i] = pn;
    if (pb) pb[i] = f2bf(pn);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ x, long n, uint16_t* __restrict__ y) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < n) y[i] = f2bf(x[i]);
}

// Training batch from rendered ranges: X = bf16(clamp((r + noise N(0,1)) / rmax)), Y = targets.
__global__ __launch_bounds__(256) void k_make_batch(const float* __restrict__ ranges, const float* __restrict__ poses,
                                                    int B, int nrays, int in_dim, float rmax, float noise, uint64_t seed,
                                                    float cx, float cy, float inv_span, uint16_t* __restrict__ X,
                                                    float* __restrict__ Y) {
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= (long)B * in_dim) return;
    const int m = (int)(i / in_dim), r = (int)(i % in_dim);
    float v = 0.f;
    if (r < nrays) {
        float rg = ranges[(long)m * nrays + r];
        if (noise > 0.f) {
            uint64_t x = pfd::sm_final(seed * pfd::kPhi + (uint64_t)i * pfd::kC1);
            const uint64_t z0 = pfd::sm_next(x), z1 = pfd::sm_next(x);
            const float u1 = ((float)(z0 >> 40) + 1.0f) * 0x1.0p-24f;
            const float u2 = (float)(z1 >> 40) * 0x1.0p-24f;
            rg += noise * sqrtf(-2.f * logf(u1)) * cosf(6.2831853f * u2);
        }
        v = fminf(fmaxf(rg / rmax, 0.f), 1.5f);
    }
    X[i] = f2bf(v);
    if (r == 0) {
        const float th = poses[3 * m + 2];
        Y[4 * m] = (poses[3 * m] - cx) * inv_span;
        Y[4 * m + 1] = (poses[3 * m + 1] - cy) * inv_span;
        Y[4 * m + 2] = sinf(th);
        Y[4 * m + 3] = cosf(th);
    }
}

// n hypotheses from the mixture of output row O (one scan): component by inverse CDF of the
// mixture weights, then mean + sigma * N(0,1) per dimension; yaw = atan2(sin, cos).
__global__ __launch_bounds__(256) void k_gmm_sample(const float* __restrict__ O, int nc, int n, uint64_t seed,
                                                    float cx, float cy, float span, float x0, float y0, float x1,
                                                    float y1, float* __restrict__ hyp) {
    const int h = blockIdx.x * 256 + threadIdx.x;
    if (h >= n) return;
    uint64_t x = pfd::sm_final(seed * pfd::kPhi + (uint64_t)h * pfd::kC1);
    float amax = -INFINITY;
    for (int k = 0; k < nc; ++k) amax = fmaxf(amax, O[k]);
    float tot = 0.f;
    for (int k = 0; k < nc; ++k) tot += expf(O[k] - amax);
    const float u = (float)(pfd::sm_next(x) >> 40) * 0x1.0p-24f * tot;
    int kc = nc - 1;
    float c = 0.f;
    for (int k = 0; k < nc; ++k) {
        c += expf(O[k] - amax);
        if (u < c) {
            kc = k;
            break;
        }
    }
    float y[4];
    for (int j = 0; j < 4; j += 2) {
        const uint64_t z0 = pfd::sm_next(x), z1 = pfd::sm_next(x);
        const float u1 = ((float)(z0 >> 40) + 1.0f) * 0x1.0p-24f;
        const float u2 = (float)(z1 >> 40) * 0x1.0p-24f;
        const float rr = sqrtf(-2.f * logf(u1));
        const float n0 = rr * cosf(6.2831853f * u2), n1 = rr * sinf(6.2831853f * u2);
        for (int q = 0; q < 2; ++q) {
            const float s = fminf(fmaxf(O[5 * nc + 4 * kc + j + q], -7.f), 3.f);
            y[j + q] = O[nc + 4 * kc + j + q] + expf(s) * (q == 0 ? n0 : n1);
        }
    }
    // keep every hypothesis inside the map rectangle (a non-finite draw lands at the centre)
    float hx = cx + span * y[0], hy = cy + span * y[1];
    hx = isfinite(hx) ? fminf(fmaxf(hx, x0), x1) : cx;
    hy = isfinite(hy) ? fminf(fmaxf(hy, y0), y1) : cy;
    const float th = atan2f(y[2], y[3]);
    hyp[3 * h] = hx;
    hyp[3 * h + 1] = hy;
    hyp[3 * h + 2] = isfinite(th) ? th : 0.f;
}

// ---------------------------------------------------------------- host side

struct Layer {
    int in = 0, out = 0;
    float *W = nullptr, *b = nullptr;
    uint16_t* Wb = nullptr;
    float *gW = nullptr, *gb = nullptr;
    float *mW = nullptr, *vW = nullptr, *mb = nullptr, *vb = nullptr;
};

struct Mlp {
    int in = 0, hid = 0, nh = 0, nc = 0, out = 0, maxB = 0;
    std::vector<Layer> L;
    uint16_t* X = nullptr;
    std::vector<float*> Z;
    std::vector<uint16_t*> H;
    float *O = nullptr, *dO = nullptr, *dA = nullptr, *dB = nullptr, *Y = nullptr;
    double* loss_part = nullptr;
    float* ws = nullptr;  // split-K partials of the weight-gradient GEMMs
    size_t ws_cap = 0;    // floats
    float* cs_part = nullptr;  // column-sum partials [kColsumSplit][max out]
    double* step_loss = nullptr;
    int step_loss_cap = 0;
    float *poses = nullptr, *ranges = nullptr;
    int ranges_cap = 0;
    hipStream_t stream = nullptr;
    long t = 0;  // Adam step
    uint64_t seed = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_propose_ms = 0.0;
    bool propose_pending = false;  // ev0/ev1 bracket a proposal not yet timed
    uint16_t* h_x = nullptr;       // pinned staging of the proposal input
    hipEvent_t ev_x = nullptr;     // recorded after the staging copy
    bool ev_x_live = false;
};

template <class TA, class TB, int EPI>
int gemm(hipStream_t st, const GemmArgs& g, float* ws = nullptr, size_t ws_cap = 0) {
    if constexpr (std::is_same<TA, uint16_t>::value && std::is_same<TB, uint16_t>::value &&
                  (EPI == EPI_HIDDEN || EPI == EPI_OUT)) {
        if (g.M <= kGemvMaxB) {
            hipLaunchKernelGGL((k_gemv<EPI>), dim3((g.N + 3) / 4), dim3(256), 0, st, g);
            ML_LAUNCH();
            return 0;
        }
    }
    GemmArgs a = g;
    const int gn = (g.N + kBN - 1) / kBN, gm = (g.M + kBM - 1) / kBM;
    int S = 1;
    if (EPI == EPI_WGRAD && ws != nullptr && gm * gn < 256) {
        // split K so that the grid has >= 512 workgroups, each with >= 4 K steps
        S = (512 + gm * gn - 1) / (gm * gn);
        const int max_s = (g.K + 4 * kBK - 1) / (4 * kBK);
        S = S < max_s ? S : max_s;
        while (S > 1 && (size_t)S * g.M * g.N > ws_cap) --S;
        S = S < 1 ? 1 : S;
    }
    a.kchunk = ((g.K + S - 1) / S + kBK - 1) / kBK * kBK;
    S = (g.K + a.kchunk - 1) / a.kchunk;
    a.ws = ws;
    const dim3 grid(gn, gm, S);
    const bool ak = g.sak == 1, bk = g.sbk == 1;
    if ((!ak && g.sam != 1) || (!bk && g.sbn != 1)) {
        fprintf(stderr, "pf_hip gemm: operands must be contiguous along one dimension\n");
        return -1;
    }
    if (ak && bk)
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI, true, true>), grid, dim3(256), 0, st, a);
    else if (ak)
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI, true, false>), grid, dim3(256), 0, st, a);
    else if (bk)
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI, false, true>), grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_gemm<TA, TB, EPI, false, false>), grid, dim3(256), 0, st, a);
    ML_LAUNCH();
    if (S > 1) {
        const long mn = (long)g.M * g.N;
        hipLaunchKernelGGL(k_splitk_sum, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, ws, S, g.M, g.N, g.C,
                           g.ldc);
        ML_LAUNCH();
    }
    return 0;
}

int forward(Mlp* n, int B, hipStream_t st) {
    for (int l = 0; l <= n->nh; ++l) {
        const Layer& ly = n->L[l];
        GemmArgs g{};
        g.M = B;
        g.N = ly.out;
        g.K = ly.in;
        g.A = l == 0 ? (const void*)n->X : (const void*)n->H[l - 1];
        g.sam = ly.in;
        g.sak = 1;
        g.B = ly.Wb;
        g.sbk = 1;
        g.sbn = ly.in;
        g.bias = ly.b;
        if (l < n->nh) {
            g.C = n->Z[l];
            g.ldc = ly.out;
            g.H = n->H[l];
            g.ldh = ly.out;
            if (gemm<uint16_t, uint16_t, EPI_HIDDEN>(st, g) != 0) return -1;
        } else {
            g.C = n->O;
            g.ldc = ly.out;
            if (gemm<uint16_t, uint16_t, EPI_OUT>(st, g) != 0) return -1;
        }
    }
    return 0;
}

// Backward from dO (already holding dLoss/dO); fills gW/gb of every layer.
int backward(Mlp* n, int B, hipStream_t st) {
    const float* dcur = n->dO;  // gradient w.r.t. this layer's pre-activation output
    float* bufs[2] = {n->dA, n->dB};
    for (int l = n->nh; l >= 0; --l) {
        Layer& ly = n->L[l];
        // weight gradient: gW[out x in] = dcur^T [out x B] * A [B x in]
        GemmArgs g{};
        g.M = ly.out;
        g.N = ly.in;
        g.K = B;
        g.A = dcur;
        g.sam = 1;
        g.sak = ly.out;
        g.B = l == 0 ? (const void*)n->X : (const void*)n->H[l - 1];
        g.sbk = ly.in;
        g.sbn = 1;
        g.C = ly.gW;
        g.ldc = ly.in;
        if (gemm<float, uint16_t, EPI_WGRAD>(st, g, n->ws, n->ws_cap) != 0) return -1;
        const int rows = (B + kColsumSplit - 1) / kColsumSplit;
        hipLaunchKernelGGL(k_colsum_part, dim3((ly.out + 255) / 256, kColsumSplit), dim3(256), 0, st, dcur, B, ly.out,
                           (long)ly.out, rows, n->cs_part);
        hipLaunchKernelGGL(k_colsum_fin, dim3((ly.out + 3) / 4), dim3(256), 0, st, n->cs_part, kColsumSplit, ly.out,
                           ly.gb);
        ML_LAUNCH();
        if (l == 0) break;
        // data gradient into the previous layer's pre-activation
        float* dnext = bufs[l & 1];
        GemmArgs d{};
        d.M = B;
        d.N = ly.in;
        d.K = ly.out;
        d.A = dcur;
        d.sam = ly.out;
        d.sak = 1;
        d.B = ly.Wb;
        d.sbk = ly.in;
        d.sbn = 1;
        d.C = dnext;
        d.ldc = ly.in;
        d.Zp = n->Z[l - 1];
        d.ldz = ly.in;
        if (gemm<float, uint16_t, EPI_DGRAD>(st, d) != 0) return -1;
        dcur = dnext;
    }
    return 0;
}

int head(Mlp* n, int B, hipStream_t st) {
    const int blocks = (B + 255) / 256;
    hipLaunchKernelGGL(k_gmm_head, dim3(blocks), dim3(256), 0, st, n->O, n->out, n->Y, B, n->nc, 1.0f / (float)B, n->dO,
                       n->loss_part);
    ML_LAUNCH();
    return 0;
}

int adam(Mlp* n, float lr, hipStream_t st) {
    n->t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.f - powf(b1, (float)n->t), bc2 = 1.f - powf(b2, (float)n->t);
    for (auto& ly : n->L) {
        const long nw = (long)ly.out * ly.in;
        hipLaunchKernelGGL(k_adam, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, st, ly.W, ly.mW, ly.vW, ly.gW,
                           ly.Wb, nw, lr, b1, b2, eps, bc1, bc2);
        hipLaunchKernelGGL(k_adam, dim3((ly.out + 255) / 256), dim3(256), 0, st, ly.b, ly.mb, ly.vb, ly.gb,
                           (uint16_t*)nullptr, (long)ly.out, lr, b1, b2, eps, bc1, bc2);
    }
    ML_LAUNCH();
    return 0;
}

int refresh_bf16(Mlp* n) {
    for (auto& ly : n->L) {
        const long nw = (long)ly.out * ly.in;
        hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, n->stream, ly.W, nw, ly.Wb);
    }
    ML_LAUNCH();
    return 0;
}

int upload_inputs(Mlp* n, const float* x, const float* y, int B) {
    const long nx = (long)B * n->in;
    float* tmp = nullptr;
    ML_CHECK(hipMallocAsync((void**)&tmp, sizeof(float) * nx, n->stream));
    ML_CHECK(hipMemcpyAsync(tmp, x, sizeof(float) * nx, hipMemcpyHostToDevice, n->stream));
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, n->stream, tmp, nx, n->X);
    ML_LAUNCH();
    ML_CHECK(hipFreeAsync(tmp, n->stream));
    if (y) ML_CHECK(hipMemcpyAsync(n->Y, y, sizeof(float) * 4 * (size_t)B, hipMemcpyHostToDevice, n->stream));
    return 0;
}

int read_loss(Mlp* n, int B, double* loss) {
    const int nparts = (B + 255) / 256;
    std::vector<double> part((size_t)nparts);
    ML_CHECK(hipMemcpyAsync(part.data(), n->loss_part, sizeof(double) * nparts, hipMemcpyDeviceToHost, n->stream));
    ML_CHECK(hipStreamSynchronize(n->stream));
    double s = 0.0;
    for (double v : part) s += v;
    if (loss) *loss = s / (double)B;
    return 0;
}

void free_mlp(Mlp* n) {
    if (n->stream) (void)hipStreamSynchronize(n->stream);
    if (n->propose_pending) (void)hipEventSynchronize(n->ev1);
    if (n->ev_x_live) (void)hipEventSynchronize(n->ev_x);
    if (n->h_x) (void)hipHostFree(n->h_x);
    if (n->ev_x) (void)hipEventDestroy(n->ev_x);
    for (auto& ly : n->L) {
        void* ps[] = {ly.W, ly.b, ly.Wb, ly.gW, ly.gb, ly.mW, ly.vW, ly.mb, ly.vb};
        for (void* p : ps) (void)hipFree(p);
    }
    for (auto* p : n->Z) (void)hipFree(p);
    for (auto* p : n->H) (void)hipFree(p);
    void* ps[] = {n->X,         n->O,         n->dO,    n->dA,     n->dB,   n->Y,
                  n->loss_part, n->step_loss, n->poses, n->ranges, n->ws,   n->cs_part};
    for (void* p : ps) (void)hipFree(p);
    if (n->ev0) (void)hipEventDestroy(n->ev0);
    if (n->ev1) (void)hipEventDestroy(n->ev1);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
}

}  // namespace mlk

using namespace mlk;

extern "C" {

void* pfl_create(int in_dim, int hid, int n_hidden, int ncomp, int max_batch, uint64_t seed) {
    if (in_dim <= 0 || hid <= 0 || n_hidden <= 0 || ncomp <= 0 || ncomp > 256 || max_batch <= 0) return nullptr;
    Mlp* n = new (std::nothrow) Mlp();
    if (!n) return nullptr;
    n->in = in_dim;
    n->hid = hid;
    n->nh = n_hidden;
    n->nc = ncomp;
    n->out = 9 * ncomp;
    n->maxB = max_batch;
    n->seed = seed;
    bool ok = hipSetDevice(0) == hipSuccess && hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && hipEventCreate(&n->ev0) == hipSuccess && hipEventCreate(&n->ev1) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&n->ev_x, hipEventDisableTiming) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&n->h_x, 2 * (size_t)in_dim, hipHostMallocDefault) == hipSuccess;
    n->L.resize(n_hidden + 1);
    pfd::dquat unused{};
    (void)unused;
    uint64_t rs = seed ^ 0x5bd1e9955bd1e995ull;
    for (int l = 0; l <= n_hidden && ok; ++l) {
        Layer& ly = n->L[l];
        ly.in = l == 0 ? in_dim : hid;
        ly.out = l == n_hidden ? n->out : hid;
        const size_t nw = (size_t)ly.in * ly.out;
        ok = ok && hipMalloc((void**)&ly.W, 4 * nw) == hipSuccess && hipMalloc((void**)&ly.Wb, 2 * nw) == hipSuccess;
        ok = ok && hipMalloc((void**)&ly.gW, 4 * nw) == hipSuccess && hipMalloc((void**)&ly.mW, 4 * nw) == hipSuccess;
        ok = ok && hipMalloc((void**)&ly.vW, 4 * nw) == hipSuccess;
        ok = ok && hipMalloc((void**)&ly.b, 4 * (size_t)ly.out) == hipSuccess &&
             hipMalloc((void**)&ly.gb, 4 * (size_t)ly.out) == hipSuccess;
        ok = ok && hipMalloc((void**)&ly.mb, 4 * (size_t)ly.out) == hipSuccess &&
             hipMalloc((void**)&ly.vb, 4 * (size_t)ly.out) == hipSuccess;
        if (!ok) break;
        // torch.nn.Linear default initialisation: U(-1/sqrt(in), 1/sqrt(in)) for W and b
        const float bound = 1.0f / sqrtf((float)ly.in);
        std::vector<float> w(nw), b((size_t)ly.out);
        for (auto& v : w) v = (float)((pfd::sm_next(rs) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * bound;
        for (auto& v : b) v = (float)((pfd::sm_next(rs) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * bound;
        ok = ok && hipMemcpy(ly.W, w.data(), 4 * nw, hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemcpy(ly.b, b.data(), 4 * (size_t)ly.out, hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemset(ly.mW, 0, 4 * nw) == hipSuccess && hipMemset(ly.vW, 0, 4 * nw) == hipSuccess;
        ok = ok && hipMemset(ly.mb, 0, 4 * (size_t)ly.out) == hipSuccess &&
             hipMemset(ly.vb, 0, 4 * (size_t)ly.out) == hipSuccess;
    }
    const size_t mb = (size_t)max_batch;
    ok = ok && hipMalloc((void**)&n->X, 2 * mb * in_dim) == hipSuccess;
    n->Z.assign(n_hidden, nullptr);
    n->H.assign(n_hidden, nullptr);
    for (int l = 0; l < n_hidden && ok; ++l)
        ok = hipMalloc((void**)&n->Z[l], 4 * mb * hid) == hipSuccess &&
             hipMalloc((void**)&n->H[l], 2 * mb * hid) == hipSuccess;
    const int wide = hid > in_dim ? hid : in_dim;
    ok = ok && hipMalloc((void**)&n->O, 4 * mb * n->out) == hipSuccess;
    ok = ok && hipMalloc((void**)&n->dO, 4 * mb * n->out) == hipSuccess;
    ok = ok && hipMalloc((void**)&n->dA, 4 * mb * wide) == hipSuccess;
    ok = ok && hipMalloc((void**)&n->dB, 4 * mb * wide) == hipSuccess;
    ok = ok && hipMalloc((void**)&n->Y, 4 * mb * 4) == hipSuccess;
    ok = ok && hipMalloc((void**)&n->loss_part, sizeof(double) * ((mb + 255) / 256)) == hipSuccess;
    ok = ok && hipMalloc((void**)&n->poses, 4 * mb * 3) == hipSuccess;
    size_t max_w = 0;
    int max_out = 0;
    for (auto& ly : n->L) {
        max_w = std::max(max_w, (size_t)ly.in * ly.out);
        max_out = std::max(max_out, ly.out);
    }
    n->ws_cap = 8 * max_w;
    ok = ok && hipMalloc((void**)&n->ws, sizeof(float) * n->ws_cap) == hipSuccess;
    ok = ok && hipMalloc((void**)&n->cs_part, sizeof(float) * kColsumSplit * (size_t)max_out) == hipSuccess;
    if (ok) ok = refresh_bf16(n) == 0 && hipStreamSynchronize(n->stream) == hipSuccess;
    if (!ok) {
        fprintf(stderr, "pf_hip pfl_create: device setup failed\n");
        free_mlp(n);
        return nullptr;
    }
    return n;
}

int pfl_destroy(void* h) {
    if (h) free_mlp(static_cast<Mlp*>(h));
    return 0;
}

int64_t pfl_param_count(void* h) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n) return -1;
    int64_t c = 0;
    for (auto& ly : n->L) c += (int64_t)ly.in * ly.out + ly.out;
    return c;
}

int pfl_get_params(void* h, float* out) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n || !out) return -1;
    ML_CHECK(hipStreamSynchronize(n->stream));
    size_t o = 0;
    for (auto& ly : n->L) {
        const size_t nw = (size_t)ly.in * ly.out;
        ML_CHECK(hipMemcpy(out + o, ly.W, 4 * nw, hipMemcpyDeviceToHost));
        o += nw;
        ML_CHECK(hipMemcpy(out + o, ly.b, 4 * (size_t)ly.out, hipMemcpyDeviceToHost));
        o += ly.out;
    }
    return 0;
}

int pfl_set_params(void* h, const float* in) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n || !in) return -1;
    ML_CHECK(hipStreamSynchronize(n->stream));
    size_t o = 0;
    for (auto& ly : n->L) {
        const size_t nw = (size_t)ly.in * ly.out;
        ML_CHECK(hipMemcpy(ly.W, in + o, 4 * nw, hipMemcpyHostToDevice));
        o += nw;
        ML_CHECK(hipMemcpy(ly.b, in + o, 4 * (size_t)ly.out, hipMemcpyHostToDevice));
        o += ly.out;
    }
    if (refresh_bf16(n) != 0) return -1;
    ML_CHECK(hipStreamSynchronize(n->stream));
    return 0;
}

int pfl_forward(void* h, const float* x, int B, float* out) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n || !x || !out || B <= 0 || B > n->maxB) return -1;
    if (upload_inputs(n, x, nullptr, B) != 0) return -1;
    if (forward(n, B, n->stream) != 0) return -1;
    ML_CHECK(hipMemcpyAsync(out, n->O, sizeof(float) * (size_t)B * n->out, hipMemcpyDeviceToHost, n->stream));
    ML_CHECK(hipStreamSynchronize(n->stream));
    return 0;
}

int pfl_grad(void* h, const float* x, const float* y, int B, float* grads, double* loss) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n || !x || !y || B <= 0 || B > n->maxB) return -1;
    if (upload_inputs(n, x, y, B) != 0) return -1;
    if (forward(n, B, n->stream) != 0 || head(n, B, n->stream) != 0 || backward(n, B, n->stream) != 0) return -1;
    if (grads) {
        size_t o = 0;
        for (auto& ly : n->L) {
            const size_t nw = (size_t)ly.in * ly.out;
            ML_CHECK(hipMemcpyAsync(grads + o, ly.gW, 4 * nw, hipMemcpyDeviceToHost, n->stream));
            o += nw;
            ML_CHECK(hipMemcpyAsync(grads + o, ly.gb, 4 * (size_t)ly.out, hipMemcpyDeviceToHost, n->stream));
            o += ly.out;
        }
    }
    return read_loss(n, B, loss);
}

int pfl_train_step(void* h, const float* x, const float* y, int B, float lr, double* loss) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n || !x || !y || B <= 0 || B > n->maxB) return -1;
    if (upload_inputs(n, x, y, B) != 0) return -1;
    if (forward(n, B, n->stream) != 0 || head(n, B, n->stream) != 0 || backward(n, B, n->stream) != 0) return -1;
    if (adam(n, lr, n->stream) != 0) return -1;
    return read_loss(n, B, loss);
}

int pfl_train_on_map(void* h, void* gl, int64_t n_samples, int batch, float lr, int nrays, float rmax, float ang_span,
                     float noise_sigma, uint64_t seed, double* mean_loss2) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n || !gl || batch <= 0 || batch > n->maxB || nrays != n->in || n_samples < batch) return -1;
    double mi[6];
    if (pfglx_map_info(gl, mi) != 0) return -1;
    const float cx = (float)(mi[3] + 0.5 * mi[0] * mi[2]), cy = (float)(mi[4] + 0.5 * mi[1] * mi[2]);
    const float span = (float)(std::max(mi[0], mi[1]) * mi[2]);
    const int steps = (int)(n_samples / batch);
    if (n->ranges_cap < batch * nrays) {
        (void)hipFree(n->ranges);
        n->ranges = nullptr;
        ML_CHECK(hipMalloc((void**)&n->ranges, sizeof(float) * (size_t)batch * nrays));
        n->ranges_cap = batch * nrays;
    }
    if (n->step_loss_cap < steps) {
        (void)hipFree(n->step_loss);
        n->step_loss = nullptr;
        ML_CHECK(hipMalloc((void**)&n->step_loss, sizeof(double) * (size_t)steps));
        n->step_loss_cap = steps;
    }
    const long nx = (long)batch * n->in;
    for (int s = 0; s < steps; ++s) {
        const uint64_t bs = pfd::sm_final(seed * pfd::kPhi + (uint64_t)s * pfd::kC1);
        if (pfglx_sample_poses_device(gl, batch, bs, n->poses, n->stream) != 0) return -1;
        if (pfglx_render_device(gl, n->poses, batch, nrays, rmax, ang_span, n->ranges, n->stream) != 0) return -1;
        hipLaunchKernelGGL(k_make_batch, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, n->stream, n->ranges,
                           n->poses, batch, nrays, n->in, rmax, noise_sigma, bs ^ 0xA5A5A5A5ull, cx, cy, 1.0f / span,
                           n->X, n->Y);
        ML_LAUNCH();
        if (forward(n, batch, n->stream) != 0 || head(n, batch, n->stream) != 0 ||
            backward(n, batch, n->stream) != 0 || adam(n, lr, n->stream) != 0)
            return -1;
        hipLaunchKernelGGL(k_loss_reduce, dim3(1), dim3(64), 0, n->stream, n->loss_part, (batch + 255) / 256,
                           n->step_loss, s);
        ML_LAUNCH();
    }
    std::vector<double> sl((size_t)steps);
    ML_CHECK(hipMemcpyAsync(sl.data(), n->step_loss, sizeof(double) * steps, hipMemcpyDeviceToHost, n->stream));
    ML_CHECK(hipStreamSynchronize(n->stream));
    if (mean_loss2) {
        const int w = std::max(1, steps / 20);
        double a = 0.0, b = 0.0;
        for (int i = 0; i < w; ++i) {
            a += sl[i];
            b += sl[steps - 1 - i];
        }
        mean_loss2[0] = a / w / batch;
        mean_loss2[1] = b / w / batch;
    }
    return 0;
}

int pfl_propose(void* h, void* gl, const float* scan, int nrays, float rmax, int n_mix, int n_uniform, uint64_t seed) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n || !gl || !scan || nrays != n->in || n_mix < 0 || n_uniform < 0 || n_mix + n_uniform <= 0) return -1;
    // capacity check before anything is written into the hypothesis buffer
    if (pfglx_set_count(gl, n_mix + n_uniform) != 0) {
        fprintf(stderr, "pf_hip pfl_propose: %d hypotheses exceed the pfgl capacity\n", n_mix + n_uniform);
        return -1;
    }
    double mi[6];
    if (pfglx_map_info(gl, mi) != 0) return -1;
    const float cx = (float)(mi[3] + 0.5 * mi[0] * mi[2]), cy = (float)(mi[4] + 0.5 * mi[1] * mi[2]);
    const float span = (float)(std::max(mi[0], mi[1]) * mi[2]);
    hipStream_t st = static_cast<hipStream_t>(pfglx_stream(gl));
    if (n->ev_x_live) ML_CHECK(hipEventSynchronize(n->ev_x));  // staging free again
    if (n->propose_pending) (void)pfl_last_propose_ms(h);      // settle the previous timing
    for (int r = 0; r < n->in; ++r) n->h_x[r] = f2bf(std::min(std::max(scan[r] / rmax, 0.0f), 1.5f));
    ML_CHECK(hipEventRecord(n->ev0, st));
    ML_CHECK(hipMemcpyAsync(n->X, n->h_x, 2 * (size_t)n->in, hipMemcpyHostToDevice, st));
    ML_CHECK(hipEventRecord(n->ev_x, st));
    n->ev_x_live = true;
    if (forward(n, 1, st) != 0) return -1;
    float* hyp = pfglx_hypotheses_device(gl);
    if (n_mix > 0) {
        const float x0 = (float)mi[3], y0 = (float)mi[4];
        const float x1 = (float)(mi[3] + mi[0] * mi[2]) - 1e-4f, y1 = (float)(mi[4] + mi[1] * mi[2]) - 1e-4f;
        hipLaunchKernelGGL(k_gmm_sample, dim3((n_mix + 255) / 256), dim3(256), 0, st, n->O, n->nc, n_mix, seed, cx, cy,
                           span, x0, y0, x1, y1, hyp);
        ML_LAUNCH();
    }
    if (pfglx_generate_range(gl, n_mix, n_uniform, seed ^ 0x9E3779B97F4A7C15ull) != 0) return -1;
    ML_CHECK(hipEventRecord(n->ev1, st));
    n->propose_pending = true;  // returns without waiting; the hypotheses are ordered on the pfgl stream
    return 0;
}

double pfl_last_propose_ms(void* h) {
    Mlp* n = static_cast<Mlp*>(h);
    if (!n) return -1.0;
    if (n->propose_pending) {
        float ms = 0.f;
        if (hipEventSynchronize(n->ev1) != hipSuccess || hipEventElapsedTime(&ms, n->ev0, n->ev1) != hipSuccess)
            return -1.0;
        n->last_propose_ms = ms;
        n->propose_pending = false;
    }
    return n->last_propose_ms;
}

}  // extern "C"

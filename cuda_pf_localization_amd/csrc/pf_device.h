// pf_device.h -- device helpers shared by the particle-filter and CBGL kernels (gfx950).
//
// Deterministic-draw mode.  Process noise is a pure function of (generation seed,
// particle index): the particle key is the splitmix64 finaliser of
// gseed*phi + i*C1 (crates/pf_core/src/cpu.rs:151-161), its stream is the splitmix64
// sequence of src/cuda/pf_kernels.cu:323-328, and normals come from a Box-Muller
// transform whose log, sqrt and sincos are built from IEEE fma/mul/add and integer
// operations only.  Every
// operation here is explicit (the translation units are compiled with
// -ffp-contract=off), so a CPU evaluation of the same sequence is bit-identical.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>

namespace pfd {

constexpr uint64_t kPhi = 0x9E3779B97F4A7C15ull;
constexpr uint64_t kC1 = 0xBF58476D1CE4E5B9ull;
constexpr uint64_t kC2 = 0x94D049BB133111EBull;
constexpr uint64_t kGenBump = 0x9E37797979797979ull;  // cpu.rs:136

__host__ __device__ inline uint64_t sm_final(uint64_t z) {
    z = (z ^ (z >> 30)) * kC1;
    z = (z ^ (z >> 27)) * kC2;
    return z ^ (z >> 31);
}
__host__ __device__ inline uint64_t particle_key(uint64_t gseed, uint64_t i) {
    return sm_final(gseed * kPhi + i * kC1);
}
__host__ __device__ inline uint64_t sm_next(uint64_t& x) {
    x += kPhi;
    return sm_final(x);
}
// A 64-bit constant held in a scalar register pair: the empty asm hides its value, so a
// Horner step fma(p, x, c) becomes one v_fma_f64 with c as its scalar operand instead of two
// v_mov_b32 that rebuild c in vector registers before a v_fmac_f64 (same operands, same bits).
// Used only in kernels with scalar registers to spare: in the shared draw and sincos helpers it
// pushed the CBGL kernels to 103 scalar registers and made them 20 % slower.
__device__ __forceinline__ double sconst(double c) {
    asm("" : "+s"(c));
    return c;
}

__device__ inline double u01_co(uint64_t z) { return (double)(z >> 11) * 0x1.0p-53; }
// sin and cos of x radians: pi/2 Cody-Waite reduction, Taylor on |r| <= pi/4.
__device__ inline void det_sincos(double x, double& so, double& co) {
    const double k = floor(x * 0.63661977236758138 + 0.5);
    double r = __fma_rn(-k, 1.5707963267948966, x);
    r = __fma_rn(-k, 6.123233995736766e-17, r);
    const double r2 = r * r;
    double ps = 1.0 / 355687428096000.0;
    ps = __fma_rn(ps, r2, -1.0 / 1307674368000.0);
    ps = __fma_rn(ps, r2, 1.0 / 6227020800.0);
    ps = __fma_rn(ps, r2, -1.0 / 39916800.0);
    ps = __fma_rn(ps, r2, 1.0 / 362880.0);
    ps = __fma_rn(ps, r2, -1.0 / 5040.0);
    ps = __fma_rn(ps, r2, 1.0 / 120.0);
    ps = __fma_rn(ps, r2, -1.0 / 6.0);
    const double sn = __fma_rn(r * r2, ps, r);
    double pc = 1.0 / 6402373705728000.0;
    pc = __fma_rn(pc, r2, -1.0 / 20922789888000.0);
    pc = __fma_rn(pc, r2, 1.0 / 87178291200.0);
    pc = __fma_rn(pc, r2, -1.0 / 479001600.0);
    pc = __fma_rn(pc, r2, 1.0 / 3628800.0);
    pc = __fma_rn(pc, r2, -1.0 / 40320.0);
    pc = __fma_rn(pc, r2, 1.0 / 720.0);
    pc = __fma_rn(pc, r2, -1.0 / 24.0);
    pc = __fma_rn(pc, r2, 0.5);
    const double cs = __fma_rn(-r2, pc, 1.0);
    const long q = (long)k & 3;
    if (q == 0) {
        so = sn;
        co = cs;
    } else if (q == 1) {
        so = cs;
        co = -sn;
    } else if (q == 2) {
        so = -sn;
        co = -cs;
    } else {
        so = -cs;
        co = sn;
    }
}

// cos(h) and sin(h)/h from h2 = h^2.  Small angles (h2 <= 2^-10, every predict draw at
// rotation noise up to ~0.05 rad): five Taylor terms each, whose first omitted terms are below
// 1e-21 relative.  Otherwise Taylor to h^20 after halving h2 <= 1, then doubling.
__device__ inline void det_cos_sinc(double h2, double& co, double& sinc) {
    if (h2 <= 0x1.0p-10) {
        double c = __fma_rn(h2, -1.0 / 3628800.0, 1.0 / 40320.0);
        c = __fma_rn(c, -h2, 1.0 / 720.0);
        c = __fma_rn(c, -h2, 1.0 / 24.0);
        c = __fma_rn(c, -h2, 0.5);
        co = __fma_rn(c, -h2, 1.0);
        double s = __fma_rn(h2, -1.0 / 39916800.0, 1.0 / 362880.0);
        s = __fma_rn(s, -h2, 1.0 / 5040.0);
        s = __fma_rn(s, -h2, 1.0 / 120.0);
        s = __fma_rn(s, -h2, 1.0 / 6.0);
        sinc = __fma_rn(s, -h2, 1.0);
        return;
    }
    int m = 0;
    while (h2 > 1.0 && m < 64) {
        h2 *= 0.25;
        ++m;
    }
    double c = 1.0 / 2432902008176640000.0;
    c = __fma_rn(c, -h2, 1.0 / 6402373705728000.0);
    c = __fma_rn(c, -h2, 1.0 / 20922789888000.0);
    c = __fma_rn(c, -h2, 1.0 / 87178291200.0);
    c = __fma_rn(c, -h2, 1.0 / 479001600.0);
    c = __fma_rn(c, -h2, 1.0 / 3628800.0);
    c = __fma_rn(c, -h2, 1.0 / 40320.0);
    c = __fma_rn(c, -h2, 1.0 / 720.0);
    c = __fma_rn(c, -h2, 1.0 / 24.0);
    c = __fma_rn(c, -h2, 0.5);
    c = __fma_rn(c, -h2, 1.0);
    double s = 1.0 / 51090942171709440000.0;
    s = __fma_rn(s, -h2, 1.0 / 121645100408832000.0);
    s = __fma_rn(s, -h2, 1.0 / 355687428096000.0);
    s = __fma_rn(s, -h2, 1.0 / 1307674368000.0);
    s = __fma_rn(s, -h2, 1.0 / 6227020800.0);
    s = __fma_rn(s, -h2, 1.0 / 39916800.0);
    s = __fma_rn(s, -h2, 1.0 / 362880.0);
    s = __fma_rn(s, -h2, 1.0 / 5040.0);
    s = __fma_rn(s, -h2, 1.0 / 120.0);
    s = __fma_rn(s, -h2, 1.0 / 6.0);
    s = __fma_rn(s, -h2, 1.0);
    for (int i = 0; i < m; ++i) {
        s = s * c;
        c = __fma_rn(2.0 * c, c, -1.0);
    }
    co = c;
    sinc = s;
}

// Motion-noise draws.  The normals perturb poses stored as f32 (the reference's GPU path draws
// f32 normals), so the Box-Muller transform runs in f32 arithmetic, which gfx950 issues at twice
// the f64 rate: -2 ln, the square root and sine/cosine are short fmaf/mul/add sequences
// (~1e-7 relative) evaluated from the integer words.  Every operation is an IEEE f32 operation
// (correctly rounded fmaf, *, +, conversions), so a CPU evaluation of the same sequence is
// bit-identical.

// -2 ln((a + 1) 2^-32) in f32: k = (float)a + 1 = 2^e m, m in [sqrt(1/2), sqrt(2)),
// ln m = 2 atanh(s) with s = (m - 1)/(m + 1); 1/(m + 1) from a linear minimax seed (relative
// error 0.015) and two Newton steps; the atanh series to s^9.
__device__ inline float mot_neg2ln(uint32_t a) {
    const float k = __fadd_rn(__uint2float_rn(a), 1.0f);
    const uint32_t b = __float_as_uint(k);
    int e = (int)(b >> 23) - 127;
    float m = __uint_as_float((b & 0x007FFFFFu) | 0x3F800000u);
    if (m > 0x1.6a09e6p+0f) {  // f32 sqrt(2)
        m = __fmul_rn(m, 0.5f);
        e += 1;
    }
    const float f = __fsub_rn(m, 1.0f);
    const float d = __fadd_rn(m, 1.0f);
    float y = __fmaf_rn(d, -0.23901600f, 0.98506150f);
    y = __fmaf_rn(y, __fmaf_rn(-d, y, 1.0f), y);
    y = __fmaf_rn(y, __fmaf_rn(-d, y, 1.0f), y);
    const float s = __fmul_rn(f, y);
    const float s2 = __fmul_rn(s, s);
    float p = __fmaf_rn(s2, 1.0f / 9.0f, 1.0f / 7.0f);
    p = __fmaf_rn(p, s2, 1.0f / 5.0f);
    p = __fmaf_rn(p, s2, 1.0f / 3.0f);
    const float h = __fmaf_rn(__fmul_rn(s, s2), p, s);  // ln(m) / 2
    return __fmaf_rn(__int2float_rn(32 - e), 0x1.62e430p+0f, __fmul_rn(-4.0f, h));  // f32 2 ln 2
}

// sqrt(x), x in [0, 45], in f32: three Newton steps on the reciprocal square root from the
// integer seed (~1e-7 relative after f32 rounding); x = 0 gives 0 (the seed stays finite).
__device__ inline float mot_sqrt(float x) {
    float y = __uint_as_float(0x5F3759DFu - (__float_as_uint(x) >> 1));
    const float h = __fmul_rn(0.5f, x);
#pragma unroll
    for (int it = 0; it < 3; ++it) {
        const float t = __fmul_rn(h, y);
        const float r = __fmaf_rn(-t, y, 0.5f);
        y = __fmaf_rn(y, r, y);
    }
    return __fmul_rn(x, y);
}

// sin and cos of the angle 2 pi b / 2^32 in f32: octant o = b >> 29, the 29-bit remainder
// reflected in odd octants, x = remainder * pi / 2^31 in [0, pi/4]; Taylor sin to x^9, cos to
// x^8 (first omitted terms below 3e-8); the octant picks, swaps and negates.
__device__ inline void mot_sincos_turn(uint32_t b, float& so, float& co) {
    const uint32_t o = b >> 29;
    const uint32_t fr = b & 0x1FFFFFFFu;
    const uint32_t odd = o & 1u;
    const uint32_t fx = odd ? 0x20000000u - fr : fr;
    const float x = __fmul_rn(__uint2float_rn(fx), 0x1.921fb6p-30f);  // f32 pi / 2^31
    const float x2 = __fmul_rn(x, x);
    float ps = __fmaf_rn(x2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = __fmaf_rn(ps, x2, 1.0f / 120.0f);
    ps = __fmaf_rn(ps, x2, -1.0f / 6.0f);
    const float sx = __fmaf_rn(__fmul_rn(x, x2), ps, x);
    float pc = __fmaf_rn(x2, 1.0f / 40320.0f, -1.0f / 720.0f);
    pc = __fmaf_rn(pc, x2, 1.0f / 24.0f);
    pc = __fmaf_rn(pc, x2, -0.5f);
    const float cx = __fmaf_rn(x2, pc, 1.0f);
    const uint32_t q = o >> 1;
    const uint32_t swap = odd ^ (q & 1u);
    const float sv = swap ? cx : sx;
    const float cv = swap ? sx : cx;
    so = (q >> 1) ? -sv : sv;
    co = ((q ^ (q >> 1)) & 1u) ? -cv : cv;
}

// Six standard normals of particle i: the particle key and the next two outputs of its
// splitmix64 stream give six 32-bit words; Box-Muller pair p takes its radius from the high
// half of value p, its angle from the low half.  The f32 products are widened to f64 exactly.
__device__ inline void particle_normals(uint64_t gseed, uint64_t i, double n6[6]) {
    uint64_t x = particle_key(gseed, i);
    uint64_t z[3];
    z[0] = x;
    z[1] = sm_next(x);
    z[2] = sm_next(x);
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        // The radius word as an opaque 32-bit register: the empty asm emits nothing, but without
        // it the compiler sees (float)(uint32_t)(z >> 32) as a 64-bit conversion and expands it to
        // five instructions instead of one v_cvt_f32_u32 (same value either way).
        uint32_t hi = (uint32_t)(z[p] >> 32);
        asm("" : "+v"(hi));
        const float r = mot_sqrt(mot_neg2ln(hi));
        float sn, cs;
        mot_sincos_turn((uint32_t)z[p], sn, cs);
        n6[2 * p] = (double)__fmul_rn(r, cs);
        n6[2 * p + 1] = (double)__fmul_rn(r, sn);
    }
}

struct dquat {
    double w, x, y, z;
};

// Hamilton product a (x) b, each component one product and three fmas.
__device__ inline dquat qmul(const dquat& a, const dquat& b) {
    dquat r;
    r.w = __fma_rn(a.w, b.w, -__fma_rn(a.x, b.x, __fma_rn(a.y, b.y, a.z * b.z)));
    r.x = __fma_rn(a.w, b.x, __fma_rn(a.x, b.w, __fma_rn(a.y, b.z, -(a.z * b.y))));
    r.y = __fma_rn(a.w, b.y, __fma_rn(-a.x, b.z, __fma_rn(a.y, b.w, a.z * b.x)));
    r.z = __fma_rn(a.w, b.z, __fma_rn(a.x, b.y, __fma_rn(-a.y, b.x, a.z * b.w)));
    return r;
}

// v' = v + 2 (w c + u x c) with c = u x v, u the vector part: the rotation formula of the
// reference (pf_kernels.cu:119-131, nalgebra UnitQuaternion::transform_vector, t = 2 c), in fmas.
__device__ inline void qrot(const dquat& q, double vx, double vy, double vz, double& ox, double& oy,
                            double& oz) {
    const double cx = __fma_rn(q.y, vz, -(q.z * vy));
    const double cy = __fma_rn(q.z, vx, -(q.x * vz));
    const double cz = __fma_rn(q.x, vy, -(q.y * vx));
    const double dx = __fma_rn(q.y, cz, -(q.z * cy));
    const double dy = __fma_rn(q.z, cx, -(q.x * cz));
    const double dz = __fma_rn(q.x, cy, -(q.y * cx));
    ox = __fma_rn(2.0, __fma_rn(q.w, cx, dx), vx);
    oy = __fma_rn(2.0, __fma_rn(q.w, cy, dy), vy);
    oz = __fma_rn(2.0, __fma_rn(q.w, cz, dz), vz);
}

// Axis-angle noise quaternion exp(a/2) = (cos h, a sinc(h)/2), |a| = 2h (cpu.rs:30-42).
__device__ inline dquat noise_dq(double sx, double sy, double sz, const double* n3) {
    const double ax = sx * n3[0], ay = sy * n3[1], az = sz * n3[2];
    const double ang2 = __fma_rn(ax, ax, __fma_rn(ay, ay, az * az));
    dquat d{1.0, 0.0, 0.0, 0.0};
    if (ang2 > 1e-24) {
        double c, s;
        det_cos_sinc(0.25 * ang2, c, s);
        const double hs = 0.5 * s;
        d.w = c;
        d.x = ax * hs;
        d.y = ay * hs;
        d.z = az * hs;
    }
    return d;
}

// ---------------------------------------------------------------- wave64 reductions

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}
// Tree sum landing in lane 0 with the order l += l+32, +16, ..., +1.
__device__ inline float wave_sum_down(float v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = v + __shfl_down(v, off, 64);
    return v;
}


// Fixed-point systematic-resampling weight floor(w * 2^56), w clamped to [0, 64).
__device__ inline uint64_t qweight(float wf) {
    const double w = (double)wf;
    if (!(w > 0.0)) return 0;
    if (w >= 64.0) return (uint64_t)1 << 62;
    return (uint64_t)(w * 0x1.0p56);
}

// Jacobi rotation (c, s) that zeroes a_pq of a symmetric matrix: tan(2 phi) = 2 a_pq / (a_qq -
// a_pp), the smaller root; the identity when a_pq = 0.
__host__ __device__ inline void jacobi_cs(double app, double aqq, double apq, double& c, double& s) {
    if (apq == 0.0) {
        c = 1.0;
        s = 0.0;
        return;
    }
    const double tau = (aqq - app) / (2.0 * apq);
    const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
    c = 1.0 / sqrt(1.0 + t * t);
    s = t * c;
}

// A <- J^T A J and V <- V J for the rotation J of plane (p, r).
template <int P, int R>
__host__ __device__ inline void jacobi_apply(double (&a)[4][4], double (&v)[4][4], double c, double s) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double kp = a[k][P], kr = a[k][R];
        a[k][P] = c * kp - s * kr;
        a[k][R] = s * kp + c * kr;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double pk = a[P][k], rk = a[R][k];
        a[P][k] = c * pk - s * rk;
        a[R][k] = s * pk + c * rk;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double kp = v[k][P], kr = v[k][R];
        v[k][P] = c * kp - s * kr;
        v[k][R] = s * kp + c * kr;
    }
}

// One round of two rotations in disjoint planes (p1, r1) and (p2, r2): both angles are taken from
// the same matrix (neither rotation changes the entries the other one's angle depends on), so
// the two chains of divisions and square roots are independent.
template <int P1, int R1, int P2, int R2>
__host__ __device__ inline void jacobi_round(double (&a)[4][4], double (&v)[4][4]) {
    double c1, s1, c2, s2;
    jacobi_cs(a[P1][P1], a[R1][R1], a[P1][R1], c1, s1);
    jacobi_cs(a[P2][P2], a[R2][R2], a[P2][R2], c2, s2);
    jacobi_apply<P1, R1>(a, v, c1, s1);
    jacobi_apply<P2, R2>(a, v, c2, s2);
}

// Dominant unit eigenvector of (M + M^T)/2 (4x4, row-major) by Jacobi sweeps in the parallel
// ordering (three rounds of two disjoint rotations cover the six planes), sign chosen so that
// w >= 0.  Returns -1 for a zero eigenvector.
__host__ __device__ inline int quat_mean4_jacobi(const double* M, double* q) {
    // Every loop below is unrolled so that all indices are compile-time constants and the
    // 4x4 arrays stay in registers on the device (one thread runs this per filter).
    double a[4][4], v[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            a[r][c] = 0.5 * (M[4 * r + c] + M[4 * c + r]);
            v[r][c] = r == c ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0, diag = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            diag += a[r][r] * a[r][r];
#pragma unroll
            for (int c = r + 1; c < 4; ++c) off += a[r][c] * a[r][c];
        }
        if (off <= 1e-32 * diag || off == 0.0) break;
        jacobi_round<0, 1, 2, 3>(a, v);
        jacobi_round<0, 2, 1, 3>(a, v);
        jacobi_round<0, 3, 1, 2>(a, v);
    }
    // column of the largest diagonal entry (first one on ties), picked by selects
    double amax = a[0][0];
    double vb[4] = {v[0][0], v[1][0], v[2][0], v[3][0]};
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        const bool take = a[i][i] > amax;
        amax = take ? a[i][i] : amax;
#pragma unroll
        for (int r = 0; r < 4; ++r) vb[r] = take ? v[r][i] : vb[r];
    }
    double nrm = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) nrm += vb[r] * vb[r];
    nrm = sqrt(nrm);
    if (!(nrm > 0.0)) return -1;
    const double sg = vb[0] < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) q[r] = sg * vb[r] / nrm;
    return 0;
}

// The Markley quaternion mean (types.rs:128-144): the dominant unit eigenvector of the moment
// matrix S = (M + M^T)/2 (4x4, row-major, positive semi-definite), sign chosen so that w >= 0.
// S is scaled to unit trace and squared repeatedly, P <- P^2 / tr(P^2): the eigenvalue ratios
// square at every step, so P tends to the projector v v^T, and it has converged when tr(P^2) = 1
// to rounding (a unit-trace PSD matrix with tr(P^2) = 1 has one non-zero eigenvalue); v is then
// the column of the largest diagonal entry, normalised.  Only multiplications, additions, one
// division per step and one square root: a short dependence chain for the single lane that runs
// it on the device.  A spectrum whose two largest eigenvalues lie within ~1e-6 of each other does
// not converge in 24 squarings and is left to the Jacobi sweeps.  Host and device run this same
// function; it uses only IEEE-exact operations (+ - * / sqrt), so both give the same bits under
// -ffp-contract=off.  Returns -1 for a zero eigenvector.
__host__ __device__ __forceinline__ int quat_mean4(const double* M, double* q) {
    double p[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) p[r][c] = 0.5 * (M[4 * r + c] + M[4 * c + r]);
    const double t = (p[0][0] + p[1][1]) + (p[2][2] + p[3][3]);
    if (!(t > 0.0) || !(t < INFINITY)) return quat_mean4_jacobi(M, q);
    const double it = 1.0 / t;
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) p[r][c] *= it;
    bool done = false;
    for (int k = 0; k < 24 && !done; ++k) {
        double s2[4][4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = r; c < 4; ++c) {
                const double x = (p[r][0] * p[0][c] + p[r][1] * p[1][c]) + (p[r][2] * p[2][c] + p[r][3] * p[3][c]);
                s2[r][c] = x;
                s2[c][r] = x;
            }
        const double t2 = (s2[0][0] + s2[1][1]) + (s2[2][2] + s2[3][3]);
        if (!(t2 > 0.0)) return quat_mean4_jacobi(M, q);
        done = t2 >= 1.0 - 1e-15;
        const double i2 = 1.0 / t2;
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < 4; ++c) p[r][c] = s2[r][c] * i2;
    }
    if (!done) return quat_mean4_jacobi(M, q);
    // column of the largest diagonal entry (first one on ties), picked by selects
    double amax = p[0][0];
    double vb[4] = {p[0][0], p[1][0], p[2][0], p[3][0]};
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        const bool take = p[i][i] > amax;
        amax = take ? p[i][i] : amax;
#pragma unroll
        for (int r = 0; r < 4; ++r) vb[r] = take ? p[r][i] : vb[r];
    }
    double nrm = 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) nrm += vb[r] * vb[r];
    nrm = sqrt(nrm);
    if (!(nrm > 0.0)) return -1;
    const double sg = vb[0] < 0.0 ? -1.0 : 1.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) q[r] = sg * vb[r] / nrm;
    return 0;
}

}  // namespace pfd

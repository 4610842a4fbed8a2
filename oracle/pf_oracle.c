/*
 * pf_oracle.c -- CPU oracle for the MI355X particle-filter / CBGL library.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (cuda_pf_localization_amd/) links,
 * loads or calls this file.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker or the timed CPU baseline.
 *
 * It restates, in C with f64 arithmetic, the reference's CPU backend and CPU
 * global-localisation twin (rus1ru/cuda_pf_localization):
 *   - crates/pf_core/src/cpu.rs          (reinit, predict, weight, softmax, resample, estimate)
 *   - crates/pf_core/src/types.rs:117-144 (Markley quaternion mean; own 4x4 Jacobi here)
 *   - crates/pf_core/src/global_loc.rs:136-221, 356-398 (DDA, map_scan, CAER, dispersal)
 * plus the formulas the reference GPU library fixes where the CPU backend has none
 * (src/cuda/pf_kernels.cu: k_inject 333-352, resample LCG 652-657, f32 DDA 784-808).
 *
 * Deterministic-draw mode.  The reference draws its process noise from rand's
 * SmallRng/StandardNormal (cpu.rs:22-28,151-161), whose streams cannot be pinned here.
 * Both this oracle and the GPU library instead consume one specified stream:
 *   key(i)  = cpu.rs:153-159 (splitmix64 finaliser of gseed*phi + i*C1), gseed bumped as
 *             cpu.rs:136 before every reinit/predict;
 *   words   = key, then the splitmix64 outputs of key + phi, key + 2 phi (cu:323-328 mixer),
 *             each split into its high and low 32 bits;
 *   normals = Box-Muller on the three (hi, lo) pairs: radius sqrt(-2 ln((hi + 1) 2^-32)),
 *             angle 2 pi lo 2^-32, in f32 arithmetic with the short fmaf-only log, sqrt and
 *             sincos below (~1e-7 relative), widened to f64, so oracle and GPU produce
 *             bit-identical draws.
 * With f32 storage on (the GPU's storage precision), poses and weights are rounded to
 * f32 at the end of every stage, so per-stage parity can be checked against the GPU on
 * identical inputs.
 *
 * Systematic resampling is restated with fixed-point weights (q_i = floor(w_i * 2^56))
 * and exact integer pointer arithmetic: ancestor(i) = first j with
 * cum_j * n * 2^32 >= T * (U32 + i * 2^32), T = sum q, U32 = top 32 bits of the
 * cu:653 LCG.  This is cpu.rs:264-289 with t_i = u0 + i*T/n evaluated exactly.
 * or_resample_f64_literal keeps the literal f64 form of cpu.rs for comparison.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int32_t id, mode;
    float dx, dy, dz, range;
} or_obs;

/* ---------------------------------------------------------------- draws */

static const uint64_t PHI = 0x9E3779B97F4A7C15ull;
static const uint64_t C1 = 0xBF58476D1CE4E5B9ull;
static const uint64_t C2 = 0x94D049BB133111EBull;

static inline uint64_t sm_final(uint64_t z) {
    z = (z ^ (z >> 30)) * C1;
    z = (z ^ (z >> 27)) * C2;
    return z ^ (z >> 31);
}

/* cpu.rs:151-161: per-particle key for generation seed gseed. */
static inline uint64_t particle_key(uint64_t gseed, uint64_t i) {
    return sm_final(gseed * PHI + i * C1);
}

/* cu:323-328: one splitmix64 step. */
static inline uint64_t sm64_next(uint64_t *x) {
    *x += PHI;
    return sm_final(*x);
}

static inline double u01_closed_open(uint64_t z) { return (double)(z >> 11) * 0x1.0p-53; }

/* fma-only sin/cos of x radians (Cody-Waite pi/2 reduction + Taylor on |r| <= pi/4). */
void or_det_sincos(double x, double *so, double *co) {
    double k = floor(x * 0.63661977236758138 + 0.5);
    double r = fma(-k, 1.5707963267948966, x);
    r = fma(-k, 6.123233995736766e-17, r);
    double r2 = r * r;
    double ps = 1.0 / 355687428096000.0; /* 1/17! */
    ps = fma(ps, r2, -1.0 / 1307674368000.0);
    ps = fma(ps, r2, 1.0 / 6227020800.0);
    ps = fma(ps, r2, -1.0 / 39916800.0);
    ps = fma(ps, r2, 1.0 / 362880.0);
    ps = fma(ps, r2, -1.0 / 5040.0);
    ps = fma(ps, r2, 1.0 / 120.0);
    ps = fma(ps, r2, -1.0 / 6.0);
    double sn = fma(r * r2, ps, r);
    double pc = 1.0 / 6402373705728000.0; /* 1/18! */
    pc = fma(pc, r2, -1.0 / 20922789888000.0);
    pc = fma(pc, r2, 1.0 / 87178291200.0);
    pc = fma(pc, r2, -1.0 / 479001600.0);
    pc = fma(pc, r2, 1.0 / 3628800.0);
    pc = fma(pc, r2, -1.0 / 40320.0);
    pc = fma(pc, r2, 1.0 / 720.0);
    pc = fma(pc, r2, -1.0 / 24.0);
    pc = fma(pc, r2, 0.5);
    double cs = fma(-r2, pc, 1.0);
    long q = (long)k & 3;
    switch (q) {
        case 0: *so = sn; *co = cs; break;
        case 1: *so = cs; *co = -sn; break;
        case 2: *so = -sn; *co = -cs; break;
        default: *so = -cs; *co = sn; break;
    }
}

/* cos(h) and sin(h)/h from h2 = h*h.  h2 <= 2^-10: five Taylor terms each (first omitted terms
 * below 1e-21 relative); otherwise Taylor to h^20 after halving to h2 <= 1, then doubling. */
void or_det_cos_sinc(double h2, double *co, double *sinc) {
    if (h2 <= 0x1.0p-10) {
        double c = fma(h2, -1.0 / 3628800.0, 1.0 / 40320.0);
        c = fma(c, -h2, 1.0 / 720.0);
        c = fma(c, -h2, 1.0 / 24.0);
        c = fma(c, -h2, 0.5);
        *co = fma(c, -h2, 1.0);
        double s = fma(h2, -1.0 / 39916800.0, 1.0 / 362880.0);
        s = fma(s, -h2, 1.0 / 5040.0);
        s = fma(s, -h2, 1.0 / 120.0);
        s = fma(s, -h2, 1.0 / 6.0);
        *sinc = fma(s, -h2, 1.0);
        return;
    }
    int m = 0;
    while (h2 > 1.0 && m < 64) {
        h2 *= 0.25;
        ++m;
    }
    double c = 1.0 / 2432902008176640000.0; /* 1/20! */
    c = fma(c, -h2, 1.0 / 6402373705728000.0);
    c = fma(c, -h2, 1.0 / 20922789888000.0);
    c = fma(c, -h2, 1.0 / 87178291200.0);
    c = fma(c, -h2, 1.0 / 479001600.0);
    c = fma(c, -h2, 1.0 / 3628800.0);
    c = fma(c, -h2, 1.0 / 40320.0);
    c = fma(c, -h2, 1.0 / 720.0);
    c = fma(c, -h2, 1.0 / 24.0);
    c = fma(c, -h2, 0.5);
    c = fma(c, -h2, 1.0);
    double s = 1.0 / 51090942171709440000.0; /* 1/21! */
    s = fma(s, -h2, 1.0 / 121645100408832000.0);
    s = fma(s, -h2, 1.0 / 355687428096000.0);
    s = fma(s, -h2, 1.0 / 1307674368000.0);
    s = fma(s, -h2, 1.0 / 6227020800.0);
    s = fma(s, -h2, 1.0 / 39916800.0);
    s = fma(s, -h2, 1.0 / 362880.0);
    s = fma(s, -h2, 1.0 / 5040.0);
    s = fma(s, -h2, 1.0 / 120.0);
    s = fma(s, -h2, 1.0 / 6.0);
    s = fma(s, -h2, 1.0);
    for (int i = 0; i < m; ++i) {
        s = s * c;
        c = fma(2.0 * c, c, -1.0);
    }
    *co = c;
    *sinc = s;
}

/* The motion draws' Box-Muller transform in f32 arithmetic (each step one IEEE f32 operation:
 * fmaf, *, +, -, conversions; C evaluates float expressions in float on x86-64).
 * -2 ln((a + 1) 2^-32) for a 32-bit word a, ~1e-7 relative: k = (float)a + 1 = 2^e m with m in
 * [sqrt(1/2), sqrt(2)), ln m = 2 atanh(s), s = (m - 1)/(m + 1); the reciprocal from a linear
 * minimax seed on [1 + sqrt(1/2), 1 + sqrt(2)] (relative error 0.015) and two Newton steps, the
 * atanh series to s^9. */
float or_mot_neg2ln(uint32_t a) {
    float k = (float)a + 1.0f;
    uint32_t b;
    memcpy(&b, &k, 4);
    int e = (int)(b >> 23) - 127;
    b = (b & 0x007FFFFFu) | 0x3F800000u;
    float m;
    memcpy(&m, &b, 4);
    if (m > 0x1.6a09e6p+0f) { /* f32 sqrt(2) */
        m = m * 0.5f;
        e += 1;
    }
    float f = m - 1.0f;
    float d = m + 1.0f;
    float y = fmaf(d, -0.23901600f, 0.98506150f);
    y = fmaf(y, fmaf(-d, y, 1.0f), y);
    y = fmaf(y, fmaf(-d, y, 1.0f), y);
    float s = f * y;
    float s2 = s * s;
    float p = fmaf(s2, 1.0f / 9.0f, 1.0f / 7.0f);
    p = fmaf(p, s2, 1.0f / 5.0f);
    p = fmaf(p, s2, 1.0f / 3.0f);
    float h = fmaf(s * s2, p, s); /* ln(m) / 2 */
    return fmaf((float)(32 - e), 0x1.62e430p+0f, -4.0f * h); /* f32 2 ln 2 */
}

/* sqrt(x) for x in [0, 45] in f32 by three Newton steps on the reciprocal square root from the
 * integer seed (~1e-7 relative); x = 0 gives 0 (the seed stays finite). */
float or_mot_sqrt(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    b = 0x5F3759DFu - (b >> 1);
    float y;
    memcpy(&y, &b, 4);
    float h = 0.5f * x;
    for (int it = 0; it < 3; ++it) {
        float t = h * y;
        float r = fmaf(-t, y, 0.5f);
        y = fmaf(y, r, y);
    }
    return x * y;
}

/* sin and cos of the turn fraction b 2^-32 (angle 2 pi b / 2^32) in f32 from the word's bits:
 * octant o = b >> 29, the 29-bit remainder reflected in odd octants, x in [0, pi/4] by one
 * product; Taylor sin to x^9, cos to x^8 (first omitted terms below 3e-8). */
void or_mot_sincos_turn(uint32_t b, float *so, float *co) {
    uint32_t o = b >> 29;
    uint32_t fr = b & 0x1FFFFFFFu;
    uint32_t odd = o & 1u;
    uint32_t fx = odd ? 0x20000000u - fr : fr;
    float x = (float)fx * 0x1.921fb6p-30f; /* f32 pi / 2^31 */
    float x2 = x * x;
    float ps = fmaf(x2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = fmaf(ps, x2, 1.0f / 120.0f);
    ps = fmaf(ps, x2, -1.0f / 6.0f);
    float sx = fmaf(x * x2, ps, x);
    float pc = fmaf(x2, 1.0f / 40320.0f, -1.0f / 720.0f);
    pc = fmaf(pc, x2, 1.0f / 24.0f);
    pc = fmaf(pc, x2, -0.5f);
    float cx = fmaf(x2, pc, 1.0f);
    uint32_t q = o >> 1;
    uint32_t swap = odd ^ (q & 1u);
    float sv = swap ? cx : sx;
    float cv = swap ? sx : cx;
    *so = (q >> 1) ? -sv : sv;
    *co = ((q ^ (q >> 1)) & 1u) ? -cv : cv;
}

/* The 6 standard normals of particle i at generation seed gseed (motion noise): the particle
 * key (cpu.rs:151-161) and the next two outputs of its splitmix64 stream (cu:323-328) give six
 * 32-bit words; Box-Muller pair p takes its radius from word 2p (hi half of a 64-bit value) and
 * its angle from word 2p + 1 (lo half). */
void or_particle_normals(uint64_t gseed, uint64_t i, double n6[6]) {
    uint64_t x = particle_key(gseed, i);
    uint64_t z[3];
    z[0] = x;
    z[1] = sm64_next(&x);
    z[2] = sm64_next(&x);
    for (int p = 0; p < 3; ++p) {
        float r = or_mot_sqrt(or_mot_neg2ln((uint32_t)(z[p] >> 32)));
        float s, c;
        or_mot_sincos_turn((uint32_t)z[p], &s, &c);
        n6[2 * p] = (double)(r * c);
        n6[2 * p + 1] = (double)(r * s);
    }
}

/* ------------------------------------------------------------ quaternions */

typedef struct { double w, x, y, z; } quat;

static inline quat qmul(quat a, quat b) {
    quat r;
    r.w = fma(a.w, b.w, -fma(a.x, b.x, fma(a.y, b.y, a.z * b.z)));
    r.x = fma(a.w, b.x, fma(a.x, b.w, fma(a.y, b.z, -(a.z * b.y))));
    r.y = fma(a.w, b.y, fma(-a.x, b.z, fma(a.y, b.w, a.z * b.x)));
    r.z = fma(a.w, b.z, fma(a.x, b.y, fma(-a.y, b.x, a.z * b.w)));
    return r;
}

/* v' = v + 2 (w c + u x c), c = u x v (cu:119-131; nalgebra transform_vector with t = 2 c). */
static inline void qrot(quat q, const double v[3], double o[3]) {
    double cx = fma(q.y, v[2], -(q.z * v[1]));
    double cy = fma(q.z, v[0], -(q.x * v[2]));
    double cz = fma(q.x, v[1], -(q.y * v[0]));
    double dx = fma(q.y, cz, -(q.z * cy));
    double dy = fma(q.z, cx, -(q.x * cz));
    double dz = fma(q.x, cy, -(q.y * cx));
    o[0] = fma(2.0, fma(q.w, cx, dx), v[0]);
    o[1] = fma(2.0, fma(q.w, cy, dy), v[1]);
    o[2] = fma(2.0, fma(q.w, cz, dz), v[2]);
}

/* cpu.rs:30-42 noise_dq in the deterministic formulation (cos h, a*sinc(h)/2). */
static inline quat noise_dq(const double std3[3], const double n3[3]) {
    double ax = std3[0] * n3[0], ay = std3[1] * n3[1], az = std3[2] * n3[2];
    double ang2 = fma(ax, ax, fma(ay, ay, az * az));
    quat d = {1.0, 0.0, 0.0, 0.0};
    if (ang2 > 1e-24) {
        double c, s;
        or_det_cos_sinc(0.25 * ang2, &c, &s);
        double hs = 0.5 * s;
        d.w = c;
        d.x = ax * hs;
        d.y = ay * hs;
        d.z = az * hs;
    }
    return d;
}

/* ------------------------------------------------------------------ filter */

typedef struct {
    int n;
    int f32_storage;
    uint64_t gseed, gen, lcg;
    double *px, *py, *pz, *qw, *qx, *qy, *qz, *w, *ll;
    int32_t *anc;
} OrFilter;

static inline double st(const OrFilter *f, double v) { return f->f32_storage ? (double)(float)v : v; }

OrFilter *or_create(int n, uint64_t seed, int f32_storage) {
    if (n <= 0) return NULL;
    OrFilter *f = (OrFilter *)calloc(1, sizeof(OrFilter));
    f->n = n;
    f->f32_storage = f32_storage;
    f->gseed = seed ? seed : 42;
    f->gen = 0;
    f->lcg = 0x243F6A8885A308D3ull; /* cu:116 */
    double **arr[] = {&f->px, &f->py, &f->pz, &f->qw, &f->qx, &f->qy, &f->qz, &f->w, &f->ll};
    for (int k = 0; k < 9; ++k) *arr[k] = (double *)calloc((size_t)n, sizeof(double));
    f->anc = (int32_t *)calloc((size_t)n, sizeof(int32_t));
    for (int i = 0; i < n; ++i) {
        f->qw[i] = 1.0;
        f->w[i] = st(f, 1.0 / n);
        f->anc[i] = i;
    }
    return f;
}

void or_destroy(OrFilter *f) {
    if (!f) return;
    double *arr[] = {f->px, f->py, f->pz, f->qw, f->qx, f->qy, f->qz, f->w, f->ll};
    for (int k = 0; k < 9; ++k) free(arr[k]);
    free(f->anc);
    free(f);
}

void or_set_rng(OrFilter *f, uint64_t gseed, uint64_t gen, uint64_t lcg) {
    f->gseed = gseed;
    f->gen = gen;
    f->lcg = lcg;
}
void or_get_rng(const OrFilter *f, uint64_t out3[3]) {
    out3[0] = f->gseed;
    out3[1] = f->gen;
    out3[2] = f->lcg;
}

void or_set_state(OrFilter *f, const double *pos3n, const double *quat4n, const double *w) {
    for (int i = 0; i < f->n; ++i) {
        f->px[i] = pos3n[3 * i];
        f->py[i] = pos3n[3 * i + 1];
        f->pz[i] = pos3n[3 * i + 2];
        f->qw[i] = quat4n[4 * i];
        f->qx[i] = quat4n[4 * i + 1];
        f->qy[i] = quat4n[4 * i + 2];
        f->qz[i] = quat4n[4 * i + 3];
        f->w[i] = w ? w[i] : st(f, 1.0 / f->n);
    }
}

void or_get_state(const OrFilter *f, double *pos3n, double *quat4n, double *w, double *ll) {
    for (int i = 0; i < f->n; ++i) {
        if (pos3n) {
            pos3n[3 * i] = f->px[i];
            pos3n[3 * i + 1] = f->py[i];
            pos3n[3 * i + 2] = f->pz[i];
        }
        if (quat4n) {
            quat4n[4 * i] = f->qw[i];
            quat4n[4 * i + 1] = f->qx[i];
            quat4n[4 * i + 2] = f->qy[i];
            quat4n[4 * i + 3] = f->qz[i];
        }
        if (w) w[i] = f->w[i];
        if (ll) ll[i] = f->ll[i];
    }
}

void or_get_ancestors(const OrFilter *f, int32_t *anc) { memcpy(anc, f->anc, sizeof(int32_t) * (size_t)f->n); }

static void set_uniform(OrFilter *f) {
    double u = st(f, 1.0 / f->n);
    for (int i = 0; i < f->n; ++i) f->w[i] = u;
}

/* cpu.rs:180-195 */
void or_reinit(OrFilter *f, const double pose7[7], const double pos_std[3], const double rot_std[3]) {
    f->gseed += 0x9E37797979797979ull; /* cpu.rs:136 */
    f->gen += 1;                       /* cu:572 */
    quat q0 = {pose7[3], pose7[4], pose7[5], pose7[6]};
    const uint64_t gs = f->gseed;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < f->n; ++i) {
        double nn[6];
        or_particle_normals(gs, (uint64_t)i, nn);
        double p0 = fma(pos_std[0], nn[0], pose7[0]);
        double p1 = fma(pos_std[1], nn[1], pose7[1]);
        double p2 = fma(pos_std[2], nn[2], pose7[2]);
        quat dq = noise_dq(rot_std, nn + 3);
        quat q = qmul(q0, dq);
        f->px[i] = st(f, p0);
        f->py[i] = st(f, p1);
        f->pz[i] = st(f, p2);
        f->qw[i] = st(f, q.w);
        f->qx[i] = st(f, q.x);
        f->qy[i] = st(f, q.y);
        f->qz[i] = st(f, q.z);
    }
    set_uniform(f);
}

/* cpu.rs:197-218 */
void or_predict(OrFilter *f, const double t[3], const double qodo[4], const double tstd[3], const double rstd[3]) {
    f->gseed += 0x9E37797979797979ull;
    f->gen += 1;
    quat qo = {qodo[0], qodo[1], qodo[2], qodo[3]};
    const uint64_t gs = f->gseed;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < f->n; ++i) {
        double nn[6];
        or_particle_normals(gs, (uint64_t)i, nn);
        double body[3] = {fma(tstd[0], nn[0], t[0]), fma(tstd[1], nn[1], t[1]), fma(tstd[2], nn[2], t[2])};
        quat q = {f->qw[i], f->qx[i], f->qy[i], f->qz[i]};
        double wd[3];
        qrot(q, body, wd);
        double p0 = f->px[i] + wd[0];
        double p1 = f->py[i] + wd[1];
        double p2 = f->pz[i] + wd[2];
        quat dq = noise_dq(rstd, nn + 3);
        quat qn = qmul(qmul(q, qo), dq);
        f->px[i] = st(f, p0);
        f->py[i] = st(f, p1);
        f->pz[i] = st(f, p2);
        f->qw[i] = st(f, qn.w);
        f->qx[i] = st(f, qn.x);
        f->qy[i] = st(f, qn.y);
        f->qz[i] = st(f, qn.z);
    }
}

/* cpu.rs:44-97 MeasModel + cpu.rs:220-260 weight().  Observations whose id is outside
 * [0, lm_count) are dropped (cu:242; cuda.rs:230 filters by map.has before the call). */
int or_weight(OrFilter *f, const or_obs *obs, int nobs, const float *lm, int lm_count, double sigma_r,
              double sigma_b, double gate) {
    int nv = 0;
    for (int k = 0; k < nobs; ++k)
        if (obs[k].id >= 0 && obs[k].id < lm_count) ++nv;
    if (nobs <= 0 || nv == 0) {
        for (int i = 0; i < f->n; ++i) f->ll[i] = 0.0;
        set_uniform(f);
        return 0;
    }
    const double sr = sigma_r > 1e-6 ? sigma_r : 1e-6;
    const double sb = sigma_b > 1e-6 ? sigma_b : 1e-6;
    const double g2 = gate * gate;
    const double inv2rr = 1.0 / (2.0 * sr * sr), inv2bb = 1.0 / (2.0 * sb * sb);
    const double cart_gate2 = g2 * 3.0 * sr * sr, rng_gate2 = g2 * sr * sr, brg_gate2 = g2 * sb * sb;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < f->n; ++i) {
        quat qc = {f->qw[i], -f->qx[i], -f->qy[i], -f->qz[i]};
        double acc = 0.0;
        for (int k = 0; k < nobs; ++k) {
            const or_obs *o = &obs[k];
            if (o->id < 0 || o->id >= lm_count) continue;
            double d[3] = {(double)lm[3 * o->id] - f->px[i], (double)lm[3 * o->id + 1] - f->py[i],
                           (double)lm[3 * o->id + 2] - f->pz[i]};
            double pred[3];
            qrot(qc, d, pred);
            double vx = o->dx, vy = o->dy, vz = o->dz;
            if (o->mode == 0) {
                double ex = pred[0] - vx, ey = pred[1] - vy, ez = pred[2] - vz;
                double r2 = ex * ex + ey * ey + ez * ez;
                acc += (gate > 0.0 && r2 > cart_gate2) ? -0.5 * gate * gate * 3.0 : -r2 * inv2rr;
            } else {
                double rn = sqrt(pred[0] * pred[0] + pred[1] * pred[1] + pred[2] * pred[2]);
                if (rn < 1e-9) rn = 1e-9;
                double er = rn - (double)o->range;
                double dot = (-pred[0] / rn) * vx + (-pred[1] / rn) * vy + (-pred[2] / rn) * vz;
                if (dot < -1.0) dot = -1.0;
                if (dot > 1.0) dot = 1.0;
                double eb = acos(dot);
                double rl = (gate > 0.0 && er * er > rng_gate2) ? -0.5 * gate * gate : -er * er * inv2rr;
                double bl = (gate > 0.0 && eb * eb > brg_gate2) ? -0.5 * gate * gate : -eb * eb * inv2bb;
                acc += rl + bl;
            }
        }
        f->ll[i] = acc;
    }
    double mx = -INFINITY;
    for (int i = 0; i < f->n; ++i)
        if (f->ll[i] > mx) mx = f->ll[i];
    double sum = 0.0;
    for (int i = 0; i < f->n; ++i) sum += exp(f->ll[i] - mx);
    if (!(sum > 0.0) || !isfinite(sum)) {
        set_uniform(f);
    } else {
        for (int i = 0; i < f->n; ++i) f->w[i] = st(f, exp(f->ll[i] - mx) / sum);
    }
    return 0;
}

/* Fixed-point weight: floor(w * 2^56) for w in [0, 64); non-positive / NaN -> 0. */
static inline uint64_t qweight(double w) {
    if (!(w > 0.0)) return 0;
    if (w >= 64.0) return (uint64_t)1 << 62;
    return (uint64_t)(w * 0x1.0p56);
}

/* cu:323-352 injection picks with the pick-ordered f64 accumulation of cpu.rs:291-302. */
static void inject(OrFilter *f, double *opx, double *opy, double *opz, int k) {
    for (int j = 0; j < k; ++j) {
        uint64_t z = f->lcg ^ (f->gen * PHI) ^ (((uint64_t)j + 1) * C1);
        if (z == 0) z = 0x243F6A8885A308D3ull;
        double u = u01_closed_open(sm64_next(&z));
        int idx = (int)(u * f->n);
        if (idx >= f->n) idx = f->n - 1;
        double jx = u01_closed_open(sm64_next(&z)) - 0.5;
        double jy = u01_closed_open(sm64_next(&z)) - 0.5;
        double jz = u01_closed_open(sm64_next(&z)) - 0.5;
        opx[idx] += jx;
        opy[idx] += jy;
        opz[idx] += jz;
    }
}

static int inject_count(float ratio, int n) {
    if (!(ratio > 0.f)) return 0;
    int k = (int)(ratio * (float)n); /* cu:667-671 (f32 product) */
    return k > n ? n : k;
}

/* Systematic resampling with exact integer pointers (see header) + injection. */
int or_resample(OrFilter *f, float ratio) {
    const int n = f->n;
    uint64_t *cum = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    uint64_t T = 0;
    for (int i = 0; i < n; ++i) {
        T += qweight(f->w[i]);
        cum[i] = T;
    }
    if (T == 0) { /* cu:647-650 */
        free(cum);
        set_uniform(f);
        return 0;
    }
    f->lcg = f->lcg * 6364136223846793005ull + 1442695040888963407ull; /* cu:653 */
    const uint64_t U32 = f->lcg >> 32;
    const unsigned __int128 A = (unsigned __int128)((uint64_t)n << 32);
    double *o[7];
    for (int c = 0; c < 7; ++c) o[c] = (double *)malloc(sizeof(double) * (size_t)n);
    int j = 0;
    for (int i = 0; i < n; ++i) {
        unsigned __int128 rhs = (unsigned __int128)T * (unsigned __int128)(U32 + ((uint64_t)i << 32));
        while (j < n - 1 && (unsigned __int128)cum[j] * A < rhs) ++j;
        f->anc[i] = j;
        o[0][i] = f->px[j];
        o[1][i] = f->py[j];
        o[2][i] = f->pz[j];
        o[3][i] = f->qw[j];
        o[4][i] = f->qx[j];
        o[5][i] = f->qy[j];
        o[6][i] = f->qz[j];
    }
    inject(f, o[0], o[1], o[2], inject_count(ratio, n));
    double *dst[7] = {f->px, f->py, f->pz, f->qw, f->qx, f->qy, f->qz};
    for (int c = 0; c < 7; ++c) {
        for (int i = 0; i < n; ++i) dst[c][i] = st(f, o[c][i]);
        free(o[c]);
    }
    free(cum);
    set_uniform(f);
    return 0;
}

/* Literal cpu.rs:264-289 (f64 cumulative sum, f64 pointers) with a caller-given
 * u0 fraction in [0,1): ancestors only, for comparison with the integer form. */
void or_resample_f64_literal(const double *w, int n, double ufrac, int32_t *anc) {
    double *cum = (double *)malloc(sizeof(double) * (size_t)n);
    cum[0] = w[0];
    for (int i = 1; i < n; ++i) cum[i] = cum[i - 1] + w[i];
    double total = cum[n - 1];
    double step = total / n;
    double u0 = ufrac * step;
    int j = 0;
    for (int i = 0; i < n; ++i) {
        double t = u0 + i * step;
        while (j < n - 1 && cum[j] < t) ++j;
        anc[i] = j;
    }
    free(cum);
}

/* Integer systematic ancestors for given weights and U32 (no state change). */
void or_systematic_int(const double *w, int n, uint32_t U32, int32_t *anc) {
    uint64_t *cum = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)n);
    uint64_t T = 0;
    for (int i = 0; i < n; ++i) {
        T += qweight(w[i]);
        cum[i] = T;
    }
    const unsigned __int128 A = (unsigned __int128)((uint64_t)n << 32);
    int j = 0;
    for (int i = 0; i < n; ++i) {
        unsigned __int128 rhs = (unsigned __int128)T * (unsigned __int128)((uint64_t)U32 + ((uint64_t)i << 32));
        while (j < n - 1 && (unsigned __int128)cum[j] * A < rhs) ++j;
        anc[i] = j;
    }
    free(cum);
}

/* pfc_est_sums semantics (cu:354-404 layout), f64 sequential. */
void or_est_sums(const OrFilter *f, double out[14]) {
    for (int k = 0; k < 14; ++k) out[k] = 0.0;
    for (int i = 0; i < f->n; ++i) {
        double w = f->w[i];
        double q[4] = {f->qw[i], f->qx[i], f->qy[i], f->qz[i]};
        out[0] += w * f->px[i];
        out[1] += w * f->py[i];
        out[2] += w * f->pz[i];
        int e = 3;
        for (int r = 0; r < 4; ++r)
            for (int c = r; c < 4; ++c) out[e++] += w * q[r] * q[c];
        out[13] += w * w;
    }
}

/* pfc_covs semantics with cpu.rs:322-356 arithmetic. */
void or_covs(const OrFilter *f, const double mp[3], const double mq[4], double out[12]) {
    for (int k = 0; k < 12; ++k) out[k] = 0.0;
    quat mc = {mq[0], -mq[1], -mq[2], -mq[3]};
    for (int i = 0; i < f->n; ++i) {
        double w = f->w[i];
        double dp[3] = {f->px[i] - mp[0], f->py[i] - mp[1], f->pz[i] - mp[2]};
        int e = 0;
        for (int r = 0; r < 3; ++r)
            for (int c = r; c < 3; ++c) out[e++] += w * dp[r] * dp[c];
        quat qi = {f->qw[i], f->qx[i], f->qy[i], f->qz[i]};
        quat d = qmul(mc, qi);
        if (d.w < 0.0) {
            d.w = -d.w;
            d.x = -d.x;
            d.y = -d.y;
            d.z = -d.z;
        }
        double hv = sqrt(d.x * d.x + d.y * d.y + d.z * d.z);
        double ang = 2.0 * atan2(hv, d.w);
        double a[3] = {0.0, 0.0, 0.0};
        if (hv > 1e-12) {
            a[0] = d.x * ang / hv;
            a[1] = d.y * ang / hv;
            a[2] = d.z * ang / hv;
        }
        e = 6;
        for (int r = 0; r < 3; ++r)
            for (int c = r; c < 3; ++c) out[e++] += w * a[r] * a[c];
    }
}

/* Symmetric 4x4 Jacobi eigen-solver: largest-eigenvalue eigenvector of (M+M^T)/2,
 * normalised, sign w >= 0 (types.rs:128-144). */
void or_quat_mean(const double M[16], double q[4]) {
    double a[4][4], v[4][4];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            a[r][c] = 0.5 * (M[4 * r + c] + M[4 * c + r]);
            v[r][c] = (r == c) ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 64; ++sweep) {
        double off = 0.0;
        for (int r = 0; r < 4; ++r)
            for (int c = r + 1; c < 4; ++c) off += a[r][c] * a[r][c];
        if (off < 1e-300) break;
        for (int p = 0; p < 3; ++p)
            for (int qq = p + 1; qq < 4; ++qq) {
                if (a[p][qq] == 0.0) continue;
                double theta = (a[qq][qq] - a[p][p]) / (2.0 * a[p][qq]);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 4; ++k) {
                    double akp = a[k][p], akq = a[k][qq];
                    a[k][p] = c * akp - s * akq;
                    a[k][qq] = s * akp + c * akq;
                }
                for (int k = 0; k < 4; ++k) {
                    double apk = a[p][k], aqk = a[qq][k];
                    a[p][k] = c * apk - s * aqk;
                    a[qq][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 4; ++k) {
                    double vkp = v[k][p], vkq = v[k][qq];
                    v[k][p] = c * vkp - s * vkq;
                    v[k][qq] = s * vkp + c * vkq;
                }
            }
    }
    int best = 0;
    for (int i = 1; i < 4; ++i)
        if (a[i][i] > a[best][best]) best = i;
    double nrm = 0.0;
    for (int r = 0; r < 4; ++r) nrm += v[r][best] * v[r][best];
    nrm = sqrt(nrm);
    double sg = v[0][best] < 0.0 ? -1.0 : 1.0;
    for (int r = 0; r < 4; ++r) q[r] = sg * v[r][best] / nrm;
}

/* Full cpu.rs:308-368 estimate().  est[0..2] mean pos, [3..6] mean quat (wxyz),
 * [7..15] pos_cov row-major, [16..24] att_cov row-major, [25] ess. */
void or_estimate(const OrFilter *f, double est[26]) {
    double s[14];
    or_est_sums(f, s);
    double M[16];
    int e = 3;
    for (int r = 0; r < 4; ++r)
        for (int c = r; c < 4; ++c) {
            M[4 * r + c] = s[e];
            M[4 * c + r] = s[e];
            ++e;
        }
    double q[4];
    or_quat_mean(M, q);
    double cv[12];
    or_covs(f, s, q, cv);
    est[0] = s[0];
    est[1] = s[1];
    est[2] = s[2];
    for (int k = 0; k < 4; ++k) est[3 + k] = q[k];
    const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            est[7 + 3 * r + c] = cv[idx[r][c]];
            est[16 + 3 * r + c] = cv[6 + idx[r][c]];
        }
    est[25] = s[13] > 0.0 ? 1.0 / s[13] : 0.0;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* ================================================================== CBGL */

typedef struct {
    int W, H;
    double res, ox, oy;
    const uint8_t *cells;
} or_grid;

/* global_loc.rs:136-149 */
static int w2c64(const or_grid *g, double wx, double wy, long *cx, long *cy) {
    double fx = (wx - g->ox) / g->res, fy = (wy - g->oy) / g->res;
    if (fx < 0.0 || fy < 0.0) return 0;
    long x = (long)fx, y = (long)fy;
    if (x < g->W && y < g->H) {
        *cx = x;
        *cy = y;
        return 1;
    }
    return 0;
}

/* global_loc.rs:152-196 (f64 DDA). */
/* The f64 DDA; *steps (when non-NULL) receives the number of cell steps the walk took. */
static double cast_ray64_steps(int W, int H, double res, double ox, double oy, const uint8_t *cells, double px,
                               double py, double dirx, double diry, double rmax, long *steps);

double or_cast_ray64(int W, int H, double res, double ox, double oy, const uint8_t *cells, double px, double py,
                     double dirx, double diry, double rmax) {
    return cast_ray64_steps(W, H, res, ox, oy, cells, px, py, dirx, diry, rmax, NULL);
}

/* Algorithmic work of a CBGL scoring pass: the total number of DDA cell steps of the f64 twin
 * (global_loc.rs:151-196) over n hypotheses x nrays rays. */
unsigned long long or_dda64_steps(int W, int H, double res, double ox, double oy, const uint8_t *cells,
                                  const double *hyp, int n, int nrays, double span, double rmax) {
    unsigned long long total = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : total)
    for (int h = 0; h < n; ++h)
        for (int r = 0; r < nrays; ++r) {
            double a = hyp[3 * h + 2] + (nrays > 1 ? span * (double)r / (double)(nrays - 1) : 0.0);
            long st = 0;
            cast_ray64_steps(W, H, res, ox, oy, cells, hyp[3 * h], hyp[3 * h + 1], cos(a), sin(a), rmax, &st);
            total += (unsigned long long)st;
        }
    return total;
}

static double cast_ray64_steps(int W, int H, double res, double ox, double oy, const uint8_t *cells, double px,
                               double py, double dirx, double diry, double rmax, long *steps) {
    or_grid g = {W, H, res, ox, oy, cells};
    long cx, cy;
    if (steps) *steps = 0;
    if (!w2c64(&g, px, py, &cx, &cy)) return rmax;
    long sx = dirx > 0.0 ? 1 : -1, sy = diry > 0.0 ? 1 : -1;
    double adx = fabs(dirx), ady = fabs(diry);
    double tdx = adx > 1e-9 ? res / adx : INFINITY;
    double tdy = ady > 1e-9 ? res / ady : INFINITY;
    double frx = (px - ox) / res - (double)cx;
    double fry = (py - oy) / res - (double)cy;
    double tmx = adx > 1e-9 ? (dirx > 0.0 ? 1.0 - frx : frx) * tdx : INFINITY;
    double tmy = ady > 1e-9 ? (diry > 0.0 ? 1.0 - fry : fry) * tdy : INFINITY;
    for (int s = 0; s < 8192; ++s) {
        double dist;
        if (steps) *steps = s + 1;
        if (tmx <= tmy) {
            dist = tmx;
            tmx += tdx;
            cx += sx;
        } else {
            dist = tmy;
            tmy += tdy;
            cy += sy;
        }
        if (dist > rmax) return rmax;
        if (cx < 0 || cy < 0 || cx >= W || cy >= H) return rmax;
        if (cells[cy * W + cx] == 1) return dist < rmax ? dist : rmax;
    }
    return rmax;
}

/* global_loc.rs:200-214 */
void or_map_scan64(int W, int H, double res, double ox, double oy, const uint8_t *cells, double x, double y,
                   double theta, int nrays, double span, double rmax, double *out) {
    for (int r = 0; r < nrays; ++r) {
        double a = theta + (nrays > 1 ? span * (double)r / (double)(nrays - 1) : 0.0);
        out[r] = or_cast_ray64(W, H, res, ox, oy, cells, x, y, cos(a), sin(a), rmax);
    }
}

/* global_loc.rs:217-221 */
double or_caer64(int W, int H, double res, double ox, double oy, const uint8_t *cells, const double *scan, int nrays,
                 double x, double y, double th, double span, double rmax) {
    double s = 0.0;
    for (int r = 0; r < nrays; ++r) {
        double a = th + (nrays > 1 ? span * (double)r / (double)(nrays - 1) : 0.0);
        double v = or_cast_ray64(W, H, res, ox, oy, cells, x, y, cos(a), sin(a), rmax);
        s += fabs(scan[r] - v);
    }
    return s;
}

/* global_loc.rs:357-382 (xorshift dispersal); hyp[3n] out. Returns 0, -1 if no free cell. */
int or_disperse64(int W, int H, double res, double ox, double oy, const uint8_t *cells, uint64_t seed, int n,
                  double *hyp) {
    int nfree = 0;
    for (int i = 0; i < W * H; ++i)
        if (cells[i] == 0) ++nfree;
    if (nfree == 0) return -1;
    int *fx = (int *)malloc(sizeof(int) * (size_t)nfree), *fy = (int *)malloc(sizeof(int) * (size_t)nfree);
    int m = 0;
    for (int i = 0; i < W * H; ++i)
        if (cells[i] == 0) {
            fx[m] = i % W;
            fy[m] = i / W;
            ++m;
        }
    uint64_t s = seed | 1;
#define NEXTU() (s ^= s << 13, s ^= s >> 7, s ^= s << 17, (double)(s % 1000000ull) / 1000000.0)
    for (int h = 0; h < n; ++h) {
        double u = NEXTU();
        long idx = (long)(u * (double)nfree);
        if (idx > nfree - 1) idx = nfree - 1;
        double ux = NEXTU();
        double uy = NEXTU();
        double ut = NEXTU();
        hyp[3 * h] = ox + ((double)fx[idx] + ux) * res;
        hyp[3 * h + 1] = oy + ((double)fy[idx] + uy) * res;
        hyp[3 * h + 2] = -3.141592653589793 + ut * 6.283185307179586;
    }
#undef NEXTU
    free(fx);
    free(fy);
    return 0;
}

/* CAER of hypotheses hyp[3n] (f64 twin), OpenMP over hypotheses (rayon par_iter). */
void or_caer64_batch(int W, int H, double res, double ox, double oy, const uint8_t *cells, const double *scan,
                     int nrays, double span, double rmax, const double *hyp, int n, double *caer) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int h = 0; h < n; ++h)
        caer[h] = or_caer64(W, H, res, ox, oy, cells, scan, nrays, hyp[3 * h], hyp[3 * h + 1], hyp[3 * h + 2], span,
                            rmax);
}

/* ---- device-semantics (f32) twin of the GPU path ---- */

/* Hypothesis h in deterministic-draw mode: cell = floor(u53 * nfree), x/y offsets and
 * yaw from 24-bit f32 uniforms, arithmetic in f32 like cu:817-822. */
void or_gen_hyp32(const int32_t *free_xy, int nfree, float res, float ox, float oy, uint64_t seed, int n,
                  float *hyp) {
    for (int h = 0; h < n; ++h) {
        uint64_t x = sm_final(seed * PHI + (uint64_t)h * C1);
        uint64_t z0 = sm64_next(&x), z1 = sm64_next(&x);
        uint64_t z2 = sm64_next(&x), z3 = sm64_next(&x);
        long c = (long)(u01_closed_open(z0) * (double)nfree);
        if (c > nfree - 1) c = nfree - 1;
        float ux = (float)(z1 >> 40) * 0x1.0p-24f;
        float uy = (float)(z2 >> 40) * 0x1.0p-24f;
        float ut = (float)(z3 >> 40) * 0x1.0p-24f;
        float fxc = (float)free_xy[2 * c] + ux;
        float fyc = (float)free_xy[2 * c + 1] + uy;
        float fxr = fxc * res;
        float fyr = fyc * res;
        hyp[3 * h] = ox + fxr;
        hyp[3 * h + 1] = oy + fyr;
        float tt = 6.28318530f * ut;
        hyp[3 * h + 2] = -3.14159265f + tt;
    }
}

/* f32 DDA with the operation sequence of cu:784-808. */
float or_cast_ray32(int W, int H, float res, float ox, float oy, const uint8_t *cells, float px, float py, float dirx,
                    float diry, float rmax) {
    float fx = (px - ox) / res, fy = (py - oy) / res;
    int cx = (int)floorf(fx), cy = (int)floorf(fy);
    if (!(cx >= 0 && cy >= 0 && cx < W && cy < H)) return rmax;
    int sx = dirx > 0.f ? 1 : -1, sy = diry > 0.f ? 1 : -1;
    float adx = fabsf(dirx), ady = fabsf(diry);
    float tdx = adx > 1e-9f ? res / adx : INFINITY;
    float tdy = ady > 1e-9f ? res / ady : INFINITY;
    float frx = fx - (float)cx, fry = fy - (float)cy;
    float tmx = adx > 1e-9f ? ((dirx > 0.f ? 1.f - frx : frx) * tdx) : INFINITY;
    float tmy = ady > 1e-9f ? ((diry > 0.f ? 1.f - fry : fry) * tdy) : INFINITY;
    for (int s = 0; s < 8192; ++s) {
        float dist;
        if (tmx <= tmy) {
            dist = tmx;
            tmx += tdx;
            cx += sx;
        } else {
            dist = tmy;
            tmy += tdy;
            cy += sy;
        }
        if (dist > rmax) return rmax;
        if (cx < 0 || cy < 0 || cx >= W || cy >= H) return rmax;
        if (cells[cy * W + cx]) return dist;
    }
    return rmax;
}

/* Ray-angle table: cos/sin of span*r/(nrays-1) in f64 (libm). */
void or_ray_table(int nrays, float span, double *cphi, double *sphi) {
    for (int r = 0; r < nrays; ++r) {
        double phi = nrays > 1 ? (double)span * (double)r / (double)(nrays - 1) : 0.0;
        cphi[r] = cos(phi);
        sphi[r] = sin(phi);
    }
}

/* Device-semantics CAER: direction of ray r = rotation of the table entry by the
 * deterministic sincos of theta, rounded to f32; |real - virt| summed per lane
 * (r mod 64, ascending) then by the wave64 tree 32,16,8,4,2,1. */
void or_caer32_batch(int W, int H, float res, float ox, float oy, const uint8_t *cells, const float *scan, int nrays,
                     float span, float rmax, const float *hyp, int n, float *caer) {
    double *cphi = (double *)malloc(sizeof(double) * (size_t)nrays);
    double *sphi = (double *)malloc(sizeof(double) * (size_t)nrays);
    or_ray_table(nrays, span, cphi, sphi);
#pragma omp parallel for schedule(dynamic, 64)
    for (int h = 0; h < n; ++h) {
        float x = hyp[3 * h], y = hyp[3 * h + 1];
        double st, ct;
        or_det_sincos((double)hyp[3 * h + 2], &st, &ct);
        float lane[64];
        for (int l = 0; l < 64; ++l) lane[l] = 0.f;
        for (int r = 0; r < nrays; ++r) {
            float dx = (float)(ct * cphi[r] - st * sphi[r]);
            float dy = (float)(st * cphi[r] + ct * sphi[r]);
            float vm = or_cast_ray32(W, H, res, ox, oy, cells, x, y, dx, dy, rmax);
            lane[r & 63] += fabsf(scan[r] - vm);
        }
        for (int off = 32; off > 0; off >>= 1)
            for (int l = 0; l < off; ++l) lane[l] = lane[l] + lane[l + off];
        caer[h] = lane[0];
    }
    free(cphi);
    free(sphi);
}

/* Device-semantics render of scans for poses[3n] -> out[n*nrays]. */
void or_render32(int W, int H, float res, float ox, float oy, const uint8_t *cells, int nrays, float span, float rmax,
                 const float *poses, int n, float *out) {
    double *cphi = (double *)malloc(sizeof(double) * (size_t)nrays);
    double *sphi = (double *)malloc(sizeof(double) * (size_t)nrays);
    or_ray_table(nrays, span, cphi, sphi);
#pragma omp parallel for schedule(static)
    for (int h = 0; h < n; ++h) {
        double st, ct;
        or_det_sincos((double)poses[3 * h + 2], &st, &ct);
        for (int r = 0; r < nrays; ++r) {
            float dx = (float)(ct * cphi[r] - st * sphi[r]);
            float dy = (float)(st * cphi[r] + ct * sphi[r]);
            out[(size_t)h * nrays + r] =
                or_cast_ray32(W, H, res, ox, oy, cells, poses[3 * h], poses[3 * h + 1], dx, dy, rmax);
        }
    }
    free(cphi);
    free(sphi);
}

/*
 * mlp_oracle.c -- CPU f64 restatement of the learned-proposal network (TEST INFRASTRUCTURE).
 *
 * Only tests/ (and bench.py's CPU baseline leg) may load this.  It is the checker for the
 * HIP library's pfl_* entry points, which have no counterpart in the reference
 * (rus1ru/cuda_pf_localization disperses hypotheses uniformly, global_loc.rs:356-382).
 *
 * Network: in -> hid (x n_hidden, exact GELU) -> 9*nc.  Output layout: mixture logits
 * [0,nc), means [nc,5nc) component-major, log standard deviations [5nc,9nc) clamped to
 * [-7,3] (zero gradient outside the open interval).  Loss = mean over the batch of
 * -log sum_k pi_k prod_j N(y_j; mu_kj, sigma_kj).
 *
 * emulate_bf16 = 1 rounds exactly the operands the GPU rounds to bfloat16 (round to
 * nearest even of the f32 value): network inputs, weights, hidden activations after GELU,
 * and the upstream gradient entering each weight-gradient and data-gradient product.  All
 * sums are f64.  emulate_bf16 = 0 is plain f64 back-propagation (pinned against torch
 * autograd and finite differences by the tests).
 *
 * Parameter and gradient layout: layer by layer, W (out x in, row-major) then b (out).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double rbf(double x, int em) {
    if (!em) return x;
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double gelu(double z) { return 0.5 * z * (1.0 + erf(z * 0.70710678118654752440)); }
static double gelu_grad(double z) {
    return 0.5 * (1.0 + erf(z * 0.70710678118654752440)) + z * 0.39894228042143772 * exp(-0.5 * z * z);
}

typedef struct {
    int nl;          /* layers = n_hidden + 1 */
    int din[16], dout[16];
    long woff[16], boff[16];
} net_t;

static int net_init(net_t* n, int in, int hid, int nh, int nc) {
    if (nh + 1 > 16 || nh < 1) return -1;
    n->nl = nh + 1;
    long o = 0;
    for (int l = 0; l < n->nl; ++l) {
        n->din[l] = l == 0 ? in : hid;
        n->dout[l] = l == nh ? 9 * nc : hid;
        n->woff[l] = o;
        o += (long)n->din[l] * n->dout[l];
        n->boff[l] = o;
        o += n->dout[l];
    }
    return 0;
}

long om_param_count(int in, int hid, int nh, int nc) {
    net_t n;
    if (net_init(&n, in, hid, nh, nc)) return -1;
    return n.boff[n.nl - 1] + n.dout[n.nl - 1];
}

/* Forward; A[l] = input of layer l (rounded as the GPU rounds it), Z[l] = pre-activation. */
static void fwd(const net_t* n, const float* P, const float* x, int B, int em, double** A, double** Z) {
    for (long i = 0; i < (long)B * n->din[0]; ++i) A[0][i] = rbf(x[i], em);
    for (int l = 0; l < n->nl; ++l) {
        const int di = n->din[l], dn = n->dout[l];
        const float* W = P + n->woff[l];
        const float* b = P + n->boff[l];
#pragma omp parallel for schedule(static)
        for (int m = 0; m < B; ++m) {
            const double* a = A[l] + (long)m * di;
            for (int o = 0; o < dn; ++o) {
                const float* w = W + (long)o * di;
                double s = 0.0;
                for (int i = 0; i < di; ++i) s += a[i] * rbf(w[i], em);
                const double z = s + b[o];
                Z[l][(long)m * dn + o] = z;
                if (l + 1 < n->nl) A[l + 1][(long)m * dn + o] = rbf(gelu(z), em);
            }
        }
    }
}

static double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Loss of one output row and d(loss)/d(row) scaled by `scale`. */
static double head_row(const double* o, const double* y, int nc, double scale, double* d) {
    double amax = -INFINITY;
    for (int k = 0; k < nc; ++k) amax = fmax(amax, o[k]);
    double as = 0.0;
    for (int k = 0; k < nc; ++k) as += exp(o[k] - amax);
    const double lse_a = amax + log(as);
    double* lp = (double*)malloc(sizeof(double) * (size_t)nc);
    double lmax = -INFINITY;
    for (int k = 0; k < nc; ++k) {
        double v = o[k] - lse_a;
        for (int j = 0; j < 4; ++j) {
            const double s = clampd(o[5 * nc + 4 * k + j], -7.0, 3.0);
            const double z = (y[j] - o[nc + 4 * k + j]) * exp(-s);
            v += -0.5 * z * z - s - 0.91893853320467274178;
        }
        lp[k] = v;
        lmax = fmax(lmax, v);
    }
    double ls = 0.0;
    for (int k = 0; k < nc; ++k) ls += exp(lp[k] - lmax);
    const double ll = lmax + log(ls);
    if (d) {
        for (int k = 0; k < nc; ++k) {
            const double r = exp(lp[k] - ll);
            const double pi = exp(o[k] - lse_a);
            d[k] = (pi - r) * scale;
            for (int j = 0; j < 4; ++j) {
                const double raw = o[5 * nc + 4 * k + j];
                const double s = clampd(raw, -7.0, 3.0);
                const double is = exp(-s);
                const double z = (y[j] - o[nc + 4 * k + j]) * is;
                d[nc + 4 * k + j] = -r * z * is * scale;
                d[5 * nc + 4 * k + j] = (raw > -7.0 && raw < 3.0) ? r * (1.0 - z * z) * scale : 0.0;
            }
        }
    }
    free(lp);
    return -ll;
}

static double** alloc_layers(const net_t* n, int B, int inputs) {
    double** v = (double**)calloc((size_t)n->nl + 1, sizeof(double*));
    for (int l = 0; l <= n->nl; ++l) {
        int d;
        if (inputs)
            d = l == 0 ? n->din[0] : n->dout[l - 1];
        else
            d = l < n->nl ? n->dout[l] : 1;
        v[l] = (double*)calloc((size_t)B * (size_t)d, sizeof(double));
    }
    return v;
}

static void free_layers(const net_t* n, double** v) {
    for (int l = 0; l <= n->nl; ++l) free(v[l]);
    free(v);
}

int om_forward(int in, int hid, int nh, int nc, const float* P, const float* x, int B, int em, double* out) {
    net_t n;
    if (net_init(&n, in, hid, nh, nc) || B <= 0) return -1;
    double** A = alloc_layers(&n, B, 1);
    double** Z = alloc_layers(&n, B, 0);
    fwd(&n, P, x, B, em, A, Z);
    memcpy(out, Z[n.nl - 1], sizeof(double) * (size_t)B * (size_t)(9 * nc));
    free_layers(&n, A);
    free_layers(&n, Z);
    return 0;
}

/* Mean mixture NLL of raw output rows o[B x 9nc] against targets y[B x 4]; optional dO. */
double om_head(const double* o, const double* y, int B, int nc, double* dO) {
    double s = 0.0;
    for (int m = 0; m < B; ++m)
        s += head_row(o + (long)m * 9 * nc, y + 4 * m, nc, 1.0 / B, dO ? dO + (long)m * 9 * nc : NULL);
    return s / B;
}

int om_loss_grad(int in, int hid, int nh, int nc, const float* P, const float* x, const float* y, int B, int em,
                 double* grads, double* loss) {
    net_t n;
    if (net_init(&n, in, hid, nh, nc) || B <= 0) return -1;
    double** A = alloc_layers(&n, B, 1);
    double** Z = alloc_layers(&n, B, 0);
    fwd(&n, P, x, B, em, A, Z);
    const int no = 9 * nc;
    double* dcur = (double*)malloc(sizeof(double) * (size_t)B * (size_t)no);
    double lsum = 0.0;
    for (int m = 0; m < B; ++m) {
        double yy[4] = {y[4 * m], y[4 * m + 1], y[4 * m + 2], y[4 * m + 3]};
        lsum += head_row(Z[n.nl - 1] + (long)m * no, yy, nc, 1.0 / B, dcur + (long)m * no);
    }
    if (loss) *loss = lsum / B;
    for (int l = n.nl - 1; l >= 0; --l) {
        const int di = n.din[l], dn = n.dout[l];
        const float* W = P + n.woff[l];
        double* gW = grads + n.woff[l];
        double* gb = grads + n.boff[l];
#pragma omp parallel for schedule(static)
        for (int o = 0; o < dn; ++o) {
            double sb = 0.0;
            for (int m = 0; m < B; ++m) sb += dcur[(long)m * dn + o];
            gb[o] = sb;
            for (int i = 0; i < di; ++i) {
                double s = 0.0;
                for (int m = 0; m < B; ++m) s += rbf(dcur[(long)m * dn + o], em) * A[l][(long)m * di + i];
                gW[(long)o * di + i] = s;
            }
        }
        if (l == 0) break;
        double* dprev = (double*)malloc(sizeof(double) * (size_t)B * (size_t)di);
#pragma omp parallel for schedule(static)
        for (int m = 0; m < B; ++m)
            for (int i = 0; i < di; ++i) {
                double s = 0.0;
                for (int o = 0; o < dn; ++o) s += rbf(dcur[(long)m * dn + o], em) * rbf(W[(long)o * di + i], em);
                dprev[(long)m * di + i] = s * gelu_grad(Z[l - 1][(long)m * di + i]);
            }
        free(dcur);
        dcur = dprev;
    }
    free(dcur);
    free_layers(&n, A);
    free_layers(&n, Z);
    return 0;
}

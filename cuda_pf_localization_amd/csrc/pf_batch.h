// pf_batch.h -- particle-store layout, per-filter parameter blocks and handle state.
//
// One handle holds nf >= 1 filters back to back in the same SoA streams.  Filter f owns
// particles [off_f, off_f + n_f) of every stream, off_f a multiple of kTile, so a
// 256-particle block or a 1024-particle tile never straddles two filters.  The reference
// ABI (pfc_*) drives a handle with nf = 1; the batched ABI (pfb_*) drives nf > 1 with the
// same kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstddef>
#include <vector>

namespace pfk {

constexpr int kTB = 256;     // threads per block
constexpr int kTile = 1024;  // filter alignment, scan tile and estimate tile
constexpr int kEstN = 14;    // pfc_est_sums outputs (pf_kernels.cu:71)
constexpr int kCovN = 12;    // pfc_covs outputs (pf_kernels.cu:72)

struct PoseSet {
    float* c[7];  // px py pz qw qx qy qz
};
struct PoseBufs {
    PoseSet s[2];  // ping-pong: F[f].parity selects the current set
};

// Device-resident per-filter state (scalars of the last weight / scan / estimate passes).
struct FilterDev {
    double max_ll, sum_e, sum_e2;  // last weight pass: max ll, sum exp(ll-max), sum exp^2
    double sum_w2, ess;             // last normalisation: sum of squared stored weights, 1/sum
    int uniform;      // last weight pass degenerate -> weights are 1/n
    int parity;       // current pose set
    int do_resample;  // resample predicate of the current cycle
    uint32_t u32;     // top 32 bits of the advanced LCG (systematic offset)
    uint64_t total;   // fixed-point weight total of the last scan
    uint64_t lcg;     // resampling LCG state (pf_kernels.cu:116, 653)
    double sums[kEstN];
    double covs[kCovN];
    double qmean[4];  // device Markley mean of the last estimate (wxyz, w >= 0)
    int qvalid;       // 0: degenerate moment matrix (qmean = identity)
    // 1: the last resampling gather left the estimate-sum partials of the current poses and
    // weights (one per 1024-particle tile); cleared by every pass that changes poses or weights
    int est_src;
};

struct Layout {
    const int* blk_filter;  // filter of every 256-particle block
    const int* f_n;
    const int* f_off;
    const int* f_padn;      // n rounded up to kTile
    const float* f_winv;    // (float)(1/n)
    int nf;                 // filters (1: every block belongs to filter 0, no lookup)
};

// Observation joined with its landmark, read as wave-uniform scalar loads (64 B).  For a
// filter on the scaled range-bearing path (WeightParams.fast_rb) the range-bearing records
// hold v / (2 u_gate) and 4 range instead of v and range.
struct ObsRec {
    double lx, ly, lz;
    double vx, vy, vz;
    double range;
    int32_t mode, pad;
};

struct MotionParams {
    double t[3];    // predict: body translation; reinit: prior position
    double q[4];    // predict: odometry rotation; reinit: prior attitude (wxyz)
    double s_t[3];  // translation / position std
    double s_r[3];  // rotation std
    uint64_t gseed; // generation seed of this call
    float winv;     // reinit: 1/n
    int pad;
};

struct WeightParams {
    double inv2rr, inv2bb, gate;
    double cart_gate2, brg_gate2;  // squared-residual gate thresholds (+inf: no gate)
    double rng_gate;               // |range residual| gate g sr (+inf: no gate)
    double u_gate;                            // 1 - cos of the bearing gate (+inf: none)
    // Scaled range-bearing path (u_gate <= 0.005): residuals formed from 2 b, the bearing as
    // us = u / u_gate so that the gate is the [0, 1] clamp, eb^2 = us (c0 + c1 us + c2 us^2),
    // the range residual at 4x scale against rng_gate4 = 4 g sr.
    int fast_rb, pad_fast;
    double cg;                 // 1 / u_gate
    double c0, c1, c2;         // series coefficients in us
    double rng_gate4;
    const ObsRec* obs;
    int nobs;  // joined records
    int nrb;   // of which range-bearing (they come first)
};

struct InjectParams {
    uint64_t gen;  // generation counter (pf_kernels.cu:337)
    int k;         // number of picks
    int pad;
};

struct CovParams {
    double mp[3];
    double mq[4];
};

// Parameters by value (one filter) or from a device array indexed by filter.
template <class T>
struct ParamRef {
    const T* dev;
    T val;
    __device__ T get(int f) const { return dev ? dev[f] : val; }
    static ParamRef value(const T& v) {
        ParamRef r;
        r.dev = nullptr;
        r.val = v;
        return r;
    }
    static ParamRef array(const T* d) {
        ParamRef r;
        r.dev = d;
        r.val = T{};
        return r;
    }
};

struct ProfEvent {
    int stage;
    hipEvent_t a, b;
};

struct PfState {
    hipStream_t stream = nullptr;
    int obs_cap = 0;
    // layout
    int nf = 0, NP = 0;
    std::vector<int> n, off, padn;
    int *d_blk_filter = nullptr, *d_fn = nullptr, *d_foff = nullptr, *d_fpadn = nullptr;
    float* d_fwinv = nullptr;
    Layout L{};
    // particle streams
    PoseBufs pose{};
    float* w = nullptr;
    double* ll = nullptr;
    uint64_t* cum = nullptr;
    int32_t* anc = nullptr;
    int32_t *inj_next = nullptr, *inj_head = nullptr;  // injection pick lists (head -1: empty)
    double* inj_jit = nullptr;
    int32_t* d_inj_blk = nullptr;  // particle-layout blocks that can hold injection picks
    int inj_nblk = 0;
    // reduction partials / scan state
    double *part_w = nullptr, *part_n = nullptr, *part_est = nullptr, *part_cov = nullptr;
    double* cov_fold = nullptr;  // staged covariance finalize: one fold per lane group and filter
    uint64_t* part_q = nullptr;                // fixed-point weight aggregate of every 1024-particle tile
    uint64_t* tcum = nullptr;                  // cumulative fixed-point weight at every tile's last particle
    uint64_t* tpre = nullptr;                  // exclusive prefix of every tile's aggregate (large filters)
    uint64_t* scan_st = nullptr;               // look-back status words, two per 1024-particle tile
    unsigned long long* scan_ctl = nullptr;     // status epoch of the look-back words
    int scan_spin_limit = 64;                   // reads of a predecessor's status before summing it here
    FilterDev* fdev = nullptr;
    // observations (reference ABI path)
    ObsRec* d_obs = nullptr;
    ObsRec* h_obs = nullptr;  // pinned staging
    hipEvent_t obs_ev = nullptr;
    // pinned landing buffer and completion event of the ABI's small synchronous reads
    double* h_small = nullptr;
    hipEvent_t small_ev = nullptr;
    std::vector<float> lm;  // host mirror of the dense landmark table
    int lm_cap = 0;
    // deterministic-draw state (single-filter ABI)
    uint64_t gseed = 42, gen = 0;
    // launch shape of the likelihood kernel (particles per lane: 1, 2 or 4)
    int weight_ppt = 1;
    // part_q holds the fixed-point tile aggregates of the current weights (left by the last
    // normalisation; cleared by every launch or copy that writes the weights otherwise), so the
    // ABI's resampling scan can start from them instead of looking back
    bool agg_ok = false;
    // profiling
    bool prof_on = false;
    std::vector<ProfEvent> prof_pending;
};

int state_alloc(PfState* s, const std::vector<int>& ns);
void state_free_particles(PfState* s);
void prof_begin(PfState* s, int stage);
void prof_end(PfState* s);
// dst_flip = 1 (predict only): the new poses go to the idle pose set, leaving the current set
// untouched until the likelihood launch's finalize makes the new set current (WeightDefer).
int launch_motion(PfState* s, ParamRef<MotionParams> A, int reinit, int dst_flip = 0);
// Work of a staged cycle carried into the next cycle's likelihood launch, which nothing in that
// cycle before it needs.  flip: predict wrote the idle pose set; the likelihood reads it and its
// softmax finalize makes it current.  cov (non-null, requires flip): the previous cycle's
// covariance pass, run by extra likelihood workgroups on the still-current pose set and weights
// with that cycle's covariance parameter blocks, and its finalize, run by extra softmax-finalize
// workgroups and completed by the normalisation launch (launch_normalize's cov_out), into
// FilterDev.covs and the results slot out.
struct WeightDefer {
    int flip = 0;
    const CovParams* cov = nullptr;
    FilterDev* out = nullptr;
};
// Likelihood pass and its softmax finalize.
int launch_weight(PfState* s, ParamRef<WeightParams> A, const WeightDefer& d = {});
// Normalised weights, ESS and (with d_ratio) the resample predicate; with scan (fused cycle,
// d_ratio required) the normalisation pass also leaves every tile's fixed-point aggregate and
// draws the injection picks inj of a possible resample, and the resampling scan of every filter
// follows, its last tile forming ESS, the predicate and the LCG advance.
// cov_out (staged cycles, with scan): extra workgroups finish the previous cycle's deferred
// covariance finalize into FilterDev.covs and cov_out.
int launch_normalize(PfState* s, int force_uniform, const float* d_ratio, bool scan = false,
                     ParamRef<InjectParams> inj = ParamRef<InjectParams>::value({}), FilterDev* cov_out = nullptr);
// Scan (unless launch_normalize already scanned), injection picks, gather (with the picks' jitter
// and the estimate-sum partials of every filter); with flip the resampled set then becomes
// current (else the fused cycle's sums finalize does it).
int launch_resample(PfState* s, ParamRef<InjectParams> inj, const int32_t* inj_tab, int inj_nblk, bool flip,
                    bool scan = true);
// Block table of the injection kernels from each filter's largest pick count.
int set_inject_blocks(PfState* s, const std::vector<int>& kmax);
// Estimate sums (the partials of a preceding resampling gather when it left them); with d_cov
// non-null also the Markley mean of every filter's sums on the device -> FilterDev.qmean and
// the covariance pass parameters d_cov[f]; with flip the resampled set becomes current.
// pass = false: the gather of the same cycle left the partials of every filter.
int launch_est_sums(PfState* s, CovParams* d_cov, bool flip, bool pass = true);
// Covariances; with results non-null every filter's scalars are also copied to results[f].
int launch_est_covs(PfState* s, ParamRef<CovParams> A, FilterDev* results = nullptr);
int launch_predicate(PfState* s, const float* d_ratio, int force);
WeightParams make_weight_params(double sigma_r, double sigma_b, double gate);
// Joins observations with their landmarks (unknown ids dropped); range-bearing records
// first, then Cartesian; *nrb = number of range-bearing records.  The range-bearing records
// are stored scaled for the path p selects (p.fast_rb).
int join_obs(const struct pfc_obs* obs, int nobs, const float* lm, int lm_cap, const WeightParams& p, ObsRec* out,
             int* nrb);

// Device address of a FilterDev member of filter f.
template <class M>
inline M* fdev_field(PfState* s, int f, size_t offset) {
    return reinterpret_cast<M*>(reinterpret_cast<char*>(s->fdev + f) + offset);
}

}  // namespace pfk

// pf_batched.hip -- batched filters (pfb_* C ABI): nf independent particle filters of uneven
// sizes share the SoA streams of one handle and advance with one launch per stage.
//
// A cycle (particle_filter.rs:110-124 for every filter at once):
//   host   : per-filter odometry / weight / injection parameters and the observations joined
//            with each filter's landmark table -> one pinned block, one H2D copy
//   device : motion -> likelihood (+ softmax partials) -> normalise (+ ESS of the stored
//            weights and the resample predicate ESS < ratio n) -> scan -> gather ->
//            injection -> estimate sums + Markley mean (types.rs:128-144) -> estimate
//            covariances -> one D2H of every filter's results  [one graph when enabled]
// The resample decision never leaves the device; filters that do not resample skip the
// resampling kernels tile by tile.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/pf_batch_api.h"
#include "pf_batch.h"
#include "pf_device.h"

#define PB_CHECK(x)                                                                              \
    do {                                                                                         \
        hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                  \
            fprintf(stderr, "pf_hip %s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            return -1;                                                                           \
        }                                                                                        \
    } while (0)


namespace {

using namespace pfk;

struct ParamBlock {  // one contiguous pinned / device block per cycle
    MotionParams* motion;
    WeightParams* weight;
    InjectParams* inject;
    CovParams* cov;
};

struct Batch {
    PfState s;
    int nf = 0, obs_cap = 0;
    std::vector<pfb_filter_cfg> cfg;
    std::vector<std::vector<float>> lm;
    std::vector<int> lm_cap;
    std::vector<uint64_t> gseed, gen;
    // device / pinned parameter blocks
    char* d_params = nullptr;
    char* h_params = nullptr;
    size_t params_bytes = 0;
    ParamBlock dp{}, hp{};
    float* d_ratio = nullptr;
    size_t stage_bytes = 0;   // parameter blocks + observation capacity (one allocation per side)
    size_t stage_used = 0;    // parameter blocks + this cycle's joined observations
    ObsRec* d_obs = nullptr;  // inside the d_params allocation, after the parameter blocks
    ObsRec* h_obs = nullptr;  // inside h_params
    hipEvent_t ev_done = nullptr;  // end of a cycle's copy-back (polled)
    FilterDev* h_fdev = nullptr;
    // graphs
    bool use_graph = false;
    bool captured = false;
    hipGraphExec_t gA = nullptr;  // the captured cycle
    // staged cycles (inputs resident on the device before the run)
    int st_cycles = 0;
    bool st_pending = false;
    char* d_staged = nullptr;  // st_cycles staging blocks of stage_bytes each
    char* h_staged = nullptr;
    FilterDev* d_results = nullptr;  // st_cycles x nf filter scalars after each cycle
    FilterDev* h_results = nullptr;
    hipEvent_t st_ev[2] = {nullptr, nullptr};  // device time of the last staged run
    float st_ms = 0.f;
    bool prof_req = false;  // per-stage events requested (recorded in eager cycles and staged runs)
};

ParamBlock carve(char* base, int nf) {
    ParamBlock b;
    size_t o = 0;
    auto take = [&](size_t sz) {
        char* p = base + o;
        o += (sz + 255) & ~size_t(255);
        return p;
    };
    b.motion = reinterpret_cast<MotionParams*>(take(sizeof(MotionParams) * nf));
    b.weight = reinterpret_cast<WeightParams*>(take(sizeof(WeightParams) * nf));
    b.inject = reinterpret_cast<InjectParams*>(take(sizeof(InjectParams) * nf));
    b.cov = reinterpret_cast<CovParams*>(take(sizeof(CovParams) * nf));
    return b;
}

size_t params_size(int nf) {
    auto r = [](size_t sz) { return (sz + 255) & ~size_t(255); };
    return r(sizeof(MotionParams) * nf) + r(sizeof(WeightParams) * nf) + r(sizeof(InjectParams) * nf) +
           r(sizeof(CovParams) * nf);
}

void free_batch(Batch* b) {
    if (b->s.stream) (void)hipStreamSynchronize(b->s.stream);
    if (b->gA) (void)hipGraphExecDestroy(b->gA);
    (void)hipFree(b->d_staged);
    (void)hipHostFree(b->h_staged);
    (void)hipFree(b->d_results);
    (void)hipHostFree(b->h_results);
    state_free_particles(&b->s);
    (void)hipFree(b->d_params);  // parameter blocks and observations share one allocation
    (void)hipHostFree(b->h_params);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    for (hipEvent_t e : b->st_ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(b->d_ratio);
    (void)hipHostFree(b->h_fdev);
    for (auto& p : b->s.prof_pending) {
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    if (b->s.stream) (void)hipStreamDestroy(b->s.stream);
    delete b;
}

// The cycle's packed inputs (parameter blocks, then the joined observations of this cycle) to
// the device: one copy of exactly the bytes in use, ahead of the cycle's graph.
int stage_in(Batch* b) {
    PB_CHECK(hipMemcpyAsync(b->d_params, b->h_params, b->stage_used, hipMemcpyHostToDevice, b->s.stream));
    return 0;
}

// Every device stage of a cycle (one graph): predict, weight, normalise (ESS + resample
// predicate), resample (skipped on the device when the predicate is false), estimate sums with
// the Markley mean, covariances, results out.
int stages_all(Batch* b) {
    PfState* s = &b->s;
    if (launch_motion(s, ParamRef<MotionParams>::array(b->dp.motion), 0) != 0) return -1;
    if (launch_weight(s, ParamRef<WeightParams>::array(b->dp.weight)) != 0) return -1;
    // + the scan of every filter and its injection picks
    if (launch_normalize(s, 0, b->d_ratio, true, ParamRef<InjectParams>::array(b->dp.inject)) != 0) return -1;
    if (launch_resample(s, ParamRef<InjectParams>::array(b->dp.inject), s->d_inj_blk, s->inj_nblk, false, false) !=
        0)
        return -1;
    // the gather left every filter's sum partials; + the resampled set becomes current
    if (launch_est_sums(s, b->dp.cov, true, false) != 0) return -1;
    if (launch_est_covs(s, ParamRef<CovParams>::array(b->dp.cov)) != 0) return -1;
    PB_CHECK(hipMemcpyAsync(b->h_fdev, s->fdev, sizeof(FilterDev) * (size_t)b->nf, hipMemcpyDeviceToHost, s->stream));
    return 0;
}

int capture(Batch* b, int (*fn)(Batch*), hipGraphExec_t* out) {
    hipGraph_t g = nullptr;
    PB_CHECK(hipStreamBeginCapture(b->s.stream, hipStreamCaptureModeThreadLocal));
    const int rc = fn(b);
    hipError_t e = hipStreamEndCapture(b->s.stream, &g);
    if (rc != 0 || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        fprintf(stderr, "pf_hip graph capture failed\n");
        return -1;
    }
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    PB_CHECK(e);
    return 0;
}

int run_cycle(Batch* b) {
    if (!b->use_graph) return stages_all(b);
    if (!b->gA && capture(b, stages_all, &b->gA) != 0) return -1;
    PB_CHECK(hipGraphLaunch(b->gA, b->s.stream));
    return 0;
}

void drop_graphs(Batch* b) {
    if (b->gA) (void)hipGraphExecDestroy(b->gA);
    b->gA = nullptr;
}

// Injection picks per resample of a filter (cu:667-671); fixed by the filter's configuration.
int inject_count(const pfb_filter_cfg& c) {
    const float r = c.inject_ratio;
    if (!(r > 0.f)) return 0;
    const int k = (int)(r * (float)c.n);
    return k > c.n ? c.n : k;
}

// Host half of a cycle before the device stages: the parameter blocks hp (host) and the joined
// observations h_obs (host), whose device copy will sit at d_obs; advances every filter's draw
// generation.  Returns the bytes in use (parameter blocks + observations).
size_t pack_inputs(Batch* b, const pfb_cycle_in* in, const ParamBlock& hp, ObsRec* h_obs, ObsRec* d_obs) {
    int total = 0;
    for (int f = 0; f < b->nf; ++f) {
        const pfb_cycle_in& c = in[f];
        b->gseed[f] += pfd::kGenBump;  // cpu.rs:198
        b->gen[f] += 1;                // cu:588
        MotionParams& m = hp.motion[f];
        memset(&m, 0, sizeof(m));
        for (int k = 0; k < 3; ++k) {
            m.t[k] = c.t[k];
            m.s_t[k] = c.tstd[k];
            m.s_r[k] = c.rstd[k];
        }
        for (int k = 0; k < 4; ++k) m.q[k] = c.q[k];
        m.gseed = b->gseed[f];
        const int nobs = c.nobs > b->obs_cap ? b->obs_cap : c.nobs;
        int mobs = 0, nrb = 0;
        WeightParams wp = make_weight_params(b->cfg[f].sigma_r, b->cfg[f].sigma_b, b->cfg[f].gate);
        if (nobs > 0 && c.obs)
            mobs = join_obs(c.obs, nobs, b->lm[f].data(), b->lm_cap[f], wp, h_obs + total, &nrb);
        wp.obs = d_obs + total;
        wp.nobs = mobs;
        wp.nrb = nrb;
        hp.weight[f] = wp;
        total += mobs;
        InjectParams& ip = hp.inject[f];
        ip.gen = b->gen[f];
        ip.k = inject_count(b->cfg[f]);
        ip.pad = 0;
    }
    return b->params_bytes + sizeof(ObsRec) * (size_t)total;
}

int upload_inputs(Batch* b, const pfb_cycle_in* in) {
    b->stage_used = pack_inputs(b, in, b->hp, b->h_obs, b->d_obs);
    return 0;
}

// Estimate of one filter from its device scalars (Estimate, types.rs:92-103).
void fill_estimate(const FilterDev& fd, pfb_estimate& o) {
    static const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    for (int k = 0; k < 3; ++k) o.pos[k] = fd.sums[k];
    for (int k = 0; k < 4; ++k) o.quat[k] = fd.qmean[k];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            o.pos_cov[3 * r + c] = fd.covs[idx[r][c]];
            o.att_cov[3 * r + c] = fd.covs[6 + idx[r][c]];
        }
    o.ess = fd.sums[13] > 0.0 ? 1.0 / fd.sums[13] : 0.0;
    o.resampled = fd.do_resample;
    o.valid = fd.qvalid;
}

// Device stages of one staged cycle: parameters read in place from its staging block.  The
// cycle's covariances are left pending: nothing in the next cycle needs them, so the next
// cycle's likelihood launch computes them (its predict stores the new poses in the idle pose
// set, leaving this cycle's poses and weights in place until the likelihood has read them) and
// its softmax finalize and normalisation launches reduce them into the results slot.  prev: the previous cycle's
// parameter blocks and results slot (null for the first cycle of a run, whose predecessor
// finished on its own); the run's last cycle finishes with a covariance pass of its own.
int stages_staged(Batch* b, const ParamBlock& dp, const ParamBlock* prev, FilterDev* prev_results) {
    PfState* s = &b->s;
    if (launch_motion(s, ParamRef<MotionParams>::array(dp.motion), 0, 1) != 0) return -1;
    WeightDefer d;
    d.flip = 1;
    if (prev) {
        d.cov = prev->cov;
        d.out = prev_results;
    }
    if (launch_weight(s, ParamRef<WeightParams>::array(dp.weight), d) != 0) return -1;
    if (launch_normalize(s, 0, b->d_ratio, true, ParamRef<InjectParams>::array(dp.inject), prev ? prev_results : nullptr) !=
        0)
        return -1;
    if (launch_resample(s, ParamRef<InjectParams>::array(dp.inject), s->d_inj_blk, s->inj_nblk, false, false) != 0)
        return -1;
    return launch_est_sums(s, dp.cov, true, false);
}

void free_staged(Batch* b) {
    (void)hipFree(b->d_staged);
    (void)hipHostFree(b->h_staged);
    (void)hipFree(b->d_results);
    (void)hipHostFree(b->h_results);
    b->d_staged = b->h_staged = nullptr;
    b->d_results = b->h_results = nullptr;
    b->st_cycles = 0;
    b->st_pending = false;
}

}  // namespace


extern "C" {

void* pfb_create(int nf, const pfb_filter_cfg* cfg, int obs_cap) {
    if (nf <= 0 || !cfg) return nullptr;
    for (int f = 0; f < nf; ++f)
        if (cfg[f].n <= 0) return nullptr;
    Batch* b = new (std::nothrow) Batch();
    if (!b) return nullptr;
    if (obs_cap < 1 || obs_cap > 4096) obs_cap = 64;
    b->nf = nf;
    b->obs_cap = obs_cap;
    b->cfg.assign(cfg, cfg + nf);
    b->lm.resize(nf);
    b->lm_cap.assign(nf, 0);
    b->gseed.resize(nf);
    b->gen.assign(nf, 0);
    std::vector<int> ns(nf), kinj(nf);
    std::vector<float> ratio(nf);
    for (int f = 0; f < nf; ++f) {
        ns[f] = cfg[f].n;
        kinj[f] = inject_count(cfg[f]);
        b->gseed[f] = cfg[f].seed ? (uint64_t)cfg[f].seed : 42ull;
        ratio[f] = cfg[f].resample_ess_ratio;
    }
    b->params_bytes = params_size(nf);
    bool ok = hipSetDevice(0) == hipSuccess &&
              hipStreamCreateWithFlags(&b->s.stream, hipStreamNonBlocking) == hipSuccess;
    ok = ok && state_alloc(&b->s, ns) == 0;
    ok = ok && set_inject_blocks(&b->s, kinj) == 0;
    // one staging allocation per side: parameter blocks, then the joined observations, so a
    // cycle's inputs go over in one copy
    b->stage_bytes = b->params_bytes + sizeof(ObsRec) * (size_t)nf * obs_cap;
    ok = ok && hipMalloc((void**)&b->d_params, b->stage_bytes) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&b->h_params, b->stage_bytes, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipMalloc((void**)&b->d_ratio, sizeof(float) * (size_t)nf) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming) == hipSuccess;
    if (ok) {
        b->d_obs = reinterpret_cast<ObsRec*>(b->d_params + b->params_bytes);
        b->h_obs = reinterpret_cast<ObsRec*>(b->h_params + b->params_bytes);
    }
    ok = ok && hipHostMalloc((void**)&b->h_fdev, sizeof(FilterDev) * (size_t)nf, hipHostMallocDefault) == hipSuccess;
    ok = ok && hipMemcpy(b->d_ratio, ratio.data(), sizeof(float) * (size_t)nf, hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        fprintf(stderr, "pf_hip pfb_create: device setup failed\n");
        free_batch(b);
        return nullptr;
    }
    memset(b->h_params, 0, b->params_bytes);
    b->dp = carve(b->d_params, nf);
    b->hp = carve(b->h_params, nf);
    return b;
}

int pfb_destroy(void* h) {
    if (h) free_batch(static_cast<Batch*>(h));
    return 0;
}

int pfb_filter_count(void* h) { return h ? static_cast<Batch*>(h)->nf : -1; }

int pfb_upload_landmarks(void* h, int f, const float* xyz, int count) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || f < 0 || f >= b->nf || count < 0 || (count > 0 && !xyz)) return -1;
    b->lm[f].assign(xyz, xyz + 3 * (size_t)count);
    b->lm_cap[f] = count;
    return 0;
}

int pfb_reinit(void* h, const float* pose7, const float* pos_std, const float* rot_std) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !pose7 || !pos_std || !rot_std) return -1;
    for (int f = 0; f < b->nf; ++f) {
        b->gseed[f] += pfd::kGenBump;  // cpu.rs:181
        b->gen[f] += 1;
        MotionParams& m = b->hp.motion[f];
        memset(&m, 0, sizeof(m));
        for (int k = 0; k < 3; ++k) {
            m.t[k] = pose7[7 * f + k];
            m.s_t[k] = pos_std[3 * f + k];
            m.s_r[k] = rot_std[3 * f + k];
        }
        for (int k = 0; k < 4; ++k) m.q[k] = pose7[7 * f + 3 + k];
        m.gseed = b->gseed[f];
        m.winv = (float)(1.0 / (double)b->cfg[f].n);
    }
    PB_CHECK(hipMemcpyAsync(b->d_params, b->h_params, b->params_bytes, hipMemcpyHostToDevice, b->s.stream));
    if (launch_motion(&b->s, ParamRef<MotionParams>::array(b->dp.motion), 1) != 0) return -1;
    PB_CHECK(hipStreamSynchronize(b->s.stream));
    return 0;
}

int pfb_cycle(void* h, const pfb_cycle_in* in, pfb_estimate* out) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !in) return -1;
    PfState* s = &b->s;
    if (upload_inputs(b, in) != 0 || stage_in(b) != 0) return -1;
    if (run_cycle(b) != 0) return -1;
    PB_CHECK(hipEventRecord(b->ev_done, s->stream));
    // poll rather than block: the cycle is short and a sleeping wait adds wake-up latency
    hipError_t q;
    while ((q = hipEventQuery(b->ev_done)) == hipErrorNotReady) {
    }
    PB_CHECK(q);
    if (out)
        for (int f = 0; f < b->nf; ++f) fill_estimate(b->h_fdev[f], out[f]);
    return 0;
}

int pfb_stage_cycles(void* h, int ncycles, const pfb_cycle_in* in) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !in || ncycles <= 0) return -1;
    PB_CHECK(hipStreamSynchronize(b->s.stream));
    free_staged(b);
    const size_t sb = b->stage_bytes, nfd = (size_t)ncycles * b->nf;
    if (hipMalloc((void**)&b->d_staged, sb * ncycles) != hipSuccess ||
        hipHostMalloc((void**)&b->h_staged, sb * ncycles, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&b->d_results, sizeof(FilterDev) * nfd) != hipSuccess ||
        hipHostMalloc((void**)&b->h_results, sizeof(FilterDev) * nfd, hipHostMallocDefault) != hipSuccess) {
        fprintf(stderr, "pf_hip pfb_stage_cycles: allocation of %d cycles failed\n", ncycles);
        free_staged(b);
        return -1;
    }
    memset(b->h_staged, 0, sb * ncycles);
    for (int k = 0; k < ncycles; ++k) {
        char* hb = b->h_staged + (size_t)k * sb;
        ObsRec* h_obs = reinterpret_cast<ObsRec*>(hb + b->params_bytes);
        ObsRec* d_obs = reinterpret_cast<ObsRec*>(b->d_staged + (size_t)k * sb + b->params_bytes);
        (void)pack_inputs(b, in + (size_t)k * b->nf, carve(hb, b->nf), h_obs, d_obs);
    }
    PB_CHECK(hipMemcpy(b->d_staged, b->h_staged, sb * ncycles, hipMemcpyHostToDevice));
    b->st_cycles = ncycles;
    b->st_pending = true;
    return 0;
}

int pfb_run_staged(void* h, pfb_estimate* out) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !b->st_pending) return -1;
    b->st_pending = false;
    PfState* s = &b->s;
    const bool prof_saved = s->prof_on;
    s->prof_on = b->prof_req;
    // Stream launches, one cycle after another with no synchronisation: the host enqueues a
    // cycle far faster than the device runs it, so the device never waits.  (One graph holding
    // every cycle stalled for milliseconds between groups of cycles when replayed.)
    for (hipEvent_t& e : b->st_ev)
        if (!e) PB_CHECK(hipEventCreate(&e));
    PB_CHECK(hipEventRecord(b->st_ev[0], s->stream));
    ParamBlock prev{};
    for (int k = 0; k < b->st_cycles; ++k) {
        const ParamBlock dp = carve(b->d_staged + (size_t)k * b->stage_bytes, b->nf);
        FilterDev* prev_out = k > 0 ? b->d_results + (size_t)(k - 1) * b->nf : nullptr;
        if (stages_staged(b, dp, k > 0 ? &prev : nullptr, prev_out) != 0) return -1;
        prev = dp;
    }
    if (launch_est_covs(s, ParamRef<CovParams>::array(prev.cov), b->d_results + (size_t)(b->st_cycles - 1) * b->nf) !=
        0)
        return -1;
    PB_CHECK(hipEventRecord(b->st_ev[1], s->stream));
    s->prof_on = prof_saved;
    const size_t nfd = (size_t)b->st_cycles * b->nf;
    PB_CHECK(hipMemcpyAsync(b->h_results, b->d_results, sizeof(FilterDev) * nfd, hipMemcpyDeviceToHost, s->stream));
    PB_CHECK(hipStreamSynchronize(s->stream));
    PB_CHECK(hipEventElapsedTime(&b->st_ms, b->st_ev[0], b->st_ev[1]));
    if (out)
        for (size_t i = 0; i < nfd; ++i) fill_estimate(b->h_results[i], out[i]);
    return 0;
}

int pfb_staged_device_ms(void* h, float* ms) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !ms) return -1;
    *ms = b->st_ms;
    return 0;
}

int pfb_use_graph(void* h, int enable) {
    Batch* b = static_cast<Batch*>(h);
    if (!b) return -1;
    b->use_graph = enable != 0;
    if (!b->use_graph) drop_graphs(b);
    b->s.prof_on = b->prof_req && !b->use_graph;  // events are not recorded inside captured graphs
    return 0;
}

int pfb_autotune(void* h, const pfb_cycle_in* in) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || !in) return -1;
    PfState* s = &b->s;
    const bool prof = s->prof_on;
    s->prof_on = false;
    // parameters of a cycle (without advancing the filters' draw state afterwards)
    std::vector<uint64_t> gs = b->gseed, gn = b->gen;
    if (upload_inputs(b, in) != 0 || stage_in(b) != 0) return -1;
    b->gseed = gs;
    b->gen = gn;
    hipEvent_t e0, e1;
    PB_CHECK(hipEventCreate(&e0));
    PB_CHECK(hipEventCreate(&e1));
    const int cand[3] = {1, 2, 4};  // particles per lane of the likelihood kernel
    float best = 1e30f;
    int best_ppt = s->weight_ppt;
    for (int c = 0; c < 3; ++c) {
        s->weight_ppt = cand[c];
        if (launch_weight(s, ParamRef<WeightParams>::array(b->dp.weight)) != 0) return -1;  // warm-up
        PB_CHECK(hipEventRecord(e0, s->stream));
        for (int r = 0; r < 3; ++r)
            if (launch_weight(s, ParamRef<WeightParams>::array(b->dp.weight)) != 0) return -1;
        PB_CHECK(hipEventRecord(e1, s->stream));
        PB_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        PB_CHECK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            best_ppt = cand[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    s->weight_ppt = best_ppt;
    s->prof_on = prof;
    drop_graphs(b);
    return best_ppt;
}

int pfb_set_weight_ppt(void* h, int ppt) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || (ppt != 1 && ppt != 2 && ppt != 4)) return -1;
    PB_CHECK(hipStreamSynchronize(b->s.stream));
    b->s.weight_ppt = ppt;
    drop_graphs(b);  // (a captured cycle holds the launch shape)
    return 0;
}

int pfb_get_state(void* h, int f, float* pos, float* quat, float* w) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || f < 0 || f >= b->nf) return -1;
    PfState* s = &b->s;
    PB_CHECK(hipStreamSynchronize(s->stream));
    int par = 0;
    PB_CHECK(hipMemcpy(&par, fdev_field<int>(s, f, offsetof(FilterDev, parity)), sizeof(int), hipMemcpyDeviceToHost));
    const int n = s->n[f], off = s->off[f];
    std::vector<float> tmp((size_t)n);
    for (int c = 0; c < 7; ++c) {
        if ((c < 3 && !pos) || (c >= 3 && !quat)) continue;
        PB_CHECK(hipMemcpy(tmp.data(), s->pose.s[par].c[c] + off, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
        if (c < 3)
            for (int i = 0; i < n; ++i) pos[3 * i + c] = tmp[i];
        else
            for (int i = 0; i < n; ++i) quat[4 * i + (c - 3)] = tmp[i];
    }
    if (w) PB_CHECK(hipMemcpy(w, s->w + off, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost));
    return 0;
}

int pfb_get_ancestors(void* h, int f, int32_t* anc) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || f < 0 || f >= b->nf || !anc) return -1;
    PB_CHECK(hipStreamSynchronize(b->s.stream));
    PB_CHECK(hipMemcpy(anc, b->s.anc + b->s.off[f], sizeof(int32_t) * (size_t)b->s.n[f], hipMemcpyDeviceToHost));
    return 0;
}

int pfb_get_cum(void* h, int f, uint64_t* cum) {
    Batch* b = static_cast<Batch*>(h);
    if (!b || f < 0 || f >= b->nf || !cum) return -1;
    PB_CHECK(hipStreamSynchronize(b->s.stream));
    PB_CHECK(hipMemcpy(cum, b->s.cum + b->s.off[f], sizeof(uint64_t) * (size_t)b->s.n[f], hipMemcpyDeviceToHost));
    return 0;
}

int pfb_profile_enable(void* h, int enable) {
    Batch* b = static_cast<Batch*>(h);
    if (!b) return -1;
    b->prof_req = enable != 0;
    b->s.prof_on = b->prof_req && !b->use_graph;
    return 0;
}

int pfb_profile_read(void* h, double* ms, int* count) {
    Batch* b = static_cast<Batch*>(h);
    if (!b) return -1;
    PfState* s = &b->s;
    PB_CHECK(hipStreamSynchronize(s->stream));
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

}  // extern "C"

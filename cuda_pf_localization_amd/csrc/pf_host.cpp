// pf_host.cpp -- C++ host layer over the pfc_* / pfgl_* ABI (the role the reference's Rust
// crate pf_core plays above libpf_kernels.so).
//
//   HipBackend      crates/pf_core/src/cuda.rs:97-357 (f64->f32 casts, map.has filtering,
//                   OBS_CAP truncation, landmark upload only on change, host Markley mean,
//                   f32 round trip of the mean position, packed-triangle unpacking)
//   ParticleFilter  crates/pf_core/src/particle_filter.rs:72-133 (update cycle order)
//   quat_mean       crates/pf_core/src/types.rs:128-144 (largest eigenvector of sym(M);
//                   cyclic Jacobi here, nalgebra is not available)
//   GlobalLoc       crates/pf_core/src/global_loc.rs:274-332 (generate -> score -> topk)
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <array>
#include <optional>
#include <vector>
#include <sstream>
#include <string>
#include <cctype>
#include <climits>
#include <algorithm>

#include "pf_gl.h"
#include "pf_host.h"
#include "pf_kernels.h"
#include "pf_learned.h"

namespace pfh {

// ------------------------------------------------------------------ Markley mean

// Markley mean (types.rs:128-144): the solver lives in libpf_kernels_hip.so
// (pfx_quat_mean) so the single-filter and batched paths share one implementation.
int quat_mean(const double* M, double* q) { return pfx_quat_mean(M, q); }

// ------------------------------------------------------------------ landmark map

class LandmarkMap {
   public:
    void add(int id, double x, double y, double z) {
        if (id < 0) return;
        if ((int)pos_.size() <= id) pos_.resize((size_t)id + 1);
        pos_[(size_t)id] = std::array<double, 3>{x, y, z};
    }
    bool has(int id) const { return id >= 0 && id < (int)pos_.size() && pos_[(size_t)id].has_value(); }
    int capacity() const { return (int)pos_.size(); }
    const std::vector<std::optional<std::array<double, 3>>>& entries() const { return pos_; }

   private:
    std::vector<std::optional<std::array<double, 3>>> pos_;
};

// ------------------------------------------------------------------ backend

struct Estimate {
    double pos[3]{0, 0, 0}, quat[4]{1, 0, 0, 0};
    double pos_cov[3][3]{{1, 1, 1}, {1, 1, 1}, {1, 1, 1}}, att_cov[3][3]{{1, 1, 1}, {1, 1, 1}, {1, 1, 1}};
    double ess = 0.0;
    bool valid = false;  // Estimate::default (types.rs:105-115)
};

class HipBackend {
   public:
    explicit HipBackend(int obs_cap) : obs_cap_(obs_cap) {}
    ~HipBackend() {
        if (h_) pfc_destroy(h_);
    }
    bool create() {
        if (pfc_device_available() != 1) return false;  // cuda.rs:113-115
        h_ = pfc_create(obs_cap_);
        return h_ != nullptr;
    }
    bool init(int64_t n, const pfh_config& cfg) {  // cuda.rs:180-185
        n_ = n;
        cfg_ = cfg;
        return pfc_init(h_, (int)n, (unsigned)cfg.seed) == 0;
    }
    int reinit(const double* pose7, const double* pos_std, const double* rot_std) {  // cuda.rs:187-205
        const float ps[3] = {(float)pos_std[0], (float)pos_std[1], (float)pos_std[2]};
        const float rs[3] = {(float)rot_std[0], (float)rot_std[1], (float)rot_std[2]};
        return pfc_reinit(h_, (float)pose7[0], (float)pose7[1], (float)pose7[2], (float)pose7[3], (float)pose7[4],
                          (float)pose7[5], (float)pose7[6], ps, rs);
    }
    int predict(const pfh_odom& o) {  // cuda.rs:207-225
        const float t[3] = {(float)o.trans_noise[0], (float)o.trans_noise[1], (float)o.trans_noise[2]};
        const float r[3] = {(float)o.rot_noise[0], (float)o.rot_noise[1], (float)o.rot_noise[2]};
        return pfc_predict(h_, (float)o.t[0], (float)o.t[1], (float)o.t[2], (float)o.q[0], (float)o.q[1],
                           (float)o.q[2], (float)o.q[3], t, r);
    }
    int weight(const pfh_obs* obs, int nobs, const LandmarkMap& map) {  // cuda.rs:227-256
        const float sr = (float)cfg_.range_noise, sb = (float)cfg_.bearing_noise, g = (float)cfg_.gate_sigma;
        cobs_.clear();
        for (int k = 0; k < nobs; ++k) {
            if (!map.has(obs[k].id)) continue;
            if ((int)cobs_.size() >= obs_cap_) break;  // take(OBS_CAP), cuda.rs:159
            pfc_obs c;
            c.id = obs[k].id;
            c.mode = obs[k].mode == 0 ? 0 : 1;
            c.dx = (float)obs[k].vx;
            c.dy = (float)obs[k].vy;
            c.dz = (float)obs[k].vz;
            c.range = (float)obs[k].range;
            cobs_.push_back(c);
        }
        if (cobs_.empty()) return pfc_weight(h_, nullptr, 0, sr, sb, g);
        if (sync_landmarks(map) != 0) return -1;
        return pfc_weight(h_, cobs_.data(), (int)cobs_.size(), sr, sb, g);
    }
    int resample(double ratio) { return pfc_resample(h_, (float)ratio); }  // cuda.rs:258-262

    // pass 1 only (sums; ESS)
    int sums(double* s14) { return pfc_est_sums(h_, s14); }

    // cuda.rs:264-329, optionally reusing the sums of a preceding pass-1 call
    int estimate(Estimate& e, const double* sums_in = nullptr) {
        if (n_ == 0) {
            e = Estimate();
            return 0;
        }
        double s[14];
        if (sums_in)
            memcpy(s, sums_in, sizeof(s));
        else if (pfc_est_sums(h_, s) != 0)
            return -1;
        double M[16];
        for (int r = 0; r < 4; ++r)
            for (int c = r; c < 4; ++c) {
                const int off = 4 * r - r * (r - 1) / 2 + (c - r);
                M[4 * r + c] = s[3 + off];
                M[4 * c + r] = s[3 + off];
            }
        double q[4];
        if (quat_mean(M, q) != 0) return -1;
        double cv[12];
        if (pfc_covs(h_, (float)s[0], (float)s[1], (float)s[2], (float)q[0], (float)q[1], (float)q[2], (float)q[3],
                     cv) != 0)
            return -1;
        auto sym3 = [&](int o, double (&m)[3][3]) {
            for (int r = 0; r < 3; ++r)
                for (int c = r; c < 3; ++c) {
                    const double x = cv[o + 3 * r - r * (r - 1) / 2 + (c - r)];
                    m[r][c] = x;
                    m[c][r] = x;
                }
        };
        for (int k = 0; k < 3; ++k) e.pos[k] = s[k];
        for (int k = 0; k < 4; ++k) e.quat[k] = q[k];
        sym3(0, e.pos_cov);
        sym3(6, e.att_cov);
        e.ess = s[13] > 0.0 ? 1.0 / s[13] : 0.0;
        e.valid = true;
        return 0;
    }
    int snapshot(double* pos, double* quat, double* w) {  // cuda.rs:331-357
        std::vector<float> p((size_t)n_ * 3), q((size_t)n_ * 4);
        if (pfc_snapshot(h_, p.data(), q.data()) != 0) return -1;
        for (int64_t i = 0; i < n_; ++i) {
            for (int c = 0; c < 3; ++c) pos[3 * i + c] = p[3 * i + c];
            double qq[4], nrm = 0.0;
            for (int c = 0; c < 4; ++c) {
                qq[c] = q[4 * i + c];
                nrm += qq[c] * qq[c];
            }
            nrm = std::sqrt(nrm);  // UnitQuaternion::from_quaternion normalises
            for (int c = 0; c < 4; ++c) quat[4 * i + c] = nrm > 0.0 ? qq[c] / nrm : qq[c];
            if (w) w[i] = 1.0 / (double)n_;
        }
        return 0;
    }
    void* handle() const { return h_; }
    int64_t n() const { return n_; }

   private:
    int sync_landmarks(const LandmarkMap& map) {  // cuda.rs:140-154
        const int cap = map.capacity() > 1 ? map.capacity() : 1;
        std::vector<float> flat((size_t)cap * 3, 0.f);
        const auto& ent = map.entries();
        for (size_t id = 0; id < ent.size(); ++id)
            if (ent[id])
                for (int c = 0; c < 3; ++c) flat[3 * id + c] = (float)(*ent[id])[c];
        if (flat != lm_cache_) {
            if (pfc_upload_landmarks(h_, flat.data(), cap) != 0) return -1;
            lm_cache_.swap(flat);
        }
        return 0;
    }
    void* h_ = nullptr;
    int obs_cap_;
    int64_t n_ = 0;
    pfh_config cfg_{};
    std::vector<float> lm_cache_;
    std::vector<pfc_obs> cobs_;
};

class ParticleFilter {
   public:
    explicit ParticleFilter(const pfh_config& cfg) : cfg_(cfg), be_(cfg.obs_cap > 0 ? cfg.obs_cap : 256) {}
    bool create() { return be_.create() && be_.init(cfg_.particle_count, cfg_); }
    LandmarkMap& map() { return map_; }
    int reinitialize(const double* pose7, const double* ps, const double* rs) {  // pf.rs:93-96
        last_ = Estimate();
        return be_.reinit(pose7, ps, rs);
    }
    // pf.rs:110-124: predict -> weight -> estimate -> [resample -> estimate]
    int update(const pfh_odom& odom, const pfh_obs* obs, int nobs, Estimate& est, bool& resampled) {
        resampled = false;
        if (be_.predict(odom) != 0) return -1;
        if (be_.weight(obs, nobs, map_) != 0) return -1;
        const double thr = cfg_.resample_ess_ratio * (double)cfg_.particle_count;
        if (cfg_.fast) {
            // The first estimate's mean and covariances are discarded when the filter
            // resamples; only its ESS is read, so compute pass 1 alone first.
            double s[14];
            if (be_.sums(s) != 0) return -1;
            const double ess = s[13] > 0.0 ? 1.0 / s[13] : 0.0;
            if (cfg_.resample_ess_ratio > 0.0 && ess < thr) {
                if (be_.resample(cfg_.random_inject_ratio) != 0) return -1;
                resampled = true;
                if (be_.estimate(est) != 0) return -1;
            } else if (be_.estimate(est, s) != 0) {
                return -1;
            }
        } else {
            if (be_.estimate(est) != 0) return -1;
            if (cfg_.resample_ess_ratio > 0.0 && est.ess < thr) {
                if (be_.resample(cfg_.random_inject_ratio) != 0) return -1;
                resampled = true;
                if (be_.estimate(est) != 0) return -1;
            }
        }
        last_ = est;
        return 0;
    }
    int estimate(Estimate& e) { return be_.estimate(e); }
    HipBackend& backend() { return be_; }

   private:
    pfh_config cfg_;
    HipBackend be_;
    LandmarkMap map_;
    Estimate last_;
};

// ------------------------------------------------------------------ global localisation

class GlobalLoc {
   public:
    ~GlobalLoc() {
        if (h_) pfgl_destroy(h_);
    }
    bool create(int W, int H, double res, double ox, double oy, const uint8_t* cells, const pfh_gl_params& p) {
        p_ = p;
        h_ = pfgl_create(W, H, (float)res, (float)ox, (float)oy, cells, p.hypotheses);  // global_loc.rs:287-303
        return h_ != nullptr;
    }
    // global_loc.rs:306-327 (the CAER of each winner is returned as well)
    int localize(const float* scan, int nrays, double* out4k) {
        if (pfgl_generate(h_, p_.hypotheses, (unsigned long long)p_.seed) != 0) return -1;
        if (pfgl_score(h_, scan, nrays, (float)p_.rmax, (float)p_.ang_span) != 0) return -1;
        std::vector<float> xyz((size_t)p_.topk * 3), caer((size_t)p_.topk);
        const int k = pfglx_topk_ex(h_, p_.topk, xyz.data(), caer.data(), nullptr);
        if (k <= 0) return -1;
        for (int i = 0; i < k; ++i) {
            out4k[4 * i] = xyz[3 * i];
            out4k[4 * i + 1] = xyz[3 * i + 1];
            out4k[4 * i + 2] = xyz[3 * i + 2];
            out4k[4 * i + 3] = caer[i];
        }
        return k;
    }
    // Learned proposal first, the reference's uniform dispersal as fallback.
    int localize_learned(void* mlp, const float* scan, int nrays, const pfh_learned_params& lp, double* out4k,
                         pfh_learned_info* info) {
        const float rmax = (float)p_.rmax, span = (float)p_.ang_span;
        // device-resident chain: one scan upload, no waits until the final top-k
        if (pfglx_set_scan(h_, scan, nrays, span) != 0) return -1;
        if (pfl_propose(mlp, h_, scan, nrays, rmax, lp.n_mix, lp.n_uniform, lp.seed) != 0) return -1;
        if (pfglx_score_async(h_, nrays, rmax) != 0) return -1;
        double sxy = lp.refine_sigma_xy, syaw = lp.refine_sigma_yaw;
        for (int r = 0; r < lp.refine_rounds; ++r) {
            const int kp = pfglx_topk_device(h_, lp.refine_parents);
            if (kp <= 0) return -1;
            if (pfglx_refine(h_, kp, lp.refine_children, (float)sxy, (float)syaw,
                             lp.seed * 0x9E3779B97F4A7C15ull + (uint64_t)r + 1) != 0)
                return -1;
            if (pfglx_score_async(h_, nrays, rmax) != 0) return -1;
            sxy *= lp.refine_shrink;
            syaw *= lp.refine_shrink;
        }
        std::vector<float> xyz((size_t)p_.topk * 3), caer((size_t)p_.topk);
        int k = pfglx_topk_ex(h_, p_.topk, xyz.data(), caer.data(), nullptr);
        if (k <= 0) return -1;
        const double first = (double)caer[0] / nrays;
        int fell_back = 0;
        if (first > lp.fallback_caer_per_ray) {
            fell_back = 1;
            if (pfgl_generate(h_, lp.fallback_hypotheses, (unsigned long long)(lp.seed ^ 0xD1B54A32D192ED03ull)) != 0)
                return -1;
            if (pfgl_score(h_, scan, nrays, rmax, span) != 0) return -1;
            k = pfglx_topk_ex(h_, p_.topk, xyz.data(), caer.data(), nullptr);
            if (k <= 0) return -1;
        }
        for (int i = 0; i < k; ++i) {
            out4k[4 * i] = xyz[3 * i];
            out4k[4 * i + 1] = xyz[3 * i + 1];
            out4k[4 * i + 2] = xyz[3 * i + 2];
            out4k[4 * i + 3] = caer[i];
        }
        if (info) {
            info->used_fallback = fell_back;
            info->proposal_best_caer_per_ray = first;
            info->best_caer_per_ray = (double)caer[0] / nrays;
        }
        return k;
    }
    void set_seed(uint64_t s) { p_.seed = s; }
    void* handle() const { return h_; }

   private:
    void* h_ = nullptr;
    pfh_gl_params p_{};
};

void to_c(const Estimate& e, bool resampled, pfh_estimate* out) {
    for (int k = 0; k < 3; ++k) out->pos[k] = e.pos[k];
    for (int k = 0; k < 4; ++k) out->quat[k] = e.quat[k];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            out->pos_cov[3 * r + c] = e.pos_cov[r][c];
            out->att_cov[3 * r + c] = e.att_cov[r][c];
        }
    out->ess = e.ess;
    out->valid = e.valid ? 1 : 0;
    out->resampled = resampled ? 1 : 0;
}

}  // namespace pfh

using namespace pfh;

extern "C" {

void* pfh_filter_new(const pfh_config* cfg) {
    // The device filter indexes particles with int (pfc_init takes int n, cu:534) and its
    // resampling merge path 2n items: at most 2^30 particles.
    if (!cfg || cfg->particle_count <= 0 || cfg->particle_count > (int64_t)1 << 30) return nullptr;
    ParticleFilter* f = new (std::nothrow) ParticleFilter(*cfg);
    if (!f) return nullptr;
    if (!f->create()) {
        delete f;
        return nullptr;
    }
    return f;
}

void pfh_filter_free(void* f) { delete static_cast<ParticleFilter*>(f); }

int pfh_filter_set_map(void* fp, const int32_t* ids, const double* xyz, int n) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    if (!f || (n > 0 && (!ids || !xyz))) return -1;
    f->map() = LandmarkMap();
    for (int i = 0; i < n; ++i) {
        if (ids[i] < 0) return -1;  // landmark_map.rs:19
        f->map().add(ids[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    }
    return 0;
}

int pfh_filter_reinitialize(void* fp, const double pose7[7], const double pos_std[3], const double rot_std[3]) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    if (!f || !pose7 || !pos_std || !rot_std) return -1;
    return f->reinitialize(pose7, pos_std, rot_std);
}

int pfh_filter_update(void* fp, const pfh_odom* odom, const pfh_obs* obs, int nobs, pfh_estimate* out) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    if (!f || !odom || (nobs > 0 && !obs)) return -1;
    Estimate e;
    bool rs = false;
    if (f->update(*odom, obs, nobs, e, rs) != 0) return -1;
    if (out) to_c(e, rs, out);
    return 0;
}

int pfh_filter_estimate(void* fp, pfh_estimate* out) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    if (!f || !out) return -1;
    Estimate e;
    if (f->estimate(e) != 0) return -1;
    to_c(e, false, out);
    return 0;
}

int pfh_filter_snapshot(void* fp, double* pos3n, double* quat4n, double* w) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    if (!f || !pos3n || !quat4n) return -1;
    return f->backend().snapshot(pos3n, quat4n, w);
}

void* pfh_filter_kernels(void* fp) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    return f ? f->backend().handle() : nullptr;
}

int pfh_filter_reinit_from_gl(void* fp, const double* xyt, double z, const double* pos_std, const double* rot_std) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    if (!f || !xyt || !pos_std || !rot_std) return -1;
    const double half = 0.5 * xyt[2];
    const double pose7[7] = {xyt[0], xyt[1], z, std::cos(half), 0.0, 0.0, std::sin(half)};
    return f->reinitialize(pose7, pos_std, rot_std);
}

namespace {
constexpr char kCkptMagic[8] = {'P', 'F', 'H', 'C', 'K', 'P', 'T', '1'};
}

int pfh_filter_save(void* fp, const char* path) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    if (!f || !path) return -1;
    void* h = f->backend().handle();
    const int64_t n = f->backend().n();
    std::vector<float> pos((size_t)n * 3), quat((size_t)n * 4), w((size_t)n);
    uint64_t gseed = 0, gen = 0, lcg = 0;
    if (pfx_get_state(h, pos.data(), quat.data(), w.data()) != 0 || pfx_get_rng(h, &gseed, &gen) != 0 ||
        pfx_get_lcg(h, &lcg) != 0)
        return -1;
    FILE* fo = fopen(path, "wb");
    if (!fo) {
        fprintf(stderr, "pf_host: cannot write %s\n", path);
        return -1;
    }
    bool ok = fwrite(kCkptMagic, 1, 8, fo) == 8 && fwrite(&n, sizeof(n), 1, fo) == 1 &&
              fwrite(&gseed, 8, 1, fo) == 1 && fwrite(&gen, 8, 1, fo) == 1 && fwrite(&lcg, 8, 1, fo) == 1;
    ok = ok && fwrite(pos.data(), 4, pos.size(), fo) == pos.size() &&
         fwrite(quat.data(), 4, quat.size(), fo) == quat.size() && fwrite(w.data(), 4, w.size(), fo) == w.size();
    ok = (fclose(fo) == 0) && ok;
    return ok ? 0 : -1;
}

int pfh_filter_load(void* fp, const char* path) {
    ParticleFilter* f = static_cast<ParticleFilter*>(fp);
    if (!f || !path) return -1;
    FILE* fi = fopen(path, "rb");
    if (!fi) {
        fprintf(stderr, "pf_host: cannot read %s\n", path);
        return -1;
    }
    char magic[8];
    int64_t n = 0;
    uint64_t gseed = 0, gen = 0, lcg = 0;
    bool ok = fread(magic, 1, 8, fi) == 8 && memcmp(magic, kCkptMagic, 8) == 0 && fread(&n, sizeof(n), 1, fi) == 1 &&
              fread(&gseed, 8, 1, fi) == 1 && fread(&gen, 8, 1, fi) == 1 && fread(&lcg, 8, 1, fi) == 1;
    if (!ok || n != f->backend().n()) {
        fclose(fi);
        fprintf(stderr, "pf_host: %s is not a checkpoint of a %lld-particle filter\n", path,
                (long long)f->backend().n());
        return -1;
    }
    std::vector<float> pos((size_t)n * 3), quat((size_t)n * 4), w((size_t)n);
    ok = fread(pos.data(), 4, pos.size(), fi) == pos.size() && fread(quat.data(), 4, quat.size(), fi) == quat.size() &&
         fread(w.data(), 4, w.size(), fi) == w.size();
    fclose(fi);
    if (!ok) return -1;
    void* h = f->backend().handle();
    if (pfx_set_state(h, pos.data(), quat.data(), w.data()) != 0 || pfx_set_rng(h, gseed, gen) != 0 ||
        pfx_set_lcg(h, lcg) != 0)
        return -1;
    return 0;
}

namespace {
void set_err(char* err, int len, const std::string& msg) {
    if (err && len > 0) {
        const size_t k = std::min((size_t)len - 1, msg.size());
        memcpy(err, msg.data(), k);
        err[k] = 0;
    }
}

// Last value of `key: <number>` in the YAML text (comments stripped), global_loc.rs:35-45.
bool yaml_number(const std::string& yaml, const std::string& key, double* out) {
    bool hit = false;
    std::istringstream in(yaml);
    std::string line;
    while (std::getline(in, line)) {
        std::string t = line.substr(0, line.find('#'));
        const size_t b = t.find_first_not_of(" \t\r");
        if (b == std::string::npos) continue;
        t = t.substr(b);
        t.erase(t.find_last_not_of(" \t\r") + 1);
        if (t.compare(0, key.size(), key) != 0) continue;
        std::string rest = t.substr(key.size());
        const size_t c = rest.find_first_not_of(" \t");
        if (c == std::string::npos || rest[c] != ':') continue;
        rest = rest.substr(c + 1);
        const size_t v0 = rest.find_first_not_of(" \t");
        if (v0 == std::string::npos) continue;
        rest = rest.substr(v0);
        char* end = nullptr;
        const double v = strtod(rest.c_str(), &end);
        if (end && *end == 0 && end != rest.c_str()) {
            *out = v;
            hit = true;
        } else {
            hit = false;  // the reference keeps the last matching line and fails its parse
        }
    }
    return hit;
}
}  // namespace

int pfh_map_from_pgm_yaml(const uint8_t* pgm, size_t pgm_len, const char* yaml_c, int* W, int* H, double* res,
                          double* ox, double* oy, uint8_t* cells, size_t cells_cap, char* err, int err_len) {
    if (!pgm || !yaml_c || !W || !H || !res || !ox || !oy) return -1;
    const std::string yaml(yaml_c);
    if (!yaml_number(yaml, "resolution", res)) {
        set_err(err, err_len, "yaml: no resolution");
        return -1;
    }
    // origin: first line starting with "origin", values after '['
    {
        std::istringstream in(yaml);
        std::string line;
        bool found = false;
        while (std::getline(in, line)) {
            const size_t b = line.find_first_not_of(" \t");
            if (b == std::string::npos || line.compare(b, 6, "origin") != 0) continue;
            const size_t br = line.find('[');
            if (br == std::string::npos) break;
            std::string body = line.substr(br + 1);
            const size_t close = body.find(']');
            if (close != std::string::npos) body = body.substr(0, close);
            std::vector<double> vals;
            std::istringstream vs(body);
            std::string tok;
            while (std::getline(vs, tok, ',')) {
                const size_t t0 = tok.find_first_not_of(" \t");
                if (t0 == std::string::npos) continue;
                const size_t t1 = tok.find_last_not_of(" \t\r");
                tok = tok.substr(t0, t1 - t0 + 1);
                char* end = nullptr;
                const double v = strtod(tok.c_str(), &end);
                if (end && *end == 0 && end != tok.c_str()) vals.push_back(v);
            }
            *ox = vals.size() > 0 ? vals[0] : 0.0;
            *oy = vals.size() > 1 ? vals[1] : 0.0;
            found = true;
            break;
        }
        if (!found) {
            set_err(err, err_len, "yaml: no origin");
            return -1;
        }
    }
    double occ_th = 0.65;
    (void)yaml_number(yaml, "occupied_thresh", &occ_th);
    // P5 header tokens (whitespace separated, '#' comments), one whitespace after maxval
    size_t idx = 0;
    auto token = [&](std::string& out) -> bool {
        for (;;) {
            while (idx < pgm_len && isspace(pgm[idx])) ++idx;
            if (idx < pgm_len && pgm[idx] == '#') {
                while (idx < pgm_len && pgm[idx] != '\n') ++idx;
                continue;
            }
            break;
        }
        const size_t start = idx;
        while (idx < pgm_len && !isspace(pgm[idx])) ++idx;
        if (start == idx) return false;
        out.assign(reinterpret_cast<const char*>(pgm + start), idx - start);
        return true;
    };
    std::string tok;
    if (!token(tok)) {
        set_err(err, err_len, "pgm: truncated");
        return -1;
    }
    if (tok != "P5") {
        set_err(err, err_len, "pgm: expected P5, got " + tok);
        return -1;
    }
    long dims[3];
    const char* what[3] = {"pgm: bad width", "pgm: bad height", "pgm: bad maxval"};
    for (int k = 0; k < 3; ++k) {
        if (!token(tok)) {
            set_err(err, err_len, "pgm: truncated");
            return -1;
        }
        char* end = nullptr;
        dims[k] = strtol(tok.c_str(), &end, 10);
        if (!end || *end != 0 || dims[k] < 0) {
            set_err(err, err_len, what[k]);
            return -1;
        }
    }
    idx += 1;
    // Width and height are bounded before their product is formed, so nothing can wrap: both in
    // [1, INT_MAX] (the grid is handed on as int W, H) and w * h no larger than the body.
    if (dims[0] < 1 || dims[0] > INT_MAX || dims[1] < 1 || dims[1] > INT_MAX) {
        set_err(err, err_len, "pgm: width and height must lie in [1, INT_MAX]");
        return -1;
    }
    const size_t w = (size_t)dims[0], h = (size_t)dims[1];
    const size_t have = pgm_len > idx ? pgm_len - idx : 0;
    if (w > have / h) {
        set_err(err, err_len,
                "pgm: header says " + std::to_string(w) + " x " + std::to_string(h) + " px, body has " +
                    std::to_string(have));
        return -1;
    }
    const size_t need = w * h;
    *W = (int)w;
    *H = (int)h;
    if (!cells) return 0;
    if (cells_cap < need) {
        set_err(err, err_len, "output buffer too small");
        return -1;
    }
    for (size_t y = 0; y < h; ++y)
        for (size_t x = 0; x < w; ++x) {
            const double occ_p = (double)(255 - pgm[idx + y * w + x]) / 255.0;
            cells[(h - 1 - y) * w + x] = occ_p > occ_th ? 1 : 0;  // rows top-down in the file
        }
    return 0;
}

int pfh_landmarks_from_text(const char* text, int32_t* ids, double* xyz, int cap, char* err, int err_len) {
    if (!text) return -1;
    std::istringstream in(text);
    std::string line;
    int count = 0, lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        line = line.substr(0, line.find('#'));
        for (char& c : line)
            if (c == '[' || c == ']' || c == ':' || c == ',') c = ' ';
        std::istringstream ls(line);
        std::vector<std::string> nums;
        std::string t;
        while (ls >> t) nums.push_back(t);
        if (nums.empty()) continue;
        if (nums.size() != 4) {
            set_err(err, err_len, "line " + std::to_string(lineno) + ": expected 'id: [x, y, z]'");
            return -1;
        }
        double v[4];
        for (int k = 0; k < 4; ++k) {
            char* end = nullptr;
            v[k] = strtod(nums[k].c_str(), &end);
            if (!end || *end != 0) {
                set_err(err, err_len, "line " + std::to_string(lineno) + ": invalid float literal");
                return -1;
            }
        }
        // `as i32` of the parsed f64: truncation toward zero, saturating, NaN -> 0
        const double iv = v[0] != v[0] ? 0.0 : std::trunc(v[0]);
        const int32_t id = iv >= 2147483647.0 ? INT32_MAX : (iv <= -2147483648.0 ? INT32_MIN : (int32_t)iv);
        if (id < 0) {
            set_err(err, err_len, "line " + std::to_string(lineno) + ": landmark ids must be non-negative");
            return -1;
        }
        if (count < cap && ids && xyz) {
            ids[count] = id;
            xyz[3 * count] = v[1];
            xyz[3 * count + 1] = v[2];
            xyz[3 * count + 2] = v[3];
        }
        ++count;
    }
    return count;
}

int pfh_quat_mean(const double M[16], double q[4]) {
    if (!M || !q) return -1;
    return quat_mean(M, q);
}

void* pfh_gl_new(int W, int H, double res, double ox, double oy, const uint8_t* cells, const pfh_gl_params* p) {
    if (!p || !cells) return nullptr;
    GlobalLoc* g = new (std::nothrow) GlobalLoc();
    if (!g) return nullptr;
    if (!g->create(W, H, res, ox, oy, cells, *p)) {
        delete g;
        return nullptr;
    }
    return g;
}

void pfh_gl_free(void* g) { delete static_cast<GlobalLoc*>(g); }

int pfh_gl_localize(void* gp, const float* scan, int nrays, double* out4k) {
    GlobalLoc* g = static_cast<GlobalLoc*>(gp);
    if (!g || !scan || !out4k || nrays <= 0) return -1;
    return g->localize(scan, nrays, out4k);
}

int pfh_gl_set_seed(void* gp, uint64_t seed) {
    GlobalLoc* g = static_cast<GlobalLoc*>(gp);
    if (!g) return -1;
    g->set_seed(seed);
    return 0;
}

int pfh_gl_localize_learned(void* gp, void* mlp, const float* scan, int nrays, const pfh_learned_params* lp,
                            double* out4k, pfh_learned_info* info) {
    GlobalLoc* g = static_cast<GlobalLoc*>(gp);
    if (!g || !mlp || !scan || !lp || !out4k || nrays <= 0) return -1;
    return g->localize_learned(mlp, scan, nrays, *lp, out4k, info);
}

void* pfh_gl_kernels(void* gp) {
    GlobalLoc* g = static_cast<GlobalLoc*>(gp);
    return g ? g->handle() : nullptr;
}

}  // extern "C"

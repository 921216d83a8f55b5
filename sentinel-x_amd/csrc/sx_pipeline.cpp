// sx_pipeline.cpp -- host orchestration of the filter cascade of score_target
// (sentinel-x src/hmmer/pipeline.rs:159-189) on the device: per-profile pass
// thresholds, the batched device-resident chain (optionally captured in a HIP graph)
// and the plain one-profile-at-a-time form.
#include "sx_pipeline.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <atomic>
#include <cstring>
#include <numeric>
#include <thread>

#include "../../include/sx_hmmer.h"

namespace sx {

// pipeline.rs:87-95
double gumbel_surv(double x, double mu, double lambda) {
    const double y = lambda * (x - mu);
    const double ey = -std::exp(-y);
    if (std::fabs(ey) < 1e-5) return -ey;
    return 1.0 - std::exp(ey);
}

// pipeline.rs:98-104
double exp_logsurv(double x, double tau, double lambda) {
    if (x < tau) return 0.0;
    return -lambda * (x - tau);
}

namespace {

int64_t f32_ord(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? -(int64_t)(b & 0x7fffffffu) : (int64_t)b;
}

float ord_f32(int64_t o) {
    uint32_t b = o >= 0 ? (uint32_t)o : (0x80000000u | (uint32_t)(-o));
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

}  // namespace

// Smallest f32 y (from -inf to +inf) that `rejects` lets through; rejects must be true
// up to some point and false after it (each reference tail test is monotone in y).
// Checks the 256 neighbours on each side of the boundary and counts any that disagree.
float pass_threshold(const std::function<bool(float)>& rejects, int* violations) {
    int64_t lo = f32_ord(-INFINITY), hi = f32_ord(INFINITY);
    if (!rejects(-INFINITY)) return -INFINITY;
    if (rejects(INFINITY)) return INFINITY;
    while (hi - lo > 1) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (rejects(ord_f32(mid)))
            lo = mid;
        else
            hi = mid;
    }
    if (violations) {
        for (int64_t o = std::max(hi - 256, f32_ord(-INFINITY)); o < hi; o++)
            if (!rejects(ord_f32(o))) ++*violations;
        for (int64_t o = hi; o < std::min(hi + 256, f32_ord(INFINITY)); o++)
            if (rejects(ord_f32(o))) ++*violations;
    }
    return ord_f32(hi);
}

// The least value the int16 MSV bound can end at for a target of length l once one of its
// cells has reached 32767 (32 nats): the row's xJ is then at least 32767 * 4096 + q(-ln2)
// and falls by at most q(tloop) per row, and the result adds q(tmove).  Targets longer
// than SX_Q_MAX_L get the least int64 (every pair a candidate).
static int64_t q_saturated_floor(int32_t l) {
    if (l > SX_Q_MAX_L) return INT64_MIN / 4;
    float tloop, tmove;
    msv_length((size_t)l, &tloop, &tmove);
    return 32767LL * 4096 + q_const(-LN2F) + (int64_t)l * std::min<int64_t>(q_const(tloop), 0) + q_const(tmove);
}

// Model-state odds of the two-state composition filter for every code
// (pipeline.rs:116-131); the background state's odds are 1 for every code.
static void bias_odds(const Profile& p, double* eo1) {
    for (int x = 0; x < K; x++) eo1[x] = (double)p.compo[x] / (double)BG[x];
    for (int code = 20; code <= 23; code++) {
        const int* mem;
        const int nm = degenerate_members(code, &mem);
        double e1 = 0.0, z = 0.0;
        for (int j = 0; j < nm; j++) {
            e1 += (double)p.compo[mem[j]];
            z += (double)BG[mem[j]];
        }
        eo1[code] = e1 / z;
    }
    eo1[CODE_STAR] = 1.0;
}

// bias_filter_score (pipeline.rs:107-150): the reference's f64 operations in its order,
// with glibc's log.  The host recomputes here the scores whose f32 rounding the device's
// error bound cannot certify (sxc_bias).
float bias_filter_score(const Profile& p, const uint8_t* dsq, int l) {
    double eo1[NCODE];
    bias_odds(p, eo1);
    const double l0 = 400.0;
    const double l1 = (double)p.m / 8.0;
    const double t00 = l0 / (l0 + 1.0), t01 = 1.0 / (l0 + 1.0);
    const double t10 = 1.0 / (l1 + 1.0), t11 = l1 / (l1 + 1.0);
    double logsc = 0.0;
    double d0 = 1.0 * 0.999, d1 = eo1[dsq[0]] * 0.001;
    double mx = std::max(std::max(d0, d1), DBL_MIN);
    d0 /= mx;
    d1 /= mx;
    logsc += std::log(mx);
    for (int i = 1; i < l; i++) {
        const int x = dsq[i];
        // the background state's odds are 1 for every code
        const double n0 = d0 * t00 + d1 * t10;
        const double n1 = (d0 * t01 + d1 * t11) * eo1[x];
        mx = std::max(std::max(n0, n1), DBL_MIN);
        d0 = n0 / mx;
        d1 = n1 / mx;
        logsc += std::log(mx);
    }
    logsc += std::log(d0 + d1);
    const double p1 = (double)l / ((double)l + 1.0);
    return (float)(logsc + (double)l * std::log(p1) + std::log(1.0 - p1));
}

}  // namespace sx

extern "C" {
__global__ void sxc_msv_pass(int, const float*, const float*, const float*, const int32_t*, const int32_t*, int64_t,
                             int64_t, int32_t*, const int64_t*, int64_t, int32_t*, int32_t*, int32_t*, float*);
__global__ void sxc_scan(const int32_t*, int64_t, const int64_t*, int64_t*);
__global__ void sxc_bias(const int64_t*, const int32_t*, const int32_t*, const float*, const double*, const int32_t*,
                         const float*, SxTargets, float*, uint8_t*, uint8_t*, int32_t*, int32_t*, int32_t, int32_t);
__global__ void sxc_compact(int, const uint8_t*, const int64_t*, const int32_t*, const int32_t*, int32_t*,
                            const int64_t*, int32_t*, int32_t*);
__global__ void sxc_tiles(int, int64_t, const int32_t*, const int32_t*, const int32_t*, int32_t*, const int64_t*,
                          const int64_t*, SxTile*);
__global__ void sxc_gather_t(const int64_t*, const int32_t*, const int32_t*, int32_t*);
__global__ void sxc_vit_pass(const int64_t*, const int32_t*, const float*, const float*, const int32_t*, const float*,
                             float*, int32_t*, uint8_t*);
__global__ void sxc_fwd_pass(const int64_t*, const int32_t*, const double*, const float*, const int32_t*,
                             const float*, double*, int32_t*);
__global__ void sxc_res_init(const int64_t*, int64_t, float*, double*, int32_t*);
__global__ void sxc_zero_i32(int32_t*, int64_t);
__global__ void sxc_msvq_cand(int, const int32_t*, const int64_t*, const int64_t*, const int64_t*, const uint8_t*,
                              const int32_t*, const int32_t*, int64_t, int64_t, int32_t*, const int64_t*, int64_t,
                              int32_t*, int32_t*, int32_t*);
__global__ void sxc_count_sorted(const int64_t*, int64_t, const int32_t*, int64_t, int32_t*);
__global__ void sxc_msv_pass_list(int, const int64_t*, const int32_t*, const int32_t*, const float*, const float*,
                                  const float*, const int32_t*, const int32_t*, int32_t*, const int64_t*, int64_t,
                                  int32_t*, int32_t*, int32_t*, float*);
__global__ void sxc_clamp(const int64_t*, int64_t, int64_t*);
__global__ void sxc_vitq_cand(const int64_t*, const int32_t*, const int32_t*, const float*, const int32_t*,
                              const int32_t*, const float*, const uint8_t*, const int64_t*, int32_t*, uint8_t*);
__global__ void sxc_records(const int64_t*, int64_t, const int32_t*, const int32_t*, const float*, const float*,
                            const float*, const double*, const int32_t*, const uint8_t*, void*);
}

namespace sx {

namespace {

const int CB = 256;  // cascade kernels' block size
// capacity of the device list of bias scores handed to the host; past it the host scans
// every record's flag
const int32_t BIAS_LIST_CAP = 1 << 16;
// an MSV launch with fewer tiles than this many per workgroup of a full grid spends a good
// part of its time on its last tiles: the cascade's MSV launches then run side by side
const int MSV_SIDE_TILES = 16;

int64_t nblk(int64_t n) { return (n + CB - 1) / CB; }

template <class T>
T* dptr(DeviceBuf& b, size_t n) {
    return (T*)b.ensure(std::max<size_t>(n, 1) * sizeof(T));
}

}  // namespace

struct CascadeState {
    // identity of what the plan was built for
    size_t nprof = 0;
    uint32_t ntarg = 0;
    double f1 = -1, f2 = -1, f3 = -1;
    int bound = -1;   // the plan runs the int16 MSV bound before the f32 MSV
    int vbound = -1;  // the plan runs the int16 Viterbi bound before the f32 Viterbi
    int64_t cap1 = 0;
    // records of the last run (one per MSV survivor), kept for sx_cascade_records
    std::vector<sx_pair_scores> last;
    // host-side tables
    std::vector<float> thr_msv, thr_vit, thr_fwd, nullsc;
    // what the thresholds were computed for (profile set, F1-F3, uncalibrated profiles let
    // through): a new target set keeps them
    uint64_t thr_version = ~0ull;
    double thr_f[3] = {-1, -1, -1};
    bool thr_uncal = false;
    std::vector<int32_t> perm;           // profile ids by increasing model length
    int threshold_violations = 0;
    // the int16 MSV bound's thresholds in its units (per length-sorted target, per
    // profile) and the profiles it skips (every pair a candidate)
    std::vector<int64_t> nullq, capq, thrq;
    std::vector<uint8_t> noq;
    // the int16 Viterbi bound: profiles it skips, per-target saturation floors (by id)
    std::vector<uint8_t> noqv;
    std::vector<int64_t> capv;
    // device tables
    DeviceBuf d_thr_msv, d_thr_vit, d_thr_fwd, d_nullsc, d_perm, d_eo1, d_ng_vit, d_ng_fwd;
    DeviceBuf d_nullq, d_capq, d_thrq, d_noq, d_ng_msv, d_noqv, d_capv, d_ng_vq;
    std::vector<DeviceBuf> msv_plists;
    struct Group {
        Shape s;
        int32_t* plist;
        int n;
        uint64_t cells;  // sum of M * target length over the group's pairs
    };
    std::vector<Group> msv_groups;  // f32 MSV over all pairs, or the int16 bound's groups
    struct Run {
        Shape s;
        int64_t a, b;  // profile positions [a, b)
    };
    std::vector<Run> vit_runs, fwd_runs;
    std::vector<Run> msv_runs;  // f32 MSV over the bound's candidates
    std::vector<Run> vq_runs;   // the int16 Viterbi bound (profiles it covers)
    // device work buffers
    DeviceBuf usc, cnt1, off1, s_p, s_pi, s_t, s_usc, s_bias, flag, res_vit, res_fwd, res_stage;
    // bias scores for the host to recompute: per-survivor flag, count, index list
    DeviceBuf bflag, bl_n, bl_list;
    DeviceBuf cnt2, off2, list2, per2, tcount, toff2, eoff2, tiles2, tids2, vsc, flag3;
    DeviceBuf cnt3, off3, list3, per3, toff3, eoff3, tiles3, tids3, fsc, recs;
    // the bound's candidates (pi, j, target id), their f32 MSV scores and tiles; clamped
    // list counts (survivors, candidates) for the kernels that read the lists
    DeviceBuf c_pi, c_j, c_t, cusc, cntq, offq, perq, toffq, eoffq, tilesq, nclamp;
    const int64_t* d_n1 = nullptr;     // raw survivor count (may exceed cap1)
    const int64_t* d_ncand = nullptr;  // raw candidate count
    // the Viterbi bound's results, candidate flags, candidate list and its tiles
    DeviceBuf vq, flagq, cntc, offc, listc, perc, toffc, eoffc, tilesc, tidsc;
    const int64_t* d_nvc = nullptr;    // Viterbi candidates
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t evq = nullptr;   // after the int16 MSV bound's launches (eager runs)
    hipEvent_t evvq = nullptr;  // after the int16 Viterbi bound's launches (eager runs)
    // an event before (lev0) and after (lev) every kernel launch of the three DP stages,
    // on the launch's own stream (eager runs only)
    std::vector<hipEvent_t> lev0, lev;
    std::vector<sx_launch_stat> launches;  // per-launch timing of the last eager run
    ~CascadeState() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        if (evq) (void)hipEventDestroy(evq);
        if (evvq) (void)hipEventDestroy(evvq);
        for (auto& e : lev)
            if (e) (void)hipEventDestroy(e);
        for (auto& e : lev0)
            if (e) (void)hipEventDestroy(e);
    }
};

void cascade_release(Ctx& c) { c.cascade.reset(); }

// A record whose bias score the device could not certify (sxc_bias): the score is
// recomputed in the reference's order and the pair re-decided (pipeline.rs:170-183) from
// the scores the device computed with the lower f32 of the bound.  With that lower score
// every test passes at least the pairs the reference passes, so each score the reference
// computes is there; scores it would not compute are cleared.  Returns the stage the
// device gave the record.
static int recheck_bias(const Ctx& c, const CascadeState& S, sx_pair_scores& r) {
    const int before = r.stage;
    const float fs = bias_filter_score(c.profs[r.profile], c.h_codes.data() + c.h_toff[r.target], c.tlen[r.target]);
    r.bias = fs;
    r.flags |= SX_REC_BIAS_HOST;
    if (before < 2) return before;
    if (!((r.msv - fs) / LN2F >= S.thr_msv[r.profile])) {
        r.stage = 1;
        r.vit = r.fwd = NAN;
        r.fwd64 = NAN;
        return before;
    }
    r.stage = 2;
    if (before < 3 || !((r.vit - fs) / LN2F >= S.thr_vit[r.profile])) {
        r.fwd = NAN;
        r.fwd64 = NAN;
        return before;
    }
    r.stage = 3;
    if (before == 4 && (r.fwd - fs) / LN2F >= S.thr_fwd[r.profile]) r.stage = 4;
    return before;
}

// recheck_bias over a list of record indices, on host threads when there are many;
// adjusts the bias / Viterbi pass counts for the records whose decision changed.
static void recheck_records(const Ctx& c, const CascadeState& S, std::vector<sx_pair_scores>& recs,
                            const std::vector<int64_t>& idx, int64_t* n_bias_pass, int64_t* n_vit_pass) {
    std::vector<int8_t> before(idx.size());
    const size_t nth = idx.size() < 4096 ? 1 : std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    std::atomic<size_t> next{0};
    auto work = [&] {
        for (size_t k; (k = next.fetch_add(256)) < idx.size();)
            for (size_t j = k; j < std::min(idx.size(), k + 256); j++)
                before[j] = (int8_t)recheck_bias(c, S, recs[(size_t)idx[j]]);
    };
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nth; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (size_t j = 0; j < idx.size(); j++) {
        const int after = recs[(size_t)idx[j]].stage;
        if (before[j] >= 2 && after < 2) --*n_bias_pass;
        if (before[j] >= 3 && after < 3) --*n_vit_pass;
    }
}

static void build_plan(Ctx& c, CascadeState& S, double f1, double f2, double f3, int bounds) {
    const int bound = (bounds & SX_BOUND_MSV) ? 1 : 0, vbound = (bounds & SX_BOUND_VITERBI) ? 1 : 0;
    const size_t P = c.profs.size();
    const uint32_t T = c.nt;
    for (const auto& p : c.profs)
        if (!p.stats.calibrated() && !c.cascade_pass_uncalibrated)
            throw SxError(SX_ERR_STATE, "profile '" + p.name +
                                            "' has no STATS lines; the filter cascade needs calibrated models");
    S.nprof = P;
    S.ntarg = T;
    S.f1 = f1;
    S.f2 = f2;
    S.f3 = f3;
    S.bound = bound;
    S.vbound = vbound;
    const bool thr_ok = S.thr_version == c.profiles_version && S.thr_f[0] == f1 && S.thr_f[1] == f2 &&
                        S.thr_f[2] == f3 && S.thr_uncal == c.cascade_pass_uncalibrated && S.thr_msv.size() == P;
    if (!thr_ok) {
        S.thr_version = c.profiles_version;
        S.thr_f[0] = f1;
        S.thr_f[1] = f2;
        S.thr_f[2] = f3;
        S.thr_uncal = c.cascade_pass_uncalibrated;
    }
    S.thr_msv.resize(P);
    S.thr_vit.resize(P);
    S.thr_fwd.resize(P);
    if (!thr_ok) S.threshold_violations = 0;
    for (size_t p = 0; p < P && !thr_ok; p++) {
        const Stats& st = c.profs[p].stats;
        if (!st.calibrated()) {
            // the reference runs no filter for an uncalibrated model (pipeline.rs:164)
            S.thr_msv[p] = S.thr_vit[p] = S.thr_fwd[p] = NEG_INF;
            continue;
        }
        S.thr_msv[p] = pass_threshold(
            [&](float y) { return gumbel_surv((double)y, st.msv[0], st.msv[1]) > f1; }, &S.threshold_violations);
        S.thr_vit[p] = pass_threshold(
            [&](float y) { return gumbel_surv((double)y, st.vit[0], st.vit[1]) > f2; }, &S.threshold_violations);
        S.thr_fwd[p] = pass_threshold(
            [&](float y) { return std::exp(exp_logsurv((double)y, st.fwd[0], st.fwd[1])) > f3; },
            &S.threshold_violations);
    }
    S.nullsc.resize(T);
    for (uint32_t t = 0; t < T; t++) S.nullsc[t] = null_score((size_t)c.tlen[t]);
    std::vector<double> eo1(P * 25);
    for (size_t p = 0; p < P; p++) bias_odds(c.profs[p], eo1.data() + p * 25);
    S.perm.resize(P);
    std::iota(S.perm.begin(), S.perm.end(), 0);
    std::stable_sort(S.perm.begin(), S.perm.end(), [&](int a, int b) { return c.profs[a].m < c.profs[b].m; });
    std::vector<int32_t> ngv(P), ngf(P), ngm(P), ngq(P, 64);
    S.vit_runs.clear();
    S.fwd_runs.clear();
    S.msv_runs.clear();
    S.vq_runs.clear();
    // the Viterbi bound skips models beyond its shapes, profiles without a finite
    // threshold, and profiles with a positive insert score (an insert state could then
    // climb past the M states it comes from, and the saturation floor would not hold)
    S.noqv.assign(P, 1);
    if (vbound)
        for (size_t p = 0; p < P; p++) {
            const Profile& pr = c.profs[p];
            bool ok = pr.m <= SX_Q_MAX_M && std::isfinite(S.thr_vit[p]);
            for (int k = 1; k <= pr.m && ok; k++)
                if (q_trans(pr.t_row(SX_TMI)[k]) > 0 || q_trans(pr.t_row(SX_TII)[k]) > 0) ok = false;
            S.noqv[p] = ok ? 0 : 1;
        }
    for (size_t i = 0; i < P; i++) {
        const int m = c.profs[S.perm[i]].m;
        const Shape sv = c.shape_for(1, SX_FORM_BATCHED, m), sf = c.shape_for(2, SX_FORM_BATCHED, m);
        const Shape sm = c.shape_for(0, SX_FORM_BATCHED, m);
        ngv[i] = c.kernel(1, sv.q, sv.g, false, sv.hoist).block / sv.g;
        ngf[i] = c.kernel(2, sf.q, sf.g, false, sf.hoist).block / sf.g;
        ngm[i] = c.kernel(0, sm.q, sm.g, false, sm.hoist).block / sm.g;
        auto add = [&](std::vector<CascadeState::Run>& runs, Shape s) {
            if (!runs.empty() && runs.back().s.q == s.q && runs.back().s.g == s.g && runs.back().s.hoist == s.hoist)
                runs.back().b = (int64_t)i + 1;
            else
                runs.push_back({s, (int64_t)i, (int64_t)i + 1});
        };
        add(S.vit_runs, sv);
        add(S.fwd_runs, sf);
        if (bound) add(S.msv_runs, sm);
        if (vbound && !S.noqv[S.perm[i]]) {
            const Shape sq = c.shape_for(SX_STAGE_VIT16, SX_FORM_BATCHED, m);
            ngq[i] = c.kernel(SX_STAGE_VIT16, sq.q, sq.g, false, sq.hoist).block / sq.g;
            // a run never spans a skipped profile: its tiles are not the bound's to score
            if (!S.vq_runs.empty() && S.vq_runs.back().b == (int64_t)i && S.vq_runs.back().s.q == sq.q &&
                S.vq_runs.back().s.g == sq.g && S.vq_runs.back().s.hoist == sq.hoist)
                S.vq_runs.back().b = (int64_t)i + 1;
            else
                S.vq_runs.push_back({sq, (int64_t)i, (int64_t)i + 1});
        }
    }
    S.capv.assign(T, 0);
    if (vbound)
        for (uint32_t t = 0; t < T; t++) {
            const int32_t l = c.tlen[t];
            if (l > SX_Q_MAX_L) {
                S.capv[t] = INT64_MIN / 4;
                continue;
            }
            const Xsc x = Xsc::make((size_t)l, true);
            S.capv[t] = 32767LL * 4096 + q_const(-LN2F) + (int64_t)l * std::min<int64_t>(q_const(x.n_loop), 0) +
                        q_const(x.n_move);
        }
    // The int16 bound's thresholds.  The f32 test passes u only if
    // fl(fl(u - null) / ln2) >= thr, so u >= null + ln2 * thr - delta, where delta = 2^-21
    // (|null| + |ln2 * thr|) covers the two roundings; 1e-6 (1 + |.|) per term exceeds it.
    // Floors of the two scaled terms sum to at most the scaled bound.
    S.noq.assign(P, 0);
    S.thrq.assign(P, 0);
    S.nullq.assign(T, 0);
    S.capq.assign(T, 0);
    if (bound) {
        const double F = SX_Q_FINE;
        for (size_t p = 0; p < P; p++) {
            const float thr = S.thr_msv[p];
            if (c.profs[p].m > SX_Q_MAX_M || !std::isfinite(thr)) {
                S.noq[p] = 1;  // beyond the bound's shapes, or no threshold: all candidates
                if (thr == INFINITY) S.noq[p] = 0, S.thrq[p] = INT64_MAX / 4;  // nothing passes
                continue;
            }
            const double ct = (double)LN2F * (double)thr;
            S.thrq[p] = (int64_t)std::floor(F * ct - F * 1e-6 * (1.0 + std::fabs(ct)));
        }
        for (uint32_t j = 0; j < T; j++) {
            const double n = (double)S.nullsc[c.sorted[j]];
            S.nullq[j] = (int64_t)std::floor(F * n - F * 1e-6 * (1.0 + std::fabs(n)));
            S.capq[j] = q_saturated_floor(c.tlen[c.sorted[j]]);
        }
    }
    // MSV groups (all pairs, implicit tiles): f32 MSV, or the int16 bound (stage 4) for
    // the profiles it covers
    std::map<std::tuple<int, int, int>, std::vector<int32_t>> g;
    for (size_t p = 0; p < P; p++) {
        if (bound && S.noq[p]) continue;
        if (bound && S.thrq[p] == INT64_MAX / 4) continue;
        const Shape s = c.shape_for(bound ? SX_STAGE_MSV16 : 0, SX_FORM_BATCHED, c.profs[p].m);
        g[{s.q, s.g, s.hoist}].push_back((int32_t)p);
    }
    S.msv_plists.clear();
    S.msv_plists.reserve(g.size());
    S.msv_groups.clear();
    for (auto& kv : g) {
        S.msv_plists.emplace_back();
        int32_t* d = dptr<int32_t>(S.msv_plists.back(), kv.second.size());
        HIPCHK(hipMemcpy(d, kv.second.data(), kv.second.size() * 4, hipMemcpyHostToDevice));
        uint64_t cells = 0;
        for (int32_t p : kv.second) cells += (uint64_t)c.profs[p].m * (uint64_t)c.total_len;
        S.msv_groups.push_back({Shape{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first)}, d,
                                (int)kv.second.size(), cells});
    }
    auto up = [](DeviceBuf& b, const void* src, size_t bytes) {
        b.ensure(std::max<size_t>(bytes, 16));
        HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
    };
    up(S.d_thr_msv, S.thr_msv.data(), P * 4);
    up(S.d_thr_vit, S.thr_vit.data(), P * 4);
    up(S.d_thr_fwd, S.thr_fwd.data(), P * 4);
    {
        // null scores in the length-sorted order the MSV test walks
        std::vector<float> ns(T);
        for (uint32_t j = 0; j < T; j++) ns[j] = S.nullsc[c.sorted[j]];
        up(S.d_nullsc, ns.data(), (size_t)T * 4);
    }
    up(S.d_perm, S.perm.data(), P * 4);
    up(S.d_eo1, eo1.data(), eo1.size() * 8);
    up(S.d_ng_vit, ngv.data(), P * 4);
    up(S.d_ng_fwd, ngf.data(), P * 4);
    up(S.d_ng_msv, ngm.data(), P * 4);
    up(S.d_nullq, S.nullq.data(), (size_t)T * 8);
    up(S.d_capq, S.capq.data(), (size_t)T * 8);
    up(S.d_thrq, S.thrq.data(), P * 8);
    up(S.d_noq, S.noq.data(), P);
    up(S.d_noqv, S.noqv.data(), P);
    up(S.d_capv, S.capv.data(), (size_t)T * 8);
    up(S.d_ng_vq, ngq.data(), P * 4);
    for (auto& e : S.ev)
        if (!e) HIPCHK(hipEventCreate(&e));
    if (!S.evq) HIPCHK(hipEventCreate(&S.evq));
    if (!S.evvq) HIPCHK(hipEventCreate(&S.evvq));
    const size_t nl =
        S.msv_groups.size() + S.msv_runs.size() + S.vq_runs.size() + S.vit_runs.size() + S.fwd_runs.size();
    while (S.lev.size() < nl) {
        S.lev.push_back(nullptr);
        HIPCHK(hipEventCreate(&S.lev.back()));
        S.lev0.push_back(nullptr);
        HIPCHK(hipEventCreate(&S.lev0.back()));
    }
    S.launches.clear();
}

static void alloc_buffers(Ctx& c, CascadeState& S, int64_t cap1) {
    const int64_t P = (int64_t)c.profs.size(), T = c.nt;
    S.cap1 = cap1;
    const int64_t nb1 = nblk(P * T), nbc = nblk(cap1);
    dptr<float>(S.usc, (size_t)(P * T));  // f32 MSV scores, or the int16 bound's int32 results
    dptr<int32_t>(S.cnt1, std::max(nb1, nbc));
    dptr<int64_t>(S.off1, std::max(nb1, nbc) + 1);
    dptr<int64_t>(S.nclamp, 2);
    if (S.bound) {
        for (DeviceBuf* b : {&S.c_pi, &S.c_j, &S.c_t}) dptr<int32_t>(*b, cap1);
        dptr<float>(S.cusc, cap1);
        const int64_t nbq = (P * T + (int64_t)CB * SX_QCAND_ITEMS - 1) / ((int64_t)CB * SX_QCAND_ITEMS);
        dptr<int32_t>(S.cntq, nbq);
        dptr<int64_t>(S.offq, nbq + 1);
        dptr<int32_t>(S.perq, P);
        for (DeviceBuf* b : {&S.toffq, &S.eoffq}) dptr<int64_t>(*b, P + 1);
        dptr<SxTile>(S.tilesq, cap1 + P);
    }
    if (S.vbound) {
        dptr<int32_t>(S.vq, cap1);
        dptr<uint8_t>(S.flagq, cap1);
        dptr<int32_t>(S.cntc, nbc);
        dptr<int64_t>(S.offc, nbc + 1);
        for (DeviceBuf* b : {&S.listc, &S.tidsc}) dptr<int32_t>(*b, cap1);
        dptr<int32_t>(S.perc, P);
        for (DeviceBuf* b : {&S.toffc, &S.eoffc}) dptr<int64_t>(*b, P + 1);
        dptr<SxTile>(S.tilesc, cap1 + P);
    }
    dptr<int32_t>(S.s_p, cap1);
    dptr<int32_t>(S.s_pi, cap1);
    dptr<int32_t>(S.s_t, cap1);
    dptr<float>(S.s_usc, cap1);
    dptr<float>(S.s_bias, cap1);
    dptr<uint8_t>(S.flag, cap1);
    dptr<uint8_t>(S.bflag, cap1);
    dptr<int32_t>(S.bl_n, 1);
    dptr<int32_t>(S.bl_list, BIAS_LIST_CAP);
    dptr<float>(S.res_vit, cap1);
    dptr<double>(S.res_fwd, cap1);
    dptr<int32_t>(S.res_stage, cap1);
    for (DeviceBuf* b : {&S.cnt2, &S.cnt3}) dptr<int32_t>(*b, nbc);
    for (DeviceBuf* b : {&S.off2, &S.off3}) dptr<int64_t>(*b, nbc + 1);
    for (DeviceBuf* b : {&S.list2, &S.list3, &S.tids2, &S.tids3}) dptr<int32_t>(*b, cap1);
    for (DeviceBuf* b : {&S.per2, &S.per3, &S.tcount}) dptr<int32_t>(*b, P);
    for (DeviceBuf* b : {&S.toff2, &S.eoff2, &S.toff3, &S.eoff3}) dptr<int64_t>(*b, P + 1);
    for (DeviceBuf* b : {&S.tiles2, &S.tiles3}) dptr<SxTile>(*b, cap1 + P);
    dptr<float>(S.vsc, cap1);
    dptr<uint8_t>(S.flag3, cap1);
    dptr<double>(S.fsc, cap1);
    S.recs.ensure(std::max<int64_t>(cap1, 1) * sizeof(sx_pair_scores));
    if (S.gexec) {
        (void)hipGraphExecDestroy(S.gexec);
        S.gexec = nullptr;
    }
}

// The batched chain.  Everything is enqueued on c.stream with no host synchronisation,
// so it can be captured.  with_events records stage boundaries for the timing split.
static void enqueue_chain(Ctx& c, CascadeState& S, bool with_events) {
    const int64_t P = (int64_t)c.profs.size(), T = c.nt, cap = S.cap1;
    const int64_t nbc = nblk(cap), nbp = nblk(P);
    hipStream_t st = c.stream;
    SxSpecialsOut nosp{};
    if (with_events) HIPCHK(hipEventRecord(S.ev[0], st));
    int slot = 0;
    int lslot = 0;
    int32_t *s_p = (int32_t*)S.s_p.p, *s_pi = (int32_t*)S.s_pi.p, *s_t = (int32_t*)S.s_t.p;
    float *s_usc = (float*)S.s_usc.p, *s_bias = (float*)S.s_bias.p;
    auto tiles = [&](DeviceBuf& per, DeviceBuf& ng, DeviceBuf& toff, DeviceBuf& eoff, DeviceBuf& tl,
                     const int64_t* n_list, DeviceBuf* list, DeviceBuf* tids) {
        hipLaunchKernelGGL(sxc_tiles, dim3(nbp), dim3(CB), 0, st, 0, P, (const int32_t*)per.p,
                           (const int32_t*)ng.p, (const int32_t*)S.d_perm.p, (int32_t*)S.tcount.p,
                           (const int64_t*)toff.p, (const int64_t*)eoff.p, (SxTile*)tl.p);
        hipLaunchKernelGGL(sxc_scan, dim3(1), dim3(1024), 0, st, (const int32_t*)S.tcount.p, P,
                           (const int64_t*)nullptr, (int64_t*)toff.p);
        hipLaunchKernelGGL(sxc_scan, dim3(1), dim3(1024), 0, st, (const int32_t*)per.p, P, (const int64_t*)nullptr,
                           (int64_t*)eoff.p);
        hipLaunchKernelGGL(sxc_tiles, dim3(nbp), dim3(CB), 0, st, 1, P, (const int32_t*)per.p,
                           (const int32_t*)ng.p, (const int32_t*)S.d_perm.p, (int32_t*)S.tcount.p,
                           (const int64_t*)toff.p, (const int64_t*)eoff.p, (SxTile*)tl.p);
        if (list)
            hipLaunchKernelGGL(sxc_gather_t, dim3(nbc), dim3(CB), 0, st, n_list, (const int32_t*)list->p,
                               (const int32_t*)s_t, (int32_t*)tids->p);
    };
    // A stage's runs cover disjoint tiles of a list: they go out side by side on the side
    // streams (the small runs of a stage no longer wait for the large ones), each with its
    // own head slot (slots repeat only every 32 launches, on the same side stream).  The
    // runs of the longest models go first: on a small cascade their longest pairs set the
    // stage's time, and a side stream would otherwise start them last.  Launch events keep
    // the runs' order.
    auto stage_runs = [&](int stage, std::vector<CascadeState::Run>& runs, DeviceBuf& tl, DeviceBuf& tids,
                          DeviceBuf& toff, void* out) {
        Fanout fo(c, st);
        const int lslot0 = lslot;
        for (size_t ri = runs.size(); ri-- > 0;) {
            const auto& r = runs[ri];
            const int ls_ev = lslot0 + (int)ri;
            const KernelInfo& k = c.kernel(stage, r.s.q, r.s.g, false, r.s.hoist);
            SxWork w{};
            w.tiles = (const SxTile*)tl.p;
            w.tids = (const int32_t*)tids.p;
            w.dyn_lo = (const int64_t*)toff.p + r.a;
            w.dyn_hi = (const int64_t*)toff.p + r.b;
            w.ntiles = 0;
            const int grid = c.persistent_grid(k, Ctx::lds_bytes(k), 1 << 30);
            hipStream_t ls = fo.next();
            if (with_events) HIPCHK(hipEventRecord(S.lev0[ls_ev], ls));
            c.launch(k, stage, w, out, nosp, grid, true, slot++ % 32 + 32, ls);
            if (with_events) HIPCHK(hipEventRecord(S.lev[ls_ev], ls));
        }
        lslot = lslot0 + (int)runs.size();
        fo.join();
    };
    // 1. MSV over all pairs (the f32 scores, or the int16 bound of each pair).  When every
    // launch has many tiles for each workgroup of a full grid, the launches run one after
    // the other on the chain's stream, each filling the device on its own.  Otherwise
    // (fewer profiles or targets) a launch's last tiles leave the device idle for a good
    // part of its time, so all of them go out side by side on the side streams: launch i
    // on side stream i mod 4 with head slot i mod 32, so launches that share a slot share
    // a stream.
    const int msv_stage = S.bound ? SX_STAGE_MSV16 : 0;
    std::vector<SxWork> msv_work;
    bool msv_side = false;
    for (auto& gr : S.msv_groups) {
        const KernelInfo& k = c.kernel(msv_stage, gr.s.q, gr.s.g, false, gr.s.hoist);
        SxWork w{};
        w.profs = gr.plist;
        w.tids = c.d_sorted;
        w.chunk = k.block / gr.s.g;
        w.ilv = w.chunk % 64 == 0 ? 1 : 0;
        w.ntids = (int32_t)T;
        w.nchunks = (int32_t)((T + w.chunk - 1) / w.chunk);
        w.ntiles = (int64_t)w.nchunks * gr.n;
        w.out_stride = T;
        w.out_by_pos = 1;  // usc[profile][length-sorted position]: contiguous per tile
        const int full = c.persistent_grid(k, Ctx::lds_bytes(k), (int64_t)1 << 40);
        if (w.ntiles < (int64_t)MSV_SIDE_TILES * full) msv_side = true;
        msv_work.push_back(w);
    }
    {
        Fanout fo(c, st);
        for (size_t i = 0; i < S.msv_groups.size(); i++) {
            const auto& gr = S.msv_groups[i];
            const KernelInfo& k = c.kernel(msv_stage, gr.s.q, gr.s.g, false, gr.s.hoist);
            const SxWork& w = msv_work[i];
            const int grid = c.persistent_grid(k, Ctx::lds_bytes(k), w.ntiles);
            hipStream_t ls = msv_side ? fo.side(slot) : st;
            if (with_events) HIPCHK(hipEventRecord(S.lev0[lslot], ls));
            c.launch(k, msv_stage, w, S.usc.p, nosp, grid, true, slot % 32, ls);
            if (with_events) HIPCHK(hipEventRecord(S.lev[lslot], ls));
            slot++;
            lslot++;
        }
        fo.join();
    }
    int64_t* nclamp = (int64_t*)S.nclamp.p;
    const float* usc = (const float*)S.usc.p;
    if (S.bound) {
        if (with_events) HIPCHK(hipEventRecord(S.evq, st));
        // 1b. candidates of the bound, in (profile position, length-sorted target) order
        const int32_t* qres = (const int32_t*)S.usc.p;
        int32_t *c_pi = (int32_t*)S.c_pi.p, *c_j = (int32_t*)S.c_j.p, *c_t = (int32_t*)S.c_t.p;
        const int64_t nbq = (P * T + (int64_t)CB * SX_QCAND_ITEMS - 1) / ((int64_t)CB * SX_QCAND_ITEMS);
        for (int phase = 0; phase < 2; phase++) {
            hipLaunchKernelGGL(sxc_msvq_cand, dim3(nbq), dim3(CB), 0, st, phase, qres,
                               (const int64_t*)S.d_nullq.p, (const int64_t*)S.d_capq.p,
                               (const int64_t*)S.d_thrq.p, (const uint8_t*)S.d_noq.p,
                               (const int32_t*)S.d_perm.p, c.d_sorted, P, T, (int32_t*)S.cntq.p,
                               (const int64_t*)S.offq.p, cap, c_pi, c_j, c_t);
            if (phase == 0)
                hipLaunchKernelGGL(sxc_scan, dim3(1), dim3(1024), 0, st, (const int32_t*)S.cntq.p, nbq,
                                   (const int64_t*)nullptr, (int64_t*)S.offq.p);
        }
        S.d_ncand = (const int64_t*)S.offq.p + nbq;
        hipLaunchKernelGGL(sxc_clamp, dim3(1), dim3(64), 0, st, S.d_ncand, cap, nclamp + 1);
        // 1c. the f32 MSV of the candidates, tiled per profile as the later stages
        hipLaunchKernelGGL(sxc_count_sorted, dim3(nbp), dim3(CB), 0, st, (const int64_t*)(nclamp + 1), cap,
                           (const int32_t*)c_pi, P, (int32_t*)S.perq.p);
        tiles(S.perq, S.d_ng_msv, S.toffq, S.eoffq, S.tilesq, nclamp + 1, nullptr, nullptr);
        stage_runs(0, S.msv_runs, S.tilesq, S.c_t, S.toffq, S.cusc.p);
        if (with_events) HIPCHK(hipEventRecord(S.ev[1], st));
        // 2. the MSV test over the candidates, in list order
        for (int phase = 0; phase < 2; phase++) {
            hipLaunchKernelGGL(sxc_msv_pass_list, dim3(nbc), dim3(CB), 0, st, phase, (const int64_t*)(nclamp + 1),
                               (const int32_t*)c_pi, (const int32_t*)c_j, (const float*)S.cusc.p,
                               (const float*)S.d_nullsc.p, (const float*)S.d_thr_msv.p, (const int32_t*)S.d_perm.p,
                               c.d_sorted, (int32_t*)S.cnt1.p, (const int64_t*)S.off1.p, cap, s_p, s_pi, s_t, s_usc);
            if (phase == 0)
                hipLaunchKernelGGL(sxc_scan, dim3(1), dim3(1024), 0, st, (const int32_t*)S.cnt1.p, nbc,
                                   (const int64_t*)nullptr, (int64_t*)S.off1.p);
        }
        S.d_n1 = (const int64_t*)S.off1.p + nbc;
    } else {
        S.d_ncand = nullptr;
        if (with_events) HIPCHK(hipEventRecord(S.ev[1], st));
        // 2. MSV test and ordered compaction
        const int64_t nbm = (P * T + (int64_t)CB * SX_QCAND_ITEMS - 1) / ((int64_t)CB * SX_QCAND_ITEMS);
        hipLaunchKernelGGL(sxc_msv_pass, dim3(nbm), dim3(CB), 0, st, 0, usc, (const float*)S.d_nullsc.p,
                           (const float*)S.d_thr_msv.p, (const int32_t*)S.d_perm.p, c.d_sorted, P, T,
                           (int32_t*)S.cnt1.p, (const int64_t*)S.off1.p, cap, s_p, s_pi, s_t, s_usc);
        hipLaunchKernelGGL(sxc_scan, dim3(1), dim3(1024), 0, st, (const int32_t*)S.cnt1.p, nbm,
                           (const int64_t*)nullptr, (int64_t*)S.off1.p);
        hipLaunchKernelGGL(sxc_msv_pass, dim3(nbm), dim3(CB), 0, st, 1, usc, (const float*)S.d_nullsc.p,
                           (const float*)S.d_thr_msv.p, (const int32_t*)S.d_perm.p, c.d_sorted, P, T,
                           (int32_t*)S.cnt1.p, (const int64_t*)S.off1.p, cap, s_p, s_pi, s_t, s_usc);
        S.d_n1 = (const int64_t*)S.off1.p + nbm;
    }
    // the survivor count the later kernels read: at most the lists' capacity
    hipLaunchKernelGGL(sxc_clamp, dim3(1), dim3(64), 0, st, S.d_n1, cap, nclamp);
    const int64_t* n1 = nclamp;
    // 3. bias filter and its test
    hipLaunchKernelGGL(sxc_res_init, dim3(nbc), dim3(CB), 0, st, n1, cap, (float*)S.res_vit.p, (double*)S.res_fwd.p,
                       (int32_t*)S.res_stage.p);
    hipLaunchKernelGGL(sxc_zero_i32, dim3(1), dim3(CB), 0, st, (int32_t*)S.bl_n.p, (int64_t)1);
    hipLaunchKernelGGL(sxc_bias, dim3(nbc), dim3(CB), 0, st, n1, (const int32_t*)s_p, (const int32_t*)s_t,
                       (const float*)s_usc, (const double*)S.d_eo1.p, (const int32_t*)c.d_m,
                       (const float*)S.d_thr_msv.p, c.dtarg(), s_bias, (uint8_t*)S.flag.p, (uint8_t*)S.bflag.p,
                       (int32_t*)S.bl_n.p, (int32_t*)S.bl_list.p, BIAS_LIST_CAP, (int32_t)c.bias_recheck_all);
    // 4. survivors -> Viterbi tiles
    auto compact = [&](const uint8_t* flag, const int64_t* n_in, const int32_t* src, DeviceBuf& cnt, DeviceBuf& off,
                       DeviceBuf& list, DeviceBuf& per) {
        hipLaunchKernelGGL(sxc_zero_i32, dim3(nbp), dim3(CB), 0, st, (int32_t*)per.p, P);
        hipLaunchKernelGGL(sxc_compact, dim3(nbc), dim3(CB), 0, st, 0, flag, n_in, src, (const int32_t*)s_pi,
                           (int32_t*)cnt.p, (const int64_t*)off.p, (int32_t*)list.p, (int32_t*)per.p);
        hipLaunchKernelGGL(sxc_scan, dim3(1), dim3(1024), 0, st, (const int32_t*)cnt.p, nbc, (const int64_t*)nullptr,
                           (int64_t*)off.p);
        hipLaunchKernelGGL(sxc_compact, dim3(nbc), dim3(CB), 0, st, 1, flag, n_in, src, (const int32_t*)s_pi,
                           (int32_t*)cnt.p, (const int64_t*)off.p, (int32_t*)list.p, (int32_t*)per.p);
        return (const int64_t*)off.p + nbc;
    };
    const int64_t* n2 = compact((const uint8_t*)S.flag.p, n1, nullptr, S.cnt2, S.off2, S.list2, S.per2);
    // the Viterbi stage's tiles: sized for the bound's shapes when it runs first
    tiles(S.per2, S.vbound ? S.d_ng_vq : S.d_ng_vit, S.toff2, S.eoff2, S.tiles2, n2, &S.list2, &S.tids2);
    if (with_events) HIPCHK(hipEventRecord(S.ev[2], st));
    // 5. Viterbi (after the int16 Viterbi bound and its candidate test) and its test
    const int64_t* nv = n2;
    DeviceBuf* vlist = &S.list2;
    if (S.vbound) {
        stage_runs(SX_STAGE_VIT16, S.vq_runs, S.tiles2, S.tids2, S.toff2, S.vq.p);
        if (with_events) HIPCHK(hipEventRecord(S.evvq, st));
        hipLaunchKernelGGL(sxc_vitq_cand, dim3(nbc), dim3(CB), 0, st, n2, (const int32_t*)S.list2.p,
                           (const int32_t*)S.vq.p, (const float*)s_bias, (const int32_t*)s_p, (const int32_t*)s_t,
                           (const float*)S.d_thr_vit.p, (const uint8_t*)S.d_noqv.p, (const int64_t*)S.d_capv.p,
                           (int32_t*)S.res_stage.p, (uint8_t*)S.flagq.p);
        nv = compact((const uint8_t*)S.flagq.p, n2, (const int32_t*)S.list2.p, S.cntc, S.offc, S.listc, S.perc);
        S.d_nvc = nv;
        tiles(S.perc, S.d_ng_vit, S.toffc, S.eoffc, S.tilesc, nv, &S.listc, &S.tidsc);
        stage_runs(1, S.vit_runs, S.tilesc, S.tidsc, S.toffc, S.vsc.p);
        vlist = &S.listc;
    } else {
        S.d_nvc = nullptr;
        stage_runs(1, S.vit_runs, S.tiles2, S.tids2, S.toff2, S.vsc.p);
    }
    hipLaunchKernelGGL(sxc_vit_pass, dim3(nbc), dim3(CB), 0, st, nv, (const int32_t*)vlist->p,
                       (const float*)S.vsc.p, (const float*)s_bias, (const int32_t*)s_p,
                       (const float*)S.d_thr_vit.p, (float*)S.res_vit.p, (int32_t*)S.res_stage.p,
                       (uint8_t*)S.flag3.p);
    const int64_t* n3 =
        compact((const uint8_t*)S.flag3.p, nv, (const int32_t*)vlist->p, S.cnt3, S.off3, S.list3, S.per3);
    tiles(S.per3, S.d_ng_fwd, S.toff3, S.eoff3, S.tiles3, n3, &S.list3, &S.tids3);
    if (with_events) HIPCHK(hipEventRecord(S.ev[3], st));
    // 6. Forward and its test
    stage_runs(2, S.fwd_runs, S.tiles3, S.tids3, S.toff3, S.fsc.p);
    hipLaunchKernelGGL(sxc_fwd_pass, dim3(nbc), dim3(CB), 0, st, n3, (const int32_t*)S.list3.p,
                       (const double*)S.fsc.p, (const float*)s_bias, (const int32_t*)s_p,
                       (const float*)S.d_thr_fwd.p, (double*)S.res_fwd.p, (int32_t*)S.res_stage.p);
    hipLaunchKernelGGL(sxc_records, dim3(nbc), dim3(CB), 0, st, n1, cap, (const int32_t*)s_p, (const int32_t*)s_t,
                       (const float*)s_usc, (const float*)s_bias, (const float*)S.res_vit.p,
                       (const double*)S.res_fwd.p, (const int32_t*)S.res_stage.p, (const uint8_t*)S.bflag.p,
                       S.recs.p);
    if (with_events) HIPCHK(hipEventRecord(S.ev[4], st));
    HIPCHK(hipGetLastError());
}

static float ev_ms(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    return ms;
}

// Batched form.  Returns the number of MSV survivors; records copied to out (<= cap).
static uint64_t run_batched(Ctx& c, sx_cascade_stats* st, sx_pair_scores* out, uint64_t cap) {
    if (!c.cascade) c.cascade = std::make_shared<CascadeState>();
    CascadeState& S = *c.cascade;
    const int bound = (st->bounds & SX_BOUND_MSV) ? 1 : 0, vbound = (st->bounds & SX_BOUND_VITERBI) ? 1 : 0;
    if (S.nprof != c.profs.size() || S.ntarg != c.nt || S.f1 != st->f1 || S.f2 != st->f2 || S.f3 != st->f3 ||
        S.bound != bound || S.vbound != vbound || S.thr_uncal != c.cascade_pass_uncalibrated ||
        c.cascade_dirty) {
        build_plan(c, S, st->f1, st->f2, st->f3, st->bounds);
        const int64_t P = (int64_t)c.profs.size(), T = c.nt;
        // profiles the bound skips make every pair a candidate
        int64_t nskip = 0;
        for (uint8_t v : S.noq) nskip += v;
        alloc_buffers(c, S,
                      std::max<int64_t>(1 << 16, (int64_t)((double)P * T * std::max(st->f1, 1e-4) * 2.0) + nskip * T));
        c.cascade_dirty = false;
    }
    for (int attempt = 0; attempt < 2; attempt++) {
        double total = 0.0;
        if (st->use_graph) {
            if (!S.gexec) {
                hipGraph_t graph;
                HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
                try {
                    enqueue_chain(c, S, false);
                } catch (...) {
                    hipGraph_t g2;
                    (void)hipStreamEndCapture(c.stream, &g2);
                    throw;
                }
                HIPCHK(hipStreamEndCapture(c.stream, &graph));
                HIPCHK(hipGraphInstantiate(&S.gexec, graph, nullptr, nullptr, 0));
                (void)hipGraphDestroy(graph);
            }
            HIPCHK(hipEventRecord(S.ev[0], c.stream));
            HIPCHK(hipGraphLaunch(S.gexec, c.stream));
            HIPCHK(hipEventRecord(S.ev[4], c.stream));
            HIPCHK(hipEventSynchronize(S.ev[4]));
            total = ev_ms(S.ev[0], S.ev[4]);
            st->ms_msv = st->ms_bias = st->ms_vit = st->ms_fwd = st->ms_msv_bound = st->ms_vit_bound = 0.0;
        } else {
            enqueue_chain(c, S, true);
            HIPCHK(hipEventSynchronize(S.ev[4]));
            st->ms_msv = ev_ms(S.ev[0], S.ev[1]);
            st->ms_msv_bound = S.bound ? ev_ms(S.ev[0], S.evq) : 0.0;
            st->ms_vit_bound = S.vbound ? ev_ms(S.ev[2], S.evvq) : 0.0;
            st->ms_bias = ev_ms(S.ev[1], S.ev[2]);
            st->ms_vit = ev_ms(S.ev[2], S.ev[3]);
            st->ms_fwd = ev_ms(S.ev[3], S.ev[4]);
            total = ev_ms(S.ev[0], S.ev[4]);
        }
        st->ms_total = total;
        const int64_t P = (int64_t)c.profs.size(), T = c.nt;
        const int64_t nbc = nblk(S.cap1);
        int64_t n1 = 0, n2 = 0, n3 = 0, ncand = 0;
        HIPCHK(hipMemcpy(&n1, S.d_n1, 8, hipMemcpyDeviceToHost));
        if (S.d_ncand) HIPCHK(hipMemcpy(&ncand, S.d_ncand, 8, hipMemcpyDeviceToHost));
        if (n1 > S.cap1 || ncand > S.cap1) {
            // a list overflowed its capacity: grow to the exact size and run again
            alloc_buffers(c, S, std::max(n1, ncand) + 1024);
            continue;
        }
        st->n_msv_cand = (uint64_t)ncand;
        int64_t nvc = 0;
        if (S.d_nvc) HIPCHK(hipMemcpy(&nvc, S.d_nvc, 8, hipMemcpyDeviceToHost));
        st->n_vit_cand = (uint64_t)nvc;
        st->cells_msv_exact = 0;
        std::vector<uint64_t> cand_run_cells(S.msv_runs.size(), 0);
        if (S.bound && !st->use_graph && ncand > 0) {
            // cells the f32 MSV rescored, per run (eager runs only: a copy of the list)
            std::vector<int32_t> cpi((size_t)ncand), ct((size_t)ncand);
            HIPCHK(hipMemcpy(cpi.data(), S.c_pi.p, (size_t)ncand * 4, hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(ct.data(), S.c_t.p, (size_t)ncand * 4, hipMemcpyDeviceToHost));
            size_t r = 0;
            for (int64_t i = 0; i < ncand; i++) {
                while (r + 1 < S.msv_runs.size() && cpi[i] >= S.msv_runs[r].b) r++;
                const uint64_t cells =
                    (uint64_t)c.profs[S.perm[cpi[i]]].m * (uint64_t)c.tlen[ct[i]];
                cand_run_cells[r] += cells;
                st->cells_msv_exact += cells;
            }
        }
        HIPCHK(hipMemcpy(&n2, (int64_t*)S.off2.p + nbc, 8, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(&n3, (int64_t*)S.off3.p + nbc, 8, hipMemcpyDeviceToHost));
        std::vector<sx_pair_scores>& recs = S.last;
        recs.resize((size_t)n1);
        if (n1) HIPCHK(hipMemcpy(recs.data(), S.recs.p, (size_t)n1 * sizeof(sx_pair_scores), hipMemcpyDeviceToHost));
        {
            // the bias scores the device could not certify: recomputed on the host
            const auto tr = std::chrono::steady_clock::now();
            int32_t nbl = 0;
            HIPCHK(hipMemcpy(&nbl, S.bl_n.p, 4, hipMemcpyDeviceToHost));
            std::vector<int64_t> bl;
            if (nbl <= BIAS_LIST_CAP) {
                std::vector<int32_t> b32((size_t)nbl);
                if (nbl) HIPCHK(hipMemcpy(b32.data(), S.bl_list.p, (size_t)nbl * 4, hipMemcpyDeviceToHost));
                bl.assign(b32.begin(), b32.end());
            } else {
                for (int64_t i = 0; i < n1; i++)
                    if (recs[(size_t)i].flags & SX_REC_BIAS_HOST) bl.push_back(i);
            }
            recheck_records(c, S, recs, bl, &n2, &n3);
            st->n_bias_recheck = (uint64_t)bl.size();
            st->ms_recheck = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr).count();
            st->ms_total += st->ms_recheck;
        }
        st->n_pairs = (uint64_t)(P * T);
        st->n_msv_pass = (uint64_t)n1;
        st->n_bias_pass = (uint64_t)n2;
        st->n_vit_pass = (uint64_t)n3;
        uint64_t npass = 0, cv = 0, cf = 0;
        for (const auto& r : recs) {
            const uint64_t cells = (uint64_t)c.profs[r.profile].m * (uint64_t)c.tlen[r.target];
            if (r.stage >= 2) cv += cells;
            if (r.stage >= 3) cf += cells;
            if (r.stage == 4) npass++;
        }
        st->n_fwd_pass = npass;
        st->cells_vit = cv;
        st->cells_fwd = cf;
        S.launches.clear();
        if (!st->use_graph) {
            // per-launch times between each launch's own events; the Viterbi and Forward
            // cells of each run come from the records of the pairs that reached the stage
            std::vector<int32_t> pos(c.profs.size());
            for (size_t i = 0; i < S.perm.size(); i++) pos[S.perm[i]] = (int32_t)i;
            auto run_cells = [&](const std::vector<CascadeState::Run>& runs, int min_stage) {
                std::vector<uint64_t> v(runs.size(), 0);
                for (const auto& r : recs) {
                    // min_stage 5: the records that reached Viterbi and got an f32 score
                    if (min_stage == 5 ? (r.stage < 2 || std::isnan(r.vit)) : r.stage < min_stage) continue;
                    const int64_t pi = pos[r.profile];
                    for (size_t k = 0; k < runs.size(); k++)
                        if (pi >= runs[k].a && pi < runs[k].b) {
                            v[k] += (uint64_t)c.profs[r.profile].m * (uint64_t)c.tlen[r.target];
                            break;
                        }
                }
                return v;
            };
            // the f32 Viterbi scores the bound's candidates only: the records with a score
            const std::vector<uint64_t> vc = run_cells(S.vit_runs, S.vbound ? 5 : 2), fc = run_cells(S.fwd_runs, 3);
            const std::vector<uint64_t> qc = run_cells(S.vq_runs, 2);
            size_t e = 0;
            auto add = [&](int stage, const Shape& sh, uint64_t cells) {
                sx_launch_stat l{};
                l.stage = stage;
                l.q = sh.q;
                l.g = sh.g;
                l.hoist = sh.hoist;
                l.ms = ev_ms(S.lev0[e], S.lev[e]);
                l.cells = cells;
                S.launches.push_back(l);
                e++;
            };
            for (size_t i = 0; i < S.msv_groups.size(); i++)
                add(S.bound ? SX_STAGE_MSV16 : 0, S.msv_groups[i].s, S.msv_groups[i].cells);
            for (size_t i = 0; i < S.msv_runs.size(); i++) add(0, S.msv_runs[i].s, cand_run_cells[i]);
            for (size_t i = 0; i < S.vq_runs.size(); i++) add(SX_STAGE_VIT16, S.vq_runs[i].s, qc[i]);
            for (size_t i = 0; i < S.vit_runs.size(); i++) add(1, S.vit_runs[i].s, vc[i]);
            for (size_t i = 0; i < S.fwd_runs.size(); i++) add(2, S.fwd_runs[i].s, fc[i]);
        }
        if (out) std::memcpy(out, recs.data(), std::min<uint64_t>(cap, (uint64_t)n1) * sizeof(sx_pair_scores));
        return (uint64_t)n1;
    }
    throw SxError(SX_ERR_STATE, "cascade survivor list did not fit after growing it");
}

// Plain form: one profile at a time, one launch per stage, tests on the host.
static uint64_t run_plain(Ctx& c, sx_cascade_stats* st, sx_pair_scores* out, uint64_t cap) {
    if (!c.cascade) c.cascade = std::make_shared<CascadeState>();
    CascadeState& S = *c.cascade;
    const int bound = (st->bounds & SX_BOUND_MSV) ? 1 : 0, vbound = (st->bounds & SX_BOUND_VITERBI) ? 1 : 0;
    if (S.nprof != c.profs.size() || S.ntarg != c.nt || S.f1 != st->f1 || S.f2 != st->f2 || S.f3 != st->f3 ||
        S.bound != bound || S.vbound != vbound || S.thr_uncal != c.cascade_pass_uncalibrated ||
        c.cascade_dirty) {
        build_plan(c, S, st->f1, st->f2, st->f3, st->bounds);
        alloc_buffers(c, S, 1 << 16);
        c.cascade_dirty = false;
    }
    c.ensure_pidx();
    const uint32_t T = c.nt;
    std::vector<sx_pair_scores> recs;
    std::vector<float> urow(T), vrow;
    std::vector<double> frow;
    float* d_u = (float*)c.scratch_out.ensure((size_t)T * 8);
    DeviceBuf d_pp, d_pt, d_pu, d_pb, d_fl, d_n, d_bf;
    uint64_t nrecheck = 0;
    double ms_recheck = 0.0;
    std::vector<double> eo1(25);
    double ms_total = 0.0;
    uint64_t n1 = 0, n2 = 0, n3 = 0, n4 = 0, cv = 0, cf = 0, ncand = 0, cq = 0, nvc = 0;
    // the bound's null thresholds by target id (S.nullq is in length-sorted order)
    std::vector<int64_t> nullq_t, capq_t;
    if (bound) {
        nullq_t.resize(T);
        capq_t.resize(T);
        for (uint32_t j = 0; j < T; j++) {
            nullq_t[c.sorted[j]] = S.nullq[j];
            capq_t[c.sorted[j]] = S.capq[j];
        }
    }
    std::vector<int32_t> qrow(bound ? T : 0);
    HIPCHK(hipEventRecord(S.ev[0], c.stream));
    for (size_t p = 0; p < c.profs.size(); p++) {
        const Profile& pr = c.profs[p];
        std::vector<int32_t> one{(int32_t)p};
        std::vector<uint32_t> sp_t;
        if (bound && !S.noq[p]) {
            if (S.thrq[p] == INT64_MAX / 4) continue;  // no score passes
            // the int16 bound: one launch over every target, then the f32 MSV of the
            // targets whose bound reaches the threshold (one launch for this profile)
            c.run_implicit(SX_STAGE_MSV16, SX_FORM_PLAIN, one, c.d_identity, (int)T, d_u, 0);
            HIPCHK(hipMemcpyAsync(qrow.data(), d_u, (size_t)T * 4, hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            std::vector<uint32_t> cp, ct;
            for (uint32_t t = 0; t < T; t++)
                if ((int64_t)qrow[t] >= std::min(nullq_t[t] + S.thrq[p], capq_t[t])) {
                    cp.push_back((uint32_t)p);
                    ct.push_back(t);
                }
            if (ct.empty()) continue;
            ncand += ct.size();
            for (uint32_t t : ct) cq += (uint64_t)pr.m * (uint64_t)c.tlen[t];
            c.run_pairs(0, SX_FORM_PLAIN, cp.data(), ct.data(), ct.size(), d_u, false, nullptr, nullptr, nullptr);
            std::vector<float> tmp(ct.size());
            HIPCHK(hipMemcpyAsync(tmp.data(), d_u, ct.size() * 4, hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            for (size_t i = 0; i < ct.size(); i++) urow[ct[c.last_order[i]]] = tmp[i];
            for (uint32_t t : ct) {
                const float y = (urow[t] - S.nullsc[t]) / LN2F;
                if (y >= S.thr_msv[p]) sp_t.push_back(t);
            }
        } else {
            // MSV: one launch over every target, in input order
            c.run_implicit(0, SX_FORM_PLAIN, one, c.d_identity, (int)T, d_u, 0);
            HIPCHK(hipMemcpyAsync(urow.data(), d_u, (size_t)T * 4, hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            for (uint32_t t = 0; t < T; t++) {
                const float y = (urow[t] - S.nullsc[t]) / LN2F;
                if (y >= S.thr_msv[p]) sp_t.push_back(t);
            }
        }
        if (sp_t.empty()) continue;
        n1 += sp_t.size();
        // bias filter on the device for this profile's survivors
        const int64_t k = (int64_t)sp_t.size();
        std::vector<int32_t> pp(k, (int32_t)p), pt(sp_t.begin(), sp_t.end());
        std::vector<float> pu(k);
        for (int64_t i = 0; i < k; i++) pu[i] = urow[sp_t[i]];
        int64_t kk = k;
        int32_t* dpp = dptr<int32_t>(d_pp, k);
        int32_t* dpt = dptr<int32_t>(d_pt, k);
        float* dpu = dptr<float>(d_pu, k);
        float* dpb = dptr<float>(d_pb, k);
        uint8_t* dfl = dptr<uint8_t>(d_fl, k);
        int64_t* dn = dptr<int64_t>(d_n, 1);
        HIPCHK(hipMemcpyAsync(dpp, pp.data(), k * 4, hipMemcpyHostToDevice, c.stream));
        HIPCHK(hipMemcpyAsync(dpt, pt.data(), k * 4, hipMemcpyHostToDevice, c.stream));
        HIPCHK(hipMemcpyAsync(dpu, pu.data(), k * 4, hipMemcpyHostToDevice, c.stream));
        HIPCHK(hipMemcpyAsync(dn, &kk, 8, hipMemcpyHostToDevice, c.stream));
        uint8_t* dbf = dptr<uint8_t>(d_bf, k);
        hipLaunchKernelGGL(sxc_zero_i32, dim3(1), dim3(CB), 0, c.stream, (int32_t*)S.bl_n.p, (int64_t)1);
        hipLaunchKernelGGL(sxc_bias, dim3(nblk(k)), dim3(CB), 0, c.stream, (const int64_t*)dn, (const int32_t*)dpp,
                           (const int32_t*)dpt, (const float*)dpu, (const double*)S.d_eo1.p, (const int32_t*)c.d_m,
                           (const float*)S.d_thr_msv.p, c.dtarg(), dpb, dfl, dbf, (int32_t*)S.bl_n.p,
                           (int32_t*)S.bl_list.p, BIAS_LIST_CAP, (int32_t)c.bias_recheck_all);
        std::vector<float> pb(k);
        std::vector<uint8_t> fl(k), bf(k);
        HIPCHK(hipMemcpyAsync(pb.data(), dpb, k * 4, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipMemcpyAsync(fl.data(), dfl, k, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipMemcpyAsync(bf.data(), dbf, k, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        {
            // the bias scores the device could not certify, recomputed on the host before
            // any decision uses them
            const auto tr = std::chrono::steady_clock::now();
            for (int64_t i = 0; i < k; i++) {
                if (!bf[i]) continue;
                pb[i] = bias_filter_score(pr, c.h_codes.data() + c.h_toff[sp_t[i]], c.tlen[sp_t[i]]);
                fl[i] = (pu[i] - pb[i]) / LN2F >= S.thr_msv[p] ? 1 : 0;
                nrecheck++;
            }
            ms_recheck += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr).count();
        }
        std::vector<uint32_t> vp, vt;
        std::vector<size_t> vidx;
        const size_t base = recs.size();
        for (int64_t i = 0; i < k; i++) {
            sx_pair_scores r{};
            r.profile = (uint32_t)p;
            r.target = sp_t[i];
            r.msv = pu[i];
            r.bias = pb[i];
            r.vit = NAN;
            r.fwd = NAN;
            r.fwd64 = NAN;
            r.stage = 1;
            r.flags = bf[i] ? SX_REC_BIAS_HOST : 0;
            recs.push_back(r);
            if (fl[i]) {
                vp.push_back((uint32_t)p);
                vt.push_back(sp_t[i]);
                vidx.push_back(base + i);
            }
        }
        if (vp.empty()) continue;
        n2 += vp.size();
        for (size_t i : vidx) recs[i].stage = 2;
        for (uint32_t t : vt) cv += (uint64_t)pr.m * (uint64_t)c.tlen[t];  // every pair entering Viterbi
        if (vbound && !S.noqv[p]) {
            // the int16 Viterbi bound over this profile's survivors (one launch), then the
            // f32 Viterbi on the pairs whose bound reaches the threshold
            void* dq = c.scratch_x5.ensure(vp.size() * 8);
            c.run_pairs(SX_STAGE_VIT16, SX_FORM_PLAIN, vp.data(), vt.data(), vp.size(), dq, false, nullptr, nullptr,
                        nullptr);
            std::vector<int32_t> qv(vp.size()), tmp(vp.size());
            HIPCHK(hipMemcpyAsync(tmp.data(), dq, vp.size() * 4, hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            for (size_t i = 0; i < vp.size(); i++) qv[c.last_order[i]] = tmp[i];
            const double F = SX_Q_FINE, ct = (double)LN2F * (double)S.thr_vit[p];
            std::vector<uint32_t> kp, kt;
            std::vector<size_t> kidx;
            for (size_t i = 0; i < vp.size(); i++) {
                const double b = (double)recs[vidx[i]].bias;
                const int64_t th = (int64_t)std::floor(F * (b + ct) - F * 1e-6 * (1.0 + std::fabs(b) + std::fabs(ct)));
                if ((int64_t)qv[i] >= std::min(th, S.capv[vt[i]])) {
                    kp.push_back(vp[i]);
                    kt.push_back(vt[i]);
                    kidx.push_back(vidx[i]);
                }
            }
            nvc += kp.size();
            vp.swap(kp);
            vt.swap(kt);
            vidx.swap(kidx);
            if (vp.empty()) continue;
        } else if (vbound) {
            nvc += vp.size();
        }
        // Viterbi: one launch for this profile's survivors
        void* dv = c.scratch_x5.ensure(vp.size() * 8);
        c.run_pairs(1, SX_FORM_PLAIN, vp.data(), vt.data(), vp.size(), dv, false, nullptr, nullptr, nullptr);
        vrow.resize(vp.size());
        {
            std::vector<float> tmp(vp.size());
            HIPCHK(hipMemcpyAsync(tmp.data(), dv, vp.size() * 4, hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            for (size_t i = 0; i < vp.size(); i++) vrow[c.last_order[i]] = tmp[i];
        }
        std::vector<uint32_t> fp, ft;
        std::vector<size_t> fidx;
        for (size_t i = 0; i < vp.size(); i++) {
            sx_pair_scores& r = recs[vidx[i]];
            r.vit = vrow[i];
            r.stage = 2;
            const float y = (vrow[i] - r.bias) / LN2F;
            if (y >= S.thr_vit[p]) {
                r.stage = 3;
                fp.push_back((uint32_t)p);
                ft.push_back(vt[i]);
                fidx.push_back(vidx[i]);
            }
        }
        if (fp.empty()) continue;
        n3 += fp.size();
        void* df = c.scratch_x5.ensure(fp.size() * 8);
        c.run_pairs(2, SX_FORM_PLAIN, fp.data(), ft.data(), fp.size(), df, false, nullptr, nullptr, nullptr);
        frow.resize(fp.size());
        {
            std::vector<double> tmp(fp.size());
            HIPCHK(hipMemcpyAsync(tmp.data(), df, fp.size() * 8, hipMemcpyDeviceToHost, c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            for (size_t i = 0; i < fp.size(); i++) frow[c.last_order[i]] = tmp[i];
        }
        for (size_t i = 0; i < fp.size(); i++) {
            sx_pair_scores& r = recs[fidx[i]];
            r.fwd64 = frow[i];
            r.fwd = (float)frow[i];
            cf += (uint64_t)pr.m * (uint64_t)c.tlen[ft[i]];
            const float y = (r.fwd - r.bias) / LN2F;
            if (y >= S.thr_fwd[p]) {
                r.stage = 4;
                n4++;
            }
        }
    }
    HIPCHK(hipEventRecord(S.ev[4], c.stream));
    HIPCHK(hipEventSynchronize(S.ev[4]));
    ms_total = ev_ms(S.ev[0], S.ev[4]);
    st->ms_total = ms_total;
    st->n_bias_recheck = nrecheck;
    st->ms_recheck = ms_recheck;
    st->ms_msv = st->ms_bias = st->ms_vit = st->ms_fwd = 0.0;
    st->n_pairs = (uint64_t)c.profs.size() * c.nt;
    st->n_msv_pass = n1;
    st->n_bias_pass = n2;
    st->n_vit_pass = n3;
    st->n_fwd_pass = n4;
    st->cells_vit = cv;
    st->cells_fwd = cf;
    st->n_msv_cand = ncand;
    st->cells_msv_exact = cq;
    st->ms_msv_bound = 0.0;
    st->n_vit_cand = nvc;
    st->ms_vit_bound = 0.0;
    if (out) std::memcpy(out, recs.data(), std::min<uint64_t>(cap, recs.size()) * sizeof(sx_pair_scores));
    S.last.swap(recs);
    return S.last.size();
}

}  // namespace sx

namespace sx {

std::vector<PassPair> cascade_pass_pairs(Ctx& c, double f1, double f2, double f3) {
    sx_cascade_stats st{};
    st.f1 = f1;
    st.f2 = f2;
    st.f3 = f3;
    st.use_graph = 1;
    st.bounds = SX_BOUND_MSV;  // the int16 Viterbi bound is no faster than the f32 Viterbi
    c.ensure_pidx();
    // the plan is rebuilt whenever this flag differs from the one it was built with, so
    // repeated searches reuse it (and its captured graph) and the next cascade rebuilds it
    c.cascade_pass_uncalibrated = true;
    try {
        (void)run_batched(c, &st, nullptr, 0);
    } catch (...) {
        c.cascade_pass_uncalibrated = false;
        throw;
    }
    c.cascade_pass_uncalibrated = false;
    std::vector<PassPair> out;
    for (const auto& r : c.cascade->last)
        if (r.stage == 4) out.push_back({r.profile, r.target, r.fwd});
    std::sort(out.begin(), out.end(), [](const PassPair& a, const PassPair& b) {
        return a.profile != b.profile ? a.profile < b.profile : a.target < b.target;
    });
    return out;
}

}  // namespace sx

extern "C" uint64_t sx_cascade_records_impl(sx::Ctx* c, sx_pair_scores* out, uint64_t cap) {
    if (!c->cascade) return 0;
    const std::vector<sx_pair_scores>& r = c->cascade->last;
    if (out) std::memcpy(out, r.data(), std::min<uint64_t>(cap, r.size()) * sizeof(sx_pair_scores));
    return r.size();
}

extern "C" uint32_t sx_cascade_launches_impl(sx::Ctx* c, sx_launch_stat* out, uint32_t cap) {
    if (!c->cascade) return 0;
    const std::vector<sx_launch_stat>& v = c->cascade->launches;
    if (out) std::memcpy(out, v.data(), std::min<size_t>(cap, v.size()) * sizeof(sx_launch_stat));
    return (uint32_t)v.size();
}

extern "C" void sx_cascade_records_clear(sx::Ctx* c) {
    if (c->cascade) c->cascade->last.clear();
}

extern "C" int sx_run_cascade_impl(sx::Ctx* c, int form, sx_cascade_stats* st, sx_pair_scores* out, uint64_t cap,
                                   uint64_t* n_out) {
    uint64_t cells = 0;
    for (const auto& p : c->profs) cells += (uint64_t)p.m * (uint64_t)c->total_len;
    st->cells_msv = cells;
    const uint64_t n = form == SX_FORM_PLAIN ? sx::run_plain(*c, st, out, cap) : sx::run_batched(*c, st, out, cap);
    if (n_out) *n_out = n;
    return SX_OK;
}

// sx_search.cpp -- hmmsearch-equivalent search over the device cascade (sentinel-x
// src/hmmer/search.rs:131-271): pairs that pass the filters get Forward, Backward,
// posterior decoding and region finding on the device (sxr_regions), the rest of
// score_target runs on host threads from the regions on (sx_domain.cpp), and `report`
// applies the E-value / bit / GA-NC-TC thresholds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "sx_search.h"

#include "sx_domain.h"
#include "sx_ingest.h"
#include "sx_pipeline.h"

extern "C" {
__global__ void sxr_regions(int64_t, const SxRegionIn*, const float*, const double*, const float*, const double*,
                            float*, float*, SxRegionRes*, SxRegion*, int32_t*, int32_t);
__global__ void sxc_zero_i32(int32_t*, int64_t);
__global__ void sxe_rescore(const SxEnvJob*, SxProfiles, SxTargets, float*, double*, SxEnvRes*);
__global__ void sxe_rescore_p2(const SxEnvJob*, SxProfiles, SxTargets, float*, double*, SxEnvRes*);
__global__ void sxe_rescore_p2t(const SxEnvJob*, SxProfiles, SxTargets, float*, double*, SxEnvRes*);
}

namespace sx {

namespace {

struct Threshold {
    bool by_evalue;
    double seq_e, dom_e;
    float seq_t, dom_t;
};

// thresholds (search.rs:131-148)
std::vector<Threshold> resolve_thresholds(const std::vector<Profile>& profs, const sx_search_cfg& cfg) {
    std::vector<Threshold> out;
    for (const Profile& p : profs) {
        Threshold t{};
        switch (cfg.cutoff) {
            case SX_CUT_EVALUE:
                t = {true, cfg.seq_e, cfg.dom_e, 0.0f, 0.0f};
                break;
            case SX_CUT_BITS:
                t = {false, 0.0, 0.0, cfg.seq_t, cfg.dom_t};
                break;
            case SX_CUT_GA:
            case SX_CUT_NC:
            case SX_CUT_TC: {
                const auto& c = cfg.cutoff == SX_CUT_GA ? p.ga : cfg.cutoff == SX_CUT_NC ? p.nc : p.tc;
                const char* nm = cfg.cutoff == SX_CUT_GA ? "GA" : cfg.cutoff == SX_CUT_NC ? "NC" : "TC";
                if (!c) throw SxError(SX_ERR_CUTOFF, std::string(nm) + " bit thresholds unavailable on model " + p.name);
                t = {false, 0.0, 0.0, c->first, c->second};
                break;
            }
            default:
                throw SxError(SX_ERR_ARG, "unknown cutoff kind");
        }
        out.push_back(t);
    }
    return out;
}

// body(i, ws, out) for pairs 0..n-1 on host worker threads; ok[i] = its result.
template <class Body>
void parallel_pairs(uint64_t n, int threads, std::vector<RawHit>& hit, std::vector<char>& ok, Body&& body) {
    hit.assign(n, RawHit{});
    ok.assign(n, 0);
    if (n == 0) return;
    int nth = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    nth = (int)std::min<uint64_t>((uint64_t)nth, n);
    std::atomic<uint64_t> next{0};
    std::vector<std::string> err((size_t)nth);
    auto worker = [&](int w) {
        TailWork ws;
        try {
            for (uint64_t i; (i = next.fetch_add(1)) < n;) ok[i] = body(i, ws, hit[i]);
        } catch (const std::exception& e) {
            err[(size_t)w] = e.what();
            next.store(n);
        }
    };
    std::vector<std::thread> pool;
    for (int w = 1; w < nth; w++) pool.emplace_back(worker, w);
    worker(0);
    for (auto& t : pool) t.join();
    for (const auto& e : err)
        if (!e.empty()) throw SxError(SX_ERR_STATE, "domain definition: " + e);
}

// The host part of score_target for pairs whose special rows are given: fsp / bsp hold
// 5 x (L+1) rows per pair starting at row_off[i].
void run_tails(const std::vector<Profile>& profs, const std::vector<const uint8_t*>& seq, const std::vector<int>& len,
               const uint32_t* prof, const uint32_t* tgt, uint64_t n, const double* fwd, const double* bwd,
               const float* fsp, const float* bsp, const std::vector<int64_t>& row_off, bool null2, int threads,
               std::vector<RawHit>& hit, std::vector<char>& ok) {
    parallel_pairs(n, threads, hit, ok, [&](uint64_t i, TailWork& ws, RawHit& out) {
        const int L = len[tgt[i]];
        const size_t r0 = (size_t)row_off[i] * 5, W = (size_t)L + 1;
        SpecRows f, b;
        auto load = [&](const float* src, SpecRows& d) {
            d.n.assign(src + r0, src + r0 + W);
            d.b.assign(src + r0 + W, src + r0 + 2 * W);
            d.e.assign(src + r0 + 2 * W, src + r0 + 3 * W);
            d.j.assign(src + r0 + 3 * W, src + r0 + 4 * W);
            d.c.assign(src + r0 + 4 * W, src + r0 + 5 * W);
        };
        load(fsp, f);
        load(bsp, b);
        return score_tail(profs[prof[i]], seq[tgt[i]], L, null2, (float)fwd[i], (float)bwd[i], f, b, ws, out);
    });
}

// The host part of score_target from the regions on, for pairs whose regions the device
// found.
void run_region_tails(const std::vector<Profile>& profs, const std::vector<const uint8_t*>& seq,
                      const std::vector<int>& len, const uint32_t* prof, const uint32_t* tgt, uint64_t n,
                      const float* fwd, const std::vector<TailRegions>& regs, bool null2, int threads,
                      std::vector<RawHit>& hit, std::vector<char>& ok) {
    parallel_pairs(n, threads, hit, ok, [&](uint64_t i, TailWork& ws, RawHit& out) {
        return score_regions(profs[prof[i]], seq[tgt[i]], len[tgt[i]], null2, fwd[i], regs[i], ws, out);
    });
}

// Bytes of device workspace per launch of the envelope rescoring (matrices of 16 bytes a
// cell); a single larger envelope gets a launch of its own, up to SXE_MAX_JOB_BYTES.
constexpr size_t SXE_BATCH_BYTES = (size_t)4 << 30;
constexpr size_t SXE_MAX_JOB_BYTES = (size_t)64 << 30;
// LDS a launch may take for rows and transition rows (of the 160 KB of a CU)
constexpr size_t SXE_LDS = (size_t)156 * 1024;

// Envelope rescoring on the device (sxe_rescore, one wave per envelope): jobs carry
// profile, target, rows and the null2 flag; the length model and workspace offsets are
// filled in here.  Results in job order.
//
// A launch's LDS is sized by its widest model (8, 14 or 22 floats a model position), and an
// envelope's time by its cells, so the jobs of a workspace batch are grouped by model
// width (powers of two) into launches of their own, which run side by side on the side
// streams, and each launch takes its largest envelopes first.
void device_rescore(Ctx& c, std::vector<SxEnvJob>& jobs, std::vector<SxEnvRes>& out) {
    out.assign(jobs.size(), SxEnvRes{});
    if (jobs.empty()) return;
    const size_t nj = jobs.size();
    std::vector<size_t> W(nj), R(nj);
    for (size_t b = 0; b < nj; b++) {
        SxEnvJob& J = jobs[b];
        if (J.prof < 0 || (size_t)J.prof >= c.profs.size() || J.target < 0 || (uint32_t)J.target >= c.nt)
            throw SxError(SX_ERR_ARG, "envelope names an unknown profile or target");
        const int L = c.tlen[(size_t)J.target];
        if (J.i < 1 || J.j < J.i || J.j > L) throw SxError(SX_ERR_ARG, "envelope outside its target");
        W[b] = (size_t)c.profs[(size_t)J.prof].m + 1;
        R[b] = (size_t)(J.j - J.i) + 2;
        if ((4 * R[b] * W[b] + 14 * R[b]) * 4 + 2 * R[b] * 8 > SXE_MAX_JOB_BYTES)
            throw SxError(SX_ERR_ARG, "envelope too large for device rescoring");
        const Xsc xs = Xsc::make((size_t)L, false);
        J.lloop = xs.n_loop;
        J.lmove = xs.n_move;
        J.ploop = xs.pn_loop;
        J.pmove = xs.pn_move;
        J.pad = 0;
    }
    auto bucket = [&](size_t b) {
        int e = 6;
        while (((size_t)1 << e) < W[b]) e++;
        return e;
    };
    // launch order: by width bucket, then most cells first
    std::vector<size_t> order(nj);
    for (size_t b = 0; b < nj; b++) order[b] = b;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        const int bx = bucket(x), by = bucket(y);
        if (bx != by) return bx < by;
        return R[x] * W[x] > R[y] * W[y];
    });
    std::vector<SxEnvJob> batch;
    std::vector<SxEnvRes> bres;
    size_t a = 0;
    while (a < nj) {
        size_t nf = 0, nd = 0, b = a;
        batch.clear();
        for (; b < nj; b++) {
            const size_t o = order[b];
            const size_t f = 4 * R[o] * W[o] + 14 * R[o], d = 2 * R[o];
            if (b > a && (nf + f) * 4 + (nd + d) * 8 > SXE_BATCH_BYTES) break;
            SxEnvJob J = jobs[o];
            J.wsoff = (int64_t)nf;
            J.dwsoff = (int64_t)nd;
            nf += f;
            nd += d;
            batch.push_back(J);
        }
        const size_t n = b - a;
        auto* djobs = (SxEnvJob*)c.env_jobs.ensure(n * sizeof(SxEnvJob));
        auto* dres = (SxEnvRes*)c.env_res.ensure(n * sizeof(SxEnvRes));
        float* dws = (float*)c.env_ws.ensure(nf * sizeof(float));
        double* ddws = (double*)c.env_dws.ensure(nd * sizeof(double));
        HIPCHK(hipMemcpyAsync(djobs, batch.data(), n * sizeof(SxEnvJob), hipMemcpyHostToDevice, c.stream));
        SxProfiles P = c.dprof();
        SxTargets T = c.dtarg();
        Fanout fo(c, c.stream);
        for (size_t g0 = 0; g0 < n;) {
            const int bk = bucket(order[a + g0]);
            size_t g1 = g0;
            size_t maxw = 1;
            while (g1 < n && bucket(order[a + g1]) == bk) maxw = std::max(maxw, W[order[a + g1++]]);
            // the widest form the launch's widest model lets fit: two waves with the
            // transition rows in LDS, two waves, or one wave
            const size_t p2t = (22 * maxw + 4) * sizeof(float), p2 = (14 * maxw + 4) * sizeof(float);
            const size_t one = 8 * maxw * sizeof(float);
            const void* fn;
            size_t lds;
            int threads = 128;
            if (p2t <= SXE_LDS) {
                fn = reinterpret_cast<const void*>(&sxe_rescore_p2t);
                lds = p2t;
            } else if (p2 <= SXE_LDS) {
                fn = reinterpret_cast<const void*>(&sxe_rescore_p2);
                lds = p2;
            } else {
                fn = reinterpret_cast<const void*>(&sxe_rescore);
                lds = one;
                threads = 64;
            }
            if (lds > 64 * 1024) HIPCHK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            const SxEnvJob* jp = djobs + g0;
            SxEnvRes* rp = dres + g0;
            void* args[] = {&jp, &P, &T, &dws, &ddws, &rp};
            HIPCHK(hipLaunchKernel(fn, dim3((unsigned)(g1 - g0)), dim3(threads), args, lds, fo.next()));
            g0 = g1;
        }
        fo.join();
        bres.resize(n);
        HIPCHK(hipMemcpyAsync(bres.data(), dres, n * sizeof(SxEnvRes), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        for (size_t q = 0; q < n; q++) out[order[a + q]] = bres[q];
        a = b;
    }
}

DeviceEnvelope to_host(const SxEnvRes& r) {
    DeviceEnvelope d{};
    d.envsc = r.envsc;
    d.status = r.status;
    d.oa = r.oa;
    d.ali = {r.iali, r.jali, r.kali, r.mali};
    for (int x = 0; x < K; x++) d.null2raw[x] = r.null2[x];
    return d;
}

// The host part of score_target from the regions on, with envelope rescoring on the
// device: multi-domain regions are resolved on host threads (stochastic traces and
// clustering), then every envelope (the single-domain regions and the ensembles'
// envelopes) is rescored on the device in one batch, and the host finishes each pair
// (null2 correction, Viterbi-trace fallback, scores).  *ms_dev = the device part.
void run_region_tails_device(Ctx& c, const std::vector<Profile>& profs,
                             const std::vector<const uint8_t*>& seq, const std::vector<int>& len, const uint32_t* prof,
                             const uint32_t* tgt, uint64_t n, const float* fwd, const std::vector<TailRegions>& regs,
                             bool null2, int threads, std::vector<RawHit>& hit, std::vector<char>& ok, double* ms_dev,
                             uint64_t* n_env) {
    // per pair: null2 corrections and the ensembles' envelopes of its multi-domain regions
    std::vector<std::vector<float>> n2sc(n);
    std::vector<std::vector<std::vector<std::pair<int, int>>>> ens(n);
    std::vector<uint64_t> multi;
    for (uint64_t i = 0; i < n; i++) {
        for (const RegionSpan& g : regs[i].spans) {
            const int L = len[tgt[i]];
            if (g.i < 1 || g.j < g.i || g.j > L) throw SxError(SX_ERR_STATE, "region outside the target");
            if (g.multi) {
                multi.push_back(i);
                break;
            }
        }
    }
    {
        std::vector<RawHit> unused_h;
        std::vector<char> unused_ok;
        parallel_pairs(multi.size(), threads, unused_h, unused_ok, [&](uint64_t m, TailWork& ws, RawHit&) {
            const uint64_t i = multi[m];
            const int L = len[tgt[i]];
            n2sc[i].assign((size_t)L + 1, 0.0f);
            for (const RegionSpan& g : regs[i].spans)
                ens[i].push_back(g.multi ? region_ensemble(profs[prof[i]], seq[tgt[i]], g.i, g.j, L, ws, n2sc[i])
                                         : std::vector<std::pair<int, int>>{});
            return true;
        });
    }
    // the envelopes in (pair, region, envelope) order
    std::vector<SxEnvJob> jobs;
    std::vector<uint64_t> first(n + 1, 0);
    for (uint64_t i = 0; i < n; i++) {
        first[i] = jobs.size();
        const int L = len[tgt[i]];
        for (size_t r = 0; r < regs[i].spans.size(); r++) {
            const RegionSpan& g = regs[i].spans[r];
            if (!g.multi) {
                jobs.push_back({(int32_t)prof[i], (int32_t)tgt[i], g.i, g.j, null2 ? 1 : 0});
            } else {
                for (const auto& ij : ens[i][r])
                    if (ij.first >= 1 && ij.second >= ij.first && ij.second <= L)
                        jobs.push_back({(int32_t)prof[i], (int32_t)tgt[i], ij.first, ij.second, 0});
            }
        }
    }
    first[n] = jobs.size();
    std::vector<SxEnvRes> er;
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    device_rescore(c, jobs, er);
    *ms_dev += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    *n_env += jobs.size();
    parallel_pairs(n, threads, hit, ok, [&](uint64_t i, TailWork&, RawHit& out) {
        const Profile& p = profs[prof[i]];
        const uint8_t* dsq = seq[tgt[i]];
        const int L = len[tgt[i]];
        if (L == 0 || p.m == 0 || !std::isfinite(fwd[i])) return false;
        std::vector<float>& n2 = n2sc[i];
        if (n2.empty()) n2.assign((size_t)L + 1, 0.0f);
        std::vector<Envelope> envs;
        uint64_t q = first[i];
        for (size_t r = 0; r < regs[i].spans.size(); r++) {
            const RegionSpan& g = regs[i].spans[r];
            Envelope env;
            if (!g.multi) {
                if (envelope_from_device(p, dsq, g.i, g.j, L, null2, false, to_host(er[q++]), n2, env))
                    envs.push_back(env);
            } else {
                for (const auto& ij : ens[i][r]) {
                    if (!(ij.first >= 1 && ij.second >= ij.first && ij.second <= L)) continue;
                    if (envelope_from_device(p, dsq, ij.first, ij.second, L, null2, true, to_host(er[q++]), n2, env))
                        envs.push_back(env);
                }
            }
        }
        if (q != first[i + 1]) throw std::runtime_error("envelope bookkeeping mismatch");
        return score_envelopes(p, L, null2, fwd[i], regs[i].nexpected, (int)regs[i].spans.size(), n2, envs, out);
    });
}

// Device buffers of the region scan, kept across the chunks of one search.
struct RegionBufs {
    DeviceBuf bx5, bls, in, bt, et, res, regs, cnt;
};

// Forward and Backward of n pairs into device buffers (rows at roff[i], rows in all),
// then posterior decoding and region finding on the device (sxr_regions).  fwd_in (may be
// null) is the Forward score the tail goes on with.  Fills regs[i] and, from the device's
// special rows, fsc[i] = the Forward score; returns the bytes copied to the host.
uint64_t device_regions(Ctx& c, RegionBufs& B, const uint32_t* prof, const uint32_t* tgt, uint64_t n,
                        const std::vector<int64_t>& roff, int64_t rows, const float* fwd_in,
                        std::vector<TailRegions>& regs, std::vector<float>& fsc) {
    regs.assign(n, TailRegions{});
    fsc.assign(n, NEG_INF);
    if (n == 0) return 0;
    float* fx5 = (float*)c.scratch_x5.ensure((size_t)rows * 5 * sizeof(float));
    double* fls = (double*)c.scratch_ls.ensure((size_t)rows * sizeof(double));
    float* bx5 = (float*)B.bx5.ensure((size_t)rows * 5 * sizeof(float));
    double* bls = (double*)B.bls.ensure((size_t)rows * sizeof(double));
    void* d = c.scratch_out.ensure(n * 8);
    c.run_pairs(SX_STAGE_FORWARD, SX_FORM_BATCHED, prof, tgt, n, d, true, &roff, fx5, fls);
    c.run_pairs(SX_STAGE_BACKWARD, SX_FORM_BATCHED, prof, tgt, n, d, true, &roff, bx5, bls);
    std::vector<SxRegionIn> in(n);
    for (uint64_t i = 0; i < n; i++) {
        const int L = c.tlen[tgt[i]];
        const Xsc xs = Xsc::make((size_t)L, true);
        in[i] = {roff[i], L, xs.n_loop, xs.n_move, xs.pc_move, fwd_in ? fwd_in[i] : 0.0f, fwd_in ? 1 : 0};
    }
    // a region takes at least two rows (the row that triggers it and the row that closes
    // it), so a pair of length L has at most L/2
    const int64_t cap = rows / 2 + (int64_t)n;
    if (cap > INT32_MAX || n > INT32_MAX) throw SxError(SX_ERR_ARG, "too many special rows in one region scan");
    auto* din = (SxRegionIn*)B.in.ensure(n * sizeof(SxRegionIn));
    auto* dres = (SxRegionRes*)B.res.ensure(n * sizeof(SxRegionRes));
    auto* dreg = (SxRegion*)B.regs.ensure((size_t)cap * sizeof(SxRegion));
    auto* dcnt = (int32_t*)B.cnt.ensure(sizeof(int32_t));
    float* dbt = (float*)B.bt.ensure((size_t)rows * sizeof(float));
    float* det = (float*)B.et.ensure((size_t)rows * sizeof(float));
    HIPCHK(hipMemcpyAsync(din, in.data(), n * sizeof(SxRegionIn), hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(sxc_zero_i32, dim3(1), dim3(256), 0, c.stream, dcnt, (int64_t)1);
    // one wave per pair
    hipLaunchKernelGGL(sxr_regions, dim3((unsigned)n), dim3(64), 0, c.stream, (int64_t)n, (const SxRegionIn*)din,
                       (const float*)fx5, (const double*)fls, (const float*)bx5, (const double*)bls, dbt, det, dres,
                       dreg, dcnt, (int32_t)cap);
    HIPCHK(hipGetLastError());
    std::vector<SxRegionRes> res(n);
    int32_t cnt = 0;
    HIPCHK(hipMemcpyAsync(res.data(), dres, n * sizeof(SxRegionRes), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipMemcpyAsync(&cnt, dcnt, sizeof(int32_t), hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    if (cnt < 0 || cnt > cap) throw SxError(SX_ERR_STATE, "region list overflow");
    std::vector<SxRegion> list((size_t)cnt);
    if (cnt > 0) {
        HIPCHK(hipMemcpyAsync(list.data(), dreg, (size_t)cnt * sizeof(SxRegion), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    }
    for (const SxRegion& g : list) {
        if (g.pair < 0 || (uint64_t)g.pair >= n) throw SxError(SX_ERR_STATE, "region of an unknown pair");
        regs[(size_t)g.pair].spans.push_back({g.i, g.j, g.multi != 0});
    }
    for (uint64_t i = 0; i < n; i++) {
        auto& sp = regs[i].spans;
        // one thread writes a pair's regions in order; the list keeps them by slot
        std::sort(sp.begin(), sp.end(), [](const RegionSpan& a, const RegionSpan& b) { return a.j < b.j; });
        if ((int32_t)sp.size() != res[i].nreg) throw SxError(SX_ERR_STATE, "region count mismatch");
        regs[i].nexpected = res[i].nexp;
        fsc[i] = res[i].fsc;
    }
    return n * sizeof(SxRegionRes) + sizeof(int32_t) + (uint64_t)cnt * sizeof(SxRegion);
}

// report (search.rs:160-210) for one profile; raws in target order
void report(uint32_t pi, const Profile& p, const std::vector<std::pair<uint32_t, const RawHit*>>& raws,
            const std::vector<int>& len, const Threshold& th, double z, sx_results& res) {
    const size_t first_hit = res.hits.size();
    for (const auto& tr : raws) {
        const RawHit& r = *tr.second;
        const double evalue = std::exp(r.lnp) * z;
        const bool keep = th.by_evalue ? evalue <= th.seq_e : r.bits >= th.seq_t;
        if (!keep) continue;
        sx_hit h{};
        h.profile = pi;
        h.target = tr.first;
        h.qlen = (uint32_t)p.m;
        h.tlen = (uint32_t)len[tr.first];
        h.score = r.bits;
        h.bias = r.bias;
        h.evalue = evalue;
        h.nexpected = r.nexpected;
        h.ndom = (uint32_t)r.domains.size();
        h.dom_first = res.doms.size();
        for (const RawDomain& d : r.domains) {
            sx_domain o{};
            o.score = d.bits;
            o.bias = d.bias;
            o.c_evalue = std::numeric_limits<double>::quiet_NaN();
            o.i_evalue = std::exp(d.lnp) * z;
            o.hmm_from = (uint32_t)d.kali;
            o.hmm_to = (uint32_t)d.mali;
            o.ali_from = (uint32_t)d.iali;
            o.ali_to = (uint32_t)d.jali;
            o.env_from = (uint32_t)d.ienv;
            o.env_to = (uint32_t)d.jenv;
            o.reported = 0;
            res.doms.push_back(o);
        }
        res.hits.push_back(h);
    }
    const double dom_z = (double)std::max<size_t>(res.hits.size() - first_hit, 1);
    for (size_t hi = first_hit; hi < res.hits.size(); hi++) {
        const sx_hit& h = res.hits[hi];
        for (uint64_t di = h.dom_first; di < h.dom_first + h.ndom; di++) {
            sx_domain& d = res.doms[di];
            d.c_evalue = d.i_evalue / z * dom_z;
            d.reported = th.by_evalue ? d.c_evalue <= th.dom_e : d.score >= th.dom_t;
        }
    }
}

// Tails and reports for sorted pairs whose scores and special rows are given.
void tails_and_report(const std::vector<Profile>& profs, const std::vector<int>& len,
                      const std::vector<Threshold>& ths, double z, const std::vector<uint32_t>& pp, const std::vector<uint32_t>& pt,
                      const std::vector<RawHit>& hit, const std::vector<char>& ok, sx_results& res) {
    size_t i = 0;
    while (i < pp.size()) {
        const uint32_t p = pp[i];
        std::vector<std::pair<uint32_t, const RawHit*>> raws;
        for (; i < pp.size() && pp[i] == p; i++)
            if (ok[i]) raws.push_back({pt[i], &hit[i]});
        report(p, profs[p], raws, len, ths[p], z, res);
    }
}

}  // namespace

void check_search_cfg(const sx_search_cfg* cfg) {
    if (!cfg) throw SxError(SX_ERR_ARG, "search configuration missing");
    if (cfg->cutoff < SX_CUT_EVALUE || cfg->cutoff > SX_CUT_TC) throw SxError(SX_ERR_ARG, "unknown cutoff kind");
}

// The device search: cascade, Forward / Backward and region finding on the device in
// chunks, host tails from the regions on, report.
void search(Ctx& c, const sx_search_cfg& cfg, sx_results& res) {
    using clk = std::chrono::steady_clock;
    auto ms_since = [](clk::time_point t0) {
        return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
    };
    check_search_cfg(&cfg);
    const std::vector<Threshold> ths = resolve_thresholds(c.profs, cfg);
    const double z = std::max(cfg.z > 0.0 ? cfg.z : (double)c.nt, 1.0);
    res.models.clear();
    for (const Profile& p : c.profs) res.models.push_back({p.name, p.acc, p.m});
    if (c.profs.empty() || c.nt == 0) return;
    // with the filters on, the tail goes on with the Forward score that decided the F3
    // test (score_target computes Forward once, pipeline.rs:180); with them off, with the
    // score of the special-row Forward
    std::vector<PassPair> pairs;
    auto t0 = clk::now();
    if (cfg.filters) {
        pairs = cascade_pass_pairs(c, cfg.f1, cfg.f2, cfg.f3);
        res.stats.ms_filters = ms_since(t0);
    } else {
        pairs.reserve((size_t)c.profs.size() * c.nt);
        for (uint32_t p = 0; p < c.profs.size(); p++)
            for (uint32_t t = 0; t < c.nt; t++) pairs.push_back({p, t, NAN});
    }
    std::vector<const uint8_t*> seq(c.nt);
    for (uint32_t t = 0; t < c.nt; t++) seq[t] = c.h_codes.data() + c.h_toff[t];
    const std::vector<int>& len = c.tlen;
    // chunks of at most 16M special rows per direction (320 MB of floats each, on the
    // device)
    const int64_t max_rows = 16LL << 20;
    size_t a = 0;
    std::vector<uint32_t> cp, ct;
    std::vector<float> fin, fsc;
    std::vector<int64_t> roff;
    std::vector<TailRegions> regs;
    std::vector<RawHit> hit;
    std::vector<char> ok;
    std::vector<uint32_t> all_p, all_t;
    std::vector<RawHit> all_hit;
    std::vector<char> all_ok;
    RegionBufs bufs;
    while (a < pairs.size()) {
        cp.clear();
        ct.clear();
        roff.clear();
        fin.clear();
        int64_t rows = 0;
        size_t b = a;
        for (; b < pairs.size(); b++) {
            const int64_t r = (int64_t)len[pairs[b].target] + 1;
            if (b > a && rows + r > max_rows) break;
            cp.push_back(pairs[b].profile);
            ct.push_back(pairs[b].target);
            fin.push_back(pairs[b].fwd);
            roff.push_back(rows);
            rows += r;
        }
        const uint64_t n = cp.size();
        t0 = clk::now();
        res.stats.bytes_to_host += device_regions(c, bufs, cp.data(), ct.data(), n, roff, rows,
                                                  cfg.filters ? fin.data() : nullptr, regs, fsc);
        res.stats.ms_specials += ms_since(t0);
        t0 = clk::now();
        const float* fw = cfg.filters ? fin.data() : fsc.data();
        if (cfg.host_rescore) {
            run_region_tails(c.profs, seq, len, cp.data(), ct.data(), n, fw, regs, cfg.null2 != 0, cfg.threads, hit,
                             ok);
        } else {
            double ms_dev = 0.0;
            run_region_tails_device(c, c.profs, seq, len, cp.data(), ct.data(), n, fw, regs, cfg.null2 != 0,
                                    cfg.threads, hit, ok, &ms_dev, &res.stats.n_envelopes);
            res.stats.ms_rescore += ms_dev;
            res.stats.ms_tail -= ms_dev;
        }
        res.stats.ms_tail += ms_since(t0);
        res.stats.n_pairs += n;
        res.stats.n_rows += (uint64_t)rows;
        for (const TailRegions& r : regs) res.stats.n_regions += r.spans.size();
        all_p.insert(all_p.end(), cp.begin(), cp.end());
        all_t.insert(all_t.end(), ct.begin(), ct.end());
        all_hit.insert(all_hit.end(), std::make_move_iterator(hit.begin()), std::make_move_iterator(hit.end()));
        all_ok.insert(all_ok.end(), ok.begin(), ok.end());
        a = b;
    }
    t0 = clk::now();
    tails_and_report(c.profs, len, ths, z, all_p, all_t, all_hit, all_ok, res);
    res.stats.ms_report = ms_since(t0);
}

void rescore_envelopes(Ctx& c, const sx_envelope* env, uint64_t n, bool on_host, sx_envelope_score* out) {
    for (uint64_t e = 0; e < n; e++) {
        const sx_envelope& v = env[e];
        if (v.profile >= c.profs.size() || v.target >= c.nt)
            throw SxError(SX_ERR_ARG, "envelope names an unknown profile or target");
        if (v.i < 1 || v.j < v.i || v.j > c.tlen[v.target]) throw SxError(SX_ERR_ARG, "envelope outside its target");
    }
    auto put = [&](uint64_t e, const DeviceEnvelope& d, float bcksc) {
        sx_envelope_score& o = out[e];
        o = sx_envelope_score{};
        o.envsc = d.envsc;
        o.bcksc = bcksc;
        o.status = d.status;
        o.oa = d.status == 0 ? d.oa : 0;
        if (o.oa) {
            o.iali = env[e].i - 1 + d.ali.iali;
            o.jali = env[e].i - 1 + d.ali.jali;
            o.kali = d.ali.kali;
            o.mali = d.ali.mali;
        }
        if (d.status == 0 && env[e].null2)
            for (int x = 0; x < K; x++) o.null2[x] = d.null2raw[x];
    };
    if (on_host) {
        std::vector<RawHit> h;
        std::vector<char> ok;
        parallel_pairs(n, 0, h, ok, [&](uint64_t e, TailWork& ws, RawHit&) {
            const sx_envelope& v = env[e];
            float b = NEG_INF;
            const DeviceEnvelope d = rescore_raw(c.profs[v.profile], c.h_codes.data() + c.h_toff[v.target], v.i, v.j,
                                                 c.tlen[v.target], v.null2 != 0, ws, &b);
            put(e, d, b);
            return true;
        });
        return;
    }
    std::vector<SxEnvJob> jobs(n);
    for (uint64_t e = 0; e < n; e++)
        jobs[e] = {(int32_t)env[e].profile, (int32_t)env[e].target, env[e].i, env[e].j, env[e].null2 ? 1 : 0};
    std::vector<SxEnvRes> er;
    device_rescore(c, jobs, er);
    for (uint64_t e = 0; e < n; e++) put(e, to_host(er[e]), er[e].bcksc);
}

void search_hmm_file(Ctx& c, const std::string& path, const sx_search_cfg& cfg, uint32_t batch, sx_results& res) {
    using clk = std::chrono::steady_clock;
    check_search_cfg(&cfg);
    // the reference's batch is 256 models (search.rs:257); a device launch over so few
    // profiles leaves most of the chip idle, so the default batch is larger (the results
    // do not depend on it)
    if (batch == 0) batch = 2048;
    const int threads = cfg.threads > 0 ? cfg.threads : (int)std::max(1u, std::thread::hardware_concurrency());
    // two batches parsed ahead of the one on the device
    auto stream = ProfileStream::open(path, threads, (size_t)batch * 2);
    res = sx_results{};
    std::vector<Profile> b;
    uint32_t base = 0;
    for (;;) {
        b.clear();
        const auto t0 = clk::now();
        if (stream->next(batch, b) == 0) break;
        res.stats.ms_ingest += std::chrono::duration<double, std::milli>(clk::now() - t0).count();
        const uint32_t nb = (uint32_t)b.size();
        c.set_profiles(std::move(b));
        sx_results r;
        search(c, cfg, r);
        const uint64_t dbase = res.doms.size();
        for (sx_hit h : r.hits) {
            h.profile += base;
            h.dom_first += dbase;
            res.hits.push_back(h);
        }
        res.doms.insert(res.doms.end(), r.doms.begin(), r.doms.end());
        res.models.insert(res.models.end(), r.models.begin(), r.models.end());
        sx_search_stats& s = res.stats;
        s.n_pairs += r.stats.n_pairs;
        s.n_rows += r.stats.n_rows;
        s.ms_filters += r.stats.ms_filters;
        s.ms_specials += r.stats.ms_specials;
        s.ms_tail += r.stats.ms_tail;
        s.ms_report += r.stats.ms_report;
        s.n_regions += r.stats.n_regions;
        s.ms_rescore += r.stats.ms_rescore;
        s.n_envelopes += r.stats.n_envelopes;
        s.bytes_to_host += r.stats.bytes_to_host;
        base += nb;
    }
    res.stats.n_models = base;
    res.stats.ms_ingest_wait = stream->ms_wait();
}

void search_host(const sx_hmm* hmms, uint32_t np, const uint8_t* codes, const uint64_t* offsets, uint32_t nt,
                 const sx_search_cfg& cfg, uint64_t npairs, const uint32_t* prof, const uint32_t* tgt,
                 const double* fwd, const double* bwd, const float* fsp, const float* bsp, sx_results& res) {
    check_search_cfg(&cfg);
    std::vector<Profile> profs;
    profs.reserve(np);
    for (uint32_t i = 0; i < np; i++) profs.push_back(configure(hmm_from_c(hmms[i])));
    const std::vector<Threshold> ths = resolve_thresholds(profs, cfg);
    const double z = std::max(cfg.z > 0.0 ? cfg.z : (double)nt, 1.0);
    std::vector<const uint8_t*> seq(nt);
    std::vector<int> len(nt);
    for (uint32_t t = 0; t < nt; t++) {
        if (offsets[t + 1] < offsets[t]) throw SxError(SX_ERR_ARG, "target offsets must be non-decreasing");
        seq[t] = codes + offsets[t];
        len[t] = (int)(offsets[t + 1] - offsets[t]);
        for (int k = 0; k < len[t]; k++)
            if (seq[t][k] >= NCODE) throw SxError(SX_ERR_ARG, "residue code out of range (expected 0..24)");
    }
    // pairs in (profile, target) order, as the reference regroups them
    std::vector<uint64_t> order(npairs);
    for (uint64_t i = 0; i < npairs; i++) {
        if (prof[i] >= np || tgt[i] >= nt) throw SxError(SX_ERR_ARG, "pair index out of range");
        order[i] = i;
    }
    std::stable_sort(order.begin(), order.end(), [&](uint64_t x, uint64_t y) {
        return prof[x] != prof[y] ? prof[x] < prof[y] : tgt[x] < tgt[y];
    });
    std::vector<int64_t> roff_in(npairs);
    int64_t rows = 0;
    for (uint64_t i = 0; i < npairs; i++) {
        roff_in[i] = rows;
        rows += (int64_t)len[tgt[i]] + 1;
    }
    std::vector<uint32_t> sp(npairs), st(npairs);
    std::vector<double> sf(npairs), sb(npairs);
    std::vector<int64_t> roff(npairs);
    for (uint64_t i = 0; i < npairs; i++) {
        const uint64_t o = order[i];
        sp[i] = prof[o];
        st[i] = tgt[o];
        sf[i] = fwd[o];
        sb[i] = bwd[o];
        roff[i] = roff_in[o];
    }
    std::vector<RawHit> hit;
    std::vector<char> ok;
    run_tails(profs, seq, len, sp.data(), st.data(), npairs, sf.data(), sb.data(), fsp, bsp, roff, cfg.null2 != 0,
              cfg.threads, hit, ok);
    tails_and_report(profs, len, ths, z, sp, st, hit, ok, res);
}

}  // namespace sx

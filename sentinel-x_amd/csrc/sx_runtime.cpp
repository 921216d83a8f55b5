// sx_runtime.cpp -- device context, data upload, launch planning (plain and batched
// forms) and the C ABI declared in include/sx_hmmer.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/sx_hmmer.h"
#include "sx_dev.h"
#include "sx_model.h"
#include "sx_runtime.h"
#include "sx_pipeline.h"
#include "sx_ingest.h"
#include "sx_search.h"

struct SxKernelEntry {
    int stage, q, g, spec, hoist, wid, block, lds;
    const void* fn;
};
extern "C" const SxKernelEntry* sx_kernel_table(int* n);

namespace sx {

static const size_t LDS_MAX = 160 * 1024;

template <class T>
static void dfree(T*& p) {
    if (p) (void)hipFree((void*)p);
    p = nullptr;
}

template <class T>
static T* dupload(const std::vector<T>& v) {
    T* d = nullptr;
    const size_t bytes = std::max<size_t>(v.size() * sizeof(T), 16);
    HIPCHK(hipMalloc((void**)&d, bytes));
    if (!v.empty()) HIPCHK(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

DeviceBuf::~DeviceBuf() { release(); }
void DeviceBuf::release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
}
void* DeviceBuf::ensure(size_t n) {
    if (n <= bytes) return p;
    release();
    const size_t nb = std::max<size_t>(n, 256);
    HIPCHK(hipMalloc(&p, nb));
    bytes = nb;
    return p;
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
    int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = 0ull;
}

// The device memory of sx_score_pairs_noalloc: queue heads, the pair list's target ids,
// its tiles and the results, for at most SX_NOALLOC_MAX_PAIRS pairs.  Each device holds
// its own copy until the process ends; host code reaches it only through the address
// hipGetSymbolAddress gives for the device (noalloc_scratch), never through the name.
constexpr size_t NA_HEADS = 0, NA_TIDS = 1024;
constexpr size_t NA_TILES = NA_TIDS + (size_t)SX_NOALLOC_MAX_PAIRS * sizeof(int32_t);
constexpr size_t NA_OUT = NA_TILES + (size_t)SX_NOALLOC_MAX_PAIRS * sizeof(SxTile);
constexpr size_t NA_BYTES = NA_OUT + (size_t)SX_NOALLOC_MAX_PAIRS * sizeof(double);
static_assert(NA_TILES % 256 == 0 && NA_OUT % 256 == 0, "scratch regions stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// The scratch array's address on device dev (which must be current), looked up once per
// device.
static void* noalloc_scratch(int dev) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it != addr.end()) return it->second;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        throw SxError(SX_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    addr[dev] = ptr;
    return ptr;
}

Ctx::Ctx(int dev) : device(dev) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw SxError(SX_ERR_NODEVICE, "no HIP device available");
    if (dev < 0 || dev >= n) throw SxError(SX_ERR_NODEVICE, "HIP device index out of range");
    HIPCHK(hipSetDevice(dev));
    HIPCHK(hipStreamCreate(&stream));
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    for (int i = 0; i < NSIDE; i++) {
        HIPCHK(hipStreamCreate(&side[i]));
        HIPCHK(hipEventCreateWithFlags(&join_ev[i], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    n_cu = prop.multiProcessorCount;
    HIPCHK(hipMalloc((void**)&d_heads, sizeof(unsigned long long) * 64));
    // test setting: every bias score of the cascade goes through the host recomputation
    if (const char* e = std::getenv("SX_BIAS_RECHECK_ALL")) bias_recheck_all = e[0] == '1';
    int nk = 0;
    const SxKernelEntry* tab = sx_kernel_table(&nk);
    for (int i = 0; i < nk; i++) {
        KernelInfo ki;
        ki.stage = tab[i].stage;
        ki.q = tab[i].q;
        ki.g = tab[i].g;
        ki.spec = tab[i].spec;
        ki.hoist = tab[i].hoist;
        ki.wid = tab[i].wid;
        ki.block = tab[i].block;
        ki.lds = tab[i].lds;
        ki.fn = tab[i].fn;
        ki.index = i;
        // every (stage, Q, G, special rows, table mode) names exactly one kernel
        if (!kernels.emplace(key(ki.stage, ki.q, ki.g, ki.spec, ki.hoist), ki).second)
            throw SxError(SX_ERR_STATE, "kernel table has two entries for stage " + std::to_string(ki.stage) +
                                            " Q=" + std::to_string(ki.q) + " G=" + std::to_string(ki.g));
    }
    // the allocation-free pair scoring: the device's scratch array and the host staging
    na_dev = noalloc_scratch(dev);
    na_order.resize(SX_NOALLOC_MAX_PAIRS);
    na_tids.resize(SX_NOALLOC_MAX_PAIRS);
    na_tiles.resize(SX_NOALLOC_MAX_PAIRS);
    na_res.resize(SX_NOALLOC_MAX_PAIRS);
    na_cnt.assign((size_t)nk, 0);
    na_kern.assign((size_t)nk, nullptr);
    for (const auto& kv : kernels) na_kern[(size_t)kv.second.index] = &kv.second;
}

Ctx::~Ctx() {
    (void)hipSetDevice(device);
    free_profiles();
    free_targets();
    dfree(d_heads);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    for (int i = 0; i < NSIDE; i++) {
        if (side[i]) (void)hipStreamSynchronize(side[i]);
        if (side[i]) (void)hipStreamDestroy(side[i]);
        if (join_ev[i]) (void)hipEventDestroy(join_ev[i]);
    }
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (stream) (void)hipStreamDestroy(stream);
}

void Ctx::free_profiles() {
    dfree(d_msc); dfree(d_mo); dfree(d_tsc); dfree(d_tp);
    dfree(d_off); dfree(d_toff); dfree(d_m); dfree(d_tbm); dfree(d_vmo);
    dfree(d_qsc); dfree(d_qtbm); dfree(d_qtsc);
    dfree(d_gmsc); dfree(d_gmo); dfree(d_goff);
}

void Ctx::free_targets() {
    dfree(d_codes); dfree(d_soff); dfree(d_len); dfree(d_lm); dfree(d_sorted); dfree(d_identity);
    dfree(d_ilv); dfree(d_ilvoff);
}

const KernelInfo& Ctx::kernel(int stage, int q, int g, bool spec, int hoist) const {
    auto it = kernels.find(key(stage, q, g, spec ? 1 : 0, hoist));
    if (it == kernels.end())
        throw SxError(SX_ERR_STATE, "no kernel for stage " + std::to_string(stage) + " Q=" + std::to_string(q) +
                                        " G=" + std::to_string(g) +
                                        (hoist == 2   ? " (global emissions)"
                                         : hoist >= 3 ? " (aligned lane blocks)"
                                         : hoist      ? " (registers)"
                                                      : " (LDS)"));
    return it->second;
}

// ---------------------------------------------------------------------------
// data upload
// ---------------------------------------------------------------------------

void Ctx::set_targets(const uint8_t* codes, const uint64_t* offsets, uint32_t n) {
    cascade_dirty = true;
    free_targets();
    nt = n;
    tlen.assign(n, 0);
    std::vector<int64_t> off(n);
    std::vector<SxLen> lm(n);
    size_t pos = 64;  // front pad: the reverse residue stream may read one chunk before a target
    for (uint32_t i = 0; i < n; i++) {
        if (offsets[i + 1] < offsets[i]) throw SxError(SX_ERR_ARG, "target offsets must be non-decreasing");
        const uint64_t L = offsets[i + 1] - offsets[i];
        if (L > (uint64_t)INT32_MAX / 2) throw SxError(SX_ERR_ARG, "target too long");
        tlen[i] = (int32_t)L;
        off[i] = (int64_t)pos;
        pos += (L + 15) & ~(size_t)15;
        SxLen& s = lm[i];
        msv_length(L, &s.msv_tloop, &s.msv_tmove);
        const Xsc x = Xsc::make(L, true);
        s.lloop = x.n_loop;
        s.lmove = x.n_move;
        s.ploop = x.pn_loop;
        s.pmove = x.pn_move;
        s.q_tloop = q_const(s.msv_tloop);
        s.q_tmove = q_const(s.msv_tmove);
    }
    pos += 64;  // back pad: prefetch of the chunk after a target's last one
    std::vector<uint8_t> packed(pos, 0);
    for (uint32_t i = 0; i < n; i++) {
        for (int32_t k = 0; k < tlen[i]; k++) {
            const uint8_t c = codes[offsets[i] + k];
            if (c >= SX_NCODE) throw SxError(SX_ERR_ARG, "residue code out of range (expected 0..24)");
            packed[off[i] + k] = c;
        }
    }
    h_toff = off;
    d_codes = dupload(packed);
    h_codes.swap(packed);  // host copy, read by the host-side domain definition
    d_soff = dupload(off);
    d_len = dupload(tlen);
    d_lm = dupload(lm);
    // processing order of the batched form: longest targets first (stable), so that a
    // persistent launch ends on its shortest tiles and drains with a short tail
    sorted.resize(n);
    std::iota(sorted.begin(), sorted.end(), 0);
    std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) { return tlen[a] > tlen[b]; });
    d_sorted = dupload(sorted);
    // the interleaved copy of the residues over the sorted order (SxTargets::ilv)
    {
        const size_t nc = ((size_t)n + 63) / 64;
        std::vector<int64_t> coff(nc + 1, 0);
        for (size_t c = 0; c < nc; c++) {
            const int64_t lc = tlen[sorted[c * 64]];  // the chunk's longest target
            coff[c + 1] = coff[c] + ((lc + 3) / 4 + 4) * 64;
        }
        std::vector<uint32_t> ilv((size_t)coff[nc] + 64, 0u);
        for (size_t c = 0; c < nc; c++) {
            for (int j = 0; j < 64 && c * 64 + j < n; j++) {
                const int t = sorted[c * 64 + j];
                const uint8_t* src = h_codes.data() + h_toff[t];
                for (int32_t k = 0; 4 * k < tlen[t]; k++) {
                    uint32_t wv = 0;
                    for (int b = 0; b < 4 && 4 * k + b < tlen[t]; b++) wv |= (uint32_t)src[4 * k + b] << (8 * b);
                    ilv[(size_t)coff[c] + (size_t)k * 64 + j] = wv;
                }
            }
        }
        d_ilv = dupload(ilv);
        d_ilvoff = dupload(coff);
    }
    std::vector<int32_t> ident(n);
    std::iota(ident.begin(), ident.end(), 0);
    d_identity = dupload(ident);
    total_len = 0;
    for (auto L : tlen) total_len += L;
}

void Ctx::set_profiles(std::vector<Profile>&& ps) {
    for (const auto& p : ps)
        if (p.m > SX_MAX_M)
            throw SxError(SX_ERR_ARG, "model '" + p.name + "' has M = " + std::to_string(p.m) +
                                          "; the longest supported model has " + std::to_string(SX_MAX_M) +
                                          " positions");
    cascade_dirty = true;
    profiles_version++;
    free_profiles();
    profs = std::move(ps);
    std::vector<float> msc, mo, tsc, tp, tbm;
    std::vector<int64_t> off, toff;
    std::vector<int32_t> ms, vmo, qtbm;
    std::vector<int16_t> qsc, qtsc;
    for (const auto& p : profs) {
        // D(k) = max(M(k-1) + tMD, D(k-1) + tDD) with D(1) = -inf: when no tMD, tDD is
        // positive, every D value is at most an earlier M value of the same row (f32
        // rounding is monotone and x + t <= x for t <= 0), so max over M and D of a row
        // equals max over M.  NaN fails the test and keeps D in the maximum.
        bool m_only = true;
        for (int k = 0; k <= p.m; k++) {
            const float md = p.t_row(SX_TMD)[k], dd = p.t_row(SX_TDD)[k];
            if (!(md <= 0.0f) || !(dd <= 0.0f)) m_only = false;
        }
        vmo.push_back(m_only ? 1 : 0);
        off.push_back((int64_t)msc.size());
        toff.push_back((int64_t)tsc.size());
        ms.push_back(p.m);
        tbm.push_back(p.msv_tbm);
        qtbm.push_back(q_const(p.msv_tbm));
        for (float v : p.msc) qsc.push_back(q_emit(v));
        for (int r = 0; r < SX_NT; r++)
            for (int k = 0; k <= p.m; k++) qtsc.push_back(k == 0 ? (int16_t)-32768 : q_trans(p.t_row(r)[k]));
        msc.insert(msc.end(), p.msc.begin(), p.msc.end());
        mo.insert(mo.end(), p.mo.begin(), p.mo.end());
        tsc.insert(tsc.end(), p.tsc.begin(), p.tsc.end());
        tp.insert(tp.end(), p.tp.begin(), p.tp.end());
    }
    d_msc = dupload(msc);
    d_mo = dupload(mo);
    d_tsc = dupload(tsc);
    d_tp = dupload(tp);
    d_off = dupload(off);
    d_toff = dupload(toff);
    d_m = dupload(ms);
    d_tbm = dupload(tbm);
    d_vmo = dupload(vmo);
    d_qsc = dupload(qsc);
    d_qtbm = dupload(qtbm);
    d_qtsc = dupload(qtsc);
    // padded emission rows of the models over 512 positions (read from global memory by
    // the global-emission kernels)
    std::vector<float> gmsc, gmo;
    std::vector<int64_t> goff(profs.size(), -1);
    for (size_t i = 0; i < profs.size(); i++) {
        const Profile& p = profs[i];
        if (p.m <= SX_GEM_COPY_MIN_M) continue;
        goff[i] = (int64_t)gmsc.size();
        const size_t base = gmsc.size(), w = (size_t)p.m + 1;
        gmsc.resize(base + (size_t)SX_NCODE * SX_GEM_WIDTH, -std::numeric_limits<float>::infinity());
        gmo.resize(base + (size_t)SX_NCODE * SX_GEM_WIDTH, 0.0f);
        for (int x = 0; x < SX_NCODE; x++)
            for (int k = 1; k <= p.m; k++) {
                gmsc[base + (size_t)x * SX_GEM_WIDTH + k + 3] = p.msc[x * w + k];
                gmo[base + (size_t)x * SX_GEM_WIDTH + k + 3] = p.mo[x * w + k];
            }
    }
    if (!gmsc.empty()) {
        d_gmsc = dupload(gmsc);
        d_gmo = dupload(gmo);
    }
    d_goff = dupload(goff);
}

SxProfiles Ctx::dprof() const {
    SxProfiles P;
    P.msc = d_msc; P.mo = d_mo; P.tsc = d_tsc; P.tp = d_tp;
    P.off = d_off; P.toff = d_toff; P.m = d_m; P.msv_tbm = d_tbm; P.vit_m_only = d_vmo;
    P.gmsc = d_gmsc; P.gmo = d_gmo; P.goff = d_goff; P.gw = SX_GEM_WIDTH; P.q_tej = q_const(-LN2F);
    P.qsc = d_qsc; P.qtbm = d_qtbm; P.qtsc = d_qtsc;
    return P;
}

SxTargets Ctx::dtarg() const {
    SxTargets T;
    T.codes = d_codes; T.off = d_soff; T.len = d_len; T.lm = d_lm;
    T.ilv = d_ilv; T.ilv_off = d_ilvoff;
    return T;
}

// ---------------------------------------------------------------------------
// shapes
// ---------------------------------------------------------------------------

// Built-in lane-group shape for a model length when no tuned table is installed: the
// smallest capacity Q*G >= M among shapes whose LDS tables fit, preferring (at equal
// capacity) register-resident transitions and then fewer lanes per pair.  Viterbi,
// Forward and Backward avoid Q > 16 per lane, whose three register rows cut occupancy.
// MSV models longer than 128 take 8-lane (M <= 320) or 16-lane groups over aligned lane
// blocks with 512-thread workgroups, which measured fastest from M = 160 to 1024.
Shape Ctx::gem_shape(int m) {
    for (int q : {24, 32, 40, 48, 64})
        if (q * 64 >= m) return {q, 64, 2};
    throw SxError(SX_ERR_ARG, "model length " + std::to_string(m) + " exceeds the longest supported (" +
                                  std::to_string(SX_MAX_M) + ")");
}

Shape Ctx::default_shape(int stage, int form, int m) const {
    if (stage == SX_STAGE_VIT16) {
        // the quantised Viterbi bound: one lane per pair up to 32 positions, then 8-, 16-
        // and 32-lane groups over aligned lane blocks
        if (m > SX_Q_MAX_M)
            throw SxError(SX_ERR_ARG, "model length " + std::to_string(m) + " exceeds the quantised Viterbi bound's " +
                                          std::to_string(SX_Q_MAX_M));
        if (m <= 16) return {16, 1, 0};
        if (m <= 32) return {32, 1, 0};
        for (int q : {16, 24, 32})
            if (8 * q >= m) return {q, 8, 3};
        for (int q : {24, 32})
            if (16 * q >= m) return {q, 16, 3};
        return {32, 32, 3};
    }
    if (stage == SX_STAGE_MSV16) {
        // the quantised MSV bound: one lane per pair up to 256 positions, then 8- and
        // 16-lane groups over aligned lane blocks
        if (m > SX_Q_MAX_M)
            throw SxError(SX_ERR_ARG, "model length " + std::to_string(m) + " exceeds the quantised MSV bound's " +
                                          std::to_string(SX_Q_MAX_M));
        for (int q : {16, 24, 32, 48, 64, 80, 96, 112, 128, 144, 160, 176, 192, 208, 224, 240, 256})
            if (q >= m) return {q, 1, 0};
        for (int q : {40, 48, 56, 64, 72, 80, 88})
            if (8 * q >= m) return {q, 8, 3};
        for (int q : {48, 56, 64})
            if (16 * q >= m) return {q, 16, 3};
    }
    if (m > SX_GEM_MIN_M) return gem_shape(m);
    if (form == SX_FORM_PLAIN) {
        for (int q : {4, 8, 16})
            if (q * 64 >= m) return {q, 64, stage != 0 && q <= 8 ? 1 : 0};
    }
    if (stage == 0 && m > 128) {
        // MSV past the one-lane shapes: 8- then 16-lane groups over aligned lane blocks
        for (int q : {20, 24, 32, 40})
            if (8 * q >= m && kernels.count(key(0, q, 8, 0, 3))) return {q, 8, 3};
        for (int q : {24, 32, 40, 48, 64})
            if (16 * q >= m && kernels.count(key(0, q, 16, 0, 3))) return {q, 16, 3};
    }
    Shape best{0, 0, 0};
    long best_cap = 1L << 40;
    for (const auto& kv : kernels) {
        const KernelInfo& k = kv.second;
        if (k.stage != stage || k.spec || k.hoist >= 2) continue;
        const long cap = (long)k.q * k.g;
        if (cap < m) continue;
        if (lds_bytes(k) > LDS_MAX) continue;
        if (stage != 0 && k.q > 16 && cap < 1024) continue;
        const bool better = cap < best_cap ||
                            (cap == best_cap && (k.hoist > best.hoist || (k.hoist == best.hoist && k.g < best.g)));
        if (better) {
            best_cap = cap;
            best = {k.q, k.g, k.hoist};
        }
    }
    if (!best.q) throw SxError(SX_ERR_ARG, "no kernel shape covers model length " + std::to_string(m));
    return best;
}

Shape Ctx::shape_for(int stage, int form, int m) const {
    const bool quant = stage == SX_STAGE_MSV16 || stage == SX_STAGE_VIT16;
    if (quant && m > SX_Q_MAX_M) return default_shape(stage, form, m);  // throws
    if (!quant && m > SX_GEM_MIN_M) return gem_shape(m);  // long models: one fixed shape
    if (form == SX_FORM_BATCHED) {
        auto gem_ok = [&](const Shape& s) {
            // a global-emission shape reads the padded copy, which only longer models have
            if (s.hoist == 2 && m <= SX_GEM_COPY_MIN_M)
                throw SxError(SX_ERR_ARG, "a global-emission shape (table mode 2) covers models of over " +
                                              std::to_string(SX_GEM_COPY_MIN_M) + " positions only");
            return s;
        };
        if (forced.count(stage)) return gem_ok(forced.at(stage));
        auto it = tuned.find(stage);
        if (it != tuned.end()) {
            for (const auto& e : it->second)
                if (m <= e.first) return gem_ok(e.second);
        }
    }
    return default_shape(stage, form, m);
}

Shape Ctx::shape_for_spec(int stage, int m) const {
    if (m > SX_GEM_MIN_M) return gem_shape(m);
    // special-row writers: Q in {4, 8} with register-resident transitions, {16, 32, 48} in LDS.
    // They run on the few pairs that pass the filters, whose longest pairs set the time, so
    // the shape takes the fewest positions a lane (the shortest row per lane) that 64 lanes
    // cover, then the fewest lanes of that Q.
    for (int q : {4, 8, 16, 32, 48})
        for (int g : {1, 2, 4, 8, 16, 32, 64}) {
            if ((long)q * g < m) continue;
            if (g == 1 && q > 16) continue;
            auto it = kernels.find(key(stage, q, g, 1, q <= 8 ? 1 : 0));
            if (it == kernels.end()) continue;
            if (lds_bytes(it->second) > LDS_MAX) continue;
            return {q, g, q <= 8 ? 1 : 0};
        }
    throw SxError(SX_ERR_ARG, "no special-row kernel covers model length " + std::to_string(m));
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------

int Ctx::persistent_grid(const KernelInfo& k, size_t lds, int64_t ntiles) const {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k.fn, k.block, lds) != hipSuccess || per_cu < 1)
        per_cu = 1;
    int64_t g = (int64_t)per_cu * n_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(g, ntiles));
}

Fanout::Fanout(Ctx& c_, hipStream_t o) : c(c_), origin(o) { HIPCHK(hipEventRecord(c.fork_ev, origin)); }

hipStream_t Fanout::next() { return side(n++); }

hipStream_t Fanout::pick(bool exclusive, int* slot) {
    constexpr int R = 64 / Ctx::NSIDE;  // slots per stream
    int i;
    if (exclusive) {
        i = 0;
        *slot = Ctx::NSIDE * (n_excl++ % R);
    } else {
        i = 1 + n_other % (Ctx::NSIDE - 1);
        *slot = i + Ctx::NSIDE * ((n_other++ / (Ctx::NSIDE - 1)) % R);
    }
    return side(i);
}

hipStream_t Fanout::side(int k) {
    static_assert(Ctx::NSIDE == sizeof(used) / sizeof(used[0]), "one flag per side stream");
    const int i = k % Ctx::NSIDE;
    if (!used[i]) {
        HIPCHK(hipStreamWaitEvent(c.side[i], c.fork_ev, 0));
        used[i] = true;
    }
    return c.side[i];
}

Fanout::~Fanout() {
    for (int i = 0; i < Ctx::NSIDE; i++)
        if (used[i]) {
            (void)hipEventRecord(c.join_ev[i], c.side[i]);
            (void)hipStreamWaitEvent(origin, c.join_ev[i], 0);
        }
}

void Fanout::join() {
    for (int i = 0; i < Ctx::NSIDE; i++)
        if (used[i]) {
            HIPCHK(hipEventRecord(c.join_ev[i], c.side[i]));
            HIPCHK(hipStreamWaitEvent(origin, c.join_ev[i], 0));
            used[i] = false;
        }
}

void Ctx::launch(const KernelInfo& k, int stage, SxWork w, void* out, SxSpecialsOut sp, int grid, bool persistent,
                 int head_slot, hipStream_t s, unsigned long long* heads) {
    if (!s) s = stream;
    const size_t lds = lds_bytes(k);
    if (lds > LDS_MAX) throw SxError(SX_ERR_ARG, "profile tables exceed the LDS budget");
    if (lds > 64 * 1024) HIPCHK(hipFuncSetAttribute(k.fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (persistent) {
        w.head = (heads ? heads : d_heads) + head_slot;
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, w.head, 1);
        HIPCHK(hipGetLastError());
    } else {
        w.head = nullptr;
    }
    SxProfiles P = dprof();
    SxTargets T = dtarg();
    void* args[] = {&P, &T, &w, &out, &sp};
    HIPCHK(hipLaunchKernel(k.fn, dim3(grid), dim3(k.block), args, lds, s));
}

// Implicit all-pairs launches for a set of profiles against a target-id list (device
// array of ntids ids).  Results go to out[profile * out_stride + target].
void Ctx::run_implicit(int stage, int form, const std::vector<int32_t>& plist, const int32_t* dtids, int ntids,
                       void* dout, int64_t out_stride) {
    SxSpecialsOut sp{};
    if (form == SX_FORM_PLAIN) {
        // one launch per profile; each workgroup takes block/G pairs in input order
        for (int32_t p : plist) {
            const Shape s = shape_for(stage, form, profs[p].m);
            const KernelInfo& k = kernel(stage, s.q, s.g, false, s.hoist);
            SxWork w{};
            w.profs = d_pidx + p;
            w.tids = dtids;
            w.chunk = k.block / s.g;
            w.ntids = ntids;
            w.nchunks = (ntids + w.chunk - 1) / w.chunk;
            w.ntiles = w.nchunks;
            w.out_stride = out_stride;
            if (w.ntiles > 0) launch(k, stage, w, dout, sp, (int)w.ntiles, false, 0);
        }
        return;
    }
    // batched: profiles grouped by shape, one persistent launch per shape over
    // length-sorted target chunks
    std::map<std::tuple<int, int, int>, std::vector<int32_t>> groups;
    for (int32_t p : plist) {
        const Shape s = shape_for(stage, form, profs[p].m);
        groups[{s.q, s.g, s.hoist}].push_back(p);
    }
    // every group's profile list goes up in one copy, which completes before the host
    // lists go out of scope (an async copy reads its host source when it executes)
    std::vector<int32_t> flat;
    std::vector<size_t> goff;
    for (auto& kv : groups) {
        goff.push_back(flat.size());
        flat.insert(flat.end(), kv.second.begin(), kv.second.end());
    }
    pgroups_keep.emplace_back();
    DeviceBuf& pb = pgroups_keep.back();
    pb.ensure(flat.size() * sizeof(int32_t));
    HIPCHK(hipMemcpyAsync(pb.p, flat.data(), flat.size() * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    size_t gi = 0;
    // The shape groups side by side, except the long-model groups (tables from global
    // memory, transition records filling 48-111 KB of LDS): those go one after another
    // (measured on 200 models of 1,500-2,500 positions x 400 proteins of 8,000-35,000
    // residues: the 32- and 40-position lane groups 2.79 s side by side, 2.21 s one after
    // the other).
    Fanout fo(*this, stream);
    for (auto& kv : groups) {
        const int q = std::get<0>(kv.first), g = std::get<1>(kv.first), h = std::get<2>(kv.first);
        const KernelInfo& k = kernel(stage, q, g, false, h);
        SxWork w{};
        w.profs = (const int32_t*)pb.p + goff[gi++];
        w.tids = dtids;
        w.chunk = k.block / g;
        w.ilv = dtids == d_sorted && ntids == (int)nt && w.chunk % 64 == 0 ? 1 : 0;
        w.ntids = ntids;
        w.nchunks = (ntids + w.chunk - 1) / w.chunk;
        w.ntiles = (int64_t)w.nchunks * (int64_t)kv.second.size();
        w.out_stride = out_stride;
        if (w.ntiles == 0) continue;
        const int grid = persistent_grid(k, lds_bytes(k), w.ntiles);
        // A workgroup restages the profile tables when its next tile is another profile's.
        // With fewer tiles per profile than workgroups that happens on nearly every tile
        // anyway, so Viterbi, Forward and Backward then take the targets longest first
        // across all profiles (the launch does not end on one profile's longest targets);
        // otherwise, and for the MSV stages, a profile's tiles stay together.
        w.chunk_major = stage != SX_STAGE_MSV && stage != SX_STAGE_MSV16 && w.nchunks < grid ? 1 : 0;
        int slot;
        hipStream_t ls = fo.pick(h == 2, &slot);
        launch(k, stage, w, dout, sp, grid, true, slot, ls);
    }
    fo.join();
}

// All pairs of every profile against every target.  out is a device buffer of P*T
// elements indexed profile*T + target.
void Ctx::run_all(int stage, int form, void* dout) {
    std::vector<int32_t> plist(profs.size());
    std::iota(plist.begin(), plist.end(), 0);
    while (pgroups_keep.size() > 256) pgroups_keep.pop_front();
    run_implicit(stage, form, plist, form == SX_FORM_PLAIN ? d_identity : d_sorted, (int)nt, dout, nt);
}

// Explicit pair list: pairs grouped by profile (and, for the batched form, ordered by
// target length inside a profile) into tiles of one lane group per pair.
void Ctx::run_pairs(int stage, int form, const uint32_t* prof, const uint32_t* tgt, uint64_t n, void* dout,
                    bool spec, const std::vector<int64_t>* row_off, float* dx5, double* dls) {
    std::vector<uint64_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    for (uint64_t i = 0; i < n; i++) {
        if (prof[i] >= profs.size() || tgt[i] >= nt) throw SxError(SX_ERR_ARG, "pair index out of range");
    }
    if (form == SX_FORM_BATCHED)
        std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) {
            if (prof[a] != prof[b]) return prof[a] < prof[b];
            return tlen[tgt[a]] > tlen[tgt[b]];
        });
    else
        std::stable_sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return prof[a] < prof[b]; });
    std::map<std::tuple<int, int, int>, std::vector<SxTile>> tiles;
    std::vector<int32_t> tids(n);
    std::vector<int64_t> roff;
    for (uint64_t i = 0; i < n; i++) tids[i] = (int32_t)tgt[order[i]];
    uint64_t i = 0;
    while (i < n) {
        const uint32_t p = prof[order[i]];
        uint64_t e = i;
        while (e < n && prof[order[e]] == p) e++;
        const Shape s = spec ? shape_for_spec(stage, profs[p].m) : shape_for(stage, form, profs[p].m);
        const int ng = kernel(stage, s.q, s.g, spec, s.hoist).block / s.g;
        for (uint64_t a = i; a < e; a += ng) {
            SxTile t;
            t.prof = (int32_t)p;
            t.count = (int32_t)std::min<uint64_t>(ng, e - a);
            t.first = (int64_t)a;
            t.out = (int64_t)a;
            tiles[{s.q, s.g, s.hoist}].push_back(t);
        }
        i = e;
    }
    scratch_tids.ensure(n * sizeof(int32_t));
    HIPCHK(hipMemcpyAsync(scratch_tids.p, tids.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    SxSpecialsOut sp{};
    if (spec) {
        roff.resize(n);
        for (uint64_t a = 0; a < n; a++) roff[a] = (*row_off)[order[a]];
        scratch_roff.ensure(n * sizeof(int64_t));
        HIPCHK(hipMemcpyAsync(scratch_roff.p, roff.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, stream));
        sp.x5 = dx5;
        sp.ls = dls;
        sp.row_off = (const int64_t*)scratch_roff.p;
    }
    // results come back in the sorted order; remember it for the caller
    last_order = order;
    std::vector<SxTile> flat;
    for (auto& kv : tiles) flat.insert(flat.end(), kv.second.begin(), kv.second.end());
    scratch_tiles.ensure(flat.size() * sizeof(SxTile));
    HIPCHK(hipMemcpyAsync(scratch_tiles.p, flat.data(), flat.size() * sizeof(SxTile), hipMemcpyHostToDevice, stream));
    // the host arrays above die with this call: the copies complete before it returns
    HIPCHK(hipStreamSynchronize(stream));
    size_t toff = 0;
    Fanout fo(*this, stream);  // batched: the shape groups side by side, long-model groups in turn
    for (auto& kv : tiles) {
        const int q = std::get<0>(kv.first), g = std::get<1>(kv.first);
        const int h = std::get<2>(kv.first);
        const KernelInfo& k = kernel(stage, q, g, spec, h);
        SxTile* dt = (SxTile*)scratch_tiles.p + toff;
        toff += kv.second.size();
        SxWork w{};
        w.tiles = dt;
        w.tids = (const int32_t*)scratch_tids.p;
        w.ntiles = (int64_t)kv.second.size();
        if (form == SX_FORM_BATCHED) {
            const int grid = persistent_grid(k, lds_bytes(k), w.ntiles);
            int slot;
            hipStream_t ls = fo.pick(h == 2, &slot);
            launch(k, stage, w, dout, sp, grid, true, slot, ls);
        } else {
            launch(k, stage, w, dout, sp, (int)w.ntiles, false, 0);
        }
    }
    fo.join();
}

// run_pairs (scores only) with no allocation: the pairs are ordered, tiled and grouped by
// shape in host staging reserved at creation (std::sort on indices with the index as the
// last key gives the stable order of run_pairs, and a counting pass over the kernel table
// groups the tiles), and every device buffer is a region of the scratch array.  The tiles
// and results therefore match run_pairs's, pair for pair.
void Ctx::score_pairs_noalloc(int stage, int form, const uint32_t* prof, const uint32_t* tgt, uint64_t n, void* out,
                              double* ms) {
    if (n > SX_NOALLOC_MAX_PAIRS) throw SxError(SX_ERR_ARG, "pair count exceeds SX_NOALLOC_MAX_PAIRS");
    for (uint64_t i = 0; i < n; i++)
        if (prof[i] >= profs.size() || tgt[i] >= nt) throw SxError(SX_ERR_ARG, "pair index out of range");
    if (n == 0) {
        if (ms) *ms = 0.0;
        return;
    }
    uint32_t* order = na_order.data();
    for (uint32_t i = 0; i < (uint32_t)n; i++) order[i] = i;
    if (form == SX_FORM_BATCHED)
        std::sort(order, order + n, [&](uint32_t a, uint32_t b) {
            if (prof[a] != prof[b]) return prof[a] < prof[b];
            if (tlen[tgt[a]] != tlen[tgt[b]]) return tlen[tgt[a]] > tlen[tgt[b]];
            return a < b;
        });
    else
        std::sort(order, order + n, [&](uint32_t a, uint32_t b) { return prof[a] != prof[b] ? prof[a] < prof[b] : a < b; });
    for (uint64_t i = 0; i < n; i++) na_tids[i] = (int32_t)tgt[order[i]];
    // tiles per kernel, then their offsets, then the tiles grouped by kernel
    std::fill(na_cnt.begin(), na_cnt.end(), 0);
    auto each_run = [&](auto&& f) {
        for (uint64_t i = 0, e; i < n; i = e) {
            const uint32_t p = prof[order[i]];
            for (e = i; e < n && prof[order[e]] == p; e++) {
            }
            const Shape s = shape_for(stage, form, profs[p].m);
            f(kernel(stage, s.q, s.g, false, s.hoist), p, i, e);
        }
    };
    each_run([&](const KernelInfo& k, uint32_t, uint64_t i, uint64_t e) {
        const uint64_t ng = (uint64_t)(k.block / k.g);
        na_cnt[(size_t)k.index] += (int64_t)((e - i + ng - 1) / ng);
    });
    int64_t tot = 0;
    for (int64_t& v : na_cnt) {
        const int64_t c = v;
        v = tot;
        tot += c;
    }
    each_run([&](const KernelInfo& k, uint32_t p, uint64_t i, uint64_t e) {
        const uint64_t ng = (uint64_t)(k.block / k.g);
        for (uint64_t a = i; a < e; a += ng) {
            SxTile& t = na_tiles[(size_t)na_cnt[(size_t)k.index]++];
            t.prof = (int32_t)p;
            t.count = (int32_t)std::min<uint64_t>(ng, e - a);
            t.first = (int64_t)a;
            t.out = (int64_t)a;
        }
    });
    // na_cnt[k] now ends kernel k's tiles
    uint8_t* base = (uint8_t*)na_dev;
    unsigned long long* heads = (unsigned long long*)(base + NA_HEADS);
    int32_t* d_tids = (int32_t*)(base + NA_TIDS);
    SxTile* d_tiles = (SxTile*)(base + NA_TILES);
    void* d_out = base + NA_OUT;
    HIPCHK(hipMemcpyAsync(d_tids, na_tids.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_tiles, na_tiles.data(), (size_t)tot * sizeof(SxTile), hipMemcpyHostToDevice, stream));
    HIPCHK(hipEventRecord(ev0, stream));
    int slot = 0;
    {
        Fanout fo(*this, stream);  // batched: the shape groups side by side
        int64_t start = 0;
        for (size_t ki = 0; ki < na_cnt.size(); ki++) {
            const int64_t end = na_cnt[ki], cnt = end - start;
            if (cnt > 0) {
                const KernelInfo& k = *na_kern[ki];
                SxWork w{};
                w.tiles = d_tiles + start;
                w.tids = d_tids;
                w.ntiles = cnt;
                SxSpecialsOut sp{};
                if (form == SX_FORM_BATCHED) {
                    const int grid = persistent_grid(k, lds_bytes(k), w.ntiles);
                    launch(k, stage, w, d_out, sp, grid, true, slot++ % 64, fo.next(), heads);
                } else {
                    launch(k, stage, w, d_out, sp, (int)w.ntiles, false, 0);
                }
            }
            start = end;
        }
        fo.join();
    }
    HIPCHK(hipEventRecord(ev1, stream));
    const size_t es = stage == SX_STAGE_FORWARD || stage == SX_STAGE_BACKWARD ? 8 : 4;
    HIPCHK(hipMemcpyAsync(na_res.data(), d_out, n * es, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (ms) {
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, ev0, ev1));
        *ms = t;
    }
    if (out) {
        const uint8_t* r = (const uint8_t*)na_res.data();
        for (uint64_t i = 0; i < n; i++) std::memcpy((uint8_t*)out + (size_t)order[i] * es, r + i * es, es);
    }
}

// ---------------------------------------------------------------------------
// autotuner: for each model-length bucket present in the profile set, time every
// kernel shape that covers the bucket's longest model on a sample of that bucket's
// profiles against a length-stratified sample of the targets, and keep the fastest.
// ---------------------------------------------------------------------------
std::vector<TuneRecord> Ctx::autotune(int stage, double budget_ms) {
    static const std::vector<int> coarse = {8,   16,  24,  32,  48,   64,   96,   128,  160,  192,  256, 320,
                                            384, 512, 640, 768, 1024, 1280, 1536, 2048, 2560, 3072, 4096};
    // the MSV bound and Viterbi have finer menus of capacities (their time follows the
    // capacity)
    static const std::vector<int> fine = {16,  24,  32,  48,  64,  80,  96,  112, 128, 144, 160, 176, 192, 208, 224, 240,
                                          256, 288, 320, 352, 384, 416, 448, 512, 576, 640, 704, 768, 896, 1024};
    const std::vector<int>& edges = stage == SX_STAGE_MSV16 || stage == SX_STAGE_VITERBI ? fine : coarse;
    std::vector<TuneRecord> log;
    if (profs.empty() || nt == 0) return log;
    ensure_pidx();
    // target sample: every k-th id of the length-sorted order (all of them up to 2^20):
    // a small sample would leave the persistent grid under-filled for shapes with many
    // pairs per tile and misjudge them
    const int nts = (int)std::min<uint32_t>(nt, 1u << 20);
    std::vector<int32_t> ts(nts);
    int64_t sum_l = 0;
    for (int i = 0; i < nts; i++) {
        ts[i] = sorted[(size_t)((double)i * nt / nts)];
        sum_l += tlen[ts[i]];
    }
    DeviceBuf dts;
    dts.ensure(ts.size() * sizeof(int32_t));
    HIPCHK(hipMemcpy(dts.p, ts.data(), ts.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    DeviceBuf dout;
    dout.ensure((size_t)nt * 8);
    std::vector<std::pair<int, Shape>> table;
    // buckets present in this profile set share the budget equally
    std::vector<int> present;
    {
        int lo0 = 0;
        for (int hi : edges) {
            for (const auto& p : profs)
                if (p.m > lo0 && p.m <= hi) {
                    present.push_back(hi);
                    break;
                }
            lo0 = hi;
        }
    }
    const double per_bucket = budget_ms / std::max<size_t>(1, present.size());
    int lo = 0;
    for (int hi : edges) {
        const int blo = lo;  // the bucket holds models of blo+1 .. hi positions
        std::vector<int32_t> inb;
        int mmax = 0;
        for (size_t p = 0; p < profs.size(); p++)
            if (profs[p].m > lo && profs[p].m <= hi) {
                inb.push_back((int32_t)p);
                mmax = std::max(mmax, profs[p].m);
            }
        lo = hi;
        if (inb.empty()) continue;
        const bool quant = stage == SX_STAGE_MSV16 || stage == SX_STAGE_VIT16;
        if (mmax > (quant ? SX_Q_MAX_M : SX_GEM_MIN_M)) break;  // no tuned shape beyond
        // profiles spread over the bucket, repeated if the bucket is small: about 4e10
        // cells (up to 16 profiles) for the later stages; three times that (up to 64) for
        // MSV, whose one-lane shapes run a whole pair per lane, so that a tile lasts long
        // enough for the persistent grid's drain tail to swing a short sample's time
        std::vector<int32_t> sample;
        const bool msv_like = stage == 0 || stage == SX_STAGE_MSV16;
        const double cells_want = msv_like ? 1.2e11 : 4e10;
        const int want = (int)std::clamp<int64_t>((int64_t)(cells_want / ((double)mmax * (double)sum_l)), 1,
                                                  msv_like ? 64 : 16);
        for (int i = 0; i < want; i++) sample.push_back(inb[(size_t)((double)i * inb.size() / want)]);
        int64_t cells = 0;
        for (int32_t p : sample) cells += (int64_t)profs[p].m * sum_l;
        // the same profiles twice: twice the cells with the same mix of lengths
        std::vector<int32_t> sample2(sample);
        sample2.insert(sample2.end(), sample.begin(), sample.end());
        // candidates: every shape covering the bucket's longest model with at most twice its
        // capacity (at least the three tightest), tightest first, then fewer lanes per pair
        std::vector<Shape> cands;
        for (const auto& kv : kernels) {
            const KernelInfo& k = kv.second;
            if (k.stage != stage || k.spec) continue;
            // global-emission shapes: Viterbi and Forward buckets whose models all have the
            // padded copy
            if (k.hoist == 2 && !((stage == SX_STAGE_VITERBI || stage == SX_STAGE_FORWARD) &&
                                  blo >= SX_GEM_COPY_MIN_M && k.q <= 16))
                continue;
            if ((long)k.q * k.g < mmax) continue;
            // an MSV shape with an aligned-block twin: the twin reads LDS without bank
            // conflicts and is always the faster of the two
            if (stage == 0 && k.hoist == 0 && kernels.count(key(0, k.q, k.g, 0, 3))) continue;
            if (lds_bytes(k) > LDS_MAX) continue;
            cands.push_back({k.q, k.g, k.hoist});
        }
        std::sort(cands.begin(), cands.end(), [](const Shape& a, const Shape& b) {
            const long ca = (long)a.q * a.g, cb = (long)b.q * b.g;
            if (ca != cb) return ca < cb;
            if (a.g != b.g) return a.g < b.g;
            return a.hoist > b.hoist;
        });
        {
            std::vector<Shape> keep;
            for (size_t i = 0; i < cands.size(); i++)
                if (i < 3 || (long)cands[i].q * cands[i].g <= 2L * std::max(mmax, 16)) keep.push_back(cands[i]);
            cands.swap(keep);
        }
        Shape best = default_shape(stage, SX_FORM_BATCHED, mmax);
        // the built-in choice is timed first: when the budget runs out on a bucket whose
        // tightest shapes are slow ones, the table still holds a measured good shape
        {
            auto it = std::find_if(cands.begin(), cands.end(), [&](const Shape& x) {
                return x.q == best.q && x.g == best.g && x.hoist == best.hoist;
            });
            if (it != cands.end()) std::rotate(cands.begin(), it, it + 1);
        }
        double best_ms = 1e300, spent = 0.0;
        for (const Shape& s : cands) {
            if (spent > per_bucket && best_ms < 1e300) break;
            forced[stage] = s;
            try {
                const double t0 = timed([&] {  // warm-up
                    run_implicit(stage, SX_FORM_BATCHED, sample, (const int32_t*)dts.p, nts, dout.p, 0);
                });
                spent += t0;
                if (t0 > 3.0 * best_ms) {
                    // far slower than the best so far even counting first-launch costs:
                    // logged from the warm-up alone, not timed again
                    log.push_back({hi, s, t0, cells / t0 / 1e6});
                    forced.erase(stage);
                    continue;
                }
                const double t1 = timed([&] {
                    run_implicit(stage, SX_FORM_BATCHED, sample, (const int32_t*)dts.p, nts, dout.p, 0);
                });
                double ms = t1;
                spent += t1;
                if (msv_like) {
                    // The MSV launches of a search cover thousands of profiles x every
                    // target: rank shapes by the marginal time of a second copy of the
                    // sample, in which the launch, the stream synchronisation and the
                    // persistent grid's drain tail cancel.  The later stages' launches
                    // cover survivor lists about the size of the sample, where the tail
                    // is part of the cost, so they are ranked by the plain time.
                    const double t2 = timed([&] {
                        run_implicit(stage, SX_FORM_BATCHED, sample2, (const int32_t*)dts.p, nts, dout.p, 0);
                    });
                    spent += t2;
                    ms = (t2 - t1 > 0.25 * t1) ? t2 - t1 : t2 / 2;
                }
                log.push_back({hi, s, ms, cells / ms / 1e6});
                // a shape with more slots than the best so far must beat it by more than
                // the timing noise (3%): it does more work on every model of the bucket
                const bool looser = (long)s.q * s.g > (long)best.q * best.g;
                if (ms < best_ms * (looser && best_ms < 1e300 ? 0.97 : 1.0)) {
                    best_ms = ms;
                    best = s;
                }
            } catch (...) {
                forced.erase(stage);
                throw;
            }
            forced.erase(stage);
            while (pgroups_keep.size() > 64) pgroups_keep.pop_front();
        }
        // a row never claims more than the shape covers; longer models fall through
        table.push_back({std::min<int>(hi, best.q * best.g), best});
    }
    pgroups_keep.clear();
    tuned[stage] = table;
    cascade_dirty = true;
    return log;
}

void Ctx::ensure_pidx() {
    if (d_pidx_n == profs.size() && d_pidx) return;
    dfree(d_pidx);
    std::vector<int32_t> v(profs.size());
    std::iota(v.begin(), v.end(), 0);
    d_pidx = dupload(v);
    d_pidx_n = profs.size();
}

double Ctx::timed(const std::function<void()>& f) {
    HIPCHK(hipEventRecord(ev0, stream));
    f();
    HIPCHK(hipEventRecord(ev1, stream));
    HIPCHK(hipEventSynchronize(ev1));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    return ms;
}

// Forward or Backward special rows of a pair list: the device writes the scaled special
// values and the f64 row scale; the host takes logs exactly as the reference does.
void Ctx::special_rows(int stage, const uint32_t* prof, const uint32_t* tgt, uint64_t n, float* sp, double* score) {
    Ctx& c = *this;
    if (n == 0) return;
    std::vector<int64_t> roff(n);
    int64_t rows = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (tgt[i] >= c.nt || prof[i] >= c.profs.size()) throw SxError(SX_ERR_ARG, "pair index out of range");
        roff[i] = rows;
        rows += c.tlen[tgt[i]] + 1;
    }
    void* d = c.scratch_out.ensure(n * 8);
    float* dx5 = (float*)c.scratch_x5.ensure((size_t)rows * 5 * sizeof(float));
    double* dls = (double*)c.scratch_ls.ensure((size_t)rows * sizeof(double));
    c.run_pairs(stage, SX_FORM_BATCHED, prof, tgt, n, d, true, &roff, dx5, dls);
    HIPCHK(hipStreamSynchronize(c.stream));
    std::vector<float> x5((size_t)rows * 5);
    std::vector<double> ls((size_t)rows);
    std::vector<double> dev(n);
    HIPCHK(hipMemcpy(x5.data(), dx5, x5.size() * 4, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ls.data(), dls, ls.size() * 8, hipMemcpyDeviceToHost));
    {
        std::vector<double> tmp(n);
        HIPCHK(hipMemcpy(tmp.data(), d, n * 8, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < n; i++) dev[c.last_order[i]] = tmp[i];
    }
    // ln of the stored values plus the row scale, with the host libm, as
    // dp.rs:284-292 (Forward) and dp.rs:372-377, 383-390 (Backward) do
    auto lnf = [](float x) { return x > 0.0f ? std::log(x) : sx::NEG_INF; };
    for (uint64_t i = 0; i < n; i++) {
        const int L = c.tlen[tgt[i]];
        const float* r = x5.data() + (size_t)roff[i] * 5;
        const double* lr = ls.data() + roff[i];
        float* o = sp ? sp + (size_t)roff[i] * 5 : nullptr;
        if (L == 0) {
            if (o)
                for (int q = 0; q < 5; q++) o[q] = sx::NEG_INF;
            if (score) score[i] = -INFINITY;
            continue;
        }
        const sx::Xsc xs = sx::Xsc::make((size_t)L, true);
        if (o) {
            for (int row = 0; row <= L; row++) {
                const float l = (float)lr[row];
                for (int q = 0; q < 5; q++) o[(size_t)q * (L + 1) + row] = lnf(r[row * 5 + q]) + l;
            }
            if (stage == SX_STAGE_FORWARD) {
                o[0] = 0.0f;  // N(0) = ln 1
                for (int q = 2; q < 5; q++) o[(size_t)q * (L + 1)] = sx::NEG_INF;
            } else {
                o[(size_t)2 * (L + 1)] = o[(size_t)3 * (L + 1)] = o[(size_t)4 * (L + 1)] = sx::NEG_INF;
                o[L] = sx::NEG_INF;                                   // N(L)
                o[(size_t)1 * (L + 1) + L] = sx::NEG_INF;             // B(L)
                o[(size_t)3 * (L + 1) + L] = sx::NEG_INF;             // J(L)
                const float n1 = L > 1 ? o[1] : sx::NEG_INF;
                o[0] = sx::logsum(n1 + xs.n_loop, o[(size_t)1 * (L + 1)] + xs.n_move);
            }
        }
        if (score) {
            if (stage == SX_STAGE_FORWARD) {
                const float fin = r[L * 5 + 4] * xs.pc_move;
                score[i] = (double)lnf(fin) + lr[L];
            } else {
                score[i] = o ? (double)o[0] : dev[i];
            }
        }
    }
    }

}  // namespace sx

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using sx::Ctx;
using sx::SxError;

struct sx_ctx {
    Ctx* c = nullptr;
    std::string err;
};

template <class F>
static int guard(sx_ctx* ctx, F&& f) {
    try {
        if (ctx && ctx->c) (void)hipSetDevice(ctx->c->device);
        f();
        return SX_OK;
    } catch (const SxError& e) {
        if (ctx) ctx->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        if (ctx) ctx->err = "out of host memory";
        return SX_ERR_NOMEM;
    } catch (const std::exception& e) {
        if (ctx) ctx->err = e.what();
        return SX_ERR_PARSE;
    }
}

static std::string g_create_err;

extern "C" int sx_create(int device, sx_ctx** out) {
    if (!out) return SX_ERR_ARG;
    *out = nullptr;
    try {
        auto* c = new sx_ctx;
        try {
            c->c = new Ctx(device);
        } catch (...) {
            delete c;
            throw;
        }
        *out = c;
        return SX_OK;
    } catch (const SxError& e) {
        g_create_err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        g_create_err = e.what();
        return SX_ERR_HIP;
    }
}

extern "C" void sx_destroy(sx_ctx* ctx) {
    if (!ctx) return;
    delete ctx->c;
    delete ctx;
}

extern "C" const char* sx_last_error(const sx_ctx* ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

extern "C" int sx_set_targets(sx_ctx* ctx, const uint8_t* codes, const uint64_t* offsets, uint32_t n) {
    if (!ctx || (!offsets && n)) return SX_ERR_ARG;
    return guard(ctx, [&] { ctx->c->set_targets(codes, offsets, n); });
}

sx::Hmm sx::hmm_from_c(const sx_hmm& h) {
    sx::Hmm o;
    if (h.m <= 0 || !h.mat || !h.t) throw SxError(SX_ERR_ARG, "model needs m > 0, mat and t");
    o.name = h.name ? h.name : "";
    if (h.acc) o.acc = std::string(h.acc);
    if (h.desc) o.desc = std::string(h.desc);
    o.m = h.m;
    o.mat.resize((size_t)h.m + 1);
    o.t.resize((size_t)h.m + 1);
    for (int k = 0; k <= h.m; k++) {
        for (int x = 0; x < sx::K; x++) o.mat[k][x] = h.mat[(size_t)k * sx::K + x];
        for (int j = 0; j < 7; j++) o.t[k][j] = h.t[(size_t)k * 7 + j];
    }
    if (h.compo) {
        std::array<float, sx::K> c;
        for (int x = 0; x < sx::K; x++) c[x] = h.compo[x];
        o.compo = c;
    }
    if (h.has_ga) o.ga = std::make_pair(h.ga[0], h.ga[1]);
    if (h.has_tc) o.tc = std::make_pair(h.tc[0], h.tc[1]);
    if (h.has_nc) o.nc = std::make_pair(h.nc[0], h.nc[1]);
    o.stats.has_msv = h.has_stats & 1;
    o.stats.has_vit = (h.has_stats >> 1) & 1;
    o.stats.has_fwd = (h.has_stats >> 2) & 1;
    o.stats.msv[0] = h.stats[0];
    o.stats.msv[1] = h.stats[1];
    o.stats.vit[0] = h.stats[2];
    o.stats.vit[1] = h.stats[3];
    o.stats.fwd[0] = h.stats[4];
    o.stats.fwd[1] = h.stats[5];
    return o;
}

extern "C" int sx_set_profiles(sx_ctx* ctx, const sx_hmm* hmms, uint32_t n) {
    if (!ctx || (!hmms && n)) return SX_ERR_ARG;
    return guard(ctx, [&] {
        std::vector<sx::Profile> ps;
        ps.reserve(n);
        for (uint32_t i = 0; i < n; i++) ps.push_back(sx::configure(sx::hmm_from_c(hmms[i])));
        ctx->c->set_profiles(std::move(ps));
    });
}

static int host_threads() { return (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

extern "C" int sx_load_hmm_file(sx_ctx* ctx, const char* path, uint32_t* n_loaded) {
    if (!ctx || !path) return SX_ERR_ARG;
    return guard(ctx, [&] {
        std::vector<sx::Profile> ps = ctx->c->profs;
        std::vector<sx::Profile> add = sx::load_profiles(path, host_threads());
        const uint32_t k = (uint32_t)add.size();
        for (auto& p : add) ps.push_back(std::move(p));
        ctx->c->set_profiles(std::move(ps));
        if (n_loaded) *n_loaded = k;
    });
}

extern "C" int sx_write_profile_cache(const sx_ctx* ctx, const char* path) {
    if (!ctx || !path) return SX_ERR_ARG;
    return guard(const_cast<sx_ctx*>(ctx), [&] {
        sx::CacheWriter w(path);
        for (const auto& p : ctx->c->profs) w.add(p);
        w.close();
    });
}

extern "C" int sx_convert_hmm_file(const char* hmm_path, const char* cache_path, int threads, uint32_t* n_models) {
    if (!hmm_path || !cache_path) return SX_ERR_ARG;
    try {
        auto s = sx::ProfileStream::open(hmm_path, threads > 0 ? threads : host_threads(), 1024);
        sx::CacheWriter w(cache_path);
        std::vector<sx::Profile> b;
        uint32_t n = 0;
        for (;;) {
            b.clear();
            if (s->next(256, b) == 0) break;
            for (const auto& p : b) w.add(p);
            n += (uint32_t)b.size();
        }
        w.close();
        if (n_models) *n_models = n;
        return SX_OK;
    } catch (const SxError& e) {
        g_create_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_create_err = "out of host memory";
        return SX_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_create_err = e.what();
        return SX_ERR_PARSE;
    }
}

extern "C" int sx_search_hmm_file(sx_ctx* ctx, const char* path, const sx_search_cfg* cfg, uint32_t batch,
                                  sx_results** out) {
    if (!ctx || !path || !cfg || !out) return SX_ERR_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        auto r = std::make_unique<sx_results>();
        sx::search_hmm_file(*ctx->c, path, *cfg, batch, *r);
        *out = r.release();
    });
}

extern "C" int sx_results_models(const sx_results* r, uint32_t* n) {
    if (!r || !n) return SX_ERR_ARG;
    *n = (uint32_t)r->models.size();
    return SX_OK;
}

extern "C" int sx_results_model(const sx_results* r, uint32_t i, const char** name, const char** acc, int32_t* m) {
    if (!r || i >= r->models.size()) return SX_ERR_ARG;
    const auto& x = r->models[i];
    if (name) *name = x.name.c_str();
    if (acc) *acc = x.acc ? x.acc->c_str() : nullptr;
    if (m) *m = x.m;
    return SX_OK;
}

extern "C" int sx_encode(const uint8_t* seq, size_t n, uint8_t* out, size_t* n_out) {
    if ((n && (!seq || !out)) || !n_out) return SX_ERR_ARG;
    size_t k = 0;
    for (size_t i = 0; i < n; i++) {
        const uint8_t b = seq[i];
        if (b == ' ' || b == '\t' || b == '\n' || b == '\r' || b == '\f') continue;  // is_ascii_whitespace
        out[k++] = sx::encode_residue(b);
    }
    *n_out = k;
    return SX_OK;
}

extern "C" int sx_profile_info(const sx_ctx* ctx, uint32_t p, const char** name, const char** acc, const char** desc,
                               int32_t* m) {
    if (!ctx || p >= ctx->c->profs.size()) return SX_ERR_ARG;
    const sx::Profile& pr = ctx->c->profs[p];
    if (name) *name = pr.name.c_str();
    if (acc) *acc = pr.acc ? pr.acc->c_str() : nullptr;
    if (desc) *desc = pr.desc ? pr.desc->c_str() : nullptr;
    if (m) *m = pr.m;
    return SX_OK;
}

extern "C" uint32_t sx_num_profiles(const sx_ctx* ctx) { return ctx ? (uint32_t)ctx->c->profs.size() : 0; }
extern "C" uint32_t sx_num_targets(const sx_ctx* ctx) { return ctx ? ctx->c->nt : 0; }

extern "C" int sx_profile_tables(const sx_ctx* ctx, uint32_t p, int32_t* m, float* msc, float* tsc, float* mo,
                                 float* tp, float* tbm) {
    if (!ctx || p >= ctx->c->profs.size()) return SX_ERR_ARG;
    const sx::Profile& pr = ctx->c->profs[p];
    if (m) *m = pr.m;
    if (msc) std::memcpy(msc, pr.msc.data(), pr.msc.size() * 4);
    if (tsc) std::memcpy(tsc, pr.tsc.data(), pr.tsc.size() * 4);
    if (mo) std::memcpy(mo, pr.mo.data(), pr.mo.size() * 4);
    if (tp) std::memcpy(tp, pr.tp.data(), pr.tp.size() * 4);
    if (tbm) *tbm = pr.msv_tbm;
    return SX_OK;
}

static size_t elem_size(int stage) { return stage == SX_STAGE_FORWARD || stage == SX_STAGE_BACKWARD ? 8 : 4; }  // MSV16 / VIT16: int32

extern "C" int sx_score_all(sx_ctx* ctx, int stage, int form, void* out, double* ms) {
    if (!ctx || stage < 0 || stage > SX_STAGE_VIT16 || (form != SX_FORM_PLAIN && form != SX_FORM_BATCHED)) return SX_ERR_ARG;
    return guard(ctx, [&] {
        Ctx& c = *ctx->c;
        if (c.profs.empty() || c.nt == 0) {
            if (ms) *ms = 0.0;
            return;
        }
        c.ensure_pidx();
        const size_t n = c.profs.size() * (size_t)c.nt;
        void* d = c.scratch_out.ensure(n * elem_size(stage));
        const double t = c.timed([&] { c.run_all(stage, form, d); });
        if (ms) *ms = t;
        if (out) HIPCHK(hipMemcpy(out, d, n * elem_size(stage), hipMemcpyDeviceToHost));
    });
}

extern "C" int sx_score_pairs(sx_ctx* ctx, int stage, int form, const uint32_t* prof, const uint32_t* tgt, uint64_t n,
                              void* out, double* ms) {
    if (!ctx || stage < 0 || stage > SX_STAGE_VIT16 || (n && (!prof || !tgt))) return SX_ERR_ARG;
    return guard(ctx, [&] {
        Ctx& c = *ctx->c;
        if (n == 0) {
            if (ms) *ms = 0.0;
            return;
        }
        const size_t es = elem_size(stage);
        void* d = c.scratch_out.ensure(n * es);
        const double t = c.timed([&] { c.run_pairs(stage, form, prof, tgt, n, d, false, nullptr, nullptr, nullptr); });
        if (ms) *ms = t;
        if (out) {
            std::vector<uint8_t> tmp(n * es);
            HIPCHK(hipMemcpy(tmp.data(), d, n * es, hipMemcpyDeviceToHost));
            for (uint64_t i = 0; i < n; i++)
                std::memcpy((uint8_t*)out + c.last_order[i] * es, tmp.data() + i * es, es);
        }
    });
}

extern "C" int sx_score_pairs_noalloc(sx_ctx* ctx, int stage, int form, const uint32_t* prof, const uint32_t* tgt,
                                      uint64_t n, void* out, double* ms) {
    if (!ctx || stage < 0 || stage > SX_STAGE_VIT16 || (form != SX_FORM_PLAIN && form != SX_FORM_BATCHED) ||
        (n && (!prof || !tgt)))
        return SX_ERR_ARG;
    return guard(ctx, [&] { ctx->c->score_pairs_noalloc(stage, form, prof, tgt, n, out, ms); });
}

extern "C" int sx_specials(sx_ctx* ctx, int stage, const uint32_t* prof, const uint32_t* tgt, uint64_t n, float* sp,
                           double* score) {
    if (!ctx || (stage != SX_STAGE_FORWARD && stage != SX_STAGE_BACKWARD) || (n && (!prof || !tgt))) return SX_ERR_ARG;
    return guard(ctx, [&] { ctx->c->special_rows(stage, prof, tgt, n, sp, score); });
}

extern "C" int sx_shape_for(const sx_ctx* ctx, int stage, int form, int32_t m, int32_t* q, int32_t* g,
                            int32_t* hoist) {
    if (!ctx || !q || !g || stage < 0 || stage > SX_STAGE_VIT16) return SX_ERR_ARG;
    try {
        const sx::Shape s = ctx->c->shape_for(stage, form, m);
        *q = s.q;
        *g = s.g;
        if (hoist) *hoist = s.hoist;
        return SX_OK;
    } catch (const SxError& e) {
        return e.code;
    }
}

extern "C" int sx_set_shapes(sx_ctx* ctx, int stage, uint32_t n, const int32_t* max_m, const int32_t* q,
                             const int32_t* g, const int32_t* hoist) {
    if (!ctx || stage < 0 || stage > SX_STAGE_VIT16 || (n && (!max_m || !q || !g))) return SX_ERR_ARG;
    return guard(ctx, [&] {
        Ctx& c = *ctx->c;
        std::vector<std::pair<int, sx::Shape>> tab;
        for (uint32_t i = 0; i < n; i++) {
            const int h = hoist ? hoist[i] : 0;
            if (h < 0 || h > 4 || (h == 2 && stage != SX_STAGE_VITERBI && stage != SX_STAGE_FORWARD))
                throw SxError(SX_ERR_ARG, "table mode must be 0, 1, 3 or 4 (2 for Viterbi and Forward only; long "
                                          "models have a fixed shape)");
            // a mode-2 row may only serve models over 512 positions (the lookup checks it)
            if (h == 2 && max_m[i] <= SX_GEM_COPY_MIN_M)
                throw SxError(SX_ERR_ARG, "a global-emission row must cover models over " +
                                              std::to_string(SX_GEM_COPY_MIN_M) + " positions");
            const bool quant = stage == SX_STAGE_MSV16 || stage == SX_STAGE_VIT16;
            const int top = quant ? SX_Q_MAX_M : SX_GEM_MIN_M;
            if (max_m[i] > top)
                throw SxError(SX_ERR_ARG, "rows cover models up to " + std::to_string(top) +
                                              (quant ? "; the quantised bounds have no longer shape"
                                                     : "; longer models use a fixed shape"));
            const sx::KernelInfo& k = c.kernel(stage, q[i], g[i], false, h);
            if ((long)q[i] * g[i] < max_m[i])
                throw SxError(SX_ERR_ARG, "shape capacity below its max_m");
            if (Ctx::lds_bytes(k) > 160 * 1024) throw SxError(SX_ERR_ARG, "shape exceeds the LDS budget");
            if (i && max_m[i] <= max_m[i - 1]) throw SxError(SX_ERR_ARG, "max_m must increase");
            tab.push_back({max_m[i], sx::Shape{q[i], g[i], h}});
        }
        if (n == 0)
            c.tuned.erase(stage);
        else
            c.tuned[stage] = tab;
        c.cascade_dirty = true;
    });
}

extern "C" int sx_get_shapes(sx_ctx* ctx, int stage, uint32_t cap, uint32_t* n, int32_t* max_m, int32_t* q,
                             int32_t* g, int32_t* hoist) {
    if (!ctx || !n || stage < 0 || stage > SX_STAGE_VIT16) return SX_ERR_ARG;
    return guard(ctx, [&] {
        Ctx& c = *ctx->c;
        auto it = c.tuned.find(stage);
        *n = it == c.tuned.end() ? 0 : (uint32_t)it->second.size();
        if (it == c.tuned.end()) return;
        for (uint32_t i = 0; i < *n && i < cap; i++) {
            if (max_m) max_m[i] = it->second[i].first;
            if (q) q[i] = it->second[i].second.q;
            if (g) g[i] = it->second[i].second.g;
            if (hoist) hoist[i] = it->second[i].second.hoist;
        }
    });
}

extern "C" int sx_autotune(sx_ctx* ctx, int stage, double budget_ms) {
    if (!ctx || stage < 0 || stage > SX_STAGE_VIT16) return SX_ERR_ARG;
    return guard(ctx, [&] { ctx->c->last_tune[stage] = ctx->c->autotune(stage, budget_ms); });
}

extern "C" int sx_autotune_log(sx_ctx* ctx, int stage, uint32_t cap, uint32_t* n, int32_t* max_m, int32_t* q,
                               int32_t* g, int32_t* hoist, double* gcups) {
    if (!ctx || !n || stage < 0 || stage > SX_STAGE_VIT16) return SX_ERR_ARG;
    const auto& L = ctx->c->last_tune[stage];
    *n = (uint32_t)L.size();
    for (uint32_t i = 0; i < *n && i < cap; i++) {
        if (max_m) max_m[i] = L[i].max_m;
        if (q) q[i] = L[i].shape.q;
        if (g) g[i] = L[i].shape.g;
        if (hoist) hoist[i] = L[i].shape.hoist;
        if (gcups) gcups[i] = L[i].gcups;
    }
    return SX_OK;
}

extern "C" int sx_run_cascade(sx_ctx* ctx, int form, sx_cascade_stats* st, sx_pair_scores* out, uint64_t cap,
                              uint64_t* n_out) {
    if (!ctx || !st || (form != SX_FORM_PLAIN && form != SX_FORM_BATCHED) || (cap && !out)) return SX_ERR_ARG;
    return guard(ctx, [&] {
        if (ctx->c->profs.empty() || ctx->c->nt == 0) {
            sx_cascade_records_clear(ctx->c);
            if (n_out) *n_out = 0;
            return;
        }
        ctx->c->ensure_pidx();
        (void)sx_run_cascade_impl(ctx->c, form, st, out, cap, n_out);
    });
}

extern "C" int sx_cascade_launches(sx_ctx* ctx, uint32_t cap, uint32_t* n, sx_launch_stat* out) {
    if (!ctx || (cap && !out)) return SX_ERR_ARG;
    return guard(ctx, [&] {
        const uint32_t k = sx_cascade_launches_impl(ctx->c, out, cap);
        if (n) *n = k;
    });
}

extern "C" int sx_cascade_records(sx_ctx* ctx, sx_pair_scores* out, uint64_t cap, uint64_t* n_out) {
    if (!ctx || (cap && !out)) return SX_ERR_ARG;
    return guard(ctx, [&] {
        const uint64_t n = sx_cascade_records_impl(ctx->c, out, cap);
        if (n_out) *n_out = n;
    });
}

extern "C" int sx_search(sx_ctx* ctx, const sx_search_cfg* cfg, sx_results** out) {
    if (!ctx || !cfg || !out) return SX_ERR_ARG;
    *out = nullptr;
    return guard(ctx, [&] {
        auto r = std::make_unique<sx_results>();
        sx::search(*ctx->c, *cfg, *r);
        *out = r.release();
    });
}

extern "C" int sx_rescore(sx_ctx* ctx, const sx_envelope* env, uint64_t n, sx_envelope_score* out) {
    if (!ctx || (n && (!env || !out))) return SX_ERR_ARG;
    return guard(ctx, [&] { sx::rescore_envelopes(*ctx->c, env, n, false, out); });
}

extern "C" int sx_rescore_host(sx_ctx* ctx, const sx_envelope* env, uint64_t n, sx_envelope_score* out) {
    if (!ctx || (n && (!env || !out))) return SX_ERR_ARG;
    return guard(ctx, [&] { sx::rescore_envelopes(*ctx->c, env, n, true, out); });
}

extern "C" int sx_search_host(const sx_hmm* hmms, uint32_t np, const uint8_t* codes, const uint64_t* offsets,
                              uint32_t nt, const sx_search_cfg* cfg, uint64_t npairs, const uint32_t* prof,
                              const uint32_t* tgt, const double* fwd, const double* bwd, const float* fsp,
                              const float* bsp, sx_results** out) {
    if (!out || !cfg || (np && !hmms) || !offsets || (npairs && (!prof || !tgt || !fwd || !bwd || !fsp || !bsp)))
        return SX_ERR_ARG;
    *out = nullptr;
    try {
        auto r = std::make_unique<sx_results>();
        sx::search_host(hmms, np, codes, offsets, nt, *cfg, npairs, prof, tgt, fwd, bwd, fsp, bsp, *r);
        *out = r.release();
        return SX_OK;
    } catch (const SxError& e) {
        g_create_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_create_err = "out of host memory";
        return SX_ERR_NOMEM;
    } catch (const std::exception& e) {
        g_create_err = e.what();
        return SX_ERR_PARSE;
    }
}

extern "C" int sx_results_size(const sx_results* r, uint64_t* nhits, uint64_t* ndomains) {
    if (!r) return SX_ERR_ARG;
    if (nhits) *nhits = r->hits.size();
    if (ndomains) *ndomains = r->doms.size();
    return SX_OK;
}

extern "C" int sx_results_stats(const sx_results* r, sx_search_stats* st) {
    if (!r || !st) return SX_ERR_ARG;
    *st = r->stats;
    return SX_OK;
}

extern "C" int sx_results_copy(const sx_results* r, sx_hit* hits, sx_domain* domains) {
    if (!r) return SX_ERR_ARG;
    if (hits && !r->hits.empty()) std::memcpy(hits, r->hits.data(), r->hits.size() * sizeof(sx_hit));
    if (domains && !r->doms.empty()) std::memcpy(domains, r->doms.data(), r->doms.size() * sizeof(sx_domain));
    return SX_OK;
}

extern "C" void sx_results_free(sx_results* r) { delete r; }

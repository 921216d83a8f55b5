// sx_runtime.h -- host-side device context for the MI355X Plan7 search path.
#pragma once
#include <hip/hip_runtime.h>

#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <tuple>
#include <stdexcept>
#include <string>
#include <vector>

#include "sx_dev.h"
#include "sx_model.h"
#include "../../include/sx_hmmer.h"

namespace sx {

struct CascadeState;

struct SxError : std::runtime_error {
    int code;
    SxError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// throws SxError(SX_ERR_HIP) with the failing call's text when a HIP call fails
#define HIPCHK(expr)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            throw SxError(SX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
    } while (0)


struct DeviceBuf {
    void* p = nullptr;
    size_t bytes = 0;
    void* ensure(size_t n);
    void release();
    DeviceBuf() = default;
    DeviceBuf(const DeviceBuf&) = delete;
    DeviceBuf& operator=(const DeviceBuf&) = delete;
    DeviceBuf(DeviceBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    ~DeviceBuf();
};

struct KernelInfo {
    int stage = 0, q = 0, g = 0, spec = 0, hoist = 0, wid = 0;
    int block = 256;  // threads per workgroup; a tile holds block / g pairs
    int lds = 0;      // bytes of dynamic LDS per workgroup (the staged profile tables)
    int index = 0;    // position in the kernel table
    const void* fn = nullptr;
};

struct Shape {
    int q, g;
    // table mode: 0 = tables in LDS; 1 = transitions kept in registers across rows;
    // 2 = long models, emission rows read from global memory (transitions in LDS);
    // 3, 4 = MSV with G = 8 or G >= 16: 16-byte-aligned lane blocks in LDS, 512 / 1024 threads
    int hoist;
};

struct TuneRecord {
    int max_m;
    Shape shape;
    double ms;
    double gcups;
};

struct Ctx;

// Fan independent launches out over the side streams after everything already queued
// on `origin`, and join them back: launches on different side streams may run at the
// same time.  Works eagerly and under stream capture (the joins become graph edges).
// The i-th launch goes to side stream i % NSIDE, so launches i and i + NSIDE*k share a
// stream and run in order: head slots need only differ modulo NSIDE*k.
struct Fanout {
    Fanout(Ctx& c, hipStream_t origin);
    ~Fanout();  // joins what join() has not (an exception between launches)
    Fanout(const Fanout&) = delete;
    Fanout& operator=(const Fanout&) = delete;
    hipStream_t next();
    hipStream_t side(int i);  // side stream i mod NSIDE (ordered after the fork)
    // The stream and head slot (0-63) of a launch that does not use next(): exclusive
    // launches one after another on side stream 0, the others in turn on the other side
    // streams.  A slot is congruent to its stream's index modulo NSIDE, so launches that
    // share a slot share a stream.  For persistent kernels whose tables take most of a
    // CU's LDS: side by side, each one's resident workgroups keep the other's off their
    // CUs until it ends.
    hipStream_t pick(bool exclusive, int* slot);
    void join();
    Ctx& c;
    hipStream_t origin;
    int n = 0;
    int n_excl = 0, n_other = 0;
    bool used[4] = {};
};

struct Ctx {
    explicit Ctx(int dev);
    ~Ctx();

    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // side streams (default priority) on which a stage's independent per-shape launches
    // run side by side; fork_ev / join_ev order them after / before the main stream
    static constexpr int NSIDE = 4;
    hipStream_t side[NSIDE] = {};
    hipEvent_t fork_ev = nullptr, join_ev[NSIDE] = {};
    int n_cu = 0;
    std::map<long, KernelInfo> kernels;
    // distinct fields: stage | Q (20 bits) | G (16 bits) | special rows (1 bit) | table mode (3 bits)
    static long key(int st, int q, int g, int sp, int h) {
        return ((long)st << 40) | ((long)q << 20) | ((long)g << 4) | ((long)sp << 3) | (long)h;
    }
    const KernelInfo& kernel(int stage, int q, int g, bool spec, int hoist) const;
    static size_t lds_bytes(const KernelInfo& k) { return (size_t)k.lds; }

    // the search's cascade lets profiles without STATS lines through every filter
    bool cascade_pass_uncalibrated = false;
    // every bias score of the cascade is recomputed on the host (set from the environment
    // variable SX_BIAS_RECHECK_ALL=1 at creation; a test of the host recomputation)
    bool bias_recheck_all = false;

    // profiles
    std::vector<Profile> profs;
    float *d_msc = nullptr, *d_mo = nullptr, *d_tsc = nullptr, *d_tp = nullptr, *d_tbm = nullptr;
    float *d_gmsc = nullptr, *d_gmo = nullptr;  // padded emission rows of long models
    int64_t* d_goff = nullptr;
    int64_t *d_off = nullptr, *d_toff = nullptr;
    int32_t* d_m = nullptr;
    int32_t* d_vmo = nullptr;  // SxProfiles::vit_m_only
    int16_t* d_qsc = nullptr;  // SxProfiles::qsc, qtbm (quantised MSV bound)
    int32_t* d_qtbm = nullptr;
    int16_t* d_qtsc = nullptr;  // SxProfiles::qtsc (quantised Viterbi bound)
    int32_t* d_pidx = nullptr;
    size_t d_pidx_n = 0;
    void set_profiles(std::vector<Profile>&& ps);
    void free_profiles();
    void ensure_pidx();

    // targets
    uint32_t nt = 0;
    std::vector<int32_t> tlen;
    std::vector<int64_t> h_toff;   // start of each target in h_codes
    std::vector<uint8_t> h_codes;  // host copy of the packed codes (the device buffer's layout)
    std::vector<int32_t> sorted;
    int64_t total_len = 0;
    uint8_t* d_codes = nullptr;
    int64_t* d_soff = nullptr;
    int32_t* d_len = nullptr;
    SxLen* d_lm = nullptr;
    int32_t* d_sorted = nullptr;
    int32_t* d_identity = nullptr;
    uint32_t* d_ilv = nullptr;    // SxTargets::ilv
    int64_t* d_ilvoff = nullptr;  // SxTargets::ilv_off
    void set_targets(const uint8_t* codes, const uint64_t* offsets, uint32_t n);
    void free_targets();

    SxProfiles dprof() const;
    SxTargets dtarg() const;

    // shapes: tuned[stage] = sorted (max M, shape) table
    std::map<int, std::vector<std::pair<int, Shape>>> tuned;
    std::map<int, Shape> forced;  // autotuner: every model of the stage uses this shape
    std::vector<TuneRecord> autotune(int stage, double budget_ms);
    std::vector<TuneRecord> last_tune[6];
    Shape default_shape(int stage, int form, int m) const;
    Shape shape_for(int stage, int form, int m) const;
    Shape shape_for_spec(int stage, int m) const;
    // the fixed shape of a long model (M > SX_GEM_MIN_M): 64 lanes, the smallest Q that covers M
    static Shape gem_shape(int m);

    // launches
    unsigned long long* d_heads = nullptr;
    DeviceBuf scratch_out, scratch_tids, scratch_tiles, scratch_roff, scratch_x5, scratch_ls;
    // envelope rescoring: jobs, results and the matrices' workspace, kept across searches
    DeviceBuf env_jobs, env_res, env_ws, env_dws;
    std::deque<DeviceBuf> pgroups_keep;
    std::vector<uint64_t> last_order;
    int persistent_grid(const KernelInfo& k, size_t lds, int64_t ntiles) const;
    // s = nullptr: the context's stream.  Two launches that may run at the same time
    // need different head slots.
    void launch(const KernelInfo& k, int stage, SxWork w, void* out, SxSpecialsOut sp, int grid, bool persistent,
                int head_slot, hipStream_t s = nullptr, unsigned long long* heads = nullptr);
    void run_all(int stage, int form, void* dout);
    void run_implicit(int stage, int form, const std::vector<int32_t>& plist, const int32_t* dtids, int ntids,
                      void* dout, int64_t out_stride);
    void run_pairs(int stage, int form, const uint32_t* prof, const uint32_t* tgt, uint64_t n, void* dout, bool spec,
                   const std::vector<int64_t>* row_off, float* dx5, double* dls);
    double timed(const std::function<void()>& f);
    // sx_score_pairs_noalloc: run_pairs with every device buffer in the device's static
    // scratch array (na_dev, from hipGetSymbolAddress at creation) and host staging
    // reserved at creation, so that a call allocates nothing
    void* na_dev = nullptr;
    std::vector<uint32_t> na_order;
    std::vector<int32_t> na_tids;
    std::vector<SxTile> na_tiles;
    std::vector<double> na_res;
    std::vector<int64_t> na_cnt;             // tiles per kernel-table index
    std::vector<const KernelInfo*> na_kern;  // kernel-table index -> kernel
    void score_pairs_noalloc(int stage, int form, const uint32_t* prof, const uint32_t* tgt, uint64_t n, void* out,
                             double* ms);
    // Forward (stage 2) or Backward (stage 3) special rows of n pairs: sp receives
    // 5 x (L+1) floats per pair (rows N B E J C, pairs back to back), score one value each
    void special_rows(int stage, const uint32_t* prof, const uint32_t* tgt, uint64_t n, float* sp, double* score);

    // device filter cascade (sx_pipeline.cpp); rebuilt when profiles, targets or shapes change
    std::shared_ptr<CascadeState> cascade;
    bool cascade_dirty = true;
    // counts set_profiles calls: the cascade keeps the per-profile pass thresholds, which
    // depend on the profiles and F1-F3 only, across changes of targets and shapes
    uint64_t profiles_version = 0;
};

}  // namespace sx

// sx_model.h -- host-side Plan7 model for the MI355X search path: digital alphabet
// (sentinel-x src/hmmer/alphabet.rs), HMMER3 text reader (parse.rs), local/multihit
// profile configuration and per-target length model (model.rs).  The configured
// tables are what the device kernels consume, so they must be bit-identical to the
// reference's: same f32/f64 operation order, glibc log/exp (what Rust's f32::ln,
// f64::ln and f32::exp call on x86_64-linux-gnu), no FP contraction.
#pragma once
#include <stdint.h>

#include <array>
#include <memory>
#include <optional>
#include <string>
#include <vector>

namespace sx {

constexpr int K = 20;
constexpr int NCODE = 25;
constexpr uint8_t CODE_STAR = 24;
constexpr float NEG_INF = -__builtin_inff();
constexpr float LN2F = 0.693147182464599609375f;

extern const float BG[K];                            // alphabet.rs:36-39
int degenerate_members(int code, const int** mem);   // alphabet.rs:42-54
uint8_t encode_residue(uint8_t b);                   // alphabet.rs:58-88 (sx_encode: alphabet.rs:91-93)

struct Stats {  // model.rs:27-40
    bool has_msv = false, has_vit = false, has_fwd = false;
    double msv[2] = {0, 0}, vit[2] = {0, 0}, fwd[2] = {0, 0};
    bool calibrated() const { return has_msv && has_vit && has_fwd; }
};

struct Hmm {  // model.rs:54-71
    std::string name;
    std::optional<std::string> acc, desc;
    int m = 0;
    std::vector<std::array<float, K>> mat;  // (M+1) rows, row 0 unused
    std::vector<std::array<float, 7>> t;    // (M+1) rows, MM MI MD IM II DM DD
    std::optional<std::array<float, K>> compo;
    std::optional<std::pair<float, float>> ga, tc, nc;
    Stats stats;
};

struct Profile {  // model.rs:75-113 (packed: 8 transition rows MM MI MD IM II DM DD BM)
    std::string name;
    std::optional<std::string> acc, desc;
    int m = 0;
    std::vector<float> msc, tsc, mo, tp;  // NCODE*(M+1), 8*(M+1), NCODE*(M+1), 8*(M+1)
    float msv_tbm = 0.0f;
    std::array<float, K> compo{};
    std::optional<std::pair<float, float>> ga, tc, nc;
    Stats stats;
    const float* msc_row(int x) const { return msc.data() + (size_t)x * (m + 1); }
    const float* mo_row(int x) const { return mo.data() + (size_t)x * (m + 1); }
    const float* t_row(int r) const { return tsc.data() + (size_t)r * (m + 1); }
    const float* p_row(int r) const { return tp.data() + (size_t)r * (m + 1); }
};

Profile configure(const Hmm& h);  // model.rs:146-245

struct Xsc {  // model.rs:266-312 (scores) with XscP (model.rs:322-331)
    float n_loop, n_move, j_loop, j_move, c_loop, c_move, e_move, e_loop;
    float pn_loop, pn_move, pj_loop, pj_move, pc_loop, pc_move, pe_move, pe_loop;
    static Xsc make(size_t l, bool multihit);
};

float null_score(size_t l);  // model.rs:334-337
float logsum(float a, float b);  // dp.rs:32-40

// MSV length model in f32 (dp.rs:98-100).
void msv_length(size_t l, float* tloop, float* tmove);

// Quantisation of the int16 MSV bound (no reference counterpart; the bound is a
// conservative prefilter of dp.rs:92-145): an emission score rounded up to int16
// units (-inf -> -32768, clamped to the int16 range), and a special-state score
// rounded up to int32 fine units.
int16_t q_emit(float sc);
int32_t q_const(float v);
// A transition score for the int16 Viterbi bound: q_emit, except that a score <= 0 stays
// <= 0 (fl(a + t) <= a for t <= 0, so 0 is a valid upper rounding).
int16_t q_trans(float t);

// HMMER3/b..f reader (parse.rs:42-298): plain text or gzip by magic; throws
// std::runtime_error with the reference's messages on malformed input.
class HmmReader {
public:
    static std::unique_ptr<HmmReader> open(const std::string& path);
    static std::unique_ptr<HmmReader> from_string(const std::string& text);
    // A piece of a file: messages name `source` and count lines from line0 + 1.
    static std::unique_ptr<HmmReader> from_string(std::string text, const std::string& source, size_t line0);
    // Returns false at end of input.
    bool next(Hmm& out);
    virtual ~HmmReader();

private:
    struct Impl;
    std::unique_ptr<Impl> impl_;
    HmmReader();
};

std::optional<std::pair<float, float>> cutoff_pair(const std::string& rest);  // parse.rs:260-266

}  // namespace sx

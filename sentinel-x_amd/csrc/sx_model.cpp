// sx_model.cpp -- alphabet, HMMER3 reader, profile configuration and length model
// (sentinel-x src/hmmer/alphabet.rs, parse.rs, model.rs) for the MI355X search path.
#include "sx_model.h"

#include <zlib.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace sx {

const float BG[K] = {
    0.0787945f, 0.0151600f, 0.0535222f, 0.0668298f, 0.0397062f, 0.0695071f, 0.0229198f,
    0.0590092f, 0.0594422f, 0.0963728f, 0.0237718f, 0.0414386f, 0.0482904f, 0.0395639f,
    0.0540978f, 0.0683364f, 0.0540687f, 0.0673417f, 0.0114135f, 0.0304133f,
};

int degenerate_members(int code, const int** mem) {
    static const int B[2] = {2, 11}, J[2] = {7, 9}, Z[2] = {3, 13};
    static const int X[20] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19};
    switch (code) {
        case 20: *mem = B; return 2;
        case 21: *mem = J; return 2;
        case 22: *mem = Z; return 2;
        case 23: *mem = X; return 20;
        default: *mem = nullptr; return 0;
    }
}

uint8_t encode_residue(uint8_t b) {
    static const char* AA = "ACDEFGHIKLMNPQRSTVWY";
    if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 32);
    switch (b) {
        case 'U': return 1;
        case 'O': return 8;
        case 'B': return 20;
        case 'J': return 21;
        case 'Z': return 22;
        case 'X': return 23;
        case '*': case '-': case '.': case '_': case '~': return CODE_STAR;
        default: break;
    }
    const char* p = std::strchr(AA, (char)b);
    if (b != 0 && p) return (uint8_t)(p - AA);
    if (b >= 'A' && b <= 'Z') return 23;
    return CODE_STAR;
}

static inline float lnp32(float p) { return p > 0.0f ? (float)std::log((double)p) : NEG_INF; }

Profile configure(const Hmm& h) {
    const int m = h.m;
    const int w = m + 1;
    Profile p;
    p.name = h.name;
    p.acc = h.acc;
    p.desc = h.desc;
    p.m = m;
    p.msc.assign((size_t)NCODE * w, NEG_INF);
    for (int k = 1; k <= m; k++) {
        float sc[NCODE];
        for (int x = 0; x < K; x++) {
            const double pr = (double)h.mat[k][x];
            sc[x] = pr > 0.0 ? (float)std::log(pr / (double)BG[x]) : NEG_INF;
        }
        for (int code = 20; code <= 23; code++) {
            const int* mem;
            const int nm = degenerate_members(code, &mem);
            float num = 0.0f, den = 0.0f;
            for (int j = 0; j < nm; j++) {
                num += sc[mem[j]] * BG[mem[j]];
                den += BG[mem[j]];
            }
            sc[code] = den > 0.0f ? num / den : 0.0f;
        }
        sc[CODE_STAR] = NEG_INF;
        for (int x = 0; x < NCODE; x++) p.msc[(size_t)x * w + k] = sc[x];
    }
    p.tsc.assign((size_t)8 * w, NEG_INF);
    for (int k = 1; k < m; k++)
        for (int j = 0; j < 7; j++) p.tsc[(size_t)j * w + k] = lnp32(h.t[k][j]);
    if (m >= 1) {
        p.tsc[(size_t)5 * w + 1] = NEG_INF;  // D_1 does not exist
        p.tsc[(size_t)6 * w + 1] = NEG_INF;
    }
    // occupancy (model.rs:248-260)
    std::vector<float> occ((size_t)w, 0.0f);
    if (m > 0) {
        occ[1] = h.t[0][1] + h.t[0][0];
        for (int k = 2; k <= m; k++) {
            const auto& tk = h.t[k - 1];
            occ[k] = occ[k - 1] * (tk[0] + tk[1]) + (1.0f - occ[k - 1]) * tk[5];
        }
    }
    float z = 0.0f;
    for (int k = 1; k <= m; k++) z += occ[k] * (float)(m - k + 1);
    for (int k = 1; k <= m; k++)
        p.tsc[(size_t)7 * w + k] = (z > 0.0f && occ[k] > 0.0f) ? (float)std::log((double)(occ[k] / z)) : NEG_INF;
    p.msv_tbm = std::log(2.0f / ((float)m * ((float)m + 1.0f)));
    p.mo.resize(p.msc.size());
    for (size_t i = 0; i < p.msc.size(); i++) p.mo[i] = std::exp(p.msc[i]);
    p.tp.resize(p.tsc.size());
    for (size_t i = 0; i < p.tsc.size(); i++) p.tp[i] = std::exp(p.tsc[i]);
    if (h.compo) {
        p.compo = *h.compo;
    } else {
        for (int x = 0; x < K; x++) p.compo[x] = BG[x];
    }
    p.ga = h.ga;
    p.tc = h.tc;
    p.nc = h.nc;
    p.stats = h.stats;
    return p;
}

Xsc Xsc::make(size_t l, bool multihit) {
    const float nj = multihit ? 1.0f : 0.0f;
    const float pmove = (2.0f + nj) / ((float)l + 2.0f + nj);
    const float ploop = 1.0f - pmove;
    const float lloop = (float)std::log((double)ploop);
    const float lmove = (float)std::log((double)pmove);
    Xsc x;
    x.n_loop = x.j_loop = x.c_loop = lloop;
    x.n_move = x.j_move = x.c_move = lmove;
    x.e_move = multihit ? -LN2F : 0.0f;
    x.e_loop = multihit ? -LN2F : NEG_INF;
    x.pn_loop = x.pj_loop = x.pc_loop = ploop;
    x.pn_move = x.pj_move = x.pc_move = pmove;
    x.pe_move = multihit ? 0.5f : 1.0f;
    x.pe_loop = multihit ? 0.5f : 0.0f;
    return x;
}

float null_score(size_t l) {
    const float p1 = (float)l / ((float)l + 1.0f);
    return (float)((double)l * std::log((double)p1) + std::log(1.0 - (double)p1));
}

float logsum(float a, float b) {
    const float hi = a > b ? a : b, lo = a > b ? b : a;
    const float d = hi - lo;
    if (!(d < 30.0f)) return hi;
    return hi + std::log1p(std::exp(-d));
}

void msv_length(size_t l, float* tloop, float* tmove) {
    const float lf = (float)l;
    *tloop = std::log(lf / (lf + 3.0f));
    *tmove = std::log(3.0f / (lf + 3.0f));
}

// The bound's units (see sx_dev.h): 1024 emission units per nat and 2^22 fine units per
// nat; the slack is the f32 rounding allowance of a cell (512 nats) or of an N/J/B step
// (64 nats).  The products and the slack are exact in double, so the ceiling is the
// rounding.
int16_t q_emit(float sc) {
    if (sc == NEG_INF) return -32768;
    const double v = std::ceil((double)sc * 1024.0 + 0.03125);
    return (int16_t)std::min(32767.0, std::max(-32768.0, v));
}

int32_t q_const(float v) { return (int32_t)std::ceil((double)v * 4194304.0 + 16.0); }

int16_t q_trans(float t) {
    const int16_t q = q_emit(t);
    return (t <= 0.0f && q > 0) ? (int16_t)0 : q;
}

// ---------------------------------------------------------------------------
// HMMER3 reader
// ---------------------------------------------------------------------------

namespace {

std::vector<std::string> split_ws(const std::string& s) {
    std::vector<std::string> out;
    size_t i = 0, n = s.size();
    while (i < n) {
        while (i < n && std::isspace((unsigned char)s[i])) i++;
        if (i >= n) break;
        size_t j = i;
        while (j < n && !std::isspace((unsigned char)s[j])) j++;
        out.push_back(s.substr(i, j - i));
        i = j;
    }
    return out;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

std::string trim_start(const std::string& s) {
    size_t a = 0;
    while (a < s.size() && std::isspace((unsigned char)s[a])) a++;
    return s.substr(a);
}

bool parse_f64(const std::string& t, double* v) {
    if (t.empty() || t.find_first_of("xXpP") != std::string::npos) return false;
    errno = 0;
    char* end = nullptr;
    *v = std::strtod(t.c_str(), &end);
    return end && *end == 0;
}

bool parse_f32(const std::string& t, float* v) {
    if (t.empty() || t.find_first_of("xXpP") != std::string::npos) return false;
    char* end = nullptr;
    *v = std::strtof(t.c_str(), &end);
    return end && *end == 0;
}

bool parse_usize(const std::string& t, long long* v) {
    if (t.empty()) return false;
    size_t i = t[0] == '+' ? 1 : 0;
    if (i >= t.size()) return false;
    long long r = 0;
    for (; i < t.size(); i++) {
        if (t[i] < '0' || t[i] > '9') return false;
        r = r * 10 + (t[i] - '0');
        if (r > (1LL << 40)) return false;
    }
    *v = r;
    return true;
}

float tok_prob(const std::string& tok) {
    if (tok == "*") return 0.0f;
    double v;
    if (!parse_f64(tok, &v)) throw std::runtime_error("non-numeric probability token \"" + tok + "\"");
    if (!std::isfinite(v) || v < 0.0) throw std::runtime_error("invalid -ln(p) value \"" + tok + "\"");
    return (float)std::exp(-v);
}

template <size_t N>
void fill(const std::vector<std::string>& toks, size_t from, std::array<float, N>& out, const char* what) {
    if (toks.size() < from + N)
        throw std::runtime_error(std::string("expected ") + std::to_string(N) + " " + what + ", found " +
                                 std::to_string(toks.size() - std::min(toks.size(), from)));
    for (size_t i = 0; i < N; i++) out[i] = tok_prob(toks[from + i]);
}

}  // namespace

std::optional<std::pair<float, float>> cutoff_pair(const std::string& rest) {
    std::string c = rest;
    for (auto& ch : c)
        if (ch == ';') ch = ' ';
    auto toks = split_ws(c);
    if (toks.empty()) return std::nullopt;
    float a, b;
    if (!parse_f32(toks[0], &a)) return std::nullopt;
    if (toks.size() < 2 || !parse_f32(toks[1], &b)) b = a;
    return std::make_pair(a, b);
}

struct HmmReader::Impl {
    gzFile gz = nullptr;
    std::string text;
    size_t tpos = 0;
    bool from_text = false;
    std::string source = "<stream>";
    size_t lineno = 0;
    bool done = false;

    bool next_line(std::string& line) {
        line.clear();
        if (from_text) {
            if (tpos >= text.size()) return false;
            size_t e = text.find('\n', tpos);
            if (e == std::string::npos) e = text.size();
            line = text.substr(tpos, e - tpos);
            tpos = e + 1;
        } else {
            char buf[65536];
            bool any = false;
            for (;;) {
                if (!gzgets(gz, buf, sizeof buf)) {
                    // end of input, or a read / decompression error (a damaged or cut
                    // gzip stream), which must not pass for the end of the file
                    int err = Z_OK;
                    const char* msg = gzerror(gz, &err);
                    if (err != Z_OK) throw std::runtime_error(source + ": reading: " + (msg ? msg : "error"));
                    break;
                }
                any = true;
                line += buf;
                if (!line.empty() && line.back() == '\n') break;
            }
            if (!any) return false;
        }
        lineno++;
        while (!line.empty() && (line.back() == '\n' || line.back() == '\r')) line.pop_back();
        return true;
    }
    std::string need_line(const char* what) {
        std::string l;
        if (!next_line(l)) throw std::runtime_error(source + ": unexpected end of file (" + what + ")");
        return l;
    }
    [[noreturn]] void fail(const std::string& msg) {
        throw std::runtime_error(source + " line " + std::to_string(lineno) + ": " + msg);
    }

    bool record(Hmm& h) {  // parse.rs:107-236
        std::string first;
        for (;;) {
            if (!next_line(first)) return false;
            if (!trim(first).empty()) break;
        }
        if (first.rfind("HMMER3", 0) != 0) {
            if (first.rfind("HMMER2", 0) == 0) fail("HMMER2 format is not supported; convert with `hmmconvert`");
            fail("expected 'HMMER3/x' header, found \"" + first + "\"");
        }
        h = Hmm();
        bool alph_ok = false;
        long long m = 0;
        for (;;) {
            std::string line = need_line("header");
            std::string tr = trim_start(line);
            if (tr.rfind("HMM ", 0) == 0 || tr == "HMM" || tr.rfind("HMM\t", 0) == 0) break;
            size_t sp = 0;
            while (sp < tr.size() && !std::isspace((unsigned char)tr[sp])) sp++;
            const std::string tag = tr.substr(0, sp);
            const std::string rest = sp < tr.size() ? trim(tr.substr(sp + 1)) : std::string();
            try {
                if (tag == "NAME") {
                    h.name = rest;
                } else if (tag == "ACC") {
                    h.acc = rest;
                } else if (tag == "DESC") {
                    h.desc = rest;
                } else if (tag == "LENG") {
                    if (!parse_usize(rest, &m)) fail("bad LENG value \"" + rest + "\"");
                } else if (tag == "ALPH") {
                    std::string r = rest;
                    for (auto& ch : r) ch = (char)std::tolower((unsigned char)ch);
                    if (r != "amino")
                        fail("profile '" + h.name + "' has alphabet '" + rest +
                             "'; only protein (amino) profiles are supported");
                    alph_ok = true;
                } else if (tag == "GA" || tag == "TC" || tag == "NC") {
                    auto cp = cutoff_pair(rest);
                    if (!cp) fail("bad " + tag + " line");
                    (tag == "GA" ? h.ga : tag == "TC" ? h.tc : h.nc) = cp;
                } else if (tag == "STATS") {
                    auto f = split_ws(rest);
                    if (f.size() >= 4) {
                        double a, b;
                        if (!parse_f64(f[2], &a) || !parse_f64(f[3], &b)) fail("bad STATS value");
                        if (f[1] == "MSV") {
                            h.stats.has_msv = true;
                            h.stats.msv[0] = a;
                            h.stats.msv[1] = b;
                        } else if (f[1] == "VITERBI") {
                            h.stats.has_vit = true;
                            h.stats.vit[0] = a;
                            h.stats.vit[1] = b;
                        } else if (f[1] == "FORWARD") {
                            h.stats.has_fwd = true;
                            h.stats.fwd[0] = a;
                            h.stats.fwd[1] = b;
                        }
                    }
                }
            } catch (std::runtime_error&) {
                throw;
            }
        }
        if (m == 0) fail("profile '" + h.name + "' has no LENG");
        if (!alph_ok) fail("profile '" + h.name + "' has no ALPH line");
        h.m = (int)m;
        need_line("transition header");
        h.mat.assign((size_t)m + 1, std::array<float, K>{});
        h.t.assign((size_t)m + 1, std::array<float, 7>{});
        std::string l = need_line("model body");
        std::string body_first;
        if (trim_start(l).rfind("COMPO", 0) == 0) {
            auto toks = split_ws(l);
            std::array<float, K> c{};
            try {
                fill(toks, 1, c, "values");
            } catch (std::runtime_error& e) {
                fail(std::string("COMPO: ") + e.what());
            }
            h.compo = c;
            body_first = need_line("node-0 insert line");
        } else {
            body_first = l;
        }
        if (split_ws(body_first).size() < (size_t)K) fail("truncated node-0 insert line");
        {
            auto toks = split_ws(need_line("node-0 transitions"));
            try {
                fill(toks, 0, h.t[0], "transition values");
            } catch (std::runtime_error& e) {
                fail(std::string("node 0 transitions: ") + e.what());
            }
        }
        for (long long k = 1; k <= m; k++) {
            auto toks = split_ws(need_line("match line"));
            long long idx;
            if (toks.empty() || !parse_usize(toks[0], &idx)) fail("bad match-line node index");
            if (idx != k) fail("node index " + std::to_string(idx) + ", expected " + std::to_string(k));
            if (toks.size() < (size_t)(1 + K)) fail("truncated match emissions at node " + std::to_string(k));
            try {
                fill(toks, 1, h.mat[k], "values");
            } catch (std::runtime_error& e) {
                fail("node " + std::to_string(k) + ": " + e.what());
            }
            if (split_ws(need_line("insert line")).size() < (size_t)K)
                fail("truncated insert emissions at node " + std::to_string(k));
            auto tt = split_ws(need_line("transition line"));
            try {
                fill(tt, 0, h.t[k], "transition values");
            } catch (std::runtime_error& e) {
                fail("node " + std::to_string(k) + " transitions: " + e.what());
            }
        }
        for (;;) {
            std::string x;
            if (!next_line(x)) break;
            const std::string tx = trim(x);
            if (tx.empty()) continue;
            if (tx == "//") break;
            fail("expected '//', found \"" + x + "\"");
        }
        return true;
    }
};

HmmReader::HmmReader() : impl_(new Impl) {}
HmmReader::~HmmReader() {
    if (impl_ && impl_->gz) gzclose(impl_->gz);
}

std::unique_ptr<HmmReader> HmmReader::open(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::runtime_error("opening HMM file " + path + ": " + std::strerror(errno));
    std::fclose(f);
    std::unique_ptr<HmmReader> r(new HmmReader);
    // zlib reads plain files transparently and concatenated gzip members in sequence
    r->impl_->gz = gzopen(path.c_str(), "rb");
    if (!r->impl_->gz) throw std::runtime_error("opening HMM file " + path);
    gzbuffer(r->impl_->gz, 1 << 20);
    r->impl_->source = path;
    return r;
}

std::unique_ptr<HmmReader> HmmReader::from_string(const std::string& text) {
    std::unique_ptr<HmmReader> r(new HmmReader);
    r->impl_->from_text = true;
    r->impl_->text = text;
    return r;
}

std::unique_ptr<HmmReader> HmmReader::from_string(std::string text, const std::string& source, size_t line0) {
    std::unique_ptr<HmmReader> r(new HmmReader);
    r->impl_->from_text = true;
    r->impl_->text = std::move(text);
    r->impl_->source = source;
    r->impl_->lineno = line0;
    return r;
}

bool HmmReader::next(Hmm& out) {
    if (impl_->done) return false;
    try {
        if (!impl_->record(out)) {
            impl_->done = true;
            return false;
        }
        return true;
    } catch (...) {
        impl_->done = true;
        throw;
    }
}

}  // namespace sx

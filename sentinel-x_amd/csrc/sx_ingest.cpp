// sx_ingest.cpp -- streamed parsing and configuration of model files, and the packed
// profile cache (sentinel-x src/hmmer/search.rs:252-271, parse.rs, model.rs:146-245).
#include "sx_ingest.h"

#include <zlib.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <thread>

namespace sx {

namespace {

constexpr uint32_t CACHE_VERSION = 1;
constexpr uint32_t ABSENT = 0xFFFFFFFFu;

void put(std::vector<uint8_t>& b, const void* p, size_t n) {
    const uint8_t* s = static_cast<const uint8_t*>(p);
    b.insert(b.end(), s, s + n);
}
template <class T>
void put_v(std::vector<uint8_t>& b, const T& v) {
    put(b, &v, sizeof v);
}

struct Cursor {
    const uint8_t* p;
    const uint8_t* end;
    bool take(void* dst, size_t n) {
        if ((size_t)(end - p) < n) return false;
        std::memcpy(dst, p, n);
        p += n;
        return true;
    }
};

bool only_space(const std::string& s) {
    for (char c : s)
        if (!std::isspace((unsigned char)c)) return false;
    return true;
}

bool is_end_line(const std::string& line) {  // trimmed "//", as the reader's record end
    size_t a = 0, b = line.size();
    while (a < b && std::isspace((unsigned char)line[a])) a++;
    while (b > a && std::isspace((unsigned char)line[b - 1])) b--;
    return b - a == 2 && line[a] == '/' && line[a + 1] == '/';
}

// ---------------------------------------------------------------------------
// packed profile cache
// ---------------------------------------------------------------------------

class CacheStream : public ProfileStream {
public:
    explicit CacheStream(const std::string& path) : path_(path) {
        f_ = std::fopen(path.c_str(), "rb");
        if (!f_) throw std::runtime_error("opening profile cache " + path + ": " + std::strerror(errno));
        char magic[8];
        uint32_t ver = 0, zero = 0;
        if (std::fread(magic, 1, 8, f_) != 8 || std::memcmp(magic, CACHE_MAGIC, 8) != 0 ||
            std::fread(&ver, 4, 1, f_) != 1 || std::fread(&zero, 4, 1, f_) != 1 || std::fread(&count_, 8, 1, f_) != 1)
            throw std::runtime_error(path + ": not a profile cache");
        if (ver != CACHE_VERSION) throw std::runtime_error(path + ": profile cache version " + std::to_string(ver));
    }
    ~CacheStream() override {
        if (f_) std::fclose(f_);
    }
    size_t next(size_t n, std::vector<Profile>& out) override {
        size_t k = 0;
        std::vector<uint8_t> buf;
        for (; k < n && done_ < count_; k++, done_++) {
            uint32_t sz = 0;
            if (std::fread(&sz, 4, 1, f_) != 1 || sz > (1u << 30)) damaged();
            buf.resize(sz);
            if (std::fread(buf.data(), 1, sz, f_) != sz) damaged();
            out.push_back(decode(buf));
        }
        return k;
    }
    double ms_wait() const override { return 0.0; }

private:
    [[noreturn]] void damaged() const {
        throw std::runtime_error(path_ + ": truncated or damaged profile cache at profile " + std::to_string(done_));
    }
    Profile decode(const std::vector<uint8_t>& b) const {
        Cursor c{b.data(), b.data() + b.size()};
        Profile p;
        int32_t m;
        uint32_t ln[3];
        uint8_t fl[4];
        float ga[2], tc[2], nc[2];
        if (!c.take(&m, 4) || !c.take(&p.msv_tbm, 4) || !c.take(ln, 12) || !c.take(fl, 4) || !c.take(ga, 8) ||
            !c.take(tc, 8) || !c.take(nc, 8) || !c.take(p.stats.msv, 16) || !c.take(p.stats.vit, 16) ||
            !c.take(p.stats.fwd, 16) || !c.take(p.compo.data(), 4 * K) || m < 1 || m > (1 << 20))
            damaged();
        p.m = m;
        size_t sl = 0;
        for (uint32_t l : ln) {
            if (l != ABSENT && l > (1u << 20)) damaged();
            sl += l == ABSENT ? 0 : l;
        }
        std::string s[3];
        for (int i = 0; i < 3; i++) {
            if (ln[i] == ABSENT) continue;
            s[i].resize(ln[i]);
            if (!c.take(s[i].data(), ln[i])) damaged();
        }
        const size_t padn = (4 - sl % 4) % 4;
        char pad[4];
        if (!c.take(pad, padn)) damaged();
        p.name = s[0];
        if (ln[1] != ABSENT) p.acc = s[1];
        if (ln[2] != ABSENT) p.desc = s[2];
        if (fl[0] & 1) p.ga = std::make_pair(ga[0], ga[1]);
        if (fl[0] & 2) p.tc = std::make_pair(tc[0], tc[1]);
        if (fl[0] & 4) p.nc = std::make_pair(nc[0], nc[1]);
        p.stats.has_msv = fl[0] & 8;
        p.stats.has_vit = fl[0] & 16;
        p.stats.has_fwd = fl[0] & 32;
        const size_t w = (size_t)m + 1;
        p.msc.resize(NCODE * w);
        p.tsc.resize(8 * w);
        p.mo.resize(NCODE * w);
        p.tp.resize(8 * w);
        if (!c.take(p.msc.data(), 4 * p.msc.size()) || !c.take(p.tsc.data(), 4 * p.tsc.size()) ||
            !c.take(p.mo.data(), 4 * p.mo.size()) || !c.take(p.tp.data(), 4 * p.tp.size()) || c.p != c.end)
            damaged();
        return p;
    }

    std::string path_;
    FILE* f_ = nullptr;
    uint64_t count_ = 0, done_ = 0;
};

// ---------------------------------------------------------------------------
// HMMER3 text / gzip: records split by a reader thread, parsed and configured by a pool
// ---------------------------------------------------------------------------

class TextStream : public ProfileStream {
public:
    TextStream(const std::string& path, int threads, size_t ahead)
        : path_(path), ahead_(std::max<size_t>(ahead, 1)) {
        FILE* f = std::fopen(path.c_str(), "rb");
        if (!f) throw std::runtime_error("opening HMM file " + path + ": " + std::strerror(errno));
        std::fclose(f);
        gz_ = gzopen(path.c_str(), "rb");  // plain files are read transparently
        if (!gz_) throw std::runtime_error("opening HMM file " + path);
        gzbuffer(gz_, 1 << 20);
        const int nw = std::max(1, threads);
        reader_ = std::thread([this] { read_records(); });
        for (int i = 0; i < nw; i++) workers_.emplace_back([this] { work(); });
    }
    ~TextStream() override { stop(); }

    size_t next(size_t n, std::vector<Profile>& out) override {
        size_t k = 0;
        while (k < n) {
            if (seq_) {  // sequential after a record the pool could not parse
                Hmm h;
                if (!seq_->next(h)) break;
                out.push_back(configure(h));
                k++;
                continue;
            }
            Result r;
            const auto t0 = std::chrono::steady_clock::now();
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_res_.wait(lk, [&] { return results_.count(taken_) || (reader_done_ && taken_ >= n_records_); });
                if (!results_.count(taken_)) {
                    wait_ms_ += ms_since(t0);
                    break;  // end of file
                }
                r = std::move(results_[taken_]);
                results_.erase(taken_);
            }
            wait_ms_ += ms_since(t0);
            if (!r.ok) {
                // The reference reads the file in one pass; its message (and whether the
                // file is in fact malformed, should a record break where a piece ends)
                // comes from reading it that way from this record on.
                fall_back(taken_);
                continue;
            }
            {
                std::lock_guard<std::mutex> lk(mu_);
                taken_++;
                cv_space_.notify_all();
            }
            if (!r.have) continue;  // a trailing piece with no record
            out.push_back(std::move(r.prof));
            k++;
        }
        return k;
    }
    double ms_wait() const override { return wait_ms_; }

private:
    struct Work {
        size_t idx;
        std::string text;
        size_t line0;
        bool failed;  // the file could not be read past the previous record
    };
    struct Result {
        bool ok = false, have = false;
        Profile prof;
    };

    static double ms_since(std::chrono::steady_clock::time_point t0) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }

    void read_records() {
        std::string rec, line;
        size_t lineno = 0, line0 = 0, idx = 0;
        char buf[1 << 16];
        auto push = [&](std::string&& text, size_t first, bool failed = false) {
            std::unique_lock<std::mutex> lk(mu_);
            cv_space_.wait(lk, [&] { return stop_ || idx - taken_ < ahead_; });
            if (stop_) return false;
            work_.push_back({idx++, std::move(text), first, failed});
            cv_work_.notify_one();
            return true;
        };
        bool ok = true;
        for (;;) {
            line.clear();
            bool any = false;
            while (gzgets(gz_, buf, sizeof buf)) {
                any = true;
                line += buf;
                if (line.back() == '\n') break;
            }
            if (!any) break;
            lineno++;
            rec += line;
            if (is_end_line(line)) {
                if (!(ok = push(std::move(rec), line0))) break;
                rec.clear();
                line0 = lineno;
            }
        }
        int err = Z_OK;
        gzerror(gz_, &err);
        if (ok && err != Z_OK) {
            // a read or decompression error is not the end of the file: the record that
            // follows fails, so the one-pass reader takes over there and reports it
            push(std::string(), line0, true);
        } else if (ok && !only_space(rec)) {
            push(std::move(rec), line0);
        }
        std::lock_guard<std::mutex> lk(mu_);
        reader_done_ = true;
        n_records_ = idx;
        cv_work_.notify_all();
        cv_res_.notify_all();
    }

    void work() {
        for (;;) {
            Work w;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_work_.wait(lk, [&] { return stop_ || !work_.empty() || reader_done_; });
                if (stop_ || (work_.empty() && reader_done_)) return;
                w = std::move(work_.front());
                work_.pop_front();
            }
            Result r;
            if (w.failed) {
                std::lock_guard<std::mutex> lk(mu_);
                results_[w.idx] = std::move(r);  // not ok
                cv_res_.notify_all();
                continue;
            }
            try {
                auto rd = HmmReader::from_string(std::move(w.text), path_, w.line0);
                Hmm h;
                r.have = rd->next(h);
                if (r.have) r.prof = configure(h);
                r.ok = true;
            } catch (...) {
                r.ok = false;
            }
            std::lock_guard<std::mutex> lk(mu_);
            results_[w.idx] = std::move(r);
            cv_res_.notify_all();
        }
    }

    void stop() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            cv_space_.notify_all();
            cv_work_.notify_all();
            cv_res_.notify_all();
        }
        if (reader_.joinable()) reader_.join();
        for (auto& t : workers_)
            if (t.joinable()) t.join();
        workers_.clear();
        if (gz_) gzclose(gz_);
        gz_ = nullptr;
    }

    void fall_back(size_t idx) {
        stop();
        seq_ = HmmReader::open(path_);
        Hmm h;
        for (size_t i = 0; i < idx; i++)
            if (!seq_->next(h)) break;
    }

    std::string path_;
    size_t ahead_;
    gzFile gz_ = nullptr;
    std::thread reader_;
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_space_, cv_work_, cv_res_;
    std::deque<Work> work_;
    std::map<size_t, Result> results_;
    size_t taken_ = 0, n_records_ = 0;
    bool reader_done_ = false, stop_ = false;
    std::unique_ptr<HmmReader> seq_;
    double wait_ms_ = 0.0;
};

}  // namespace

bool is_profile_cache(const std::string& path) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    char m[8];
    const bool ok = std::fread(m, 1, 8, f) == 8 && std::memcmp(m, CACHE_MAGIC, 8) == 0;
    std::fclose(f);
    return ok;
}

CacheWriter::CacheWriter(const std::string& path) : path_(path) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("creating profile cache " + path + ": " + std::strerror(errno));
    f_ = f;
    const uint32_t ver = CACHE_VERSION, zero = 0;
    const uint64_t n = 0;
    if (std::fwrite(CACHE_MAGIC, 1, 8, f) != 8 || std::fwrite(&ver, 4, 1, f) != 1 || std::fwrite(&zero, 4, 1, f) != 1 ||
        std::fwrite(&n, 8, 1, f) != 1)
        throw std::runtime_error("writing profile cache " + path);
}

void CacheWriter::add(const Profile& p) {
    std::vector<uint8_t> b;
    const int32_t m = p.m;
    put_v(b, m);
    put_v(b, p.msv_tbm);
    const uint32_t ln[3] = {(uint32_t)p.name.size(), p.acc ? (uint32_t)p.acc->size() : ABSENT,
                            p.desc ? (uint32_t)p.desc->size() : ABSENT};
    put(b, ln, sizeof ln);
    const uint8_t fl[4] = {(uint8_t)((p.ga ? 1 : 0) | (p.tc ? 2 : 0) | (p.nc ? 4 : 0) | (p.stats.has_msv ? 8 : 0) |
                                     (p.stats.has_vit ? 16 : 0) | (p.stats.has_fwd ? 32 : 0)),
                           0, 0, 0};
    put(b, fl, 4);
    auto pair = [&](const std::optional<std::pair<float, float>>& c) {
        const float v[2] = {c ? c->first : 0.0f, c ? c->second : 0.0f};
        put(b, v, 8);
    };
    pair(p.ga);
    pair(p.tc);
    pair(p.nc);
    put(b, p.stats.msv, 16);
    put(b, p.stats.vit, 16);
    put(b, p.stats.fwd, 16);
    put(b, p.compo.data(), 4 * K);
    put(b, p.name.data(), p.name.size());
    if (p.acc) put(b, p.acc->data(), p.acc->size());
    if (p.desc) put(b, p.desc->data(), p.desc->size());
    const size_t sl = p.name.size() + (p.acc ? p.acc->size() : 0) + (p.desc ? p.desc->size() : 0);
    const uint8_t zero[4] = {0, 0, 0, 0};
    put(b, zero, (4 - sl % 4) % 4);
    put(b, p.msc.data(), 4 * p.msc.size());
    put(b, p.tsc.data(), 4 * p.tsc.size());
    put(b, p.mo.data(), 4 * p.mo.size());
    put(b, p.tp.data(), 4 * p.tp.size());
    const uint32_t sz = (uint32_t)b.size();
    FILE* f = static_cast<FILE*>(f_);
    if (std::fwrite(&sz, 4, 1, f) != 1 || std::fwrite(b.data(), 1, b.size(), f) != b.size())
        throw std::runtime_error("writing profile cache " + path_);
    n_++;
}

void CacheWriter::close() {
    FILE* f = static_cast<FILE*>(f_);
    if (!f) return;
    const bool ok = std::fseek(f, 16, SEEK_SET) == 0 && std::fwrite(&n_, 8, 1, f) == 1;
    const bool closed = std::fclose(f) == 0;
    f_ = nullptr;
    if (!ok || !closed) {
        std::remove(path_.c_str());
        throw std::runtime_error("writing profile cache " + path_);
    }
}

CacheWriter::~CacheWriter() {
    if (f_) {
        std::fclose(static_cast<FILE*>(f_));
        std::remove(path_.c_str());
    }
}

std::unique_ptr<ProfileStream> ProfileStream::open(const std::string& path, int threads, size_t ahead) {
    if (is_profile_cache(path)) return std::make_unique<CacheStream>(path);
    return std::make_unique<TextStream>(path, threads, ahead);
}

std::vector<Profile> load_profiles(const std::string& path, int threads) {
    auto s = ProfileStream::open(path, threads, 1024);
    std::vector<Profile> out;
    while (s->next(1024, out) > 0) {
    }
    return out;
}

}  // namespace sx

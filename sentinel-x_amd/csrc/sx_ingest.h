// sx_ingest.h -- HMM library ingest for the MI355X search path: the reference streams a
// model file 256 profiles at a time, parsing and configuring each batch before its search
// (sentinel-x src/hmmer/search.rs:252-271, parse.rs + model.rs:146).  Here a reader thread
// splits the file into model records and a pool of host threads parses and configures them
// ahead of the device, in file order, so that parsing overlaps the search of the previous
// batch.  A packed profile cache holds configured profiles as `configure` builds them, so
// that a library used again is loaded without parsing or configuration.
#pragma once
#include <stdint.h>

#include <memory>
#include <string>
#include <vector>

#include "sx_model.h"

namespace sx {

// Packed profile cache.  Little-endian; a header, then one record per profile:
//   header: "SXPCACHE" (8 bytes), uint32 version = 1, uint32 0, uint64 profile count
//   record: uint32 payload bytes, then the payload:
//     int32 m, float msv_tbm, uint32 name / acc / desc lengths (0xFFFFFFFF: absent),
//     uint8 flags (bit 0 GA, 1 TC, 2 NC, 3 STATS MSV, 4 STATS VITERBI, 5 STATS FORWARD),
//     3 bytes 0, float ga[2], tc[2], nc[2], double msv[2], vit[2], fwd[2], float compo[20],
//     the three strings, zero-padded to a multiple of 4 bytes,
//     float msc[25 (m+1)], tsc[8 (m+1)], mo[25 (m+1)], tp[8 (m+1)]  (Profile's layout)
constexpr char CACHE_MAGIC[8] = {'S', 'X', 'P', 'C', 'A', 'C', 'H', 'E'};
bool is_profile_cache(const std::string& path);

class CacheWriter {
public:
    explicit CacheWriter(const std::string& path);
    void add(const Profile& p);
    void close();  // writes the count; the destructor of an unclosed writer removes the file
    ~CacheWriter();

private:
    std::string path_;
    void* f_ = nullptr;
    uint64_t n_ = 0;
};

// Configured profiles of a model file, in file order: HMMER3 text or gzip (records parsed
// and configured on `threads` host threads, at most `ahead` records in flight), or a
// packed profile cache.  Errors are the reference's parse errors, with its line numbers.
class ProfileStream {
public:
    static std::unique_ptr<ProfileStream> open(const std::string& path, int threads, size_t ahead);
    // Appends up to n profiles to out; returns how many (0 at the end of the file).
    virtual size_t next(size_t n, std::vector<Profile>& out) = 0;
    // Time next() spent waiting for records to be parsed and configured, ms.
    virtual double ms_wait() const = 0;
    virtual ~ProfileStream() = default;
};

// Every profile of a file (sx_load_hmm_file).
std::vector<Profile> load_profiles(const std::string& path, int threads);

}  // namespace sx

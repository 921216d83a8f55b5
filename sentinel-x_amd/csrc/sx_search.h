// sx_search.h -- hmmsearch-equivalent search (sentinel-x src/hmmer/search.rs:131-271) over the
// device cascade and the host-side domain definition.
#pragma once
#include <optional>
#include <string>
#include <vector>

#include "../../include/sx_hmmer.h"
#include "sx_runtime.h"

struct sx_results {
    std::vector<sx_hit> hits;
    std::vector<sx_domain> doms;
    sx_search_stats stats{};
    // the searched models in profile-index order (ModelInfo, search.rs:245-249)
    struct Model {
        std::string name;
        std::optional<std::string> acc;
        int32_t m;
    };
    std::vector<Model> models;
};

namespace sx {

Hmm hmm_from_c(const sx_hmm& h);  // sx_hmm (caller's model view) -> parsed model
void check_search_cfg(const sx_search_cfg* cfg);
// Device search of the context's profiles against its targets.
void search(Ctx& c, const sx_search_cfg& cfg, sx_results& res);
// search_hmm_file_models (search.rs:252-271): the models of a file (HMMER3 text, gzip or
// a packed profile cache) searched `batch` at a time, each batch parsed and configured on
// host threads while the device searches the one before.  The context's profile set
// becomes the last batch.
void search_hmm_file(Ctx& c, const std::string& path, const sx_search_cfg& cfg, uint32_t batch, sx_results& res);
// The host part alone for pairs whose scores and special rows are given.
void search_host(const sx_hmm* hmms, uint32_t np, const uint8_t* codes, const uint64_t* offsets, uint32_t nt,
                 const sx_search_cfg& cfg, uint64_t npairs, const uint32_t* prof, const uint32_t* tgt,
                 const double* fwd, const double* bwd, const float* fsp, const float* bsp, sx_results& res);

// Envelope rescoring of n envelopes on the device (sxe_rescore) or on the host
// (rescore_raw on host threads); results as sx_envelope_score.
void rescore_envelopes(Ctx& c, const sx_envelope* env, uint64_t n, bool on_host, sx_envelope_score* out);

}  // namespace sx

// sx_pipeline.h -- filter cascade of score_target (pipeline.rs:159-189) on the device.
#pragma once
#include <functional>
#include <utility>
#include <vector>

#include "../../include/sx_hmmer.h"
#include "sx_runtime.h"

namespace sx {

double gumbel_surv(double x, double mu, double lambda);   // pipeline.rs:87-95
double exp_logsurv(double x, double tau, double lambda);  // pipeline.rs:98-104

// Smallest f32 value that a monotone reject-test lets through (see sx_pipeline.cpp).
float pass_threshold(const std::function<bool(float)>& rejects, int* violations);

// bias_filter_score (pipeline.rs:107-150) on the host, in the reference's operation order
// with glibc: the recomputation of the scores the device's bound cannot certify.
float bias_filter_score(const Profile& p, const uint8_t* dsq, int l);

// A pair that passed every filter, with the Forward score that decided its F3 test
// (pipeline.rs:179-183), the score score_target goes on with.
struct PassPair {
    uint32_t profile, target;
    float fwd;
};

// The batched cascade over every pair, with no filter for profiles lacking STATS lines
// (as score_target, pipeline.rs:164); returns the pairs that pass every filter, sorted.
std::vector<PassPair> cascade_pass_pairs(Ctx& c, double f1, double f2, double f3);

}  // namespace sx

extern "C" int sx_run_cascade_impl(sx::Ctx* c, int form, sx_cascade_stats* st, sx_pair_scores* out, uint64_t cap,
                                   uint64_t* n_out);
// Records of the last cascade run on c: copies up to cap of them, returns how many there are.
extern "C" uint64_t sx_cascade_records_impl(sx::Ctx* c, sx_pair_scores* out, uint64_t cap);
extern "C" void sx_cascade_records_clear(sx::Ctx* c);
// Per-launch timing of the last eager batched run: copies up to cap, returns the count.
extern "C" uint32_t sx_cascade_launches_impl(sx::Ctx* c, sx_launch_stat* out, uint32_t cap);

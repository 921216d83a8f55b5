// sx_domain.h -- host-side remainder of score_target (sentinel-x src/hmmer/pipeline.rs:186-294):
// posterior domain decoding from the Forward/Backward special rows, region finding,
// multi-domain resolution by a stochastic-trace ensemble, envelope rescoring with
// full-matrix Forward/Backward, null2 and optimal-accuracy coordinates, and the
// sequence / domain scores.  This part runs on the host, as in the reference.  The device
// search hands it the regions (posterior decoding and region finding run on the device,
// next to Backward); sx_search_host starts from special rows and finds the regions here.
//
// Arithmetic follows the reference's f32/f64 operation order (built without FP
// contraction, glibc logf/expf/log/exp) so that scores agree to f32 rounding and
// coordinates agree exactly.
#pragma once
#include <stdint.h>

#include <vector>

#include "sx_model.h"

namespace sx {

// Special-state rows N, B, E, J, C for rows 0..L as natural-log values (dp.rs:211-226).
struct SpecRows {
    std::vector<float> n, b, e, j, c;
    void reset(int l);
};

// Full DP matrix in probability space, rows 0..L x columns 0..M, with the cumulative
// ln scale per row, the log special rows, and the scaled special values and row
// scale factor that the stochastic traceback reads (dp.rs:398-440).
struct DpMatrix {
    int m = 0, l = 0;
    std::vector<float> mm, im, dm;
    std::vector<double> lnscale;
    SpecRows sp;
    std::vector<float> xn, xb, xe, xj, xc, scale;
    void reset(int m, int l);
};

float forward_matrix(const Profile& p, const uint8_t* dsq, int l, const Xsc& xs, DpMatrix& fx);   // dp.rs:443-522
float backward_matrix(const Profile& p, const uint8_t* dsq, int l, const Xsc& xs, DpMatrix& bx);  // dp.rs:525-601

struct Posterior {  // dp.rs:1262-1272
    int m = 0, l = 0;
    std::vector<float> pm, pi, pn, pj, pc;
};
void posterior_decode(const Profile& p, const Xsc& xs, const DpMatrix& fx, const DpMatrix& bx, Posterior& pp);
void null2_from_posterior(const Profile& p, const Posterior& pp, float out[NCODE]);
// null2_from_posterior in two steps: the expected emission odds of the 20 residues
// (dp.rs:1310-1345), then the degenerate codes and the logarithm (dp.rs:1346-1350).
void null2_raw(const Profile& p, const Posterior& pp, float raw[K]);
void null2_log(const float raw[K], float out[NCODE]);

struct AliCoords {
    int iali, jali, kali, mali;
};
bool optimal_accuracy_coords(const Profile& p, const Xsc& xs, const Posterior& pp, AliCoords& out);
std::vector<AliCoords> viterbi_trace(const Profile& p, const uint8_t* dsq, int l, const Xsc& xs);

// Easel's fast generator (esl_randomness_CreateFast): 32-bit LCG x = 69069 x + 1,
// seeded through esl_mix3 (dp.rs:962-1024).
class EaselRng {
public:
    explicit EaselRng(uint32_t seed);
    void reseed();
    double next();

private:
    uint32_t seed_, x_;
};

struct SampledDomain {
    int sqfrom = 0, sqto = 0, hmmfrom = 0, hmmto = 0;
    std::vector<int> nodes;  // model node of each M or I emission
};
std::vector<SampledDomain> stochastic_trace(const Profile& p, const DpMatrix& fx, const Xsc& xs, EaselRng& rng);

struct RawDomain {
    int ienv, jenv, iali, jali, kali, mali;
    float bits, bias;
    double lnp;
};
struct RawHit {
    float bits, bias;
    double lnp;
    float nexpected;
    int nregions;
    std::vector<RawDomain> domains;
};

// Scratch reused across pairs by one worker.
struct TailWork {
    DpMatrix fx, bx;
    Posterior pp;
};

// A rescored envelope (pipeline.rs:538-539): rows i..j, its unihit Forward score, the
// summed null2 correction over its rows and the alignment coordinates.
struct Envelope {
    int i, j;
    float envsc, domcorr;
    AliCoords ali;
};

// What the device rescoring of one envelope returns: status 0 = rescored, else Forward
// (1) or Backward (2) not finite; the optimal-accuracy coordinates relative to the
// envelope when oa is set; the raw null2 odds (null2_raw) when they were asked for.
struct DeviceEnvelope {
    float envsc;
    int status, oa;
    AliCoords ali;
    float null2raw[K];
};

// The multi-domain resolution of region ireg..jreg (pipeline.rs:298-536): 200 stochastic
// traces of a multihit Forward matrix of the region, clustered; returns the envelopes and
// sets n2sc[ireg..jreg] to the traces' null2 correction.
std::vector<std::pair<int, int>> region_ensemble(const Profile& p, const uint8_t* dsq, int ireg, int jreg, int l,
                                                 TailWork& ws, std::vector<float>& n2sc);

// The steps of rescore (pipeline.rs:545-591) that the device rescoring replaces, on the
// host (its parity reference): unihit full-matrix Forward and Backward of envelope i..j,
// posterior decoding, the raw null2 odds (when null2) and the optimal-accuracy
// coordinates; *bcksc = the Backward score.
DeviceEnvelope rescore_raw(const Profile& p, const uint8_t* dsq, int i, int j, int l, bool null2, TailWork& ws,
                           float* bcksc);

// The rest of rescore (pipeline.rs:568-590) given the device's result for envelope i..j:
// the null2 correction (written to n2sc[i..j] unless null2_done, then summed), and the
// coordinates (the Viterbi-trace fallback on the host when the optimal-accuracy trace
// has no M state).  Returns false where rescore returns None.
bool envelope_from_device(const Profile& p, const uint8_t* dsq, int i, int j, int l, bool null2, bool null2_done,
                          const DeviceEnvelope& r, std::vector<float>& n2sc, Envelope& env);

// The sequence and domain scores of score_target (pipeline.rs:243-293) from its envelopes.
bool score_envelopes(const Profile& p, int l, bool null2, float fwdsc, float nexpected, int nregions,
                     const std::vector<float>& n2sc, const std::vector<Envelope>& envs, RawHit& out);

// The regions that posterior decoding finds (pipeline.rs:194-242): rows i..j (1-based,
// inclusive) in order, each with the outcome of the multi-domain test, and the expected
// number of domains btot[L].
struct RegionSpan {
    int i, j;
    bool multi;
};
struct TailRegions {
    float nexpected = 0.0f;
    std::vector<RegionSpan> spans;
};

// pipeline.rs:194-242 on the log special rows, overall = the Backward score (the Forward
// score if that is not finite).
void find_regions(int l, float overall, const SpecRows& f, const SpecRows& b, TailRegions& out);

// pipeline.rs:208-294 from the regions on: multi-domain resolution, envelope rescoring and
// the scores.  Returns false where score_target returns None.
bool score_regions(const Profile& p, const uint8_t* dsq, int l, bool null2, float fwdsc, const TailRegions& r,
                   TailWork& ws, RawHit& out);

// pipeline.rs:186-294, given the whole-sequence Forward and Backward scores and their
// special rows: find_regions, then score_regions.
bool score_tail(const Profile& p, const uint8_t* dsq, int l, bool null2, float fwdsc, float bcksc,
                const SpecRows& f, const SpecRows& b, TailWork& ws, RawHit& out);

}  // namespace sx

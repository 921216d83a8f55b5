/*
 * sx_hmmer.h -- C ABI of the MI355X implementation of sentinel-x's src/hmmer Plan7
 * search path: the MSV -> bias -> Viterbi -> Forward filter cascade over every
 * (profile, target) pair, the DP stages on their own (MSV, Viterbi, Forward, Backward)
 * and the Forward/Backward special-state rows that the host domain definition reads.
 *
 * The reference is a Rust module with no FFI (src/hmmer/mod.rs:23-30).  Each entry
 * point below names the reference interface it replaces; a Rust caller binds them with
 * an `extern "C"` block.  Conventions: plain C types, caller-owned
 * buffers, status codes (0 = SX_OK, negative = error, message via sx_last_error), no
 * exceptions across the boundary.  A context owns one HIP device, one stream and the
 * device-resident profile and target tables; it is used by one host thread at a time.
 */
#ifndef SX_HMMER_H
#define SX_HMMER_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SX_OK 0
#define SX_ERR_ARG (-1)
#define SX_ERR_HIP (-2)
#define SX_ERR_NOMEM (-3)
#define SX_ERR_PARSE (-4)
#define SX_ERR_STATE (-5)
#define SX_ERR_CUTOFF (-6)
#define SX_ERR_NODEVICE (-7)

typedef struct sx_ctx sx_ctx;

/* DP stages (dp.rs:69 msv, :152 viterbi_score, :229 forward, :301 backward).
 * SX_STAGE_MSV16 has no reference counterpart: the int16 MSV prefilter.  It computes an
 * upper bound of the f32 MSV score with saturating packed int16 cells, as an int32 in
 * units of 2^-22 nat: score(MSV16) >= 2^22 * score(MSV) for every pair whose row maxima
 * stay below 32 nats (INT32_MIN for an empty target); once one reaches 32 nats the result
 * stays at or above 32767*4096 + q(-ln 2) + L*min(q(tloop), 0) + q(tmove), q() being
 * the bound's rounding of a score.  Cells are kept relative to a floor at or above the
 * row's B entry (moved to B when B rises above it and every 16th row), so a segment may
 * start up to 16 rows of B's fall (|tloop| a row) above B.  Models up to 1024
 * positions. */
/* SX_STAGE_VIT16, likewise without a reference counterpart, is the int16 Viterbi
 * prefilter: an upper bound of the f32 Viterbi score in the same units, with the same
 * floor once a cell reaches 32 nats (with the Viterbi length model's loop and move
 * scores).  Profiles with a positive insert-state score are not bounded. */
enum {
    SX_STAGE_MSV = 0,
    SX_STAGE_VITERBI = 1,
    SX_STAGE_FORWARD = 2,
    SX_STAGE_BACKWARD = 3,
    SX_STAGE_MSV16 = 4,
    SX_STAGE_VIT16 = 5
};
/* Launch forms: PLAIN = one launch per profile over blocks of targets in input order,
 * one wavefront per pair; BATCHED = one persistent launch per stage and lane-group
 * shape, length-bucketed tiles from a device-resident queue. */
enum { SX_FORM_PLAIN = 0, SX_FORM_BATCHED = 1 };

/* A parsed HMM in probability space, as parse.rs:107-236 produces it (model.rs:54-71). */
typedef struct {
    const char* name;
    const char* acc;     /* NULL if absent */
    const char* desc;    /* NULL if absent */
    int32_t m;
    const float* mat;    /* (m+1) x 20 match emission probabilities, row 0 unused */
    const float* t;      /* (m+1) x 7 transitions MM MI MD IM II DM DD */
    const float* compo;  /* 20 values, or NULL (then the background is used) */
    int32_t has_ga, has_tc, has_nc;
    float ga[2], tc[2], nc[2];
    int32_t has_stats;   /* bit 0 MSV, bit 1 VITERBI, bit 2 FORWARD */
    double stats[6];     /* msv mu, lambda; viterbi mu, lambda; forward tau, lambda */
} sx_hmm;

/* Create a context on HIP device `device`.  Fails with SX_ERR_NODEVICE when no HIP
 * device is usable; there is no CPU fallback. */
int sx_create(int device, sx_ctx** out);
void sx_destroy(sx_ctx* ctx);
const char* sx_last_error(const sx_ctx* ctx);

/* Targets::new / Targets::from_pairs (search.rs:56-71) for already-digitised codes
 * (alphabet.rs:91-93): target i is codes[offsets[i] .. offsets[i+1]).  The codes stay
 * resident on the device until the next call. */
int sx_set_targets(sx_ctx* ctx, const uint8_t* codes, const uint64_t* offsets, uint32_t n_targets);

/* alphabet::encode (alphabet.rs:58-93), the digitisation Targets::new applies to each
 * sequence (search.rs:56-67): ASCII residues seq[0 .. n) to codes 0..24 (case-insensitive;
 * U -> C, O -> K, B J Z X degenerate, other letters X, gap and stop characters and any
 * other byte '*'), ASCII whitespace dropped.  out holds at least n bytes; *n_out receives
 * the number of codes written.  Needs no device. */
int sx_encode(const uint8_t* seq, size_t n, uint8_t* out, size_t* n_out);

/* Hmm::configure (model.rs:146-245) for each model, then upload.  Replaces the
 * per-batch `configure()` of search_hmm_file_models (search.rs:257-263). */
int sx_set_profiles(sx_ctx* ctx, const sx_hmm* hmms, uint32_t n);
/* Append the models of a HMMER3 file, text or gzip (parse.rs:42-54 open_hmm + HmmReader;
 * records parsed and configured on host threads, in file order, with the reference's
 * messages on malformed input), or of a packed profile cache (sx_write_profile_cache). */
int sx_load_hmm_file(sx_ctx* ctx, const char* path, uint32_t* n_loaded);
uint32_t sx_num_profiles(const sx_ctx* ctx);
uint32_t sx_num_targets(const sx_ctx* ctx);
/* ModelInfo (search.rs:245-249) of profile p, as search_hmm_file_models returns it for
 * every model in file order (search.rs:264), and the profile / acc fields of Hit
 * (search.rs:101-116): name, accession and description (NULL when the model has none)
 * and the model length.  Any output pointer may be NULL.  The strings stay valid until
 * the context's profiles change. */
int sx_profile_info(const sx_ctx* ctx, uint32_t p, const char** name, const char** acc, const char** desc,
                    int32_t* m);
/* Copy out the configured tables of profile p (for checking against the oracle):
 * msc, mo: 25*(M+1); tsc, tp: 8*(M+1) in the order MM MI MD IM II DM DD BM. */
int sx_profile_tables(const sx_ctx* ctx, uint32_t p, int32_t* m, float* msc, float* tsc, float* mo,
                      float* tp, float* msv_tbm);

/* One DP stage over every (profile, target) pair, the pair loop of search_profiles
 * (search.rs:217-224) with dp.rs:69/152/229/301 per pair.  out: P*T values in
 * (profile, target) order, float for MSV/VITERBI, double for FORWARD/BACKWARD.
 * ms (may be NULL) receives the device time of the stage's launches. */
int sx_score_all(sx_ctx* ctx, int stage, int form, void* out, double* ms);

/* The same stage over an explicit pair list (prof[i], tgt[i]). */
int sx_score_pairs(sx_ctx* ctx, int stage, int form, const uint32_t* prof, const uint32_t* tgt, uint64_t n,
                   void* out, double* ms);

/* sx_score_pairs without allocating memory.  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, kernels and results as
 * sx_score_pairs, for at most SX_NOALLOC_MAX_PAIRS pairs (SX_ERR_ARG beyond).  The pair
 * list, its tiles, the device results and the queue heads live in a static device array
 * of about 9 MiB that each device holds from library load until the process ends; the host
 * staging is reserved when the context is created.  Every call on one device shares that
 * array, so calls that overlap on one device (from two contexts or two threads) race. */
#define SX_NOALLOC_MAX_PAIRS (1u << 18)
int sx_score_pairs_noalloc(sx_ctx* ctx, int stage, int form, const uint32_t* prof, const uint32_t* tgt, uint64_t n,
                           void* out, double* ms);

/* Forward (dp.rs:229-298) or Backward (dp.rs:301-392) special-state rows for a pair
 * list, as score_target consumes them (pipeline.rs:180-206).  sp receives 5*(L+1)
 * floats per pair (rows N, B, E, J, C; pairs back to back in list order), score one
 * double per pair.  stage is SX_STAGE_FORWARD or SX_STAGE_BACKWARD. */
int sx_specials(sx_ctx* ctx, int stage, const uint32_t* prof, const uint32_t* tgt, uint64_t n, float* sp,
                double* score);

/* Filter cascade of score_target (pipeline.rs:159-189) over every (profile, target)
 * pair: MSV on all pairs, the bias filter (pipeline.rs:107-150) on MSV survivors,
 * Viterbi on bias-filter survivors, Forward on Viterbi survivors, with the thresholds
 * f1, f2, f3 of PipelineConfig (pipeline.rs:27-42).  Every profile must carry STATS
 * (the reference runs the filters only for calibrated models, pipeline.rs:164);
 * otherwise SX_ERR_STATE.
 *
 * form = SX_FORM_BATCHED keeps the whole chain on the device: survivor lists are
 * compacted in (model length, target length) order, the next stage's tiles are built
 * on the device and consumed by persistent kernels from device-resident queues, and
 * with st->use_graph the chain is captured once in a HIP graph and replayed.
 * form = SX_FORM_PLAIN runs it one profile at a time (one launch per stage and
 * profile, survivor tests on the host between launches).
 *
 * out receives one record per MSV survivor (pair that reached the bias filter), up to
 * cap records, in no particular order (sort on (profile, target) for a canonical one);
 * *n_out is the total number, which may exceed cap (then only cap are written).  stage = 1 rejected by the bias-corrected MSV
 * test, 2 by Viterbi, 3 by Forward, 4 passed every filter; vit / fwd are NaN for
 * stages not reached.
 *
 * Every msv, bias and vit value, and every decision, is the reference's bit for bit.
 * The device's bias filter keeps the reference's f64 recurrence but takes its logarithm
 * differently, under an error bound: where the bound cannot certify the f32 rounding of
 * the score (a few pairs in a million), the host recomputes that score in the
 * reference's order with glibc and re-decides the pair, and the record's flags carry
 * SX_REC_BIAS_HOST. */
#define SX_REC_BIAS_HOST 1
typedef struct {
    uint32_t profile, target;
    float msv, bias, vit;
    float fwd;       /* f32 score as the reference returns it */
    double fwd64;    /* the same before the final f32 rounding */
    int32_t stage;
    int32_t flags;   /* SX_REC_BIAS_HOST: bias score recomputed on the host */
} sx_pair_scores;

typedef struct {
    double f1, f2, f3;          /* in: PipelineConfig thresholds */
    int32_t use_graph;          /* in: batched form replays a captured HIP graph */
    /* in: integer bounds in front of the exact stages, a mask of SX_BOUND_*:
     * SX_BOUND_MSV: the MSV stage first runs the int16 MSV bound (SX_STAGE_MSV16) on every
     *   pair and the f32 MSV only on the pairs whose bound reaches the pair's threshold;
     * SX_BOUND_VITERBI: the Viterbi stage first runs the int16 Viterbi bound
     *   (SX_STAGE_VIT16) and the f32 Viterbi only on its candidates.
     * A bound never lies below the f32 score, so every decision, the survivors and their
     * scores are the same as with 0; the one difference is that a record rejected by
     * the Viterbi bound carries vit = NaN (its f32 Viterbi score was not computed). */
    int32_t bounds;
    uint64_t n_pairs, n_msv_pass, n_bias_pass, n_vit_pass, n_fwd_pass;  /* out */
    double ms_msv, ms_bias, ms_vit, ms_fwd, ms_total;                   /* out: device time */
    uint64_t cells_msv, cells_vit, cells_fwd;                           /* out: DP cells */
    /* out (SX_BOUND_MSV): pairs whose bound reached the threshold (rescored in f32),
     * device time of the bound's launches alone (eager runs), and the rescored cells
     * (eager runs) */
    uint64_t n_msv_cand;
    double ms_msv_bound;
    uint64_t cells_msv_exact;
    /* out (SX_BOUND_VITERBI): the same for the Viterbi stage */
    uint64_t n_vit_cand;
    double ms_vit_bound;
    /* out: bias scores the host recomputed (SX_REC_BIAS_HOST records) and the host time
     * that took, in ms; ms_total includes it */
    uint64_t n_bias_recheck;
    double ms_recheck;
} sx_cascade_stats;
#define SX_BOUND_MSV 1
#define SX_BOUND_VITERBI 2

int sx_run_cascade(sx_ctx* ctx, int form, sx_cascade_stats* st, sx_pair_scores* out, uint64_t cap,
                   uint64_t* n_out);
/* The records of the last sx_run_cascade on ctx, without running it again: up to cap
 * of them into out; *n_out is the total.  A caller that does not know the survivor
 * count runs the cascade with cap = 0, sizes out from *n_out and fetches them here. */
int sx_cascade_records(sx_ctx* ctx, sx_pair_scores* out, uint64_t cap, uint64_t* n_out);

/* Device time of every DP-stage kernel launch of the last batched run made without the
 * graph (st->use_graph = 0), in launch order: the MSV launches (one per shape group,
 * each over all its profiles x all targets), then the Viterbi and Forward launches (one
 * per run of consecutive same-shape profiles, over the survivors).  ms is the time
 * between HIP events recorded on the launch stream around the launch; cells counts
 * M x L over the pairs the launch scored.  n receives the count (0 after a graph run).
 * No counterpart in the reference (a CPU loop has no launches); used for measurement. */
typedef struct {
    int32_t stage, q, g, hoist;
    double ms;
    uint64_t cells;
} sx_launch_stat;
int sx_cascade_launches(sx_ctx* ctx, uint32_t cap, uint32_t* n, sx_launch_stat* out);

/* hmmsearch-equivalent search (search.rs:131-271): the filter cascade on the device;
 * for the pairs that pass it, Forward, Backward, posterior decoding and region finding
 * on the device (pipeline.rs:186-242); then the rest of score_target on the host
 * (pipeline.rs:208-294 from the regions on: stochastic-trace clustering of multi-domain
 * regions, envelope rescoring, null2, optimal-accuracy coordinates, sequence and domain
 * scores) and
 * `report` (search.rs:160-210: sequence E-value = P x Z, domain c-Evalue = P x the
 * profile's number of reported hits, inclusion by E-value, bits or GA/NC/TC).
 *
 * With filters = 0, or for a profile without STATS lines, every pair goes through
 * Forward and the domain definition, as score_target does (pipeline.rs:164, 184-186).
 * Hits come out ordered by profile, then target (search.rs:225-233).  A GA/NC/TC
 * cutoff on a profile that lacks it fails with SX_ERR_CUTOFF (search.rs:131-148). */
enum { SX_CUT_EVALUE = 0, SX_CUT_BITS = 1, SX_CUT_GA = 2, SX_CUT_NC = 3, SX_CUT_TC = 4 };

typedef struct {
    int32_t cutoff;        /* SX_CUT_* (Cutoff, search.rs:26-32) */
    int32_t filters;       /* PipelineConfig.filters (pipeline.rs:27-42) */
    int32_t null2;         /* PipelineConfig.null2 */
    int32_t threads;       /* host threads for the domain definition; 0 = all cores */
    double seq_e, dom_e;   /* SX_CUT_EVALUE thresholds */
    float seq_t, dom_t;    /* SX_CUT_BITS thresholds */
    double f1, f2, f3;     /* filter thresholds */
    double z;              /* database size for E-values (hmmsearch -Z); <= 0: number of targets */
    int32_t host_rescore;  /* 1: envelope rescoring on host threads instead of the device */
    int32_t pad;
} sx_search_cfg;

typedef struct {           /* Hit (search.rs:101-116), by index */
    uint32_t profile, target;
    uint32_t qlen, tlen;
    float score, bias;     /* bits */
    double evalue;
    float nexpected;
    uint32_t ndom;         /* this hit's domains: domains[dom_first .. dom_first + ndom) */
    uint64_t dom_first;
} sx_hit;

typedef struct {           /* DomainHit (search.rs:84-97) */
    float score, bias;
    double c_evalue, i_evalue;
    uint32_t hmm_from, hmm_to, ali_from, ali_to, env_from, env_to;
    int32_t reported;
    int32_t pad;
} sx_domain;

typedef struct sx_results sx_results;

/* Where an sx_search call spent its time (host wall clock, ms) and how much reached the
 * host part: n_pairs pairs got Forward, Backward and domain definition, n_rows = their
 * summed L+1 (special rows computed and decoded on the device), n_regions = the regions
 * that posterior decoding found on the device and handed to the host tail, bytes_to_host
 * = what those steps copied back (per-pair scores and the region list; no special rows).
 * No counterpart in the reference (it does not time its phases). */
typedef struct {
    uint64_t n_pairs, n_rows;
    double ms_filters;   /* the device cascade (0 with filters off) */
    double ms_specials;  /* device Forward + Backward and region finding, with their copies */
    double ms_tail;      /* the rest of score_target on the host threads, from the regions on */
    double ms_report;    /* report: thresholds, E-values, domain flags */
    uint64_t n_regions, bytes_to_host;
    /* sx_search_hmm_file: models read, time spent obtaining each batch (parsing and
     * configuring, or reading a cache, past what overlapped the previous batch's search)
     * and the part of it spent waiting on the parser threads */
    uint64_t n_models;
    double ms_ingest, ms_ingest_wait;
    /* envelope rescoring on the device: its time (ms, with copies; not part of ms_tail)
     * and the envelopes rescored */
    double ms_rescore;
    uint64_t n_envelopes;
} sx_search_stats;

/* Search the context's profiles against its targets. */
int sx_search(sx_ctx* ctx, const sx_search_cfg* cfg, sx_results** out);

/* The host part alone, for callers that already hold the device results: the rest of
 * score_target and `report` for the listed pairs, given their whole-sequence Forward
 * and Backward scores and special rows (5 x (L+1) floats per pair, rows N B E J C,
 * pairs back to back, as sx_specials writes them).  Models are configured from hmms;
 * targets are codes[offsets[i] .. offsets[i+1]).  Needs no device.  Errors are read
 * with sx_last_error(NULL). */
int sx_search_host(const sx_hmm* hmms, uint32_t np, const uint8_t* codes, const uint64_t* offsets, uint32_t nt,
                   const sx_search_cfg* cfg, uint64_t npairs, const uint32_t* prof, const uint32_t* tgt,
                   const double* fwd, const double* bwd, const float* fsp, const float* bsp, sx_results** out);

/* Envelope rescoring (rescore_isolated_domain, pipeline.rs:545-591) of target rows
 * i..j (1-based, inclusive) against a profile, with the unihit length model of the whole
 * target: full-matrix Forward and Backward over the envelope (dp.rs:442-601), posterior
 * decoding (dp.rs:1274-1306), the null2 odds (dp.rs:1310-1345, asked for with null2 = 1)
 * and the optimal-accuracy coordinates (dp.rs:1355-1564).  sx_search runs it for every
 * envelope of the pairs it scores; these entry points expose it on its own.  sx_rescore
 * runs on the device, one wave per envelope; sx_rescore_host runs the same steps on the
 * host (forward_matrix ... optimal_accuracy_coords in the reference's operation order). */
typedef struct {
    uint32_t profile, target;
    int32_t i, j;
    int32_t null2;
} sx_envelope;

typedef struct {
    float envsc;        /* the envelope's Forward score (nats), pipeline.rs:561 */
    float bcksc;        /* its Backward score */
    int32_t status;     /* 0 rescored; 1 Forward not finite; 2 Backward not finite (rescore returns None) */
    int32_t oa;         /* 1: the optimal-accuracy trace has an M state and gives the coordinates below;
                           0: the reference falls back to a Viterbi trace (pipeline.rs:580-588) */
    int32_t iali, jali; /* target rows (absolute) */
    int32_t kali, mali; /* model positions */
    float null2[20];    /* null2 odds of the 20 residues before the degenerate codes and the log */
} sx_envelope_score;

int sx_rescore(sx_ctx* ctx, const sx_envelope* env, uint64_t n, sx_envelope_score* out);
int sx_rescore_host(sx_ctx* ctx, const sx_envelope* env, uint64_t n, sx_envelope_score* out);

/* search_hmm_file_models (search.rs:252-271): every model of an HMMER3 file (text or
 * gzip, or a packed profile cache written by sx_write_profile_cache / sx_convert_hmm_file)
 * searched against the context's targets, `batch` models at a time (the reference takes
 * 256; 0 here means 2048, since a device launch over a few hundred profiles leaves most
 * of the chip idle; the results do not depend on it).  While the device searches a batch, host
 * threads parse and configure the next ones.  Hits refer to models by their index in the
 * file (sx_results_model); the context's profile set is the last batch afterwards.
 * Parse errors are the reference's messages (parse.rs), SX_ERR_PARSE. */
int sx_search_hmm_file(sx_ctx* ctx, const char* path, const sx_search_cfg* cfg, uint32_t batch, sx_results** out);

/* The models a search covered, by profile index: the ModelInfo list (name, acc, m) of
 * search_hmm_file_models (search.rs:245-271).  acc is NULL when absent; the strings live
 * as long as the results. */
int sx_results_models(const sx_results* r, uint32_t* n);
int sx_results_model(const sx_results* r, uint32_t i, const char** name, const char** acc, int32_t* m);

/* Packed profile cache: the context's configured profiles (what configure builds from the
 * parsed models, model.rs:146-245), written so that sx_load_hmm_file and
 * sx_search_hmm_file read them back without parsing or configuring.  No counterpart in
 * the reference.  sx_convert_hmm_file does the same from a model file without a device
 * (threads: parser threads, 0 = all cores; errors with sx_last_error(NULL)). */
int sx_write_profile_cache(const sx_ctx* ctx, const char* path);
int sx_convert_hmm_file(const char* hmm_path, const char* cache_path, int threads, uint32_t* n_models);

int sx_results_size(const sx_results* r, uint64_t* nhits, uint64_t* ndomains);
int sx_results_stats(const sx_results* r, sx_search_stats* st);
int sx_results_copy(const sx_results* r, sx_hit* hits, sx_domain* domains);
void sx_results_free(sx_results* r);

/* Lane-group shapes.  A pair is scored by a group of G lanes of one wavefront, each lane
 * holding Q consecutive model positions of the DP row in registers (capacity Q*G >= M).
 * `hoist` names the table mode: 0 = profile tables in LDS; 1 = the transition scores
 * stay in registers across rows instead of being re-read from LDS; 2 = long models
 * (M > 1024, fixed shapes), emission rows read from global memory; 3 and 4 = MSV with
 * G = 8 or G >= 16, emission rows in LDS as 16-byte-aligned lane blocks read four
 * positions per load, 512 (3) or 1024 (4) threads per workgroup.  The batched form looks a model's shape up in a per-stage
 * table of rows (max_m[i], q[i], g[i], hoist[i]) in increasing max_m; a model of length
 * M uses the first row with M <= max_m.  These calls have no counterpart in the
 * reference (a CPU loop has no launch shape).
 *
 * sx_autotune: build the table for `stage` by timing every covering shape on the
 * current profiles and targets, per model-length bucket; budget_ms bounds the time. */
int sx_autotune(sx_ctx* ctx, int stage, double budget_ms);
/* Install a table (n = 0 restores the built-in choice).  Fails if a row names a shape
 * with no kernel, whose capacity q*g is below max_m, or whose tables exceed LDS. */
int sx_set_shapes(sx_ctx* ctx, int stage, uint32_t n, const int32_t* max_m, const int32_t* q, const int32_t* g,
                  const int32_t* hoist);
/* Read the installed table: up to cap rows; n receives the row count. */
int sx_get_shapes(sx_ctx* ctx, int stage, uint32_t cap, uint32_t* n, int32_t* max_m, int32_t* q, int32_t* g,
                  int32_t* hoist);
/* The measurements of the last sx_autotune of `stage`: one row per (bucket, shape)
 * timed, with the bucket's max_m and the measured GCUPS; n receives the row count. */
int sx_autotune_log(sx_ctx* ctx, int stage, uint32_t cap, uint32_t* n, int32_t* max_m, int32_t* q, int32_t* g,
                    int32_t* hoist, double* gcups);
/* Shape used for a model length (form = SX_FORM_PLAIN or SX_FORM_BATCHED). */
int sx_shape_for(const sx_ctx* ctx, int stage, int form, int32_t m, int32_t* q, int32_t* g, int32_t* hoist);

#ifdef __cplusplus
}
#endif
#endif

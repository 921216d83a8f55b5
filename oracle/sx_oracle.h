/*
 * sx_oracle.h -- CPU restatement of sentinel-x's src/hmmer Plan7 search path.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP kernels
 * of the sentinel-x_amd package.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only to check or to time a
 * CPU baseline; no product path calls it.
 *
 * Every function restates one reference function in the same operation
 * order, in IEEE single precision where the reference uses f32, compiled with
 * -ffp-contract=off (Rust never fuses a*b+c), and with glibc logf/log/expf/exp
 * (what Rust's f32::ln / f64::ln / exp lower to on x86_64-linux-gnu).
 *
 * Layouts shared with the product library:
 *   msc, mo : NCODE x (M+1) floats, residue-major, column 0 = -inf / 0.
 *   tsc, tp : 8 x (M+1) floats in the order MM MI MD IM II DM DD BM.
 *   xsc     : 16 floats: n_loop n_move j_loop j_move c_loop c_move e_move e_loop
 *             then the same eight as probabilities (XscP).
 *   sp      : 5 x (L+1) floats: N, B, E, J, C rows.
 *
 * Pinning (what checks each function independently of this restatement) is
 * listed per function below.  Where the reference holds no golden vector for a
 * value, the function is pinned by the independent checks named there.
 */
#ifndef SX_ORACLE_H
#define SX_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SXO_K 20
#define SXO_NCODE 25
#define SXO_CODE_STAR 24

enum { SXO_TMM = 0, SXO_TMI, SXO_TMD, SXO_TIM, SXO_TII, SXO_TDM, SXO_TDD, SXO_TBM, SXO_NT };

/* alphabet.rs:58-93 -- returns the number of codes written (whitespace dropped). */
size_t sxo_encode(const uint8_t* seq, size_t n, uint8_t* out);
/* alphabet.rs:36-39 */
const float* sxo_background(void);

/* model.rs:146-245 + occupancy model.rs:248-260.
 * mat: (M+1) x 20 probabilities (row 0 unused), t: (M+1) x 7 transitions.
 * Pinned by model.rs:364-393 (entry normalisation, flat model => 0 scores,
 * '*' = -inf) and by the golden .hmm fixtures end to end. */
void sxo_configure(int m, const float* mat, const float* t,
                   float* msc, float* tsc, float* mo, float* tp, float* msv_tbm);

/* model.rs:284-312 (Xsc::new) */
void sxo_xsc(int64_t l, int multihit, float* xsc16);
/* model.rs:334-337 */
float sxo_null_score(int64_t l);

/* dp.rs:92-145 (msv_kernel).  No golden vector in the reference; pinned by a
 * generic max-product over the explicit MSV state graph, the flat-model closed
 * form and '*'-only targets (tests). */
float sxo_msv(int m, const float* msc, float msv_tbm, const uint8_t* dsq, int64_t l);

/* Quantised MSV upper bound: the product's int16 MSV prefilter, which has no
 * counterpart in the reference (its MSV is the f32 dp.rs:92-145).  Emission scores
 * are rounded up to int16 units of 1/1024 nat (-inf -> -32768) and the cells use
 * saturating int16 arithmetic; N/J/B and the length-model scores are int32 in units
 * of 2^-22 nat, rounded up.  Each rounding adds the f32 rounding allowance of the
 * value it feeds (magnitude <= 512 nats for a cell, <= 64 nats for N/J/B, which holds
 * for M <= 1024 and L <= 1e6), so the result bounds the f32 MSV from above:
 * 2^22 * sxo_msv() <= sxo_msv16() unless a row maximum reached 32767 (32 nats), in
 * which case the result is at least 32767*4096 + q(-ln2) + L*min(q(tloop),0) + q(tmove).
 * Cells are kept relative to a frame floor F >= the row's B entry, so that the max with
 * B is implied by the saturating add; F moves to B when B rises above it and on every
 * 16th row (B falls with xN), so a segment starts at most 16 rows of that fall
 * (16 * ceil(|q(tloop)| / 4096) units of 1/1024 nat) above B.  L = 0 gives INT32_MIN.  Parity of the device kernel with this function
 * is bit-exact; this function is pinned against sxo_msv by the bound and its
 * tightness (tests). */
int32_t sxo_msv16(int m, const float* msc, float msv_tbm, const uint8_t* dsq, int64_t l);
/* The quantisation steps of sxo_msv16. */
int16_t sxo_q_emit(float sc);
int32_t sxo_q_const(float v);

/* Quantised Viterbi upper bound: the product's int16 Viterbi prefilter, no counterpart
 * in the reference (its Viterbi is the f32 dp.rs:152-189).  The units and roundings of
 * sxo_msv16; transitions use sxo_q_trans (a score <= 0 stays <= 0), the transitions out
 * of node 0 are -32768 (the reference's k = 0 states are -inf), insert states are not
 * masked (at k = M or on '*', where the reference sets -inf), and the row maximum is over
 * M and D.  So the result bounds the f32 Viterbi score from above, as sxo_msv16 bounds
 * MSV, until a cell reaches 32 nats; then it stays at or above 32767*4096 + q(-ln2) +
 * L*min(q(n_loop),0) + q(n_move).  xsc16 as sxo_viterbi (multihit: J and C coincide). */
int32_t sxo_vit16(int m, const float* msc, const float* tsc, const float* xsc16, const uint8_t* dsq, int64_t l);
int16_t sxo_q_trans(float t);

/* dp.rs:152-189.  No golden vector in the reference; pinned by a generic
 * max-product over the explicit Plan7 state graph and brute-force path
 * enumeration on tiny cases (tests). */
float sxo_viterbi(int m, const float* msc, const float* tsc, const float* xsc16,
                  const uint8_t* dsq, int64_t l);

/* dp.rs:229-298.  Returns lnf(xC*c_move) + lnscale in double; the reference's
 * f32 score is (float) of it.  sp (5 x (L+1)) may be NULL. */
double sxo_forward(int m, const float* mo, const float* tp, const float* xsc16,
                   const uint8_t* dsq, int64_t l, float* sp);

/* dp.rs:301-392.  sp (5 x (L+1)) must be given. */
float sxo_backward(int m, const float* mo, const float* tp, const float* xsc16,
                   const uint8_t* dsq, int64_t l, float* sp);

/* FP64 twins of forward/backward: the same recurrences with every cell in
 * double over the same f32 parameter tables (promoted), rescaled the same
 * way.  They are the error reference for the f32 GPU Forward. */
double sxo_forward64(int m, const float* mo, const float* tp, const float* xsc16,
                     const uint8_t* dsq, int64_t l);
double sxo_backward64(int m, const float* mo, const float* tp, const float* xsc16,
                      const uint8_t* dsq, int64_t l);

/* pipeline.rs:107-150 (bias_filter_score).  No golden vector in the
 * reference; pinned by an independent log-space two-state forward (tests). */
float sxo_bias_filter(int m, const float* compo, const uint8_t* dsq, int64_t l);

/* pipeline.rs:87-95, 98-104 */
double sxo_gumbel_surv(double x, double mu, double lambda);
double sxo_exp_logsurv(double x, double tau, double lambda);
/* dp.rs:32-40 */
float sxo_logsum(float a, float b);

/* Filter cascade of score_target, pipeline.rs:159-189.
 * stats: msv mu, lambda, viterbi mu, lambda, forward tau, lambda.
 * out: [0] msv, [1] bias filtersc, [2] viterbi, [3] forward (f32 value),
 * returns the stage reached: 0 rejected by MSV, 1 by the bias-corrected MSV
 * test, 2 by Viterbi, 3 by Forward, 4 passed every filter. */
int sxo_cascade(int m, const float* msc, const float* tsc, const float* mo, const float* tp,
                float msv_tbm, const float* compo, const double* stats6,
                double f1, double f2, double f3,
                const uint8_t* dsq, int64_t l, float* out4);

/* Many-pair drivers over packed profiles and targets, used as the CPU
 * baseline (the reference fans (profile, target) pairs out over rayon,
 * search.rs:213-224).  Pairs are the flat index p*nt+t in [pair_begin,
 * pair_end); workers take chunks dynamically.  out is indexed by the flat
 * pair index minus pair_begin.
 *   prof_m[p], prof_off[p]: offset (in floats) of profile p's (M+1) columns
 *   inside the packed NCODE-row / 8-row tables (see sx_oracle.c). */
typedef struct {
    int np;
    const int32_t* m;
    const int64_t* off;      /* per-profile offset into msc/mo (NCODE*(M+1) each) */
    const int64_t* toff;     /* per-profile offset into tsc/tp (8*(M+1) each) */
    const float* msc;
    const float* tsc;
    const float* mo;
    const float* tp;
    const float* msv_tbm;
    int nt;
    const int64_t* seq_off;  /* nt+1 offsets into codes */
    const uint8_t* codes;
} sxo_batch;

int sxo_batch_msv(const sxo_batch* b, int64_t pair_begin, int64_t pair_end, float* out, int nthreads);
int sxo_batch_viterbi(const sxo_batch* b, int64_t pair_begin, int64_t pair_end, float* out, int nthreads);
int sxo_batch_forward(const sxo_batch* b, int64_t pair_begin, int64_t pair_end, double* out, int nthreads);
int sxo_batch_forward64(const sxo_batch* b, int64_t pair_begin, int64_t pair_end, double* out, int nthreads);
int sxo_batch_msv16(const sxo_batch* b, int64_t pair_begin, int64_t pair_end, int32_t* out, int nthreads);
int sxo_batch_vit16(const sxo_batch* b, int64_t pair_begin, int64_t pair_end, int32_t* out, int nthreads);
int sxo_batch_backward64(const sxo_batch* b, int64_t pair_begin, int64_t pair_end, double* out, int nthreads);
/* sxo_cascade over many pairs (the reference's search_profiles pair loop with
 * filters on, search.rs:217-224 + pipeline.rs:159-189).  stats: 6 doubles per profile;
 * compo: 20 floats per profile.  stage_out: one int per pair; sc_out: 4 floats per pair. */
int sxo_batch_cascade(const sxo_batch* b, const double* stats, const float* compo, double f1, double f2, double f3,
                      int64_t pair_begin, int64_t pair_end, int32_t* stage_out, float* sc_out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif

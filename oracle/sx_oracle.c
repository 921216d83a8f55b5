/*
 * sx_oracle.c -- CPU restatement of sentinel-x src/hmmer (alphabet.rs,
 * model.rs, dp.rs, pipeline.rs filter cascade).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the sentinel-x_amd HIP kernels
 * and the timed CPU baseline of bench.py.  No product path calls into it.
 *
 * Build: see oracle/Makefile (gcc -O3 -mavx2 -ffp-contract=off, matching the
 * reference's AVX2 auto-vectorised MSV, dp.rs:81-85, and Rust's no-fusion
 * floating point).  See sx_oracle.h for layouts and the pin of each function.
 */
#include "sx_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

#define NEG_INF (-INFINITY)
#define LN2F ((float)0.69314718055994530942)
#define RESCALE 1.0e4f

/* alphabet.rs:36-39 -- HMMER3 amino background. */
static const float BG[SXO_K] = {
    0.0787945f, 0.0151600f, 0.0535222f, 0.0668298f, 0.0397062f, 0.0695071f, 0.0229198f,
    0.0590092f, 0.0594422f, 0.0963728f, 0.0237718f, 0.0414386f, 0.0482904f, 0.0395639f,
    0.0540978f, 0.0683364f, 0.0540687f, 0.0673417f, 0.0114135f, 0.0304133f,
};

const float* sxo_background(void) { return BG; }

/* alphabet.rs:42-54 */
static int degenerate_members(int code, const int** mem) {
    static const int B[2] = {2, 11}, J[2] = {7, 9}, Z[2] = {3, 13};
    static const int X[20] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19};
    switch (code) {
        case 20: *mem = B; return 2;
        case 21: *mem = J; return 2;
        case 22: *mem = Z; return 2;
        case 23: *mem = X; return 20;
        default: *mem = NULL; return 0;
    }
}

/* alphabet.rs:58-88 */
static uint8_t encode_residue(uint8_t b) {
    if (b >= 'a' && b <= 'z') b = (uint8_t)(b - 'a' + 'A');
    switch (b) {
        case 'A': return 0;
        case 'C': case 'U': return 1;
        case 'D': return 2;
        case 'E': return 3;
        case 'F': return 4;
        case 'G': return 5;
        case 'H': return 6;
        case 'I': return 7;
        case 'K': case 'O': return 8;
        case 'L': return 9;
        case 'M': return 10;
        case 'N': return 11;
        case 'P': return 12;
        case 'Q': return 13;
        case 'R': return 14;
        case 'S': return 15;
        case 'T': return 16;
        case 'V': return 17;
        case 'W': return 18;
        case 'Y': return 19;
        case 'B': return 20;
        case 'J': return 21;
        case 'Z': return 22;
        case 'X': return 23;
        case '*': case '-': case '.': case '_': case '~': return 24;
        default: break;
    }
    if (b >= 'A' && b <= 'Z') return 23;
    return 24;
}

/* alphabet.rs:91-93 -- ASCII whitespace (Rust is_ascii_whitespace) is dropped. */
size_t sxo_encode(const uint8_t* seq, size_t n, uint8_t* out) {
    size_t o = 0;
    for (size_t i = 0; i < n; i++) {
        uint8_t b = seq[i];
        if (b == ' ' || b == '\t' || b == '\n' || b == '\r' || b == '\f') continue;
        out[o++] = encode_residue(b);
    }
    return o;
}

static inline float fmaxs(float a, float b) { return a > b ? a : b; } /* dp.rs:21-27 */
static inline float lnf(float x) { return x > 0.0f ? logf(x) : NEG_INF; } /* dp.rs:200-206 */
static inline double lnd(double x) { return x > 0.0 ? log(x) : -INFINITY; }

/* model.rs:248-260 */
static void occupancy(int m, const float* t, float* occ) {
    for (int k = 0; k <= m; k++) occ[k] = 0.0f;
    if (m == 0) return;
    occ[1] = t[0 * 7 + 1] + t[0 * 7 + 0]; /* T_MI + T_MM */
    for (int k = 2; k <= m; k++) {
        const float* tk = t + (size_t)(k - 1) * 7;
        occ[k] = occ[k - 1] * (tk[0] + tk[1]) + (1.0f - occ[k - 1]) * tk[5];
    }
}

static inline float lnp32(float p) { return p > 0.0f ? (float)log((double)p) : NEG_INF; }

/* model.rs:146-245 */
void sxo_configure(int m, const float* mat, const float* t, float* msc, float* tsc, float* mo,
                   float* tp, float* msv_tbm) {
    const int w = m + 1;
    for (int i = 0; i < SXO_NCODE * w; i++) msc[i] = NEG_INF;
    for (int k = 1; k <= m; k++) {
        float sc[SXO_NCODE];
        for (int x = 0; x < SXO_K; x++) {
            double p = (double)mat[(size_t)k * SXO_K + x];
            sc[x] = p > 0.0 ? (float)log(p / (double)BG[x]) : NEG_INF;
        }
        for (int code = 20; code <= 23; code++) {
            const int* mem;
            int nm = degenerate_members(code, &mem);
            float num = 0.0f, den = 0.0f;
            for (int j = 0; j < nm; j++) {
                num += sc[mem[j]] * BG[mem[j]];
                den += BG[mem[j]];
            }
            sc[code] = den > 0.0f ? num / den : 0.0f;
        }
        sc[SXO_CODE_STAR] = NEG_INF;
        for (int x = 0; x < SXO_NCODE; x++) msc[(size_t)x * w + k] = sc[x];
    }
    for (int i = 0; i < SXO_NT * w; i++) tsc[i] = NEG_INF;
    /* node transitions for k = 1..M-1; on-disk order MM MI MD IM II DM DD
     * equals the packed order of the first seven rows. */
    for (int k = 1; k < m; k++)
        for (int j = 0; j < 7; j++) tsc[(size_t)j * w + k] = lnp32(t[(size_t)k * 7 + j]);
    if (m >= 1) {
        tsc[(size_t)SXO_TDM * w + 1] = NEG_INF;
        tsc[(size_t)SXO_TDD * w + 1] = NEG_INF;
    }
    float* occ = (float*)malloc(sizeof(float) * (size_t)w);
    occupancy(m, t, occ);
    float z = 0.0f;
    for (int k = 1; k <= m; k++) z += occ[k] * (float)(m - k + 1);
    for (int k = 1; k <= m; k++)
        tsc[(size_t)SXO_TBM * w + k] =
            (z > 0.0f && occ[k] > 0.0f) ? (float)log((double)(occ[k] / z)) : NEG_INF;
    free(occ);
    *msv_tbm = logf(2.0f / ((float)m * ((float)m + 1.0f)));
    for (int i = 0; i < SXO_NCODE * w; i++) mo[i] = expf(msc[i]);
    for (int i = 0; i < SXO_NT * w; i++) tp[i] = expf(tsc[i]);
}

/* model.rs:284-312 */
void sxo_xsc(int64_t l, int multihit, float* x) {
    float nj = multihit ? 1.0f : 0.0f;
    float pmove = (2.0f + nj) / ((float)l + 2.0f + nj);
    float ploop = 1.0f - pmove;
    float lloop = (float)log((double)ploop);
    float lmove = (float)log((double)pmove);
    x[0] = lloop; x[1] = lmove; x[2] = lloop; x[3] = lmove; x[4] = lloop; x[5] = lmove;
    x[6] = multihit ? -LN2F : 0.0f;
    x[7] = multihit ? -LN2F : NEG_INF;
    x[8] = ploop; x[9] = pmove; x[10] = ploop; x[11] = pmove; x[12] = ploop; x[13] = pmove;
    x[14] = multihit ? 0.5f : 1.0f;
    x[15] = multihit ? 0.5f : 0.0f;
}

/* model.rs:334-337 */
float sxo_null_score(int64_t l) {
    float p1 = (float)l / ((float)l + 1.0f);
    return (float)((double)l * log((double)p1) + log(1.0 - (double)p1));
}

/* dp.rs:32-40 */
float sxo_logsum(float a, float b) {
    float hi = a > b ? a : b, lo = a > b ? b : a;
    float d = hi - lo;
    if (!(d < 30.0f)) return hi;
    return hi + log1pf(expf(-d));
}

static double logsum64(double a, double b) {
    double hi = a > b ? a : b, lo = a > b ? b : a;
    double d = hi - lo;
    if (!(d < 30.0)) return hi;
    return hi + log1p(exp(-d));
}

/* dp.rs:92-145 (msv_kernel, with its 8-wide chunking of the inner loop). */
float sxo_msv(int m, const float* msc, float tbmk, const uint8_t* dsq, int64_t l) {
    if (l == 0 || m == 0) return NEG_INF;
    const int w = m + 1;
    float lf = (float)l;
    float tloop = logf(lf / (lf + 3.0f));
    float tmove = logf(3.0f / (lf + 3.0f));
    const float tec = -LN2F, tej = -LN2F;
    float* prev = (float*)malloc(sizeof(float) * (size_t)w * 2);
    float* cur = prev + w;
    for (int k = 0; k < 2 * w; k++) prev[k] = NEG_INF;
    float xn = 0.0f, xb = tmove, xj = NEG_INF, xc = NEG_INF;
    for (int64_t i = 0; i < l; i++) {
        const float* rs = msc + (size_t)dsq[i] * w + 1;
        float be = xb + tbmk;
        float e8[8];
        for (int j = 0; j < 8; j++) e8[j] = NEG_INF;
        int nch = m / 8;
        for (int c = 0; c < nch; c++) {
            for (int j = 0; j < 8; j++) {
                int k = c * 8 + j;
                float v = rs[k] + fmaxs(prev[k], be);
                cur[k + 1] = v;
                e8[j] = fmaxs(e8[j], v);
            }
        }
        for (int k = nch * 8; k < m; k++) {
            float v = rs[k] + fmaxs(prev[k], be);
            cur[k + 1] = v;
            e8[0] = fmaxs(e8[0], v);
        }
        cur[0] = NEG_INF;
        float xe = NEG_INF;
        for (int j = 0; j < 8; j++) xe = fmaxs(xe, e8[j]);
        xj = fmaxs(xj + tloop, xe + tej);
        xc = fmaxs(xc + tloop, xe + tec);
        xn += tloop;
        xb = fmaxs(xn + tmove, xj + tmove);
        float* s = prev; prev = cur; cur = s;
    }
    free(prev < cur ? prev : cur);
    return xc + tmove;
}

/* Quantised MSV upper bound (see sx_oracle.h).  No reference counterpart: the
 * recurrence is dp.rs:92-145 (as sxo_msv) with every score rounded up. */
#define SXQ_S 1024.0          /* emission units per nat */
#define SXQ_F 4194304.0       /* 2^22 special-state units per nat (SXQ_S * 4096) */
#define SXQ_SLACK_S 0.03125   /* SXQ_S * 2^-24 * 512: f32 rounding allowance of a cell */
#define SXQ_SLACK_F 16.0      /* SXQ_F * 2^-24 * 64: of an N/J/B step */

int16_t sxo_q_emit(float sc) {
    if (sc == NEG_INF) return -32768;
    double v = ceil((double)sc * SXQ_S + SXQ_SLACK_S);
    if (v < -32768.0) v = -32768.0;
    if (v > 32767.0) v = 32767.0;
    return (int16_t)v;
}

int32_t sxo_q_const(float v) { return (int32_t)ceil((double)v * SXQ_F + SXQ_SLACK_F); }

static inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
static inline int32_t sat16(int32_t v) { return v < -32768 ? -32768 : v > 32767 ? 32767 : v; }
/* ceil(v / 4096) for any sign */
static inline int32_t ceil4096(int32_t v) { return (int32_t)floor(((double)v + 4095.0) / 4096.0); }

/* Cells hold t = s - F - 32768 for a frame floor F >= the row's B entry, so
 * max(s[k-1], B) + e[k] is the saturating sum t[k-1] + e[k]; F moves to B on rows where
 * B exceeds it and on every 16th row (i % 16 == 0) where B is below it, shifting every
 * cell and position 0 (the entry to position 1) by the move. */
static inline void q_shift(int32_t* v, int n, int32_t d) {
    const int32_t d1 = d > 32767 ? 32767 : d, d2 = d - d1;
    for (int k = 0; k < n; k++) v[k] = sat16(v[k] - d1);
    if (d2 > 0)
        for (int k = 0; k < n; k++) v[k] = sat16(v[k] - d2);
}

int32_t sxo_msv16(int m, const float* msc, float tbmk, const uint8_t* dsq, int64_t l) {
    if (l == 0) return INT32_MIN;
    const int w = m + 1;
    const float lf = (float)l;
    const int32_t tl = sxo_q_const(logf(lf / (lf + 3.0f)));
    const int32_t tm = sxo_q_const(logf(3.0f / (lf + 3.0f)));
    const int32_t tb = sxo_q_const(tbmk);
    const int32_t te = sxo_q_const(-LN2F);
    int16_t* qe = (int16_t*)malloc(sizeof(int16_t) * (size_t)w * SXO_NCODE);
    for (int k = 0; k < w * SXO_NCODE; k++) qe[k] = sxo_q_emit(msc[k]);
    /* positions 0..M in frame units; position 0 is the floor, shifted with the rest */
    int32_t* prev = (int32_t*)malloc(sizeof(int32_t) * (size_t)w * 2);
    int32_t* cur = prev + w;
    for (int k = 0; k < 2 * w; k++) prev[k] = -32768;
    int32_t xn = 0, xj = -(1 << 30), xb = imax(xn, xj) + tm;
    int32_t fl = ceil4096(xb + tb);
    for (int64_t i = 0; i < l; i++) {
        const int16_t* rs = qe + (size_t)dsq[i] * w;
        const int32_t be = ceil4096(xb + tb);
        if (be > fl || (i % 16 == 0 && be < fl)) {
            q_shift(prev, w, be - fl);
            fl = be;
        }
        int32_t xe = -32768;
        for (int k = 1; k <= m; k++) {
            const int32_t v = sat16(prev[k - 1] + rs[k]);
            cur[k] = v;
            xe = imax(xe, v);
        }
        cur[0] = -32768;
        xe = xe + fl + 32768;
        if (xe > 32767) xe = 32767;
        xj = imax(xj + tl, xe * 4096 + te);
        xn += tl;
        xb = imax(xn, xj) + tm;
        int32_t* s = prev; prev = cur; cur = s;
    }
    free(prev < cur ? prev : cur);
    free(qe);
    return xj + tm;
}

/* Quantised Viterbi upper bound (see sx_oracle.h); the recurrence of dp.rs:152-189 with
 * every score rounded up. */
int16_t sxo_q_trans(float t) {
    const int16_t q = sxo_q_emit(t);
    /* a score <= 0 cannot raise a value: fl(a + t) <= a, so 0 is a valid rounding */
    return (t <= 0.0f && q > 0) ? 0 : q;
}

static inline int32_t q_lconst(float v) { return (int32_t)ceil((double)v * SXQ_F + SXQ_SLACK_F); }

int32_t sxo_vit16(int m, const float* msc, const float* tsc, const float* xs, const uint8_t* dsq, int64_t l) {
    if (l == 0) return INT32_MIN;
    const int w = m + 1;
    const int32_t tl = q_lconst(xs[0]), tm = q_lconst(xs[1]), te = sxo_q_const(-LN2F);
    /* transitions by row MM MI MD IM II DM DD BM; out of node 0 (MM MD IM DM DD at k = 0,
     * read by position 1) the reference's states are -inf, so those scores are -32768 */
    int16_t* qt = (int16_t*)malloc(sizeof(int16_t) * (size_t)8 * w);
    for (int r = 0; r < 8; r++)
        for (int k = 0; k <= m; k++) qt[r * w + k] = (k == 0) ? -32768 : sxo_q_trans(tsc[r * w + k]);
    const int16_t *tmm = qt, *tmi = qt + w, *tmd = qt + 2 * w, *tim = qt + 3 * w, *tii = qt + 4 * w,
                  *tdm = qt + 5 * w, *tdd = qt + 6 * w, *tbm = qt + 7 * w;
    int32_t* buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)w * 6);
    for (int k = 0; k < 6 * w; k++) buf[k] = -32768;
    int32_t *mp = buf, *ip = buf + w, *dp = buf + 2 * w, *mc = buf + 3 * w, *ic = buf + 4 * w, *dc = buf + 5 * w;
    int32_t xn = 0, xj = -(1 << 30), xb = imax(xn, xj) + tm;
    for (int64_t i = 0; i < l; i++) {
        const float* rs = msc + (size_t)dsq[i] * w;
        const int32_t bq = imax(ceil4096(xb), -32768);
        int32_t xe = -32768;
        mc[0] = ic[0] = dc[0] = -32768;
        for (int k = 1; k <= m; k++) {
            const int32_t a = sat16(mp[k - 1] + tmm[k - 1]), b = sat16(ip[k - 1] + tim[k - 1]);
            const int32_t c = sat16(bq + tbm[k]), d = sat16(dp[k - 1] + tdm[k - 1]);
            const int32_t mv = sat16(imax(imax(a, b), imax(c, d)) + sxo_q_emit(rs[k]));
            mc[k] = mv;
            ic[k] = imax(sat16(mp[k] + tmi[k]), sat16(ip[k] + tii[k]));
            const int32_t dv = imax(sat16(mc[k - 1] + tmd[k - 1]), sat16(dc[k - 1] + tdd[k - 1]));
            dc[k] = dv;
            xe = imax(xe, imax(mv, dv));
        }
        xj = imax(xj + tl, xe * 4096 + te);
        xn += tl;
        xb = imax(xn, xj) + tm;
        int32_t* t;
        t = mp; mp = mc; mc = t;
        t = ip; ip = ic; ic = t;
        t = dp; dp = dc; dc = t;
    }
    free(buf);
    free(qt);
    return xj + tm;
}

/* dp.rs:152-189 */
float sxo_viterbi(int m, const float* msc, const float* tsc, const float* xs, const uint8_t* dsq,
                  int64_t l) {
    if (l == 0 || m == 0) return NEG_INF;
    const int w = m + 1;
    const float *t_mm = tsc, *t_mi = tsc + w, *t_md = tsc + 2 * w, *t_im = tsc + 3 * w;
    const float *t_ii = tsc + 4 * w, *t_dm = tsc + 5 * w, *t_dd = tsc + 6 * w, *t_bm = tsc + 7 * w;
    float* buf = (float*)malloc(sizeof(float) * (size_t)w * 6);
    for (int k = 0; k < 6 * w; k++) buf[k] = NEG_INF;
    float *mp = buf, *ip = buf + w, *dp = buf + 2 * w, *mc = buf + 3 * w, *ic = buf + 4 * w,
          *dc = buf + 5 * w;
    float xn = 0.0f, xb = xs[1], xj = NEG_INF, xc = NEG_INF;
    for (int64_t i = 0; i < l; i++) {
        const uint8_t x = dsq[i];
        const float* rs = msc + (size_t)x * w;
        int ins_ok = x != SXO_CODE_STAR;
        mc[0] = NEG_INF; ic[0] = NEG_INF; dc[0] = NEG_INF;
        float xe = NEG_INF;
        for (int k = 1; k <= m; k++) {
            float sc = fmaxs(fmaxs(mp[k - 1] + t_mm[k - 1], ip[k - 1] + t_im[k - 1]),
                             fmaxs(xb + t_bm[k], dp[k - 1] + t_dm[k - 1]));
            float mv = sc + rs[k];
            mc[k] = mv;
            ic[k] = (k < m && ins_ok) ? fmaxs(mp[k] + t_mi[k], ip[k] + t_ii[k]) : NEG_INF;
            float dv = fmaxs(mc[k - 1] + t_md[k - 1], dc[k - 1] + t_dd[k - 1]);
            dc[k] = dv;
            xe = fmaxs(xe, fmaxs(mv, dv));
        }
        xj = fmaxs(xj + xs[2], xe + xs[7]);
        xc = fmaxs(xc + xs[4], xe + xs[6]);
        xn += xs[0];
        xb = fmaxs(xn + xs[1], xj + xs[3]);
        float* s;
        s = mp; mp = mc; mc = s;
        s = ip; ip = ic; ic = s;
        s = dp; dp = dc; dc = s;
    }
    free(buf);
    return xc + xs[5];
}

/* dp.rs:229-298 */
double sxo_forward(int m, const float* mo, const float* tp, const float* xs, const uint8_t* dsq,
                   int64_t l, float* sp) {
    if (l == 0 || m == 0) return -INFINITY;
    const int w = m + 1;
    const float* xp = xs + 8;
    const float *p_mm = tp, *p_mi = tp + w, *p_md = tp + 2 * w, *p_im = tp + 3 * w;
    const float *p_ii = tp + 4 * w, *p_dm = tp + 5 * w, *p_dd = tp + 6 * w, *p_bm = tp + 7 * w;
    float *sn = NULL, *sb = NULL, *se = NULL, *sj = NULL, *sc = NULL;
    if (sp) {
        for (int64_t i = 0; i < 5 * (l + 1); i++) sp[i] = NEG_INF;
        sn = sp; sb = sp + (l + 1); se = sp + 2 * (l + 1); sj = sp + 3 * (l + 1); sc = sp + 4 * (l + 1);
    }
    float* buf = (float*)calloc((size_t)w * 6, sizeof(float));
    float *mp = buf, *ip = buf + w, *dp = buf + 2 * w, *mc = buf + 3 * w, *ic = buf + 4 * w,
          *dc = buf + 5 * w;
    float xn = 1.0f, xb = xp[1], xj = 0.0f, xc = 0.0f;
    double lnscale = 0.0;
    if (sp) { sn[0] = 0.0f; sb[0] = lnf(xb); }
    for (int64_t i0 = 0; i0 < l; i0++) {
        const uint8_t x = dsq[i0];
        const float* eo = mo + (size_t)x * w;
        float ins = x != SXO_CODE_STAR ? 1.0f : 0.0f;
        float xe = 0.0f;
        for (int k = 1; k <= m; k++) {
            float mv = (xb * p_bm[k] + mp[k - 1] * p_mm[k - 1] + ip[k - 1] * p_im[k - 1] +
                        dp[k - 1] * p_dm[k - 1]) * eo[k];
            mc[k] = mv;
            ic[k] = k < m ? (mp[k] * p_mi[k] + ip[k] * p_ii[k]) * ins : 0.0f;
            float dv = mc[k - 1] * p_md[k - 1] + dc[k - 1] * p_dd[k - 1];
            dc[k] = dv;
            xe += mv + dv;
        }
        xj = xj * xp[2] + xe * xp[7];
        xc = xc * xp[4] + xe * xp[6];
        xn *= xp[0];
        xb = xn * xp[1] + xj * xp[3];
        if (xe > RESCALE) {
            float inv = 1.0f / xe;
            for (int k = 1; k <= m; k++) { mc[k] *= inv; ic[k] *= inv; dc[k] *= inv; }
            xn *= inv; xb *= inv; xj *= inv; xc *= inv;
            lnscale += log((double)xe);
            xe = 1.0f;
        }
        if (sp) {
            int64_t i = i0 + 1;
            float ls = (float)lnscale;
            sn[i] = lnf(xn) + ls; sb[i] = lnf(xb) + ls; se[i] = lnf(xe) + ls;
            sj[i] = lnf(xj) + ls; sc[i] = lnf(xc) + ls;
        }
        float* s;
        s = mp; mp = mc; mc = s;
        s = ip; ip = ic; ic = s;
        s = dp; dp = dc; dc = s;
    }
    free(buf);
    return (double)lnf(xc * xp[5]) + lnscale;
}

/* dp.rs:301-392 */
float sxo_backward(int m, const float* mo, const float* tp, const float* xs, const uint8_t* dsq,
                   int64_t l, float* sp) {
    for (int64_t i = 0; i < 5 * (l + 1); i++) sp[i] = NEG_INF;
    if (l == 0 || m == 0) return NEG_INF;
    float *sn = sp, *sb = sp + (l + 1), *se = sp + 2 * (l + 1), *sj = sp + 3 * (l + 1),
          *sc = sp + 4 * (l + 1);
    const int w = m + 1;
    const float* xp = xs + 8;
    const float *p_mm = tp, *p_mi = tp + w, *p_md = tp + 2 * w, *p_im = tp + 3 * w;
    const float *p_ii = tp + 4 * w, *p_dm = tp + 5 * w, *p_dd = tp + 6 * w, *p_bm = tp + 7 * w;
    float* buf = (float*)calloc((size_t)w * 6, sizeof(float));
    float *mn = buf, *inx = buf + w, *dn = buf + 2 * w, *mc = buf + 3 * w, *ic = buf + 4 * w,
          *dc = buf + 5 * w;
    float xc = xp[5];
    float xe = xc * xp[6];
    float xj = 0.0f, xn = 0.0f, xb;
    mn[m] = xe;
    dn[m] = xe;
    for (int k = m - 1; k >= 1; k--) {
        mn[k] = xe + dn[k + 1] * p_md[k];
        dn[k] = xe + dn[k + 1] * p_dd[k];
    }
    double lnscale = 0.0;
    sc[l] = lnf(xc);
    se[l] = lnf(xe);
    for (int64_t i = l - 1; i >= 1; i--) {
        const uint8_t xnext = dsq[i];
        const float* eo = mo + (size_t)xnext * w;
        float ins = xnext != SXO_CODE_STAR ? 1.0f : 0.0f;
        float b = 0.0f;
        for (int k = 1; k <= m; k++) b += mn[k] * p_bm[k] * eo[k];
        xb = b;
        xj = xj * xp[2] + xb * xp[3];
        xc *= xp[4];
        xe = xj * xp[7] + xc * xp[6];
        xn = xn * xp[0] + xb * xp[1];
        mc[m] = xe;
        dc[m] = xe;
        ic[m] = 0.0f;
        float rowmax = fmaxf(fmaxf(fmaxf(fmaxf(xb, xe), xn), xj), xc);
        for (int k = m - 1; k >= 1; k--) {
            float mnext = mn[k + 1] * eo[k + 1];
            float iv = inx[k] * ins;
            float mv = mnext * p_mm[k] + iv * p_mi[k] + xe + dc[k + 1] * p_md[k];
            mc[k] = mv;
            ic[k] = mnext * p_im[k] + iv * p_ii[k];
            dc[k] = mnext * p_dm[k] + dc[k + 1] * p_dd[k] + xe;
            if (mv > rowmax) rowmax = mv;
        }
        if (rowmax > RESCALE) {
            float inv = 1.0f / rowmax;
            for (int k = 1; k <= m; k++) { mc[k] *= inv; ic[k] *= inv; dc[k] *= inv; }
            xb *= inv; xe *= inv; xn *= inv; xj *= inv; xc *= inv;
            lnscale += log((double)rowmax);
        }
        float ls = (float)lnscale;
        sn[i] = lnf(xn) + ls; sb[i] = lnf(xb) + ls; se[i] = lnf(xe) + ls;
        sj[i] = lnf(xj) + ls; sc[i] = lnf(xc) + ls;
        float* s;
        s = mn; mn = mc; mc = s;
        s = inx; inx = ic; ic = s;
        s = dn; dn = dc; dc = s;
    }
    const float* eo = mo + (size_t)dsq[0] * w;
    float b = 0.0f;
    for (int k = 1; k <= m; k++) b += mn[k] * p_bm[k] * eo[k];
    float ls = (float)lnscale;
    sb[0] = lnf(b) + ls;
    sn[0] = sxo_logsum(sn[1] + xs[0], sb[0] + xs[1]);
    free(buf);
    return sn[0];
}

/* FP64 twin of dp.rs:229-298: same recurrence, every cell in double. */
double sxo_forward64(int m, const float* mo, const float* tp, const float* xs, const uint8_t* dsq,
                     int64_t l) {
    if (l == 0 || m == 0) return -INFINITY;
    const int w = m + 1;
    const float* xp = xs + 8;
    const float *p_mm = tp, *p_mi = tp + w, *p_md = tp + 2 * w, *p_im = tp + 3 * w;
    const float *p_ii = tp + 4 * w, *p_dm = tp + 5 * w, *p_dd = tp + 6 * w, *p_bm = tp + 7 * w;
    double* buf = (double*)calloc((size_t)w * 6, sizeof(double));
    double *mp = buf, *ip = buf + w, *dp = buf + 2 * w, *mc = buf + 3 * w, *ic = buf + 4 * w,
           *dc = buf + 5 * w;
    double xn = 1.0, xb = xp[1], xj = 0.0, xc = 0.0, lnscale = 0.0;
    for (int64_t i0 = 0; i0 < l; i0++) {
        const uint8_t x = dsq[i0];
        const float* eo = mo + (size_t)x * w;
        double ins = x != SXO_CODE_STAR ? 1.0 : 0.0;
        double xe = 0.0;
        for (int k = 1; k <= m; k++) {
            double mv = ((double)xb * p_bm[k] + mp[k - 1] * p_mm[k - 1] + ip[k - 1] * p_im[k - 1] +
                         dp[k - 1] * p_dm[k - 1]) * eo[k];
            mc[k] = mv;
            ic[k] = k < m ? (mp[k] * p_mi[k] + ip[k] * p_ii[k]) * ins : 0.0;
            double dv = mc[k - 1] * p_md[k - 1] + dc[k - 1] * p_dd[k - 1];
            dc[k] = dv;
            xe += mv + dv;
        }
        xj = xj * xp[2] + xe * xp[7];
        xc = xc * xp[4] + xe * xp[6];
        xn *= xp[0];
        xb = xn * xp[1] + xj * xp[3];
        if (xe > 1.0e4) {
            double inv = 1.0 / xe;
            for (int k = 1; k <= m; k++) { mc[k] *= inv; ic[k] *= inv; dc[k] *= inv; }
            xn *= inv; xb *= inv; xj *= inv; xc *= inv;
            lnscale += log(xe);
        }
        double* s;
        s = mp; mp = mc; mc = s;
        s = ip; ip = ic; ic = s;
        s = dp; dp = dc; dc = s;
    }
    free(buf);
    return lnd(xc * xp[5]) + lnscale;
}

/* FP64 twin of dp.rs:301-392 (score only). */
double sxo_backward64(int m, const float* mo, const float* tp, const float* xs, const uint8_t* dsq,
                      int64_t l) {
    if (l == 0 || m == 0) return -INFINITY;
    const int w = m + 1;
    const float* xp = xs + 8;
    const float *p_mm = tp, *p_mi = tp + w, *p_md = tp + 2 * w, *p_im = tp + 3 * w;
    const float *p_ii = tp + 4 * w, *p_dm = tp + 5 * w, *p_dd = tp + 6 * w, *p_bm = tp + 7 * w;
    double* buf = (double*)calloc((size_t)w * 6, sizeof(double));
    double *mn = buf, *inx = buf + w, *dn = buf + 2 * w, *mc = buf + 3 * w, *ic = buf + 4 * w,
           *dc = buf + 5 * w;
    double xc = xp[5];
    double xe = xc * xp[6];
    double xj = 0.0, xn = 0.0, xb;
    mn[m] = xe;
    dn[m] = xe;
    for (int k = m - 1; k >= 1; k--) {
        mn[k] = xe + dn[k + 1] * p_md[k];
        dn[k] = xe + dn[k + 1] * p_dd[k];
    }
    double lnscale = 0.0, ln_n1 = -INFINITY;
    for (int64_t i = l - 1; i >= 1; i--) {
        const uint8_t xnext = dsq[i];
        const float* eo = mo + (size_t)xnext * w;
        double ins = xnext != SXO_CODE_STAR ? 1.0 : 0.0;
        double b = 0.0;
        for (int k = 1; k <= m; k++) b += mn[k] * p_bm[k] * eo[k];
        xb = b;
        xj = xj * xp[2] + xb * xp[3];
        xc *= xp[4];
        xe = xj * xp[7] + xc * xp[6];
        xn = xn * xp[0] + xb * xp[1];
        mc[m] = xe;
        dc[m] = xe;
        ic[m] = 0.0;
        double rowmax = fmax(fmax(fmax(fmax(xb, xe), xn), xj), xc);
        for (int k = m - 1; k >= 1; k--) {
            double mnext = mn[k + 1] * eo[k + 1];
            double iv = inx[k] * ins;
            double mv = mnext * p_mm[k] + iv * p_mi[k] + xe + dc[k + 1] * p_md[k];
            mc[k] = mv;
            ic[k] = mnext * p_im[k] + iv * p_ii[k];
            dc[k] = mnext * p_dm[k] + dc[k + 1] * p_dd[k] + xe;
            if (mv > rowmax) rowmax = mv;
        }
        if (rowmax > 1.0e4) {
            double inv = 1.0 / rowmax;
            for (int k = 1; k <= m; k++) { mc[k] *= inv; ic[k] *= inv; dc[k] *= inv; }
            xb *= inv; xe *= inv; xn *= inv; xj *= inv; xc *= inv;
            lnscale += log(rowmax);
        }
        if (i == 1) ln_n1 = lnd(xn) + lnscale;
        double* s;
        s = mn; mn = mc; mc = s;
        s = inx; inx = ic; ic = s;
        s = dn; dn = dc; dc = s;
    }
    const float* eo = mo + (size_t)dsq[0] * w;
    double b = 0.0;
    for (int k = 1; k <= m; k++) b += mn[k] * p_bm[k] * eo[k];
    double b0 = lnd(b) + lnscale;
    free(buf);
    return logsum64(ln_n1 + (double)xs[0], b0 + (double)xs[1]);
}

/* pipeline.rs:107-150 */
float sxo_bias_filter(int m, const float* compo, const uint8_t* dsq, int64_t l) {
    const double l0 = 400.0;
    const double l1 = (double)m / 8.0;
    const double t00 = l0 / (l0 + 1.0), t01 = 1.0 / (l0 + 1.0);
    const double t10 = 1.0 / (l1 + 1.0), t11 = l1 / (l1 + 1.0);
    const double pi0 = 0.999, pi1 = 0.001;
    double eo[SXO_NCODE][2];
    for (int x = 0; x < SXO_NCODE; x++) { eo[x][0] = 1.0; eo[x][1] = 1.0; }
    for (int x = 0; x < SXO_K; x++) {
        eo[x][0] = 1.0;
        eo[x][1] = (double)compo[x] / (double)BG[x];
    }
    for (int code = 20; code <= 23; code++) {
        const int* mem;
        int nm = degenerate_members(code, &mem);
        double e0 = 0, e1 = 0, z = 0;
        for (int j = 0; j < nm; j++) {
            e0 += (double)BG[mem[j]];
            e1 += (double)compo[mem[j]];
            z += (double)BG[mem[j]];
        }
        eo[code][0] = e0 / z;
        eo[code][1] = e1 / z;
    }
    eo[SXO_CODE_STAR][0] = 1.0;
    eo[SXO_CODE_STAR][1] = 1.0;
    double logsc = 0.0;
    int x1 = dsq[0];
    double d0 = eo[x1][0] * pi0, d1 = eo[x1][1] * pi1;
    double mx = fmax(fmax(d0, d1), 2.2250738585072014e-308);
    d0 /= mx; d1 /= mx;
    logsc += log(mx);
    for (int64_t i = 1; i < l; i++) {
        int x = dsq[i];
        double n0 = (d0 * t00 + d1 * t10) * eo[x][0];
        double n1 = (d0 * t01 + d1 * t11) * eo[x][1];
        mx = fmax(fmax(n0, n1), 2.2250738585072014e-308);
        d0 = n0 / mx; d1 = n1 / mx;
        logsc += log(mx);
    }
    logsc += log(d0 + d1);
    double p1 = (double)l / ((double)l + 1.0);
    return (float)(logsc + (double)l * log(p1) + log(1.0 - p1));
}

/* pipeline.rs:87-95 */
double sxo_gumbel_surv(double x, double mu, double lambda) {
    double y = lambda * (x - mu);
    double ey = -exp(-y);
    if (fabs(ey) < 1e-5) return -ey;
    return 1.0 - exp(ey);
}

/* pipeline.rs:98-104 */
double sxo_exp_logsurv(double x, double tau, double lambda) {
    if (x < tau) return 0.0;
    return -lambda * (x - tau);
}

/* SXO_PLANT selects a deliberate mistake in the cascade's decision logic, for the
 * mutation tests only (tests/test_oracle_cascade.py builds each into a scratch library
 * and checks that the closed-form boundary tests catch it): 1 swaps F2 and F3, 2 uses the
 * null score where the bias filter score belongs, 3 rejects on >= instead of >.  Every
 * real build has 0. */
#ifndef SXO_PLANT
#define SXO_PLANT 0
#endif
#if SXO_PLANT == 3
#define SXO_REJECTS(p, f) ((p) >= (f))
#else
#define SXO_REJECTS(p, f) ((p) > (f))
#endif

/* pipeline.rs:159-189 (the cfg.filters && calibrated branch). */
int sxo_cascade(int m, const float* msc, const float* tsc, const float* mo, const float* tp,
                float msv_tbm, const float* compo, const double* st, double f1, double f2, double f3,
                const uint8_t* dsq, int64_t l, float* out) {
    out[0] = out[1] = out[2] = out[3] = NAN;
    if (l == 0 || m == 0) return 0;
    if (SXO_PLANT == 1) {
        const double f = f2;
        f2 = f3;
        f3 = f;
    }
    float nullsc = sxo_null_score(l);
    float xs[16];
    sxo_xsc(l, 1, xs);
    float usc = sxo_msv(m, msc, msv_tbm, dsq, l);
    out[0] = usc;
    if (SXO_REJECTS(sxo_gumbel_surv((double)((usc - nullsc) / LN2F), st[0], st[1]), f1)) return 0;
    float filtersc = sxo_bias_filter(m, compo, dsq, l);
    out[1] = filtersc;
    if (SXO_PLANT == 2) filtersc = nullsc;
    if (SXO_REJECTS(sxo_gumbel_surv((double)((usc - filtersc) / LN2F), st[0], st[1]), f1)) return 1;
    float vsc = sxo_viterbi(m, msc, tsc, xs, dsq, l);
    out[2] = vsc;
    if (SXO_REJECTS(sxo_gumbel_surv((double)((vsc - filtersc) / LN2F), st[2], st[3]), f2)) return 2;
    float fwd = (float)sxo_forward(m, mo, tp, xs, dsq, l, NULL);
    out[3] = fwd;
    if (SXO_REJECTS(exp(sxo_exp_logsurv((double)((fwd - filtersc) / LN2F), st[4], st[5])), f3)) return 3;
    return 4;
}

/* ---- many-pair drivers (CPU baseline) ---- */

typedef struct {
    const sxo_batch* b;
    int64_t begin, end;
    atomic_llong next;
    int stage;
    void* out;
    const double* stats;
    const float* compo;
    double f[3];
    float* sc;
} pool_job;

#define CHUNK 16

static void run_pair(const pool_job* j, const sxo_batch* b, int stage, int64_t idx, int64_t begin, void* out) {
    int p = (int)(idx / b->nt), t = (int)(idx % b->nt);
    int m = b->m[p];
    const uint8_t* dsq = b->codes + b->seq_off[t];
    int64_t l = b->seq_off[t + 1] - b->seq_off[t];
    float xs[16];
    sxo_xsc(l, 1, xs);
    switch (stage) {
        case 0:
            ((float*)out)[idx - begin] = sxo_msv(m, b->msc + b->off[p], b->msv_tbm[p], dsq, l);
            break;
        case 1:
            ((float*)out)[idx - begin] = sxo_viterbi(m, b->msc + b->off[p], b->tsc + b->toff[p], xs, dsq, l);
            break;
        case 2:
            ((double*)out)[idx - begin] = sxo_forward(m, b->mo + b->off[p], b->tp + b->toff[p], xs, dsq, l, NULL);
            break;
        case 3:
            ((double*)out)[idx - begin] = sxo_forward64(m, b->mo + b->off[p], b->tp + b->toff[p], xs, dsq, l);
            break;
        case 5:
            ((int32_t*)out)[idx - begin] = sxo_msv16(m, b->msc + b->off[p], b->msv_tbm[p], dsq, l);
            break;
        case 6:
            ((int32_t*)out)[idx - begin] = sxo_vit16(m, b->msc + b->off[p], b->tsc + b->toff[p], xs, dsq, l);
            break;
        case 7:
            ((double*)out)[idx - begin] = sxo_backward64(m, b->mo + b->off[p], b->tp + b->toff[p], xs, dsq, l);
            break;
        default:
            ((int32_t*)out)[idx - begin] =
                sxo_cascade(m, b->msc + b->off[p], b->tsc + b->toff[p], b->mo + b->off[p], b->tp + b->toff[p],
                            b->msv_tbm[p], j->compo + (size_t)p * SXO_K, j->stats + (size_t)p * 6, j->f[0], j->f[1],
                            j->f[2], dsq, l, j->sc + 4 * (idx - begin));
            break;
    }
}

static void* pool_worker(void* arg) {
    pool_job* j = (pool_job*)arg;
    for (;;) {
        int64_t s = atomic_fetch_add(&j->next, CHUNK);
        if (s >= j->end) break;
        int64_t e = s + CHUNK < j->end ? s + CHUNK : j->end;
        for (int64_t i = s; i < e; i++) run_pair(j, j->b, j->stage, i, j->begin, j->out);
    }
    return NULL;
}

static int run_pool_ex(const sxo_batch* b, int64_t begin, int64_t end, void* out, int nthreads, int stage,
                       const double* stats, const float* compo, const double* f, float* sc) {
    if (end <= begin) return 0;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pool_job j;
    j.b = b; j.begin = begin; j.end = end; j.stage = stage; j.out = out;
    j.stats = stats; j.compo = compo; j.sc = sc;
    j.f[0] = f ? f[0] : 0; j.f[1] = f ? f[1] : 0; j.f[2] = f ? f[2] : 0;
    atomic_init(&j.next, begin);
    pthread_t th[256];
    for (int i = 1; i < nthreads; i++) pthread_create(&th[i], NULL, pool_worker, &j);
    pool_worker(&j);
    for (int i = 1; i < nthreads; i++) pthread_join(th[i], NULL);
    return 0;
}

static int run_pool(const sxo_batch* b, int64_t begin, int64_t end, void* out, int nthreads, int stage) {
    return run_pool_ex(b, begin, end, out, nthreads, stage, NULL, NULL, NULL, NULL);
}

int sxo_batch_cascade(const sxo_batch* b, const double* stats, const float* compo, double f1, double f2, double f3,
                      int64_t pb, int64_t pe, int32_t* stage_out, float* sc_out, int nt) {
    const double f[3] = {f1, f2, f3};
    return run_pool_ex(b, pb, pe, stage_out, nt, 4, stats, compo, f, sc_out);
}

int sxo_batch_msv(const sxo_batch* b, int64_t pb, int64_t pe, float* out, int nt) { return run_pool(b, pb, pe, out, nt, 0); }
int sxo_batch_viterbi(const sxo_batch* b, int64_t pb, int64_t pe, float* out, int nt) { return run_pool(b, pb, pe, out, nt, 1); }
int sxo_batch_forward(const sxo_batch* b, int64_t pb, int64_t pe, double* out, int nt) { return run_pool(b, pb, pe, out, nt, 2); }
int sxo_batch_forward64(const sxo_batch* b, int64_t pb, int64_t pe, double* out, int nt) { return run_pool(b, pb, pe, out, nt, 3); }
int sxo_batch_msv16(const sxo_batch* b, int64_t pb, int64_t pe, int32_t* out, int nt) { return run_pool(b, pb, pe, out, nt, 5); }
int sxo_batch_vit16(const sxo_batch* b, int64_t pb, int64_t pe, int32_t* out, int nt) { return run_pool(b, pb, pe, out, nt, 6); }
int sxo_batch_backward64(const sxo_batch* b, int64_t pb, int64_t pe, double* out, int nt) { return run_pool(b, pb, pe, out, nt, 7); }

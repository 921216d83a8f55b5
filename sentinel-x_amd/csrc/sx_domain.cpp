// sx_domain.cpp -- host-side domain definition, envelope rescoring and scoring of one
// (profile, target) pair (sentinel-x src/hmmer/pipeline.rs:186-617, dp.rs:395-1564).
#include "sx_domain.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace sx {

double exp_logsurv(double x, double tau, double lambda);  // pipeline.rs:98-104 (sx_pipeline.cpp)

namespace {

enum { TMM = 0, TMI, TMD, TIM, TII, TDM, TDD, TBM };

constexpr float RESCALE = 1.0e4f;

inline float lnf(float x) { return x > 0.0f ? std::log(x) : NEG_INF; }

// Degenerate codes and '*' of a null2 vector from its 20 residue entries (averages of
// the member entries, summed in member order; '*' = 1).
template <class T>
void fill_degenerate(T* v) {
    for (int code = 20; code <= 23; code++) {
        const int* mem;
        const int nm = degenerate_members(code, &mem);
        T s = 0;
        for (int t = 0; t < nm; t++) s += v[mem[t]];
        v[code] = s / (T)nm;
    }
    v[CODE_STAR] = 1;
}

}  // namespace

void SpecRows::reset(int l) {
    for (auto* v : {&n, &b, &e, &j, &c}) v->assign((size_t)l + 1, NEG_INF);
}

void DpMatrix::reset(int m_, int l_) {
    m = m_;
    l = l_;
    const size_t n = (size_t)(l + 1) * (m + 1);
    for (auto* v : {&mm, &im, &dm}) v->assign(n, 0.0f);
    lnscale.assign((size_t)l + 2, 0.0);
    sp.reset(l);
    for (auto* v : {&xn, &xb, &xe, &xj, &xc}) v->assign((size_t)l + 1, 0.0f);
    scale.assign((size_t)l + 1, 1.0f);
}

// ---------------------------------------------------------------------------
// full-matrix Forward / Backward (dp.rs:443-601)
// ---------------------------------------------------------------------------

float forward_matrix(const Profile& p, const uint8_t* dsq, int l, const Xsc& xs, DpMatrix& fx) {
    const int m = p.m;
    fx.reset(m, l);
    if (l == 0 || m == 0) return NEG_INF;
    const int w = m + 1;
    const float *pmm = p.p_row(TMM), *pmi = p.p_row(TMI), *pmd = p.p_row(TMD), *pim = p.p_row(TIM),
                *pii = p.p_row(TII), *pdm = p.p_row(TDM), *pdd = p.p_row(TDD), *pbm = p.p_row(TBM);
    float xn = 1.0f, xb = xs.pn_move, xj = 0.0f, xc = 0.0f;
    double lnscale = 0.0;
    fx.sp.n[0] = 0.0f;
    fx.sp.b[0] = lnf(xb);
    fx.xn[0] = xn;
    fx.xb[0] = xb;
    for (int i = 1; i <= l; i++) {
        const uint8_t x = dsq[i - 1];
        const float* eo = p.mo_row(x);
        const float ins = x != CODE_STAR ? 1.0f : 0.0f;
        float* cm = &fx.mm[(size_t)i * w];
        float* ci = &fx.im[(size_t)i * w];
        float* cd = &fx.dm[(size_t)i * w];
        const float* pm = &fx.mm[(size_t)(i - 1) * w];
        const float* pi = &fx.im[(size_t)(i - 1) * w];
        const float* pd = &fx.dm[(size_t)(i - 1) * w];
        float xe = 0.0f;
        for (int k = 1; k <= m; k++) {
            // summation order B, M, I, D
            const float mv = (xb * pbm[k] + pm[k - 1] * pmm[k - 1] + pi[k - 1] * pim[k - 1] + pd[k - 1] * pdm[k - 1]) *
                             eo[k];
            cm[k] = mv;
            ci[k] = k < m ? (pm[k] * pmi[k] + pi[k] * pii[k]) * ins : 0.0f;
            const float dv = cm[k - 1] * pmd[k - 1] + cd[k - 1] * pdd[k - 1];
            cd[k] = dv;
            xe += mv + dv;
        }
        xj = xj * xs.pj_loop + xe * xs.pe_loop;
        xc = xc * xs.pc_loop + xe * xs.pe_move;
        xn *= xs.pn_loop;
        xb = xn * xs.pn_move + xj * xs.pj_move;
        if (xe > RESCALE) {
            const float inv = 1.0f / xe;
            for (int k = 1; k <= m; k++) {
                cm[k] *= inv;
                ci[k] *= inv;
                cd[k] *= inv;
            }
            xn *= inv;
            xb *= inv;
            xj *= inv;
            xc *= inv;
            lnscale += std::log((double)xe);
            fx.scale[i] = xe;
            xe = 1.0f;
        }
        fx.xn[i] = xn;
        fx.xb[i] = xb;
        fx.xe[i] = xe;
        fx.xj[i] = xj;
        fx.xc[i] = xc;
        fx.lnscale[i] = lnscale;
        const float ls = (float)lnscale;
        fx.sp.n[i] = lnf(xn) + ls;
        fx.sp.b[i] = lnf(xb) + ls;
        fx.sp.e[i] = lnf(xe) + ls;
        fx.sp.j[i] = lnf(xj) + ls;
        fx.sp.c[i] = lnf(xc) + ls;
    }
    return (float)((double)lnf(xc * xs.pc_move) + lnscale);
}

float backward_matrix(const Profile& p, const uint8_t* dsq, int l, const Xsc& xs, DpMatrix& bx) {
    const int m = p.m;
    bx.reset(m, l);
    if (l == 0 || m == 0) return NEG_INF;
    const int w = m + 1;
    const float *pmm = p.p_row(TMM), *pmi = p.p_row(TMI), *pmd = p.p_row(TMD), *pim = p.p_row(TIM),
                *pii = p.p_row(TII), *pdm = p.p_row(TDM), *pdd = p.p_row(TDD), *pbm = p.p_row(TBM);
    float xc = xs.pc_move;
    float xe = xc * xs.pe_move;
    float xj = 0.0f, xn = 0.0f;
    {
        float* lm = &bx.mm[(size_t)l * w];
        float* ld = &bx.dm[(size_t)l * w];
        lm[m] = xe;
        ld[m] = xe;
        for (int k = m - 1; k >= 1; k--) {
            lm[k] = xe + ld[k + 1] * pmd[k];
            ld[k] = xe + ld[k + 1] * pdd[k];
        }
    }
    double lnscale = 0.0;
    bx.lnscale[l] = 0.0;
    bx.sp.c[l] = lnf(xc);
    bx.sp.e[l] = lnf(xe);
    for (int i = l - 1; i >= 1; i--) {
        const uint8_t xnext = dsq[i];
        const float* eo = p.mo_row(xnext);
        const float ins = xnext != CODE_STAR ? 1.0f : 0.0f;
        const float* nm = &bx.mm[(size_t)(i + 1) * w];
        const float* ni = &bx.im[(size_t)(i + 1) * w];
        float* cm = &bx.mm[(size_t)i * w];
        float* ci = &bx.im[(size_t)i * w];
        float* cd = &bx.dm[(size_t)i * w];
        float xb = 0.0f;
        for (int k = 1; k <= m; k++) xb += nm[k] * pbm[k] * eo[k];
        xj = xj * xs.pj_loop + xb * xs.pj_move;
        xc *= xs.pc_loop;
        xe = xj * xs.pe_loop + xc * xs.pe_move;
        xn = xn * xs.pn_loop + xb * xs.pn_move;
        cm[m] = xe;
        cd[m] = xe;
        ci[m] = 0.0f;
        float rowmax = std::max(std::max(std::max(std::max(xb, xe), xn), xj), xc);
        for (int k = m - 1; k >= 1; k--) {
            const float mnext = nm[k + 1] * eo[k + 1];
            const float iv = ni[k] * ins;
            const float mv = mnext * pmm[k] + iv * pmi[k] + xe + cd[k + 1] * pmd[k];
            cm[k] = mv;
            ci[k] = mnext * pim[k] + iv * pii[k];
            cd[k] = mnext * pdm[k] + cd[k + 1] * pdd[k] + xe;
            if (mv > rowmax) rowmax = mv;
        }
        if (rowmax > RESCALE) {
            const float inv = 1.0f / rowmax;
            for (int k = 1; k <= m; k++) {
                cm[k] *= inv;
                ci[k] *= inv;
                cd[k] *= inv;
            }
            xb *= inv;
            xe *= inv;
            xn *= inv;
            xj *= inv;
            xc *= inv;
            lnscale += std::log((double)rowmax);
        }
        bx.lnscale[i] = lnscale;
        const float ls = (float)lnscale;
        bx.sp.n[i] = lnf(xn) + ls;
        bx.sp.b[i] = lnf(xb) + ls;
        bx.sp.e[i] = lnf(xe) + ls;
        bx.sp.j[i] = lnf(xj) + ls;
        bx.sp.c[i] = lnf(xc) + ls;
    }
    const float* eo = p.mo_row(dsq[0]);
    const float* r1 = &bx.mm[(size_t)w];
    float b0 = 0.0f;
    for (int k = 1; k <= m; k++) b0 += r1[k] * pbm[k] * eo[k];
    bx.lnscale[0] = lnscale;
    bx.sp.b[0] = lnf(b0) + (float)lnscale;
    bx.sp.n[0] = logsum(bx.sp.n[1] + xs.n_loop, bx.sp.b[0] + xs.n_move);
    return bx.sp.n[0];
}

// ---------------------------------------------------------------------------
// posterior decoding and null2 (dp.rs:1274-1350)
// ---------------------------------------------------------------------------

void posterior_decode(const Profile& p, const Xsc& xs, const DpMatrix& fx, const DpMatrix& bx, Posterior& pp) {
    const int m = p.m, l = fx.l, w = m + 1;
    pp.m = m;
    pp.l = l;
    pp.pm.assign((size_t)(l + 1) * w, 0.0f);
    pp.pi.assign((size_t)(l + 1) * w, 0.0f);
    for (auto* v : {&pp.pn, &pp.pj, &pp.pc}) v->assign((size_t)l + 1, 0.0f);
    const double overall = (double)bx.sp.n[0];
    for (int i = 1; i <= l; i++) {
        const size_t row = (size_t)i * w;
        const double fac = std::exp(fx.lnscale[i] + bx.lnscale[i] - overall);
        for (int k = 1; k <= m; k++) {
            pp.pm[row + k] = (float)((double)fx.mm[row + k] * (double)bx.mm[row + k] * fac);
            if (k < m) pp.pi[row + k] = (float)((double)fx.im[row + k] * (double)bx.im[row + k] * fac);
        }
        pp.pn[i] = (float)std::exp((double)(fx.sp.n[i - 1] + bx.sp.n[i] + xs.n_loop) - overall);
        pp.pj[i] = (float)std::exp((double)(fx.sp.j[i - 1] + bx.sp.j[i] + xs.j_loop) - overall);
        pp.pc[i] = (float)std::exp((double)(fx.sp.c[i - 1] + bx.sp.c[i] + xs.c_loop) - overall);
    }
}

void null2_raw(const Profile& p, const Posterior& pp, float raw[K]) {
    const int m = p.m, l = pp.l, w = m + 1;
    std::vector<float> sm((size_t)w, 0.0f);
    float si = 0.0f, sn = 0.0f, sj = 0.0f, sc = 0.0f;
    for (int i = 1; i <= l; i++) {
        const size_t row = (size_t)i * w;
        for (int k = 1; k <= m; k++) {
            sm[k] += pp.pm[row + k];
            si += pp.pi[row + k];
        }
        sn += pp.pn[i];
        sj += pp.pj[i];
        sc += pp.pc[i];
    }
    const float norm = 1.0f / (float)std::max(l, 1);
    const float xfactor = sn * norm + sc * norm + sj * norm;
    const float isum = si * norm;
    for (int x = 0; x < K; x++) {
        const float* row = p.mo_row(x);
        float v = 0.0f;
        for (int k = 1; k <= m; k++) v += sm[k] * norm * row[k];
        raw[x] = v + isum + xfactor;
    }
}

void null2_log(const float raw[K], float out[NCODE]) {
    float null2[NCODE];
    for (int x = 0; x < K; x++) null2[x] = raw[x];
    fill_degenerate(null2);
    for (int x = 0; x < NCODE; x++) out[x] = null2[x] > 0.0f ? std::log(null2[x]) : NEG_INF;
}

void null2_from_posterior(const Profile& p, const Posterior& pp, float out[NCODE]) {
    float raw[K];
    null2_raw(p, pp, raw);
    null2_log(raw, out);
}

// ---------------------------------------------------------------------------
// optimal-accuracy alignment of one envelope (dp.rs:1355-1564)
// ---------------------------------------------------------------------------

namespace {
// the order in which the striped implementation visits model positions
inline int stripe_count(int m) { return std::max(2, (m - 1) / 4 + 1); }
}  // namespace

bool optimal_accuracy_coords(const Profile& p, const Xsc& xs, const Posterior& pp, AliCoords& out) {
    const int m = p.m, l = pp.l;
    if (l == 0 || m == 0) return false;
    const int w = m + 1;
    const float *pmm = p.p_row(TMM), *pmi = p.p_row(TMI), *pmd = p.p_row(TMD), *pim = p.p_row(TIM),
                *pii = p.p_row(TII), *pdm = p.p_row(TDM), *pdd = p.p_row(TDD), *pbm = p.p_row(TBM);
    const size_t n = (size_t)(l + 1) * w;
    std::vector<float> om(n, NEG_INF), oi(n, NEG_INF), od(n, NEG_INF);
    std::vector<float> xe((size_t)l + 1, NEG_INF), xn(xe), xj(xe), xb(xe), xc(xe);
    xn[0] = 0.0f;
    xb[0] = 0.0f;
    // an edge with zero probability contributes 0 (not -inf) to the max
    auto gate = [](float t, float v) { return t > 0.0f ? v : 0.0f; };
    for (int i = 1; i <= l; i++) {
        const size_t pr = (size_t)(i - 1) * w, cr = (size_t)i * w;
        float emax = NEG_INF;
        for (int k = 1; k <= m; k++) {
            float sv = gate(pbm[k], xb[i - 1]);
            if (k >= 2) {
                sv = std::max(sv, gate(pmm[k - 1], om[pr + k - 1]));
                sv = std::max(sv, gate(pim[k - 1], oi[pr + k - 1]));
                sv = std::max(sv, gate(pdm[k - 1], od[pr + k - 1]));
            } else {
                sv = std::max(sv, 0.0f);
            }
            const float mv = sv + pp.pm[cr + k];
            om[cr + k] = mv;
            emax = std::max(emax, mv);
            oi[cr + k] = std::max(gate(pmi[k], om[pr + k]), gate(pii[k], oi[pr + k])) + pp.pi[cr + k];
        }
        od[cr + 1] = NEG_INF;
        for (int k = 2; k <= m; k++) {
            od[cr + k] = std::max(gate(pmd[k - 1], om[cr + k - 1]), gate(pdd[k - 1], od[cr + k - 1]));
            emax = std::max(emax, od[cr + k]);
        }
        xe[i] = emax;
        xj[i] = std::max(xs.pj_loop == 0.0f ? 0.0f : xj[i - 1] + pp.pj[i], xs.pe_loop == 0.0f ? 0.0f : xe[i]);
        xc[i] = std::max(xs.pc_loop == 0.0f ? 0.0f : xc[i - 1] + pp.pc[i], xs.pe_move == 0.0f ? 0.0f : xe[i]);
        xn[i] = xs.pn_loop == 0.0f ? 0.0f : xn[i - 1] + pp.pn[i];
        xb[i] = std::max(xs.pn_move == 0.0f ? 0.0f : xn[i], xs.pj_move == 0.0f ? 0.0f : xj[i]);
    }
    // traceback from C; the first and last M states give the coordinates
    enum St { C, E, M, I, D, B, J, N };
    const int qs = stripe_count(m);
    auto dis = [](float t, float v) { return t == 0.0f ? NEG_INF : v; };
    St s0 = C;
    int i = l, k = 0;
    bool have = false;
    int fi = 0, fk = 0, li = 0, lk = 0;
    const long limit = 8L * (l + 2) * (m + 2) + 64;
    for (long guard = 1; guard <= limit; guard++) {
        St s1;
        if (s0 == C) {
            if (i == 0) break;
            const float a = xs.pc_loop == 0.0f ? NEG_INF : xc[i - 1] + pp.pc[i];
            const float b = xs.pe_move == 0.0f ? NEG_INF : xe[i];
            s1 = a > b ? C : E;
        } else if (s0 == E) {
            const size_t row = (size_t)i * w;
            float mx = NEG_INF;
            int pk = -1;
            St pst = M;
            for (int q = 0; q < qs; q++) {
                for (int r = 0; r < 4; r++) {
                    const int kk = r * qs + q + 1;
                    if (kk <= m && om[row + kk] >= mx) {
                        mx = om[row + kk];
                        pk = kk;
                        pst = M;
                    }
                }
                for (int r = 0; r < 4; r++) {
                    const int kk = r * qs + q + 1;
                    if (kk <= m && od[row + kk] > mx) {
                        mx = od[row + kk];
                        pk = kk;
                        pst = D;
                    }
                }
            }
            if (pk < 0) break;
            k = pk;
            s1 = pst;
        } else if (s0 == M) {
            if (!have) {
                li = i;
                lk = k;
                have = true;
            }
            fi = i;
            fk = k;
            const size_t prow = (size_t)(i - 1) * w;
            float mv = NEG_INF, iv = NEG_INF, dv = NEG_INF;
            if (k >= 2) {
                mv = dis(pmm[k - 1], om[prow + k - 1]);
                iv = dis(pim[k - 1], oi[prow + k - 1]);
                dv = dis(pdm[k - 1], od[prow + k - 1]);
            }
            const float path[4] = {mv, iv, dv, dis(pbm[k], xb[i - 1])};
            int best = 0;
            for (int t = 1; t < 4; t++)
                if (path[t] > path[best]) best = t;
            k--;
            i--;
            s1 = best == 0 ? M : best == 1 ? I : best == 2 ? D : B;
        } else if (s0 == D) {
            const size_t row = (size_t)i * w;
            float a = NEG_INF, b = NEG_INF;
            if (k >= 2) {
                a = dis(pmd[k - 1], om[row + k - 1]);
                b = dis(pdd[k - 1], od[row + k - 1]);
            }
            k--;
            s1 = a >= b ? M : D;
        } else if (s0 == I) {
            const size_t prow = (size_t)(i - 1) * w;
            const float a = dis(pmi[k], om[prow + k]);
            const float b = dis(pii[k], oi[prow + k]);
            i--;
            s1 = a >= b ? M : I;
        } else if (s0 == B) {
            const float a = xs.pn_move == 0.0f ? NEG_INF : xn[i];
            const float b = xs.pj_move == 0.0f ? NEG_INF : xj[i];
            if (a > b)
                s1 = N;
            else
                break;  // a domain begins here; a unihit trace has one
        } else {
            break;  // J or N
        }
        if ((s1 == N || s1 == J || s1 == C) && s1 == s0) i--;
        if (s1 == N) break;
        s0 = s1;
    }
    if (!have) return false;
    out = {fi, li, fk, lk};
    return true;
}

// ---------------------------------------------------------------------------
// Viterbi with traceback (dp.rs:691-904)
// ---------------------------------------------------------------------------

std::vector<AliCoords> viterbi_trace(const Profile& p, const uint8_t* dsq, int l, const Xsc& xs) {
    std::vector<AliCoords> out;
    const int m = p.m;
    if (l == 0 || m == 0) return out;
    const int w = m + 1;
    const float *tmm = p.t_row(TMM), *tmi = p.t_row(TMI), *tmd = p.t_row(TMD), *tim = p.t_row(TIM),
                *tii = p.t_row(TII), *tdm = p.t_row(TDM), *tdd = p.t_row(TDD), *tbm = p.t_row(TBM);
    const size_t n = (size_t)(l + 1) * w;
    std::vector<float> vm(n, NEG_INF), vi(n, NEG_INF), vd(n, NEG_INF);
    enum : uint8_t { FROM_M = 0, FROM_I = 1, FROM_D = 2, FROM_B = 3 };
    std::vector<uint8_t> pm(n, 0), pi(n, 0), pd(n, 0);
    std::vector<float> xn((size_t)l + 1, NEG_INF), xb(xn), xe(xn), xj(xn), xc(xn);
    std::vector<int> e_k((size_t)l + 1, 0);
    std::vector<uint8_t> e_del((size_t)l + 1, 0), b_from_j((size_t)l + 1, 0), j_from_e((size_t)l + 1, 0),
        c_from_e((size_t)l + 1, 0);
    xn[0] = 0.0f;
    xb[0] = xs.n_move;
    for (int i = 1; i <= l; i++) {
        const uint8_t x = dsq[i - 1];
        const float* rs = p.msc_row(x);
        const bool ins_ok = x != CODE_STAR;
        float best_e = NEG_INF;
        int bk = 0;
        bool bdel = false;
        for (int k = 1; k <= m; k++) {
            const size_t pc = (size_t)(i - 1) * w + k, cc = (size_t)i * w + k;
            float sc = vm[pc - 1] + tmm[k - 1];
            uint8_t ptr = FROM_M;
            const float c1 = vi[pc - 1] + tim[k - 1], c2 = vd[pc - 1] + tdm[k - 1], c3 = xb[i - 1] + tbm[k];
            if (c1 > sc) {
                sc = c1;
                ptr = FROM_I;
            }
            if (c2 > sc) {
                sc = c2;
                ptr = FROM_D;
            }
            if (c3 > sc) {
                sc = c3;
                ptr = FROM_B;
            }
            vm[cc] = sc + rs[k];
            pm[cc] = ptr;
            if (k < m && ins_ok) {
                const float a = vm[pc] + tmi[k], b = vi[pc] + tii[k];
                if (a >= b) {
                    vi[cc] = a;
                    pi[cc] = FROM_M;
                } else {
                    vi[cc] = b;
                    pi[cc] = FROM_I;
                }
            }
            const float a = vm[cc - 1] + tmd[k - 1], b = vd[cc - 1] + tdd[k - 1];
            if (a >= b) {
                vd[cc] = a;
                pd[cc] = 0;
            } else {
                vd[cc] = b;
                pd[cc] = 1;
            }
            if (vm[cc] > best_e) {
                best_e = vm[cc];
                bk = k;
                bdel = false;
            }
            if (vd[cc] > best_e) {
                best_e = vd[cc];
                bk = k;
                bdel = true;
            }
        }
        xe[i] = best_e;
        e_k[i] = bk;
        e_del[i] = bdel;
        const float jl = xj[i - 1] + xs.j_loop, je = best_e + xs.e_loop;
        if (je > jl) {
            xj[i] = je;
            j_from_e[i] = 1;
        } else {
            xj[i] = jl;
        }
        const float cl = xc[i - 1] + xs.c_loop, ce = best_e + xs.e_move;
        if (ce > cl) {
            xc[i] = ce;
            c_from_e[i] = 1;
        } else {
            xc[i] = cl;
        }
        xn[i] = xn[i - 1] + xs.n_loop;
        const float bn = xn[i] + xs.n_move, bj = xj[i] + xs.j_move;
        if (bj > bn) {
            xb[i] = bj;
            b_from_j[i] = 1;
        } else {
            xb[i] = bn;
        }
    }
    if (!(xc[l] > NEG_INF)) return out;

    enum St { C, E, M, I, D, B, J, N };
    St st = C;
    int i = l, k = 0;
    bool open = false;
    int cj = 0, cm = 0;  // end of the domain being traced (jali, mali)
    const long limit = 4L * (l + 2) * (m + 2);
    for (long guard = 1; guard <= limit; guard++) {
        if (st == C) {
            if (i == 0) break;
            if (c_from_e[i])
                st = E;
            else
                i--;
        } else if (st == E) {
            if (e_k[i] == 0) break;
            k = e_k[i];
            st = e_del[i] ? D : M;
            open = true;
            cj = i;
            cm = k;
        } else if (st == M) {
            const uint8_t ptr = pm[(size_t)i * w + k];
            if (ptr == FROM_M) {
                st = M;
            } else if (ptr == FROM_I) {
                st = I;
            } else if (ptr == FROM_D) {
                st = D;
            } else {
                if (open) out.push_back({i, cj, k, cm});
                open = false;
                st = B;
                i--;
                continue;
            }
            k--;
            i--;
        } else if (st == I) {
            st = pi[(size_t)i * w + k] == FROM_M ? M : I;
            i--;
        } else if (st == D) {
            st = pd[(size_t)i * w + k] == 0 ? M : D;
            k--;
        } else if (st == B) {
            st = b_from_j[i] ? J : N;
        } else if (st == J) {
            if (j_from_e[i])
                st = E;
            else
                i--;
        } else {
            break;
        }
    }
    std::reverse(out.begin(), out.end());
    return out;
}

// ---------------------------------------------------------------------------
// Easel RNG and stochastic traceback (dp.rs:958-1258)
// ---------------------------------------------------------------------------

namespace {
uint32_t esl_mix3(uint32_t a, uint32_t b, uint32_t c) {
    a -= b; a -= c; a ^= c >> 13;
    b -= c; b -= a; b ^= a << 8;
    c -= a; c -= b; c ^= b >> 13;
    a -= b; a -= c; a ^= c >> 12;
    b -= c; b -= a; b ^= a << 16;
    c -= a; c -= b; c ^= b >> 5;
    a -= b; a -= c; a ^= c >> 3;
    b -= c; b -= a; b ^= a << 10;
    c -= a; c -= b; c ^= b >> 15;
    return c;
}

// esl_vec_FNorm: Kahan-summed normalisation of a small probability vector
template <size_t N>
void fnorm(float (&v)[N]) {
    float sum = 0.0f, comp = 0.0f;
    for (size_t t = 0; t < N; t++) {
        const float y = v[t] - comp;
        const float s = sum + y;
        comp = (s - sum) - y;
        sum = s;
    }
    if (sum != 0.0f) {
        for (size_t t = 0; t < N; t++) v[t] /= sum;
    } else {
        for (size_t t = 0; t < N; t++) v[t] = 1.0f / (float)N;
    }
}

// esl_rnd_FChoose
template <size_t N>
int fchoose(EaselRng& rng, const float (&v)[N]) {
    double norm = 0.0;
    for (size_t t = 0; t < N; t++) norm += (double)v[t];
    const double roll = rng.next();
    double sum = 0.0;
    for (size_t t = 0; t < N; t++) {
        sum += (double)v[t];
        if (roll < sum / norm) return (int)t;
    }
    return (int)N - 1;
}
}  // namespace

EaselRng::EaselRng(uint32_t seed) : seed_(seed), x_(0) { reseed(); }

void EaselRng::reseed() {
    x_ = esl_mix3(seed_, 87654321u, 12345678u);
    if (x_ == 0) x_ = 42;
}

double EaselRng::next() {
    x_ = x_ * 69069u + 1u;
    return (double)x_ / 4294967296.0;
}

std::vector<SampledDomain> stochastic_trace(const Profile& p, const DpMatrix& fx, const Xsc& xs, EaselRng& rng) {
    const int m = p.m, w = m + 1;
    const float *pmm = p.p_row(TMM), *pmi = p.p_row(TMI), *pmd = p.p_row(TMD), *pim = p.p_row(TIM),
                *pii = p.p_row(TII), *pdm = p.p_row(TDM), *pdd = p.p_row(TDD), *pbm = p.p_row(TBM);
    const int qs = stripe_count(std::max(m, 1));
    std::vector<SampledDomain> doms;
    SampledDomain cur;
    bool open = false;
    enum St { C, E, M, I, D, B, J, N };
    St s0 = C;
    int i = fx.l, k = 0;
    const long limit = 8L * (fx.l + 2) * (m + 2) + 64;
    for (long guard = 1; guard <= limit; guard++) {
        St s1;
        if (s0 == C) {
            if (i == 0) break;
            float path[2] = {fx.xc[i - 1] * xs.pc_loop, fx.xe[i] * xs.pe_move * fx.scale[i]};
            fnorm(path);
            s1 = fchoose(rng, path) == 0 ? C : E;
        } else if (s0 == E) {
            const double roll = rng.next();
            const float norm = (float)(1.0 / (double)fx.xe[i]);
            const size_t row = (size_t)i * w;
            double sum = 0.0;
            int pk = -1;
            St pst = M;
            for (int passes = 0; pk < 0 && passes < 1000; passes++) {
                for (int q = 0; q < qs && pk < 0; q++) {
                    for (int r = 0; r < 4; r++) {
                        const int kk = r * qs + q + 1;
                        if (kk <= m) {
                            sum += (double)(fx.mm[row + kk] * norm);
                            if (roll < sum) {
                                pk = kk;
                                pst = M;
                                break;
                            }
                        }
                    }
                    if (pk >= 0) break;
                    for (int r = 0; r < 4; r++) {
                        const int kk = r * qs + q + 1;
                        if (kk <= m) {
                            sum += (double)(fx.dm[row + kk] * norm);
                            if (roll < sum) {
                                pk = kk;
                                pst = D;
                                break;
                            }
                        }
                    }
                }
                if (pk < 0 && sum <= 0.0) break;
            }
            if (pk < 0) break;
            k = pk;
            cur = SampledDomain();
            open = true;
            s1 = pst;
        } else if (s0 == M) {
            if (open) {
                if (cur.sqto == 0) {
                    cur.sqto = i;
                    cur.hmmto = k;
                }
                cur.sqfrom = i;
                cur.hmmfrom = k;
                cur.nodes.push_back(k);
            }
            const size_t prow = (size_t)(i - 1) * w;
            float mpv = 0.0f, ipv = 0.0f, dpv = 0.0f, tm = 0.0f, ti = 0.0f, td = 0.0f;
            if (k >= 2) {
                mpv = fx.mm[prow + k - 1];
                ipv = fx.im[prow + k - 1];
                dpv = fx.dm[prow + k - 1];
                tm = pmm[k - 1];
                ti = pim[k - 1];
                td = pdm[k - 1];
            }
            float path[4] = {fx.xb[i - 1] * pbm[k], mpv * tm, ipv * ti, dpv * td};
            fnorm(path);
            const int c = fchoose(rng, path);
            k--;
            i--;
            s1 = c == 0 ? B : c == 1 ? M : c == 2 ? I : D;
        } else if (s0 == D) {
            const size_t row = (size_t)i * w;
            float mpv = 0.0f, dpv = 0.0f, tmd_ = 0.0f, tdd_ = 0.0f;
            if (k >= 2) {
                mpv = fx.mm[row + k - 1];
                dpv = fx.dm[row + k - 1];
                tmd_ = pmd[k - 1];
                tdd_ = pdd[k - 1];
            }
            float path[2] = {mpv * tmd_, dpv * tdd_};
            fnorm(path);
            const int c = fchoose(rng, path);
            k--;
            s1 = c == 0 ? M : D;
        } else if (s0 == I) {
            if (open) cur.nodes.push_back(k);
            const size_t prow = (size_t)(i - 1) * w;
            float path[2] = {fx.mm[prow + k] * pmi[k], fx.im[prow + k] * pii[k]};
            fnorm(path);
            const int c = fchoose(rng, path);
            i--;
            s1 = c == 0 ? M : I;
        } else if (s0 == B) {
            if (open) {
                doms.push_back(std::move(cur));
                cur = SampledDomain();
                open = false;
            }
            float path[2] = {fx.xn[i] * xs.pn_move, fx.xj[i] * xs.pj_move};
            fnorm(path);
            s1 = fchoose(rng, path) == 0 ? N : J;
        } else if (s0 == J) {
            if (i == 0) break;
            float path[2] = {fx.xj[i - 1] * xs.pj_loop, fx.xe[i] * xs.pe_loop * fx.scale[i]};
            fnorm(path);
            s1 = fchoose(rng, path) == 0 ? J : E;
        } else {  // N
            if (i == 0) break;
            s1 = N;
        }
        if ((s1 == N || s1 == J || s1 == C) && s1 == s0) i--;
        s0 = s1;
    }
    std::reverse(doms.begin(), doms.end());
    return doms;
}

// ---------------------------------------------------------------------------
// region ensemble, clustering and envelope rescoring (pipeline.rs:298-591)
// ---------------------------------------------------------------------------

namespace {

struct SpCoord {
    int idx, i, j, k, m;
};

// link_spsamples with min overlap 0.8 of the shorter and max diagonal difference 4
bool sp_link(const SpCoord& a, const SpCoord& b) {
    long nov = (long)std::min(a.j, b.j) - (long)std::max(a.i, b.i) + 1;
    float n = (float)std::min(a.j - a.i + 1, b.j - b.i + 1);
    if ((float)nov / n < 0.8f) return false;
    nov = (long)std::min(a.m, b.m) - (long)std::max(a.k, b.k);
    n = (float)std::min(a.m - a.k + 1, b.m - b.k + 1);
    if ((float)nov / n < 0.8f) return false;
    if (std::labs(((long)a.i - a.k) - ((long)b.i - b.k)) <= 4) return true;
    return std::labs(((long)a.j - a.m) - ((long)b.j - b.m)) <= 4;
}

// esl_cluster_SingleLinkage traversal (cluster ids assigned in the same order)
int single_linkage(const std::vector<SpCoord>& v, std::vector<int>& assign) {
    const int n = (int)v.size();
    std::vector<int> todo((size_t)n), stack;
    for (int x = 0; x < n; x++) todo[x] = n - x - 1;
    assign.assign((size_t)n, 0);
    int nc = 0;
    while (!todo.empty()) {
        stack.push_back(todo.back());
        todo.pop_back();
        while (!stack.empty()) {
            const int x = stack.back();
            stack.pop_back();
            assign[x] = nc;
            for (size_t ii = todo.size(); ii-- > 0;) {
                if (sp_link(v[x], v[todo[ii]])) {
                    const int y = todo[ii];
                    todo[ii] = todo.back();
                    todo.pop_back();
                    stack.push_back(y);
                }
            }
        }
        nc++;
    }
    return nc;
}

// first (or, from the right, last) endpoint value reached by thr samples, else the mode
int endpoint(const std::vector<int>& vals, bool from_left, int thr) {
    const auto mm = std::minmax_element(vals.begin(), vals.end());
    const int lo = *mm.first, hi = *mm.second;
    std::vector<int> cnt((size_t)(hi - lo + 1), 0);
    for (int v : vals) cnt[v - lo]++;
    if (from_left) {
        for (size_t o = 0; o < cnt.size(); o++)
            if (cnt[o] >= thr) return lo + (int)o;
    } else {
        for (size_t o = cnt.size(); o-- > 0;)
            if (cnt[o] >= thr) return lo + (int)o;
    }
    size_t best = 0;
    for (size_t o = 1; o < cnt.size(); o++)
        if (cnt[o] > cnt[best]) best = o;
    return lo + (int)best;
}

bool is_multidomain(const std::vector<float>& btot, const std::vector<float>& etot, int i, int j, float rt3) {
    float mx = -1.0f;
    for (int jj = i; jj <= j; jj++) {
        const float v = std::min(etot[jj] - etot[i - 1], btot[j] - btot[jj - 1]);
        if (v > mx) mx = v;
    }
    return mx >= rt3;
}

}  // namespace

std::vector<std::pair<int, int>> region_ensemble(const Profile& p, const uint8_t* dsq, int ireg, int jreg, int l,
                                                 TailWork& ws, std::vector<float>& n2sc) {
    constexpr int NS = 200;
    const uint8_t* sub = dsq + (ireg - 1);
    const int lr = jreg - ireg + 1;
    const Xsc xs = Xsc::make((size_t)l, true);
    forward_matrix(p, sub, lr, xs, ws.fx);
    EaselRng rng(42);
    for (int v = ireg; v <= jreg; v++) n2sc[v] = 0.0f;
    std::vector<SpCoord> sp;
    const int w = p.m + 1;
    float null2[NCODE] = {};
    for (int t = 0; t < NS; t++) {
        const std::vector<SampledDomain> doms = stochastic_trace(p, ws.fx, xs, rng);
        int pos = 1;
        for (const SampledDomain& d : doms) {
            sp.push_back({t, d.sqfrom + ireg - 1, d.sqto + ireg - 1, d.hmmfrom, d.hmmto});
            // null2 by trace: mean match-emission odds over the domain's emitting nodes
            const float ld = (float)std::max<size_t>(d.nodes.size(), 1);
            for (int x = 0; x < K; x++) {
                const float* row = p.mo.data() + (size_t)x * w;
                float v = 0.0f;
                for (int kk : d.nodes) v += row[kk];
                null2[x] = v / ld;
            }
            fill_degenerate(null2);
            for (; pos <= d.sqfrom; pos++) n2sc[ireg + pos - 1] += 1.0f;
            for (; pos <= d.sqto; pos++) n2sc[ireg + pos - 1] += null2[dsq[ireg + pos - 2]];
        }
        for (; pos <= lr; pos++) n2sc[ireg + pos - 1] += 1.0f;
    }
    for (int v = ireg; v <= jreg; v++) n2sc[v] = std::log(n2sc[v] / (float)NS);

    std::vector<int> assign;
    const int nc = single_linkage(sp, assign);
    struct Sig {
        int i, j;
        float prob;
    };
    std::vector<Sig> sig;
    for (int c = 0; c < nc; c++) {
        int ninc = 0, last = -1;
        std::vector<int> vi, vj, vk, vm;
        for (size_t h = 0; h < sp.size(); h++) {
            if (assign[h] != c) continue;
            if (last != sp[h].idx) ninc++;
            last = sp[h].idx;
            vi.push_back(sp[h].i);
            vj.push_back(sp[h].j);
            vk.push_back(sp[h].k);
            vm.push_back(sp[h].m);
        }
        if ((float)ninc / (float)NS < 0.25f) continue;
        const int thr = (int)std::ceil((float)ninc * 0.02f);
        const int bi = endpoint(vi, true, thr), bk = endpoint(vk, true, thr);
        const int bj = endpoint(vj, false, thr), bm = endpoint(vm, false, thr);
        if (bi > bj || bk > bm) continue;
        sig.push_back({bi, bj, (float)ninc / (float)NS});
    }
    std::stable_sort(sig.begin(), sig.end(), [](const Sig& a, const Sig& b) { return a.i < b.i; });
    std::vector<char> dominated(sig.size(), 0);
    for (size_t d = 0; d < sig.size(); d++) {
        for (size_t d2 = d + 1; d2 < sig.size(); d2++) {
            const long nov = (long)std::min(sig[d].j, sig[d2].j) - (long)std::max(sig[d].i, sig[d2].i) + 1;
            if (nov == 0) break;
            const float n = (float)std::min(sig[d].j - sig[d].i + 1, sig[d2].j - sig[d2].i + 1);
            if ((float)nov / n >= 0.8f) {
                if (sig[d].prob > sig[d2].prob)
                    dominated[d2] = 1;
                else
                    dominated[d] = 1;
            }
        }
    }
    std::vector<std::pair<int, int>> out;
    for (size_t d = 0; d < sig.size(); d++)
        if (!dominated[d]) out.push_back({sig[d].i, sig[d].j});
    return out;
}

namespace {

// rescore_isolated_domain: envelope i..j (1-based, inclusive) in unihit mode
bool rescore(const Profile& p, const uint8_t* dsq, int i, int j, int l, bool null2, TailWork& ws,
             std::vector<float>& n2sc, bool null2_done, Envelope& env) {
    if (i == 0 || j < i || j > l) return false;
    const uint8_t* sub = dsq + (i - 1);
    const int ls = j - i + 1;
    const Xsc xs = Xsc::make((size_t)l, false);
    const float envsc = forward_matrix(p, sub, ls, xs, ws.fx);
    if (!std::isfinite(envsc)) return false;
    const float bsc = backward_matrix(p, sub, ls, xs, ws.bx);
    if (!std::isfinite(bsc)) return false;
    posterior_decode(p, xs, ws.fx, ws.bx, ws.pp);
    float domcorr = 0.0f;
    if (null2) {
        if (!null2_done) {
            float n2[NCODE];
            null2_from_posterior(p, ws.pp, n2);
            for (int pos = i; pos <= j; pos++) {
                const float v = n2[dsq[pos - 1]];
                n2sc[pos] = std::isfinite(v) ? v : 0.0f;
            }
        }
        for (int pos = i; pos <= j; pos++) domcorr += n2sc[pos];
    }
    AliCoords a;
    if (optimal_accuracy_coords(p, xs, ws.pp, a)) {
        env.ali = {i - 1 + a.iali, i - 1 + a.jali, a.kali, a.mali};
    } else {
        const std::vector<AliCoords> vt = viterbi_trace(p, sub, ls, xs);
        if (!vt.empty())
            env.ali = {i - 1 + vt[0].iali, i - 1 + vt[0].jali, vt[0].kali, vt[0].mali};
        else
            env.ali = {i, j, 1, p.m};
    }
    env.i = i;
    env.j = j;
    env.envsc = envsc;
    env.domcorr = domcorr;
    return true;
}

}  // namespace

DeviceEnvelope rescore_raw(const Profile& p, const uint8_t* dsq, int i, int j, int l, bool null2, TailWork& ws,
                           float* bcksc) {
    DeviceEnvelope r{};
    r.status = 1;
    r.envsc = NEG_INF;
    *bcksc = NEG_INF;
    const uint8_t* sub = dsq + (i - 1);
    const int ls = j - i + 1;
    const Xsc xs = Xsc::make((size_t)l, false);
    r.envsc = forward_matrix(p, sub, ls, xs, ws.fx);
    if (!std::isfinite(r.envsc)) return r;
    *bcksc = backward_matrix(p, sub, ls, xs, ws.bx);
    if (!std::isfinite(*bcksc)) {
        r.status = 2;
        return r;
    }
    r.status = 0;
    posterior_decode(p, xs, ws.fx, ws.bx, ws.pp);
    if (null2) null2_raw(p, ws.pp, r.null2raw);
    r.oa = optimal_accuracy_coords(p, xs, ws.pp, r.ali) ? 1 : 0;
    return r;
}

bool envelope_from_device(const Profile& p, const uint8_t* dsq, int i, int j, int l, bool null2, bool null2_done,
                          const DeviceEnvelope& r, std::vector<float>& n2sc, Envelope& env) {
    if (r.status != 0) return false;  // Forward or Backward not finite (pipeline.rs:561-567)
    float domcorr = 0.0f;
    if (null2) {
        if (!null2_done) {
            float n2[NCODE];
            null2_log(r.null2raw, n2);
            for (int pos = i; pos <= j; pos++) {
                const float v = n2[dsq[pos - 1]];
                n2sc[pos] = std::isfinite(v) ? v : 0.0f;
            }
        }
        for (int pos = i; pos <= j; pos++) domcorr += n2sc[pos];
    }
    if (r.oa) {
        env.ali = {i - 1 + r.ali.iali, i - 1 + r.ali.jali, r.ali.kali, r.ali.mali};
    } else {
        // no M state on the optimal-accuracy trace: the reference falls back to a Viterbi
        // trace of the envelope (pipeline.rs:580-588)
        const Xsc xs = Xsc::make((size_t)l, false);
        const std::vector<AliCoords> vt = viterbi_trace(p, dsq + (i - 1), j - i + 1, xs);
        if (!vt.empty())
            env.ali = {i - 1 + vt[0].iali, i - 1 + vt[0].jali, vt[0].kali, vt[0].mali};
        else
            env.ali = {i, j, 1, p.m};
    }
    env.i = i;
    env.j = j;
    env.envsc = r.envsc;
    env.domcorr = domcorr;
    return true;
}

// ---------------------------------------------------------------------------
// pipeline.rs:186-294
// ---------------------------------------------------------------------------

void find_regions(int l, float overall, const SpecRows& f, const SpecRows& b, TailRegions& out) {
    const Xsc xs = Xsc::make((size_t)l, true);
    // posterior decoding of begin, end and core occupancy per residue
    std::vector<float> btot((size_t)l + 1, 0.0f), etot(btot), mocc(btot);
    for (int i = 1; i <= l; i++) {
        btot[i] = btot[i - 1] + std::exp(f.b[i - 1] + b.b[i - 1] - overall);
        etot[i] = etot[i - 1] + std::exp(f.e[i] + b.e[i] - overall);
        const float njc = std::exp(f.n[i - 1] + b.n[i] + xs.n_loop - overall) +
                          std::exp(f.j[i - 1] + b.j[i] + xs.j_loop - overall) +
                          std::exp(f.c[i - 1] + b.c[i] + xs.c_loop - overall);
        mocc[i] = 1.0f - njc;
    }
    out.nexpected = btot[l];
    out.spans.clear();

    // regions by posterior heuristics
    const float rt1 = 0.25f, rt2 = 0.10f, rt3 = 0.20f;
    int i = -1;
    bool triggered = false;
    for (int j = 1; j <= l; j++) {
        if (!triggered) {
            if (mocc[j] - (btot[j] - btot[j - 1]) < rt2 || i == -1) i = j;
            if (mocc[j] >= rt1) triggered = true;
        } else if (mocc[j] - (etot[j] - etot[j - 1]) < rt2) {
            const int ri = std::max(i, 1);
            out.spans.push_back({ri, j, is_multidomain(btot, etot, ri, j, rt3)});
            i = -1;
            triggered = false;
        }
    }
}

bool score_tail(const Profile& p, const uint8_t* dsq, int l, bool null2, float fwdsc, float bcksc,
                const SpecRows& f, const SpecRows& b, TailWork& ws, RawHit& out) {
    if (l == 0 || p.m == 0 || !std::isfinite(fwdsc)) return false;
    TailRegions r;
    find_regions(l, std::isfinite(bcksc) ? bcksc : fwdsc, f, b, r);
    return score_regions(p, dsq, l, null2, fwdsc, r, ws, out);
}

bool score_regions(const Profile& p, const uint8_t* dsq, int l, bool null2, float fwdsc, const TailRegions& r,
                   TailWork& ws, RawHit& out) {
    if (l == 0 || p.m == 0 || !std::isfinite(fwdsc)) return false;
    const float nexpected = r.nexpected;
    std::vector<float> n2sc((size_t)l + 1, 0.0f);
    std::vector<Envelope> envs;
    const int nregions = (int)r.spans.size();
    for (const RegionSpan& g : r.spans) {
        Envelope env;
        if (g.i < 1 || g.j < g.i || g.j > l) throw std::runtime_error("region outside the target");
        if (g.multi) {
            for (const auto& ij : region_ensemble(p, dsq, g.i, g.j, l, ws, n2sc))
                if (rescore(p, dsq, ij.first, ij.second, l, null2, ws, n2sc, true, env)) envs.push_back(env);
        } else if (rescore(p, dsq, g.i, g.j, l, null2, ws, n2sc, false, env)) {
            envs.push_back(env);
        }
    }
    if (nregions == 0 || envs.empty()) return false;
    return score_envelopes(p, l, null2, fwdsc, nexpected, nregions, n2sc, envs, out);
}

bool score_envelopes(const Profile& p, int l, bool null2, float fwdsc, float nexpected, int nregions,
                     const std::vector<float>& n2sc, const std::vector<Envelope>& envs, RawHit& out) {
    if (nregions == 0 || envs.empty()) return false;
    const float nullsc = null_score((size_t)l);

    // sequence and domain scores
    const float omega_ln = std::log(1.0f / 256.0f);
    float seqbias = 0.0f;
    if (null2) {
        float s = 0.0f;
        for (int v = 1; v <= l; v++) s += n2sc[v];
        seqbias = logsum(0.0f, omega_ln + s);
    }
    float pre_score = (fwdsc - nullsc) / LN2F;
    float seq_score = (fwdsc - (nullsc + seqbias)) / LN2F;
    const float ln_multi_loop = std::log((float)l / ((float)l + 3.0f));
    float sum = 0.0f, sum_bias = 0.0f;
    int ld_tot = 0;
    for (const Envelope& e : envs) {
        const bool keep = null2 ? e.envsc - e.domcorr > 0.0f : e.envsc > 0.0f;
        if (keep) {
            sum += e.envsc;
            ld_tot += e.j - e.i + 1;
            sum_bias += e.domcorr;
        }
    }
    const float sum_seqbias = null2 ? logsum(0.0f, omega_ln + sum_bias) : 0.0f;
    sum += (float)(l - std::min(ld_tot, l)) * ln_multi_loop;
    const float pre2 = (sum - nullsc) / LN2F;
    const float sum_score = (sum - (nullsc + sum_seqbias)) / LN2F;
    if (ld_tot > 0 && sum_score > seq_score) {
        seq_score = sum_score;
        pre_score = pre2;
    }
    const bool has_fwd = p.stats.has_fwd;
    const double ftau = p.stats.fwd[0], flam = p.stats.fwd[1];
    out.bits = seq_score;
    out.bias = pre_score - seq_score;
    out.lnp = has_fwd ? exp_logsurv((double)seq_score, ftau, flam) : std::numeric_limits<double>::quiet_NaN();
    out.nexpected = nexpected;
    out.nregions = nregions;
    out.domains.clear();
    for (const Envelope& e : envs) {
        const int ld = e.j - e.i + 1;
        const float dombias = null2 ? logsum(0.0f, omega_ln + e.domcorr) : 0.0f;
        const float bits = (e.envsc + (float)(l - ld) * ln_multi_loop - (nullsc + dombias)) / LN2F;
        RawDomain d;
        d.ienv = e.i;
        d.jenv = e.j;
        d.iali = e.ali.iali;
        d.jali = e.ali.jali;
        d.kali = e.ali.kali;
        d.mali = e.ali.mali;
        d.bits = bits;
        d.bias = dombias / LN2F;
        d.lnp = has_fwd ? exp_logsurv((double)bits, ftau, flam) : std::numeric_limits<double>::quiet_NaN();
        out.domains.push_back(d);
    }
    return true;
}

}  // namespace sx

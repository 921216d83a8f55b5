// sx_envelope.hip -- envelope rescoring on the device (sentinel-x src/hmmer
// pipeline.rs:545-591, rescore_isolated_domain): for each envelope, full-matrix Forward
// and Backward over the envelope's rows with the unihit length model (dp.rs:442-601),
// posterior decoding (dp.rs:1274-1306), the null2 sums (dp.rs:1310-1350) and the
// optimal-accuracy alignment with its traceback (dp.rs:1355-1564).
//
// One wave64 per envelope.  A DP row is processed in chunks of 64 model positions, one
// position per lane (k = 64c + lane + 1), so every global access is coalesced.  The
// previous and current rows sit in LDS.  The serial chains of a row are resolved across
// the lanes of a chunk by scans: the D chain of Forward and Backward is affine in the
// value entering the chunk (D_k = A_k + B_k * D_k-1), and the D chain of the
// optimal-accuracy DP is a max over a gated run (od_k = max(c_k, p_k ? od_k-1 : -inf)).
// Row sums and maxima are wave reductions.  Forward / Backward sums therefore run in
// another order than the reference's serial loops (their contract is an error bound, as
// for the filter kernels); the optimal-accuracy recurrence uses only maxima and single
// additions, and its traceback follows the reference's tie rules step for step.
//
// The matrices live in a device workspace: Forward M and I rows, Backward M and I rows
// (four floats a cell).  The decoding pass reads them row by row and overwrites them
// with the optimal-accuracy M, I and D rows, which the traceback then reads.
#include <hip/hip_runtime.h>

#include "sx_dev.h"

#define SXE_RESCALE 1.0e4f

namespace sxe {

__device__ __forceinline__ float ninf() { return -__builtin_inff(); }

// Cross-lane steps by DPP (gfx9): row shifts inside 16-lane rows, whole-wave shifts by
// one lane, and v_readlane for the row totals.  A lane with no source keeps `old`.
template <int CTRL>
__device__ __forceinline__ float dppf(float old, float src) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ float dppm(float src) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(src), CTRL, 0xF, 0xF, true));
}
__device__ __forceinline__ float rdl(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Wave sum / maximum; every lane gets the same value (butterflies inside a row: quad
// xor 1, quad xor 2, half-row mirror, row mirror; then the four row totals).
__device__ __forceinline__ float wsum(float v) {
    v += dppm<0xB1>(v);
    v += dppm<0x4E>(v);
    v += dppm<0x141>(v);
    v += dppm<0x140>(v);
    return (rdl(v, 0) + rdl(v, 16)) + (rdl(v, 32) + rdl(v, 48));
}
__device__ __forceinline__ float wmax(float v) {
    v = fmaxf(v, dppm<0xB1>(v));
    v = fmaxf(v, dppm<0x4E>(v));
    v = fmaxf(v, dppm<0x141>(v));
    v = fmaxf(v, dppm<0x140>(v));
    return fmaxf(fmaxf(rdl(v, 0), rdl(v, 16)), fmaxf(rdl(v, 32), rdl(v, 48)));
}
__device__ __forceinline__ int wmaxi(int v) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v = max(v, __shfl_xor(v, s, 64));
    return v;
}
__device__ __forceinline__ int wmini(int v) {
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) v = min(v, __shfl_xor(v, s, 64));
    return v;
}

// lane l receives lane l-1's v, lane 0 receives b (wave_shr:1)
__device__ __forceinline__ float from_prev(float v, float b) { return dppf<0x138>(b, v); }
// lane l receives lane l+1's v, lane 63 receives b (wave_shl:1)
__device__ __forceinline__ float from_next(float v, float b) { return dppf<0x130>(b, v); }

template <int CTRL>
__device__ __forceinline__ void affine_step(float& A, float& B) {
    const float a = dppf<CTRL>(0.0f, A), b = dppf<CTRL>(1.0f, B);  // identity (0, 1) past the row edge
    A = fmaf(B, a, A);
    B = B * b;
}

// Inclusive scan of affine maps D -> A + B*D over the wave, left to right: lane l ends
// with the maps of lanes 0..l composed (lane 0 applied first).
__device__ __forceinline__ void scan_affine_up(float& A, float& B, int lane) {
    affine_step<0x111>(A, B);  // row_shr:1
    affine_step<0x112>(A, B);
    affine_step<0x114>(A, B);
    affine_step<0x118>(A, B);
    // rows 1..3 take the composition of the rows before them (row totals in lanes 15, 31, 47)
    const float a0 = rdl(A, 15), b0 = rdl(B, 15), a1 = rdl(A, 31), b1 = rdl(B, 31), a2 = rdl(A, 47), b2 = rdl(B, 47);
    const float e1a = a0, e1b = b0;
    const float e2a = fmaf(b1, e1a, a1), e2b = b1 * e1b;
    const float e3a = fmaf(b2, e2a, a2), e3b = b2 * e2b;
    const int row = lane >> 4;
    const float ea = row == 0 ? 0.0f : row == 1 ? e1a : row == 2 ? e2a : e3a;
    const float eb = row == 0 ? 1.0f : row == 1 ? e1b : row == 2 ? e2b : e3b;
    A = fmaf(B, ea, A);
    B = B * eb;
}
// The same, right to left: lane l ends with lanes l..63 composed (lane 63 applied first).
__device__ __forceinline__ void scan_affine_down(float& A, float& B, int lane) {
    affine_step<0x101>(A, B);  // row_shl:1
    affine_step<0x102>(A, B);
    affine_step<0x104>(A, B);
    affine_step<0x108>(A, B);
    // rows 0..2 take the composition of the rows after them (row totals in lanes 16, 32, 48)
    const float a1 = rdl(A, 16), b1 = rdl(B, 16), a2 = rdl(A, 32), b2 = rdl(B, 32), a3 = rdl(A, 48), b3 = rdl(B, 48);
    const float f3a = a3, f3b = b3;
    const float f2a = fmaf(b2, f3a, a2), f2b = b2 * f3b;
    const float f1a = fmaf(b1, f2a, a1), f1b = b1 * f2b;
    const int row = lane >> 4;
    const float ea = row == 3 ? 0.0f : row == 2 ? f3a : row == 1 ? f2a : f1a;
    const float eb = row == 3 ? 1.0f : row == 2 ? f3b : row == 1 ? f2b : f1b;
    A = fmaf(B, ea, A);
    B = B * eb;
}

template <int CTRL>
__device__ __forceinline__ void gmax_step(float& C, float& P) {
    const float c = dppf<CTRL>(ninf(), C), p = dppf<CTRL>(1.0f, P);  // identity (-inf, pass)
    C = fmaxf(C, P != 0.0f ? c : ninf());
    P = P * p;
}
// Inclusive scan of the gated maxima od -> max(C, P ? od : -inf) (P = 1 or 0), left to
// right.
__device__ __forceinline__ void scan_gmax_up(float& C, float& P, int lane) {
    gmax_step<0x111>(C, P);
    gmax_step<0x112>(C, P);
    gmax_step<0x114>(C, P);
    gmax_step<0x118>(C, P);
    const float c0 = rdl(C, 15), p0 = rdl(P, 15), c1 = rdl(C, 31), p1 = rdl(P, 31), c2 = rdl(C, 47), p2 = rdl(P, 47);
    const float e1c = c0, e1p = p0;
    const float e2c = fmaxf(c1, p1 != 0.0f ? e1c : ninf()), e2p = p1 * e1p;
    const float e3c = fmaxf(c2, p2 != 0.0f ? e2c : ninf()), e3p = p2 * e2p;
    const int row = lane >> 4;
    const float ec = row == 0 ? ninf() : row == 1 ? e1c : row == 2 ? e2c : e3c;
    const float ep = row == 0 ? 1.0f : row == 1 ? e1p : row == 2 ? e2p : e3p;
    C = fmaxf(C, P != 0.0f ? ec : ninf());
    P = P * ep;
}

// dst[k] = src[k] for k = 1..M, lane-strided (lane l copies k = 1 + l + 64u, the positions
// the chunk loops give it), eight loads in flight at a time
__device__ __forceinline__ void stage_row(float* dst, const float* __restrict__ src, int M, int lane) {
    int k = lane + 1;
    for (; k + 7 * 64 <= M; k += 8 * 64) {
        float v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = src[k + 64 * u];
#pragma unroll
        for (int u = 0; u < 8; u++) dst[k + 64 * u] = v[u];
    }
    for (; k <= M; k += 64) dst[k] = src[k];
}

// Orders one wave's LDS accesses across its lanes (a wave's LDS operations complete in
// order; this keeps the compiler from moving them across the point either).  The passes
// that one wave runs alone use it instead of a workgroup barrier.
__device__ __forceinline__ void wsync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

__device__ __forceinline__ float lnf(float x) { return x > 0.0f ? logf(x) : ninf(); }

// dp.rs:32-40
__device__ __forceinline__ float logsum(float a, float b) {
    const float hi = a > b ? a : b, lo = a > b ? b : a;
    const float d = hi - lo;
    if (!(d < 30.0f)) return hi;
    return hi + log1pf(expf(-d));
}

// an edge with zero probability contributes 0 to the max (dp.rs:1380-1420)
__device__ __forceinline__ float gate(float t, float v) { return t > 0.0f ? v : 0.0f; }
// an edge with zero probability is not a traceback candidate
__device__ __forceinline__ float dis(float t, float v) { return t == 0.0f ? ninf() : v; }

}  // namespace sxe

using namespace sxe;

// LDS, in rows of M+1 floats: previous and current M, I, D rows, the null2 column sums and
// the staged emission row (8 rows); with TL also the profile's eight transition rows and
// the four Forward / Backward matrix rows the decoding pass reads (20 rows).  Each row's
// global operands are staged by a copy loop with several loads in flight, so the chunk
// loops (serial through their lane scans) wait on LDS, not on memory.
//
// PAR: a workgroup of two waves, the Forward pass on wave 0 and the Backward pass on wave 1
// at the same time (they are independent), each with its own rows in LDS; after one
// workgroup barrier wave 0 runs the decoding pass and the traceback alone.  LDS then holds
// Forward rows (7 x (M+1) floats), Backward rows (7), with TL the transition rows (8), and
// one exchange slot; the decoding pass reuses the first 11 rows.
template <bool TL, bool PAR>
__device__ __forceinline__ void rescore_body(const SxEnvJob* __restrict__ jobs, const SxProfiles& P,
                                             const SxTargets& T, float* __restrict__ ws, double* __restrict__ dws,
                                             SxEnvRes* __restrict__ res, float* lds) {
    const SxEnvJob job = jobs[blockIdx.x];
    const int lane = (int)(threadIdx.x & 63);
    const int wave = (int)(threadIdx.x >> 6);
    const int M = P.m[job.prof];
    const int W = M + 1;
    const int Lr = job.j - job.i + 1;
    const int R = Lr + 1;
    const int nch = (M + 63) >> 6;
    const float* __restrict__ mo = P.mo + P.off[job.prof];
    const float* __restrict__ tg = P.tp + P.toff[job.prof];
    const int tof = PAR ? 14 * W : 8 * W;  // the staged transition rows (TL)
    if constexpr (TL) {
        float* tl = lds + tof;
        for (int k = (int)threadIdx.x; k < 8 * W; k += (int)blockDim.x) tl[k] = tg[k];
        __syncthreads();
    }
    float* xslot = lds + (TL ? tof + 8 * W : tof);  // PAR: the Backward score, wave 1 -> wave 0
    const float* __restrict__ tp = TL ? lds + tof : tg;
    const float* __restrict__ pmm = tp + SX_TMM * W;
    const float* __restrict__ pmi = tp + SX_TMI * W;
    const float* __restrict__ pmd = tp + SX_TMD * W;
    const float* __restrict__ pim = tp + SX_TIM * W;
    const float* __restrict__ pii = tp + SX_TII * W;
    const float* __restrict__ pdm = tp + SX_TDM * W;
    const float* __restrict__ pdd = tp + SX_TDD * W;
    const float* __restrict__ pbm = tp + SX_TBM * W;
    const uint8_t* __restrict__ sub = T.codes + T.off[job.target] + (job.i - 1);
    // unihit length model: N, J, C loops and moves share one probability; E -> C = 1,
    // E -> J = 0
    const float ploop = job.ploop, pmove = job.pmove;
    const float pe_move = 1.0f, pe_loop = 0.0f;
    const float lloop = job.lloop, lmove = job.lmove;

    const size_t RW = (size_t)R * W;
    float* FM = ws + job.wsoff;
    float* FI = FM + RW;
    float* BM = FI + RW;
    float* BI = BM + RW;
    float* rowf = BI + RW;
    float* fsn = rowf;
    float* fsj = fsn + R;
    float* fsc = fsj + R;
    float* bsn = fsc + R;
    float* bsj = bsn + R;
    float* bsc = bsj + R;
    float* ppn = bsc + R;
    float* ppj = ppn + R;
    float* ppc = ppj + R;
    float* oxe = ppc + R;
    float* oxn = oxe + R;
    float* oxj = oxn + R;
    float* oxb = oxj + R;
    float* oxc = oxb + R;
    double* fls = dws + job.dwsoff;
    double* bls = fls + R;

    // the pass's rows: previous and current M, I, D, and its staged emission row
    float* rb = PAR && wave == 1 ? lds + 7 * W : lds;
    float* aM = rb;
    float* aI = rb + W;
    float* aD = rb + 2 * W;
    float* cM = rb + 3 * W;
    float* cI = rb + 4 * W;
    float* cD = rb + 5 * W;
    float* eL = rb + (PAR ? 6 * W : 7 * W);
    float* sm = lds + 6 * W;  // decoding pass: null2 column sums
    float* sFM = lds + (PAR ? 7 * W : 16 * W);  // with TL: staged Forward / Backward rows of the decoding pass
    float* sFI = sFM + W;
    float* sBM = sFM + 2 * W;
    float* sBI = sFM + 3 * W;

    SxEnvRes out;
    out.envsc = ninf();
    out.bcksc = ninf();
    out.status = 1;
    out.oa = 0;
    out.iali = out.jali = out.kali = out.mali = 0;
    for (int x = 0; x < 20; x++) out.null2[x] = 0.0f;

    float xn = 1.0f, xb = pmove, xj = 0.0f, xc = 0.0f;
    // ---------------------------------------------------------------- Forward (dp.rs:443-522)
    if (!PAR || wave == 0) {
    for (int k = lane; k < W; k += 64) {
        aM[k] = aI[k] = aD[k] = 0.0f;
        cM[k] = cI[k] = cD[k] = 0.0f;
    }
    wsync();
    double lnscale = 0.0;
    if (lane == 0) {
        fsn[0] = 0.0f;
        fsj[0] = ninf();
        fsc[0] = ninf();
        fls[0] = 0.0;
    }
    for (int r = 1; r <= Lr; r++) {
        const int x = sub[r - 1];
        stage_row(eL, mo + (size_t)x * W, M, lane);
        wsync();
        const float* eo = eL;
        const float ins = x != SX_CODE_STAR ? 1.0f : 0.0f;
        float carry_m = 0.0f, carry_d = 0.0f, part = 0.0f;
        for (int c = 0; c < nch; c++) {
            const int k = (c << 6) + lane + 1;
            const bool valid = k <= M;
            float mv = 0.0f, iv = 0.0f, A = 0.0f, B = 1.0f;
            if (valid) {
                // summation order B, M, I, D
                mv = (xb * pbm[k] + aM[k - 1] * pmm[k - 1] + aI[k - 1] * pim[k - 1] + aD[k - 1] * pdm[k - 1]) * eo[k];
                iv = k < M ? (aM[k] * pmi[k] + aI[k] * pii[k]) * ins : 0.0f;
                B = pdd[k - 1];
            }
            const float mprev = from_prev(mv, carry_m);
            if (valid) A = mprev * pmd[k - 1];
            scan_affine_up(A, B, lane);
            const float dv = fmaf(B, carry_d, A);
            if (valid) {
                cM[k] = mv;
                cI[k] = iv;
                cD[k] = dv;
                part += mv + dv;
            }
            carry_m = rdl(mv, 63);
            carry_d = rdl(dv, 63);
        }
        float xe = wsum(part);
        xj = xj * ploop + xe * pe_loop;
        xc = xc * ploop + xe * pe_move;
        xn *= ploop;
        xb = xn * pmove + xj * pmove;
        float sc = 1.0f;
        if (xe > SXE_RESCALE) {
            sc = 1.0f / xe;
            xn *= sc;
            xb *= sc;
            xj *= sc;
            xc *= sc;
            lnscale += log((double)xe);
            xe = 1.0f;
        }
        float* gM = FM + (size_t)r * W;
        float* gI = FI + (size_t)r * W;
        for (int c = 0; c < nch; c++) {
            const int k = (c << 6) + lane + 1;
            if (k <= M) {
                float vm = cM[k], vi = cI[k];
                if (sc != 1.0f) {
                    vm *= sc;
                    vi *= sc;
                    cM[k] = vm;
                    cI[k] = vi;
                    cD[k] *= sc;
                }
                gM[k] = vm;
                gI[k] = vi;
            }
        }
        if (lane == 0) {
            const float ls = (float)lnscale;
            fls[r] = lnscale;
            fsn[r] = lnf(xn) + ls;
            fsj[r] = lnf(xj) + ls;
            fsc[r] = lnf(xc) + ls;
        }
        wsync();
        float* t;
        t = aM; aM = cM; cM = t;
        t = aI; aI = cI; cI = t;
        t = aD; aD = cD; cD = t;
    }
    out.envsc = (float)((double)lnf(xc * pmove) + lnscale);
    }
    if constexpr (!PAR) {
        if (!isfinite(out.envsc)) {
            if (lane == 0) res[blockIdx.x] = out;
            return;
        }
    }

    // ---------------------------------------------------------------- Backward (dp.rs:525-601)
    if (!PAR || wave == 1) {
        double lnscale = 0.0;
        xc = pmove;
        float xe = xc * pe_move;
        xj = 0.0f;
        xn = 0.0f;
        // row Lr: M and D from E; I = 0
        float carry_d = 0.0f;
        float* gM = BM + (size_t)Lr * W;
        float* gI = BI + (size_t)Lr * W;
        for (int c = nch - 1; c >= 0; c--) {
            const int k = (c << 6) + lane + 1;
            const bool valid = k <= M;
            float A = 0.0f, B = 1.0f;
            if (valid) {
                A = xe;
                B = k < M ? pdd[k] : 0.0f;
            }
            scan_affine_down(A, B, lane);
            const float dv = fmaf(B, carry_d, A);
            const float dnext = from_next(dv, carry_d);
            if (valid) {
                const float mv = k == M ? xe : xe + dnext * pmd[k];
                cM[k] = mv;
                cI[k] = 0.0f;
                cD[k] = dv;
                gM[k] = mv;
                gI[k] = 0.0f;
            }
            carry_d = rdl(dv, 0);
        }
        lnscale = 0.0;
        if (lane == 0) {
            bls[Lr] = 0.0;
            bsn[Lr] = ninf();
            bsj[Lr] = ninf();
            bsc[Lr] = lnf(xc);
        }
        wsync();
        float* t;
        t = aM; aM = cM; cM = t;
        t = aI; aI = cI; cI = t;
        t = aD; aD = cD; cD = t;
        float sn1 = ninf();  // N of row 1 (row Lr's N is -inf)
        for (int r = Lr - 1; r >= 1; r--) {
            const int xnext = sub[r];
            stage_row(eL, mo + (size_t)xnext * W, M, lane);
            wsync();
            const float* eo = eL;
            const float ins = xnext != SX_CODE_STAR ? 1.0f : 0.0f;
            float part = 0.0f;
                for (int c = 0; c < nch; c++) {
                const int k = (c << 6) + lane + 1;
                if (k <= M) part += aM[k] * pbm[k] * eo[k];
            }
            float xbv = wsum(part);
            xj = xj * ploop + xbv * pmove;
            xc *= ploop;
            xe = xj * pe_loop + xc * pe_move;
            xn = xn * ploop + xbv * pmove;
            float pmx = ninf();
            carry_d = 0.0f;
                for (int c = nch - 1; c >= 0; c--) {
                const int k = (c << 6) + lane + 1;
                const bool valid = k <= M;
                float A = 0.0f, B = 1.0f, mnext = 0.0f, ivv = 0.0f;
                if (valid) {
                    if (k < M) {
                        mnext = aM[k + 1] * eo[k + 1];
                        ivv = aI[k] * ins;
                        A = mnext * pdm[k] + xe;
                        B = pdd[k];
                    } else {
                        A = xe;
                        B = 0.0f;
                    }
                }
                scan_affine_down(A, B, lane);
                const float dv = fmaf(B, carry_d, A);
                const float dnext = from_next(dv, carry_d);
                if (valid) {
                    float mv, iv;
                    if (k < M) {
                        mv = mnext * pmm[k] + ivv * pmi[k] + xe + dnext * pmd[k];
                        iv = mnext * pim[k] + ivv * pii[k];
                    } else {
                        mv = xe;
                        iv = 0.0f;
                    }
                    cM[k] = mv;
                    cI[k] = iv;
                    cD[k] = dv;
                    pmx = fmaxf(pmx, mv);
                }
                carry_d = rdl(dv, 0);
            }
            const float rowmax = fmaxf(fmaxf(fmaxf(fmaxf(fmaxf(xbv, xe), xn), xj), xc), wmax(pmx));
            float sc = 1.0f;
            if (rowmax > SXE_RESCALE) {
                sc = 1.0f / rowmax;
                xbv *= sc;
                xe *= sc;
                xn *= sc;
                xj *= sc;
                xc *= sc;
                lnscale += log((double)rowmax);
            }
            gM = BM + (size_t)r * W;
            gI = BI + (size_t)r * W;
            for (int c = 0; c < nch; c++) {
                const int k = (c << 6) + lane + 1;
                if (k <= M) {
                    float vm = cM[k], vi = cI[k];
                    if (sc != 1.0f) {
                        vm *= sc;
                        vi *= sc;
                        cM[k] = vm;
                        cI[k] = vi;
                        cD[k] *= sc;
                    }
                    gM[k] = vm;
                    gI[k] = vi;
                }
            }
            const float ls = (float)lnscale;
            if (r == 1) sn1 = lnf(xn) + ls;
            if (lane == 0) {
                bls[r] = lnscale;
                bsn[r] = lnf(xn) + ls;
                bsj[r] = lnf(xj) + ls;
                bsc[r] = lnf(xc) + ls;
            }
            wsync();
            t = aM; aM = cM; cM = t;
            t = aI; aI = cI; cI = t;
            t = aD; aD = cD; cD = t;
        }
        // row 0: B from row 1 and residue 1; the score is N(0)
        stage_row(eL, mo + (size_t)sub[0] * W, M, lane);
        wsync();
        const float* eo = eL;
        float part = 0.0f;
        for (int c = 0; c < nch; c++) {
            const int k = (c << 6) + lane + 1;
            if (k <= M) part += aM[k] * pbm[k] * eo[k];
        }
        const float b0 = wsum(part);
        const float sb0 = lnf(b0) + (float)lnscale;
        out.bcksc = logsum(sn1 + lloop, sb0 + lmove);
        if (PAR && lane == 0) xslot[0] = out.bcksc;
    }
    if constexpr (PAR) {
        // the two passes' rows and scores in global memory and the exchange slot become
        // visible to wave 0, which goes on alone
        __threadfence();
        __syncthreads();
        if (wave == 1) return;
        __threadfence();
        out.bcksc = xslot[0];
        if (!isfinite(out.envsc)) {
            if (lane == 0) res[blockIdx.x] = out;
            return;
        }
    }
    if (!isfinite(out.bcksc)) {
        out.status = 2;
        if (lane == 0) res[blockIdx.x] = out;
        return;
    }
    out.status = 0;
    // the decoding pass's rows: the first ones of LDS
    aM = lds;
    aI = lds + W;
    aD = lds + 2 * W;
    cM = lds + 3 * W;
    cI = lds + 4 * W;
    cD = lds + 5 * W;
    wsync();

    // ---------------------------------------------------------------- posterior decoding,
    // null2 sums and the optimal-accuracy DP (dp.rs:1274-1306, 1310-1350, 1355-1460)
    const double overall = (double)out.bcksc;
    for (int k = lane; k < W; k += 64) {
        aM[k] = aI[k] = aD[k] = ninf();
        sm[k] = 0.0f;
    }
    wsync();
    float si = 0.0f, sn = 0.0f, sj = 0.0f, scs = 0.0f;
    float oxn_p = 0.0f, oxb_p = 0.0f, oxj_p = ninf(), oxc_p = ninf();
    if (lane == 0) {
        oxe[0] = ninf();
        oxn[0] = 0.0f;
        oxj[0] = ninf();
        oxb[0] = 0.0f;
        oxc[0] = ninf();
        ppn[0] = ppj[0] = ppc[0] = 0.0f;
    }
    for (int c = 0; c < nch; c++) {
        const int k = (c << 6) + lane + 1;
        if (k <= M) FM[k] = FI[k] = BM[k] = ninf();  // row 0 of the optimal-accuracy matrices
    }
    for (int r = 1; r <= Lr; r++) {
        const double fac = exp(fls[r] + bls[r] - overall);
        const float pn = (float)exp((double)(fsn[r - 1] + bsn[r] + lloop) - overall);
        const float pj = (float)exp((double)(fsj[r - 1] + bsj[r] + lloop) - overall);
        const float pc = (float)exp((double)(fsc[r - 1] + bsc[r] + lloop) - overall);
        float* gM = FM + (size_t)r * W;
        float* gI = FI + (size_t)r * W;
        float* gD = BM + (size_t)r * W;
        const float* fM = gM;
        const float* fI = gI;
        const float* hM = BM + (size_t)r * W;
        const float* hI = BI + (size_t)r * W;
        if constexpr (TL) {
            stage_row(sFM, gM, M, lane);
            stage_row(sFI, gI, M, lane);
            stage_row(sBM, hM, M, lane);
            stage_row(sBI, hI, M, lane);
            wsync();
            fM = sFM;
            fI = sFI;
            hM = sBM;
            hI = sBI;
        }
        float carry_m = ninf(), carry_d = ninf(), emax = ninf();
        for (int c = 0; c < nch; c++) {
            const int k = (c << 6) + lane + 1;
            const bool valid = k <= M;
            float mv = ninf(), ov = ninf(), C = ninf();
            float Pg = 1.0f;
            if (valid) {
                const float ppm = (float)((double)fM[k] * (double)hM[k] * fac);
                const float ppi = k < M ? (float)((double)fI[k] * (double)hI[k] * fac) : 0.0f;
                sm[k] += ppm;
                si += ppi;
                float sv = gate(pbm[k], oxb_p);
                if (k >= 2) {
                    sv = fmaxf(sv, gate(pmm[k - 1], aM[k - 1]));
                    sv = fmaxf(sv, gate(pim[k - 1], aI[k - 1]));
                    sv = fmaxf(sv, gate(pdm[k - 1], aD[k - 1]));
                } else {
                    sv = fmaxf(sv, 0.0f);
                }
                mv = sv + ppm;
                ov = fmaxf(gate(pmi[k], aM[k]), gate(pii[k], aI[k])) + ppi;
            }
            // od_1 = -inf; od_k = max(gate(MD, om_k-1), gate(DD, od_k-1)) for k >= 2
            const float mprev = from_prev(mv, carry_m);
            if (valid && k >= 2) {
                const bool pass = pdd[k - 1] > 0.0f;
                const float g = gate(pmd[k - 1], mprev);
                C = pass ? g : fmaxf(g, 0.0f);
                Pg = pass ? 1.0f : 0.0f;
            } else if (valid) {
                C = ninf();
                Pg = 0.0f;
            }
            scan_gmax_up(C, Pg, lane);
            const float dv = fmaxf(C, Pg != 0.0f ? carry_d : ninf());
            if (valid) {
                cM[k] = mv;
                cI[k] = ov;
                cD[k] = dv;
                gM[k] = mv;
                gI[k] = ov;
                gD[k] = dv;
                emax = fmaxf(emax, fmaxf(mv, dv));
            }
            carry_m = rdl(mv, 63);
            carry_d = rdl(dv, 63);
        }
        const float xe = wmax(emax);
        const float xjv = fmaxf(ploop == 0.0f ? 0.0f : oxj_p + pj, pe_loop == 0.0f ? 0.0f : xe);
        const float xcv = fmaxf(ploop == 0.0f ? 0.0f : oxc_p + pc, pe_move == 0.0f ? 0.0f : xe);
        const float xnv = ploop == 0.0f ? 0.0f : oxn_p + pn;
        const float xbv = fmaxf(pmove == 0.0f ? 0.0f : xnv, pmove == 0.0f ? 0.0f : xjv);
        if (lane == 0) {
            oxe[r] = xe;
            oxj[r] = xjv;
            oxc[r] = xcv;
            oxn[r] = xnv;
            oxb[r] = xbv;
            ppc[r] = pc;
        }
        oxj_p = xjv;
        oxc_p = xcv;
        oxn_p = xnv;
        oxb_p = xbv;
        sn += pn;
        sj += pj;
        scs += pc;
        wsync();
        float* t;
        t = aM; aM = cM; cM = t;
        t = aI; aI = cI; cI = t;
        t = aD; aD = cD; cD = t;
    }

    // null2 odds (dp.rs:1310-1350): expected emissions of the envelope's posterior
    if (job.null2) {
        const float norm = 1.0f / (float)(Lr > 1 ? Lr : 1);
        const float xfactor = sn * norm + scs * norm + sj * norm;
        const float isum = wsum(si) * norm;
        for (int x = 0; x < 20; x++) {
            const float* __restrict__ row = mo + (size_t)x * W;
            float v = 0.0f;
            for (int c = 0; c < nch; c++) {
                const int k = (c << 6) + lane + 1;
                if (k <= M) v += sm[k] * norm * row[k];
            }
            out.null2[x] = wsum(v) + isum + xfactor;
        }
    }

    // ---------------------------------------------------------------- traceback
    // (dp.rs:1462-1564).  Every lane follows the same path (the reads are broadcasts);
    // the E state's choice of the row's best M or D cell is a wave reduction with the
    // reference's visiting order and tie rules.
    __threadfence();
    wsync();
    enum { ST_C, ST_E, ST_M, ST_I, ST_D, ST_B, ST_J, ST_N };
    const int qs = max(2, (M - 1) / 4 + 1);
    int s0 = ST_C, i = Lr, k = 0;
    bool have = false;
    int fi = 0, fk = 0, li = 0, lk = 0;
    const long limit = 8L * (Lr + 2) * (M + 2) + 64;
    for (long guard = 1; guard <= limit; guard++) {
        int s1;
        if (s0 == ST_C) {
            if (i <= 0) break;
            const float a = ploop == 0.0f ? ninf() : oxc[i - 1] + ppc[i];
            const float b = pe_move == 0.0f ? ninf() : oxe[i];
            s1 = a > b ? ST_C : ST_E;
        } else if (s0 == ST_E) {
            // visiting order: stripe q = 0..qs-1, then M cells of the four segments
            // (kk = r*qs + q + 1), then their D cells; an M cell wins ties against what
            // came before (>=), a D cell does not (>)
            const float* rm = FM + (size_t)i * W;
            const float* rd = BM + (size_t)i * W;
            float mx = ninf();
            for (int c = 0; c < nch; c++) {
                const int kk = (c << 6) + lane + 1;
                if (kk <= M) mx = fmaxf(mx, fmaxf(rm[kk], rd[kk]));
            }
            mx = wmax(mx);
            int bm = -1, bd = 0x7fffffff;
            for (int c = 0; c < nch; c++) {
                const int kk = (c << 6) + lane + 1;
                if (kk <= M) {
                    const int q = (kk - 1) % qs, rr = (kk - 1) / qs;
                    if (rm[kk] == mx) bm = max(bm, q * 8 + rr);
                    if (rd[kk] == mx) bd = min(bd, q * 8 + 4 + rr);
                }
            }
            bm = wmaxi(bm);
            bd = wmini(bd);
            if (bm >= 0) {
                k = (bm & 7) * qs + (bm >> 3) + 1;
                s1 = ST_M;
            } else if (bd != 0x7fffffff) {
                k = ((bd & 7) - 4) * qs + (bd >> 3) + 1;
                s1 = ST_D;
            } else {
                break;
            }
        } else if (s0 == ST_M) {
            // a consistent trace reaches M only at i >= 1, k >= 1 (row 0 and position 0
            // hold -inf); the guard keeps every read inside the matrices
            if (i < 1 || k < 1) break;
            if (!have) {
                li = i;
                lk = k;
                have = true;
            }
            fi = i;
            fk = k;
            const size_t prow = (size_t)(i - 1) * W;
            float mv = ninf(), iv = ninf(), dv = ninf();
            if (k >= 2) {
                mv = dis(pmm[k - 1], FM[prow + k - 1]);
                iv = dis(pim[k - 1], FI[prow + k - 1]);
                dv = dis(pdm[k - 1], BM[prow + k - 1]);
            }
            const float bv = dis(pbm[k], oxb[i - 1]);
            int best = 0;
            float bvv = mv;
            if (iv > bvv) { best = 1; bvv = iv; }
            if (dv > bvv) { best = 2; bvv = dv; }
            if (bv > bvv) { best = 3; bvv = bv; }
            k--;
            i--;
            s1 = best == 0 ? ST_M : best == 1 ? ST_I : best == 2 ? ST_D : ST_B;
        } else if (s0 == ST_D) {
            if (i < 1 || k < 1) break;
            const size_t row = (size_t)i * W;
            float a = ninf(), b = ninf();
            if (k >= 2) {
                a = dis(pmd[k - 1], FM[row + k - 1]);
                b = dis(pdd[k - 1], BM[row + k - 1]);
            }
            k--;
            s1 = a >= b ? ST_M : ST_D;
        } else if (s0 == ST_I) {
            if (i < 1 || k < 1 || k > M) break;
            const size_t prow = (size_t)(i - 1) * W;
            const float a = dis(pmi[k], FM[prow + k]);
            const float b = dis(pii[k], FI[prow + k]);
            i--;
            s1 = a >= b ? ST_M : ST_I;
        } else if (s0 == ST_B) {
            if (i < 0) break;
            const float a = pmove == 0.0f ? ninf() : oxn[i];
            const float b = pmove == 0.0f ? ninf() : oxj[i];
            if (a > b)
                s1 = ST_N;
            else
                break;  // a domain begins here; a unihit trace has one
        } else {
            break;  // J or N
        }
        if ((s1 == ST_N || s1 == ST_J || s1 == ST_C) && s1 == s0) i--;
        if (s1 == ST_N) break;
        s0 = s1;
    }
    if (have) {
        out.oa = 1;
        out.iali = fi;
        out.jali = li;
        out.kali = fk;
        out.mali = lk;
    }
    if (lane == 0) res[blockIdx.x] = out;
}

// one wave per envelope, transitions through the caches (8 x (M+1) floats of LDS)
extern "C" __global__ void __launch_bounds__(64)
sxe_rescore(const SxEnvJob* __restrict__ jobs, SxProfiles P, SxTargets T, float* __restrict__ ws,
            double* __restrict__ dws, SxEnvRes* __restrict__ res) {
    extern __shared__ float lds[];
    rescore_body<false, false>(jobs, P, T, ws, dws, res, lds);
}

// two waves per envelope (Forward and Backward side by side), transitions through the
// caches (14 x (M+1) + 1 floats of LDS)
extern "C" __global__ void __launch_bounds__(128)
sxe_rescore_p2(const SxEnvJob* __restrict__ jobs, SxProfiles P, SxTargets T, float* __restrict__ ws,
               double* __restrict__ dws, SxEnvRes* __restrict__ res) {
    extern __shared__ float lds[];
    rescore_body<false, true>(jobs, P, T, ws, dws, res, lds);
}

// two waves per envelope, transitions and the decoding pass's rows staged in LDS
// (22 x (M+1) + 1 floats of LDS)
extern "C" __global__ void __launch_bounds__(128)
sxe_rescore_p2t(const SxEnvJob* __restrict__ jobs, SxProfiles P, SxTargets T, float* __restrict__ ws,
                double* __restrict__ dws, SxEnvRes* __restrict__ res) {
    extern __shared__ float lds[];
    rescore_body<true, true>(jobs, P, T, ws, dws, res, lds);
}

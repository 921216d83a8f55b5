// sx_cascade.hip -- device side of the filter cascade of score_target
// (sentinel-x src/hmmer/pipeline.rs:159-189): pass tests, order-preserving survivor
// compaction, the two-state composition (bias) filter (pipeline.rs:107-150), tile
// construction for the next stage, and the final record assembly.
//
// Pass tests.  The reference rejects a pair when a tail probability of the stage's
// bit score y = ((score - ref) / ln 2) exceeds a threshold (gumbel_surv for MSV and
// Viterbi, exp(exp_logsurv) for Forward).  Each of those tails is monotone in y, so
// the host finds, per profile and with the same libm, the smallest f32 y that passes;
// the device computes y with the reference's f32 operations and compares.  The
// decision is therefore the reference's whenever the score is.
//
// Survivor lists keep the order (profile position, target length): every list is built
// by a count / exclusive-scan / write sequence, never by atomic appends, so a stage's
// tiles are contiguous runs of one profile.  Built without -ffp-contract, so the bias
// filter's f64 products, sums and quotients round as the reference's; its logarithm is
// taken differently, under an error bound that certifies the f32 score or hands the pair
// to the host (sxc_bias).
#include <hip/hip_runtime.h>

#include "../../include/sx_hmmer.h"
#include "sx_dev.h"

#define SXC_LN2F 0.693147182464599609375f
#define SXC_BLOCK 256

namespace sxc {

__device__ __forceinline__ int block_rank(bool pass, int* s_wave, int* total) {
    // exclusive rank of this thread's pass flag inside the block, and the block total
    const unsigned long long m = __ballot(pass ? 1 : 0);
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int below = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    int base = 0, tot = 0;
    for (int w = 0; w < SXC_BLOCK / 64; w++) {
        if (w < wave) base += s_wave[w];
        tot += s_wave[w];
    }
    *total = tot;
    return base + below;
}

}  // namespace sxc

// MSV test over every pair, enumerated as (profile position pi, length-sorted target j).
// usc is indexed [profile][j] and nullsc [j], so the reads are coalesced.
// Phase 0 writes block counts, phase 1 writes the survivors at their scanned offsets.  A
// block tests SX_QCAND_ITEMS x SXC_BLOCK consecutive pairs, as sxc_msvq_cand does.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_msv_pass(int phase, const float* __restrict__ usc, const float* __restrict__ nullsc,
             const float* __restrict__ thr, const int32_t* __restrict__ perm, const int32_t* __restrict__ sorted,
             int64_t P, int64_t T, int32_t* counts, const int64_t* offsets, int64_t cap, int32_t* s_p, int32_t* s_pi,
             int32_t* s_t, float* s_usc) {
    __shared__ int s_wave[SXC_BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * SXC_BLOCK * SX_QCAND_ITEMS;
    int64_t run = phase ? offsets[blockIdx.x] : 0;
    int32_t total = 0;
    for (int k = 0; k < SX_QCAND_ITEMS; k++) {
        const int64_t pos = base + (int64_t)k * SXC_BLOCK + threadIdx.x;
        bool pass = false;
        int32_t p = 0, pi = 0, t = 0;
        float u = 0.0f;
        if (pos < P * T) {
            pi = (int32_t)(pos / T);
            const int64_t j = pos - (int64_t)pi * T;
            p = perm[pi];
            t = sorted[j];
            u = usc[(int64_t)p * T + j];
            const float y = (u - nullsc[j]) / SXC_LN2F;
            pass = y >= thr[p];
        }
        int tot;
        const int r = sxc::block_rank(pass, s_wave, &tot);
        if (phase == 1 && pass) {
            const int64_t o = run + r;
            if (o < cap) {
                s_p[o] = p;
                s_pi[o] = pi;
                s_t[o] = t;
                s_usc[o] = u;
            }
        }
        run += tot;
        total += tot;
        __syncthreads();  // every lane has read s_wave before the next item's rank rewrites it
    }
    if (phase == 0 && threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Candidates of the quantised MSV bound: every pair (profile position pi, length-sorted
// target j) whose int16 bound reaches the pair's pass threshold in the bound's units.
// The exact test passes a pair only if its f32 score u satisfies
// fl(fl(u - null) / ln2) >= thr, hence u >= null + ln2 * thr - delta with delta
// covering those two roundings; nullq[j] and thrq[p] hold floor(2^22 * (null - ...)) and
// floor(2^22 * (ln2 * thr - ...)), so a bound below their sum proves the rejection.
// A bound that saturated (a cell reached 32 nats) is no longer an upper bound but stays
// at or above capq[j], which the host derives from the target's length model; the
// threshold is capped there, so such a pair is a candidate.  Profiles with noq set
// (longer than the bound's shapes, or without a finite threshold) make every pair a
// candidate; a profile whose threshold no score passes (thrq at or above SX_Q_NOPASS) has
// no bound launch, so its rows of qsc hold stale values and none of its pairs is one.
// Phase 0 writes block counts, phase 1 the candidates (pi, j, target id) at their
// scanned offsets, in (pi, j) order.  A block tests SX_QCAND_ITEMS x SXC_BLOCK
// consecutive pairs, a block's width at a time, so that the scan between the phases runs
// over that many times fewer block counts.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_msvq_cand(int phase, const int32_t* __restrict__ qsc, const int64_t* __restrict__ nullq,
              const int64_t* __restrict__ capq, const int64_t* __restrict__ thrq, const uint8_t* __restrict__ noq,
              const int32_t* __restrict__ perm, const int32_t* __restrict__ sorted, int64_t P, int64_t T,
              int32_t* counts, const int64_t* offsets, int64_t cap, int32_t* c_pi, int32_t* c_j, int32_t* c_t) {
    __shared__ int s_wave[SXC_BLOCK / 64];
    const int64_t base = (int64_t)blockIdx.x * SXC_BLOCK * SX_QCAND_ITEMS;
    int64_t run = phase ? offsets[blockIdx.x] : 0;
    int32_t total = 0;
    for (int k = 0; k < SX_QCAND_ITEMS; k++) {
        const int64_t pos = base + (int64_t)k * SXC_BLOCK + threadIdx.x;
        bool pass = false;
        int32_t pi = 0, j = 0;
        if (pos < P * T) {
            pi = (int32_t)(pos / T);
            j = (int32_t)(pos - (int64_t)pi * T);
            const int32_t p = perm[pi];
            const int64_t tq = thrq[p];
            const int64_t th = nullq[j] + tq;
            pass = noq[p] != 0 ||
                   (tq < SX_Q_NOPASS && (int64_t)qsc[(int64_t)p * T + j] >= (th < capq[j] ? th : capq[j]));
        }
        int tot;
        const int r = sxc::block_rank(pass, s_wave, &tot);
        if (phase == 1 && pass) {
            const int64_t o = run + r;
            if (o < cap) {
                c_pi[o] = pi;
                c_j[o] = j;
                c_t[o] = sorted[j];
            }
        }
        run += tot;
        total += tot;
        __syncthreads();  // every lane has read s_wave before the next item's rank rewrites it
    }
    if (phase == 0 && threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Entries per profile position of a list sorted by position: per[pi] = the length of
// pi's run in sorted_pi[0 .. *n_dev), by two binary searches.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_count_sorted(const int64_t* __restrict__ n_dev, int64_t cap, const int32_t* __restrict__ sorted_pi, int64_t P,
                 int32_t* per) {
    const int64_t pi = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (pi >= P) return;
    const int64_t n = *n_dev < cap ? *n_dev : cap;
    auto lower = [&](int64_t v) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if ((int64_t)sorted_pi[mid] < v)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo;
    };
    per[pi] = (int32_t)(lower(pi + 1) - lower(pi));
}

// The MSV test (pipeline.rs:166-169) over a candidate list with its f32 MSV scores usc[i]:
// the same decision as sxc_msv_pass, and the survivors come out in list order, which is
// the (pi, j) order sxc_msv_pass produces.  Phase 0 counts per block, phase 1 writes.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_msv_pass_list(int phase, const int64_t* __restrict__ n_dev, const int32_t* __restrict__ c_pi,
                  const int32_t* __restrict__ c_j, const float* __restrict__ usc, const float* __restrict__ nullsc,
                  const float* __restrict__ thr, const int32_t* __restrict__ perm, const int32_t* __restrict__ sorted,
                  int32_t* counts, const int64_t* offsets, int64_t cap, int32_t* s_p, int32_t* s_pi, int32_t* s_t,
                  float* s_usc) {
    __shared__ int s_wave[SXC_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    bool pass = false;
    int32_t p = 0, pi = 0, t = 0;
    float u = 0.0f;
    if (i < *n_dev) {
        pi = c_pi[i];
        const int32_t j = c_j[i];
        p = perm[pi];
        t = sorted[j];
        u = usc[i];
        const float y = (u - nullsc[j]) / SXC_LN2F;
        pass = y >= thr[p];
    }
    int tot;
    const int r = sxc::block_rank(pass, s_wave, &tot);
    if (phase == 0) {
        if (threadIdx.x == 0) counts[blockIdx.x] = tot;
        return;
    }
    if (pass) {
        const int64_t o = offsets[blockIdx.x] + r;
        if (o < cap) {
            s_p[o] = p;
            s_pi[o] = pi;
            s_t[o] = t;
            s_usc[o] = u;
        }
    }
}

// Exclusive scan of n int32 counts into n+1 int64 offsets (out[n] = total), one
// workgroup of 1024 threads; n may be a device value.
extern "C" __global__ void __launch_bounds__(1024)
sxc_scan(const int32_t* __restrict__ in, int64_t n_host, const int64_t* n_dev, int64_t* out) {
    __shared__ int64_t s_part[1024];
    const int64_t n = n_dev ? *n_dev : n_host;
    const int tid = (int)threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t a = (int64_t)tid * per, b = a + per < n ? a + per : n;
    int64_t sum = 0;
    for (int64_t i = a; i < b; i++) sum += in[i];
    s_part[tid] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int64_t v = tid >= off ? s_part[tid - off] : 0;
        __syncthreads();
        s_part[tid] += v;
        __syncthreads();
    }
    int64_t run = s_part[tid] - sum;
    for (int64_t i = a; i < b; i++) {
        out[i] = run;
        run += in[i];
    }
    if (tid == 1023) out[n] = s_part[1023];
}

// Two-state composition filter (pipeline.rs:107-150) on each MSV survivor, then the
// bias-corrected MSV test (pipeline.rs:170-173).  eo1[p*25 + x] holds the model-state
// emission odds computed on the host exactly as the reference (the background state's
// odds are 1 for every code, so the reference's product with them is exact).
//
// The two forward values d0, d1 and each row maximum m_i follow the reference's f64
// operations in its order (products, sums, max, division by the maximum; one of the
// two quotients is m_i / m_i = 1 exactly), so they are bit-identical to the host's.
// The reference then sums ln(m_i) with glibc's log, one call a residue; here the m_i are
// multiplied into a running product kept as a mantissa in [0.5, 1) and an integer
// exponent (exact renormalisation), and one device log is taken at the end.  With
// u = 2^-53, the two f64 results differ by at most
//   beta = u (1 + 1/16) [ (L+1) + 2|ex| ln2 + 2k |ln mant| + 2g A + sum_i A_i
//                         + (2(k+g+1) + 1) L |ln p1| + 3 |s1| + 2(k+g) |ln(1-p1)| + 6 |v| ]
// where the device log is within k = 4 ulps and glibc's within g = 2 ulps of the true
// logarithm (both claim 1 ulp or less), A_i = sum_{j<=i} |ln m_j| (each term bounded
// above by |m_j - 1| / min(m_j, 1), evaluated in f32 with a 1% margin and summed in
// f64), A = A_{L+1}, s1 the partial sum before ln(1-p1) and v the result.  The terms are
// the product's roundings (one per factor), the exponent's ln2, the device log and the
// rounding of its sum, glibc's per-residue errors and the host sum's roundings
// (u |partial sum| each), the tail's logs and roundings, and those of v -/+ beta.
// If v - beta and v + beta round to the same f32, that f32 is the reference's score
// (rounding is monotone).  Otherwise the pair is flagged for the host, which recomputes it
// in the reference's order with glibc, and the lower f32 is kept meanwhile: each later
// test passes more pairs with a lower bias score, so every pair the reference passes
// still reaches the next stage and the host can re-decide it from the scores computed.
// With recheck_all every pair is flagged (a test of that host path).
// The flagged survivor indices are appended to fl_list (fl_n[0] counts them; entries
// past fl_cap are dropped, and the host then scans every record's flag instead).
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_bias(const int64_t* __restrict__ n1, const int32_t* __restrict__ s_p, const int32_t* __restrict__ s_t,
         const float* __restrict__ s_usc, const double* __restrict__ eo1, const int32_t* __restrict__ pm,
         const float* __restrict__ thr, SxTargets T, float* s_bias, uint8_t* flag, uint8_t* bflag,
         int32_t* fl_n, int32_t* fl_list, int32_t fl_cap, int32_t recheck_all) {
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (i >= *n1) return;
    const int32_t p = s_p[i], t = s_t[i];
    // targets start 16-byte aligned with 64 bytes of padding after the last: whole dwords
    const uint32_t* wp = reinterpret_cast<const uint32_t*>(T.codes + T.off[t]);
    const int L = T.len[t];
    const double* e1 = eo1 + (int64_t)p * 25;
    const double l0 = 400.0;
    const double l1 = (double)pm[p] / 8.0;
    const double t00 = l0 / (l0 + 1.0), t01 = 1.0 / (l0 + 1.0);
    const double t10 = 1.0 / (l1 + 1.0), t11 = l1 / (l1 + 1.0);
    const double MINPOS = 2.2250738585072014e-308;
    double mant = 1.0;  // running product of the row maxima = mant * 2^ex
    int ex = 0;
    double a_run = 0.0, a_sum = 0.0;  // A_i and sum_i A_i (upper bounds)
    double d0, d1;
    // one row maximum: normalise (d0, d1) by it as the reference does and fold it into
    // the product and the bound sums
    auto take = [&](double n0, double nn1) {
        const double mx = fmax(fmax(n0, nn1), MINPOS);
        if (mx == n0 || mx == nn1) {
            // x / x = 1 exactly: one division for the smaller value
            const double q = (mx == n0 ? nn1 : n0) / mx;
            d0 = mx == n0 ? 1.0 : q;
            d1 = mx == n0 ? q : 1.0;
        } else {
            d0 = n0 / mx;
            d1 = nn1 / mx;
        }
        const double pr = mant * mx;
        ex += __builtin_amdgcn_frexp_exp(pr);
        mant = __builtin_amdgcn_frexp_mant(pr);
        const float mf = (float)mx;
        a_run += (double)((fabsf(mf - 1.0f) + 1e-7f) * __builtin_amdgcn_rcpf(fminf(mf, 1.0f)) * 1.01f);
        a_sum += a_run;
    };
    uint32_t w = wp[0];
    take(1.0 * 0.999, e1[w & 0xFFu] * 0.001);
    for (int k = 1; k < L; k++) {
        if ((k & 3) == 0) w = wp[k >> 2];
        const int x = (int)((w >> ((k & 3) * 8)) & 0xFFu);
        const double n0 = d0 * t00 + d1 * t10;
        const double nn1 = (d0 * t01 + d1 * t11) * e1[x];
        take(n0, nn1);
    }
    // the reference's final ln(d0 + d1) is one more factor of the product
    const double s = d0 + d1;
    {
        const double pr = mant * s;
        ex += __builtin_amdgcn_frexp_exp(pr);
        mant = __builtin_amdgcn_frexp_mant(pr);
        const float sf = (float)s;
        a_run += (double)((fabsf(sf - 1.0f) + 1e-7f) * __builtin_amdgcn_rcpf(fminf(sf, 1.0f)) * 1.01f);
        a_sum += a_run;
    }
    const double lmant = log(mant);
    const double logsc = (double)ex * 0.6931471805599453 + lmant;
    const double p1 = (double)L / ((double)L + 1.0);
    const double lp1 = log(p1), lq1 = log(1.0 - p1);
    const double s1 = logsc + (double)L * lp1;
    const double v = s1 + lq1;
    const double kd = 4.0, kg = 2.0;
    const double beta = 0x1p-53 * (17.0 / 16.0) *
                        ((double)(L + 1) + 2.0 * fabs((double)ex) * 0.6931471805599453 + 2.0 * kd * fabs(lmant) +
                         2.0 * kg * a_run + a_sum + (2.0 * (kd + kg + 1.0) + 1.0) * (double)L * fabs(lp1) +
                         3.0 * fabs(s1) + 2.0 * (kd + kg) * fabs(lq1) + 6.0 * fabs(v));
    const float lo = (float)(v - beta), hi = (float)(v + beta);
    const bool amb = recheck_all != 0 || lo != hi;
    s_bias[i] = lo;
    bflag[i] = amb ? 1 : 0;
    if (amb) {
        const int slot = atomicAdd(fl_n, 1);
        if (slot < fl_cap) fl_list[slot] = (int32_t)i;
    }
    const float y = (s_usc[i] - lo) / SXC_LN2F;
    flag[i] = y >= thr[p] ? 1 : 0;
}

// Order-preserving compaction of a flagged list: out[k] = src[i] for the k-th set flag.
// Phase 0 counts per block, phase 1 writes and counts survivors per profile position.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_compact(int phase, const uint8_t* __restrict__ flag, const int64_t* __restrict__ n_dev,
            const int32_t* __restrict__ src, const int32_t* __restrict__ s1_pi, int32_t* counts,
            const int64_t* offsets, int32_t* out, int32_t* per_pi) {
    __shared__ int s_wave[SXC_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    const bool pass = i < *n_dev && flag[i] != 0;
    int tot;
    const int r = sxc::block_rank(pass, s_wave, &tot);
    if (phase == 0) {
        if (threadIdx.x == 0) counts[blockIdx.x] = tot;
        return;
    }
    if (pass) {
        const int32_t s1 = src ? src[i] : (int32_t)i;
        out[offsets[blockIdx.x] + r] = s1;
        atomicAdd(&per_pi[s1_pi[s1]], 1);
    }
}

// Tiles for a stage: profile position pi gets ceil(cnt[pi] / ng[pi]) tiles of ng[pi]
// consecutive list entries.  Phase 0 writes the per-position tile counts; phase 1
// (after both scans) writes the tiles and the stage's target-id list.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_tiles(int phase, int64_t P, const int32_t* __restrict__ cnt, const int32_t* __restrict__ ng,
          const int32_t* __restrict__ perm, int32_t* tcount, const int64_t* __restrict__ t_off,
          const int64_t* __restrict__ e_off, SxTile* tiles) {
    const int64_t pi = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (pi >= P) return;
    const int c = cnt[pi], g = ng[pi];
    if (phase == 0) {
        tcount[pi] = (c + g - 1) / g;
        return;
    }
    int64_t w = t_off[pi];
    for (int a = 0; a < c; a += g) {
        SxTile t;
        t.prof = perm[pi];
        t.count = c - a < g ? c - a : g;
        t.first = e_off[pi] + a;
        t.out = t.first;
        tiles[w++] = t;
    }
}

// Target ids of a list of MSV-survivor indices.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_gather_t(const int64_t* __restrict__ n_dev, const int32_t* __restrict__ list, const int32_t* __restrict__ s_t,
             int32_t* tids) {
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (i < *n_dev) tids[i] = s_t[list[i]];
}

// Viterbi test (pipeline.rs:174-178) on the Viterbi scores of list entries.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_vit_pass(const int64_t* __restrict__ n_dev, const int32_t* __restrict__ list, const float* __restrict__ vsc,
             const float* __restrict__ s_bias, const int32_t* __restrict__ s_p, const float* __restrict__ thr,
             float* res_vit, int32_t* res_stage, uint8_t* flag) {
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (i >= *n_dev) return;
    const int32_t s1 = list[i];
    const float v = vsc[i];
    const float y = (v - s_bias[s1]) / SXC_LN2F;
    const bool pass = y >= thr[s_p[s1]];
    res_vit[s1] = v;
    res_stage[s1] = pass ? 3 : 2;
    flag[i] = pass ? 1 : 0;
}

// Candidates of the quantised Viterbi bound among the Viterbi stage's pairs (list entries
// i < *n_dev, survivor indices list[i]).  The Viterbi test (pipeline.rs:174-178) passes a
// pair only if fl(fl(v - bias) / ln2) >= thr, hence v >= bias + ln2 * thr - delta with
// delta = 2^-21 (|bias| + |ln2 * thr|); 1e-6 (1 + |bias| + |ln2 * thr|) exceeds it.  The
// threshold in the bound's units is capped at the target's saturation floor capv[t].
// Every entry is marked as having reached the Viterbi stage (res_stage 2); the f32 test
// of the candidates overwrites it.  Profiles with noqv set are candidates throughout.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_vitq_cand(const int64_t* __restrict__ n_dev, const int32_t* __restrict__ list, const int32_t* __restrict__ vq,
              const float* __restrict__ s_bias, const int32_t* __restrict__ s_p, const int32_t* __restrict__ s_t,
              const float* __restrict__ thr, const uint8_t* __restrict__ noqv, const int64_t* __restrict__ capv,
              int32_t* res_stage, uint8_t* flag) {
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (i >= *n_dev) return;
    const int32_t s1 = list[i];
    const int32_t p = s_p[s1];
    res_stage[s1] = 2;
    bool pass = true;
    if (!noqv[p]) {
        const double b = (double)s_bias[s1];
        const double ct = (double)SXC_LN2F * (double)thr[p];
        const double F = SX_Q_FINE;
        const int64_t th = (int64_t)floor(F * (b + ct) - F * 1e-6 * (1.0 + fabs(b) + fabs(ct)));
        const int64_t cap = capv[s_t[s1]];
        pass = (int64_t)vq[i] >= (th < cap ? th : cap);
    }
    flag[i] = pass ? 1 : 0;
}

// Forward test (pipeline.rs:179-183): exp(exp_logsurv(y)) <= F3 passes.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_fwd_pass(const int64_t* __restrict__ n_dev, const int32_t* __restrict__ list, const double* __restrict__ fsc,
             const float* __restrict__ s_bias, const int32_t* __restrict__ s_p, const float* __restrict__ thr,
             double* res_fwd, int32_t* res_stage) {
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (i >= *n_dev) return;
    const int32_t s1 = list[i];
    const double f = fsc[i];
    const float ff = (float)f;
    const float y = (ff - s_bias[s1]) / SXC_LN2F;
    res_fwd[s1] = f;
    res_stage[s1] = y >= thr[s_p[s1]] ? 4 : 3;
}

// Initialise the per-survivor result slots.
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_res_init(const int64_t* __restrict__ n_dev, int64_t cap, float* res_vit, double* res_fwd, int32_t* res_stage) {
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (i >= cap) return;
    res_vit[i] = __builtin_nanf("");
    res_fwd[i] = __builtin_nan("");
    res_stage[i] = 1;
}

// out = min(*n, cap): the entry count downstream kernels may read from a list whose
// true length *n may exceed its capacity (the host then grows the lists and runs again).
extern "C" __global__ void sxc_clamp(const int64_t* __restrict__ n, int64_t cap, int64_t* out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = *n < cap ? *n : cap;
}

// Zero an int32 array (graph-safe: no memset nodes).
extern "C" __global__ void __launch_bounds__(SXC_BLOCK) sxc_zero_i32(int32_t* a, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (i < n) a[i] = 0;
}

struct SxPairRec {
    uint32_t profile, target;
    float msv, bias, vit, fwd;
    double fwd64;
    int32_t stage, flags;
};

// One record per MSV survivor, in list order; flags marks the bias scores the host
// must recompute (sxc_bias).
extern "C" __global__ void __launch_bounds__(SXC_BLOCK)
sxc_records(const int64_t* __restrict__ n_dev, int64_t cap, const int32_t* __restrict__ s_p,
            const int32_t* __restrict__ s_t, const float* __restrict__ s_usc, const float* __restrict__ s_bias,
            const float* __restrict__ res_vit, const double* __restrict__ res_fwd,
            const int32_t* __restrict__ res_stage, const uint8_t* __restrict__ bflag, SxPairRec* out) {
    const int64_t i = (int64_t)blockIdx.x * SXC_BLOCK + threadIdx.x;
    if (i >= *n_dev || i >= cap) return;
    SxPairRec r;
    r.profile = (uint32_t)s_p[i];
    r.target = (uint32_t)s_t[i];
    r.msv = s_usc[i];
    r.bias = s_bias[i];
    r.vit = res_vit[i];
    r.fwd64 = res_fwd[i];
    r.fwd = (float)res_fwd[i];
    r.stage = res_stage[i];
    r.flags = bflag[i] ? SX_REC_BIAS_HOST : 0;
    out[i] = r;
}

// ---------------------------------------------------------------------------
// Posterior decoding and region finding next to Forward / Backward
// (pipeline.rs:186-242), one wave per pair.
// ---------------------------------------------------------------------------

// dp.rs:32-40 in f32
__device__ __forceinline__ float sxr_logsum(float a, float b) {
    const float hi = a > b ? a : b, lo = a > b ? b : a;
    const float d = hi - lo;
    if (!(d < 30.0f)) return hi;
    return hi + log1pf(expf(-d));
}

__device__ __forceinline__ float sxr_lnf(float x) { return x > 0.0f ? logf(x) : -__builtin_inff(); }

// For pair p: the scaled special values x5 (N B E J C, 5 floats a row) and row log scales
// ls of its Forward (f*) and Backward (b*) passes start at row in[p].roff (rows 0..L).
// The log special rows are lnf(x) + (f32)ls as dp.rs:284-292 and :372-390 store them
// (Forward row 0: N = 0, E = J = C = -inf; Backward row 0: E = J = C = -inf; row L:
// N = B = J = -inf).  With overall = the Backward score (the given Forward score if that
// is not finite, pipeline.rs:190), the posterior sums of pipeline.rs:194-206 and the
// region scan of :208-242 run over the rows; btot / etot are kept in bt / et (rows 0..L
// at roff) for the multi-domain test (is_multidomain, :527-536) of each region as it
// closes.  res[p] gets the Forward score of the rows, the Backward score, nexpected =
// btot[L] and the region count; each region goes to the list regs as (pair, i, j, multi)
// at a slot taken from *nreg_all (a pair's regions keep their order; slots past cap are
// counted, not written).
//
// One wave per pair.  A row's posterior terms (its B and E posteriors and the N/J/C
// occupancy, five logs and five exponentials) depend only on rows r-1 and r of the two
// passes, so the 64 lanes compute 64 rows' terms at once; the running sums, their stores
// and the region scan then take the rows in order, uniformly across the wave, with lane
// 0 storing btot / etot and reading them back for the multi-domain test (one thread's
// own stores, in program order).  Every value is the one-thread scan's: the same
// operations on the same operands, the sums in row order.
extern "C" __global__ void __launch_bounds__(64)
sxr_regions(int64_t n, const SxRegionIn* __restrict__ in, const float* __restrict__ fx5,
            const double* __restrict__ fls, const float* __restrict__ bx5, const double* __restrict__ bls,
            float* bt, float* et, SxRegionRes* __restrict__ res, SxRegion* __restrict__ regs, int32_t* nreg_all,
            int32_t cap) {
    const int64_t p = blockIdx.x;
    const int lane = (int)threadIdx.x;
    if (p >= n) return;
    const SxRegionIn a = in[p];
    const int L = a.len;
    const float NI = -__builtin_inff();
    SxRegionRes o;
    if (L == 0) {
        if (lane == 0) {
            o.fsc = o.bsc = NI;
            o.nexp = 0.0f;
            o.nreg = 0;
            res[p] = o;
        }
        return;
    }
    const float* F = fx5 + a.roff * 5;
    const float* B = bx5 + a.roff * 5;
    const double* FL = fls + a.roff;
    const double* BL = bls + a.roff;
    // the scores: Forward lnf(xC(L) * c_move) + ls(L) (dp.rs:293-297); Backward N(0)
    const float fs = (float)((double)sxr_lnf(F[(int64_t)L * 5 + 4] * a.pmove) + FL[L]);
    const float bn1 = L > 1 ? sxr_lnf(B[5 + 0]) + (float)BL[1] : NI;
    const float bb0 = sxr_lnf(B[1]) + (float)BL[0];
    const float bs = sxr_logsum(bn1 + a.lloop, bb0 + a.lmove);
    const float overall = isfinite(bs) ? bs : (a.has_fwd ? a.fwd : fs);
    float btot = 0.0f, etot = 0.0f;
    float* BT = bt + a.roff;
    float* ET = et + a.roff;
    if (lane == 0) {
        BT[0] = 0.0f;
        ET[0] = 0.0f;
    }
    const float rt1 = 0.25f, rt2 = 0.10f, rt3 = 0.20f;
    int i = -1, nr = 0;
    bool triggered = false;
    for (int base = 1; base <= L; base += 64) {
        // this lane's row r: its terms from rows r-1 and r
        const int r = base + lane;
        float tb = 0.0f, te = 0.0f, mocc = 0.0f;
        if (r <= L) {
            const float fl = (float)FL[r], bl = (float)BL[r];
            const float* fr = F + (int64_t)r * 5;
            const float* br = B + (int64_t)r * 5;
            const float bnr = r < L ? sxr_lnf(br[0]) + bl : NI;
            const float ber = sxr_lnf(br[2]) + bl;
            const float bjr = r < L ? sxr_lnf(br[3]) + bl : NI;
            const float bcr = sxr_lnf(br[4]) + bl;
            const float fer = sxr_lnf(fr[2]) + fl;
            // row r-1's Forward N B J C and Backward B (row 0: N = 0, J = C = -inf)
            float fn, fb, fj, fc, bb;
            if (r == 1) {
                fn = 0.0f;
                fb = sxr_lnf(F[1]) + (float)FL[0];
                fj = NI;
                fc = NI;
                bb = bb0;
            } else {
                const float fl1 = (float)FL[r - 1];
                const float* fp = fr - 5;
                fn = sxr_lnf(fp[0]) + fl1;
                fb = sxr_lnf(fp[1]) + fl1;
                fj = sxr_lnf(fp[3]) + fl1;
                fc = sxr_lnf(fp[4]) + fl1;
                bb = sxr_lnf(br[1 - 5]) + (float)BL[r - 1];
            }
            tb = expf(fb + bb - overall);
            te = expf(fer + ber - overall);
            const float njc = expf(fn + bnr + a.lloop - overall) + expf(fj + bjr + a.lloop - overall) +
                              expf(fc + bcr + a.lloop - overall);
            mocc = 1.0f - njc;
        }
        const int rows = L - base + 1 < 64 ? L - base + 1 : 64;
        for (int k = 0; k < rows; k++) {
            const int rr = base + k;
            const float tbk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tb), k));
            const float tek = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(te), k));
            const float mk = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mocc), k));
            const float bt_prev = btot, et_prev = etot;
            btot = btot + tbk;
            etot = etot + tek;
            if (lane == 0) {
                BT[rr] = btot;
                ET[rr] = etot;
            }
            if (!triggered) {
                if (mk - (btot - bt_prev) < rt2 || i == -1) i = rr;
                if (mk >= rt1) triggered = true;
            } else if (mk - (etot - et_prev) < rt2) {
                const int ri = i > 1 ? i : 1;
                if (lane == 0) {
                    float mx = -1.0f;
                    for (int jj = ri; jj <= rr; jj++) {
                        const float v = fminf(ET[jj] - ET[ri - 1], btot - BT[jj - 1]);
                        if (v > mx) mx = v;
                    }
                    const int32_t slot = atomicAdd(nreg_all, 1);
                    if (slot < cap) {
                        SxRegion g;
                        g.pair = (int32_t)p;
                        g.i = ri;
                        g.j = rr;
                        g.multi = mx >= rt3 ? 1 : 0;
                        regs[slot] = g;
                    }
                }
                nr++;
                i = -1;
                triggered = false;
            }
        }
    }
    if (lane == 0) {
        o.fsc = fs;
        o.bsc = bs;
        o.nexp = btot;
        o.nreg = nr;
        res[p] = o;
    }
}

// sx_kernels.hip -- MI355X (gfx950) kernels for the Plan7 filter cascade of
// sentinel-x src/hmmer: MSV (dp.rs:92-145), Viterbi score (dp.rs:152-189),
// scaled-probability Forward (dp.rs:229-298) and Backward (dp.rs:301-392).
//
// Decomposition.  A (profile, target) pair is owned by a group of G lanes of one
// wavefront (G = 1, 2, 4, ..., 64).  Lane j of the group keeps the model positions
// k = j*Q+1 .. j*Q+Q of the current DP row in registers (Q per lane, fully unrolled), so
// a group covers Q*G >= M positions; positions past M are padding that the tables make
// inert (-inf scores, zero probabilities).  Rows (residues) are processed in order; the
// row-global dependency xB(i) <- xE(i-1) is a group reduction per row (DPP butterflies
// inside 16-lane rows, v_readlane across rows), the k-1 neighbour at a lane boundary is
// one DPP shift per row, and the within-row D chain crossing lane boundaries is resolved
// exactly by lazy-F propagation (Viterbi, max-plus) or by an affine DPP scan
// (Forward/Backward, sum-product).  A workgroup stages one profile's emission and
// transition tables in LDS and runs 256/G pairs of that profile at once, one target per
// group.
//
// Exactness.  MSV and Viterbi use only f32 additions and maxima in the reference's
// per-path order, so every value is bit-identical to the serial loops: rounding is
// monotone, hence fl(max(a,b)+t) = max(fl(a+t), fl(b+t)), and a max may be taken in any
// order.  v_max_f32 can differ from the reference's select (a > b ? a : b) only in the
// sign of a zero result; no NaN or +inf can arise (all inputs are finite or -inf).
// Forward/Backward reorder sums (FMA, tree reductions), so their contract is an error
// bound against the FP64 oracle; the row log-scale accumulates in f64.
#include <hip/hip_runtime.h>

#include "sx_dev.h"

#define SX_NEG_INF (-__builtin_inff())
#define SX_LN2F 0.693147182464599609375f
#define SX_RESCALE 1.0e4f

namespace sxk {

// ---------------------------------------------------------------------------
// cross-lane helpers (wave64, gfx9 DPP)
// ---------------------------------------------------------------------------

template <int CTRL, int RM = 0xF, int BM = 0xF, bool BC = false>
__device__ __forceinline__ float dppf(float old, float src) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, RM, BM, BC));
}

// Full-permutation DPP read (every lane has a source lane): no old value, so the
// compiler can fold it into the consuming VALU instruction as a DPP operand.
template <int CTRL>
__device__ __forceinline__ float dppm(float src) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(src), CTRL, 0xF, 0xF, true));
}

__device__ __forceinline__ float rdlane(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

__device__ __forceinline__ int lane64() { return (int)(threadIdx.x & 63); }

// Hides a value's loop invariance from the optimiser.  Used on the lane's LDS column
// offset so that per-row transition reads stay LDS reads instead of being hoisted into
// 8*Q registers; kernels instantiated with HOIST keep the transitions in registers.
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}


struct OpMax {
    __device__ __forceinline__ float operator()(float a, float b) const { return fmaxf(a, b); }
};
struct OpAdd {
    __device__ __forceinline__ float operator()(float a, float b) const { return a + b; }
};

// Reduction over the G lanes of a group; every lane of the group receives the same value
// (each butterfly stage pairs lanes symmetrically and + / max are commutative).
template <int G, class Op>
__device__ __forceinline__ float grp_reduce(float v, Op op) {
    if constexpr (G >= 2) v = op(v, dppm<0xB1>(v));   // quad_perm [1,0,3,2]: xor 1
    if constexpr (G >= 4) v = op(v, dppm<0x4E>(v));   // quad_perm [2,3,0,1]: xor 2
    if constexpr (G >= 8) v = op(v, dppm<0x141>(v));  // row_half_mirror
    if constexpr (G >= 16) v = op(v, dppm<0x140>(v)); // row_mirror
    if constexpr (G == 32) {
        float a = rdlane(v, 0), b = rdlane(v, 16), c = rdlane(v, 32), d = rdlane(v, 48);
        v = lane64() < 32 ? op(a, b) : op(c, d);
    } else if constexpr (G == 64) {
        float a = rdlane(v, 0), b = rdlane(v, 16), c = rdlane(v, 32), d = rdlane(v, 48);
        v = op(op(a, b), op(c, d));
    }
    return v;
}

// Lane j of the group receives lane j-1's v; lane 0 receives b.
template <int G>
__device__ __forceinline__ float grp_from_prev(float v, float b) {
    if constexpr (G == 1) {
        return b;
    } else if constexpr (G <= 16) {
        float s = dppf<0x111>(b, v);  // row_shr:1, row lane 0 keeps b
        if constexpr (G < 16) s = ((lane64() & (G - 1)) == 0) ? b : s;
        return s;
    } else {
        float s = dppf<0x138>(b, v);  // wave_shr:1, lane 0 keeps b
        if constexpr (G == 32) s = ((lane64() & 31) == 0) ? b : s;
        return s;
    }
}

// grp_from_prev(v, 0): the shifts read with bound_ctrl, so an invalid source lane reads
// 0 and no old value has to be set first.
template <int G>
__device__ __forceinline__ float grp_from_prev0(float v) {
    if constexpr (G == 1) {
        return 0.0f;
    } else if constexpr (G <= 16) {
        float s = dppf<0x111, 0xF, 0xF, true>(0.0f, v);  // row_shr:1
        if constexpr (G < 16) s = ((lane64() & (G - 1)) == 0) ? 0.0f : s;
        return s;
    } else {
        float s = dppf<0x138, 0xF, 0xF, true>(0.0f, v);  // wave_shr:1
        if constexpr (G == 32) s = ((lane64() & 31) == 0) ? 0.0f : s;
        return s;
    }
}

// Lane j of the group receives lane j+1's v; lane G-1 receives b.
template <int G>
__device__ __forceinline__ float grp_from_next(float v, float b) {
    if constexpr (G == 1) {
        return b;
    } else if constexpr (G <= 16) {
        float s = dppf<0x101>(b, v);  // row_shl:1, row lane 15 keeps b
        if constexpr (G < 16) s = ((lane64() & (G - 1)) == G - 1) ? b : s;
        return s;
    } else {
        float s = dppf<0x130>(b, v);  // wave_shl:1, lane 63 keeps b
        if constexpr (G == 32) s = ((lane64() & 31) == 31) ? b : s;
        return s;
    }
}

template <int S>
struct RowShr;
template <>
struct RowShr<1> { static constexpr int c = 0x111; };
template <>
struct RowShr<2> { static constexpr int c = 0x112; };
template <>
struct RowShr<4> { static constexpr int c = 0x114; };
template <>
struct RowShr<8> { static constexpr int c = 0x118; };
template <int S>
struct RowShl;
template <>
struct RowShl<1> { static constexpr int c = 0x101; };
template <>
struct RowShl<2> { static constexpr int c = 0x102; };
template <>
struct RowShl<4> { static constexpr int c = 0x104; };
template <>
struct RowShl<8> { static constexpr int c = 0x108; };

// One step of an inclusive (Hillis-Steele) scan of affine maps D -> A + B*D over the
// lanes of a group; lanes with no partner S lanes away combine with the identity (0, 1).
// A's partner value is read with bound_ctrl (an invalid source lane reads 0, the identity's
// A), so no old value has to be set first.  Within a 16-lane row (G <= 16) a lane's B only
// matters while its composite has not yet reached the group's first lane: once it has, its
// partner A is 0 from then on and its B multiplies only the B of later lanes' composites,
// which have reached it too.  So B's partner value is read with bound_ctrl as well, and
// lanes of the row's other groups are not masked out of it.  Wider groups combine the
// rows' total B values, so there B keeps the identity 1.
template <int G, int S, bool FROM_LEFT>
__device__ __forceinline__ void scan_step(float& A, float& B) {
    if constexpr (G > S && S < 16) {
        constexpr int C = FROM_LEFT ? RowShr<S>::c : RowShl<S>::c;
        float a2 = dppf<C, 0xF, 0xF, true>(0.0f, A);
        const float b2 = G <= 16 ? dppf<C, 0xF, 0xF, true>(0.0f, B) : dppf<C>(1.0f, B);
        if constexpr (G < 16) {
            const int jg = lane64() & (G - 1);
            const bool ok = FROM_LEFT ? (jg >= S) : (jg + S < G);
            a2 = ok ? a2 : 0.0f;
        }
        A = fmaf(B, a2, A);
        B = B * b2;
    }
}

// Inclusive scan over the group of the affine maps (A, B).  FROM_LEFT: lane j ends with
// the composition of lanes 0..j applied to 0; otherwise lanes j..G-1 applied to 0.  Only
// A is meaningful afterwards.
template <int G, bool FROM_LEFT>
__device__ __forceinline__ void grp_affine_scan(float& A, float B) {
    scan_step<G, 1, FROM_LEFT>(A, B);
    scan_step<G, 2, FROM_LEFT>(A, B);
    scan_step<G, 4, FROM_LEFT>(A, B);
    scan_step<G, 8, FROM_LEFT>(A, B);
    if constexpr (G >= 32) {
        const int row = lane64() >> 4;
        if constexpr (FROM_LEFT) {
            // row totals sit in lane 15 of each 16-lane row
            float a0 = rdlane(A, 15), a1 = rdlane(A, 31), b1 = rdlane(B, 31), a2 = rdlane(A, 47), b2 = rdlane(B, 47);
            float e;
            if constexpr (G == 64) {
                float e1 = a0;
                float e2 = fmaf(b1, e1, a1);
                float e3 = fmaf(b2, e2, a2);
                e = row == 0 ? 0.0f : row == 1 ? e1 : row == 2 ? e2 : e3;
            } else {
                e = row == 1 ? a0 : row == 3 ? a2 : 0.0f;
            }
            A = fmaf(B, e, A);
        } else {
            // suffix totals sit in lane 0 of each row
            float a1 = rdlane(A, 16), b1 = rdlane(B, 16), a2 = rdlane(A, 32), b2 = rdlane(B, 32), a3 = rdlane(A, 48);
            float e;
            if constexpr (G == 64) {
                float e2 = a3;
                float e1 = fmaf(b2, e2, a2);
                float e0 = fmaf(b1, e1, a1);
                e = row == 0 ? e0 : row == 1 ? e1 : row == 2 ? e2 : 0.0f;
            } else {
                e = row == 0 ? a1 : row == 2 ? a3 : 0.0f;
            }
            A = fmaf(B, e, A);
        }
    }
}

// ---------------------------------------------------------------------------
// residue streams: dword loads with a two-dword prefetch ring; a residue is one byte
// taken from the current dword by a shift (no run-time selection among vector words:
// hipcc for gfx950 miscompiled a four-way select over the components of a uint4 here,
// returning the first component for the fourth).  Targets start 16-byte aligned and the
// code buffer is padded by 64 bytes on both ends, so the prefetch may read past a
// target's first or last dword.
// ---------------------------------------------------------------------------

struct FwdStream {
    const uint32_t* p;
    uint32_t cur, n1, n2;
    int r;
    __device__ __forceinline__ void init(const uint8_t* base) {
        p = reinterpret_cast<const uint32_t*>(base);
        cur = p[0];
        n1 = p[1];
        n2 = p[2];
        r = 0;
    }
    __device__ __forceinline__ int next() {
        const int x = (int)((cur >> (r * 8)) & 0xFFu);
        if (++r == 4) {
            r = 0;
            cur = n1;
            n1 = n2;
            n2 = p[3];
            ++p;
        }
        return x;
    }
};

// Reads positions L-1, L-2, ..., 0 of a target.
struct RevStream {
    const uint32_t* p;
    uint32_t cur, n1, n2;
    int r;
    __device__ __forceinline__ void init(const uint8_t* base, int L) {
        const int last = L - 1;
        p = reinterpret_cast<const uint32_t*>(base) + (last >> 2);
        cur = p[0];
        n1 = p[-1];
        n2 = p[-2];
        r = last & 3;
    }
    __device__ __forceinline__ int next() {
        const int x = (int)((cur >> (r * 8)) & 0xFFu);
        if (--r < 0) {
            r = 3;
            cur = n1;
            n1 = n2;
            n2 = p[-3];
            --p;
        }
        return x;
    }
};

// LDS layout of a profile table row for a (Q, G) shape.  Lane j of a group owns the
// Q+1 slots [j*(Q+1), (j+1)*(Q+1)): slot s holds model position k = j*Q + s, so slot 0
// duplicates the previous lane's last position (the k-1 operand of the lane's first
// cell) and slots 1..Q are the lane's own positions.  The odd stride Q+1 puts the lanes
// of a group on distinct banks when they read the same slot (a stride of Q, always even
// here, made them collide).  Two extra slots let the last lane read position k+1, and
// the row width is odd so that groups reading different residue rows spread over banks.
template <int Q, int G>
struct Geo {
    static constexpr int CAP = Q * G;
    static constexpr int STRIDE = Q + 1;
    static constexpr int WID = ((G * STRIDE + 2) & 1) ? (G * STRIDE + 2) : (G * STRIDE + 3);
};

// Model position stored at LDS slot `pos` of a row.
template <int Q, int G>
__device__ __forceinline__ int slot_k(int pos) {
    const int j = pos / Geo<Q, G>::STRIDE;
    return j * Q + (pos - j * Geo<Q, G>::STRIDE);
}

// Table modes of the Viterbi/Forward/Backward kernels (MSV has no transition rows):
//   TM_LDS  emission and transition rows staged in LDS, transitions re-read every row;
//   TM_REG  as TM_LDS, but the lane's transitions stay in registers across rows;
//   TM_GEM  long models: emission rows read from a padded copy in global memory
//           (through L1/L2), transitions staged in LDS and re-read every row.
// and two modes of MSV and Viterbi for groups of G = 8 or G >= 16 lanes:
//   TM_ALN  profile rows in LDS as 16-byte-aligned lane blocks (GeoA), read four
//           positions per load; 512-thread workgroups share one staged table;
//   TM_ALN1K  as TM_ALN with 1024-thread workgroups.
enum { TM_LDS = 0, TM_REG = 1, TM_GEM = 2, TM_ALN = 3, TM_ALN1K = 4 };

template <int TM>
constexpr bool tm_aligned() {
    return TM == TM_ALN || TM == TM_ALN1K;
}
// threads per workgroup of a table mode
template <int TM>
constexpr int tm_block() {
    return TM == TM_ALN ? 512 : TM == TM_ALN1K ? 1024 : 256;
}

// Aligned LDS layout (TM_ALN*): lane j of a group owns the STRIDE slots
// [j*STRIDE, (j+1)*STRIDE) of a row, slot s < Q holding model position k = j*Q + 1 + s
// (the rest is padding, never read).  A 128-bit LDS read serves 16 lanes per cycle
// from 64 four-byte banks, so its 16 lanes must start on 16 distinct multiples of
// four banks.  With STRIDE/4 odd, the lanes j..j+15 of one group do (j*STRIDE mod 64
// runs over all 16 slots); a group of G >= 16 lanes fills the 16 lanes with one residue
// row, so every read is conflict-free whatever the residues of the other groups.
// Groups of 8 lanes share a quarter-wave in pairs that read different rows: the table
// is then kept twice, the second copy 32 banks (SKEW words) further round, rows are
// padded to a multiple of 64 floats so the row index moves no bank, and the odd group
// of each pair reads the second copy.  The 8 slots {j*STRIDE mod 64 : j < 8} and the
// same set turned 32 banks are disjoint (j - j' would have to be 8 mod 16).  Groups of 4
// lanes share a quarter-wave in fours: the table is kept four times, each copy 16 banks
// further round, with STRIDE/4 = 1 (mod 4), so lane j of copy c starts on slot
// 4c + j*STRIDE/4 (mod 16): j is that slot mod 4 and then c the rest, 16 distinct slots.
template <int Q, int G>
struct GeoA {
    static_assert(Q % 4 == 0 && (G == 4 || G == 8 || G >= 16), "aligned lane blocks need Q % 4 == 0 and G = 4, 8 or >= 16");
    static constexpr int STRIDE = G == 4 ? Q + ((4 - Q % 16) + 16) % 16 : (Q % 8 == 0) ? Q + 4 : Q + 8;
    static constexpr int COPIES = G >= 16 ? 1 : G == 8 ? 2 : 4;
    static constexpr int WID = COPIES == 1 ? G * STRIDE : (G * STRIDE + 63) / 64 * 64;
    static constexpr int SKEW = COPIES == 1 ? 0 : 64 / COPIES;
    static_assert(G != 4 || STRIDE % 16 == 4, "four-lane stride");
    static constexpr int COPY = SX_NCODE * WID + SKEW;        // floats from one copy to the next
    static constexpr int TAB = COPIES * SX_NCODE * WID + (COPIES - 1) * SKEW;  // emission floats staged
    // Viterbi's eight transition rows follow the emission rows, one copy (every group of
    // a quarter-wave reads the same addresses of them: a broadcast), TWID floats apart.
    static constexpr int TWID = G * STRIDE;
};

// LDS row width of a shape in a table mode
template <int Q, int G, int TM>
constexpr int tm_wid() {
    if constexpr (tm_aligned<TM>())
        return GeoA<Q, G>::WID;
    else
        return Geo<Q, G>::WID;
}

// Padded global emission rows: position k of residue row x sits at x * ew + k + 3, so a
// lane's block of Q positions (Q a multiple of 4) starts 16-byte aligned; positions 0
// and beyond M hold the pad value (-inf scores, 0 odds).
constexpr int GEM_SHIFT = 3;

// Emission row access for one lane: row(x)[q] is the lane's q-th position of residue
// row x; row(x)[NEXT] is the first position of the next lane of the group.
template <int Q, int G, int TM>
struct EmRows {
    static constexpr int NEXT = TM == TM_GEM ? Q : Q + 1;
    const float* base;
    int width;
    __device__ __forceinline__ EmRows(const float* em, int ew, int j) {
        if constexpr (TM == TM_GEM) {
            base = em + j * Q + 1 + GEM_SHIFT;
            width = ew;
        } else if constexpr (tm_aligned<TM>()) {
            // the groups sharing a quarter-wave read different copies
            const int c = (int)((threadIdx.x / G) & (GeoA<Q, G>::COPIES - 1));
            base = em + c * GeoA<Q, G>::COPY + j * GeoA<Q, G>::STRIDE;
            width = GeoA<Q, G>::WID;
        } else {
            base = em + j * Geo<Q, G>::STRIDE + 1;
            width = Geo<Q, G>::WID;
        }
    }
    __device__ __forceinline__ const float* row(int x) const {
        if constexpr (TM == TM_GEM)
            return static_cast<const float*>(__builtin_assume_aligned(base + x * width, 16));
        else if constexpr (tm_aligned<TM>())
            return static_cast<const float*>(__builtin_assume_aligned(base + x * GeoA<Q, G>::WID, 16));
        else
            return base + x * Geo<Q, G>::WID;
    }
};

// Forward rows two at a time (one loop iteration) for the shapes where that measured
// faster (5-10%; others by as much slower, or spilling).  In a one-row loop the compiler
// copies the row's registers at every iteration's end (the values of a row land in other
// registers than the previous row's: about 130 of 500 instructions a row of 24 positions
// of the long-model form); over two rows it keeps them in place.
template <int Q, int TM>
constexpr bool fwd_rows2() {
    return (TM == TM_GEM && (Q == 12 || Q == 24)) || (TM == TM_LDS && Q <= 32) || (TM == TM_REG && Q <= 8);
}

// Forward's transition records: for cell q of lane j (model position k = j*Q + 1 + q) the
// eight transition probabilities it reads, in one 32-byte record
//   {MM(k-1), IM(k-1), DM(k-1), BM(k)} {MI(k), II(k), MD(k), DD(k)},
// so a cell's operands are two ds_read_b128 at immediate offsets from the lane's base
// (positions 0 and past M hold 0).  Lane blocks are 8Q + 4 floats apart: an odd number of
// 16-byte slots, so the 16 lanes of a quarter-wave that read different positions start on
// 16 distinct four-bank slots; groups of a workgroup read the same addresses (the records
// do not depend on the residue), which the LDS broadcasts.  After the records, each lane
// block's D-chain products: PC[q] = DD(k0) ... DD(k0+q-1) for q = 0..Q (PC[0] = 1, PC[Q]
// the block's slope) and their sum PC[0] + ... + PC[Q-1], in blocks of PSTRIDE floats (an
// odd number of 16-byte slots).
template <int Q, int G>
struct FwdRec {
    static constexpr int RSTRIDE = 8 * Q + 4;
    static constexpr int PSTRIDE = ((Q + 4) / 4) % 2 ? Q + 4 : Q + 8;
    static constexpr int REC = G * RSTRIDE;
    static constexpr int TAB = REC + G * PSTRIDE;
};
constexpr int round4(int n) { return (n + 3) & ~3; }

// ---------------------------------------------------------------------------
// MSV (dp.rs:92-145)
// ---------------------------------------------------------------------------
template <int Q, int G, int TM>
__device__ __forceinline__ float msv_pair(const float* __restrict__ em, int ew, float tbm, const uint8_t* seq, int L,
                                          float tloop, float tmove) {
    const int j = (int)(threadIdx.x & (G - 1));
    const EmRows<Q, G, TM> er(em, ew, j);
    float row[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) row[q] = SX_NEG_INF;
    // dp.rs:102 gives E->J and E->C the same score (-ln 2) and J, C start at -inf, so
    // xJ and xC follow one recurrence and are equal on every row: xC is not kept.
    // xB = max(xN + tmove, xJ + tmove) is computed as max(xN, xJ) + tmove, which is the
    // same f32 value because rounding is monotone.
    float xn = 0.0f, xb = tmove, xj = SX_NEG_INF;
    auto step = [&](int x) {
        const float* __restrict__ sc = er.row(x);
        const float be = xb + tbm;
        const float left = grp_from_prev<G>(row[Q - 1], SX_NEG_INF);
        float e = SX_NEG_INF;
#pragma unroll
        for (int q = Q - 1; q >= 1; q--) {
            const float v = sc[q] + fmaxf(row[q - 1], be);
            row[q] = v;
            e = fmaxf(e, v);
        }
        {
            const float v = sc[0] + fmaxf(left, be);
            row[0] = v;
            e = fmaxf(e, v);
        }
        const float xe = grp_reduce<G>(e, OpMax());
        xj = fmaxf(xj + tloop, xe - SX_LN2F);
        xn = xn + tloop;
        xb = fmaxf(xn, xj) + tmove;
    };
    // Residues come one dword (four rows) at a time with the next dword loaded a dword
    // ahead; targets start 16-byte aligned and the code buffer has 64 bytes of padding
    // after the last one, so the look-ahead stays inside it.
    const uint32_t* wp = reinterpret_cast<const uint32_t*>(seq);
    if constexpr (Q < 96) {
        // four rows per loop trip: the byte position is fixed in each copy of the step
        uint32_t nxt = wp[0];
        for (int i = 0;;) {
            const uint32_t w = nxt;
            nxt = wp[1];
            ++wp;
            step((int)(w & 0xFFu));
            if (++i == L) break;
            step((int)((w >> 8) & 0xFFu));
            if (++i == L) break;
            step((int)((w >> 16) & 0xFFu));
            if (++i == L) break;
            step((int)(w >> 24));
            if (++i == L) break;
        }
    } else {
        // one row per trip: four copies of a wide row would not fit the register file
        FwdStream rs;
        rs.init(seq);
        for (int i = 0; i < L; i++) step(rs.next());
    }
    return xj + tmove;
}

// ---------------------------------------------------------------------------
// Quantised MSV upper bound: the int16 prefilter of the MSV stage.  No counterpart in
// the reference, whose MSV is the f32 recurrence above; this one bounds it from above.
// ---------------------------------------------------------------------------
// Two model positions per 32-bit register, as int16 halves: lane j's register r holds
// slot j*Q + 1 + r in the low half and j*Q + 1 + Q/2 + r in the high half, so one
// packed instruction advances two cells and the k-1 operand of every half is the same
// half of register r-1 (register 0 takes the previous lane's last high half and its own
// last low half, one alignbit).
//
// Frame.  A cell holds t = s - F - 32768, s its score in 1/1024 nat and F the pair's
// frame floor, which is never below the row's B entry: t = -32768 stands for F.  The
// recurrence max(s[k-1], B) + e[k] then needs no max with B: every t is >= -32768, so
// s[k-1] >= F >= B, and the saturating add t[k-1] + e[k] (v_pk_add_i16 clamp) only
// raises a sum below -32768 to F.  One add and one max (the row maximum) per two
// cells.  F follows B: it moves up to B on any row whose B exceeds it (rare: xJ rising),
// and down to B on every 16th row (B falls with xN by |tloop| a row), so F - B stays
// below 16 rows of that drift.  A move by d shifts every t by -d (saturating; above
// 32767 in two steps).  Slots before position 1 hold -32768 scores: after each row they
// hold exactly -32768, so position 1 sees F, or after a move the shifted floor, which
// the value injected on the left of the group's first slot reproduces: the result does
// not depend on the shape.  Scores are rounded up on the host (q_emit, q_const) with an
// allowance for the f32 rounding of each step, so by induction every cell, the row
// maximum xE, xJ, xN, xB and the result are at least 2^22 (fine units) or 1024 (cells)
// times their f32 counterparts of dp.rs:92-145, until a cell reaches 32767; from then
// on the result stays above a floor the host knows.
typedef short sx_v2i16 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t pk_max16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(sx_v2i16, a),
                                                                  __builtin_bit_cast(sx_v2i16, b)));
}
__device__ __forceinline__ uint32_t pk_adds16(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_add_sat(__builtin_bit_cast(sx_v2i16, a),
                                                                      __builtin_bit_cast(sx_v2i16, b)));
}

// Lane j of the group receives lane j-1's v; lane 0 receives b (bitwise, as grp_from_prev).
template <int G>
__device__ __forceinline__ uint32_t grp_from_prev_u(uint32_t v, uint32_t b) {
    if constexpr (G == 1) {
        return b;
    } else if constexpr (G <= 16) {
        uint32_t s = (uint32_t)__builtin_amdgcn_update_dpp((int)b, (int)v, 0x111, 0xF, 0xF, false);
        if constexpr (G < 16) s = ((lane64() & (G - 1)) == 0) ? b : s;
        return s;
    } else {
        uint32_t s = (uint32_t)__builtin_amdgcn_update_dpp((int)b, (int)v, 0x138, 0xF, 0xF, false);
        if constexpr (G == 32) s = ((lane64() & 31) == 0) ? b : s;
        return s;
    }
}

template <int CTRL>
__device__ __forceinline__ int dppm_i(int src) {
    return __builtin_amdgcn_mov_dpp(src, CTRL, 0xF, 0xF, true);
}

// Integer maximum over the G lanes of a group (the butterfly of grp_reduce).
template <int G>
__device__ __forceinline__ int grp_max_i(int v) {
    if constexpr (G >= 2) v = max(v, dppm_i<0xB1>(v));
    if constexpr (G >= 4) v = max(v, dppm_i<0x4E>(v));
    if constexpr (G >= 8) v = max(v, dppm_i<0x141>(v));
    if constexpr (G >= 16) v = max(v, dppm_i<0x140>(v));
    if constexpr (G == 32) {
        const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
        const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
        v = lane64() < 32 ? max(a, b) : max(c, d);
    } else if constexpr (G == 64) {
        const int a = __builtin_amdgcn_readlane(v, 0), b = __builtin_amdgcn_readlane(v, 16);
        const int c = __builtin_amdgcn_readlane(v, 32), d = __builtin_amdgcn_readlane(v, 48);
        v = max(max(a, b), max(c, d));
    }
    return v;
}

// LDS layout of the packed int16 table (dwords).  One-lane groups: residue row x is
// WID dwords, dword r packing positions r+1 and Q/2+r+1; WID is odd, so the 64 lanes of a
// wave, each on its own target, read distinct banks unless they read the same address.
// Groups of 8 or >= 16 lanes: the aligned lane blocks of GeoA over Q/2 dwords per lane.
template <int Q>
struct GeoQ1 {
    static constexpr int H = Q / 2;
    static constexpr int WID = (H & 1) ? H : H + 1;
};

template <int Q, int G, int TM>
constexpr int q_tab_dwords() {
    if constexpr (tm_aligned<TM>())
        return GeoA<Q / 2, G>::TAB;
    else
        return SX_NCODE * GeoQ1<Q>::WID;
}

// Residue words of a target, word k at wp[k * ws]: ws = 1 for the packed codes, 64 for the
// interleaved copy (SxTargets::ilv).
struct WordStream {
    const uint32_t* p;
    int ws;
    uint32_t cur, n1, n2;
    int r;
    __device__ __forceinline__ void init(const uint32_t* base, int stride) {
        p = base;
        ws = stride;
        cur = p[0];
        n1 = p[ws];
        n2 = p[2 * ws];
        r = 0;
    }
    __device__ __forceinline__ int next() {
        const int x = (int)((cur >> (r * 8)) & 0xFFu);
        if (++r == 4) {
            r = 0;
            cur = n1;
            n1 = n2;
            n2 = p[3 * ws];
            p += ws;
        }
        return x;
    }
};

template <int Q, int G, int TM>
__device__ __forceinline__ int32_t msvq_pair(const uint32_t* __restrict__ tab, int tbm, int tej,
                                             const uint32_t* __restrict__ words, int ws, int L, int tloop, int tmove) {
    static_assert(Q % 2 == 0, "two positions per register");
    static_assert(G == 1 || tm_aligned<TM>(), "the int16 table is laid out for one-lane or aligned groups");
    constexpr int H = Q / 2;
    const int j = (int)(threadIdx.x & (G - 1));
    const uint32_t* base;
    int width;
    if constexpr (tm_aligned<TM>()) {
        using GA = GeoA<H, G>;
        const int c = (int)((threadIdx.x / G) & (GA::COPIES - 1));
        base = tab + c * GA::COPY + j * GA::STRIDE;
        width = GA::WID;
    } else {
        base = tab;
        width = GeoQ1<Q>::WID;
    }
    uint32_t row[H];
#pragma unroll
    for (int r = 0; r < H; r++) row[r] = 0x80008000u;
    // As msv_pair: xC = xJ on every row and xB = max(xN, xJ) + tmove.  The B entry of a
    // row is ceil((xB + tbm) / 4096) = max(xn', xj') >> 12 with the offsets folded in:
    // xn' = xN + c, xj' = xJ + c, c = tmove + tbm + 4095.
    const int c = tmove + tbm + ((1 << SX_Q_SHIFT) - 1);
    const int ce = tej + c;
    int xn = c, xj = -(1 << 30);
    int be = c >> SX_Q_SHIFT;
    int fl = be;  // the frame floor F
    // sched: a row on which the frame may move down (every 16th)
    auto step = [&](int x, bool sched) {
        const uint32_t* __restrict__ sc = base + x * width;
        if constexpr (tm_aligned<TM>()) sc = static_cast<const uint32_t*>(__builtin_assume_aligned(sc, 16));
        uint32_t inj = 0x80000000u;  // the floor, in the high half
        const bool mv = be > fl || (sched && be < fl);
        if (__any(mv)) {
            // d < 65535: be < 32767 (xE is capped there) and F > -32768
            const int d = mv ? be - fl : 0;
            fl += d;
            const int d1 = min(d, 32767);
            const uint32_t n1 = __builtin_amdgcn_perm((uint32_t)(-d1), (uint32_t)(-d1), 0x05040504u);
#pragma unroll
            for (int r = 0; r < H; r++) row[r] = pk_adds16(row[r], n1);
            if (__any(d > 32767)) {
                const int d2 = d - d1;
                const uint32_t n2 = __builtin_amdgcn_perm((uint32_t)(-d2), (uint32_t)(-d2), 0x05040504u);
#pragma unroll
                for (int r = 0; r < H; r++) row[r] = pk_adds16(row[r], n2);
            }
            inj = (uint32_t)max(-32768 - d, -32768) << 16;  // -d < 32768: no upper clamp
        }
        const uint32_t pl = grp_from_prev_u<G>(row[H - 1], inj);
        const uint32_t left = __builtin_amdgcn_alignbit(row[H - 1], pl, 16);
        // the row maximum in NACC independent chains: one chain serialises every
        // register's max behind the previous one, four cost more merging than they hide
        // (measured: 2 chains 1% faster than 4 and 1.5% than 1)
        constexpr int NACC = H >= 8 ? 2 : 1;
        uint32_t ea[NACC];
#pragma unroll
        for (int a = 0; a < NACC; a++) ea[a] = 0x80008000u;
#pragma unroll
        for (int r = H - 1; r >= 1; r--) {
            const uint32_t v = pk_adds16(row[r - 1], sc[r]);
            row[r] = v;
            ea[r % NACC] = pk_max16(ea[r % NACC], v);
        }
        {
            const uint32_t v = pk_adds16(left, sc[0]);
            row[0] = v;
            ea[0] = pk_max16(ea[0], v);
        }
        uint32_t e = ea[0];
#pragma unroll
        for (int a = 1; a < NACC; a++) e = pk_max16(e, ea[a]);
        e = pk_max16(e, __builtin_amdgcn_alignbit(e, e, 16));  // both halves: the lane's maximum
        int xe = (int)(short)(e & 0xFFFFu);
        // capped at 32767 as the absolute cells of a frameless bound are: the floor of a
        // saturated bound is then the same (F >= -29.7 nats > -32 for L <= 10^6, M <= 1024,
        // so a cell at 32767 reaches the cap)
        xe = min(grp_max_i<G>(xe) + fl + 32768, 32767);
        xj = max(xj + tloop, xe * (1 << SX_Q_SHIFT) + ce);
        xn = xn + tloop;
        be = max(xn, xj) >> SX_Q_SHIFT;
    };
    // Four rows per loaded word: the compiler then schedules a row's LDS reads ahead into
    // the previous row, which pays on the aligned lane blocks and short one-lane rows.  On
    // one-lane rows of 96 positions or more it costs occupancy (190 VGPRs against 99 at
    // Q = 128) and the residue stream is faster (measured 3-15%).
    const uint32_t* wp = words;
    if constexpr (G > 1 || H < 48) {
        uint32_t nxt = wp[0];
        for (int i = 0;;) {
            const uint32_t w = nxt;
            nxt = wp[ws];
            wp += ws;
            step((int)(w & 0xFFu), (i & 15) == 0);
            if (++i == L) break;
            step((int)((w >> 8) & 0xFFu), false);
            if (++i == L) break;
            step((int)((w >> 16) & 0xFFu), false);
            if (++i == L) break;
            step((int)(w >> 24), false);
            if (++i == L) break;
        }
    } else {
        WordStream rs;
        rs.init(words, ws);
        for (int i = 0; i < L; i++) step(rs.next(), (i & 15) == 0);
    }
    return xj - tbm - ((1 << SX_Q_SHIFT) - 1);  // xJ + tmove
}

// ---------------------------------------------------------------------------
// Quantised Viterbi upper bound: the int16 prefilter of the Viterbi stage (no
// counterpart in the reference, whose Viterbi is vit_pair's f32 recurrence).  The units,
// roundings, packing and folded special states of msvq_pair, with:
//   - transition rows in LDS after the emission rows, staged shifted so that every
//     operand of a register sits in the same register slot (MM IM DM MD DD read at k-1,
//     MI II BM at k); out of node 0 and in the padding slots every score is -32768, so
//     padding cells stay exactly at -32768;
//   - insert states not masked (the reference sets them -inf at k = M and on '*'), which
//     only raises values;
//   - the D chain by lazy-F: a first pass from M alone, then the value entering each half
//     block (the previous lane's high half, or the lane's own low half) is propagated
//     until nothing changes, at most 2G - 1 times;
//   - the row maximum over M and D (rounding up can make a D->D score of -1e-6 nat +1
//     unit, so M alone is no longer an upper bound of D).
// Cells hold absolute scores (no frame: B enters through the BM max).  Transition scores
// <= 0 are rounded to <= 0 (q_trans) and the host skips profiles with a positive insert
// score, so an insert cell reaches 32767 only after an M cell has: the result then stays
// above the saturation floor of msvq_pair.
template <int Q, int G>
struct GeoQV1 {  // one-lane groups: transition rows 16-byte aligned after the odd emission rows
    static constexpr int H = Q / 2;
    static constexpr int TBASE = (SX_NCODE * GeoQ1<Q>::WID + 3) & ~3;
    static constexpr int TW = (H + 3) & ~3;
};

template <int Q, int G, int TM>
constexpr int vq_tab_dwords() {
    if constexpr (tm_aligned<TM>())
        return GeoA<Q / 2, G>::TAB + SX_NT * GeoA<Q / 2, G>::TWID;
    else
        return GeoQV1<Q, G>::TBASE + SX_NT * GeoQV1<Q, G>::TW;
}

template <int Q, int G, int TM>
__device__ __forceinline__ int32_t vitq_pair(const uint32_t* __restrict__ tab, int tej, const uint8_t* seq, int L,
                                             int qloop, int qmove) {
    static_assert(Q % 2 == 0, "two positions per register");
    static_assert(G == 1 || tm_aligned<TM>(), "the int16 tables are laid out for one-lane or aligned groups");
    constexpr int H = Q / 2;
    // transitions are re-read from LDS every row: keeping them in registers (8 x H) cost
    // more occupancy than the reads (129 VGPRs at Q = 16: three waves per SIMD)
    constexpr bool HOIST = false;
    const int j = (int)(threadIdx.x & (G - 1));
    const uint32_t* base;
    int width, tofs, twid;
    if constexpr (tm_aligned<TM>()) {
        using GA = GeoA<H, G>;
        const int c = (int)((threadIdx.x / G) & (GA::COPIES - 1));
        base = tab + c * GA::COPY + j * GA::STRIDE;
        width = GA::WID;
        tofs = GA::TAB + j * GA::STRIDE;
        twid = GA::TWID;
    } else {
        base = tab;
        width = GeoQ1<Q>::WID;
        tofs = GeoQV1<Q, G>::TBASE;
        twid = GeoQV1<Q, G>::TW;
    }
    uint32_t mr[H], ir[H], dr[H];
#pragma unroll
    for (int r = 0; r < H; r++) mr[r] = ir[r] = dr[r] = 0x80008000u;
    // B entry in cell units: max(xn', xj') >> 12 with xn' = xN + c, xj' = xJ + c, c = move + 4095
    const int c0 = qmove + ((1 << SX_Q_SHIFT) - 1);
    const int ce = tej + c0;
    int xn = c0, xj = -(1 << 30);
    int xbq = max(c0 >> SX_Q_SHIFT, -32768);
    auto step = [&](int x) {
        const uint32_t* __restrict__ sc = base + x * width;
        // larger blocks re-read the transitions every row (hidden from the optimiser so
        // that the loads are not hoisted into registers)
        const uint32_t* __restrict__ tr = tab + (HOIST ? tofs : opaque(tofs));
        if constexpr (tm_aligned<TM>()) {
            sc = static_cast<const uint32_t*>(__builtin_assume_aligned(sc, 16));
        }
        tr = static_cast<const uint32_t*>(__builtin_assume_aligned(tr, 16));
        const uint32_t* __restrict__ tmm = tr + SX_TMM * twid;
        const uint32_t* __restrict__ tmi = tr + SX_TMI * twid;
        const uint32_t* __restrict__ tmd = tr + SX_TMD * twid;
        const uint32_t* __restrict__ tim = tr + SX_TIM * twid;
        const uint32_t* __restrict__ tii = tr + SX_TII * twid;
        const uint32_t* __restrict__ tdm = tr + SX_TDM * twid;
        const uint32_t* __restrict__ tdd = tr + SX_TDD * twid;
        const uint32_t* __restrict__ tbm = tr + SX_TBM * twid;
        const uint32_t xb2 = __builtin_amdgcn_perm((uint32_t)xbq, (uint32_t)xbq, 0x05040504u);
        // k-1 operands of register 0: the previous lane's last high half, the lane's own
        // last low half
        const uint32_t mL = __builtin_amdgcn_alignbit(mr[H - 1], grp_from_prev_u<G>(mr[H - 1], 0x80000000u), 16);
        const uint32_t iL = __builtin_amdgcn_alignbit(ir[H - 1], grp_from_prev_u<G>(ir[H - 1], 0x80000000u), 16);
        const uint32_t dL = __builtin_amdgcn_alignbit(dr[H - 1], grp_from_prev_u<G>(dr[H - 1], 0x80000000u), 16);
        // M and I, descending so that register r-1 still holds the previous row
#pragma unroll
        for (int r = H - 1; r >= 0; r--) {
            const uint32_t mp = r ? mr[r - 1] : mL;
            const uint32_t ip = r ? ir[r - 1] : iL;
            const uint32_t dp = r ? dr[r - 1] : dL;
            const uint32_t s = pk_max16(pk_max16(pk_adds16(mp, tmm[r]), pk_adds16(ip, tim[r])),
                                        pk_max16(pk_adds16(xb2, tbm[r]), pk_adds16(dp, tdm[r])));
            const uint32_t iv = pk_max16(pk_adds16(mr[r], tmi[r]), pk_adds16(ir[r], tii[r]));
            mr[r] = pk_adds16(s, sc[r]);
            ir[r] = iv;
        }
        // D chain, ascending, from M alone; then lazy-F over the half-block boundaries
        const uint32_t mcL = __builtin_amdgcn_alignbit(mr[H - 1], grp_from_prev_u<G>(mr[H - 1], 0x80000000u), 16);
        dr[0] = pk_adds16(mcL, tmd[0]);
#pragma unroll
        for (int r = 1; r < H; r++) dr[r] = pk_max16(pk_adds16(mr[r - 1], tmd[r]), pk_adds16(dr[r - 1], tdd[r]));
        for (int it = 0; it < 2 * G - 1; it++) {
            const uint32_t din = __builtin_amdgcn_alignbit(dr[H - 1], grp_from_prev_u<G>(dr[H - 1], 0x80000000u), 16);
            const uint32_t nd = pk_max16(dr[0], pk_adds16(din, tdd[0]));
            if (!__any(nd != dr[0])) break;
            dr[0] = nd;
#pragma unroll
            for (int r = 1; r < H; r++) dr[r] = pk_max16(dr[r], pk_adds16(dr[r - 1], tdd[r]));
        }
        uint32_t e0 = 0x80008000u, e1 = 0x80008000u;
#pragma unroll
        for (int r = 0; r < H; r++) {
            if (r & 1)
                e1 = pk_max16(e1, pk_max16(mr[r], dr[r]));
            else
                e0 = pk_max16(e0, pk_max16(mr[r], dr[r]));
        }
        uint32_t e = pk_max16(e0, e1);
        e = pk_max16(e, __builtin_amdgcn_alignbit(e, e, 16));
        int xe = (int)(short)(e & 0xFFFFu);
        xe = grp_max_i<G>(xe);
        xj = max(xj + qloop, xe * (1 << SX_Q_SHIFT) + ce);
        xn = xn + qloop;
        xbq = max(max(xn, xj) >> SX_Q_SHIFT, -32768);
    };
    FwdStream rs;
    rs.init(seq);
    for (int i = 0; i < L; i++) step(rs.next());
    return xj - ((1 << SX_Q_SHIFT) - 1);  // xJ + move
}

// ---------------------------------------------------------------------------
// Viterbi score (dp.rs:152-189)
// ---------------------------------------------------------------------------
// MONLY: the profile's M->D and D->D scores are all <= 0, so no D value of a row exceeds
// the row's best M value and the row maximum xE is taken over M alone (the same value as
// over M and D; see SxProfiles::vit_m_only).
template <int Q, int G, int TM, bool MONLY>
__device__ __forceinline__ float vit_pair(const float* __restrict__ em, int ew, const float* __restrict__ tr,
                                          const uint8_t* seq, int L, float lloop, float lmove) {
    constexpr int WID = Geo<Q, G>::WID;
    constexpr bool HOIST = TM == TM_REG;
    const int j = (int)(threadIdx.x & (G - 1));
    const int kb = j * Geo<Q, G>::STRIDE + 1;  // LDS slot of the lane's first position
    const EmRows<Q, G, TM> er(em, ew, j);
    // tr rows: MM MI MD IM II DM DD BM; Tx(k) = tr[x*WID + k]
    float mr[Q], ir[Q], dr[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        mr[q] = SX_NEG_INF;
        ir[q] = SX_NEG_INF;
        dr[q] = SX_NEG_INF;
    }
    float xn = 0.0f, xb = lmove, xj = SX_NEG_INF, xc = SX_NEG_INF;
    FwdStream rs;
    rs.init(seq);
    // the insert states of a row that read '*' are -inf (dp.rs:165,175): a rare residue,
    // so they are cleared at the top of the next row, when any lane of the wave read it,
    // instead of by a select in every cell
    bool star = false;
    for (int i = 0; i < L; i++) {
        const int x = rs.next();
        if (__any(star)) {
#pragma unroll
            for (int q = 0; q < Q; q++) ir[q] = star ? SX_NEG_INF : ir[q];
        }
        star = x == SX_CODE_STAR;
        const int kbo = HOIST ? kb : opaque(kb);
        const float* __restrict__ tmm = tr + SX_TMM * WID + kbo;
        const float* __restrict__ tmi = tr + SX_TMI * WID + kbo;
        const float* __restrict__ tmd = tr + SX_TMD * WID + kbo;
        const float* __restrict__ tim = tr + SX_TIM * WID + kbo;
        const float* __restrict__ tii = tr + SX_TII * WID + kbo;
        const float* __restrict__ tdm = tr + SX_TDM * WID + kbo;
        const float* __restrict__ tdd = tr + SX_TDD * WID + kbo;
        const float* __restrict__ tbm = tr + SX_TBM * WID + kbo;
        const float* __restrict__ sc = er.row(x);
        const float mL = grp_from_prev<G>(mr[Q - 1], SX_NEG_INF);
        const float iL = grp_from_prev<G>(ir[Q - 1], SX_NEG_INF);
        const float dL = grp_from_prev<G>(dr[Q - 1], SX_NEG_INF);
        // M and I, descending so that position q-1 still holds the previous row
#pragma unroll
        for (int q = Q - 1; q >= 0; q--) {
            const float mp = q ? mr[q - 1] : mL;
            const float ip = q ? ir[q - 1] : iL;
            const float dp = q ? dr[q - 1] : dL;
            const float s = fmaxf(fmaxf(mp + tmm[q - 1], ip + tim[q - 1]), fmaxf(xb + tbm[q], dp + tdm[q - 1]));
            const float iv = fmaxf(mr[q] + tmi[q], ir[q] + tii[q]);
            mr[q] = s + sc[q];
            ir[q] = iv;
        }
        // D chain, ascending; the value entering from the left lane is resolved below
        const float mcL = grp_from_prev<G>(mr[Q - 1], SX_NEG_INF);
        dr[0] = mcL + tmd[-1];
#pragma unroll
        for (int q = 1; q < Q; q++) dr[q] = fmaxf(mr[q - 1] + tmd[q - 1], dr[q - 1] + tdd[q - 1]);
        if constexpr (G > 1) {
            // lazy-F: re-run the chain from each lane's true incoming D until nothing
            // changes.  The first pass runs without a test (a lane's carry raises its first
            // D on nearly every row, and a pass that raises nothing changes nothing); the
            // loop then tests before each further pass
            {
                const float c = grp_from_prev<G>(dr[Q - 1], SX_NEG_INF) + tdd[-1];
                dr[0] = fmaxf(dr[0], c);
#pragma unroll
                for (int q = 1; q < Q; q++) dr[q] = fmaxf(dr[q], dr[q - 1] + tdd[q - 1]);
            }
            for (int it = 2; it < G; it++) {
                const float din = grp_from_prev<G>(dr[Q - 1], SX_NEG_INF);
                const float c = din + tdd[-1];
                if (!__any(c > dr[0])) break;
                dr[0] = fmaxf(dr[0], c);
#pragma unroll
                for (int q = 1; q < Q; q++) dr[q] = fmaxf(dr[q], dr[q - 1] + tdd[q - 1]);
            }
        }
        float e = SX_NEG_INF;
        if constexpr (MONLY) {
#pragma unroll
            for (int q = 0; q < Q; q++) e = fmaxf(e, mr[q]);
        } else {
#pragma unroll
            for (int q = 0; q < Q; q++) e = fmaxf(e, fmaxf(mr[q], dr[q]));
        }
        const float xe = grp_reduce<G>(e, OpMax());
        xj = fmaxf(xj + lloop, xe - SX_LN2F);
        xc = fmaxf(xc + lloop, xe - SX_LN2F);
        xn = xn + lloop;
        xb = fmaxf(xn + lmove, xj + lmove);
    }
    return xc + lmove;
}

// Viterbi over aligned lane blocks (TM_ALN*): the same recurrence as vit_pair, with
// every profile row read four positions per LDS load.  Transition rows are staged
// already shifted to the cell they feed (position k-1 for MM IM DM MD DD, k for MI II
// BM), so the lane's operands for cell q all sit at slot q of its block.  Nothing is
// kept in registers across rows but the three DP rows.
template <int Q, int G, int TM, bool MONLY>
__device__ __forceinline__ float vit_pair_aln(const float* __restrict__ em, const float* __restrict__ tr,
                                              const uint8_t* seq, int L, float lloop, float lmove) {
    using GA = GeoA<Q, G>;
    const int j = (int)(threadIdx.x & (G - 1));
    const EmRows<Q, G, TM> er(em, 0, j);
    const int tj = j * GA::STRIDE;
    float mr[Q], ir[Q], dr[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        mr[q] = SX_NEG_INF;
        ir[q] = SX_NEG_INF;
        dr[q] = SX_NEG_INF;
    }
    float xn = 0.0f, xb = lmove, xj = SX_NEG_INF, xc = SX_NEG_INF;
    FwdStream rs;
    rs.init(seq);
    bool star = false;  // as in vit_pair
    for (int i = 0; i < L; i++) {
        const int x = rs.next();
        if (__any(star)) {
#pragma unroll
            for (int q = 0; q < Q; q++) ir[q] = star ? SX_NEG_INF : ir[q];
        }
        star = x == SX_CODE_STAR;
        // re-read every row (an opaque offset keeps the compiler from hoisting the rows
        // into 8*Q registers, which is what costs TM_REG its occupancy)
        const float* tb = tr + opaque(tj >> 2) * 4;  // a multiple of four floats: stays aligned
        auto trow = [&](int r) {
            return static_cast<const float*>(__builtin_assume_aligned(tb + r * GA::TWID, 16));
        };
        const float* __restrict__ tmm = trow(SX_TMM);
        const float* __restrict__ tmi = trow(SX_TMI);
        const float* __restrict__ tmd = trow(SX_TMD);
        const float* __restrict__ tim = trow(SX_TIM);
        const float* __restrict__ tii = trow(SX_TII);
        const float* __restrict__ tdm = trow(SX_TDM);
        const float* __restrict__ tdd = trow(SX_TDD);
        const float* __restrict__ tbm = trow(SX_TBM);
        const float* __restrict__ sc = er.row(x);
        const float mL = grp_from_prev<G>(mr[Q - 1], SX_NEG_INF);
        const float iL = grp_from_prev<G>(ir[Q - 1], SX_NEG_INF);
        const float dL = grp_from_prev<G>(dr[Q - 1], SX_NEG_INF);
        // M and I, descending so that position q-1 still holds the previous row
#pragma unroll
        for (int q = Q - 1; q >= 0; q--) {
            const float mp = q ? mr[q - 1] : mL;
            const float ip = q ? ir[q - 1] : iL;
            const float dp = q ? dr[q - 1] : dL;
            const float s = fmaxf(fmaxf(mp + tmm[q], ip + tim[q]), fmaxf(xb + tbm[q], dp + tdm[q]));
            const float iv = fmaxf(mr[q] + tmi[q], ir[q] + tii[q]);
            mr[q] = s + sc[q];
            ir[q] = iv;
        }
        // D chain, ascending; the value entering from the left lane is resolved below
        const float mcL = grp_from_prev<G>(mr[Q - 1], SX_NEG_INF);
        dr[0] = mcL + tmd[0];
#pragma unroll
        for (int q = 1; q < Q; q++) dr[q] = fmaxf(mr[q - 1] + tmd[q], dr[q - 1] + tdd[q]);
        if constexpr (G > 1) {
            // lazy-F: re-run the chain from each lane's true incoming D until nothing
            // changes.  The first pass runs without a test (a lane's carry raises its first
            // D on nearly every row, and a pass that raises nothing changes nothing); the
            // loop then tests before each further pass
            {
                const float c = grp_from_prev<G>(dr[Q - 1], SX_NEG_INF) + tdd[0];
                dr[0] = fmaxf(dr[0], c);
#pragma unroll
                for (int q = 1; q < Q; q++) dr[q] = fmaxf(dr[q], dr[q - 1] + tdd[q]);
            }
            for (int it = 2; it < G; it++) {
                const float din = grp_from_prev<G>(dr[Q - 1], SX_NEG_INF);
                const float c = din + tdd[0];
                if (!__any(c > dr[0])) break;
                dr[0] = fmaxf(dr[0], c);
#pragma unroll
                for (int q = 1; q < Q; q++) dr[q] = fmaxf(dr[q], dr[q - 1] + tdd[q]);
            }
        }
        float e = SX_NEG_INF;
        if constexpr (MONLY) {
#pragma unroll
            for (int q = 0; q < Q; q++) e = fmaxf(e, mr[q]);
        } else {
#pragma unroll
            for (int q = 0; q < Q; q++) e = fmaxf(e, fmaxf(mr[q], dr[q]));
        }
        const float xe = grp_reduce<G>(e, OpMax());
        xj = fmaxf(xj + lloop, xe - SX_LN2F);
        xc = fmaxf(xc + lloop, xe - SX_LN2F);
        xn = xn + lloop;
        xb = fmaxf(xn + lmove, xj + lmove);
    }
    return xc + lmove;
}

// ---------------------------------------------------------------------------
// Forward (dp.rs:229-298), multihit probabilities e_move = e_loop = 1/2
// ---------------------------------------------------------------------------
// The D chain is affine across lanes: with c the D value at the lane's first position k0
// (set by the lanes before it), the lane's values are D[k0+q] = Dl[q] + c * PC[q], where
// Dl is the chain started from Dl[0] = 0, and the value it hands on is out(c) = Dl_out +
// c * PC[Q].  Each row computes M, I and Dl over the lane's cells, an affine scan of the
// lanes' maps gives every lane its c, and the row sum adds c * sum(PC) to the sum of M and
// Dl.  The stored D values stay in the (Dl, c) form: the next row corrects each one as it
// reads it (one FMA per cell), so no extra pass is needed.  With the records in registers
// (TM_REG) a row is two passes: M and I in descending position order, so that every cell
// overwrites its own registers and the loop carries no register rotation (in one
// ascending pass the compiler copies the cells' registers once a row), then the Dl chain
// ascending (8-9% faster).  With the records in LDS it is one ascending pass, each record
// read once: there the two passes measured up to 17% slower (the second pass's serial
// chain of Q FMAs, with {MD, DD} held in registers or read again), though they issue
// fewer instructions.
template <int Q, int G, bool SPEC, int TM>
__device__ __forceinline__ double fwd_pair(const float* __restrict__ eot, int ew, const float* __restrict__ trec,
                                           const uint8_t* seq, int L, float ploop, float pmove, float* x5,
                                           double* lsrow, bool writer) {
    constexpr bool HOIST = TM == TM_REG;
    // long models read their emission rows from global memory, one register a position: a
    // cell's register is refilled with the next row's value as soon as this row has used
    // it, so the load has a whole row to arrive (up to 40 positions a lane, 32 for the
    // special-row writers; wider lanes read the row inline)
    constexpr bool GEM = TM == TM_GEM && Q <= (SPEC ? 32 : 40);
    using FR = FwdRec<Q, G>;
    const int j = (int)(threadIdx.x & (G - 1));
    const EmRows<Q, G, TM> er(eot, ew, j);
    const float* __restrict__ pcb = trec + FR::REC + j * FR::PSTRIDE;
    const float blk = pcb[Q];       // the block's D-chain slope
    const float spc = pcb[Q + 1];   // sum of PC[0..Q-1]
    // TM_REG: the lane's records and D-chain products loaded once, held in registers
    // TM_REG: the lane's records and D-chain products loaded once, held in registers
    float4 hra[HOIST ? Q : 1], hrb[HOIST ? Q : 1];
    float hpc[HOIST ? Q : 1];
    // XSUM: the lane's part of the row sum, sum(M) + sum(Dl), as one weighted sum of M.
    // Dl[t] = sum_{r<t} M[r] MD[r] prod_{s=r+1}^{t-1} DD[s], so sum_t Dl[t] = sum_r M[r]
    // MD[r] W[r] with W[Q-1] = 0, W[r] = 1 + DD[r+1] W[r+1]: the weights X[r] = 1 + MD[r]
    // W[r] depend on the profile only, and the row sum costs one FMA a cell instead of two
    // adds (with the records in registers at 12 positions a lane: 3-5% faster at 8-16
    // lanes; the two-row lanes of 4 and 8 positions would lose a wave per SIMD to the Q
    // more registers, and 16 positions would spill)
    constexpr bool XSUM = HOIST && Q == 12 && !fwd_rows2<Q, TM>();
    float hx[XSUM ? Q : 1];
    if constexpr (HOIST) {
        const float4* __restrict__ r0 = reinterpret_cast<const float4*>(trec + j * FR::RSTRIDE);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            hra[q] = r0[2 * q];
            hrb[q] = r0[2 * q + 1];
            hpc[q] = pcb[q];
        }
        if constexpr (XSUM) {
            float w = 0.0f;
#pragma unroll
            for (int q = Q - 1; q >= 0; q--) {
                if (q < Q - 1) w = fmaf(hrb[q + 1].w, w, 1.0f);
                hx[q] = fmaf(hrb[q].z, w, 1.0f);
            }
        }
    }
    float mr[Q], ir[Q], dl[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        mr[q] = 0.0f;
        ir[q] = 0.0f;
        dl[q] = 0.0f;
    }
    float c = 0.0f;  // this lane's incoming D value of the stored row
    float xn = 1.0f, xb = pmove, xj = 0.0f, xc = 0.0f;
    double lnscale = 0.0;
    if constexpr (SPEC) {
        if (writer) {
            x5[0] = xn; x5[1] = xb; x5[2] = 0.0f; x5[3] = 0.0f; x5[4] = 0.0f;
            lsrow[0] = 0.0;
        }
    }
    FwdStream rs;
    rs.init(seq);
    int x = rs.next();
    // long models: this row's emission values, refilled with the next row's as they are used
    float eb[GEM ? Q : 1];
    if constexpr (GEM) {
        const float* __restrict__ e0 = er.row(x);
#pragma unroll
        for (int q = 0; q < Q; q++) eb[q] = e0[q];
    }
    bool star_prev = false;
    // one row; the loop runs two rows an iteration where fwd_rows2 says so
    auto row = [&](int i) {
        // a '*' residue emits no insert state (dp.rs:256,260): the inserts of a '*' row are
        // cleared at the top of the next row, under a wave vote, instead of a multiply in
        // every cell (the row itself reads no insert state of its own)
        if (__builtin_amdgcn_ballot_w64(star_prev) != 0ull) [[unlikely]] {
            __asm__ volatile("s_nop 0");  // a side effect: this rare path stays a branch, not selects
            if (star_prev) {
#pragma unroll
                for (int q = 0; q < Q; q++) ir[q] = 0.0f;
            }
        }
        star_prev = x == SX_CODE_STAR;
        const float* __restrict__ eo = er.row(x);
        // the next residue, read without a branch: past the last row the stream reads the
        // target's padding, and row 0 is then loaded instead (its values go unused)
        const int xs = rs.next();
        x = i + 1 < L ? xs : 0;
        const float* __restrict__ en = er.row(x);
        if constexpr (!GEM) (void)en;
        const int ro = HOIST ? 0 : opaque(0);
        const float4* __restrict__ rec = reinterpret_cast<const float4*>(trec + j * FR::RSTRIDE + ro);
        const float* __restrict__ pc = pcb + ro;
        float em = 0.0f, ed = 0.0f;
        float dnext = 0.0f;  // Dl of the next cell
        if constexpr (HOIST) {
            // M and I in descending position order, each cell overwriting its own registers
            // (cell q reads the stored row at q-1 and q only, so the loop carries no
            // register rotation), then the Dl chain ascending
            const float mpi = grp_from_prev0<G>(mr[Q - 1]);
            const float ipi = grp_from_prev0<G>(ir[Q - 1]);
            const float dpi = grp_from_prev0<G>(fmaf(c, hpc[Q - 1], dl[Q - 1]));
#pragma unroll
            for (int q = Q - 1; q >= 0; q--) {
                const float4 ta = hra[q], tb = hrb[q];  // {MM IM DM BM}, {MI II MD DD}
                const float mp = q ? mr[q - 1] : mpi;
                const float ip = q ? ir[q - 1] : ipi;
                const float dp = q ? fmaf(c, hpc[q - 1], dl[q - 1]) : dpi;  // stored D at q-1
                float s = xb * ta.w;
                s = fmaf(mp, ta.x, s);
                s = fmaf(ip, ta.y, s);
                s = fmaf(dp, ta.z, s);
                const float mv = s * eo[q];
                ir[q] = fmaf(mr[q], tb.x, ir[q] * tb.y);
                mr[q] = mv;
                if constexpr (XSUM)
                    em = q == Q - 1 ? mv * hx[q] : fmaf(mv, hx[q], em);
                else
                    em = q == Q - 1 ? mv : em + mv;
            }
#pragma unroll
            for (int q = 0; q < Q; q++) {
                dl[q] = dnext;
                if constexpr (!XSUM)
                    if (q) ed = q == 1 ? dnext : ed + dnext;
                dnext = q ? fmaf(dnext, hrb[q].w, mr[q] * hrb[q].z) : mr[q] * hrb[q].z;
            }
        } else {
            // records in LDS: one ascending pass, each record read once
            const float dlast = fmaf(c, pc[Q - 1], dl[Q - 1]);
            float mpo = grp_from_prev0<G>(mr[Q - 1]);
            float ipo = grp_from_prev0<G>(ir[Q - 1]);
            float dpo = grp_from_prev0<G>(dlast);
#pragma unroll
            for (int q = 0; q < Q; q++) {
                const float4 ta = rec[2 * q], tb = rec[2 * q + 1];  // {MM IM DM BM}, {MI II MD DD}
                const float mo = mr[q], io = ir[q];
                const float dold = q + 1 < Q ? fmaf(c, pc[q], dl[q]) : dlast;  // stored D at this cell
                float s = xb * ta.w;
                s = fmaf(mpo, ta.x, s);
                s = fmaf(ipo, ta.y, s);
                s = fmaf(dpo, ta.z, s);
                float mv;
                if constexpr (GEM) {
                    mv = s * eb[q];
                    eb[q] = en[q];
                } else {
                    mv = s * eo[q];
                }
                ir[q] = fmaf(mo, tb.x, io * tb.y);
                mr[q] = mv;
                dl[q] = dnext;
                em += mv;
                ed += dnext;
                dnext = fmaf(dnext, tb.w, mv * tb.z);
                mpo = mo;
                ipo = io;
                dpo = dold;
            }
        }
        float cn = dnext;  // the lane's map out(c) = dnext + c * blk, scanned over the group
        if constexpr (G > 1) {
            grp_affine_scan<G, true>(cn, blk);
            cn = grp_from_prev0<G>(cn);
        } else {
            cn = 0.0f;
        }
        c = cn;
        float xe = grp_reduce<G>(fmaf(c, spc, em + ed), OpAdd());
        xj = fmaf(xj, ploop, xe * 0.5f);
        xc = fmaf(xc, ploop, xe * 0.5f);
        xn = xn * ploop;
        xb = fmaf(xn, pmove, xj * pmove);
        if (xe > SX_RESCALE) {
            const float inv = 1.0f / xe;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                mr[q] *= inv;
                ir[q] *= inv;
                dl[q] *= inv;
            }
            c *= inv;
            xn *= inv;
            xb *= inv;
            xj *= inv;
            xc *= inv;
            lnscale += log((double)xe);
            xe = 1.0f;
        }
        if constexpr (SPEC) {
            if (writer) {
                float* o = x5 + 5 * (i + 1);
                o[0] = xn; o[1] = xb; o[2] = xe; o[3] = xj; o[4] = xc;
                lsrow[i + 1] = lnscale;
            }
        }
        };
    int i = 0;
    if constexpr (fwd_rows2<Q, TM>()) {
        for (; i + 2 <= L; i += 2) {
            row(i);
            row(i + 1);
        }
    }
    for (; i < L; i++) row(i);
    const float fin = xc * pmove;
    return (fin > 0.0f ? (double)logf(fin) : -(double)__builtin_inf()) + lnscale;
}

// ---------------------------------------------------------------------------
// Backward (dp.rs:301-392)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double logsum_d(double a, double b) {
    // dp.rs:32-40 in double; written without NaN compares (kernels build with -fno-honor-nans)
    const double hi = a > b ? a : b, lo = a > b ? b : a;
    if (lo == -(double)__builtin_inf() || hi - lo >= 30.0) return hi;
    return hi + log1p(exp(lo - hi));
}

template <int Q, int G, bool SPEC, int TM>
__device__ __forceinline__ double bwd_pair(const float* __restrict__ eot, int ew, const float* __restrict__ tr,
                                           const uint8_t* seq, int L, float ploop, float pmove, float lloop,
                                           float lmove, float* x5, double* lsrow, bool writer) {
    constexpr int WID = Geo<Q, G>::WID;
    constexpr bool HOIST = TM == TM_REG;
    constexpr int NX = EmRows<Q, G, TM>::NEXT;  // the next lane's first position, from row(x)
    const int j = (int)(threadIdx.x & (G - 1));
    const int kb = j * Geo<Q, G>::STRIDE + 1;  // LDS slot of the lane's first position
    const EmRows<Q, G, TM> er(eot, ew, j);
    const float* __restrict__ pmd = tr + SX_TMD * WID + kb;
    const float* __restrict__ pdd = tr + SX_TDD * WID + kb;
    const float* __restrict__ pbm = tr + SX_TBM * WID + kb;
    float blk = 1.0f;
    if constexpr (G > 1) {
#pragma unroll
        for (int q = 0; q < Q; q++) blk *= pdd[q];
    }
    float mr[Q], ir[Q], dr[Q];  // row i+1 values (mn, in, dn), overwritten by row i
    float xc = pmove;
    float xe = xc * 0.5f;
    float xj = 0.0f, xn = 0.0f, xb = 0.0f;
    // row L: dn[k] = xe + dn[k+1]*p_dd[k], mn[k] = xe + dn[k+1]*p_md[k]
    {
        dr[Q - 1] = xe;
#pragma unroll
        for (int q = Q - 2; q >= 0; q--) dr[q] = fmaf(dr[q + 1], pdd[q], xe);
        if constexpr (G > 1) {
            float A = dr[0];
            grp_affine_scan<G, false>(A, blk);
            const float din = grp_from_next<G>(A, 0.0f);
            float P = pdd[Q - 1];
            dr[Q - 1] = fmaf(din, P, dr[Q - 1]);
#pragma unroll
            for (int q = Q - 2; q >= 0; q--) {
                P *= pdd[q];
                dr[q] = fmaf(din, P, dr[q]);
            }
        }
        const float dR = grp_from_next<G>(dr[0], 0.0f);
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const float dn1 = q < Q - 1 ? dr[q + 1] : dR;
            mr[q] = fmaf(dn1, pmd[q], xe);
            ir[q] = 0.0f;
        }
    }
    double lnscale = 0.0;
    if constexpr (SPEC) {
        if (writer) {
            float* o = x5 + 5 * L;
            o[0] = 0.0f; o[1] = 0.0f; o[2] = xe; o[3] = 0.0f; o[4] = xc;
            lsrow[L] = 0.0;
        }
    }
    RevStream rs;
    rs.init(seq, L);
    int xnext = rs.next();  // residue L (position L-1)
    double n1 = -(double)__builtin_inf();
    for (int i = L - 1; i >= 1; i--) {
        const int kbo = HOIST ? kb : opaque(kb);
        const float* __restrict__ pmm = tr + SX_TMM * WID + kbo;
        const float* __restrict__ pmi = tr + SX_TMI * WID + kbo;
        const float* __restrict__ pmd = tr + SX_TMD * WID + kbo;
        const float* __restrict__ pim = tr + SX_TIM * WID + kbo;
        const float* __restrict__ pii = tr + SX_TII * WID + kbo;
        const float* __restrict__ pdm = tr + SX_TDM * WID + kbo;
        const float* __restrict__ pdd = tr + SX_TDD * WID + kbo;
        const float* __restrict__ pbm = tr + SX_TBM * WID + kbo;
        const float* __restrict__ eo = er.row(xnext);  // emissions of residue i+1
        const float ins = xnext != SX_CODE_STAR ? 1.0f : 0.0f;
        float bl = 0.0f;
#pragma unroll
        for (int q = 0; q < Q; q++) bl = fmaf(mr[q] * pbm[q], eo[q], bl);
        xb = grp_reduce<G>(bl, OpAdd());
        xj = fmaf(xj, ploop, xb * pmove);
        xc = xc * ploop;
        xe = fmaf(xj, 0.5f, xc * 0.5f);
        xn = fmaf(xn, ploop, xb * pmove);
        float rowmax = fmaxf(fmaxf(fmaxf(fmaxf(xb, xe), xn), xj), xc);
        const float mR = grp_from_next<G>(mr[0], 0.0f);  // mn at the next lane's first k
        // D chain (descending): dc[k] = mnext*p_dm[k] + dc[k+1]*p_dd[k] + xe
        {
            const float mnx = mR * eo[NX];
            dr[Q - 1] = fmaf(mnx, pdm[Q - 1], xe);
        }
#pragma unroll
        for (int q = Q - 2; q >= 0; q--) {
            const float mnx = mr[q + 1] * eo[q + 1];
            dr[q] = fmaf(dr[q + 1], pdd[q], fmaf(mnx, pdm[q], xe));
        }
        if constexpr (G > 1) {
            float A = dr[0];
            grp_affine_scan<G, false>(A, blk);
            const float din = grp_from_next<G>(A, 0.0f);
            float P = pdd[Q - 1];
            dr[Q - 1] = fmaf(din, P, dr[Q - 1]);
#pragma unroll
            for (int q = Q - 2; q >= 0; q--) {
                P *= pdd[q];
                dr[q] = fmaf(din, P, dr[q]);
            }
        }
        const float dR = grp_from_next<G>(dr[0], 0.0f);
        float lmax = SX_NEG_INF;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const float mnx = q < Q - 1 ? mr[q + 1] * eo[q + 1] : mR * eo[NX];
            const float iv = ir[q] * ins;
            const float dn1 = q < Q - 1 ? dr[q + 1] : dR;
            const float mv = fmaf(dn1, pmd[q], fmaf(mnx, pmm[q], iv * pmi[q]) + xe);
            ir[q] = fmaf(mnx, pim[q], iv * pii[q]);
            mr[q] = mv;
            lmax = fmaxf(lmax, mv);
        }
        rowmax = fmaxf(rowmax, grp_reduce<G>(lmax, OpMax()));
        if (rowmax > SX_RESCALE) {
            const float inv = 1.0f / rowmax;
#pragma unroll
            for (int q = 0; q < Q; q++) {
                mr[q] *= inv;
                ir[q] *= inv;
                dr[q] *= inv;
            }
            xb *= inv;
            xe *= inv;
            xn *= inv;
            xj *= inv;
            xc *= inv;
            lnscale += log((double)rowmax);
        }
        if constexpr (SPEC) {
            if (writer) {
                float* o = x5 + 5 * i;
                o[0] = xn; o[1] = xb; o[2] = xe; o[3] = xj; o[4] = xc;
                lsrow[i] = lnscale;
            }
        }
        if (i == 1) n1 = (xn > 0.0f ? log((double)xn) : -(double)__builtin_inf()) + lnscale;
        xnext = rs.next();
    }
    // row 0: B(0) from residue 1
    const float* __restrict__ eo = er.row(xnext);
    float bl = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; q++) bl = fmaf(mr[q] * pbm[q], eo[q], bl);
    const float b0 = grp_reduce<G>(bl, OpAdd());
    if constexpr (SPEC) {
        if (writer) {
            float* o = x5;
            o[0] = 0.0f; o[1] = b0; o[2] = 0.0f; o[3] = 0.0f; o[4] = 0.0f;
            lsrow[0] = lnscale;
        }
    }
    const double lb0 = (b0 > 0.0f ? log((double)b0) : -(double)__builtin_inf()) + lnscale;
    return logsum_d(n1 + (double)lloop, lb0 + (double)lmove);
}

// ---------------------------------------------------------------------------
// tile loop: one workgroup = one profile staged in LDS, one target per lane group
// ---------------------------------------------------------------------------
enum { ST_MSV = 0, ST_VIT = 1, ST_FWD = 2, ST_BWD = 3, ST_MSVQ = 4, ST_VITQ = 5 };

// Floats of a kernel's dynamic LDS: the staged emission rows (none for TM_GEM) and,
// past MSV, the eight transition rows after them.
template <int Q, int G, int TM>
constexpr int lds_emission_floats() {
    if constexpr (TM == TM_GEM)
        return 0;
    else if constexpr (tm_aligned<TM>())
        return GeoA<Q, G>::TAB;
    else
        return SX_NCODE * Geo<Q, G>::WID;
}
template <int Q, int G, int ST, int TM>
constexpr int lds_floats() {
    if constexpr (ST == ST_MSVQ)
        return q_tab_dwords<Q, G, TM>();
    else if constexpr (ST == ST_VITQ)
        return vq_tab_dwords<Q, G, TM>();
    else if constexpr (tm_aligned<TM>())
        return GeoA<Q, G>::TAB + (ST == ST_MSV ? 0 : SX_NT * GeoA<Q, G>::TWID);
    else if constexpr (ST == ST_FWD)
        return round4(lds_emission_floats<Q, G, TM>()) + FwdRec<Q, G>::TAB;
    else
        return lds_emission_floats<Q, G, TM>() + (ST == ST_MSV ? 0 : SX_NT * tm_wid<Q, G, TM>());
}

// LDS row width recorded in the kernel table (dwords for the int16 stage)
template <int Q, int G, int ST, int TM>
constexpr int k_wid() {
    if constexpr (ST == ST_MSVQ || ST == ST_VITQ) {
        if constexpr (tm_aligned<TM>())
            return GeoA<Q / 2, G>::WID;
        else
            return GeoQ1<Q>::WID;
    } else {
        return tm_wid<Q, G, TM>();
    }
}

struct TileInfo {
    int32_t prof, count;
    int64_t first, out;
    bool by_tid;  // result index = out + target id (else out + lane group)
};

__device__ __forceinline__ TileInfo tile_of(const SxWork& w, int64_t t) {
    TileInfo ti;
    if (w.tiles) {
        const SxTile tl = w.tiles[t];
        ti.prof = tl.prof;
        ti.count = tl.count;
        ti.first = tl.first;
        ti.out = tl.out;
        ti.by_tid = false;
    } else {
        int64_t pi;
        int c;
        if (w.chunk_major) {
            const int64_t np = w.ntiles / w.nchunks;
            c = (int)(t / np);
            pi = t - (int64_t)c * np;
        } else {
            pi = t / w.nchunks;
            c = (int)(t - pi * w.nchunks);
        }
        ti.prof = w.profs[pi];
        ti.first = (int64_t)c * w.chunk;
        const int rem = w.ntids - (int)ti.first;
        ti.count = rem < w.chunk ? rem : w.chunk;
        ti.out = (int64_t)ti.prof * w.out_stride + (w.out_by_pos ? ti.first : 0);
        ti.by_tid = !w.out_by_pos;
    }
    return ti;
}

// Waves per SIMD the compiler must leave room for: the Forward kernels with transitions in
// registers need two (at one they leave the SIMD idle through every dependent step).
// So do the long-model Forward kernels up to 40 positions a lane (at most 254 VGPRs, no
// spills); the special-row writers and the wider ones would spill at two.
template <int Q, int STAGE, bool SPEC, int TM>
constexpr int k_min_waves() {
    return STAGE == ST_FWD && (TM == TM_REG || (TM == TM_GEM && !SPEC && Q <= 40)) ? 2 : 1;
}

// Threads per workgroup: the table mode's, except the long-model Forward and Backward
// scorers of 40 and 48 positions a lane.  Their staged transition records (92-111 KB)
// leave room for one workgroup per CU, so a 256-thread workgroup would run one wave per
// SIMD; 512 threads (eight pairs of one profile) share the table and fill two.
template <int Q, int STAGE, bool SPEC, int TM>
constexpr int k_block() {
    return TM == TM_GEM && !SPEC && ((STAGE == ST_FWD && (Q == 40 || Q == 48)) || (STAGE == ST_BWD && Q == 40))
               ? 512
               : tm_block<TM>();
}

template <int Q, int G, int STAGE, bool SPEC, int TM>
__global__ void __launch_bounds__((k_block<Q, STAGE, SPEC, TM>()), (k_min_waves<Q, STAGE, SPEC, TM>()))
    k_tiles(SxProfiles P, SxTargets T, SxWork W, void* out, SxSpecialsOut SP) {
    static_assert(STAGE == ST_MSV || STAGE == ST_VIT || STAGE == ST_MSVQ || STAGE == ST_VITQ || !tm_aligned<TM>(),
                  "aligned lane blocks are an MSV and Viterbi layout");
    constexpr int WID = tm_wid<Q, G, TM>();
    constexpr int NTAB_E = lds_emission_floats<Q, G, TM>();  // emission rows staged in LDS
    extern __shared__ __align__(16) float lds[];
    __shared__ long long s_tile;
    const int gid = (int)(threadIdx.x / G);
    const int jl = (int)(threadIdx.x & (G - 1));
    int cur = -1;
    const int64_t tlo = W.dyn_lo ? *W.dyn_lo : 0;
    const int64_t thi = W.dyn_lo ? *W.dyn_hi : W.ntiles;
    int64_t t = (int64_t)blockIdx.x + W.tile_base;
    for (;;) {
        if (W.head) {
            if (threadIdx.x == 0) s_tile = (long long)atomicAdd(W.head, 1ull);
            __syncthreads();
            t = s_tile;
        }
        if (tlo + t >= thi) break;
        const TileInfo ti = tile_of(W, tlo + t);
        if (ti.prof != cur) {
            __syncthreads();
            const int m = P.m[ti.prof];
            const int w = m + 1;
            if constexpr (STAGE == ST_MSVQ || STAGE == ST_VITQ) {
                // packed int16 table: dword (lane j, register r) = slots j*Q + 1 + r
                // (low half) and j*Q + 1 + Q/2 + r (high half); slot s holds model
                // position s - (Q*G - M), the slots before position 1 the score -32768
                constexpr int H = Q / 2;
                const int16_t* src = P.qsc + P.off[ti.prof];
                uint32_t* qt = reinterpret_cast<uint32_t*>(lds);
                const int shift = Q * G - m;
                auto qv = [&](int x, int slot) -> uint32_t {
                    const int k = slot - shift;
                    return k >= 1 ? (uint32_t)(uint16_t)src[x * w + k] : 0x8000u;
                };
                if constexpr (tm_aligned<TM>()) {
                    using GA = GeoA<H, G>;
                    for (int idx = threadIdx.x; idx < GA::TAB; idx += blockDim.x) {
                        const int r = idx - (idx / GA::COPY) * GA::COPY;
                        const int x = r / GA::WID, pos = r - x * GA::WID;
                        const int jj = pos / GA::STRIDE, sl = pos - jj * GA::STRIDE;
                        uint32_t v = 0x80008000u;
                        if (x < SX_NCODE && jj < G && sl < H) {
                            const int k0 = jj * Q + 1 + sl;
                            v = qv(x, k0) | (qv(x, k0 + H) << 16);
                        }
                        qt[idx] = v;
                    }
                } else {
                    constexpr int QW = GeoQ1<Q>::WID;
                    for (int idx = threadIdx.x; idx < SX_NCODE * QW; idx += blockDim.x) {
                        const int x = idx / QW, r = idx - x * QW;
                        qt[idx] = r < H ? (qv(x, r + 1) | (qv(x, r + 1 + H) << 16)) : 0x80008000u;
                    }
                }
                if constexpr (STAGE == ST_VITQ) {
                    // transition rows, shifted to the cell they feed: position k-1 for
                    // MM MD IM DM DD, k for MI II BM; -32768 in the padding slots
                    const int16_t* tsrc = P.qtsc + P.toff[ti.prof];
                    auto tv = [&](int r, int slot) -> uint32_t {
                        const int k = slot - shift;
                        const bool at_k = r == SX_TMI || r == SX_TII || r == SX_TBM;
                        return k >= 1 ? (uint32_t)(uint16_t)tsrc[r * w + (at_k ? k : k - 1)] : 0x8000u;
                    };
                    int tb, twd, tst;
                    if constexpr (tm_aligned<TM>()) {
                        tb = GeoA<H, G>::TAB;
                        twd = GeoA<H, G>::TWID;
                        tst = GeoA<H, G>::STRIDE;
                    } else {
                        tb = GeoQV1<Q, G>::TBASE;
                        twd = GeoQV1<Q, G>::TW;
                        tst = GeoQV1<Q, G>::TW;
                    }
                    for (int idx = threadIdx.x; idx < SX_NT * twd; idx += blockDim.x) {
                        const int r = idx / twd, pos = idx - r * twd;
                        const int jj = pos / tst, sl = pos - jj * tst;
                        uint32_t v = 0x80008000u;
                        if (jj < G && sl < H) {
                            const int s0 = jj * Q + 1 + sl;
                            v = tv(r, s0) | (tv(r, s0 + H) << 16);
                        }
                        qt[tb + idx] = v;
                    }
                }
            } else if constexpr (TM == TM_GEM) {
                // emission rows stay in global memory (padded copy)
            } else if constexpr (tm_aligned<TM>()) {
                const float* src = P.msc + P.off[ti.prof];
                using GA = GeoA<Q, G>;
                for (int idx = threadIdx.x; idx < GA::TAB; idx += blockDim.x) {
                    const int r = idx - (idx / GA::COPY) * GA::COPY;  // offset inside its copy
                    const int x = r / WID, pos = r - x * WID;
                    const int j = pos / GA::STRIDE, s = pos - j * GA::STRIDE;
                    const int k = j * Q + 1 + s;
                    lds[idx] = x < SX_NCODE && j < G && s < Q && k < w ? src[x * w + k] : SX_NEG_INF;
                }
                if constexpr (STAGE == ST_VIT) {
                    // slot s of lane j's block holds position j*Q + s (the k-1 operand of
                    // the lane's cell s) in the rows read at k-1 (MM IM DM MD DD) and
                    // position j*Q + 1 + s in those read at k (MI II BM)
                    const float* tsrc = P.tsc + P.toff[ti.prof];
                    for (int idx = threadIdx.x; idx < SX_NT * GA::TWID; idx += blockDim.x) {
                        const int r = idx / GA::TWID, pos = idx - r * GA::TWID;
                        const int j = pos / GA::STRIDE, s = pos - j * GA::STRIDE;
                        const int k = j * Q + s + (r == SX_TMI || r == SX_TII || r == SX_TBM ? 1 : 0);
                        lds[GA::TAB + idx] = s < Q && k < w ? tsrc[r * w + k] : SX_NEG_INF;
                    }
                }
            } else if (STAGE == ST_MSV || STAGE == ST_VIT) {
                const float* src = P.msc + P.off[ti.prof];
                for (int idx = threadIdx.x; idx < NTAB_E; idx += blockDim.x) {
                    const int x = idx / WID, k = slot_k<Q, G>(idx - x * WID);
                    lds[idx] = k < w ? src[x * w + k] : SX_NEG_INF;
                }
            } else {
                const float* src = P.mo + P.off[ti.prof];
                for (int idx = threadIdx.x; idx < NTAB_E; idx += blockDim.x) {
                    const int x = idx / WID, k = slot_k<Q, G>(idx - x * WID);
                    lds[idx] = k < w ? src[x * w + k] : 0.0f;
                }
            }
            if constexpr (STAGE == ST_FWD) {
                // transition records (FwdRec): lane jj's cell q is model position
                // k = jj*Q + 1 + q; fields read at k-1 (MM IM DM MD DD) or k (BM MI II)
                const float* src = P.tp + P.toff[ti.prof];
                float* rec = lds + round4(NTAB_E);
                using FR = FwdRec<Q, G>;
                constexpr int RS = FR::RSTRIDE;
                for (int idx = threadIdx.x; idx < FR::REC; idx += blockDim.x) {
                    const int jj = idx / RS, r = idx - jj * RS;
                    const int q = r >> 3, c = r & 7;
                    const int k = jj * Q + 1 + q;
                    // c: 0 MM, 1 IM, 2 DM (at k-1), 3 BM, 4 MI, 5 II, 6 MD, 7 DD (at k)
                    const int row = c == 0 ? SX_TMM : c == 1 ? SX_TIM : c == 2 ? SX_TDM : c == 3 ? SX_TBM
                                  : c == 4 ? SX_TMI : c == 5 ? SX_TII : c == 6 ? SX_TMD : SX_TDD;
                    const int kk = c <= 2 ? k - 1 : k;
                    rec[idx] = q < Q && kk >= 0 && kk < w ? src[row * w + kk] : 0.0f;
                }
                // the D-chain products of each lane block, one thread a block
                for (int jj = threadIdx.x; jj < G; jj += blockDim.x) {
                    float* pcb = rec + FR::REC + jj * FR::PSTRIDE;
                    float prod = 1.0f, sum = 0.0f;
                    for (int q = 0; q < Q; q++) {
                        const int k = jj * Q + 1 + q;
                        pcb[q] = prod;
                        sum += prod;
                        prod *= k < w ? src[SX_TDD * w + k] : 0.0f;
                    }
                    pcb[Q] = prod;
                    pcb[Q + 1] = sum;
                }
            } else if (STAGE != ST_MSV && STAGE != ST_MSVQ && !tm_aligned<TM>()) {
                const float* src = (STAGE == ST_VIT ? P.tsc : P.tp) + P.toff[ti.prof];
                const float pad = STAGE == ST_VIT ? SX_NEG_INF : 0.0f;
                for (int idx = threadIdx.x; idx < SX_NT * WID; idx += blockDim.x) {
                    const int r = idx / WID, k = slot_k<Q, G>(idx - r * WID);
                    lds[NTAB_E + idx] = k < w ? src[r * w + k] : pad;
                }
            }
            cur = ti.prof;
            __syncthreads();
        } else if (W.head) {
            __syncthreads();  // s_tile is rewritten only after every lane has read it
        }
        // emission rows: LDS, or the profile's padded global copy for long models
        const float* em = lds;
        if constexpr (TM == TM_GEM)
            em = (STAGE == ST_MSV || STAGE == ST_VIT ? P.gmsc : P.gmo) + P.goff[ti.prof];
        if (gid < ti.count) {
            const int tid = W.tids[ti.first + gid];
            const int L = T.len[tid];
            const uint8_t* seq = T.codes + T.off[tid];
            const SxLen lm = T.lm[tid];
            const int64_t o = ti.out + (ti.by_tid ? (int64_t)tid : (int64_t)gid);
            if constexpr (STAGE == ST_VITQ) {
                int32_t r = INT32_MIN;
                if (L > 0) {
                    // the Viterbi length model rounded up as the host's q_const does
                    const int ql = (int)ceil((double)lm.lloop * SX_Q_FINE + SX_Q_SLACK_FINE);
                    const int qm = (int)ceil((double)lm.lmove * SX_Q_FINE + SX_Q_SLACK_FINE);
                    r = vitq_pair<Q, G, TM>(reinterpret_cast<const uint32_t*>(lds), P.q_tej, seq, L, ql, qm);
                }
                if (jl == 0) reinterpret_cast<int32_t*>(out)[o] = r;
            } else if constexpr (STAGE == ST_MSVQ) {
                int32_t r = INT32_MIN;
                // one-lane groups over the sorted order read the interleaved copy: the
                // wave's 64 targets are positions 64c .. 64c+63 (first is a multiple of
                // the chunk, itself a multiple of 64), one per lane
                const uint32_t* words = reinterpret_cast<const uint32_t*>(seq);
                int ws = 1;
                if constexpr (G == 1) {
                    if (W.ilv && !W.tiles) {
                        const int64_t s = ti.first + gid;
                        words = T.ilv + T.ilv_off[s >> 6] + (s & 63);
                        ws = 64;
                    }
                }
                if (L > 0)
                    r = msvq_pair<Q, G, TM>(reinterpret_cast<const uint32_t*>(lds), P.qtbm[ti.prof], P.q_tej, words, ws,
                                            L, lm.q_tloop, lm.q_tmove);
                if (jl == 0) reinterpret_cast<int32_t*>(out)[o] = r;
            } else if constexpr (STAGE == ST_MSV) {
                float r = SX_NEG_INF;
                if (L > 0) r = msv_pair<Q, G, TM>(em, P.gw, P.msv_tbm[ti.prof], seq, L, lm.msv_tloop, lm.msv_tmove);
                if (jl == 0) reinterpret_cast<float*>(out)[o] = r;
            } else if constexpr (STAGE == ST_VIT) {
                float r = SX_NEG_INF;
                const bool monly = __builtin_amdgcn_readfirstlane(P.vit_m_only[ti.prof]) != 0;  // tile-uniform
                if constexpr (tm_aligned<TM>()) {
                    const float* tr = lds + GeoA<Q, G>::TAB;
                    if (L > 0)
                        r = monly ? vit_pair_aln<Q, G, TM, true>(em, tr, seq, L, lm.lloop, lm.lmove)
                                  : vit_pair_aln<Q, G, TM, false>(em, tr, seq, L, lm.lloop, lm.lmove);
                } else {
                    const float* tr = lds + NTAB_E;
                    if (L > 0)
                        r = monly ? vit_pair<Q, G, TM, true>(em, P.gw, tr, seq, L, lm.lloop, lm.lmove)
                                  : vit_pair<Q, G, TM, false>(em, P.gw, tr, seq, L, lm.lloop, lm.lmove);
                }
                if (jl == 0) reinterpret_cast<float*>(out)[o] = r;
            } else if constexpr (STAGE == ST_FWD) {
                double r = -(double)__builtin_inf();
                float* x5 = nullptr;
                double* lsr = nullptr;
                if constexpr (SPEC) {
                    x5 = SP.x5 + 5 * SP.row_off[o];
                    lsr = SP.ls + SP.row_off[o];
                }
                if (L > 0)
                    r = fwd_pair<Q, G, SPEC, TM>(em, P.gw, lds + round4(NTAB_E), seq, L, lm.ploop, lm.pmove, x5,
                                                 lsr, jl == 0);
                if (jl == 0) reinterpret_cast<double*>(out)[o] = r;
            } else {
                double r = -(double)__builtin_inf();
                float* x5 = nullptr;
                double* lsr = nullptr;
                if constexpr (SPEC) {
                    x5 = SP.x5 + 5 * SP.row_off[o];
                    lsr = SP.ls + SP.row_off[o];
                }
                if (L > 0)
                    r = bwd_pair<Q, G, SPEC, TM>(em, P.gw, lds + NTAB_E, seq, L, lm.ploop, lm.pmove, lm.lloop,
                                                 lm.lmove, x5, lsr, jl == 0);
                if (jl == 0) reinterpret_cast<double*>(out)[o] = r;
            }
        }
        if (!W.head) break;
    }
}

}  // namespace sxk

// ---------------------------------------------------------------------------
// kernel registry
// ---------------------------------------------------------------------------
struct SxKernelEntry {
    int stage;
    int q;
    int g;
    int spec;
    int hoist;
    int wid;
    int block;
    int lds;  // bytes of dynamic LDS per workgroup
    const void* fn;
};

#define SXK_E(ST, Q, G, SP, H) \
    {ST, Q, G, (SP) ? 1 : 0, (int)(H), sxk::k_wid<Q, G, ST, (int)(H)>(), sxk::k_block<Q, ST, SP, (int)(H)>(), \
     4 * sxk::lds_floats<Q, G, ST, (int)(H)>(), \
     reinterpret_cast<const void*>(&sxk::k_tiles<Q, G, ST, SP, (int)(H)>)}

#define SXK_ALLG(ST, Q, SP, H)                                                                          \
    SXK_E(ST, Q, 1, SP, H), SXK_E(ST, Q, 2, SP, H), SXK_E(ST, Q, 4, SP, H), SXK_E(ST, Q, 8, SP, H),     \
        SXK_E(ST, Q, 16, SP, H), SXK_E(ST, Q, 32, SP, H), SXK_E(ST, Q, 64, SP, H)

// Viterbi / Forward / Backward shapes: transitions in registers (HOIST) for Q <= 16, in LDS
// for Q >= 16; the autotuner picks among them per model length (and MSV's wider menu).
#define SXK_DP3(ST)                                                                             \
    SXK_ALLG(ST, 4, false, true), SXK_ALLG(ST, 8, false, true), SXK_ALLG(ST, 12, false, true),  \
        SXK_ALLG(ST, 16, false, true),                                                          \
        SXK_ALLG(ST, 16, false, false), SXK_ALLG(ST, 32, false, false), SXK_ALLG(ST, 48, false, false), \
        SXK_E(ST, 64, 1, false, false), SXK_E(ST, 64, 2, false, false), SXK_E(ST, 64, 4, false, false)

#define SXK_GEM(ST)                                                                                        \
    SXK_E(ST, 24, 64, false, 2), SXK_E(ST, 32, 64, false, 2), SXK_E(ST, 40, 64, false, 2),                \
        SXK_E(ST, 48, 64, false, 2), SXK_E(ST, 64, 64, false, 2)
#define SXK_GEMS(ST)                                                                                       \
    SXK_E(ST, 24, 64, true, 2), SXK_E(ST, 32, 64, true, 2), SXK_E(ST, 40, 64, true, 2),                   \
        SXK_E(ST, 48, 64, true, 2), SXK_E(ST, 64, 64, true, 2)

#define SXK_ALN(G, Q) SXK_E(0, Q, G, false, sxk::TM_ALN), SXK_E(0, Q, G, false, sxk::TM_ALN1K)
#define SXK_QALN(G, Q) SXK_E(4, Q, G, false, sxk::TM_ALN), SXK_E(4, Q, G, false, sxk::TM_ALN1K)

static const SxKernelEntry k_table[] = {
    // MSV: one f32 row per lane
    SXK_ALLG(0, 4, false, false), SXK_ALLG(0, 8, false, false), SXK_ALLG(0, 16, false, false),
    SXK_ALLG(0, 32, false, false), SXK_ALLG(0, 64, false, false),
    SXK_E(0, 48, 1, false, false), SXK_E(0, 96, 1, false, false), SXK_E(0, 128, 1, false, false),
    SXK_E(0, 160, 1, false, false), SXK_E(0, 192, 1, false, false), SXK_E(0, 256, 1, false, false),
    SXK_E(0, 48, 2, false, false), SXK_E(0, 96, 2, false, false), SXK_E(0, 128, 2, false, false),
    SXK_E(0, 48, 4, false, false),
    // finer capacities for the common model lengths
    SXK_E(0, 40, 1, false, false), SXK_E(0, 56, 1, false, false), SXK_E(0, 72, 1, false, false),
    SXK_E(0, 80, 1, false, false), SXK_E(0, 112, 1, false, false), SXK_E(0, 144, 1, false, false),
    SXK_E(0, 176, 1, false, false), SXK_E(0, 208, 1, false, false), SXK_E(0, 224, 1, false, false),
    SXK_E(0, 24, 4, false, false), SXK_E(0, 40, 4, false, false), SXK_E(0, 56, 4, false, false),
    SXK_E(0, 72, 4, false, false), SXK_E(0, 80, 4, false, false),
    SXK_E(0, 24, 8, false, false), SXK_E(0, 40, 8, false, false), SXK_E(0, 48, 8, false, false),
    SXK_E(0, 56, 8, false, false),
    SXK_E(0, 12, 16, false, false), SXK_E(0, 20, 16, false, false), SXK_E(0, 24, 16, false, false),
    SXK_E(0, 40, 16, false, false), SXK_E(0, 48, 16, false, false),
    SXK_E(0, 24, 32, false, false), SXK_E(0, 40, 32, false, false), SXK_E(0, 48, 32, false, false),
    // aligned lane blocks, 512- and 1024-thread workgroups
    SXK_ALN(16, 12), SXK_ALN(16, 16), SXK_ALN(16, 20), SXK_ALN(16, 24), SXK_ALN(16, 32), SXK_ALN(16, 40),
    SXK_ALN(16, 48), SXK_E(0, 64, 16, false, sxk::TM_ALN), SXK_ALN(32, 16), SXK_ALN(32, 24), SXK_ALN(32, 32),
    SXK_ALN(8, 16), SXK_ALN(8, 20), SXK_ALN(8, 24), SXK_ALN(8, 32), SXK_ALN(8, 40), SXK_ALN(8, 48),
    // Viterbi, Forward, Backward: three rows per lane
    SXK_DP3(1), SXK_DP3(2), SXK_DP3(3),
    // Viterbi, finer capacities between the Q = 8 / 12 / 16 ones (transitions in registers)
    SXK_E(1, 10, 4, false, 1), SXK_E(1, 10, 8, false, 1), SXK_E(1, 10, 16, false, 1), SXK_E(1, 10, 32, false, 1),
    SXK_E(1, 14, 4, false, 1), SXK_E(1, 14, 8, false, 1), SXK_E(1, 14, 16, false, 1), SXK_E(1, 14, 32, false, 1),
    // Viterbi over aligned lane blocks, 512-thread workgroups
    SXK_E(1, 8, 8, false, sxk::TM_ALN), SXK_E(1, 12, 8, false, sxk::TM_ALN), SXK_E(1, 16, 8, false, sxk::TM_ALN),
    SXK_E(1, 8, 16, false, sxk::TM_ALN), SXK_E(1, 12, 16, false, sxk::TM_ALN), SXK_E(1, 16, 16, false, sxk::TM_ALN),
    SXK_E(1, 24, 16, false, sxk::TM_ALN), SXK_E(1, 12, 32, false, sxk::TM_ALN), SXK_E(1, 16, 32, false, sxk::TM_ALN),
    SXK_E(1, 24, 32, false, sxk::TM_ALN), SXK_E(1, 32, 32, false, sxk::TM_ALN),
    // special-row writers for the host-side domain definition
    SXK_ALLG(2, 4, true, true), SXK_ALLG(2, 8, true, true), SXK_ALLG(2, 16, true, false),
    SXK_ALLG(2, 32, true, false), SXK_ALLG(2, 48, true, false),
    SXK_ALLG(3, 4, true, true), SXK_ALLG(3, 8, true, true), SXK_ALLG(3, 16, true, false),
    SXK_ALLG(3, 32, true, false), SXK_ALLG(3, 48, true, false),
    // long models (M > 1024): 64 lanes per pair, emission rows from global memory
    SXK_GEM(0), SXK_GEM(1), SXK_GEM(2), SXK_GEM(3),
    // Viterbi and Forward of models of 513-1024 positions through the global-emission form
    SXK_E(1, 12, 64, false, 2), SXK_E(1, 16, 64, false, 2), SXK_E(2, 12, 64, false, 2), SXK_E(2, 16, 64, false, 2),
    SXK_GEMS(2), SXK_GEMS(3),
    // quantised MSV bound (stage 4): one-lane groups, and aligned lane blocks of 8, 16, 32
    SXK_E(4, 16, 1, false, 0), SXK_E(4, 24, 1, false, 0), SXK_E(4, 32, 1, false, 0), SXK_E(4, 48, 1, false, 0),
    SXK_E(4, 64, 1, false, 0), SXK_E(4, 80, 1, false, 0), SXK_E(4, 96, 1, false, 0), SXK_E(4, 128, 1, false, 0),
    SXK_E(4, 160, 1, false, 0), SXK_E(4, 192, 1, false, 0), SXK_E(4, 256, 1, false, 0),
    // finer one-lane capacities: the bound's time follows Q, not M
    SXK_E(4, 112, 1, false, 0), SXK_E(4, 144, 1, false, 0), SXK_E(4, 176, 1, false, 0), SXK_E(4, 208, 1, false, 0),
    SXK_E(4, 224, 1, false, 0), SXK_E(4, 240, 1, false, 0), SXK_E(4, 288, 1, false, 0), SXK_E(4, 320, 1, false, 0),
    SXK_QALN(8, 16), SXK_QALN(8, 24), SXK_QALN(8, 32), SXK_QALN(8, 40), SXK_QALN(8, 48), SXK_QALN(8, 56),
    SXK_QALN(8, 64), SXK_QALN(8, 72), SXK_QALN(8, 80), SXK_E(4, 88, 8, false, sxk::TM_ALN),
    SXK_E(4, 96, 8, false, sxk::TM_ALN), SXK_E(4, 128, 8, false, sxk::TM_ALN),
    SXK_QALN(16, 16), SXK_QALN(16, 24), SXK_QALN(16, 32), SXK_QALN(16, 40), SXK_QALN(16, 48), SXK_QALN(16, 56),
    SXK_QALN(16, 64),
    SXK_QALN(32, 16), SXK_QALN(32, 24), SXK_QALN(32, 32),
    // four-lane groups (four table copies) for 257-512 positions
    SXK_QALN(4, 72), SXK_QALN(4, 80), SXK_QALN(4, 88), SXK_QALN(4, 96), SXK_QALN(4, 104), SXK_QALN(4, 112),
    SXK_QALN(4, 128),
    // quantised Viterbi bound (stage 5): one-lane groups, aligned lane blocks of 8, 16, 32
    SXK_E(5, 16, 1, false, 0), SXK_E(5, 32, 1, false, 0),
    SXK_E(5, 16, 8, false, sxk::TM_ALN), SXK_E(5, 24, 8, false, sxk::TM_ALN), SXK_E(5, 32, 8, false, sxk::TM_ALN),
    SXK_E(5, 16, 16, false, sxk::TM_ALN), SXK_E(5, 24, 16, false, sxk::TM_ALN), SXK_E(5, 32, 16, false, sxk::TM_ALN),
    SXK_E(5, 16, 32, false, sxk::TM_ALN), SXK_E(5, 24, 32, false, sxk::TM_ALN), SXK_E(5, 32, 32, false, sxk::TM_ALN),
};

extern "C" const SxKernelEntry* sx_kernel_table(int* n) {
    *n = (int)(sizeof(k_table) / sizeof(k_table[0]));
    return k_table;
}

// sx_dev.h -- device-side data layout shared by the HIP kernels and the host runtime.
//
// Profiles (configured as in sentinel-x src/hmmer/model.rs:146-245) live in HBM packed
// back to back:
//   msc, mo : NCODE x (M+1) floats each, residue-major (msc[x*(M+1)+k]), column 0 = -inf / 0
//   tsc, tp : 8 x (M+1) floats each, rows MM MI MD IM II DM DD BM (t_xx[k] = edge out of node k)
// Targets (digitised as alphabet.rs:58-93) are concatenated, each starting on a 16-byte
// boundary so a group of lanes can fetch 16 residues with one 128-bit load; the buffer is
// padded by 32 bytes past the last target.  Per target the host precomputes the
// length-model scalars with glibc (model.rs:284-312, dp.rs:98-100), so device code never
// evaluates a logarithm on a bit-exact path.
#pragma once
#include <stdint.h>

#define SX_NCODE 25
#define SX_CODE_STAR 24
#define SX_NT 8
enum { SX_TMM = 0, SX_TMI, SX_TMD, SX_TIM, SX_TII, SX_TDM, SX_TDD, SX_TBM };

// Quantised MSV upper bound (the int16 prefilter of the MSV stage).  Emission scores
// are int16 in units of 1/SX_Q_EMIT nat, rounded up; the special states N, J, B and
// the length-model scores are int32 in units of 2^-SX_Q_SHIFT emission units
// (2^-22 nat), rounded up.  Each rounding adds the largest f32 rounding error of the
// value it feeds (<= 512 nats for a cell, <= 64 nats for N, J and B: true for models
// up to SX_Q_MAX_M and targets up to SX_Q_MAX_L), so the integer recurrence bounds the
// f32 MSV of dp.rs:92-145 from above until a cell reaches the top of the int16 range.
#define SX_Q_EMIT 1024.0
#define SX_Q_SHIFT 12
#define SX_Q_FINE 4194304.0      // SX_Q_EMIT * 2^SX_Q_SHIFT units per nat
#define SX_Q_SLACK_EMIT 0.03125  // SX_Q_EMIT * 2^-24 * 512
#define SX_Q_SLACK_FINE 16.0     // SX_Q_FINE * 2^-24 * 64
#define SX_Q_MAX_L 1000000       // longer targets are candidates of every profile
#define SX_QCAND_ITEMS 16        // pairs a thread tests in the candidate compaction
#define SX_Q_NOPASS (INT64_MAX / 8)  // bound thresholds at or above this: no score passes

// Per-target length model: MSV tloop/tmove (f32 logf, dp.rs:99-100), Plan7 multihit
// loop/move in nats (model.rs:286-289) and as probabilities (XscP), and the MSV
// tloop/tmove rounded up in the quantised bound's fine units.
struct SxLen {
    float msv_tloop, msv_tmove;
    float lloop, lmove;
    float ploop, pmove;
    int32_t q_tloop, q_tmove;
};

struct SxProfiles {
    const float* msc;
    const float* mo;
    const float* tsc;
    const float* tp;
    const int64_t* off;    // per profile: offset of its NCODE x (M+1) block in msc / mo
    const int64_t* toff;   // per profile: offset of its 8 x (M+1) block in tsc / tp
    const int32_t* m;
    const float* msv_tbm;
    // per profile: 1 if every M->D and D->D score is <= 0 (true of any probability
    // model), so that no D state can exceed the best M state of its row and Viterbi's
    // row maximum xE can skip the D row; 0 otherwise (then D is included, as written)
    const int32_t* vit_m_only;
    // Long models (M > SX_GEM_MIN_M): padded copies of the emission rows read straight
    // from global memory.  Row x of profile p starts at gmsc/gmo + goff[p] + x * gw, and
    // position k sits at index k + 3; positions 0 and M+1 .. gw-4 hold the pad value
    // (-inf in gmsc, 0 in gmo).  goff[p] = -1 for the other profiles.
    const float* gmsc;
    const float* gmo;
    const int64_t* goff;
    int32_t gw;
    int32_t q_tej;  // quantised MSV bound: the E->J score -ln 2 rounded up in fine units
    // Quantised MSV bound: emission scores rounded up to int16 (same layout and offsets
    // as msc; -inf -> -32768) and msv_tbm rounded up in fine units, per profile.
    const int16_t* qsc;
    const int32_t* qtbm;
    // Quantised Viterbi bound: transition scores rounded up to int16 (same layout and
    // offsets as tsc; a score <= 0 stays <= 0; column 0, out of node 0, is -32768).
    const int16_t* qtsc;
};

// Models longer than this use the global-emission kernels (64 lanes per pair); the
// longest supported model is 64 x 64 = 4096 positions.
#define SX_GEM_MIN_M 1024
// Models longer than this get the padded global copy of their emission rows: the long
// models, and those of 513-1024 positions, whose Viterbi and Forward may also run the
// global-emission kernels (64 lanes, 12 or 16 positions a lane) when the autotuner finds
// them faster.
#define SX_GEM_COPY_MIN_M 512
#define SX_MAX_M 4096
#define SX_GEM_WIDTH 4104  // >= SX_MAX_M + 2 + 3, a multiple of 8
// Longest model of the quantised MSV bound (its largest shape capacity); longer models
// skip the bound and go straight to the f32 MSV.
#define SX_Q_MAX_M 1024

struct SxTargets {
    const uint8_t* codes;
    const int64_t* off;    // start of each target in codes (16-byte aligned)
    const int32_t* len;
    const SxLen* lm;
    // The same residues interleaved for one-lane groups over the length-sorted order:
    // sorted positions 64c .. 64c+63 form chunk c, and word k (residues 4k .. 4k+3) of
    // position 64c + j sits at ilv[ilv_off[c] + 64 k + j], so a wave's 64 lanes read 256
    // contiguous bytes per four rows.  A chunk holds 4 more words a lane than its longest
    // target needs (the stream's look-ahead).
    const uint32_t* ilv;
    const int64_t* ilv_off;
};

// A tile: one profile against `count` targets taken from tids[first .. first+count).
// Result j goes to out[out + j].
struct SxTile {
    int32_t prof;
    int32_t count;
    int64_t first;
    int64_t out;
};

// Work list for one launch.  Either explicit tiles, or the implicit all-pairs form:
// tile t -> profile profs[t / nchunks], targets tids[(t % nchunks) * chunk ...],
// output row profs[..] * out_stride.
struct SxWork {
    const SxTile* tiles;      // explicit tiles, or nullptr
    const int32_t* profs;     // implicit form: profile ids
    const int32_t* tids;      // target ids in processing order
    int64_t ntiles;
    int32_t nchunks;
    int32_t chunk;
    int32_t ntids;
    int32_t out_by_pos;       // implicit form: result index = profile * out_stride + position
                              // in tids (0: + target id)
    int32_t ilv;              // implicit form over the length-sorted order with chunks a
                              // multiple of 64: one-lane groups may read SxTargets::ilv
    int64_t out_stride;
    int32_t chunk_major;      // implicit form: tile t -> chunk t / nprofiles, profile t % nprofiles
                              // (every profile's longest targets first), else profile-major
    int32_t pad0;
    unsigned long long* head; // persistent-queue head; nullptr -> tile = blockIdx.x (+ base)
    int64_t tile_base;
    // Data-dependent tile range (device cascade): when set, the launch covers explicit
    // tiles [*dyn_lo, *dyn_hi), read from device memory at kernel start.
    const int64_t* dyn_lo;
    const int64_t* dyn_hi;
};

// Raw scaled special states per row, written for survivors that need the host-side
// domain definition (pipeline.rs:190-242): xN xB xE xJ xC as stored (scaled) f32
// values plus the row's cumulative ln scale in f64.  The host applies lnf() + ls with
// glibc exactly as dp.rs:284-292 / 372-377 do.
struct SxSpecialsOut {
    float* x5;        // 5 floats per row, rows 0..L of each pair, at offsets[pair]
    double* ls;       // 1 double per row
    const int64_t* row_off;  // per pair (pair = out index) first row offset
};

// Per-pair input of the region scan (sxr_regions): first row, length, the multihit
// length-model loop / move scores and move probability of Xsc(L) (model.rs:284-312), and
// the Forward score that stands in for a non-finite Backward score when has_fwd is set.
struct SxRegionIn {
    int64_t roff;
    int32_t len;
    float lloop, lmove, pmove, fwd;
    int32_t has_fwd;
};
// Per-pair result of the region scan.
struct SxRegionRes {
    float fsc, bsc, nexp;
    int32_t nreg;
};
// One region: rows i..j (1-based, inclusive) of pair `pair`, multi = 1 when the
// multi-domain test holds.
struct SxRegion {
    int32_t pair, i, j, multi;
};

// One envelope for the device rescoring (rescore_isolated_domain, pipeline.rs:545-591):
// rows i..j (1-based, inclusive) of target `target` against profile `prof`, scored with
// the unihit length model of the whole target (Xsc::new(L, false), model.rs:284-312):
// log-space loop/move scores and their probabilities (E->C = 1, E->J = 0).  null2 = 1
// asks for the null2 sums.  The job's matrices start at ws[wsoff] (4 x (j-i+2) x (M+1)
// floats, then 14 x (j-i+2) floats of per-row values) and its row log scales at
// dws[dwsoff] (2 x (j-i+2) doubles).
struct SxEnvJob {
    int32_t prof, target, i, j;
    int32_t null2;
    float lloop, lmove, ploop, pmove;
    int32_t pad;
    int64_t wsoff, dwsoff;
};
// Result of one envelope: the unihit Forward and Backward scores, status 0 (rescored),
// 1 (Forward not finite) or 2 (Backward not finite), the optimal-accuracy coordinates
// (oa = 1; rows relative to the envelope, 1-based) and the null2 odds of the 20 residues
// before the degenerate codes and the logarithm (dp.rs:1310-1350).
struct SxEnvRes {
    float envsc, bcksc;
    int32_t status, oa;
    int32_t iali, jali, kali, mali;
    float null2[20];
};

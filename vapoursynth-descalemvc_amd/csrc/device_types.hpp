// POD descriptors shared by the host executor and the gfx950 kernels.
// Passed by value as kernel arguments; all pointers are device pointers.
#pragma once

#include <cstdint>

namespace dsmvc::hip {

// Row-record width of the packed substitution coefficients.  Row i holds
//   [0 .. 6]   lld_d(i) = L(i, i-d) * D(i-d)   for d = 1..7 (zero past W)
//   [7]        1 / (D_i + eps)
//   [8 .. 14]  ul_d(i)  = L(i+d, i)           for d = 1..7 (zero past W)
// so one 64-byte scalar load fetches everything a wavefront needs for row i.
// Bands 8..15 (blurred wide kernels) use the same layout with 15 slots per
// side: 32-float rows, the diagonal at [15], upper entries at [16 .. 30].
constexpr int kMaxBand = 7;
constexpr int kMaxWideBand = 15;

constexpr int coef_stride(int w) { return w <= kMaxBand ? 16 : 32; }

struct DevAxis {
    std::int32_t n;      // destination length
    std::int32_t m;      // source length
    std::int32_t w;      // half bandwidth
    std::int32_t nnz;
    const std::uint32_t *off;   // [n + 1] CSR of A^T by destination element
    const std::int32_t *idx;    // [nnz]   source index, ascending per row
    const float *wt;            // [nnz]
    const float *coef;          // [n * coef_stride(w)]
    // Float64 route: weights and coefficient rows in double.  For plans that
    // retain FP64 factors (zform = 1) the forward rows hold unit-L entries
    // L(i, i-d) and the diagonal is applied after the recurrence (the
    // recurrence runs on z = L^-1 b); otherwise they are the promoted F32 rows.
    const double *wt64;
    const double *coef64;
    std::int32_t zform;
    std::int32_t pad_;
};

// Partition of each vector into C chunks of exactly L elements for the chunked
// substitution kernels.  The plan is padded to C * L rows with zero rows
// (empty gather, zero coefficients): a zero row solves to 0 and no real row
// refers to it, so padding changes no result.  Per padded row:
//   ell_idx / ell_w  [C*L][K]  the CSR row packed to K taps; short rows are
//                    padded at the FRONT with weight 0 (an FMA chain that starts
//                    with 0 * y + 0 stays +0, so the sum is unchanged)
//   coef             [C*L][16] the coefficient record (kMaxBand layout)
// Per chunk q: W x W carry transfer matrices (FP64-built, stored F32)
//   tf[q][r][c] = d y[b_q - 1 - r] / d y[a_q - 1 - c]   (forward)
//   tb[q][r][c] = d x[a_q + r]     / d x[b_q + c]       (backward)
// and, for the horizontal kernel, the first source column of each block of
// block_rows destination rows (blk_lo[q][b]); the block's taps all lie in
// [blk_lo, blk_lo + 32).  Blocks are 16 rows when every window fits, else 8.
struct DevChunks {
    std::int32_t chunks;
    std::int32_t length;
    std::int32_t ell_width;
    std::int32_t blocks;
    std::int32_t block_rows;
    std::int32_t pad_;
    const float *tf;
    const float *tb;
    const std::int32_t *ell_idx;
    const float *ell_w;
    const float *coef;
    const std::int32_t *blk_lo;
};

constexpr int kBlockRowsMin = 8;   // destination rows per staging block (rows kernel)
constexpr int kStageCols = 32;     // source columns staged per block

// Windowed partition for the independent-chunk kernels.  Chunk q owns the
// destination rows [qL, qL + L); its wavefront runs the forward recurrence
// over [qL - P, qL + L + P) from a zero state and the backward recurrence
// over [qL, qL + L + P) from a zero state.  L and P are fixed per band class
// (window_length / window_warmup); a plan uses the windowed kernels only when
// the influence of the wrong starting state, computed in FP64 from its F32
// factors, has decayed below 2^-28 of the carried values by the chunk's first
// (forward) and last (backward) row.  Rows are stored padded: padded row
// r <-> plan row r - P, with zero rows before 0 and from n on, so every chunk
// visits exactly L + 2P padded rows starting at padded row qL.  Per padded
// row, for the plan's band class W and tap width K:
//   fwd       [rows][window_fwd_words(W, K)] forward record: the K weights of
//             the row's tap run, the forward factors lld_1..lld_W and the
//             inverse diagonal (zero in padding rows and unused slots)
//   bwd       [rows][window_bwd_words(W)] backward record: ul_1..ul_W
//   tap_start [rows] first source sample of the row's tap run: the taps are
//             the K consecutive samples [start, start + K), weight 0 where the
//             row has no tap, inside its staging window (both block sizes)
//   blk_lo    [rows / 8]  first staged source sample of each 8-row block (its
//             taps lie in [lo, lo + 32)); even, so that integer samples can
//             be staged two at a time with aligned loads
//   blk_lo16  [rows / 16] the same for 16-row blocks, valid when block16 = 1
//             (every 16-row block's taps fit 32 samples)
// The records are read a few rows at a time with wide scalar loads, so the
// kernels wait for scalar memory once per batch of rows, not once per row.
struct DevWindow {
    std::int32_t n;
    std::int32_t m;
    std::int32_t length;      // L
    std::int32_t warmup;      // P
    std::int32_t chunks;
    std::int32_t ell_width;   // K
    std::int32_t block16;
    std::int32_t pad_;
    const float *fwd;
    const float *bwd;
    const std::int32_t *tap_start;
    const std::int32_t *blk_lo;
    const std::int32_t *blk_lo16;
};

constexpr int kWindowBlock = 8;    // padded rows per staging block (rows kernel)

struct DevConv {
    float in_offset;
    float in_scale;
    float out_scale;
    float out_offset;
    float out_max;
    std::int32_t pad_;
};

enum class Sample : std::int32_t { f32 = 0, u8 = 1, u16 = 2, f64 = 3 };

// ---- fused 2D route (explicit operators on MFMA) ----------------------------
// The operator of an axis, M = G^-1 A^T truncated to its band, in blocks of
// kOpBlock destination indices; block b covers source indices
// [kstart[b], kstart[b] + steps * kOpStep).  A workgroup owns kOpTileCols
// destination columns (kOpTileBlocks horizontal blocks) and a run of units of
// kOpUnit destination rows (two vertical blocks); it stages kOpSlab source
// rows at a time (at most kOpStageCols columns), multiplies them by the
// horizontal operator, keeps the results in a ring of kOpRing source rows in
// LDS and multiplies the ring by the vertical operator.
constexpr int kOpBlock = 16;
constexpr int kOpStep = 32;
constexpr int kOpMaxSteps = 4;
constexpr int kOpTileBlocks = 4;
constexpr int kOpTileCols = kOpBlock * kOpTileBlocks;
constexpr int kOpUnit = 2 * kOpBlock;
constexpr int kOpSlab = 32;
// Most units of a vertical operator that complete with one slab (near the
// plane's edges, and for near-identity plans, several units' K ranges start
// in the same slab).
constexpr int kOpMaxUnitsPerSlab = 4;
constexpr int kOpStageCols = 208;
constexpr int kOpRing = 192;

struct DevOperator {
    const std::uint32_t *data;   // [blocks][steps][parts][64][4] 16-bit operand pairs
    const std::int32_t *kstart;  // [blocks]
    std::int32_t n, m, blocks, steps;
};

struct OpJob {
    DevOperator h, v;
    const void *src;
    void *dst;
    long s_row, s_frame, d_row, d_frame;
    std::int32_t frames;
    std::int32_t tiles;           // ceil(n_h / kOpTileCols)
    std::int32_t units;           // ceil(n_v / kOpUnit)
    std::int32_t units_per_item;  // units of one workgroup
    std::int32_t segments;        // ceil(units / units_per_item)
    std::int32_t vector_loads;    // source rows and base allow 16-byte loads
    long first_item;
    DevConv conv;
    // Integer output: sample = rint(clamp(x * out_mul + out_add, 0, max)) for
    // the kernel's result x = (M_v Y M_h^T) / 256 in input codes, i.e. the
    // conversion's ((256 x - in_offset) * in_scale) * out_scale + out_offset.
    float out_mul, out_add;
    // F32 samples: the FP16 pass sets *flag when a sample is non-finite or
    // above `range` in magnitude (so that the horizontal results stay within
    // FP16's range); the exact pass runs only then.
    float range;
    // A unit run in a slab step may read ring rows the step's slab
    // overwrites (some unit spans more than kOpRing - 2 kOpSlab + 1 rows,
    // its last block's K range included): a barrier separates them.
    std::int32_t unit_barrier;
    int *flag;
};

constexpr long operator_items(const OpJob &j) {
    return static_cast<long>(j.frames) * j.tiles * j.segments;
}

// One axis pass of one plan over `frames` planes, as a unit of a
// (possibly multi-plan) windowed launch.  Items are (frame, segment, group
// of 64 vectors), the group varying fastest; a segment is `span` consecutive
// chunks, which one wavefront solves in sequence (the forward recurrence runs
// on from one chunk into the next, so only the segment's first chunk pays the
// forward warm-up).  Rows pass: vectors are rows, `src` holds samples of the
// launch's kind, `dst` F32.  Columns pass: vectors are columns, `src` F32,
// `dst` samples of the launch's kind.
struct WinJob {
    DevWindow plan;
    const void *src;
    long s_row, s_frame;
    void *dst;
    long d_row, d_frame;
    std::int32_t nvec;
    std::int32_t frames;
    std::int32_t groups;      // ceil(nvec / 64)
    std::int32_t span;        // chunks per segment; 0 or 1: one
    long first_item;          // items of the jobs before this one in its table
    DevConv conv;
};

constexpr int window_span(const WinJob &j) { return j.span > 1 ? j.span : 1; }
constexpr int window_segments(const WinJob &j) {
    return (j.plan.chunks + window_span(j) - 1) / window_span(j);
}
constexpr long window_items(const WinJob &j) {
    return static_cast<long>(j.frames) * window_segments(j) * j.groups;
}

// Refinement tables of a plan that retains FP64 data: the FP64 weights of
// A^T (same CSR as DevAxis) and the FP64 normal matrix G = A^T A as rows of
// 2w+1 entries G(i, i+d), d = -w..w (zero outside the matrix), each split
// into a high F32 part and the F32 rounding of the remainder.
struct DevRefine {
    std::int32_t n;
    std::int32_t w;
    const float *w_hi;
    const float *w_lo;
    const float *g_hi;     // [n][2w+1]
    const float *g_lo;
    // The same data in FP64 and tap-major for element-parallel kernels,
    // whose lanes run along consecutive elements: row i reads source samples
    // [first[i], first[i] + k) with weights t64[q * n + i] (0 past the row's
    // taps), and G(i, i - w + d) = b64[d * n + i] (0 outside G).
    std::int32_t m;
    std::int32_t k;
    const std::int32_t *first;  // [n]
    const double *t64;          // [k][n]
    const double *b64;          // [2w+1][n]
};

// Windowed-kernel shape per band class: chunk length L and warm-up P.
constexpr int window_length(int) { return 32; }
constexpr int window_warmup(int w) { return w <= 3 ? 24 : (w <= 5 ? 32 : 64); }
// Record sizes of the windowed tables (DevWindow) in words, and the rows one
// batch of scalar loads covers (at most 64 words, a divisor of 8).
constexpr int window_fwd_words(int w, int k) { return (k + w + 1 + 3) & ~3; }
// Tap widths of the windowed F32 kernels: 6 for band classes 1 and 3, whose
// kernels at moderate ratios read at most 6 samples a row, then 8 and 16.
constexpr int window_width_for(int band, int taps) {
    return taps <= 6 && band <= 3 ? 6 : (taps <= 8 ? 8 : (taps <= 16 ? 16 : 0));
}
constexpr int window_bwd_words(int w) { return w <= 3 ? 4 : 8; }
constexpr int window_fwd_batch(int w, int k) { return window_fwd_words(w, k) <= 16 ? 4 : 2; }
constexpr int window_bwd_batch(int w) { return 64 / window_bwd_words(w); }

// Windowed FP64 partition (the Float64 route of F32 planes): the same
// padded-row layout as DevWindow with FP64 records, a warm-up long enough
// for a 2^-40 decay check and chunks short enough that a lane keeps its
// L + P values as doubles in registers without spilling at two wavefronts
// per SIMD (L = 32 for band classes up to 3, 24 above).  Forward records hold the FP64 tap
// weights, forward factors and the diagonal multiplier of the Float64 route
// (unit-L factors when zform = 1: the recurrence carries z = L^-1 b and the
// diagonal is applied to the stored value only).  Staging uses 8-row blocks.
constexpr int window_length64(int w) { return w <= 3 ? 32 : 24; }
constexpr int window_warmup64(int w) { return w <= 3 ? 32 : (w <= 5 ? 48 : 64); }
constexpr int window_fwd_batch64(int w, int k) { return k + w + 1 <= 12 ? 2 : 1; }
constexpr int window_bwd_batch64(int w) { return 32 / window_bwd_words(w); }
struct DevWindow64 {
    std::int32_t n;
    std::int32_t m;
    std::int32_t length;      // L
    std::int32_t warmup;      // P
    std::int32_t chunks;
    std::int32_t ell_width;   // K
    std::int32_t zform;
    std::int32_t pad_;
    const double *fwd;        // [rows][window_fwd_words(W, K)]
    const double *bwd;        // [rows][window_bwd_words(W)]
    const std::int32_t *tap_start;
    const std::int32_t *blk_lo;   // [rows / 8]
};

// One windowed FP64 axis pass of one plan over `frames` planes, as a unit of
// a (possibly multi-plan) FP64 launch; the FP64 counterpart of WinJob.  Items
// are (frame, segment, group of 64 vectors), the group varying fastest; a
// segment is `span` consecutive chunks solved by one wavefront in sequence,
// as in WinJob.  Rows pass: `src` F32 or integer samples, `dst` the FP64
// intermediate; columns pass: `src` FP64, `dst` samples of the launch's kind.
struct Win64Job {
    DevWindow64 plan;
    const void *src;
    long s_row, s_frame;
    void *dst;
    long d_row, d_frame;
    std::int32_t nvec;
    std::int32_t frames;
    std::int32_t groups;      // ceil(nvec / 64)
    std::int32_t span;        // chunks per segment; 0 or 1: one
    long first_item;          // items of the jobs before this one in its table
    int *flag;                // set to 1 when a stored value is not finite
    DevConv conv;
};

constexpr int window64_span(const Win64Job &j) { return j.span > 1 ? j.span : 1; }
constexpr int window64_segments(const Win64Job &j) {
    return (j.plan.chunks + window64_span(j) - 1) / window64_span(j);
}
constexpr long window64_items(const Win64Job &j) {
    return static_cast<long>(j.frames) * window64_segments(j) * j.groups;
}

// The Float64 route's input check of one F32 plane batch: items are
// (frame, row); a non-finite sample sets the job's flag.
struct CheckJob {
    const float *src;
    long s_row, s_frame;
    std::int32_t rows, cols;
    std::int32_t frames, pad_;
    long first_item;
    int *flag;
};

// One windowed FP64 axis pass over `frames` planes: rows pass F32 or integer
// source to an FP64 intermediate, columns pass FP64 intermediate to F32 or
// integer samples (`conv`).  `flag` is set when a stored value is not finite.
struct Win64Args {
    DevWindow64 plan;
    const void *src;
    long s_row, s_frame;
    void *dst;
    long d_row, d_frame;
    std::int32_t nvec;
    std::int32_t frames;
    std::int32_t groups;      // ceil(nvec / 64)
    std::int32_t pad_;
    int *flag;
    DevConv conv;
};

} // namespace dsmvc::hip

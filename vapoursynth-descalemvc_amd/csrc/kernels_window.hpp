// Windowed (independent-chunk) axis solves for gfx950, templated on the band
// class W, the tap width K and the sample type.  The chunk length L and the
// warm-up P are fixed per band class (window_length / window_warmup).
// Included by kernels_window_w{1,3,5,7}_k{6,8,16}.hip, one (band class, tap
// width) pair each, so the build compiles them in parallel.
//
// Every wavefront solves one item: 64 vectors (lanes) x one segment of
// `span` consecutive chunks of L destination rows of one plan.  It runs the
// forward recurrence from a zero state P rows before the segment, and on
// through the segment's chunks without a restart; after the forward sweep
// passes P rows beyond a chunk it runs the backward recurrence down from
// there from a zero state (the last L + P forward values are in registers),
// stores the chunk's L rows and carries the P values beyond the chunk into
// the next.  Chunks never exchange carries, so a workgroup's wavefronts are
// independent (no workgroup barrier) and one launch can hold items of many
// plans.  The four wavefronts of a workgroup take consecutive items
// (decode_item, xcd_block).
//
// Vertical pass (columns_window): lanes are 64 adjacent columns; each block
// of G rows stages the 32 source rows its taps need (256-byte runs) into the
// wavefront's LDS ring, the next block's loads in flight meanwhile.
// Horizontal pass (rows_window): lanes are 64 rows; each block of G (16 when
// the plan's windows allow it, else 8) rows stages the 32 source columns its
// taps need through the wavefront's LDS ring (transposed, pitch 33), the
// next block's loads in flight meanwhile, and results leave through the same
// ring.  Integer samples are staged two to a register and enter the solve as
// (float)x - offset, which is exact; the scale multiplies each result once
// instead of each sample (the solve is linear).
#pragma once

#include <cstdint>
#include <type_traits>

#include "kernels_common.hpp"

namespace dsmvc::hip {
namespace dev {
namespace {

// Per-wavefront LDS ring of the horizontal kernel: 64 rows (one per lane) of
// 32 staged source columns, row pitch 33 floats so that lanes reading the
// same column, and a row's 32 writers, hit 32 different banks.
constexpr int kRowPitch = kStageCols + 1;
constexpr int kWinRing = 64 * kRowPitch;

// A wave-uniform pointer passed through an empty asm that also reads
// `after`: loads through the result cannot be moved above the computation
// of `after`.
template <class T>
__device__ __forceinline__ ConstPtr<T> pinned(ConstPtr<T> p, float after) {
    auto v = reinterpret_cast<std::uintptr_t>(p);
    asm volatile("" : "+s"(v) : "v"(after));
    return reinterpret_cast<ConstPtr<T>>(v);
}

struct ItemPos {
    int frame, chunk, group;   // chunk: the first chunk of the item's segment
};

// Items run vector group fastest: a workgroup's four wavefronts usually
// solve the same segment of one plan for neighbouring vectors, so they read
// the same records through the scalar cache.  With bit 0 of `order` the
// segment varies fastest instead.
__device__ __forceinline__ ItemPos decode_item(const WinJob &jb, int local, int order) {
    const int segments = window_segments(jb), groups = jb.groups;
    ItemPos p;
    int segment;
    if (order & 1) {
        const int rest = local / segments;
        segment = local - rest * segments;
        p.frame = rest / groups;
        p.group = rest - p.frame * groups;
    } else {
        const int per_frame = segments * groups;
        p.frame = local / per_frame;
        const int r = local - p.frame * per_frame;
        segment = r / groups;
        p.group = r - segment * groups;
    }
    p.chunk = segment * window_span(jb);
    return p;
}

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD
// b % 8), each XCD with its own L2.  With bit 1 of `order`, blocks are
// renumbered so that every XCD receives a contiguous range of the item
// order: neighbouring chunks, whose source windows overlap, then share an L2.
constexpr int kXcds = 8;
__device__ __forceinline__ long xcd_block(long b, long blocks, int order) {
    if (!(order & 2)) return b;
    const long q = blocks / kXcds, r = blocks % kXcds;
    const long xcd = b % kXcds, slot = b / kXcds;
    return xcd < r ? xcd * (q + 1) + slot : r * (q + 1) + (xcd - r) * q + slot;
}

// Forward sweep of one batch of B rows whose taps are in tap[u][e]: each
// row's record (weights, factors, inverse diagonal) is one wide scalar load,
// all B issued together with the batch's LDS reads, so the wave waits once
// per batch.  The arithmetic is the ordered one: the gather FMA chain in
// ascending tap order, then the forward terms d = W..1, then the diagonal.
template <int W, int K, int B>
__device__ __forceinline__ void forward_batch(ConstPtr<float> rec, const float (&tap)[B][K],
                                              float (&st)[W], float (&out)[B]) {
    constexpr int RW = window_fwd_words(W, K);
    float c[B][K + W + 1];
#pragma unroll
    for (int u = 0; u < B; ++u)
#pragma unroll
        for (int i = 0; i < K + W + 1; ++i) c[u][i] = rec[u * RW + i];
#pragma unroll
    for (int u = 0; u < B; ++u) {
        float v = 0.0F;
#pragma unroll
        for (int e = 0; e < K; ++e) v = __builtin_fmaf(c[u][e], tap[u][e], v);
#pragma unroll
        for (int d = W; d >= 1; --d) v = __builtin_fmaf(-c[u][K + d - 1], st[d - 1], v);
        const float y = __fmul_rn(v, c[u][K + W]);
        push(st, y);
        out[u] = y;
    }
}

// Backward sweep over y[0, L + P) from the end, a batch of rows' backward
// records per group of scalar loads; record row of y[k] is k + P.  W == 3
// sums d ascending, other bands descending, as the reference solver does.
// The chunk's rows y[0, L) receive the solution; y[L, L + P) keep their
// forward values, from which the segment's next chunk continues.
template <int W, int L, int P>
__device__ __forceinline__ void backward_sweep(ConstPtr<float> bwd, float (&y)[L + P]) {
    constexpr int BW = window_bwd_words(W);
    constexpr int BB = window_bwd_batch(W);
    float st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0F;
#pragma unroll
    for (int top = L + P; top > 0; top -= BB) {
        const int count = top < BB ? top : BB;
        // The batch's records through a pointer pinned here, so that their
        // scalar loads stay with the batch: the records are invariant loads,
        // and hoisted all together they would exhaust the scalar registers.
        const auto rec = pinned(bwd, st[0]);
        float c[BB][W];
#pragma unroll
        for (int u = 0; u < BB; ++u)
            if (u < count)
#pragma unroll
                for (int d = 0; d < W; ++d) c[u][d] = rec[(top - 1 - u + P) * BW + d];
#pragma unroll
        for (int u = 0; u < BB; ++u) {
            if (u >= count) break;
            const int k = top - 1 - u;
            float v = y[k];
            if constexpr (W == 3) {
#pragma unroll
                for (int d = 1; d <= W; ++d) v = __builtin_fmaf(-c[u][d - 1], st[d - 1], v);
            } else {
#pragma unroll
                for (int d = W; d >= 1; --d) v = __builtin_fmaf(-c[u][d - 1], st[d - 1], v);
            }
            push(st, v);
            if (k < L) y[k] = v;
        }
    }
}

// Integer samples staged two to a 32-bit register: 16-bit halves for u16,
// the low two bytes for u8.  `both` loads an aligned pair, `one` a single
// sample, `join` packs two singles as `both` would.
template <class In> struct Pair;
template <> struct Pair<std::uint16_t> {
    __device__ static std::uint32_t both(__amdgpu_buffer_rsrc_t r, int v) {
        return __builtin_amdgcn_raw_buffer_load_b32(r, v, 0, 0);
    }
    __device__ static std::uint32_t one(__amdgpu_buffer_rsrc_t r, int v) {
        return __builtin_amdgcn_raw_buffer_load_b16(r, v, 0, 0);
    }
    __device__ static std::uint32_t join(std::uint32_t a, std::uint32_t b) { return a | (b << 16); }
    __device__ static float first(std::uint32_t w) { return static_cast<float>(w & 0xFFFFU); }
    __device__ static float second(std::uint32_t w) { return static_cast<float>(w >> 16); }
};
template <> struct Pair<std::uint8_t> {
    __device__ static std::uint32_t both(__amdgpu_buffer_rsrc_t r, int v) {
        return __builtin_amdgcn_raw_buffer_load_b16(r, v, 0, 0);
    }
    __device__ static std::uint32_t one(__amdgpu_buffer_rsrc_t r, int v) {
        return __builtin_amdgcn_raw_buffer_load_b8(r, v, 0, 0);
    }
    __device__ static std::uint32_t join(std::uint32_t a, std::uint32_t b) { return a | (b << 8); }
    __device__ static float first(std::uint32_t w) { return static_cast<float>(w & 0xFFU); }
    __device__ static float second(std::uint32_t w) { return static_cast<float>((w >> 8) & 0xFFU); }
};

// One converted sample into a plane through a buffer: `v` bytes of lane
// offset plus `s` bytes of wave-uniform offset.
template <class Out> struct BufferStore;
template <> struct BufferStore<float> {
    __device__ static void at(__amdgpu_buffer_rsrc_t r, int v, int s, float x, const DevConv &) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(std::uint32_t, x), r, v, s, 0);
    }
};
template <class T> __device__ __forceinline__ T to_sample(float v, const DevConv &c) {
    // Output conversion of Store<T>::f32 (kernels_common.hpp).
    float s = __fadd_rn(__fmul_rn(v, c.out_scale), c.out_offset);
    s = s < 0.0F ? 0.0F : (c.out_max < s ? c.out_max : s);
    return static_cast<T>(rintf(s));
}
template <> struct BufferStore<std::uint16_t> {
    __device__ static void at(__amdgpu_buffer_rsrc_t r, int v, int s, float x, const DevConv &c) {
        __builtin_amdgcn_raw_buffer_store_b16(to_sample<std::uint16_t>(x, c), r, v, s, 0);
    }
};
template <> struct BufferStore<std::uint8_t> {
    __device__ static void at(__amdgpu_buffer_rsrc_t r, int v, int s, float x, const DevConv &c) {
        __builtin_amdgcn_raw_buffer_store_b8(to_sample<std::uint8_t>(x, c), r, v, s, 0);
    }
};

template <int W, int K, int G, class In>
__device__ __forceinline__ void rows_item(const WinJob &jb, const ItemPos &at, float *ring,
                                          int lane) {
    constexpr int L = window_length(W);
    constexpr int P = window_warmup(W);
    constexpr int NBL = L / G;         // staging blocks per chunk
    constexpr int NBP = 2 * P / G;     // blocks before a chunk's own rows in its window
    constexpr int RW = window_fwd_words(W, K);
    constexpr bool kInt = !std::is_same_v<In, float>;
    // Rows per batch of scalar record loads; integer samples use 2 to make
    // room for their second staging buffer.
    constexpr int B = kInt ? 2 : window_fwd_batch(W, K);
    // Blocks in flight: integer samples take half the staging registers, so
    // two blocks are loaded ahead of the one being solved.
    constexpr int D = kInt ? 2 : 1;
    static_assert(L % G == 0 && (2 * P) % G == 0 && G % B == 0 && NBL % D == 0,
                  "blocks tile the window");
    const DevWindow &pl = jb.plan;
    const int nrows = jb.nvec;
    const int row0 = at.group * 64;
    const int q0 = at.chunk;
    const int q1 = min(q0 + window_span(jb), pl.chunks);
    const In *src = static_cast<const In *>(jb.src) + static_cast<long>(at.frame) * jb.s_frame;
    const auto rs = plane_rsrc(src, static_cast<long>(nrows) * jb.s_row * static_cast<long>(sizeof(In)));
    const int m = pl.m;
    const int s_row = static_cast<int>(jb.s_row);
    // Plan tables through the constant address space: the loads stay scalar
    // across the inline-asm pins below (which count as memory side effects).
    // Indexed by padded row (records, tap runs) and by block (staging starts).
    const auto ts = constant(pl.tap_start);
    const auto fwd = constant(pl.fwd);
    const auto blk = constant(G == 16 ? pl.blk_lo16 : pl.blk_lo);
    const DevConv cv = jb.conv;
    const float *mine = ring + lane * kRowPitch;   // this lane's row of the ring

    // Staging of a block's 32 source columns for the 64 rows: F32 samples
    // one per register (lane: column lane & 31 of rows 2t + lane / 32),
    // integer samples two per register (lane: columns 2 (lane & 15) + {0, 1}
    // of rows 4t + lane / 16).  Pairs are loaded whole when every pair is
    // aligned and inside its row (even width, stride and frame stride,
    // aligned base), else sample by sample.  Rows past the plane (last
    // vector group) read 0 through the buffer's range check; their results
    // are not stored.  The lane's address is rebuilt per block (pinned, so
    // the row offsets are not kept live).
    constexpr int NS = kInt ? 16 : 32;
    constexpr int kLoadRows = kInt ? 4 : 2;
    using Word = std::conditional_t<kInt, std::uint32_t, float>;
    Word staged[D][NS];
    const bool paired =
        kInt && uniform((m % 2) == 0 && (s_row % 2) == 0 && (jb.s_frame % 2) == 0
                        && (reinterpret_cast<std::uintptr_t>(jb.src) % (2 * sizeof(In))) == 0) != 0;
    const int scol = kInt ? 2 * (lane & 15) : (lane & 31);
    const int srow = kInt ? (lane >> 4) : (lane >> 5);
    const int row_step = kLoadRows * s_row * static_cast<int>(sizeof(In));
    // Blocks before padded row 0 hold zero rows: they load block 0's samples
    // (finite, never used) and skip the solve.
    auto fetch = [&](int gb, Word (&to)[NS]) {
        const int lo = blk[max(gb, 0)];
        const int rbase = (row0 + srow) * s_row;
        if constexpr (kInt) {
            if (paired) {
                int a = (rbase + min(lo + scol, m - 2)) * static_cast<int>(sizeof(In));
                asm volatile("" : "+v"(a));
#pragma unroll
                for (int t = 0; t < NS; ++t) to[t] = Pair<In>::both(rs, a + t * row_step);
            } else {
                int a0 = (rbase + min(lo + scol, m - 1)) * static_cast<int>(sizeof(In));
                int a1 = (rbase + min(lo + scol + 1, m - 1)) * static_cast<int>(sizeof(In));
                asm volatile("" : "+v"(a0), "+v"(a1));
#pragma unroll
                for (int t = 0; t < NS; ++t)
                    to[t] = Pair<In>::join(Pair<In>::one(rs, a0 + t * row_step),
                                           Pair<In>::one(rs, a1 + t * row_step));
            }
        } else {
            int a = (rbase + min(lo + scol, m - 1)) * 4;
            asm volatile("" : "+v"(a));
#pragma unroll
            for (int t = 0; t < NS; ++t) to[t] = BufferLoad<float>::at(rs, a + t * row_step, 0);
        }
    };
    // A staged block into the ring.  Pinned here, so the conversion (and
    // the wait for the loads) stays at this point instead of moving up to
    // the fetch.
    auto stage = [&](Word (&from)[NS]) {
#pragma unroll
        for (int t = 0; t < NS; ++t) {
            asm volatile("" : "+v"(from[t]));
            float *to = ring + (srow + kLoadRows * t) * kRowPitch + scol;
            if constexpr (kInt) {
                to[0] = __fsub_rn(Pair<In>::first(from[t]), cv.in_offset);
                to[1] = __fsub_rn(Pair<In>::second(from[t]), cv.in_offset);
            } else {
                to[0] = from[t];
            }
        }
    };

    float st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0F;
    float y[L + P];
#pragma unroll
    for (int k = 0; k < P; ++k) y[k] = 0.0F;
    // The segment's blocks run from `first` to `last`; block gb is staged
    // from buffer gb % D (the loop body holds a whole number of D blocks),
    // and requests block gb + D into that buffer once staged.  Row t of the
    // block is y[base + t] (rows before y[0] are warm-up, not kept).
    constexpr int WARM = (2 * P + L - 1) / L;
    const int first = (q0 - WARM) * NBL + NBP;
    const int last = (q1 - 1) * NBL + NBP + NBL - 1;
    auto block = [&](int gb, int base, Word (&buf)[NS]) {
        wave_sync();
        stage(buf);
        wave_sync();
        if (gb + D <= last) fetch(gb + D, buf);
        if (gb < 0) return;
        const int lo = blk[gb];
        const auto fr = fwd + static_cast<long>(gb) * G * RW;
#pragma unroll
        for (int j = 0; j < G; j += B) {
            // Each row's K taps are consecutive columns of this lane's ring
            // row: one address, K immediate offsets (ascending column = CSR
            // order).
            float tap[B][K];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const float *a = mine + (ts[gb * G + j + u] - lo);
#pragma unroll
                for (int e = 0; e < K; ++e) tap[u][e] = a[e];
            }
            float out[B];
            forward_batch<W, K, B>(pinned(fr + j * RW, st[0]), tap, st, out);
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (base + j + u >= 0) y[base + j + u] = out[u];
        }
    };

    const int n = pl.n;
    float *dst = static_cast<float *>(jb.dst) + static_cast<long>(at.frame) * jb.d_frame;
    const auto ws = plane_rsrc(dst, static_cast<long>(nrows) * jb.d_row * 4);
    const int d_row = static_cast<int>(jb.d_row);

    // Iteration q runs the forward sweep over padded rows [qL + 2P,
    // (q + 1)L + 2P), i.e. y[P, L + P) of chunk q, then (for q >= q0) the
    // chunk's backward sweep and stores.  The WARM iterations before q0 only
    // run the forward sweep: they cover the segment's warm-up (at least 2P
    // rows; blocks before padded row 0 are skipped).  One loop body for all
    // iterations keeps the register allocation of a single chunk.
#pragma unroll
    for (int d = 0; d < D; ++d) fetch(first + d, staged[d]);
    int gb = first;
    for (int q = q0 - WARM; q < q1; ++q) {
#pragma unroll
        for (int b = NBP; b < NBP + NBL; ++b, ++gb) block(gb, b * G - P, staged[(b - NBP) % D]);
        if (q >= q0) {
            backward_sweep<W, L, P>(constant(pl.bwd + static_cast<long>(q) * L * window_bwd_words(W)), y);

            // Results leave through the ring, 32 destination columns at a
            // time: lanes 0-31 write row t, lanes 32-63 row t + 1 (128-byte
            // runs).  Rows past the plane fall outside the store buffer's
            // range.  The lane's addresses are derived here from a pinned
            // copy of the lane index, so that they are not kept live (or
            // spilled) across the chunk loop.
            const int r0 = q * L;
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int oc = ln & 31;
            const int orow = ln >> 5;
            float *const my_row = ring + ln * kRowPitch;
#pragma unroll
            for (int kb = 0; kb < L; kb += 32) {
                wave_sync();
#pragma unroll
                for (int c = 0; c < 32; ++c)
                    if (kb + c < L) my_row[c] = kInt ? __fmul_rn(y[kb + c], cv.in_scale) : y[kb + c];
                wave_sync();
                if (kb + oc < L && r0 + kb + oc < n) {
                    const int a = ((row0 + orow) * d_row + r0 + kb + oc) * 4;
                    const float *from = ring + orow * kRowPitch + oc;
#pragma unroll 4
                    for (int t = 0; t < 64; t += 2)
                        BufferStore<float>::at(ws, a + t * d_row * 4, 0, from[t * kRowPitch], cv);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < P; ++k) y[k] = y[L + k];
    }
}

template <int W, int K, int G, class Out>
__device__ __forceinline__ void columns_item(const WinJob &jb, const ItemPos &at, float *ring,
                                             int lane) {
    constexpr int L = window_length(W);
    constexpr int P = window_warmup(W);
    constexpr int NBL = L / G;
    constexpr int NBP = 2 * P / G;
    constexpr int RW = window_fwd_words(W, K);
    constexpr int B = window_fwd_batch(W, K);
    static_assert(L % G == 0 && (2 * P) % G == 0 && G % B == 0, "blocks tile the window");
    static_assert(kStageCols * 64 <= kWinRing, "ring holds a block");
    const DevWindow &pl = jb.plan;
    const int ncols = jb.nvec;
    const int col = at.group * 64 + lane;
    const bool live = col < ncols;
    const int cc = live ? col : ncols - 1;
    const int q0 = at.chunk;
    const int q1 = min(q0 + window_span(jb), pl.chunks);
    const float *src = static_cast<const float *>(jb.src) + static_cast<long>(at.frame) * jb.s_frame;
    const auto rs = plane_rsrc(src, static_cast<long>(pl.m) * jb.s_row * 4);
    const int voff = cc * 4;
    const int row_bytes = static_cast<int>(jb.s_row * 4);
    const int m1 = pl.m - 1;
    const auto ts = constant(pl.tap_start);
    const auto fwd = constant(pl.fwd);
    const auto blk = constant(G == 16 ? pl.blk_lo16 : pl.blk_lo);
    const float *mine = ring + lane;   // this lane's column of the ring

    // Each block stages the 32 source rows its taps need (one 256-byte run
    // per row across the wavefront) into the ring, the next block's rows in
    // flight meanwhile; the taps then read the ring, each source row loaded
    // from memory once per block instead of once per tap.  Blocks before
    // padded row 0 load block 0's rows and skip the solve.
    float staged[32];
    auto fetch = [&](int gb) {
        const int lo = blk[max(gb, 0)];
#pragma unroll
        for (int t = 0; t < 32; ++t)
            staged[t] = BufferLoad<float>::at(rs, voff, min(lo + t, m1) * row_bytes);
    };

    float st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0F;
    float y[L + P];
#pragma unroll
    for (int k = 0; k < P; ++k) y[k] = 0.0F;
    constexpr int WARM = (2 * P + L - 1) / L;
    const int first = (q0 - WARM) * NBL + NBP;
    const int last = (q1 - 1) * NBL + NBP + NBL - 1;
    auto block = [&](int gb, int base) {
        wave_sync();
#pragma unroll
        for (int t = 0; t < 32; ++t) {
            asm volatile("" : "+v"(staged[t]));
            ring[t * 64 + lane] = staged[t];
        }
        wave_sync();
        if (gb < last) fetch(gb + 1);
        if (gb < 0) return;
        const int lo = blk[gb];
        const auto fr = fwd + static_cast<long>(gb) * G * RW;
#pragma unroll
        for (int j = 0; j < G; j += B) {
            float tap[B][K];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const float *row = mine + (ts[gb * G + j + u] - lo) * 64;
#pragma unroll
                for (int e = 0; e < K; ++e) tap[u][e] = row[e * 64];
            }
            float out[B];
            forward_batch<W, K, B>(pinned(fr + j * RW, st[0]), tap, st, out);
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (base + j + u >= 0) y[base + j + u] = out[u];
        }
    };

    // Results: row r0 + k of column cc, through a buffer over the output
    // plane (the row as the wave-uniform offset; rows past the plane are
    // skipped, columns past it masked).
    const int n = pl.n;
    Out *dst = static_cast<Out *>(jb.dst) + static_cast<long>(at.frame) * jb.d_frame;
    const auto ws = plane_rsrc(dst, static_cast<long>(n) * jb.d_row * static_cast<long>(sizeof(Out)));
    const int d_row_bytes = static_cast<int>(jb.d_row) * static_cast<int>(sizeof(Out));
    const int ooff = cc * static_cast<int>(sizeof(Out));
    const DevConv cv = jb.conv;

    // The iterations of rows_item: WARM forward-only iterations, then one
    // per chunk; the next block's rows are always requested before a
    // chunk's stores.
    fetch(first);
    int gb = first;
    for (int q = q0 - WARM; q < q1; ++q) {
#pragma unroll
        for (int b = NBP; b < NBP + NBL; ++b, ++gb) block(gb, b * G - P);
        if (q >= q0) {
            backward_sweep<W, L, P>(constant(pl.bwd + static_cast<long>(q) * L * window_bwd_words(W)), y);
            // Row offsets advance one row at a time through a pinned
            // register (precomputed for all L rows they would exhaust the
            // scalar registers).
            const int r0 = q * L;
            if (live) {
                const int rows = min(L, n - r0);
                int so = r0 * d_row_bytes;
#pragma unroll
                for (int k = 0; k < L; ++k) {
                    if (k < rows) BufferStore<Out>::at(ws, ooff, so, y[k], cv);
                    so += d_row_bytes;
                    asm volatile("" : "+s"(so));
                }
            }
        }
#pragma unroll
        for (int k = 0; k < P; ++k) y[k] = y[L + k];
    }
}

// A job record, read with scalar loads.
__device__ __forceinline__ WinJob load_job(const WinJob *p) {
    static_assert(sizeof(WinJob) % 4 == 0, "job record is whole words");
    struct Words {
        std::uint32_t v[sizeof(WinJob) / 4];
    } out;
    const auto words = constant(reinterpret_cast<const std::uint32_t *>(p));
#pragma unroll
    for (unsigned i = 0; i < sizeof(WinJob) / 4; ++i) out.v[i] = words[i];
    return __builtin_bit_cast(WinJob, out);
}

__device__ __forceinline__ int find_job(const WinJob *__restrict__ jobs, int count, long item) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (constant(jobs)[mid].first_item <= item) lo = mid;
        else hi = mid - 1;
    }
    return uniform(lo);
}

// One launch over a device table of jobs (one job or many): the grid covers
// every item of every job; wave w of workgroup b takes item b * kWindowWaves
// + w and finds its job by a binary search over the table.  Each wave solves
// one item and exits, so the hardware dispatcher balances plans of uneven
// size.
template <int W, int K, int G, class In>
__device__ __forceinline__ void rows_table_item(const WinJob *__restrict__ jobs, int count,
                                                long items, int order, float *lds) {
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const long item = xcd_block(blockIdx.x, gridDim.x, order) * kWindowWaves + wave;
    if (item >= items) return;
    const WinJob jb = load_job(jobs + find_job(jobs, count, item));
    rows_item<W, K, G, In>(jb, decode_item(jb, static_cast<int>(item - jb.first_item), order),
                           lds + wave * kWinRing, lane);
}

template <int W, int K, int G, class In>
__global__ void __launch_bounds__(64 * kWindowWaves)
rows_window_table(const WinJob *__restrict__ jobs, int count, long items, int order) {
    __shared__ float lds[kWindowWaves * kWinRing];
    rows_table_item<W, K, G, In>(jobs, count, items, order, lds);
}

// The same kernel held to 128 VGPRs, so that four wavefronts share each
// SIMD (the LDS ring allows four workgroups per CU).  Used where that fits
// without spilling much (see rows_four_waves).
template <int W, int K, int G, class In>
__global__ void __launch_bounds__(64 * kWindowWaves) __attribute__((amdgpu_waves_per_eu(4)))
rows_window_table4(const WinJob *__restrict__ jobs, int count, long items, int order) {
    __shared__ float lds[kWindowWaves * kWinRing];
    rows_table_item<W, K, G, In>(jobs, count, items, order, lds);
}

// 8-row blocks of u16 samples, and of band class 7 (64 warm-up rows), need
// more than 128 registers without spilling.
template <int W, class In, int G>
constexpr bool rows_four_waves = !(G == 8 && (W == 7 || std::is_same_v<In, std::uint16_t>));

template <int W, int K, int G, class Out>
__global__ void __launch_bounds__(64 * kWindowWaves)
columns_window_table(const WinJob *__restrict__ jobs, int count, long items, int order) {
    __shared__ float lds[kWindowWaves * kWinRing];
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const long item = xcd_block(blockIdx.x, gridDim.x, order) * kWindowWaves + wave;
    if (item >= items) return;
    const WinJob jb = load_job(jobs + find_job(jobs, count, item));
    columns_item<W, K, G, Out>(jb, decode_item(jb, static_cast<int>(item - jb.first_item), order),
                               lds + wave * kWinRing, lane);
}

// `block16`: every job of the table has 16-row staging blocks.
template <int W, int K, class T>
hipError_t launch_table(const WinJob *jobs, int count, long items, bool rows, bool block16,
                        int order, hipStream_t s) {
    if (items <= 0 || count <= 0) return hipSuccess;
    const long needed = (items + kWindowWaves - 1) / kWindowWaves;
    if (needed > 0x7FFFFFFFL) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(needed));
    if (!rows && block16)
        columns_window_table<W, K, 16, T><<<grid, 64 * kWindowWaves, 0, s>>>(jobs, count, items, order);
    else if (!rows)
        columns_window_table<W, K, 8, T><<<grid, 64 * kWindowWaves, 0, s>>>(jobs, count, items, order);
    else if (block16)
        rows_window_table4<W, K, 16, T><<<grid, 64 * kWindowWaves, 0, s>>>(jobs, count, items, order);
    else if constexpr (rows_four_waves<W, T, 8>)
        rows_window_table4<W, K, 8, T><<<grid, 64 * kWindowWaves, 0, s>>>(jobs, count, items, order);
    else
        rows_window_table<W, K, 8, T><<<grid, 64 * kWindowWaves, 0, s>>>(jobs, count, items, order);
    return hipGetLastError();
}

// Wavefronts of the launch's kernel the GPU holds at once (occupancy per CU
// times CUs), computed once per kernel.
template <class F> long resident_waves(F kernel) {
    int blocks = 0, device = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, 64 * kWindowWaves, 0) != hipSuccess
        || blocks <= 0)
        blocks = 1;
    if (hipGetDevice(&device) != hipSuccess
        || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess
        || cus <= 0)
        cus = 256;
    return static_cast<long>(blocks) * cus * kWindowWaves;
}

template <int W, int K, class T> long table_slots(bool rows, bool block16) {
    if (!rows && block16) {
        static const long v = resident_waves(columns_window_table<W, K, 16, T>);
        return v;
    }
    if (!rows) {
        static const long v = resident_waves(columns_window_table<W, K, 8, T>);
        return v;
    }
    if (block16) {
        static const long v = resident_waves(rows_window_table4<W, K, 16, T>);
        return v;
    }
    if constexpr (rows_four_waves<W, T, 8>) {
        static const long v = resident_waves(rows_window_table4<W, K, 8, T>);
        return v;
    } else {
        static const long v = resident_waves(rows_window_table<W, K, 8, T>);
        return v;
    }
}

} // namespace
} // namespace dev

template <int W, int K>
hipError_t launch_window_table_band(const WinJob *jobs, int count, long items, bool rows,
                                    bool block16, Sample kind, int order, hipStream_t s) {
    switch (kind) {
    case Sample::f32: return dev::launch_table<W, K, float>(jobs, count, items, rows, block16, order, s);
    case Sample::u8: return dev::launch_table<W, K, std::uint8_t>(jobs, count, items, rows, block16, order, s);
    case Sample::u16: return dev::launch_table<W, K, std::uint16_t>(jobs, count, items, rows, block16, order, s);
    default: return hipErrorInvalidValue;
    }
}

template <int W, int K>
long window_table_slots_band(bool rows, bool block16, Sample kind) {
    switch (kind) {
    case Sample::f32: return dev::table_slots<W, K, float>(rows, block16);
    case Sample::u8: return dev::table_slots<W, K, std::uint8_t>(rows, block16);
    case Sample::u16: return dev::table_slots<W, K, std::uint16_t>(rows, block16);
    default: return 0;
    }
}

} // namespace dsmvc::hip

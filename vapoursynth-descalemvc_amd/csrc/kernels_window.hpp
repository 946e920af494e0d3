// Windowed (independent-chunk) axis solves for gfx950, templated on the band
// class W, the tap width K and the sample type.  The chunk length L and the
// warm-up P are fixed per band class (window_length / window_warmup).
// Included by kernels_window_w{1,3,5,7}_k{8,16}.hip, one (band class, tap
// width) pair each, so the build compiles them in parallel.
//
// Every wavefront solves one item: 64 vectors (lanes) x one chunk of L
// destination rows of one plan.  It runs the forward recurrence over the
// L + 2P padded rows [qL - P, qL + L + P) from a zero state, keeps the last
// L + P values in registers, runs the backward recurrence down from the end
// of the window from a zero state, and stores the chunk's L rows.  Chunks
// never exchange carries, so a workgroup's wavefronts are independent (no
// workgroup barrier) and one launch can hold items of many plans.  The four
// wavefronts of a workgroup take consecutive items (decode_item, xcd_block).
//
// Vertical pass (columns_window): lanes are 64 adjacent columns; each block
// of G rows stages the 32 source rows its taps need (256-byte runs) into the
// wavefront's LDS ring, the next block's loads in flight meanwhile.
// Horizontal pass (rows_window): lanes are 64 rows; each block of G (16 when
// the plan's windows allow it, else 8) rows stages the 32 source columns its
// taps need through the wavefront's LDS ring (transposed, pitch 65), the next
// block's loads in flight meanwhile, and results leave through the same ring.
#pragma once

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

struct ItemPos {
    int frame, chunk, group;
};

// Items run vector group fastest: a workgroup's four wavefronts usually
// solve the same chunk of one plan for neighbouring vectors, so they read the
// same records through the scalar cache.  With bit 0 of `order` the chunk
// varies fastest instead.
__device__ __forceinline__ ItemPos decode_item(const WinJob &jb, int local, int order) {
    const int chunks = jb.plan.chunks, groups = jb.groups;
    ItemPos p;
    if (order & 1) {
        const int rest = local / chunks;
        p.chunk = local - rest * chunks;
        p.frame = rest / groups;
        p.group = rest - p.frame * groups;
    } else {
        const int per_frame = chunks * groups;
        p.frame = local / per_frame;
        const int r = local - p.frame * per_frame;
        p.chunk = r / groups;
        p.group = r - p.chunk * groups;
    }
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
        float c[BB][W];
#pragma unroll
        for (int u = 0; u < BB; ++u)
            if (u < count)
#pragma unroll
                for (int d = 0; d < W; ++d) c[u][d] = bwd[(top - 1 - u + P) * BW + d];
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
            y[k] = v;
        }
    }
}

template <int W, int K, int G, class In>
__device__ __forceinline__ void rows_item(const WinJob &jb, const ItemPos &at, float *ring,
                                          int lane) {
    constexpr int L = window_length(W);
    constexpr int P = window_warmup(W);
    constexpr int NB = (L + 2 * P) / G;
    constexpr int RW = window_fwd_words(W, K);
    constexpr int B = window_fwd_batch(W, K);
    static_assert((L + 2 * P) % G == 0 && L % G == 0 && G % B == 0, "blocks tile the window");
    const DevWindow &pl = jb.plan;
    const int nrows = jb.nvec;
    const int row0 = at.group * 64;
    const int r0 = at.chunk * L;
    const In *src = static_cast<const In *>(jb.src) + static_cast<long>(at.frame) * jb.s_frame;
    const auto rs = plane_rsrc(src, static_cast<long>(nrows) * jb.s_row * static_cast<long>(sizeof(In)));
    const int m1 = pl.m - 1;
    const int scol = lane & (kStageCols - 1);
    const int srow = lane >> 5;
    // Plan tables through the constant address space: the loads stay scalar
    // across the inline-asm pins below (which count as memory side effects).
    const auto ts = constant(pl.tap_start + r0);
    const auto fr = constant(pl.fwd + static_cast<long>(r0) * RW);
    const auto blk = constant((G == 16 ? pl.blk_lo16 : pl.blk_lo) + r0 / G);
    const DevConv cv = jb.conv;
    const float *mine = ring + lane * kRowPitch;   // this lane's row of the ring

    // The next block's 32 source columns are loaded into registers while the
    // current block is solved, so the loads' latency overlaps the solve.
    // Held as 32-bit words so they can be pinned in registers below.
    using Word = std::conditional_t<std::is_same_v<In, float>, float, std::uint32_t>;
    // Rows past the plane (last vector group) read 0 through the buffer's
    // range check; their results are not stored.  The lane's address is
    // rebuilt per block (pinned, so the 32 row offsets are not kept live).
    Word staged[32];
    const int row_step = 2 * static_cast<int>(jb.s_row) * static_cast<int>(sizeof(In));
    auto fetch = [&](int b) {
        const int column = min(blk[b] + scol, m1);
        int at = ((row0 + srow) * static_cast<int>(jb.s_row) + column) * static_cast<int>(sizeof(In));
        asm volatile("" : "+v"(at));
#pragma unroll
        for (int t = 0; t < 32; ++t)
            staged[t] = static_cast<Word>(BufferLoad<In>::at(rs, at + t * row_step, 0));
    };
    fetch(0);

    float st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0F;
    float y[L + P];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int lo = blk[b];
        int start[G];   // the block's tap runs, one scalar load
#pragma unroll
        for (int t = 0; t < G; ++t) start[t] = ts[b * G + t] - lo;
        wave_sync();
#pragma unroll
        for (int t = 0; t < 32; ++t) {
            // Pinned here, so the conversion (and the wait for the load)
            // stays at this point instead of moving up to the fetch.
            asm volatile("" : "+v"(staged[t]));
            ring[(srow + 2 * t) * kRowPitch + scol] = Load<In>::norm(static_cast<In>(staged[t]), cv);
        }
        wave_sync();
        if (b + 1 < NB) fetch(b + 1);
#pragma unroll
        for (int j = 0; j < G; j += B) {
            // Each row's K taps are consecutive columns of this lane's ring
            // row: one address, K immediate offsets (ascending column = CSR
            // order).
            float tap[B][K];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const float *a = mine + start[j + u];
#pragma unroll
                for (int e = 0; e < K; ++e) tap[u][e] = a[e];
            }
            float out[B];
            forward_batch<W, K, B>(fr + (b * G + j) * RW, tap, st, out);
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (b * G + j + u >= P) y[b * G + j + u - P] = out[u];
        }
    }
    backward_sweep<W, L, P>(constant(pl.bwd + static_cast<long>(r0) * window_bwd_words(W)), y);

    // Results leave through the ring, 32 destination columns at a time:
    // lanes 0-31 write row t, lanes 32-63 row t + 1 (128-byte runs).
    float *out = static_cast<float *>(jb.dst) + static_cast<long>(at.frame) * jb.d_frame + r0;
    const int oc = lane & 31;
    const int orow = lane >> 5;
    const int n = pl.n;
#pragma unroll
    for (int kb = 0; kb < L; kb += 32) {
        wave_sync();
#pragma unroll
        for (int c = 0; c < 32; ++c)
            if (kb + c < L) ring[lane * kRowPitch + c] = y[kb + c];
        wave_sync();
        const bool col_ok = kb + oc < L && r0 + kb + oc < n;
#pragma unroll 4
        for (int t = 0; t < 64; t += 2) {
            const int rr = t + orow;
            const int row = row0 + rr;
            if (col_ok && row < nrows)
                out[static_cast<long>(row) * jb.d_row + kb + oc] = ring[rr * kRowPitch + oc];
        }
    }
}

template <int W, int K, int G, class Out>
__device__ __forceinline__ void columns_item(const WinJob &jb, const ItemPos &at, float *ring,
                                             int lane) {
    constexpr int L = window_length(W);
    constexpr int P = window_warmup(W);
    constexpr int NB = (L + 2 * P) / G;
    constexpr int RW = window_fwd_words(W, K);
    constexpr int B = window_fwd_batch(W, K);
    static_assert((L + 2 * P) % G == 0 && L % G == 0 && G % B == 0, "blocks tile the window");
    static_assert(kStageCols * 64 <= kWinRing, "ring holds a block");
    const DevWindow &pl = jb.plan;
    const int ncols = jb.nvec;
    const int col = at.group * 64 + lane;
    const bool live = col < ncols;
    const int cc = live ? col : ncols - 1;
    const int r0 = at.chunk * L;
    const float *src = static_cast<const float *>(jb.src) + static_cast<long>(at.frame) * jb.s_frame;
    const auto rs = plane_rsrc(src, static_cast<long>(pl.m) * jb.s_row * 4);
    const int voff = cc * 4;
    const int row_bytes = static_cast<int>(jb.s_row * 4);
    const int m1 = pl.m - 1;
    const auto ts = constant(pl.tap_start + r0);
    const auto fr = constant(pl.fwd + static_cast<long>(r0) * RW);
    const auto blk = constant((G == 16 ? pl.blk_lo16 : pl.blk_lo) + r0 / G);
    const float *mine = ring + lane;   // this lane's column of the ring

    // Each block stages the 32 source rows its taps need (one 256-byte run
    // per row across the wavefront) into the ring, the next block's rows in
    // flight meanwhile; the taps then read the ring, each source row loaded
    // from memory once per block instead of once per tap.
    float staged[32];
    auto fetch = [&](int b) {
        const int lo = blk[b];
#pragma unroll
        for (int t = 0; t < 32; ++t)
            staged[t] = BufferLoad<float>::at(rs, voff, min(lo + t, m1) * row_bytes);
    };
    fetch(0);

    float st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0F;
    float y[L + P];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int lo = blk[b];
        int start[G];
#pragma unroll
        for (int t = 0; t < G; ++t) start[t] = ts[b * G + t] - lo;
        wave_sync();
#pragma unroll
        for (int t = 0; t < 32; ++t) {
            asm volatile("" : "+v"(staged[t]));
            ring[t * 64 + lane] = staged[t];
        }
        wave_sync();
        if (b + 1 < NB) fetch(b + 1);
#pragma unroll
        for (int j = 0; j < G; j += B) {
            float tap[B][K];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const float *row = mine + start[j + u] * 64;
#pragma unroll
                for (int e = 0; e < K; ++e) tap[u][e] = row[e * 64];
            }
            float out[B];
            forward_batch<W, K, B>(fr + (b * G + j) * RW, tap, st, out);
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (b * G + j + u >= P) y[b * G + j + u - P] = out[u];
        }
    }
    backward_sweep<W, L, P>(constant(pl.bwd + static_cast<long>(r0) * window_bwd_words(W)), y);

    Out *o = static_cast<Out *>(jb.dst) + static_cast<long>(at.frame) * jb.d_frame + cc;
    const int n = pl.n;
    const DevConv cv = jb.conv;
#pragma unroll
    for (int k = 0; k < L; ++k)
        if (r0 + k < n && live) Store<Out>::f32(o + static_cast<long>(r0 + k) * jb.d_row, y[k], cv);
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

} // namespace dsmvc::hip

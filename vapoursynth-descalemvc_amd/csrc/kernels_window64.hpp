// Windowed FP64 axis solves for the Float64 route of F32 planes (gfx950),
// templated on the band class W and the tap width K.  Included by
// kernels_window64_w{1,3,5,7}.hip, one band class each.
//
// The structure follows the F32 windowed kernels: every wavefront solves one
// item, 64 vectors (lanes) x one chunk of L destination rows of one plan,
// from a zero state with P warm-up rows on each side (DevWindow64: the plan's
// FP64 recurrence has decayed below 2^-40 over P rows).  The arithmetic is
// the Float64 route's (reference src/axis_plan.cpp:1124-1248, as the
// sequential FP64 kernel): FP64 gather FMAs in CSR order on the promoted
// samples, forward terms d = W..1, the diagonal multiplier, backward terms
// d = W..1; the carried forward state is z (unit-L FP64 factors) or y.
//
// Horizontal pass (rows_window64): lanes are 64 rows; each block of 8
// destination rows stages the 32 source columns its taps need through the
// wavefront's LDS ring (transposed, pitch 33), the next block's loads in
// flight meanwhile; the FP64 results leave through the same ring as pairs of
// 32-bit words, row-contiguous.  Vertical pass (columns_window64): lanes are
// 64 adjacent columns of the FP64 intermediate; each tap is one coalesced
// 512-byte load, the next batch's loads in flight while a batch is solved.
#pragma once

#include <type_traits>

#include "kernels_common.hpp"

namespace dsmvc::hip {
namespace dev {
namespace {

constexpr int kRowPitch64 = kStageCols + 1;
constexpr int kRing64 = 64 * kRowPitch64;   // 32-bit words per wavefront
constexpr int kWaves64 = 4;

struct Pos64 {
    int frame, chunk, group;
};

// Items run vector group fastest, then chunk, then frame.
__device__ __forceinline__ Pos64 decode64(const Win64Job &a, long item) {
    const long per_frame = static_cast<long>(a.plan.chunks) * a.groups;
    Pos64 p;
    p.frame = static_cast<int>(item / per_frame);
    const int r = static_cast<int>(item - static_cast<long>(p.frame) * per_frame);
    p.chunk = r / a.groups;
    p.group = r - p.chunk * a.groups;
    return p;
}

// Forward sweep of a batch of B rows whose taps are tap[u][e] (`sample`
// turns one into the Float64 route's FP64 sample): each row's FP64 record
// (weights, factors, diagonal multiplier) comes from scalar loads issued
// together for the batch.
template <int W, int K, int B, class Tap, class Sample64>
__device__ __forceinline__ void forward_batch64(ConstPtr<double> rec, const Tap (&tap)[B][K],
                                                Sample64 sample, bool zform, double (&st)[W],
                                                double (&out)[B]) {
    constexpr int RW = window_fwd_words(W, K);
    double c[B][K + W + 1];
#pragma unroll
    for (int u = 0; u < B; ++u)
#pragma unroll
        for (int i = 0; i < K + W + 1; ++i) c[u][i] = rec[u * RW + i];
#pragma unroll
    for (int u = 0; u < B; ++u) {
        double v = 0.0;
#pragma unroll
        for (int e = 0; e < K; ++e) v = __builtin_fma(c[u][e], sample(tap[u][e]), v);
#pragma unroll
        for (int d = W; d >= 1; --d) v = __builtin_fma(-c[u][K + d - 1], st[d - 1], v);
        const double y = __dmul_rn(v, c[u][K + W]);
        push(st, zform ? v : y);
        out[u] = y;
    }
}

// Backward sweep over y[0, L + P) from the end; record row of y[k] is k + P.
template <int W, int L, int P>
__device__ __forceinline__ void backward_sweep64(ConstPtr<double> bwd, double (&y)[L + P]) {
    constexpr int BW = window_bwd_words(W);
    constexpr int BB = window_bwd_batch64(W);
    double st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0;
#pragma unroll
    for (int top = L + P; top > 0; top -= BB) {
        const int count = top < BB ? top : BB;
        double c[BB][W];
#pragma unroll
        for (int u = 0; u < BB; ++u)
            if (u < count)
#pragma unroll
                for (int d = 0; d < W; ++d) c[u][d] = bwd[(top - 1 - u + P) * BW + d];
#pragma unroll
        for (int u = 0; u < BB; ++u) {
            if (u >= count) break;
            const int k = top - 1 - u;
            double v = y[k];
#pragma unroll
            for (int d = W; d >= 1; --d) v = __builtin_fma(-c[u][d - 1], st[d - 1], v);
            push(st, v);
            y[k] = v;
        }
    }
}

// Returns false when a stored value is not finite.  Integer samples are
// staged exactly as F32 and normalised per tap in FP64, as the Float64
// route's load does: (x - offset) * scale.
template <int W, int K, class In>
__device__ __forceinline__ bool rows64_item(const Win64Job &a, const Pos64 &at,
                                            std::uint32_t *ring, int lane) {
    constexpr int L = window_length64(W);
    constexpr int P = window_warmup64(W);
    constexpr int G = kWindowBlock;
    constexpr int NB = (L + 2 * P) / G;
    constexpr int RW = window_fwd_words(W, K);
    constexpr int B = window_fwd_batch64(W, K);
    static_assert((L + 2 * P) % G == 0 && G % B == 0, "blocks tile the window");
    const DevWindow64 &pl = a.plan;
    const int nrows = a.nvec;
    const int row0 = at.group * 64;
    const int r0 = at.chunk * L;
    const In *src = static_cast<const In *>(a.src) + static_cast<long>(at.frame) * a.s_frame;
    const auto rs = plane_rsrc(src, static_cast<long>(nrows) * a.s_row * static_cast<long>(sizeof(In)));
    const double in_offset = a.conv.in_offset, in_scale = a.conv.in_scale;
    auto sample = [&](float x) {
        if constexpr (std::is_same_v<In, float>) return static_cast<double>(x);
        else return __dmul_rn(__dsub_rn(static_cast<double>(x), in_offset), in_scale);
    };
    const int m1 = pl.m - 1;
    const int scol = lane & (kStageCols - 1);
    const int srow = lane >> 5;
    const auto ts = constant(pl.tap_start + r0);
    const auto fr = constant(pl.fwd + static_cast<long>(r0) * RW);
    const auto blk = constant(pl.blk_lo + r0 / G);
    const bool zform = pl.zform != 0;
    const float *mine = reinterpret_cast<const float *>(ring) + lane * kRowPitch64;
    float *fring = reinterpret_cast<float *>(ring);

    using Word = std::conditional_t<std::is_same_v<In, float>, float, std::uint32_t>;
    Word staged[32];
    auto fetch = [&](int b) {
        const int column = min(blk[b] + scol, m1);
#pragma unroll
        for (int t = 0; t < 32; ++t) {
            const int row = min(row0 + srow + 2 * t, nrows - 1);
            staged[t] = static_cast<Word>(BufferLoad<In>::at(
                rs, (row * static_cast<int>(a.s_row) + column) * static_cast<int>(sizeof(In)), 0));
        }
    };
    fetch(0);

    double st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0;
    double y[L + P];
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
            fring[(srow + 2 * t) * kRowPitch64 + scol] = static_cast<float>(staged[t]);
        }
        wave_sync();
        if (b + 1 < NB) fetch(b + 1);
#pragma unroll
        for (int j = 0; j < G; j += B) {
            float tap[B][K];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const float *p = mine + start[j + u];
#pragma unroll
                for (int e = 0; e < K; ++e) tap[u][e] = p[e];
            }
            double out[B];
            forward_batch64<W, K, B>(fr + (b * G + j) * RW, tap, sample, zform, st, out);
#pragma unroll
            for (int u = 0; u < B; ++u)
                if (b * G + j + u >= P) y[b * G + j + u - P] = out[u];
        }
    }
    backward_sweep64<W, L, P>(constant(pl.bwd + static_cast<long>(r0) * window_bwd_words(W)), y);
    bool finite = true;
#pragma unroll
    for (int k = 0; k < L; ++k) finite = finite && isfinite(y[k]);

    // FP64 results leave through the ring 16 at a time as 32-bit word pairs:
    // lanes 0-31 write words of row t, lanes 32-63 of row t + 1 (128 bytes).
    auto *out = static_cast<std::uint32_t *>(a.dst) + 2 * (static_cast<long>(at.frame) * a.d_frame + r0);
    const int oc = lane & 31;
    const int orow = lane >> 5;
    const int n = pl.n;
#pragma unroll
    for (int kb = 0; kb < L; kb += 16) {
        wave_sync();
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            if (kb + c >= L) break;
            const auto bits = __builtin_bit_cast(std::uint64_t, y[kb + c]);
            ring[lane * kRowPitch64 + 2 * c] = static_cast<std::uint32_t>(bits);
            ring[lane * kRowPitch64 + 2 * c + 1] = static_cast<std::uint32_t>(bits >> 32);
        }
        wave_sync();
        const int col = kb + (oc >> 1);
        const bool col_ok = col < L && r0 + col < n;
#pragma unroll 4
        for (int t = 0; t < 64; t += 2) {
            const int rr = t + orow;
            const int row = row0 + rr;
            if (col_ok && row < nrows)
                out[2 * static_cast<long>(row) * a.d_row + 2 * kb + oc] = ring[rr * kRowPitch64 + oc];
        }
    }
    return finite;
}

template <int W, int K, class Out>
__device__ __forceinline__ bool columns64_item(const Win64Job &a, const Pos64 &at, int lane) {
    constexpr int L = window_length64(W);
    constexpr int P = window_warmup64(W);
    constexpr int R = L + 2 * P;
    constexpr int RW = window_fwd_words(W, K);
    constexpr int B = window_fwd_batch64(W, K);
    static_assert(R % B == 0, "batches tile the window");
    const DevWindow64 &pl = a.plan;
    const int ncols = a.nvec;
    const int col = at.group * 64 + lane;
    const bool live = col < ncols;
    const int cc = live ? col : ncols - 1;
    const int r0 = at.chunk * L;
    const double *src = static_cast<const double *>(a.src) + static_cast<long>(at.frame) * a.s_frame;
    const auto rs = plane_rsrc(src, static_cast<long>(pl.m) * a.s_row * 8);
    const int voff = cc * 8;
    const int row_bytes = static_cast<int>(a.s_row * 8);
    const int m1 = pl.m - 1;
    const auto ts = constant(pl.tap_start + r0);
    const auto fr = constant(pl.fwd + static_cast<long>(r0) * RW);
    const bool zform = pl.zform != 0;

    // Taps of batch j: K consecutive source rows per destination row, each a
    // coalesced load across the wavefront (rows past the plane clamp to its
    // last row; their weights are 0).
    double cur[B][K], nxt[B][K];
    auto fetch = [&](int j, double (&t)[B][K]) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int s0 = ts[j + u];
#pragma unroll
            for (int e = 0; e < K; ++e) {
                const auto w = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, min(s0 + e, m1) * row_bytes, 0);
                t[u][e] = __builtin_bit_cast(double, w);
            }
        }
    };
    fetch(0, cur);

    double st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0;
    double y[L + P];
#pragma unroll
    for (int j = 0; j < R; j += B) {
        if (j + B < R) fetch(j + B, nxt);
        double out[B];
        forward_batch64<W, K, B>(fr + j * RW, cur, [](double x) { return x; }, zform, st, out);
#pragma unroll
        for (int u = 0; u < B; ++u) {
            if (j + u >= P) y[j + u - P] = out[u];
#pragma unroll
            for (int e = 0; e < K; ++e) cur[u][e] = nxt[u][e];
        }
    }
    backward_sweep64<W, L, P>(constant(pl.bwd + static_cast<long>(r0) * window_bwd_words(W)), y);

    Out *o = static_cast<Out *>(a.dst) + static_cast<long>(at.frame) * a.d_frame + cc;
    const int n = pl.n;
    const DevConv cv = a.conv;
    bool finite = true;
#pragma unroll
    for (int k = 0; k < L; ++k)
        if (r0 + k < n && live)
            finite = Store<Out>::f64(o + static_cast<long>(r0 + k) * a.d_row, y[k], cv) && finite;
    return finite;
}

// A job record from the device table, read with scalar loads.
__device__ __forceinline__ Win64Job load_job64(const Win64Job *p) {
    static_assert(sizeof(Win64Job) % 4 == 0, "job record is whole words");
    struct Words {
        std::uint32_t v[sizeof(Win64Job) / 4];
    } out;
    const auto words = constant(reinterpret_cast<const std::uint32_t *>(p));
#pragma unroll
    for (unsigned i = 0; i < sizeof(Win64Job) / 4; ++i) out.v[i] = words[i];
    return __builtin_bit_cast(Win64Job, out);
}

__device__ __forceinline__ int find_job64(const Win64Job *__restrict__ jobs, int count, long item) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (constant(jobs)[mid].first_item <= item) lo = mid;
        else hi = mid - 1;
    }
    return uniform(lo);
}

// One launch over a device table of jobs (many plans of one band class, tap
// width, direction and sample kind) or, when `jobs` is null, over the one
// job passed by value: wave w of workgroup b takes item b * kWaves64 + w and
// finds its job by a binary search over the table.  A non-finite stored
// value sets the job's flag.
template <int W, int K, class In>
__global__ void __launch_bounds__(64 * kWaves64)
rows_window64(const Win64Job *__restrict__ jobs, int count, long items, Win64Job one) {
    __shared__ std::uint32_t lds[kWaves64 * kRing64];
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const long item = static_cast<long>(blockIdx.x) * kWaves64 + wave;
    if (item >= items) return;
    const Win64Job jb = jobs ? load_job64(jobs + find_job64(jobs, count, item)) : one;
    if (!rows64_item<W, K, In>(jb, decode64(jb, item - jb.first_item), lds + wave * kRing64, lane)
        && jb.flag)
        atomicOr(jb.flag, 1);
}

template <int W, int K, class Out>
__global__ void __launch_bounds__(64 * kWaves64)
columns_window64(const Win64Job *__restrict__ jobs, int count, long items, Win64Job one) {
    const int lane = threadIdx.x & 63;
    const int wave = uniform(threadIdx.x >> 6);
    const long item = static_cast<long>(blockIdx.x) * kWaves64 + wave;
    if (item >= items) return;
    const Win64Job jb = jobs ? load_job64(jobs + find_job64(jobs, count, item)) : one;
    if (!columns64_item<W, K, Out>(jb, decode64(jb, item - jb.first_item), lane) && jb.flag)
        atomicOr(jb.flag, 1);
}

} // namespace
} // namespace dev

namespace dev {
namespace {
template <int W, int K, class T>
hipError_t launch64(const Win64Job *jobs, int count, long items, const Win64Job &one, bool rows,
                    hipStream_t s) {
    const long blocks = (items + kWaves64 - 1) / kWaves64;
    if (blocks > 0x7FFFFFFFL) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(blocks));
    const dim3 block(64 * kWaves64);
    if (rows) rows_window64<W, K, T><<<grid, block, 0, s>>>(jobs, count, items, one);
    else columns_window64<W, K, T><<<grid, block, 0, s>>>(jobs, count, items, one);
    return hipGetLastError();
}
} // namespace
} // namespace dev

// `kind`: the sample type of the rows pass's source or the columns pass's
// destination (the intermediate is FP64).  `jobs` null: the single job `one`.
template <int W>
hipError_t launch_window64_table_band(const Win64Job *jobs, int count, long items,
                                      const Win64Job &one, int width, bool rows, Sample kind,
                                      hipStream_t s) {
    if (items <= 0) return hipSuccess;
    if (width != 8 && width != 16) return hipErrorInvalidValue;
    const bool wide = width == 16;
    switch (kind) {
    case Sample::f32:
        return wide ? dev::launch64<W, 16, float>(jobs, count, items, one, rows, s)
                    : dev::launch64<W, 8, float>(jobs, count, items, one, rows, s);
    case Sample::u8:
        return wide ? dev::launch64<W, 16, std::uint8_t>(jobs, count, items, one, rows, s)
                    : dev::launch64<W, 8, std::uint8_t>(jobs, count, items, one, rows, s);
    case Sample::u16:
        return wide ? dev::launch64<W, 16, std::uint16_t>(jobs, count, items, one, rows, s)
                    : dev::launch64<W, 8, std::uint16_t>(jobs, count, items, one, rows, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace dsmvc::hip

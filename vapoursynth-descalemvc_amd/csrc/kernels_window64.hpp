// Windowed FP64 axis solves for the Float64 route of F32 planes (gfx950),
// templated on the band class W and the tap width K.  Included by
// kernels_window64_w{1,3,5,7}.hip, one band class each.
//
// The structure follows the F32 windowed kernels: every wavefront solves one
// item, 64 vectors (lanes) x one segment of `span` consecutive chunks of L
// destination rows of one plan.  The forward recurrence starts from a zero
// state P rows before the segment and runs on through its chunks; P rows
// past each chunk the backward recurrence starts from a zero state and
// walks down over the chunk (DevWindow64: the plan's FP64 recurrence has
// decayed below 2^-40 over P rows), and the P forward values beyond the
// chunk carry into the next.  The arithmetic is
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

// Items run vector group fastest, then segment, then frame.
__device__ __forceinline__ Pos64 decode64(const Win64Job &a, long item) {
    const int segments = window64_segments(a);
    const long per_frame = static_cast<long>(segments) * a.groups;
    Pos64 p;
    p.frame = static_cast<int>(item / per_frame);
    const int r = static_cast<int>(item - static_cast<long>(p.frame) * per_frame);
    const int segment = r / a.groups;
    p.group = r - segment * a.groups;
    p.chunk = segment * window64_span(a);
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
// The chunk's rows y[0, L) receive the solution; y[L, L + P) keep their
// forward values, from which the segment's next chunk continues.
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
            if (k < L) y[k] = v;
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
    constexpr int NBL = L / G;          // staging blocks per chunk
    constexpr int NBP = 2 * P / G;      // blocks before a chunk's own rows in its window
    constexpr int RW = window_fwd_words(W, K);
    constexpr int B = window_fwd_batch64(W, K);
    // Forward-only iterations before the segment's first chunk: at least 2P
    // rows (P of warm-up, P beyond the chunk's end).
    constexpr int WARM = (2 * P + L - 1) / L;
    static_assert(L % G == 0 && (2 * P) % G == 0 && G % B == 0, "blocks tile the window");
    const DevWindow64 &pl = a.plan;
    const int nrows = a.nvec;
    const int row0 = at.group * 64;
    const int q0 = at.chunk;
    const int q1 = min(q0 + window64_span(a), pl.chunks);
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
    // Plan tables by padded row (records, tap runs) and by block (staging starts).
    const auto ts = constant(pl.tap_start);
    const auto fwd = constant(pl.fwd);
    const auto blk = constant(pl.blk_lo);
    const bool zform = pl.zform != 0;
    const float *mine = reinterpret_cast<const float *>(ring) + lane * kRowPitch64;
    float *fring = reinterpret_cast<float *>(ring);

    // Block gb covers padded rows [gb G, gb G + G).  Blocks before padded
    // row 0 hold no rows: they load block 0's samples and skip the solve.
    // Lane (srow, scol) loads column scol of rows srow + 2t; rows past the
    // plane (last vector group) read 0 through the buffer's range check and
    // their results are not stored.  The lane's address is rebuilt per block
    // from a pinned value, so that per-row addresses are not kept live
    // across the chunk loop.
    using Word = std::conditional_t<std::is_same_v<In, float>, float, std::uint32_t>;
    Word staged[32];
    const int row_step = 2 * static_cast<int>(a.s_row) * static_cast<int>(sizeof(In));
    auto fetch = [&](int gb) {
        const int column = min(blk[max(gb, 0)] + scol, m1);
        int v = ((row0 + srow) * static_cast<int>(a.s_row) + column) * static_cast<int>(sizeof(In));
        asm volatile("" : "+v"(v));
#pragma unroll
        for (int t = 0; t < 32; ++t)
            staged[t] = static_cast<Word>(BufferLoad<In>::at(rs, v, t * row_step));
    };

    double st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0;
    double y[L + P];
#pragma unroll
    for (int k = 0; k < P; ++k) y[k] = 0.0;
    const int first = (q0 - WARM) * NBL + NBP;
    fetch(first);
    int gb = first;
    bool finite = true;
    auto *out = static_cast<std::uint32_t *>(a.dst) + 2 * static_cast<long>(at.frame) * a.d_frame;
    const int n = pl.n;
    // Iteration q runs the forward sweep over padded rows [qL + 2P,
    // (q + 1)L + 2P), i.e. y[P, L + P) with y[k] at destination row qL + k,
    // then (for q >= q0) chunk q's backward sweep and stores.
    // A chunk's first block is requested after the previous chunk's stores,
    // not before its backward sweep: staged samples held across the sweep
    // would need more than the 256 registers of two wavefronts per SIMD.
    for (int q = q0 - WARM; q < q1; ++q) {
        if (gb != first) fetch(gb);
#pragma unroll
        for (int b = NBP; b < NBP + NBL; ++b, ++gb) {
            wave_sync();
#pragma unroll
            for (int t = 0; t < 32; ++t) {
                asm volatile("" : "+v"(staged[t]));
                fring[(srow + 2 * t) * kRowPitch64 + scol] = static_cast<float>(staged[t]);
            }
            wave_sync();
            if (b + 1 < NBP + NBL) fetch(gb + 1);
            if (gb < 0) continue;
            const int lo = blk[gb];
            int start[G];
#pragma unroll
            for (int t = 0; t < G; ++t) start[t] = ts[gb * G + t] - lo;
#pragma unroll
            for (int j = 0; j < G; j += B) {
                float tap[B][K];
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const float *p = mine + start[j + u];
#pragma unroll
                    for (int e = 0; e < K; ++e) tap[u][e] = p[e];
                }
                double res[B];
                forward_batch64<W, K, B>(fwd + static_cast<long>(gb * G + j) * RW, tap, sample, zform,
                                         st, res);
#pragma unroll
                for (int u = 0; u < B; ++u) y[b * G + j + u - P] = res[u];
            }
        }
        if (q >= q0) {
            backward_sweep64<W, L, P>(constant(pl.bwd + static_cast<long>(q) * L * window_bwd_words(W)), y);
#pragma unroll
            for (int k = 0; k < L; ++k) finite = finite && isfinite(y[k]);
            // FP64 results leave through the ring 16 at a time as 32-bit
            // word pairs: lanes 0-31 write words of row t, lanes 32-63 of
            // row t + 1 (128 bytes).  The lane's addresses come from a
            // pinned copy of the lane index, so they are not kept live
            // across the chunk loop.
            const int r0 = q * L;
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const int oc = ln & 31;
            const int orow = ln >> 5;
            std::uint32_t *const my_row = ring + ln * kRowPitch64;
#pragma unroll
            for (int kb = 0; kb < L; kb += 16) {
                wave_sync();
#pragma unroll
                for (int c = 0; c < 16; ++c) {
                    if (kb + c >= L) break;
                    const auto bits = __builtin_bit_cast(std::uint64_t, y[kb + c]);
                    my_row[2 * c] = static_cast<std::uint32_t>(bits);
                    my_row[2 * c + 1] = static_cast<std::uint32_t>(bits >> 32);
                }
                wave_sync();
                const int col = kb + (oc >> 1);
                const bool col_ok = col < L && r0 + col < n;
#pragma unroll 4
                for (int t = 0; t < 64; t += 2) {
                    const int rr = t + orow;
                    const int row = row0 + rr;
                    if (col_ok && row < nrows)
                        out[2 * (static_cast<long>(row) * a.d_row + r0) + 2 * kb + oc] =
                            ring[rr * kRowPitch64 + oc];
                }
            }
        }
#pragma unroll
        for (int k = 0; k < P; ++k) y[k] = y[L + k];
    }
    return finite;
}

template <int W, int K, class Out>
__device__ __forceinline__ bool columns64_item(const Win64Job &a, const Pos64 &at, int lane) {
    constexpr int L = window_length64(W);
    constexpr int P = window_warmup64(W);
    constexpr int RW = window_fwd_words(W, K);
    constexpr int B = window_fwd_batch64(W, K);
    constexpr int WARM = (2 * P + L - 1) / L;
    static_assert(L % B == 0 && (2 * P) % B == 0, "batches tile the window");
    const DevWindow64 &pl = a.plan;
    const int ncols = a.nvec;
    const int col = at.group * 64 + lane;
    const bool live = col < ncols;
    const int cc = live ? col : ncols - 1;
    const int q0 = at.chunk;
    const int q1 = min(q0 + window64_span(a), pl.chunks);
    const double *src = static_cast<const double *>(a.src) + static_cast<long>(at.frame) * a.s_frame;
    const auto rs = plane_rsrc(src, static_cast<long>(pl.m) * a.s_row * 8);
    const int voff = cc * 8;
    const int row_bytes = static_cast<int>(a.s_row * 8);
    const int m1 = pl.m - 1;
    const auto ts = constant(pl.tap_start);
    const auto fwd = constant(pl.fwd);
    const bool zform = pl.zform != 0;

    // Taps of the batch at padded row pr: K consecutive source rows per
    // destination row, each a coalesced load across the wavefront (rows past
    // the plane clamp to its last row; their weights are 0).
    double cur[B][K], nxt[B][K];
    auto fetch = [&](int pr, double (&t)[B][K]) {
#pragma unroll
        for (int u = 0; u < B; ++u) {
            const int s0 = ts[pr + u];
#pragma unroll
            for (int e = 0; e < K; ++e) {
                const auto w = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, min(s0 + e, m1) * row_bytes, 0);
                t[u][e] = __builtin_bit_cast(double, w);
            }
        }
    };
    // The segment's forward sweep covers padded rows [first, end); rows
    // before padded row 0 do not exist and are skipped.
    int pr = (q0 - WARM) * L + 2 * P;
    const int end = q1 * L + 2 * P;
    fetch(max(pr, 0), cur);

    double st[W];
#pragma unroll
    for (int d = 0; d < W; ++d) st[d] = 0.0;
    double y[L + P];
#pragma unroll
    for (int k = 0; k < P; ++k) y[k] = 0.0;
    Out *o = static_cast<Out *>(a.dst) + static_cast<long>(at.frame) * a.d_frame + cc;
    const int n = pl.n;
    const DevConv cv = a.conv;
    bool finite = true;
    // The iterations of rows64_item: y[k] at destination row qL + k.
    for (int q = q0 - WARM; q < q1; ++q) {
#pragma unroll
        for (int j = 0; j < L; j += B, pr += B) {
            if (pr < 0) continue;
            if (pr + B < end) fetch(pr + B, nxt);
            double res[B];
            forward_batch64<W, K, B>(fwd + static_cast<long>(pr) * RW, cur, [](double x) { return x; },
                                     zform, st, res);
#pragma unroll
            for (int u = 0; u < B; ++u) {
                y[P + j + u] = res[u];
#pragma unroll
                for (int e = 0; e < K; ++e) cur[u][e] = nxt[u][e];
            }
        }
        if (q >= q0) {
            backward_sweep64<W, L, P>(constant(pl.bwd + static_cast<long>(q) * L * window_bwd_words(W)), y);
            const int r0 = q * L;
#pragma unroll
            for (int k = 0; k < L; ++k)
                if (r0 + k < n && live)
                    finite = Store<Out>::f64(o + static_cast<long>(r0 + k) * a.d_row, y[k], cv) && finite;
        }
#pragma unroll
        for (int k = 0; k < P; ++k) y[k] = y[L + k];
    }
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
// Held to two wavefronts per SIMD (256 registers), as the columns kernel
// reaches by itself.
template <int W, int K, class In>
__global__ void __launch_bounds__(64 * kWaves64) __attribute__((amdgpu_waves_per_eu(2)))
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

namespace dev {
namespace {
template <class F> long resident_waves64(F kernel) {
    int blocks = 0, device = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, kernel, 64 * kWaves64, 0) != hipSuccess
        || blocks <= 0)
        blocks = 1;
    if (hipGetDevice(&device) != hipSuccess
        || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess
        || cus <= 0)
        cus = 256;
    return static_cast<long>(blocks) * cus * kWaves64;
}

template <int W, int K, class T> long slots64(bool rows) {
    if (rows) {
        static const long v = resident_waves64(rows_window64<W, K, T>);
        return v;
    }
    static const long v = resident_waves64(columns_window64<W, K, T>);
    return v;
}
} // namespace
} // namespace dev

template <int W> long window64_table_slots_band(int width, bool rows, Sample kind) {
    const bool wide = width == 16;
    switch (kind) {
    case Sample::f32: return wide ? dev::slots64<W, 16, float>(rows) : dev::slots64<W, 8, float>(rows);
    case Sample::u8:
        return wide ? dev::slots64<W, 16, std::uint8_t>(rows) : dev::slots64<W, 8, std::uint8_t>(rows);
    case Sample::u16:
        return wide ? dev::slots64<W, 16, std::uint16_t>(rows) : dev::slots64<W, 8, std::uint16_t>(rows);
    default: return 0;
    }
}

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

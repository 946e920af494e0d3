// Chunked (partitioned) axis solves for gfx950, templated on the band W, the
// chunk length L and the packed tap count K.  Included by
// kernels_chunked_w{1,3,5,7}.hip, which instantiate one band each so the
// build runs in parallel.
//
// Partition: a workgroup owns 64 vectors (lanes) and C chunks of L elements
// (one wavefront per chunk).  Each wavefront
//   1. gathers rhs for its L rows into registers (straight-line code: the
//      K taps of every row are wave-uniform scalar loads, the samples are
//      64-wide loads, so many loads are in flight at once),
//   2. runs the forward recurrence from a zero carry and publishes its final
//      W states to LDS,
//   3. forms its true carry-in from the earlier chunks' states and the FP64
//      transfer matrices (S_q = e_q + T_q S_{q-1}) and re-solves the chunk,
//   4. does the same for the backward recurrence, then stores.
// The first chunk (forward) and the last (backward) start from the true zero
// carry, so their values match the sequential solve bit for bit; the others
// differ from it only by the rounding of the carried states.
//
// Vertical pass (columns_chunked): lanes are 64 adjacent columns; taps read
// source rows directly (buffer loads at a scalar row offset).
// Horizontal pass (rows_chunked): lanes are 64 rows; every G (16, or 8 for
// wide windows) destination rows the wave stages the 32 source columns their
// taps need into its LDS ring,
// transposed (pitch 65, conflict-free), and the results leave through the
// same ring transposed back, so global traffic stays row-contiguous.
#pragma once

#include "kernels_common.hpp"

namespace dsmvc::hip {
namespace dev {
namespace {

constexpr int kWaveRing = kStageCols * kRingPitch;   // floats of LDS per wavefront

template <int W, int L>
__device__ __forceinline__ void solve_chunk(float (&y)[L], const float *__restrict__ cf,
                                            float *ef, float *eb, const float *lds_base,
                                            int stride, int fwd_off, int bwd_off,
                                            const DevChunks &ch, int j, int lane) {
    const int C = ch.chunks;
    float r[W];
#pragma unroll
    for (int d = 0; d < W; ++d) r[d] = 0.0F;
    if (j < C - 1) {
#pragma unroll
        for (int k = 0; k < L; ++k) (void)forward_step<W>(cf + k * Band<W>::kStride, y[k], r);
#pragma unroll
        for (int d = 0; d < W; ++d) ef[d * 64 + lane] = r[d];
    }
    __syncthreads();
    carry_chain<W>(r, lds_base + fwd_off, stride, ch.tf, 0, j, 1, lane);
#pragma unroll
    for (int k = 0; k < L; ++k) y[k] = forward_step<W>(cf + k * Band<W>::kStride, y[k], r);
    if (j > 0) {
#pragma unroll
        for (int d = 0; d < W; ++d) r[d] = 0.0F;
#pragma unroll
        for (int k = L - 1; k >= 0; --k) (void)backward_step<W>(cf + k * Band<W>::kStride, y[k], r);
#pragma unroll
        for (int d = 0; d < W; ++d) eb[d * 64 + lane] = r[d];
    }
    __syncthreads();
    carry_chain<W>(r, lds_base + bwd_off, stride, ch.tb, C - 1, j, -1, lane);
#pragma unroll
    for (int k = L - 1; k >= 0; --k) y[k] = backward_step<W>(cf + k * Band<W>::kStride, y[k], r);
}

template <int W, int L, int K, class Out>
__global__ void __launch_bounds__(1024)
columns_chunked(DevAxis ax, DevChunks ch, const float *__restrict__ src, long s_row, long s_frame,
                Out *__restrict__ dst, long d_row, long d_frame, int ncols, DevConv cv) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int j = uniform(threadIdx.x >> 6);
    const int C = ch.chunks;
    const int a = j * L;
    const int col = blockIdx.x * 64 + lane;
    const bool live = col < ncols;
    const int cc = live ? col : ncols - 1;
    const auto rs = plane_rsrc(src + blockIdx.y * s_frame, static_cast<long>(ax.m) * s_row * 4);
    const int voff = cc * 4;
    const int row_bytes = static_cast<int>(s_row * 4);
    const std::int32_t *ei = ch.ell_idx + static_cast<long>(a) * K;
    const float *ew = ch.ell_w + static_cast<long>(a) * K;

    float y[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
        float v = 0.0F;
#pragma unroll
        for (int e = 0; e < K; ++e)
            v = __builtin_fmaf(ew[k * K + e], BufferLoad<float>::at(rs, voff, ei[k * K + e] * row_bytes), v);
        y[k] = v;
    }
    solve_chunk<W, L>(y, ch.coef + static_cast<long>(a) * Band<W>::kStride, lds + j * W * 64,
                      lds + (C + j) * W * 64, lds, W * 64, 0, C * W * 64, ch, j, lane);
    Out *o = dst + blockIdx.y * d_frame + cc;
    const int n = ax.n;
#pragma unroll
    for (int k = 0; k < L; ++k)
        if (a + k < n && live) Store<Out>::f32(o + static_cast<long>(a + k) * d_row, y[k], cv);
}

template <int W, int L, int K, int G, class In>
__global__ void __launch_bounds__(1024)
rows_chunked(DevAxis ax, DevChunks ch, const In *__restrict__ src, long s_row, long s_frame,
             float *__restrict__ dst, long d_row, long d_frame, int nrows, DevConv cv) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int j = uniform(threadIdx.x >> 6);
    const int a = j * L;
    const int row0 = blockIdx.x * 64;
    float *ring = lds + j * kWaveRing;
    const auto rs = plane_rsrc(src + blockIdx.y * s_frame,
                               static_cast<long>(nrows) * s_row * static_cast<long>(sizeof(In)));
    const int m1 = ax.m - 1;
    const int scol = lane & (kStageCols - 1);
    const int srow = lane >> 5;
    const std::int32_t *ei = ch.ell_idx + static_cast<long>(a) * K;
    const float *ew = ch.ell_w + static_cast<long>(a) * K;
    const std::int32_t *blk = ch.blk_lo + j * (L / G);

    float y[L];
#pragma unroll
    for (int b = 0; b < L / G; ++b) {
        const int lo = blk[b];
        const int column = min(lo + scol, m1);
        wave_sync();
#pragma unroll
        for (int t = 0; t < 32; ++t) {
            const int rr = srow + 2 * t;
            const int row = min(row0 + rr, nrows - 1);
            const int voff = (row * static_cast<int>(s_row) + column) * static_cast<int>(sizeof(In));
            ring[scol * kRingPitch + rr] = Load<In>::norm(BufferLoad<In>::at(rs, voff, 0), cv);
        }
        wave_sync();
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const int k = b * G + t;
            float v = 0.0F;
#pragma unroll
            for (int e = 0; e < K; ++e)
                v = __builtin_fmaf(ew[k * K + e], ring[(ei[k * K + e] - lo) * kRingPitch + lane], v);
            y[k] = v;
        }
    }
    wave_sync();
    // Chunk states live in the wave's own ring from here on.
    solve_chunk<W, L>(y, ch.coef + static_cast<long>(a) * Band<W>::kStride, ring, ring + W * 64,
                      lds, kWaveRing, 0, W * 64, ch, j, lane);
    __syncthreads();   // every wave has read the chunk states; rings are free

    // Results leave through an LDS transpose, 32 destination columns at a
    // time: lanes 0-31 write row t, lanes 32-63 row t + 1, 128-byte runs.
    float *out = dst + blockIdx.y * d_frame + a;
    const int oc = lane & 31;
    const int orow = lane >> 5;
    const int n = ax.n;
#pragma unroll
    for (int kb = 0; kb < L; kb += 32) {
#pragma unroll
        for (int c = 0; c < 32; ++c)
            if (kb + c < L) ring[c * kRingPitch + lane] = y[kb + c];
        wave_sync();
        const bool col_ok = kb + oc < L && a + kb + oc < n;
#pragma unroll 4
        for (int t = 0; t < 64; t += 2) {
            const int rr = t + orow;
            const int row = row0 + rr;
            if (col_ok && row < nrows)
                out[static_cast<long>(row) * d_row + kb + oc] = ring[oc * kRingPitch + rr];
        }
        wave_sync();
    }
}

template <int W, int L, int K, class Out>
hipError_t launch_cols_wlk(const ChunkArgs &a, hipStream_t s) {
    const dim3 grid((a.nvec + 63) / 64, a.frames);
    const std::size_t shm = 2U * a.chunks.chunks * W * 64 * sizeof(float);
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&columns_chunked<W, L, K, Out>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 2U * 16 * W * 64 * sizeof(float));
    if (attr != hipSuccess) return attr;
    columns_chunked<W, L, K, Out><<<grid, 64 * a.chunks.chunks, shm, s>>>(
        a.axis, a.chunks, static_cast<const float *>(a.src), a.s_row, a.s_frame,
        static_cast<Out *>(a.dst), a.d_row, a.d_frame, a.nvec, a.conv);
    return hipGetLastError();
}

template <int W, int L, int K, int G, class In>
hipError_t launch_rows_wlkg(const ChunkArgs &a, hipStream_t s) {
    const dim3 grid((a.nvec + 63) / 64, a.frames);
    const std::size_t shm = static_cast<std::size_t>(a.chunks.chunks) * kWaveRing * sizeof(float);
    // Up to 16 staging rings (133 KB) exceed the default dynamic-LDS limit.
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&rows_chunked<W, L, K, G, In>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 16 * kWaveRing * sizeof(float));
    if (attr != hipSuccess) return attr;
    rows_chunked<W, L, K, G, In><<<grid, 64 * a.chunks.chunks, shm, s>>>(
        a.axis, a.chunks, static_cast<const In *>(a.src), a.s_row, a.s_frame,
        static_cast<float *>(a.dst), a.d_row, a.d_frame, a.nvec, a.conv);
    return hipGetLastError();
}

template <int W, int L, int K, class In>
hipError_t launch_rows_wlk(const ChunkArgs &a, hipStream_t s) {
    return a.chunks.block_rows == 16 ? launch_rows_wlkg<W, L, K, 16, In>(a, s)
                                     : launch_rows_wlkg<W, L, K, 8, In>(a, s);
}

template <int W, int L, class T, bool Rows>
hipError_t launch_wl(const ChunkArgs &a, hipStream_t s) {
    switch (a.chunks.ell_width) {
    case 8: return Rows ? launch_rows_wlk<W, L, 8, T>(a, s) : launch_cols_wlk<W, L, 8, T>(a, s);
    case 16: return Rows ? launch_rows_wlk<W, L, 16, T>(a, s) : launch_cols_wlk<W, L, 16, T>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <int W, class T, bool Rows>
hipError_t launch_wt(const ChunkArgs &a, hipStream_t s) {
    switch (a.chunks.length) {
    case 32: return launch_wl<W, 32, T, Rows>(a, s);
    case 48: return launch_wl<W, 48, T, Rows>(a, s);
    case 64: return launch_wl<W, 64, T, Rows>(a, s);
    case 80: return launch_wl<W, 80, T, Rows>(a, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace
} // namespace dev

template <int W>
hipError_t launch_chunked_band(const ChunkArgs &a, bool rows, Sample kind, hipStream_t s) {
    using namespace dev;
    switch (kind) {
    case Sample::f32:
        return rows ? launch_wt<W, float, true>(a, s) : launch_wt<W, float, false>(a, s);
    case Sample::u8:
        return rows ? launch_wt<W, std::uint8_t, true>(a, s) : launch_wt<W, std::uint8_t, false>(a, s);
    case Sample::u16:
        return rows ? launch_wt<W, std::uint16_t, true>(a, s)
                    : launch_wt<W, std::uint16_t, false>(a, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace dsmvc::hip

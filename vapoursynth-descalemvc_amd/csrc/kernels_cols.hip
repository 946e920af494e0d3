// Vertical pass, chunked (see kernels_seq.hip for the kernel family overview).
#include "kernels_common.hpp"

namespace dsmvc::hip {
namespace dev {
namespace {

// ---- vertical pass, chunked ---------------------------------------------------
template <int W, int MAXL, class Out>
__global__ void __launch_bounds__(1024)
columns_chunked(DevAxis ax, DevChunks ch, const float *__restrict__ src, long s_row, long s_frame,
                Out *__restrict__ dst, long d_row, long d_frame, int ncols, DevConv cv) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int j = uniform(threadIdx.x >> 6);
    const int C = ch.chunks, L = ch.length, n = ax.n;
    const int a = j * L;
    const int len = min(L, n - a);
    const int col = blockIdx.x * 64 + lane;
    const bool live = col < ncols;
    const long cc = live ? col : ncols - 1;
    const float *s = src + blockIdx.y * s_frame + cc;
    float *ef = lds;                 // [C][W][64] forward chunk-end states
    float *eb = lds + C * W * 64;    // [C][W][64] backward chunk-start states

    float y[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < len) {
            const int i = a + k;
            const std::uint32_t q1 = ax.off[i + 1];
            float v = 0.0F;
            for (std::uint32_t q = ax.off[i]; q < q1; ++q)
                v = __builtin_fmaf(ax.wt[q], s[static_cast<long>(ax.idx[q]) * s_row], v);
            y[k] = v;
        }
    }
    float r[W];
#pragma unroll
    for (int d = 0; d < W; ++d) r[d] = 0.0F;
    if (j < C - 1) {
#pragma unroll
        for (int k = 0; k < MAXL; ++k)
            if (k < len) (void)forward_step<W>(ax.coef + static_cast<long>(a + k) * Band<W>::kStride, y[k], r);
#pragma unroll
        for (int d = 0; d < W; ++d) ef[(j * W + d) * 64 + lane] = r[d];
    }
    __syncthreads();
    carry_chain<W>(r, ef, W * 64, ch.tf, 0, j, 1, lane);
#pragma unroll
    for (int k = 0; k < MAXL; ++k)
        if (k < len) y[k] = forward_step<W>(ax.coef + static_cast<long>(a + k) * Band<W>::kStride, y[k], r);
    if (j > 0) {
#pragma unroll
        for (int d = 0; d < W; ++d) r[d] = 0.0F;
#pragma unroll
        for (int k = MAXL - 1; k >= 0; --k)
            if (k < len) (void)backward_step<W>(ax.coef + static_cast<long>(a + k) * Band<W>::kStride, y[k], r);
#pragma unroll
        for (int d = 0; d < W; ++d) eb[(j * W + d) * 64 + lane] = r[d];
    }
    __syncthreads();
    carry_chain<W>(r, eb, W * 64, ch.tb, C - 1, j, -1, lane);
    Out *o = dst + blockIdx.y * d_frame + cc;
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k) {
        if (k < len) {
            const float x = backward_step<W>(ax.coef + static_cast<long>(a + k) * Band<W>::kStride, y[k], r);
            if (live) Store<Out>::f32(o + static_cast<long>(a + k) * d_row, x, cv);
        }
    }
}

} // namespace
} // namespace dev

using namespace dev;

namespace {

template <int W, int MAXL, class Out>
hipError_t launch_cols_wl(const ChunkArgs &a, hipStream_t s) {
    const dim3 grid((a.nvec + 63) / 64, a.frames);
    const std::size_t shm = 2U * a.chunks.chunks * W * 64 * sizeof(float);
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&columns_chunked<W, MAXL, Out>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 2U * 16 * W * 64 * sizeof(float));
    if (attr != hipSuccess) return attr;
    columns_chunked<W, MAXL, Out><<<grid, 64 * a.chunks.chunks, shm, s>>>(
        a.axis, a.chunks, static_cast<const float *>(a.src), a.s_row, a.s_frame,
        static_cast<Out *>(a.dst), a.d_row, a.d_frame, a.nvec, a.conv);
    return hipGetLastError();
}

template <int MAXL, class Out>
hipError_t launch_cols_l(const ChunkArgs &a, hipStream_t s) {
    switch (a.axis.w) {
    case 1: return launch_cols_wl<1, MAXL, Out>(a, s);
    case 3: return launch_cols_wl<3, MAXL, Out>(a, s);
    case 5: return launch_cols_wl<5, MAXL, Out>(a, s);
    case 7: return launch_cols_wl<7, MAXL, Out>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <class Out>
hipError_t launch_cols_t(const ChunkArgs &a, hipStream_t s) {
    const int L = a.chunks.length;
    if (L <= 32) return launch_cols_l<32, Out>(a, s);
    if (L <= 48) return launch_cols_l<48, Out>(a, s);
    if (L <= 64) return launch_cols_l<64, Out>(a, s);
    if (L <= 80) return launch_cols_l<80, Out>(a, s);
    return hipErrorInvalidValue;
}

} // namespace

hipError_t launch_columns_chunked(const ChunkArgs &a, Sample out, hipStream_t s) {
    switch (out) {
    case Sample::f32: return launch_cols_t<float>(a, s);
    case Sample::u8: return launch_cols_t<std::uint8_t>(a, s);
    case Sample::u16: return launch_cols_t<std::uint16_t>(a, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace dsmvc::hip

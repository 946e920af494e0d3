// Horizontal pass, chunked (see kernels_seq.hip for the kernel family overview).
#include "kernels_common.hpp"

namespace dsmvc::hip {
namespace dev {
namespace {

// ---- horizontal pass, chunked -------------------------------------------------
template <int W, int MAXL, class In>
__global__ void __launch_bounds__(1024)
rows_chunked(DevAxis ax, DevChunks ch, const In *__restrict__ src, long s_row, long s_frame,
             float *__restrict__ dst, long d_row, long d_frame, int nrows, DevConv cv) {
    extern __shared__ float lds[];
    const int lane = threadIdx.x & 63;
    const int j = uniform(threadIdx.x >> 6);
    const int C = ch.chunks, L = ch.length, n = ax.n;
    const int a = j * L;
    const int len = min(L, n - a);
    const int row0 = blockIdx.x * 64;
    constexpr int kWaveLds = kRingSlots * kRingPitch;
    float *ring = lds + j * kWaveLds;
    const In *sf = src + blockIdx.y * s_frame;
    const int m1 = ax.m - 1;
    const int stage_col = lane & (kRingHalf - 1);
    const int stage_row = lane >> 4;

    // Stage source columns [c0, c0 + 16) of the tile's 64 rows into the ring,
    // slot = column mod 32, element (column, row) at slot * 65 + row.
    auto stage = [&](int c0) {
        const int column = min(c0 + stage_col, m1);
        float *slot = ring + ((c0 + stage_col) & (kRingSlots - 1)) * kRingPitch;
#pragma unroll
        for (int t = 0; t < 16; ++t) {
            const int rr = stage_row + 4 * t;
            const int row = min(row0 + rr, nrows - 1);
            slot[rr] = Load<In>::f32(sf + static_cast<long>(row) * s_row + column, cv);
        }
    };

    int base = uniform(ch.win[2 * j] & ~(kRingHalf - 1));
    stage(base);
    stage(base + kRingHalf);
    wave_sync();

    float y[MAXL];
#pragma unroll
    for (int k = 0; k < MAXL; ++k) {
        if (k < len) {
            const int i = a + k;
            const std::uint32_t q0 = ax.off[i], q1 = ax.off[i + 1];
            const int top = q1 > q0 ? ax.idx[q1 - 1] : base;
            while (top >= base + kRingSlots) {
                base += kRingHalf;
                stage(base + kRingHalf);
                wave_sync();
            }
            float v = 0.0F;
            for (std::uint32_t q = q0; q < q1; ++q)
                v = __builtin_fmaf(ax.wt[q], ring[(ax.idx[q] & (kRingSlots - 1)) * kRingPitch + lane], v);
            y[k] = v;
        }
    }
    wave_sync();
    // Chunk states live in the wave's own ring region from here on.
    float *ef = ring;               // [W][64]
    float *eb = ring + W * 64;      // [W][64]
    float r[W];
#pragma unroll
    for (int d = 0; d < W; ++d) r[d] = 0.0F;
    if (j < C - 1) {
#pragma unroll
        for (int k = 0; k < MAXL; ++k)
            if (k < len) (void)forward_step<W>(ax.coef + static_cast<long>(a + k) * Band<W>::kStride, y[k], r);
#pragma unroll
        for (int d = 0; d < W; ++d) ef[d * 64 + lane] = r[d];
    }
    __syncthreads();
    carry_chain<W>(r, lds, kWaveLds, ch.tf, 0, j, 1, lane);
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
        for (int d = 0; d < W; ++d) eb[d * 64 + lane] = r[d];
    }
    __syncthreads();
    carry_chain<W>(r, lds + W * 64, kWaveLds, ch.tb, C - 1, j, -1, lane);
#pragma unroll
    for (int k = MAXL - 1; k >= 0; --k)
        if (k < len) y[k] = backward_step<W>(ax.coef + static_cast<long>(a + k) * Band<W>::kStride, y[k], r);
    __syncthreads();   // every wave has read the chunk states; rings are free

    // Results leave through an LDS transpose, 32 destination columns at a
    // time: lanes 0-31 write row t, lanes 32-63 row t + 1, 128 B runs each.
    float *out = dst + blockIdx.y * d_frame + a;
    const int oc = lane & 31;
    const int orow = lane >> 5;
#pragma unroll
    for (int kb = 0; kb < MAXL; kb += 32) {
        if (kb < len) {
#pragma unroll
            for (int c = 0; c < 32; ++c)
                if (kb + c < MAXL && kb + c < len) ring[c * kRingPitch + lane] = y[kb + c];
            wave_sync();
            const bool col_ok = kb + oc < len;
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
}

} // namespace
} // namespace dev

using namespace dev;

namespace {

template <int W, int MAXL, class In>
hipError_t launch_rows_wl(const ChunkArgs &a, hipStream_t s) {
    const dim3 grid((a.nvec + 63) / 64, a.frames);
    const std::size_t shm =
        static_cast<std::size_t>(a.chunks.chunks) * kRingSlots * kRingPitch * sizeof(float);
    // Up to 16 staging rings (133 KB) exceed the default dynamic-LDS limit.
    static const hipError_t attr = hipFuncSetAttribute(
        reinterpret_cast<const void *>(&rows_chunked<W, MAXL, In>),
        hipFuncAttributeMaxDynamicSharedMemorySize, 16 * kRingSlots * kRingPitch * sizeof(float));
    if (attr != hipSuccess) return attr;
    rows_chunked<W, MAXL, In><<<grid, 64 * a.chunks.chunks, shm, s>>>(
        a.axis, a.chunks, static_cast<const In *>(a.src), a.s_row, a.s_frame,
        static_cast<float *>(a.dst), a.d_row, a.d_frame, a.nvec, a.conv);
    return hipGetLastError();
}

template <int MAXL, class In>
hipError_t launch_rows_l(const ChunkArgs &a, hipStream_t s) {
    switch (a.axis.w) {
    case 1: return launch_rows_wl<1, MAXL, In>(a, s);
    case 3: return launch_rows_wl<3, MAXL, In>(a, s);
    case 5: return launch_rows_wl<5, MAXL, In>(a, s);
    case 7: return launch_rows_wl<7, MAXL, In>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template <class In>
hipError_t launch_rows_t(const ChunkArgs &a, hipStream_t s) {
    const int L = a.chunks.length;
    if (L <= 32) return launch_rows_l<32, In>(a, s);
    if (L <= 48) return launch_rows_l<48, In>(a, s);
    if (L <= 64) return launch_rows_l<64, In>(a, s);
    if (L <= 80) return launch_rows_l<80, In>(a, s);
    return hipErrorInvalidValue;
}

} // namespace

int chunked_max_length() { return 80; }

bool chunked_band_supported(int w) { return w == 1 || w == 3 || w == 5 || w == 7; }

int rows_ring_span() { return kRingHalf; }

hipError_t launch_rows_chunked(const ChunkArgs &a, Sample in, hipStream_t s) {
    switch (in) {
    case Sample::f32: return launch_rows_t<float>(a, s);
    case Sample::u8: return launch_rows_t<std::uint8_t>(a, s);
    case Sample::u16: return launch_rows_t<std::uint16_t>(a, s);
    default: return hipErrorInvalidValue;
    }
}

} // namespace dsmvc::hip

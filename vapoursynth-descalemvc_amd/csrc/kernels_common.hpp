// Device helpers shared by the gfx950 kernel translation units:
// sample conversion, coefficient-record layout, the recurrence steps and the
// chunk carry chain.  Included by the kernel translation units (the
// sequential, chunked, windowed and refinement kernels) only.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device_types.hpp"
#include "kernels.hpp"

namespace dsmvc::hip {
namespace dev {

constexpr int kRingPitch = 65;     // +1 float: conflict-free transposed LDS access

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Tables that no kernel writes while it runs, read through the constant
// address space: a wave-uniform read is then a scalar load even after the
// kernel's own global stores.
template <class T> using ConstPtr = const __attribute__((address_space(4))) T *;
template <class T> __device__ __forceinline__ ConstPtr<T> constant(const T *p) {
    return (ConstPtr<T>)(p);   // address-space cast: flat to constant
}
// Device-memory pointers read from a table (not kernel arguments) are flat
// pointers to the compiler; casting them to the global address space lets
// wave-uniform reads through them become scalar loads.
template <class T> using GlobalPtr = const __attribute__((address_space(1))) T *;
template <class T> __device__ __forceinline__ GlobalPtr<T> global(const T *p) {
    return (GlobalPtr<T>)(p);   // address-space cast: flat to global
}

// Raw buffer over a plane: loads take a per-lane byte offset and a scalar
// (wave-uniform) byte offset, and read 0 past `bytes`.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *base, long bytes) {
    const int n = bytes > 0x7FFFFFF0L ? 0x7FFFFFF0 : static_cast<int>(bytes);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, n, 0x00020000);
}

template <class T> struct BufferLoad;
template <> struct BufferLoad<float> {
    __device__ static float at(__amdgpu_buffer_rsrc_t r, int v, int s) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, v, s, 0));
    }
};
template <> struct BufferLoad<std::uint16_t> {
    __device__ static std::uint16_t at(__amdgpu_buffer_rsrc_t r, int v, int s) {
        return __builtin_amdgcn_raw_buffer_load_b16(r, v, s, 0);
    }
};
template <> struct BufferLoad<std::uint8_t> {
    __device__ static std::uint8_t at(__amdgpu_buffer_rsrc_t r, int v, int s) {
        return __builtin_amdgcn_raw_buffer_load_b8(r, v, s, 0);
    }
};

// ---- sample conversion ----------------------------------------------------
// Integer input is normalised as ((float)x - offset) * scale with two
// roundings, output as clamp(v * scale + offset, 0, max) rounded to nearest
// even (reference src/cpu_executor.cpp:1756-1763, 1825-1831).
template <class In> struct Load;
template <> struct Load<float> {
    __device__ static float f32(const float *p, const DevConv &) { return *p; }
    __device__ static float norm(float x, const DevConv &) { return x; }
    __device__ static double f64(const float *p, const DevConv &) { return static_cast<double>(*p); }
};
template <> struct Load<double> {
    __device__ static double f64(const double *p, const DevConv &) { return *p; }
};
template <> struct Load<std::uint8_t> {
    __device__ static float f32(const std::uint8_t *p, const DevConv &c) {
        return __fmul_rn(__fsub_rn(static_cast<float>(*p), c.in_offset), c.in_scale);
    }
    __device__ static float norm(std::uint8_t x, const DevConv &c) {
        return __fmul_rn(__fsub_rn(static_cast<float>(x), c.in_offset), c.in_scale);
    }
    __device__ static double f64(const std::uint8_t *p, const DevConv &c) {
        return __dmul_rn(__dsub_rn(static_cast<double>(*p), static_cast<double>(c.in_offset)),
                         static_cast<double>(c.in_scale));
    }
};
template <> struct Load<std::uint16_t> {
    __device__ static float f32(const std::uint16_t *p, const DevConv &c) {
        return __fmul_rn(__fsub_rn(static_cast<float>(*p), c.in_offset), c.in_scale);
    }
    __device__ static float norm(std::uint16_t x, const DevConv &c) {
        return __fmul_rn(__fsub_rn(static_cast<float>(x), c.in_offset), c.in_scale);
    }
    __device__ static double f64(const std::uint16_t *p, const DevConv &c) {
        return __dmul_rn(__dsub_rn(static_cast<double>(*p), static_cast<double>(c.in_offset)),
                         static_cast<double>(c.in_scale));
    }
};

template <class Out> struct Store;
template <> struct Store<float> {
    __device__ static void f32(float *p, float v, const DevConv &) { *p = v; }
    __device__ static bool f64(float *p, double v, const DevConv &) {
        const float f = static_cast<float>(v);
        *p = f;
        return isfinite(v) && isfinite(f);
    }
};
template <> struct Store<double> {
    __device__ static bool f64(double *p, double v, const DevConv &) {
        *p = v;
        return isfinite(v);
    }
};
template <class T> struct StoreInt {
    __device__ static void f32(T *p, float v, const DevConv &c) {
        float s = __fadd_rn(__fmul_rn(v, c.out_scale), c.out_offset);
        s = s < 0.0F ? 0.0F : (c.out_max < s ? c.out_max : s);
        *p = static_cast<T>(rintf(s));
    }
    __device__ static bool f64(T *p, double v, const DevConv &c) {
        const double s = __dadd_rn(__dmul_rn(v, static_cast<double>(c.out_scale)),
                                   static_cast<double>(c.out_offset));
        if (!isfinite(v) || !isfinite(s)) {
            *p = T{0};
            return false;
        }
        const double mx = static_cast<double>(c.out_max);
        *p = static_cast<T>(rint(s < 0.0 ? 0.0 : (mx < s ? mx : s)));
        return true;
    }
};
template <> struct Store<std::uint8_t> : StoreInt<std::uint8_t> {};
template <> struct Store<std::uint16_t> : StoreInt<std::uint16_t> {};

// ---- recurrence steps -------------------------------------------------------
// Coefficient record of a band class: bands up to 7 use 16-float rows, wider
// bands (up to 15) 32-float rows; see device_types.hpp.
template <int W> struct Band {
    static constexpr int kSlots = W <= kMaxBand ? kMaxBand : kMaxWideBand;
    static constexpr int kStride = 2 * kSlots + 2;
    static constexpr int kInv = kSlots;
    static constexpr int kUp = kSlots + 1;
};

template <int W, class T>
__device__ __forceinline__ void push(T (&r)[W], T v) {
#pragma unroll
    for (int d = W - 1; d > 0; --d) r[d] = r[d - 1];
    r[0] = v;
}

// Forward + diagonal for row i.  r[d-1] holds y_{i-d}.  Terms with i-d < 0
// meet zero coefficients and a zero ring, which leaves the FMA chain's value
// unchanged, so applying all W terms equals the reference's truncated loop.
template <int W, class Coef>
__device__ __forceinline__ float forward_step_at(Coef c, float rhs, float (&r)[W]) {
    float v = rhs;
#pragma unroll
    for (int d = W; d >= 1; --d) v = __builtin_fmaf(-c[d - 1], r[d - 1], v);
    const float y = __fmul_rn(v, c[Band<W>::kInv]);
    push(r, y);
    return y;
}

// Backward for row i.  r[d-1] holds x_{i+d}.  W == 3 sums d ascending, other
// bands descending, as the reference solver does.
template <int W, class Coef>
__device__ __forceinline__ float backward_step_at(Coef c, float y, float (&r)[W]) {
    float v = y;
    if constexpr (W == 3) {
#pragma unroll
        for (int d = 1; d <= W; ++d) v = __builtin_fmaf(-c[Band<W>::kUp + d - 1], r[d - 1], v);
    } else {
#pragma unroll
        for (int d = W; d >= 1; --d) v = __builtin_fmaf(-c[Band<W>::kUp + d - 1], r[d - 1], v);
    }
    push(r, v);
    return v;
}

template <int W>
__device__ __forceinline__ float forward_step(const float *__restrict__ c, float rhs, float (&r)[W]) {
    return forward_step_at<W>(c, rhs, r);
}
template <int W>
__device__ __forceinline__ float forward_step(ConstPtr<float> c, float rhs, float (&r)[W]) {
    return forward_step_at<W>(c, rhs, r);
}
template <int W>
__device__ __forceinline__ float backward_step(const float *__restrict__ c, float y, float (&r)[W]) {
    return backward_step_at<W>(c, y, r);
}
template <int W>
__device__ __forceinline__ float backward_step(ConstPtr<float> c, float y, float (&r)[W]) {
    return backward_step_at<W>(c, y, r);
}
template <int W>
__device__ __forceinline__ float forward_step(GlobalPtr<float> __restrict__ c, float rhs, float (&r)[W]) {
    return forward_step_at<W>(c, rhs, r);
}
template <int W>
__device__ __forceinline__ float backward_step(GlobalPtr<float> __restrict__ c, float y, float (&r)[W]) {
    return backward_step_at<W>(c, y, r);
}

// ---- chunk carry chain --------------------------------------------------------
// s <- e_q + T_q s for q in [first, last) (step +1) or (last, first] (step -1);
// e_q lives in LDS as [W][64] floats at base + q * stride.
template <int W>
__device__ __forceinline__ void carry_chain(float (&s)[W], const float *__restrict__ lds_e,
                                            int stride, const float *__restrict__ T, int q_begin,
                                            int q_end, int step, int lane) {
#pragma unroll
    for (int d = 0; d < W; ++d) s[d] = 0.0F;
    for (int q = q_begin; q != q_end; q += step) {
        const float *tq = T + q * W * W;
        const float *eq = lds_e + q * stride;
        float t[W];
#pragma unroll
        for (int r = 0; r < W; ++r) {
            float acc = eq[r * 64 + lane];
#pragma unroll
            for (int c = 0; c < W; ++c) acc = __builtin_fmaf(tq[r * W + c], s[c], acc);
            t[r] = acc;
        }
#pragma unroll
        for (int r = 0; r < W; ++r) s[r] = t[r];
    }
}

} // namespace dev
} // namespace dsmvc::hip

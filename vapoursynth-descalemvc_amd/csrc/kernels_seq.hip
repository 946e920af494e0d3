// gfx950 kernels of the HIP descale executor: sequential solves and the
// finiteness check.
//
// One axis pass computes, for every vector (a row for the horizontal pass, a
// column for the vertical pass), x = G^-1 A^T y with G = L D L^T banded:
//   rhs_i = sum_k w_k y[idx_k]                       (CSR gather, FMA in CSR order)
//   y_i   = (rhs_i - sum_d lld_d(i) y_{i-d}) / D_i   (forward + diagonal)
//   x_i   = y_i - sum_d ul_d(i) x_{i+d}              (backward)
// with single-rounding FMAs in the order of the reference's ordered solver
// (reference src/axis_plan.cpp:1068-1122), so a vector solved in one piece
// is bit-identical to it.
//
// Layout rule everywhere: a wavefront's 64 lanes own 64 vectors, so the plan
// row (CSR entries, coefficient record) is wave-uniform and arrives through
// scalar loads, and every global access is one contiguous 64-wide run.
//
// Kernel families
//   solve_seq        one lane per vector, whole vector in sequence; the
//                    exact route and the Float64 route.  Forward results go
//                    to a vector-contiguous scratch and are read back by the
//                    backward sweep.
//   columns_chunked  vertical pass.  Workgroup = 64 columns x C wavefronts;
//                    wavefront j owns rows [jL, jL+L) of those columns in
//                    registers.  Each wave solves its chunk from a zero carry,
//                    the W-element carries are chained across waves through
//                    LDS with the chunk transfer matrices, and each wave
//                    re-solves its chunk with the true carry.  Source read
//                    once, destination written once.
//   rows_chunked     horizontal pass, same partition with lanes = 64 rows.
//                    Source columns are staged through a per-wave LDS ring
//                    (transposed, pitch 65) and results leave through an LDS
//                    transpose, so HBM traffic stays coalesced.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "kernels_common.hpp"

namespace dsmvc::hip {
namespace dev {
namespace {

// ---- sequential solve (exact route, Float64 route) --------------------------
template <int W, class In, class Out, class T>
__global__ void __launch_bounds__(256)
solve_seq(DevAxis ax, const In *__restrict__ src, long s_vec, long s_el, long s_frame,
          Out *__restrict__ dst, long d_vec, long d_el, long d_frame, T *__restrict__ fwd,
          long f_vec, long f_el, long f_frame, int nvec, DevConv cv, int *__restrict__ flag) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvec) return;
    src += blockIdx.y * s_frame + v * s_vec;
    dst += blockIdx.y * d_frame + v * d_vec;
    fwd += blockIdx.y * f_frame + v * f_vec;
    const int n = ax.n;
    bool finite = true;
    T r[W];
#pragma unroll
    for (int d = 0; d < W; ++d) r[d] = T(0);
    for (int i = 0; i < n; ++i) {
        T acc = T(0);
        const std::uint32_t q1 = ax.off[i + 1];
        for (std::uint32_t q = ax.off[i]; q < q1; ++q) {
            const long at = static_cast<long>(ax.idx[q]) * s_el;
            if constexpr (sizeof(T) == 4) {
                acc = __builtin_fmaf(ax.wt[q], Load<In>::f32(src + at, cv), acc);
            } else {
                const double x = Load<In>::f64(src + at, cv);
                finite = finite && isfinite(x);
                acc = __builtin_fma(ax.wt64[q], x, acc);
            }
        }
        if constexpr (sizeof(T) == 4) {
            fwd[i * f_el] = forward_step<W>(ax.coef + static_cast<long>(i) * Band<W>::kStride, acc, r);
        } else {
            const double *c = ax.coef64 + static_cast<long>(i) * Band<W>::kStride;
#pragma unroll
            for (int d = W; d >= 1; --d) acc = __builtin_fma(-c[d - 1], r[d - 1], acc);
            const double y = __dmul_rn(acc, c[Band<W>::kInv]);
            fwd[i * f_el] = y;
            push(r, ax.zform ? acc : y);
        }
    }
#pragma unroll
    for (int d = 0; d < W; ++d) r[d] = T(0);
    for (int i = n - 1; i >= 0; --i) {
        const T y = fwd[i * f_el];
        if constexpr (sizeof(T) == 4) {
            const float x = backward_step<W>(ax.coef + static_cast<long>(i) * Band<W>::kStride, y, r);
            Store<Out>::f32(dst + i * d_el, x, cv);
        } else {
            const double *c = ax.coef64 + static_cast<long>(i) * Band<W>::kStride + Band<W>::kUp;
            double x = y;
#pragma unroll
            for (int d = W; d >= 1; --d) x = __builtin_fma(-c[d - 1], r[d - 1], x);
            push(r, x);
            finite = Store<Out>::f64(dst + i * d_el, x, cv) && finite;
        }
    }
    if (!finite && flag) atomicOr(flag, 1);
}

// Sets *flag when any element of a (rows x cols) plane is NaN or infinite.
// The Float64 route rejects such input, referenced or not, before solving.
template <class T>
__global__ void __launch_bounds__(256)
check_finite(const T *__restrict__ src, long s_row, long s_frame, int rows, int cols,
             int *__restrict__ flag) {
    const long total = static_cast<long>(rows) * cols;
    const T *f = src + blockIdx.y * s_frame;
    bool ok = true;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < total; e += 256L * gridDim.x) {
        const long r = e / cols, c = e - r * cols;
        ok = ok && isfinite(f[r * s_row + c]);
    }
    if (!ok) atomicOr(flag, 1);
}

// One workgroup per (plan batch, frame, row) of a table of F32 plane batches.
__global__ void __launch_bounds__(256)
check_finite_table(const CheckJob *__restrict__ jobs, int count) {
    const long item = blockIdx.x;
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (constant(jobs)[mid].first_item <= item) lo = mid;
        else hi = mid - 1;
    }
    const CheckJob jb = jobs[uniform(lo)];
    const long local = item - jb.first_item;
    const long frame = local / jb.rows, row = local - frame * jb.rows;
    const float *r = jb.src + frame * jb.s_frame + row * jb.s_row;
    bool ok = true;
    for (int c = static_cast<int>(threadIdx.x); c < jb.cols; c += 256) ok = ok && isfinite(r[c]);
    if (!ok) atomicOr(jb.flag, 1);
}

} // namespace
} // namespace dev

using namespace dev;

namespace {

template <int W, class In, class Out, class T>
hipError_t launch_seq_w(const SeqArgs &a, hipStream_t stream) {
    const dim3 grid((a.nvec + 255) / 256, a.frames);
    solve_seq<W, In, Out, T><<<grid, 256, 0, stream>>>(
        a.axis, static_cast<const In *>(a.src), a.s_vec, a.s_el, a.s_frame,
        static_cast<Out *>(a.dst), a.d_vec, a.d_el, a.d_frame, static_cast<T *>(a.fwd), a.f_vec,
        a.f_el, a.f_frame, a.nvec, a.conv, a.flag);
    return hipGetLastError();
}

// Coefficient slots past the plan's band are zero, so a band is solved by the
// next instantiated width of its class: the extra terms come first in the
// descending FMA chains and leave the value unchanged.  Band 2 has its own
// instantiation because the width-3 backward chain runs ascending.  (A
// one-sample destination has band 0 and runs as 1.)
template <class In, class Out, class T>
hipError_t launch_seq_t(const SeqArgs &a, hipStream_t s) {
    const int w = a.axis.w;
    if (w <= 1) return launch_seq_w<1, In, Out, T>(a, s);
    if (w == 2) return launch_seq_w<2, In, Out, T>(a, s);   // keeps descending order
    if (w == 3) return launch_seq_w<3, In, Out, T>(a, s);
    if (w <= 5) return launch_seq_w<5, In, Out, T>(a, s);
    if (w <= 7) return launch_seq_w<7, In, Out, T>(a, s);
    if (w <= 9) return launch_seq_w<9, In, Out, T>(a, s);
    if (w <= 11) return launch_seq_w<11, In, Out, T>(a, s);
    if (w <= 13) return launch_seq_w<13, In, Out, T>(a, s);
    if (w <= 15) return launch_seq_w<15, In, Out, T>(a, s);
    return hipErrorInvalidValue;
}

} // namespace

bool seq_band_supported(int w) { return w >= 0 && w <= kMaxWideBand; }

bool chunked_band_supported(int w) { return w == 1 || w == 3 || w == 5 || w == 7; }

int chunked_length_for(int needed) {
    for (const int l : {32, 48, 64, 80})
        if (needed <= l) return l;
    return 0;
}

int chunked_width_for(int taps) {
    for (const int k : {8, 16})
        if (taps <= k) return k;
    return 0;
}

int window_band_class(int w) {
    if (w < 0) return 0;
    if (w <= 1) return 1;
    if (w <= 3) return 3;
    if (w <= 5) return 5;
    if (w <= 7) return 7;
    return 0;
}

namespace {
hipError_t window64_dispatch(const Win64Job *jobs, int count, long items, const Win64Job &one,
                             int band, int width, bool rows, Sample kind, hipStream_t s) {
    switch (band) {
    case 1: return launch_window64_table_band<1>(jobs, count, items, one, width, rows, kind, s);
    case 3: return launch_window64_table_band<3>(jobs, count, items, one, width, rows, kind, s);
    case 5: return launch_window64_table_band<5>(jobs, count, items, one, width, rows, kind, s);
    case 7: return launch_window64_table_band<7>(jobs, count, items, one, width, rows, kind, s);
    default: return hipErrorInvalidValue;
    }
}
} // namespace

hipError_t launch_window64(const Win64Args &a, int band, int width, bool rows, Sample kind,
                           hipStream_t s) {
    Win64Job one{};
    one.plan = a.plan;
    one.src = a.src;
    one.s_row = a.s_row;
    one.s_frame = a.s_frame;
    one.dst = a.dst;
    one.d_row = a.d_row;
    one.d_frame = a.d_frame;
    one.nvec = a.nvec;
    one.frames = a.frames;
    one.groups = a.groups;
    one.flag = a.flag;
    one.conv = a.conv;
    const long items = assign_window64_spans(&one, 1, window64_table_slots(band, width, rows, kind));
    return window64_dispatch(nullptr, 0, items, one, band, width, rows, kind, s);
}

hipError_t launch_window64_table(const Win64Job *jobs, int count, long items, int band, int width,
                                 bool rows, Sample kind, hipStream_t s) {
    if (!jobs || count <= 0) return items > 0 ? hipErrorInvalidValue : hipSuccess;
    return window64_dispatch(jobs, count, items, Win64Job{}, band, width, rows, kind, s);
}

hipError_t launch_check_finite_table(const CheckJob *jobs, int count, long items, hipStream_t s) {
    if (items <= 0 || count <= 0) return hipSuccess;
    if (items > 0x7FFFFFFFL) return hipErrorInvalidValue;
    check_finite_table<<<dim3(static_cast<unsigned>(items)), 256, 0, s>>>(jobs, count);
    return hipGetLastError();
}

hipError_t launch_window_table(const WinJob *jobs, int count, long items, int band, int width,
                               bool rows, bool block16, Sample kind, int order, hipStream_t s) {
    switch (band * 100 + width) {
    case 106: return launch_window_table_band<1, 6>(jobs, count, items, rows, block16, kind, order, s);
    case 108: return launch_window_table_band<1, 8>(jobs, count, items, rows, block16, kind, order, s);
    case 116: return launch_window_table_band<1, 16>(jobs, count, items, rows, block16, kind, order, s);
    case 306: return launch_window_table_band<3, 6>(jobs, count, items, rows, block16, kind, order, s);
    case 308: return launch_window_table_band<3, 8>(jobs, count, items, rows, block16, kind, order, s);
    case 316: return launch_window_table_band<3, 16>(jobs, count, items, rows, block16, kind, order, s);
    case 508: return launch_window_table_band<5, 8>(jobs, count, items, rows, block16, kind, order, s);
    case 516: return launch_window_table_band<5, 16>(jobs, count, items, rows, block16, kind, order, s);
    case 708: return launch_window_table_band<7, 8>(jobs, count, items, rows, block16, kind, order, s);
    case 716: return launch_window_table_band<7, 16>(jobs, count, items, rows, block16, kind, order, s);
    default: return hipErrorInvalidValue;
    }
}

long window_table_slots(int band, int width, bool rows, bool block16, Sample kind) {
    switch (band * 100 + width) {
    case 106: return window_table_slots_band<1, 6>(rows, block16, kind);
    case 108: return window_table_slots_band<1, 8>(rows, block16, kind);
    case 116: return window_table_slots_band<1, 16>(rows, block16, kind);
    case 306: return window_table_slots_band<3, 6>(rows, block16, kind);
    case 308: return window_table_slots_band<3, 8>(rows, block16, kind);
    case 316: return window_table_slots_band<3, 16>(rows, block16, kind);
    case 508: return window_table_slots_band<5, 8>(rows, block16, kind);
    case 516: return window_table_slots_band<5, 16>(rows, block16, kind);
    case 708: return window_table_slots_band<7, 8>(rows, block16, kind);
    case 716: return window_table_slots_band<7, 16>(rows, block16, kind);
    default: return 0;
    }
}

// A longer span saves the forward warm-up (2P rows per segment) but leaves
// fewer wavefronts.  The span minimises a model of the launch: the rounds of
// `slots` resident wavefronts it takes, times the mean wavefront cost in
// forward rows (a backward row counts 0.3 of one).  DSMVC_HIP_WINDOW_SPAN
// (> 0) fixes the span instead.
namespace {
// Spans (chunks per wavefront) of a launch's jobs.  For a common target
// segment length of t chunks, each job splits its chunks into
// ceil(chunks / t) segments of near-equal length, so that a job with fewer
// chunks than another in the same launch is not left with a short last
// segment that still pays a full warm-up.  The target chosen minimises
// max(rounds of resident waves x mean work per wave, longest wave), with
// each segment's forward warm-up (2P rows) and backward warm-ups (P per
// chunk) counted as work.  `fixed` > 0 sets every job's span instead.
template <class Job, class Items>
long assign_spans(Job *jobs, int count, long slots, int fixed, Items items_of) {
    auto span_for = [](int chunks, int target) {
        const int segments = (chunks + target - 1) / target;
        return (chunks + segments - 1) / segments;
    };
    int best = fixed > 0 ? fixed : 1;
    if (fixed <= 0 && count > 0) {
        int top = 1;
        for (int i = 0; i < count; ++i) top = std::max(top, jobs[i].plan.chunks);
        top = std::min(top, 64);
        const double resident = static_cast<double>(std::max(slots, 1L));
        double best_time = 0.0;
        for (int target = 1; target <= top; ++target) {
            double waves = 0.0, work = 0.0, longest = 0.0;
            for (int i = 0; i < count; ++i) {
                const auto &p = jobs[i].plan;
                const int span = span_for(p.chunks, target);
                const int segments = (p.chunks + span - 1) / span;
                const double per = static_cast<double>(jobs[i].frames) * jobs[i].groups * segments;
                const double chunks = static_cast<double>(p.chunks) / segments;   // mean per segment
                const double each = chunks * p.length + 2.0 * p.warmup
                                    + 0.3 * chunks * (p.length + p.warmup);
                waves += per;
                work += per * each;
                longest = std::max(longest, span * p.length + 2.0 * p.warmup
                                                + 0.3 * span * (p.length + p.warmup));
            }
            if (waves <= 0.0) break;
            const double time = std::max(std::ceil(waves / resident) * (work / waves), longest);
            if (target == 1 || time < best_time * 0.999) {
                best_time = time;
                best = target;
            }
        }
    }
    long items = 0;
    for (int i = 0; i < count; ++i) {
        jobs[i].span = fixed > 0 ? fixed : span_for(jobs[i].plan.chunks, best);
        jobs[i].first_item = items;
        items += items_of(jobs[i]);
    }
    return items;
}

int span_override(const char *name) {
    const char *env = std::getenv(name);
    return env ? std::atoi(env) : 0;
}
} // namespace

long assign_window_spans(WinJob *jobs, int count, long slots) {
    static const int fixed = span_override("DSMVC_HIP_WINDOW_SPAN");
    return assign_spans(jobs, count, slots, fixed, [](const WinJob &j) { return window_items(j); });
}

long assign_window64_spans(Win64Job *jobs, int count, long slots) {
    static const int fixed = span_override("DSMVC_HIP_WINDOW64_SPAN");
    return assign_spans(jobs, count, slots, fixed, [](const Win64Job &j) { return window64_items(j); });
}

long window64_table_slots(int band, int width, bool rows, Sample kind) {
    switch (band) {
    case 1: return window64_table_slots_band<1>(width, rows, kind);
    case 3: return window64_table_slots_band<3>(width, rows, kind);
    case 5: return window64_table_slots_band<5>(width, rows, kind);
    case 7: return window64_table_slots_band<7>(width, rows, kind);
    default: return 0;
    }
}

hipError_t launch_chunked(const ChunkArgs &a, bool rows, Sample kind, hipStream_t s) {
    switch (a.axis.w) {
    case 1: return launch_chunked_band<1>(a, rows, kind, s);
    case 3: return launch_chunked_band<3>(a, rows, kind, s);
    case 5: return launch_chunked_band<5>(a, rows, kind, s);
    case 7: return launch_chunked_band<7>(a, rows, kind, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_seq(const SeqArgs &a, Sample in, Sample out, bool f64, hipStream_t s) {
    if (!f64) {
        if (out == Sample::f32) {
            switch (in) {
            case Sample::f32: return launch_seq_t<float, float, float>(a, s);
            case Sample::u8: return launch_seq_t<std::uint8_t, float, float>(a, s);
            case Sample::u16: return launch_seq_t<std::uint16_t, float, float>(a, s);
            default: break;
            }
        } else if (in == Sample::f32) {
            if (out == Sample::u8) return launch_seq_t<float, std::uint8_t, float>(a, s);
            if (out == Sample::u16) return launch_seq_t<float, std::uint16_t, float>(a, s);
        }
        return hipErrorInvalidValue;
    }
    if (out == Sample::f64) {
        switch (in) {
        case Sample::f32: return launch_seq_t<float, double, double>(a, s);
        case Sample::u8: return launch_seq_t<std::uint8_t, double, double>(a, s);
        case Sample::u16: return launch_seq_t<std::uint16_t, double, double>(a, s);
        default: return hipErrorInvalidValue;
        }
    }
    if (in == Sample::f64) {
        switch (out) {
        case Sample::f32: return launch_seq_t<double, float, double>(a, s);
        case Sample::u8: return launch_seq_t<double, std::uint8_t, double>(a, s);
        case Sample::u16: return launch_seq_t<double, std::uint16_t, double>(a, s);
        default: return hipErrorInvalidValue;
        }
    }
    if (in == Sample::f32 && out == Sample::f32) return launch_seq_t<float, float, double>(a, s);
    return hipErrorInvalidValue;
}

hipError_t launch_check_finite(const void *src, Sample in, long s_row, long s_frame, int rows,
                               int cols, int frames, int *flag, hipStream_t s) {
    const long total = static_cast<long>(rows) * cols;
    const int blocks = static_cast<int>(std::min<long>((total + 255) / 256, 1024));
    const dim3 grid(blocks > 0 ? blocks : 1, frames);
    if (in == Sample::f32)
        check_finite<float><<<grid, 256, 0, s>>>(static_cast<const float *>(src), s_row, s_frame,
                                                 rows, cols, flag);
    else if (in == Sample::f64)
        check_finite<double><<<grid, 256, 0, s>>>(static_cast<const double *>(src), s_row,
                                                  s_frame, rows, cols, flag);
    else
        return hipSuccess;   // integer samples are always finite
    return hipGetLastError();
}

} // namespace dsmvc::hip

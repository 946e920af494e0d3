// inverse_2d_batch without allocation (HipExecutor::inverse_2d_batch_noalloc).
//
// Some callers cannot allocate device memory during a call: code under HIP
// graph capture, tight serving loops, hosts that forbid it.  For them every
// buffer the 2D route needs beyond the caller's input and output -- the
// pass-to-pass intermediate, the sequential kernels' forward sweeps, the
// windowed job records and the Float64 route's non-finite flag -- lives in
// one fixed device array per GPU, `scratch` below, held until the process
// ends.  Its device address is looked up once per device, with that device
// current, when an executor is created; host code uses only that address.
//
// Layout of the array (all offsets 256-byte aligned):
//   [0, 256)                 the Float64 route's flag (one int)
//   [256, 256 + 2 jobs)      the windowed job records of the two passes
//   [kNoallocDataOffset, ..) the intermediate, then the forward sweeps
//
// Calls on one GPU share the array and it keeps the last call's contents:
// the flag is cleared by a fill kernel on each Float64-route call (no
// memset node, so the call can be captured into a graph), job records are
// written by a kernel that takes them by value, and calls that overlap on
// the device race.
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>

#include "device_plan.hpp"
#include "executor.hpp"
#include "kernels.hpp"

namespace dsmvc {

namespace {

constexpr std::size_t kNoallocScratchBytes = std::size_t{64} << 20;   // 64 MiB per GPU
constexpr std::size_t kNoallocJobOffset = 256;
constexpr std::size_t kNoallocDataOffset = 4096;
static_assert(kNoallocJobOffset + 2 * sizeof(hip::WinJob) <= kNoallocDataOffset,
              "two job records fit before the data region");

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_setup_mutex;
std::array<void *, kMaxDevices> g_scratch{};   // device address of `scratch` per device

void check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("dsmvc: ") + what + ": " + hipGetErrorString(e));
}

std::size_t align256(std::size_t v) { return (v + 255U) & ~std::size_t{255}; }

// Writes the two passes' job records into the array: taken by value as
// kernel arguments, so a captured graph replays them without host memory.
__global__ void __launch_bounds__(64)
store_jobs(hip::WinJob rows, hip::WinJob columns, hip::WinJob *out) {
    static_assert(sizeof(hip::WinJob) % 4 == 0, "job record is whole words");
    constexpr int words = static_cast<int>(sizeof(hip::WinJob) / 4);
    const auto *a = reinterpret_cast<const std::uint32_t *>(&rows);
    const auto *b = reinterpret_cast<const std::uint32_t *>(&columns);
    auto *o = reinterpret_cast<std::uint32_t *>(out);
    for (int i = static_cast<int>(threadIdx.x); i < words; i += 64) {
        o[i] = a[i];
        o[words + i] = b[i];
    }
}

__global__ void __launch_bounds__(64) clear_flag(int *flag) {
    if (threadIdx.x == 0) *flag = 0;
}

hip::DevConv conversion_of(const IntegerConversion *c) {
    hip::DevConv d{};
    if (c) {
        d.in_offset = c->input_offset;
        d.in_scale = c->input_scale;
        d.out_scale = c->output_scale;
        d.out_offset = c->output_offset;
        d.out_max = static_cast<float>(c->output_maximum);
    } else {
        d.in_scale = 1.0F;
        d.out_scale = 1.0F;
    }
    return d;
}

// The array's address on `device` (made current here).  Host code never
// uses the array's name itself, which is only a host-side stand-in.
void *lookup_scratch(int device) {
    int current = 0;
    check(hipGetDevice(&current), "hipGetDevice failed");
    check(hipSetDevice(device), "hipSetDevice failed");
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    (void)hipSetDevice(current);
    check(e, "no-allocation scratch lookup failed");
    if (!ptr) throw std::runtime_error("dsmvc: no-allocation scratch lookup returned null");
    return ptr;
}

hip::Sample sample_of(SampleKind k) {
    switch (k) {
    case SampleKind::u8: return hip::Sample::u8;
    case SampleKind::u16: return hip::Sample::u16;
    case SampleKind::float32: break;
    }
    return hip::Sample::f32;
}

} // namespace

void *HipExecutor::noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices)
        throw std::invalid_argument("dsmvc: HIP device index out of range");
    std::lock_guard lock(g_setup_mutex);
    if (!g_scratch[static_cast<std::size_t>(device)])
        g_scratch[static_cast<std::size_t>(device)] = lookup_scratch(device);
    return g_scratch[static_cast<std::size_t>(device)];
}

void HipExecutor::inverse_2d_batch_noalloc(const AxisPlan &h, const AxisPlan &v, SampleKind kind,
                                           const PlaneBatch &b,
                                           const IntegerConversion *conversion,
                                           hipStream_t stream, bool async) {
    const bool integer = kind != SampleKind::float32;
    if (!h.valid() || !v.valid() || !b.input || !b.output || b.frames < 1
        || b.input_row_stride < h.source_size || b.output_row_stride < h.destination_size
        || (b.frames > 1 && (b.input_frame_stride < b.input_row_stride * v.source_size
                             || b.output_frame_stride < b.output_row_stride * v.destination_size)))
        throw std::invalid_argument("dsmvc: invalid 2D executor arguments");
    if (integer) {
        if (!conversion) throw std::invalid_argument("dsmvc: integer 2D execution needs a conversion");
        const std::uint32_t top = kind == SampleKind::u8 ? 255U : 65535U;
        if (!(conversion->input_scale > 0.0F) || !(conversion->output_scale > 0.0F)
            || conversion->output_maximum == 0U || conversion->output_maximum > top)
            throw std::invalid_argument("dsmvc: invalid integer 2D executor arguments");
    }
    if (mode_ == SolveMode::refined)
        throw std::invalid_argument("dsmvc: the no-allocation call does not run refined solves");
    std::lock_guard lock(mutex_);
    check(hipSetDevice(device_), "hipSetDevice failed");
    hip::DevicePlan *hp = prepared_plan(h), *vp = prepared_plan(v);
    if (!hp || !vp)
        throw std::invalid_argument("dsmvc: the no-allocation call needs both plans prepared on "
                                    "this executor (dsmvc_hip_executor_prepare)");
    const bool f64 = float64_route(h) || float64_route(v);
    if (f64 && async)
        throw std::invalid_argument("dsmvc: the no-allocation call takes the Float64 route "
                                    "synchronously only (it reports non-finite results on return)");
    const bool windows = automatic_like(mode_) || mode_ == SolveMode::windowed;

    // Sizes of the scratch regions, checked against the array before any
    // launch: the intermediate [frame][v.m][pitch(h.n)] and, for sequential
    // passes, the larger forward sweep.
    const std::size_t elem = f64 ? 8 : 4;
    const std::ptrdiff_t inter_row = (static_cast<std::ptrdiff_t>(h.destination_size) + 63) & ~std::ptrdiff_t{63};
    const std::ptrdiff_t inter_frame = inter_row * v.source_size;
    const std::size_t inter_bytes = align256(static_cast<std::size_t>(inter_frame) * b.frames * elem);
    const std::size_t fwd_bytes = align256(
        std::max(static_cast<std::size_t>(v.source_size) * h.destination_size,
                 static_cast<std::size_t>(h.destination_size) * v.destination_size)
        * b.frames * elem);
    const hip::DevWindow *wh = (!f64 && windows) ? hp->built_window(true) : nullptr;
    const hip::DevWindow *wv = (!f64 && windows) ? vp->built_window(false) : nullptr;
    const bool windowed = wh && wv;
    const bool need_fwd = !windowed;
    const std::size_t needed = kNoallocDataOffset + inter_bytes + (need_fwd ? fwd_bytes : 0);
    if (needed > kNoallocScratchBytes)
        throw std::invalid_argument("dsmvc: the no-allocation call needs " + std::to_string(needed)
                                    + " scratch bytes; its fixed array holds "
                                    + std::to_string(kNoallocScratchBytes));
    if (mode_ == SolveMode::windowed && !windowed)
        throw std::invalid_argument("dsmvc: plan cannot use the windowed HIP solve");

    auto *base = static_cast<unsigned char *>(noalloc_scratch_);
    int *flag = reinterpret_cast<int *>(base);
    auto *jobs = reinterpret_cast<hip::WinJob *>(base + kNoallocJobOffset);
    void *inter = base + kNoallocDataOffset;
    void *fwd = base + kNoallocDataOffset + inter_bytes;
    const hip::DevConv conv = conversion_of(conversion);
    const hip::Sample in_kind = sample_of(kind);

    if (windowed) {
        hip::WinJob a{};
        a.plan = *wh;
        a.src = b.input;
        a.s_row = b.input_row_stride;
        a.s_frame = b.input_frame_stride;
        a.dst = inter;
        a.d_row = inter_row;
        a.d_frame = inter_frame;
        a.nvec = v.source_size;
        a.frames = b.frames;
        a.groups = (a.nvec + 63) / 64;
        a.conv = conv;
        hip::WinJob c{};
        c.plan = *wv;
        c.src = inter;
        c.s_row = inter_row;
        c.s_frame = inter_frame;
        c.dst = b.output;
        c.d_row = b.output_row_stride;
        c.d_frame = b.output_frame_stride;
        c.nvec = h.destination_size;
        c.frames = b.frames;
        c.groups = (c.nvec + 63) / 64;
        c.conv = conv;
        const int band_h = hip::window_band_class(h.half_bandwidth);
        const int band_v = hip::window_band_class(v.half_bandwidth);
        const long items_h = hip::assign_window_spans(
            &a, 1, hip::window_table_slots(band_h, a.plan.ell_width, true, a.plan.block16 != 0, in_kind));
        const long items_v = hip::assign_window_spans(
            &c, 1, hip::window_table_slots(band_v, c.plan.ell_width, false, c.plan.block16 != 0, in_kind));
        store_jobs<<<1, 64, 0, stream>>>(a, c, jobs);
        check(hipGetLastError(), "job record launch failed");
        constexpr int order = 2;   // XCD-contiguous workgroups, vector group fastest
        check(hip::launch_window_table(jobs, 1, items_h, band_h, a.plan.ell_width, true,
                                       a.plan.block16 != 0, in_kind, order, stream),
              "windowed launch failed");
        check(hip::launch_window_table(jobs + 1, 1, items_v, band_v, c.plan.ell_width, false,
                                       c.plan.block16 != 0, in_kind, order, stream),
              "windowed launch failed");
    } else {
        // Sequential kernels (the exact route; the Float64 route's
        // sequential kernels), forward sweeps in the scratch array.
        if (!hip::seq_band_supported(h.half_bandwidth) || !hip::seq_band_supported(v.half_bandwidth))
            throw std::invalid_argument("dsmvc: half bandwidth exceeds the HIP executor's limit");
        if (f64) {
            clear_flag<<<1, 64, 0, stream>>>(flag);
            check(hipGetLastError(), "flag clear launch failed");
            check(hip::launch_check_finite(b.input, in_kind, b.input_row_stride,
                                           b.input_frame_stride, v.source_size, h.source_size,
                                           b.frames, flag, stream),
                  "check launch failed");
        }
        const hip::Sample inter_kind = f64 ? hip::Sample::f64 : hip::Sample::f32;
        // The Float64 route of an F32 (or integer) 2D plane: each pass on the
        // windowed FP64 kernels where its plan's prepared tables take it.
        const hip::DevWindow64 *h64 = (f64 && windows) ? hp->built_window64() : nullptr;
        const hip::DevWindow64 *v64 = (f64 && windows) ? vp->built_window64() : nullptr;
        auto pass = [&](const AxisPlan &plan, hip::DevicePlan &dp, const hip::DevWindow64 *w64,
                        bool rows, const void *src, hip::Sample sk, long s_row, long s_frame,
                        void *dst, hip::Sample dk, long d_row, long d_frame, int nvec) {
            if (w64) {
                hip::Win64Args a{};
                a.plan = *w64;
                a.src = src;
                a.s_row = s_row;
                a.s_frame = s_frame;
                a.dst = dst;
                a.d_row = d_row;
                a.d_frame = d_frame;
                a.nvec = nvec;
                a.frames = b.frames;
                a.groups = (nvec + 63) / 64;
                a.flag = flag;
                a.conv = conv;
                check(hip::launch_window64(a, hip::window_band_class(plan.half_bandwidth),
                                           w64->ell_width, rows, rows ? sk : dk, stream),
                      "windowed Float64 launch failed");
                return;
            }
            hip::SeqArgs a{};
            a.axis = dp.axis();
            a.src = src;
            a.dst = dst;
            a.nvec = nvec;
            a.frames = b.frames;
            a.conv = conv;
            a.flag = f64 ? flag : nullptr;
            const long n = plan.destination_size;
            if (rows) {
                a.s_vec = s_row; a.s_el = 1; a.s_frame = s_frame;
                a.d_vec = d_row; a.d_el = 1; a.d_frame = d_frame;
            } else {
                a.s_vec = 1; a.s_el = s_row; a.s_frame = s_frame;
                a.d_vec = 1; a.d_el = d_row; a.d_frame = d_frame;
            }
            a.fwd = fwd;
            a.f_vec = 1;
            a.f_el = nvec;
            a.f_frame = static_cast<long>(nvec) * n;
            check(hip::launch_seq(a, sk, dk, f64, stream), "sequential solve launch failed");
        };
        pass(h, *hp, h64, true, b.input, in_kind, b.input_row_stride, b.input_frame_stride,
             inter, inter_kind, inter_row, inter_frame, v.source_size);
        pass(v, *vp, v64, false, inter, inter_kind, inter_row, inter_frame, b.output, in_kind,
             b.output_row_stride, b.output_frame_stride, h.destination_size);
    }
    check(hipGetLastError(), "kernel launch failed");
    if (async) return;
    check(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize(), "execution failed");
    if (f64) {
        // On the call's stream: a null-stream copy would wait for every other
        // blocking stream's work too.
        int bad = 0;
        if (stream) {
            check(hipMemcpyAsync(&bad, flag, sizeof bad, hipMemcpyDeviceToHost, stream),
                  "flag readback failed");
            check(hipStreamSynchronize(stream), "flag readback failed");
        } else {
            check(hipMemcpy(&bad, flag, sizeof bad, hipMemcpyDeviceToHost), "flag readback failed");
        }
        if (bad) throw std::runtime_error("dsmvc: HIP Float64 2D execution produced NaN or infinity");
    }
}

} // namespace dsmvc

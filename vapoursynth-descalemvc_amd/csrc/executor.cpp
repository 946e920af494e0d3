// HIP executor: route selection, scratch management and launches.
#include "executor.hpp"
#include "exported.hpp"
#include "trace.hpp"

#include <algorithm>
#include <bit>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <stdexcept>
#include <string>

#include "batch.hpp"
#include "kernels.hpp"

namespace dsmvc {

using hip::DevConv;
using hip::Sample;

namespace {

void check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("dsmvc: ") + what + ": " + hipGetErrorString(e));
}

std::size_t align256(std::size_t v) { return (v + 255U) & ~std::size_t{255}; }

// Reads a Float64 flag (into pinned `host`) once its stream has finished.
// On the stream itself: a null-stream copy would first wait for the work of
// every other blocking stream (the other execution slots).
int read_flag(hipStream_t stream, const int *flag, int *host) {
    if (stream) {
        check(hipMemcpyAsync(host, flag, sizeof *host, hipMemcpyDeviceToHost, stream),
              "flag readback failed");
        check(hipStreamSynchronize(stream), "flag readback failed");
    } else {
        check(hipMemcpy(host, flag, sizeof *host, hipMemcpyDeviceToHost), "flag readback failed");
    }
    return *host;
}

std::ptrdiff_t sample_size(hip::Sample kind) {
    switch (kind) {
    case hip::Sample::u8: return 1;
    case hip::Sample::u16: return 2;
    case hip::Sample::f64: return 8;
    case hip::Sample::f32: break;
    }
    return 4;
}

std::ptrdiff_t padded_width(std::int32_t n) { return (static_cast<std::ptrdiff_t>(n) + 63) & ~std::ptrdiff_t{63}; }

Sample sample_of(SampleKind k) {
    switch (k) {
    case SampleKind::float32: return Sample::f32;
    case SampleKind::u8: return Sample::u8;
    case SampleKind::u16: return Sample::u16;
    }
    return Sample::f32;
}

DevConv make_conv(const IntegerConversion *c) {
    DevConv d{};
    if (c) {
        d.in_offset = c->input_offset;
        d.in_scale = c->input_scale;
        d.out_scale = c->output_scale;
        d.out_offset = c->output_offset;
        d.out_max = static_cast<float>(c->output_maximum);
    } else {
        d.in_scale = 1.0F;
        d.out_scale = 1.0F;
        d.out_max = 0.0F;
    }
    return d;
}

bool same_plan(const AxisPlan &a, const AxisPlan &b) {
    // Built plans carry a unique serial: equal serials, equal plans.
    if (a.serial != 0 || b.serial != 0) return a.serial == b.serial;
    return a.source_size == b.source_size && a.destination_size == b.destination_size
        && a.half_bandwidth == b.half_bandwidth
        && a.transpose_indices.size() == b.transpose_indices.size()
        && std::bit_cast<std::uint64_t>(a.normal_rcond) == std::bit_cast<std::uint64_t>(b.normal_rcond)
        && std::bit_cast<std::uint64_t>(a.active_length) == std::bit_cast<std::uint64_t>(b.active_length)
        && std::bit_cast<std::uint64_t>(a.shift) == std::bit_cast<std::uint64_t>(b.shift)
        && a.requires_float64() == b.requires_float64()
        && a.transpose_weights == b.transpose_weights && a.inverse_diagonal == b.inverse_diagonal;
}

void validate_conversion(const IntegerConversion &c, std::uint32_t sample_max, const char *what) {
    if (!(c.input_scale > 0.0F) || !(c.output_scale > 0.0F) || !std::isfinite(c.input_offset)
        || !std::isfinite(c.input_scale) || !std::isfinite(c.output_scale)
        || !std::isfinite(c.output_offset) || c.output_maximum == 0U || c.output_maximum > sample_max)
        throw std::invalid_argument(what);
}

} // namespace

bool hip_available() noexcept {
    int count = 0;
    return hipGetDeviceCount(&count) == hipSuccess && count > 0;
}

HipExecutor::HipExecutor(int device, SolveMode mode) : device_(device), mode_(mode) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        throw std::runtime_error("dsmvc: backend hip is unavailable: no HIP device");
    if (device_ < 0) check(hipGetDevice(&device_), "hipGetDevice failed");
    if (device_ >= count) throw std::invalid_argument("dsmvc: HIP device index out of range");
    check(hipSetDevice(device_), "hipSetDevice failed");
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, device_) == hipSuccess && prop.multiProcessorCount > 0)
        compute_units_ = prop.multiProcessorCount;
    if (const char *env = std::getenv("DSMVC_HIP_AUTOTUNE")) autotune_ = std::string(env) != "0";
    if (const char *env = std::getenv("DSMVC_HIP_FUSED")) fused_auto_ = std::string(env) != "0";
    if (const char *env = std::getenv("DSMVC_HIP_PLAN_CACHE_MB")) {
        const long mb = std::strtol(env, nullptr, 10);
        if (mb > 0) plan_cache_limit_ = static_cast<std::size_t>(mb) << 20;
    }
    noalloc_scratch_ = noalloc_setup(device_);
}

HipExecutor::~HipExecutor() {
    // The per-stream entries are keyed by streams this executor does not
    // own; their owners may have destroyed them already, so only the device
    // as a whole is waited for.
    (void)hipSetDevice(device_);
    (void)hipDeviceSynchronize();
    imports_.reset();
    plane_tables_.clear();
    for (auto &entry : scratch_) {
        if (entry.second.buffer) (void)hipFree(entry.second.buffer);
        if (entry.second.flag) (void)hipFree(entry.second.flag);
        if (entry.second.flag_host) (void)hipHostFree(entry.second.flag_host);
    }
    scratch_.clear();
    for (Slot &slot : slots_) {
        if (slot.stream) (void)hipStreamDestroy(slot.stream);
        if (slot.pinned) (void)hipHostFree(slot.pinned);
        if (slot.device) (void)hipFree(slot.device);
    }
    plans_.clear();
}

void HipExecutor::release_stream(hipStream_t stream) {
    std::lock_guard lock(mutex_);
    check(hipSetDevice(device_), "hipSetDevice failed");
    check(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize(), "stream synchronize failed");
    plane_tables_.erase(stream);
    if (auto it = scratch_.find(stream); it != scratch_.end()) {
        if (it->second.buffer) (void)hipFree(it->second.buffer);
        if (it->second.flag) (void)hipFree(it->second.flag);
        if (it->second.flag_host) (void)hipHostFree(it->second.flag_host);
        scratch_.erase(it);
    }
}

void HipExecutor::prepare(const std::shared_ptr<const AxisPlan> &plan) {
    if (!plan || !plan->valid())
        throw std::invalid_argument("dsmvc: cannot prepare an invalid HIP axis plan");
    std::lock_guard lock(mutex_);
    const auto image = device_plan(*plan);
    hip::DevicePlan &dp = *image;
    // The windowed tables too, so that a prepared plan runs without any
    // further device allocation (inverse_2d_batch_noalloc relies on this).
    if (automatic_like(mode_) || mode_ == SolveMode::windowed) {
        (void)dp.window(false);
        if (plan->requires_float64()) (void)dp.window64();
    }
}

void HipExecutor::seal() {
    std::lock_guard lock(mutex_);
    sealed_ = true;
}

std::uint64_t HipExecutor::plan_key(const AxisPlan &plan) {
    // Built and imported plans carry a unique serial; other plans are keyed
    // by address (top bit set) and checked by content on every hit.
    if (plan.serial != 0) return plan.serial & ~(std::uint64_t{1} << 63);
    return static_cast<std::uint64_t>(reinterpret_cast<std::uintptr_t>(&plan)) | (std::uint64_t{1} << 63);
}

std::shared_ptr<hip::DevicePlan> HipExecutor::device_plan(const AxisPlan &plan) {
    const trace::Range range(trace::Label::plan_cache_lookup);
    const std::uint64_t key = plan_key(plan);
    auto it = plans_.find(key);
    if (it != plans_.end() && same_plan(it->second.image->host(), plan)) {
        lru_.splice(lru_.begin(), lru_, it->second.lru);
        ++plan_stats_.hits;
        trace::mark("dsmvc.hip: plan cache hit");
        return it->second.image;
    }
    if (it != plans_.end()) {
        // The address was reused by another plan: the old image goes (once
        // the device has finished with it) and is rebuilt.
        (void)hipDeviceSynchronize();
        const hip::DevicePlan *stale = it->second.image.get();
        std::erase_if(tuned_, [stale](const auto &entry) { return entry.first.plan == stale; });
        lru_.erase(it->second.lru);
        plans_.erase(it);
    }
    check(hipSetDevice(device_), "hipSetDevice failed");
    trace::mark("dsmvc.hip: plan cache miss");
    const trace::Range upload(trace::Label::plan_upload);
    auto image = std::make_shared<hip::DevicePlan>(plan, device_, cache_bytes_);
    lru_.push_front(key);
    plans_[key] = PlanEntry{image, lru_.begin()};
    ++plan_stats_.misses;
    evict_plans();
    return image;
}

// Once the executor's images exceed the cache limit, drops least recently
// used images until they are down to three quarters of it, skipping images
// someone holds (a call in progress, a batch, a served request).  The device
// is synchronized first, so no queued kernel still reads an image that is
// freed; evicting a quarter of the limit at a time keeps those
// synchronizations rare.  The byte total is kept by the images themselves,
// so the check costs the same however many images are cached.
void HipExecutor::evict_plans() {
    std::size_t total = cache_bytes_->load(std::memory_order_relaxed);
    if (total <= plan_cache_limit_) return;
    const std::size_t target = plan_cache_limit_ / 4 * 3;
    std::vector<std::uint64_t> victims;
    for (auto k = lru_.rbegin(); k != lru_.rend() && total > target; ++k) {
        const PlanEntry &e = plans_.at(*k);
        if (e.image.use_count() > 1) continue;
        total -= e.image->total_bytes();
        victims.push_back(*k);
    }
    if (victims.empty()) return;
    const trace::Range range(trace::Label::plan_evict);
    (void)hipDeviceSynchronize();
    for (const std::uint64_t key : victims) {
        auto it = plans_.find(key);
        const hip::DevicePlan *gone = it->second.image.get();
        std::erase_if(tuned_, [gone](const auto &entry) { return entry.first.plan == gone; });
        lru_.erase(it->second.lru);
        plans_.erase(it);
        ++plan_stats_.evictions;
    }
}

HipExecutor::PlanCacheStats HipExecutor::plan_cache_stats() {
    std::lock_guard lock(mutex_);
    PlanCacheStats out = plan_stats_;
    out.entries = plans_.size();
    out.bytes = 0;
    for (auto &entry : plans_) out.bytes += entry.second.image->total_bytes();
    out.limit = plan_cache_limit_;
    return out;
}

void HipExecutor::set_plan_cache_limit(std::size_t bytes) {
    if (bytes == 0) throw std::invalid_argument("dsmvc: the plan cache limit must be positive");
    std::lock_guard lock(mutex_);
    check(hipSetDevice(device_), "hipSetDevice failed");
    plan_cache_limit_ = bytes;
    evict_plans();
}

// ---------------------------------------------------------------------------
// Execution slots
// ---------------------------------------------------------------------------
HipExecutor::Slot *HipExecutor::acquire_slot() {
    const trace::Range range(trace::Label::slot_wait);
    std::unique_lock lock(slot_mutex_);
    for (;;) {
        for (Slot &slot : slots_) {
            if (slot.busy) continue;
            if (!slot.stream) {
                // Default priority and flags: ordered with the null stream,
                // as the synchronous calls it stands in for were.
                check(hipSetDevice(device_), "hipSetDevice failed");
                check(hipStreamCreate(&slot.stream), "slot stream creation failed");
            }
            slot.busy = true;
            return &slot;
        }
        slot_cv_.wait(lock);
    }
}

void HipExecutor::release_slot(Slot *slot) {
    {
        std::lock_guard lock(slot_mutex_);
        slot->busy = false;
    }
    slot_cv_.notify_one();
}

void HipExecutor::slot_buffers(Slot &slot, std::size_t pinned, std::size_t device) {
    // The slot is idle here (its last call waited for its stream).
    check(hipSetDevice(device_), "hipSetDevice failed");
    if (slot.pinned_bytes < pinned) {
        if (slot.pinned) check(hipHostFree(slot.pinned), "staging release failed");
        slot.pinned = nullptr;
        slot.pinned_bytes = 0;
        const std::size_t grown = align256(std::max(pinned, slot.pinned_bytes * 3 / 2));
        check(hipHostMalloc(&slot.pinned, grown), "staging allocation failed");
        slot.pinned_bytes = grown;
    }
    if (slot.device_bytes < device) {
        if (slot.device) check(hipFree(slot.device), "staging release failed");
        slot.device = nullptr;
        slot.device_bytes = 0;
        const std::size_t grown = align256(std::max(device, slot.device_bytes * 3 / 2));
        check(hipMalloc(&slot.device, grown), "staging allocation failed");
        slot.device_bytes = grown;
    }
}

namespace {
struct SlotGuard {
    std::function<void()> release;
    ~SlotGuard() {
        if (release) release();
    }
};
} // namespace

template <class Enqueue>
void HipExecutor::run_call(hipStream_t stream, bool async, const char *what, Enqueue &&enqueue) {
    Slot *slot = (!stream && !async) ? acquire_slot() : nullptr;
    SlotGuard guard;
    if (slot) guard.release = [this, slot] { release_slot(slot); };
    const hipStream_t s = slot ? slot->stream : stream;
    Scratch *used = nullptr;
    {
        std::lock_guard lock(mutex_);
        check(hipSetDevice(device_), "hipSetDevice failed");
        used = &enqueue(s);
        check(hipGetLastError(), "kernel launch failed");
        // Tables built by this call count against the cache's limit too.
        evict_plans();
    }
    if (async) return;
    complete(s, *used, what);
}

// Waits for `stream` without the executor's lock and reports a Float64
// non-finite result.  The scratch entry belongs to the stream; a stream is
// used by one call at a time (slots are leased, caller streams are the
// caller's to order).
void HipExecutor::complete(hipStream_t stream, Scratch &s, const char *what) {
    const trace::Range range(trace::Label::stream_synchronize);
    check(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize(), "execution failed");
    if (s.flag_armed) {
        s.flag_armed = false;
        if (read_flag(stream, s.flag, s.flag_host)) throw std::runtime_error(what);
    }
}

hip::DevicePlan *HipExecutor::prepared_plan(const AxisPlan &plan) {
    // Looks the plan's device image up without building one.
    auto it = plans_.find(plan_key(plan));
    if (it == plans_.end() || !same_plan(it->second.image->host(), plan)) return nullptr;
    lru_.splice(lru_.begin(), lru_, it->second.lru);
    return it->second.image.get();
}

WindowTables &HipExecutor::stream_tables(hipStream_t stream) {
    auto &slot = plane_tables_[stream];
    if (!slot) slot = std::make_unique<WindowTables>(device_);
    return *slot;
}

std::shared_ptr<hip::DevicePlan> HipExecutor::plan_image(const AxisPlan &plan) {
    if (!plan.valid()) throw std::invalid_argument("dsmvc: invalid HIP axis plan");
    std::lock_guard lock(mutex_);
    check(hipSetDevice(device_), "hipSetDevice failed");
    return device_plan(plan);
}

HipExecutor::Scratch &HipExecutor::scratch(hipStream_t stream, std::size_t bytes) {
    Scratch &s = scratch_[stream];
    if (!s.flag) check(hipMalloc(reinterpret_cast<void **>(&s.flag), 256), "scratch allocation failed");
    if (!s.flag_host)
        check(hipHostMalloc(reinterpret_cast<void **>(&s.flag_host), sizeof(int)),
              "scratch allocation failed");
    if (s.bytes < bytes) {
        const trace::Range range(trace::Label::buffer_reserve);
        const std::size_t grown = align256(std::max(bytes, s.bytes + s.bytes / 2));
        if (s.buffer) {
            check(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize(), "stream synchronize failed");
            check(hipFree(s.buffer), "scratch release failed");
            s.buffer = nullptr;
            s.bytes = 0;
        }
        check(hipMalloc(&s.buffer, grown), "scratch allocation failed");
        s.bytes = grown;
    }
    return s;
}

void HipExecutor::finish(hipStream_t stream, Scratch &s, bool async, const char *what) {
    check(hipGetLastError(), "kernel launch failed");
    if (async) return;
    const trace::Range range(trace::Label::stream_synchronize);
    check(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize(), "execution failed");
    if (s.flag_armed) {
        s.flag_armed = false;
        if (read_flag(stream, s.flag, s.flag_host)) throw std::runtime_error(what);
    }
}

void HipExecutor::synchronize(hipStream_t stream) {
    std::lock_guard lock(mutex_);
    check(hipSetDevice(device_), "hipSetDevice failed");
    auto it = scratch_.find(stream);
    if (it == scratch_.end()) {
        check(stream ? hipStreamSynchronize(stream) : hipDeviceSynchronize(), "execution failed");
        return;
    }
    finish(stream, it->second, false, "dsmvc: HIP Float64 execution produced NaN or infinity");
}

// Launches one F32 pass on a route: a chunk count (> 0) for the carry-chain
// chunked kernels, or kWindowRoute for the windowed kernels.
hipError_t HipExecutor::launch_route(hip::DevicePlan &dp, int route, bool rows,
                                     const hip::ChunkArgs &a, hip::Sample kind, hipStream_t stream) {
    if (route == kWindowRoute) {
        const hip::DevWindow *win = dp.window(rows);
        if (!win) return hipErrorInvalidValue;
        hip::WinJob job{};
        job.plan = *win;
        job.src = a.src;
        job.s_row = a.s_row;
        job.s_frame = a.s_frame;
        job.dst = a.dst;
        job.d_row = a.d_row;
        job.d_frame = a.d_frame;
        job.nvec = a.nvec;
        job.frames = a.frames;
        job.groups = (a.nvec + 63) / 64;
        job.first_item = 0;
        job.conv = a.conv;
        // A one-job table on this stream's table ring.
        WindowTables &tables = stream_tables(stream);
        tables.clear();
        tables.add(job, hip::window_band_class(dp.axis().w), rows, kind);
        tables.upload(stream);
        tables.launch(stream);
        return hipGetLastError();
    }
    const hip::DevChunks *ch = dp.chunks(route, rows);
    if (!ch) return hipErrorInvalidValue;
    hip::ChunkArgs c = a;
    c.chunks = *ch;
    return hip::launch_chunked(c, rows, kind, stream);
}

// Picks the route of an F32 pass: the executor's tuned choice, a fresh
// measurement when the launch fills the GPU (every feasible route timed with
// HIP events on the call's own buffers, whose output the caller's launch then
// rewrites), or the default: the windowed kernels, else the largest chunked
// partition.  0 = no parallel route (sequential solve).
int HipExecutor::choose_partition(hip::DevicePlan &dp, bool rows, const hip::ChunkArgs &base,
                                  hip::Sample kind, hipStream_t stream) {
    const TuneKey key{&dp, rows, static_cast<int>(kind)};
    if (auto it = tuned_.find(key); it != tuned_.end()) return it->second;
    // The windowed route comes first: it is the default, kept unless a
    // chunked partition is 10% faster.  Its timed launch includes the job
    // table upload, which every call of that route pays.
    std::vector<int> candidates;
    if (mode_ != SolveMode::chunked && dp.window(rows)) candidates.push_back(kWindowRoute);
    if (mode_ != SolveMode::windowed) {
        std::vector<std::pair<int, int>> shapes;
        for (const int c : {hip::kMaxChunks, 12, 10, 8, 6, 4}) {
            const hip::DevChunks *ch = dp.chunks(c, rows);
            if (!ch) continue;
            const std::pair<int, int> shape{ch->chunks, ch->length};
            if (std::find(shapes.begin(), shapes.end(), shape) != shapes.end()) continue;
            shapes.push_back(shape);
            candidates.push_back(c);
        }
    }
    if (candidates.empty()) return 0;
    const long workgroups = static_cast<long>((base.nvec + 63) / 64) * base.frames;
    if (!autotune_ || candidates.size() == 1) {
        tuned_[key] = candidates.front();
        return candidates.front();
    }
    if (workgroups < 2L * compute_units_) return candidates.front();
    hipEvent_t start = nullptr, stop = nullptr;
    check(hipEventCreate(&start), "event creation failed");
    check(hipEventCreate(&stop), "event creation failed");
    std::vector<float> best_ms(candidates.size(), std::numeric_limits<float>::max());
    for (std::size_t i = 0; i < candidates.size(); ++i) {
        check(launch_route(dp, candidates[i], rows, base, kind, stream), "solve launch failed");
        for (int rep = 0; rep < 2; ++rep) {
            check(hipEventRecord(start, stream), "event record failed");
            check(launch_route(dp, candidates[i], rows, base, kind, stream), "solve launch failed");
            check(hipEventRecord(stop, stream), "event record failed");
            check(hipEventSynchronize(stop), "event synchronize failed");
            float ms = 0.0F;
            check(hipEventElapsedTime(&ms, start, stop), "event timing failed");
            best_ms[i] = std::min(best_ms[i], ms);
        }
    }
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    std::size_t best = 0;
    for (std::size_t i = 1; i < candidates.size(); ++i)
        if (best_ms[i] < 0.9F * best_ms[0] && best_ms[i] < best_ms[best]) best = i;
    tuned_[key] = candidates[best];
    return candidates[best];
}

int HipExecutor::tuned_chunks(const AxisPlan &plan, bool rows, int kind) {
    std::lock_guard lock(mutex_);
    auto it = plans_.find(plan_key(plan));
    if (it == plans_.end() || !same_plan(it->second.image->host(), plan)) return -2;
    auto t = tuned_.find(TuneKey{it->second.image.get(), rows, kind});
    return t == tuned_.end() ? 0 : t->second;
}

// One axis pass over `vectors` rows (rows = true) or columns of `frames`
// planes.  Picks the chunked kernels when the mode and plan allow, else the
// sequential solve, whose forward sweep goes to scratch at `fwd_offset`.  In
// refined mode, a plan that retains FP64 data is solved on the F32 route and
// then refined in place (the scratch at `fwd_offset` holds the residuals).
void HipExecutor::axis_pass(const AxisPlan &plan, bool rows, const void *input, Sample in_kind,
                            std::ptrdiff_t in_row, std::ptrdiff_t in_frame, void *output,
                            Sample out_kind, std::ptrdiff_t out_row, std::ptrdiff_t out_frame,
                            std::int32_t vectors, std::int32_t frames, const DevConv &conv,
                            bool f64, Scratch &s, std::size_t fwd_offset, hipStream_t stream) {
    solve_pass(plan, rows, input, in_kind, in_row, in_frame, output, out_kind, out_row, out_frame,
               vectors, frames, conv, f64, s, fwd_offset, stream);
    if (f64 || mode_ != SolveMode::refined || !plan.requires_float64()) return;
    if (in_kind != Sample::f32 || out_kind != Sample::f32)
        throw std::invalid_argument("dsmvc: refined solves need Float32 planes");
    const auto image = device_plan(plan);
    hip::DevicePlan &dp = *image;
    const hip::DevRefine *rt = dp.refine();
    if (!rt) throw std::runtime_error("dsmvc: refinement tables are missing");
    hip::RefineArgs a{};
    a.axis = dp.axis();
    a.refine = *rt;
    a.y = static_cast<const float *>(input);
    a.x = static_cast<float *>(output);
    if (rows) {
        a.y_vec = in_row; a.y_el = 1;
        a.x_vec = out_row; a.x_el = 1;
    } else {
        a.y_vec = 1; a.y_el = in_row;
        a.x_vec = 1; a.x_el = out_row;
    }
    a.y_frame = in_frame;
    a.x_frame = out_frame;
    a.scratch = reinterpret_cast<float *>(static_cast<unsigned char *>(s.buffer) + fwd_offset);
    a.nvec = vectors;
    a.frames = frames;
    const hip::DevWindow *solve = dp.solve_window(rows);
    if (!solve) {
        for (int step = 0; step < refine_steps_; ++step)
            check(hip::launch_refine_step(a, stream), "refinement launch failed");
        return;
    }
    // Residual element by element, the correction G d = r on the windowed
    // kernels (identity taps over r), then x += d.  r and d are laid out
    // vector-major for a rows pass, element-major for a columns pass.
    const std::ptrdiff_t n = plan.destination_size;
    const std::size_t region = align256(static_cast<std::size_t>(n) * vectors * frames * sizeof(float));
    float *r = a.scratch;
    float *d = reinterpret_cast<float *>(reinterpret_cast<unsigned char *>(r) + region);
    hip::WinJob job{};
    job.plan = *solve;
    job.src = r;
    job.s_row = rows ? n : vectors;
    job.s_frame = n * vectors;
    job.dst = d;
    job.d_row = rows ? n : vectors;
    job.d_frame = n * vectors;
    job.nvec = vectors;
    job.frames = frames;
    job.groups = (vectors + 63) / 64;
    job.conv = make_conv(nullptr);
    for (int step = 0; step < refine_steps_; ++step) {
        check(hip::launch_refine_residual(a, r, rows, stream), "refinement launch failed");
        WindowTables &tables = stream_tables(stream);
        tables.clear();
        tables.add(job, hip::window_band_class(dp.axis().w), rows, Sample::f32);
        tables.upload(stream);
        tables.launch(stream);
        check(hipGetLastError(), "refinement launch failed");
        check(hip::launch_refine_add(a, d, rows, stream), "refinement launch failed");
    }
}

void HipExecutor::solve_pass(const AxisPlan &plan, bool rows, const void *input, Sample in_kind,
                            std::ptrdiff_t in_row, std::ptrdiff_t in_frame, void *output,
                            Sample out_kind, std::ptrdiff_t out_row, std::ptrdiff_t out_frame,
                            std::int32_t vectors, std::int32_t frames, const DevConv &conv,
                            bool f64, Scratch &s, std::size_t fwd_offset, hipStream_t stream) {
    {
        // The kernels read other vectors' samples after writing their own
        // outputs, so the two ranges must be disjoint (the 2D entries pass
        // through an intermediate buffer and always are).
        const std::ptrdiff_t m = plan.source_size, n = plan.destination_size;
        const std::ptrdiff_t in_span = (frames - 1) * in_frame
            + (rows ? (vectors - 1) * in_row + m : (m - 1) * in_row + vectors);
        const std::ptrdiff_t out_span = (frames - 1) * out_frame
            + (rows ? (vectors - 1) * out_row + n : (n - 1) * out_row + vectors);
        const auto *a0 = static_cast<const unsigned char *>(input);
        const auto *b0 = static_cast<const unsigned char *>(output);
        const auto *a1 = a0 + in_span * sample_size(in_kind);
        const auto *b1 = b0 + out_span * sample_size(out_kind);
        if (a0 < b1 && b0 < a1)
            throw std::invalid_argument("dsmvc: input and output planes must not overlap");
    }
    const auto image = device_plan(plan);
    hip::DevicePlan &dp = *image;
    // Float64 route of a 2D solve (FP64 intermediate): the windowed FP64
    // kernels when the plan's FP64 recurrence decays within their warm-up.
    const bool f64_pair = rows ? in_kind != Sample::f64 && out_kind == Sample::f64
                               : in_kind == Sample::f64 && out_kind != Sample::f64;
    if (f64 && f64_pair && (automatic_like(mode_) || mode_ == SolveMode::windowed)) {
        if (const hip::DevWindow64 *win = dp.window64()) {
            hip::Win64Args a{};
            a.plan = *win;
            a.src = input;
            a.s_row = in_row;
            a.s_frame = in_frame;
            a.dst = output;
            a.d_row = out_row;
            a.d_frame = out_frame;
            a.nvec = vectors;
            a.frames = frames;
            a.groups = (vectors + 63) / 64;
            a.flag = s.flag;
            a.conv = conv;
            check(hip::launch_window64(a, hip::window_band_class(plan.half_bandwidth), win->ell_width,
                                       rows, rows ? in_kind : out_kind, stream),
                  "windowed Float64 launch failed");
            return;
        }
    }
    if (!f64 && mode_ != SolveMode::exact) {
        hip::ChunkArgs a{};
        a.axis = dp.axis();
        a.src = input;
        a.s_row = in_row;
        a.s_frame = in_frame;
        a.dst = output;
        a.d_row = out_row;
        a.d_frame = out_frame;
        a.nvec = vectors;
        a.frames = frames;
        a.conv = conv;
        const hip::Sample kind = rows ? in_kind : out_kind;
        int route = rows ? tuning_.rows_chunks : tuning_.columns_chunks;
        if ((route > 0 && !dp.chunks(route, rows)) || (route == kWindowRoute && !dp.window(rows))
            || route < kWindowRoute)
            route = 0;
        if (route <= 0) route = choose_partition(dp, rows, a, kind, stream);
        if (route != 0) {
            check(launch_route(dp, route, rows, a, kind, stream), "parallel solve launch failed");
            return;
        }
        if (mode_ == SolveMode::chunked)
            throw std::invalid_argument("dsmvc: plan cannot use the chunked HIP solve");
        if (mode_ == SolveMode::windowed)
            throw std::invalid_argument("dsmvc: plan cannot use the windowed HIP solve");
    }
    if (!hip::seq_band_supported(plan.half_bandwidth))
        throw std::invalid_argument("dsmvc: half bandwidth exceeds the HIP executor's limit");
    hip::SeqArgs a{};
    a.axis = dp.axis();
    a.src = input;
    a.dst = output;
    a.nvec = vectors;
    a.frames = frames;
    a.conv = conv;
    a.flag = f64 ? s.flag : nullptr;
    const std::ptrdiff_t n = plan.destination_size;
    if (rows) {
        a.s_vec = in_row; a.s_el = 1; a.s_frame = in_frame;
        a.d_vec = out_row; a.d_el = 1; a.d_frame = out_frame;
    } else {
        a.s_vec = 1; a.s_el = in_row; a.s_frame = in_frame;
        a.d_vec = 1; a.d_el = out_row; a.d_frame = out_frame;
    }
    // Forward sweep: vector-contiguous scratch, [frame][element][vector].
    a.fwd = static_cast<unsigned char *>(s.buffer) + fwd_offset;
    a.f_vec = 1;
    a.f_el = vectors;
    a.f_frame = static_cast<long>(vectors) * n;
    check(hip::launch_seq(a, in_kind, out_kind, f64, stream), "sequential solve launch failed");
}

void HipExecutor::inverse_rows(const AxisPlan &plan, const float *input,
                               std::ptrdiff_t input_row_stride, float *output,
                               std::ptrdiff_t output_row_stride, std::int32_t row_count,
                               hipStream_t stream, bool async) {
    const trace::Range range(trace::Label::execute_rows);
    if (!plan.valid() || !input || !output || input_row_stride < plan.source_size
        || output_row_stride < plan.destination_size || row_count < 0)
        throw std::invalid_argument("dsmvc: invalid row executor arguments");
    if (row_count == 0) return;
    run_call(stream, async, "dsmvc: HIP Float64 row execution produced NaN or infinity",
             [&](hipStream_t st) -> Scratch & {
        const bool f64 = float64_route(plan);
        const std::size_t elem = f64 ? 8 : 4;
        Scratch &s = scratch(st, refine_regions() * align256(static_cast<std::size_t>(row_count)
                                                             * plan.destination_size * elem));
        const DevConv conv = make_conv(nullptr);
        if (f64) {
            arm_flag(s, st);
            check(hip::launch_check_finite(input, Sample::f32, input_row_stride, 0, row_count,
                                           plan.source_size, 1, s.flag, st), "check launch failed");
        }
        axis_pass(plan, true, input, Sample::f32, input_row_stride, 0, output, Sample::f32,
                  output_row_stride, 0, row_count, 1, conv, f64, s, 0, st);
        return s;
    });
}

void HipExecutor::inverse_columns(const AxisPlan &plan, const float *input,
                                  std::ptrdiff_t input_row_stride, float *output,
                                  std::ptrdiff_t output_row_stride, std::int32_t column_count,
                                  hipStream_t stream, bool async) {
    const trace::Range range(trace::Label::execute_columns);
    if (!plan.valid() || !input || !output || column_count < 0
        || input_row_stride < column_count || output_row_stride < column_count)
        throw std::invalid_argument("dsmvc: invalid column executor arguments");
    if (column_count == 0) return;
    run_call(stream, async, "dsmvc: HIP Float64 column execution produced NaN or infinity",
             [&](hipStream_t st) -> Scratch & {
        const bool f64 = float64_route(plan);
        const std::size_t elem = f64 ? 8 : 4;
        Scratch &s = scratch(st, refine_regions() * align256(static_cast<std::size_t>(column_count)
                                                             * plan.destination_size * elem));
        const DevConv conv = make_conv(nullptr);
        if (f64) {
            arm_flag(s, st);
            check(hip::launch_check_finite(input, Sample::f32, input_row_stride, 0,
                                           plan.source_size, column_count, 1, s.flag, st),
                  "check launch failed");
        }
        axis_pass(plan, false, input, Sample::f32, input_row_stride, 0, output, Sample::f32,
                  output_row_stride, 0, column_count, 1, conv, f64, s, 0, st);
        return s;
    });
}

void HipExecutor::inverse_2d_planes(const PlaneJob *planes, std::int32_t count, SampleKind kind,
                                    hipStream_t stream, bool async) {
    const trace::Range range(trace::Label::execute_planes);
    const bool integer = kind != SampleKind::float32;
    if (count < 0 || (count > 0 && !planes))
        throw std::invalid_argument("dsmvc: invalid plane list");
    for (std::int32_t k = 0; k < count; ++k) {
        const PlaneJob &p = planes[k];
        const PlaneBatch &b = p.batch;
        if (!p.horizontal || !p.vertical || !p.horizontal->valid() || !p.vertical->valid()
            || !b.input || !b.output || b.frames < 1
            || b.input_row_stride < p.horizontal->source_size
            || b.output_row_stride < p.horizontal->destination_size
            || (b.frames > 1
                && (b.input_frame_stride < b.input_row_stride * p.vertical->source_size
                    || b.output_frame_stride < b.output_row_stride * p.vertical->destination_size)))
            throw std::invalid_argument("dsmvc: invalid plane arguments");
        if (integer) {
            if (!p.conversion) throw std::invalid_argument("dsmvc: integer planes need a conversion");
            validate_conversion(*p.conversion, kind == SampleKind::u8 ? 255U : 65535U,
                                "dsmvc: invalid plane arguments");
        }
    }
    run_call(stream, async, "dsmvc: HIP Float64 2D execution produced NaN or infinity",
             [&](hipStream_t st) -> Scratch & { return enqueue_planes(planes, count, kind, st); });
}

// Queues a planes call on `stream` (caller holds the lock); returns the
// scratch entry whose flag a Float64 plane armed.
HipExecutor::Scratch &HipExecutor::enqueue_planes(const PlaneJob *planes, std::int32_t count,
                                                  SampleKind kind, hipStream_t stream) {
    const bool windows = automatic_like(mode_) || mode_ == SolveMode::windowed;
    // The device images stay pinned until every launch of this call is queued.
    std::vector<std::shared_ptr<hip::DevicePlan>> pins;
    const auto n = static_cast<std::size_t>(count);
    std::vector<const hip::DevWindow *> wh(n, nullptr), wv(n, nullptr);
    std::vector<const hip::DevWindow64 *> wh64(n, nullptr), wv64(n, nullptr);
    std::vector<std::size_t> offset(n, 0);
    std::vector<char> fused(n, 0);
    std::size_t total = 0;
    WindowTables &tables = stream_tables(stream);
    tables.clear();
    const bool try_fused = fused_route(kind);
    // Workgroup columns of the whole call, so that a plane's columns are
    // split only when the call as a whole leaves the device short.
    long columns = 0;
    for (std::size_t k = 0; k < n; ++k)
        columns += static_cast<long>(planes[k].batch.frames)
                   * ((planes[k].horizontal->destination_size + hip::kOpTileCols - 1) / hip::kOpTileCols);
    // The F32 fused jobs' range flag, next to the stream's Float64 flag.
    int *fused_flag = kind == SampleKind::float32 ? scratch(stream, 0).flag + 1 : nullptr;
    for (std::size_t k = 0; k < n; ++k) {
        const AxisPlan &h = *planes[k].horizontal, &v = *planes[k].vertical;
        const bool f64 = float64_route(h) || float64_route(v);
        if (try_fused && !f64
            && add_fused(h, v, kind, planes[k].batch, planes[k].conversion, tables, pins, fused_flag, false,
                         columns)) {
            fused[k] = 1;
            continue;
        }
        if (!windows) continue;
        // Held by value: a reference into `pins` would dangle once it grows.
        const std::shared_ptr<hip::DevicePlan> ih = device_plan(h), iv = device_plan(v);
        pins.push_back(ih);
        pins.push_back(iv);
        std::size_t elem = 4;
        if (!f64) {
            wh[k] = ih->window(true);
            wv[k] = iv->window(false);
            if (!wh[k] || !wv[k]) {
                wh[k] = wv[k] = nullptr;
                continue;
            }
        } else {
            // The Float64 route on the windowed FP64 kernels, in the same
            // table set, when both plans' FP64 recurrences decay in time.
            wh64[k] = ih->window64();
            wv64[k] = iv->window64();
            if (!wh64[k] || !wv64[k]) {
                wh64[k] = wv64[k] = nullptr;
                continue;
            }
            elem = 8;
        }
        offset[k] = total;
        total += align256(static_cast<std::size_t>(padded_width(h.destination_size)) * v.source_size
                          * static_cast<std::size_t>(planes[k].batch.frames) * elem);
    }
    Scratch *last = nullptr;
    const bool any_fused = std::find(fused.begin(), fused.end(), 1) != fused.end();
    if (total || any_fused) {
        Scratch &s = scratch(stream, total);
        last = &s;
        if (any_fused && fused_flag) tables.clear_flags(fused_flag, 1);
        const Sample sk = sample_of(kind);
        bool armed = false;
        for (std::size_t k = 0; k < n; ++k) {
            if (!wh[k] && !wh64[k]) continue;
            const PlaneJob &p = planes[k];
            const AxisPlan &h = *p.horizontal, &v = *p.vertical;
            const std::ptrdiff_t row = padded_width(h.destination_size);
            const std::ptrdiff_t frame = row * v.source_size;
            void *inter = static_cast<unsigned char *>(s.buffer) + offset[k];
            const DevConv conv = make_conv(p.conversion);
            if (wh[k]) {
                hip::WinJob a{};
                a.plan = *wh[k];
                a.src = p.batch.input;
                a.s_row = p.batch.input_row_stride;
                a.s_frame = p.batch.input_frame_stride;
                a.dst = inter;
                a.d_row = row;
                a.d_frame = frame;
                a.nvec = v.source_size;
                a.frames = p.batch.frames;
                a.groups = (a.nvec + 63) / 64;
                a.conv = conv;
                tables.add(a, hip::window_band_class(h.half_bandwidth), true, sk);
                hip::WinJob b{};
                b.plan = *wv[k];
                b.src = inter;
                b.s_row = row;
                b.s_frame = frame;
                b.dst = p.batch.output;
                b.d_row = p.batch.output_row_stride;
                b.d_frame = p.batch.output_frame_stride;
                b.nvec = h.destination_size;
                b.frames = p.batch.frames;
                b.groups = (b.nvec + 63) / 64;
                b.conv = conv;
                tables.add(b, hip::window_band_class(v.half_bandwidth), false, sk);
                continue;
            }
            if (!armed) {
                arm_flag(s, stream);
                armed = true;
            }
            if (sk == Sample::f32) {
                hip::CheckJob c{};
                c.src = static_cast<const float *>(p.batch.input);
                c.s_row = p.batch.input_row_stride;
                c.s_frame = p.batch.input_frame_stride;
                c.rows = v.source_size;
                c.cols = h.source_size;
                c.frames = p.batch.frames;
                c.flag = s.flag;
                tables.add_check(c);
            }
            hip::Win64Job a{};
            a.plan = *wh64[k];
            a.src = p.batch.input;
            a.s_row = p.batch.input_row_stride;
            a.s_frame = p.batch.input_frame_stride;
            a.dst = inter;
            a.d_row = row;
            a.d_frame = frame;
            a.nvec = v.source_size;
            a.frames = p.batch.frames;
            a.groups = (a.nvec + 63) / 64;
            a.flag = s.flag;
            a.conv = conv;
            tables.add64(a, hip::window_band_class(h.half_bandwidth), true, sk);
            hip::Win64Job b{};
            b.plan = *wv64[k];
            b.src = inter;
            b.s_row = row;
            b.s_frame = frame;
            b.dst = p.batch.output;
            b.d_row = p.batch.output_row_stride;
            b.d_frame = p.batch.output_frame_stride;
            b.nvec = h.destination_size;
            b.frames = p.batch.frames;
            b.groups = (b.nvec + 63) / 64;
            b.flag = s.flag;
            b.conv = conv;
            tables.add64(b, hip::window_band_class(v.half_bandwidth), false, sk);
        }
        tables.upload(stream);
        tables.launch(stream);
    }
    for (std::size_t k = 0; k < n; ++k) {
        if (wh[k] || wh64[k] || fused[k]) continue;
        const PlaneJob &p = planes[k];
        last = &solve_2d(*p.horizontal, *p.vertical, kind, p.batch, p.conversion, stream);
    }
    if (!last) last = &scratch(stream, 0);
    return *last;
}

int *HipExecutor::float64_flag(hipStream_t stream) {
    std::lock_guard lock(mutex_);
    check(hipSetDevice(device_), "hipSetDevice failed");
    Scratch &s = scratch(stream, 0);
    arm_flag(s, stream);
    return s.flag;
}

void HipExecutor::inverse_2d_batch(const AxisPlan &h, const AxisPlan &v, SampleKind kind,
                                   const PlaneBatch &b, const IntegerConversion *conversion,
                                   hipStream_t stream, bool async) {
    const trace::Range range(trace::Label::execute_2d);
    const bool integer = kind != SampleKind::float32;
    if (!h.valid() || !v.valid() || !b.input || !b.output || b.frames < 1
        || b.input_row_stride < h.source_size || b.output_row_stride < h.destination_size
        || v.source_size < 1 || v.destination_size < 1
        || (b.frames > 1 && (b.input_frame_stride < b.input_row_stride * v.source_size
                             || b.output_frame_stride < b.output_row_stride * v.destination_size)))
        throw std::invalid_argument(integer ? "dsmvc: invalid integer 2D executor arguments"
                                            : "dsmvc: invalid 2D executor arguments");
    if (integer) {
        if (!conversion) throw std::invalid_argument("dsmvc: integer 2D execution needs a conversion");
        validate_conversion(*conversion, kind == SampleKind::u8 ? 255U : 65535U,
                            "dsmvc: invalid integer 2D executor arguments");
    }
    run_call(stream, async, "dsmvc: HIP Float64 2D execution produced NaN or infinity",
             [&](hipStream_t st) -> Scratch & { return solve_2d(h, v, kind, b, conversion, st); });
}

bool HipExecutor::fused_job(const AxisPlan &h, const AxisPlan &v, SampleKind kind, const PlaneBatch &b,
                            const IntegerConversion *conversion, WindowTables &tables,
                            std::vector<std::shared_ptr<hip::DevicePlan>> &pins, int *flag) {
    if (!fused_route(kind) || float64_route(h) || float64_route(v)) return false;
    std::lock_guard lock(mutex_);
    check(hipSetDevice(device_), "hipSetDevice failed");
    return add_fused(h, v, kind, b, conversion, tables, pins, flag, true);
}

bool HipExecutor::add_fused(const AxisPlan &h, const AxisPlan &v, SampleKind kind, const PlaneBatch &b,
                            const IntegerConversion *conversion, WindowTables &tables,
                            std::vector<std::shared_ptr<hip::DevicePlan>> &pins, int *flag,
                            bool whole_columns, long launch_columns) {
    const std::shared_ptr<hip::DevicePlan> ih = device_plan(h), iv = device_plan(v);
    const hip::OperatorPlan *ph = ih->operator_plan(), *pv = iv->operator_plan();
    if (!ph || !pv) return false;
    // The kernel addresses one frame's plane with 32-bit byte offsets; its
    // stores run over whole units of rows (the rows past the plane are
    // discarded by the buffer range check), so their offsets are bounded too.
    const std::size_t bytes = kind == SampleKind::float32 ? 4 : (kind == SampleKind::u16 ? 2 : 1);
    const auto span = [&](std::ptrdiff_t row, std::int32_t rows, std::int32_t cols) {
        return (static_cast<std::size_t>(rows - 1) * static_cast<std::size_t>(row) + cols) * bytes;
    };
    const std::int32_t unit_rows = (v.destination_size + hip::kOpUnit - 1) / hip::kOpUnit * hip::kOpUnit;
    if (span(b.input_row_stride, v.source_size, h.source_size) >= 0x7FFFFFF0U
        || span(b.output_row_stride, unit_rows, h.destination_size) >= 0x7FFFFFF0U)
        return false;
    const int steps = std::max({ph->steps, pv->steps, 2});
    const bool f32 = kind == SampleKind::float32;
    const hip::DevOperator *oh = ih->op(true, false, steps), *ov = iv->op(false, false, steps);
    if (!oh || !ov) return false;
    const hip::DevOperator *eh = f32 ? ih->op(true, true, steps) : nullptr;
    const hip::DevOperator *ev = f32 ? iv->op(false, true, steps) : nullptr;
    if (f32 && (!eh || !ev || !flag)) return false;
    pins.push_back(ih);
    pins.push_back(iv);
    hip::OpJob j{};
    j.h = *oh;
    j.v = *ov;
    j.src = b.input;
    j.dst = b.output;
    j.s_row = b.input_row_stride;
    j.s_frame = b.input_frame_stride;
    j.d_row = b.output_row_stride;
    j.d_frame = b.output_frame_stride;
    j.frames = b.frames;
    j.tiles = (h.destination_size + hip::kOpTileCols - 1) / hip::kOpTileCols;
    j.units = (v.destination_size + hip::kOpUnit - 1) / hip::kOpUnit;
    // Whole columns per workgroup unless that leaves the device too few
    // workgroups (counting every column of the launch); each split repeats
    // about one unit's source rows.
    const long per_column = std::max(static_cast<long>(j.frames) * j.tiles, launch_columns);
    int upi = j.units;
    while (!whole_columns && upi > 4 && per_column * ((j.units + upi - 1) / upi) < 4L * compute_units_)
        upi = (upi + 1) / 2;
    j.units_per_item = upi;
    j.segments = (j.units + upi - 1) / upi;
    // A barrier after a step's units when some unit spans more than
    // kOpRing - 2 kOpSlab + 1 ring rows (its two blocks' K ranges at
    // `steps`): only then can the step's slab overwrite rows it reads.
    int unit_span = 0;
    for (int u = 0; u < j.units; ++u)
        unit_span = std::max(unit_span, pv->kstart[2 * u + 1] - pv->kstart[2 * u] + hip::kOpStep * steps);
    j.unit_barrier = unit_span > hip::kOpRing - 2 * hip::kOpSlab + 1 ? 1 : 0;
    const std::size_t align = kind == SampleKind::u8 ? 8 : 16;
    const bool aligned = reinterpret_cast<std::uintptr_t>(b.input) % align == 0
                         && (static_cast<std::size_t>(b.input_row_stride) * bytes) % align == 0
                         && (b.frames == 1 || (static_cast<std::size_t>(b.input_frame_stride) * bytes) % align == 0);
    j.vector_loads = aligned ? 1 : 0;
    j.conv = make_conv(conversion);
    if (conversion) {
        const double scale = static_cast<double>(conversion->input_scale) * conversion->output_scale;
        j.out_mul = static_cast<float>(256.0 * scale);
        j.out_add = static_cast<float>(conversion->output_offset - conversion->input_offset * scale);
    }
    if (f32) {
        // Horizontal results stay below FP16's largest value when every
        // sample is at most this large in magnitude.
        j.range = static_cast<float>(65504.0 / (1.01 * std::max(1.0, ph->max_row_l1)));
        j.flag = flag;
        hip::OpJob e = j;
        e.h = *eh;
        e.v = *ev;
        tables.add_op(j, steps, sample_of(kind));
        tables.add_op(e, steps, sample_of(kind), true);
        return true;
    }
    tables.add_op(j, steps, sample_of(kind));
    return true;
}

// Queues one plane batch's two passes on `stream` (caller holds the lock).
HipExecutor::Scratch &HipExecutor::solve_2d(const AxisPlan &h, const AxisPlan &v, SampleKind kind,
                                            const PlaneBatch &b, const IntegerConversion *conversion,
                                            hipStream_t stream) {
    const bool f64 = float64_route(h) || float64_route(v);
    if (!f64 && fused_route(kind)) {
        std::vector<std::shared_ptr<hip::DevicePlan>> pins;
        WindowTables &tables = stream_tables(stream);
        tables.clear();
        Scratch &s0 = scratch(stream, 0);
        int *flag = kind == SampleKind::float32 ? s0.flag + 1 : nullptr;
        if (add_fused(h, v, kind, b, conversion, tables, pins, flag)) {
            if (flag) tables.clear_flags(flag, 1);
            tables.upload(stream);
            tables.launch(stream);
            return s0;
        }
    }
    const std::size_t elem = f64 ? 8 : 4;
    const std::ptrdiff_t inter_row = padded_width(h.destination_size);
    const std::ptrdiff_t inter_frame = inter_row * v.source_size;
    const std::size_t inter_bytes = align256(static_cast<std::size_t>(inter_frame) * b.frames * elem);
    // Sequential forward sweeps, the larger of the two passes.
    const std::size_t fwd_bytes = refine_regions() * align256(
        std::max(static_cast<std::size_t>(v.source_size) * h.destination_size,
                 static_cast<std::size_t>(h.destination_size) * v.destination_size)
        * b.frames * elem);
    Scratch &s = scratch(stream, inter_bytes + fwd_bytes);
    const DevConv conv = make_conv(conversion);
    const Sample in_kind = sample_of(kind);
    if (f64) {
        arm_flag(s, stream);
        check(hip::launch_check_finite(b.input, in_kind, b.input_row_stride, b.input_frame_stride,
                                       v.source_size, h.source_size, b.frames, s.flag, stream),
              "check launch failed");
    }
    void *inter = s.buffer;
    const Sample inter_kind = f64 ? Sample::f64 : Sample::f32;
    axis_pass(h, true, b.input, in_kind, b.input_row_stride, b.input_frame_stride, inter,
              inter_kind, inter_row, inter_frame, v.source_size, b.frames, conv, f64, s,
              inter_bytes, stream);
    axis_pass(v, false, inter, inter_kind, inter_row, inter_frame, b.output, in_kind,
              b.output_row_stride, b.output_frame_stride, h.destination_size, b.frames, conv,
              f64, s, inter_bytes, stream);
    return s;
}

// Clears the stream's Float64 flag unless an earlier, unreported call armed
// it (its failure must survive until finish or synchronize reads it).
void HipExecutor::arm_flag(Scratch &s, hipStream_t stream) {
    if (!s.flag_armed) check(hipMemsetAsync(s.flag, 0, sizeof(int), stream), "flag reset failed");
    s.flag_armed = true;
}

void HipExecutor::inverse_rows_batch(const AxisPlan &plan, SampleKind in_kind,
                                     const PlaneBatch &b, std::int32_t rows,
                                     const IntegerConversion *conversion, hipStream_t stream,
                                     bool async) {
    const trace::Range range(trace::Label::execute_rows);
    if (!plan.valid() || !b.input || !b.output || b.frames < 1 || rows < 1
        || b.input_row_stride < plan.source_size || b.output_row_stride < plan.destination_size
        || (b.frames > 1 && (b.input_frame_stride < b.input_row_stride * rows
                             || b.output_frame_stride < b.output_row_stride * rows)))
        throw std::invalid_argument("dsmvc: invalid row executor arguments");
    if (float64_route(plan))
        throw std::invalid_argument("dsmvc: batched single-axis passes take Float32 plans; "
                                    "use the 2D entry for Float64 plans");
    if (in_kind != SampleKind::float32) {
        if (!conversion) throw std::invalid_argument("dsmvc: integer rows need a conversion");
        validate_conversion(*conversion, in_kind == SampleKind::u8 ? 255U : 65535U,
                            "dsmvc: invalid row executor arguments");
    }
    run_call(stream, async, "dsmvc: HIP row execution failed", [&](hipStream_t st) -> Scratch & {
        Scratch &s = scratch(st, refine_regions() * align256(static_cast<std::size_t>(rows)
                                                             * plan.destination_size * b.frames * 4));
        axis_pass(plan, true, b.input, sample_of(in_kind), b.input_row_stride,
                  b.input_frame_stride, b.output, Sample::f32, b.output_row_stride,
                  b.output_frame_stride, rows, b.frames, make_conv(conversion), false, s, 0, st);
        return s;
    });
}

void HipExecutor::inverse_columns_batch(const AxisPlan &plan, SampleKind out_kind,
                                        const PlaneBatch &b, std::int32_t columns,
                                        const IntegerConversion *conversion, hipStream_t stream,
                                        bool async) {
    const trace::Range range(trace::Label::execute_columns);
    if (!plan.valid() || !b.input || !b.output || b.frames < 1 || columns < 1
        || b.input_row_stride < columns || b.output_row_stride < columns
        || (b.frames > 1 && (b.input_frame_stride < b.input_row_stride * plan.source_size
                             || b.output_frame_stride < b.output_row_stride * plan.destination_size)))
        throw std::invalid_argument("dsmvc: invalid column executor arguments");
    if (float64_route(plan))
        throw std::invalid_argument("dsmvc: batched single-axis passes take Float32 plans; "
                                    "use the 2D entry for Float64 plans");
    if (out_kind != SampleKind::float32) {
        if (!conversion) throw std::invalid_argument("dsmvc: integer columns need a conversion");
        validate_conversion(*conversion, out_kind == SampleKind::u8 ? 255U : 65535U,
                            "dsmvc: invalid column executor arguments");
    }
    run_call(stream, async, "dsmvc: HIP column execution failed", [&](hipStream_t st) -> Scratch & {
        Scratch &s = scratch(st, refine_regions() * align256(static_cast<std::size_t>(columns)
                                                             * plan.destination_size * b.frames * 4));
        axis_pass(plan, false, b.input, Sample::f32, b.input_row_stride, b.input_frame_stride,
                  b.output, sample_of(out_kind), b.output_row_stride, b.output_frame_stride,
                  columns, b.frames, make_conv(conversion), false, s, 0, st);
        return s;
    });
}

// ---------------------------------------------------------------------------
// Host frames
// ---------------------------------------------------------------------------
namespace {
// Copies `rows` rows of `width` samples between strided and packed layouts.
void pack_rows(void *dst, std::ptrdiff_t dst_stride, const void *src, std::ptrdiff_t src_stride,
               std::int64_t rows, std::size_t width, std::size_t sample) {
    auto *d = static_cast<unsigned char *>(dst);
    const auto *q = static_cast<const unsigned char *>(src);
    for (std::int64_t r = 0; r < rows; ++r)
        std::memcpy(d + r * dst_stride * static_cast<std::ptrdiff_t>(sample),
                    q + r * src_stride * static_cast<std::ptrdiff_t>(sample), width * sample);
}
} // namespace

void HipExecutor::inverse_2d_host(const AxisPlan &h, const AxisPlan &v, SampleKind kind,
                                  const PlaneBatch &b, const IntegerConversion *conversion) {
    const trace::Range range(trace::Label::execute_host_frame);
    const bool integer = kind != SampleKind::float32;
    if (!h.valid() || !v.valid() || !b.input || !b.output || b.frames < 1
        || b.input_row_stride < h.source_size || b.output_row_stride < h.destination_size
        || (b.frames > 1 && (b.input_frame_stride < b.input_row_stride * v.source_size
                             || b.output_frame_stride < b.output_row_stride * v.destination_size)))
        throw std::invalid_argument(integer ? "dsmvc: invalid integer 2D executor arguments"
                                            : "dsmvc: invalid 2D executor arguments");
    if (integer) {
        if (!conversion) throw std::invalid_argument("dsmvc: integer 2D execution needs a conversion");
        validate_conversion(*conversion, kind == SampleKind::u8 ? 255U : 65535U,
                            "dsmvc: invalid integer 2D executor arguments");
    }
    const std::size_t sample = static_cast<std::size_t>(sample_size(sample_of(kind)));
    const std::size_t in_plane = static_cast<std::size_t>(h.source_size) * v.source_size;
    const std::size_t out_plane = static_cast<std::size_t>(h.destination_size) * v.destination_size;
    const std::size_t in_bytes = align256(in_plane * b.frames * sample);
    const std::size_t out_bytes = align256(out_plane * b.frames * sample);
    Slot *slot = acquire_slot();
    SlotGuard guard{[this, slot] { release_slot(slot); }};
    slot_buffers(*slot, std::max(in_bytes, out_bytes), in_bytes + out_bytes);
    {
        const trace::Range pack(trace::Label::host_pack);
        for (std::int32_t f = 0; f < b.frames; ++f)
            pack_rows(static_cast<unsigned char *>(slot->pinned) + f * in_plane * sample,
                      h.source_size,
                      static_cast<const unsigned char *>(b.input)
                          + f * b.input_frame_stride * static_cast<std::ptrdiff_t>(sample),
                      b.input_row_stride, v.source_size, static_cast<std::size_t>(h.source_size),
                      sample);
    }
    void *dev_in = slot->device;
    void *dev_out = static_cast<unsigned char *>(slot->device) + in_bytes;
    const hipStream_t st = slot->stream;
    Scratch *used = nullptr;
    {
        std::lock_guard lock(mutex_);
        check(hipSetDevice(device_), "hipSetDevice failed");
        check(hipMemcpyAsync(dev_in, slot->pinned, in_plane * b.frames * sample,
                             hipMemcpyHostToDevice, st), "host frame upload failed");
        const PlaneBatch dev{dev_in, h.source_size, static_cast<std::ptrdiff_t>(in_plane), dev_out,
                             h.destination_size, static_cast<std::ptrdiff_t>(out_plane), b.frames};
        used = &solve_2d(h, v, kind, dev, conversion, st);
        check(hipMemcpyAsync(slot->pinned, dev_out, out_plane * b.frames * sample,
                             hipMemcpyDeviceToHost, st), "host frame download failed");
        check(hipGetLastError(), "kernel launch failed");
    }
    complete(st, *used, "dsmvc: HIP Float64 2D execution produced NaN or infinity");
    const trace::Range unpack(trace::Label::host_unpack);
    for (std::int32_t f = 0; f < b.frames; ++f)
        pack_rows(static_cast<unsigned char *>(b.output)
                      + f * b.output_frame_stride * static_cast<std::ptrdiff_t>(sample),
                  b.output_row_stride,
                  static_cast<const unsigned char *>(slot->pinned) + f * out_plane * sample,
                  h.destination_size, v.destination_size,
                  static_cast<std::size_t>(h.destination_size), sample);
}

void HipExecutor::inverse_rows_host(const AxisPlan &plan, const float *input,
                                    std::ptrdiff_t input_row_stride, float *output,
                                    std::ptrdiff_t output_row_stride, std::int32_t row_count) {
    const trace::Range range(trace::Label::execute_host_frame);
    if (!plan.valid() || !input || !output || input_row_stride < plan.source_size
        || output_row_stride < plan.destination_size || row_count < 0)
        throw std::invalid_argument("dsmvc: invalid row executor arguments");
    if (row_count == 0) return;
    const std::size_t m = plan.source_size, n = plan.destination_size;
    const std::size_t in_bytes = align256(m * row_count * 4), out_bytes = align256(n * row_count * 4);
    Slot *slot = acquire_slot();
    SlotGuard guard{[this, slot] { release_slot(slot); }};
    slot_buffers(*slot, std::max(in_bytes, out_bytes), in_bytes + out_bytes);
    pack_rows(slot->pinned, static_cast<std::ptrdiff_t>(m), input, input_row_stride, row_count, m, 4);
    auto *dev_in = static_cast<float *>(slot->device);
    auto *dev_out = reinterpret_cast<float *>(static_cast<unsigned char *>(slot->device) + in_bytes);
    const hipStream_t st = slot->stream;
    Scratch *used = nullptr;
    {
        std::lock_guard lock(mutex_);
        check(hipSetDevice(device_), "hipSetDevice failed");
        check(hipMemcpyAsync(dev_in, slot->pinned, m * row_count * 4, hipMemcpyHostToDevice, st),
              "host frame upload failed");
        const bool f64 = float64_route(plan);
        Scratch &s = scratch(st, refine_regions() * align256(static_cast<std::size_t>(row_count) * n
                                                             * (f64 ? 8 : 4)));
        if (f64) {
            arm_flag(s, st);
            check(hip::launch_check_finite(dev_in, Sample::f32, static_cast<long>(m), 0, row_count,
                                           plan.source_size, 1, s.flag, st), "check launch failed");
        }
        axis_pass(plan, true, dev_in, Sample::f32, static_cast<std::ptrdiff_t>(m), 0, dev_out,
                  Sample::f32, static_cast<std::ptrdiff_t>(n), 0, row_count, 1, make_conv(nullptr),
                  f64, s, 0, st);
        check(hipMemcpyAsync(slot->pinned, dev_out, n * row_count * 4, hipMemcpyDeviceToHost, st),
              "host frame download failed");
        check(hipGetLastError(), "kernel launch failed");
        used = &s;
    }
    complete(st, *used, "dsmvc: HIP Float64 row execution produced NaN or infinity");
    pack_rows(output, output_row_stride, slot->pinned, static_cast<std::ptrdiff_t>(n), row_count, n, 4);
}

void HipExecutor::inverse_columns_host(const AxisPlan &plan, const float *input,
                                       std::ptrdiff_t input_row_stride, float *output,
                                       std::ptrdiff_t output_row_stride, std::int32_t column_count) {
    const trace::Range range(trace::Label::execute_host_frame);
    if (!plan.valid() || !input || !output || column_count < 0
        || input_row_stride < column_count || output_row_stride < column_count)
        throw std::invalid_argument("dsmvc: invalid column executor arguments");
    if (column_count == 0) return;
    const std::size_t m = plan.source_size, n = plan.destination_size, c = column_count;
    const std::size_t in_bytes = align256(m * c * 4), out_bytes = align256(n * c * 4);
    Slot *slot = acquire_slot();
    SlotGuard guard{[this, slot] { release_slot(slot); }};
    slot_buffers(*slot, std::max(in_bytes, out_bytes), in_bytes + out_bytes);
    pack_rows(slot->pinned, static_cast<std::ptrdiff_t>(c), input, input_row_stride,
              static_cast<std::int64_t>(m), c, 4);
    auto *dev_in = static_cast<float *>(slot->device);
    auto *dev_out = reinterpret_cast<float *>(static_cast<unsigned char *>(slot->device) + in_bytes);
    const hipStream_t st = slot->stream;
    Scratch *used = nullptr;
    {
        std::lock_guard lock(mutex_);
        check(hipSetDevice(device_), "hipSetDevice failed");
        check(hipMemcpyAsync(dev_in, slot->pinned, m * c * 4, hipMemcpyHostToDevice, st),
              "host frame upload failed");
        const bool f64 = float64_route(plan);
        Scratch &s = scratch(st, refine_regions() * align256(c * n * (f64 ? 8 : 4)));
        if (f64) {
            arm_flag(s, st);
            check(hip::launch_check_finite(dev_in, Sample::f32, static_cast<long>(c), 0,
                                           plan.source_size, column_count, 1, s.flag, st),
                  "check launch failed");
        }
        axis_pass(plan, false, dev_in, Sample::f32, static_cast<std::ptrdiff_t>(c), 0, dev_out,
                  Sample::f32, static_cast<std::ptrdiff_t>(c), 0, column_count, 1,
                  make_conv(nullptr), f64, s, 0, st);
        check(hipMemcpyAsync(slot->pinned, dev_out, n * c * 4, hipMemcpyDeviceToHost, st),
              "host frame download failed");
        check(hipGetLastError(), "kernel launch failed");
        used = &s;
    }
    complete(st, *used, "dsmvc: HIP Float64 column execution produced NaN or infinity");
    pack_rows(output, output_row_stride, slot->pinned, static_cast<std::ptrdiff_t>(c),
              static_cast<std::int64_t>(n), c, 4);
}

void HipExecutor::inverse_2d(const AxisPlan &h, const AxisPlan &v, const float *input,
                             std::ptrdiff_t is, float *output, std::ptrdiff_t os,
                             hipStream_t stream, bool async) {
    inverse_2d_batch(h, v, SampleKind::float32, {input, is, 0, output, os, 0, 1}, nullptr, stream, async);
}

void HipExecutor::inverse_2d_u8(const AxisPlan &h, const AxisPlan &v, const std::uint8_t *input,
                                std::ptrdiff_t is, std::uint8_t *output, std::ptrdiff_t os,
                                const IntegerConversion &c, hipStream_t stream, bool async) {
    inverse_2d_batch(h, v, SampleKind::u8, {input, is, 0, output, os, 0, 1}, &c, stream, async);
}

void HipExecutor::inverse_2d_u16(const AxisPlan &h, const AxisPlan &v, const std::uint16_t *input,
                                 std::ptrdiff_t is, std::uint16_t *output, std::ptrdiff_t os,
                                 const IntegerConversion &c, hipStream_t stream, bool async) {
    inverse_2d_batch(h, v, SampleKind::u16, {input, is, 0, output, os, 0, 1}, &c, stream, async);
}

// ---------------------------------------------------------------------------
// Backend names (reference src/backend.cpp:107-154, plus "hip").
// ---------------------------------------------------------------------------
BackendKind parse_backend(std::string_view name) {
    if (name.empty() || name == "auto" || name == "automatic") return BackendKind::automatic;
    if (name == "cpu") return BackendKind::cpu;
    if (name == "metal") return BackendKind::metal;
    if (name == "vulkan") return BackendKind::vulkan;
    if (name == "cuda") return BackendKind::cuda;
    if (name == "hip") return BackendKind::hip;
    throw std::invalid_argument("backend must be auto, cpu, metal, vulkan, cuda, or hip");
}

BackendKind resolve_backend(BackendKind requested) {
    switch (requested) {
    case BackendKind::automatic:
    case BackendKind::cpu:
        return BackendKind::cpu;
    case BackendKind::hip:
        if (!hip_available()) throw std::runtime_error("backend hip is unavailable");
        return BackendKind::hip;
    case BackendKind::metal:
    case BackendKind::vulkan:
    case BackendKind::cuda:
        throw std::runtime_error(std::string("backend ") + backend_name(requested)
                                 + " is unavailable");
    }
    throw std::invalid_argument("backend is invalid");
}

const char *backend_name(BackendKind kind) noexcept {
    switch (kind) {
    case BackendKind::automatic: return "auto";
    case BackendKind::cpu: return "cpu";
    case BackendKind::metal: return "metal";
    case BackendKind::vulkan: return "vulkan";
    case BackendKind::cuda: return "cuda";
    case BackendKind::hip: return "hip";
    }
    return "unknown";
}

} // namespace dsmvc

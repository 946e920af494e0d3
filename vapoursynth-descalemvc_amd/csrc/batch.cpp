// Multi-plan batches over windowed job tables (see batch.hpp).
#include "batch.hpp"
#include "trace.hpp"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <cstdlib>
#include <exception>
#include <stdexcept>
#include <string>
#include <system_error>
#include <thread>

#include "device_plan.hpp"
#include "executor.hpp"
#include "kernels.hpp"

namespace dsmvc {
namespace {

void check(hipError_t e, const char *what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string("dsmvc: ") + what + ": " + hipGetErrorString(e));
}

std::size_t align256(std::size_t v) { return (v + 255U) & ~std::size_t{255}; }

std::ptrdiff_t pitch_of(std::int32_t n) { return (static_cast<std::ptrdiff_t>(n) + 63) & ~std::ptrdiff_t{63}; }

hip::DevConv identity_conv() {
    hip::DevConv c{};
    c.in_scale = 1.0F;
    c.out_scale = 1.0F;
    return c;
}

} // namespace

// ---------------------------------------------------------------------------
// WindowTables
// ---------------------------------------------------------------------------
// Launches use XCD-contiguous workgroups with the vector group fastest
// (launch_window_table order 2); DSMVC_HIP_WINDOW_ORDER (0-3) overrides it.
WindowTables::WindowTables(int device) : device_(device) {
    if (const char *env = std::getenv("DSMVC_HIP_WINDOW_ORDER")) order_ = std::atoi(env) & 3;
}

WindowTables::~WindowTables() {
    (void)hipSetDevice(device_);
    for (Slot &slot : slots_) {
        if (slot.done) {
            if (slot.pending) (void)hipEventSynchronize(slot.done);
            (void)hipEventDestroy(slot.done);
        }
        if (slot.device) (void)hipFree(slot.device);
        if (slot.host) (void)hipHostFree(slot.host);
    }
}

void WindowTables::clear() {
    groups_.clear();
    flags_ = nullptr;
    flag_count_ = 0;
}

void WindowTables::clear_flags(int *flags, int count) {
    flags_ = flags;
    flag_count_ = flags ? count : 0;
}

WindowTables::Group &WindowTables::group_for(Kind type, int band, int width, bool rows,
                                             bool block16, hip::Sample kind) {
    auto it = std::find_if(groups_.begin(), groups_.end(), [&](const Group &g) {
        return g.type == type && g.rows == rows && g.band == band && g.width == width
            && g.block16 == block16 && g.kind == kind;
    });
    if (it != groups_.end()) return *it;
    Group g;
    g.type = type;
    g.band = band;
    g.width = width;
    g.rows = rows;
    g.block16 = block16;
    g.kind = kind;
    groups_.push_back(std::move(g));
    return groups_.back();
}

void WindowTables::add(const hip::WinJob &job, int band, bool rows, hip::Sample kind) {
    Group &g = group_for(Kind::f32, band, job.plan.ell_width, rows, job.plan.block16 != 0, kind);
    hip::WinJob j = job;
    j.first_item = g.items;
    g.items += hip::window_items(j);   // until upload() sets the launch's spans
    g.jobs.push_back(j);
}

void WindowTables::add64(const hip::Win64Job &job, int band, bool rows, hip::Sample kind) {
    Group &g = group_for(Kind::f64, band, job.plan.ell_width, rows, false, kind);
    hip::Win64Job j = job;
    j.first_item = g.items;
    g.items += hip::window64_items(j);   // until upload() sets the launch's spans
    g.jobs64.push_back(j);
}

void WindowTables::add_check(const hip::CheckJob &job) {
    Group &g = group_for(Kind::check, 0, 0, true, false, hip::Sample::f32);
    hip::CheckJob j = job;
    j.first_item = g.items;
    g.items += static_cast<long>(j.frames) * j.rows;
    g.checks.push_back(j);
}

void WindowTables::add_op(const hip::OpJob &job, int steps, hip::Sample kind, bool exact) {
    // `block16` marks the launches whose jobs all allow wide source loads.
    Group &g = group_for(Kind::op, exact ? 1 : 0, steps, true, job.vector_loads != 0, kind);
    hip::OpJob j = job;
    j.first_item = g.items;
    g.items += hip::operator_items(j);
    g.ops.push_back(j);
}

// Tables go through a ring of slots, each a pinned host buffer, a device
// buffer and an event recorded after the launches that read it: an upload
// reuses the oldest slot once its event has passed, so uploads and launches
// stay asynchronous (a copy from pageable memory would wait for the stream).
int WindowTables::upload(hipStream_t stream) {
    const trace::Range range(trace::Label::table_upload);
    // Checks first, then horizontal groups: every vertical pass reads an
    // intermediate some horizontal group writes.
    std::stable_sort(groups_.begin(), groups_.end(), [](const Group &a, const Group &b) {
        auto rank = [](const Group &g) {
            if (g.type == Kind::check) return 0;
            if (g.type == Kind::op) return g.band;   // first passes, then exact passes
            return g.rows ? 2 : 3;
        };
        return rank(a) < rank(b);
    });
    check(hipSetDevice(device_), "hipSetDevice failed");
    // Each launch's span (chunks per wavefront), from its kernel's
    // occupancy; the jobs' first items follow.
    for (Group &g : groups_) {
        if (g.type == Kind::f32 && !g.jobs.empty())
            g.items = hip::assign_window_spans(
                g.jobs.data(), static_cast<int>(g.jobs.size()),
                hip::window_table_slots(g.band, g.width, g.rows, g.block16, g.kind));
        if (g.type == Kind::f64 && !g.jobs64.empty())
            g.items = hip::assign_window64_spans(
                g.jobs64.data(), static_cast<int>(g.jobs64.size()),
                hip::window64_table_slots(g.band, g.width, g.rows, g.kind));
    }
    std::size_t bytes = 0;
    for (Group &g : groups_) {
        g.offset = bytes;
        bytes += (g.bytes() + 15) & ~std::size_t{15};
    }
    current_ = (current_ + 1) % kSlots;
    Slot &slot = slots_[current_];
    if (slot.pending) {
        check(hipEventSynchronize(slot.done), "table slot wait failed");
        slot.pending = false;
    }
    if (!slot.done) check(hipEventCreateWithFlags(&slot.done, hipEventDisableTiming), "event creation failed");
    if (bytes > slot.capacity) {
        const std::size_t want = std::max<std::size_t>(bytes, slot.capacity * 2);
        if (slot.device) check(hipFree(slot.device), "job table release failed");
        if (slot.host) check(hipHostFree(slot.host), "job table release failed");
        slot.device = nullptr;
        slot.host = nullptr;
        slot.capacity = 0;
        check(hipHostMalloc(reinterpret_cast<void **>(&slot.host), want), "job table allocation failed");
        check(hipMalloc(reinterpret_cast<void **>(&slot.device), want), "job table allocation failed");
        slot.capacity = want;
    }
    for (const Group &g : groups_) {
        unsigned char *at = slot.host + g.offset;
        if (!g.jobs.empty()) std::memcpy(at, g.jobs.data(), g.jobs.size() * sizeof(hip::WinJob));
        if (!g.jobs64.empty()) std::memcpy(at, g.jobs64.data(), g.jobs64.size() * sizeof(hip::Win64Job));
        if (!g.checks.empty()) std::memcpy(at, g.checks.data(), g.checks.size() * sizeof(hip::CheckJob));
        if (!g.ops.empty()) std::memcpy(at, g.ops.data(), g.ops.size() * sizeof(hip::OpJob));
    }
    if (bytes) {
        if (stream)
            check(hipMemcpyAsync(slot.device, slot.host, bytes, hipMemcpyHostToDevice, stream),
                  "job table upload failed");
        else
            check(hipMemcpy(slot.device, slot.host, bytes, hipMemcpyHostToDevice),
                  "job table upload failed");
    }
    record_ = stream != nullptr;
    return static_cast<int>(groups_.size());
}

void WindowTables::launch(hipStream_t stream) {
    Slot &slot = slots_[current_];
    if (flag_count_ > 0) check(hip::launch_fill_i32(flags_, flag_count_, 0, stream), "flag reset failed");
    for (const Group &g : groups_) {
        if (g.items == 0) continue;
        // Checks and horizontal groups come first (upload() sorts them).
        const trace::Range range(g.type == Kind::op
                                     ? trace::Label::fused_stage
                                     : (g.rows || g.type == Kind::check ? trace::Label::horizontal_stage
                                                                        : trace::Label::vertical_stage));
        unsigned char *at = slot.device + g.offset;
        const int count = static_cast<int>(g.count());
        hipError_t e = hipSuccess;
        switch (g.type) {
        case Kind::check:
            e = hip::launch_check_finite_table(reinterpret_cast<const hip::CheckJob *>(at), count,
                                               g.items, stream);
            break;
        case Kind::f32:
            e = hip::launch_window_table(reinterpret_cast<const hip::WinJob *>(at), count, g.items,
                                         g.band, g.width, g.rows, g.block16, g.kind, order_, stream);
            break;
        case Kind::f64:
            e = hip::launch_window64_table(reinterpret_cast<const hip::Win64Job *>(at), count,
                                           g.items, g.band, g.width, g.rows, g.kind, stream);
            break;
        case Kind::op: {
            // The exact pass reads one flag first when all its jobs share it.
            const int *shared = g.ops.empty() ? nullptr : g.ops.front().flag;
            for (const hip::OpJob &j : g.ops)
                if (j.flag != shared) shared = nullptr;
            e = hip::launch_operator_table(reinterpret_cast<const hip::OpJob *>(at), count, g.items,
                                           g.width, g.kind, g.band != 0, g.block16, shared, stream);
            break;
        }
        }
        check(e, "table launch failed");
    }
    if (record_) {
        check(hipEventRecord(slot.done, stream), "event record failed");
        slot.pending = true;
    }
}

std::int64_t WindowTables::items() const {
    std::int64_t total = 0;
    for (const Group &g : groups_) total += g.items;
    return total;
}

// ---------------------------------------------------------------------------
// HipBatch
// ---------------------------------------------------------------------------
HipBatch::HipBatch(HipExecutor &executor, std::vector<BatchJob> jobs, bool graph)
    : executor_(executor), jobs_(std::move(jobs)), tables_(executor.device()) {
    const trace::Range range(trace::Label::batch_setup);
    for (const BatchJob &j : jobs_) {
        const AxisPlan *h = j.horizontal.get(), *v = j.vertical.get();
        if (!h || !v || !h->valid() || !v->valid() || !j.input || !j.output || j.frames < 1
            || j.input_row_stride < h->source_size || j.output_row_stride < h->destination_size
            || (j.frames > 1 && (j.input_frame_stride < j.input_row_stride * v->source_size
                                 || j.output_frame_stride < j.output_row_stride * v->destination_size)))
            throw std::invalid_argument("dsmvc: invalid batch job");
    }
    check(hipSetDevice(executor_.device()), "hipSetDevice failed");
    const bool parallel = automatic_like(executor_.mode())
                          || executor_.mode() == SolveMode::windowed;
    // Intermediate layout of the table jobs: [frame][source row][pitch(n_h)].
    std::vector<std::size_t> offset(jobs_.size(), 0);
    std::vector<const hip::DevWindow *> wh(jobs_.size(), nullptr), wv(jobs_.size(), nullptr);
    // Device images of the table jobs, held (pinned in the executor's plan
    // cache) for the batch's lifetime: the tables point into them.
    std::vector<std::shared_ptr<hip::DevicePlan>> ih(jobs_.size()), iv(jobs_.size());
    // Jobs whose operators fit the fused 2D route (one kernel, no
    // intermediate); their images are built with the others.
    const bool fused = executor_.fused_enabled(SampleKind::float32);
    std::vector<char> fit(jobs_.size(), 0);
    if (parallel) {
        // Device images and window tables of every job, built on host
        // threads (a scan brings hundreds of new plans).  Each task writes
        // only its own job's slots; the executor and each device plan lock
        // their own state.
        std::vector<std::exception_ptr> failed(jobs_.size());
        std::atomic<std::size_t> next{0};
        auto work = [&] {
            for (std::size_t k = next++; k < jobs_.size(); k = next++) {
                const AxisPlan &h = *jobs_[k].horizontal, &v = *jobs_[k].vertical;
                if (h.requires_float64() || v.requires_float64()) continue;
                try {
                    ih[k] = executor_.plan_image(h);
                    iv[k] = executor_.plan_image(v);
                    if (fused) {
                        const hip::OperatorPlan *oh = ih[k]->operator_plan(), *ov = iv[k]->operator_plan();
                        if (oh && ov) {
                            const int steps = std::max({oh->steps, ov->steps, 2});
                            fit[k] = ih[k]->op(true, false, steps) && iv[k]->op(false, false, steps)
                                     && ih[k]->op(true, true, steps) && iv[k]->op(false, true, steps);
                        }
                        if (fit[k]) continue;
                    }
                    wh[k] = ih[k]->window(true);
                    wv[k] = iv[k]->window(false);
                } catch (...) {
                    failed[k] = std::current_exception();
                }
            }
        };
        const std::size_t workers = std::min<std::size_t>(
            jobs_.size(), std::min(16U, std::max(1U, std::thread::hardware_concurrency())));
        std::vector<std::thread> pool;
        // Fewer threads than asked for (thread creation failed) only means
        // less parallelism: the calling thread takes whatever is left.
        for (std::size_t t = 1; t < workers; ++t) {
            try {
                pool.emplace_back(work);
            } catch (const std::system_error &) {
                break;
            }
        }
        work();
        for (auto &t : pool) t.join();
        for (const auto &e : failed)
            if (e) std::rethrow_exception(e);
    }
    if (std::find(fit.begin(), fit.end(), 1) != fit.end())
        check(hipMalloc(reinterpret_cast<void **>(&flags_), jobs_.size() * sizeof(int)),
              "batch flag allocation failed");
    std::size_t total = 0;
    for (std::size_t k = 0; k < jobs_.size(); ++k) {
        const BatchJob &j = jobs_[k];
        const AxisPlan &h = *j.horizontal, &v = *j.vertical;
        if (fit[k]) {
            // Whole columns per workgroup: the table holds many jobs.
            if (executor_.fused_job(h, v, SampleKind::float32,
                                    PlaneBatch{j.input, j.input_row_stride, j.input_frame_stride, j.output,
                                               j.output_row_stride, j.output_frame_stride, j.frames},
                                    nullptr, tables_, images_, flags_ + k)) {
                ++stats_.fused_jobs;
                ++stats_.table_jobs;
                continue;
            }
            fit[k] = 0;
            // Not taken after all: the windowed tables, if they fit.
            wh[k] = ih[k]->window(true);
            wv[k] = iv[k]->window(false);
        }
        if (!wh[k] || !wv[k]) {
            if (executor_.mode() == SolveMode::windowed)
                throw std::invalid_argument("dsmvc: batch job cannot use the windowed HIP solve");
            fallback_.push_back(k);
            continue;
        }
        images_.push_back(std::move(ih[k]));
        images_.push_back(std::move(iv[k]));
        offset[k] = total;
        total += align256(static_cast<std::size_t>(pitch_of(h.destination_size)) * v.source_size
                          * static_cast<std::size_t>(j.frames) * sizeof(float));
    }
    if (total) check(hipMalloc(&intermediate_, total), "batch intermediate allocation failed");
    if (flags_) tables_.clear_flags(flags_, static_cast<int>(jobs_.size()));
    for (std::size_t k = 0; k < jobs_.size(); ++k) {
        // Jobs with either pass outside the windowed kernels went to the
        // fallback list above and are solved there only; fused jobs are in
        // the tables already.
        if (fit[k] || !wh[k] || !wv[k]) continue;
        const BatchJob &j = jobs_[k];
        const AxisPlan &h = *j.horizontal, &v = *j.vertical;
        const std::ptrdiff_t row = pitch_of(h.destination_size);
        const std::ptrdiff_t frame = row * v.source_size;
        float *inter = reinterpret_cast<float *>(static_cast<unsigned char *>(intermediate_) + offset[k]);
        hip::WinJob a{};
        a.plan = *wh[k];
        a.src = j.input;
        a.s_row = j.input_row_stride;
        a.s_frame = j.input_frame_stride;
        a.dst = inter;
        a.d_row = row;
        a.d_frame = frame;
        a.nvec = v.source_size;
        a.frames = j.frames;
        a.groups = (a.nvec + 63) / 64;
        a.conv = identity_conv();
        tables_.add(a, hip::window_band_class(h.half_bandwidth), true, hip::Sample::f32);
        hip::WinJob b{};
        b.plan = *wv[k];
        b.src = inter;
        b.s_row = row;
        b.s_frame = frame;
        b.dst = j.output;
        b.d_row = j.output_row_stride;
        b.d_frame = j.output_frame_stride;
        b.nvec = h.destination_size;
        b.frames = j.frames;
        b.groups = (b.nvec + 63) / 64;
        b.conv = identity_conv();
        tables_.add(b, hip::window_band_class(v.half_bandwidth), false, hip::Sample::f32);
        ++stats_.table_jobs;
    }
    stats_.jobs = static_cast<std::int32_t>(jobs_.size());
    stats_.fallback_jobs = static_cast<std::int32_t>(fallback_.size());
    stats_.launches = tables_.upload(nullptr);
    stats_.items = tables_.items();
    stats_.intermediate_bytes = static_cast<std::int64_t>(total);
    if (graph && stats_.launches > 0) {
        hipStream_t capture = nullptr;
        check(hipStreamCreateWithFlags(&capture, hipStreamNonBlocking), "stream creation failed");
        hipError_t e = hipStreamBeginCapture(capture, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            try {
                tables_.launch(capture);
            } catch (...) {
                (void)hipStreamEndCapture(capture, &graph_);
                (void)hipStreamDestroy(capture);
                throw;
            }
            e = hipStreamEndCapture(capture, &graph_);
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0);
        (void)hipStreamDestroy(capture);
        check(e, "batch graph capture failed");
        stats_.graph = true;
    }
}

HipBatch::~HipBatch() {
    (void)hipSetDevice(executor_.device());
    (void)hipDeviceSynchronize();
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (intermediate_) (void)hipFree(intermediate_);
    if (flags_) (void)hipFree(flags_);
}

void HipBatch::run(hipStream_t stream, bool async) {
    const trace::Range range(trace::Label::batch_run);
    check(hipSetDevice(executor_.device()), "hipSetDevice failed");
    if (graph_exec_) check(hipGraphLaunch(graph_exec_, stream), "batch graph launch failed");
    else tables_.launch(stream);
    for (const std::size_t k : fallback_) {
        const BatchJob &j = jobs_[k];
        executor_.inverse_2d_batch(*j.horizontal, *j.vertical, SampleKind::float32,
                                   PlaneBatch{j.input, j.input_row_stride, j.input_frame_stride,
                                              j.output, j.output_row_stride, j.output_frame_stride,
                                              j.frames},
                                   nullptr, stream, true);
    }
    if (!async) executor_.synchronize(stream);
}

} // namespace dsmvc

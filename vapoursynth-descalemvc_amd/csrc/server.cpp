// Continuous-batching request loop (see server.hpp).
#include "server.hpp"
#include "trace.hpp"

#include <algorithm>
#include <array>
#include <future>
#include <set>
#include <stdexcept>
#include <system_error>

#include "batch.hpp"
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

void require(bool ok, const char *what) {
    if (!ok) throw std::invalid_argument(what);
}

} // namespace

struct HipServer::Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::unique_ptr<WindowTables> tables;
    std::vector<std::shared_ptr<Entry>> round;
    bool fallback = false;   // the round solved planes through the executor
    bool float64 = false;    // the round ran Float64 table jobs (per-request flags)
    bool busy = false;
    // One Float64 failure flag per request of a round: set by its table
    // jobs, copied back with the round, read when it completes.  Then one
    // range flag per request for its fused F32 planes (the FP16 pass sets
    // it, the exact pass reads it).
    int *flags = nullptr;         // device, 2 x max_round_requests
    int *flags_host = nullptr;    // pinned
};

HipServer::HipServer(HipExecutor &executor, const ServerConfig &config)
    : executor_(executor), config_(config), epoch_(std::chrono::steady_clock::now()) {
    require(config_.memory_budget >= 256 && config_.slots >= 1 && config_.slots <= 16
                && config_.max_round_requests >= 1,
            "dsmvc: invalid server configuration");
    config_.memory_budget &= ~std::size_t{255};
    check(hipSetDevice(executor_.device()), "hipSetDevice failed");
    check(hipMalloc(&arena_, config_.memory_budget), "server memory allocation failed");
    free_[0] = config_.memory_budget;
    for (int k = 0; k < config_.slots; ++k) {
        auto slot = std::make_unique<Slot>();
        check(hipStreamCreateWithFlags(&slot->stream, hipStreamNonBlocking), "stream creation failed");
        check(hipEventCreateWithFlags(&slot->done, hipEventDisableTiming), "event creation failed");
        slot->tables = std::make_unique<WindowTables>(executor_.device());
        const std::size_t flag_bytes = static_cast<std::size_t>(config_.max_round_requests) * sizeof(int);
        check(hipMalloc(reinterpret_cast<void **>(&slot->flags), 2 * flag_bytes), "flag allocation failed");
        check(hipHostMalloc(reinterpret_cast<void **>(&slot->flags_host), flag_bytes),
              "flag allocation failed");
        slots_.push_back(std::move(slot));
    }
    worker_ = std::thread([this] { loop(); });
    if (config_.fused == ServeFused::warm && executor_.fused_enabled(SampleKind::float32)) {
        try {
            builder_ = std::thread([this] { builder(); });
        } catch (const std::system_error &) {
            // No builder: planes keep the windowed route.
        }
    }
}

HipServer::~HipServer() {
    {
        std::lock_guard lock(build_mutex_);
        build_stop_ = true;
    }
    build_cv_.notify_all();
    if (builder_.joinable()) builder_.join();
    {
        std::lock_guard lock(mutex_);
        stop_ = true;
    }
    work_cv_.notify_all();
    if (worker_.joinable()) worker_.join();
    (void)hipSetDevice(executor_.device());
    for (auto &slot : slots_) {
        if (slot->stream) {
            (void)hipStreamSynchronize(slot->stream);
            try {
                executor_.release_stream(slot->stream);   // Float64 planes ran on it
            } catch (...) {
            }
        }
        slot->tables.reset();
        if (slot->flags) (void)hipFree(slot->flags);
        if (slot->flags_host) (void)hipHostFree(slot->flags_host);
        if (slot->done) (void)hipEventDestroy(slot->done);
        if (slot->stream) (void)hipStreamDestroy(slot->stream);
    }
    if (arena_) (void)hipFree(arena_);
}

void HipServer::request_operators(const std::shared_ptr<hip::DevicePlan> &h,
                                  const std::shared_ptr<hip::DevicePlan> &v) {
    std::lock_guard lock(build_mutex_);
    if (!builder_.joinable() || build_stop_ || build_queue_.size() >= 1024) return;
    if (!build_pending_.insert({h.get(), v.get()}).second) return;
    build_queue_.emplace_back(h, v);
    build_cv_.notify_one();
}

// Builds the operator images (FP16 and bfloat16 parts, at the pair's common
// step count) of queued plan pairs; a build that fails leaves the pair on
// the windowed route.
void HipServer::builder() {
    for (;;) {
        std::pair<std::shared_ptr<hip::DevicePlan>, std::shared_ptr<hip::DevicePlan>> job;
        {
            std::unique_lock lock(build_mutex_);
            build_cv_.wait(lock, [&] { return build_stop_ || !build_queue_.empty(); });
            if (build_stop_) return;
            job = std::move(build_queue_.front());
            build_queue_.pop_front();
        }
        try {
            check(hipSetDevice(executor_.device()), "hipSetDevice failed");
            const hip::OperatorPlan *oh = job.first->operator_plan(), *ov = job.second->operator_plan();
            if (oh && ov) {
                const int steps = std::max({oh->steps, ov->steps, 2});
                if (job.first->op(true, false, steps) && job.second->op(false, false, steps)) {
                    (void)job.first->op(true, true, steps);
                    (void)job.second->op(false, true, steps);
                }
            }
        } catch (...) {
        }
        std::lock_guard lock(build_mutex_);
        build_pending_.erase({job.first.get(), job.second.get()});
    }
}

double HipServer::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - epoch_).count();
}

std::uint64_t HipServer::submit(const ServeRequest &r) {
    const double t0 = now();
    const trace::Range range(trace::Label::server_submit);
    require(r.planes == 1 || r.planes == 3, "dsmvc: a request has 1 or 3 planes");
    require(r.src_width > 0 && r.src_height > 0 && r.dst_width > 0 && r.dst_height > 0,
            "dsmvc: invalid request geometry");
    const int ssw = r.planes == 3 ? r.subsampling_w : 0, ssh = r.planes == 3 ? r.subsampling_h : 0;
    require(ssw >= 0 && ssw <= 2 && ssh >= 0 && ssh <= 2, "dsmvc: invalid chroma subsampling");
    require((r.dst_width % (1 << ssw)) == 0 && (r.dst_height % (1 << ssh)) == 0,
            "dsmvc: output width is incompatible with subsampling");

    // Requests per plane kind (reference prepare_filter_requests,
    // src/vs_plugin.cpp:656-705, with src_left = src_top = 0 and the
    // active region equal to the output size).
    AxisRequest base;
    base.kernel = r.kernel;
    base.border = r.border;
    base.f64_mode = r.f64_mode;
    auto axis = [&](int source, int destination, double active, double shift) {
        AxisRequest q = base;
        q.source_size = source;
        q.destination_size = destination;
        q.active_length = active;
        q.shift = shift;
        return q;
    };
    auto plane_size = [&](int p, int &sw, int &sh, int &dw, int &dh) {
        const bool chroma = p > 0;
        sw = chroma ? r.src_width >> ssw : r.src_width;
        sh = chroma ? r.src_height >> ssh : r.src_height;
        dw = chroma ? r.dst_width >> ssw : r.dst_width;
        dh = chroma ? r.dst_height >> ssh : r.dst_height;
    };
    for (int p = 0; p < r.planes; ++p) {
        int sw, sh, dw, dh;
        plane_size(p, sw, sh, dw, dh);
        require(sw > 0 && sh > 0 && r.src[p] && r.dst[p] && r.src_stride[p] >= sw
                    && r.dst_stride[p] >= dw,
                "dsmvc: invalid request plane");
    }
    // Axis requests: luma horizontal, luma vertical, then chroma.
    std::array<AxisRequest, 4> wanted;
    int kinds = 2;
    wanted[0] = axis(r.src_width, r.dst_width, r.dst_width, 0.0);
    wanted[1] = axis(r.src_height, r.dst_height, r.dst_height, 0.0);
    if (r.planes == 3) {
        int sw, sh, dw, dh;
        plane_size(1, sw, sh, dw, dh);
        const double cw = sw, fw = r.src_width;
        wanted[2] = ssw ? axis(sw, dw, double(r.dst_width) * cw / fw, 0.25 - 0.25 * double(r.dst_width) / fw)
                        : axis(sw, dw, dw, 0.0);
        wanted[3] = ssh ? axis(sh, dh, double(r.dst_height) * double(sh) / double(r.src_height), 0.0)
                        : axis(sh, dh, dh, 0.0);
        kinds = 4;
    }
    const bool parallel = automatic_like(executor_.mode())
                          || executor_.mode() == SolveMode::windowed;
    // Device image plus the windowed tables the round (or, for a Float64
    // pair, the executor) will use.
    auto prepare = [&](const AxisPlan &plan, bool rows,
                       bool f64_pair) -> std::shared_ptr<hip::DevicePlan> {
        std::shared_ptr<hip::DevicePlan> image = executor_.plan_image(plan);
        if (parallel) {
            if (f64_pair) (void)image->window64();
            else (void)image->window(rows);
        }
        return image;
    };
    // Cached plans are used as they are.  A request with several new plans
    // (a new geometry) builds them, and then their device tables, on
    // concurrent tasks: for a 4K source that is several milliseconds of FP64
    // host work per plan, which the request's latency includes.
    std::array<std::shared_ptr<const AxisPlan>, 4> plans;
    int missing = 0;
    for (int k = 0; k < kinds; ++k) {
        plans[static_cast<std::size_t>(k)] = find_axis_plan(wanted[static_cast<std::size_t>(k)]);
        missing += plans[static_cast<std::size_t>(k)] ? 0 : 1;
    }
    // A thread that cannot be started (std::system_error from std::async)
    // leaves the remaining work to the inline path below.
    if (missing > 1) try {
        std::array<std::future<std::shared_ptr<const AxisPlan>>, 4> built;
        for (int k = 0; k < kinds; ++k)
            if (!plans[static_cast<std::size_t>(k)])
                built[static_cast<std::size_t>(k)] = std::async(
                    std::launch::async, [q = wanted[static_cast<std::size_t>(k)]] { return get_or_build_axis_plan(q); });
        for (int k = 0; k < kinds; ++k)
            if (!plans[static_cast<std::size_t>(k)]) plans[static_cast<std::size_t>(k)] = built[static_cast<std::size_t>(k)].get();
        std::array<std::future<void>, 4> images;
        for (int k = 0; k < kinds; ++k) {
            const AxisPlan &plan = *plans[static_cast<std::size_t>(k)];
            const AxisPlan &pair = *plans[static_cast<std::size_t>(k ^ 1)];
            const bool f64_pair = plan.requires_float64() || pair.requires_float64();
            images[static_cast<std::size_t>(k)] = std::async(
                std::launch::async, [&prepare, &plan, k, f64_pair] { (void)prepare(plan, (k & 1) == 0, f64_pair); });
        }
        for (int k = 0; k < kinds; ++k) images[static_cast<std::size_t>(k)].get();
    } catch (const std::system_error &) {
    }
    for (int k = 0; k < kinds; ++k)
        if (!plans[static_cast<std::size_t>(k)])
            plans[static_cast<std::size_t>(k)] = get_or_build_axis_plan(wanted[static_cast<std::size_t>(k)]);

    auto entry = std::make_shared<Entry>();
    entry->submitted = t0;
    for (int p = 0; p < r.planes; ++p) {
        int sw, sh, dw, dh;
        plane_size(p, sw, sh, dw, dh);
        const std::size_t kind = p > 0 ? 2 : 0;
        Plane plane;
        plane.h = plans[kind];
        plane.v = plans[kind + 1];
        plane.src = r.src[p];
        plane.src_stride = r.src_stride[p];
        plane.dst = r.dst[p];
        plane.dst_stride = r.dst_stride[p];
        // Device images are uploaded here, on the caller's thread, so the
        // loop only builds tables and launches.
        const bool f64_pair = plane.h->requires_float64() || plane.v->requires_float64();
        plane.ih = prepare(*plane.h, true, f64_pair);
        plane.iv = prepare(*plane.v, false, f64_pair);
        if (parallel && !f64_pair) {
            plane.wh = plane.ih->window(true);
            plane.wv = plane.iv->window(false);
        } else if (parallel) {
            plane.wh64 = plane.ih->window64();
            plane.wv64 = plane.iv->window64();
        }
        plane.table = plane.wh && plane.wv;
        plane.table64 = plane.wh64 && plane.wv64;
        // The fused 2D route once the pair's operator images exist (built
        // off this path by the builder thread, or here); no intermediate then.
        if (!f64_pair && config_.fused != ServeFused::off && executor_.fused_enabled(SampleKind::float32)) {
            if (config_.fused == ServeFused::always) {
                const hip::OperatorPlan *oh = plane.ih->operator_plan(), *ov = plane.iv->operator_plan();
                if (oh && ov) {
                    const int steps = std::max({oh->steps, ov->steps, 2});
                    if (plane.ih->op(true, false, steps) && plane.iv->op(false, false, steps)) {
                        (void)plane.ih->op(true, true, steps);
                        (void)plane.iv->op(false, true, steps);
                    }
                }
            }
            const int a = plane.ih->built_op_steps(), b = plane.iv->built_op_steps();
            if (a > 0 && b > 0) {
                const int steps = std::max({a, b, 2});
                plane.fused = plane.ih->built_op(true, false, steps) && plane.iv->built_op(false, false, steps)
                              && plane.ih->built_op(true, true, steps) && plane.iv->built_op(false, true, steps);
            }
            if (!plane.fused && a != 0 && b != 0) request_operators(plane.ih, plane.iv);
        }
        if (!plane.fused && (plane.table || plane.table64))
            plane.inter_bytes = align256(static_cast<std::size_t>(pitch_of(dw)) * sh
                                         * (plane.table64 ? sizeof(double) : sizeof(float)));
        entry->bytes += plane.inter_bytes;
        entry->planes.push_back(std::move(plane));
    }
    require(entry->bytes <= config_.memory_budget,
            "dsmvc: request needs more device memory than the server's budget");
    std::uint64_t id;
    {
        std::lock_guard lock(mutex_);
        require(!stop_, "dsmvc: server is stopping");
        id = next_id_++;
        entry->id = id;
        entries_[id] = entry;
        queue_.push_back(entry);
        ++stats_.submitted;
    }
    work_cv_.notify_one();
    return id;
}

RequestState HipServer::wait(std::uint64_t id, std::string *error, double *submitted,
                             double *completed) {
    std::unique_lock lock(mutex_);
    auto it = entries_.find(id);
    if (it == entries_.end()) throw std::invalid_argument("dsmvc: unknown request id");
    std::shared_ptr<Entry> e = it->second;
    done_cv_.wait(lock, [&] { return e->state != RequestState::pending; });
    if (error) *error = e->error;
    if (submitted) *submitted = e->submitted;
    if (completed) *completed = e->completed;
    return e->state;
}

RequestState HipServer::poll(std::uint64_t id) {
    std::lock_guard lock(mutex_);
    auto it = entries_.find(id);
    if (it == entries_.end()) throw std::invalid_argument("dsmvc: unknown request id");
    return it->second->state;
}

void HipServer::release(std::uint64_t id) {
    std::lock_guard lock(mutex_);
    auto it = entries_.find(id);
    if (it != entries_.end() && it->second->state != RequestState::pending) entries_.erase(it);
}

ServerStats HipServer::stats() const {
    std::lock_guard lock(mutex_);
    return stats_;
}

bool HipServer::arena_alloc(std::size_t bytes, std::size_t &offset) {
    if (bytes == 0) {
        offset = 0;
        return true;
    }
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < bytes) continue;
        offset = it->first;
        const std::size_t rest = it->second - bytes;
        free_.erase(it);
        if (rest) free_[offset + bytes] = rest;
        in_use_ += bytes;
        stats_.memory_high_water = std::max(stats_.memory_high_water, in_use_);
        return true;
    }
    return false;
}

void HipServer::arena_free(std::size_t offset, std::size_t bytes) {
    if (bytes == 0) return;
    in_use_ -= bytes;
    auto it = free_.emplace(offset, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
        it->second += next->second;
        free_.erase(next);
    }
    if (it != free_.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            free_.erase(it);
        }
    }
}

// Takes every queued request that fits (in arrival order, stopping at the
// first that does not) and launches them as one round on `slot`.
bool HipServer::admit_round(Slot &slot, std::unique_lock<std::mutex> &lock) {
    std::vector<std::shared_ptr<Entry>> round;
    while (!queue_.empty() && static_cast<int>(round.size()) < config_.max_round_requests) {
        const std::shared_ptr<Entry> &e = queue_.front();
        if (!arena_alloc(e->bytes, e->offset)) break;
        round.push_back(e);
        queue_.pop_front();
    }
    if (round.empty()) return false;
    const trace::Range range(trace::Label::server_round);
    slot.busy = true;
    slot.round = round;
    ++stats_.rounds;
    stats_.largest_round = std::max<std::uint64_t>(stats_.largest_round, round.size());
    lock.unlock();
    bool fallback = false, float64_round = false;
    std::string failure;
    std::uint64_t table_planes = 0, fallback_planes = 0, fused_planes = 0;
    try {
        check(hipSetDevice(executor_.device()), "hipSetDevice failed");
        WindowTables &tables = *slot.tables;
        tables.clear();
        bool float64 = false;
        int *fused_flags = slot.flags + config_.max_round_requests;
        bool any_fused = false;
        // Fused planes the executor did not take after all: solved with the
        // other executor-route planes after the tables.
        std::set<const Plane *> untaken;
        std::vector<std::shared_ptr<hip::DevicePlan>> pins;   // the planes hold their images
        for (std::size_t r = 0; r < round.size(); ++r) {
            const auto &e = round[r];
            std::size_t at = e->offset;
            for (const Plane &p : e->planes) {
                if (p.fused) {
                    if (executor_.fused_job(*p.h, *p.v, SampleKind::float32,
                                            PlaneBatch{p.src, p.src_stride, 0, p.dst, p.dst_stride, 0, 1},
                                            nullptr, tables, pins, fused_flags + r)) {
                        any_fused = true;
                        ++table_planes;
                        ++fused_planes;
                    } else {
                        untaken.insert(&p);
                    }
                    continue;
                }
                if (!p.table && !p.table64) continue;
                const AxisPlan &h = *p.h, &v = *p.v;
                const std::ptrdiff_t row = pitch_of(h.destination_size);
                void *inter = static_cast<unsigned char *>(arena_) + at;
                at += p.inter_bytes;
                if (p.table64) {
                    // The Float64 route in the round's tables: the input
                    // check and both passes report into the request's flag.
                    int *flag = slot.flags + r;
                    hip::CheckJob c{};
                    c.src = p.src;
                    c.s_row = p.src_stride;
                    c.rows = v.source_size;
                    c.cols = h.source_size;
                    c.frames = 1;
                    c.flag = flag;
                    tables.add_check(c);
                    hip::Win64Job a{};
                    a.plan = *p.wh64;
                    a.src = p.src;
                    a.s_row = p.src_stride;
                    a.dst = inter;
                    a.d_row = row;
                    a.nvec = v.source_size;
                    a.frames = 1;
                    a.groups = (a.nvec + 63) / 64;
                    a.flag = flag;
                    a.conv = identity_conv();
                    tables.add64(a, hip::window_band_class(h.half_bandwidth), true, hip::Sample::f32);
                    hip::Win64Job b{};
                    b.plan = *p.wv64;
                    b.src = inter;
                    b.s_row = row;
                    b.dst = p.dst;
                    b.d_row = p.dst_stride;
                    b.nvec = h.destination_size;
                    b.frames = 1;
                    b.groups = (b.nvec + 63) / 64;
                    b.flag = flag;
                    b.conv = identity_conv();
                    tables.add64(b, hip::window_band_class(v.half_bandwidth), false, hip::Sample::f32);
                    float64 = true;
                    ++table_planes;
                    continue;
                }
                hip::WinJob a{};
                a.plan = *p.wh;
                a.src = p.src;
                a.s_row = p.src_stride;
                a.dst = inter;
                a.d_row = row;
                a.nvec = v.source_size;
                a.frames = 1;
                a.groups = (a.nvec + 63) / 64;
                a.conv = identity_conv();
                tables.add(a, hip::window_band_class(h.half_bandwidth), true, hip::Sample::f32);
                hip::WinJob b{};
                b.plan = *p.wv;
                b.src = inter;
                b.s_row = row;
                b.dst = p.dst;
                b.d_row = p.dst_stride;
                b.nvec = h.destination_size;
                b.frames = 1;
                b.groups = (b.nvec + 63) / 64;
                b.conv = identity_conv();
                tables.add(b, hip::window_band_class(v.half_bandwidth), false, hip::Sample::f32);
                ++table_planes;
            }
        }
        if (float64)
            check(hipMemsetAsync(slot.flags, 0, round.size() * sizeof(int), slot.stream),
                  "flag reset failed");
        if (any_fused) tables.clear_flags(fused_flags, static_cast<int>(round.size()));
        tables.upload(slot.stream);
        tables.launch(slot.stream);
        if (float64)
            check(hipMemcpyAsync(slot.flags_host, slot.flags, round.size() * sizeof(int),
                                 hipMemcpyDeviceToHost, slot.stream),
                  "flag readback failed");
        for (const auto &e : round)
            for (const Plane &p : e->planes) {
                if (p.fused ? untaken.count(&p) == 0 : (p.table || p.table64)) continue;
                executor_.inverse_2d_batch(*p.h, *p.v, SampleKind::float32,
                                           PlaneBatch{p.src, p.src_stride, 0, p.dst, p.dst_stride, 0, 1},
                                           nullptr, slot.stream, true);
                fallback = true;
                ++fallback_planes;
            }
        float64_round = float64;
        check(hipEventRecord(slot.done, slot.stream), "event record failed");
    } catch (const std::exception &ex) {
        failure = ex.what();
    }
    lock.lock();
    slot.fallback = fallback;
    slot.float64 = float64_round;
    stats_.table_planes += table_planes;
    stats_.fused_planes += fused_planes;
    stats_.fallback_planes += fallback_planes;
    if (!failure.empty()) {
        // Nothing of this round can be trusted: wait for whatever was queued,
        // then fail its requests.
        lock.unlock();
        (void)hipStreamSynchronize(slot.stream);
        lock.lock();
        const double t = now();
        for (const auto &e : round) {
            arena_free(e->offset, e->bytes);
            e->state = RequestState::failed;
            e->error = failure;
            e->completed = t;
            ++stats_.failed;
        }
        slot.busy = false;
        slot.round.clear();
        done_cv_.notify_all();
    }
    return true;
}

void HipServer::finish_round(Slot &slot, std::unique_lock<std::mutex> &lock, std::string device_error) {
    std::string failure;
    if (slot.fallback && device_error.empty()) {
        lock.unlock();
        try {
            executor_.synchronize(slot.stream);   // reports a Float64 non-finite result
        } catch (const std::exception &ex) {
            failure = ex.what();
        }
        lock.lock();
    }
    const double t = now();
    for (std::size_t r = 0; r < slot.round.size(); ++r) {
        const auto &e = slot.round[r];
        arena_free(e->offset, e->bytes);
        bool uses_fallback = false, uses_table64 = false;
        for (const Plane &p : e->planes) {
            uses_fallback |= !p.fused && !p.table && !p.table64;
            uses_table64 |= p.table64;
        }
        // The request's own Float64 flag (copied back with the round).
        const bool nonfinite = slot.float64 && uses_table64 && slot.flags_host[r] != 0;
        if (!device_error.empty() || (!failure.empty() && uses_fallback) || nonfinite) {
            e->state = RequestState::failed;
            e->error = !device_error.empty() ? device_error
                       : nonfinite ? "dsmvc: HIP Float64 2D execution produced NaN or infinity"
                                   : failure;
            ++stats_.failed;
        } else {
            e->state = RequestState::done;
            ++stats_.completed;
        }
        e->completed = t;
    }
    slot.round.clear();
    slot.busy = false;
    slot.fallback = false;
    slot.float64 = false;
    done_cv_.notify_all();
}

void HipServer::loop() {
    std::unique_lock lock(mutex_);
    for (;;) {
        bool progressed = false, busy = false;
        for (auto &slot : slots_) {
            if (!slot->busy) continue;
            const hipError_t q = hipEventQuery(slot->done);
            if (q == hipErrorNotReady) {
                busy = true;
                continue;
            }
            finish_round(*slot, lock, q == hipSuccess ? std::string()
                                                      : std::string("dsmvc: round failed: ")
                                                            + hipGetErrorString(q));
            progressed = true;
        }
        for (auto &slot : slots_) {
            if (queue_.empty()) break;
            if (slot->busy) continue;
            if (!admit_round(*slot, lock)) break;
            progressed = true;
            busy = true;
        }
        if (stop_ && queue_.empty() && !busy) {
            bool any = false;
            for (auto &slot : slots_) any |= slot->busy;
            if (!any) break;
        }
        if (progressed) continue;
        if (busy) work_cv_.wait_for(lock, std::chrono::microseconds(20));
        else work_cv_.wait(lock, [&] { return stop_ || !queue_.empty(); });
    }
}

} // namespace dsmvc

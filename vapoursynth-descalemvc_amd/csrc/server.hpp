// Continuous-batching request loop for descale requests of mixed geometry.
//
// The reference serves concurrent frame requests through VapourSynth's
// fmParallel getFrame calls, each a synchronous executor call bounded by the
// CUDA executor's execution slots (src/cuda/cuda_executor.cpp:1102-1157,
// 1462-1503), with plans cached per executor (acquire_packed_plan,
// 1936-2020).  Here one worker thread owns the GPU side:
//   * submit() plans the request on the caller's thread (host planner LRU,
//     keyed by kernel, geometry, padding and precision), uploads the plan
//     images once (the executor's device plan cache) and queues it;
//   * admission: a queued request starts when a slot (stream) is free and
//     its intermediates fit the device-memory budget (a first-fit arena);
//   * every round takes all admitted requests that arrived while earlier
//     rounds were on the GPU, puts their F32 planes into one set of windowed
//     table launches (horizontal then vertical) on the slot's stream, and
//     solves Float64-route planes through the executor after them; rounds on
//     different slots overlap on the device;
//   * completion is observed per round (HIP event), which releases the
//     memory and wakes waiters.
// A request's result does not depend on what else is in its round: each
// windowed item is computed the same way whatever the table holds, and
// Float64 planes take the executor's own route.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "device_types.hpp"
#include "engine.hpp"

namespace dsmvc::hip {
class DevicePlan;
}

namespace dsmvc {

class HipExecutor;
class WindowTables;

// One request: 1 plane (gray) or 3 planes (YUV with chroma subsampled by
// 2^subsampling in each direction), F32 samples in device memory.
struct ServeRequest {
    std::int32_t planes = 1;
    std::int32_t subsampling_w = 0;
    std::int32_t subsampling_h = 0;
    std::int32_t src_width = 0, src_height = 0;
    std::int32_t dst_width = 0, dst_height = 0;
    KernelSpec kernel{};
    BorderMode border = BorderMode::symmetric;
    F64Mode f64_mode = F64Mode::automatic;
    const float *src[3] = {};
    std::ptrdiff_t src_stride[3] = {};
    float *dst[3] = {};
    std::ptrdiff_t dst_stride[3] = {};
};

// When an F32 plane takes the fused 2D route (for an executor that takes
// it): once its plan pair's operator images exist (built off the submit path
// after the pair's first request), always (built at submit), or never.
enum class ServeFused : std::uint8_t { warm, always, off };

struct ServerConfig {
    std::size_t memory_budget = std::size_t{1} << 30;   // device bytes for in-flight intermediates
    std::int32_t slots = 4;                             // rounds in flight (one stream each)
    std::int32_t max_round_requests = 256;
    ServeFused fused = ServeFused::warm;
};

struct ServerStats {
    std::uint64_t submitted = 0, completed = 0, failed = 0;
    std::uint64_t rounds = 0;
    std::uint64_t table_planes = 0, fallback_planes = 0;
    std::uint64_t largest_round = 0;
    std::size_t memory_high_water = 0;
    std::uint64_t fused_planes = 0;    // of the table planes, those on the fused 2D route
};

enum class RequestState : std::int32_t { pending = 0, done = 1, failed = -1 };

class HipServer {
public:
    HipServer(HipExecutor &executor, const ServerConfig &config);
    ~HipServer();   // finishes every accepted request, then stops
    HipServer(const HipServer &) = delete;
    HipServer &operator=(const HipServer &) = delete;

    // Validates, plans and queues a request; returns its id.  Throws
    // std::invalid_argument for a malformed request or one whose
    // intermediates exceed the whole memory budget.
    std::uint64_t submit(const ServeRequest &request);
    // Blocks until the request has finished.  Returns its state; `error`
    // receives a failure's message.  Times in seconds since the server
    // started: when submit() was entered and when completion was observed.
    RequestState wait(std::uint64_t id, std::string *error, double *submitted, double *completed);
    RequestState poll(std::uint64_t id);
    // Forgets a finished request's record.
    void release(std::uint64_t id);
    [[nodiscard]] ServerStats stats() const;

private:
    struct Plane {
        std::shared_ptr<const AxisPlan> h, v;
        const float *src = nullptr;
        std::ptrdiff_t src_stride = 0;
        float *dst = nullptr;
        std::ptrdiff_t dst_stride = 0;
        bool table = false;          // windowed F32 table launch
        bool table64 = false;        // windowed Float64 table launch (else executor route)
        bool fused = false;          // fused 2D table launch (operator images built)
        std::size_t inter_bytes = 0;
        // The plans' device images, held so that the executor's plan cache
        // keeps them (and their window tables) until the request is gone.
        std::shared_ptr<hip::DevicePlan> ih, iv;
        const hip::DevWindow64 *wh64 = nullptr;
        const hip::DevWindow64 *wv64 = nullptr;
        const hip::DevWindow *wh = nullptr;
        const hip::DevWindow *wv = nullptr;
    };
    struct Entry {
        std::uint64_t id = 0;
        std::vector<Plane> planes;
        std::size_t bytes = 0;       // arena bytes for its intermediates
        std::size_t offset = 0;      // arena offset once admitted
        RequestState state = RequestState::pending;
        std::string error;
        double submitted = 0.0, completed = 0.0;
    };
    struct Slot;

    void loop();
    bool admit_round(Slot &slot, std::unique_lock<std::mutex> &lock);
    void finish_round(Slot &slot, std::unique_lock<std::mutex> &lock, std::string device_error);
    bool arena_alloc(std::size_t bytes, std::size_t &offset);
    void arena_free(std::size_t offset, std::size_t bytes);
    double now() const;

    HipExecutor &executor_;
    ServerConfig config_;
    std::chrono::steady_clock::time_point epoch_;
    void *arena_ = nullptr;
    std::map<std::size_t, std::size_t> free_;   // offset -> size
    std::size_t in_use_ = 0;

    mutable std::mutex mutex_;
    std::condition_variable work_cv_, done_cv_;
    std::deque<std::shared_ptr<Entry>> queue_;
    std::unordered_map<std::uint64_t, std::shared_ptr<Entry>> entries_;
    std::vector<std::unique_ptr<Slot>> slots_;
    std::uint64_t next_id_ = 1;
    bool stop_ = false;
    ServerStats stats_{};
    std::thread worker_;

    // Operator images of F32 plan pairs, built off the submit path by one
    // builder thread; a plane takes the fused route once they exist.
    void builder();
    void request_operators(const std::shared_ptr<hip::DevicePlan> &h, const std::shared_ptr<hip::DevicePlan> &v);
    std::mutex build_mutex_;
    std::condition_variable build_cv_;
    std::deque<std::pair<std::shared_ptr<hip::DevicePlan>, std::shared_ptr<hip::DevicePlan>>> build_queue_;
    std::set<std::pair<const hip::DevicePlan *, const hip::DevicePlan *>> build_pending_;
    bool build_stop_ = false;
    std::thread builder_;
};

} // namespace dsmvc

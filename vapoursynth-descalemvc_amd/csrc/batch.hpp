// Multi-plan batches: many 2D descale jobs of uneven geometry (different
// plans, sizes and buffers) solved by a few launches over device-resident job
// tables.
//
// The reference has no multi-plan launch: its executors solve one plane per
// call (src/vs_plugin.cpp:1636-1745), and its Metal route batches at most 48
// axis jobs per dispatch from a descriptor table (metal_routes.metal:25-35,
// 363-408).  Here every job's horizontal pass goes into one table launch per
// (band class, tap width) group, then every vertical pass likewise, on the
// windowed kernels.  Jobs those kernels cannot take (Float64 route, bands
// above 7, slowly decaying plans) are solved one by one through the
// executor after the table launches.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <vector>

#include "device_types.hpp"
#include "engine.hpp"

namespace dsmvc {

class HipExecutor;
namespace hip {
class DevicePlan;
}

// One 2D job: `frames` F32 planes of the horizontal plan's source width and
// the vertical plan's source height, into planes of the destination size.
struct BatchJob {
    std::shared_ptr<const AxisPlan> horizontal;
    std::shared_ptr<const AxisPlan> vertical;
    const float *input = nullptr;
    std::ptrdiff_t input_row_stride = 0;
    std::ptrdiff_t input_frame_stride = 0;
    float *output = nullptr;
    std::ptrdiff_t output_row_stride = 0;
    std::ptrdiff_t output_frame_stride = 0;
    std::int32_t frames = 1;
};

struct BatchStats {
    std::int32_t jobs = 0;
    std::int32_t table_jobs = 0;       // solved by the table launches
    std::int32_t fallback_jobs = 0;    // solved one by one
    std::int32_t fused_jobs = 0;       // of the table jobs, those on the fused 2D route
    std::int32_t launches = 0;         // table launches per run
    std::int64_t items = 0;            // wavefront items per run
    std::int64_t intermediate_bytes = 0;
    bool graph = false;                // table launches replayed from a HIP graph
};

// Groups windowed axis jobs by (direction, band class, tap width, staging
// block size, sample kind) into one device buffer of job tables and launches
// them: the Float64 route's input checks first, then all horizontal groups,
// then all vertical groups, on one stream.  F32 jobs (WinJob) and Float64
// jobs (Win64Job) form separate groups; Float64 jobs and checks report a
// non-finite value through their own flag pointers.  Reusable:
// `clear` then `add*` lay out a new set of jobs, `upload` (growing the
// buffers) and `launch` enqueue them.
class WindowTables {
public:
    explicit WindowTables(int device);
    ~WindowTables();
    WindowTables(const WindowTables &) = delete;
    WindowTables &operator=(const WindowTables &) = delete;

    void clear();
    // `kind`: input samples of a horizontal job, output samples of a vertical one.
    void add(const hip::WinJob &job, int band, bool rows, hip::Sample kind);
    void add64(const hip::Win64Job &job, int band, bool rows, hip::Sample kind);
    void add_check(const hip::CheckJob &job);
    // A fused 2D job (kernels_operator.hip) whose operator images were
    // packed at `steps` K steps; `kind`: its input and output samples.
    // `exact`: an F32 job's second pass (bfloat16 images), launched after
    // every first pass.
    void add_op(const hip::OpJob &job, int steps, hip::Sample kind, bool exact = false);
    // Device flags [flags, flags + count) that every launch() sets to zero
    // (with a kernel) before its first table launch: the F32 fused jobs'
    // range flags.  clear() forgets them.
    void clear_flags(int *flags, int count);
    // Uploads the tables (stream-ordered when `stream` is given, else
    // synchronously) and returns the number of launches `launch` makes.
    int upload(hipStream_t stream);
    // Launches the uploaded tables.  After a synchronous upload the launches
    // may be captured into a graph and replayed: that slot is never reused
    // by this object unless upload() is called again.
    void launch(hipStream_t stream);
    [[nodiscard]] std::int64_t items() const;

private:
    enum class Kind : std::uint8_t { check = 0, f32 = 1, f64 = 2, op = 3 };
    struct Group {
        Kind type = Kind::f32;
        int band = 0, width = 0;
        bool rows = false;
        bool block16 = false;       // 16-row staging blocks (F32); wide source loads (fused jobs)
        hip::Sample kind = hip::Sample::f32;
        std::vector<hip::WinJob> jobs;
        std::vector<hip::Win64Job> jobs64;
        std::vector<hip::CheckJob> checks;
        std::vector<hip::OpJob> ops;
        std::size_t offset = 0;     // byte offset of the first record in the buffers
        long items = 0;
        [[nodiscard]] std::size_t count() const {
            return jobs.size() + jobs64.size() + checks.size() + ops.size();
        }
        [[nodiscard]] std::size_t bytes() const {
            return jobs.size() * sizeof(hip::WinJob) + jobs64.size() * sizeof(hip::Win64Job)
                   + checks.size() * sizeof(hip::CheckJob) + ops.size() * sizeof(hip::OpJob);
        }
    };
    int *flags_ = nullptr;
    int flag_count_ = 0;
    struct Slot {
        unsigned char *host = nullptr;     // pinned
        unsigned char *device = nullptr;
        std::size_t capacity = 0;          // bytes
        hipEvent_t done = nullptr;
        bool pending = false;
    };
    Group &group_for(Kind type, int band, int width, bool rows, bool block16, hip::Sample kind);
    static constexpr int kSlots = 3;
    int device_;
    std::vector<Group> groups_;
    Slot slots_[kSlots];
    int current_ = 0;
    bool record_ = false;
    int order_ = 2;   // item order of the launches (launch_window_table)
};

class HipBatch {
public:
    // Validates and uploads everything a run needs.  `graph`: capture the
    // table launches into a HIP graph once and replay it on every run.
    // The executor must outlive the batch.
    HipBatch(HipExecutor &executor, std::vector<BatchJob> jobs, bool graph);
    ~HipBatch();
    HipBatch(const HipBatch &) = delete;
    HipBatch &operator=(const HipBatch &) = delete;

    // Enqueues one run on `stream`; waits for it unless `async`.
    void run(hipStream_t stream, bool async);
    [[nodiscard]] const BatchStats &stats() const { return stats_; }

private:
    HipExecutor &executor_;
    std::vector<BatchJob> jobs_;
    std::vector<std::size_t> fallback_;
    std::vector<std::shared_ptr<hip::DevicePlan>> images_;
    WindowTables tables_;
    void *intermediate_ = nullptr;
    int *flags_ = nullptr;             // the fused jobs' range flags, one per job
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    BatchStats stats_{};
};

} // namespace dsmvc

// HIP executor: applies axis plans to device-resident planes on one GPU.
//
// Public surface mirrors the reference's GPU executor and the backend-neutral
// Executor (reference include/dsmvc/engine.hpp:363-504,
// src/cuda/cuda_executor.hpp:11-74): prepare / seal, inverse_rows,
// inverse_columns, inverse_2d, inverse_2d_u8, inverse_2d_u16, with strides in
// elements, the output's logical width only written, std::invalid_argument
// for bad arguments and std::runtime_error for device failures or non-finite
// Float64 results.  Unlike the reference, buffers are device pointers (HBM
// resident) and batches of frames are first class.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <condition_variable>
#include <cstdint>
#include <list>
#include <memory>
#include <mutex>
#include <unordered_map>

#include "device_plan.hpp"
#include "device_types.hpp"
#include "engine.hpp"

namespace dsmvc {
namespace hip {
struct ChunkArgs;
}
class WindowTables;

enum class SolveMode : std::uint8_t {
    automatic = 0,   // chunked or windowed where the plan allows it (autotuned), else sequential
    exact = 1,       // sequential: bit-identical to the ordered CPU solver
    chunked = 2,     // carry-chain chunked kernels only; throws when the plan cannot use them
    windowed = 3,    // windowed kernels only; throws when the plan cannot use them
    refined = 4,     // as automatic, but plans that retain FP64 data take the F32 route
                     // plus refinement steps instead of the Float64 route
    fused = 5,       // as automatic, and 2D calls take the fused route wherever both plans
                     // fit it (automatic does too unless DSMVC_HIP_FUSED=0)
};

// Modes whose solves choose among the parallel kernels by themselves.
constexpr bool automatic_like(SolveMode m) { return m == SolveMode::automatic || m == SolveMode::fused; }

enum class SampleKind : std::uint8_t { float32 = 0, u8 = 1, u16 = 2 };

// A batch of `frames` planes: frame f's row r starts at base + f * frame_stride
// + r * row_stride (elements).
struct PlaneBatch {
    const void *input = nullptr;
    std::ptrdiff_t input_row_stride = 0;
    std::ptrdiff_t input_frame_stride = 0;
    void *output = nullptr;
    std::ptrdiff_t output_row_stride = 0;
    std::ptrdiff_t output_frame_stride = 0;
    std::int32_t frames = 1;
};

// Per-launch tuning the executor chooses by itself unless overridden.
// Per direction: a chunk count (carry-chain kernels), -1 (windowed kernels)
// or 0 (automatic: autotuned when enabled).
struct LaunchTuning {
    std::int32_t rows_chunks = 0;
    std::int32_t columns_chunks = 0;
};

class HipExecutor {
public:
    explicit HipExecutor(int device = -1, SolveMode mode = SolveMode::automatic);
    ~HipExecutor();
    HipExecutor(const HipExecutor &) = delete;
    HipExecutor &operator=(const HipExecutor &) = delete;

    [[nodiscard]] int device() const noexcept { return device_; }
    [[nodiscard]] SolveMode mode() const noexcept { return mode_; }
    [[nodiscard]] const char *name() const noexcept { return "hip"; }

    void prepare(const std::shared_ptr<const AxisPlan> &plan);
    void seal();
    void set_tuning(const LaunchTuning &t) { tuning_ = t; }
    // Autotuning of the chunked kernels' partition (on by default; the
    // environment variable DSMVC_HIP_AUTOTUNE=0 turns it off).  The first
    // chunked pass of a (plan, direction, sample kind) that fills the GPU
    // (at least two workgroups per compute unit) times every feasible chunk
    // count once on the call's own buffers and keeps the fastest; a smaller
    // count must win by 10% to replace the largest.  Passes before that use
    // the largest partition.  The choice is fixed for the executor's lifetime,
    // so every later result of that key is bit-reproducible within it; the
    // partition affects the last bits of Float32 results (carry rounding).
    void set_autotune(bool enabled) { autotune_ = enabled; }
    // Refinement steps per axis in refined mode (1..3).
    void set_refinement(int steps) {
        if (steps < 1 || steps > 3) throw std::invalid_argument("dsmvc: refinement steps must be 1..3");
        refine_steps_ = steps;
    }
    // The tuned route of a key: a chunk count, kWindowRoute (-1), 0 before
    // tuning, -2 for a plan this executor has not seen.
    [[nodiscard]] int tuned_chunks(const AxisPlan &plan, bool rows, int kind);

    // Float32 planes (device pointers).  Synchronous unless `async`.
    void inverse_rows(const AxisPlan &plan, const float *input, std::ptrdiff_t input_row_stride,
                      float *output, std::ptrdiff_t output_row_stride, std::int32_t row_count,
                      hipStream_t stream = nullptr, bool async = false);
    void inverse_columns(const AxisPlan &plan, const float *input,
                         std::ptrdiff_t input_row_stride, float *output,
                         std::ptrdiff_t output_row_stride, std::int32_t column_count,
                         hipStream_t stream = nullptr, bool async = false);
    void inverse_2d(const AxisPlan &horizontal, const AxisPlan &vertical, const float *input,
                    std::ptrdiff_t input_row_stride, float *output,
                    std::ptrdiff_t output_row_stride, hipStream_t stream = nullptr,
                    bool async = false);
    void inverse_2d_u8(const AxisPlan &horizontal, const AxisPlan &vertical,
                       const std::uint8_t *input, std::ptrdiff_t input_row_stride,
                       std::uint8_t *output, std::ptrdiff_t output_row_stride,
                       const IntegerConversion &conversion, hipStream_t stream = nullptr,
                       bool async = false);
    void inverse_2d_u16(const AxisPlan &horizontal, const AxisPlan &vertical,
                        const std::uint16_t *input, std::ptrdiff_t input_row_stride,
                        std::uint16_t *output, std::ptrdiff_t output_row_stride,
                        const IntegerConversion &conversion, hipStream_t stream = nullptr,
                        bool async = false);

    // The general 2D entry: a batch of frames of one sample kind.
    void inverse_2d_batch(const AxisPlan &horizontal, const AxisPlan &vertical, SampleKind kind,
                          const PlaneBatch &batch, const IntegerConversion *conversion,
                          hipStream_t stream, bool async);

    // Several planes in one set of launches -- a YUV frame batch's luma and
    // chroma, each with its own plans and conversion: every plane's
    // horizontal pass goes into one windowed table launch per kernel class,
    // then every vertical pass.  Planes the windowed kernels cannot take
    // (Float64 route, refined, exact or chunked modes, wide bands) are solved
    // one by one after them.
    struct PlaneJob {
        const AxisPlan *horizontal = nullptr;
        const AxisPlan *vertical = nullptr;
        PlaneBatch batch{};
        const IntegerConversion *conversion = nullptr;
    };
    void inverse_2d_planes(const PlaneJob *planes, std::int32_t count, SampleKind kind,
                           hipStream_t stream, bool async);

    // One axis over a batch, for F32 plans: rows of `kind` samples into a
    // float32 intermediate, or columns of a float32 intermediate into `kind`
    // samples -- the two halves of inverse_2d_batch, callable on their own.
    void inverse_rows_batch(const AxisPlan &plan, SampleKind in_kind, const PlaneBatch &batch,
                            std::int32_t row_count, const IntegerConversion *conversion,
                            hipStream_t stream, bool async);
    void inverse_columns_batch(const AxisPlan &plan, SampleKind out_kind, const PlaneBatch &batch,
                               std::int32_t column_count, const IntegerConversion *conversion,
                               hipStream_t stream, bool async);

    // inverse_2d_batch without any allocation, for callers that cannot
    // allocate (code under HIP graph capture, tight serving loops): every
    // buffer beyond the caller's input and output lives in a fixed device
    // array of kNoallocScratchBytes per GPU (noalloc.cpp).  Both plans must
    // have been prepared on this executor (prepare builds their device
    // images and windowed tables).  Routes: the windowed kernels where the
    // prepared tables allow, else the sequential kernels (always in exact
    // mode); no autotuning, and refined mode is refused.  A request whose
    // intermediate and forward sweeps exceed the array is refused.  The
    // Float64 route needs a synchronous call (it reports non-finite results
    // on return).  Calls on one GPU share the array: calls that overlap on
    // the device (different streams) race.
    void inverse_2d_batch_noalloc(const AxisPlan &horizontal, const AxisPlan &vertical,
                                  SampleKind kind, const PlaneBatch &batch,
                                  const IntegerConversion *conversion, hipStream_t stream,
                                  bool async);

    // Waits for `stream` and reports a deferred Float64 non-finite failure.
    void synchronize(hipStream_t stream);
    // The device flag of `stream`'s Float64 failures, armed: cleared on the
    // stream unless an unreported failure is pending.  Table launches of
    // Float64 jobs made by other components on that stream may point their
    // jobs at it; synchronize(stream) reports it.
    int *float64_flag(hipStream_t stream);
    // Waits for `stream` and frees the scratch and job tables kept for it.
    // Call before destroying a stream this executor has run work on (the
    // executor keeps per-stream state until then, or until it is destroyed).
    void release_stream(hipStream_t stream);

    // Host-frame entries: the reference's executors take host buffers
    // (include/dsmvc/engine.hpp:455-490) and stage them through pinned
    // memory on a stream of their own (src/cuda/cuda_executor.cpp:2798-2822).
    // These do the same on an execution slot: the input is packed into the
    // slot's pinned buffer, copied to the device, solved, copied back and
    // unpacked into the caller's rows (logical width only).  Synchronous.
    void inverse_2d_host(const AxisPlan &horizontal, const AxisPlan &vertical, SampleKind kind,
                         const PlaneBatch &host, const IntegerConversion *conversion);
    void inverse_rows_host(const AxisPlan &plan, const float *input,
                           std::ptrdiff_t input_row_stride, float *output,
                           std::ptrdiff_t output_row_stride, std::int32_t row_count);
    void inverse_columns_host(const AxisPlan &plan, const float *input,
                              std::ptrdiff_t input_row_stride, float *output,
                              std::ptrdiff_t output_row_stride, std::int32_t column_count);

    // GPU frames exported by another API (reference include/dsmvc/engine.hpp:
    // 306-326, 414-442: VapourSynth's GPU core exports a frame's memory as a
    // POSIX file descriptor, src/cuda/cuda_executor.cpp:1180-1300,
    // 3640-3900).  The allocation is imported once per memory_id
    // (hipImportExternalMemory, then mapped) and kept in a bounded LRU; the
    // plane lives at `offset` with `pitch_bytes` between rows.  The handle is
    // consumed by every call, whatever the outcome: a successful import owns
    // it, any other path closes it.  An input whose producer pair names a
    // semaphore imported with import_exported_semaphore is waited for on the
    // device (hipWaitExternalSemaphoresAsync); otherwise the caller has
    // waited.  Synchronous: the output is written when the call returns.
    void inverse_2d_exported(const AxisPlan &horizontal, const AxisPlan &vertical, SampleKind kind,
                             const ExportedPlane &input, const ExportedPlane &output,
                             const IntegerConversion *conversion);
    void inverse_rows_exported(const AxisPlan &plan, const ExportedPlane &input,
                               const ExportedPlane &output, std::int32_t row_count);
    void inverse_columns_exported(const AxisPlan &plan, const ExportedPlane &input,
                                  const ExportedPlane &output, std::int32_t column_count);
    // A timeline semaphore exported by the frame producer, keyed by its
    // producer-side handle value; returns whether it could be imported (the
    // handle is consumed either way).
    bool import_exported_semaphore(std::uint64_t semaphore, int handle_type, std::intptr_t handle);
    [[nodiscard]] bool has_exported_semaphore(std::uint64_t semaphore);
    struct ImportStats {
        std::uint64_t entries = 0, imports = 0, hits = 0, evictions = 0;
    };
    ImportStats import_stats();
    // Imported exported allocations and semaphores (exported.hpp).
    class Imports;

    // The device image of a plan (uploaded on first use), for callers that
    // launch the kernels themselves (multi-plan batches, the request loop).
    // The cache never evicts an image while a returned pointer to it lives.
    std::shared_ptr<hip::DevicePlan> plan_image(const AxisPlan &plan);

    // Device plan cache: images keyed by the plan's build serial (one per
    // planner request key), least recently used first out once their bytes
    // (device image plus every table built for it) exceed the limit --
    // DSMVC_HIP_PLAN_CACHE_MB, default 4096 MiB (the reference bounds its
    // packed-plan LRU the same way, src/cuda/cuda_executor.cpp:219-235,
    // 1936-2020).  Images in use (held by a call, a batch or a served
    // request) are never evicted.
    struct PlanCacheStats {
        std::uint64_t entries = 0, bytes = 0, limit = 0;
        std::uint64_t hits = 0, misses = 0, evictions = 0;
    };
    PlanCacheStats plan_cache_stats();
    void set_plan_cache_limit(std::size_t bytes);

    // Whether 2D calls of `kind` take the fused route where it fits.
    [[nodiscard]] bool fused_enabled(SampleKind kind) const { return fused_route(kind); }
    // add_fused (below) under the executor's lock, for batches and served
    // rounds, whose workgroups take whole columns (many jobs per launch).
    // False when this executor does not take the fused route for `kind` or
    // the plans do not fit it.
    bool fused_job(const AxisPlan &h, const AxisPlan &v, SampleKind kind, const PlaneBatch &b,
                   const IntegerConversion *conversion, WindowTables &tables,
                   std::vector<std::shared_ptr<hip::DevicePlan>> &pins, int *flag);

private:
    struct Scratch {
        void *buffer = nullptr;
        std::size_t bytes = 0;
        int *flag = nullptr;
        int *flag_host = nullptr;   // pinned: the flag's readback lands here
        bool flag_armed = false;
    };

    // The device address of this GPU's no-allocation scratch array,
    // resolved once per device when an executor is created (noalloc.cpp).
    static void *noalloc_setup(int device);
    hip::DevicePlan *prepared_plan(const AxisPlan &plan);
    void *noalloc_scratch_ = nullptr;

    // Execution slots: synchronous calls made without a stream each take a
    // slot -- a stream of the executor's own (blocking with respect to the
    // null stream, default priority) with its host staging buffers -- so that
    // concurrent callers (fmParallel getFrame threads) overlap on the device
    // and wait for their own work without holding the executor's lock
    // (the reference bounds concurrency the same way,
    // src/cuda/cuda_executor.cpp:1102-1157, 1462-1503).
    struct Slot {
        hipStream_t stream = nullptr;
        bool busy = false;
        void *pinned = nullptr;     // host staging (input, then output)
        std::size_t pinned_bytes = 0;
        void *device = nullptr;     // device copies of host frames
        std::size_t device_bytes = 0;
    };
    static constexpr int kMaxSlots = 8;
    std::array<Slot, kMaxSlots> slots_{};
    std::mutex slot_mutex_;
    std::condition_variable slot_cv_;
    Slot *acquire_slot();
    void release_slot(Slot *slot);
    void slot_buffers(Slot &slot, std::size_t pinned, std::size_t device);
    // Runs `enqueue(stream) -> Scratch &` under the lock on the caller's
    // stream, or on a slot for a synchronous call without one, then waits
    // without the lock and reports a Float64 failure.
    template <class Enqueue>
    void run_call(hipStream_t stream, bool async, const char *what, Enqueue &&enqueue);
    void complete(hipStream_t stream, Scratch &s, const char *what);

    std::shared_ptr<hip::DevicePlan> device_plan(const AxisPlan &plan);
    void evict_plans();
    Scratch &scratch(hipStream_t stream, std::size_t bytes);
    void finish(hipStream_t stream, Scratch &s, bool async, const char *what);
    // Refined solves keep the residual and its correction side by side in
    // the scratch region that otherwise holds one forward sweep.
    std::size_t refine_regions() const { return mode_ == SolveMode::refined ? 2 : 1; }
    bool float64_route(const AxisPlan &plan) const {
        return plan.requires_float64() && mode_ != SolveMode::refined;
    }
    void solve_pass(const AxisPlan &plan, bool rows, const void *input, hip::Sample in_kind,
                    std::ptrdiff_t in_row, std::ptrdiff_t in_frame, void *output,
                    hip::Sample out_kind, std::ptrdiff_t out_row, std::ptrdiff_t out_frame,
                    std::int32_t vectors, std::int32_t frames, const hip::DevConv &conv, bool f64,
                    Scratch &s, std::size_t fwd_offset, hipStream_t stream);
    Scratch &solve_2d(const AxisPlan &h, const AxisPlan &v, SampleKind kind, const PlaneBatch &b,
                      const IntegerConversion *conversion, hipStream_t stream);
    Scratch &enqueue_planes(const PlaneJob *planes, std::int32_t count, SampleKind kind,
                            hipStream_t stream);
    // True when this executor's 2D calls take the fused route where it fits:
    // in fused mode, and in automatic mode unless DSMVC_HIP_FUSED=0.
    bool fused_route(SampleKind) const {
        return mode_ == SolveMode::fused || (mode_ == SolveMode::automatic && fused_auto_);
    }
    // Adds one plane batch to `tables` as a fused 2D job when both plans fit
    // the fused kernel (and pins their images); false otherwise.  F32
    // planes add two jobs (the FP16 pass and the exact pass it may need)
    // sharing `flag`, which the tables clear before their launches
    // (WindowTables::clear_flags).  `whole_columns`: a workgroup takes a
    // whole column of the plane; otherwise columns are split until the
    // launch fills the device, `launch_columns` counting the workgroup
    // columns of every plane the launch takes.
    bool add_fused(const AxisPlan &h, const AxisPlan &v, SampleKind kind, const PlaneBatch &b,
                   const IntegerConversion *conversion, WindowTables &tables,
                   std::vector<std::shared_ptr<hip::DevicePlan>> &pins, int *flag,
                   bool whole_columns = false, long launch_columns = 0);
    void arm_flag(Scratch &s, hipStream_t stream);
    WindowTables &stream_tables(hipStream_t stream);
    void axis_pass(const AxisPlan &plan, bool rows, const void *input, hip::Sample in_kind,
                   std::ptrdiff_t in_row, std::ptrdiff_t in_frame, void *output,
                   hip::Sample out_kind, std::ptrdiff_t out_row, std::ptrdiff_t out_frame,
                   std::int32_t vectors, std::int32_t frames, const hip::DevConv &conv,
                   bool f64, Scratch &s, std::size_t fwd_offset, hipStream_t stream);

    static constexpr int kWindowRoute = -1;
    int choose_partition(hip::DevicePlan &dp, bool rows, const hip::ChunkArgs &base,
                         hip::Sample kind, hipStream_t stream);
    hipError_t launch_route(hip::DevicePlan &dp, int route, bool rows, const hip::ChunkArgs &a,
                            hip::Sample kind, hipStream_t stream);

    struct TuneKey {
        const hip::DevicePlan *plan;
        bool rows;
        int kind;
        bool operator==(const TuneKey &o) const {
            return plan == o.plan && rows == o.rows && kind == o.kind;
        }
    };
    struct TuneKeyHash {
        std::size_t operator()(const TuneKey &k) const noexcept {
            return std::hash<const void *>()(k.plan) ^ (static_cast<std::size_t>(k.rows) << 1)
                ^ (static_cast<std::size_t>(k.kind) << 2);
        }
    };

    int device_;
    SolveMode mode_;
    LaunchTuning tuning_{};
    bool autotune_ = true;
    bool fused_auto_ = true;
    int refine_steps_ = 1;
    int compute_units_ = 256;
    std::unordered_map<TuneKey, int, TuneKeyHash> tuned_;
    std::unordered_map<hipStream_t, std::unique_ptr<WindowTables>> plane_tables_;
    std::mutex mutex_;
    struct PlanEntry {
        std::shared_ptr<hip::DevicePlan> image;
        std::list<std::uint64_t>::iterator lru;
    };
    static std::uint64_t plan_key(const AxisPlan &plan);
    std::unordered_map<std::uint64_t, PlanEntry> plans_;
    std::list<std::uint64_t> lru_;   // most recent first
    std::size_t plan_cache_limit_ = std::size_t{4096} << 20;
    // Device bytes of every image this executor built that is still alive
    // (cached, or evicted but held by a call, batch or served request).
    hip::ByteAccount cache_bytes_ = std::make_shared<std::atomic<std::size_t>>(0);
    PlanCacheStats plan_stats_{};
    std::unordered_map<hipStream_t, Scratch> scratch_;
    bool sealed_ = false;

    std::unique_ptr<Imports> imports_;
    Imports &imports();
    std::mutex imports_mutex_;
};

} // namespace dsmvc

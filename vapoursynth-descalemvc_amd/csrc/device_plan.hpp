// Device images of axis plans and chunk partitions.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "device_types.hpp"
#include "engine.hpp"
#include "operator_plan.hpp"

namespace dsmvc::hip {

// Host-side packing of one AxisPlan into the layouts the kernels read:
// CSR arrays as-is, coefficient rows (see device_types.hpp), and the FP64
// route's rows.  Pure host code, unit-testable without a device.
struct PackedPlan {
    std::int32_t n = 0, m = 0, w = 0, nnz = 0;
    std::int32_t coef_stride = 16;
    bool zform = false;
    std::vector<float> coef;         // [n * coef_stride]
    std::vector<double> coef64;      // [n * coef_stride]
    std::vector<double> wt64;        // [nnz]
};
PackedPlan pack_plan(const AxisPlan &plan);

// A chunk partition of a plan for the chunked kernels (layout in
// device_types.hpp, DevChunks): padded ELL taps and coefficient rows, the
// FP64-built transfer matrices (stored F32) and the horizontal kernel's
// staging windows.
struct ChunkPlan {
    std::int32_t chunks = 0;
    std::int32_t length = 0;
    std::int32_t width = 0;          // packed taps per row
    std::int32_t blocks = 0;         // staging blocks per chunk
    std::int32_t block_rows = 8;     // destination rows per staging block
    bool rows_feasible = false;      // every 8-row block's taps fit 32 columns
    std::vector<float> tf, tb;       // [chunks][W][W]
    std::vector<std::int32_t> ell_idx;
    std::vector<float> ell_w;
    std::vector<float> coef;         // [chunks * length][16]
    std::vector<std::int32_t> blk_lo;
};
// Partition into at most `max_chunks` chunks; false when the plan cannot use
// the chunked kernels (band, length or tap count out of range).
bool build_chunk_plan(const AxisPlan &plan, int max_chunks, ChunkPlan &out);

// A windowed partition (layout in device_types.hpp, DevWindow and
// DevWindow64): padded coefficient rows and tap runs with P zero rows on each
// side, the staging windows and the warm-up check, in F32 (T = float) or for
// the Float64 route (T = double).
template <class T> struct WindowPlanT {
    std::int32_t band = 0;           // band class the kernels run (1, 3, 5, 7)
    std::int32_t chunks = 0;
    std::int32_t length = 0;
    std::int32_t warmup = 0;
    std::int32_t width = 0;          // tap run length K
    bool rows_feasible = false;      // every 8-row block's taps fit 32 samples
    bool zform = false;              // forward state is z = L^-1 b (FP64 factors)
    double residual = 0.0;           // worst relative influence of a chunk's zero start
    std::vector<T> coef;             // [rows][16] kMaxBand records (warm-up check)
    std::vector<T> fwd;              // [rows][window_fwd_words(band, width)]
    std::vector<T> bwd;              // [rows][window_bwd_words(band)]
    std::vector<std::int32_t> blk_lo; // [rows / 8]
    std::vector<std::int32_t> blk_lo16; // [rows / 16]
    bool block16 = false;            // every 16-row block's taps fit 32 samples
    std::vector<std::int32_t> tap_start; // [rows] contiguous tap runs
    std::vector<T> tap_w;                // [rows][K] (copied into fwd)
};
using WindowPlan = WindowPlanT<float>;
using WindowPlan64 = WindowPlanT<double>;
// Largest influence (relative) of a wrong starting state allowed at a
// windowed chunk's first forward / last backward row: 2^-28 for the F32
// kernels, 2^-40 for the Float64 route (so that rounding its result to F32
// almost never differs from the sequential Float64 solve).
constexpr double kWindowTolerance = 1.0 / 268435456.0;
constexpr double kWindowTolerance64 = 1.0 / 1099511627776.0;
// False when the plan cannot use the windowed kernels: band above 7, a tap
// run longer than 16 or outside its 32-sample staging window, or a warm-up
// longer than the band class's P would be needed.
bool build_window_plan(const AxisPlan &plan, WindowPlan &out);
// The same for the Float64 route, from the plan's FP64 packing.
bool build_window_plan64(const AxisPlan &plan, const PackedPlan &packed, WindowPlan64 &out);
// The plan's system G d = r with identity taps (row i reads element i of r),
// for the correction solve of refinement: the windowed kernels then solve a
// given right-hand side of n elements per vector.
bool build_solve_window_plan(const AxisPlan &plan, WindowPlan &out);

// A running total of device bytes, shared by the images that report to it.
using ByteAccount = std::shared_ptr<std::atomic<std::size_t>>;

// One device allocation holding a plan's image (and its chunk partitions).
// Every byte it allocates, now or in a later table build, is added to
// `account` (when given) and taken off again when the image is destroyed.
class DevicePlan {
public:
    DevicePlan(const AxisPlan &plan, int device, ByteAccount account = {});
    ~DevicePlan();
    DevicePlan(const DevicePlan &) = delete;
    DevicePlan &operator=(const DevicePlan &) = delete;

    [[nodiscard]] const DevAxis &axis() const { return axis_; }
    [[nodiscard]] const AxisPlan &host() const { return *plan_; }
    // Chunk partition (at most `max_chunks` chunks) on the device; nullptr when
    // the plan cannot use the chunked kernels in that direction.  Built
    // lazily, cached.
    const DevChunks *chunks(int max_chunks, bool rows);
    // Windowed partition on the device; nullptr when the plan cannot use the
    // windowed kernels in that direction.  Built lazily, cached.
    const DevWindow *window(bool rows);
    // Windowed FP64 partition (Float64 route); nullptr when the plan's FP64
    // recurrence does not decay fast enough.  Built lazily, cached.
    const DevWindow64 *window64();
    // Windowed tables of build_solve_window_plan (source size n); nullptr
    // when the windowed kernels cannot take the plan in that direction.
    // Built lazily, cached.
    const DevWindow *solve_window(bool rows);
    // Refinement tables (FP64 weights and normal matrix split into F32
    // pairs, and tap-major FP64 copies); nullptr for a plan without FP64
    // data.  Built lazily, cached.
    const DevRefine *refine();
    // The plan's explicit operator for the fused 2D route, on the host;
    // nullptr when no direction can take the route.  Built lazily, cached.
    const OperatorPlan *operator_plan();
    // Its MFMA operand image packed at `steps` K steps, FP16 (bf16 = false)
    // or bfloat16 parts; nullptr when the plan cannot take the route in the
    // given direction at that step count.  Built lazily, cached.
    const DevOperator *op(bool rows, bool bf16, int steps);
    // The explicit operator's K steps if operator_plan() was built already
    // (0: no direction can take the route); -1 if it was not built yet, or
    // another thread holds the plan.  Never builds and never waits.
    int built_op_steps();
    // op(rows, bf16, steps) only if an earlier call built it; nullptr
    // otherwise, or while another thread holds the plan.  Never builds and
    // never waits.
    const DevOperator *built_op(bool rows, bool bf16, int steps);
    [[nodiscard]] std::size_t device_bytes() const { return bytes_; }
    // Device bytes of the image and of every table built for it so far.
    [[nodiscard]] std::size_t total_bytes() const { return total_bytes_.load(std::memory_order_relaxed); }
    // The windowed tables only if an earlier window() / window64() call
    // built them; never builds or allocates (nullptr otherwise).
    const DevWindow *built_window(bool rows);
    const DevWindow64 *built_window64();

private:
    struct ChunkImage {
        DevChunks dev{};
        void *memory = nullptr;
        std::size_t bytes = 0;
        bool ok = false;
        bool rows_feasible = false;
    };
    std::shared_ptr<const AxisPlan> keep_;
    const AxisPlan *plan_;
    int device_;
    void *memory_ = nullptr;
    std::size_t bytes_ = 0;
    DevAxis axis_{};
    std::mutex mutex_;
    std::map<int, ChunkImage> chunk_images_;
    struct WindowImage {
        DevWindow dev{};
        void *memory = nullptr;
        std::size_t bytes = 0;
        bool built = false;
        bool ok = false;
        bool rows_feasible = false;
    };
    void *upload_window(const WindowPlan &wp, DevWindow &dev, std::size_t &bytes);
    WindowImage window_image_;
    WindowImage solve_image_;
    struct Window64Image {
        DevWindow64 dev{};
        void *memory = nullptr;
        std::size_t bytes = 0;
        bool built = false;
        bool ok = false;
    };
    Window64Image window64_image_;
    struct OperatorImage {
        DevOperator dev{};
        void *memory = nullptr;
        bool ok = false;
    };
    std::unique_ptr<OperatorPlan> op_plan_;
    bool op_built_ = false;
    std::map<int, OperatorImage> op_images_;   // key: bf16 * 8 + steps
    const OperatorPlan *operator_plan_locked();
    DevRefine refine_{};
    void *refine_memory_ = nullptr;
    std::size_t refine_bytes_ = 0;
    std::atomic<std::size_t> total_bytes_{0};
    ByteAccount account_;
    void grow(std::size_t bytes);
    bool refine_built_ = false;
};

} // namespace dsmvc::hip

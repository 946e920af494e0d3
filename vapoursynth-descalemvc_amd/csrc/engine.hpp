// Host-side engine types of the HIP descale executor.
//
// The value types keep the reference engine's names, fields and numbering
// (reference include/dsmvc/engine.hpp:16-173) so that a caller of the
// reference's AxisRequest / AxisPlan / IntegerConversion can switch to this
// library unchanged.  Everything below them (planner internals, device plan
// images, executors) is this library's own design.
#pragma once

#include <atomic>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <vector>

namespace dsmvc {

enum class KernelKind : std::uint8_t {
    bilinear, bicubic, lanczos, spline16, spline36, spline64, custom,
};

// Reference BackendKind (engine.hpp:25-31) plus `hip`, the backend this
// library provides.
enum class BackendKind : std::uint8_t {
    automatic, cpu, metal, vulkan, cuda, hip,
};

enum class BorderMode : std::uint8_t {
    zero = 0, repeat = 1, reflect101 = 2, symmetric = 3, mirror = 4,
};

enum class F64Mode : std::uint8_t {
    automatic = 0, float32_only = 1, float64_only = 2,
};

struct KernelSpec {
    KernelKind kind = KernelKind::bicubic;
    std::int32_t taps = 3;
    double b = 0.0;
    double c = 0.5;
    double blur = 1.0;
};

struct AxisRequest {
    std::int32_t source_size = 0;
    std::int32_t destination_size = 0;
    double active_length = 0.0;
    double shift = 0.0;
    KernelSpec kernel{};
    BorderMode border = BorderMode::symmetric;
    F64Mode f64_mode = F64Mode::automatic;
};

// Immutable per-axis plan: CSR of A^T by destination element, the banded
// LDL^T factors packed for F32 substitution, and -- for plans that need the
// Float64 route -- the retained FP64 data.  Field meaning follows the
// reference AxisPlan (engine.hpp:82-107).
struct AxisPlan {
    std::int32_t source_size = 0;
    std::int32_t destination_size = 0;
    std::int32_t support = 0;
    std::int32_t half_bandwidth = 0;
    double active_length = 0.0;
    double shift = 0.0;

    std::vector<std::uint32_t> transpose_offsets;  // [n + 1]
    std::vector<std::int32_t> transpose_indices;   // [nnz], ascending per row
    std::vector<float> transpose_weights;          // [nnz]
    std::vector<float> lower_ld;                   // [(d - 1) * n + i] = L(i, i-d) D(i-d)
    std::vector<float> upper_l;                    // [(d - 1) * n + i] = L(i+d, i)
    std::vector<float> inverse_diagonal;           // [n]

    double normal_rcond = 1.0;
    double normal_inf_norm = 0.0;
    std::vector<double> transpose_weights_f64;
    std::vector<double> normal_bands_f64;          // [d * n + i] = G(i, i+d)
    std::vector<double> ldlt_bands_f64;            // [0 * n + i] = D_i, [d * n + i] = L(i+d, i)
    std::vector<double> inverse_diagonal_f64;

    // Nonzero for plans made by build_axis_plan: unique per build (copies
    // keep it), so two plans with the same nonzero serial hold the same
    // arrays.  Plans are immutable once built.
    std::uint64_t serial = 0;

    // The result of the last full check, kept by the plan object itself and
    // not by its copies (a copy starts unchecked); a plan is not changed
    // after it has been used (reference engine.hpp:109-137 memoises the same
    // way).
    struct CheckedFlag {
        mutable std::atomic<std::uint8_t> state{0};   // 0 unchecked, 1 valid, 2 invalid
        CheckedFlag() noexcept = default;
        CheckedFlag(const CheckedFlag &) noexcept {}
        CheckedFlag &operator=(const CheckedFlag &) noexcept {
            state.store(0, std::memory_order_relaxed);
            return *this;
        }
    };
    CheckedFlag checked;

    [[nodiscard]] bool valid() const noexcept;
    // The full check behind valid(), every call.
    [[nodiscard]] bool check_arrays() const noexcept;
    [[nodiscard]] bool requires_float64() const noexcept { return !ldlt_bands_f64.empty(); }
    [[nodiscard]] std::size_t storage_bytes() const noexcept;
};

struct IntegerConversion {
    float input_offset = 0.0F;
    float input_scale = 1.0F;
    float output_scale = 1.0F;
    float output_offset = 0.0F;
    std::uint32_t output_maximum = 255U;
};

struct PlannerCacheStats {
    std::uint64_t plan_hits = 0;
    std::uint64_t plan_builds = 0;
    std::size_t plan_entries = 0;
    std::size_t plan_resident_bytes = 0;
    std::uint64_t geometry_hits = 0;
    std::uint64_t geometry_builds = 0;
    std::size_t geometry_entries = 0;
    std::size_t geometry_resident_bytes = 0;
};

// Planner (host, FP64).  Throws std::invalid_argument for a malformed request
// and std::runtime_error / std::length_error for a plan that cannot be built,
// with the reference's messages.
// Custom kernels (KernelKind::custom, support = kernel.taps) are evaluated
// through the callback at |distance| on the host; their plans are never
// cached (reference include/dsmvc/engine.hpp:78, src/axis_plan.cpp:535-599,
// 1367-1372).
using CustomKernel = std::function<double(double)>;
[[nodiscard]] AxisPlan build_axis_plan(const AxisRequest &request,
                                       const CustomKernel &custom_kernel = {});
[[nodiscard]] std::shared_ptr<const AxisPlan> get_or_build_axis_plan(
    const AxisRequest &request, const CustomKernel &custom_kernel = {});
// A plan from caller-supplied arrays (for example the reference planner's
// AxisPlan), checked by valid(); throws std::invalid_argument when invalid.
// The returned plan has a fresh serial.
[[nodiscard]] std::shared_ptr<const AxisPlan> import_axis_plan(AxisPlan plan);
[[nodiscard]] std::uint64_t next_plan_serial() noexcept;
// The cached plan for `request` when it is built, else nullptr; never builds.
[[nodiscard]] std::shared_ptr<const AxisPlan> find_axis_plan(const AxisRequest &request);
[[nodiscard]] PlannerCacheStats planner_cache_stats();

// One plane of a GPU frame exported by its producer for a foreign API
// (reference include/dsmvc/engine.hpp:306-326): the whole allocation is
// imported once (keyed by memory_id) and the plane lives at `offset`.
struct ExportedPlane {
    std::uint64_t memory_id = 0;     // stable for the allocation's lifetime
    std::uint64_t memory_size = 0;
    std::uint64_t offset = 0;
    std::uint64_t plane_bytes = 0;   // pitch * height
    int handle_type = 0;             // VkExternalMemoryHandleTypeFlagBits (0x1: opaque fd)
    std::intptr_t handle = -1;
    std::ptrdiff_t pitch_bytes = 0;
    std::uint64_t ready_semaphore = 0;   // producer's timeline semaphore, 0: ready now
    std::uint64_t ready_value = 0;
};
// Closes the handle of an exported plane that was never passed to an executor.
void close_exported_plane(const ExportedPlane &plane) noexcept;
void clear_planner_caches();

[[nodiscard]] std::int32_t padding_index(std::int64_t index, std::int32_t size, BorderMode mode);

[[nodiscard]] BackendKind parse_backend(std::string_view name);
[[nodiscard]] BackendKind resolve_backend(BackendKind requested);
[[nodiscard]] const char *backend_name(BackendKind kind) noexcept;
[[nodiscard]] bool hip_available() noexcept;

} // namespace dsmvc

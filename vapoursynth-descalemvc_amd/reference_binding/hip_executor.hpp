// The HIP executor as a backend of the reference engine.
//
// This file and hip_executor.cpp are the reference-side half of the drop-in
// boundary: they compile against the reference's own public header
// (include/dsmvc/engine.hpp) and give the reference's Executor::Impl
// (src/executor.cpp:39-83) a member with the surface of its CUDA executor
// (src/cuda/cuda_executor.hpp:11-74, host-buffer entries), implemented
// entirely through the C ABI of libdsmvc_hip.so (include/dsmvc_hip.h).  A
// maintainer adds them under src/hip/, links libdsmvc_hip.so, and routes
// BackendKind::hip to this class as the CUDA branch routes to CudaExecutor.
//
// The reference's plans are used as they are: each AxisPlan is imported into
// the HIP library (dsmvc_hip_plan_import), once for plans registered through
// prepare() -- kept alive by the shared_ptr, as the CUDA executor's packed
// plan cache keeps its plans -- and per call for borrowed plans, which the
// reference's executors also copy per call (src/cpu_executor.cpp:553-558).
// Buffers are host memory; the HIP library stages them through pinned
// memory on one of its execution slots.  Errors keep the reference's
// exception kinds: std::invalid_argument for bad arguments, std::runtime_error
// for device failures and non-finite Float64 results.
#pragma once

#include <dsmvc/engine.hpp>

#include <cstdint>
#include <memory>

namespace dsmvc::hip_detail {

[[nodiscard]] bool backend_available() noexcept;

class HipExecutor {
public:
    // solve_mode: DSMVC_HIP_SOLVE_* (0 automatic, 1 exact, ...); device -1 is
    // the current HIP device.
    explicit HipExecutor(int solve_mode = 0, int device = -1);
    ~HipExecutor();
    HipExecutor(const HipExecutor &) noexcept = default;
    HipExecutor &operator=(const HipExecutor &) noexcept = default;
    HipExecutor(HipExecutor &&) noexcept = default;
    HipExecutor &operator=(HipExecutor &&) noexcept = default;

    [[nodiscard]] const char *name() const noexcept { return "hip"; }
    void prepare(std::shared_ptr<const AxisPlan> plan) const;
    void seal() const;

    void inverse_rows(const AxisPlan &plan, const float *input, std::ptrdiff_t input_row_stride,
                      float *output, std::ptrdiff_t output_row_stride, std::int32_t row_count,
                      std::shared_ptr<const void> input_lifetime) const;
    void inverse_columns(const AxisPlan &plan, const float *input,
                         std::ptrdiff_t input_row_stride, float *output,
                         std::ptrdiff_t output_row_stride, std::int32_t column_count,
                         std::shared_ptr<const void> input_lifetime) const;
    void inverse_2d(const AxisPlan &horizontal, const AxisPlan &vertical, const float *input,
                    std::ptrdiff_t input_row_stride, float *output,
                    std::ptrdiff_t output_row_stride,
                    std::shared_ptr<const void> input_lifetime) const;
    void inverse_2d_u8(const AxisPlan &horizontal, const AxisPlan &vertical,
                       const std::uint8_t *input, std::ptrdiff_t input_row_stride,
                       std::uint8_t *output, std::ptrdiff_t output_row_stride,
                       const IntegerConversion &conversion,
                       std::shared_ptr<const void> input_lifetime) const;
    void inverse_2d_u16(const AxisPlan &horizontal, const AxisPlan &vertical,
                        const std::uint16_t *input, std::ptrdiff_t input_row_stride,
                        std::uint16_t *output, std::ptrdiff_t output_row_stride,
                        const IntegerConversion &conversion,
                        std::shared_ptr<const void> input_lifetime) const;

    // GPU frames exported by the core (the CUDA executor's *_exported
    // surface, src/cuda/cuda_executor.hpp; include/dsmvc/engine.hpp:414-442):
    // the planes' descriptors are imported by the HIP library, which consumes
    // them whatever the outcome.
    [[nodiscard]] bool has_exported_semaphore(std::uint64_t semaphore) const;
    [[nodiscard]] bool import_exported_semaphore(std::uint64_t semaphore, int handle_type,
                                                 std::intptr_t handle) const;
    void inverse_rows_exported(const AxisPlan &plan, const ExportedPlane &input,
                               const ExportedPlane &output, std::int32_t row_count) const;
    void inverse_columns_exported(const AxisPlan &plan, const ExportedPlane &input,
                                  const ExportedPlane &output, std::int32_t column_count) const;
    void inverse_2d_exported(const AxisPlan &horizontal, const AxisPlan &vertical,
                             const ExportedPlane &input, const ExportedPlane &output) const;
    void inverse_2d_u8_exported(const AxisPlan &horizontal, const AxisPlan &vertical,
                                const ExportedPlane &input, const ExportedPlane &output,
                                const IntegerConversion &conversion) const;
    void inverse_2d_u16_exported(const AxisPlan &horizontal, const AxisPlan &vertical,
                                 const ExportedPlane &input, const ExportedPlane &output,
                                 const IntegerConversion &conversion) const;

private:
    struct State;
    std::shared_ptr<State> state_;
};

} // namespace dsmvc::hip_detail

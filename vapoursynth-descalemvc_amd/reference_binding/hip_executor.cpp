// The HIP executor as a backend of the reference engine (see hip_executor.hpp).
#include "hip_executor.hpp"

#include "dsmvc_hip.h"

#include <mutex>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace dsmvc::hip_detail {
namespace {

// Status codes back to the reference's exception kinds.
void check(int status) {
    if (status == DSMVC_HIP_OK) return;
    const std::string message = dsmvc_hip_last_error();
    if (status == DSMVC_HIP_INVALID_ARGUMENT) throw std::invalid_argument(message);
    throw std::runtime_error(message);
}

// A HIP-library plan for one reference AxisPlan; releases it when dropped.
struct Imported {
    dsmvc_hip_plan *handle = nullptr;
    Imported() = default;
    explicit Imported(const AxisPlan &plan) {
        if (!plan.valid()) throw std::invalid_argument("dsmvc: invalid HIP axis plan");
        dsmvc_hip_plan_info info{};
        info.source_size = plan.source_size;
        info.destination_size = plan.destination_size;
        info.support = plan.support;
        info.half_bandwidth = plan.half_bandwidth;
        info.active_length = plan.active_length;
        info.shift = plan.shift;
        info.normal_rcond = plan.normal_rcond;
        info.normal_inf_norm = plan.normal_inf_norm;
        info.requires_float64 = plan.requires_float64() ? 1 : 0;
        info.nnz = static_cast<std::int32_t>(plan.transpose_indices.size());
        dsmvc_hip_plan_data data{};
        data.offsets = plan.transpose_offsets.data();
        data.indices = plan.transpose_indices.data();
        data.weights = plan.transpose_weights.data();
        data.lower_ld = plan.lower_ld.data();
        data.upper_l = plan.upper_l.data();
        data.inverse_diagonal = plan.inverse_diagonal.data();
        data.weights_f64 = plan.transpose_weights_f64.data();
        data.normal_bands_f64 = plan.normal_bands_f64.data();
        data.ldlt_bands_f64 = plan.ldlt_bands_f64.data();
        data.inverse_diagonal_f64 = plan.inverse_diagonal_f64.data();
        check(dsmvc_hip_plan_import(&info, &data, &handle));
    }
    Imported(const Imported &) = delete;
    Imported &operator=(const Imported &) = delete;
    Imported(Imported &&o) noexcept : handle(o.handle) { o.handle = nullptr; }
    ~Imported() {
        if (handle) dsmvc_hip_plan_release(handle);
    }
};

} // namespace

bool backend_available() noexcept {
    std::int32_t count = 0;
    return dsmvc_hip_device_count(&count) == DSMVC_HIP_OK && count > 0;
}

struct HipExecutor::State {
    dsmvc_hip_executor *executor = nullptr;
    std::mutex mutex;
    struct Prepared {
        std::shared_ptr<const AxisPlan> plan;   // keeps the address valid
        std::shared_ptr<Imported> imported;
    };
    std::unordered_map<const AxisPlan *, Prepared> prepared;
    bool sealed = false;

    ~State() {
        if (executor) dsmvc_hip_executor_destroy(executor);
    }

    // The imported plan of a prepared plan, or a per-call import of a
    // borrowed one.
    std::shared_ptr<Imported> plan(const AxisPlan &p) {
        {
            std::lock_guard lock(mutex);
            auto it = prepared.find(&p);
            if (it != prepared.end()) return it->second.imported;
        }
        return std::make_shared<Imported>(p);
    }
};

HipExecutor::HipExecutor(int solve_mode, int device) : state_(std::make_shared<State>()) {
    check(dsmvc_hip_executor_create(device, solve_mode, &state_->executor));
}

HipExecutor::~HipExecutor() = default;

void HipExecutor::prepare(std::shared_ptr<const AxisPlan> plan) const {
    if (!plan) throw std::invalid_argument("dsmvc: cannot prepare a null HIP axis plan");
    auto imported = std::make_shared<Imported>(*plan);
    check(dsmvc_hip_executor_prepare(state_->executor, imported->handle));
    std::lock_guard lock(state_->mutex);
    if (state_->sealed) throw std::logic_error("dsmvc: executor is sealed");
    const AxisPlan *key = plan.get();
    state_->prepared[key] = State::Prepared{std::move(plan), std::move(imported)};
}

void HipExecutor::seal() const {
    check(dsmvc_hip_executor_seal(state_->executor));
    std::lock_guard lock(state_->mutex);
    state_->sealed = true;
}

void HipExecutor::inverse_rows(const AxisPlan &plan, const float *input,
                               std::ptrdiff_t input_row_stride, float *output,
                               std::ptrdiff_t output_row_stride, std::int32_t row_count,
                               std::shared_ptr<const void>) const {
    const auto p = state_->plan(plan);
    check(dsmvc_hip_inverse_rows_host(state_->executor, p->handle, input, input_row_stride,
                                      output, output_row_stride, row_count));
}

void HipExecutor::inverse_columns(const AxisPlan &plan, const float *input,
                                  std::ptrdiff_t input_row_stride, float *output,
                                  std::ptrdiff_t output_row_stride, std::int32_t column_count,
                                  std::shared_ptr<const void>) const {
    const auto p = state_->plan(plan);
    check(dsmvc_hip_inverse_columns_host(state_->executor, p->handle, input, input_row_stride,
                                         output, output_row_stride, column_count));
}

void HipExecutor::inverse_2d(const AxisPlan &horizontal, const AxisPlan &vertical,
                             const float *input, std::ptrdiff_t input_row_stride, float *output,
                             std::ptrdiff_t output_row_stride,
                             std::shared_ptr<const void>) const {
    const auto h = state_->plan(horizontal), v = state_->plan(vertical);
    check(dsmvc_hip_inverse_2d_host(state_->executor, h->handle, v->handle, DSMVC_HIP_SAMPLE_F32,
                                    input, input_row_stride, 0, output, output_row_stride, 0, 1,
                                    nullptr));
}

namespace {
dsmvc_hip_integer_conversion to_c(const IntegerConversion &c) {
    return dsmvc_hip_integer_conversion{c.input_offset, c.input_scale, c.output_scale,
                                        c.output_offset, c.output_maximum};
}
} // namespace

void HipExecutor::inverse_2d_u8(const AxisPlan &horizontal, const AxisPlan &vertical,
                                const std::uint8_t *input, std::ptrdiff_t input_row_stride,
                                std::uint8_t *output, std::ptrdiff_t output_row_stride,
                                const IntegerConversion &conversion,
                                std::shared_ptr<const void>) const {
    const auto h = state_->plan(horizontal), v = state_->plan(vertical);
    const dsmvc_hip_integer_conversion c = to_c(conversion);
    check(dsmvc_hip_inverse_2d_host(state_->executor, h->handle, v->handle, DSMVC_HIP_SAMPLE_U8,
                                    input, input_row_stride, 0, output, output_row_stride, 0, 1,
                                    &c));
}

void HipExecutor::inverse_2d_u16(const AxisPlan &horizontal, const AxisPlan &vertical,
                                 const std::uint16_t *input, std::ptrdiff_t input_row_stride,
                                 std::uint16_t *output, std::ptrdiff_t output_row_stride,
                                 const IntegerConversion &conversion,
                                 std::shared_ptr<const void>) const {
    const auto h = state_->plan(horizontal), v = state_->plan(vertical);
    const dsmvc_hip_integer_conversion c = to_c(conversion);
    check(dsmvc_hip_inverse_2d_host(state_->executor, h->handle, v->handle, DSMVC_HIP_SAMPLE_U16,
                                    input, input_row_stride, 0, output, output_row_stride, 0, 1,
                                    &c));
}

namespace {
dsmvc_hip_exported_plane to_c(const ExportedPlane &p) {
    dsmvc_hip_exported_plane out{};
    out.memory_id = p.memory_id;
    out.memory_size = p.memory_size;
    out.offset = p.offset;
    out.plane_bytes = p.plane_bytes;
    out.handle_type = p.handle_type;
    out.handle = static_cast<std::int64_t>(p.handle);
    out.pitch_bytes = static_cast<std::int64_t>(p.pitch_bytes);
    out.ready_semaphore = p.ready_semaphore;
    out.ready_value = p.ready_value;
    return out;
}
} // namespace

bool HipExecutor::has_exported_semaphore(std::uint64_t semaphore) const {
    std::int32_t imported = 0;
    check(dsmvc_hip_has_exported_semaphore(state_->executor, semaphore, &imported));
    return imported != 0;
}

bool HipExecutor::import_exported_semaphore(std::uint64_t semaphore, int handle_type,
                                            std::intptr_t handle) const {
    std::int32_t imported = 0;
    check(dsmvc_hip_import_exported_semaphore(state_->executor, semaphore, handle_type,
                                              static_cast<std::int64_t>(handle), &imported));
    return imported != 0;
}

void HipExecutor::inverse_rows_exported(const AxisPlan &plan, const ExportedPlane &input,
                                        const ExportedPlane &output, std::int32_t row_count) const {
    const dsmvc_hip_exported_plane in = to_c(input), out = to_c(output);
    std::shared_ptr<Imported> p;
    try {
        p = state_->plan(plan);
    } catch (...) {
        dsmvc_hip_close_exported_plane(&in);
        dsmvc_hip_close_exported_plane(&out);
        throw;
    }
    check(dsmvc_hip_inverse_rows_exported(state_->executor, p->handle, &in, &out, row_count));
}

void HipExecutor::inverse_columns_exported(const AxisPlan &plan, const ExportedPlane &input,
                                           const ExportedPlane &output,
                                           std::int32_t column_count) const {
    const dsmvc_hip_exported_plane in = to_c(input), out = to_c(output);
    std::shared_ptr<Imported> p;
    try {
        p = state_->plan(plan);
    } catch (...) {
        dsmvc_hip_close_exported_plane(&in);
        dsmvc_hip_close_exported_plane(&out);
        throw;
    }
    check(dsmvc_hip_inverse_columns_exported(state_->executor, p->handle, &in, &out,
                                             column_count));
}

namespace {
template <class State, class Plan>
void exported_2d(State &state, const Plan &horizontal, const Plan &vertical, std::int32_t kind,
                 const ExportedPlane &input, const ExportedPlane &output,
                 const dsmvc_hip_integer_conversion *conversion) {
    const dsmvc_hip_exported_plane in = to_c(input), out = to_c(output);
    decltype(state.plan(horizontal)) h, v;
    try {
        h = state.plan(horizontal);
        v = state.plan(vertical);
    } catch (...) {
        dsmvc_hip_close_exported_plane(&in);
        dsmvc_hip_close_exported_plane(&out);
        throw;
    }
    check(dsmvc_hip_inverse_2d_exported(state.executor, h->handle, v->handle, kind, &in, &out,
                                        conversion));
}
} // namespace

void HipExecutor::inverse_2d_exported(const AxisPlan &horizontal, const AxisPlan &vertical,
                                      const ExportedPlane &input,
                                      const ExportedPlane &output) const {
    exported_2d(*state_, horizontal, vertical, DSMVC_HIP_SAMPLE_F32, input, output, nullptr);
}

void HipExecutor::inverse_2d_u8_exported(const AxisPlan &horizontal, const AxisPlan &vertical,
                                         const ExportedPlane &input, const ExportedPlane &output,
                                         const IntegerConversion &conversion) const {
    const dsmvc_hip_integer_conversion c = to_c(conversion);
    exported_2d(*state_, horizontal, vertical, DSMVC_HIP_SAMPLE_U8, input, output, &c);
}

void HipExecutor::inverse_2d_u16_exported(const AxisPlan &horizontal, const AxisPlan &vertical,
                                          const ExportedPlane &input, const ExportedPlane &output,
                                          const IntegerConversion &conversion) const {
    const dsmvc_hip_integer_conversion c = to_c(conversion);
    exported_2d(*state_, horizontal, vertical, DSMVC_HIP_SAMPLE_U16, input, output, &c);
}

} // namespace dsmvc::hip_detail

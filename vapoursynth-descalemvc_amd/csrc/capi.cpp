// The C ABI (include/dsmvc_hip.h) over the C++ planner and HIP executor.
// Exceptions never cross it: invalid_argument / out_of_range -> 1, other
// exceptions -> 2, missing device -> 3; the message is kept per thread.
#include "dsmvc_hip.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "batch.hpp"
#include "device_plan.hpp"
#include "operator_plan.hpp"
#include "executor.hpp"
#include "server.hpp"

struct dsmvc_hip_plan {
    std::shared_ptr<const dsmvc::AxisPlan> plan;
};

struct dsmvc_hip_executor {
    std::unique_ptr<dsmvc::HipExecutor> executor;
};

struct dsmvc_hip_batch {
    std::unique_ptr<dsmvc::HipBatch> batch;
};

struct dsmvc_hip_server {
    std::unique_ptr<dsmvc::HipServer> server;
};

namespace {

thread_local std::string g_error;

struct Unavailable : std::runtime_error {
    using std::runtime_error::runtime_error;
};

template <class F>
int guarded(F &&f) noexcept {
    try {
        f();
        g_error.clear();
        return DSMVC_HIP_OK;
    } catch (const Unavailable &e) {
        g_error = e.what();
        return DSMVC_HIP_UNAVAILABLE;
    } catch (const std::invalid_argument &e) {
        g_error = e.what();
        return DSMVC_HIP_INVALID_ARGUMENT;
    } catch (const std::out_of_range &e) {
        g_error = e.what();
        return DSMVC_HIP_INVALID_ARGUMENT;
    } catch (const std::exception &e) {
        g_error = e.what();
        return DSMVC_HIP_RUNTIME_ERROR;
    } catch (...) {
        g_error = "dsmvc: unknown error";
        return DSMVC_HIP_RUNTIME_ERROR;
    }
}

void require(bool ok, const char *what) {
    if (!ok) throw std::invalid_argument(what);
}

dsmvc::AxisRequest to_request(const dsmvc_hip_axis_request &r, bool custom = false) {
    dsmvc::AxisRequest q;
    q.source_size = r.source_size;
    q.destination_size = r.destination_size;
    q.active_length = r.active_length;
    q.shift = r.shift;
    if (r.kernel == DSMVC_HIP_CUSTOM && !custom)
        throw std::invalid_argument("custom kernel callback is missing");
    require(r.kernel >= 0 && r.kernel <= DSMVC_HIP_CUSTOM, "kernel is invalid");
    q.kernel.kind = static_cast<dsmvc::KernelKind>(r.kernel);
    q.kernel.taps = r.taps;
    q.kernel.b = r.b;
    q.kernel.c = r.c;
    q.kernel.blur = r.blur;
    require(r.border >= 0 && r.border <= 4, "padding mode is invalid");
    q.border = static_cast<dsmvc::BorderMode>(r.border);
    require(r.f64_mode >= 0 && r.f64_mode <= 2, "f64 mode is invalid");
    q.f64_mode = static_cast<dsmvc::F64Mode>(r.f64_mode);
    return q;
}

dsmvc::IntegerConversion to_conversion(const dsmvc_hip_integer_conversion &c) {
    dsmvc::IntegerConversion r;
    r.input_offset = c.input_offset;
    r.input_scale = c.input_scale;
    r.output_scale = c.output_scale;
    r.output_offset = c.output_offset;
    r.output_maximum = c.output_maximum;
    return r;
}

dsmvc::ExportedPlane to_exported(const dsmvc_hip_exported_plane &p) {
    dsmvc::ExportedPlane e;
    e.memory_id = p.memory_id;
    e.memory_size = p.memory_size;
    e.offset = p.offset;
    e.plane_bytes = p.plane_bytes;
    e.handle_type = p.handle_type;
    e.handle = static_cast<std::intptr_t>(p.handle);
    e.pitch_bytes = static_cast<std::ptrdiff_t>(p.pitch_bytes);
    e.ready_semaphore = p.ready_semaphore;
    e.ready_value = p.ready_value;
    return e;
}

// Runs an exported call; the planes' handles are consumed on every path,
// including arguments rejected before the executor sees them.
template <class F>
int exported_call(const dsmvc_hip_exported_plane *input, const dsmvc_hip_exported_plane *output,
                  F &&f) {
    return guarded([&] {
        if (!input || !output) {
            if (input) dsmvc::close_exported_plane(to_exported(*input));
            if (output) dsmvc::close_exported_plane(to_exported(*output));
            throw std::invalid_argument("dsmvc: null argument");
        }
        f(to_exported(*input), to_exported(*output));
    });
}

bool usable(const dsmvc_hip_executor *e) { return e && e->executor; }
bool usable(const dsmvc_hip_plan *p) { return p && p->plan; }

dsmvc::HipExecutor &exec_of(dsmvc_hip_executor *e) {
    require(e && e->executor, "dsmvc: executor is null");
    return *e->executor;
}

const dsmvc::AxisPlan &plan_of(const dsmvc_hip_plan *p) {
    require(p && p->plan, "dsmvc: plan is null");
    return *p->plan;
}

hipStream_t stream_of(void *s) { return static_cast<hipStream_t>(s); }

bool is_async(uint32_t flags) { return (flags & DSMVC_HIP_ASYNC) != 0U; }

} // namespace

extern "C" {

const char *dsmvc_hip_last_error(void) { return g_error.c_str(); }

int dsmvc_hip_plan_build(const dsmvc_hip_axis_request *request, dsmvc_hip_plan **plan) {
    return guarded([&] {
        require(request && plan, "dsmvc: null argument");
        auto handle = std::make_unique<dsmvc_hip_plan>();
        handle->plan = dsmvc::get_or_build_axis_plan(to_request(*request));
        *plan = handle.release();
    });
}

int dsmvc_hip_plan_build_many(const dsmvc_hip_axis_request *requests, int32_t count,
                              int32_t threads, dsmvc_hip_plan **plans) {
    return guarded([&] {
        require(count >= 0 && (count == 0 || (requests && plans)), "dsmvc: null argument");
        const auto n = static_cast<std::size_t>(count);
        for (std::size_t k = 0; k < n; ++k) plans[k] = nullptr;
        std::vector<dsmvc::AxisRequest> wanted;
        wanted.reserve(n);
        for (std::size_t k = 0; k < n; ++k) wanted.push_back(to_request(requests[k]));
        std::vector<std::shared_ptr<const dsmvc::AxisPlan>> built(n);
        std::vector<std::exception_ptr> failed(n);
        std::atomic<std::size_t> next{0};
        auto work = [&] {
            for (std::size_t k = next++; k < n; k = next++) {
                try {
                    built[k] = dsmvc::get_or_build_axis_plan(wanted[k]);
                } catch (...) {
                    failed[k] = std::current_exception();
                }
            }
        };
        const unsigned hardware = std::max(1U, std::thread::hardware_concurrency());
        const std::size_t workers = std::min<std::size_t>(
            n, threads > 0 ? static_cast<std::size_t>(threads) : std::min(hardware, 16U));
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
        for (std::size_t k = 0; k < n; ++k)
            if (failed[k]) std::rethrow_exception(failed[k]);
        std::vector<std::unique_ptr<dsmvc_hip_plan>> handles(n);
        for (std::size_t k = 0; k < n; ++k) {
            handles[k] = std::make_unique<dsmvc_hip_plan>();
            handles[k]->plan = std::move(built[k]);
        }
        for (std::size_t k = 0; k < n; ++k) plans[k] = handles[k].release();
    });
}

int dsmvc_hip_plan_build_custom(const dsmvc_hip_axis_request *request, dsmvc_hip_kernel_fn kernel,
                                void *user, dsmvc_hip_plan **plan) {
    return guarded([&] {
        require(request && plan, "dsmvc: null argument");
        require(request->kernel == DSMVC_HIP_CUSTOM, "kernel is invalid");
        if (!kernel) throw std::invalid_argument("custom kernel callback is missing");
        auto handle = std::make_unique<dsmvc_hip_plan>();
        handle->plan = dsmvc::get_or_build_axis_plan(
            to_request(*request, true), [kernel, user](double x) { return kernel(x, user); });
        *plan = handle.release();
    });
}

void dsmvc_hip_plan_release(dsmvc_hip_plan *plan) { delete plan; }

int dsmvc_hip_plan_import(const dsmvc_hip_plan_info *info, const dsmvc_hip_plan_data *data,
                          dsmvc_hip_plan **plan) {
    return guarded([&] {
        require(info && data && plan, "dsmvc: null argument");
        require(info->source_size > 0 && info->destination_size > 0 && info->nnz >= 0
                    && info->half_bandwidth >= 0 && info->half_bandwidth < info->destination_size,
                "dsmvc: imported axis plan is invalid");
        const auto n = static_cast<std::size_t>(info->destination_size);
        const auto nnz = static_cast<std::size_t>(info->nnz);
        const auto factors = static_cast<std::size_t>(info->half_bandwidth) * n;
        const auto bands = factors + n;
        const bool f64 = info->requires_float64 != 0;
        require(data->offsets && data->inverse_diagonal && (nnz == 0 || (data->indices && data->weights))
                    && (factors == 0 || (data->lower_ld && data->upper_l)),
                "dsmvc: imported axis plan is missing an array");
        require(!f64 || (data->normal_bands_f64 && data->ldlt_bands_f64 && data->inverse_diagonal_f64
                         && (nnz == 0 || data->weights_f64)),
                "dsmvc: imported axis plan is missing an FP64 array");
        dsmvc::AxisPlan p;
        p.source_size = info->source_size;
        p.destination_size = info->destination_size;
        p.support = info->support;
        p.half_bandwidth = info->half_bandwidth;
        p.active_length = info->active_length;
        p.shift = info->shift;
        p.normal_rcond = info->normal_rcond;
        p.normal_inf_norm = info->normal_inf_norm;
        p.transpose_offsets.assign(data->offsets, data->offsets + n + 1);
        if (nnz) {
            p.transpose_indices.assign(data->indices, data->indices + nnz);
            p.transpose_weights.assign(data->weights, data->weights + nnz);
        }
        if (factors) {
            p.lower_ld.assign(data->lower_ld, data->lower_ld + factors);
            p.upper_l.assign(data->upper_l, data->upper_l + factors);
        }
        p.inverse_diagonal.assign(data->inverse_diagonal, data->inverse_diagonal + n);
        if (f64) {
            if (nnz) p.transpose_weights_f64.assign(data->weights_f64, data->weights_f64 + nnz);
            p.normal_bands_f64.assign(data->normal_bands_f64, data->normal_bands_f64 + bands);
            p.ldlt_bands_f64.assign(data->ldlt_bands_f64, data->ldlt_bands_f64 + bands);
            p.inverse_diagonal_f64.assign(data->inverse_diagonal_f64, data->inverse_diagonal_f64 + n);
        }
        auto handle = std::make_unique<dsmvc_hip_plan>();
        handle->plan = dsmvc::import_axis_plan(std::move(p));
        *plan = handle.release();
    });
}

int dsmvc_hip_plan_info_get(const dsmvc_hip_plan *plan, dsmvc_hip_plan_info *info) {
    return guarded([&] {
        const auto &p = plan_of(plan);
        require(info, "dsmvc: null argument");
        info->source_size = p.source_size;
        info->destination_size = p.destination_size;
        info->support = p.support;
        info->half_bandwidth = p.half_bandwidth;
        info->active_length = p.active_length;
        info->shift = p.shift;
        info->normal_rcond = p.normal_rcond;
        info->normal_inf_norm = p.normal_inf_norm;
        info->requires_float64 = p.requires_float64() ? 1 : 0;
        info->nnz = static_cast<int32_t>(p.transpose_indices.size());
    });
}

int dsmvc_hip_plan_window_info(const dsmvc_hip_plan *plan, dsmvc_hip_window_info *info) {
    return guarded([&] {
        const auto &p = plan_of(plan);
        require(info, "dsmvc: null argument");
        dsmvc::hip::WindowPlan wp;
        const bool ok = dsmvc::hip::build_window_plan(p, wp);
        *info = dsmvc_hip_window_info{};
        info->feasible = ok ? 1 : 0;
        info->rows_feasible = ok && wp.rows_feasible ? 1 : 0;
        info->band = wp.band;
        info->chunks = wp.chunks;
        info->length = wp.length;
        info->warmup = wp.warmup;
        info->width = wp.width;
        info->residual = wp.residual;
        dsmvc::hip::WindowPlan64 wp64;
        info->float64_feasible = dsmvc::hip::build_window_plan64(p, dsmvc::hip::pack_plan(p), wp64) ? 1 : 0;
    });
}

int dsmvc_hip_plan_operator(const dsmvc_hip_plan *plan, dsmvc_hip_operator_info *info, int32_t *lo,
                            int32_t *count, double *values, int64_t capacity) {
    return guarded([&] {
        const auto &p = plan_of(plan);
        require(info, "dsmvc: null argument");
        dsmvc::hip::OperatorPlan op;
        dsmvc::hip::build_operator_plan(p, op);
        *info = dsmvc_hip_operator_info{};
        info->rows_feasible = op.h_feasible ? 1 : 0;
        info->columns_feasible = op.v_feasible ? 1 : 0;
        info->blocks = op.blocks;
        info->steps = op.steps;
        for (int c : op.count) info->max_band = std::max(info->max_band, c);
        info->values = static_cast<int64_t>(op.values.size());
        info->dropped = op.dropped;
        info->max_abs = op.max_abs;
        if (lo) std::memcpy(lo, op.lo.data(), op.lo.size() * sizeof(int32_t));
        if (count) std::memcpy(count, op.count.data(), op.count.size() * sizeof(int32_t));
        if (values) {
            require(capacity >= static_cast<int64_t>(op.values.size()), "dsmvc: operator buffer too small");
            std::memcpy(values, op.values.data(), op.values.size() * sizeof(double));
        }
    });
}

int dsmvc_hip_plan_arrays(const dsmvc_hip_plan *plan, uint32_t *offsets, int32_t *indices,
                          float *weights, float *lower_ld, float *upper_l,
                          float *inverse_diagonal, double *weights_f64,
                          double *normal_bands_f64, double *ldlt_bands_f64,
                          double *inverse_diagonal_f64) {
    return guarded([&] {
        const auto &p = plan_of(plan);
        auto copy = [](auto *dst, const auto &src) {
            if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
        };
        copy(offsets, p.transpose_offsets);
        copy(indices, p.transpose_indices);
        copy(weights, p.transpose_weights);
        copy(lower_ld, p.lower_ld);
        copy(upper_l, p.upper_l);
        copy(inverse_diagonal, p.inverse_diagonal);
        copy(weights_f64, p.transpose_weights_f64);
        copy(normal_bands_f64, p.normal_bands_f64);
        copy(ldlt_bands_f64, p.ldlt_bands_f64);
        copy(inverse_diagonal_f64, p.inverse_diagonal_f64);
    });
}

int dsmvc_hip_planner_cache_stats(dsmvc_hip_planner_stats *stats) {
    return guarded([&] {
        require(stats, "dsmvc: null argument");
        const auto s = dsmvc::planner_cache_stats();
        stats->plan_hits = s.plan_hits;
        stats->plan_builds = s.plan_builds;
        stats->plan_entries = s.plan_entries;
        stats->plan_resident_bytes = s.plan_resident_bytes;
        stats->geometry_hits = s.geometry_hits;
        stats->geometry_builds = s.geometry_builds;
        stats->geometry_entries = s.geometry_entries;
        stats->geometry_resident_bytes = s.geometry_resident_bytes;
    });
}

void dsmvc_hip_clear_planner_caches(void) { dsmvc::clear_planner_caches(); }

int dsmvc_hip_padding_index(int64_t index, int32_t size, int32_t mode, int32_t *mapped) {
    return guarded([&] {
        require(mapped, "dsmvc: null argument");
        require(mode >= 0 && mode <= 4, "padding mode is invalid");
        *mapped = dsmvc::padding_index(index, size, static_cast<dsmvc::BorderMode>(mode));
    });
}

int dsmvc_hip_parse_backend(const char *name, int32_t *kind) {
    return guarded([&] {
        require(name && kind, "dsmvc: null argument");
        *kind = static_cast<int32_t>(dsmvc::parse_backend(name));
    });
}

int dsmvc_hip_resolve_backend(int32_t requested, int32_t *resolved) {
    return guarded([&] {
        require(resolved && requested >= 0 && requested <= 5, "backend is invalid");
        const auto kind = static_cast<dsmvc::BackendKind>(requested);
        if (kind == dsmvc::BackendKind::hip && !dsmvc::hip_available())
            throw Unavailable("backend hip is unavailable");
        *resolved = static_cast<int32_t>(dsmvc::resolve_backend(kind));
    });
}

int dsmvc_hip_device_count(int32_t *count) {
    return guarded([&] {
        require(count, "dsmvc: null argument");
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
    });
}

int dsmvc_hip_executor_create(int32_t device, int32_t solve_mode, dsmvc_hip_executor **executor) {
    return guarded([&] {
        require(executor && solve_mode >= 0 && solve_mode <= 5, "dsmvc: invalid executor arguments");
        if (!dsmvc::hip_available()) throw Unavailable("dsmvc: backend hip is unavailable: no HIP device");
        auto handle = std::make_unique<dsmvc_hip_executor>();
        handle->executor = std::make_unique<dsmvc::HipExecutor>(
            device, static_cast<dsmvc::SolveMode>(solve_mode));
        *executor = handle.release();
    });
}

void dsmvc_hip_executor_destroy(dsmvc_hip_executor *executor) { delete executor; }

int dsmvc_hip_executor_prepare(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan) {
    return guarded([&] {
        require(plan && plan->plan, "dsmvc: plan is null");
        exec_of(executor).prepare(plan->plan);
    });
}

int dsmvc_hip_executor_seal(dsmvc_hip_executor *executor) {
    return guarded([&] { exec_of(executor).seal(); });
}

int dsmvc_hip_executor_set_refinement(dsmvc_hip_executor *executor, int32_t steps) {
    return guarded([&] { exec_of(executor).set_refinement(steps); });
}

int dsmvc_hip_executor_tuned_route(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                   int32_t rows, int32_t sample_kind, int32_t *route) {
    return guarded([&] {
        require(route && sample_kind >= 0 && sample_kind <= 2, "dsmvc: invalid route query");
        const int r = exec_of(executor).tuned_chunks(plan_of(plan), rows != 0, sample_kind);
        *route = r < -1 ? 0 : r;
    });
}

int dsmvc_hip_executor_tune(dsmvc_hip_executor *executor, int32_t rows_chunks,
                            int32_t columns_chunks) {
    return guarded([&] {
        require(rows_chunks >= -1 && rows_chunks <= 16 && columns_chunks >= -1 && columns_chunks <= 16,
                "dsmvc: routes must be -1 (windowed), 0 (automatic) or a chunk count up to 16");
        exec_of(executor).set_tuning({rows_chunks, columns_chunks});
    });
}

int dsmvc_hip_inverse_rows(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                           const float *input, ptrdiff_t input_row_stride, float *output,
                           ptrdiff_t output_row_stride, int32_t row_count, void *stream,
                           uint32_t flags) {
    return guarded([&] {
        exec_of(executor).inverse_rows(plan_of(plan), input, input_row_stride, output,
                                       output_row_stride, row_count, stream_of(stream),
                                       is_async(flags));
    });
}

int dsmvc_hip_inverse_columns(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                              const float *input, ptrdiff_t input_row_stride, float *output,
                              ptrdiff_t output_row_stride, int32_t column_count, void *stream,
                              uint32_t flags) {
    return guarded([&] {
        exec_of(executor).inverse_columns(plan_of(plan), input, input_row_stride, output,
                                          output_row_stride, column_count, stream_of(stream),
                                          is_async(flags));
    });
}

int dsmvc_hip_inverse_2d(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                         const dsmvc_hip_plan *vertical, const float *input,
                         ptrdiff_t input_row_stride, float *output, ptrdiff_t output_row_stride,
                         void *stream, uint32_t flags) {
    return guarded([&] {
        exec_of(executor).inverse_2d(plan_of(horizontal), plan_of(vertical), input,
                                     input_row_stride, output, output_row_stride,
                                     stream_of(stream), is_async(flags));
    });
}

int dsmvc_hip_inverse_2d_u8(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                            const dsmvc_hip_plan *vertical, const uint8_t *input,
                            ptrdiff_t input_row_stride, uint8_t *output,
                            ptrdiff_t output_row_stride,
                            const dsmvc_hip_integer_conversion *conversion, void *stream,
                            uint32_t flags) {
    return guarded([&] {
        require(conversion, "dsmvc: conversion is null");
        exec_of(executor).inverse_2d_u8(plan_of(horizontal), plan_of(vertical), input,
                                        input_row_stride, output, output_row_stride,
                                        to_conversion(*conversion), stream_of(stream),
                                        is_async(flags));
    });
}

int dsmvc_hip_inverse_2d_u16(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                             const dsmvc_hip_plan *vertical, const uint16_t *input,
                             ptrdiff_t input_row_stride, uint16_t *output,
                             ptrdiff_t output_row_stride,
                             const dsmvc_hip_integer_conversion *conversion, void *stream,
                             uint32_t flags) {
    return guarded([&] {
        require(conversion, "dsmvc: conversion is null");
        exec_of(executor).inverse_2d_u16(plan_of(horizontal), plan_of(vertical), input,
                                         input_row_stride, output, output_row_stride,
                                         to_conversion(*conversion), stream_of(stream),
                                         is_async(flags));
    });
}

int dsmvc_hip_inverse_2d_batch(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                               const dsmvc_hip_plan *vertical, int32_t sample_kind,
                               const void *input, ptrdiff_t input_row_stride,
                               ptrdiff_t input_frame_stride, void *output,
                               ptrdiff_t output_row_stride, ptrdiff_t output_frame_stride,
                               int32_t frames, const dsmvc_hip_integer_conversion *conversion,
                               void *stream, uint32_t flags) {
    return guarded([&] {
        require(sample_kind >= 0 && sample_kind <= 2, "dsmvc: sample kind is invalid");
        dsmvc::IntegerConversion conv;
        if (conversion) conv = to_conversion(*conversion);
        const dsmvc::PlaneBatch batch{input, input_row_stride, input_frame_stride, output,
                                      output_row_stride, output_frame_stride, frames};
        exec_of(executor).inverse_2d_batch(plan_of(horizontal), plan_of(vertical),
                                           static_cast<dsmvc::SampleKind>(sample_kind), batch,
                                           conversion ? &conv : nullptr, stream_of(stream),
                                           is_async(flags));
    });
}

int dsmvc_hip_inverse_2d_batch_noalloc(dsmvc_hip_executor *executor,
                                       const dsmvc_hip_plan *horizontal,
                                       const dsmvc_hip_plan *vertical, int32_t sample_kind,
                                       const void *input, ptrdiff_t input_row_stride,
                                       ptrdiff_t input_frame_stride, void *output,
                                       ptrdiff_t output_row_stride,
                                       ptrdiff_t output_frame_stride, int32_t frames,
                                       const dsmvc_hip_integer_conversion *conversion,
                                       void *stream, uint32_t flags) {
    return guarded([&] {
        require(sample_kind >= 0 && sample_kind <= 2, "dsmvc: sample kind is invalid");
        dsmvc::IntegerConversion conv;
        if (conversion) conv = to_conversion(*conversion);
        const dsmvc::PlaneBatch batch{input, input_row_stride, input_frame_stride, output,
                                      output_row_stride, output_frame_stride, frames};
        exec_of(executor).inverse_2d_batch_noalloc(
            plan_of(horizontal), plan_of(vertical), static_cast<dsmvc::SampleKind>(sample_kind),
            batch, conversion ? &conv : nullptr, stream_of(stream), is_async(flags));
    });
}

int dsmvc_hip_inverse_rows_host(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                const float *input, ptrdiff_t input_row_stride, float *output,
                                ptrdiff_t output_row_stride, int32_t row_count) {
    return guarded([&] {
        exec_of(executor).inverse_rows_host(plan_of(plan), input, input_row_stride, output,
                                            output_row_stride, row_count);
    });
}

int dsmvc_hip_inverse_columns_host(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                   const float *input, ptrdiff_t input_row_stride,
                                   float *output, ptrdiff_t output_row_stride,
                                   int32_t column_count) {
    return guarded([&] {
        exec_of(executor).inverse_columns_host(plan_of(plan), input, input_row_stride, output,
                                               output_row_stride, column_count);
    });
}

int dsmvc_hip_inverse_2d_host(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                              const dsmvc_hip_plan *vertical, int32_t sample_kind,
                              const void *input, ptrdiff_t input_row_stride,
                              ptrdiff_t input_frame_stride, void *output,
                              ptrdiff_t output_row_stride, ptrdiff_t output_frame_stride,
                              int32_t frames, const dsmvc_hip_integer_conversion *conversion) {
    return guarded([&] {
        require(sample_kind >= 0 && sample_kind <= 2, "dsmvc: sample kind is invalid");
        dsmvc::IntegerConversion conv;
        if (conversion) conv = to_conversion(*conversion);
        const dsmvc::PlaneBatch batch{input, input_row_stride, input_frame_stride, output,
                                      output_row_stride, output_frame_stride, frames};
        exec_of(executor).inverse_2d_host(plan_of(horizontal), plan_of(vertical),
                                          static_cast<dsmvc::SampleKind>(sample_kind), batch,
                                          conversion ? &conv : nullptr);
    });
}

int dsmvc_hip_inverse_2d_exported(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                                  const dsmvc_hip_plan *vertical, int32_t sample_kind,
                                  const dsmvc_hip_exported_plane *input,
                                  const dsmvc_hip_exported_plane *output,
                                  const dsmvc_hip_integer_conversion *conversion) {
    return exported_call(input, output, [&](const dsmvc::ExportedPlane &in,
                                            const dsmvc::ExportedPlane &out) {
        if (!usable(executor) || !usable(horizontal) || !usable(vertical) || sample_kind < 0
            || sample_kind > 2) {
            dsmvc::close_exported_plane(in);
            dsmvc::close_exported_plane(out);
            require(sample_kind >= 0 && sample_kind <= 2, "dsmvc: sample kind is invalid");
            throw std::invalid_argument("dsmvc: null argument");
        }
        dsmvc::IntegerConversion conv;
        if (conversion) conv = to_conversion(*conversion);
        exec_of(executor).inverse_2d_exported(plan_of(horizontal), plan_of(vertical),
                                              static_cast<dsmvc::SampleKind>(sample_kind), in, out,
                                              conversion ? &conv : nullptr);
    });
}

int dsmvc_hip_inverse_rows_exported(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                    const dsmvc_hip_exported_plane *input,
                                    const dsmvc_hip_exported_plane *output, int32_t row_count) {
    return exported_call(input, output, [&](const dsmvc::ExportedPlane &in,
                                            const dsmvc::ExportedPlane &out) {
        if (!usable(executor) || !usable(plan)) {
            dsmvc::close_exported_plane(in);
            dsmvc::close_exported_plane(out);
            throw std::invalid_argument("dsmvc: null argument");
        }
        exec_of(executor).inverse_rows_exported(plan_of(plan), in, out, row_count);
    });
}

int dsmvc_hip_inverse_columns_exported(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                       const dsmvc_hip_exported_plane *input,
                                       const dsmvc_hip_exported_plane *output,
                                       int32_t column_count) {
    return exported_call(input, output, [&](const dsmvc::ExportedPlane &in,
                                            const dsmvc::ExportedPlane &out) {
        if (!usable(executor) || !usable(plan)) {
            dsmvc::close_exported_plane(in);
            dsmvc::close_exported_plane(out);
            throw std::invalid_argument("dsmvc: null argument");
        }
        exec_of(executor).inverse_columns_exported(plan_of(plan), in, out, column_count);
    });
}

int dsmvc_hip_import_exported_semaphore(dsmvc_hip_executor *executor, uint64_t semaphore,
                                        int32_t handle_type, int64_t handle, int32_t *imported) {
    return guarded([&] {
        if (!usable(executor) || !imported) {
            dsmvc::ExportedPlane p;
            p.handle_type = handle_type;
            p.handle = static_cast<std::intptr_t>(handle);
            dsmvc::close_exported_plane(p);
            throw std::invalid_argument("dsmvc: null argument");
        }
        *imported = exec_of(executor).import_exported_semaphore(
                        semaphore, handle_type, static_cast<std::intptr_t>(handle))
                        ? 1 : 0;
    });
}

int dsmvc_hip_has_exported_semaphore(dsmvc_hip_executor *executor, uint64_t semaphore,
                                     int32_t *imported) {
    return guarded([&] {
        require(imported, "dsmvc: null argument");
        *imported = exec_of(executor).has_exported_semaphore(semaphore) ? 1 : 0;
    });
}

void dsmvc_hip_close_exported_plane(const dsmvc_hip_exported_plane *plane) {
    if (plane) dsmvc::close_exported_plane(to_exported(*plane));
}

int dsmvc_hip_executor_import_stats(dsmvc_hip_executor *executor, dsmvc_hip_import_stats *stats) {
    return guarded([&] {
        require(stats, "dsmvc: null argument");
        const auto s = exec_of(executor).import_stats();
        stats->entries = s.entries;
        stats->imports = s.imports;
        stats->hits = s.hits;
        stats->evictions = s.evictions;
    });
}

int dsmvc_hip_executor_set_plan_cache(dsmvc_hip_executor *executor, int64_t limit_bytes) {
    return guarded([&] {
        require(limit_bytes > 0, "dsmvc: the plan cache limit must be positive");
        exec_of(executor).set_plan_cache_limit(static_cast<std::size_t>(limit_bytes));
    });
}

int dsmvc_hip_executor_plan_cache_stats(dsmvc_hip_executor *executor,
                                        dsmvc_hip_plan_cache_stats *stats) {
    return guarded([&] {
        require(stats, "dsmvc: null argument");
        const auto s = exec_of(executor).plan_cache_stats();
        stats->entries = s.entries;
        stats->bytes = s.bytes;
        stats->limit = s.limit;
        stats->hits = s.hits;
        stats->misses = s.misses;
        stats->evictions = s.evictions;
    });
}

int dsmvc_hip_inverse_rows_batch(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                 int32_t input_kind, const void *input,
                                 ptrdiff_t input_row_stride, ptrdiff_t input_frame_stride,
                                 float *output, ptrdiff_t output_row_stride,
                                 ptrdiff_t output_frame_stride, int32_t rows, int32_t frames,
                                 const dsmvc_hip_integer_conversion *conversion, void *stream,
                                 uint32_t flags) {
    return guarded([&] {
        require(input_kind >= 0 && input_kind <= 2, "dsmvc: sample kind is invalid");
        dsmvc::IntegerConversion conv;
        if (conversion) conv = to_conversion(*conversion);
        const dsmvc::PlaneBatch batch{input, input_row_stride, input_frame_stride, output,
                                      output_row_stride, output_frame_stride, frames};
        exec_of(executor).inverse_rows_batch(plan_of(plan),
                                             static_cast<dsmvc::SampleKind>(input_kind), batch,
                                             rows, conversion ? &conv : nullptr,
                                             stream_of(stream), is_async(flags));
    });
}

int dsmvc_hip_inverse_columns_batch(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                    const float *input, ptrdiff_t input_row_stride,
                                    ptrdiff_t input_frame_stride, int32_t output_kind,
                                    void *output, ptrdiff_t output_row_stride,
                                    ptrdiff_t output_frame_stride, int32_t columns,
                                    int32_t frames, const dsmvc_hip_integer_conversion *conversion,
                                    void *stream, uint32_t flags) {
    return guarded([&] {
        require(output_kind >= 0 && output_kind <= 2, "dsmvc: sample kind is invalid");
        dsmvc::IntegerConversion conv;
        if (conversion) conv = to_conversion(*conversion);
        const dsmvc::PlaneBatch batch{input, input_row_stride, input_frame_stride, output,
                                      output_row_stride, output_frame_stride, frames};
        exec_of(executor).inverse_columns_batch(plan_of(plan),
                                                static_cast<dsmvc::SampleKind>(output_kind),
                                                batch, columns, conversion ? &conv : nullptr,
                                                stream_of(stream), is_async(flags));
    });
}

int dsmvc_hip_inverse_2d_planes(dsmvc_hip_executor *executor, const dsmvc_hip_plane_job *planes,
                                int32_t count, int32_t sample_kind, void *stream, uint32_t flags) {
    return guarded([&] {
        require((planes || count == 0) && count >= 0 && sample_kind >= 0 && sample_kind <= 2,
                "dsmvc: invalid plane arguments");
        std::vector<dsmvc::IntegerConversion> conv(static_cast<std::size_t>(count));
        std::vector<dsmvc::HipExecutor::PlaneJob> list(static_cast<std::size_t>(count));
        for (int32_t k = 0; k < count; ++k) {
            const dsmvc_hip_plane_job &p = planes[k];
            require(p.horizontal && p.vertical, "dsmvc: invalid plane arguments");
            auto &j = list[static_cast<std::size_t>(k)];
            j.horizontal = &plan_of(p.horizontal);
            j.vertical = &plan_of(p.vertical);
            j.batch = dsmvc::PlaneBatch{p.input, p.input_row_stride, p.input_frame_stride, p.output,
                                        p.output_row_stride, p.output_frame_stride, p.frames};
            if (p.conversion) {
                conv[static_cast<std::size_t>(k)] = to_conversion(*p.conversion);
                j.conversion = &conv[static_cast<std::size_t>(k)];
            }
        }
        exec_of(executor).inverse_2d_planes(list.data(), count,
                                            static_cast<dsmvc::SampleKind>(sample_kind),
                                            static_cast<hipStream_t>(stream),
                                            (flags & DSMVC_HIP_ASYNC) != 0);
    });
}

int dsmvc_hip_batch_create(dsmvc_hip_executor *executor, const dsmvc_hip_batch_job *jobs,
                           int32_t count, uint32_t flags, dsmvc_hip_batch **batch) {
    return guarded([&] {
        require(batch && (jobs || count == 0) && count >= 0, "dsmvc: invalid batch arguments");
        require((flags & ~DSMVC_HIP_BATCH_GRAPH) == 0, "dsmvc: unknown batch flags");
        auto &ex = exec_of(executor);
        std::vector<dsmvc::BatchJob> list;
        list.reserve(static_cast<std::size_t>(count));
        for (int32_t k = 0; k < count; ++k) {
            const dsmvc_hip_batch_job &j = jobs[k];
            require(j.horizontal && j.vertical, "dsmvc: invalid batch job");
            dsmvc::BatchJob b;
            b.horizontal = j.horizontal->plan;
            b.vertical = j.vertical->plan;
            b.input = j.input;
            b.input_row_stride = j.input_row_stride;
            b.input_frame_stride = j.input_frame_stride;
            b.output = j.output;
            b.output_row_stride = j.output_row_stride;
            b.output_frame_stride = j.output_frame_stride;
            b.frames = j.frames;
            list.push_back(std::move(b));
        }
        auto handle = std::make_unique<dsmvc_hip_batch>();
        handle->batch = std::make_unique<dsmvc::HipBatch>(ex, std::move(list),
                                                          (flags & DSMVC_HIP_BATCH_GRAPH) != 0);
        *batch = handle.release();
    });
}

int dsmvc_hip_batch_run(dsmvc_hip_batch *batch, void *stream, uint32_t flags) {
    return guarded([&] {
        require(batch && batch->batch, "dsmvc: null batch");
        batch->batch->run(static_cast<hipStream_t>(stream), (flags & DSMVC_HIP_ASYNC) != 0);
    });
}

int dsmvc_hip_batch_info_get(const dsmvc_hip_batch *batch, dsmvc_hip_batch_info *info) {
    return guarded([&] {
        require(batch && batch->batch && info, "dsmvc: null argument");
        const dsmvc::BatchStats &s = batch->batch->stats();
        *info = dsmvc_hip_batch_info{};
        info->jobs = s.jobs;
        info->table_jobs = s.table_jobs;
        info->fallback_jobs = s.fallback_jobs;
        info->launches = s.launches;
        info->items = s.items;
        info->intermediate_bytes = s.intermediate_bytes;
        info->graph = s.graph ? 1 : 0;
        info->fused_jobs = s.fused_jobs;
    });
}

void dsmvc_hip_batch_destroy(dsmvc_hip_batch *batch) { delete batch; }

int dsmvc_hip_server_create(dsmvc_hip_executor *executor, const dsmvc_hip_server_config *config,
                            dsmvc_hip_server **server) {
    return guarded([&] {
        require(config && server && config->memory_budget > 0, "dsmvc: invalid server arguments");
        dsmvc::ServerConfig c;
        c.memory_budget = static_cast<std::size_t>(config->memory_budget);
        c.slots = config->slots;
        c.max_round_requests = config->max_round_requests;
        require(config->fused >= DSMVC_HIP_SERVE_FUSED_WARM && config->fused <= DSMVC_HIP_SERVE_FUSED_OFF,
                "dsmvc: invalid server arguments");
        c.fused = config->fused == DSMVC_HIP_SERVE_FUSED_ALWAYS ? dsmvc::ServeFused::always
                  : config->fused == DSMVC_HIP_SERVE_FUSED_OFF ? dsmvc::ServeFused::off
                                                               : dsmvc::ServeFused::warm;
        auto handle = std::make_unique<dsmvc_hip_server>();
        handle->server = std::make_unique<dsmvc::HipServer>(exec_of(executor), c);
        *server = handle.release();
    });
}

int dsmvc_hip_server_submit(dsmvc_hip_server *server, const dsmvc_hip_serve_request *request,
                            uint64_t *id) {
    return guarded([&] {
        require(server && server->server && request && id, "dsmvc: null argument");
        const dsmvc_hip_serve_request &r = *request;
        require(r.kernel >= 0 && r.kernel <= 5, "kernel is invalid");
        require(r.border >= 0 && r.border <= 4, "padding mode is invalid");
        require(r.f64_mode >= 0 && r.f64_mode <= 2, "f64 mode is invalid");
        dsmvc::ServeRequest q;
        q.planes = r.planes;
        q.subsampling_w = r.subsampling_w;
        q.subsampling_h = r.subsampling_h;
        q.src_width = r.src_width;
        q.src_height = r.src_height;
        q.dst_width = r.dst_width;
        q.dst_height = r.dst_height;
        q.kernel.kind = static_cast<dsmvc::KernelKind>(r.kernel);
        q.kernel.taps = r.taps;
        q.kernel.b = r.b;
        q.kernel.c = r.c;
        q.kernel.blur = r.blur;
        q.border = static_cast<dsmvc::BorderMode>(r.border);
        q.f64_mode = static_cast<dsmvc::F64Mode>(r.f64_mode);
        for (int p = 0; p < 3; ++p) {
            q.src[p] = r.src[p];
            q.src_stride[p] = r.src_stride[p];
            q.dst[p] = r.dst[p];
            q.dst_stride[p] = r.dst_stride[p];
        }
        *id = server->server->submit(q);
    });
}

int dsmvc_hip_server_wait(dsmvc_hip_server *server, uint64_t id, double *submitted,
                          double *completed) {
    std::string failure;
    dsmvc::RequestState state = dsmvc::RequestState::pending;
    const int status = guarded([&] {
        require(server && server->server, "dsmvc: null server");
        state = server->server->wait(id, &failure, submitted, completed);
    });
    if (status != DSMVC_HIP_OK) return status;
    if (state == dsmvc::RequestState::failed) {
        g_error = failure;
        return DSMVC_HIP_RUNTIME_ERROR;
    }
    return DSMVC_HIP_OK;
}

int dsmvc_hip_server_poll(dsmvc_hip_server *server, uint64_t id, int32_t *state) {
    return guarded([&] {
        require(server && server->server && state, "dsmvc: null argument");
        *state = static_cast<int32_t>(server->server->poll(id));
    });
}

int dsmvc_hip_server_release(dsmvc_hip_server *server, uint64_t id) {
    return guarded([&] {
        require(server && server->server, "dsmvc: null server");
        server->server->release(id);
    });
}

int dsmvc_hip_server_stats_get(const dsmvc_hip_server *server, dsmvc_hip_server_stats *stats) {
    return guarded([&] {
        require(server && server->server && stats, "dsmvc: null argument");
        const dsmvc::ServerStats s = server->server->stats();
        stats->submitted = s.submitted;
        stats->completed = s.completed;
        stats->failed = s.failed;
        stats->rounds = s.rounds;
        stats->table_planes = s.table_planes;
        stats->fallback_planes = s.fallback_planes;
        stats->largest_round = s.largest_round;
        stats->memory_high_water = static_cast<int64_t>(s.memory_high_water);
        stats->fused_planes = s.fused_planes;
    });
}

void dsmvc_hip_server_destroy(dsmvc_hip_server *server) { delete server; }

int dsmvc_hip_executor_release_stream(dsmvc_hip_executor *executor, void *stream) {
    return guarded([&] { exec_of(executor).release_stream(static_cast<hipStream_t>(stream)); });
}

int dsmvc_hip_synchronize(dsmvc_hip_executor *executor, void *stream) {
    return guarded([&] { exec_of(executor).synchronize(stream_of(stream)); });
}

} // extern "C"

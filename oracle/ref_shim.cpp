// Test infrastructure only: a C shim over the reference descale engine, built
// from the reference's own sources by oracle/Makefile into oracle/_ref/.  It is
// used by tests (to pin the C restatement in oracle/dsmvc_oracle.c at the
// real workload sizes) and by bench.py's cpu_baseline leg.  Nothing in the
// product links or loads it.
//
// Interfaces wrapped (reference paths):
//   dsmvc::build_axis_plan                    include/dsmvc/engine.hpp:175
//   dsmvc::inverse_axis_f32                   src/axis_plan.cpp:1409
//   dsmvc::detail::inverse_axis_f32_ordered   src/axis_plan.cpp:1068
//   dsmvc::detail::inverse_axis_f64_ordered   src/axis_plan.cpp:1124
//   dsmvc::CpuExecutor::inverse_2d / _u8 / _u16 / _streamed
//                                             src/cpu_executor.cpp:1941-2398
#include <dsmvc/engine.hpp>

#include "axis_plan_internal.hpp"

#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>

namespace {

thread_local std::string last_error;

struct RefPlan {
    dsmvc::AxisPlan plan;
};

struct RefExecutor {
    dsmvc::CpuExecutor executor;
    explicit RefExecutor(dsmvc::CpuPath path) : executor(path) {}
};

template <class F>
int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::invalid_argument &e) {
        last_error = e.what();
        return 1;
    } catch (const std::exception &e) {
        last_error = e.what();
        return 2;
    }
}

} // namespace

extern "C" {

struct ref_request {
    int32_t source_size;
    int32_t destination_size;
    double active_length;
    double shift;
    int32_t kernel;   // dsmvc::KernelKind
    int32_t taps;
    double b;
    double c;
    double blur;
    int32_t border;   // dsmvc::BorderMode
    int32_t f64_mode; // dsmvc::F64Mode
};

struct ref_plan_info {
    int32_t source_size;
    int32_t destination_size;
    int32_t support;
    int32_t half_bandwidth;
    double active_length;
    double shift;
    double normal_rcond;
    double normal_inf_norm;
    int32_t requires_float64;
    int32_t nnz;
};

const char *ref_last_error(void) { return last_error.c_str(); }

int ref_plan_build(const ref_request *r, void **out) {
    return guarded([&] {
        dsmvc::AxisRequest request;
        request.source_size = r->source_size;
        request.destination_size = r->destination_size;
        request.active_length = r->active_length;
        request.shift = r->shift;
        request.kernel.kind = static_cast<dsmvc::KernelKind>(r->kernel);
        request.kernel.taps = r->taps;
        request.kernel.b = r->b;
        request.kernel.c = r->c;
        request.kernel.blur = r->blur;
        request.border = static_cast<dsmvc::BorderMode>(r->border);
        request.f64_mode = static_cast<dsmvc::F64Mode>(r->f64_mode);
        auto holder = std::make_unique<RefPlan>();
        holder->plan = dsmvc::build_axis_plan(request);
        *out = holder.release();
    });
}

// dsmvc::build_axis_plan(request, custom_kernel) (include/dsmvc/engine.hpp:175)
// with the callback called at each tap's |distance|.
int ref_plan_build_custom(const ref_request *r, double (*kernel)(double, void *), void *user,
                          void **out) {
    return guarded([&] {
        dsmvc::AxisRequest request;
        request.source_size = r->source_size;
        request.destination_size = r->destination_size;
        request.active_length = r->active_length;
        request.shift = r->shift;
        request.kernel.kind = dsmvc::KernelKind::custom;
        request.kernel.taps = r->taps;
        request.kernel.b = r->b;
        request.kernel.c = r->c;
        request.kernel.blur = r->blur;
        request.border = static_cast<dsmvc::BorderMode>(r->border);
        request.f64_mode = static_cast<dsmvc::F64Mode>(r->f64_mode);
        auto holder = std::make_unique<RefPlan>();
        holder->plan = dsmvc::build_axis_plan(request, [kernel, user](double x) {
            return kernel(x, user);
        });
        *out = holder.release();
    });
}

void ref_plan_free(void *p) { delete static_cast<RefPlan *>(p); }

void ref_plan_info_get(const void *p, ref_plan_info *info) {
    const auto &plan = static_cast<const RefPlan *>(p)->plan;
    info->source_size = plan.source_size;
    info->destination_size = plan.destination_size;
    info->support = plan.support;
    info->half_bandwidth = plan.half_bandwidth;
    info->active_length = plan.active_length;
    info->shift = plan.shift;
    info->normal_rcond = plan.normal_rcond;
    info->normal_inf_norm = plan.normal_inf_norm;
    info->requires_float64 = plan.requires_float64() ? 1 : 0;
    info->nnz = static_cast<int32_t>(plan.transpose_indices.size());
}

// Copies every plan array; F64 arrays only when the plan retains them.
void ref_plan_arrays(const void *p, uint32_t *offsets, int32_t *indices,
                     float *weights, float *lower_ld, float *upper_l,
                     float *inverse_diagonal, double *weights_f64,
                     double *normal_bands_f64, double *ldlt_bands_f64,
                     double *inverse_diagonal_f64) {
    const auto &plan = static_cast<const RefPlan *>(p)->plan;
    const auto copy = [](auto *dst, const auto &src) {
        if (dst && !src.empty()) std::memcpy(dst, src.data(), src.size() * sizeof(src[0]));
    };
    copy(offsets, plan.transpose_offsets);
    copy(indices, plan.transpose_indices);
    copy(weights, plan.transpose_weights);
    copy(lower_ld, plan.lower_ld);
    copy(upper_l, plan.upper_l);
    copy(inverse_diagonal, plan.inverse_diagonal);
    copy(weights_f64, plan.transpose_weights_f64);
    copy(normal_bands_f64, plan.normal_bands_f64);
    copy(ldlt_bands_f64, plan.ldlt_bands_f64);
    copy(inverse_diagonal_f64, plan.inverse_diagonal_f64);
}

int ref_inverse_axis_f32(const void *p, const float *in, ptrdiff_t in_stride,
                         float *out, ptrdiff_t out_stride) {
    return guarded([&] {
        dsmvc::inverse_axis_f32(static_cast<const RefPlan *>(p)->plan, in,
                                in_stride, out, out_stride);
    });
}

void ref_inverse_axis_f32_ordered(const void *p, const float *in,
                                  ptrdiff_t in_stride, float *out,
                                  ptrdiff_t out_stride) {
    dsmvc::detail::inverse_axis_f32_ordered(
        static_cast<const RefPlan *>(p)->plan, in, in_stride, out, out_stride);
}

void ref_inverse_axis_f64_ordered(const void *p, const double *in,
                                  ptrdiff_t in_stride, double *out,
                                  ptrdiff_t out_stride) {
    dsmvc::detail::inverse_axis_f64_ordered(
        static_cast<const RefPlan *>(p)->plan, in, in_stride, out, out_stride);
}

void ref_inverse_axis_f64(const void *p, const double *in, ptrdiff_t in_stride,
                          double *out, ptrdiff_t out_stride) {
    dsmvc::detail::inverse_axis_f64(static_cast<const RefPlan *>(p)->plan, in,
                                    in_stride, out, out_stride);
}

// path: dsmvc::CpuPath (0 automatic, 1 scalar, 2 avx2, 4 avx512)
int ref_executor_create(int32_t path, void **out) {
    return guarded([&] {
        *out = new RefExecutor(static_cast<dsmvc::CpuPath>(path));
    });
}

void ref_executor_free(void *e) { delete static_cast<RefExecutor *>(e); }

const char *ref_executor_name(const void *e) {
    return static_cast<const RefExecutor *>(e)->executor.name();
}

int ref_inverse_rows(const void *e, const void *p, const float *in,
                     ptrdiff_t in_stride, float *out, ptrdiff_t out_stride,
                     int32_t rows) {
    return guarded([&] {
        static_cast<const RefExecutor *>(e)->executor.inverse_rows(
            static_cast<const RefPlan *>(p)->plan, in, in_stride, out,
            out_stride, rows);
    });
}

int ref_inverse_columns(const void *e, const void *p, const float *in,
                        ptrdiff_t in_stride, float *out, ptrdiff_t out_stride,
                        int32_t columns) {
    return guarded([&] {
        static_cast<const RefExecutor *>(e)->executor.inverse_columns(
            static_cast<const RefPlan *>(p)->plan, in, in_stride, out,
            out_stride, columns);
    });
}

int ref_inverse_2d(const void *e, const void *h, const void *v,
                   const float *in, ptrdiff_t in_stride, float *out,
                   ptrdiff_t out_stride) {
    return guarded([&] {
        static_cast<const RefExecutor *>(e)->executor.inverse_2d(
            static_cast<const RefPlan *>(h)->plan,
            static_cast<const RefPlan *>(v)->plan, in, in_stride, out,
            out_stride);
    });
}

struct ref_conversion {
    float input_offset;
    float input_scale;
    float output_scale;
    float output_offset;
    uint32_t output_maximum;
};

static dsmvc::IntegerConversion to_conversion(const ref_conversion *c) {
    dsmvc::IntegerConversion r;
    r.input_offset = c->input_offset;
    r.input_scale = c->input_scale;
    r.output_scale = c->output_scale;
    r.output_offset = c->output_offset;
    r.output_maximum = c->output_maximum;
    return r;
}

int ref_inverse_2d_u16(const void *e, const void *h, const void *v,
                       const uint16_t *in, ptrdiff_t in_stride, uint16_t *out,
                       ptrdiff_t out_stride, const ref_conversion *c,
                       int32_t streamed) {
    return guarded([&] {
        const auto &ex = static_cast<const RefExecutor *>(e)->executor;
        const auto &hp = static_cast<const RefPlan *>(h)->plan;
        const auto &vp = static_cast<const RefPlan *>(v)->plan;
        if (streamed) {
            ex.inverse_2d_u16_streamed(hp, vp, in, in_stride, out, out_stride,
                                       to_conversion(c));
        } else {
            ex.inverse_2d_u16(hp, vp, in, in_stride, out, out_stride,
                              to_conversion(c));
        }
    });
}

int ref_inverse_2d_u8(const void *e, const void *h, const void *v,
                      const uint8_t *in, ptrdiff_t in_stride, uint8_t *out,
                      ptrdiff_t out_stride, const ref_conversion *c,
                      int32_t streamed) {
    return guarded([&] {
        const auto &ex = static_cast<const RefExecutor *>(e)->executor;
        const auto &hp = static_cast<const RefPlan *>(h)->plan;
        const auto &vp = static_cast<const RefPlan *>(v)->plan;
        if (streamed) {
            ex.inverse_2d_u8_streamed(hp, vp, in, in_stride, out, out_stride,
                                      to_conversion(c));
        } else {
            ex.inverse_2d_u8(hp, vp, in, in_stride, out, out_stride,
                             to_conversion(c));
        }
    });
}

} // extern "C"

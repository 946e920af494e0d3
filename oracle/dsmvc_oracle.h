/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.
 *
 * A plain-C restatement of the reference's CPU descale path: the FP64 axis
 * planner, the scalar F32 axis solver (production and ordered-FMA forms), the
 * Float64 route, and the scalar 2D / integer drivers.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product (libdsmvc_hip.so) never links it.
 *
 * Built with -ffp-contract=off -fno-fast-math, as the reference's numerical
 * contract requires (reference CMakeLists.txt:361-371), so `a += w * x` is a
 * rounded product followed by a rounded sum and `fma()` is the only fused
 * operation.
 *
 * Pinning (what fixes each function independently of this file):
 *   orc_plan_build          the reference's seven f32_plan_hash /
 *                           complete_plan_hash goldens
 *                           (tests/numerical_conformance.hpp:126-164), the
 *                           long-double normal-matrix audit, the padding-index
 *                           table (tests/engine_tests.cpp:1349-1368), and the
 *                           reference engine built from its sources
 *                           (oracle/_ref) at every workload geometry.
 *   orc_axis_f32            production_output_hash goldens + oracle/_ref.
 *   orc_axis_f32_ordered    ordered_output_hash goldens + oracle/_ref.
 *   orc_axis_f64_ordered    F64 ordered goldens, the QR anchor
 *                           (tests/numerical_contract_tests.cpp:697-726),
 *                           dense numpy float64 solves on small systems.
 *   orc_2d / orc_2d_u8/u16  the mixed safe/risky 2D hash a9dcffe6649bd96e and
 *                           oracle/_ref's CpuExecutor scalar routes.
 */
#ifndef DSMVC_ORACLE_H
#define DSMVC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors dsmvc::KernelKind / BorderMode / F64Mode numbering
 * (reference include/dsmvc/engine.hpp:16-58). */
enum { ORC_BILINEAR = 0, ORC_BICUBIC = 1, ORC_LANCZOS = 2, ORC_SPLINE16 = 3,
       ORC_SPLINE36 = 4, ORC_SPLINE64 = 5 };
enum { ORC_ZERO = 0, ORC_REPEAT = 1, ORC_REFLECT101 = 2, ORC_SYMMETRIC = 3,
       ORC_MIRROR = 4 };
enum { ORC_F64_AUTO = 0, ORC_F64_NEVER = 1, ORC_F64_ALWAYS = 2 };

/* Status: 0 ok, 1 invalid argument, 2 runtime error (message in
 * orc_last_error()). */
typedef struct orc_request {
    int32_t source_size;
    int32_t destination_size;
    double active_length;
    double shift;
    int32_t kernel;
    int32_t taps;
    double b;
    double c;
    double blur;
    int32_t border;
    int32_t f64_mode;
} orc_request;

typedef struct orc_plan {
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
    uint32_t *offsets;          /* [n + 1] */
    int32_t *indices;           /* [nnz], ascending per row */
    float *weights;             /* [nnz] */
    float *lower_ld;            /* [W * n] */
    float *upper_l;             /* [W * n] */
    float *inverse_diagonal;    /* [n] */
    double *weights_f64;        /* [nnz] or NULL */
    double *normal_bands_f64;   /* [(W + 1) * n] or NULL */
    double *ldlt_bands_f64;     /* [(W + 1) * n] or NULL */
    double *inverse_diagonal_f64; /* [n] or NULL */
} orc_plan;

typedef struct orc_conversion {
    float input_offset;
    float input_scale;
    float output_scale;
    float output_offset;
    uint32_t output_maximum;
} orc_conversion;

const char *orc_last_error(void);

int32_t orc_padding_index(int64_t index, int32_t size, int32_t mode);

int orc_plan_build(const orc_request *request, orc_plan *plan);
void orc_plan_free(orc_plan *plan);

/* One vector.  Strides are in elements. */
int orc_axis_f32(const orc_plan *plan, const float *in, ptrdiff_t in_stride,
                 float *out, ptrdiff_t out_stride);
void orc_axis_f32_ordered(const orc_plan *plan, const float *in,
                          ptrdiff_t in_stride, float *out, ptrdiff_t out_stride);
void orc_axis_f64_ordered(const orc_plan *plan, const double *in,
                          ptrdiff_t in_stride, double *out,
                          ptrdiff_t out_stride);
void orc_axis_f64(const orc_plan *plan, const double *in, ptrdiff_t in_stride,
                  double *out, ptrdiff_t out_stride);

/* Planes.  Row strides in elements. */
int orc_rows(const orc_plan *plan, const float *in, ptrdiff_t in_stride,
             float *out, ptrdiff_t out_stride, int32_t rows);
int orc_columns(const orc_plan *plan, const float *in, ptrdiff_t in_stride,
                float *out, ptrdiff_t out_stride, int32_t columns);
int orc_2d(const orc_plan *h, const orc_plan *v, const float *in,
           ptrdiff_t in_stride, float *out, ptrdiff_t out_stride);
/* Ordered-FMA 2D: ordered rows then ordered columns (F32 plans only). */
int orc_2d_ordered(const orc_plan *h, const orc_plan *v, const float *in,
                   ptrdiff_t in_stride, float *out, ptrdiff_t out_stride);
int orc_2d_u8(const orc_plan *h, const orc_plan *v, const uint8_t *in,
              ptrdiff_t in_stride, uint8_t *out, ptrdiff_t out_stride,
              const orc_conversion *conversion);
int orc_2d_u16(const orc_plan *h, const orc_plan *v, const uint16_t *in,
               ptrdiff_t in_stride, uint16_t *out, ptrdiff_t out_stride,
               const orc_conversion *conversion);
/* FP64 accuracy reference: both axes solved in double with the plans' own
 * factors (use plans built with f64_mode = ORC_F64_ALWAYS), double output. */
int orc_2d_f64_exact(const orc_plan *h, const orc_plan *v, const double *in,
                     ptrdiff_t in_stride, double *out, ptrdiff_t out_stride);

#ifdef __cplusplus
}
#endif

#endif

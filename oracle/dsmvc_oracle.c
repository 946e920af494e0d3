/*
 * TEST INFRASTRUCTURE ONLY -- the oracle (see dsmvc_oracle.h for what it is,
 * who may load it and what pins each function).
 *
 * Each function cites the reference function whose arithmetic it restates;
 * paths are relative to the reference repository.  Every floating-point
 * expression keeps the reference's operand order, because the plan arrays
 * and scalar outputs are compared bit for bit.
 */
#include "dsmvc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_INVALID 1
#define ORC_RUNTIME 2

static _Thread_local char orc_error[256];

const char *orc_last_error(void) { return orc_error; }

static int fail(int status, const char *fmt, ...) {
    va_list args;
    va_start(args, fmt);
    vsnprintf(orc_error, sizeof orc_error, fmt, args);
    va_end(args);
    return status;
}

static int32_t imin(int32_t a, int32_t b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* Planner                                                                  */
/* ------------------------------------------------------------------------ */

/* src/axis_plan.cpp:253-256 */
static double round_half_up(double value) {
    return value < 0.0 ? floor(value + 0.5) : floor(value + 0.49999999999999994);
}

static double square(double v) { return v * v; }
static double cube(double v) { return v * v * v; }

/* src/axis_plan.cpp:266-270 (std::numbers::pi) */
static double sinc(double value) {
    if (value == 0.0) return 1.0;
    const double scaled = 3.141592653589793 * value;
    return sin(scaled) / scaled;
}

/* src/axis_plan.cpp:272-344 */
static double filter_weight(const orc_request *r, double distance) {
    double x = fabs(distance);
    const double b = r->b, c = r->c;
    switch (r->kernel) {
    case ORC_BILINEAR:
        return (1.0 - x) > 0.0 ? (1.0 - x) : 0.0;
    case ORC_BICUBIC:
        if (x < 1.0) {
            return ((12.0 - 9.0 * b - 6.0 * c) * cube(x)
                    + (-18.0 + 12.0 * b + 6.0 * c) * square(x)
                    + (6.0 - 2.0 * b)) / 6.0;
        }
        if (x < 2.0) {
            return ((-b - 6.0 * c) * cube(x)
                    + (6.0 * b + 30.0 * c) * square(x)
                    + (-12.0 * b - 48.0 * c) * x
                    + (8.0 * b + 24.0 * c)) / 6.0;
        }
        return 0.0;
    case ORC_LANCZOS:
        return r->taps > 0 && x < (double)r->taps
            ? sinc(x) * sinc(x / (double)r->taps) : 0.0;
    case ORC_SPLINE16:
        if (x < 1.0) return 1.0 - x / 5.0 - 9.0 * square(x) / 5.0 + cube(x);
        if (x < 2.0) {
            x -= 1.0;
            return -7.0 * x / 15.0 + 4.0 * square(x) / 5.0 - cube(x) / 3.0;
        }
        return 0.0;
    case ORC_SPLINE36:
        if (x < 1.0) {
            return 1.0 - 3.0 * x / 209.0 - 453.0 * square(x) / 209.0
                + 13.0 * cube(x) / 11.0;
        }
        if (x < 2.0) {
            x -= 1.0;
            return -156.0 * x / 209.0 + 270.0 * square(x) / 209.0
                - 6.0 * cube(x) / 11.0;
        }
        if (x < 3.0) {
            x -= 2.0;
            return 26.0 * x / 209.0 - 45.0 * square(x) / 209.0 + cube(x) / 11.0;
        }
        return 0.0;
    case ORC_SPLINE64:
        if (x < 1.0) {
            return 1.0 - 3.0 * x / 2911.0 - 6387.0 * square(x) / 2911.0
                + 49.0 * cube(x) / 41.0;
        }
        if (x < 2.0) {
            x -= 1.0;
            return -2328.0 * x / 2911.0 + 4032.0 * square(x) / 2911.0
                - 24.0 * cube(x) / 41.0;
        }
        if (x < 3.0) {
            x -= 2.0;
            return 582.0 * x / 2911.0 - 1008.0 * square(x) / 2911.0
                + 6.0 * cube(x) / 41.0;
        }
        if (x < 4.0) {
            x -= 3.0;
            return -97.0 * x / 2911.0 + 168.0 * square(x) / 2911.0
                - cube(x) / 41.0;
        }
        return 0.0;
    }
    return 0.0;
}

/* src/axis_plan.cpp:1024-1066; returns -2 for an invalid mode or size. */
int32_t orc_padding_index(int64_t index, int32_t size, int32_t mode) {
    if (size <= 0) return -2;
    if (mode < ORC_ZERO || mode > ORC_MIRROR) return -2;
    const int64_t extent = size;
    if (index >= 0 && index < extent) return (int32_t)index;
    switch (mode) {
    case ORC_ZERO:
        return -1;
    case ORC_REPEAT:
        return index < 0 ? 0 : size - 1;
    case ORC_REFLECT101: {
        if (size == 1) return 0;
        const int64_t period = 2 * (extent - 1);
        int64_t mapped = index % period;
        if (mapped < 0) mapped += period;
        return (int32_t)(mapped < extent ? mapped : period - mapped);
    }
    case ORC_SYMMETRIC: {
        const int64_t period = 2 * extent;
        int64_t mapped = index % period;
        if (mapped < 0) mapped += period;
        return (int32_t)(mapped < extent ? mapped : period - 1 - mapped);
    }
    case ORC_MIRROR:
        if (index < -extent || index >= 2 * extent) return -1;
        return (int32_t)(index < 0 ? -index - 1 : 2 * extent - index - 1);
    }
    return -2;
}

/* src/axis_plan.cpp:370-424; returns support or a negative status. */
static int filter_support(const orc_request *r, int32_t *support) {
    if (r->source_size <= 0 || r->destination_size <= 0)
        return fail(ORC_INVALID, "axis dimensions must be positive");
    if (!(r->active_length > 0.0) || !isfinite(r->active_length)
        || !isfinite(r->shift))
        return fail(ORC_INVALID,
                    "active length and shift must be finite, with positive active length");
    if (r->border < ORC_ZERO || r->border > ORC_MIRROR)
        return fail(ORC_INVALID, "padding mode is invalid");
    if (r->f64_mode < ORC_F64_AUTO || r->f64_mode > ORC_F64_ALWAYS)
        return fail(ORC_INVALID, "f64 mode is invalid");
    if (!(r->blur > 0.0) || !isfinite(r->blur))
        return fail(ORC_INVALID, "blur must be finite and greater than zero");
    int32_t base = 0;
    switch (r->kernel) {
    case ORC_BILINEAR: base = 1; break;
    case ORC_BICUBIC:
    case ORC_SPLINE16: base = 2; break;
    case ORC_SPLINE36: base = 3; break;
    case ORC_SPLINE64: base = 4; break;
    case ORC_LANCZOS: base = r->taps; break;
    default: return fail(ORC_INVALID, "kernel is invalid");
    }
    const int32_t maximum_support = (2147483647 - 1) / 2;
    if (base <= 0 || base > maximum_support)
        return fail(ORC_INVALID, "filter support is invalid or too large");
    const double scaled = (double)base * r->blur;
    if (!isfinite(scaled) || scaled > (double)maximum_support)
        return fail(ORC_RUNTIME, "effective filter support is too large");
    const double effective = ceil(scaled);
    if (!(effective >= 1.0) || effective > (double)maximum_support)
        return fail(ORC_RUNTIME, "effective filter support is too large");
    *support = (int32_t)effective;
    return ORC_OK;
}

typedef struct {
    uint32_t *offsets;
    int32_t *indices;
    double *weights;
    size_t nnz;
} dcsr;

static void dcsr_free(dcsr *m) {
    free(m->offsets);
    free(m->indices);
    free(m->weights);
    memset(m, 0, sizeof *m);
}

static int cmp_i32(const void *a, const void *b) {
    const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/*
 * Geometry + descale matrix A (source rows, destination columns):
 * src/axis_plan.cpp:456-518 (build_geometry) and 535-599
 * (make_descale_matrix), fused per source row.
 */
static int make_descale_matrix(const orc_request *r, int32_t support, dcsr *out) {
    const int32_t taps = 2 * support;
    const int32_t m = r->source_size;
    memset(out, 0, sizeof *out);
    out->offsets = calloc((size_t)m + 1, sizeof(uint32_t));
    out->indices = malloc((size_t)m * (size_t)taps * sizeof(int32_t));
    out->weights = malloc((size_t)m * (size_t)taps * sizeof(double));
    int32_t *tap_index = malloc((size_t)taps * sizeof(int32_t));
    int32_t *unique = malloc((size_t)taps * sizeof(int32_t));
    double *distance = malloc((size_t)taps * sizeof(double));
    double *tap_weight = malloc((size_t)taps * sizeof(double));
    double *coalesced = malloc((size_t)taps * sizeof(double));
    unsigned char *seen = malloc((size_t)taps);
    int status = ORC_OK;
    if (!out->offsets || !out->indices || !out->weights || !tap_index || !unique
        || !distance || !tap_weight || !coalesced || !seen) {
        status = fail(ORC_RUNTIME, "out of memory");
        goto done;
    }
    const double ratio = (double)r->source_size / r->active_length;
    const int unity_blur = r->blur == 1.0;
    size_t nnz = 0;
    for (int32_t row = 0; row < m; ++row) {
        const double position = ((double)row + 0.5) / ratio + r->shift;
        const double begin = round_half_up(position - (double)support) + 0.5;
        int32_t unique_count = 0;
        for (int32_t tap = 0; tap < taps; ++tap) {
            const double center = begin + (double)tap;
            distance[tap] = center - position;
            const double fl = floor(center);
            if (!isfinite(fl) || fl < -0x1p63 || fl >= 0x1p63) {
                status = fail(ORC_RUNTIME,
                              "shift places filter support outside the 64-bit pixel grid");
                goto done;
            }
            const int32_t index = orc_padding_index((int64_t)fl, r->destination_size,
                                                    r->border);
            tap_index[tap] = index;
            if (index >= 0) {
                int found = 0;
                for (int32_t k = 0; k < unique_count; ++k) found |= unique[k] == index;
                if (!found) unique[unique_count++] = index;
            }
        }
        qsort(unique, (size_t)unique_count, sizeof(int32_t), cmp_i32);

        double total = 0.0;
        for (int32_t tap = 0; tap < taps; ++tap) {
            const double d = unity_blur ? distance[tap] : distance[tap] / r->blur;
            const double w = filter_weight(r, d);
            if (!isfinite(w)) {
                status = fail(ORC_RUNTIME, "filter returned a non-finite value");
                goto done;
            }
            tap_weight[tap] = w;
            total += w;
        }
        if (!isfinite(total) || total == 0.0) {
            status = fail(ORC_RUNTIME, "filter produced a zero or non-finite weight sum");
            goto done;
        }
        for (int32_t k = 0; k < unique_count; ++k) {
            coalesced[k] = 0.0;
            seen[k] = 0;
        }
        for (int32_t tap = 0; tap < taps; ++tap) {
            const int32_t index = tap_index[tap];
            if (index < 0) continue;
            int32_t slot = 0;
            while (unique[slot] != index) ++slot;
            const double w = tap_weight[tap] / total;
            if (w == 0.0) continue;
            coalesced[slot] += w;
            seen[slot] = 1;
        }
        for (int32_t k = 0; k < unique_count; ++k) {
            if (!seen[k]) continue;
            out->indices[nnz] = unique[k];
            out->weights[nnz] = coalesced[k];
            ++nnz;
        }
        out->offsets[row + 1] = (uint32_t)nnz;
    }
    out->nnz = nnz;
done:
    free(tap_index);
    free(unique);
    free(distance);
    free(tap_weight);
    free(coalesced);
    free(seen);
    if (status != ORC_OK) dcsr_free(out);
    return status;
}

/* src/axis_plan.cpp:601-626 */
static int transpose_csr(const dcsr *in, int32_t rows, int32_t columns, dcsr *out) {
    memset(out, 0, sizeof *out);
    out->offsets = calloc((size_t)columns + 1, sizeof(uint32_t));
    out->indices = malloc((in->nnz ? in->nnz : 1) * sizeof(int32_t));
    out->weights = malloc((in->nnz ? in->nnz : 1) * sizeof(double));
    uint32_t *cursor = malloc(((size_t)columns + 1) * sizeof(uint32_t));
    if (!out->offsets || !out->indices || !out->weights || !cursor) {
        free(cursor);
        dcsr_free(out);
        return fail(ORC_RUNTIME, "out of memory");
    }
    for (size_t k = 0; k < in->nnz; ++k) ++out->offsets[in->indices[k] + 1];
    for (int32_t c = 0; c < columns; ++c) out->offsets[c + 1] += out->offsets[c];
    memcpy(cursor, out->offsets, ((size_t)columns + 1) * sizeof(uint32_t));
    for (int32_t row = 0; row < rows; ++row) {
        for (uint32_t k = in->offsets[row]; k < in->offsets[row + 1]; ++k) {
            const uint32_t target = cursor[in->indices[k]]++;
            out->indices[target] = row;
            out->weights[target] = in->weights[k];
        }
    }
    out->nnz = in->nnz;
    free(cursor);
    return ORC_OK;
}

/* src/axis_plan.cpp:628-668: bands[d * n + i] = G(i, i + d). */
static void form_normal_bands(const dcsr *t, int32_t n, int32_t w, double *bands) {
    for (int32_t left_row = 0; left_row < n; ++left_row) {
        const int32_t maximum_distance = imin(w, n - left_row - 1);
        for (int32_t d = 0; d <= maximum_distance; ++d) {
            const int32_t right_row = left_row + d;
            uint32_t left = t->offsets[left_row];
            const uint32_t left_end = t->offsets[left_row + 1];
            uint32_t right = t->offsets[right_row];
            const uint32_t right_end = t->offsets[right_row + 1];
            double sum = 0.0;
            while (left < left_end && right < right_end) {
                const int32_t li = t->indices[left], ri = t->indices[right];
                if (li < ri) {
                    ++left;
                } else if (ri < li) {
                    ++right;
                } else {
                    sum += t->weights[left] * t->weights[right];
                    ++left;
                    ++right;
                }
            }
            bands[(size_t)d * (size_t)n + (size_t)left_row] = sum;
        }
    }
}

/* src/axis_plan.cpp:675-712 */
static void band_condition_bounds(const double *bands, int32_t n, int32_t w,
                                  double *work, double *one_norm,
                                  double *dominance_rcond) {
    for (int32_t i = 0; i < n; ++i) work[i] = 0.0;
    for (int32_t col = 0; col < n; ++col) {
        work[col] += fabs(bands[col]);
        const int32_t available = imin(w, n - col - 1);
        for (int32_t d = 1; d <= available; ++d) {
            const double value = fabs(bands[(size_t)d * (size_t)n + (size_t)col]);
            work[col] += value;
            work[col + d] += value;
        }
    }
    double norm = n > 0 ? work[0] : 0.0;
    for (int32_t i = 1; i < n; ++i)
        if (norm < work[i]) norm = work[i];
    double margin = INFINITY;
    for (int32_t row = 0; row < n; ++row) {
        const double candidate = 2.0 * fabs(bands[row]) - work[row];
        if (candidate < margin) margin = candidate;
    }
    *one_norm = norm;
    *dominance_rcond = margin > 0.0 && norm > 0.0 ? margin / norm : 0.0;
}

static double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

/* src/axis_plan.cpp:714-764 */
static double rcond_lower_bound(double matrix_norm, const double *f, int32_t n,
                                int32_t w, double *work) {
    const double eps = DBL_EPSILON;
    for (int32_t i = 0; i < n; ++i) work[i] = 1.0;
    double inv_l_inf = 1.0, inv_diag = 0.0;
    for (int32_t row = 0; row < n; ++row) {
        double sum = 1.0;
        const int32_t available = imin(w, row);
        for (int32_t d = 1; d <= available; ++d) {
            const int32_t col = row - d;
            sum += fabs(f[(size_t)d * (size_t)n + (size_t)col]) * work[col];
        }
        work[row] = sum;
        if (inv_l_inf < sum) inv_l_inf = sum;
        const double candidate = 1.0 / fabs(f[row] + eps);
        if (inv_diag < candidate) inv_diag = candidate;
    }
    for (int32_t i = 0; i < n; ++i) work[i] = 1.0;
    double inv_l_one = 1.0;
    for (int32_t col = n - 1; col >= 0; --col) {
        double sum = 1.0;
        const int32_t available = imin(w, n - col - 1);
        for (int32_t d = 1; d <= available; ++d) {
            sum += fabs(f[(size_t)d * (size_t)n + (size_t)col]) * work[col + d];
        }
        work[col] = sum;
        if (inv_l_one < sum) inv_l_one = sum;
    }
    const double bound = inv_l_one * inv_diag * inv_l_inf;
    if (!(matrix_norm > 0.0) || !(bound > 0.0) || !isfinite(matrix_norm)
        || !isfinite(bound))
        return 0.0;
    return clamp01(1.0 / (matrix_norm * bound));
}

/* src/axis_plan.cpp:766-791 */
static void factor_banded_ldlt(double *bands, int32_t n, int32_t w) {
    const double eps = DBL_EPSILON;
    const size_t width = (size_t)n;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t end = imin(w + 1, n - i);
        const double pivot = bands[i] + eps;
        for (int32_t d = 1; d < end; ++d) {
            const double multiplier = bands[(size_t)d * width + (size_t)i] / pivot;
            for (int32_t o = 0; o < end - d; ++o) {
                bands[(size_t)o * width + (size_t)(i + d)] -=
                    multiplier * bands[(size_t)(d + o) * width + (size_t)i];
            }
        }
        const double inverse_pivot = 1.0 / pivot;
        for (int32_t d = 1; d < end; ++d) bands[(size_t)d * width + (size_t)i] *= inverse_pivot;
    }
}

/* src/axis_plan.cpp:793-823 */
static void solve_banded_ldlt(const double *f, int32_t n, int32_t w, double *v,
                              ptrdiff_t s) {
    const double eps = DBL_EPSILON;
    const size_t width = (size_t)n;
    for (int32_t i = 0; i < n; ++i) {
        double value = v[(ptrdiff_t)i * s];
        const int32_t available = imin(w, i);
        for (int32_t d = available; d >= 1; --d)
            value -= f[(size_t)d * width + (size_t)(i - d)] * v[(ptrdiff_t)(i - d) * s];
        v[(ptrdiff_t)i * s] = value;
    }
    for (int32_t i = 0; i < n; ++i) v[(ptrdiff_t)i * s] /= f[i] + eps;
    for (int32_t i = n - 2; i >= 0; --i) {
        double value = v[(ptrdiff_t)i * s];
        const int32_t available = imin(w, n - i - 1);
        for (int32_t d = available; d >= 1; --d)
            value -= f[(size_t)d * width + (size_t)i] * v[(ptrdiff_t)(i + d) * s];
        v[(ptrdiff_t)i * s] = value;
    }
}

/* src/axis_plan.cpp:825-873 (Hager 1-norm estimate, 8 iterations) */
static double rcond_estimate(double matrix_norm, const double *f, int32_t n,
                             int32_t w) {
    const size_t size = (size_t)n;
    double *x = malloc(size * sizeof(double));
    double *y = malloc(size * sizeof(double));
    double *z = malloc(size * sizeof(double));
    double estimate = 0.0;
    size_t previous = size;
    for (size_t i = 0; i < size; ++i) x[i] = 1.0 / (double)n;
    for (int it = 0; it < 8; ++it) {
        memcpy(y, x, size * sizeof(double));
        solve_banded_ldlt(f, n, w, y, 1);
        double next = 0.0;
        for (size_t i = 0; i < size; ++i) next = next + fabs(y[i]);
        if (estimate < next) estimate = next;
        for (size_t i = 0; i < size; ++i) z[i] = y[i] < 0.0 ? -1.0 : 1.0;
        solve_banded_ldlt(f, n, w, z, 1);
        size_t index = 0;
        for (size_t i = 1; i < size; ++i)
            if (fabs(z[index]) < fabs(z[i])) index = i;
        double dot = 0.0;
        for (size_t i = 0; i < size; ++i) dot = dot + z[i] * x[i];
        if (index == previous || fabs(z[index]) <= dot) break;
        for (size_t i = 0; i < size; ++i) x[i] = 0.0;
        x[index] = 1.0;
        previous = index;
    }
    free(x);
    free(y);
    free(z);
    const double inverse_norm = estimate;
    if (!(matrix_norm > 0.0) || !(inverse_norm > 0.0) || !isfinite(matrix_norm)
        || !isfinite(inverse_norm))
        return 0.0;
    return clamp01(1.0 / (matrix_norm * inverse_norm));
}

void orc_plan_free(orc_plan *p) {
    if (!p) return;
    free(p->offsets);
    free(p->indices);
    free(p->weights);
    free(p->lower_ld);
    free(p->upper_l);
    free(p->inverse_diagonal);
    free(p->weights_f64);
    free(p->normal_bands_f64);
    free(p->ldlt_bands_f64);
    free(p->inverse_diagonal_f64);
    memset(p, 0, sizeof *p);
}

static int all_finite_f(const float *v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(v[i])) return 0;
    return 1;
}

static int all_finite_d(const double *v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(v[i])) return 0;
    return 1;
}

/* src/axis_plan.cpp:875-970 (build_axis_plan_impl) */
int orc_plan_build(const orc_request *r, orc_plan *p) {
    memset(p, 0, sizeof *p);
    int32_t support = 0;
    int status = filter_support(r, &support);
    if (status != ORC_OK) return status;
    dcsr matrix, t;
    status = make_descale_matrix(r, support, &matrix);
    if (status != ORC_OK) return status;
    status = transpose_csr(&matrix, r->source_size, r->destination_size, &t);
    dcsr_free(&matrix);
    if (status != ORC_OK) return status;

    const int32_t n = r->destination_size;
    const int32_t w = imin(2 * support - 1, n - 1);
    const size_t width = (size_t)n;
    const size_t bands = ((size_t)w + 1) * width;
    double *normal = calloc(bands, sizeof(double));
    double *factors = malloc(bands * sizeof(double));
    double *work = malloc(width * sizeof(double));
    form_normal_bands(&t, n, w, normal);
    memcpy(factors, normal, bands * sizeof(double));
    factor_banded_ldlt(factors, n, w);

    const double threshold = 1.0e-4;
    double one_norm = 0.0, rcond = 0.0;
    band_condition_bounds(normal, n, w, work, &one_norm, &rcond);
    if (rcond < threshold) {
        rcond = rcond_lower_bound(one_norm, factors, n, w, work);
        if (rcond < threshold) rcond = rcond_estimate(one_norm, factors, n, w);
    }
    free(work);
    const int f64 = r->f64_mode == ORC_F64_ALWAYS
        || (r->f64_mode == ORC_F64_AUTO && rcond < threshold);

    p->source_size = r->source_size;
    p->destination_size = n;
    p->support = support;
    p->half_bandwidth = w;
    p->active_length = r->active_length;
    p->shift = r->shift;
    p->normal_rcond = rcond;
    p->requires_float64 = f64;
    p->nnz = (int32_t)t.nnz;
    p->offsets = t.offsets;
    p->indices = t.indices;
    p->weights = malloc((t.nnz ? t.nnz : 1) * sizeof(float));
    for (size_t k = 0; k < t.nnz; ++k) p->weights[k] = (float)t.weights[k];
    const double eps = DBL_EPSILON;
    if (f64) {
        p->normal_inf_norm = one_norm;
        p->weights_f64 = t.weights;
        p->normal_bands_f64 = normal;
        p->ldlt_bands_f64 = factors;
        p->inverse_diagonal_f64 = malloc(width * sizeof(double));
        for (int32_t i = 0; i < n; ++i) p->inverse_diagonal_f64[i] = 1.0 / (factors[i] + eps);
    } else {
        free(t.weights);
        free(normal);
    }
    const size_t count = (size_t)w * width;
    p->lower_ld = calloc(count ? count : 1, sizeof(float));
    p->upper_l = calloc(count ? count : 1, sizeof(float));
    p->inverse_diagonal = malloc(width * sizeof(float));
    for (int32_t i = 0; i < n; ++i) {
        const double diagonal = factors[i];
        p->inverse_diagonal[i] = (float)(1.0 / (diagonal + eps));
        const int32_t available = imin(w, n - i - 1);
        for (int32_t d = 1; d <= available; ++d) {
            const double l = factors[(size_t)d * width + (size_t)i];
            p->upper_l[(size_t)(d - 1) * width + (size_t)i] = (float)l;
            p->lower_ld[(size_t)(d - 1) * width + (size_t)(i + d)] = (float)(l * diagonal);
        }
    }
    if (!f64) free(factors);

    /* AxisPlan::valid (src/axis_plan.cpp:1259-1341): finiteness. */
    int valid = isfinite(p->normal_rcond) && p->normal_rcond >= 0.0
        && p->normal_rcond <= 1.0 && all_finite_f(p->weights, t.nnz)
        && all_finite_f(p->lower_ld, count) && all_finite_f(p->upper_l, count)
        && all_finite_f(p->inverse_diagonal, width);
    if (valid && f64) {
        valid = p->normal_inf_norm > 0.0 && all_finite_d(p->weights_f64, t.nnz)
            && all_finite_d(p->normal_bands_f64, bands)
            && all_finite_d(p->ldlt_bands_f64, bands)
            && all_finite_d(p->inverse_diagonal_f64, width);
    }
    if (!valid) {
        orc_plan_free(p);
        return fail(ORC_RUNTIME, "failed to build inverse axis plan");
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Axis solvers                                                             */
/* ------------------------------------------------------------------------ */

/* src/axis_plan.cpp:972-1018 (inverse_axis_impl<B>), F32 plans. */
static void axis_f32_production(const orc_plan *p, const float *in, ptrdiff_t is,
                                float *out, ptrdiff_t os) {
    const int32_t n = p->destination_size, w = p->half_bandwidth;
    const size_t width = (size_t)n;
    for (int32_t i = 0; i < n; ++i) {
        float sum = 0.0F;
        for (uint32_t k = p->offsets[i]; k < p->offsets[i + 1]; ++k)
            sum += p->weights[k] * in[(ptrdiff_t)p->indices[k] * is];
        const int32_t available = imin(w, i);
        for (int32_t d = available; d >= 1; --d)
            sum -= p->lower_ld[(size_t)(d - 1) * width + (size_t)i]
                * out[(ptrdiff_t)(i - d) * os];
        out[(ptrdiff_t)i * os] = sum * p->inverse_diagonal[i];
    }
    for (int32_t i = n - 2; i >= 0; --i) {
        float sum = 0.0F;
        const int32_t available = imin(w, n - i - 1);
        if (w == 3) {
            for (int32_t d = 1; d <= available; ++d)
                sum += p->upper_l[(size_t)(d - 1) * width + (size_t)i]
                    * out[(ptrdiff_t)(i + d) * os];
        } else {
            for (int32_t d = available; d >= 1; --d)
                sum += p->upper_l[(size_t)(d - 1) * width + (size_t)i]
                    * out[(ptrdiff_t)(i + d) * os];
        }
        out[(ptrdiff_t)i * os] -= sum;
    }
}

/* src/axis_plan.cpp:1193-1248 (detail::inverse_axis_f64) */
void orc_axis_f64(const orc_plan *p, const double *in, ptrdiff_t is, double *out,
                  ptrdiff_t os) {
    const int32_t n = p->destination_size, w = p->half_bandwidth;
    const size_t width = (size_t)n;
    if (p->requires_float64) {
        for (int32_t i = 0; i < n; ++i) {
            double sum = 0.0;
            for (uint32_t k = p->offsets[i]; k < p->offsets[i + 1]; ++k)
                sum += p->weights_f64[k] * in[(ptrdiff_t)p->indices[k] * is];
            out[(ptrdiff_t)i * os] = sum;
        }
        solve_banded_ldlt(p->ldlt_bands_f64, n, w, out, os);
        return;
    }
    for (int32_t i = 0; i < n; ++i) {
        double sum = 0.0;
        for (uint32_t k = p->offsets[i]; k < p->offsets[i + 1]; ++k)
            sum += (double)p->weights[k] * in[(ptrdiff_t)p->indices[k] * is];
        const int32_t available = imin(w, i);
        for (int32_t d = available; d >= 1; --d)
            sum -= (double)p->lower_ld[(size_t)(d - 1) * width + (size_t)i]
                * out[(ptrdiff_t)(i - d) * os];
        out[(ptrdiff_t)i * os] = sum * (double)p->inverse_diagonal[i];
    }
    for (int32_t i = n - 2; i >= 0; --i) {
        double sum = 0.0;
        const int32_t available = imin(w, n - i - 1);
        for (int32_t d = available; d >= 1; --d)
            sum += (double)p->upper_l[(size_t)(d - 1) * width + (size_t)i]
                * out[(ptrdiff_t)(i + d) * os];
        out[(ptrdiff_t)i * os] -= sum;
    }
}

/* src/axis_plan.cpp:1409-1449 (dsmvc::inverse_axis_f32) */
int orc_axis_f32(const orc_plan *p, const float *in, ptrdiff_t is, float *out,
                 ptrdiff_t os) {
    if (!in || !out || is == 0 || os == 0)
        return fail(ORC_INVALID, "invalid inverse axis arguments");
    if (!p->requires_float64) {
        axis_f32_production(p, in, is, out, os);
        return ORC_OK;
    }
    double *source = malloc((size_t)p->source_size * sizeof(double));
    double *destination = malloc((size_t)p->destination_size * sizeof(double));
    int status = ORC_OK;
    for (int32_t i = 0; i < p->source_size; ++i) {
        const float value = in[(ptrdiff_t)i * is];
        if (!isfinite(value)) {
            status = fail(ORC_RUNTIME, "Float64 inverse axis input contains NaN or infinity");
            goto done;
        }
        source[i] = (double)value;
    }
    orc_axis_f64(p, source, 1, destination, 1);
    for (int32_t i = 0; i < p->destination_size; ++i) {
        const double value = destination[i];
        const float converted = (float)value;
        if (!isfinite(value) || !isfinite(converted)) {
            status = fail(ORC_RUNTIME, "Float64 inverse axis produced NaN or infinity");
            goto done;
        }
        out[(ptrdiff_t)i * os] = converted;
    }
done:
    free(source);
    free(destination);
    return status;
}

/* src/axis_plan.cpp:1068-1122 (detail::inverse_axis_f32_ordered) */
void orc_axis_f32_ordered(const orc_plan *p, const float *in, ptrdiff_t is,
                          float *out, ptrdiff_t os) {
    const int32_t n = p->destination_size, w = p->half_bandwidth;
    const size_t width = (size_t)n;
    for (int32_t i = 0; i < n; ++i) {
        float value = 0.0F;
        for (uint32_t k = p->offsets[i]; k < p->offsets[i + 1]; ++k)
            value = fmaf(p->weights[k], in[(ptrdiff_t)p->indices[k] * is], value);
        const int32_t available = imin(w, i);
        for (int32_t d = available; d >= 1; --d)
            value = fmaf(-p->lower_ld[(size_t)(d - 1) * width + (size_t)i],
                         out[(ptrdiff_t)(i - d) * os], value);
        out[(ptrdiff_t)i * os] = value * p->inverse_diagonal[i];
    }
    for (int32_t i = n - 2; i >= 0; --i) {
        float value = out[(ptrdiff_t)i * os];
        const int32_t available = imin(w, n - i - 1);
        if (w == 3) {
            for (int32_t d = 1; d <= available; ++d)
                value = fmaf(-p->upper_l[(size_t)(d - 1) * width + (size_t)i],
                             out[(ptrdiff_t)(i + d) * os], value);
        } else {
            for (int32_t d = available; d >= 1; --d)
                value = fmaf(-p->upper_l[(size_t)(d - 1) * width + (size_t)i],
                             out[(ptrdiff_t)(i + d) * os], value);
        }
        out[(ptrdiff_t)i * os] = value;
    }
}

/* src/axis_plan.cpp:1124-1191 (detail::inverse_axis_f64_ordered) */
void orc_axis_f64_ordered(const orc_plan *p, const double *in, ptrdiff_t is,
                          double *out, ptrdiff_t os) {
    const int32_t n = p->destination_size, w = p->half_bandwidth;
    const size_t width = (size_t)n;
    const int retained = p->requires_float64;
    for (int32_t i = 0; i < n; ++i) {
        double value = 0.0;
        for (uint32_t k = p->offsets[i]; k < p->offsets[i + 1]; ++k) {
            const double weight = retained ? p->weights_f64[k] : (double)p->weights[k];
            value = fma(weight, in[(ptrdiff_t)p->indices[k] * is], value);
        }
        const int32_t available = imin(w, i);
        for (int32_t d = available; d >= 1; --d) {
            const double factor = retained
                ? p->ldlt_bands_f64[(size_t)d * width + (size_t)(i - d)]
                : (double)p->lower_ld[(size_t)(d - 1) * width + (size_t)i];
            value = fma(-factor, out[(ptrdiff_t)(i - d) * os], value);
        }
        out[(ptrdiff_t)i * os] = retained ? value : value * (double)p->inverse_diagonal[i];
    }
    if (retained)
        for (int32_t i = 0; i < n; ++i) out[(ptrdiff_t)i * os] *= p->inverse_diagonal_f64[i];
    for (int32_t i = n - 2; i >= 0; --i) {
        double value = out[(ptrdiff_t)i * os];
        const int32_t available = imin(w, n - i - 1);
        for (int32_t d = available; d >= 1; --d) {
            const double factor = retained
                ? p->ldlt_bands_f64[(size_t)d * width + (size_t)i]
                : (double)p->upper_l[(size_t)(d - 1) * width + (size_t)i];
            value = fma(-factor, out[(ptrdiff_t)(i + d) * os], value);
        }
        out[(ptrdiff_t)i * os] = value;
    }
}

/* ------------------------------------------------------------------------ */
/* Plane drivers (src/cpu_executor.cpp scalar routes)                        */
/* ------------------------------------------------------------------------ */

static int finite_matrix(const float *v, int32_t rows, int32_t cols, ptrdiff_t s) {
    for (int32_t r = 0; r < rows; ++r)
        for (int32_t c = 0; c < cols; ++c)
            if (!isfinite(v[(ptrdiff_t)r * s + c])) return 0;
    return 1;
}

/* CpuExecutor::inverse_rows scalar branch + inverse_rows_f64
 * (src/cpu_executor.cpp:903-1171, 1417-1448). */
int orc_rows(const orc_plan *p, const float *in, ptrdiff_t is, float *out,
             ptrdiff_t os, int32_t rows) {
    if (!in || !out || is < p->source_size || os < p->destination_size || rows < 0)
        return fail(ORC_INVALID, "invalid row executor arguments");
    if (!p->requires_float64) {
        for (int32_t r = 0; r < rows; ++r)
            axis_f32_production(p, in + (ptrdiff_t)r * is, 1, out + (ptrdiff_t)r * os, 1);
        return ORC_OK;
    }
    if (!finite_matrix(in, rows, p->source_size, is))
        return fail(ORC_RUNTIME, "CPU Float64 row input contains NaN or infinity");
    double *source = malloc((size_t)p->source_size * sizeof(double));
    double *destination = malloc((size_t)p->destination_size * sizeof(double));
    for (int32_t r = 0; r < rows; ++r) {
        for (int32_t c = 0; c < p->source_size; ++c)
            source[c] = (double)in[(ptrdiff_t)r * is + c];
        orc_axis_f64(p, source, 1, destination, 1);
        for (int32_t c = 0; c < p->destination_size; ++c)
            out[(ptrdiff_t)r * os + c] = (float)destination[c];
    }
    free(source);
    free(destination);
    if (!finite_matrix(out, rows, p->destination_size, os))
        return fail(ORC_RUNTIME, "CPU Float64 row execution produced NaN or infinity");
    return ORC_OK;
}

/* CpuExecutor::inverse_columns scalar branch + inverse_columns_f64
 * (src/cpu_executor.cpp:1391-1394, 1450-1485). */
int orc_columns(const orc_plan *p, const float *in, ptrdiff_t is, float *out,
                ptrdiff_t os, int32_t columns) {
    if (!in || !out || columns < 0)
        return fail(ORC_INVALID, "invalid column executor arguments");
    if (!p->requires_float64) {
        for (int32_t c = 0; c < columns; ++c)
            axis_f32_production(p, in + c, is, out + c, os);
        return ORC_OK;
    }
    if (!finite_matrix(in, p->source_size, columns, is))
        return fail(ORC_RUNTIME, "CPU Float64 column input contains NaN or infinity");
    double *source = malloc((size_t)p->source_size * sizeof(double));
    double *destination = malloc((size_t)p->destination_size * sizeof(double));
    for (int32_t c = 0; c < columns; ++c) {
        for (int32_t r = 0; r < p->source_size; ++r)
            source[r] = (double)in[(ptrdiff_t)r * is + c];
        orc_axis_f64(p, source, 1, destination, 1);
        for (int32_t r = 0; r < p->destination_size; ++r)
            out[(ptrdiff_t)r * os + c] = (float)destination[r];
    }
    free(source);
    free(destination);
    if (!finite_matrix(out, p->destination_size, columns, os))
        return fail(ORC_RUNTIME, "CPU Float64 column execution produced NaN or infinity");
    return ORC_OK;
}

/*
 * inverse_2d_f64<Sample> (src/cpu_executor.cpp:1528-1651): a double
 * intermediate between the axes.  kind 0 = float, 1 = u8, 2 = u16.
 */
static int twod_f64(const orc_plan *h, const orc_plan *v, const void *in,
                    ptrdiff_t is, void *out, ptrdiff_t os,
                    const orc_conversion *conv, int kind) {
    const int32_t ist = h->destination_size;
    double *inter = malloc((size_t)v->source_size * (size_t)ist * sizeof(double));
    double *source = malloc((size_t)h->source_size * sizeof(double));
    double *destination = malloc((size_t)v->destination_size * sizeof(double));
    int nonfinite = 0;
    for (int32_t r = 0; r < v->source_size; ++r) {
        for (int32_t c = 0; c < h->source_size; ++c) {
            const ptrdiff_t at = (ptrdiff_t)r * is + c;
            if (kind == 0) {
                source[c] = (double)((const float *)in)[at];
            } else {
                const double x = kind == 1 ? (double)((const uint8_t *)in)[at]
                                           : (double)((const uint16_t *)in)[at];
                source[c] = (x - (double)conv->input_offset) * (double)conv->input_scale;
            }
        }
        orc_axis_f64(h, source, 1, inter + (ptrdiff_t)r * ist, 1);
    }
    for (int32_t c = 0; c < h->destination_size; ++c) {
        orc_axis_f64(v, inter + c, ist, destination, 1);
        for (int32_t r = 0; r < v->destination_size; ++r) {
            const double value = destination[r];
            const ptrdiff_t at = (ptrdiff_t)r * os + c;
            if (kind == 0) {
                const float converted = (float)value;
                if (!isfinite(value) || !isfinite(converted)) nonfinite = 1;
                ((float *)out)[at] = converted;
                continue;
            }
            const double scaled = value * (double)conv->output_scale
                + (double)conv->output_offset;
            if (!isfinite(value) || !isfinite(scaled)) {
                nonfinite = 1;
                if (kind == 1) ((uint8_t *)out)[at] = 0;
                else ((uint16_t *)out)[at] = 0;
                continue;
            }
            const double maximum = (double)conv->output_maximum;
            const double clamped = scaled < 0.0 ? 0.0 : (maximum < scaled ? maximum : scaled);
            const double rounded = nearbyint(clamped);
            if (kind == 1) ((uint8_t *)out)[at] = (uint8_t)rounded;
            else ((uint16_t *)out)[at] = (uint16_t)rounded;
        }
    }
    free(inter);
    free(source);
    free(destination);
    if (nonfinite) return fail(ORC_RUNTIME, "CPU Float64 2D execution produced NaN or infinity");
    return ORC_OK;
}

/*
 * CpuExecutor::inverse_2d scalar (src/cpu_executor.cpp:1941-2049).  The
 * streamed scalar route (forward_2d_rhs_scalar + backward_rhs_scalar) is bit
 * identical to rows-then-columns (reference tests/engine_tests.cpp:318-325),
 * so the restatement uses the two-pass form.
 */
int orc_2d(const orc_plan *h, const orc_plan *v, const float *in, ptrdiff_t is,
           float *out, ptrdiff_t os) {
    if (!in || !out || is < h->source_size || os < h->destination_size
        || v->source_size < 1 || v->destination_size < 1)
        return fail(ORC_INVALID, "invalid 2D executor arguments");
    if (h->requires_float64 || v->requires_float64) {
        if (!finite_matrix(in, v->source_size, h->source_size, is))
            return fail(ORC_RUNTIME, "CPU Float64 2D input contains NaN or infinity");
        return twod_f64(h, v, in, is, out, os, NULL, 0);
    }
    const int32_t ist = h->destination_size;
    float *inter = malloc((size_t)v->source_size * (size_t)ist * sizeof(float));
    for (int32_t r = 0; r < v->source_size; ++r)
        axis_f32_production(h, in + (ptrdiff_t)r * is, 1, inter + (ptrdiff_t)r * ist, 1);
    for (int32_t c = 0; c < h->destination_size; ++c)
        axis_f32_production(v, inter + c, ist, out + c, os);
    free(inter);
    return ORC_OK;
}

int orc_2d_ordered(const orc_plan *h, const orc_plan *v, const float *in,
                   ptrdiff_t is, float *out, ptrdiff_t os) {
    if (h->requires_float64 || v->requires_float64)
        return fail(ORC_INVALID, "ordered F32 2D needs F32 plans");
    const int32_t ist = h->destination_size;
    float *inter = malloc((size_t)v->source_size * (size_t)ist * sizeof(float));
    for (int32_t r = 0; r < v->source_size; ++r)
        orc_axis_f32_ordered(h, in + (ptrdiff_t)r * is, 1, inter + (ptrdiff_t)r * ist, 1);
    for (int32_t c = 0; c < h->destination_size; ++c)
        orc_axis_f32_ordered(v, inter + c, ist, out + c, os);
    free(inter);
    return ORC_OK;
}

static int conversion_valid(const orc_conversion *c, uint32_t sample_max) {
    return c->input_scale > 0.0F && c->output_scale > 0.0F
        && isfinite(c->input_offset) && isfinite(c->input_scale)
        && isfinite(c->output_scale) && isfinite(c->output_offset)
        && c->output_maximum != 0U && c->output_maximum <= sample_max;
}

/*
 * CpuExecutor::inverse_2d_integer scalar (src/cpu_executor.cpp:2051-2257):
 * normalise each source row ((float)x - offset) * scale, horizontal
 * inverse_axis_f32, vertical RHS accumulated in source order (the same
 * order as the CSR gather, so the two forms are identical), solve_rhs_scalar
 * per column, then convert_rhs_to_integer_scalar (1918-1937).
 */
static int twod_integer(const orc_plan *h, const orc_plan *v, const void *in,
                        ptrdiff_t is, void *out, ptrdiff_t os,
                        const orc_conversion *conv, int kind) {
    const uint32_t sample_max = kind == 1 ? 255U : 65535U;
    if (!in || !out || is < h->source_size || os < h->destination_size
        || v->source_size < 1 || v->destination_size < 1
        || !conversion_valid(conv, sample_max))
        return fail(ORC_INVALID, "invalid integer 2D executor arguments");
    if (h->requires_float64 || v->requires_float64)
        return twod_f64(h, v, in, is, out, os, conv, kind);
    const int32_t ist = h->destination_size;
    float *row = malloc((size_t)h->source_size * sizeof(float));
    float *inter = malloc((size_t)v->source_size * (size_t)ist * sizeof(float));
    float *result = malloc((size_t)v->destination_size * (size_t)ist * sizeof(float));
    for (int32_t r = 0; r < v->source_size; ++r) {
        for (int32_t c = 0; c < h->source_size; ++c) {
            const ptrdiff_t at = (ptrdiff_t)r * is + c;
            const float x = kind == 1 ? (float)((const uint8_t *)in)[at]
                                      : (float)((const uint16_t *)in)[at];
            row[c] = (x - conv->input_offset) * conv->input_scale;
        }
        axis_f32_production(h, row, 1, inter + (ptrdiff_t)r * ist, 1);
    }
    for (int32_t c = 0; c < ist; ++c)
        axis_f32_production(v, inter + c, ist, result + c, ist);
    const float maximum = (float)conv->output_maximum;
    for (int32_t r = 0; r < v->destination_size; ++r) {
        for (int32_t c = 0; c < ist; ++c) {
            float scaled = result[(ptrdiff_t)r * ist + c] * conv->output_scale
                + conv->output_offset;
            /* std::clamp(v, 0, max): NaN passes through unchanged. */
            if (scaled < 0.0F) scaled = 0.0F;
            else if (maximum < scaled) scaled = maximum;
            const float rounded = nearbyintf(scaled);
            const ptrdiff_t at = (ptrdiff_t)r * os + c;
            if (kind == 1) ((uint8_t *)out)[at] = (uint8_t)rounded;
            else ((uint16_t *)out)[at] = (uint16_t)rounded;
        }
    }
    free(row);
    free(inter);
    free(result);
    return ORC_OK;
}

int orc_2d_u8(const orc_plan *h, const orc_plan *v, const uint8_t *in,
              ptrdiff_t is, uint8_t *out, ptrdiff_t os, const orc_conversion *conv) {
    return twod_integer(h, v, in, is, out, os, conv, 1);
}

int orc_2d_u16(const orc_plan *h, const orc_plan *v, const uint16_t *in,
               ptrdiff_t is, uint16_t *out, ptrdiff_t os, const orc_conversion *conv) {
    return twod_integer(h, v, in, is, out, os, conv, 2);
}

int orc_2d_f64_exact(const orc_plan *h, const orc_plan *v, const double *in,
                     ptrdiff_t is, double *out, ptrdiff_t os) {
    if (!h->requires_float64 || !v->requires_float64)
        return fail(ORC_INVALID, "the FP64 reference needs forced-F64 plans");
    const int32_t ist = h->destination_size;
    double *inter = malloc((size_t)v->source_size * (size_t)ist * sizeof(double));
    for (int32_t r = 0; r < v->source_size; ++r)
        orc_axis_f64(h, in + (ptrdiff_t)r * is, 1, inter + (ptrdiff_t)r * ist, 1);
    for (int32_t c = 0; c < ist; ++c) orc_axis_f64(v, inter + c, ist, out + c, os);
    free(inter);
    return ORC_OK;
}

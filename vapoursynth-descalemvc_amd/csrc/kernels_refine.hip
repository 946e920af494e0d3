// Iterative refinement of F32 axis solves for plans that retain FP64 data.
//
// The F32 route solves G x = A^T y with F32 factors; its error grows with
// the plan's condition number.  One refinement step computes the residual
// of the normal system,
//     r = A^T y - G x   (= A^T (y - A x)),
// in compensated F32: every product w*v with an FP64 coefficient w split as
// w_hi + w_lo is formed exactly as p + e by a two-product (e = fma(w_hi, v,
// -p)), the running sum keeps its rounding error by a two-sum, and w_lo*v
// joins the error term.  The difference of the two sums is rounded to F32
// once.  Then G d = r is solved with the F32 factors and d is added to x.
// The FP64 route (reference src/axis_plan.cpp:1124-1248) solves the same
// system in double; refinement converges to its solution at F32 rounding.
//
// Lanes own vectors (one lane per vector, like the sequential solve); the
// band width is a runtime value.  This is the precision route, not the
// throughput route.
#include "kernels_common.hpp"

namespace dsmvc::hip {
namespace dev {
namespace {

// s + e == a + b exactly (Knuth's two-sum).
__device__ __forceinline__ void two_sum(float a, float b, float &s, float &e) {
    s = __fadd_rn(a, b);
    const float bb = __fsub_rn(s, a);
    e = __fadd_rn(__fsub_rn(a, __fsub_rn(s, bb)), __fsub_rn(b, bb));
}

// (sum, err) += (hi + lo) * v, compensated.
__device__ __forceinline__ void accumulate(float &sum, float &err, float hi, float lo, float v) {
    const float p = __fmul_rn(hi, v);
    const float pe = __builtin_fmaf(hi, v, -p);
    float s, e;
    two_sum(sum, p, s, e);
    sum = s;
    err = __fadd_rn(err, __fadd_rn(__fadd_rn(e, pe), __fmul_rn(lo, v)));
}

__global__ void __launch_bounds__(256) normal_residual(RefineArgs a) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nvec) return;
    const long f = blockIdx.y;
    const float *y = a.y + f * a.y_frame + v * a.y_vec;
    const float *x = a.x + f * a.x_frame + v * a.x_vec;
    const int n = a.axis.n, w = a.refine.w, span = 2 * a.refine.w + 1;
    float *r = a.scratch + f * static_cast<long>(n) * a.nvec + v;
    for (int i = 0; i < n; ++i) {
        float bs = 0.0F, be = 0.0F;
        for (std::uint32_t q = a.axis.off[i]; q < a.axis.off[i + 1]; ++q)
            accumulate(bs, be, a.refine.w_hi[q], a.refine.w_lo[q], y[static_cast<long>(a.axis.idx[q]) * a.y_el]);
        float gs = 0.0F, ge = 0.0F;
        const float *gh = a.refine.g_hi + static_cast<long>(i) * span;
        const float *gl = a.refine.g_lo + static_cast<long>(i) * span;
        const int d0 = i - w < 0 ? w - i : 0;
        const int d1 = i + w >= n ? w + (n - 1 - i) : 2 * w;
        for (int d = d0; d <= d1; ++d)
            accumulate(gs, ge, gh[d], gl[d], x[static_cast<long>(i - w + d) * a.x_el]);
        float s, e;
        two_sum(bs, -gs, s, e);
        r[static_cast<long>(i) * a.nvec] = __fadd_rn(s, __fadd_rn(e, __fsub_rn(be, ge)));
    }
}

// Solves G d = r with the F32 factors (the residual is overwritten by the
// forward sweep) and adds d to x.
__global__ void __launch_bounds__(256) refine_correct(RefineArgs a) {
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= a.nvec) return;
    const long f = blockIdx.y;
    float *x = a.x + f * a.x_frame + v * a.x_vec;
    const int n = a.axis.n, w = a.axis.w;
    const int stride = coef_stride(w);
    const int slots = (stride - 2) / 2;
    float *r = a.scratch + f * static_cast<long>(n) * a.nvec + v;
    const long step = a.nvec;
    for (int i = 0; i < n; ++i) {
        const float *c = a.axis.coef + static_cast<long>(i) * stride;
        float acc = r[i * step];
        for (int d = w; d >= 1; --d)
            if (i - d >= 0) acc = __builtin_fmaf(-c[d - 1], r[(i - d) * step], acc);
        r[i * step] = __fmul_rn(acc, c[slots]);
    }
    for (int i = n - 1; i >= 0; --i) {
        const float *c = a.axis.coef + static_cast<long>(i) * stride + slots + 1;
        float acc = r[i * step];
        for (int d = w; d >= 1; --d)
            if (i + d < n) acc = __builtin_fmaf(-c[d - 1], r[(i + d) * step], acc);
        r[i * step] = acc;
        x[static_cast<long>(i) * a.x_el] = __fadd_rn(x[static_cast<long>(i) * a.x_el], acc);
    }
}

} // namespace
} // namespace dev

hipError_t launch_refine_step(const RefineArgs &a, hipStream_t s) {
    if (a.nvec <= 0 || a.frames <= 0) return hipSuccess;
    const dim3 grid((a.nvec + 255) / 256, a.frames);
    dev::normal_residual<<<grid, 256, 0, s>>>(a);
    dev::refine_correct<<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

} // namespace dsmvc::hip

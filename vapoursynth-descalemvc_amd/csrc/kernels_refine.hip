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
// Two forms: the residual element by element in FP64 with the correction
// solved by the windowed kernels (launch_refine_residual / launch_refine_add,
// used when the plan's windowed tables exist), or one lane per vector walking
// the whole vector for both, with the compensated F32 residual above
// (launch_refine_step).  The band width is a runtime value.
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

// The residual element by element.  r is laid out [frame][vector][element]
// for a rows pass and [frame][element][vector] for a columns pass.  The sums
// run in FP64 over the tap-major FP64 tables (the FP64 vector rate of this GPU,
// 78.6 TFLOP/s, is half its FP32 rate, ample for a pass bound by memory), and
// r is rounded to F32 once.  Indices past a row's
// samples are clamped; their weights are 0, so the terms are exactly 0 (the
// inputs are finite).
//
// Rows pass: a block covers 256 consecutive elements of V vectors; each
// thread sums one element of V vectors, so one coefficient load serves V
// sums.  Elements run along the lanes, so the tap-major tables are read
// coalesced.  The block first copies the sample spans its elements read
// (source samples [min first, max first + k), solution elements i0 - w ..
// i0 + 255 + w, both clamped to the vector) of its V vectors into LDS, so
// every sample leaves HBM once per block; a block whose source span does not
// fit the LDS rows reads its samples from global memory instead.
template <int V>
__global__ void __launch_bounds__(256) residual_rows(RefineArgs a, float *r) {
    constexpr int kCapY = 640, kCapX = 256 + 2 * 15 + 2;
    __shared__ float ys[V][kCapY];
    __shared__ float xs[V][kCapX];
    __shared__ int bounds[2][4];
    const int n = a.axis.n, w = a.refine.w, k = a.refine.k, span = 2 * w + 1;
    const int iblocks = (n + 255) / 256;
    const int i0 = static_cast<int>(blockIdx.x % iblocks) * 256;
    const int tid = static_cast<int>(threadIdx.x);
    const int i = i0 + tid;
    const int v0 = static_cast<int>(blockIdx.x / iblocks) * V;
    const long f = blockIdx.y;
    const float *y[V];
    const float *x[V];
#pragma unroll
    for (int u = 0; u < V; ++u) {
        const long v = min(v0 + u, a.nvec - 1);
        y[u] = a.y + f * a.y_frame + v * a.y_vec;
        x[u] = a.x + f * a.x_frame + v * a.x_vec;
    }
    const int m1 = a.refine.m - 1;
    const int j0 = a.refine.first[min(i, n - 1)];
    // The block's source span: min and max of first[] over its elements.
    int lo = j0, hi = j0;
    for (int off = 32; off >= 1; off >>= 1) {
        lo = min(lo, __shfl_xor(lo, off));
        hi = max(hi, __shfl_xor(hi, off));
    }
    if ((tid & 63) == 0) {
        bounds[0][tid >> 6] = lo;
        bounds[1][tid >> 6] = hi;
    }
    __syncthreads();
    lo = min(min(bounds[0][0], bounds[0][1]), min(bounds[0][2], bounds[0][3]));
    hi = max(max(bounds[1][0], bounds[1][1]), max(bounds[1][2], bounds[1][3])) + k;
    const bool staged = hi - lo <= kCapY && 256 + 2 * w <= kCapX;
    if (staged) {
        const int ylen = hi - lo, xlen = 256 + 2 * w;
#pragma unroll
        for (int u = 0; u < V; ++u) {
            for (int e = tid; e < ylen; e += 256) ys[u][e] = y[u][min(lo + e, m1)];
            for (int e = tid; e < xlen; e += 256) xs[u][e] = x[u][min(max(i0 - w + e, 0), n - 1)];
        }
        __syncthreads();
    }
    if (i >= n) return;
    const double *t = a.refine.t64 + i;
    const double *g = a.refine.b64 + i;
    double s[V];
#pragma unroll
    for (int u = 0; u < V; ++u) s[u] = 0.0;
    if (staged) {
        const int jb = j0 - lo;
#pragma unroll 2
        for (int q = 0; q < k; ++q) {
            const double c = t[static_cast<long>(q) * n];
#pragma unroll
            for (int u = 0; u < V; ++u) s[u] = __builtin_fma(c, static_cast<double>(ys[u][jb + q]), s[u]);
        }
#pragma unroll 2
        for (int d = 0; d < span; ++d) {
            const double c = -g[static_cast<long>(d) * n];
#pragma unroll
            for (int u = 0; u < V; ++u) s[u] = __builtin_fma(c, static_cast<double>(xs[u][tid + d]), s[u]);
        }
    } else {
#pragma unroll 2
        for (int q = 0; q < k; ++q) {
            const double c = t[static_cast<long>(q) * n];
            const int j = min(j0 + q, m1);
#pragma unroll
            for (int u = 0; u < V; ++u) s[u] = __builtin_fma(c, static_cast<double>(y[u][j]), s[u]);
        }
#pragma unroll 2
        for (int d = 0; d < span; ++d) {
            const double c = -g[static_cast<long>(d) * n];
            const int j = min(max(i - w + d, 0), n - 1);
#pragma unroll
            for (int u = 0; u < V; ++u) s[u] = __builtin_fma(c, static_cast<double>(x[u][j]), s[u]);
        }
    }
    const long per = static_cast<long>(n) * a.nvec;
#pragma unroll
    for (int u = 0; u < V; ++u)
        if (v0 + u < a.nvec)
            r[f * per + static_cast<long>(v0 + u) * n + i] = static_cast<float>(s[u]);
}

// Columns pass: the residual of element i for one vector (lanes = adjacent
// vectors), coefficients as workgroup-uniform scalar loads.  Two FMA chains
// keep loads in flight.
__device__ __forceinline__ double column_residual(const RefineArgs &a, int i, const float *y,
                                                  const float *x) {
    const int n = a.axis.n, w = a.refine.w;
    const int j0 = a.refine.first[i], m1 = a.refine.m - 1;
    const double *t = a.refine.t64 + i;
    const double *g = a.refine.b64 + i;
    const int k = a.refine.k, span = 2 * w + 1;
    double s0 = 0.0, s1 = 0.0;
    int q = 0;
#pragma unroll 2
    for (; q + 1 < k; q += 2) {
        s0 = __builtin_fma(t[static_cast<long>(q) * n],
                           static_cast<double>(y[min(j0 + q, m1) * a.y_el]), s0);
        s1 = __builtin_fma(t[static_cast<long>(q + 1) * n],
                           static_cast<double>(y[min(j0 + q + 1, m1) * a.y_el]), s1);
    }
    if (q < k)
        s0 = __builtin_fma(t[static_cast<long>(q) * n], static_cast<double>(y[min(j0 + q, m1) * a.y_el]), s0);
    int d = 0;
#pragma unroll 2
    for (; d + 1 < span; d += 2) {
        s0 = __builtin_fma(-g[static_cast<long>(d) * n],
                           static_cast<double>(x[min(max(i - w + d, 0), n - 1) * a.x_el]), s0);
        s1 = __builtin_fma(-g[static_cast<long>(d + 1) * n],
                           static_cast<double>(x[min(max(i - w + d + 1, 0), n - 1) * a.x_el]), s1);
    }
    if (d < span)
        s0 = __builtin_fma(-g[static_cast<long>(d) * n],
                           static_cast<double>(x[min(max(i - w + d, 0), n - 1) * a.x_el]), s0);
    return s0 + s1;
}

// Columns pass: a block covers I consecutive elements of 256 adjacent vectors.
// Neighbouring elements' sample runs overlap almost entirely, so the block
// first writes its coefficients to LDS as dense rows over the union of the
// runs (0 where an element has no weight), then each thread reads every
// sample row of the union once and adds it to all I sums.  A block whose
// union does not fit the LDS rows sums element by element instead.
template <int I>
__global__ void __launch_bounds__(256) residual_columns(RefineArgs a, float *r) {
    constexpr int kCapY = 64, kCapX = 2 * 15 + I + 2;
    __shared__ double cy[I][kCapY];
    __shared__ double cx[I][kCapX];
    const int n = a.axis.n, w = a.refine.w, k = a.refine.k, span = 2 * w + 1;
    const int vblocks = (a.nvec + 255) / 256;
    const int i0 = static_cast<int>(blockIdx.x) / vblocks * I;
    const int v = static_cast<int>(blockIdx.x % vblocks) * 256 + static_cast<int>(threadIdx.x);
    const long f = blockIdx.y;
    const float *y = a.y + f * a.y_frame + min(v, a.nvec - 1);
    const float *x = a.x + f * a.x_frame + min(v, a.nvec - 1);
    const long plane = static_cast<long>(n) * a.nvec;
    int i[I], j0[I];
    int lo = 0x7fffffff, hi = -0x7fffffff;
#pragma unroll
    for (int u = 0; u < I; ++u) {
        i[u] = min(i0 + u, n - 1);
        j0[u] = a.refine.first[i[u]];
        lo = min(lo, j0[u]);
        hi = max(hi, j0[u] + k);
    }
    const int ylen = hi - lo, xlen = i[I - 1] - i[0] + span;
    if (ylen > kCapY || xlen > kCapX) {
        if (v >= a.nvec) return;
#pragma unroll
        for (int u = 0; u < I; ++u)
            if (i0 + u < n)
                r[f * plane + static_cast<long>(i0 + u) * a.nvec + v]
                    = static_cast<float>(column_residual(a, i[u], y, x));
        return;
    }
    for (int e = static_cast<int>(threadIdx.x); e < I * kCapY; e += 256) {
        const int u = e / kCapY, jj = e % kCapY;
        const int q = lo + jj - j0[u];
        cy[u][jj] = (jj < ylen && q >= 0 && q < k) ? a.refine.t64[static_cast<long>(q) * n + i[u]] : 0.0;
    }
    for (int e = static_cast<int>(threadIdx.x); e < I * kCapX; e += 256) {
        const int u = e / kCapX, dd = e % kCapX;
        const int d = dd - (i[u] - i[0]);
        cx[u][dd] = (dd < xlen && d >= 0 && d < span) ? -a.refine.b64[static_cast<long>(d) * n + i[u]] : 0.0;
    }
    __syncthreads();
    if (v >= a.nvec) return;
    double s[I];
#pragma unroll
    for (int u = 0; u < I; ++u) s[u] = 0.0;
    const int m1 = a.refine.m - 1;
#pragma unroll 4
    for (int jj = 0; jj < ylen; ++jj) {
        const double yv = static_cast<double>(y[min(lo + jj, m1) * a.y_el]);
#pragma unroll
        for (int u = 0; u < I; ++u) s[u] = __builtin_fma(cy[u][jj], yv, s[u]);
    }
#pragma unroll 4
    for (int dd = 0; dd < xlen; ++dd) {
        const double xv = static_cast<double>(x[min(max(i[0] - w + dd, 0), n - 1) * a.x_el]);
#pragma unroll
        for (int u = 0; u < I; ++u) s[u] = __builtin_fma(cx[u][dd], xv, s[u]);
    }
#pragma unroll
    for (int u = 0; u < I; ++u)
        if (i0 + u < n) r[f * plane + static_cast<long>(i0 + u) * a.nvec + v] = static_cast<float>(s[u]);
}

// x += d, d laid out as the residual kernels lay out r.  One thread per
// element, threads along the contiguous dimension; strides are fixed per
// pass and a plane is indexed in 32 bits.
template <bool Rows>
__global__ void __launch_bounds__(256) add_correction(RefineArgs a, const float *d) {
    const unsigned per = static_cast<unsigned>(a.axis.n) * static_cast<unsigned>(a.nvec);
    const unsigned idx = blockIdx.x * 256U + threadIdx.x;
    if (idx >= per) return;
    const long f = blockIdx.y;
    const auto un = static_cast<unsigned>(a.axis.n), uv = static_cast<unsigned>(a.nvec);
    const long v = Rows ? idx / un : idx % uv;
    const long i = Rows ? idx % un : idx / uv;
    float *x = a.x + f * a.x_frame + (Rows ? v * a.x_vec + i : i * a.x_el + v);
    *x = __fadd_rn(*x, d[f * static_cast<long>(per) + idx]);
}

} // namespace
} // namespace dev

hipError_t launch_refine_residual(const RefineArgs &a, float *r, bool rows, hipStream_t s) {
    const long per = static_cast<long>(a.axis.n) * a.nvec;
    if (per <= 0 || a.frames <= 0) return hipSuccess;
    constexpr int V = 4, I = 4;
    const long blocks = rows ? static_cast<long>((a.axis.n + 255) / 256) * ((a.nvec + V - 1) / V)
                             : static_cast<long>((a.axis.n + I - 1) / I) * ((a.nvec + 255) / 256);
    if (blocks >= (1L << 31)) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(blocks), a.frames);
    if (rows) dev::residual_rows<V><<<grid, 256, 0, s>>>(a, r);
    else dev::residual_columns<I><<<grid, 256, 0, s>>>(a, r);
    return hipGetLastError();
}

hipError_t launch_refine_add(const RefineArgs &a, const float *d, bool rows, hipStream_t s) {
    const long per = static_cast<long>(a.axis.n) * a.nvec;
    if (per <= 0 || a.frames <= 0) return hipSuccess;
    if (per >= (1L << 31)) return hipErrorInvalidValue;   // the kernel indexes a plane in 32 bits
    const dim3 grid(static_cast<unsigned>((per + 255) / 256), a.frames);
    if (rows) dev::add_correction<true><<<grid, 256, 0, s>>>(a, d);
    else dev::add_correction<false><<<grid, 256, 0, s>>>(a, d);
    return hipGetLastError();
}

hipError_t launch_refine_step(const RefineArgs &a, hipStream_t s) {
    if (a.nvec <= 0 || a.frames <= 0) return hipSuccess;
    const dim3 grid((a.nvec + 255) / 256, a.frames);
    dev::normal_residual<<<grid, 256, 0, s>>>(a);
    dev::refine_correct<<<grid, 256, 0, s>>>(a);
    return hipGetLastError();
}

} // namespace dsmvc::hip

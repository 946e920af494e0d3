// Launch interface of the gfx950 kernels (kernels.hip).  Host code builds the
// argument blocks; every launcher returns the launch's hipError_t.
#pragma once

#include <hip/hip_runtime.h>

#include "device_types.hpp"

namespace dsmvc::hip {

// Sequential solve of `nvec` vectors per frame.  Element e of vector v of
// frame f lives at base + f * frame + v * vec + e * el (all in elements).
// `fwd` holds the forward sweep (type of the compute precision).
struct SeqArgs {
    DevAxis axis;
    const void *src;
    long s_vec, s_el, s_frame;
    void *dst;
    long d_vec, d_el, d_frame;
    void *fwd;
    long f_vec, f_el, f_frame;
    int nvec;
    int frames;
    DevConv conv;
    int *flag;   // set to 1 by the Float64 route on a non-finite value
};

// Chunked solve: rows (horizontal pass, vectors are rows) or columns
// (vertical pass, vectors are columns) of `frames` planes.
struct ChunkArgs {
    DevAxis axis;
    DevChunks chunks;
    const void *src;
    long s_row, s_frame;
    void *dst;
    long d_row, d_frame;
    int nvec;
    int frames;
    DevConv conv;
};

bool seq_band_supported(int w);

// Chunked kernels: bands 1, 3, 5, 7; chunk lengths 32, 48, 64, 80; packed
// tap counts 8, 16; at most 16 chunks.
bool chunked_band_supported(int w);
int chunked_length_for(int needed);   // smallest supported length >= needed, or 0
int chunked_width_for(int taps);      // smallest supported tap count >= taps, or 0
constexpr int kMaxChunks = 16;

hipError_t launch_seq(const SeqArgs &a, Sample in, Sample out, bool f64, hipStream_t s);
// rows: horizontal pass, `kind` is the input sample; else vertical pass,
// `kind` is the output sample.
hipError_t launch_chunked(const ChunkArgs &a, bool rows, Sample kind, hipStream_t s);
template <int W>
hipError_t launch_chunked_band(const ChunkArgs &a, bool rows, Sample kind, hipStream_t s);
extern template hipError_t launch_chunked_band<1>(const ChunkArgs &, bool, Sample, hipStream_t);
extern template hipError_t launch_chunked_band<3>(const ChunkArgs &, bool, Sample, hipStream_t);
extern template hipError_t launch_chunked_band<5>(const ChunkArgs &, bool, Sample, hipStream_t);
extern template hipError_t launch_chunked_band<7>(const ChunkArgs &, bool, Sample, hipStream_t);
// Windowed (independent-chunk) kernels.  The band class is 1, 3, 5 or 7
// (smaller bands run in the next class with zero coefficients); the tap
// width 6 (band classes 1 and 3), 8 or 16.  `launch_window_table` runs
// `count` jobs from a device-resident table (`items` items in all) in one
// launch, all jobs sharing the band class, tap width, direction, staging
// block size and sample kind (the input kind of a horizontal pass, the
// output kind of a vertical pass).  A single pass is a one-job table.
int window_band_class(int w);   // 1, 3, 5, 7, or 0 when unsupported
constexpr int kWindowWaves = 4;   // wavefronts (one item each) per workgroup
// `order`: bit 0 segment-fastest item order, bit 1 XCD-contiguous workgroups.
hipError_t launch_window_table(const WinJob *jobs, int count, long items, int band, int width,
                               bool rows, bool block16, Sample kind, int order, hipStream_t s);
// Wavefronts of that launch's kernel the GPU holds at once.
long window_table_slots(int band, int width, bool rows, bool block16, Sample kind);
// Sets the span (chunks per wavefront) and first item of a launch's `count`
// jobs for `slots` resident wavefronts; returns the launch's item count.
long assign_window_spans(WinJob *jobs, int count, long slots);
template <int W, int K>
hipError_t launch_window_table_band(const WinJob *jobs, int count, long items, bool rows,
                                    bool block16, Sample kind, int order, hipStream_t s);
template <int W, int K> long window_table_slots_band(bool rows, bool block16, Sample kind);
#define DSMVC_WINDOW_EXTERN(W, K)                                                                 \
    extern template hipError_t launch_window_table_band<W, K>(const WinJob *, int, long, bool,    \
                                                              bool, Sample, int, hipStream_t);    \
    extern template long window_table_slots_band<W, K>(bool, bool, Sample);
DSMVC_WINDOW_EXTERN(1, 6)
DSMVC_WINDOW_EXTERN(3, 6)
DSMVC_WINDOW_EXTERN(1, 8)
DSMVC_WINDOW_EXTERN(1, 16)
DSMVC_WINDOW_EXTERN(3, 8)
DSMVC_WINDOW_EXTERN(3, 16)
DSMVC_WINDOW_EXTERN(5, 8)
DSMVC_WINDOW_EXTERN(5, 16)
DSMVC_WINDOW_EXTERN(7, 8)
DSMVC_WINDOW_EXTERN(7, 16)
#undef DSMVC_WINDOW_EXTERN

// Windowed FP64 pass of the Float64 route (kernels_window64.hpp): rows
// F32/u8/u16 source -> FP64 intermediate, columns FP64 intermediate ->
// F32/u8/u16, instantiated per band class in kernels_window64_w{1,3,5,7}.hip.
// `kind`: sample type of the rows pass's source or the columns pass's output.
// launch_window64 runs one pass (its job passed by value; a non-finite
// stored value sets a.flag); the table form runs `count` jobs from a device
// table, `items` items in all, every job of one band class, tap width,
// direction and sample kind (a non-finite stored value sets its job's flag).
hipError_t launch_window64(const Win64Args &a, int band, int width, bool rows, Sample kind,
                           hipStream_t s);
hipError_t launch_window64_table(const Win64Job *jobs, int count, long items, int band, int width,
                                 bool rows, Sample kind, hipStream_t s);
// Wavefronts of an FP64 launch's kernel the GPU holds at once, and the span
// (chunks per wavefront) and first items of a launch's jobs for that many
// resident wavefronts; returns the launch's item count.
long window64_table_slots(int band, int width, bool rows, Sample kind);
long assign_window64_spans(Win64Job *jobs, int count, long slots);
template <int W>
hipError_t launch_window64_table_band(const Win64Job *jobs, int count, long items,
                                      const Win64Job &one, int width, bool rows, Sample kind,
                                      hipStream_t s);
template <int W> long window64_table_slots_band(int width, bool rows, Sample kind);
#define DSMVC_WINDOW64_EXTERN(W)                                                                 \
    extern template hipError_t launch_window64_table_band<W>(                                     \
        const Win64Job *, int, long, const Win64Job &, int, bool, Sample, hipStream_t);           \
    extern template long window64_table_slots_band<W>(int, bool, Sample);
DSMVC_WINDOW64_EXTERN(1)
DSMVC_WINDOW64_EXTERN(3)
DSMVC_WINDOW64_EXTERN(5)
DSMVC_WINDOW64_EXTERN(7)
#undef DSMVC_WINDOW64_EXTERN
// The Float64 route's input check over `count` F32 plane batches (items =
// frames x rows in all): a non-finite sample sets its job's flag.
hipError_t launch_check_finite_table(const CheckJob *jobs, int count, long items, hipStream_t s);

// Refinement of an F32 axis solve x (in place): `steps` times,
//   r = A^T y - G x   in compensated F32 (two-product and two-sum, with the
//                     FP64 weights and normal matrix split into F32 pairs),
//   solve G d = r     with the plan's F32 factors,
//   x += d.
// y: source vectors (element e of vector v of frame f at
// y + f*y_frame + v*y_vec + e*y_el), x likewise; `scratch` holds
// nvec * n * frames floats.
struct RefineArgs {
    DevAxis axis;
    DevRefine refine;
    const float *y;
    long y_vec, y_el, y_frame;
    float *x;
    long x_vec, x_el, x_frame;
    float *scratch;
    int nvec;
    int frames;
};
hipError_t launch_refine_step(const RefineArgs &a, hipStream_t s);
// The residual r (n * nvec * frames floats) element by element, laid out
// [frame][vector][element] for a rows pass and [frame][element][vector] for a
// columns pass; then x += d for a correction d laid out the same way.
hipError_t launch_refine_residual(const RefineArgs &a, float *r, bool rows, hipStream_t s);
hipError_t launch_refine_add(const RefineArgs &a, const float *d, bool rows, hipStream_t s);

hipError_t launch_check_finite(const void *src, Sample in, long s_row, long s_frame, int rows,
                               int cols, int frames, int *flag, hipStream_t s);

// Fused 2D route (kernels_operator.hip): one launch over `count` jobs of one
// sample kind whose operator images were packed at `steps` K steps (2..4)
// (FP16 parts).  `exact`: the F32 jobs' second pass on bfloat16 images,
// which returns at once unless the first pass set the jobs' flag.
// `vector`: every job's source base and strides allow 16-byte loads (8 for
// u8 samples), i.e. each has vector_loads set.  `shared_flag` (exact pass):
// the flag every job of the table uses, or nullptr when they differ.
hipError_t launch_operator_table(const OpJob *table, int count, long items, int steps, Sample kind,
                                 bool exact, bool vector, const int *shared_flag, hipStream_t stream);

// p[0 .. n) = value, as a kernel (a graph may capture it; no memset node).
hipError_t launch_fill_i32(int *p, int n, int value, hipStream_t stream);

} // namespace dsmvc::hip

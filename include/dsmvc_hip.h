/*
 * dsmvc_hip.h -- C ABI of the MI355X (gfx950) descale executor.
 *
 * The drop-in boundary for `backend="hip"`: plain C types, pointers and sizes,
 * int status codes, no exceptions and no C++/torch types across it.  Every
 * entry point names the reference interface it replaces (paths relative to
 * the reference repository, AmusementClub/vapoursynth-descalemvc v0.1.4).
 *
 * Buffers passed to the inverse_* calls are DEVICE pointers on the
 * executor's GPU.  Strides are in elements (samples), as in the reference's
 * buffer contract (include/dsmvc/engine.hpp:231-235): a stride covers the
 * logical width, only the logical width of each output row is written, and
 * the last row needs no pitch padding.  Calls are synchronous -- they return
 * after the output is written, like the reference's executors
 * (src/cuda/cuda_executor.cpp:2968-2990) -- unless DSMVC_HIP_ASYNC is set, in
 * which case work is queued on `stream` and dsmvc_hip_synchronize() reports
 * any deferred Float64 non-finite failure.  `stream` is a hipStream_t (NULL:
 * the null stream).
 */
#ifndef DSMVC_HIP_H
#define DSMVC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  Messages: dsmvc_hip_last_error() (thread-local). */
#define DSMVC_HIP_OK 0
#define DSMVC_HIP_INVALID_ARGUMENT 1 /* reference: std::invalid_argument */
#define DSMVC_HIP_RUNTIME_ERROR 2    /* reference: std::runtime_error   */
#define DSMVC_HIP_UNAVAILABLE 3      /* no HIP device                   */

/* Flags of the inverse_* calls. */
#define DSMVC_HIP_ASYNC 1u

/* Kernel kinds, border (padding) modes, f64mode values and backends keep the
 * reference numbering (include/dsmvc/engine.hpp:16-58); backend 5 is hip. */
enum { DSMVC_HIP_BILINEAR = 0, DSMVC_HIP_BICUBIC = 1, DSMVC_HIP_LANCZOS = 2,
       DSMVC_HIP_SPLINE16 = 3, DSMVC_HIP_SPLINE36 = 4, DSMVC_HIP_SPLINE64 = 5,
       DSMVC_HIP_CUSTOM = 6 };
enum { DSMVC_HIP_SOLVE_AUTO = 0, DSMVC_HIP_SOLVE_EXACT = 1, DSMVC_HIP_SOLVE_CHUNKED = 2,
       DSMVC_HIP_SOLVE_WINDOWED = 3, DSMVC_HIP_SOLVE_REFINED = 4, DSMVC_HIP_SOLVE_FUSED = 5 };
enum { DSMVC_HIP_SAMPLE_F32 = 0, DSMVC_HIP_SAMPLE_U8 = 1, DSMVC_HIP_SAMPLE_U16 = 2 };

/* dsmvc::AxisRequest with KernelSpec flattened (include/dsmvc/engine.hpp:60-78). */
typedef struct dsmvc_hip_axis_request {
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
} dsmvc_hip_axis_request;

/* dsmvc::IntegerConversion (include/dsmvc/engine.hpp:167-173). */
typedef struct dsmvc_hip_integer_conversion {
    float input_offset;
    float input_scale;
    float output_scale;
    float output_offset;
    uint32_t output_maximum;
} dsmvc_hip_integer_conversion;

/* Scalar fields of dsmvc::AxisPlan (include/dsmvc/engine.hpp:82-107). */
typedef struct dsmvc_hip_plan_info {
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
} dsmvc_hip_plan_info;

/* Windowed partition of a plan (see dsmvc_hip_plan_window_info). */
typedef struct dsmvc_hip_window_info {
    int32_t feasible;        /* 1: the windowed kernels can solve this plan */
    int32_t rows_feasible;   /* 1: also in the horizontal pass (staging fits) */
    int32_t band;            /* band class the kernels run: 1, 3, 5 or 7 */
    int32_t chunks;
    int32_t length;          /* destination rows per chunk (L) */
    int32_t warmup;          /* warm-up rows on each side (P) */
    int32_t width;           /* packed taps per row: 8 or 16 */
    int32_t float64_feasible; /* 1: the windowed FP64 kernels take this plan on
                                 the Float64 route of F32 planes (2^-40 check) */
    double residual;         /* worst influence of a chunk's zero start, relative */
} dsmvc_hip_window_info;

/* Explicit operator of a plan for the fused 2D route (see
 * dsmvc_hip_plan_operator). */
typedef struct dsmvc_hip_operator_info {
    int32_t rows_feasible;    /* 1: the fused route can take the plan horizontally */
    int32_t columns_feasible; /* 1: ... vertically */
    int32_t blocks;           /* blocks of 16 destination indices (padded) */
    int32_t steps;            /* K steps of 32 source indices per block */
    int32_t max_band;         /* longest kept row */
    int32_t reserved;
    int64_t values;           /* kept entries over all rows */
    double dropped;           /* largest L1 share truncation dropped from a row */
    double max_abs;
} dsmvc_hip_operator_info;

typedef struct dsmvc_hip_planner_stats {
    uint64_t plan_hits;
    uint64_t plan_builds;
    uint64_t plan_entries;
    uint64_t plan_resident_bytes;
    uint64_t geometry_hits;      /* sampling geometries reused (any kernel or blur of equal support) */
    uint64_t geometry_builds;
    uint64_t geometry_entries;
    uint64_t geometry_resident_bytes;
} dsmvc_hip_planner_stats;

typedef struct dsmvc_hip_plan dsmvc_hip_plan;         /* immutable axis plan */
typedef struct dsmvc_hip_executor dsmvc_hip_executor; /* one GPU            */

const char *dsmvc_hip_last_error(void);

/* ---- planner ------------------------------------------------------------ */
/* Replaces dsmvc::get_or_build_axis_plan (include/dsmvc/engine.hpp:177-178,
 * src/axis_plan.cpp:1367-1383): single-flight LRU keyed by the request's bits. */
int dsmvc_hip_plan_build(const dsmvc_hip_axis_request *request, dsmvc_hip_plan **plan);
/* No single counterpart in the reference, which plans one request at a time
 * (src/axis_plan.cpp:1367-1383); for scans over many geometries.  Builds
 * `count` plans through the same cache on up to `threads` host threads
 * (<= 0: one per hardware thread, at most 16).  Every request is validated
 * before any is built.  On failure no handle is returned (all of `plans` are
 * NULL) and the error is that of the lowest failing index.  Each plan equals
 * what dsmvc_hip_plan_build returns for its request. */
int dsmvc_hip_plan_build_many(const dsmvc_hip_axis_request *requests, int32_t count,
                              int32_t threads, dsmvc_hip_plan **plans);
/* dsmvc::get_or_build_axis_plan(request, custom_kernel) with a custom kernel
 * (include/dsmvc/engine.hpp:78, 177-178; src/axis_plan.cpp:535-599): request
 * kernel DSMVC_HIP_CUSTOM, support = request taps (times blur, rounded up).
 * `kernel(x, user)` is called on the host, during this call only, at each
 * tap's |distance| (blur applied); a non-finite value fails the build, as a
 * zero weight sum does.  Plans of custom kernels are built on every call and
 * never cached, as in the reference (src/axis_plan.cpp:1367-1372); the plugin
 * passes a VapourSynth function here (src/vs_plugin.cpp:628-653). */
typedef double (*dsmvc_hip_kernel_fn)(double x, void *user);
int dsmvc_hip_plan_build_custom(const dsmvc_hip_axis_request *request, dsmvc_hip_kernel_fn kernel,
                                void *user, dsmvc_hip_plan **plan);
void dsmvc_hip_plan_release(dsmvc_hip_plan *plan);
/* AxisPlan scalar fields. */
int dsmvc_hip_plan_info_get(const dsmvc_hip_plan *plan, dsmvc_hip_plan_info *info);
/* Copies AxisPlan's arrays (engine.hpp:90-107) into caller buffers sized from
 * the info: offsets n+1, indices/weights nnz, lower_ld/upper_l W*n,
 * inverse_diagonal n; the F64 arrays (nnz, (W+1)*n, (W+1)*n, n) only when
 * requires_float64.  NULL skips an array. */
int dsmvc_hip_plan_arrays(const dsmvc_hip_plan *plan, uint32_t *offsets, int32_t *indices,
                          float *weights, float *lower_ld, float *upper_l,
                          float *inverse_diagonal, double *weights_f64,
                          double *normal_bands_f64, double *ldlt_bands_f64,
                          double *inverse_diagonal_f64);
/* The arrays of an AxisPlan made elsewhere (reference include/dsmvc/engine.hpp:
 * 90-107), sized as in dsmvc_hip_plan_arrays: offsets n+1, indices/weights
 * nnz, lower_ld/upper_l W*n, inverse_diagonal n; the four F64 arrays (nnz,
 * (W+1)*n, (W+1)*n, n) are read only when info->requires_float64. */
typedef struct dsmvc_hip_plan_data {
    const uint32_t *offsets;
    const int32_t *indices;
    const float *weights;
    const float *lower_ld;
    const float *upper_l;
    const float *inverse_diagonal;
    const double *weights_f64;
    const double *normal_bands_f64;
    const double *ldlt_bands_f64;
    const double *inverse_diagonal_f64;
} dsmvc_hip_plan_data;
/* Takes a plan built by the reference's own planner (or any caller) instead
 * of building one: the reference executors accept any valid
 * `const AxisPlan &` (include/dsmvc/engine.hpp:455-490) and copy borrowed
 * plans per call (src/cpu_executor.cpp:553-558).  The scalar fields come from
 * `info` (its nnz and requires_float64 size the arrays); the arrays are
 * copied.  The plan must pass the reference's AxisPlan::valid() rules
 * (src/axis_plan.cpp:1252-1341), else DSMVC_HIP_INVALID_ARGUMENT. */
int dsmvc_hip_plan_import(const dsmvc_hip_plan_info *info, const dsmvc_hip_plan_data *data,
                          dsmvc_hip_plan **plan);
/* No counterpart in the reference: reports how the windowed (independent
 * chunk) kernels would partition this plan, computed on the host. */
int dsmvc_hip_plan_window_info(const dsmvc_hip_plan *plan, dsmvc_hip_window_info *info);
/* No counterpart in the reference: the plan's explicit operator
 * M = G^-1 A^T (G = A^T A), built on the host in FP64 and truncated to the
 * band holding all but 2^-28 of each row's L1 mass at either end, which the
 * fused 2D route multiplies on MFMA.  `lo` / `count` ([destination_size],
 * optional) receive each row's first source index and length, `values`
 * (optional, `capacity` entries) the rows one after another. */
int dsmvc_hip_plan_operator(const dsmvc_hip_plan *plan, dsmvc_hip_operator_info *info, int32_t *lo,
                            int32_t *count, double *values, int64_t capacity);
/* dsmvc::planner_cache_stats / clear_planner_caches (engine.hpp:140-149,
 * 179-180): the plan LRU and the sampling-geometry LRU beneath it
 * (src/axis_plan.cpp:520-533). */
int dsmvc_hip_planner_cache_stats(dsmvc_hip_planner_stats *stats);
void dsmvc_hip_clear_planner_caches(void);
/* dsmvc::detail::padding_index (src/axis_plan.cpp:1024-1066). */
int dsmvc_hip_padding_index(int64_t index, int32_t size, int32_t mode, int32_t *mapped);

/* ---- backends ----------------------------------------------------------- */
/* dsmvc::parse_backend / resolve_backend (src/backend.cpp:107-142): "hip" is
 * accepted; resolving hip without a device fails, never falls back. */
int dsmvc_hip_parse_backend(const char *name, int32_t *kind);
int dsmvc_hip_resolve_backend(int32_t requested, int32_t *resolved);
int dsmvc_hip_device_count(int32_t *count);

/* ---- executor ----------------------------------------------------------- */
/* dsmvc::Executor(BackendKind::hip) (engine.hpp:363-372, src/executor.cpp:39-59);
 * device -1 = current device.  solve_mode: DSMVC_HIP_SOLVE_*:
 *   AUTO      F32 plans: the windowed or chunked (partitioned) kernels, chosen
 *             per plan and direction by timing; Float64 route of F32 2D
 *             planes: the windowed FP64 kernels where the plan's FP64
 *             recurrence decays within their warm-up; otherwise sequential.
 *   EXACT     sequential kernels, bit-identical to the reference's ordered
 *             F32 solver (and its Float64 route).
 *   CHUNKED / WINDOWED  force that F32 family (plans it cannot take fail).
 *   REFINED   AUTO, with plans that retain FP64 data solved in F32 plus
 *             refinement instead of the Float64 route.
 *   FUSED     AUTO, and every 2D call (F32, u8, u16 planes on the F32 route)
 *             takes the fused route wherever both plans' explicit operators
 *             fit it (dsmvc_hip_plan_operator): X = M_v Y M_h^T in one launch
 *             on MFMA with 16-bit split operands, no intermediate in memory
 *             (F32 planes: FP16 parts, and an exact bfloat16 pass when a
 *             sample is beyond FP16's range).  AUTO takes it the same way
 *             unless DSMVC_HIP_FUSED=0. */
int dsmvc_hip_executor_create(int32_t device, int32_t solve_mode, dsmvc_hip_executor **executor);
void dsmvc_hip_executor_destroy(dsmvc_hip_executor *executor);
/* Executor::prepare / seal (src/executor.cpp:313-349): uploads the plan image. */
int dsmvc_hip_executor_prepare(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan);
int dsmvc_hip_executor_seal(dsmvc_hip_executor *executor);
/* No counterpart in the reference (which solves such plans in double only,
 * src/axis_plan.cpp:1124-1248): in DSMVC_HIP_SOLVE_REFINED mode, plans that
 * retain FP64 data are solved in F32 and refined `steps` times (1..3): the
 * residual A^T y - G x from the FP64 weights and normal matrix (in FP64 for
 * plans the windowed kernels take, else in compensated F32 with two-product
 * and two-sum), rounded to F32, the correction solved with the F32 factors
 * and added.  Float32 planes only. */
int dsmvc_hip_executor_set_refinement(dsmvc_hip_executor *executor, int32_t steps);
/* No counterpart in the reference: the route the autotuner chose for a
 * plan, direction (rows != 0: horizontal) and sample kind -- a chunk count,
 * -1 for the windowed kernels, 0 when not tuned yet. */
int dsmvc_hip_executor_tuned_route(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                   int32_t rows, int32_t sample_kind, int32_t *route);
/* Route of the parallel F32 solves per direction: a chunk count for the
 * carry-chain chunked kernels, -1 for the windowed kernels, 0 = automatic
 * (autotuned).  A route the plan cannot use falls back to automatic. */
int dsmvc_hip_executor_tune(dsmvc_hip_executor *executor, int32_t rows_chunks,
                            int32_t columns_chunks);

/* Executor::inverse_rows (engine.hpp:468-472, src/executor.cpp:351-377). */
int dsmvc_hip_inverse_rows(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                           const float *input, ptrdiff_t input_row_stride, float *output,
                           ptrdiff_t output_row_stride, int32_t row_count, void *stream,
                           uint32_t flags);
/* Executor::inverse_columns (engine.hpp:474-478, src/executor.cpp:379-404). */
int dsmvc_hip_inverse_columns(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                              const float *input, ptrdiff_t input_row_stride, float *output,
                              ptrdiff_t output_row_stride, int32_t column_count, void *stream,
                              uint32_t flags);
/* Executor::inverse_2d (engine.hpp:480-484, src/executor.cpp:420-445). */
int dsmvc_hip_inverse_2d(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                         const dsmvc_hip_plan *vertical, const float *input,
                         ptrdiff_t input_row_stride, float *output, ptrdiff_t output_row_stride,
                         void *stream, uint32_t flags);
/* Executor::inverse_2d_u8 / inverse_2d_u16 (engine.hpp:486-498,
 * src/executor.cpp:447-501); the streamed variants (503-557) compute the same
 * values and map here too. */
int dsmvc_hip_inverse_2d_u8(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                            const dsmvc_hip_plan *vertical, const uint8_t *input,
                            ptrdiff_t input_row_stride, uint8_t *output,
                            ptrdiff_t output_row_stride,
                            const dsmvc_hip_integer_conversion *conversion, void *stream,
                            uint32_t flags);
int dsmvc_hip_inverse_2d_u16(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                             const dsmvc_hip_plan *vertical, const uint16_t *input,
                             ptrdiff_t input_row_stride, uint16_t *output,
                             ptrdiff_t output_row_stride,
                             const dsmvc_hip_integer_conversion *conversion, void *stream,
                             uint32_t flags);
/* A batch of `frames` equally shaped planes in one submission: the reference
 * runs one plane per call (src/vs_plugin.cpp:1636-1745); this entry keeps its
 * arithmetic and adds a frame stride (elements).  sample_kind:
 * DSMVC_HIP_SAMPLE_*; conversion is read for integer samples only. */
int dsmvc_hip_inverse_2d_batch(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                               const dsmvc_hip_plan *vertical, int32_t sample_kind,
                               const void *input, ptrdiff_t input_row_stride,
                               ptrdiff_t input_frame_stride, void *output,
                               ptrdiff_t output_row_stride, ptrdiff_t output_frame_stride,
                               int32_t frames, const dsmvc_hip_integer_conversion *conversion,
                               void *stream, uint32_t flags);
/* No counterpart in the reference.  dsmvc_hip_inverse_2d_batch for callers
 * that cannot allocate device memory during a call (code under HIP graph
 * capture, tight serving loops, hosts that forbid allocation): it calls no
 * hipMalloc, hipFree or allocating library routine.  Everything beyond the
 * caller's input and output -- the intermediate, the sequential kernels'
 * forward sweeps, the job records, the Float64 flag -- lives in a fixed
 * 64 MiB device array per GPU, held until the process ends; a request that
 * needs more (about a 4K F32 frame on the sequential route) fails with
 * DSMVC_HIP_INVALID_ARGUMENT.  Both plans must have been prepared on this
 * executor first (dsmvc_hip_executor_prepare); runs the windowed kernels
 * where the prepared tables allow, else the sequential ones (always in EXACT
 * mode), without autotuning; REFINED executors are refused; the Float64
 * route needs a synchronous call.  All calls on one GPU share the array:
 * calls that overlap on the device (other streams, other executors) race. */
int dsmvc_hip_inverse_2d_batch_noalloc(dsmvc_hip_executor *executor,
                                       const dsmvc_hip_plan *horizontal,
                                       const dsmvc_hip_plan *vertical, int32_t sample_kind,
                                       const void *input, ptrdiff_t input_row_stride,
                                       ptrdiff_t input_frame_stride, void *output,
                                       ptrdiff_t output_row_stride,
                                       ptrdiff_t output_frame_stride, int32_t frames,
                                       const dsmvc_hip_integer_conversion *conversion,
                                       void *stream, uint32_t flags);
/* The two halves of the 2D route as separate batched calls (Float32 plans):
 * the reference's GPU pipeline runs the horizontal stage into an intermediate
 * and the vertical stage out of it (src/cuda/cuda_executor.cpp:2866-2949).
 * rows: `rows` rows of `input_kind` samples -> float32 output rows;
 * columns: `columns` columns of float32 input -> `output_kind` samples.
 * Integer kinds need `conversion` (input normalisation / output rounding). */
int dsmvc_hip_inverse_rows_batch(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                 int32_t input_kind, const void *input,
                                 ptrdiff_t input_row_stride, ptrdiff_t input_frame_stride,
                                 float *output, ptrdiff_t output_row_stride,
                                 ptrdiff_t output_frame_stride, int32_t rows, int32_t frames,
                                 const dsmvc_hip_integer_conversion *conversion, void *stream,
                                 uint32_t flags);
int dsmvc_hip_inverse_columns_batch(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                    const float *input, ptrdiff_t input_row_stride,
                                    ptrdiff_t input_frame_stride, int32_t output_kind,
                                    void *output, ptrdiff_t output_row_stride,
                                    ptrdiff_t output_frame_stride, int32_t columns,
                                    int32_t frames, const dsmvc_hip_integer_conversion *conversion,
                                    void *stream, uint32_t flags);

/* ---- host frames ----------------------------------------------------------
 * The reference's executors take host buffers (include/dsmvc/engine.hpp:
 * 455-490); its CUDA executor stages them through pinned memory on a stream
 * of its own (src/cuda/cuda_executor.cpp:2798-2822).  These entries take HOST
 * pointers and do the same on one of the executor's execution slots: the
 * input is packed into the slot's pinned buffer, copied to the device,
 * solved, copied back and written to the caller's rows (logical width only).
 * Synchronous; concurrent callers overlap (up to 8 slots).  Same arguments
 * and errors as the device-pointer entries.
 * Executor::inverse_rows / inverse_columns (engine.hpp:468-478). */
int dsmvc_hip_inverse_rows_host(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                const float *input, ptrdiff_t input_row_stride, float *output,
                                ptrdiff_t output_row_stride, int32_t row_count);
int dsmvc_hip_inverse_columns_host(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                   const float *input, ptrdiff_t input_row_stride,
                                   float *output, ptrdiff_t output_row_stride,
                                   int32_t column_count);
/* Executor::inverse_2d / inverse_2d_u8 / inverse_2d_u16 (engine.hpp:480-498)
 * by sample kind, for `frames` planes (frame strides read when frames > 1). */
int dsmvc_hip_inverse_2d_host(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                              const dsmvc_hip_plan *vertical, int32_t sample_kind,
                              const void *input, ptrdiff_t input_row_stride,
                              ptrdiff_t input_frame_stride, void *output,
                              ptrdiff_t output_row_stride, ptrdiff_t output_frame_stride,
                              int32_t frames, const dsmvc_hip_integer_conversion *conversion);

/* ---- GPU frames exported by another API ----------------------------------
 * One plane of a GPU frame exported by its producer (VapourSynth's GPU core)
 * for a foreign API (reference include/dsmvc/engine.hpp:306-323,
 * ExportedPlane).  handle_type 0x1 (VkExternalMemoryHandleTypeFlagBits
 * OPAQUE_FD): `handle` is a POSIX file descriptor of the whole allocation,
 * imported with hipImportExternalMemory once per memory_id (the producer
 * never reuses one) and kept in a bounded LRU of the executor; the plane
 * lives at `offset`, rows `pitch_bytes` apart.  Every call consumes both
 * planes' handles, whatever the outcome.  ready_semaphore / ready_value: the
 * input's producer pair, waited for on the device when that semaphore was
 * imported (dsmvc_hip_import_exported_semaphore), else 0 or already waited
 * for by the caller. */
typedef struct dsmvc_hip_exported_plane {
    uint64_t memory_id;
    uint64_t memory_size;
    uint64_t offset;
    uint64_t plane_bytes;    /* pitch_bytes * height */
    int32_t handle_type;
    int32_t reserved;
    int64_t handle;
    int64_t pitch_bytes;
    uint64_t ready_semaphore;
    uint64_t ready_value;
} dsmvc_hip_exported_plane;

typedef struct dsmvc_hip_import_stats {
    uint64_t entries;        /* allocations currently imported */
    uint64_t imports;
    uint64_t hits;
    uint64_t evictions;
} dsmvc_hip_import_stats;

/* Executor::inverse_2d_exported / inverse_2d_u8_exported /
 * inverse_2d_u16_exported (engine.hpp:432-442) by sample kind.  Synchronous:
 * the output plane is written when the call returns.  A plane outside its
 * allocation, a pitch that is not a whole number of samples or shorter than
 * a row: DSMVC_HIP_INVALID_ARGUMENT; a handle HIP cannot import:
 * DSMVC_HIP_RUNTIME_ERROR. */
int dsmvc_hip_inverse_2d_exported(dsmvc_hip_executor *executor, const dsmvc_hip_plan *horizontal,
                                  const dsmvc_hip_plan *vertical, int32_t sample_kind,
                                  const dsmvc_hip_exported_plane *input,
                                  const dsmvc_hip_exported_plane *output,
                                  const dsmvc_hip_integer_conversion *conversion);
/* Executor::inverse_rows_exported / inverse_columns_exported
 * (engine.hpp:426-431): Float32 planes. */
int dsmvc_hip_inverse_rows_exported(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                    const dsmvc_hip_exported_plane *input,
                                    const dsmvc_hip_exported_plane *output, int32_t row_count);
int dsmvc_hip_inverse_columns_exported(dsmvc_hip_executor *executor, const dsmvc_hip_plan *plan,
                                       const dsmvc_hip_exported_plane *input,
                                       const dsmvc_hip_exported_plane *output,
                                       int32_t column_count);
/* Executor::import_exported_semaphore / has_exported_semaphore
 * (engine.hpp:423-425): a timeline semaphore of the producer, keyed by its
 * producer-side handle value `semaphore`; *imported = 1 when HIP imported it
 * (hipImportExternalSemaphore).  The handle is consumed either way. */
int dsmvc_hip_import_exported_semaphore(dsmvc_hip_executor *executor, uint64_t semaphore,
                                        int32_t handle_type, int64_t handle, int32_t *imported);
int dsmvc_hip_has_exported_semaphore(dsmvc_hip_executor *executor, uint64_t semaphore,
                                     int32_t *imported);
/* dsmvc::close_exported_plane (engine.hpp:325-326): closes the handle of a
 * plane that was never passed to an executor. */
void dsmvc_hip_close_exported_plane(const dsmvc_hip_exported_plane *plane);
/* No counterpart in the reference: the executor's import cache counters. */
int dsmvc_hip_executor_import_stats(dsmvc_hip_executor *executor, dsmvc_hip_import_stats *stats);

/* ---- device plan cache ----------------------------------------------------
 * The executor keeps each plan's device image (and the tables built for it)
 * in an LRU keyed by the plan's build serial -- one per planner request key
 * (kernel, geometry, padding, precision) -- bounded by bytes: 4096 MiB by
 * default, DSMVC_HIP_PLAN_CACHE_MB, or this call.  The reference bounds its
 * packed-plan cache the same way (src/cuda/cuda_executor.cpp:219-235,
 * 1936-2020).  Images in use by a call, a batch or a served request are
 * never evicted. */
typedef struct dsmvc_hip_plan_cache_stats {
    uint64_t entries, bytes, limit;
    uint64_t hits, misses, evictions;
} dsmvc_hip_plan_cache_stats;
int dsmvc_hip_executor_set_plan_cache(dsmvc_hip_executor *executor, int64_t limit_bytes);
int dsmvc_hip_executor_plan_cache_stats(dsmvc_hip_executor *executor,
                                        dsmvc_hip_plan_cache_stats *stats);

/* Waits for `stream`; reports a deferred Float64 non-finite failure. */
/* One plane of a multi-plane call (dsmvc_hip_inverse_2d_planes). */
typedef struct dsmvc_hip_plane_job {
    const dsmvc_hip_plan *horizontal;
    const dsmvc_hip_plan *vertical;
    const void *input;               /* device, samples of the call's kind */
    ptrdiff_t input_row_stride;      /* samples */
    ptrdiff_t input_frame_stride;    /* samples; read when frames > 1 */
    void *output;
    ptrdiff_t output_row_stride;
    ptrdiff_t output_frame_stride;
    int32_t frames;
    int32_t reserved;
    const dsmvc_hip_integer_conversion *conversion;   /* integer kinds only */
} dsmvc_hip_plane_job;

/* Every plane of a frame batch in one set of launches.  The reference's
 * filter_get_frame solves a frame plane by plane, one Executor::inverse_2d /
 * inverse_2d_u16 call each (src/vs_plugin.cpp:1636-1745); here the
 * horizontal passes of all planes (luma and chroma, each with its own plans
 * and conversion) share one windowed launch per kernel class, then the
 * vertical passes.  Planes those kernels cannot take are solved one by one.
 * sample_kind: DSMVC_HIP_SAMPLE_*; results match per-plane calls within the
 * F32 route's tolerance. */
int dsmvc_hip_inverse_2d_planes(dsmvc_hip_executor *executor, const dsmvc_hip_plane_job *planes,
                                int32_t count, int32_t sample_kind, void *stream, uint32_t flags);

/* ---- multi-plan batches --------------------------------------------------
 * No single counterpart in the reference, whose executors solve one plane per
 * call (src/vs_plugin.cpp:1636-1745); its closest analogue is the Metal
 * route's axis-job descriptor table, at most 48 jobs per dispatch
 * (src/metal/metal_routes.metal:25-35, 363-408).  A batch holds many 2D F32
 * jobs of uneven geometry; each run solves all of them with one launch per
 * (direction, band class, tap width) over device-resident job tables, and
 * solves jobs those kernels cannot take (Float64 route, bands above 7) one by
 * one after them.  Buffers are fixed at creation; a run rewrites the outputs.
 * Job outputs must not overlap any job's input or output.  The executor must
 * outlive the batch. */
typedef struct dsmvc_hip_batch_job {
    const dsmvc_hip_plan *horizontal;
    const dsmvc_hip_plan *vertical;
    const float *input;              /* device, F32 */
    ptrdiff_t input_row_stride;      /* samples */
    ptrdiff_t input_frame_stride;    /* samples; read when frames > 1 */
    float *output;                   /* device, F32 */
    ptrdiff_t output_row_stride;
    ptrdiff_t output_frame_stride;
    int32_t frames;
    int32_t reserved;
} dsmvc_hip_batch_job;

typedef struct dsmvc_hip_batch_info {
    int32_t jobs;
    int32_t table_jobs;              /* solved by the table launches */
    int32_t fallback_jobs;           /* solved one by one */
    int32_t launches;                /* table launches per run */
    int64_t items;                   /* wavefront work items per run */
    int64_t intermediate_bytes;
    int32_t graph;                   /* 1: runs replay a captured HIP graph */
    int32_t fused_jobs;              /* of the table jobs, those on the fused 2D route */
} dsmvc_hip_batch_info;

typedef struct dsmvc_hip_batch dsmvc_hip_batch;

/* flags for dsmvc_hip_batch_create: capture the table launches into a HIP
 * graph at creation and replay it on every run. */
#define DSMVC_HIP_BATCH_GRAPH 2u

int dsmvc_hip_batch_create(dsmvc_hip_executor *executor, const dsmvc_hip_batch_job *jobs,
                           int32_t count, uint32_t flags, dsmvc_hip_batch **batch);
/* flags: DSMVC_HIP_ASYNC to return once the run is queued on `stream`. */
int dsmvc_hip_batch_run(dsmvc_hip_batch *batch, void *stream, uint32_t flags);
int dsmvc_hip_batch_info_get(const dsmvc_hip_batch *batch, dsmvc_hip_batch_info *info);
void dsmvc_hip_batch_destroy(dsmvc_hip_batch *batch);

/* ---- request loop (continuous batching) ------------------------------------
 * Counterpart of serving concurrent getFrame calls through the reference's
 * executors: fmParallel filters (src/vs_plugin.cpp:1955-1957) whose CUDA
 * calls are bounded by execution slots (src/cuda/cuda_executor.cpp:1102-1157,
 * 1462-1503) with a per-executor plan cache (1936-2020).  Requests of any
 * geometry are planned on submit (plans cached by kernel, geometry, padding
 * and precision), admitted when a slot and device memory for their
 * intermediates are free, and solved together with everything else admitted
 * in the same round; rounds on different slots overlap.  A request's result
 * is the same alone and in any mix.  Planes are F32 device buffers that must
 * stay valid until the request finishes. */
typedef struct dsmvc_hip_serve_request {
    int32_t planes;                  /* 1 (gray) or 3 (YUV) */
    int32_t subsampling_w;           /* log2 chroma subsampling, YUV only */
    int32_t subsampling_h;
    int32_t src_width, src_height;
    int32_t dst_width, dst_height;
    int32_t kernel;                  /* DSMVC_HIP_BILINEAR ... */
    int32_t taps;
    int32_t border;                  /* padding mode 0..4 */
    int32_t f64_mode;                /* 0 automatic, 1 float32 only, 2 float64 */
    int32_t reserved;
    double b, c, blur;
    const float *src[3];
    ptrdiff_t src_stride[3];         /* samples */
    float *dst[3];
    ptrdiff_t dst_stride[3];
} dsmvc_hip_serve_request;

/* fused: when an F32 plane (of an executor that takes the fused 2D route)
 * runs on the fused route instead of the windowed tables.
 * DSMVC_HIP_SERVE_FUSED_WARM (0, the default): once the plan pair's
 * operator images exist, which a builder thread prepares off the submit
 * path after the pair's first request; DSMVC_HIP_SERVE_FUSED_ALWAYS: the
 * images are built at submit (a new geometry's request waits for them), so
 * the route never changes; DSMVC_HIP_SERVE_FUSED_OFF: never. */
#define DSMVC_HIP_SERVE_FUSED_WARM 0
#define DSMVC_HIP_SERVE_FUSED_ALWAYS 1
#define DSMVC_HIP_SERVE_FUSED_OFF 2

typedef struct dsmvc_hip_server_config {
    int64_t memory_budget;           /* device bytes for in-flight intermediates */
    int32_t slots;                   /* rounds in flight, 1..16 */
    int32_t max_round_requests;
    int32_t fused;                   /* DSMVC_HIP_SERVE_FUSED_* */
    int32_t reserved;
} dsmvc_hip_server_config;

typedef struct dsmvc_hip_server_stats {
    uint64_t submitted, completed, failed, rounds;
    uint64_t table_planes, fallback_planes, largest_round;
    int64_t memory_high_water;
    uint64_t fused_planes;           /* of the table planes, those on the fused 2D route */
} dsmvc_hip_server_stats;

typedef struct dsmvc_hip_server dsmvc_hip_server;

int dsmvc_hip_server_create(dsmvc_hip_executor *executor, const dsmvc_hip_server_config *config,
                            dsmvc_hip_server **server);
/* Plans and queues a request (thread-safe); *id identifies it. */
int dsmvc_hip_server_submit(dsmvc_hip_server *server, const dsmvc_hip_serve_request *request,
                            uint64_t *id);
/* Waits for a request; returns its status (DSMVC_HIP_OK or the failure's
 * code).  Times are seconds since the server started: submission and the
 * moment completion was observed (either pointer may be NULL). */
int dsmvc_hip_server_wait(dsmvc_hip_server *server, uint64_t id, double *submitted,
                          double *completed);
/* *state: 0 pending, 1 done, -1 failed. */
int dsmvc_hip_server_poll(dsmvc_hip_server *server, uint64_t id, int32_t *state);
/* Forgets a finished request. */
int dsmvc_hip_server_release(dsmvc_hip_server *server, uint64_t id);
int dsmvc_hip_server_stats_get(const dsmvc_hip_server *server, dsmvc_hip_server_stats *stats);
/* Finishes every accepted request, then stops. */
void dsmvc_hip_server_destroy(dsmvc_hip_server *server);

/* No counterpart in the reference: waits for `stream` and frees the scratch
 * and job tables the executor keeps for it.  Call before destroying a stream
 * the executor has run work on; otherwise that state lives until
 * dsmvc_hip_executor_destroy. */
int dsmvc_hip_executor_release_stream(dsmvc_hip_executor *executor, void *stream);
int dsmvc_hip_synchronize(dsmvc_hip_executor *executor, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* DSMVC_HIP_H */

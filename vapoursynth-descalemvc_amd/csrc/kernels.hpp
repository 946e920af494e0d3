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
bool chunked_band_supported(int w);
int chunked_max_length();
int rows_ring_span();   // widest CSR row (last - first source index) rows_chunked can stage

hipError_t launch_seq(const SeqArgs &a, Sample in, Sample out, bool f64, hipStream_t s);
hipError_t launch_rows_chunked(const ChunkArgs &a, Sample in, hipStream_t s);
hipError_t launch_columns_chunked(const ChunkArgs &a, Sample out, hipStream_t s);
hipError_t launch_check_finite(const void *src, Sample in, long s_row, long s_frame, int rows,
                               int cols, int frames, int *flag, hipStream_t s);

} // namespace dsmvc::hip

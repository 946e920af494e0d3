// Chunked axis kernels for half bandwidth 5 (see kernels_chunked.hpp).
#include "kernels_chunked.hpp"

namespace dsmvc::hip {
template hipError_t launch_chunked_band<5>(const ChunkArgs &, bool, Sample, hipStream_t);
} // namespace dsmvc::hip

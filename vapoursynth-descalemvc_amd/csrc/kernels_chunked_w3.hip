// Chunked axis kernels for half bandwidth 3 (see kernels_chunked.hpp).
#include "kernels_chunked.hpp"

namespace dsmvc::hip {
template hipError_t launch_chunked_band<3>(const ChunkArgs &, bool, Sample, hipStream_t);
} // namespace dsmvc::hip

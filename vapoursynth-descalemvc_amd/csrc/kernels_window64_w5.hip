// Windowed FP64 kernels (Float64 route), band class 5: rows and columns,
// tap widths 8 and 16, F32 and integer samples (kernels_window64.hpp).
#include "kernels_window64.hpp"

namespace dsmvc::hip {
template hipError_t launch_window64_table_band<5>(const Win64Job *, int, long, const Win64Job &,
                                                     int, bool, Sample, hipStream_t);
template long window64_table_slots_band<5>(int, bool, Sample);
} // namespace dsmvc::hip

// Windowed axis kernels for band class 7, 8 taps per row (see kernels_window.hpp).
#include "kernels_window.hpp"

namespace dsmvc::hip {
template hipError_t launch_window_table_band<7, 8>(const WinJob *, int, long, bool, bool,
                                                        Sample, int, hipStream_t);
template long window_table_slots_band<7, 8>(bool, bool, Sample);
} // namespace dsmvc::hip

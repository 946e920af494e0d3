// Windowed axis kernels for band class 3, 6 taps per row (see kernels_window.hpp).
#include "kernels_window.hpp"

namespace dsmvc::hip {
template hipError_t launch_window_table_band<3, 6>(const WinJob *, int, long, bool, bool,
                                                        Sample, int, hipStream_t);
template long window_table_slots_band<3, 6>(bool, bool, Sample);
} // namespace dsmvc::hip

// fft_line32.hip -- FP32 instantiation of the line passes (fft_line.h).
#include "fft_line.h"

namespace fts {
template hipError_t launch_line<float>(const FftGeom&, int, int, const void*, const LineJob<float>*, int, hipStream_t);
template hipError_t launch_line_table<float>(const FftGeom&, int, int, const void*, const LineJob<float>*, int, hipStream_t);
} // namespace fts

// fft_line64.hip -- FP64 instantiation of the line passes (fft_line.h).
#include "fft_line.h"

namespace fts {
template hipError_t launch_line<double>(const FftGeom&, int, int, const void*, const LineJob<double>*, int, hipStream_t);
template hipError_t launch_line_table<double>(const FftGeom&, int, int, const void*, const LineJob<double>*, int, hipStream_t);
} // namespace fts

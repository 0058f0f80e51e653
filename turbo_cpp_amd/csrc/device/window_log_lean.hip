// Window decoder kernels, Log-MAP in the log domain, for launches whose every
// job has the lean geometry (one code block per wave in 64 segments, segment
// length and warm-up multiples of 8: the long blocks, K >= 2048 and a
// multiple of 512).  The algorithm is in window_kernel.h (siso_half_lean);
// the results equal those of window_log.hip's kernels bit for bit.
#include "window_kernel.h"

namespace tcdev {

__global__ __launch_bounds__(64, 2) void window_decode_log_lean_kernel(WinArgs args) {
  window_decode<LogDomain, false, _Float16, true>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_log_lean_es_kernel(WinArgs args) {
  window_decode<LogDomain, true, _Float16, true>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_log_lean_f32_kernel(WinArgs args) {
  window_decode<LogDomain, false, float, true>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_log_lean_es_f32_kernel(WinArgs args) {
  window_decode<LogDomain, true, float, true>(args);
}

}  // namespace tcdev

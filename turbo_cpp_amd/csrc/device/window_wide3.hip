// Window decoder kernels whose jobs span 3 waves of one work-group: code
// blocks of up to 192 segments (e.g. K = 6144 at segment length 32), and
// 192 lanes of shorter blocks per job.  Log and probability domain,
// extrinsics held as fp16 or fp32, without and with CRC24B early stop.  The
// algorithm is in window_kernel.h (Wide<D, 3>); two waves per SIMD, as the
// one-wave kernels.
#include "window_kernel.h"

namespace tcdev {

__global__ __launch_bounds__(192, 2) void window_decode_log_w3_kernel(WinArgs args) {
  window_decode<Wide<LogDomain, 3>, false, _Float16>(args);
}
__global__ __launch_bounds__(192, 2) void window_decode_log_es_w3_kernel(WinArgs args) {
  window_decode<Wide<LogDomain, 3>, true, _Float16>(args);
}
__global__ __launch_bounds__(192, 2) void window_decode_log_f32_w3_kernel(WinArgs args) {
  window_decode<Wide<LogDomain, 3>, false, float>(args);
}
__global__ __launch_bounds__(192, 2) void window_decode_log_es_f32_w3_kernel(WinArgs args) {
  window_decode<Wide<LogDomain, 3>, true, float>(args);
}
__global__ __launch_bounds__(192, 2) void window_decode_lin_w3_kernel(WinArgs args) {
  window_decode<Wide<LinDomain, 3>, false, _Float16>(args);
}
__global__ __launch_bounds__(192, 2) void window_decode_lin_es_w3_kernel(WinArgs args) {
  window_decode<Wide<LinDomain, 3>, true, _Float16>(args);
}
__global__ __launch_bounds__(192, 2) void window_decode_lin_f32_w3_kernel(WinArgs args) {
  window_decode<Wide<LinDomain, 3>, false, float>(args);
}
__global__ __launch_bounds__(192, 2) void window_decode_lin_es_f32_w3_kernel(WinArgs args) {
  window_decode<Wide<LinDomain, 3>, true, float>(args);
}

}  // namespace tcdev

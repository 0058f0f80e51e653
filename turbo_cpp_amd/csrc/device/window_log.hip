// Window decoder kernels, Log-MAP in the log domain (max* on v_exp_f32 / v_log_f32): extrinsics held as fp16 or fp32 between
// half-iterations, without and with CRC24B early stop.  The algorithm is in
// window_kernel.h.
#include "window_kernel.h"

namespace tcdev {

__global__ __launch_bounds__(64, 2) void window_decode_log_kernel(WinArgs args) {
  window_decode<LogDomain, false, _Float16>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_log_es_kernel(WinArgs args) {
  window_decode<LogDomain, true, _Float16>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_log_f32_kernel(WinArgs args) {
  window_decode<LogDomain, false, float>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_log_es_f32_kernel(WinArgs args) {
  window_decode<LogDomain, true, float>(args);
}

}  // namespace tcdev

// Window decoder kernels, MAP in the probability domain (power-of-two rescaling): extrinsics held as fp16 or fp32 between
// half-iterations, without and with CRC24B early stop.  The algorithm is in
// window_kernel.h.
#include "window_kernel.h"

namespace tcdev {

__global__ __launch_bounds__(64, 2) void window_decode_lin_kernel(WinArgs args) {
  window_decode<LinDomain, false, _Float16>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_lin_es_kernel(WinArgs args) {
  window_decode<LinDomain, true, _Float16>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_lin_f32_kernel(WinArgs args) {
  window_decode<LinDomain, false, float>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_lin_es_f32_kernel(WinArgs args) {
  window_decode<LinDomain, true, float>(args);
}

}  // namespace tcdev

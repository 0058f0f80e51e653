// Exact decoder kernels: the reference's arithmetic (lte_turbo.cpp:410-486)
// computed segment-parallel with exact segment boundaries (siso_half_exact in
// window_kernel.h) and fp32-held extrinsics, without and with CRC24B early
// stop.  Posteriors equal the serial exact decoder's (exact_kernel.hip) bit
// for bit.
#include "window_kernel.h"

namespace tcdev {

__global__ __launch_bounds__(64, 2) void window_decode_exact_kernel(WinArgs args) {
  window_decode<ExactDomain, false, float>(args);
}
__global__ __launch_bounds__(64, 2) void window_decode_exact_es_kernel(WinArgs args) {
  window_decode<ExactDomain, true, float>(args);
}

}  // namespace tcdev

// VALU issue calibration for the roofline of the decoder kernels (bench.py
// --valu-calibration; not part of the decoder library or its C ABI).
//
// Each wave runs `iters` rounds of 16 independent FP32 operations of one kind
// (op 0: v_add_f32, op 1: v_exp_f32) and records its shader-clock cycles
// (s_memtime) and its start and end on the 100 MHz real-time counter
// (s_memrealtime).  From these the host derives the shader clock, the
// average number of resident waves per SIMD (wave lifetimes over SIMDs x
// kernel span) and so the SIMD cycles one wave64 instruction of that kind
// occupies: a wave's cycles per instruction divided by the waves sharing its
// SIMD.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

namespace {

template <int OP>
__global__ __launch_bounds__(64) void valu_rate_kernel(float* out, uint64_t* rec, int iters) {
  float r[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) r[j] = -0.5f - 1e-3f * static_cast<float>(threadIdx.x + j);
  const uint64_t c0 = __builtin_amdgcn_s_memtime();
  const uint64_t w0 = __builtin_amdgcn_s_memrealtime();
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int j = 0; j < 16; ++j) r[j] = OP == 0 ? r[j] + 1.0f : __builtin_amdgcn_exp2f(r[j]);
  }
  const uint64_t c1 = __builtin_amdgcn_s_memtime();
  const uint64_t w1 = __builtin_amdgcn_s_memrealtime();
  float s = 0.0f;
#pragma unroll
  for (int j = 0; j < 16; ++j) s += r[j];
  out[blockIdx.x * 64 + threadIdx.x] = s;
  if (threadIdx.x == 0) {
    rec[3 * blockIdx.x] = c1 - c0;
    rec[3 * blockIdx.x + 1] = w0;
    rec[3 * blockIdx.x + 2] = w1;
  }
}

}  // namespace

// op: 0 add, 1 exp; waves: one-wave work-groups to launch; results: shader
// clock (MHz), average resident waves per SIMD, a wave's cycles per
// instruction, SIMD cycles per instruction.  Returns 0 on success.
extern "C" int tc_calib_valu(int op, int waves, int iters, double* clock_mhz, double* waves_per_simd,
                             double* wave_cycles_per_instr, double* simd_cycles_per_instr) {
  if (op < 0 || op > 1 || waves <= 0 || iters <= 0) return 1;
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 2;
  float* out = nullptr;
  uint64_t* rec = nullptr;
  if (hipMalloc(&out, sizeof(float) * 64 * static_cast<size_t>(waves)) != hipSuccess) return 3;
  if (hipMalloc(&rec, sizeof(uint64_t) * 3 * static_cast<size_t>(waves)) != hipSuccess) {
    (void)hipFree(out);
    return 3;
  }
  auto kern = op == 0 ? valu_rate_kernel<0> : valu_rate_kernel<1>;
  int rc = 0;
  for (int rep = 0; rep < 2 && rc == 0; ++rep) {  // the first launch raises the clock
    hipLaunchKernelGGL(kern, dim3(waves), dim3(64), 0, 0, out, rec, iters);
    if (hipDeviceSynchronize() != hipSuccess) rc = 4;
  }
  std::vector<uint64_t> h(3 * static_cast<size_t>(waves));
  if (rc == 0 && hipMemcpy(h.data(), rec, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost) != hipSuccess) rc = 4;
  (void)hipFree(out);
  (void)hipFree(rec);
  if (rc) return rc;
  double cyc = 0.0, life = 0.0;
  uint64_t first = h[1], last = h[2];
  for (int i = 0; i < waves; ++i) {
    cyc += static_cast<double>(h[3 * i]);
    life += static_cast<double>(h[3 * i + 2] - h[3 * i + 1]);
    first = std::min(first, h[3 * i + 1]);
    last = std::max(last, h[3 * i + 2]);
  }
  const double simds = 4.0 * prop.multiProcessorCount;
  *clock_mhz = cyc / life * 100.0;  // shader cycles per 100 MHz tick
  *waves_per_simd = life / (static_cast<double>(last - first) * simds);
  *wave_cycles_per_instr = cyc / waves / (16.0 * iters);
  *simd_cycles_per_instr = *wave_cycles_per_instr / *waves_per_simd;
  return 0;
}

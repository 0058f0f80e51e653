// Device-side trellis of the 8-state RSC constituent code and the math
// primitives shared by the decoder kernels (gfx950).
//
// RSC (lte_turbo.cpp:333-341): state bits (s2 s1 s0), s2 newest;
// feedback = u ^ s1 ^ s2, parity = feedback ^ s0 ^ s2,
// next = (feedback << 2) | (s2 << 1) | s1.
// Consequences used below: the predecessors of next state n are
// 2(n & 3) and 2(n & 3) + 1, both through the same input u(n) = n2^n1^n0;
// the parity bit from predecessor 2(n&3)+s0 is n2 ^ n1 ^ s0.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace tcdev {

constexpr float kLog2e = 1.44269504088896340736f;
constexpr float kLn2 = 0.693147180559945309417f;
constexpr float kHalfLog2e = 0.721347520444481703680f;
constexpr float kNeg = -1.0e30f;          // finite stand-in for -inf in the window kernels
constexpr float kFltMin = 1.17549435e-38f;

struct Tr {
  unsigned parity, next;
};
__host__ __device__ constexpr Tr rsc(unsigned s, unsigned u) {
  const unsigned s0 = s & 1u, s1 = (s >> 1) & 1u, s2 = (s >> 2) & 1u;
  const unsigned fb = (u ^ s1 ^ s2) & 1u;
  return {fb ^ s0 ^ s2, (fb << 2) | (s2 << 1) | s1};
}
__host__ __device__ constexpr unsigned in_of(unsigned n) { return ((n >> 2) ^ (n >> 1) ^ n) & 1u; }
__host__ __device__ constexpr unsigned pred0(unsigned n) { return (n & 3u) << 1; }
__host__ __device__ constexpr unsigned par0(unsigned n) { return ((n >> 2) ^ (n >> 1)) & 1u; }
__host__ __device__ constexpr unsigned term_in(unsigned s) { return ((s >> 1) ^ (s >> 2)) & 1u; }

// Raw hardware transcendentals: v_exp_f32 (2^x) and v_log_f32 (log2 x).
__device__ __forceinline__ float hw_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ float hw_log2(float x) { return __builtin_amdgcn_logf(x); }

// max*(x, y) = hi + log1p(exp(lo - hi)) of two natural-log metrics (the
// reference's max_star for finite operands, lte_turbo.cpp:380-390), evaluated
// as hi + ln2 * log2(1 + 2^((lo - hi) log2e)) on v_exp_f32 / v_log_f32, the
// last multiply-add fused.  Both
// exact decoders use this one function, so their metrics agree bit for bit;
// against the reference's expf/log1pf it differs by rounding (~1e-7 per call).
__device__ __forceinline__ float jac_ref(float x, float y) {
#pragma clang fp contract(off)
  const float hi = fmaxf(x, y), lo = fminf(x, y);
  return __builtin_fmaf(hw_log2(1.0f + hw_exp2((lo - hi) * kLog2e)), kLn2, hi);
}

// Branch metric of input u and parity c from g1 = x + y, g2 = x - y
// (x = (Ls+La)/2, y = Lp/2): +-g1 for (1,1)/(0,0), +-g2 for (1,0)/(0,1).
__device__ __forceinline__ float gsel(unsigned u, unsigned c, float g1, float g2) {
  return u ? (c ? g1 : g2) : (c ? -g2 : -g1);
}

}  // namespace tcdev

// Window-parallel turbo decoder for gfx950: the throughput path.
//
// Work decomposition.  A job is what one 64-lane wavefront decodes from start
// to finish (all iterations): nblk code blocks of one K, each cut into M
// segments of S = K / M steps; lane = block * M + segment.  Each lane runs
// both SISO halves of every iteration over its own segment, lane-serially,
// with all 8 state metrics in VGPRs.  Blocks of more than 64 segments run on
// jobs of two or three waves of one work-group (Wide<D, NW>, below).
//
// Segment boundaries.  A segment's initial alpha (beta) comes from a g-step
// warm-up recursion over the neighbouring segment's last (first) g steps,
// itself started from the metrics that neighbour exported in the previous
// iteration (next-iteration initialisation).  Segment 0 starts from
// alpha_0 = state 0 and the last segment from the reference's tail beta
// (lte_turbo.cpp:357-378, tail read-out :758-773), so M == 1 is a plain
// full-length pass.  The exported metrics move between lanes with DPP
// wave_shr:1 / wave_shl:1 (on several waves: through the lanes' records).
//
// Memory.  The wavefront's extrinsics live in LDS for the whole decode, one
// fp16 per bit (fp32-held variants of the log and exact decoders: in the
// wave's scratch slot, see window_job) in a transposed layout: bit i = q*S + t of block b sits at
// row t, column b*M + q (rows of 64).  The upper half reads and writes
// its own column; the lower half reads and writes bit pi(j) for its steps
// j = q*S + t.  Because S divides K, pi(q*S + t) mod S = pi(t) mod S for
// every q (the QPP is contention-free), so at each step all lanes hit one
// row and distinct columns: no bank conflicts.  Row pi(t) mod S and the
// column offset pi(t) div S come from a small per-K table; the per-lane
// part of the column is (A_q + B_q t) mod M with A_q = (f1 q + f2 q^2 S)
// mod M and B_q = 2 f2 q mod M.  The channel LLRs (d0, d1 for the upper
// half; d0[pi(j)], d2 for the lower) are staged once per job into a
// per-wave scratch area in [t][lane] order (one plane of floats per input
// stream), so every per-step load of a wavefront is one contiguous 256-byte
// row per plane.  Each lower half leaves the
// posteriors (fp32) in the same scratch slot at the bits' extrinsic
// addresses; the job ends with one copy in natural order.
//
// Within a segment the backward pass stores beta only at sub-window
// boundaries (every 8 steps, in the wave's scratch slot); the forward pass
// recomputes each sub-window's beta into registers before computing alpha
// and the extrinsics.  The
// recomputation repeats the same instructions on the same inputs, so it is
// bit-identical to the first pass.
//
// Arithmetic: either max* in the log domain or the probability domain with
// power-of-two renormalisation; oracle/turbo_oracle.c (or_decode_window)
// states both step by step and the test suite compares the two.
#pragma once

#include <hip/hip_runtime.h>

#include "decoder_types.h"
#include "trellis.h"

#pragma clang fp contract(off)

namespace tcdev {
namespace {

__device__ __forceinline__ float dpp_from_prev_lane(float v) {  // wave_shr:1
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float dpp_from_next_lane(float v) {  // wave_shl:1
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}

// Two waves share each SIMD (the kernels take the 256 VGPRs a lane may have
// at two waves per SIMD).  The SIMD's issue arbitration favours the older of
// the two, which then finishes each job about 30% sooner and, at the end of
// the queue, leaves the younger one to run its remainder alone, at a lower
// throughput than the pair.  Alternating the two waves' priority on a global
// time slice of 16384 ticks of the 100 MHz reference clock (164 us, about a
// twelfth of a job), by the parity of their wave slots (slots 0 and 1 of the
// SIMD), keeps their progress even: 4096 blocks of K=6144 decode 6% faster,
// where the tail of the queue was 9% of the launch.  Called at every chunk of
// both recursion passes (set in one pass only, a priority would hold through
// the other).  The log-domain decoder only (D::kShareSimd).
__device__ __forceinline__ void share_simd() {
  const uint32_t slot = __builtin_amdgcn_s_getreg((3 << 11) | 4) & 1u;  // HW_ID, wave slot bits [3:0]
  const uint32_t t = (uint32_t)(__builtin_amdgcn_s_memrealtime() >> 14);
  if ((t ^ slot) & 1u) {
    __builtin_amdgcn_s_setprio(1);
  } else {
    __builtin_amdgcn_s_setprio(0);
  }
}

// 8 metrics of one lane as two float4 (per-lane records keep every access a
// single base register plus an immediate offset)
__device__ __forceinline__ void store8(float* __restrict__ p, const float* v) {
  reinterpret_cast<float4*>(p)[0] = make_float4(v[0], v[1], v[2], v[3]);
  reinterpret_cast<float4*>(p)[1] = make_float4(v[4], v[5], v[6], v[7]);
}
__device__ __forceinline__ void load8(const float* __restrict__ p, float* v) {
  const float4 a = reinterpret_cast<const float4*>(p)[0], b = reinterpret_cast<const float4*>(p)[1];
  v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
}

__device__ __forceinline__ float max8(const float* v) {
  return fmaxf(fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3])), fmaxf(fmaxf(v[4], v[5]), fmaxf(v[6], v[7])));
}

// ------------------------------------------------------------ log domain --
// Metrics in base-2 units (natural-log metric * log2e), so max* is
// max(a,b) + log2(1 + 2^-|a-b|) with no scaling inside the transcendentals.
struct LogDomain {
  struct Br {
    float g1, g2;  // branch metrics: +-g1 for (u,c) = (1,1)/(0,0), +-g2 for (1,0)/(0,1)
  };
  static constexpr float kUniform = 0.0f;
  static constexpr float kZero = kNeg;
  static constexpr bool kShareSimd = true;  // alternate priority with the SIMD's other wave
  static constexpr bool kExact = false;
  static constexpr int kWaves = 1;  // waves per job (Wide<D, NW>: NW)
  // fp32-held extrinsics of one-wave jobs in the scratch slot rather than LDS
  // (eight waves per CU instead of six: 4096 x 6144 in 4.6 instead of 5.3 ms)
  static constexpr bool kF32HeldInSlot = true;
  __device__ static __forceinline__ float from_log(float v) { return v * kLog2e; }
  // the staged parity LLR, as the branch metrics take it: scaled once per job
  // (the same product the branch metric would form at every use)
  __device__ static __forceinline__ float stage_par(float par) { return par * kHalfLog2e; }
  __device__ static __forceinline__ Br branch(float sa, float y) {
    const float x = sa * kHalfLog2e;
    return {x + y, x - y};
  }
  __device__ static __forceinline__ float maxstar(float a, float b) {
    return fmaxf(a, b) + hw_log2(1.0f + hw_exp2(-fabsf(a - b)));
  }
  // metrics relative to state 0 (any reference state keeps max* exact; the
  // spread between states stays bounded because every state reaches every
  // other in three steps).  Done once per chunk of kSubWin steps, not every
  // step (nrm): the metrics then drift by at most kSubWin branch metrics
  // between renormalisations, and the chunk-aligned positions (the chunk's
  // last beta step, its kSubWin-th alpha step) keep the beta recompute
  // bit-identical to the first pass
  __device__ static __forceinline__ void norm(float* v) {
    const float m = v[0];
#pragma unroll
    for (int i = 1; i < 8; ++i) v[i] = v[i] - m;
    v[0] = 0.0f;
  }
  // beta_t from beta_{t+1}
  __device__ static __forceinline__ void bstep(float* b, const Br& r, bool nrm = true) {
    float nb[8];
#pragma unroll
    for (unsigned p = 0; p < 8; ++p) {
      const Tr t0 = rsc(p, 0), t1 = rsc(p, 1);
      nb[p] = maxstar(b[t0.next] + gsel(0, t0.parity, r.g1, r.g2), b[t1.next] + gsel(1, t1.parity, r.g1, r.g2));
    }
    if (nrm) norm(nb);
#pragma unroll
    for (int i = 0; i < 8; ++i) b[i] = nb[i];
  }
  // alpha_{t+1} from alpha_t; returns the extrinsic of step t when WANT
  template <bool WANT>
  __device__ static __forceinline__ float astep(float* a, const Br& r, float sa, float /*sys*/, float /*ap*/,
                                                const float* bn, bool nrm = true) {
    float na[8];
#pragma unroll
    for (unsigned n = 0; n < 8; ++n) {
      const unsigned u = in_of(n), p = pred0(n), c = par0(n);
      na[n] = maxstar(a[p] + gsel(u, c, r.g1, r.g2), a[p + 1] + gsel(u, c ^ 1u, r.g1, r.g2));
    }
    float ext = 0.0f;
    if constexpr (WANT) {
      float tv[8];
#pragma unroll
      for (int n = 0; n < 8; ++n) tv[n] = na[n] + bn[n];
      // u(n) = 1 for n in {1,2,4,7}, 0 for {0,3,5,6}
      const float m1 = fmaxf(fmaxf(fmaxf(fmaxf(kNeg, tv[1]), tv[2]), tv[4]), tv[7]);
      const float m0 = fmaxf(fmaxf(fmaxf(fmaxf(kNeg, tv[0]), tv[3]), tv[5]), tv[6]);
      // 0 + x == x for the non-negative x here: summed as 0.0f + x + ...
      float s0 = hw_exp2(tv[0] - m0);
      float s1 = hw_exp2(tv[1] - m1);
      s1 += hw_exp2(tv[2] - m1);
      s0 += hw_exp2(tv[3] - m0);
      s1 += hw_exp2(tv[4] - m1);
      s0 += hw_exp2(tv[5] - m0);
      s0 += hw_exp2(tv[6] - m0);
      s1 += hw_exp2(tv[7] - m1);
      const float l1 = m1 + hw_log2(s1);
      const float l0 = m0 + hw_log2(s0);
      ext = (l1 - l0) * kLn2 - sa;
    }
    if (nrm) norm(na);
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = na[i];
    return ext;
  }
};

// ---------------------------------------------------- probability domain --
struct LinDomain {
  struct Br {
    float u1, u0, q1, q0;  // branch weight exp(gamma - |x| - |y|) = U_u * Q_c
  };
  static constexpr float kUniform = 1.0f;
  static constexpr float kZero = 0.0f;
  // measured slower with the alternating priority (3.23 -> 3.33 ms on 4096 x 6144)
  static constexpr bool kShareSimd = false;
  static constexpr bool kExact = false;
  static constexpr int kWaves = 1;  // waves per job (Wide<D, NW>: NW)
  // fp32-held extrinsics stay in LDS: with them in the scratch slot at eight
  // waves per CU this kernel ran 4.26 instead of 3.2 ms (its waves already
  // wait on memory for a third of their cycles)
  static constexpr bool kF32HeldInSlot = false;
  __device__ static __forceinline__ float from_log(float v) { return hw_exp2(v * kLog2e); }
  // The parity weight depends on the channel only, so it is staged once per
  // job: e^-|par| (y = par/2, y + y = par exactly) with the sign bit set
  // when y < 0; the sign bit, not the value, carries it (the weight may
  // underflow to zero).
  __device__ static __forceinline__ float stage_par(float par) {
    const float y = 0.5f * par;
    const float ey = hw_exp2(-fabsf(y + y) * kLog2e);
    return (y >= 0.0f) ? ey : -ey;
  }
  __device__ static __forceinline__ Br branch(float sa, float sp) {
    const float x = 0.5f * sa;
    const float ex = hw_exp2(-fabsf(x + x) * kLog2e);
    const bool yp = (__builtin_bit_cast(uint32_t, sp) >> 31) == 0u;
    return {(x >= 0.0f) ? 1.0f : ex, (x >= 0.0f) ? ex : 1.0f, yp ? 1.0f : -sp, yp ? sp : 1.0f};
  }
  // rescale by 2^-e so that the largest entry lies in [0.5, 1) (no branch:
  // an all-zero vector has e = 0 and stays as it is)
  __device__ static __forceinline__ void norm(float* v) {
    const int e = __builtin_amdgcn_frexp_expf(max8(v));
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = __builtin_amdgcn_ldexpf(v[i], -e);
  }
  // rescaled every step (nrm ignored): products of probabilities under- or
  // overflow within a few steps otherwise
  __device__ static __forceinline__ void bstep(float* b, const Br& r, bool /*nrm*/ = true) {
    const float w[2][2] = {{r.u0 * r.q0, r.u0 * r.q1}, {r.u1 * r.q0, r.u1 * r.q1}};
    float nb[8];
#pragma unroll
    for (unsigned p = 0; p < 8; ++p) {
      const Tr t0 = rsc(p, 0), t1 = rsc(p, 1);
      nb[p] = __builtin_fmaf(w[1][t1.parity], b[t1.next], w[0][t0.parity] * b[t0.next]);
    }
    norm(nb);
#pragma unroll
    for (int i = 0; i < 8; ++i) b[i] = nb[i];
  }
  template <bool WANT>
  __device__ static __forceinline__ float astep(float* a, const Br& r, float /*sa*/, float /*sys*/, float /*ap*/,
                                                const float* bn, bool /*nrm*/ = true) {
    const float q[2] = {r.q0, r.q1}, uu[2] = {r.u0, r.u1};
    float na[8];
    float s1 = 0.0f, s0 = 0.0f;
#pragma unroll
    for (unsigned n = 0; n < 8; ++n) {
      const unsigned u = in_of(n), p = pred0(n), c = par0(n);
      const float bb = __builtin_fmaf(a[p + 1], q[c ^ 1u], a[p] * q[c]);
      if constexpr (WANT) {
        if (u) s1 = __builtin_fmaf(bb, bn[n], s1);
        else s0 = __builtin_fmaf(bb, bn[n], s0);
      }
      na[n] = uu[u] * bb;
    }
    float ext = 0.0f;
    if constexpr (WANT) {
      s1 = fmaxf(s1, kFltMin);
      s0 = fmaxf(s0, kFltMin);
      ext = kLn2 * (hw_log2(s1) - hw_log2(s0));
    }
    norm(na);
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = na[i];
    return ext;
  }
};

// ------------------------------------------------ the reference's arithmetic --
// Natural-log metrics with the operations of lte_turbo.cpp:410-486 in their
// order: branch metric 0.5 (+-(sys + ap) +- par), candidates alpha + gamma
// (beta + gamma) combined by max* (jac_ref: hi + log1p(exp(lo - hi))),
// metrics renormalised by their maximum every step, the LLR of a step as the
// sequential max* accumulation of (alpha + gamma) + beta over (state, input)
// in the reference's order, extrinsic = ((L1 - L0) - sys) - ap.  A finite
// stand-in kNeg replaces the reference's -inf for the states alpha_0 ..
// alpha_2 cannot reach: max* with it returns the other operand (its exp2
// underflows to 0) and subtracting the maximum leaves it a stand-in, so every
// finite metric, and every extrinsic, is the one the reference's rules give
// with this max*.  Segment boundaries are exact (siso_half_exact).
struct ExactDomain {
  struct Br {
    float g1, g2;  // 0.5 (sa + par), 0.5 (sa - par): gamma = +-g1 for (u,c) = (1,1)/(0,0), +-g2 for (1,0)/(0,1)
  };
  static constexpr float kUniform = 0.0f;
  static constexpr float kZero = kNeg;
  static constexpr bool kShareSimd = true;
  static constexpr bool kExact = true;
  static constexpr int kWaves = 1;  // waves per job (Wide<D, NW>: NW)
  // fp32 extrinsics in the scratch slot (eight waves per CU; 12.5 instead of
  // 13.3 ms on 4096 x 6144)
  static constexpr bool kF32HeldInSlot = true;
  __device__ static __forceinline__ float from_log(float v) { return v; }
  __device__ static __forceinline__ float stage_par(float par) { return par; }
  __device__ static __forceinline__ Br branch(float sa, float par) { return {0.5f * (sa + par), 0.5f * (sa - par)}; }
  __device__ static __forceinline__ void norm(float* v) {
    const float m = max8(v);
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = v[i] - m;
  }
  // renormalised every step, as the reference does (nrm ignored)
  __device__ static __forceinline__ void bstep(float* b, const Br& r, bool /*nrm*/ = true) {
    float nb[8];
#pragma unroll
    for (unsigned p = 0; p < 8; ++p) {
      const Tr t0 = rsc(p, 0), t1 = rsc(p, 1);
      nb[p] = jac_ref(b[t0.next] + gsel(0, t0.parity, r.g1, r.g2), b[t1.next] + gsel(1, t1.parity, r.g1, r.g2));
    }
    norm(nb);
#pragma unroll
    for (int i = 0; i < 8; ++i) b[i] = nb[i];
  }
  template <bool WANT>
  __device__ static __forceinline__ float astep(float* a, const Br& r, float /*sa*/, float sys, float ap,
                                                const float* bn, bool /*nrm*/ = true) {
    float ext = 0.0f;
    if constexpr (WANT) {
      float l[2];
#pragma unroll
      for (unsigned s = 0; s < 8; ++s) {
#pragma unroll
        for (unsigned u = 0; u < 2; ++u) {
          const Tr tr = rsc(s, u);
          const float m = (a[s] + gsel(u, tr.parity, r.g1, r.g2)) + bn[tr.next];
          l[u] = (s == 0) ? m : jac_ref(l[u], m);
        }
      }
      ext = ((l[1] - l[0]) - sys) - ap;
    }
    float na[8];
#pragma unroll
    for (unsigned n = 0; n < 8; ++n) {
      const unsigned u = in_of(n), p = pred0(n), c = par0(n);
      na[n] = jac_ref(a[p] + gsel(u, c, r.g1, r.g2), a[p + 1] + gsel(u, c ^ 1u, r.g1, r.g2));
    }
    norm(na);
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = na[i];
    return ext;
  }
};

// A job on NW waves of one work-group (code blocks of more than 64
// segments): D's arithmetic, with the job's lanes numbered 0 .. 64 NW - 1
// across the waves.  Every row of the job's LDS extrinsics, of its staged
// inputs and of its posteriors then holds 64 NW columns.  Waves are not in
// lockstep, so the boundary metrics move between neighbouring segments
// through their records in memory instead of DPP, and the passes that read a
// neighbour's extrinsics finish before any lane overwrites its own.
template <class D, int NW>
struct Wide : D {
  static constexpr int kWaves = NW;
};

// Tail beta in the log domain, the reference's rules (lte_turbo.cpp:357-378):
// every state's forced three-step path to state 0, renormalised to max 0.
__device__ __forceinline__ void tail_beta(float ts0, float ts1, float ts2, float tp0, float tp1, float tp2, float* beta) {
  const float ts[3] = {ts0, ts1, ts2}, tp[3] = {tp0, tp1, tp2};
#pragma unroll
  for (unsigned s = 0; s < 8; ++s) {
    unsigned cur = s;
    float metric = 0.0f;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const unsigned u = term_in(cur);
      const Tr tr = rsc(cur, u);
      metric += 0.5f * ((u ? 1.0f : -1.0f) * ts[t]) + 0.5f * ((tr.parity ? 1.0f : -1.0f) * tp[t]);
      cur = tr.next;
    }
    beta[s] = metric;  // every state reaches 0 in three forced steps
  }
  const float m = max8(beta);
#pragma unroll
  for (int i = 0; i < 8; ++i) beta[i] -= m;
}

struct Lane {
  uint32_t S, M, g, nW, K;
  uint32_t q, colb, le;
  bool first, last, active;
  uint32_t f1, f2, qtab_off;
};

// (A_q + B_q t) mod M: the per-lane column offset of pi(q S + t).
__device__ __forceinline__ uint32_t qpp_col0(const Lane& L, uint32_t qq, uint32_t t) {
  const uint32_t A = L.f1 * qq + L.f2 * qq * qq * L.S;
  const uint32_t B = 2u * L.f2 * qq;
  return (A % L.M + (B % L.M) * t) % L.M;
}
__device__ __forceinline__ uint32_t qpp_stride(const Lane& L, uint32_t qq) { return (2u * L.f2 * qq) % L.M; }

// The per-K table is read-only for the whole launch: read through the
// constant address space, so that a wave-uniform index becomes a scalar load
// (waited on with lgkmcnt) instead of a vector load, whose wait would also
// drain every prefetched vector load behind it.
typedef __attribute__((address_space(4))) const uint32_t ConstU32;
__device__ __forceinline__ uint32_t qtab_at(const uint32_t* qtab, uint32_t i) {
  return ((ConstU32*)qtab)[i];
}

// u: the table entry of step t (lower half only); R: entries per row (the
// job's lanes)
template <int H, int R = kExRow>
__device__ __forceinline__ uint32_t lds_addr_q(const Lane& L, uint32_t u, uint32_t qq, uint32_t t, uint32_t y) {
  if constexpr (H == 0) {
    return t * R + L.colb + qq;
  } else {
    uint32_t c = (u & 0xFFFFu) + y;
    c = (c >= L.M) ? c - L.M : c;
    return (u >> 16) * R + L.colb + c;
  }
}
template <int H, int R = kExRow>
__device__ __forceinline__ uint32_t lds_addr(const Lane& L, const uint32_t* __restrict__ qtab, uint32_t qq,
                                             uint32_t t, uint32_t y) {
  return lds_addr_q<H, R>(L, H == 1 ? qtab_at(qtab, L.qtab_off + t) : 0u, qq, t, y);
}

// Extrinsics held in LDS between half-iterations: X = _Float16 (saturated
// at +-65504 and rounded to nearest even, v_cvt_f16_f32 in the default
// rounding mode) or float (held exactly).
template <class X>
__device__ __forceinline__ X to_ex(float v) {
  if constexpr (sizeof(X) == 2) {
    return static_cast<X>(__builtin_amdgcn_fmed3f(v, -65504.0f, 65504.0f));
  } else {
    return v;
  }
}

__device__ __forceinline__ uint32_t y_back(uint32_t y, uint32_t b, uint32_t m) { return (y >= b) ? y - b : y + m - b; }
__device__ __forceinline__ uint32_t y_fwd(uint32_t y, uint32_t b, uint32_t m) {
  y += b;
  return (y >= m) ? y - m : y;
}

// Loads of one chunk of up to kSubWin consecutive steps [lo, lo+len) of
// segment qq: the staged channel pair from the wave's scratch (global) and
// the extrinsic from LDS.  Each pass issues the next chunk's loads before it
// computes the current one.  FULL: len == kSubWin, no per-step guards.
struct Chunk {
  float2 v[kSubWin];
  float e[kSubWin];
};

typedef __attribute__((address_space(1))) const float GlobalFloat;
// A half's staged inputs are two planes of [t][lane] floats, kPlane apart:
// systematic (+ a-priori source) and parity.  A job on NW waves has rows of
// 64 NW floats and planes NW times as large (plane_of<R>).
constexpr uint32_t kPlane = kMaxSeg * kWaveLanes;
template <int R>
constexpr uint32_t plane_of() {
  return kMaxSeg * R;
}

template <bool FULL = false, int R = kWaveLanes>
__device__ __forceinline__ void load_chunk(const float* __restrict__ sc, uint32_t col, uint32_t lo, uint32_t len,
                                           Chunk& c) {
  const GlobalFloat* base = (const GlobalFloat*)(sc + lo * R + col);
  // Opaque to the optimiser: without this the per-lane 64-bit address of
  // every (chunk, step, column) combination is hoisted out of the loops and
  // spilled.  One address pair per chunk; steps use immediate offsets.  The
  // pointer stays in the global address space: flat loads would also count
  // against the LDS wait counter, so every LDS wait of the sub-window would
  // drain the next sub-window's prefetch.
  asm volatile("" : "+v"(base));
#pragma unroll
  for (int j = 0; j < kSubWin; ++j)
    if (FULL || (uint32_t)j < len) c.v[j] = make_float2(base[j * R], base[plane_of<R>() + j * R]);
}

// LDS extrinsics of the chunk, issued in the order the pass consumes them;
// addr (may be null) keeps the addresses for the write-back.
template <int H, bool BACKWARD, bool FULL, class X, int R = kExRow>
__device__ __forceinline__ void load_chunk_lds(const Lane& L, const X* __restrict__ ex,
                                               const uint32_t* __restrict__ qtab, uint32_t qq, uint32_t by,
                                               uint32_t lo, uint32_t len, Chunk& c, uint32_t* addr) {
  uint32_t y = (H == 1) ? qpp_col0(L, qq, BACKWARD ? lo + len - 1 : lo) : 0u;
  // the chunk's kSubWin table entries first, in one scalar batch (the table
  // is padded by kSubWin entries), then the LDS reads back to back
  uint32_t qv[kSubWin];
  if constexpr (H == 1) {
    // opaque per chunk, so the compiler does not hoist (and then spill) the
    // table addresses of every chunk out of the loops
    uint32_t off = __builtin_amdgcn_readfirstlane(L.qtab_off + lo);  // wave-uniform
    asm volatile("" : "+s"(off));
    const uint32_t* qt = qtab + off;
#pragma unroll
    for (int j = 0; j < kSubWin; ++j) qv[j] = qtab_at(qt, j);
  }
#pragma unroll
  for (int jj = 0; jj < kSubWin; ++jj) {
    const int j = BACKWARD ? kSubWin - 1 - jj : jj;
    if (FULL || (uint32_t)j < len) {
      const uint32_t a = lds_addr_q<H, R>(L, qv[j], qq, lo + j, y);
      c.e[j] = static_cast<float>(ex[a]);
      if (addr) addr[j] = a;
      if constexpr (H == 1) y = BACKWARD ? y_back(y, by, L.M) : y_fwd(y, by, L.M);
    }
  }
}

// Backward recursion of beta over steps [lo, hi) of segment qq (chunks of
// kSubWin aligned at lo, last chunk first).  on_chunk(c, beta) runs before
// chunk c; on_hand(beta) right after step hand_t (beta then is beta_hand_t;
// hand_t outside [lo, hi): never).  A chunk of kSubWin steps that does not
// hold hand_t strictly inside runs without per-step guards.
template <class D, int H, class X, class OnChunk, class OnHand>
__device__ __forceinline__ void beta_pass(const Lane& L, const X* __restrict__ ex,
                                          const float* __restrict__ sc, const uint32_t* __restrict__ qtab,
                                          uint32_t qq, uint32_t col, uint32_t lo, uint32_t hi, float* bc,
                                          uint32_t hand_t, OnChunk&& on_chunk, OnHand&& on_hand) {
  if (hi <= lo) return;
  constexpr int R = kWaveLanes * D::kWaves;
  const uint32_t n = (hi - lo + kSubWin - 1) / kSubWin;
  const uint32_t by = (H == 1) ? qpp_stride(L, qq) : 0u;
  Chunk nxt;
  {
    const uint32_t c0 = lo + (n - 1) * kSubWin;
    load_chunk<false, R>(sc, col, c0, hi - c0, nxt);
    load_chunk_lds<H, true, false, X, R>(L, ex, qtab, qq, by, c0, hi - c0, nxt, nullptr);
  }
  for (int c = (int)n - 1; c >= 0; --c) {
    if constexpr (D::kShareSimd) share_simd();
    const uint32_t clo = lo + c * kSubWin;
    const uint32_t len = min((uint32_t)kSubWin, hi - clo);
    const Chunk cur = nxt;
    if (c > 0) {
      load_chunk<true, R>(sc, col, clo - kSubWin, kSubWin, nxt);
      load_chunk_lds<H, true, true, X, R>(L, ex, qtab, qq, by, clo - kSubWin, kSubWin, nxt, nullptr);
    }
    on_chunk((uint32_t)c, bc);
    if (len == kSubWin && (hand_t <= clo || hand_t >= clo + kSubWin)) {
#pragma unroll
      for (int j = kSubWin - 1; j >= 0; --j) {
        __builtin_amdgcn_sched_barrier(0);
        D::bstep(bc, D::branch(cur.v[j].x + cur.e[j], cur.v[j].y), j == 0);
      }
      if (hand_t == clo) on_hand(bc);
    } else {
#pragma unroll
      for (int j = kSubWin - 1; j >= 0; --j) {
        if ((uint32_t)j < len) {
          __builtin_amdgcn_sched_barrier(0);
          D::bstep(bc, D::branch(cur.v[j].x + cur.e[j], cur.v[j].y), j == 0);
          if (clo + j == hand_t) on_hand(bc);
        }
      }
    }
  }
}

// Forward recursion of alpha without extrinsics (the warm-up) over steps
// [lo, hi) of segment qq, chunks of kSubWin aligned at lo.
template <class D, int H, class X>
__device__ __forceinline__ void alpha_warmup(const Lane& L, const X* __restrict__ ex,
                                             const float* __restrict__ sc, const uint32_t* __restrict__ qtab,
                                             uint32_t qq, uint32_t col, uint32_t lo, uint32_t hi, float* ac) {
  if (hi <= lo) return;
  constexpr int R = kWaveLanes * D::kWaves;
  const uint32_t by = (H == 1) ? qpp_stride(L, qq) : 0u;
  const uint32_t n = (hi - lo + kSubWin - 1) / kSubWin;
  Chunk nxt;
  load_chunk<false, R>(sc, col, lo, min((uint32_t)kSubWin, hi - lo), nxt);
  load_chunk_lds<H, false, false, X, R>(L, ex, qtab, qq, by, lo, min((uint32_t)kSubWin, hi - lo), nxt, nullptr);
  for (uint32_t c = 0; c < n; ++c) {
    const uint32_t clo = lo + c * kSubWin;
    const uint32_t len = min((uint32_t)kSubWin, hi - clo);
    const Chunk cur = nxt;
    if (c + 1 < n) {
      const uint32_t nlen = min((uint32_t)kSubWin, hi - clo - kSubWin);
      load_chunk<false, R>(sc, col, clo + kSubWin, nlen, nxt);
      load_chunk_lds<H, false, false, X, R>(L, ex, qtab, qq, by, clo + kSubWin, nlen, nxt, nullptr);
    }
    if (len == kSubWin) {
#pragma unroll
      for (int j = 0; j < kSubWin; ++j) {
        __builtin_amdgcn_sched_barrier(0);
        const float sa = cur.v[j].x + cur.e[j];
        D::template astep<false>(ac, D::branch(sa, cur.v[j].y), sa, 0.0f, 0.0f, nullptr, j == kSubWin - 1);
      }
    } else {
#pragma unroll
      for (int j = 0; j < kSubWin; ++j) {
        if ((uint32_t)j < len) {
          __builtin_amdgcn_sched_barrier(0);
          const float sa = cur.v[j].x + cur.e[j];
          D::template astep<false>(ac, D::branch(sa, cur.v[j].y), sa, 0.0f, 0.0f, nullptr, j == kSubWin - 1);
        }
      }
    }
  }
}

// CRC24B early stop (lte_turbo.cpp:792-804) state of this lane's block.  A
// block whose check passed is done: its lanes stop writing (their stores go
// to per-lane dummy slots), so its posteriors stay those of the iteration
// that passed.
struct EarlyStop {
  bool on;
  bool done;
};
static_assert(kExRow == 64, "extrinsic address = row * 64 + column");

// Sub-window w of the forward pass: recompute beta from the checkpoint bb,
// then alpha + extrinsic per step.  Writes go to index wa(j) = the bit's
// extrinsic address, or a per-lane dummy slot for lanes that must not write:
// the fp16 extrinsic to LDS (upper half: E1 of the bit; lower: E2) and, in
// the lower half, the fp32 posterior to post_sc.
template <class D, int H, bool ES, bool FULL, class X>
__device__ __forceinline__ void alpha_subwindow(const Lane& L, X* __restrict__ ex, const Chunk& cur,
                                                const uint32_t* addr, float* bb, float* ac, uint32_t t0,
                                                uint32_t len, uint32_t hand_t, float* __restrict__ nii,
                                                uint32_t dummy, bool wr, float* __restrict__ post_sc,
                                                const EarlyStop& es) {
  float bs[kSubWin][8];
  typename D::Br br[kSubWin];
  float isa[kSubWin];
#pragma unroll
  for (int j = kSubWin - 1; j >= 0; --j) {
    if (FULL || (uint32_t)j < len) {
      isa[j] = cur.v[j].x + cur.e[j];
      br[j] = D::branch(isa[j], cur.v[j].y);
#pragma unroll
      for (int i = 0; i < 8; ++i) bs[j][i] = bb[i];
      __builtin_amdgcn_sched_barrier(0);
      if (j > 0) D::bstep(bb, br[j], false);  // steps 1.. of the chunk: not renormalised in the first pass either
    }
  }
  auto hand = [&]() {
    // uniform: hand alpha at step S-g to the next segment
    if constexpr (D::kWaves == 1) {
      float na[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) na[i] = dpp_from_prev_lane(ac[i]);
      if (!L.first) store8(nii, na);
    } else {
      if (L.active && !L.last) store8(nii, ac);  // nii: the next segment's record
    }
  };
  if (FULL && hand_t == t0) hand();
#pragma unroll
  for (int j = 0; j < kSubWin; ++j) {
    if (FULL || (uint32_t)j < len) {
      if (!FULL && t0 + j == hand_t) hand();
      __builtin_amdgcn_sched_barrier(0);  // keep steps in order: bounds register pressure
      const float ext = D::template astep<true>(ac, br[j], isa[j], cur.v[j].x, cur.e[j], bs[j], j == kSubWin - 1);
      const uint32_t a = wr ? addr[j] : dummy;
      ex[a] = to_ex<X>(ext);
      if constexpr (H == 1) {
        const float p = isa[j] + ext;
        post_sc[a] = p;
      }
    }
  }
}

// Per sub-window of this lane's segment, from alpha ac at its first step:
// recompute beta from the checkpoint ck[w] (beta after the sub-window's last
// step), then alpha + extrinsic per step (alpha_subwindow).  hand_t: the step
// whose alpha goes to the next segment's record (kNever: none).
template <class D, int H, bool ES, class X>
__device__ __forceinline__ void extrinsic_pass(const Lane& L, X* __restrict__ ex, const float* __restrict__ sc,
                                               const uint32_t* __restrict__ qtab, const float* __restrict__ ck,
                                               float* ac, uint32_t hand_t, float* __restrict__ nii,
                                               float* __restrict__ post_sc, const EarlyStop& es) {
  constexpr int R = kWaveLanes * D::kWaves;
  const uint32_t S = L.S;
  const uint32_t lane = threadIdx.x;
  const uint32_t by = (H == 1) ? qpp_stride(L, L.q) : 0u;
  const uint32_t dummy = S * R + lane;
  const bool wr = L.active && !es.done;
  // global loads (staged LLRs, beta checkpoint) run one sub-window ahead;
  // the LDS extrinsics of a sub-window are read at its start
  Chunk nxt;
  float nck[8];
  load_chunk<false, R>(sc, L.le, 0, min((uint32_t)kSubWin, S), nxt);
  load8(ck, nck);
  for (uint32_t w = 0; w < L.nW; ++w) {
    if constexpr (D::kShareSimd) share_simd();
    const uint32_t t0 = w * kSubWin;
    const uint32_t len = min((uint32_t)kSubWin, S - t0);
    Chunk cur;
#pragma unroll
    for (int j = 0; j < kSubWin; ++j) cur.v[j] = nxt.v[j];
    float bb[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) bb[i] = nck[i];
    if (w + 1 < L.nW) {
      load_chunk<false, R>(sc, L.le, t0 + kSubWin, min((uint32_t)kSubWin, S - t0 - kSubWin), nxt);
      load8(ck + 8 * (w + 1), nck);
    }
    uint32_t addr[kSubWin];
    if (len == kSubWin && (hand_t <= t0 || hand_t >= t0 + kSubWin)) {
      load_chunk_lds<H, true, true, X, R>(L, ex, qtab, L.q, by, t0, len, cur, addr);
      alpha_subwindow<D, H, ES, true>(L, ex, cur, addr, bb, ac, t0, len, hand_t, nii, dummy, wr, post_sc, es);
    } else {
      load_chunk_lds<H, true, false, X, R>(L, ex, qtab, L.q, by, t0, len, cur, addr);
      alpha_subwindow<D, H, ES, false>(L, ex, cur, addr, bb, ac, t0, len, hand_t, nii, dummy, wr, post_sc, es);
    }
  }
}

// One SISO half (H = 0 upper, 1 lower) for this lane's segment.
// nii: this lane's boundary metrics of this half, [alpha 8 | beta 8]
// (alpha at step -g, beta at step S+g of the segment); ck: this lane's beta
// checkpoints at sub-window boundaries, [w][8].  The lower half also leaves
// each bit's posterior in post_sc (fp32, at the bit's extrinsic address).
// A job on several waves (D::kWaves > 1) reads its boundary metrics from nii
// and hands its own to the neighbours' records at push (the records of the
// next iteration, kRecFloats apart per lane, at this half's offset).
constexpr uint32_t kRecFloats = 32;  // per-lane record: [half][alpha 8 | beta 8]
template <class D, int H, bool ES, class X>
__device__ __forceinline__ void siso_half(const Lane& L, X* __restrict__ ex, const float* __restrict__ sc,
                                          const uint32_t* __restrict__ qtab, float* __restrict__ nii,
                                          float* __restrict__ ck, float* __restrict__ post_sc,
                                          const EarlyStop& es, float* __restrict__ push = nullptr) {
  const uint32_t S = L.S, g = L.g;
  const uint32_t lane = threadIdx.x;
  constexpr uint32_t kNever = 0xFFFFFFFFu;
  // where this lane's hand-offs go: alpha to the next segment, beta to the
  // previous one (one wave: its own record, the value pulled by DPP)
  float* a_to = nii;
  float* b_to = nii + 8;
  if constexpr (D::kWaves > 1) {
    a_to = L.last ? nii : push + (L.le + 1) * kRecFloats;
    b_to = L.first ? nii + 8 : push + (L.le - 1) * kRecFloats + 8;
  }
  // ---------------- backward pass: warm-up over the next segment's first g
  // steps, then the segment, storing beta at sub-window boundaries
  float bc[8];
  load8(nii + 8, bc);
  if (g > 0) {
    const uint32_t qr = L.last ? L.q : L.q + 1;
    const uint32_t col = L.last ? L.le : L.le + 1;
    beta_pass<D, H>(L, ex, sc, qtab, qr, col, 0, g, bc, kNever, [](uint32_t, float*) {}, [](const float*) {});
    if (L.last) load8(nii + 8, bc);
  }
  beta_pass<D, H>(
      L, ex, sc, qtab, L.q, L.le, 0, S, bc, g > 0 ? g : kNever,
      [&](uint32_t w, float* b) { store8(ck + 8 * w, b); },
      [&](const float* b) {
        // the whole wave hands beta at step g to the previous segment (used
        // next iteration); the last segment keeps its tail beta
        if constexpr (D::kWaves == 1) {
          float nb[8];
#pragma unroll
          for (int i = 0; i < 8; ++i) nb[i] = dpp_from_next_lane(b[i]);
          if (!L.last) store8(b_to, nb);
        } else {
          if (L.active && !L.first) store8(b_to, b);
        }
      });
  // ---------------- forward pass: warm-up over the previous segment's last g
  // steps, then per sub-window: recompute beta, alpha + extrinsic
  float ac[8];
  load8(nii, ac);
  if (g > 0) {
    const uint32_t ql = L.first ? L.q : L.q - 1;
    const uint32_t col = L.first ? L.le : L.le - 1;
    alpha_warmup<D, H>(L, ex, sc, qtab, ql, col, S - g, S, ac);
    if (L.first) load8(nii, ac);
  }
  // several waves: every warm-up has read its neighbours' extrinsics before
  // any lane overwrites its own (one wave runs in lockstep)
  if constexpr (D::kWaves > 1) __syncthreads();
  extrinsic_pass<D, H, ES>(L, ex, sc, qtab, ck, ac, g > 0 ? S - g : kNever, a_to, post_sc, es);
}

// ------------------------------------------------------------- lean path --
// The geometry of the long blocks: one block per wave in M = 64 segments
// (every K >= 2048; S and the warm-up g multiples of kSubWin at K a multiple
// of 512 and warm-ups 8, 16, 24, ...).  Every
// chunk is then full and every hand-off falls on a chunk edge, so the passes
// below run without per-step guards; a pass's chunks alternate between two
// register sets (no copies at the loop edge); and the lower half's extrinsic
// column (c_t + A_q + B_q t) mod 64 needs no conditional subtract.  The
// arithmetic is siso_half's, step for step, so the results are bit-identical.

// One segment's extrinsic addressing as this lane sees it (its own or a
// neighbour's), in bytes of the LDS rows of X (XB bytes per entry): the upper
// half reads column q, byte (t * 64 + q) XB; the lower half reads bit
// pi(q S + t) at row r_t, column (c_t + A_q + B_q t) mod 64 (r_t, c_t from
// the per-K table), byte ((y + XB c_t) & (64 XB - 1)) | r_t 64 XB with
// y = XB (A_q + B_q t) kept unreduced: 2^32 is a multiple of 64 XB.
struct LeanSeg {
  uint32_t col;  // column of the segment in the staged planes
  uint32_t qb;   // XB q
  uint32_t y0;   // XB A_q (mod 2^32)
  uint32_t dy;   // XB B_q (mod 2^32)
};

template <class X>
__device__ __forceinline__ LeanSeg lean_seg(const Lane& L, uint32_t qq, uint32_t col) {
  constexpr uint32_t XB = sizeof(X);
  return {col, XB * qq, XB * (L.f1 * qq + L.f2 * qq * qq * L.S), XB * (2u * L.f2 * qq)};
}

// The extrinsics of chunk [lo, lo + kSubWin) of segment G, issued in the
// order the pass consumes them (BACKWARD: last step first); ab (may be null)
// keeps their byte addresses for the write-back.
template <int H, bool BACKWARD, class X>
__device__ __forceinline__ void lean_lds(const X* __restrict__ ex, const uint32_t* __restrict__ qtab, const Lane& L,
                                         const LeanSeg& G, uint32_t lo, Chunk& c, uint32_t* ab) {
  constexpr uint32_t XB = sizeof(X);
  constexpr uint32_t kRowB = kExRow * XB;
  const char* base = reinterpret_cast<const char*>(ex);
  if constexpr (H == 0) {
    const uint32_t a0 = lo * kRowB + G.qb;
#pragma unroll
    for (int jj = 0; jj < kSubWin; ++jj) {
      const int j = BACKWARD ? kSubWin - 1 - jj : jj;
      const uint32_t a = a0 + j * kRowB;
      c.e[j] = static_cast<float>(*reinterpret_cast<const X*>(base + a));
      if (ab) ab[j] = a;
    }
  } else {
    // the chunk's table entries in one scalar batch (opaque per chunk, as in
    // load_chunk_lds, so the addresses of every chunk are not hoisted)
    uint32_t off = __builtin_amdgcn_readfirstlane(L.qtab_off + lo);
    asm volatile("" : "+s"(off));
    const uint32_t* qt = qtab + off;
    uint32_t qv[kSubWin];
#pragma unroll
    for (int j = 0; j < kSubWin; ++j) qv[j] = qtab_at(qt, j);
    uint32_t y = G.y0 + (lo + (BACKWARD ? kSubWin - 1 : 0)) * G.dy;
#pragma unroll
    for (int jj = 0; jj < kSubWin; ++jj) {
      const int j = BACKWARD ? kSubWin - 1 - jj : jj;
      const uint32_t a = ((y + XB * (qv[j] & 0xFFFFu)) & (kRowB - 1u)) | ((qv[j] >> 16) * kRowB);
      c.e[j] = static_cast<float>(*reinterpret_cast<const X*>(base + a));
      if (ab) ab[j] = a;
      y = BACKWARD ? y - G.dy : y + G.dy;
    }
  }
}

template <int H, bool BACKWARD, class X>
__device__ __forceinline__ void lean_load(const X* __restrict__ ex, const float* __restrict__ sc,
                                          const uint32_t* __restrict__ qtab, const Lane& L, const LeanSeg& G,
                                          uint32_t c, Chunk& x) {
  load_chunk<true>(sc, G.col, c * kSubWin, kSubWin, x);
  lean_lds<H, BACKWARD>(ex, qtab, L, G, c * kSubWin, x, nullptr);
}

// beta down chunks n-1 .. 0 of segment G from bc; on_chunk(c, beta) before
// chunk c, on_hand(beta) after chunk hand_c (beta then is beta at its first step)
template <class D, int H, class X, class OnChunk, class OnHand>
__device__ __forceinline__ void lean_beta(const Lane& L, const X* __restrict__ ex, const float* __restrict__ sc,
                                          const uint32_t* __restrict__ qtab, const LeanSeg& G, uint32_t n, float* bc,
                                          uint32_t hand_c, OnChunk&& on_chunk, OnHand&& on_hand) {
  auto run = [&](uint32_t c, const Chunk& x) {
    if constexpr (D::kShareSimd) share_simd();
    on_chunk(c, bc);
#pragma unroll
    for (int j = kSubWin - 1; j >= 0; --j) {
      __builtin_amdgcn_sched_barrier(0);
      D::bstep(bc, D::branch(x.v[j].x + x.e[j], x.v[j].y), j == 0);
    }
    if (c == hand_c) on_hand(bc);
  };
  Chunk P, Q;
  lean_load<H, true>(ex, sc, qtab, L, G, n - 1, P);
  for (uint32_t c = n - 1;; c -= 2) {
    if (c >= 1) lean_load<H, true>(ex, sc, qtab, L, G, c - 1, Q);
    run(c, P);
    if (c == 0) break;
    if (c >= 2) lean_load<H, true>(ex, sc, qtab, L, G, c - 2, P);
    run(c - 1, Q);
    if (c == 1) break;
  }
}

// alpha up chunks c0 .. c0+n-1 of segment G (the warm-up: no extrinsics)
template <class D, int H, class X>
__device__ __forceinline__ void lean_alpha(const Lane& L, const X* __restrict__ ex, const float* __restrict__ sc,
                                           const uint32_t* __restrict__ qtab, const LeanSeg& G, uint32_t c0,
                                           uint32_t n, float* ac) {
  auto run = [&](const Chunk& x) {
#pragma unroll
    for (int j = 0; j < kSubWin; ++j) {
      __builtin_amdgcn_sched_barrier(0);
      const float sa = x.v[j].x + x.e[j];
      D::template astep<false>(ac, D::branch(sa, x.v[j].y), sa, 0.0f, 0.0f, nullptr, j == kSubWin - 1);
    }
  };
  Chunk P, Q;
  lean_load<H, false>(ex, sc, qtab, L, G, c0, P);
  for (uint32_t c = 0;; c += 2) {
    if (c + 1 < n) lean_load<H, false>(ex, sc, qtab, L, G, c0 + c + 1, Q);
    run(P);
    if (c + 1 >= n) break;
    if (c + 2 < n) lean_load<H, false>(ex, sc, qtab, L, G, c0 + c + 2, P);
    run(Q);
    if (c + 2 >= n) break;
  }
}

// Sub-window of the forward pass (alpha_subwindow's full-chunk path with byte
// addresses ab): recompute beta from the checkpoint bb, then alpha +
// extrinsic per step; the extrinsic to LDS, in the lower half also the
// posterior to post_sc (float index = byte address / XB).
template <class D, int H, class X>
__device__ __forceinline__ void lean_subwindow(const Lane& L, X* __restrict__ ex, const float2* v, const float* e,
                                               const uint32_t* ab, float* bb, float* ac, bool hand,
                                               float* __restrict__ nii, uint32_t dummy_b, bool wr,
                                               float* __restrict__ post_sc) {
  constexpr uint32_t XB = sizeof(X);
  float bs[kSubWin][8];
  typename D::Br br[kSubWin];
  float isa[kSubWin];
#pragma unroll
  for (int j = kSubWin - 1; j >= 0; --j) {
    isa[j] = v[j].x + e[j];
    br[j] = D::branch(isa[j], v[j].y);
#pragma unroll
    for (int i = 0; i < 8; ++i) bs[j][i] = bb[i];
    __builtin_amdgcn_sched_barrier(0);
    if (j > 0) D::bstep(bb, br[j], false);
  }
  if (hand) {
    // uniform: hand alpha at step S-g to the next segment
    float na[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) na[i] = dpp_from_prev_lane(ac[i]);
    if (!L.first) store8(nii, na);
  }
  char* exb = reinterpret_cast<char*>(ex);
  char* pb = reinterpret_cast<char*>(post_sc);
#pragma unroll
  for (int j = 0; j < kSubWin; ++j) {
    __builtin_amdgcn_sched_barrier(0);
    const float ext = D::template astep<true>(ac, br[j], isa[j], v[j].x, e[j], bs[j], j == kSubWin - 1);
    const uint32_t a = wr ? ab[j] : dummy_b;
    *reinterpret_cast<X*>(exb + a) = to_ex<X>(ext);
    if constexpr (H == 1) *reinterpret_cast<float*>(pb + a * (4 / XB)) = isa[j] + ext;
  }
}

// Forward pass of the lane's own segment G: per sub-window, the staged pair
// and the beta checkpoint one sub-window ahead (alternating register sets),
// the extrinsics at its start; hand_c: the sub-window whose first alpha goes
// to the next segment's record.
template <class D, int H, bool ES, class X>
__device__ __forceinline__ void lean_extrinsic(const Lane& L, X* __restrict__ ex, const float* __restrict__ sc,
                                               const uint32_t* __restrict__ qtab, const LeanSeg& G,
                                               const float* __restrict__ ck, float* ac, uint32_t hand_c,
                                               float* __restrict__ nii, float* __restrict__ post_sc,
                                               const EarlyStop& es) {
  const uint32_t n = L.nW;
  const uint32_t dummy_b = (L.S * kExRow + threadIdx.x) * sizeof(X);
  const bool wr = L.active && !es.done;
  struct In {
    float2 v[kSubWin];
    float k[8];
  };
  auto load = [&](uint32_t w, In& in) {
    const GlobalFloat* p = (const GlobalFloat*)(sc + w * kSubWin * kWaveLanes + G.col);
    asm volatile("" : "+v"(p));
#pragma unroll
    for (int j = 0; j < kSubWin; ++j) in.v[j] = make_float2(p[j * kWaveLanes], p[kPlane + j * kWaveLanes]);
    load8(ck + 8 * w, in.k);
  };
  auto run = [&](uint32_t w, In& in) {
    if constexpr (D::kShareSimd) share_simd();
    Chunk x;
    uint32_t ab[kSubWin];
    lean_lds<H, true>(ex, qtab, L, G, w * kSubWin, x, ab);
    lean_subwindow<D, H>(L, ex, in.v, x.e, ab, in.k, ac, w == hand_c, nii, dummy_b, wr, post_sc);
  };
  In P, Q;
  load(0, P);
  for (uint32_t w = 0;; w += 2) {
    if (w + 1 < n) load(w + 1, Q);
    run(w, P);
    if (w + 1 >= n) break;
    if (w + 2 < n) load(w + 2, P);
    run(w + 1, Q);
    if (w + 2 >= n) break;
  }
}

// siso_half on the lean geometry (same records, same order of operations).
template <class D, int H, bool ES, class X>
__device__ __forceinline__ void siso_half_lean(const Lane& L, X* __restrict__ ex, const float* __restrict__ sc,
                                               const uint32_t* __restrict__ qtab, float* __restrict__ nii,
                                               float* __restrict__ ck, float* __restrict__ post_sc,
                                               const EarlyStop& es) {
  const uint32_t nW = L.nW, ng = L.g / kSubWin;
  constexpr uint32_t kNever = 0xFFFFFFFFu;
  const LeanSeg own = lean_seg<X>(L, L.q, L.le);
  // backward pass: warm-up over the next segment's first g steps, then the
  // segment, beta checkpoints at chunk starts, beta at step g handed back
  float bc[8];
  load8(nii + 8, bc);
  {
    const LeanSeg nx = L.last ? own : lean_seg<X>(L, L.q + 1, L.le + 1);
    lean_beta<D, H>(L, ex, sc, qtab, nx, ng, bc, kNever, [](uint32_t, float*) {}, [](const float*) {});
    if (L.last) load8(nii + 8, bc);
  }
  lean_beta<D, H>(
      L, ex, sc, qtab, own, nW, bc, ng, [&](uint32_t w, float* b) { store8(ck + 8 * w, b); },
      [&](const float* b) {
        float nb[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) nb[i] = dpp_from_next_lane(b[i]);
        if (!L.last) store8(nii + 8, nb);
      });
  // forward pass: warm-up over the previous segment's last g steps, then per
  // sub-window beta recomputed, alpha + extrinsic
  float ac[8];
  load8(nii, ac);
  {
    const LeanSeg pv = L.first ? own : lean_seg<X>(L, L.q - 1, L.le - 1);
    lean_alpha<D, H>(L, ex, sc, qtab, pv, nW - ng, ng, ac);
    if (L.first) load8(nii, ac);
  }
  lean_extrinsic<D, H, ES>(L, ex, sc, qtab, own, ck, ac, nW - ng, nii, post_sc, es);
}

// ---------------------------------------------- exact segment boundaries --
__device__ __forceinline__ bool wave_any(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0; }

// the 8 metrics are the same bit patterns
__device__ __forceinline__ bool same8(const float* a, const float* b) {
  uint32_t d = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) d |= __float_as_uint(a[i]) ^ __float_as_uint(b[i]);
  return d == 0;
}

// beta down / alpha up one chunk of len <= kSubWin steps
template <class D>
__device__ __forceinline__ void beta_chunk(const Chunk& c, uint32_t len, float* b) {
#pragma unroll
  for (int j = kSubWin - 1; j >= 0; --j) {
    if ((uint32_t)j < len) {
      __builtin_amdgcn_sched_barrier(0);
      D::bstep(b, D::branch(c.v[j].x + c.e[j], c.v[j].y));
    }
  }
}
template <class D>
__device__ __forceinline__ void alpha_chunk(const Chunk& c, uint32_t len, float* a) {
#pragma unroll
  for (int j = 0; j < kSubWin; ++j) {
    if ((uint32_t)j < len) {
      __builtin_amdgcn_sched_barrier(0);
      const float sa = c.v[j].x + c.e[j];
      D::template astep<false>(a, D::branch(sa, c.v[j].y), sa, 0.0f, 0.0f, nullptr);
    }
  }
}

// alpha over this lane's segment from ac, keeping ack[w] = alpha at the
// first step of chunk w; ac ends as alpha after the segment's last step.
template <class D, int H, class X>
__device__ __forceinline__ void alpha_pass_ck(const Lane& L, const X* __restrict__ ex, const float* __restrict__ sc,
                                              const uint32_t* __restrict__ qtab, float* ac, float* __restrict__ ack) {
  const uint32_t S = L.S, n = L.nW;
  const uint32_t by = (H == 1) ? qpp_stride(L, L.q) : 0u;
  Chunk nxt;
  load_chunk(sc, L.le, 0, min((uint32_t)kSubWin, S), nxt);
  load_chunk_lds<H, false, false>(L, ex, qtab, L.q, by, 0, min((uint32_t)kSubWin, S), nxt, nullptr);
  for (uint32_t c = 0; c < n; ++c) {
    if constexpr (D::kShareSimd) share_simd();
    const uint32_t clo = c * kSubWin;
    const uint32_t len = min((uint32_t)kSubWin, S - clo);
    const Chunk cur = nxt;
    if (c + 1 < n) {
      const uint32_t nlen = min((uint32_t)kSubWin, S - clo - kSubWin);
      load_chunk(sc, L.le, clo + kSubWin, nlen, nxt);
      load_chunk_lds<H, false, false>(L, ex, qtab, L.q, by, clo + kSubWin, nlen, nxt, nullptr);
    }
    store8(ack + 8 * c, ac);
    alpha_chunk<D>(cur, len, ac);
  }
}

// One SISO half with exact segment boundaries (the reference's recursions,
// lte_turbo.cpp:410-486, computed segment-parallel; or_decode_exact_seg
// states the schedule serially):
//  1. speculation: beta down and alpha up this lane's segment from its
//     record nii (the exact boundary metrics of the previous iteration;
//     uniform at first; alpha_0 and the tail beta are exact), keeping
//     checkpoints ck[w] (beta after chunk w) and ack[w] (alpha before it);
//  2. fix-up rounds: each lane takes its neighbours' edge metrics as they
//     stand (wave_shr / wave_shl) and, where one differs from the value its
//     own pass started from, re-runs chunk by chunk until its metrics equal
//     the stored checkpoint bit for bit -- from there on the recursion is the
//     same function of the same inputs, so nothing further changes.  A lane
//     that never meets its checkpoints has a new edge value; rounds repeat
//     until no edge changes.  Every checkpoint then holds the metric of the
//     full-length serial recursion;
//  3. extrinsics per chunk from the checkpoints (extrinsic_pass).
// The extrinsics held between halves are fp32, so the decoder computes the
// serial decoder's numbers bit for bit.
template <int H, bool ES, class X>
__device__ __forceinline__ void siso_half_exact(const Lane& L, X* __restrict__ ex, const float* __restrict__ sc,
                                                const uint32_t* __restrict__ qtab, float* __restrict__ nii,
                                                float* __restrict__ ck, float* __restrict__ ack,
                                                float* __restrict__ post_sc, const EarlyStop& es) {
  using D = ExactDomain;
  const uint32_t S = L.S, nW = L.nW;
  constexpr uint32_t kNever = 0xFFFFFFFFu;
  const uint32_t by = (H == 1) ? qpp_stride(L, L.q) : 0u;
  // 1. speculation: beta down and alpha up in one loop, one chunk of each per
  // trip (two independent chains), inputs loaded a trip ahead
  float bc[8], ac[8];
  load8(nii + 8, bc);
  load8(nii, ac);
  {
    Chunk nb, na;
    const uint32_t lb0 = S - (nW - 1) * kSubWin, la0 = min((uint32_t)kSubWin, S);
    load_chunk(sc, L.le, (nW - 1) * kSubWin, lb0, nb);
    load_chunk_lds<H, true, false>(L, ex, qtab, L.q, by, (nW - 1) * kSubWin, lb0, nb, nullptr);
    load_chunk(sc, L.le, 0, la0, na);
    load_chunk_lds<H, false, false>(L, ex, qtab, L.q, by, 0, la0, na, nullptr);
    for (uint32_t c = 0; c < nW; ++c) {
      if constexpr (D::kShareSimd) share_simd();
      const uint32_t wb = nW - 1 - c;
      const uint32_t lb = min((uint32_t)kSubWin, S - wb * kSubWin), la = min((uint32_t)kSubWin, S - c * kSubWin);
      const Chunk xb = nb, xa = na;
      if (c + 1 < nW) {
        const uint32_t wb1 = wb - 1, wa1 = c + 1;
        const uint32_t lb1 = min((uint32_t)kSubWin, S - wb1 * kSubWin), la1 = min((uint32_t)kSubWin, S - wa1 * kSubWin);
        load_chunk(sc, L.le, wb1 * kSubWin, lb1, nb);
        load_chunk_lds<H, true, false>(L, ex, qtab, L.q, by, wb1 * kSubWin, lb1, nb, nullptr);
        load_chunk(sc, L.le, wa1 * kSubWin, la1, na);
        load_chunk_lds<H, false, false>(L, ex, qtab, L.q, by, wa1 * kSubWin, la1, na, nullptr);
      }
      store8(ck + 8 * wb, bc);
      store8(ack + 8 * c, ac);
#pragma unroll
      for (int j = 0; j < kSubWin; ++j) {
        const int jb = kSubWin - 1 - j;
        __builtin_amdgcn_sched_barrier(0);
        if ((uint32_t)jb < lb) D::bstep(bc, D::branch(xb.v[jb].x + xb.e[jb], xb.v[jb].y));
        if ((uint32_t)j < la) {
          const float sa = xa.v[j].x + xa.e[j];
          D::template astep<false>(ac, D::branch(sa, xa.v[j].y), sa, 0.0f, 0.0f, nullptr);
        }
      }
    }
  }
  // 2. fix-up rounds (idle lanes and the exact ends never need one).  After
  // round r the segments within r of an exact end hold exact metrics, so
  // M + 1 rounds always suffice; the bound only guarantees every wave leaves.
  const bool may_b = L.active && !L.last, may_a = L.active && !L.first;
  for (uint32_t round = 0; round <= L.M; ++round) {
    float b[8], a[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      b[i] = dpp_from_next_lane(bc[i]);
      a[i] = dpp_from_prev_lane(ac[i]);
    }
    bool db = may_b, da = may_a;
    for (uint32_t c = 0; c < nW; ++c) {
      if constexpr (D::kShareSimd) share_simd();
      const uint32_t wb = nW - 1 - c;
      if (db) {
        float k8[8];
        load8(ck + 8 * wb, k8);
        if (same8(k8, b)) {
          db = false;
        } else {
          store8(ck + 8 * wb, b);
        }
      }
      if (da) {
        float k8[8];
        load8(ack + 8 * c, k8);
        if (same8(k8, a)) {
          da = false;
        } else {
          store8(ack + 8 * c, a);
        }
      }
      if (!wave_any(db || da)) break;
      const uint32_t lb = min((uint32_t)kSubWin, S - wb * kSubWin), la = min((uint32_t)kSubWin, S - c * kSubWin);
      Chunk xb, xa;
      load_chunk(sc, L.le, wb * kSubWin, lb, xb);
      load_chunk(sc, L.le, c * kSubWin, la, xa);
      load_chunk_lds<H, true, false>(L, ex, qtab, L.q, by, wb * kSubWin, lb, xb, nullptr);
      load_chunk_lds<H, false, false>(L, ex, qtab, L.q, by, c * kSubWin, la, xa, nullptr);
      float nb[8], na[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        nb[i] = b[i];
        na[i] = a[i];
      }
      beta_chunk<D>(xb, lb, nb);
      alpha_chunk<D>(xa, la, na);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        b[i] = db ? nb[i] : b[i];
        a[i] = da ? na[i] : a[i];
      }
    }
    bool changed = false;
    if (db && !same8(b, bc)) {
      changed = true;
#pragma unroll
      for (int i = 0; i < 8; ++i) bc[i] = b[i];
    }
    if (da && !same8(a, ac)) {
      changed = true;
#pragma unroll
      for (int i = 0; i < 8; ++i) ac[i] = a[i];
    }
    if (!wave_any(changed)) break;
  }
  // the exact edge metrics: next iteration's starting values
  float t8[8];
  load8(ck + 8 * (nW - 1), t8);
  store8(nii + 8, t8);
  load8(ack, ac);
  store8(nii, ac);
  // 3. extrinsics
  extrinsic_pass<D, H, ES>(L, ex, sc, qtab, ck, ac, kNever, nii, post_sc, es);
}

// remainder test of CRC24B (generator of lte_turbo.cpp:108-111) on the
// hard decisions (posterior >= 0) of bits [F, K) of every block of the job,
// read from the posteriors the lower half left in the wave's scratch slot
// (bit i = q S + t at row t, column b M + q): bit i contributes D^(K-1-i) mod
// G.  Each lane folds its own segment (natural order: r(i+1) = r(i) * D^-1
// mod G), then the first lane of each block XORs its M partials.  Returns this
// lane's block verdict.  Reading the posteriors after the half keeps the
// early-stop work out of the recursion loops.
template <int R = kExRow>
__device__ __forceinline__ bool crc24b_job(const Lane& L, const float* __restrict__ post_sc,
                                           uint32_t* __restrict__ partial, const uint32_t* __restrict__ crc_pow,
                                           uint32_t filler) {
  const uint32_t S = L.S, K = L.K, lane = threadIdx.x;
  uint32_t x = 0;
  uint32_t r = crc_pow[K - 1 - L.q * S];
  // kCrcBatch posteriors in flight per lane (the registers of the recursions
  // are free here): two round trips to the cache for S = 96 instead of one
  // per 8 steps, whose latency the fold could not cover
  constexpr uint32_t kCrcBatch = 48;
  const float* base = post_sc + L.colb + L.q;
  for (uint32_t t0 = 0; t0 < S; t0 += kCrcBatch) {
    // every load unconditional (rows past the segment re-read row S-1 and
    // are masked below): a load under a per-step guard becomes a branch that
    // waits for it, one round trip per step
    float pv[kCrcBatch];
#pragma unroll
    for (uint32_t j = 0; j < kCrcBatch; ++j) pv[j] = base[min(t0 + j, S - 1) * R];
#pragma unroll
    for (uint32_t j = 0; j < kCrcBatch; ++j) {
      const bool in = t0 + j < S;  // uniform
      const uint32_t i = L.q * S + t0 + j;
      x ^= (in && i >= filler && pv[j] >= 0.0f) ? r : 0u;
      const uint32_t rn = (r >> 1) ^ ((0u - (r & 1u)) & ((0xC00063u | (1u << 24)) >> 1));  // r * D^-1 mod G
      r = in ? rn : r;
    }
  }
  partial[lane] = x;
  __syncthreads();
  uint32_t y = 0;
  for (uint32_t m = 0; m < L.M; ++m) y ^= partial[L.colb + m];
  return y == 0 && K - filler >= 24;
}

// One block's view of the received LLRs for staging-time de-rate matching
// (scalars only, so it lives in registers).
struct RxBlk {
  uint64_t in_off;
  uint32_t e, nv;
  const uint16_t* rank;  // rank of each [d0|d1|d2] position, kNullRank = NULL entry
  uint32_t n00, n01, n02, n03;  // first rank of each transmission
};

__device__ __forceinline__ uint32_t start_of(const RxBlk& B, int t) {
  return t == 0 ? B.n00 : (t == 1 ? B.n01 : (t == 2 ? B.n02 : B.n03));
}

__device__ __forceinline__ void rx_block(const RxIn& X, uint32_t cb, RxBlk& r) {
  const DerateCb* c = X.dcb + cb;
  r.in_off = c->in_off;
  r.e = c->e;
  r.nv = c->n_valid;
  r.rank = X.rank + c->rank_off;
  const uint32_t own = c->rv;
  r.n00 = c->start[(X.rv[0] == kBlockRv ? own : X.rv[0]) & 3u];
  r.n01 = c->start[(X.rv[1] == kBlockRv ? own : X.rv[1]) & 3u];
  r.n02 = c->start[(X.rv[2] == kBlockRv ? own : X.rv[2]) & 3u];
  r.n03 = c->start[(X.rv[3] == kBlockRv ? own : X.rv[3]) & 3u];
}

// kStageSteps steps of the three streams at once: all rank loads, then all
// value loads, so a lane keeps that many gathers in flight while staging.
constexpr uint32_t kStageSteps = 8;
// posteriors copied out per lane and batch
constexpr uint32_t kCopyBatch = 16;
// Soft de-rate values of N decoder-input positions (pos = s (K+4) + i) of
// one block, every one in range (callers clamp the slots they drop): the
// received values landing on each position's circular-buffer rank summed in
// reception order per transmission, transmissions added in order onto the
// held input or the first of them -- the operations and order of
// tc_derate_plan_run, so the result is bit-identical to it.  The loads of a
// batch are unconditional and issued together (a load under a per-lane
// condition compiles to a branch that waits for it): the ranks, the held
// input when it is added, then per transmission one received value per
// position at a clamped index, selected afterwards.  Repetition (E >
// N_valid, several values per rank) keeps a per-position loop.
template <uint32_t N>
__device__ __forceinline__ void rx_batch(const RxIn& X, const RxBlk& B, uint64_t llr_off, const uint32_t (&pos)[N],
                                         float (&out)[N]) {
  uint32_t r[N];
#pragma unroll
  for (uint32_t u = 0; u < N; ++u) r[u] = B.rank[pos[u]];
  if (X.base) {
    const float* keep = X.keep + llr_off;
#pragma unroll
    for (uint32_t u = 0; u < N; ++u) out[u] = keep[pos[u]];
  } else {
#pragma unroll
    for (uint32_t u = 0; u < N; ++u) out[u] = 0.0f;
  }
#pragma unroll
  for (int t = 0; t < kMaxTx; ++t) {
    if ((uint32_t)t >= X.n_tx) break;
    const uint32_t n0 = start_of(B, t);
    const float* in = X.rx[t] + B.in_off;
    const bool first = t == 0 && !X.base;
    if (B.e <= B.nv) {
      float a[N];
#pragma unroll
      for (uint32_t u = 0; u < N; ++u) {
        const uint32_t m = r[u] >= n0 ? r[u] - n0 : r[u] + B.nv - n0;
        a[u] = in[min(m, B.e - 1)];
      }
#pragma unroll
      for (uint32_t u = 0; u < N; ++u) {
        const uint32_t m = r[u] >= n0 ? r[u] - n0 : r[u] + B.nv - n0;
        const float x = 0.0f + a[u];  // the de-rate kernel's sum of one value (-0 -> +0)
        const float sum = first ? x : out[u] + x;
        // a select, not a branch: under a branch the load would be sunk into it
        out[u] = (r[u] != kNullRank && m < B.e) ? sum : out[u];
      }
    } else {
#pragma unroll
      for (uint32_t u = 0; u < N; ++u) {
        if (r[u] == kNullRank) continue;
        const uint32_t m = r[u] >= n0 ? r[u] - n0 : r[u] + B.nv - n0;
        if (m < B.e) {
          float x = 0.0f;
          for (uint32_t k = m; k < B.e; k += B.nv) x += in[k];
          out[u] = first ? x : out[u] + x;
        }
      }
    }
  }
}

// Transport-block assembly of one decoded block (lte_turbo.cpp:958-979), run
// by its wave right after the block's posteriors are written: hard decisions
// (LLR >= 0), filler strip, CRC24B over the block when it carries one, the
// block's share of its TB's CRC24A (XOR of D^(len-1-p) mod G over its ones at
// TB positions p), payload bits out.  Per-TB partial remainders, block counts
// and iteration maxima meet in acc[] through atomics; the last block of a TB
// finalises it.
__device__ __forceinline__ void tb_epilogue(const WinArgs& A, uint32_t cb, const float* post, uint32_t K,
                                            uint32_t filler) {
  const TbOut& T = A.tb;
  const uint32_t lane = threadIdx.x;
  const TbCbMap m = T.map[cb];
  const TbInfo t = T.tbs[m.tb];
  const uint32_t n = K - filler, len = t.b + 24;
  uint32_t cbx = 0, tbx = 0;
  // kCopyBatch bits per lane and batch: all loads first, then the XORs.  The
  // loads are unconditional, at indices clamped into range (n >= 1), and
  // masked afterwards: a load under a per-lane condition becomes a branch
  // that waits for it.
  for (uint32_t base = 0; base < n; base += kCopyBatch * kWaveLanes) {
    float pv[kCopyBatch];
    uint32_t pb[kCopyBatch], pa[kCopyBatch];
#pragma unroll
    for (uint32_t u = 0; u < kCopyBatch; ++u) {
      const uint32_t q = base + u * kWaveLanes + lane;
      const uint32_t qn = min(q, n - 1);
      pv[u] = post[filler + qn];
      pb[u] = T.pow24b[n - 1 - qn];
      pa[u] = T.pow24a[m.np ? len - 1 - (m.pos + min(q, m.np - 1)) : 0u];
    }
#pragma unroll
    for (uint32_t u = 0; u < kCopyBatch; ++u) {
      const uint32_t q = base + u * kWaveLanes + lane;
      const bool in = q < n, in_tb = q < m.np;
      pv[u] = in ? pv[u] : -1.0f;
      pb[u] = (in && m.seg) ? pb[u] : 0u;
      pa[u] = in_tb ? pa[u] : 0u;
    }
#pragma unroll
    for (uint32_t u = 0; u < kCopyBatch; ++u) {
      const uint32_t q = base + u * kWaveLanes + lane;
      const bool one = pv[u] >= 0.0f;
      if (one) {
        cbx ^= pb[u];
        tbx ^= pa[u];
      }
      if (q < m.np) {
        const uint32_t p = m.pos + q;
        if (T.bits && p < t.b) T.bits[t.bits_off + p] = one ? 1 : 0;
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    cbx ^= __shfl_xor(cbx, o, 64);
    tbx ^= __shfl_xor(tbx, o, 64);
  }
  if (lane == 0) {
    const bool cb_pass = m.seg ? (cbx == 0 && n >= 24) : true;
    if (T.cb_ok) T.cb_ok[cb] = cb_pass ? 1 : 0;
    uint32_t* acc = T.acc + 3 * m.tb;
    if (t.cb_crc) tbx = cb_pass ? 0u : 1u;  // a code-block-CRC descriptor is checked by its CRC24B
    // order-free device-scope atomics; tb_finalize_kernel reads them after this launch.  No
    // fence: on this GPU a device-scope fence writes back and invalidates the XCD's L2
    // (buffer_wbl2 / buffer_inv), which at two per block cost 8 ms on 65536 blocks
    atomicXor(acc, tbx);
    atomicMax(acc + 2, A.iters ? A.iters[cb] : 0u);
  }
  (void)len;
}

// Jobs are taken from a counter rather than by a fixed stride: waves that
// share a SIMD run slower than waves alone on one, and early-stopped blocks
// finish early, so a fixed assignment leaves the fast waves idle at the end.
template <int NW = 1>
__device__ __forceinline__ uint32_t next_job(uint32_t* sched) {
  if constexpr (NW == 1) {
    uint32_t j = 0;
    if (threadIdx.x == 0) j = atomicAdd(sched, 1u);
    return __builtin_amdgcn_readfirstlane(j);
  } else {
    // one ticket per work-group, passed to its waves through LDS
    __shared__ uint32_t ticket;
    __syncthreads();  // every wave has read the previous ticket
    if (threadIdx.x == 0) ticket = atomicAdd(sched, 1u);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(ticket);
  }
}
// Every wave has taken its last (empty) ticket before it counts itself out,
// so the last one out can reset the queue for the next launch.
__device__ __forceinline__ void leave_queue(uint32_t* sched) {
  if (threadIdx.x == 0 && atomicAdd(sched + 1, 1u) == gridDim.x - 1) {
    atomicExch(sched, 0u);
    atomicExch(sched + 1, 0u);
  }
}

// Transposes between a block's natural order and the wave's [t][lane]
// scratch rows go through the job's LDS (the extrinsic area, idle outside
// the iterations, (S + 1) * 32 floats with the dummy slots): R = ceil(S/2)
// rows of 64 floats per pass, column c of row u at u*64 + (c ^ (u mod 64)),
// so that a run of consecutive rows of one column (the natural order) and a
// row across the columns are both conflict-free.  Global accesses are then
// runs of consecutive floats on the natural side and whole rows on the
// scratch side.
template <int R = kWaveLanes>
__device__ __forceinline__ uint32_t tile_at(uint32_t u, uint32_t c) {
  return u * R + (c ^ (u & 63u));
}
constexpr uint32_t kTileBatch = 16;

// One-block job (M segments): the decoder input of the block into the scratch
// planes: d0 and d1 into the upper half's, d2 and d0[pi] (gathered from the
// tile while it holds d0 rows) into the lower half's; fused de-rate matching
// when A.rx.n_tx > 0 (then also the held input, written in natural order).
template <class D>
__device__ __forceinline__ void stage_tiled(const WinArgs& A, const Lane& L, const CbInfo& ci, const RxBlk& B,
                                            bool fused, const uint32_t* __restrict__ qtab, float* __restrict__ tile,
                                            float* __restrict__ up, float* __restrict__ lo) {
  constexpr int R = kWaveLanes * D::kWaves;  // the job's lanes: columns of a row
  const uint32_t lane = threadIdx.x, S = L.S, M = L.M, D4 = L.K + 4;
  const uint32_t rows = (S + 1) / 2;
  float* keep = (fused && A.rx.keep) ? A.rx.keep + ci.llr_off : nullptr;
  const uint32_t by = qpp_stride(L, L.q);
  for (uint32_t sgm = 0; sgm < 3; ++sgm) {
    for (uint32_t r0 = 0; r0 < S; r0 += rows) {
      const uint32_t rp = min(rows, S - r0);
      const uint32_t mg = 0xFFFFFFFFu / rp + 1u;  // ceil(2^32 / rp): exact idx / rp for idx < 2^16
      const uint32_t n = M * rp;
      for (uint32_t base = 0; base < n; base += kTileBatch * R) {
        uint32_t pos[kTileBatch], tix[kTileBatch], mask = 0;
#pragma unroll
        for (uint32_t k = 0; k < kTileBatch; ++k) {
          const uint32_t idx = base + k * R + lane;
          const uint32_t q = __umulhi(idx, mg), u = idx - q * rp;
          pos[k] = sgm * D4 + q * S + r0 + u;
          tix[k] = tile_at<R>(u, q);
          mask |= (idx < n ? 1u : 0u) << k;
        }
        float v[kTileBatch];
        if (fused) {
          uint32_t posl[kTileBatch];  // in range: a slot past the pass takes the pass's last entry
#pragma unroll
          for (uint32_t k = 0; k < kTileBatch; ++k) {
            const uint32_t idx = min(base + k * R + lane, n - 1);
            const uint32_t q = __umulhi(idx, mg), u = idx - q * rp;
            posl[k] = sgm * D4 + q * S + r0 + u;
          }
          rx_batch(A.rx, B, ci.llr_off, posl, v);
        } else {
          // unconditional loads (a slot past the pass re-reads its last
          // entry; the mask drops it below): a load under a per-lane
          // condition becomes a branch that waits for it
          const uint32_t nlast = n - 1;
#pragma unroll
          for (uint32_t k = 0; k < kTileBatch; ++k) {
            const uint32_t idx = min(base + k * R + lane, nlast);
            const uint32_t q = __umulhi(idx, mg), u = idx - q * rp;
            v[k] = A.llr[ci.llr_off + sgm * D4 + q * S + r0 + u];
          }
        }
#pragma unroll
        for (uint32_t k = 0; k < kTileBatch; ++k) {
          if ((mask >> k) & 1u) {
            tile[tix[k]] = v[k];
            if (keep) keep[pos[k]] = v[k];
          }
        }
      }
      __syncthreads();
      // rows r0 .. r0+rp-1 of this lane's column
      float* dst = (sgm == 0) ? up : (sgm == 1 ? up + plane_of<R>() : lo + plane_of<R>());
      for (uint32_t u0 = 0; u0 < rp; u0 += kSubWin) {
        float v[kSubWin];
#pragma unroll
        for (uint32_t k = 0; k < kSubWin; ++k) v[k] = tile[tile_at<R>(min(u0 + k, rp - 1), lane)];
#pragma unroll
        for (uint32_t k = 0; k < kSubWin; ++k) {
          if (u0 + k < rp && L.active)
            dst[(r0 + u0 + k) * R + lane] = (sgm == 0) ? v[k] : D::stage_par(v[k]);
        }
      }
      if (sgm == 0) {
        // lower-half systematic: step t of segment q reads d0 at pi(q S + t),
        // i.e. row pi(t) mod S (uniform) at this lane's permuted column.
        // kSubWin steps at a time: their table entries in one scalar batch
        // (the table is padded by kSubWin entries), then every tile read at
        // a row clamped into the pass, then the stores of the steps whose
        // row the pass holds -- not one dependent round trip per step
        uint32_t y = qpp_col0(L, L.q, 0);
        for (uint32_t t0 = 0; t0 < S; t0 += kSubWin) {
          uint32_t off = __builtin_amdgcn_readfirstlane(L.qtab_off + t0);  // wave-uniform
          asm volatile("" : "+s"(off));
          const uint32_t* qt = qtab + off;
          uint32_t qv[kSubWin];
#pragma unroll
          for (uint32_t j = 0; j < kSubWin; ++j) qv[j] = qtab_at(qt, j);
          float x[kSubWin];
#pragma unroll
          for (uint32_t j = 0; j < kSubWin; ++j) {
            uint32_t c = (qv[j] & 0xFFFFu) + y;
            c = (c >= M) ? c - M : c;
            x[j] = tile[tile_at<R>(min((qv[j] >> 16) - r0, rp - 1), c)];
            y = y_fwd(y, by, M);
          }
#pragma unroll
          for (uint32_t j = 0; j < kSubWin; ++j) {
            const uint32_t row = qv[j] >> 16;
            if (t0 + j < S && row >= r0 && row < r0 + rp && L.active) lo[(t0 + j) * R + lane] = x[j];
          }
        }
      }
      __syncthreads();
    }
  }
}

// One-block job: posteriors from the scratch rows (post_sc, R per row) to
// dst in natural order, zero when no iteration ran.
template <int R = kWaveLanes>
__device__ __forceinline__ void copy_out_tiled(const Lane& L, bool any, const float* __restrict__ post_sc,
                                               float* __restrict__ tile, float* __restrict__ dst) {
  const uint32_t lane = threadIdx.x, S = L.S, M = L.M;
  const uint32_t rows = (S + 1) / 2;
  for (uint32_t r0 = 0; r0 < S; r0 += rows) {
    const uint32_t rp = min(rows, S - r0);
    for (uint32_t u0 = 0; u0 < rp; u0 += kTileBatch) {
      float v[kTileBatch];
#pragma unroll
      for (uint32_t k = 0; k < kTileBatch; ++k) v[k] = post_sc[(r0 + min(u0 + k, rp - 1)) * R + lane];
#pragma unroll
      for (uint32_t k = 0; k < kTileBatch; ++k)
        if (u0 + k < rp) tile[tile_at<R>(u0 + k, lane)] = v[k];
    }
    __syncthreads();
    const uint32_t mg = 0xFFFFFFFFu / rp + 1u;
    const uint32_t n = M * rp;
    for (uint32_t base = 0; base < n; base += kTileBatch * R) {
      float v[kTileBatch];  // read at clamped slots, so no read waits under a branch
#pragma unroll
      for (uint32_t k = 0; k < kTileBatch; ++k) {
        const uint32_t idx = min(base + k * R + lane, n - 1);
        const uint32_t q = __umulhi(idx, mg), u = idx - q * rp;
        v[k] = tile[tile_at<R>(u, q)];
      }
#pragma unroll
      for (uint32_t k = 0; k < kTileBatch; ++k) {
        const uint32_t idx = base + k * R + lane;
        const uint32_t q = __umulhi(idx, mg), u = idx - q * rp;
        if (idx < n) dst[q * S + r0 + u] = any ? v[k] : 0.0f;
      }
    }
    __syncthreads();
  }
}

// One job: stage its inputs, run every iteration, write its outputs.
// LEAN: every job of the launch has the lean geometry (siso_half_lean); the
// host picks such a kernel only for plans whose jobs all have it, so that
// neither path's registers constrain the other's.
// A job on several waves (D::kWaves > 1) keeps two sets of boundary records,
// [lane][kRecFloats] each, from recs: iteration it reads set it mod 2 and hands
// its metrics to the other (nii: this lane's record in set 0).
template <class D, bool ES, class X, bool LEAN = false>
__device__ __forceinline__ void window_job(const WinArgs& A, const WinJob& J, X* __restrict__ ex,
                                           float* __restrict__ sc_up, float* __restrict__ sc_lo,
                                           float* __restrict__ nii, float* __restrict__ ck,
                                           float* __restrict__ post_sc, float* __restrict__ ack,
                                           float* __restrict__ recs = nullptr, X* __restrict__ ex_slot = nullptr) {
  constexpr int R = kWaveLanes * D::kWaves;  // the job's lanes
  constexpr uint32_t kRecSet = R * kRecFloats;
  const uint32_t lane = threadIdx.x;
  const uint32_t* __restrict__ cb_index = A.cb_index;
  const CbInfo* __restrict__ cbs = A.cbs;
  const uint32_t* __restrict__ qtab = A.qtab;
  const bool es_on = ES && A.max_iterations > 0;
  Lane L;
  L.S = J.seg;
  L.M = J.nseg;
  L.g = J.warm;
  L.K = J.k;
  L.nW = (J.seg + kSubWin - 1) / kSubWin;
  L.f1 = J.f1;
  L.f2 = J.f2;
  L.qtab_off = J.qtab_off;
  const uint32_t used = J.nblk * J.nseg;
  L.active = lane < used;
  L.le = L.active ? lane : used - 1;  // idle lanes shadow the last busy lane
  const uint32_t b = L.le / L.M;
  L.q = L.le - b * L.M;
  L.colb = b * L.M;
  L.first = L.q == 0;
  L.last = L.q == L.M - 1;
  const uint32_t K = J.k, S = J.seg, M = J.nseg;
  const uint32_t my_cb = cb_index[J.cb_off + b];
  const CbInfo ci = cbs[my_cb];
  const bool fused = A.rx.n_tx > 0;
  // LDS: extrinsics [S][R] + one dummy slot per lane, then the CRC
  // partials (early stop only)
  const uint32_t ex_slots = S * R + R;
  // fp32-held extrinsics of a one-wave job (the exact boundaries' and the
  // log decoder's fp32 variant, D::kF32HeldInSlot): in the wave's scratch slot (ex_slot),
  // LDS only for the staging tile, in which the CRC partials sit during the
  // iterations -- so eight waves fit a CU instead of six
  constexpr bool kHeldInSlot = D::kWaves == 1 && sizeof(X) == 4 && D::kF32HeldInSlot;
  X* const exh = kHeldInSlot ? ex_slot : ex;
  uint32_t* partial = kHeldInSlot ? reinterpret_cast<uint32_t*>(ex) : reinterpret_cast<uint32_t*>(ex + ex_slots);
  RxBlk Bm{};
  if (fused) rx_block(A.rx, my_cb, Bm);
  const uint32_t D4 = K + 4;

  // ---- the decoder input into the wave's scratch planes in [t][lane]
  // order: d0 and d1 in the upper half's, d2 in the lower half's parity
  // plane.  One-block jobs transpose through LDS;
  // otherwise each lane stages its own segment.  Idle lanes (shadowing the
  // last busy lane) neither read nor write the held input.
  if (J.nblk == 1) {
    stage_tiled<D>(A, L, ci, Bm, fused, qtab, reinterpret_cast<float*>(ex), sc_up, sc_lo);
  } else {
    float* keep = (fused && A.rx.keep) ? A.rx.keep + ci.llr_off : nullptr;
    for (uint32_t t0 = 0; t0 < S; t0 += kStageSteps) {
      constexpr uint32_t N = 3 * kStageSteps;
      uint32_t pos[N], mask = 0;
#pragma unroll
      for (uint32_t u = 0; u < kStageSteps; ++u) {
        const uint32_t i = L.q * S + t0 + u;
        pos[3 * u] = i;
        pos[3 * u + 1] = D4 + i;
        pos[3 * u + 2] = 2 * D4 + i;
        mask |= ((L.active && t0 + u < S) ? 7u : 0u) << (3 * u);
      }
      // loads at in-range positions (steps past the segment re-read its last
      // step; the mask drops them below), unconditional
      uint32_t posl[N];
#pragma unroll
      for (uint32_t u = 0; u < kStageSteps; ++u) {
        const uint32_t i = L.q * S + min(t0 + u, S - 1);
        posl[3 * u] = i;
        posl[3 * u + 1] = D4 + i;
        posl[3 * u + 2] = 2 * D4 + i;
      }
      float v[N];
      if (fused) {
        rx_batch(A.rx, Bm, ci.llr_off, posl, v);
      } else {
#pragma unroll
        for (uint32_t u = 0; u < N; ++u) v[u] = A.llr[ci.llr_off + posl[u]];
      }
#pragma unroll
      for (uint32_t u = 0; u < kStageSteps; ++u) {
        const uint32_t t = t0 + u;
        if (((mask >> (3 * u)) & 1u) == 0) continue;
        sc_up[t * R + lane] = v[3 * u];
        sc_up[plane_of<R>() + t * R + lane] = D::stage_par(v[3 * u + 1]);
        sc_lo[plane_of<R>() + t * R + lane] = D::stage_par(v[3 * u + 2]);
        if (keep) {
          keep[pos[3 * u]] = v[3 * u];
          keep[pos[3 * u + 1]] = v[3 * u + 1];
          keep[pos[3 * u + 2]] = v[3 * u + 2];
        }
      }
    }
  }
  // ---- boundary metrics: alpha_0 / tail beta / uniform ----
  float tl[3][4];  // d0, d1, d2 at K .. K+3
  {
    uint32_t tp[12];
    float tv[12];
#pragma unroll
    for (uint32_t u = 0; u < 12; ++u) tp[u] = (u / 4) * D4 + K + u % 4;
    if (fused) {
      rx_batch(A.rx, Bm, ci.llr_off, tp, tv);
    } else {
#pragma unroll
      for (uint32_t u = 0; u < 12; ++u) tv[u] = A.llr[ci.llr_off + tp[u]];
    }
#pragma unroll
    for (uint32_t u = 0; u < 12; ++u) tl[u / 4][u % 4] = tv[u];
  }
  {
    float tb[2][8];
    // decoder tail read-out of lte_turbo.cpp:758-773
    tail_beta(tl[0][0], tl[0][2], 0.0f, tl[1][0], tl[0][1], tl[1][2], tb[0]);
    tail_beta(tl[1][3], tl[2][0], tl[2][2], tl[2][1], tl[0][3], tl[2][3], tb[1]);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        nii[h * 16 + s] = L.first ? (s == 0 ? D::kUniform : D::kZero) : D::kUniform;
        nii[h * 16 + 8 + s] = L.last ? D::from_log(tb[h][s]) : D::kUniform;
        if constexpr (D::kWaves > 1) {  // both record sets
          nii[kRecSet + h * 16 + s] = nii[h * 16 + s];
          nii[kRecSet + h * 16 + 8 + s] = nii[h * 16 + 8 + s];
        }
      }
    }
  }
  __syncthreads();  // every lane's d0 in sc_up before the gather below
  if (fused && A.rx.keep && L.active && L.first) {
    // every lane of the block has read the tails; one writes them
    for (uint32_t sgm = 0; sgm < 3; ++sgm)
      for (uint32_t j = 0; j < 4; ++j) A.rx.keep[ci.llr_off + sgm * D4 + K + j] = tl[sgm][j];
  }
  // ---- several-block jobs: interleaved systematic, the lower half's plane
  // at step t = d0[pi(q S + t)], which sits in the upper half's plane at the
  // extrinsic address of that bit ----
  if (J.nblk > 1) {
    uint32_t y = qpp_col0(L, L.q, 0);
    const uint32_t by = qpp_stride(L, L.q);
    for (uint32_t t0 = 0; t0 < S; t0 += kSubWin) {
      float xp[kSubWin];
#pragma unroll
      for (uint32_t u = 0; u < kSubWin; ++u) {
        if (t0 + u < S) {
          xp[u] = sc_up[lds_addr<1, R>(L, qtab, L.q, t0 + u, y)];
          y = y_fwd(y, by, M);
        }
      }
#pragma unroll
      for (uint32_t u = 0; u < kSubWin; ++u)
        if (t0 + u < S && L.active) sc_lo[(t0 + u) * R + lane] = xp[u];
    }
  }
  {
    uint32_t* exw = reinterpret_cast<uint32_t*>(exh);
    for (uint32_t i = lane; i < ex_slots * sizeof(X) / 4; i += R) exw[i] = 0u;
  }
  __syncthreads();

  EarlyStop es{es_on, false};
  uint32_t used_it = A.max_iterations;
  for (uint32_t it = 0; it < A.max_iterations; ++it) {
    // several waves: this iteration's records and the next one's
    float* const own = (D::kWaves > 1 && (it & 1u)) ? nii + kRecSet : nii;
    float* const push = (D::kWaves > 1) ? recs + ((it & 1u) ? 0u : kRecSet) : nullptr;
    if constexpr (D::kExact) {
      siso_half_exact<0, ES>(L, exh, sc_up, qtab, nii, ck, ack, post_sc, es);
    } else if constexpr (LEAN) {
      siso_half_lean<D, 0, ES>(L, exh, sc_up, qtab, nii, ck, post_sc, es);
    } else {
      siso_half<D, 0, ES>(L, exh, sc_up, qtab, own, ck, post_sc, es, push);
    }
    __syncthreads();
    // every lower half leaves the posteriors in post_sc; a block that has
    // stopped no longer overwrites its own
    if constexpr (D::kExact) {
      siso_half_exact<1, ES>(L, exh, sc_lo, qtab, nii + 16, ck, ack, post_sc, es);
    } else if constexpr (LEAN) {
      siso_half_lean<D, 1, ES>(L, exh, sc_lo, qtab, nii + 16, ck, post_sc, es);
    } else {
      siso_half<D, 1, ES>(L, exh, sc_lo, qtab, own + 16, ck, post_sc, es, D::kWaves > 1 ? push + 16 : nullptr);
    }
    __syncthreads();
    if (es_on) {
      const bool ok = crc24b_job<R>(L, post_sc, partial, A.crc_pow, ci.filler);
      if (ok && !es.done) {
        es.done = true;
        used_it = it + 1;
        if (L.active && L.first) {
          if (A.crc_ok) A.crc_ok[my_cb] = 1;
          if (A.iters) A.iters[my_cb] = it + 1;
        }
      }
      if (__syncthreads_and(es.done || !L.active)) break;
    }
  }

  // ---- posterior out, natural order (zero when no iteration ran), with
  // kCopyBatch loads in flight per lane ----
  if (J.nblk == 1)
    copy_out_tiled<R>(L, A.max_iterations > 0, post_sc, reinterpret_cast<float*>(ex), A.post + ci.post_off);
  else for (uint32_t bb = 0; bb < J.nblk; ++bb) {
    float* dst = A.post + cbs[cb_index[J.cb_off + bb]].post_off;
    for (uint32_t base = 0; base < K; base += kCopyBatch * R) {
      float v[kCopyBatch];
#pragma unroll
      for (uint32_t u = 0; u < kCopyBatch; ++u) {
        const uint32_t i = base + u * R + lane;
        const uint32_t qq = __umulhi(i, J.div_magic);
        const uint32_t t = i - qq * S;
        v[u] = (i < K && A.max_iterations) ? post_sc[t * R + bb * M + qq] : 0.0f;
      }
#pragma unroll
      for (uint32_t u = 0; u < kCopyBatch; ++u) {
        const uint32_t i = base + u * R + lane;
        if (i < K) dst[i] = v[u];
      }
    }
  }
  if (L.active && L.first && !es.done) {
    if (A.crc_ok) A.crc_ok[my_cb] = 0;
    if (A.iters) A.iters[my_cb] = used_it;
  }
  if (A.tb.on) {
    __threadfence_block();  // this wave's posterior and iteration stores before it reads them back
    __syncthreads();
    // one wave assembles (the job's first, when it has several)
    if (D::kWaves == 1 || lane < (uint32_t)kWaveLanes) {
      for (uint32_t bb = 0; bb < J.nblk; ++bb) {
        const uint32_t cbb = cb_index[J.cb_off + bb];
        const CbInfo cbi = cbs[cbb];
        tb_epilogue(A, cbb, A.post + cbi.post_off, K, cbi.filler);
      }
    }
  }
  __syncthreads();
}

// A work-group of D::kWaves waves runs one job at a time; its scratch slot is
// kSlotFloat2 per wave.
template <class D, bool ES, class X, bool LEAN = false>
__device__ __forceinline__ void window_decode(const WinArgs& A) {
  extern __shared__ float lds[];
  constexpr int NW = D::kWaves, R = kWaveLanes * NW;
  const uint32_t lane = threadIdx.x;
  // staged inputs: two planes per half
  float* sc_up = reinterpret_cast<float*>(A.scratch + (size_t)blockIdx.x * NW * kSlotFloat2);
  float* sc_lo = sc_up + 2 * plane_of<R>();
  // per-lane records: boundary metrics [lane][half][16] (two sets on several
  // waves), checkpoints [lane][w][8]
  constexpr uint32_t kRecs = (NW > 1 ? 2 : 1) * R * kRecFloats;
  float* recs = sc_lo + 2 * plane_of<R>();
  float* nii = recs + lane * kRecFloats;
  float* ck = recs + kRecs + lane * (8 * kMaxSubWins);
  // posteriors [S][R] + one dummy slot per lane
  float* post_sc = recs + kRecs + R * 8 * kMaxSubWins;
  // alpha checkpoints [lane][w][8] and the held extrinsics [S][64] + one
  // dummy slot per lane (exact segment boundaries; one wave only)
  float* ack = post_sc + (kMaxSeg + 1) * R + lane * (8 * kMaxSubWins);
  X* ex_slot = reinterpret_cast<X*>(post_sc + (kMaxSeg + 1) * R + R * 8 * kMaxSubWins);
  static_assert(4 * plane_of<R>() + kRecs + R * 8 * kMaxSubWins + (kMaxSeg + 1) * R +
                        (NW == 1 ? R * 8 * kMaxSubWins + (kMaxSeg + 1) * R : 0) <= 2 * NW * kSlotFloat2,
                "the job's scratch fits its waves' slots");
  for (uint32_t jb = next_job<NW>(A.sched); jb < A.njobs; jb = next_job<NW>(A.sched)) {
    const WinJob J = A.jobs[jb];
    window_job<D, ES, X, LEAN>(A, J, reinterpret_cast<X*>(lds), sc_up, sc_lo, nii, ck, post_sc, ack, recs, ex_slot);
  }
  leave_queue(A.sched);
}

}  // namespace
}  // namespace tcdev

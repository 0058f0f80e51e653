// Exact Log-MAP turbo decoder, serial form: full-length forward and backward
// recursions, one per code block, evaluated in the reference's operation
// order (lte_turbo.cpp:380-486, 736-813) so that posterior LLRs agree with
// turbo_decode_code_block_cpu to the rounding of max* (jac_ref against
// expf/log1pf).
//
// TC_DECODER_EXACT_SERIAL and the tile8 entry point run it; the
// segment-parallel exact decoder (window_kernel.h, siso_half_exact) computes
// the same metrics bit for bit and backs GpuStage1Exact.
//
// Layout: one wave decodes up to kExactBlocksPerWave code blocks of one K,
// kExactLanes lanes per block; the first eight lanes of a block carry one
// state each through the recursions, and all of the block's lanes share its
// extrinsic pass.  Each recursion lane carries its own state's metric
// through the recursions: a step reads the two predecessor (successor)
// metrics of the block's other lanes with ds_bpermute and evaluates exactly
// the reference's expression for its state (the two transitions in order of
// increasing source state, the reference's max* with its infinity rules), and
// the renormalisation takes the block's maximum across its eight lanes with
// DPP.  The alpha and beta recursions of a half-iteration are independent and
// run in the same loop (alpha at step i, beta at step K-1-i), storing every
// step's metrics; the extrinsics then follow in a pass parallel over steps,
// each lane doing the reference's sequential LLR accumulation for its steps.
// Two blocks per wave (rather than eight) give two waves per SIMD over a
// 4096-block batch and four times the lanes per block in the extrinsic pass.
//
// Work space per code block (float offsets from work_off):
//   alpha (K+1)*8   forward metrics, all steps
//   beta  (K+1)*8   backward metrics, all steps
//   e1    K         upper extrinsic, natural order
//   e2    K         lower extrinsic, natural order (= next a-priori)
//   si    K         systematic LLRs in interleaved order
//   ia    K         upper extrinsic in interleaved order
#include <hip/hip_runtime.h>

#include "decoder_types.h"
#include "trellis.h"

#pragma clang fp contract(off)

namespace tcdev {
namespace {

constexpr float kInf = __builtin_inff();

// Two waves share each SIMD; the issue arbitration favours the older one, so
// it finishes its blocks first and leaves the other to run alone, which for
// this latency-bound recursion halves the SIMD's throughput.  The two
// alternate priority on a global slice of 16384 ticks of the 100 MHz
// reference clock, by the parity of their wave slots (0 and 1 of the SIMD).
__device__ __forceinline__ void share_simd() {
  const uint32_t slot = __builtin_amdgcn_s_getreg((3 << 11) | 4) & 1u;  // HW_ID, wave slot bits [3:0]
  const uint32_t t = (uint32_t)(__builtin_amdgcn_s_memrealtime() >> 14);
  if ((t ^ slot) & 1u) {
    __builtin_amdgcn_s_setprio(1);
  } else {
    __builtin_amdgcn_s_setprio(0);
  }
}

__device__ __forceinline__ bool fin(float v) { return __builtin_isfinite(v); }

// max* with the reference's infinity rules (lte_turbo.cpp:380-390): y when
// x is not finite, else x when y is not finite, else jac_ref (trellis.h).
// Written without branches (the Jacobian term is evaluated in every case and
// selected away when an input is not finite) so that the alpha and beta steps
// of a loop iteration, two independent chains, can interleave.
__device__ __forceinline__ float jac(float x, float y) {
  const float r = jac_ref(x, y);
  return !fin(x) ? y : (!fin(y) ? x : r);
}

__device__ __forceinline__ float bm(unsigned u, unsigned p, float sa, float lp) {
  return 0.5f * ((u ? 1.0f : -1.0f) * sa + (p ? 1.0f : -1.0f) * lp);
}

// metric of lane src (a lane of the same wave)
__device__ __forceinline__ float from_lane(float v, uint32_t src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(static_cast<int>(src << 2), __builtin_bit_cast(int, v)));
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
}
// maximum over the eight lanes of a block (xor 1, xor 2 within quads, then
// the mirror within eight lanes)
__device__ __forceinline__ float group_max(float v) {
  v = fmaxf(v, dpp<0xB1>(v));   // quad_perm [1,0,3,2]
  v = fmaxf(v, dpp<0x4E>(v));   // quad_perm [2,3,0,1]
  v = fmaxf(v, dpp<0x141>(v));  // row_half_mirror
  return v;
}
// lte_turbo.cpp:392-402: subtract the maximum from the finite metrics
__device__ __forceinline__ float renorm(float m) {
  const float mx = group_max(m);
  return (fin(mx) && fin(m)) ? m - mx : m;
}

constexpr uint32_t kExactLanes = 64 / kExactBlocksPerWave;

// QPP Pi(t) for t = t0, t0 + L, t0 + 2L, ... (L = kExactLanes): Pi(t + L) -
// Pi(t) = L f1 + f2 (2 L t + L^2), which grows by 2 L^2 f2 per stride (all
// mod K; K <= 6144 keeps every product within 32 bits).
struct QppStride {
  uint32_t pi, d, dd, k;
  __device__ QppStride(uint32_t t0, uint32_t kk, uint32_t f1, uint32_t f2) : k(kk) {
    constexpr uint32_t L = kExactLanes;
    pi = (f1 % kk * t0 % kk + f2 % kk * t0 % kk * t0) % kk;
    d = (L * (f1 % kk) % kk + f2 % kk * ((2u * L * t0 + L * L) % kk)) % kk;
    dd = 2u * L * L % kk * (f2 % kk) % kk;
  }
  __device__ void next() {
    pi += d;
    pi = pi >= k ? pi - k : pi;
    d += dd;
    d = d >= k ? d - k : d;
  }
};

// Tail beta of state st (lte_turbo.cpp:357-378) from three systematic/parity
// tail LLRs, renormalised across the block.
__device__ float tail_beta(unsigned st, const float* ts, const float* tp) {
  unsigned cur = st;
  float metric = 0.0f;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const unsigned u = term_in(cur);
    const Tr tr = rsc(cur, u);
    metric += 0.5f * ((u ? 1.0f : -1.0f) * ts[t]) + 0.5f * ((tr.parity ? 1.0f : -1.0f) * tp[t]);
    cur = tr.next;
  }
  return renorm((cur == 0) ? metric : -kInf);
}

// One SISO half for this lane's block.  sys[t], ap[t] (t < K), par[t]; the
// extrinsic of step t is written to ext[scatter ? Pi(t) : t].
// r: the lane's index within its block; lanes r < 8 (state st = r) run the
// recursions.
__device__ void siso(unsigned r, uint32_t gbase, const float* sys, const float* ap, const float* par, uint32_t k,
                     const float* ts, const float* tp, float* alpha, float* beta, float* ext, bool scatter,
                     uint32_t f1, uint32_t f2, bool active) {
  const unsigned st = r & 7u;
  const bool rec = active && r < 8;  // stores of the recursions
  // this lane's transitions: backward from st to its successors n0 (u=0)
  // and n1 (u=1); forward into st from predecessors p, p+1 with input u
  const Tr t0 = rsc(st, 0), t1 = rsc(st, 1);
  const unsigned u = in_of(st), p = pred0(st), c = par0(st);
  float b = tail_beta(st, ts, tp);
  float a = (st == 0) ? 0.0f : -kInf;
  if (rec) {
    beta[8 * (size_t)k + st] = b;
    alpha[st] = a;
  }
  // inputs by chunks of kExactLanes steps, loaded one chunk ahead, one step
  // per lane of the block (alpha side: step i0 + r; beta side: step
  // k - 1 - i0 - r), and handed to the recursion lanes with ds_bpermute
  float ca_s, ca_p, cb_s, cb_p, na_s = 0.0f, na_p = 0.0f, nb_s = 0.0f, nb_p = 0.0f;
  auto load_chunk = [&](uint32_t i0, float& as, float& apv, float& bs, float& bp) {
    const uint32_t i = i0 + r;
    if (i < k) {
      const uint32_t tb = k - 1 - i;
      as = sys[i] + ap[i];
      apv = par[i];
      bs = sys[tb] + ap[tb];
      bp = par[tb];
    } else {
      as = apv = bs = bp = 0.0f;
    }
  };
  load_chunk(0, ca_s, ca_p, cb_s, cb_p);
  for (uint32_t i0 = 0; i0 < k; i0 += kExactLanes) {
    share_simd();
    if (i0 + kExactLanes < k) load_chunk(i0 + kExactLanes, na_s, na_p, nb_s, nb_p);
    const uint32_t n = min(kExactLanes, k - i0);
    for (uint32_t j = 0; j < n; ++j) {
      const uint32_t i = i0 + j, tb = k - 1 - i, ta = i;
      const float saa = from_lane(ca_s, gbase + j), lpa = from_lane(ca_p, gbase + j);
      const float sab = from_lane(cb_s, gbase + j), lpb = from_lane(cb_p, gbase + j);
      // beta_tb(st) from beta_{tb+1}; the reference skips non-finite successors
      const float b0 = from_lane(b, gbase + t0.next), b1 = from_lane(b, gbase + t1.next);
      const float cb0 = b0 + bm(0, t0.parity, sab, lpb);
      const float cb1 = b1 + bm(1, t1.parity, sab, lpb);
      // alpha_{ta+1}(st) from alpha_ta
      const float a0 = from_lane(a, gbase + p), a1 = from_lane(a, gbase + p + 1);
      const float ca0 = a0 + bm(u, c, saa, lpa);
      const float ca1 = a1 + bm(u, c ^ 1u, saa, lpa);
      b = renorm(jac(fin(b0) ? cb0 : -kInf, fin(b1) ? cb1 : -kInf));
      a = renorm(jac(fin(a0) ? ca0 : -kInf, fin(a1) ? ca1 : -kInf));
      if (rec) {
        beta[8 * (size_t)tb + st] = b;
        alpha[8 * (size_t)(ta + 1) + st] = a;
      }
    }
    ca_s = na_s, ca_p = na_p, cb_s = nb_s, cb_p = nb_p;
  }
  __syncthreads();  // every lane's metrics stored before the extrinsic pass reads them
  // extrinsics: this lane takes steps r, r + L, ...; the reference's
  // accumulation order over (state, input) within each step
  QppStride q(r, k, f1, f2);
  for (uint32_t t = r; t < k; t += kExactLanes, q.next()) {
    const float4* ap4 = reinterpret_cast<const float4*>(alpha + 8 * (size_t)t);
    const float4* bp4 = reinterpret_cast<const float4*>(beta + 8 * (size_t)(t + 1));
    const float4 al = ap4[0], ah = ap4[1], bl = bp4[0], bh = bp4[1];
    const float av[8] = {al.x, al.y, al.z, al.w, ah.x, ah.y, ah.z, ah.w};
    const float bn[8] = {bl.x, bl.y, bl.z, bl.w, bh.x, bh.y, bh.z, bh.w};
    const float sa = sys[t] + ap[t];
    const float lp = par[t];
    float l[2] = {-kInf, -kInf};
#pragma unroll
    for (unsigned s = 0; s < 8; ++s) {
#pragma unroll
      for (unsigned uu = 0; uu < 2; ++uu) {
        const Tr tr = rsc(s, uu);
        const float m = av[s] + bm(uu, tr.parity, sa, lp) + bn[tr.next];
        l[uu] = fin(l[uu]) ? jac(l[uu], m) : m;
      }
    }
    if (active) ext[scatter ? q.pi : t] = (l[1] - l[0]) - sys[t] - ap[t];
  }
  __syncthreads();
}

// CRC24B early-stop test on hard decisions of the posterior, filler bits
// forced to 0 and excluded (lte_turbo.cpp:792-797, generator of :108-111).
__device__ bool crc24b_ok(const float* post, uint32_t k, uint32_t filler) {
  if (k - filler < 24) return false;
  uint32_t r = 0;
  const uint32_t msg_end = k - 24;
  for (uint32_t i = filler; i < msg_end; ++i) {
    const uint32_t bit = post[i] >= 0.0f;
    const uint32_t fb = ((r >> 23) & 1u) ^ bit;
    r = (r << 1) & 0xFFFFFFu;
    if (fb) r ^= 0xC00063u;
  }
  for (uint32_t j = 0; j < 24; ++j) {
    const uint32_t bit = post[msg_end + j] >= 0.0f;
    if (bit != ((r >> (23 - j)) & 1u)) return false;
  }
  return true;
}

}  // namespace

// groups: kExactBlocksPerWave code-block indices per wave (kNoBlock where
// the group has fewer), all of one K.
__global__ __launch_bounds__(64, 2) void exact_decode_kernel(const CbInfo* __restrict__ cbs,
                                                          const uint32_t* __restrict__ groups, uint32_t n_groups,
                                                          const float* __restrict__ llr, float* __restrict__ post,
                                                          float* __restrict__ work, int32_t* __restrict__ crc_ok,
                                                          uint32_t* __restrict__ iters, uint32_t max_iterations,
                                                          int32_t early_stop) {
  constexpr uint32_t B = kExactBlocksPerWave;
  const uint32_t lane = threadIdx.x, g = lane / kExactLanes, r = lane % kExactLanes, gbase = g * kExactLanes;
  for (uint32_t gi = blockIdx.x; gi < n_groups; gi += gridDim.x) {
    const uint32_t id = groups[B * gi + g];
    const bool active = id != kNoBlock;
    const CbInfo cb = cbs[active ? id : groups[B * gi]];  // idle blocks shadow the group's first
    const uint32_t k = cb.k;
    const float* d0 = llr + cb.llr_off;
    const float* d1 = d0 + (k + 4);
    const float* d2 = d1 + (k + 4);
    float* out = post + cb.post_off;
    float* alpha = work + cb.work_off;
    float* beta = alpha + 8 * (size_t)(k + 1);
    float* e1 = beta + 8 * (size_t)(k + 1);
    float* e2 = e1 + k;
    float* si = e2 + k;
    float* ia = si + k;
    // decoder tail read-out of lte_turbo.cpp:758-773
    const float us[3] = {d0[k + 0], d0[k + 2], 0.0f};
    const float up[3] = {d1[k + 0], d0[k + 1], d1[k + 2]};
    const float ls[3] = {d1[k + 3], d2[k + 0], d2[k + 2]};
    const float lp[3] = {d2[k + 1], d0[k + 3], d2[k + 3]};
    if (active) {
      QppStride q(r, k, cb.f1, cb.f2);
      for (uint32_t t = r; t < k; t += kExactLanes, q.next()) {
        si[t] = d0[q.pi];
        e2[t] = 0.0f;
        out[t] = 0.0f;
      }
    }
    __syncthreads();
    // a block whose check passed is done: its lanes stop writing, so its
    // posteriors stay those of the passing iteration, while the wave goes on
    // for its other blocks
    bool ok = false, done = false;
    uint32_t used = 0;
    for (uint32_t it = 0; it < max_iterations; ++it) {
      const bool wr = active && !done;
      siso(r, gbase, d0, e2, d1, k, us, up, alpha, beta, e1, false, cb.f1, cb.f2, wr);
      if (wr) {
        QppStride q(r, k, cb.f1, cb.f2);
        for (uint32_t t = r; t < k; t += kExactLanes, q.next()) ia[t] = e1[q.pi];
      }
      __syncthreads();
      siso(r, gbase, si, ia, d2, k, ls, lp, alpha, beta, e2, true, cb.f1, cb.f2, wr);
      if (wr)
        for (uint32_t t = r; t < k; t += kExactLanes) out[t] = d0[t] + e1[t] + e2[t];
      __syncthreads();
      if (!done) used = it + 1;
      if (early_stop) {
        // one lane per block runs the reference's bitwise division; the
        // verdict goes to all eight lanes of the block
        const bool mine = (r == 0 && wr) ? crc24b_ok(out, k, cb.filler) : false;
        if (from_lane(mine ? 1.0f : 0.0f, gbase) != 0.0f) ok = done = true;
        if (__syncthreads_and(done || !active)) break;
      }
    }
    if (active && r == 0) {
      if (crc_ok) crc_ok[id] = ok ? 1 : 0;
      if (iters) iters[id] = used;
    }
    __syncthreads();
  }
}

}  // namespace tcdev

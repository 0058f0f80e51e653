// Exact Log-MAP turbo decoder: one lane per code block, full-length forward
// and backward recursions with no windowing, evaluated in the reference's
// operation order (lte_turbo.cpp:380-486, 736-813) so that posterior LLRs
// agree with turbo_decode_code_block_cpu to rounding of expf/log1pf.
//
// This is the parity path behind GpuStage1Exact; the window kernel is the
// throughput path.  Work space per code block (float offsets from work_off):
//   beta  (K+1)*8   backward metrics, all steps (read back by the forward pass)
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

__device__ __forceinline__ bool fin(float v) { return __builtin_isfinite(v); }

// max* with the reference's infinity rules (lte_turbo.cpp:380-390).
__device__ __forceinline__ float jac(float x, float y) {
  if (!fin(x)) return y;
  if (!fin(y)) return x;
  const float hi = (x < y) ? y : x;
  const float lo = (y < x) ? y : x;
  return hi + log1pf(expf(lo - hi));
}

// lte_turbo.cpp:392-402
__device__ __forceinline__ void renorm(float* m) {
  float mx = m[0];
#pragma unroll
  for (int i = 1; i < 8; ++i) mx = (mx < m[i]) ? m[i] : mx;
  if (!fin(mx)) return;
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (fin(m[i])) m[i] -= mx;
}

__device__ __forceinline__ float bm(unsigned u, unsigned p, float sa, float lp) {
  return 0.5f * ((u ? 1.0f : -1.0f) * sa + (p ? 1.0f : -1.0f) * lp);
}

// Tail beta (lte_turbo.cpp:357-378) from three systematic/parity tail LLRs.
__device__ void tail_beta(const float* ts, const float* tp, float* beta) {
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
    beta[s] = (cur == 0) ? metric : -kInf;
  }
  renorm(beta);
}

// One SISO half.  sys[t], ap[t] (t < K), par[t]; the extrinsic of step t is
// written to ext[perm ? perm(t) : t] where perm is the QPP when interleaved.
__device__ void siso(const float* sys, const float* ap, const float* par, uint32_t k,
                     const float* ts, const float* tp, float* beta, float* ext, bool scatter,
                     uint32_t f1, uint32_t f2) {
  float b[8];
  tail_beta(ts, tp, b);
  float4* bst = reinterpret_cast<float4*>(beta);
  bst[2 * k] = make_float4(b[0], b[1], b[2], b[3]);
  bst[2 * k + 1] = make_float4(b[4], b[5], b[6], b[7]);
  for (uint32_t t = k; t-- > 0;) {
    const float sa = sys[t] + ap[t];
    const float lp = par[t];
    float nb[8];
#pragma unroll
    for (unsigned s = 0; s < 8; ++s) {
      const Tr t0 = rsc(s, 0), t1 = rsc(s, 1);
      const float c0 = b[t0.next] + bm(0, t0.parity, sa, lp);
      const float c1 = b[t1.next] + bm(1, t1.parity, sa, lp);
      // the reference skips non-finite successors; jac() does the same
      nb[s] = jac(fin(b[t0.next]) ? c0 : -kInf, fin(b[t1.next]) ? c1 : -kInf);
    }
    renorm(nb);
#pragma unroll
    for (int s = 0; s < 8; ++s) b[s] = nb[s];
    bst[2 * t] = make_float4(b[0], b[1], b[2], b[3]);
    bst[2 * t + 1] = make_float4(b[4], b[5], b[6], b[7]);
  }
  float a[8] = {0.0f, -kInf, -kInf, -kInf, -kInf, -kInf, -kInf, -kInf};
  // forward: alpha, LLR; QPP by recurrence pi(t+1) = pi(t) + d(t), d(t+1) = d(t) + 2 f2
  uint32_t pi = 0, d = (f1 + f2) % k;
  const uint32_t d2 = (2 * f2) % k;
  for (uint32_t t = 0; t < k; ++t) {
    const float sa = sys[t] + ap[t];
    const float lp = par[t];
    const float4 bl = bst[2 * (t + 1)], bh = bst[2 * (t + 1) + 1];
    const float bn[8] = {bl.x, bl.y, bl.z, bl.w, bh.x, bh.y, bh.z, bh.w};
    float l[2] = {-kInf, -kInf};
#pragma unroll
    for (unsigned s = 0; s < 8; ++s) {
#pragma unroll
      for (unsigned u = 0; u < 2; ++u) {
        const Tr tr = rsc(s, u);
        const float m = a[s] + bm(u, tr.parity, sa, lp) + bn[tr.next];
        l[u] = fin(l[u]) ? jac(l[u], m) : m;
      }
    }
    ext[scatter ? pi : t] = (l[1] - l[0]) - sys[t] - ap[t];
    float na[8];
#pragma unroll
    for (unsigned n = 0; n < 8; ++n) {
      const unsigned u = in_of(n), p = pred0(n), c = par0(n);
      const float ca = a[p] + bm(u, c, sa, lp);
      const float cb = a[p + 1] + bm(u, c ^ 1u, sa, lp);
      na[n] = jac(fin(a[p]) ? ca : -kInf, fin(a[p + 1]) ? cb : -kInf);
    }
    renorm(na);
#pragma unroll
    for (int s = 0; s < 8; ++s) a[s] = na[s];
    pi += d;
    if (pi >= k) pi -= k;
    d += d2;
    if (d >= k) d -= k;
  }
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

__global__ __launch_bounds__(64) void exact_decode_kernel(const CbInfo* __restrict__ cbs, uint32_t n_cb,
                                                          const float* __restrict__ llr, float* __restrict__ post,
                                                          float* __restrict__ work, int32_t* __restrict__ crc_ok,
                                                          uint32_t* __restrict__ iters, uint32_t max_iterations,
                                                          int32_t early_stop) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cb) return;
  const CbInfo cb = cbs[i];
  const uint32_t k = cb.k;
  const float* d0 = llr + cb.llr_off;
  const float* d1 = d0 + (k + 4);
  const float* d2 = d1 + (k + 4);
  float* out = post + cb.post_off;
  float* beta = work + cb.work_off;
  float* e1 = beta + 8 * (size_t)(k + 1);
  float* e2 = e1 + k;
  float* si = e2 + k;
  float* ia = si + k;
  // decoder tail read-out of lte_turbo.cpp:758-773
  const float us[3] = {d0[k + 0], d0[k + 2], 0.0f};
  const float up[3] = {d1[k + 0], d0[k + 1], d1[k + 2]};
  const float ls[3] = {d1[k + 3], d2[k + 0], d2[k + 2]};
  const float lp[3] = {d2[k + 1], d0[k + 3], d2[k + 3]};
  {
    uint32_t pi = 0, d = (cb.f1 + cb.f2) % k;
    const uint32_t dd = (2 * cb.f2) % k;
    for (uint32_t t = 0; t < k; ++t) {
      si[t] = d0[pi];
      e2[t] = 0.0f;
      out[t] = 0.0f;
      pi += d;
      if (pi >= k) pi -= k;
      d += dd;
      if (d >= k) d -= k;
    }
  }
  bool ok = false;
  uint32_t used = 0;
  for (uint32_t it = 0; it < max_iterations; ++it) {
    siso(d0, e2, d1, k, us, up, beta, e1, false, cb.f1, cb.f2);
    {
      uint32_t pi = 0, d = (cb.f1 + cb.f2) % k;
      const uint32_t dd = (2 * cb.f2) % k;
      for (uint32_t t = 0; t < k; ++t) {
        ia[t] = e1[pi];
        pi += d;
        if (pi >= k) pi -= k;
        d += dd;
        if (d >= k) d -= k;
      }
    }
    siso(si, ia, d2, k, ls, lp, beta, e2, true, cb.f1, cb.f2);
    for (uint32_t t = 0; t < k; ++t) out[t] = d0[t] + e1[t] + e2[t];
    used = it + 1;
    if (early_stop) {
      ok = crc24b_ok(out, k, cb.filler);
      if (ok) break;
    }
  }
  if (crc_ok) crc_ok[i] = ok ? 1 : 0;
  if (iters) iters[i] = used;
}

}  // namespace tcdev

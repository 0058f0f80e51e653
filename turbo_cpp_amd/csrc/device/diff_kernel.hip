// Differentiable turbo decoder: forward pass and hand-written adjoint.
//
// The decoder is a sliding-window Log-MAP turbo decoder with a learnable
// scale per half-iteration: the extrinsic output of half h is multiplied by
// w[h] wherever it is used (the next half's a-priori input and, for the last
// two halves, the posterior).  With w = 1 and one segment per block it is the
// reference's iteration (lte_turbo.cpp:776-790) with its exact SISO
// (lte_turbo.cpp:410-486): gamma = (x_u (sys + la) + z_p par) / 2,
// alpha/beta by the Jacobian logarithm, ext = L1 - L0 - sys - la.
//
// Parallel layout: one wave per job of nblk code blocks of one K; each block
// is cut into M = K / S segments, lane = block * M + segment.  Segment q owns
// steps [qS, qS + S): its alpha starts from zero metrics g steps early (from
// state 0 when that reaches step 0) and its beta from zero metrics g steps
// late (from the tail beta when that reaches K).  A segment therefore
// depends only on the half-iteration's inputs in [qS - g, qS + S + g), so
// each half-iteration is a function of the previous one's extrinsics and the
// adjoint is local too: a segment's backward recomputes its alpha and beta
// over that window (the checkpoints are the segment boundaries; nothing but
// the extrinsics of every half is stored by the forward pass), then runs
//   - the output adjoint of the extrinsic (softmax weights of L1 and L0),
//   - the alpha adjoint backwards in time (weights of each state's two
//     incoming transitions),
//   - the beta adjoint forwards in time (weights of each state's two
//     outgoing transitions),
// and reduces the branch-metric adjoints onto sys + la and par.  Warm-up
// steps belong to neighbouring segments; their contributions are added in a
// fixed order (own, next segment's alpha warm-up, previous segment's beta
// warm-up), so the result is deterministic.  Per-step renormalisation does
// not change any output (the LLRs are invariant to adding a constant to all
// eight metrics of a step), so it has no adjoint.
//
// Per-lane scratch is laid out [field][j][lane] (j = local step), so lanes in
// lockstep access consecutive addresses.  Templated on the real type: the
// float build is the fast one, the double build serves finite-difference
// checks.
#include <hip/hip_runtime.h>

#include "diff_types.h"

namespace tcdev {
namespace {

// The recursions run in base 2: state metrics, branch metrics and LLR terms
// are the natural-log quantities times log2(e), so each Jacobian logarithm is
// one exp2 and one log2 with no scaling; an LLR difference is taken back to
// natural units once (times ln 2).  The max* weights exp(x - y) / (1 + ...)
// are the same numbers in either base, so the adjoints, which are built from
// those weights and from the natural-unit derivatives of gamma (+-1/2 per
// input), stay in natural units.
template <typename R>
struct Num;
// float: the hardware exp2 / log2 / reciprocal (about 1 ulp); double: libm
// and IEEE division, for the finite-difference and FP64 reference checks.
template <>
struct Num<float> {
  static __device__ float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
  static __device__ float lg2(float x) { return __builtin_amdgcn_logf(x); }
  static __device__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};
template <>
struct Num<double> {
  static __device__ double ex2(double x) { return exp2(x); }
  static __device__ double lg2(double x) { return log2(x); }
  static __device__ double rcp(double x) { return 1.0 / x; }
};
// half of log2(e): a branch metric's input scale in base 2 (gamma = +-x/2)
template <typename R>
constexpr R kHalfL2e = R(0.72134752044448170368);
template <typename R>
constexpr R kLn2 = R(0.69314718055994530942);

constexpr float kNegF = -1.0e30f;

// RSC transition (lte_turbo.cpp:333-341): feedback = u ^ s1 ^ s2,
// parity = feedback ^ s0 ^ s2, next = (feedback << 2) | (s2 << 1) | s1.
__host__ __device__ constexpr unsigned nxt(unsigned s, unsigned u) {
  return ((u ^ (s >> 1) ^ (s >> 2)) & 1u) << 2 | ((s >> 2) & 1u) << 1 | ((s >> 1) & 1u);
}
__host__ __device__ constexpr unsigned par(unsigned s, unsigned u) {
  return ((u ^ (s >> 1) ^ (s >> 2)) ^ s ^ (s >> 2)) & 1u;
}
// input bit of the termination step from state s (lte_turbo.cpp:343-353)
__host__ __device__ constexpr unsigned term_in(unsigned s) { return ((s >> 1) ^ (s >> 2)) & 1u; }

// Jacobian logarithm and the weights of its two arguments (max and |a - b|:
// a - b and b - a differ only in sign, so this is the same value as taking
// the larger minus the smaller).
template <typename R>
__device__ inline R jac(R a, R b) {
  const R m = fmax(a, b);
  const R d = fabs(a - b);
  return m + Num<R>::lg2(R(1) + Num<R>::ex2(-d));
}
template <typename R>
__device__ inline void jac_w(R a, R b, R& wa, R& wb) {
  const R d = fabs(a - b);
  const R e = Num<R>::ex2(-d);
  const R big = Num<R>::rcp(R(1) + e);
  const R small = e * big;
  wa = a >= b ? big : small;
  wb = a >= b ? small : big;
}

template <typename R>
__device__ inline R gam(unsigned u, unsigned p, R hs, R hp) {
  return (u ? hs : -hs) + (p ? hp : -hp);
}

template <typename R>
__device__ inline void normalise(R (&v)[8]) {
  R m = v[0];
#pragma unroll
  for (int s = 1; s < 8; ++s) m = fmax(v[s], m);
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] -= m;
}

// alpha_{t+1} from alpha_t: each state's two incoming transitions in order
// of increasing predecessor state.  Metrics are renormalised (maximum 0) only
// where nrm is set: when the step ends a chunk of 8 local steps, so every
// value a recompute from a checkpoint reproduces was computed the same way
// the first time (the LLRs and the adjoints do not depend on where the
// constant offsets were taken out; within 8 steps the metrics drift by at most
// 8 branch metrics).
template <typename R>
__device__ inline void alpha_step(const R (&a)[8], R hs, R hp, R (&o)[8], bool nrm) {
  bool seen[8] = {false, false, false, false, false, false, false, false};
#pragma unroll
  for (unsigned s = 0; s < 8; ++s)
#pragma unroll
    for (unsigned u = 0; u < 2; ++u) {
      const unsigned n = nxt(s, u);
      const R c = a[s] + gam(u, par(s, u), hs, hp);
      o[n] = seen[n] ? jac(o[n], c) : c;
      seen[n] = true;
    }
  if (nrm) normalise(o);
}

template <typename R>
__device__ inline void beta_step(const R (&b)[8], R hs, R hp, R (&o)[8], bool nrm) {
#pragma unroll
  for (unsigned s = 0; s < 8; ++s)
    o[s] = jac(b[nxt(s, 0)] + gam(0u, par(s, 0), hs, hp), b[nxt(s, 1)] + gam(1u, par(s, 1), hs, hp));
  if (nrm) normalise(o);
}

// L_u = log2 sum_s 2^(alpha(s) + gamma(s,u) + beta(next(s,u))) (base-2
// metrics) and, when P is given, the weights P[u][s] = 2^(m(s,u) - L_u).
template <typename R>
__device__ inline void llr_terms(const R (&a)[8], const R (&b)[8], R hs, R hp, R (&L)[2], R (*P)[8]) {
#pragma unroll
  for (unsigned u = 0; u < 2; ++u) {
    R m[8];
    R mx = R(kNegF);
#pragma unroll
    for (unsigned s = 0; s < 8; ++s) {
      m[s] = a[s] + gam(u, par(s, u), hs, hp) + b[nxt(s, u)];
      mx = fmax(m[s], mx);
    }
    R sum = 0;
#pragma unroll
    for (unsigned s = 0; s < 8; ++s) {
      m[s] = Num<R>::ex2(m[s] - mx);
      sum += m[s];
    }
    L[u] = mx + Num<R>::lg2(sum);
    if (P) {
      const R inv = Num<R>::rcp(sum);
#pragma unroll
      for (unsigned s = 0; s < 8; ++s) P[u][s] = m[s] * inv;
    }
  }
}

// Tail beta (lte_turbo.cpp:357-378): three forced steps to state 0.
template <typename R>
__device__ inline void tail_beta(const R (&ts)[3], const R (&tp)[3], R (&o)[8]) {
#pragma unroll
  for (unsigned s = 0; s < 8; ++s) {
    unsigned cur = s;
    R m = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const unsigned u = term_in(cur);
      m += gam(u, par(cur, u), kHalfL2e<R> * ts[j], kHalfL2e<R> * tp[j]);
      cur = nxt(cur, u);
    }
    o[s] = m;  // every state reaches 0 in three termination steps
  }
  normalise(o);
}

// QPP address Pi(i) = (f1 i + f2 i^2) mod K, walked step by step:
// Pi(i+1) - Pi(i) = f1 + f2 (2i + 1) and that difference grows by 2 f2, all
// mod K (K <= 6144, so every product below fits 32 bits).  One modulo per
// walk instead of a 64-bit modulo per step.
struct QppWalk {
  uint32_t pi, d, d2, k;  // Pi(i), Pi(i+1) - Pi(i), 2 f2, all mod k
  __device__ QppWalk(uint32_t i, uint32_t kk, uint32_t f1, uint32_t f2) : k(kk) {
    const uint32_t im = i % kk;
    pi = (f1 % kk * im % kk + f2 % kk * im % kk * im) % kk;
    d = (f1 % kk + f2 % kk * ((2 * im + 1) % kk)) % kk;
    d2 = 2 * f2 % kk;
  }
  __device__ void next() {
    pi += d;
    pi = pi >= k ? pi - k : pi;
    d += d2;
    d = d >= k ? d - k : d;
  }
  // next() when c, else nothing (no branch)
  __device__ void next_if(bool c) {
    uint32_t p = pi + d, dd = d + d2;
    p = p >= k ? p - k : p;
    dd = dd >= k ? dd - k : dd;
    pi = c ? p : pi;
    d = c ? dd : d;
  }
  __device__ void prev() {
    d = d >= d2 ? d - d2 : d + k - d2;
    pi = pi >= d ? pi - d : pi + k - d;
  }
};

// Per-lane scratch fields, each [j][lane] with j < its length.
struct Layout {
  uint32_t P, S, g;
  uint32_t sa, pr, al, be, bo, dsa, dpr, tl, total;
  __device__ Layout(uint32_t s, uint32_t gw) : P(s + 2 * gw), S(s), g(gw) {
    sa = 0;
    pr = sa + P;
    al = pr + P;                 // (g + S + 1) x 8
    be = al + (g + S + 1) * 8;   // (S + g + 1) x 8
    bo = be + (S + g + 1) * 8;   // S x 8
    dsa = bo + S * 8;
    dpr = dsa + P;
    tl = dpr + P;                // 6
    total = tl + 6;
  }
};

// Address-space-qualified pointers, for code where the optimiser cannot see
// that a pointer is global or LDS memory (a generic pointer would make every
// access a flat one).
template <typename R>
using GlobalPtr = __attribute__((address_space(1))) R*;
template <typename R>
using LdsPtr = __attribute__((address_space(3))) R*;

template <typename R>
struct Lane {
  uint32_t K, S, M, g, q, f1, f2;
  bool active;
  uint32_t lane;
  R* sc;  // this wave's scratch (uniform), element (field + j) * 64 + lane
  // staged inputs: the block's LDS pairs (lds; local step j is position
  // p = qS + j, pair at p + p / 32) or the per-lane sa / pr fields
  bool lds;
  LdsPtr<R> xs;
  // 32-bit byte offsets from the wave's base (a slot is far below 4 GiB):
  // one 32-bit offset per field and the steps as immediate offsets, not a
  // 64-bit address per access
  __device__ R& nb(uint32_t field, uint32_t j, uint32_t other) const {
    return *reinterpret_cast<R*>(reinterpret_cast<char*>(sc) + (uint32_t)(((field + j) * 64u + other) * sizeof(R)));
  }
  __device__ R& at(uint32_t field, uint32_t j) const { return nb(field, j, lane); }
  // local step j covers t = qS - g + j
  __device__ int t_of(uint32_t j) const { return (int)(q * S) - (int)g + (int)j; }
};

// The lane with its lane index made opaque to the optimiser.  Each pass starts from such a copy, so the per-step scratch
// offsets of a pass are not hoisted out of the half-iteration loop (where
// they would hold registers through every other pass).
template <typename R>
__device__ __forceinline__ Lane<R> fresh(const Lane<R>& L0) {
  Lane<R> L = L0;
  asm volatile("" : "+v"(L.lane));
  return L;
}

// Inlining: everything is inlined into the kernels.  A job is one wave and
// the kernels run one wave per SIMD, so a lane has 512 registers (256 VGPRs
// and 256 AGPRs); a call would instead save and restore the callee's
// registers through private memory on every half-iteration.
//
// Memory order and control flow of every sequential pass below.  A job is
// one wave and a batch rarely has more jobs than SIMDs, so no other wave
// hides a wait.  Each pass therefore walks its steps in chunks of 8 with two
// input buffers used in turn: the loads of chunk c + 1 go out before chunk c
// is computed and stored, and no register copy sits between a load and its
// use (a copy would wait for the load at the copy).  The vector-memory
// counter retires loads and stores in issue order, so a load issued behind a
// store would also wait for that store; issued ahead, it waits only for
// stores of earlier chunks.  Loop bounds are wave-uniform (all lanes of a job
// share K, S and g); steps that are outside the block for some lanes
// (t < 0 in the first segment, t >= K in the last) are computed by every
// lane and their results selected away, so no per-lane branch splits a
// chunk.  Values computed that way are only ever written to positions that
// no later pass reads.

// Runs body(j0, buf) for j0 = lo, lo + 8, ... < hi with load(j0, buf) one
// chunk ahead, alternating two buffers.
template <class Buf, class Load, class Body>
__device__ __forceinline__ void chunks_up(uint32_t lo, uint32_t hi, Load&& load, Body&& body) {
  if (lo >= hi) return;
  Buf b0, b1;
  load(lo, b0);
  for (uint32_t j0 = lo; j0 < hi; j0 += 16) {
    if (j0 + 8 < hi) load(j0 + 8, b1);
    body(j0, b0);
    if (j0 + 8 >= hi) break;
    if (j0 + 16 < hi) load(j0 + 16, b0);
    body(j0 + 8, b1);
  }
}
// The same from the top down: body(top, buf) covers steps top - 1 down to
// max(top - 8, lo).
template <class Buf, class Load, class Body>
__device__ __forceinline__ void chunks_down(int lo, int hi, Load&& load, Body&& body) {
  if (hi <= lo) return;
  Buf b0, b1;
  load(hi, b0);
  for (int top = hi; top > lo; top -= 16) {
    if (top - 8 > lo) load(top - 8, b1);
    body(top, b0);
    if (top - 8 <= lo) break;
    if (top - 16 > lo) load(top - 16, b0);
    body(top - 8, b1);
  }
}

template <typename R>
__device__ __forceinline__ void select8(bool c, const R (&x)[8], R (&v)[8]) {
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] = c ? x[s] : v[s];
}

// Staged inputs (sys + la, par) of up to 8 consecutive local steps: entry jj
// is step j0 + jj (ascending) or j0 - jj (descending).
template <typename R>
struct In8 {
  R sa[8], pr[8];
};
template <typename R, bool DOWN = false>
__device__ __forceinline__ void load_in8(const Lane<R>& L, const Layout& Y, int j0, int n, In8<R>& c) {
  if (L.lds) {
    const uint32_t pb = L.q * L.S;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj)
      if (jj < n) {
        const uint32_t p = pb + (uint32_t)(DOWN ? j0 - jj : j0 + jj);
        const uint32_t i = 2 * (p + (p >> 5));
        c.sa[jj] = L.xs[i];
        c.pr[jj] = L.xs[i + 1];
      }
  } else {
#pragma unroll
    for (int jj = 0; jj < 8; ++jj)
      if (jj < n) {
        const uint32_t j = (uint32_t)(DOWN ? j0 - jj : j0 + jj);
        c.sa[jj] = L.at(Y.sa, j);
        c.pr[jj] = L.at(Y.pr, j);
      }
  }
}

// Stage sys + la and par of one half over the lane's window.  Upper half
// (h even): natural order, la = w[h-1] e_{h-1}[t].  Lower half: position i
// reads d0[Pi(i)], d2[i] and la = w[h-1] e_{h-1}[Pi(i)] (extrinsics are kept
// in natural order).  Steps outside [0, K) stage zeros (their loads read a
// clamped, valid position).
template <typename R>
struct StageIn {
  R s[8], e[8], p[8];  // d0, e_{h-1} (at t or Pi(t)), parity
};

template <typename R>
__device__ __forceinline__ void stage(const Lane<R> L0, const Layout Y, uint32_t h, const R* d0, const R* d1,
                                      const R* d2, const R* sv, const R* scales) {
  const Lane<R> L = fresh(L0);
  const R w = h ? scales[h - 1] : R(0);
  const R* eprev = sv + (size_t)(h ? h - 1 : 0) * L.K;
  const bool lower = (h & 1u) != 0;
  const R* dp = lower ? d2 : d1;
  const int t0 = L.t_of(0);
  QppWalk qw(t0 > 0 ? (uint32_t)t0 : 0u, L.K, L.f1, L.f2);
  auto in_block = [&](uint32_t j) {
    const int t = L.t_of(j);
    return t >= 0 && t < (int)L.K;
  };
  auto load = [&](uint32_t j0, StageIn<R>& c) {
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      if (j0 + jj < Y.P) {
        const int t = L.t_of(j0 + jj);
        const uint32_t tc = (uint32_t)min(max(t, 0), (int)L.K - 1);
        const uint32_t i = lower ? qw.pi : tc;
        c.s[jj] = d0[i];
        c.e[jj] = h ? eprev[i] : R(0);
        c.p[jj] = dp[tc];
        qw.next_if(in_block(j0 + jj));
      }
    }
  };
  auto body = [&](uint32_t j0, const StageIn<R>& c) {
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      const uint32_t j = j0 + jj;
      if (j < Y.P) {
        const R sa = lower ? c.s[jj] + w * c.e[jj] : c.s[jj] + (h ? w * c.e[jj] : R(0));
        const bool v = in_block(j);
        L.at(Y.sa, j) = v ? sa : R(0);
        L.at(Y.pr, j) = v ? c.p[jj] : R(0);
      }
    }
  };
  chunks_up<StageIn<R>>(0, Y.P, load, body);
}

// LDS copies of per-block arrays of the job (all lanes take part in filling
// and writing back, in natural order, so every global access of a fill is a
// contiguous 256-byte row per wave).  Used when the plan's jobs fit (use_lds);
// then staging and the adjoint combine read and update these copies instead
// of gathering with a stride of S elements between lanes.
template <typename R>
struct BlockLds {
  R* base;
  uint32_t kp, nblk;  // diff_lds_stride(K), blocks of the job
  uint32_t kq;        // diff_stage_pairs(K, g)
  __device__ uint32_t idx(uint32_t arr, uint32_t b, uint32_t i) const { return (arr * nblk + b) * kp + i + (i >> 5); }
  __device__ R& at(uint32_t arr, uint32_t b, uint32_t i) const { return base[idx(arr, b, i)]; }
  // first element of block b's staged pairs (after arrays 0 and 1)
  __device__ uint32_t stage_base(uint32_t b) const { return 2 * nblk * kp + 2 * b * kq; }
};
template <typename R>
__device__ __forceinline__ GlobalPtr<R> global_view(R* p) {
  return (GlobalPtr<R>)p;
}
template <typename R>
__device__ __forceinline__ GlobalPtr<const R> global_view(const R* p) {
  return (GlobalPtr<const R>)p;
}
template <typename R>
__device__ __forceinline__ LdsPtr<R> lds_view(const BlockLds<R>& X) {
  return (LdsPtr<R>)X.base;
}

template <typename R>
__device__ __forceinline__ BlockLds<R> block_lds(R* base, const DiffJob& J) {
  return BlockLds<R>{base, diff_lds_stride(J.k), J.nblk, diff_stage_pairs(J.k, J.warm)};
}

constexpr uint32_t kFillBatch = 16;  // loads in flight per lane and array

// Fill of the staging arrays: 0 = sys + la at natural position i (the lower
// half reads it at Pi(t)), 1 = the half's parity stream.  Same arithmetic as
// the per-lane staging.  Two blocks per batch of loads, so a job of two
// blocks waits for memory once.
template <typename R>
__device__ __forceinline__ void stage_fill(const DiffArgs<R> A, const DiffJob J, uint32_t h, const BlockLds<R> X) {
  const R w = h ? A.scales[h - 1] : R(0);
  const bool lower = (h & 1u) != 0;
  const uint32_t K = J.k;
  const LdsPtr<R> xl = lds_view(X);
  for (uint32_t b0 = 0; b0 < J.nblk; b0 += 2) {
    const uint32_t nb = min(2u, J.nblk - b0);
    GlobalPtr<const R> d0[2], dp[2], ep[2];
#pragma unroll
    for (uint32_t x = 0; x < 2; ++x) {
      const DiffCb C = A.cbs[J.cb_first + b0 + (x < nb ? x : 0u)];
      d0[x] = global_view(A.llr + C.llr_off);
      dp[x] = d0[x] + (size_t)(lower ? 2 : 1) * (K + 4);
      ep[x] = global_view(A.saved + C.save_off + (size_t)(h ? h - 1 : 0) * K);
    }
    for (uint32_t i0 = threadIdx.x; i0 < K; i0 += 64 * kFillBatch) {
      R sv[2][kFillBatch], ev[2][kFillBatch], pv[2][kFillBatch];
#pragma unroll
      for (uint32_t x = 0; x < 2; ++x)
#pragma unroll
        for (uint32_t u = 0; u < kFillBatch; ++u) {
          const uint32_t i = i0 + 64 * u;
          if (x < nb && i < K) {
            sv[x][u] = d0[x][i];
            ev[x][u] = h ? ep[x][i] : R(0);
            pv[x][u] = dp[x][i];
          }
        }
#pragma unroll
      for (uint32_t x = 0; x < 2; ++x)
#pragma unroll
        for (uint32_t u = 0; u < kFillBatch; ++u) {
          const uint32_t i = i0 + 64 * u;
          if (x < nb && i < K) {
            xl[X.idx(0, b0 + x, i)] = lower ? sv[x][u] + w * ev[x][u] : sv[x][u] + (h ? w * ev[x][u] : R(0));
            xl[X.idx(1, b0 + x, i)] = pv[x][u];
          }
        }
    }
  }
}

// Staged pairs of the half from arrays 0 and 1 (see stage for the values):
// pair p of block b holds (sys + la, parity) at position t = p - g of the
// half's order, zeros outside [0, K).  Each lane takes a run of consecutive
// positions (the lower half's interleaver walked by its recurrence).
template <typename R>
__device__ __forceinline__ void stage_gather(const DiffJob J, uint32_t h, const BlockLds<R> X) {
  const LdsPtr<R> xl = lds_view(X);
  const bool lower = (h & 1u) != 0;
  const uint32_t K = J.k, g = J.warm, n = K + 2 * g;
  const uint32_t run = (n + 63) / 64, p0 = threadIdx.x * run;
  const int t0 = (int)p0 - (int)g;
  for (uint32_t b = 0; b < J.nblk; ++b) {
    const uint32_t sb = X.stage_base(b);
    QppWalk qw(t0 > 0 ? (uint32_t)t0 : 0u, K, J.f1, J.f2);
    for (uint32_t u = 0; u < run; ++u) {
      const uint32_t p = p0 + u;
      if (p < n) {
        const int t = (int)p - (int)g;
        const bool v = t >= 0 && t < (int)K;
        const uint32_t tc = v ? (uint32_t)t : 0u;
        const R sa = xl[X.idx(0, b, lower ? qw.pi : tc)];
        const R pr = xl[X.idx(1, b, tc)];
        const uint32_t i = sb + 2 * (p + (p >> 5));
        xl[i] = v ? sa : R(0);
        xl[i + 1] = v ? pr : R(0);
        qw.next_if(v);
      }
    }
  }
}

template <typename R>
__device__ void tail_inputs(uint32_t h, const R* d0, const R* d1, const R* d2, uint32_t K, R (&ts)[3], R (&tp)[3]) {
  // the decoder's tail read-out (lte_turbo.cpp:758-773)
  if ((h & 1u) == 0) {
    ts[0] = d0[K];
    ts[1] = d0[K + 2];
    ts[2] = 0;
    tp[0] = d1[K];
    tp[1] = d0[K + 1];
    tp[2] = d1[K + 2];
  } else {
    ts[0] = d1[K + 3];
    ts[1] = d2[K];
    ts[2] = d2[K + 2];
    tp[0] = d2[K + 1];
    tp[1] = d0[K + 3];
    tp[2] = d2[K + 3];
  }
}

// Runs body(c, buf) for c = hi - 1 down to lo with load(c, buf) one ahead,
// alternating two buffers.
template <class Buf, class Load, class Body>
__device__ __forceinline__ void subwindows_down(int lo, int hi, Load&& load, Body&& body) {
  if (hi <= lo) return;
  Buf b0, b1;
  load(hi - 1, b0);
  for (int c = hi - 1; c >= lo; c -= 2) {
    if (c - 1 >= lo) load(c - 1, b1);
    body(c, b0);
    if (c - 1 < lo) break;
    if (c - 2 >= lo) load(c - 2, b0);
    body(c - 1, b1);
  }
}

// Forward beta pass over the own steps that also emits their extrinsics:
// beta from local g + S (b: beta there, after the warm-up of
// alpha_ckpt_beta_warm_pass) down to local g, in sub-windows [8c, 8c + 8) of
// local steps; alpha of each sub-window is recomputed from its checkpoint
// (al[c], alpha at local 8c) into registers first.  Past the block end
// (t >= K) beta stays at its start value; beta is not stored.
template <typename R>
struct BetaExtIn {
  In8<R> in;  // steps 8c + jj
  R a0[8];    // alpha at local 8c
};

template <typename R>
__device__ __forceinline__ void beta_ext_pass(const Lane<R>& L0, const Layout& Y, R (&b)[8], uint32_t h, R* e) {
  const Lane<R> L = fresh(L0);
  const int own_end = (int)(L.g + L.S);  // own steps: g <= j < g + S
  const int g = (int)L.g;
  const uint32_t st = Y.g + Y.S + 1;
  // Pi of the own steps, from the last one down
  QppWalk qw((uint32_t)L.t_of((uint32_t)own_end - 1), L.K, L.f1, L.f2);
  auto load = [&](int c, BetaExtIn<R>& x) {
    const int j0 = 8 * c;
    load_in8(L, Y, j0, min(8, own_end - j0), x.in);
    load8(L, Y.al, st, (uint32_t)c, x.a0);
  };
  auto body = [&](int c, const BetaExtIn<R>& x) {
    const int j0 = 8 * c;
    R areg[8][8];
    {
      R a[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) a[s] = x.a0[s];
#pragma unroll
      for (int jj = 0; jj < 8; ++jj) {
#pragma unroll
        for (int s = 0; s < 8; ++s) areg[jj][s] = a[s];
        if (jj + 1 < 8 && j0 + jj + 1 < own_end) {
          R o[8];
          alpha_step(a, kHalfL2e<R> * x.in.sa[jj], kHalfL2e<R> * x.in.pr[jj], o, false);
          select8(L.t_of((uint32_t)(j0 + jj)) >= 0, o, a);
        }
      }
    }
#pragma unroll
    for (int jj = 7; jj >= 0; --jj) {
      const int j = j0 + jj;
      if (j >= g && j < own_end) {
        const R sa = x.in.sa[jj], pr = x.in.pr[jj];
        const int t = L.t_of((uint32_t)j);
        {
          R Lu[2];
          llr_terms(areg[jj], b, kHalfL2e<R> * sa, kHalfL2e<R> * pr, Lu, (R(*)[8]) nullptr);
          e[(h & 1u) ? qw.pi : (uint32_t)t] = (Lu[1] - Lu[0]) * kLn2<R> - sa;
          if (j > g) qw.prev();
        }
        R o[8];
        beta_step(b, kHalfL2e<R> * sa, kHalfL2e<R> * pr, o, ((j - g) & 7) == 0);
        select8(t < (int)L.K, o, b);
      }
    }
  };
  subwindows_down<BetaExtIn<R>>(g / 8, (own_end + 7) / 8, load, body);
}

template <typename R>
__device__ __forceinline__ void load8(const Lane<R>& L, uint32_t field, uint32_t stride, uint32_t j, R (&v)[8]) {
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] = L.at(field + s * stride, j);
}

// Alpha over local steps 0 .. g + S with only every 8th stored (al[c] =
// alpha at local 8c; alpha stays at its start value before step 0 of the
// block, t < 0) and, in the same loop, beta over the warm-up above the own
// steps: local 2g + S - 1 down to g + S, from bcur (beta at local 2g + S) to
// beta at local g + S, with beta_ck(i, beta) called before each beta step
// (i = j + 1 - g for step j) to keep checkpoints.  The two recursions are
// independent; a job is one wave and rarely shares its SIMD, so each chunk
// interleaves them to give the SIMD two dependency chains to issue from.
// Every value is the one the two recursions compute run one after the other.
template <typename R>
struct AlphaBetaIn {
  In8<R> up;  // alpha inputs, local j0 + jj
  In8<R> dn;  // beta inputs, local 2g + S - 1 - (j0 + jj)
};

template <typename R, class Ck>
__device__ __forceinline__ void alpha_ckpt_beta_warm_pass(const Lane<R>& L0, const Layout& Y, R (&bcur)[8],
                                                          Ck&& beta_ck) {
  const Lane<R> L = fresh(L0);
  R a[8];
  const bool from_zero_state = (int)(L.q * L.S) - (int)L.g <= 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) a[s] = from_zero_state ? (s == 0 ? R(0) : R(kNegF)) : R(0);
  const uint32_t n = L.g + L.S;
  const int top = (int)(2 * L.g + L.S);
  auto load = [&](uint32_t j0, AlphaBetaIn<R>& c) {
    load_in8(L, Y, (int)j0, (int)min(8u, n - j0), c.up);
    if (j0 < L.g) load_in8<R, true>(L, Y, top - 1 - (int)j0, min(8, (int)L.g - (int)j0), c.dn);
  };
  auto body = [&](uint32_t j0, const AlphaBetaIn<R>& c) {
#pragma unroll
    for (int s = 0; s < 8; ++s) L.at(Y.al + s * (Y.g + Y.S + 1), j0 / 8) = a[s];
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      const uint32_t j = j0 + jj;
      if (j < n) {
        R o[8];
        alpha_step(a, kHalfL2e<R> * c.up.sa[jj], kHalfL2e<R> * c.up.pr[jj], o, ((j + 1) & 7u) == 0);
        select8(L.t_of(j) >= 0, o, a);
      }
      if (j < L.g) {
        const int jb = top - 1 - (int)j;
        beta_ck((uint32_t)jb + 1 - L.g, bcur);
        R o[8];
        beta_step(bcur, kHalfL2e<R> * c.dn.sa[jj], kHalfL2e<R> * c.dn.pr[jj], o, (((uint32_t)jb - L.g) & 7u) == 0);
        select8(L.t_of((uint32_t)jb) < (int)L.K, o, bcur);
      }
    }
  };
  chunks_up<AlphaBetaIn<R>>(0, n, load, body);
}

// One step of the alpha adjoint at local j (alpha_j = a, beta_{j+1} = b,
// half gammas hs, hp): the output adjoint of the extrinsic e when the step is
// the lane's own (then the beta output adjoints of the step go to bo[j - g]),
// then the adjoint of alpha_{j+1} pulled back through its two incoming
// transitions; ab holds the adjoint of alpha_{j+1} on entry and of alpha_j on
// exit; ds, dp: the adjoints of sys + la and par of the step.
template <typename R>
__device__ inline void alpha_adjoint_step(const R (&a)[8], const R (&b)[8], R hs, R hp, R e, bool own, R (&ab)[8],
                                          R& ds_out, R& dp_out, const Lane<R>& L, const Layout& Y, uint32_t j) {
  R dg[8][2];
#pragma unroll
  for (int s = 0; s < 8; ++s) dg[s][0] = dg[s][1] = 0;
  R abn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  R dsa_direct = 0;
  if (own) {
    R Lu[2], P[2][8];
    llr_terms(a, b, hs, hp, Lu, P);
    R bo[8] = {0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
    for (unsigned s = 0; s < 8; ++s)
#pragma unroll
      for (unsigned u = 0; u < 2; ++u) {
        const R dm = u ? e * P[1][s] : -e * P[0][s];
        dg[s][u] += dm;
        abn[s] += dm;
        bo[nxt(s, u)] += dm;
      }
#pragma unroll
    for (int s = 0; s < 8; ++s) L.at(Y.bo + s * Y.S, j - L.g) = bo[s];
    dsa_direct = -e;
  }
  // alpha_{t+1}(n) = jac over its two incoming transitions
  bool seen[8] = {false, false, false, false, false, false, false, false};
  unsigned fs[8], fu[8];
#pragma unroll
  for (unsigned s = 0; s < 8; ++s)
#pragma unroll
    for (unsigned u = 0; u < 2; ++u) {
      const unsigned n = nxt(s, u);
      if (!seen[n]) {
        fs[n] = s;
        fu[n] = u;
        seen[n] = true;
      } else {
        const unsigned s0 = fs[n], u0 = fu[n];
        R w0, w1;
        jac_w(a[s0] + gam(u0, par(s0, u0), hs, hp), a[s] + gam(u, par(s, u), hs, hp), w0, w1);
        dg[s0][u0] += ab[n] * w0;
        abn[s0] += ab[n] * w0;
        dg[s][u] += ab[n] * w1;
        abn[s] += ab[n] * w1;
      }
    }
#pragma unroll
  for (int s = 0; s < 8; ++s) ab[s] = abn[s];
  R ds = dsa_direct, dp = 0;
#pragma unroll
  for (unsigned s = 0; s < 8; ++s)
#pragma unroll
    for (unsigned u = 0; u < 2; ++u) {
      ds += u ? R(0.5) * dg[s][u] : R(-0.5) * dg[s][u];
      dp += par(s, u) ? R(0.5) * dg[s][u] : R(-0.5) * dg[s][u];
    }
  ds_out = ds;
  dp_out = dp;
}

// One step of the beta adjoint at local j (beta_{j+1} = b): bb holds the
// adjoint of beta_j on entry (bn: the step's output adjoints of beta_{j+1})
// and of beta_{j+1} on exit.
template <typename R>
__device__ inline void beta_adjoint_step(const R (&b)[8], R hs, R hp, R (&bb)[8], R (&bn)[8], R& ds_out,
                                         R& dp_out) {
  R ds = 0, dp = 0;
#pragma unroll
  for (unsigned s = 0; s < 8; ++s) {
    R w0, w1;
    jac_w(b[nxt(s, 0)] + gam(0u, par(s, 0), hs, hp), b[nxt(s, 1)] + gam(1u, par(s, 1), hs, hp), w0, w1);
    const R d0g = bb[s] * w0, d1g = bb[s] * w1;
    bn[nxt(s, 0)] += d0g;
    bn[nxt(s, 1)] += d1g;
    ds += R(0.5) * (d1g - d0g);
    dp += par(s, 0) ? R(0.5) * d0g : R(-0.5) * d0g;
    dp += par(s, 1) ? R(0.5) * d1g : R(-0.5) * d1g;
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) bb[s] = bn[s];
  ds_out = ds;
  dp_out = dp;
}

template <typename R>
__device__ Lane<R> make_lane(const DiffJob& J, const DiffArgs<R>& A) {
  Lane<R> L;
  L.lane = threadIdx.x;
  L.K = J.k;
  L.S = J.seg;
  L.M = J.nseg;
  L.g = J.warm;
  L.q = L.lane % J.nseg;
  L.f1 = J.f1;
  L.f2 = J.f2;
  L.active = L.lane / J.nseg < J.nblk;
  L.sc = A.scratch + (size_t)blockIdx.x * 64 * A.slot_elems;
  L.lds = false;
  L.xs = nullptr;
  return L;
}

// The lane reading its staged inputs from the block's LDS pairs.
template <typename R>
__device__ __forceinline__ Lane<R> with_lds(Lane<R> L, const BlockLds<R>& X, uint32_t blk) {
  L.lds = true;
  L.xs = lds_view(X) + X.stage_base(blk);
  return L;
}

}  // namespace


template <typename R>
__global__ __launch_bounds__(64) void diff_forward_kernel(DiffArgs<R> A) {
  extern __shared__ __attribute__((aligned(16))) char diff_lds[];
  const Layout Y(A.max_seg, A.max_warm);
  for (uint32_t jb = blockIdx.x; jb < A.njobs; jb += gridDim.x) {
    const DiffJob J = A.jobs[jb];
    const Lane<R> L0 = make_lane(J, A);
    const uint32_t blk = L0.active ? L0.lane / J.nseg : 0;
    const BlockLds<R> X = block_lds(reinterpret_cast<R*>(diff_lds), J);
    const Lane<R> L = A.use_lds ? with_lds(L0, X, blk) : L0;
    const DiffCb C = A.cbs[J.cb_first + blk];
    const R* d0 = A.llr + C.llr_off;
    const R* d1 = d0 + (L.K + 4);
    const R* d2 = d1 + (L.K + 4);
    R* sv = A.saved + C.save_off;
    for (uint32_t h = 0; h < A.halves; ++h) {
      if (A.use_lds) {
        stage_fill(A, J, h, X);
        __syncthreads();
        stage_gather(J, h, X);
        __syncthreads();
      }
      if (L.active) {
        R ts[3], tp[3];
        tail_inputs(h, d0, d1, d2, L.K, ts, tp);
        if (!A.use_lds) stage(L, Y, h, d0, d1, d2, sv, A.scales);
        R b[8];
        if (L.q * L.S + L.S + L.g >= L.K) {
          tail_beta(ts, tp, b);
        } else {
#pragma unroll
          for (int s = 0; s < 8; ++s) b[s] = 0;
        }
        alpha_ckpt_beta_warm_pass(L, Y, b, [](uint32_t, const R (&)[8]) {});
        beta_ext_pass(L, Y, b, h, sv + (size_t)h * L.K);
      }
      __syncthreads();
    }
    if (L.active) {
      const uint32_t H = A.halves;
      const R* e1 = sv + (size_t)(H - 2) * L.K;
      const R* e2 = sv + (size_t)(H - 1) * L.K;
      R* post = A.post + C.post_off;
      const R w1 = A.scales[H - 2], w2 = A.scales[H - 1];
      const uint32_t hi = L.q * L.S + L.S;
      for (uint32_t t0 = L.q * L.S; t0 < hi; t0 += 8) {
        R x[8], y1[8], y2[8];
#pragma unroll
        for (uint32_t jj = 0; jj < 8; ++jj)
          if (t0 + jj < hi) {
            x[jj] = d0[t0 + jj];
            y1[jj] = e1[t0 + jj];
            y2[jj] = e2[t0 + jj];
          }
#pragma unroll
        for (uint32_t jj = 0; jj < 8; ++jj)
          if (t0 + jj < hi) post[t0 + jj] = x[jj] + w1 * y1[jj] + w2 * y2[jj];
      }
    }
    __syncthreads();
  }
}

namespace {

// Posterior adjoint of one block segment: grad d0 = grad post, the adjoint
// of the last half's extrinsic, and this lane's terms of the last two scales'
// adjoints (post = d0 + w[H-2] e_{H-2} + w[H-1] e_{H-1}).
template <typename R>
__device__ __forceinline__ void post_adjoint(const DiffArgs<R>& A, const Lane<R>& L, const R* sv, const R* gp, R* g0,
                                             R* g1, R* g2, R* el, R& s1, R& s2) {
  const uint32_t H = A.halves, K = L.K, lo = L.q * L.S, hi = lo + L.S;
  const R* e1 = sv + (size_t)(H - 2) * K;
  const R* e2 = sv + (size_t)(H - 1) * K;
  const R wl = A.scales[H - 1];
  for (uint32_t t0 = lo; t0 < hi; t0 += 8) {
    R gv[8], x1[8], x2[8];
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj)
      if (t0 + jj < hi) {
        gv[jj] = gp[t0 + jj];
        x1[jj] = e1[t0 + jj];
        x2[jj] = e2[t0 + jj];
      }
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj)
      if (t0 + jj < hi) {
        const uint32_t t = t0 + jj;
        g0[t] = gv[jj];
        g1[t] = 0;
        g2[t] = 0;
        el[t] = wl * gv[jj];
        s1 += gv[jj] * x1[jj];
        s2 += gv[jj] * x2[jj];
      }
  }
  if (L.q == L.M - 1)
    for (uint32_t t = K; t < K + 4; ++t) g0[t] = g1[t] = g2[t] = 0;
}

// Alpha adjoint, backwards over local steps g + S - 1 .. 0 in sub-windows of
// 8 (alpha recomputed from the checkpoint), the output adjoints of the own
// steps on the way, and beta of the own steps (bcur: beta at local g + S on
// entry), whose checkpoints beta_ck keeps for the beta-adjoint sweep.
// Writes dsa / dpr of local steps below g + S (their first writer).  The
// adjoints of the own steps' extrinsics are read from ecur (natural order:
// position t, or Pi(t) in a lower half), a sub-window ahead like the other
// inputs.
template <typename R>
struct AlphaIn {
  R a0[8], e[8];
  In8<R> in;  // staged inputs (halved where used)
};

template <typename R, class Ck>
__device__ __forceinline__ void alpha_adjoint_sweep(const Lane<R>& L0, const Layout& Y, R (&bcur)[8], Ck&& beta_ck,
                                                    bool lower, const R* ecur) {
  const Lane<R> L = fresh(L0);
  R ab[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t n_al = L.g + L.S;
  const GlobalPtr<const R> eg = global_view(ecur);
  // own steps' positions, walked down from the last one (loads go in
  // descending order of steps)
  QppWalk qw((uint32_t)L.t_of(n_al - 1), L.K, L.f1, L.f2);
  auto load = [&](int c, AlphaIn<R>& x) {
    const uint32_t j0 = 8u * (uint32_t)c, len = min(8u, n_al - j0);
    load8(L, Y.al, Y.g + Y.S + 1, (uint32_t)c, x.a0);
    load_in8(L, Y, (int)j0, (int)len, x.in);
#pragma unroll
    for (int jj = 7; jj >= 0; --jj) {
      if ((uint32_t)jj < len) {
        const uint32_t j = j0 + (uint32_t)jj;
        if (j >= L.g) {
          x.e[jj] = eg[lower ? qw.pi : (uint32_t)L.t_of(j)];
          if (j > L.g) qw.prev();
        } else {
          x.e[jj] = R(0);
        }
      }
    }
  };
  auto body = [&](int c, const AlphaIn<R>& x) {
    const uint32_t j0 = 8u * (uint32_t)c, len = min(8u, n_al - j0);
    R areg[8][8];
    {
      R a[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) a[s] = x.a0[s];
#pragma unroll
      for (uint32_t jj = 0; jj < 8; ++jj) {
        if (jj < len) {
#pragma unroll
          for (int s = 0; s < 8; ++s) areg[jj][s] = a[s];
          if (jj + 1 < len) {
            R o[8];
            alpha_step(a, kHalfL2e<R> * x.in.sa[jj], kHalfL2e<R> * x.in.pr[jj], o, false);
            select8(L.t_of(j0 + jj) >= 0, o, a);
          }
        }
      }
    }
#pragma unroll
    for (int jj = 7; jj >= 0; --jj) {
      if ((uint32_t)jj < len) {
        const uint32_t j = j0 + (uint32_t)jj;
        R b[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) b[s] = bcur[s];  // beta_{j+1} while j >= g
        const R hs = kHalfL2e<R> * x.in.sa[jj], hp = kHalfL2e<R> * x.in.pr[jj];
        if (j >= L.g) {
          // own step (inside the block for every lane): checkpoint
          // beta_{j+1}, step beta to beta_j
          beta_ck(j + 1 - L.g, b);
          R o[8];
          beta_step(b, hs, hp, o, ((j - L.g) & 7u) == 0);
#pragma unroll
          for (int s = 0; s < 8; ++s) bcur[s] = o[s];
        }
        R ds, dp;
        alpha_adjoint_step(areg[jj], b, hs, hp, x.e[jj], j >= L.g, ab, ds, dp, L, Y, j);
        L.at(Y.dsa, j) = ds;
        L.at(Y.dpr, j) = dp;
      }
    }
  };
  subwindows_down<AlphaIn<R>>(0, (int)((n_al + 7) / 8), load, body);
}

// Beta adjoint, forwards over local steps g .. g + S + g - 1 in sub-windows
// of 8 (beta recomputed from the checkpoint at the top); adds to dsa / dpr.
// Returns in bb the adjoint of beta at the window top (at the block end,
// that of the tail beta).
template <typename R>
struct BetaRe {
  R b0[8];
  In8<R> in;  // staged inputs (halved where used)
};

template <typename R>
__device__ __forceinline__ void beta_adjoint_sweep(const Lane<R>& L0, const Layout& Y, R (&bb)[8]) {
  const Lane<R> L = fresh(L0);
  const uint32_t jend = 2 * L.g + L.S;
  const uint32_t bst = Y.S + Y.g + 1;
#pragma unroll
  for (int s = 0; s < 8; ++s) bb[s] = 0;
  auto load = [&](uint32_t jb0, BetaRe<R>& x) {
    const uint32_t len = min(8u, jend - jb0);
    const uint32_t itop = jb0 + len - L.g;  // beta at local jb0 + len = g + itop
    load8(L, Y.be, bst, (itop + 7) / 8, x.b0);
    load_in8(L, Y, (int)jb0, (int)len, x.in);
  };
  auto body = [&](uint32_t jb0, const BetaRe<R>& x) {
    const uint32_t len = min(8u, jend - jb0);
    // the sub-window's output adjoints and running sys/par adjoints, loaded
    // ahead of the recompute
    R bor[8][8], dsr[8], dpr[8];
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      const uint32_t j = jb0 + jj;
      if (jj < len) {
        if (j < L.g + L.S) {
          load8(L, Y.bo, Y.S, j - L.g, bor[jj]);
          dsr[jj] = L.at(Y.dsa, j);
          dpr[jj] = L.at(Y.dpr, j);
        } else {
          // above the own steps this sweep is the only writer: its sums start at zero
#pragma unroll
          for (int s = 0; s < 8; ++s) bor[jj][s] = 0;
          dsr[jj] = dpr[jj] = R(0);
        }
      }
    }
    R breg[8][8];
    {
      R b[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) b[s] = x.b0[s];
#pragma unroll
      for (int jj = 7; jj >= 0; --jj) {
        if ((uint32_t)jj < len) {
#pragma unroll
          for (int s = 0; s < 8; ++s) breg[jj][s] = b[s];  // beta at local j + 1
          if (jj > 0) {
            R o[8];
            beta_step(b, kHalfL2e<R> * x.in.sa[jj], kHalfL2e<R> * x.in.pr[jj], o, false);
            select8(L.t_of(jb0 + (uint32_t)jj) < (int)L.K, o, b);
          }
        }
      }
    }
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      if (jj < len) {
        const uint32_t j = jb0 + jj;
        R nb[8], bn[8], ds, dp;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          nb[s] = bb[s];
          bn[s] = bor[jj][s];
        }
        beta_adjoint_step(breg[jj], kHalfL2e<R> * x.in.sa[jj], kHalfL2e<R> * x.in.pr[jj], nb, bn, ds, dp);
        // past the block end the adjoint stops (beta is not stepped there)
        select8(L.t_of(j) < (int)L.K, nb, bb);
        L.at(Y.dsa, j) = dsr[jj] + ds;
        L.at(Y.dpr, j) = dpr[jj] + dp;
      }
    }
  };
  chunks_up<BetaRe<R>>(L.g, jend, load, body);
}

// Combine of one half: own window, next segment's alpha warm-up, previous
// segment's beta warm-up (in that order), added into the LLR adjoints and the
// previous half's extrinsic adjoint; returns this lane's term of the scale
// adjoint.  Read-modify-writes in chunks of 8 steps (t, or Pi(t), is a
// permutation, so no two steps share a position).
template <typename R>
struct CombIn {
  R ds[8], dp[8], dsn[8], dpn[8], dsp[8], dpp[8], gs[8], gq[8], pv[8], ep[8];
  uint32_t at[8];
};

template <typename R>
__device__ __forceinline__ R combine(const Lane<R> L0, const Layout Y, uint32_t h, R w, const R* eprev, R* g0,
                                     R* gpar, R* prev) {
  const Lane<R> L = fresh(L0);
  const bool lower = (h & 1u) != 0;
  const bool with_prev = h != 0;  // the first upper half has no a-priori input
  const uint32_t own_end = L.g + L.S;
  const bool has_next = L.q + 1 < L.M, has_prev = L.q > 0;
  const uint32_t ln = min(L.lane + 1, 63u), lp = L.lane > 0 ? L.lane - 1 : 0u;
  QppWalk qw(L.q * L.S, L.K, L.f1, L.f2);
  R wsum = 0;
  auto load = [&](uint32_t j0, CombIn<R>& x) {
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      const uint32_t j = j0 + jj;
      if (j < own_end) {
        const uint32_t t = (uint32_t)L.t_of(j);
        const uint32_t i = lower ? qw.pi : t;
        qw.next();
        x.at[jj] = i;
        x.ds[jj] = L.at(Y.dsa, j);
        x.dp[jj] = L.at(Y.dpr, j);
        if (j >= L.S) {  // t in [qS + S - g, qS + S): lane q+1's local j - S
          x.dsn[jj] = L.nb(Y.dsa, j - L.S, ln);
          x.dpn[jj] = L.nb(Y.dpr, j - L.S, ln);
        }
        if (j < 2 * L.g) {  // t in [qS, qS + g): lane q-1's local j + S
          x.dsp[jj] = L.nb(Y.dsa, j + L.S, lp);
          x.dpp[jj] = L.nb(Y.dpr, j + L.S, lp);
        }
        x.gs[jj] = g0[i];
        x.gq[jj] = gpar[t];
        if (with_prev) {
          x.pv[jj] = prev[i];
          x.ep[jj] = eprev[i];
        }
      }
    }
  };
  auto body = [&](uint32_t j0, const CombIn<R>& x) {
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      const uint32_t j = j0 + jj;
      if (j < own_end) {
        const uint32_t t = (uint32_t)L.t_of(j), i = x.at[jj];
        R ds = x.ds[jj], dp = x.dp[jj];
        if (j >= L.S) {
          ds = has_next ? ds + x.dsn[jj] : ds;
          dp = has_next ? dp + x.dpn[jj] : dp;
        }
        if (j < 2 * L.g) {
          ds = has_prev ? ds + x.dsp[jj] : ds;
          dp = has_prev ? dp + x.dpp[jj] : dp;
        }
        g0[i] = x.gs[jj] + ds;
        gpar[t] = x.gq[jj] + dp;
        if (with_prev) {
          prev[i] = x.pv[jj] + w * ds;
          wsum += ds * x.ep[jj];
        }
      }
    }
  };
  chunks_up<CombIn<R>>(L.g, own_end, load, body);
  return wsum;
}

// LDS path of the combine.  Fill: arrays 0 = grad d0, 1 = grad of the
// half's parity stream, 2 = the adjoint of e_{h-1} started as the per-lane
// path starts it (the posterior term for h - 1 == H - 2, else zero), 3 =
// e_{h-1}.  The lanes then add into the copies in the order of combine; the
// write-back stores arrays 0-2.  Two blocks per batch of loads.
template <typename R>
__device__ __forceinline__ void combine_fill(const DiffArgs<R> A, const DiffJob J, uint32_t h, const BlockLds<R> X) {
  const uint32_t H = A.halves, K = J.k;
  const bool lower = (h & 1u) != 0, with_prev = h != 0;
  const bool from_post = with_prev && h - 1 == H - 2;
  const R wp = A.scales[H - 2];
  const LdsPtr<R> xl = lds_view(X);
  for (uint32_t b0 = 0; b0 < J.nblk; b0 += 2) {
    const uint32_t nb = min(2u, J.nblk - b0);
    GlobalPtr<const R> g0[2], gq[2], gp[2], ep[2];
#pragma unroll
    for (uint32_t x = 0; x < 2; ++x) {
      const DiffCb C = A.cbs[J.cb_first + b0 + (x < nb ? x : 0u)];
      g0[x] = global_view<R>(A.grad_llr + C.llr_off);
      gq[x] = g0[x] + (size_t)(lower ? 2 : 1) * (K + 4);
      gp[x] = global_view(A.grad_post + C.post_off);
      ep[x] = global_view(A.saved + C.save_off + (size_t)(h ? h - 1 : 0) * K);
    }
    for (uint32_t i0 = threadIdx.x; i0 < K; i0 += 64 * kFillBatch) {
      R x0[2][kFillBatch], x1[2][kFillBatch], x2[2][kFillBatch], x3[2][kFillBatch];
#pragma unroll
      for (uint32_t x = 0; x < 2; ++x)
#pragma unroll
        for (uint32_t u = 0; u < kFillBatch; ++u) {
          const uint32_t i = i0 + 64 * u;
          if (x < nb && i < K) {
            x0[x][u] = g0[x][i];
            x1[x][u] = gq[x][i];
            x2[x][u] = from_post ? gp[x][i] : R(0);
            x3[x][u] = with_prev ? ep[x][i] : R(0);
          }
        }
#pragma unroll
      for (uint32_t x = 0; x < 2; ++x)
#pragma unroll
        for (uint32_t u = 0; u < kFillBatch; ++u) {
          const uint32_t i = i0 + 64 * u;
          if (x < nb && i < K) {
            xl[X.idx(0, b0 + x, i)] = x0[x][u];
            xl[X.idx(1, b0 + x, i)] = x1[x][u];
            xl[X.idx(2, b0 + x, i)] = from_post ? wp * x2[x][u] : R(0);
            xl[X.idx(3, b0 + x, i)] = x3[x][u];
          }
        }
    }
  }
}

template <typename R>
__device__ __forceinline__ void combine_store(const DiffArgs<R> A, const DiffJob J, uint32_t h, const BlockLds<R> X) {
  const uint32_t K = J.k;
  const bool lower = (h & 1u) != 0, with_prev = h != 0;
  const LdsPtr<R> xl = lds_view(X);
  for (uint32_t b = 0; b < J.nblk; ++b) {
    const DiffCb C = A.cbs[J.cb_first + b];
    const GlobalPtr<R> g0 = global_view(A.grad_llr + C.llr_off);
    const GlobalPtr<R> gq = g0 + (size_t)(lower ? 2 : 1) * (K + 4);
    const GlobalPtr<R> prev = global_view(A.ebar + C.ebar_off + (size_t)((h + 1) & 1u) * K);
    for (uint32_t i = threadIdx.x; i < K; i += 64) {
      g0[i] = xl[X.idx(0, b, i)];
      gq[i] = xl[X.idx(1, b, i)];
      if (with_prev) prev[i] = xl[X.idx(2, b, i)];
    }
  }
}

// Per-lane part of the LDS combine (the same sums, in the same order, as
// combine), with the next chunk's scratch loads in flight while a chunk is
// added into the LDS copies.
template <typename R>
struct CombLdsIn {
  R ds[8], dp[8], dsn[8], dpn[8], dsp[8], dpp[8];
};

template <typename R>
__device__ __forceinline__ R combine_lds(const Lane<R>& L0, const Layout& Y, uint32_t h, R w, const BlockLds<R>& X,
                                         uint32_t blk) {
  const Lane<R> L = fresh(L0);
  const bool lower = (h & 1u) != 0;
  const bool with_prev = h != 0;
  const uint32_t own_end = L.g + L.S;
  const bool has_next = L.q + 1 < L.M, has_prev = L.q > 0;
  const uint32_t ln = min(L.lane + 1, 63u), lp = L.lane > 0 ? L.lane - 1 : 0u;
  QppWalk qw(L.q * L.S, L.K, L.f1, L.f2);
  R wsum = 0;
  auto load = [&](uint32_t j0, CombLdsIn<R>& x) {
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      const uint32_t j = j0 + jj;
      if (j < own_end) {
        x.ds[jj] = L.at(Y.dsa, j);
        x.dp[jj] = L.at(Y.dpr, j);
        if (j >= L.S) {  // t in [qS + S - g, qS + S): lane q+1's local j - S
          x.dsn[jj] = L.nb(Y.dsa, j - L.S, ln);
          x.dpn[jj] = L.nb(Y.dpr, j - L.S, ln);
        }
        if (j < 2 * L.g) {  // t in [qS, qS + g): lane q-1's local j + S
          x.dsp[jj] = L.nb(Y.dsa, j + L.S, lp);
          x.dpp[jj] = L.nb(Y.dpr, j + L.S, lp);
        }
      }
    }
  };
  auto body = [&](uint32_t j0, const CombLdsIn<R>& x) {
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      const uint32_t j = j0 + jj;
      if (j < own_end) {
        const uint32_t t = (uint32_t)L.t_of(j);
        const uint32_t i = lower ? qw.pi : t;
        qw.next();
        R a = x.ds[jj], c = x.dp[jj];
        if (j >= L.S) {
          a = has_next ? a + x.dsn[jj] : a;
          c = has_next ? c + x.dpn[jj] : c;
        }
        if (j < 2 * L.g) {
          a = has_prev ? a + x.dsp[jj] : a;
          c = has_prev ? c + x.dpp[jj] : c;
        }
        X.at(0, blk, i) = X.at(0, blk, i) + a;
        X.at(1, blk, t) = X.at(1, blk, t) + c;
        if (with_prev) {
          X.at(2, blk, i) = X.at(2, blk, i) + w * a;
          wsum += a * X.at(3, blk, i);
        }
      }
    }
  };
  chunks_up<CombLdsIn<R>>(L.g, own_end, load, body);
  return wsum;
}

}  // namespace

template <typename R>
__global__ __launch_bounds__(64) void diff_backward_kernel(DiffArgs<R> A) {
  extern __shared__ __attribute__((aligned(16))) char diff_lds[];
  const Layout Y(A.max_seg, A.max_warm);
  const uint32_t H = A.halves;
  for (uint32_t jb = blockIdx.x; jb < A.njobs; jb += gridDim.x) {
    const DiffJob J = A.jobs[jb];
    const Lane<R> L0 = make_lane(J, A);
    const uint32_t blk = L0.active ? L0.lane / J.nseg : 0;
    const BlockLds<R> X = block_lds(reinterpret_cast<R*>(diff_lds), J);
    const Lane<R> L = A.use_lds ? with_lds(L0, X, blk) : L0;
    const DiffCb C = A.cbs[J.cb_first + blk];
    const uint32_t K = L.K, D = K + 4;
    const R* d0 = A.llr + C.llr_off;
    const R* d1 = d0 + D;
    const R* d2 = d1 + D;
    const R* sv = A.saved + C.save_off;
    const R* gp = A.grad_post + C.post_off;
    R* g0 = A.grad_llr + C.llr_off;
    R* g1 = g0 + D;
    R* g2 = g1 + D;
    R* eb[2] = {A.ebar + C.ebar_off, A.ebar + C.ebar_off + K};
    R* dsc = A.dscale_part + (size_t)jb * H;
    const uint32_t lo = L.q * L.S, hi = lo + L.S;
    {
      R s1 = 0, s2 = 0;
      if (L.active) post_adjoint(A, L, sv, gp, g0, g1, g2, eb[(H - 1) & 1u], s1, s2);
      for (int o = 32; o > 0; o >>= 1) {
        s1 += __shfl_xor(s1, o, 64);
        s2 += __shfl_xor(s2, o, 64);
      }
      if (threadIdx.x == 0) {
        for (uint32_t h = 0; h < H; ++h) dsc[h] = 0;
        dsc[H - 2] = s1;
        dsc[H - 1] = s2;
      }
    }
    __syncthreads();
    for (uint32_t h = H; h-- > 0;) {
      R* cur = eb[h & 1u];
      R* prev = eb[(h + 1) & 1u];
      const bool lower = (h & 1u) != 0;
      R ts[3], tp[3];
      if (A.use_lds) {
        stage_fill(A, J, h, X);
        __syncthreads();
        stage_gather(J, h, X);
        if (L.active) tail_inputs(h, d0, d1, d2, K, ts, tp);
        __syncthreads();
      } else if (L.active) {
        // adjoint of e_{h-1}: starts at the posterior term for h - 1 == H - 2
        // (the LDS path starts it in its combine fill)
        if (h >= 1) {
          if (h - 1 == H - 2) {
            const R wp = A.scales[H - 2];
            for (uint32_t t0 = lo; t0 < hi; t0 += 8) {
              R gv[8];
#pragma unroll
              for (uint32_t jj = 0; jj < 8; ++jj)
                if (t0 + jj < hi) gv[jj] = gp[t0 + jj];
#pragma unroll
              for (uint32_t jj = 0; jj < 8; ++jj)
                if (t0 + jj < hi) prev[t0 + jj] = wp * gv[jj];
            }
          } else {
            for (uint32_t t = lo; t < hi; ++t) prev[t] = 0;
          }
        }
        tail_inputs(h, d0, d1, d2, K, ts, tp);
        stage(L, Y, h, d0, d1, d2, sv, A.scales);
      }
      if (L.active) {
        // the alpha-adjoint sweep is the first writer of local steps below
        // g + S, the beta-adjoint sweep of those up to 2g + S; zeros beyond
        // (a plan's longer windows) keep the combine's reads defined
        for (uint32_t j = 2 * L.g + L.S; j < Y.P; ++j) L.at(Y.dsa, j) = L.at(Y.dpr, j) = 0;
        // beta over the window's warm-up above the own steps (local 2g + S - 1
        // down to g + S), run beside the alpha checkpoint pass (alpha at every
        // 8th local step only, al[c] = alpha at local 8c; the adjoint sweeps
        // recompute each sub-window's metrics into registers, which keeps the
        // wave's scratch small enough to stay in the caches).  The own steps'
        // beta runs inside the alpha-adjoint sweep, an independent chain
        // beside the adjoint recursion.  Checkpoints for the beta-adjoint
        // sweep: beta at local g + i for i a multiple of 8 or i = S + g, at
        // be[(i + 7) / 8]
        R bcur[8];
        if (lo + L.S + L.g >= K) {
          tail_beta(ts, tp, bcur);
        } else {
#pragma unroll
          for (int s = 0; s < 8; ++s) bcur[s] = 0;
        }
        const uint32_t bst = Y.S + Y.g + 1;
        auto beta_ck = [&](uint32_t i, const R (&b)[8]) {
          if ((i & 7u) == 0 || i == L.S + L.g) {
#pragma unroll
            for (int s = 0; s < 8; ++s) L.at(Y.be + s * bst, (i + 7) / 8) = b[s];
          }
        };
        alpha_ckpt_beta_warm_pass(L, Y, bcur, beta_ck);
        alpha_adjoint_sweep(L, Y, bcur, beta_ck, lower, cur);
        R bb[8];
        beta_adjoint_sweep(L, Y, bb);
        // beta of the window top: the tail beta when it reaches K
        R dts[3] = {0, 0, 0}, dtp[3] = {0, 0, 0};
        if (lo + L.S + L.g >= K) {
#pragma unroll
          for (unsigned s = 0; s < 8; ++s) {
            unsigned c = s;
#pragma unroll
            for (int k3 = 0; k3 < 3; ++k3) {
              const unsigned u = term_in(c);
              dts[k3] += u ? R(0.5) * bb[s] : R(-0.5) * bb[s];
              dtp[k3] += par(c, u) ? R(0.5) * bb[s] : R(-0.5) * bb[s];
              c = nxt(c, u);
            }
          }
        }
        for (int k3 = 0; k3 < 3; ++k3) {
          L.at(Y.tl, k3) = dts[k3];
          L.at(Y.tl, 3 + k3) = dtp[k3];
        }
      }
      __syncthreads();
      R wsum = 0;
      const R w = h ? A.scales[h - 1] : R(0);
      if (A.use_lds) {
        combine_fill(A, J, h, X);
        __syncthreads();
        if (L.active) wsum = combine_lds(L, Y, h, w, X, blk);
        __syncthreads();
        combine_store(A, J, h, X);
      } else if (L.active) {
        const R* eprev = sv + (size_t)(h ? h - 1 : 0) * K;
        wsum = combine(L, Y, h, w, eprev, g0, lower ? g2 : g1, prev);
      }
      if (L.active) {
        if (L.q == L.M - 1) {
          // tail LLR adjoints, in segment order
          R dts[3] = {0, 0, 0}, dtp[3] = {0, 0, 0};
          for (uint32_t m = 0; m < L.M; ++m) {
            const uint32_t other = L.lane - (L.M - 1) + m;
            if (m * L.S + L.S + L.g < K) continue;
            for (int k3 = 0; k3 < 3; ++k3) {
              dts[k3] += L.nb(Y.tl, k3, other);
              dtp[k3] += L.nb(Y.tl, 3 + k3, other);
            }
          }
          if (!lower) {
            g0[K] += dts[0];
            g0[K + 2] += dts[1];
            g1[K] += dtp[0];
            g0[K + 1] += dtp[1];
            g1[K + 2] += dtp[2];
          } else {
            g1[K + 3] += dts[0];
            g2[K] += dts[1];
            g2[K + 2] += dts[2];
            g2[K + 1] += dtp[0];
            g0[K + 3] += dtp[1];
            g2[K + 3] += dtp[2];
          }
        }
      }
      for (int o = 32; o > 0; o >>= 1) wsum += __shfl_xor(wsum, o, 64);
      if (threadIdx.x == 0 && h) dsc[h - 1] += wsum;
      __syncthreads();
    }
  }
}

// dscales[h] = sum over jobs of the per-job partial sums: one workgroup per
// h, thread t sums jobs t, t + 256, ... in order, then a fixed tree over
// the threads (deterministic).
template <typename R>
__global__ __launch_bounds__(256) void diff_scale_reduce_kernel(const R* part, uint32_t njobs, uint32_t halves,
                                                                R* out) {
  __shared__ R red[256];
  const uint32_t h = blockIdx.x;
  R s = 0;
  for (uint32_t j = threadIdx.x; j < njobs; j += 256) s += part[(size_t)j * halves + h];
  red[threadIdx.x] = s;
  __syncthreads();
  for (uint32_t w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[h] = red[0];
}

template __global__ void diff_forward_kernel<float>(DiffArgs<float>);
template __global__ void diff_forward_kernel<double>(DiffArgs<double>);
template __global__ void diff_backward_kernel<float>(DiffArgs<float>);
template __global__ void diff_backward_kernel<double>(DiffArgs<double>);
template __global__ void diff_scale_reduce_kernel<float>(const float*, uint32_t, uint32_t, float*);
template __global__ void diff_scale_reduce_kernel<double>(const double*, uint32_t, uint32_t, double*);

}  // namespace tcdev

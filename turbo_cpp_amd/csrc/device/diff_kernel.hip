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
// Parallel layout: a job is nblk code blocks of one K run by a workgroup of
// two waves; each block is cut into M = K / S segments and lane =
// block * M + segment in both waves.  Segment q owns steps [qS, qS + S): its
// alpha starts from zero metrics g steps early (from state 0 when that
// reaches step 0) and its beta from zero metrics g steps late (from the tail
// beta when that reaches K).  A segment therefore depends only on the
// half-iteration's inputs in [qS - g, qS + S + g), so each half-iteration is
// a function of the previous one's extrinsics and the adjoint is local too.
// Each segment has an alpha side and a beta side, one per wave; per half:
//   1. (forward) wave 0 runs alpha over the segment's window, wave 1 beta,
//      each keeping a checkpoint every 8 steps in the saved state, which the
//      adjoint reads back instead of running this phase again;
//   2. (forward) the two waves split the own steps: from the checkpoints
//      each recomputes alpha_{t+1} and beta_{t+1} of its steps and forms the
//      LLR from the eight states at t + 1 (the input bit of a transition is
//      fixed by the state it enters);
//   3. (adjoint) wave 1 runs the alpha adjoint backwards in time (weights of
//      each state's two incoming transitions), wave 0 the beta adjoint
//      forwards in time (weights of each state's two outgoing transitions),
//      both seeded at the own steps with G(t, n) = dloss/d(alpha_{t+1}(n) +
//      beta_{t+1}(n)), the LLR's share of both metrics' adjoints, which each
//      wave forms from metrics it recomputes; each writes its own sums of the
//      branch-metric adjoints on sys + la and par;
//   4. (adjoint) the two waves split the own steps again and add the sums
//      into the LLR adjoints in a fixed order (own alpha side + own beta
//      side, next segment's alpha warm-up, previous segment's beta warm-up),
//      so the result is deterministic.
// Renormalisation does not change any output (the LLRs are invariant to
// adding a constant to all eight metrics of a step), so it has no adjoint.
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
  // The walk's state as the optimiser's barrier: steps after it cannot be
  // computed before it, so a gather walked step by step computes each
  // position where it loads (otherwise a chunk's positions are computed
  // ahead and parked in spill slots, and each reload waits for every load
  // before it).
  __device__ void pin() { asm volatile("" : "+v"(pi), "+v"(d)); }
  // a step back, Pi(i-1) and its difference, when c, else nothing (no
  // branch)
  __device__ void prev_if(bool c) {
    const uint32_t dd = d >= d2 ? d - d2 : d + k - d2;
    const uint32_t p = pi >= dd ? pi - dd : pi + k - dd;
    pi = c ? p : pi;
    d = c ? dd : d;
  }
};

// Per-lane scratch fields, each [j][lane] with j < its length, and the
// fields of a lane's checkpoints of one half (alpha: cna per state, beta: cnb
// per state; kept by the forward for the adjoint).
struct Layout {
  uint32_t P, S, g;
  uint32_t sa, pr, dsa, dpr, dsb, dpb;
  uint32_t cna, cnb, cal, cbe;
  __device__ Layout(uint32_t s, uint32_t gw) : P(s + 2 * gw), S(s), g(gw) {
    cna = diff_ck_alpha(s, gw);
    cnb = diff_ck_beta(s, gw);
    cal = 0;
    cbe = 8 * cna;
    sa = 0;
    pr = sa + P;
    dsa = pr + P;                // alpha side: sys + la and par adjoints
    dpr = dsa + P;
    dsb = dpr + P;               // beta side
    dpb = dsb + P;
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
  R* sc;  // the job's scratch (uniform), element (field + j) * 64 + lane
  R* ck;  // the job's checkpoints of the current half, the same order
  // staged inputs: the block's LDS pairs (lds; local step j is position
  // p = qS + j, pair at p + p / 32) or the per-lane sa / pr fields
  bool lds;
  LdsPtr<R> xs;
  // 32-bit byte offsets from the job's base (a slot is far below 4 GiB):
  // one 32-bit offset per field and the steps as immediate offsets, not a
  // 64-bit address per access
  __device__ R& nb(uint32_t field, uint32_t j, uint32_t other) const {
    return *reinterpret_cast<R*>(reinterpret_cast<char*>(sc) + (uint32_t)(((field + j) * 64u + other) * sizeof(R)));
  }
  __device__ R& at(uint32_t field, uint32_t j) const { return nb(field, j, lane); }
  __device__ R& ckat(uint32_t field, uint32_t j) const {
    return *reinterpret_cast<R*>(reinterpret_cast<char*>(ck) + (uint32_t)(((field + j) * 64u + lane) * sizeof(R)));
  }
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

// Inlining: everything is inlined into the kernels, so no call saves and
// restores registers through private memory.
//
// Memory order and control flow of every sequential pass below.  Each pass
// walks its steps in chunks of at most 8 with two input buffers used in turn:
// the loads of the next chunk go out before a chunk is computed and stored,
// and no register copy sits between a load and its use (a copy would wait for
// the load at the copy).  The vector-memory counter retires loads and stores
// in issue order, so a load issued behind a store would also wait for that
// store; issued ahead, it waits only for stores of earlier chunks.  Loop
// bounds are wave-uniform (all lanes of a job share K, S and g); steps that
// are outside the block for some lanes (t < 0 in the first segment, t >= K
// in the last) are computed by every lane and their results selected away,
// so no per-lane branch splits a chunk.  Values computed that way are only
// ever written to positions that no later pass reads.
//
// Chunks.  Alpha runs over local steps [0, g + S) in chunks whose boundaries
// are multiples of 8 counted from the first own step (local g): chunk k
// covers [max(0, 8k - off), min(8k + 8 - off, g + S)) with off =
// (8 - g mod 8) mod 8.  Beta runs over [g, 2g + S) in chunks
// [g + 8m, min(g + 8m + 8, 2g + S)).  Over the own steps the two grids
// coincide: own chunk m is alpha chunk (g + off) / 8 + m and beta chunk m.
// Metrics are renormalised at the chunk boundaries only (alpha when it
// reaches local g + 8i, beta when it reaches local g + 8i), so a recompute
// from a checkpoint reproduces the first pass's values bit for bit.
// Checkpoints: al[k] = alpha at the start of alpha chunk k; be[m] = beta at
// local g + 8m for 0 < m < nb and be[nb] = beta at the window top 2g + S, so
// beta chunk m is recomputed downwards from be[m + 1].
// Lane src's value of v (a wave-wide read: every lane of the wave takes
// part); float and double.
__device__ __forceinline__ float shfl_r(float v, int src) { return __shfl(v, src, 64); }
__device__ __forceinline__ double shfl_r(double v, int src) { return __shfl(v, src, 64); }

// t_of(j) of a chunk's first step as a value the optimiser can neither hoist
// nor share with other uses of t_of: the base of the chunk's contiguous row of
// loads (immediate offsets from one address, where per-step positions kept
// live across the loop would take registers and come back from spill slots,
// each reload waiting for every load before it).
template <typename R>
__device__ __forceinline__ uint32_t row_t(const Lane<R>& L, uint32_t j) {
  uint32_t t = (uint32_t)L.t_of(j);
  asm volatile("" : "+v"(t));
  return t;
}

struct Chunks {
  uint32_t off, na, nb, nown;
  __device__ Chunks(uint32_t S, uint32_t g)
      : off((8u - (g & 7u)) & 7u), na((g + S + ((8u - (g & 7u)) & 7u) + 7u) >> 3), nb((g + S + 7u) >> 3),
        nown((S + 7u) >> 3) {}
  __device__ uint32_t a_lo(uint32_t k) const { return 8u * k > off ? 8u * k - off : 0u; }
  __device__ uint32_t a_hi(uint32_t k, uint32_t end) const { return min(8u * k + 8u - off, end); }
  __device__ uint32_t own_alpha(uint32_t m, uint32_t g) const { return (g + off) / 8u + m; }
};
__device__ __forceinline__ bool alpha_nrm(uint32_t j, uint32_t off) { return ((j + 1u + off) & 7u) == 0; }
__device__ __forceinline__ bool beta_nrm(uint32_t j, uint32_t g) { return ((j - g) & 7u) == 0; }

// body(k, buf) for k = lo .. hi - 1 (up) or hi - 1 .. lo (down), with
// load(k, buf) one chunk ahead, alternating two buffers.
template <class Buf, class Load, class Body>
__device__ __forceinline__ void loop_up(uint32_t lo, uint32_t hi, Load&& load, Body&& body) {
  if (lo >= hi) return;
  Buf b0, b1;
  load(lo, b0);
  for (uint32_t k = lo; k < hi; k += 2) {
    if (k + 1 < hi) load(k + 1, b1);
    body(k, b0);
    if (k + 1 >= hi) break;
    if (k + 2 < hi) load(k + 2, b0);
    body(k + 1, b1);
  }
}
template <class Buf, class Load, class Body>
__device__ __forceinline__ void loop_down(uint32_t lo, uint32_t hi, Load&& load, Body&& body) {
  if (lo >= hi) return;
  Buf b0, b1;
  load(hi - 1, b0);
  for (int k = (int)hi - 1; k >= (int)lo; k -= 2) {
    if (k - 1 >= (int)lo) load((uint32_t)k - 1, b1);
    body((uint32_t)k, b0);
    if (k - 1 < (int)lo) break;
    if (k - 2 >= (int)lo) load((uint32_t)k - 2, b0);
    body((uint32_t)k - 1, b1);
  }
}

// Runs body(j0, buf) for j0 = lo, lo + C, ... < hi with load(j0, buf) one
// chunk ahead, alternating two buffers.
template <class Buf, uint32_t C = 8, class Load, class Body>
__device__ __forceinline__ void chunks_up(uint32_t lo, uint32_t hi, Load&& load, Body&& body) {
  if (lo >= hi) return;
  Buf b0, b1;
  load(lo, b0);
  for (uint32_t j0 = lo; j0 < hi; j0 += 2 * C) {
    if (j0 + C < hi) load(j0 + C, b1);
    body(j0, b0);
    if (j0 + C >= hi) break;
    if (j0 + 2 * C < hi) load(j0 + 2 * C, b0);
    body(j0 + C, b1);
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
// Entries jj >= n repeat entry n - 1: every entry loads (one wait for the
// batch, where a branch per entry would wait at each) and no later code reads
// them.
template <typename R, bool DOWN = false>
__device__ __forceinline__ void load_in8(const Lane<R>& L, const Layout& Y, int j0, int n, In8<R>& c) {
  if (L.lds) {
    const uint32_t pb = L.q * L.S;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const int jc = jj < n ? jj : (n > 0 ? n - 1 : 0);
      const uint32_t p = pb + (uint32_t)(DOWN ? j0 - jc : j0 + jc);
      const uint32_t i = 2 * (p + (p >> 5));
      c.sa[jj] = L.xs[i];
      c.pr[jj] = L.xs[i + 1];
    }
  } else {
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const int jc = jj < n ? jj : (n > 0 ? n - 1 : 0);
      const uint32_t j = (uint32_t)(DOWN ? j0 - jc : j0 + jc);
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

// Loads in flight per thread, block and array in the LDS fills (a workgroup
// of 128 threads covers 1024 positions per batch).
constexpr uint32_t kFillBatch = 8;

// Fill of the staging arrays: 0 = sys + la at natural position i (the lower
// half reads it at Pi(t)), 1 = the half's parity stream.  Same arithmetic as
// the per-lane staging.  Two blocks per batch of loads, so a job of two
// blocks waits for memory once.
template <typename R>
__device__ __forceinline__ void stage_fill(const DiffArgs<R> A, const DiffJob J, uint32_t h, const BlockLds<R> X) {
  const R w = h ? A.scales[h - 1] : R(0);
  const bool lower = (h & 1u) != 0;
  const uint32_t K = J.k, nt = blockDim.x;
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
    for (uint32_t i0 = threadIdx.x; i0 < K; i0 += nt * kFillBatch) {
      R sv[2][kFillBatch], ev[2][kFillBatch], pv[2][kFillBatch];
#pragma unroll
      for (uint32_t x = 0; x < 2; ++x)
#pragma unroll
        for (uint32_t u = 0; u < kFillBatch; ++u) {
          const uint32_t i = i0 + nt * u;
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
          const uint32_t i = i0 + nt * u;
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
// half's order, zeros outside [0, K).  Each thread takes a run of consecutive
// positions (the lower half's interleaver walked by its recurrence).
template <typename R>
__device__ __forceinline__ void stage_gather(const DiffJob J, uint32_t h, const BlockLds<R> X) {
  const LdsPtr<R> xl = lds_view(X);
  const bool lower = (h & 1u) != 0;
  const uint32_t K = J.k, g = J.warm, n = K + 2 * g;
  const uint32_t run = (n + blockDim.x - 1) / blockDim.x, p0 = threadIdx.x * run;
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

template <typename R>
__device__ __forceinline__ void ckload8(const Lane<R>& L, uint32_t field, uint32_t stride, uint32_t j, R (&v)[8]) {
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] = L.ckat(field + s * stride, j);
}

template <typename R>
__device__ __forceinline__ void load8(const Lane<R>& L, uint32_t field, uint32_t stride, uint32_t j, R (&v)[8]) {
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] = L.at(field + s * stride, j);
}

// Input bit of the transitions into state n (n = (fb, s2, s1) with
// fb = u ^ s1 ^ s2): the parity of n's bits.
__host__ __device__ constexpr unsigned in_of(unsigned n) { return (n ^ (n >> 1) ^ (n >> 2)) & 1u; }

// Per input-bit class of the eight states at t + 1: m[u] = max T(n),
// s[u] = sum 2^(T(n) - m[u]), p[n] = 2^(T(n) - m[u(n)]).  L_u = m[u] +
// log2 s[u] is the base-2 LLR term of bit u: every transition into a state of
// class u carries input u, so summing alpha_t + gamma + beta_{t+1} over the
// 16 transitions equals summing alpha_{t+1} + beta_{t+1} over the 8 states.
template <typename R>
__device__ __forceinline__ void class_sums(const R (&T)[8], R (&m)[2], R (&s)[2], R (&p)[8]) {
  m[0] = m[1] = R(kNegF);
#pragma unroll
  for (unsigned n = 0; n < 8; ++n) m[in_of(n)] = fmax(m[in_of(n)], T[n]);
  s[0] = s[1] = R(0);
#pragma unroll
  for (unsigned n = 0; n < 8; ++n) {
    p[n] = Num<R>::ex2(T[n] - m[in_of(n)]);
    s[in_of(n)] += p[n];
  }
}

// The weight of a in jac(a, b): 2^a / (2^a + 2^b) (base-2 metrics).
template <typename R>
__device__ inline R jac_w0(R a, R b) {
  const R e = Num<R>::ex2(-fabs(a - b));
  const R big = Num<R>::rcp(R(1) + e);
  return a >= b ? big : e * big;
}

// The same weight from the Jacobian logarithm of the pair already at hand
// (j = jac(a, b), e.g. a recomputed metric before its renormalisation):
// 2^(a - j), one exp2 instead of an exp2 and a reciprocal; at most 1.
template <typename R>
__device__ inline R jac_wj(R a, R j) {
  return fmin(Num<R>::ex2(a - j), R(1));
}

// One step of the alpha adjoint at local j (alpha_j = a, half gammas hs, hp
// in base 2): ab holds the adjoint of alpha_{j+1} on entry (the LLR's share
// included) and of alpha_j on exit; ds, dp: the step's adjoints of sys + la
// and par through its transitions (gamma = +-(sys + la)/2 +- par/2).  With
// NEXT, an holds alpha_{j+1} before any renormalisation (each state's jac of
// its two candidates) and the weights come from it (jac_wj); else from the
// two candidates (jac_w0).
template <typename R, bool NEXT = false>
__device__ inline void alpha_adjoint_step(const R (&a)[8], const R (&an)[8], R hs, R hp, R (&ab)[8], R& ds_out,
                                          R& dp_out) {
  R abn[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  R ds = 0, dp = 0;
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
        const R c0 = a[s0] + gam(u0, par(s0, u0), hs, hp);
        const R w0 = NEXT ? jac_wj(c0, an[n]) : jac_w0(c0, a[s] + gam(u, par(s, u), hs, hp));
        const R d0 = ab[n] * w0, d1 = ab[n] - d0;  // the two weights add to 1
        abn[s0] += d0;
        abn[s] += d1;
        ds += u0 ? d0 : -d0;
        ds += u ? d1 : -d1;
        dp += par(s0, u0) ? d0 : -d0;
        dp += par(s, u) ? d1 : -d1;
      }
    }
#pragma unroll
  for (int s = 0; s < 8; ++s) ab[s] = abn[s];
  ds_out = R(0.5) * ds;
  dp_out = R(0.5) * dp;
}

// One step of the beta adjoint at local j (beta_{j+1} = b): bb is the
// adjoint of beta_j; bn holds the output adjoints of beta_{j+1} (the LLR's
// share) on entry and the whole adjoint of beta_{j+1} on exit.  With NEXT,
// bj holds beta_j before any renormalisation and the weights come from it.
template <typename R, bool NEXT = false>
__device__ inline void beta_adjoint_step(const R (&b)[8], const R (&bj)[8], R hs, R hp, const R (&bb)[8],
                                         R (&bn)[8], R& ds_out, R& dp_out) {
  R ds = 0, dp = 0;
#pragma unroll
  for (unsigned s = 0; s < 8; ++s) {
    const R c0 = b[nxt(s, 0)] + gam(0u, par(s, 0), hs, hp);
    const R w0 = NEXT ? jac_wj(c0, bj[s]) : jac_w0(c0, b[nxt(s, 1)] + gam(1u, par(s, 1), hs, hp));
    const R d0g = bb[s] * w0, d1g = bb[s] - d0g;
    bn[nxt(s, 0)] += d0g;
    bn[nxt(s, 1)] += d1g;
    ds += d1g - d0g;
    dp += par(s, 0) ? d0g : -d0g;
    dp += par(s, 1) ? d1g : -d1g;
  }
  ds_out = R(0.5) * ds;
  dp_out = R(0.5) * dp;
}

// Phase 1, alpha side: alpha over local steps [0, g + S) from its start
// (state 0 when the window reaches the block start, else zero metrics; alpha
// stays at its start value before step 0 of the block, t < 0), al[k] kept at
// the start of each chunk.
template <typename R>
__device__ __forceinline__ void alpha_ckpt_pass(const Lane<R>& L0, const Layout& Y, const Chunks& C) {
  const Lane<R> L = fresh(L0);
  R a[8];
  const bool from_zero_state = (int)(L.q * L.S) - (int)L.g <= 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) a[s] = from_zero_state ? (s == 0 ? R(0) : R(kNegF)) : R(0);
  const uint32_t end = L.g + L.S;
  auto load = [&](uint32_t k, In8<R>& c) {
    const uint32_t lo = C.a_lo(k);
    load_in8(L, Y, (int)lo, (int)(C.a_hi(k, end) - lo), c);
  };
  auto body = [&](uint32_t k, const In8<R>& c) {
    const uint32_t lo = C.a_lo(k), n = C.a_hi(k, end) - lo;
#pragma unroll
    for (int s = 0; s < 8; ++s) L.ckat(Y.cal + s * Y.cna, k) = a[s];
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj)
      if (jj < n) {
        const uint32_t j = lo + jj;
        R o[8];
        alpha_step(a, kHalfL2e<R> * c.sa[jj], kHalfL2e<R> * c.pr[jj], o, alpha_nrm(j, C.off));
        select8(L.t_of(j) >= 0, o, a);
      }
  };
  loop_up<In8<R>>(0, C.na, load, body);
}

// Phase 1, beta side: beta from the window top (b: beta at local 2g + S)
// down to local g; beta stays at its start value past the block end
// (t >= K).  be[m + 1] is kept at the top of each chunk m.
template <typename R>
__device__ __forceinline__ void beta_ckpt_pass(const Lane<R>& L0, const Layout& Y, const Chunks& C, R (&b)[8]) {
  const Lane<R> L = fresh(L0);
  const uint32_t top = 2 * L.g + L.S;
  auto load = [&](uint32_t m, In8<R>& c) {
    const uint32_t lo = L.g + 8u * m, hi = min(lo + 8u, top);
    load_in8(L, Y, (int)lo, (int)(hi - lo), c);
  };
  auto body = [&](uint32_t m, const In8<R>& c) {
    const uint32_t lo = L.g + 8u * m, n = min(lo + 8u, top) - lo;
#pragma unroll
    for (int s = 0; s < 8; ++s) L.ckat(Y.cbe + s * Y.cnb, m + 1) = b[s];
#pragma unroll
    for (int jj = 7; jj >= 0; --jj)
      if ((uint32_t)jj < n) {
        const uint32_t j = lo + (uint32_t)jj;
        R o[8];
        beta_step(b, kHalfL2e<R> * c.sa[jj], kHalfL2e<R> * c.pr[jj], o, beta_nrm(j, L.g));
        select8(L.t_of(j) < (int)L.K, o, b);
      }
  };
  loop_down<In8<R>>(0, C.nb, load, body);
}

// Phase 2: for own chunks m in [m0, m1), alpha_{j+1} and beta_{j+1} of each
// own step j recomputed from the checkpoints (alpha up from al, beta down
// from be[m + 1]), then term(j, T, sys + la of j) with T(n) =
// alpha_{j+1}(n) + beta_{j+1}(n), in ascending order of j.
template <typename R>
struct OwnIn {
  R a0[8], b0[8];
};

template <typename R, class Term>
__device__ __forceinline__ void own_terms(const Lane<R>& L0, const Layout& Y, const Chunks& C, uint32_t m0, uint32_t m1,
                                          Term&& term) {
  const Lane<R> L = fresh(L0);
  const uint32_t own_end = L.g + L.S, top = 2 * L.g + L.S;
  auto load = [&](uint32_t m, OwnIn<R>& x) {
    const uint32_t lo = L.g + 8u * m;
    (void)lo;
    ckload8(L, Y.cal, Y.cna, C.own_alpha(m, L.g), x.a0);
    ckload8(L, Y.cbe, Y.cnb, m + 1, x.b0);
  };
  auto body = [&](uint32_t m, const OwnIn<R>& xc) {
    const uint32_t lo = L.g + 8u * m, n = min(8u, own_end - lo), nb = min(lo + 8u, top) - lo;
    // staged inputs of steps lo .. lo + nb - 1 (LDS, or the lane's scratch)
    In8<R> in;
    load_in8(L, Y, (int)lo, (int)nb, in);

    // beta of the chunk into registers first (downwards), then alpha stepped
    // up with each step's terms formed as soon as alpha_{j+1} is known
    R breg[8][8];
    {
      R b[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) b[s] = xc.b0[s];
#pragma unroll
      for (int jj = 7; jj >= 0; --jj)
        if ((uint32_t)jj < nb) {
          if ((uint32_t)jj < n) {
#pragma unroll
            for (int s = 0; s < 8; ++s) breg[jj][s] = b[s];  // beta at local lo + jj + 1
          }
          if (jj > 0) {
            const uint32_t j = lo + (uint32_t)jj;
            R o[8];
            beta_step(b, kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], o, beta_nrm(j, L.g));
            select8(L.t_of(j) < (int)L.K, o, b);
          }
        }
    }
    R a[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) a[s] = xc.a0[s];
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj)
      if (jj < n) {
        R o[8];
        alpha_step(a, kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], o, alpha_nrm(lo + jj, C.off));
        R T[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          a[s] = o[s];  // own steps lie inside the block
          T[s] = o[s] + breg[jj][s];
        }
        term(lo + jj, T, in.sa[jj]);
      }
  };
  loop_up<OwnIn<R>>(m0, m1, load, body);
}

// G(n) of one own step, the LLR's share of the adjoints of alpha_{j+1}(n) and
// beta_{j+1}(n): from T(n) = alpha_{j+1}(n) + beta_{j+1}(n) and the adjoint e
// of the step's extrinsic, e x (+-1 by n's input bit) x the weight of n in
// its class's LLR term.
template <typename R>
__device__ __forceinline__ void llr_adjoint(const R (&T)[8], R e, R (&G)[8]) {
  R m[2], s[2], p[8];
  class_sums(T, m, s, p);
  const R f1 = e * Num<R>::rcp(s[1]), f0 = -e * Num<R>::rcp(s[0]);
#pragma unroll
  for (unsigned n = 0; n < 8; ++n) G[n] = (in_of(n) ? f1 : f0) * p[n];
}

// Phase 3, alpha side: the alpha adjoint backwards over local steps
// g + S - 1 .. 0, chunk by chunk (alpha recomputed from its checkpoint into
// registers), seeded at each own step with G, formed here from alpha_{j+1}
// and beta_{j+1} (beta recomputed downwards from its checkpoint at the
// chunk's top); writes dsa / dpr of local steps [0, g + S), with the direct
// term -e of the extrinsic (ext = ... - (sys + la)) at the own steps.  e: the
// adjoints of the own steps' extrinsics, read from ecur (natural order:
// position t, or Pi(t) in a lower half) a chunk ahead.
template <typename R>
struct AlphaIn {
  R a0[8], b0[8], e[8];
};

template <typename R>
__device__ __forceinline__ void alpha_adjoint_sweep(const Lane<R>& L0, const Layout& Y, const Chunks& C, bool lower,
                                                    const R* ecur) {
  const Lane<R> L = fresh(L0);
  R ab[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t end = L.g + L.S, top = 2 * L.g + L.S;
  const GlobalPtr<const R> eg = global_view(ecur);
  // own steps' positions, walked down from the last one (loads go in
  // descending order of steps)
  QppWalk qw((uint32_t)L.t_of(end - 1), L.K, L.f1, L.f2);
  // b0 and e are loaded in own chunks only, and nothing else defines them
  // (the body reads them at own steps only): a buffer entry that another path
  // defines too (zeros for the warm-up chunks, say) is copied where the paths
  // join, and the copy waits for its load, so the load one chunk ahead would
  // wait at once.
  auto load = [&](uint32_t k, AlphaIn<R>& x) {
    const uint32_t lo = C.a_lo(k), n = C.a_hi(k, end) - lo;
    ckload8(L, Y.cal, Y.cna, k, x.a0);
    if (lo >= L.g) {  // an own chunk (alpha chunks start at the own steps' grid)
      ckload8(L, Y.cbe, Y.cnb, (lo - L.g) / 8u + 1u, x.b0);
      if (lower) {
        // the walk stays in [0, K), so every entry loads unconditionally
        // (entries past a short chunk, jj >= n, are never used); a branch
        // per entry would make each load wait for the ones before it
#pragma unroll
        for (int jj = 7; jj >= 0; --jj) {
          qw.pin();
          x.e[jj] = eg[qw.pi];
          qw.prev_if((uint32_t)jj < n && lo + (uint32_t)jj > L.g);
        }
      } else {
        const GlobalPtr<const R> row = eg + row_t(L, lo);
#pragma unroll
        for (int jj = 7; jj >= 0; --jj)
          if ((uint32_t)jj < n) x.e[jj] = row[jj];
      }
    }
  };
  auto body = [&](uint32_t k, const AlphaIn<R>& xc) {
    const uint32_t lo = C.a_lo(k), n = C.a_hi(k, end) - lo;
    const bool own = lo >= L.g;  // alpha chunks start at the own steps' grid
    const uint32_t nin = own ? min(lo + 8u, top) - lo : n;
    In8<R> in;
    load_in8(L, Y, (int)lo, (int)nin, in);
    // areg[jj] = alpha at local lo + jj; within a chunk only its last step is
    // renormalised, so areg[jj + 1] is also alpha_{lo+jj+1} before
    // renormalisation, and anext + amx that of the last step in an own chunk
    R areg[8][8], anext[8], amx = R(0);
    {
      R a[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) a[s] = xc.a0[s];
#pragma unroll
      for (uint32_t jj = 0; jj < 8; ++jj) {
        if (jj < n) {
#pragma unroll
          for (int s = 0; s < 8; ++s) areg[jj][s] = a[s];
          if (jj + 1 < n || own) {
            R o[8];
            alpha_step(a, kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], o, false);
            if (alpha_nrm(lo + jj, C.off)) {  // normalise(), keeping the maximum it subtracts
              R m = o[0];
#pragma unroll
              for (int s = 1; s < 8; ++s) m = fmax(o[s], m);
#pragma unroll
              for (int s = 0; s < 8; ++s) o[s] -= m;
              amx = m;
            }
            select8(L.t_of(lo + jj) >= 0, o, a);
          }
        }
      }
#pragma unroll
      for (int s = 0; s < 8; ++s) anext[s] = a[s];  // alpha_{lo + n} in an own chunk
    }
    // Descending over the chunk: at an own step, G from alpha_{j+1} and b =
    // beta_{j+1} (beta recomputed downwards from the checkpoint at the chunk's
    // top, nin steps in an own chunk, so also above the last own step of a
    // short chunk) joins the adjoint, then the adjoint step of j.
    R b[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) b[s] = xc.b0[s];
#pragma unroll
    for (int jj = 7; jj >= 0; --jj) {
      if ((uint32_t)jj < nin) {
        const uint32_t j = lo + (uint32_t)jj;
        if ((uint32_t)jj < n) {
          if (own) {
            R T[8], gv[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) T[s] = (jj == 7 || (uint32_t)jj + 1 >= n ? anext[s] : areg[jj + 1 < 8 ? jj + 1 : 7][s]) + b[s];
            llr_adjoint(T, xc.e[jj], gv);
#pragma unroll
            for (int s = 0; s < 8; ++s) ab[s] += gv[s];
          }
          // weights from alpha_{j+1} before renormalisation (steps whose
          // alpha was not stepped, t < 0, write only values no later pass
          // reads)
          R ds, dp;
          if ((uint32_t)jj + 1 < n) {
            alpha_adjoint_step<R, true>(areg[jj], areg[jj < 7 ? jj + 1 : 7], kHalfL2e<R> * in.sa[jj],
                                        kHalfL2e<R> * in.pr[jj], ab, ds, dp);
          } else if (own) {
            R an[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) an[s] = anext[s] + amx;
            alpha_adjoint_step<R, true>(areg[jj], an, kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], ab, ds,
                                        dp);
          } else {
            alpha_adjoint_step(areg[jj], areg[jj], kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], ab, ds, dp);
          }
          L.at(Y.dsa, j) = own ? ds - xc.e[jj] : ds;
          L.at(Y.dpr, j) = dp;
        }
        if (own && jj > 0) {  // b: beta at local lo + jj
          R o[8];
          beta_step(b, kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], o, beta_nrm(j, L.g));
          select8(L.t_of(j) < (int)L.K, o, b);
        }
      }
    }
  };
  loop_down<AlphaIn<R>>(0, C.na, load, body);
}

// Phase 3, beta side: the beta adjoint forwards over local steps
// g .. 2g + S - 1, chunk by chunk (beta recomputed downwards from its
// checkpoint into registers), with G as the output adjoints of beta_{j+1} at
// the own steps, formed here from beta_{j+1} and alpha_{j+1} (alpha stepped
// up from its checkpoint at the chunk's start); writes dsb / dpb of local
// steps [g, 2g + S).  Returns in bb the adjoint of beta at the window top (at
// the block end, that of the tail beta).  e as for the alpha side.
template <typename R>
struct BetaRe {
  R b0[8], a0[8], e[8];
};

template <typename R>
__device__ __forceinline__ void beta_adjoint_sweep(const Lane<R>& L0, const Layout& Y, const Chunks& C, bool lower,
                                                   const R* ecur, R (&bb)[8]) {
  const Lane<R> L = fresh(L0);
  const uint32_t top = 2 * L.g + L.S, own_end = L.g + L.S;
  const GlobalPtr<const R> eg = global_view(ecur);
  // own steps' positions, walked up from the first one
  QppWalk qw(L.q * L.S, L.K, L.f1, L.f2);
#pragma unroll
  for (int s = 0; s < 8; ++s) bb[s] = 0;
  // a0 and e: own chunks only, as on the alpha side
  auto load = [&](uint32_t m, BetaRe<R>& x) {
    const uint32_t lo = L.g + 8u * m;
    ckload8(L, Y.cbe, Y.cnb, m + 1, x.b0);
    if (lo < own_end) {
      ckload8(L, Y.cal, Y.cna, C.own_alpha(m, L.g), x.a0);
      if (lower) {
        // unconditional loads, as on the alpha side (entries past the own
        // steps are never used)
#pragma unroll
        for (uint32_t jj = 0; jj < 8; ++jj) {
          qw.pin();
          x.e[jj] = eg[qw.pi];
          qw.next_if(lo + jj < own_end);
        }
      } else {
        const GlobalPtr<const R> row = eg + row_t(L, lo);
#pragma unroll
        for (uint32_t jj = 0; jj < 8; ++jj)
          if (lo + jj < own_end) x.e[jj] = row[jj];
      }
    }
  };
  auto body = [&](uint32_t m, const BetaRe<R>& xc) {
    const uint32_t lo = L.g + 8u * m, n = min(lo + 8u, top) - lo;
    const bool own = lo < own_end;
    const uint32_t nown = own ? min(8u, own_end - lo) : 0u;
    In8<R> in;
    load_in8(L, Y, (int)lo, (int)n, in);
    R breg[8][8];
    {
      R b[8];
#pragma unroll
      for (int s = 0; s < 8; ++s) b[s] = xc.b0[s];
#pragma unroll
      for (int jj = 7; jj >= 0; --jj) {
        if ((uint32_t)jj < n) {
#pragma unroll
          for (int s = 0; s < 8; ++s) breg[jj][s] = b[s];  // beta at local lo + jj + 1
          if (jj > 0) {
            const uint32_t j = lo + (uint32_t)jj;
            R o[8];
            beta_step(b, kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], o, beta_nrm(j, L.g));
            select8(L.t_of(j) < (int)L.K, o, b);
          }
        }
      }
    }
    // Ascending over the chunk: at an own step, alpha_{j+1} stepped up from
    // the checkpoint at the chunk's start and G from it and beta_{j+1} seed
    // the output adjoints of beta_{j+1}, then the adjoint step of j.
    R a[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) a[s] = xc.a0[s];
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      if (jj < n) {
        const uint32_t j = lo + jj;
        R bn[8], ds, dp;
        if (jj < nown) {
          R o[8];
          alpha_step(a, kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], o, alpha_nrm(lo + jj, C.off));
          R T[8];
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            a[s] = o[s];  // own steps lie inside the block
            T[s] = o[s] + breg[jj][s];
          }
          llr_adjoint(T, xc.e[jj], bn);
        } else {
#pragma unroll
          for (int s = 0; s < 8; ++s) bn[s] = R(0);
        }
        // weights from beta_j = breg[jj - 1]: within a chunk only beta at its
        // first step (jj = 0) is renormalised, and that one is not in breg
        // (steps past the block end, whose beta was not stepped, are dropped
        // by the select below)
        if (jj > 0) {
          beta_adjoint_step<R, true>(breg[jj], breg[jj > 0 ? jj - 1 : 0], kHalfL2e<R> * in.sa[jj],
                                     kHalfL2e<R> * in.pr[jj], bb, bn, ds, dp);
        } else {
          beta_adjoint_step(breg[jj], breg[jj], kHalfL2e<R> * in.sa[jj], kHalfL2e<R> * in.pr[jj], bb, bn, ds, dp);
        }
        // past the block end the adjoint stops (beta is not stepped there)
        select8(L.t_of(j) < (int)L.K, bn, bb);
        L.at(Y.dsb, j) = ds;
        L.at(Y.dpb, j) = dp;
      }
    }
  };
  loop_up<BetaRe<R>>(0, C.nb, load, body);
}

template <typename R>
__device__ Lane<R> make_lane(const DiffJob& J, const DiffArgs<R>& A) {
  Lane<R> L;
  L.lane = threadIdx.x & 63u;
  L.K = J.k;
  L.S = J.seg;
  L.M = J.nseg;
  L.g = J.warm;
  L.q = L.lane % J.nseg;
  L.f1 = J.f1;
  L.f2 = J.f2;
  L.active = L.lane / J.nseg < J.nblk;
  L.sc = A.scratch + (size_t)blockIdx.x * 64 * A.slot_elems;
  L.ck = nullptr;
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

// The own chunks a wave takes in phases 2 and 4: wave 0 the first half
// (rounded up), wave 1 the rest.
__device__ __forceinline__ void own_range(const Chunks& C, uint32_t wave, uint32_t& m0, uint32_t& m1) {
  const uint32_t h = (C.nown + 1) / 2;
  m0 = wave ? h : 0u;
  m1 = wave ? C.nown : h;
}

// A lane's code block: its descriptor and the pointers derived from it.
template <typename R>
struct BlockPtrs {
  DiffCb c;
  const R *d0, *d1, *d2, *sv;
  R* sv_mut;
  __device__ BlockPtrs(const DiffArgs<R>& A, const DiffJob& J, uint32_t blk) : c(A.cbs[J.cb_first + blk]) {
    const uint32_t D = J.k + 4;
    d0 = A.llr + c.llr_off;
    d1 = d0 + D;
    d2 = d1 + D;
    sv_mut = A.saved + c.save_off;
    sv = sv_mut;
  }
  __device__ R* g(const DiffArgs<R>& A, uint32_t stream, uint32_t k) const { return A.grad_llr + c.llr_off + (size_t)stream * (k + 4); }
  __device__ R* ebar(const DiffArgs<R>& A, uint32_t which, uint32_t k) const { return A.ebar + c.ebar_off + (size_t)which * k; }
};

// Phase 1 of a half on both waves: staging is done; wave 0 runs alpha, wave 1
// beta from its start (the tail beta when the window reaches the block end).
template <typename R>
__device__ __forceinline__ void checkpoint_passes(const Lane<R>& L, const Layout& Y, const Chunks& C, uint32_t wave,
                                                  uint32_t h, const R* d0, const R* d1, const R* d2) {
  if (!L.active) return;
  if (wave == 0) {
    alpha_ckpt_pass(L, Y, C);
  } else {
    R b[8];
    if (L.q * L.S + L.S + L.g >= L.K) {
      R ts[3], tp[3];
      tail_inputs(h, d0, d1, d2, L.K, ts, tp);
      tail_beta(ts, tp, b);
    } else {
#pragma unroll
      for (int s = 0; s < 8; ++s) b[s] = 0;
    }
    beta_ckpt_pass(L, Y, C, b);
  }
}

}  // namespace

// Waves per SIMD the kernels are built for: two (one job's two waves, or two
// jobs) in float, one in double (whose registers would not fit two).
template <typename R>
constexpr int kDiffWavesPerSimd = sizeof(R) == 4 ? 2 : 1;

// kLds: the plan's jobs fit their block copies in LDS (A.use_lds); the
// global-memory path is a kernel of its own, so its code does not take
// registers from the LDS path.
template <typename R, bool kLds>
__global__ __launch_bounds__(128, kDiffWavesPerSimd<R>) void diff_forward_kernel(DiffArgs<R> A) {
  extern __shared__ __attribute__((aligned(16))) char diff_lds[];
  const Layout Y(A.max_seg, A.max_warm);
  // the wave's index in the workgroup, as a scalar (uniform in a wave;
  // read through readfirstlane, everything derived from it stays in scalar
  // registers and scalar branches)
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (uint32_t jb = blockIdx.x; jb < A.njobs; jb += gridDim.x) {
    const DiffJob J = A.jobs[jb];
    const Lane<R> L0 = make_lane(J, A);
    const uint32_t blk = L0.active ? L0.lane / J.nseg : 0;
    const BlockLds<R> X = block_lds(reinterpret_cast<R*>(diff_lds), J);
    const Lane<R> L = kLds ? with_lds(L0, X, blk) : L0;
    // the lane's block, fetched where a phase needs it (its pointers then
    // hold no registers through the passes)
    auto cb = [&]() { return BlockPtrs<R>(A, J, blk); };
    const Chunks CH(L.S, L.g);
    uint32_t m0, m1;
    own_range(CH, wave, m0, m1);
    for (uint32_t h = 0; h < A.halves; ++h) {
      if constexpr (kLds) {
        stage_fill(A, J, h, X);
        __syncthreads();
        stage_gather(J, h, X);
      } else if (wave == 0 && L.active) {
        const BlockPtrs<R> B = cb();
        stage(L, Y, h, B.d0, B.d1, B.d2, B.sv, A.scales);
      }
      __syncthreads();
      // this half's checkpoints, kept for the adjoint
      Lane<R> Lh = L;
      Lh.ck = A.ckpt + ((size_t)jb * A.halves + h) * A.ckpt_job_half;
      if (L.active) {
        const BlockPtrs<R> B = cb();
        checkpoint_passes(Lh, Y, CH, wave, h, B.d0, B.d1, B.d2);
      }
      __syncthreads();
      if (L.active) {
        const bool lower = (h & 1u) != 0;
        R* e = cb().sv_mut + (size_t)h * L.K;
        QppWalk qw(L.q * L.S + 8u * m0, L.K, L.f1, L.f2);
        own_terms(Lh, Y, CH, m0, m1, [&](uint32_t j, const R (&T)[8], R sa) {
          R m[2], s[2], p[8];
          class_sums(T, m, s, p);
          const R l1 = m[1] + Num<R>::lg2(s[1]), l0 = m[0] + Num<R>::lg2(s[0]);
          e[lower ? qw.pi : (uint32_t)L.t_of(j)] = (l1 - l0) * kLn2<R> - sa;
          qw.next();
        });
      }
      __syncthreads();
    }
    if (L.active) {
      const BlockPtrs<R> B = cb();
      const uint32_t H = A.halves;
      const R* e1 = B.sv + (size_t)(H - 2) * L.K;
      const R* e2 = B.sv + (size_t)(H - 1) * L.K;
      R* post = A.post + B.c.post_off;
      const R w1 = A.scales[H - 2], w2 = A.scales[H - 1];
      const uint32_t lo = L.q * L.S + 8u * m0, hi = min(L.q * L.S + 8u * m1, L.q * L.S + L.S);
      for (uint32_t t0 = lo; t0 < hi; t0 += 8) {
        R x[8], y1[8], y2[8];
#pragma unroll
        for (uint32_t jj = 0; jj < 8; ++jj)
          if (t0 + jj < hi) {
            x[jj] = B.d0[t0 + jj];
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

// Posterior adjoint of the positions [lo, hi) of a block segment: grad d0 =
// grad post, the adjoint of the last half's extrinsic, and this lane's terms
// of the last two scales' adjoints (post = d0 + w[H-2] e_{H-2} +
// w[H-1] e_{H-1}); the tail positions are zeroed by the last segment's lane
// when tail is set.
template <typename R>
__device__ __forceinline__ void post_adjoint(const DiffArgs<R>& A, const Lane<R>& L, const R* sv, const R* gp, R* g0,
                                             R* g1, R* g2, R* el, uint32_t lo, uint32_t hi, bool tail, R& s1, R& s2) {
  const uint32_t H = A.halves, K = L.K;
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
  if (tail && L.q == L.M - 1)
    for (uint32_t t = K; t < K + 4; ++t) g0[t] = g1[t] = g2[t] = 0;
}

// Own steps [jlo, jhi) (local) of the combine of one half: the alpha- and
// beta-side sums of the own window, then the next segment's alpha warm-up
// and the previous segment's beta warm-up (in that order), added into the
// LLR adjoints and the previous half's extrinsic adjoint; returns this lane's
// term of the scale adjoint.  Read-modify-writes in chunks of 8 steps (t, or
// Pi(t), is a permutation, so no two steps share a position).
template <typename R>
struct CombIn {
  R ds[8], dp[8], dsn[8], dpn[8], dsp[8], dpp[8], gs[8], gq[8], pv[8], ep[8];
  uint32_t at[8];
};

// the own-window sums of local step j: alpha side + beta side
template <typename R>
__device__ __forceinline__ void own_sums(const Lane<R>& L, const Layout& Y, uint32_t j, R& ds, R& dp) {
  ds = L.at(Y.dsa, j) + L.at(Y.dsb, j);
  dp = L.at(Y.dpr, j) + L.at(Y.dpb, j);
}

template <typename R>
__device__ __forceinline__ R combine(const Lane<R> L0, const Layout Y, uint32_t h, R w, const R* eprev, R* g0, R* gpar,
                                     R* prev, uint32_t jlo, uint32_t jhi) {
  const Lane<R> L = fresh(L0);
  const bool lower = (h & 1u) != 0;
  const bool with_prev = h != 0;  // the first upper half has no a-priori input
  const bool has_next = L.q + 1 < L.M, has_prev = L.q > 0;
  const uint32_t ln = min(L.lane + 1, 63u), lp = L.lane > 0 ? L.lane - 1 : 0u;
  QppWalk qw((uint32_t)L.t_of(jlo), L.K, L.f1, L.f2);
  R wsum = 0;
  auto load = [&](uint32_t j0, CombIn<R>& x) {
#pragma unroll
    for (uint32_t jj = 0; jj < 8; ++jj) {
      const uint32_t j = j0 + jj;
      if (j < jhi) {
        const uint32_t t = (uint32_t)L.t_of(j);
        const uint32_t i = lower ? qw.pi : t;
        qw.next();
        x.at[jj] = i;
        own_sums(L, Y, j, x.ds[jj], x.dp[jj]);
        if (j >= L.S) {  // t in [qS + S - g, qS + S): lane q+1's local j - S
          x.dsn[jj] = L.nb(Y.dsa, j - L.S, ln);
          x.dpn[jj] = L.nb(Y.dpr, j - L.S, ln);
        }
        if (j < 2 * L.g) {  // t in [qS, qS + g): lane q-1's local j + S
          x.dsp[jj] = L.nb(Y.dsb, j + L.S, lp);
          x.dpp[jj] = L.nb(Y.dpb, j + L.S, lp);
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
      if (j < jhi) {
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
  chunks_up<CombIn<R>>(jlo, jhi, load, body);
  return wsum;
}

// LDS path of the combine.  Fill: arrays 0 = grad d0, 1 = grad of the
// half's parity stream, 2 = the adjoint of e_{h-1} started as the per-lane
// path starts it (the posterior term for h - 1 == H - 2, else zero), 3 =
// e_{h-1}.  The lanes then add into the copies in the order of combine; the
// write-back stores arrays 0-2.  Two blocks per batch of loads.
template <typename R>
__device__ __forceinline__ void combine_fill(const DiffArgs<R> A, const DiffJob J, uint32_t h, const BlockLds<R> X) {
  const uint32_t H = A.halves, K = J.k, nt = blockDim.x;
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
    for (uint32_t i0 = threadIdx.x; i0 < K; i0 += nt * kFillBatch) {
      R x0[2][kFillBatch], x1[2][kFillBatch], x2[2][kFillBatch], x3[2][kFillBatch];
#pragma unroll
      for (uint32_t x = 0; x < 2; ++x)
#pragma unroll
        for (uint32_t u = 0; u < kFillBatch; ++u) {
          const uint32_t i = i0 + nt * u;
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
          const uint32_t i = i0 + nt * u;
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
    for (uint32_t i = threadIdx.x; i < K; i += blockDim.x) {
      g0[i] = xl[X.idx(0, b, i)];
      gq[i] = xl[X.idx(1, b, i)];
      if (with_prev) prev[i] = xl[X.idx(2, b, i)];
    }
  }
}

// Per-lane part of the LDS combine over own local steps [jlo, jhi) (the same
// sums, in the same order, as combine), with the next chunk's scratch loads
// in flight while a chunk is added into the LDS copies.
// Chunks of four steps: the raw terms of two chunks in flight stay within the
// registers the kernel has (with eight, the combine's buffers made the
// allocator spill values that the sweeps then reload).
constexpr uint32_t kCombStep = 4;
template <typename R>
struct CombLdsIn {
  R dsa[kCombStep], dsb[kCombStep], dpa[kCombStep], dpb[kCombStep], dsn[kCombStep], dpn[kCombStep],
      dsp[kCombStep], dpp[kCombStep];
};

template <typename R>
__device__ __forceinline__ R combine_lds(const Lane<R>& L0, const Layout& Y, uint32_t h, R w, const BlockLds<R>& X,
                                         uint32_t blk, uint32_t jlo, uint32_t jhi) {
  const Lane<R> L = fresh(L0);
  const bool lower = (h & 1u) != 0;
  const bool with_prev = h != 0;
  const bool has_next = L.q + 1 < L.M, has_prev = L.q > 0;
  const uint32_t ln = min(L.lane + 1, 63u), lp = L.lane > 0 ? L.lane - 1 : 0u;
  QppWalk qw((uint32_t)L.t_of(jlo), L.K, L.f1, L.f2);
  R wsum = 0;
  // Loads: every entry, at steps clamped into [j0, jhi) and neighbour
  // indices clamped into their fields (entries of steps past jhi, and
  // neighbour terms a step does not take, are loaded and not used), so the
  // batch goes out without a branch per load.
  const uint32_t top = L.S + 2 * L.g;
  auto load = [&](uint32_t j0, CombLdsIn<R>& x) {
#pragma unroll
    for (uint32_t jj = 0; jj < kCombStep; ++jj) {
      const uint32_t j = min(j0 + jj, jhi - 1);
      // the own-window sums' terms, added in the body (an add here would
      // wait for the loads)
      x.dsa[jj] = L.at(Y.dsa, j);
      x.dsb[jj] = L.at(Y.dsb, j);
      x.dpa[jj] = L.at(Y.dpr, j);
      x.dpb[jj] = L.at(Y.dpb, j);
      // t in [qS + S - g, qS + S) (j >= S): lane q+1's local j - S
      const uint32_t jn = j >= L.S ? j - L.S : 0u;
      x.dsn[jj] = L.nb(Y.dsa, jn, ln);
      x.dpn[jj] = L.nb(Y.dpr, jn, ln);
      // t in [qS, qS + g) (j < 2g): lane q-1's local j + S
      const uint32_t jp = min(j + L.S, top - 1);
      x.dsp[jj] = L.nb(Y.dsb, jp, lp);
      x.dpp[jj] = L.nb(Y.dpb, jp, lp);
    }
  };
  // Each step updates its own positions (t, and Pi(t) in a lower half, a
  // permutation: no two steps share one), so a chunk reads all its LDS
  // values first and then writes them (one wait, where read-add-write per
  // step waits at every step); entries past jhi read and do not write.
  auto body = [&](uint32_t j0, const CombLdsIn<R>& x) {
    uint32_t ii[kCombStep], tt[kCombStep];
    R aa[kCombStep], cc[kCombStep];
#pragma unroll
    for (uint32_t jj = 0; jj < kCombStep; ++jj) {
      const uint32_t j = j0 + jj;
      tt[jj] = (uint32_t)L.t_of(min(j, jhi - 1));
      ii[jj] = lower ? qw.pi : tt[jj];
      qw.next_if(j < jhi);
      R a = x.dsa[jj] + x.dsb[jj], c = x.dpa[jj] + x.dpb[jj];  // own_sums
      if (j >= L.S) {
        a = has_next ? a + x.dsn[jj] : a;
        c = has_next ? c + x.dpn[jj] : c;
      }
      if (j < 2 * L.g) {
        a = has_prev ? a + x.dsp[jj] : a;
        c = has_prev ? c + x.dpp[jj] : c;
      }
      aa[jj] = a;
      cc[jj] = c;
    }
    R r0[kCombStep], r1[kCombStep], r2[kCombStep], r3[kCombStep];
#pragma unroll
    for (uint32_t jj = 0; jj < kCombStep; ++jj) {
      r0[jj] = X.at(0, blk, ii[jj]);
      r1[jj] = X.at(1, blk, tt[jj]);
      if (with_prev) {
        r2[jj] = X.at(2, blk, ii[jj]);
        r3[jj] = X.at(3, blk, ii[jj]);
      }
    }
#pragma unroll
    for (uint32_t jj = 0; jj < kCombStep; ++jj) {
      if (j0 + jj < jhi) {
        X.at(0, blk, ii[jj]) = r0[jj] + aa[jj];
        X.at(1, blk, tt[jj]) = r1[jj] + cc[jj];
        if (with_prev) {
          X.at(2, blk, ii[jj]) = r2[jj] + w * aa[jj];
          wsum += aa[jj] * r3[jj];
        }
      }
    }
  };
  chunks_up<CombLdsIn<R>, kCombStep>(jlo, jhi, load, body);
  return wsum;
}

// A workgroup barrier that orders LDS accesses only: global stores issued
// before it may still be in flight after it.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Sum over the wave, then over the workgroup's two waves (wave 1's partial
// through LDS, added to wave 0's in that order); the result is valid in
// thread 0.  kLdsOnly: its barrier orders LDS only (lds_barrier).
template <bool kLdsOnly = false, typename R>
__device__ __forceinline__ R workgroup_sum(R v, R* slot) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (threadIdx.x == 64) *slot = v;
  if constexpr (kLdsOnly)
    lds_barrier();
  else
    __syncthreads();
  return v + *slot;
}

}  // namespace

template <typename R, bool kLds>
__global__ __launch_bounds__(128, kDiffWavesPerSimd<R>) void diff_backward_kernel(DiffArgs<R> A) {
  extern __shared__ __attribute__((aligned(16))) char diff_lds[];
  __shared__ R wave1_part[2];
  const Layout Y(A.max_seg, A.max_warm);
  const uint32_t H = A.halves;
  // the wave's index in the workgroup, as a scalar (uniform in a wave;
  // read through readfirstlane, everything derived from it stays in scalar
  // registers and scalar branches)
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (uint32_t jb = blockIdx.x; jb < A.njobs; jb += gridDim.x) {
    const DiffJob J = A.jobs[jb];
    const Lane<R> L0 = make_lane(J, A);
    const uint32_t blk = L0.active ? L0.lane / J.nseg : 0;
    const BlockLds<R> X = block_lds(reinterpret_cast<R*>(diff_lds), J);
    const Lane<R> L = kLds ? with_lds(L0, X, blk) : L0;
    // the lane's block, fetched where a phase needs it (its pointers then
    // hold no registers through the passes)
    auto cb = [&]() { return BlockPtrs<R>(A, J, blk); };
    const uint32_t K = L.K;
    R* dsc = A.dscale_part + (size_t)jb * H;
    const Chunks CH(L.S, L.g);
    uint32_t m0, m1;
    own_range(CH, wave, m0, m1);
    // own local steps of this wave in phase 4, and their natural positions
    const uint32_t jlo = L.g + min(8u * m0, L.S), jhi = L.g + min(8u * m1, L.S);
    const uint32_t lo = L.q * L.S + (jlo - L.g), hi = L.q * L.S + (jhi - L.g);
    {
      R s1 = 0, s2 = 0;
      if (L.active) {
        const BlockPtrs<R> B = cb();
        post_adjoint(A, L, B.sv, A.grad_post + B.c.post_off, B.g(A, 0, K), B.g(A, 1, K), B.g(A, 2, K),
                     B.ebar(A, (H - 1) & 1u, K), lo, hi, wave == 0, s1, s2);
      }
      s1 = workgroup_sum(s1, &wave1_part[0]);
      s2 = workgroup_sum(s2, &wave1_part[1]);
      if (threadIdx.x == 0) {
        for (uint32_t h = 0; h < H; ++h) dsc[h] = 0;
        dsc[H - 2] = s1;
        dsc[H - 1] = s2;
      }
    }
    __syncthreads();
    for (uint32_t h = H; h-- > 0;) {
      const bool lower = (h & 1u) != 0;
      if constexpr (kLds) {
        stage_fill(A, J, h, X);
        __syncthreads();
        stage_gather(J, h, X);
      } else if (L.active) {
        const BlockPtrs<R> B = cb();
        // adjoint of e_{h-1}: starts at the posterior term for h - 1 == H - 2
        // (the LDS path starts it in its combine fill)
        R* prev = B.ebar(A, (h + 1) & 1u, K);
        const R* gp = A.grad_post + B.c.post_off;
        if (h >= 1) {
          if (h - 1 == H - 2) {
            const R wp = A.scales[H - 2];
            for (uint32_t t = lo; t < hi; ++t) prev[t] = wp * gp[t];
          } else {
            for (uint32_t t = lo; t < hi; ++t) prev[t] = 0;
          }
        }
        if (wave == 0) stage(L, Y, h, B.d0, B.d1, B.d2, B.sv, A.scales);
      }
      __syncthreads();
      // the forward's checkpoints of this half (phase 1 ran there)
      Lane<R> Lh = L;
      Lh.ck = A.ckpt + ((size_t)jb * H + h) * A.ckpt_job_half;
      // phase 3: the two adjoint sweeps, each forming the LLR terms G of the
      // own steps it needs.  Wave 1 takes the alpha side, whose sweep has the
      // more work: of a workgroup's two waves, wave 0 was measured slower on
      // the same code (both ways round), so the pairing that ends the two
      // together is this one.
      // The tail LLR adjoints of the block, summed in segment order into the
      // last segment's lane of wave 0 (the beta side, whose lanes hold each
      // segment's share), for phase 4.
      R tail_s[3] = {0, 0, 0}, tail_p[3] = {0, 0, 0};
      if (L.active) {
        if (wave == 1) {
          alpha_adjoint_sweep(Lh, Y, CH, lower, cb().ebar(A, h & 1u, K));
        } else {
          R bb[8];
          beta_adjoint_sweep(Lh, Y, CH, lower, cb().ebar(A, h & 1u, K), bb);
          // beta of the window top: the tail beta when it reaches K
          R dts[3] = {0, 0, 0}, dtp[3] = {0, 0, 0};
          if (L.q * L.S + L.S + L.g >= K) {
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
          // segments m0 .. M-1 reach the block end (m S + S + g >= K); lane
          // q = M - 1 of each block reads their shares from the block's
          // lanes, ascending (the other lanes' sums are not used)
          const uint32_t m0 = K > L.S + L.g ? (K - L.g - 1) / L.S : 0u;
          for (uint32_t m = m0; m < L.M; ++m) {
            const int src = (int)L.lane - (int)(L.M - 1) + (int)m;
#pragma unroll
            for (int k3 = 0; k3 < 3; ++k3) {
              tail_s[k3] += shfl_r(dts[k3], src);
              tail_p[k3] += shfl_r(dtp[k3], src);
            }
          }
        }
      }
      __syncthreads();
      // phase 4: combine
      R wsum = 0;
      const R w = h ? A.scales[h - 1] : R(0);
      if constexpr (kLds) {
        combine_fill(A, J, h, X);
        __syncthreads();
        if (L.active) wsum = combine_lds(L, Y, h, w, X, blk, jlo, jhi);
        __syncthreads();
        combine_store(A, J, h, X);
      } else if (L.active) {
        const BlockPtrs<R> B = cb();
        const R* eprev = B.sv + (size_t)(h ? h - 1 : 0) * K;
        wsum = combine(L, Y, h, w, eprev, B.g(A, 0, K), B.g(A, lower ? 2 : 1, K), B.ebar(A, (h + 1) & 1u, K), jlo,
                       jhi);
      }
      if (L.active && wave == 0 && L.q == L.M - 1) {
        const BlockPtrs<R> B = cb();
        R* g0 = B.g(A, 0, K);
        R* g1 = B.g(A, 1, K);
        R* g2 = B.g(A, 2, K);
        // tail LLR adjoints added in: every load before any store (distinct
        // positions), so the loads go out together
        if (!lower) {
          const R x0 = g0[K], x1 = g0[K + 1], x2 = g0[K + 2], y0 = g1[K], y2 = g1[K + 2];
          g0[K] = x0 + tail_s[0];
          g0[K + 2] = x2 + tail_s[1];
          g1[K] = y0 + tail_p[0];
          g0[K + 1] = x1 + tail_p[1];
          g1[K + 2] = y2 + tail_p[2];
        } else {
          const R y3 = g1[K + 3], z0 = g2[K], z1 = g2[K + 1], z2 = g2[K + 2], z3 = g2[K + 3], x3 = g0[K + 3];
          g1[K + 3] = y3 + tail_s[0];
          g2[K] = z0 + tail_s[1];
          g2[K + 2] = z2 + tail_s[2];
          g2[K + 1] = z1 + tail_p[0];
          g0[K + 3] = x3 + tail_p[1];
          g2[K + 3] = z3 + tail_p[2];
        }
      }
      // the combine's global stores need not be complete here (the next
      // half's first barrier, after its fill, waits for them): these two
      // barriers order LDS only
      wsum = workgroup_sum<true>(wsum, &wave1_part[0]);
      if (threadIdx.x == 0 && h) dsc[h - 1] += wsum;
      lds_barrier();
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

template __global__ void diff_forward_kernel<float, true>(DiffArgs<float>);
template __global__ void diff_forward_kernel<float, false>(DiffArgs<float>);
template __global__ void diff_forward_kernel<double, true>(DiffArgs<double>);
template __global__ void diff_forward_kernel<double, false>(DiffArgs<double>);
template __global__ void diff_backward_kernel<float, true>(DiffArgs<float>);
template __global__ void diff_backward_kernel<float, false>(DiffArgs<float>);
template __global__ void diff_backward_kernel<double, true>(DiffArgs<double>);
template __global__ void diff_backward_kernel<double, false>(DiffArgs<double>);
template __global__ void diff_scale_reduce_kernel<float>(const float*, uint32_t, uint32_t, float*);
template __global__ void diff_scale_reduce_kernel<double>(const double*, uint32_t, uint32_t, double*);

}  // namespace tcdev

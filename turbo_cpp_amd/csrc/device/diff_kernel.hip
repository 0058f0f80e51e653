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

template <typename R>
struct Num;
// float: the hardware exp2 / log2 / reciprocal (about 1 ulp); double: libm
// and IEEE division, for the finite-difference and FP64 reference checks.
template <>
struct Num<float> {
  static __device__ float ex(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
  static __device__ float ln(float x) { return __builtin_amdgcn_logf(x) * 0.693147180559945309f; }
  static __device__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
};
template <>
struct Num<double> {
  static __device__ double ex(double x) { return exp(x); }
  static __device__ double ln(double x) { return log(x); }
  static __device__ double rcp(double x) { return 1.0 / x; }
};

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

// Jacobian logarithm and the weights of its two arguments.
template <typename R>
__device__ inline R jac(R a, R b) {
  const R m = a > b ? a : b;
  const R d = a > b ? a - b : b - a;
  return m + Num<R>::ln(R(1) + Num<R>::ex(-d));
}
template <typename R>
__device__ inline void jac_w(R a, R b, R& wa, R& wb) {
  const R d = a > b ? a - b : b - a;
  const R e = Num<R>::ex(-d);
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
  for (int s = 1; s < 8; ++s) m = v[s] > m ? v[s] : m;
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] -= m;
}

// alpha_{t+1} from alpha_t: each state's two incoming transitions in order
// of increasing predecessor state.
template <typename R>
__device__ inline void alpha_step(const R (&a)[8], R hs, R hp, R (&o)[8]) {
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
  normalise(o);
}

template <typename R>
__device__ inline void beta_step(const R (&b)[8], R hs, R hp, R (&o)[8]) {
#pragma unroll
  for (unsigned s = 0; s < 8; ++s)
    o[s] = jac(b[nxt(s, 0)] + gam(0u, par(s, 0), hs, hp), b[nxt(s, 1)] + gam(1u, par(s, 1), hs, hp));
  normalise(o);
}

// L_u = log sum_s exp(alpha(s) + gamma(s,u) + beta(next(s,u))) and, when
// P is given, the weights P[u][s] = exp(m(s,u) - L_u).
template <typename R>
__device__ inline void llr_terms(const R (&a)[8], const R (&b)[8], R hs, R hp, R (&L)[2], R (*P)[8]) {
#pragma unroll
  for (unsigned u = 0; u < 2; ++u) {
    R m[8];
    R mx = R(kNegF);
#pragma unroll
    for (unsigned s = 0; s < 8; ++s) {
      m[s] = a[s] + gam(u, par(s, u), hs, hp) + b[nxt(s, u)];
      mx = m[s] > mx ? m[s] : mx;
    }
    R sum = 0;
#pragma unroll
    for (unsigned s = 0; s < 8; ++s) {
      m[s] = Num<R>::ex(m[s] - mx);
      sum += m[s];
    }
    L[u] = mx + Num<R>::ln(sum);
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
      m += gam(u, par(cur, u), R(0.5) * ts[j], R(0.5) * tp[j]);
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
  __device__ void prev() {
    d = d >= d2 ? d - d2 : d + k - d2;
    pi = pi >= d ? pi - d : pi + k - d;
  }
};

// Per-lane scratch fields, each [j][lane] with j < its length.
struct Layout {
  uint32_t P, S, g;
  uint32_t sa, pr, eb, al, be, bo, dsa, dpr, tl, total;
  __device__ Layout(uint32_t s, uint32_t gw) : P(s + 2 * gw), S(s), g(gw) {
    sa = 0;
    pr = sa + P;
    eb = pr + P;
    al = eb + S;                 // (g + S + 1) x 8
    be = al + (g + S + 1) * 8;   // (S + g + 1) x 8
    bo = be + (S + g + 1) * 8;   // S x 8
    dsa = bo + S * 8;
    dpr = dsa + P;
    tl = dpr + P;                // 6
    total = tl + 6;
  }
};

template <typename R>
struct Lane {
  uint32_t K, S, M, g, q, f1, f2;
  bool active;
  uint32_t lane;
  R* sc;  // this wave's scratch, element (field + j) * 64 + lane
  __device__ R& at(uint32_t field, uint32_t j) const { return sc[(size_t)(field + j) * 64 + lane]; }
  __device__ R& nb(uint32_t field, uint32_t j, uint32_t other) const { return sc[(size_t)(field + j) * 64 + other]; }
  // local step j covers t = qS - g + j
  __device__ int t_of(uint32_t j) const { return (int)(q * S) - (int)g + (int)j; }
};

// Stage sys + la and par of one half over the lane's window.  Upper half
// (h even): natural order, la = w[h-1] e_{h-1}[t].  Lower half: position i
// reads d0[Pi(i)], d2[i] and la = w[h-1] e_{h-1}[Pi(i)] (extrinsics are kept
// in natural order).
template <typename R>
__device__ void stage(const Lane<R>& L, const Layout& Y, uint32_t h, const R* d0, const R* d1, const R* d2,
                      const R* sv, const R* scales) {
  const R w = h ? scales[h - 1] : R(0);
  const R* eprev = sv + (size_t)(h ? h - 1 : 0) * L.K;
  const int t0 = L.t_of(0);
  QppWalk qw(t0 > 0 ? (uint32_t)t0 : 0u, L.K, L.f1, L.f2);
#pragma unroll 8
  for (uint32_t j = 0; j < Y.P; ++j) {
    const int t = L.t_of(j);
    R sa = 0, pr = 0;
    if (t >= 0 && t < (int)L.K) {
      if ((h & 1u) == 0) {
        sa = d0[t] + (h ? w * eprev[t] : R(0));
        pr = d1[t];
      } else {
        const uint32_t pi = qw.pi;
        sa = d0[pi] + w * eprev[pi];
        pr = d2[t];
      }
      qw.next();
    }
    L.at(Y.sa, j) = sa;
    L.at(Y.pr, j) = pr;
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

// alpha over local steps 0 .. g + S, stored at al[j] (alpha_t for t = t_of(j)).
// The next step's inputs are loaded before this step's stores, so their
// latency overlaps the recursion.
template <typename R>
__device__ void alpha_pass(const Lane<R>& L, const Layout& Y) {
  R a[8];
  const bool from_zero_state = (int)(L.q * L.S) - (int)L.g <= 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) a[s] = from_zero_state ? (s == 0 ? R(0) : R(kNegF)) : R(0);
  const uint32_t n = L.g + L.S;
  R sa_n = L.at(Y.sa, 0), pr_n = L.at(Y.pr, 0);
  for (uint32_t j = 0; j < n; ++j) {
    const R sa = sa_n, pr = pr_n;
    if (j + 1 < n) {
      sa_n = L.at(Y.sa, j + 1);
      pr_n = L.at(Y.pr, j + 1);
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) L.at(Y.al + s * (Y.g + Y.S + 1), j) = a[s];
    if (L.t_of(j) >= 0) {
      R o[8];
      alpha_step(a, R(0.5) * sa, R(0.5) * pr, o);
#pragma unroll
      for (int s = 0; s < 8; ++s) a[s] = o[s];
    }
  }
#pragma unroll
  for (int s = 0; s < 8; ++s) L.at(Y.al + s * (Y.g + Y.S + 1), n) = a[s];
}

// Forward beta pass that also emits the extrinsics of the own steps
// (alpha from the alpha pass, beta_{t+1} in registers); beta is not stored.
template <typename R>
__device__ void beta_ext_pass(const Lane<R>& L, const Layout& Y, const R (&ts)[3], const R (&tp)[3], uint32_t h,
                              R* e) {
  R b[8];
  const bool from_tail = L.q * L.S + L.S + L.g >= L.K;
  if (from_tail) {
    tail_beta(ts, tp, b);
  } else {
#pragma unroll
    for (int s = 0; s < 8; ++s) b[s] = 0;
  }
  const uint32_t top = 2 * L.g + L.S;
  const uint32_t own_end = L.g + L.S;  // own steps: g <= j < g + S
  R sa_n = L.at(Y.sa, top - 1), pr_n = L.at(Y.pr, top - 1), a_n[8];
#pragma unroll
  for (int s = 0; s < 8; ++s) a_n[s] = (top - 1 < own_end) ? L.at(Y.al + s * (Y.g + Y.S + 1), top - 1) : R(0);
  // Pi of the own steps, from the last one down
  QppWalk qw((uint32_t)L.t_of(own_end - 1), L.K, L.f1, L.f2);
  for (uint32_t j = top; j-- > L.g;) {
    const R sa = sa_n, pr = pr_n;
    R a[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) a[s] = a_n[s];
    if (j > L.g) {
      sa_n = L.at(Y.sa, j - 1);
      pr_n = L.at(Y.pr, j - 1);
      if (j - 1 < own_end) {
#pragma unroll
        for (int s = 0; s < 8; ++s) a_n[s] = L.at(Y.al + s * (Y.g + Y.S + 1), j - 1);
      }
    }
    const int t = L.t_of(j);
    if (j < own_end) {
      R Lu[2];
      llr_terms(a, b, R(0.5) * sa, R(0.5) * pr, Lu, (R(*)[8]) nullptr);
      const uint32_t pos = (h & 1u) ? qw.pi : (uint32_t)t;
      e[pos] = Lu[1] - Lu[0] - sa;
      if (j > L.g) qw.prev();
    }
    if (t < (int)L.K) {
      R o[8];
      beta_step(b, R(0.5) * sa, R(0.5) * pr, o);
#pragma unroll
      for (int s = 0; s < 8; ++s) b[s] = o[s];
    }
  }
}

template <typename R>
__device__ inline void load8(const Lane<R>& L, uint32_t field, uint32_t stride, uint32_t j, R (&v)[8]) {
#pragma unroll
  for (int s = 0; s < 8; ++s) v[s] = L.at(field + s * stride, j);
}

// alpha over local steps 0 .. g + S with only every 8th stored: al[c] =
// alpha at local 8c (the same recursion as alpha_pass).
template <typename R>
__device__ void alpha_ckpt_pass(const Lane<R>& L, const Layout& Y) {
  R a[8];
  const bool from_zero_state = (int)(L.q * L.S) - (int)L.g <= 0;
#pragma unroll
  for (int s = 0; s < 8; ++s) a[s] = from_zero_state ? (s == 0 ? R(0) : R(kNegF)) : R(0);
  const uint32_t n = L.g + L.S;
  R sa_n = L.at(Y.sa, 0), pr_n = L.at(Y.pr, 0);  // the next step's inputs, one step ahead
  for (uint32_t j = 0; j < n; ++j) {
    const R sa = sa_n, pr = pr_n;
    if (j + 1 < n) {
      sa_n = L.at(Y.sa, j + 1);
      pr_n = L.at(Y.pr, j + 1);
    }
    if ((j & 7u) == 0) {
#pragma unroll
      for (int s = 0; s < 8; ++s) L.at(Y.al + s * (Y.g + Y.S + 1), j / 8) = a[s];
    }
    if (L.t_of(j) >= 0) {
      R o[8];
      alpha_step(a, R(0.5) * sa, R(0.5) * pr, o);
#pragma unroll
      for (int s = 0; s < 8; ++s) a[s] = o[s];
    }
  }
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
  return L;
}

}  // namespace

template <typename R>
__global__ __launch_bounds__(64) void diff_forward_kernel(DiffArgs<R> A) {
  const Layout Y(A.max_seg, A.max_warm);
  for (uint32_t jb = blockIdx.x; jb < A.njobs; jb += gridDim.x) {
    const DiffJob J = A.jobs[jb];
    const Lane<R> L = make_lane(J, A);
    const DiffCb C = A.cbs[J.cb_first + (L.active ? L.lane / J.nseg : 0)];
    const R* d0 = A.llr + C.llr_off;
    const R* d1 = d0 + (L.K + 4);
    const R* d2 = d1 + (L.K + 4);
    R* sv = A.saved + C.save_off;
    for (uint32_t h = 0; h < A.halves; ++h) {
      if (L.active) {
        R ts[3], tp[3];
        tail_inputs(h, d0, d1, d2, L.K, ts, tp);
        stage(L, Y, h, d0, d1, d2, sv, A.scales);
        alpha_pass(L, Y);
        beta_ext_pass(L, Y, ts, tp, h, sv + (size_t)h * L.K);
      }
      __syncthreads();
    }
    if (L.active) {
      const uint32_t H = A.halves;
      const R* e1 = sv + (size_t)(H - 2) * L.K;
      const R* e2 = sv + (size_t)(H - 1) * L.K;
      R* post = A.post + C.post_off;
      const R w1 = A.scales[H - 2], w2 = A.scales[H - 1];
      for (uint32_t t = L.q * L.S; t < L.q * L.S + L.S; ++t) post[t] = d0[t] + w1 * e1[t] + w2 * e2[t];
    }
    __syncthreads();
  }
}

template <typename R>
__global__ __launch_bounds__(64) void diff_backward_kernel(DiffArgs<R> A) {
  const Layout Y(A.max_seg, A.max_warm);
  const uint32_t H = A.halves;
  for (uint32_t jb = blockIdx.x; jb < A.njobs; jb += gridDim.x) {
    const DiffJob J = A.jobs[jb];
    const Lane<R> L = make_lane(J, A);
    const uint32_t blk = L.active ? L.lane / J.nseg : 0;
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
    // posterior: post = d0 + w[H-2] e_{H-2} + w[H-1] e_{H-1}
    {
      R s1 = 0, s2 = 0;
      if (L.active) {
        const R* e1 = sv + (size_t)(H - 2) * K;
        const R* e2 = sv + (size_t)(H - 1) * K;
        for (uint32_t t = lo; t < hi; ++t) {
          const R gpt = gp[t];
          g0[t] = gpt;
          g1[t] = 0;
          g2[t] = 0;
          eb[(H - 1) & 1u][t] = A.scales[H - 1] * gpt;
          s1 += gpt * e1[t];
          s2 += gpt * e2[t];
        }
        if (L.q == L.M - 1)
          for (uint32_t t = K; t < D; ++t) g0[t] = g1[t] = g2[t] = 0;
      }
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
      R ts[3] = {0, 0, 0}, tp[3] = {0, 0, 0};
      bool to_tail = false;
      if (L.active) {
        // adjoint of e_{h-1}: starts at the posterior term for h - 1 == H - 2
        if (h >= 1)
          for (uint32_t t = lo; t < hi; ++t) prev[t] = (h - 1 == H - 2) ? A.scales[H - 2] * gp[t] : R(0);
        tail_inputs(h, d0, d1, d2, K, ts, tp);
        stage(L, Y, h, d0, d1, d2, sv, A.scales);
        {
          QppWalk qw(lo, K, L.f1, L.f2);
#pragma unroll 8
          for (uint32_t j = L.g; j < L.g + L.S; ++j, qw.next()) {
            const uint32_t t = (uint32_t)L.t_of(j);
            L.at(Y.eb, j - L.g) = cur[(h & 1u) ? qw.pi : t];
          }
        }
        // the alpha-adjoint sweep is the first writer of local steps below
        // g + S; the beta-adjoint sweep adds to those above
        for (uint32_t j = L.g + L.S; j < Y.P; ++j) L.at(Y.dsa, j) = L.at(Y.dpr, j) = 0;
        // alpha at every 8th local step only (checkpoints in the al field,
        // al[c] = alpha at local 8c); the adjoint sweeps recompute each
        // sub-window's metrics into registers, which keeps the wave's scratch
        // small enough to stay in the caches
        alpha_ckpt_pass(L, Y);
        // beta over the window's warm-up above the own steps (local 2g + S - 1
        // down to g + S); the own steps' beta runs inside the alpha-adjoint
        // loop below, an independent chain beside the adjoint recursion.
        // Checkpoints for the beta-adjoint sweep: beta at local g + i for i a
        // multiple of 8 or i = S + g, at be[(i + 7) / 8]
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
        R wsa = L.at(Y.sa, 2 * L.g + L.S - 1), wpr = L.at(Y.pr, 2 * L.g + L.S - 1);
        for (uint32_t j = 2 * L.g + L.S; j-- > L.g + L.S;) {
          const R sa = wsa, pr = wpr;
          if (j > L.g + L.S) {
            wsa = L.at(Y.sa, j - 1);
            wpr = L.at(Y.pr, j - 1);
          }
          beta_ck(j + 1 - L.g, bcur);
          if (L.t_of(j) < (int)K) {
            R o[8];
            beta_step(bcur, R(0.5) * sa, R(0.5) * pr, o);
#pragma unroll
            for (int s = 0; s < 8; ++s) bcur[s] = o[s];
          }
        }
        // alpha adjoint, backwards over local steps g + S - 1 .. 0 in
        // sub-windows of 8 (alpha recomputed from the checkpoint); output
        // adjoints of the own steps on the way
        R ab[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint32_t n_al = L.g + L.S;
        for (int c = (int)((n_al + 7) / 8) - 1; c >= 0; --c) {
          const uint32_t j0 = 8u * (uint32_t)c;
          const uint32_t len = min(8u, n_al - j0);
          if (L.t_of(j0 + len - 1) < 0) break;  // segment 0: nothing before t = 0
          R areg[8][8], hsr[8], hpr[8], er[8];
          {
            R a[8];
            load8(L, Y.al, Y.g + Y.S + 1, (uint32_t)c, a);
#pragma unroll
            for (uint32_t jj = 0; jj < 8; ++jj) {
              if (jj < len) {
                const uint32_t j = j0 + jj;
                hsr[jj] = R(0.5) * L.at(Y.sa, j);
                hpr[jj] = R(0.5) * L.at(Y.pr, j);
                er[jj] = (j >= L.g) ? L.at(Y.eb, j - L.g) : R(0);
#pragma unroll
                for (int s = 0; s < 8; ++s) areg[jj][s] = a[s];
                if (jj + 1 < len && L.t_of(j) >= 0) {
                  R o[8];
                  alpha_step(a, hsr[jj], hpr[jj], o);
#pragma unroll
                  for (int s = 0; s < 8; ++s) a[s] = o[s];
                }
              }
            }
          }
#pragma unroll
          for (int jj = 7; jj >= 0; --jj) {
            const uint32_t j = j0 + (uint32_t)jj;
            if ((uint32_t)jj >= len || L.t_of(j) < 0) continue;
            R b[8];
#pragma unroll
            for (int s = 0; s < 8; ++s) b[s] = bcur[s];  // beta_{j+1} while j >= g
            const R hs = hsr[jj], hp = hpr[jj], e = er[jj];
            if (j >= L.g) {
              // own step: checkpoint beta_{j+1}, step beta to beta_j
              beta_ck(j + 1 - L.g, b);
              R o[8];
              beta_step(b, hs, hp, o);
#pragma unroll
              for (int s = 0; s < 8; ++s) bcur[s] = o[s];
            }
            R ds, dp;
            alpha_adjoint_step(areg[jj], b, hs, hp, e, j >= L.g, ab, ds, dp, L, Y, j);
            L.at(Y.dsa, j) = ds;  // first writer of these positions (all zero before)
            L.at(Y.dpr, j) = dp;
          }
        }
        // beta adjoint, forwards over local steps g .. g + S + g - 1 in
        // sub-windows of 8 (beta recomputed from the checkpoint at the top)
        R bb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const uint32_t jend = 2 * L.g + L.S;
        for (uint32_t jb0 = L.g; jb0 < jend; jb0 += 8) {
          if (L.t_of(jb0) >= (int)K) break;
          const uint32_t len = min(8u, jend - jb0);
          const uint32_t itop = jb0 + len - L.g;  // beta at local jb0 + len = g + itop
          R breg[8][8], hsr[8], hpr[8];
          {
            R b[8];
            load8(L, Y.be, bst, (itop + 7) / 8, b);
#pragma unroll
            for (int jj = 7; jj >= 0; --jj) {
              if ((uint32_t)jj < len) {
                const uint32_t j = jb0 + (uint32_t)jj;
                hsr[jj] = R(0.5) * L.at(Y.sa, j);
                hpr[jj] = R(0.5) * L.at(Y.pr, j);
#pragma unroll
                for (int s = 0; s < 8; ++s) breg[jj][s] = b[s];  // beta at local j + 1
                if (jj > 0 && L.t_of(j) < (int)K) {
                  R o[8];
                  beta_step(b, hsr[jj], hpr[jj], o);
#pragma unroll
                  for (int s = 0; s < 8; ++s) b[s] = o[s];
                }
              }
            }
          }
          // the sub-window's output adjoints and running sys/par adjoints, all
          // loaded before its first step
          R bor[8][8], dsr[8], dpr[8];
#pragma unroll
          for (uint32_t jj = 0; jj < 8; ++jj) {
            const uint32_t j = jb0 + jj;
            if (jj < len) {
              if (j < L.g + L.S) {
                load8(L, Y.bo, Y.S, j - L.g, bor[jj]);
              } else {
#pragma unroll
                for (int s = 0; s < 8; ++s) bor[jj][s] = 0;
              }
              dsr[jj] = L.at(Y.dsa, j);
              dpr[jj] = L.at(Y.dpr, j);
            }
          }
#pragma unroll
          for (uint32_t jj = 0; jj < 8; ++jj) {
            const uint32_t j = jb0 + jj;
            if (jj >= len || L.t_of(j) >= (int)K) break;
            R ds, dp;
            beta_adjoint_step(breg[jj], hsr[jj], hpr[jj], bb, bor[jj], ds, dp);
            L.at(Y.dsa, j) = dsr[jj] + ds;
            L.at(Y.dpr, j) = dpr[jj] + dp;
          }
        }
        // beta of the window top: the tail beta when it reaches K
        to_tail = lo + L.S + L.g >= K;
        R dts[3] = {0, 0, 0}, dtp[3] = {0, 0, 0};
        if (to_tail) {
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
      // combine: own window, next segment's alpha warm-up, previous segment's beta warm-up
      R wsum = 0;
      if (L.active) {
        const R w = h ? A.scales[h - 1] : R(0);
        const R* eprev = sv + (size_t)(h ? h - 1 : 0) * K;
        QppWalk qw(lo, K, L.f1, L.f2);
#pragma unroll 4
        for (uint32_t j = L.g; j < L.g + L.S; ++j, qw.next()) {
          const uint32_t t = (uint32_t)L.t_of(j);
          R ds = L.at(Y.dsa, j), dp = L.at(Y.dpr, j);
          if (L.q + 1 < L.M && j >= L.S) {  // t in [qS + S - g, qS + S): lane q+1's local j - S
            ds += L.nb(Y.dsa, j - L.S, L.lane + 1);
            dp += L.nb(Y.dpr, j - L.S, L.lane + 1);
          }
          if (L.q > 0 && j < 2 * L.g) {  // t in [qS, qS + g): lane q-1's local j + S
            ds += L.nb(Y.dsa, j + L.S, L.lane - 1);
            dp += L.nb(Y.dpr, j + L.S, L.lane - 1);
          }
          if ((h & 1u) == 0) {
            g0[t] += ds;
            g1[t] += dp;
            if (h) {
              prev[t] += w * ds;
              wsum += ds * eprev[t];
            }
          } else {
            const uint32_t pi = qw.pi;
            g0[pi] += ds;
            g2[t] += dp;
            prev[pi] += w * ds;
            wsum += ds * eprev[pi];
          }
        }
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
          if ((h & 1u) == 0) {
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

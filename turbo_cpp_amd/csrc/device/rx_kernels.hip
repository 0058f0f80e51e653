// Receive-chain kernels: soft de-rate matching with HARQ combining, and
// transport-block assembly.
//
// Soft de-rate matching restates lte_turbo.cpp:280-325 for float LLRs.  The
// reference walks the circular buffer from k0(rv), skips NULL entries, adds
// +-8 per received hard bit onto the entry, then copies every entry to its
// stream position.  Since the non-NULL entries are visited in a fixed cyclic
// order, received value k lands on rank (start[rv] + k) mod n_valid; a
// workgroup per code block therefore
//   1. sums the received values per rank in k order into LDS (coalesced reads;
//      more than one term only when E > n_valid, i.e. repetition), and
//   2. writes each [d0|d1|d2] position from the LDS entry of its rank
//      (coalesced writes, one gather from LDS).
// The summation order (0 + first + second + ...) is the reference's, so the
// output is bit-identical to it for +-8 inputs.
#include <hip/hip_runtime.h>

#include "rx_types.h"

namespace tcdev {

__global__ __launch_bounds__(kDerateThreads) void derate_soft_kernel(const DerateCb* __restrict__ cbs,
                                                                     const uint16_t* __restrict__ rank,
                                                                     const float* __restrict__ rx,
                                                                     float* __restrict__ out, int32_t rv_override,
                                                                     int32_t accumulate) {
  extern __shared__ float acc[];
  const DerateCb c = cbs[blockIdx.x];
  const float* in = rx + c.in_off;
  const uint32_t nv = c.n_valid, e = c.e;
  constexpr uint32_t U = 4;  // loads in flight per thread
  if (e <= nv) {
    for (uint32_t base = 0; base < e; base += U * kDerateThreads) {
      float v[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) v[u] = in[min(base + u * kDerateThreads + threadIdx.x, e - 1)];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t m = base + u * kDerateThreads + threadIdx.x;
        if (m < e) acc[m] = 0.0f + v[u];
      }
    }
  } else {
    for (uint32_t m = threadIdx.x; m < nv; m += kDerateThreads) {
      float v = 0.0f;
      for (uint32_t k = m; k < e; k += nv) v += in[k];
      acc[m] = v;
    }
  }
  __syncthreads();
  const uint32_t rv = (rv_override >= 0 ? static_cast<uint32_t>(rv_override) : c.rv) & 3u;
  const uint32_t n0 = cbs[blockIdx.x].start[rv];  // indexed in global memory: no private copy
  const uint16_t* rk = rank + c.rank_off;
  float* o = out + c.out_off;
  const uint32_t n = 3 * c.d;
  for (uint32_t base = 0; base < n; base += U * kDerateThreads) {
    uint32_t r[U];
    float prev[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      // unconditional loads at clamped indices (the i >= n slots are skipped
      // below): a load under a per-lane condition becomes a branch and a wait
      const uint32_t i = min(base + u * kDerateThreads + threadIdx.x, n - 1);
      r[u] = rk[i];
      prev[u] = accumulate ? o[i] : 0.0f;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t i = base + u * kDerateThreads + threadIdx.x;
      if (i >= n) continue;
      const uint32_t m = (r[u] >= n0) ? r[u] - n0 : r[u] + nv - n0;
      const bool seen = (r[u] != kNullRank) && m < e;
      if (accumulate) {
        if (seen) o[i] = prev[u] + acc[m];
      } else {
        o[i] = seen ? acc[m] : 0.0f;
      }
    }
  }
}

// The received values of one transmission summed per rank in k order into
// acc[0 .. min(e, nv)) (derate_soft_kernel's first step).
__device__ __forceinline__ void derate_fill(const float* __restrict__ in, uint32_t e, uint32_t nv, float* acc) {
  constexpr uint32_t U = 4;
  if (e <= nv) {
    for (uint32_t base = 0; base < e; base += U * kDerateThreads) {
      float v[U];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) v[u] = in[min(base + u * kDerateThreads + threadIdx.x, e - 1)];
#pragma unroll
      for (uint32_t u = 0; u < U; ++u) {
        const uint32_t m = base + u * kDerateThreads + threadIdx.x;
        if (m < e) acc[m] = 0.0f + v[u];
      }
    }
  } else {
    for (uint32_t m = threadIdx.x; m < nv; m += kDerateThreads) {
      float v = 0.0f;
      for (uint32_t k = m; k < e; k += nv) v += in[k];
      acc[m] = v;
    }
  }
}

// Two transmissions (HARQ combining) in one pass: what derate_soft_kernel
// leaves after a run on rx_a (accumulating or not) followed by an
// accumulating run on rx_b, bit for bit (each position: the first value, or
// the held one plus it, then plus the second), with one write of the output
// instead of a write, a read and a second write.  LDS: the two per-rank sums.
__global__ __launch_bounds__(kDerateThreads) void derate_soft2_kernel(const DerateCb* __restrict__ cbs,
                                                                      const uint16_t* __restrict__ rank,
                                                                      const float* __restrict__ rx_a, int32_t rv_a,
                                                                      const float* __restrict__ rx_b, int32_t rv_b,
                                                                      float* __restrict__ out, int32_t accumulate) {
  extern __shared__ float acc2[];
  const DerateCb c = cbs[blockIdx.x];
  const uint32_t nv = c.n_valid, e = c.e;
  const uint32_t ne = e < nv ? e : nv;
  derate_fill(rx_a + c.in_off, e, nv, acc2);
  derate_fill(rx_b + c.in_off, e, nv, acc2 + ne);
  __syncthreads();
  const uint32_t ra = (rv_a >= 0 ? static_cast<uint32_t>(rv_a) : c.rv) & 3u;
  const uint32_t rb = (rv_b >= 0 ? static_cast<uint32_t>(rv_b) : c.rv) & 3u;
  const uint32_t na = cbs[blockIdx.x].start[ra], nb = cbs[blockIdx.x].start[rb];
  const uint16_t* rk = rank + c.rank_off;
  float* o = out + c.out_off;
  const uint32_t n = 3 * c.d;
  constexpr uint32_t U = 4;
  for (uint32_t base = 0; base < n; base += U * kDerateThreads) {
    uint32_t r[U];
    float prev[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      // unconditional loads at clamped indices (the i >= n slots are skipped
      // below): a load under a per-lane condition becomes a branch and a wait
      const uint32_t i = min(base + u * kDerateThreads + threadIdx.x, n - 1);
      r[u] = rk[i];
      prev[u] = accumulate ? o[i] : 0.0f;
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t i = base + u * kDerateThreads + threadIdx.x;
      if (i >= n) continue;
      const bool valid = r[u] != kNullRank;
      const uint32_t ma = (r[u] >= na) ? r[u] - na : r[u] + nv - na;
      const uint32_t mb = (r[u] >= nb) ? r[u] - nb : r[u] + nv - nb;
      const bool sa = valid && ma < e, sb = valid && mb < e;
      float v;
      if (accumulate) {
        v = sa ? prev[u] + acc2[ma] : prev[u];
      } else {
        v = sa ? acc2[ma] : 0.0f;
      }
      if (sb) v = v + acc2[ne + mb];
      if (accumulate && !sa && !sb) continue;  // untouched, as in the two runs
      o[i] = v;
    }
  }
}

// XOR over the workgroup (kTbThreads lanes); every thread gets the result.
__device__ uint32_t block_xor(uint32_t x, uint32_t* red) {
  for (int o = 32; o > 0; o >>= 1) x ^= __shfl_xor(x, o, 64);
  const uint32_t wave = threadIdx.x >> 6;
  __syncthreads();  // red[] free from the previous reduction
  if ((threadIdx.x & 63) == 0) red[wave] = x;
  __syncthreads();
  uint32_t y = 0;
  for (int w = 0; w < kTbThreads / 64; ++w) y ^= red[w];
  return y;
}

// Transport-block assembly (lte_turbo.cpp:958-979), one workgroup per TB:
// hard decision LLR >= 0 -> 1 (:488-494), filler bits dropped, CRC24B over
// each block when c > 1 (its 24 parity bits then dropped), CRC24A over the
// concatenation, payload = all but the last 24 bits, iterations = max over
// the blocks.  A CRC is checked as the remainder of the whole sequence
// a_0 .. a_{n-1} (a_0 the highest power): XOR of D^(n-1-i) mod G over the
// ones, from the tables pow24a / pow24b; zero remainder = pass, as the
// reference's long division (:137-155).
__global__ __launch_bounds__(kTbThreads) void tb_assemble_kernel(
    const TbInfo* __restrict__ tbs, const TbCb* __restrict__ cbs, const float* __restrict__ post,
    const uint32_t* __restrict__ cb_iters, uint32_t max_iterations, const uint32_t* __restrict__ pow24a,
    const uint32_t* __restrict__ pow24b, uint8_t* __restrict__ bits, int32_t* __restrict__ tb_ok,
    int32_t* __restrict__ cb_ok, uint32_t* __restrict__ tb_iters) {
  __shared__ uint32_t red[kTbThreads / 64];
  const TbInfo t = tbs[blockIdx.x];
  const uint32_t len = t.b + 24;  // bits under the CRC24A
  const bool seg = t.c > 1 || t.cb_crc;  // code blocks carry a CRC24B
  uint32_t tb_x = 0, pos = 0, it_max = 0;
  for (uint32_t r = 0; r < t.c; ++r) {
    const TbCb cb = cbs[t.cb_first + r];
    const uint32_t n = cb.k - cb.filler;     // after the filler strip
    const uint32_t np = seg ? n - 24 : n;    // bits that go to the TB
    const float* p = post + cb.post_off + cb.filler;
    uint32_t cb_x = 0;
    for (uint32_t i = threadIdx.x; i < n; i += kTbThreads) {
      const bool one = p[i] >= 0.0f;
      if (one && seg) cb_x ^= pow24b[n - 1 - i];
      if (i < np) {
        const uint32_t q = pos + i;
        if (one) tb_x ^= pow24a[len - 1 - q];
        if (bits && q < t.b) bits[t.bits_off + q] = one ? 1 : 0;
      }
    }
    if (seg) cb_x = block_xor(cb_x, red);
    if (t.cb_crc) tb_x = (cb_x == 0 && n >= 24) ? 0u : 1u;  // the block's CRC24B is the TB's check
    if (cb_ok && threadIdx.x == 0) cb_ok[t.cb_first + r] = seg ? (cb_x == 0 && n >= 24) : 1;
    if (cb_iters) it_max = max(it_max, cb_iters[t.cb_first + r]);
    pos += np;
  }
  if (!t.cb_crc) tb_x = block_xor(tb_x, red);
  if (threadIdx.x == 0) {
    if (tb_ok) tb_ok[blockIdx.x] = (tb_x == 0 && (pos == len || t.cb_crc)) ? 1 : 0;
    if (tb_iters) tb_iters[blockIdx.x] = max_iterations ? it_max : 0;
  }
}

// Clears the per-TB accumulators of the decoder-epilogue assembly (a kernel,
// so a captured graph records no memset).
__global__ void tb_acc_zero_kernel(uint32_t* acc, uint32_t n) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) acc[i] = 0u;
}

// TB verdicts of the decoder-epilogue assembly, after the decode: every block
// of TB i has XORed its CRC24A share (or, for a code-block-CRC descriptor,
// its CRC24B verdict) into acc[3i] and max-ed its iteration count into
// acc[3i + 2].
__global__ void tb_finalize_kernel(const uint32_t* acc, uint32_t n_tb, uint32_t max_iterations, int32_t* tb_ok,
                                   uint32_t* tb_iters) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_tb) return;
  if (tb_ok) tb_ok[i] = acc[3 * i] == 0u ? 1 : 0;
  if (tb_iters) tb_iters[i] = max_iterations ? acc[3 * i + 2] : 0u;
}

}  // namespace tcdev

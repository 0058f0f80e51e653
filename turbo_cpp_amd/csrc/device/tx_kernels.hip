// Transmit-chain kernels (data generation at scale, bit-exact with the host
// encoder): CRC24A per transport block, code-block construction with CRC24B
// and the turbo encoder in one wave per block, rate matching.
//
// Encoder: the RSC state update is linear over GF(2): next = A s + b u and
// parity = u + s0 + s1 (feedback u ^ s1 ^ s2, lte_turbo.cpp:333-341).  Each
// of the 64 lanes runs its chunk of the block from state 0 (zero-state
// response e_j) and records A^L_j (its chunk's zero-input map, as the images
// of the three basis states); lane 0 chains s_{j+1} = A^L_j s_j + e_j over
// the 64 chunks; each lane then reruns its chunk from s_j and writes the
// parities into LDS, from where the block's streams leave coalesced.
#include <hip/hip_runtime.h>

#include "rx_types.h"

namespace tcdev {
namespace {

__device__ inline unsigned rsc_next(unsigned s, unsigned u) {
  const unsigned fb = (u ^ (s >> 1) ^ (s >> 2)) & 1u;
  return (fb << 2) | (((s >> 2) & 1u) << 1) | ((s >> 1) & 1u);
}
__device__ inline unsigned rsc_par(unsigned s, unsigned u) { return (u ^ s ^ (s >> 1)) & 1u; }
__device__ inline unsigned term_input(unsigned s) { return ((s >> 1) ^ (s >> 2)) & 1u; }

__device__ uint32_t wave_xor(uint32_t x) {
  for (int o = 32; o > 0; o >>= 1) x ^= __shfl_xor(x, o, 64);
  return x;
}

// GF(2) 3x3 map given by the images of states 1, 2, 4.
__device__ inline unsigned apply(unsigned m1, unsigned m2, unsigned m4, unsigned s) {
  return ((s & 1u) ? m1 : 0u) ^ ((s & 2u) ? m2 : 0u) ^ ((s & 4u) ? m4 : 0u);
}

constexpr int kMaxK = 6144;

}  // namespace

// CRC24A of every transport block's payload (lte_turbo.cpp:118-135 with the
// :104-107 generator): remainder of P(D) D^24, as the XOR over the ones at
// position i of D^(B - 1 - i + 24) mod G.  One 256-thread workgroup per TB.
__global__ __launch_bounds__(256) void tb_crc24a_kernel(const TbInfo* __restrict__ tbs,
                                                        const uint8_t* __restrict__ payload,
                                                        const uint32_t* __restrict__ pow24a,
                                                        uint32_t* __restrict__ crc) {
  __shared__ uint32_t red[4];
  const TbInfo t = tbs[blockIdx.x];
  uint32_t x = 0;
  for (uint32_t i = threadIdx.x; i < t.b; i += 256)
    if (payload[t.bits_off + i]) x ^= pow24a[t.b - 1 - i + 24];
  x = wave_xor(x);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = x;
  __syncthreads();
  if (threadIdx.x == 0) crc[blockIdx.x] = red[0] ^ red[1] ^ red[2] ^ red[3];
}

// One wave per code block: build its bits (filler zeros, payload slice of
// the TB with CRC24A, CRC24B when segmented; lte_turbo.cpp:540-604) and
// encode them (lte_turbo.cpp:624-688, tail multiplexing of :664-677).
// Outputs: [d0|d1|d2] bytes (K + 4 each) and, if cb_bits is given, the
// code-block bits.
__global__ __launch_bounds__(64) void cb_encode_kernel(const TxCb* __restrict__ cbs, const TbInfo* __restrict__ tbs,
                                                       const uint8_t* __restrict__ payload,
                                                       const uint32_t* __restrict__ tb_crc,
                                                       const uint32_t* __restrict__ pow24b,
                                                       uint8_t* __restrict__ cb_bits, uint8_t* __restrict__ dst) {
  __shared__ uint8_t c[kMaxK];
  __shared__ uint8_t out[3][kMaxK + 4];
  __shared__ uint32_t st[2][65];            // chunk start states, upper / lower
  __shared__ uint32_t ze[2][64], mp[64][3]; // zero-state end states, chunk maps
  const TxCb b = cbs[blockIdx.x];
  const TbInfo t = tbs[b.tb];
  const uint32_t lane = threadIdx.x, K = b.k;
  const uint32_t crc_a = tb_crc[b.tb];
  // code-block bits
  uint32_t x = 0;
  for (uint32_t j = lane; j < K; j += 64) {
    uint8_t v = 0;
    if (j >= b.filler && j < b.filler + b.payload) {
      const uint32_t y = b.cursor + (j - b.filler);  // position in the TB with its CRC24A
      v = (y < t.b) ? payload[t.bits_off + y] : (uint8_t)((crc_a >> (23 - (y - t.b))) & 1u);
      if (v && b.seg) x ^= pow24b[b.payload - 1 - (j - b.filler) + 24];
    }
    c[j] = v;
  }
  if (b.seg) {
    x = wave_xor(x);
    for (uint32_t j = lane; j < 24; j += 64) c[b.filler + b.payload + j] = (uint8_t)((x >> (23 - j)) & 1u);
  }
  __syncthreads();
  if (cb_bits)
    for (uint32_t j = lane; j < K; j += 64) cb_bits[b.cb_off + j] = c[j];
  // chunks of L steps; lane j covers [jL, min(K, jL + L))
  const uint32_t L = (K + 63) / 64;
  const uint32_t lo = min(K, lane * L), hi = min(K, lo + L);
  // Pi at the chunk start: (f1 i + f2 i^2) mod K, then incremental
  auto pi_at = [&](uint32_t i) { return (uint32_t)(((uint64_t)b.f1 * i + (uint64_t)b.f2 * i % K * i) % K); };
  {
    unsigned su = 0, sl = 0, m1 = 1, m2 = 2, m4 = 4;
    uint32_t pi = pi_at(lo), dl = (uint32_t)(((uint64_t)b.f1 + (uint64_t)b.f2 * (2ull * lo + 1)) % K);
    for (uint32_t i = lo; i < hi; ++i) {
      su = rsc_next(su, c[i]);
      sl = rsc_next(sl, c[pi]);
      m1 = rsc_next(m1, 0);
      m2 = rsc_next(m2, 0);
      m4 = rsc_next(m4, 0);
      pi += dl;
      if (pi >= K) pi -= K;
      dl += 2 * b.f2 % K;
      if (dl >= K) dl -= K;
    }
    ze[0][lane] = su;
    ze[1][lane] = sl;
    mp[lane][0] = m1;
    mp[lane][1] = m2;
    mp[lane][2] = m4;
  }
  __syncthreads();
  if (lane == 0) {
    unsigned su = 0, sl = 0;
    for (int j = 0; j < 64; ++j) {
      st[0][j] = su;
      st[1][j] = sl;
      su = apply(mp[j][0], mp[j][1], mp[j][2], su) ^ ze[0][j];
      sl = apply(mp[j][0], mp[j][1], mp[j][2], sl) ^ ze[1][j];
    }
    st[0][64] = su;
    st[1][64] = sl;
  }
  __syncthreads();
  {
    unsigned su = st[0][lane], sl = st[1][lane];
    uint32_t pi = pi_at(lo), dl = (uint32_t)(((uint64_t)b.f1 + (uint64_t)b.f2 * (2ull * lo + 1)) % K);
    for (uint32_t i = lo; i < hi; ++i) {
      const unsigned u = c[i], ul = c[pi];
      const bool fill = i < b.filler;
      out[0][i] = fill ? 0 : (uint8_t)u;
      out[1][i] = fill ? 0 : (uint8_t)rsc_par(su, u);
      out[2][i] = (uint8_t)rsc_par(sl, ul);
      su = rsc_next(su, u);
      sl = rsc_next(sl, ul);
      pi += dl;
      if (pi >= K) pi -= K;
      dl += 2 * b.f2 % K;
      if (dl >= K) dl -= K;
    }
  }
  if (lane == 0) {
    // termination (lte_turbo.cpp:343-353) and tail multiplexing (:664-677)
    unsigned su = st[0][64], sl = st[1][64];
    uint8_t xu[3], zu[3], xl[3], zl[3];
    for (int i = 0; i < 3; ++i) {
      xu[i] = (uint8_t)term_input(su);
      xl[i] = (uint8_t)term_input(sl);
      zu[i] = (uint8_t)rsc_par(su, xu[i]);
      zl[i] = (uint8_t)rsc_par(sl, xl[i]);
      su = rsc_next(su, xu[i]);
      sl = rsc_next(sl, xl[i]);
    }
    const uint8_t t0[4] = {xu[0], zu[1], xu[1], zu[2]};
    const uint8_t t1[4] = {zu[0], xu[1], zu[2], xl[0]};
    const uint8_t t2[4] = {zl[0], xl[1], zl[1], xl[2]};
    for (int i = 0; i < 4; ++i) {
      out[0][K + i] = t0[i];
      out[1][K + i] = t1[i];
      out[2][K + i] = t2[i];
    }
  }
  __syncthreads();
  uint8_t* d = dst + b.d_off;
  for (uint32_t s = 0; s < 3; ++s)
    for (uint32_t i = lane; i < K + 4; i += 64) d[s * (K + 4) + i] = out[s][i];
}

// Rate matching (lte_turbo.cpp:250-278): output bit k of a block is the
// stream bit at the circular-buffer entry of rank (start[rv] + k) mod
// n_valid; valid[] lists the non-NULL entries (stream << 16 | index) in
// circular order.  One 256-thread workgroup per block.
__global__ __launch_bounds__(256) void rate_match_kernel(const TxCb* __restrict__ cbs,
                                                         const uint32_t* __restrict__ valid,
                                                         const uint8_t* __restrict__ dsrc, int32_t rv_override,
                                                         uint8_t* __restrict__ rm) {
  const TxCb b = cbs[blockIdx.x];
  const uint32_t rv = (rv_override >= 0 ? (uint32_t)rv_override : b.rv) & 3u;
  const uint32_t n0 = cbs[blockIdx.x].start[rv], nv = b.n_valid, D = b.k + 4;  // no private copy
  const uint32_t* v = valid + b.valid_off;
  const uint8_t* d = dsrc + b.d_off;
  for (uint32_t k = threadIdx.x; k < b.e; k += 256) {
    const uint32_t src = v[(uint32_t)(((uint64_t)n0 + k) % nv)];
    rm[b.rm_off + k] = d[(src >> 16) * D + (src & 0xFFFFu)];
  }
}

}  // namespace tcdev

// Channel front end on the GPU (not in the reference; data generation and
// receiver front end for the decoder): Gray QAM mapping of 3GPP TS 36.211
// 7.1 (QPSK, 16QAM, 64QAM, unit average energy), AWGN or i.i.d. Rayleigh
// fading known at the receiver, and the max-log demapper.  One thread per
// symbol; the fused kernel reads coded bits and writes LLRs (bit 1 > 0),
// nothing in between touches HBM.
//
// Demapping: z = conj(h) y / |h|^2; since |y - h s|^2 = |h|^2 |z - s|^2 + c,
// the max-log LLR of a bit carried by the I (Q) dimension is
//   |h|^2 / N0 * (min_{a: b=0} (z_I - a)^2 - min_{a: b=1} (z_I - a)^2)
// over that dimension's 2^(Qm/2) levels (bits b0, b2, b4 on I; b1, b3, b5 on Q).
// Noise: a counter-based generator (the SplitMix64 finaliser of Steele, Lea
// and Flood 2014 applied to seed and symbol index) and the Box-Muller
// transform, so a run is reproducible and independent of the launch shape.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "hip_util.h"
#include "turbo_cpp/c_api.h"

namespace tcdev {

struct ChanSeg {
  uint64_t bit_off, sym_off;
  uint32_t n_sym, qm, rayleigh;
  float n0;
};

namespace {

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// two standard normals from 64 random bits
__device__ inline void normals(uint64_t r, float& a, float& b) {
  const float u1 = (float)((r >> 40) + 1) * 5.9604644775390625e-8f;  // (0, 1]
  const float u2 = (float)(r & 0xFFFFFFull) * 5.9604644775390625e-8f;  // [0, 1)
  const float rad = sqrtf(-2.0f * logf(u1));
  float s, c;
  sincosf(6.283185307179586f * u2, &s, &c);
  a = rad * c;
  b = rad * s;
}

__device__ inline float scale_of(uint32_t qm) {
  return qm == 2 ? 0.70710678118654752f : (qm == 4 ? 0.31622776601683794f : 0.15430334996209191f);
}

// PAM level (unscaled) of one dimension from its bits, first bit = sign.
__device__ inline float pam(uint32_t nb, uint32_t b0, uint32_t b1, uint32_t b2) {
  const float s = b0 ? -1.0f : 1.0f;
  if (nb == 1) return s;
  if (nb == 2) return s * (b1 ? 3.0f : 1.0f);
  return s * (4.0f - (b1 ? -1.0f : 1.0f) * (2.0f - (b2 ? -1.0f : 1.0f)));
}

// LLRs of the nb bits of one dimension at equalised value z, weight w = |h|^2/N0.
__device__ inline void demap_dim(float z, float w, uint32_t nb, float scale, float* out, uint32_t stride) {
  const uint32_t n = 1u << nb;
  for (uint32_t j = 0; j < nb; ++j) {
    float d0 = 3.0e38f, d1 = 3.0e38f;
    for (uint32_t v = 0; v < n; ++v) {
      const uint32_t b0 = (v >> (nb - 1)) & 1u, b1 = nb > 1 ? (v >> (nb - 2)) & 1u : 0u, b2 = nb > 2 ? v & 1u : 0u;
      const float a = scale * pam(nb, b0, b1, b2);
      const float d = (z - a) * (z - a);
      const uint32_t bit = (v >> (nb - 1 - j)) & 1u;
      if (bit) d1 = fminf(d1, d);
      else d0 = fminf(d0, d);
    }
    out[j * stride] = w * (d0 - d1);
  }
}

__device__ inline void demap_symbol(float yr, float yi, float hr, float hi, float n0, uint32_t qm, float* llr) {
  const float g = hr * hr + hi * hi;
  const float zr = (hr * yr + hi * yi) / g;
  const float zi = (hr * yi - hi * yr) / g;
  const float w = g / n0;
  const uint32_t nb = qm / 2;
  const float sc = scale_of(qm);
  demap_dim(zr, w, nb, sc, llr, 2);      // b0, b2, b4
  demap_dim(zi, w, nb, sc, llr + 1, 2);  // b1, b3, b5
}

__device__ inline uint32_t find_seg(const uint64_t* __restrict__ sym_end, uint32_t n, uint64_t j) {
  uint32_t lo = 0, hi = n;  // first segment with sym_end > j
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (sym_end[mid] > j) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

}  // namespace

__global__ __launch_bounds__(256) void channel_llr_kernel(const ChanSeg* __restrict__ segs,
                                                          const uint64_t* __restrict__ sym_end, uint32_t n_seg,
                                                          uint64_t n_sym, uint64_t seed,
                                                          const uint8_t* __restrict__ bits, float* __restrict__ llr,
                                                          float* __restrict__ y_out, float* __restrict__ h_out) {
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n_sym) return;
  const ChanSeg sg = segs[find_seg(sym_end, n_seg, j)];
  const uint64_t k = j - sg.sym_off;
  const uint8_t* b = bits + sg.bit_off + k * sg.qm;
  const uint32_t nb = sg.qm / 2;
  const float sc = scale_of(sg.qm);
  const float xr = sc * pam(nb, b[0], nb > 1 ? b[2] : 0u, nb > 2 ? b[4] : 0u);
  const float xi = sc * pam(nb, b[1], nb > 1 ? b[3] : 0u, nb > 2 ? b[5] : 0u);
  const uint64_t base = seed * 0x9E3779B97F4A7C15ull + 2 * j;
  float nr, ni, hr = 1.0f, hi = 0.0f;
  normals(mix64(base), nr, ni);
  if (sg.rayleigh) {
    normals(mix64(base + 1), hr, hi);
    hr *= 0.70710678118654752f;
    hi *= 0.70710678118654752f;
  }
  const float sn = sqrtf(0.5f * sg.n0);
  const float yr = hr * xr - hi * xi + sn * nr;
  const float yi = hr * xi + hi * xr + sn * ni;
  if (y_out) {
    y_out[2 * j] = yr;
    y_out[2 * j + 1] = yi;
  }
  if (h_out) {
    h_out[2 * j] = hr;
    h_out[2 * j + 1] = hi;
  }
  demap_symbol(yr, yi, hr, hi, sg.n0, sg.qm, llr + sg.bit_off + k * sg.qm);
}

__global__ __launch_bounds__(256) void qam_demap_kernel(const ChanSeg* __restrict__ segs,
                                                        const uint64_t* __restrict__ sym_end, uint32_t n_seg,
                                                        uint64_t n_sym, const float* __restrict__ y,
                                                        const float* __restrict__ h, float* __restrict__ llr) {
  const uint64_t j = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n_sym) return;
  const ChanSeg sg = segs[find_seg(sym_end, n_seg, j)];
  const uint64_t k = j - sg.sym_off;
  demap_symbol(y[2 * j], y[2 * j + 1], h ? h[2 * j] : 1.0f, h ? h[2 * j + 1] : 0.0f, sg.n0, sg.qm,
               llr + sg.bit_off + k * sg.qm);
}

}  // namespace tcdev

using namespace tchost;

struct tc_channel_plan {
  uint32_t n_seg = 0;
  uint64_t n_bits = 0, n_sym = 0;
  tcdev::ChanSeg* d_segs = nullptr;
  uint64_t* d_end = nullptr;
  ~tc_channel_plan() {
    for (void* p : {(void*)d_segs, (void*)d_end})
      if (p) (void)hipFree(p);
  }
};

extern "C" {

tc_status tc_channel_plan_create(const tc_channel_seg* segs, size_t n_seg, tc_channel_plan** plan) {
  return guarded([&] {
    if (!plan || (n_seg && !segs)) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (n_seg > 0x7FFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "too many segments"};
    auto p = std::make_unique<tc_channel_plan>();
    std::vector<tcdev::ChanSeg> v(n_seg);
    std::vector<uint64_t> end(n_seg);
    uint64_t bo = 0, so = 0;
    for (size_t i = 0; i < n_seg; ++i) {
      const tc_channel_seg& s = segs[i];
      if (s.qm != 2 && s.qm != 4 && s.qm != 6) throw Failure{TC_ERR_INVALID_ARGUMENT, "qm must be 2, 4 or 6"};
      if (s.n_bits % s.qm) throw Failure{TC_ERR_INVALID_ARGUMENT, "n_bits must be a multiple of qm"};
      if (!(s.n0 > 0.0f)) throw Failure{TC_ERR_INVALID_ARGUMENT, "n0 must be positive"};
      if (s.n_bits / s.qm > 0xFFFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "segment too long"};
      v[i] = {bo, so, static_cast<uint32_t>(s.n_bits / s.qm), s.qm, s.rayleigh ? 1u : 0u, s.n0};
      bo += s.n_bits;
      so += s.n_bits / s.qm;
      end[i] = so;
    }
    p->n_seg = static_cast<uint32_t>(n_seg);
    p->n_bits = bo;
    p->n_sym = so;
    if (n_seg) {
      require_device();
      p->d_segs = dev_alloc<tcdev::ChanSeg>(n_seg);
      dev_upload(p->d_segs, v);
      p->d_end = dev_alloc<uint64_t>(n_seg);
      dev_upload(p->d_end, end);
    }
    uploads_done();
    *plan = p.release();
  });
}

void tc_channel_plan_destroy(tc_channel_plan* plan) { delete plan; }

tc_status tc_channel_plan_sizes(const tc_channel_plan* plan, uint64_t sizes[2]) {
  return guarded([&] {
    if (!plan || !sizes) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    sizes[0] = plan->n_bits;
    sizes[1] = plan->n_sym;
  });
}

tc_status tc_channel_llr(tc_channel_plan* plan, uint64_t seed, const uint8_t* bits_dev, float* llr_dev, float* y_dev,
                         float* h_dev, void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (plan->n_sym == 0) return;
    if (!bits_dev || !llr_dev) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    const uint64_t blocks = (plan->n_sym + 255) / 256;
    if (blocks > 0xFFFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "too many symbols"};
    hipLaunchKernelGGL(tcdev::channel_llr_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(hip_stream), plan->d_segs, plan->d_end, plan->n_seg, plan->n_sym, seed,
                       bits_dev, llr_dev, y_dev, h_dev);
    hip_check(hipGetLastError(), "channel launch");
  });
}

tc_status tc_qam_demap(tc_channel_plan* plan, const float* y_dev, const float* h_dev, float* llr_dev,
                       void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (plan->n_sym == 0) return;
    if (!y_dev || !llr_dev) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    const uint64_t blocks = (plan->n_sym + 255) / 256;
    if (blocks > 0xFFFFFFFFull) throw Failure{TC_ERR_INVALID_ARGUMENT, "too many symbols"};
    hipLaunchKernelGGL(tcdev::qam_demap_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                       static_cast<hipStream_t>(hip_stream), plan->d_segs, plan->d_end, plan->n_seg, plan->n_sym,
                       y_dev, h_dev, llr_dev);
    hip_check(hipGetLastError(), "demapper launch");
  });
}

}  // extern "C"

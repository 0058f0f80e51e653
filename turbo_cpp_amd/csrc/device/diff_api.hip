// C ABI of the differentiable decoder: plan construction (geometry, jobs,
// work space) and the forward / adjoint launches in float or double.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <vector>

#include "../host/codec_internal.h"
#include "diff_types.h"
#include "hip_util.h"
#include "turbo_cpp/c_api.h"

namespace tcdev {
template <typename R, bool kLds>
__global__ void diff_forward_kernel(DiffArgs<R>);
template <typename R, bool kLds>
__global__ void diff_backward_kernel(DiffArgs<R>);
template <typename R>
__global__ void diff_scale_reduce_kernel(const R*, uint32_t, uint32_t, R*);
}  // namespace tcdev

using namespace tchost;

struct tc_diff_plan {
  bool fp64 = false;
  size_t n_cb = 0, esize = 4;
  uint32_t halves = 0, njobs = 0, grid = 0, max_seg = 0, max_warm = 0, slot_elems = 0;
  size_t lds_bytes = 0;  // dynamic LDS per workgroup (0: no LDS path)
  uint64_t llr_elems = 0, post_elems = 0, saved_elems = 0, ebar_elems = 0;
  uint64_t ext_elems = 0, ckpt_job_half = 0;  // saved: extrinsics, then checkpoints
  tcdev::DiffJob* d_jobs = nullptr;
  tcdev::DiffCb* d_cbs = nullptr;
  void* d_ebar = nullptr;
  void* d_part = nullptr;
  void* d_scratch = nullptr;
  ~tc_diff_plan() {
    for (void* p : {(void*)d_jobs, (void*)d_cbs, d_ebar, d_part, d_scratch})
      if (p) (void)hipFree(p);
  }
};

namespace {

constexpr uint32_t kMaxHalves = 32;

// Scratch elements per lane: staged sys+la and par, the sys/par adjoints of
// the alpha and the beta side.
uint32_t slot_elems(uint32_t s, uint32_t g) {
  const uint32_t p = s + 2 * g;
  return 6 * p;
}

// A job runs on a workgroup of two waves (alpha side and beta side).
constexpr uint32_t kDiffThreads = 128;

void build(tc_diff_plan& p, const uint32_t* k, size_t n, uint32_t iterations, uint32_t min_segment, uint32_t warmup,
           bool fp64) {
  if (n && !k) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
  if (iterations == 0 || 2 * iterations > kMaxHalves)
    throw Failure{TC_ERR_INVALID_ARGUMENT, "iterations must be 1..16"};
  if (min_segment == 0) min_segment = 32;
  p.fp64 = fp64;
  p.esize = fp64 ? 8 : 4;
  p.halves = 2 * iterations;
  std::map<uint32_t, std::vector<uint32_t>> by_k;
  std::vector<uint32_t> order;
  for (size_t i = 0; i < n; ++i) {
    const int qi = turbo_cpp::detail::qpp_index(k[i]);
    if (qi < 0) throw Failure{TC_ERR_OUT_OF_RANGE, "unsupported QPP block size"};
    auto& v = by_k[k[i]];
    if (v.empty()) order.push_back(k[i]);
    v.push_back(static_cast<uint32_t>(i));
  }
  std::vector<tcdev::DiffCb> cbs(n);
  uint64_t lo = 0, po = 0, so = 0, eo = 0;
  for (size_t i = 0; i < n; ++i) {
    cbs[i] = {lo, po, so, eo};
    lo += 3ull * (k[i] + 4);
    po += k[i];
    so += (uint64_t)p.halves * k[i];
    eo += 2ull * k[i];
  }
  p.n_cb = n;
  p.llr_elems = lo;
  p.post_elems = po;
  p.saved_elems = so;
  p.ebar_elems = eo;
  // jobs: blocks of one K, 64 / M per wave; the job's blocks are consecutive
  // entries of a reordered copy of the descriptors
  std::vector<tcdev::DiffJob> jobs;
  std::vector<tcdev::DiffCb> ordered;
  size_t lds_elems = 0;
  for (uint32_t kk : order) {
    const auto& idx = by_k[kk];
    const uint32_t lo_seg = std::max<uint32_t>(min_segment, (kk + 63) / 64);
    uint32_t s = lo_seg;
    while (s < kk && kk % s != 0) ++s;
    if (s > kk) s = kk;
    const uint32_t m = kk / s;
    const uint32_t g = (m > 1) ? std::min(warmup ? warmup : 32u, s) : 0u;
    const uint32_t per = 64 / m;
    const auto& e = turbo_cpp::detail::kQppTable[turbo_cpp::detail::qpp_index(kk)];
    for (size_t b = 0; b < idx.size(); b += per) {
      const uint32_t nb = static_cast<uint32_t>(std::min<size_t>(per, idx.size() - b));
      jobs.push_back({kk, s, m, nb, g, static_cast<uint32_t>(e.f1), static_cast<uint32_t>(e.f2),
                      static_cast<uint32_t>(ordered.size())});
      for (uint32_t x = 0; x < nb; ++x) ordered.push_back(cbs[idx[b + x]]);
    }
    p.max_seg = std::max(p.max_seg, s);
    p.max_warm = std::max(p.max_warm, g);
    lds_elems = std::max<size_t>(
        lds_elems, tcdev::diff_lds_elems(kk, g, static_cast<uint32_t>(std::min<size_t>(per, idx.size()))));
  }
  p.njobs = static_cast<uint32_t>(jobs.size());
  p.slot_elems = slot_elems(p.max_seg, p.max_warm);
  // the forward's checkpoints follow the extrinsics in the saved state
  p.ext_elems = p.saved_elems;
  p.ckpt_job_half = 64ull * tcdev::diff_ck_lane_elems(p.max_seg, p.max_warm);
  p.saved_elems += (uint64_t)p.njobs * p.halves * p.ckpt_job_half;
  if (n == 0) return;
  require_device();
  int dev = 0, cus = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
  p.grid = std::max<uint32_t>(1, std::min<uint32_t>(p.njobs, 4u * static_cast<uint32_t>(cus)));
  // LDS copies of the job's blocks when they leave room for four jobs per CU
  // (the resident count above); larger jobs gather from global memory
  int lds_per_cu = 0;
  hip_check(hipDeviceGetAttribute(&lds_per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev),
            "hipDeviceGetAttribute");
  const size_t need = lds_elems * p.esize;
  p.lds_bytes = (need > 0 && need <= static_cast<size_t>(lds_per_cu) / 4) ? need : 0;
  p.d_jobs = dev_alloc<tcdev::DiffJob>(jobs.size());
  dev_upload(p.d_jobs, jobs);
  p.d_cbs = dev_alloc<tcdev::DiffCb>(ordered.size());
  dev_upload(p.d_cbs, ordered);
  p.d_ebar = dev_alloc<char>(p.ebar_elems * p.esize);
  p.d_part = dev_alloc<char>((size_t)p.njobs * p.halves * p.esize);
  p.d_scratch = dev_alloc<char>((size_t)p.grid * 64 * p.slot_elems * p.esize);
}

template <typename R>
tcdev::DiffArgs<R> args(const tc_diff_plan& p, const void* llr, const void* scales, void* post, void* saved,
                        const void* grad_post, void* grad_llr) {
  tcdev::DiffArgs<R> a{};
  a.jobs = p.d_jobs;
  a.njobs = p.njobs;
  a.halves = p.halves;
  a.cbs = p.d_cbs;
  a.llr = static_cast<const R*>(llr);
  a.scales = static_cast<const R*>(scales);
  a.post = static_cast<R*>(post);
  a.saved = static_cast<R*>(saved);
  a.grad_post = static_cast<const R*>(grad_post);
  a.grad_llr = static_cast<R*>(grad_llr);
  a.ebar = static_cast<R*>(p.d_ebar);
  a.dscale_part = static_cast<R*>(p.d_part);
  a.scratch = static_cast<R*>(p.d_scratch);
  a.ckpt = saved ? static_cast<R*>(saved) + p.ext_elems : nullptr;
  a.ckpt_job_half = p.ckpt_job_half;
  a.slot_elems = p.slot_elems;
  a.max_seg = p.max_seg;
  a.max_warm = p.max_warm;
  a.use_lds = p.lds_bytes > 0;
  return a;
}

template <typename R>
void forward(const tc_diff_plan& p, const void* llr, const void* scales, void* post, void* saved, hipStream_t s) {
  const auto a = args<R>(p, llr, scales, post, saved, nullptr, nullptr);
  if (p.lds_bytes)
    hipLaunchKernelGGL((tcdev::diff_forward_kernel<R, true>), dim3(p.grid), dim3(kDiffThreads), p.lds_bytes, s, a);
  else
    hipLaunchKernelGGL((tcdev::diff_forward_kernel<R, false>), dim3(p.grid), dim3(kDiffThreads), 0, s, a);
  hip_check(hipGetLastError(), "differentiable decoder forward launch");
}

template <typename R>
void backward(const tc_diff_plan& p, const void* llr, const void* scales, const void* saved, const void* grad_post,
              void* grad_llr, void* grad_scales, hipStream_t s) {
  const auto a = args<R>(p, llr, scales, nullptr, const_cast<void*>(saved), grad_post, grad_llr);
  if (p.lds_bytes)
    hipLaunchKernelGGL((tcdev::diff_backward_kernel<R, true>), dim3(p.grid), dim3(kDiffThreads), p.lds_bytes, s, a);
  else
    hipLaunchKernelGGL((tcdev::diff_backward_kernel<R, false>), dim3(p.grid), dim3(kDiffThreads), 0, s, a);
  hip_check(hipGetLastError(), "differentiable decoder backward launch");
  hipLaunchKernelGGL(tcdev::diff_scale_reduce_kernel<R>, dim3(p.halves), dim3(256), 0, s,
                     static_cast<const R*>(p.d_part), p.njobs, p.halves, static_cast<R*>(grad_scales));
  hip_check(hipGetLastError(), "scale gradient reduction launch");
}

}  // namespace

extern "C" {

tc_status tc_diff_plan_create(const uint32_t* k, size_t n_cb, uint32_t iterations, uint32_t min_segment,
                              uint32_t warmup, int32_t fp64, tc_diff_plan** plan) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    auto p = std::make_unique<tc_diff_plan>();
    build(*p, k, n_cb, iterations, min_segment, warmup, fp64 != 0);
    uploads_done();
    *plan = p.release();
  });
}

void tc_diff_plan_destroy(tc_diff_plan* plan) { delete plan; }

tc_status tc_diff_plan_sizes(const tc_diff_plan* plan, uint64_t sizes[4]) {
  return guarded([&] {
    if (!plan || !sizes) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    sizes[0] = plan->llr_elems;
    sizes[1] = plan->post_elems;
    sizes[2] = plan->halves;
    sizes[3] = plan->saved_elems;
  });
}

tc_status tc_diff_forward(tc_diff_plan* plan, const void* llr, const void* scales, void* post, void* saved,
                          void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (plan->n_cb == 0) return;
    if (!llr || !scales || !post || !saved) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    const auto s = static_cast<hipStream_t>(hip_stream);
    if (plan->fp64)
      forward<double>(*plan, llr, scales, post, saved, s);
    else
      forward<float>(*plan, llr, scales, post, saved, s);
  });
}

tc_status tc_diff_backward(tc_diff_plan* plan, const void* llr, const void* scales, const void* saved,
                           const void* grad_post, void* grad_llr, void* grad_scales, void* hip_stream) {
  return guarded([&] {
    if (!plan) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (plan->n_cb == 0) return;
    if (!llr || !scales || !saved || !grad_post || !grad_llr || !grad_scales)
      throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    const auto s = static_cast<hipStream_t>(hip_stream);
    if (plan->fp64)
      backward<double>(*plan, llr, scales, saved, grad_post, grad_llr, grad_scales, s);
    else
      backward<float>(*plan, llr, scales, saved, grad_post, grad_llr, grad_scales, s);
  });
}

}  // extern "C"

// C ABI of the HIP decoders: plan construction (geometry, job list, device
// work space), launches, and the per-code-block convenience entry point that
// backs turbo_decode_code_block_gpu_* in the C++ API.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../host/codec_internal.h"
#include "decoder_types.h"
#include "derate_plan.h"
#include "hip_util.h"
#include "turbo_cpp/c_api.h"

namespace tcdev {
__global__ void exact_decode_kernel(const CbInfo*, const uint32_t*, uint32_t, const float*, float*, float*, int32_t*,
                                    uint32_t*, uint32_t, int32_t);
__global__ void window_decode_log_kernel(WinArgs);
__global__ void window_decode_log_es_kernel(WinArgs);
__global__ void window_decode_log_f32_kernel(WinArgs);
__global__ void window_decode_log_es_f32_kernel(WinArgs);
__global__ void window_decode_log_lean_kernel(WinArgs);
__global__ void window_decode_log_lean_es_kernel(WinArgs);
__global__ void window_decode_log_lean_f32_kernel(WinArgs);
__global__ void window_decode_log_lean_es_f32_kernel(WinArgs);
__global__ void window_decode_lin_kernel(WinArgs);
__global__ void window_decode_lin_es_kernel(WinArgs);
__global__ void window_decode_lin_f32_kernel(WinArgs);
__global__ void window_decode_lin_es_f32_kernel(WinArgs);
__global__ void window_decode_exact_kernel(WinArgs);
__global__ void window_decode_exact_es_kernel(WinArgs);
// jobs on two and three waves (window_wide2.hip, window_wide3.hip)
#define TC_WIDE_KERNELS(W)                                  \
  __global__ void window_decode_log_##W##_kernel(WinArgs);        \
  __global__ void window_decode_log_es_##W##_kernel(WinArgs);     \
  __global__ void window_decode_log_f32_##W##_kernel(WinArgs);    \
  __global__ void window_decode_log_es_f32_##W##_kernel(WinArgs); \
  __global__ void window_decode_lin_##W##_kernel(WinArgs);        \
  __global__ void window_decode_lin_es_##W##_kernel(WinArgs);     \
  __global__ void window_decode_lin_f32_##W##_kernel(WinArgs);    \
  __global__ void window_decode_lin_es_f32_##W##_kernel(WinArgs);
TC_WIDE_KERNELS(w2)
TC_WIDE_KERNELS(w3)
#undef TC_WIDE_KERNELS
}  // namespace tcdev

namespace {

using namespace tchost;

// Waves one window job may span (window_wide2.hip, window_wide3.hip).
constexpr int kMaxJobWaves = 3;

tc_options normalised(const tc_options* o) {
  tc_options r{};
  if (o) r = *o;
  if (r.min_segment == 0) r.min_segment = 32;
  if (r.warmup == 0) r.warmup = 32;
  if (r.decoder < TC_DECODER_EXACT || r.decoder > TC_DECODER_EXACT_SERIAL)
    throw Failure{TC_ERR_INVALID_ARGUMENT, "unknown decoder"};
  // extrinsics held between half-iterations: the window decoders default to
  // fp16, the exact decoders hold them in fp32 (the reference's width)
  const bool exact = r.decoder == TC_DECODER_EXACT || r.decoder == TC_DECODER_EXACT_SERIAL;
  if (r.extrinsic_bits == 0) r.extrinsic_bits = exact ? 32 : tcdev::kExtrinsicBits;
  if (r.extrinsic_bits != 16 && r.extrinsic_bits != 32)
    throw Failure{TC_ERR_INVALID_ARGUMENT, "extrinsic_bits must be 16 or 32"};
  if (exact && r.extrinsic_bits != 32)
    throw Failure{TC_ERR_INVALID_ARGUMENT, "the exact decoders hold extrinsics in fp32"};
  if (r.max_waves_per_block == 0) r.max_waves_per_block = 1;
  if (r.max_waves_per_block > static_cast<uint32_t>(kMaxJobWaves))
    throw Failure{TC_ERR_INVALID_ARGUMENT, "max_waves_per_block must be 1 to " + std::to_string(kMaxJobWaves)};
  if (exact && r.max_waves_per_block != 1)
    throw Failure{TC_ERR_INVALID_ARGUMENT, "code blocks on several waves: window decoders (log, lin) only"};
  return r;
}

// Segment geometry of the window decoders: S = the smallest divisor of K
// with S >= max(min_segment, ceil(K / (64 n))), n = max_waves_per_block, so
// M = K / S <= 64 n lanes: ceil(M / 64) waves per block.
tc_segment_info geometry(uint32_t k, const tc_options& o) {
  if (turbo_cpp::detail::qpp_index(k) < 0) throw Failure{TC_ERR_OUT_OF_RANGE, "unsupported QPP block size"};
  const uint32_t lanes = tcdev::kWaveLanes * std::max<uint32_t>(o.max_waves_per_block, 1);
  const uint32_t lo = std::max<uint32_t>(o.min_segment, (k + lanes - 1) / lanes);
  uint32_t s = lo;
  while (s < k && k % s != 0) ++s;
  if (s > k) s = k;
  if (s > static_cast<uint32_t>(tcdev::kMaxSeg))
    throw Failure{TC_ERR_INVALID_ARGUMENT, "segment length above " + std::to_string(tcdev::kMaxSeg) + " for K=" +
                                               std::to_string(k)};
  tc_segment_info g{};
  g.k = k;
  g.segment_length = s;
  g.segments = k / s;
  g.waves_per_block = (g.segments + tcdev::kWaveLanes - 1) / tcdev::kWaveLanes;
  g.blocks_per_wave = tcdev::kWaveLanes * g.waves_per_block / g.segments;
  // the exact decoder's boundaries are fixed up exactly, with no warm-up
  g.warmup = (g.segments > 1 && o.decoder != TC_DECODER_EXACT) ? std::min(o.warmup, s - 1) : 0;
  g.extrinsic_bits = o.extrinsic_bits;
  return g;
}

// The lean geometry of the window decoders (window_kernel.h): one code block
// per job in 64 segments, S and the warm-up multiples of 8 (the K >= 2048
// that are multiples of 512, at warm-ups 8, 16, 24, ...).
bool lean_job(const tcdev::WinJob& j) {
  return j.nblk == 1 && j.nseg == static_cast<uint32_t>(tcdev::kWaveLanes) && j.seg % tcdev::kSubWin == 0 &&
         j.warm > 0 && j.warm % tcdev::kSubWin == 0;
}

using WinKernel = void (*)(tcdev::WinArgs);
// lean: every job of the plan has the lean geometry; the log-domain decoder
// then runs its lean kernels (the same results, bit for bit).  waves: the
// waves of one job (2, 3: the wide kernels).
WinKernel window_kernel(const tc_options& o, bool lean = false, uint32_t waves = 1) {
  const bool es = o.early_stop != 0, f32 = o.extrinsic_bits == 32;
  const bool log = o.decoder == TC_DECODER_WINDOW_LOG;
  if (waves == 2) {
    if (log) {
      if (f32) return es ? tcdev::window_decode_log_es_f32_w2_kernel : tcdev::window_decode_log_f32_w2_kernel;
      return es ? tcdev::window_decode_log_es_w2_kernel : tcdev::window_decode_log_w2_kernel;
    }
    if (f32) return es ? tcdev::window_decode_lin_es_f32_w2_kernel : tcdev::window_decode_lin_f32_w2_kernel;
    return es ? tcdev::window_decode_lin_es_w2_kernel : tcdev::window_decode_lin_w2_kernel;
  }
  if (waves == 3) {
    if (log) {
      if (f32) return es ? tcdev::window_decode_log_es_f32_w3_kernel : tcdev::window_decode_log_f32_w3_kernel;
      return es ? tcdev::window_decode_log_es_w3_kernel : tcdev::window_decode_log_w3_kernel;
    }
    if (f32) return es ? tcdev::window_decode_lin_es_f32_w3_kernel : tcdev::window_decode_lin_f32_w3_kernel;
    return es ? tcdev::window_decode_lin_es_w3_kernel : tcdev::window_decode_lin_w3_kernel;
  }
  if (lean && o.decoder == TC_DECODER_WINDOW_LOG) {
    if (f32) return es ? tcdev::window_decode_log_lean_es_f32_kernel : tcdev::window_decode_log_lean_f32_kernel;
    return es ? tcdev::window_decode_log_lean_es_kernel : tcdev::window_decode_log_lean_kernel;
  }
  switch (o.decoder) {
    case TC_DECODER_EXACT:
      return es ? tcdev::window_decode_exact_es_kernel : tcdev::window_decode_exact_kernel;
    case TC_DECODER_WINDOW_LOG:
      if (f32) return es ? tcdev::window_decode_log_es_f32_kernel : tcdev::window_decode_log_f32_kernel;
      return es ? tcdev::window_decode_log_es_kernel : tcdev::window_decode_log_kernel;
    default:
      if (f32) return es ? tcdev::window_decode_lin_es_f32_kernel : tcdev::window_decode_lin_f32_kernel;
      return es ? tcdev::window_decode_lin_es_kernel : tcdev::window_decode_lin_kernel;
  }
}

size_t lds_bytes_of(uint32_t seg, const tc_options& o, uint32_t waves = 1) {
  // fp32-held extrinsics of one-wave jobs live in the scratch slot (window_job,
  // kF32HeldInSlot: exact and log decoders)
  if (waves == 1 && o.extrinsic_bits == 32 && o.decoder != TC_DECODER_WINDOW_LIN)
    return static_cast<size_t>(tcdev::tile_job_lds_bytes(static_cast<int>(seg)));
  return static_cast<size_t>(
      tcdev::job_lds_bytes(static_cast<int>(seg), o.extrinsic_bits / 8, static_cast<int>(waves)));
}

}  // namespace

// What a launch reads: the device tables and work space of a plan.  A tc_plan
// owns its own (hipMalloc); the no-allocation entry point places a one-block
// plan in a static device array (noalloc.hip).
struct PlanDev {
  tc_options opts{};
  size_t n_cb = 0;
  tcdev::CbInfo* d_cbs = nullptr;
  // exact
  float* d_work = nullptr;
  uint32_t* d_groups = nullptr;
  uint32_t n_groups = 0;
  // window
  tcdev::WinJob* d_jobs = nullptr;
  uint32_t* d_cb_index = nullptr;
  uint32_t* d_qtab = nullptr;
  const uint32_t* d_crc_pow = nullptr;
  float2* d_scratch = nullptr;
  uint32_t* d_sched = nullptr;
  uint32_t njobs = 0, grid = 0;
  size_t lds_bytes = 0;
  bool lean = false;  // every job has the lean geometry (window_kernel)
  uint32_t waves = 1;  // waves per job (work-group)
};

struct tc_plan : PlanDev {
  int device = 0;
  size_t llr_floats = 0, post_floats = 0;
  std::vector<tcdev::CbInfo> cbs;

  ~tc_plan() {
    for (const void* p : {(const void*)d_cbs, (const void*)d_work, (const void*)d_groups, (const void*)d_jobs,
                          (const void*)d_cb_index, (const void*)d_qtab, (const void*)d_crc_pow,
                          (const void*)d_scratch, (const void*)d_sched})
      if (p) (void)hipFree(const_cast<void*>(p));
  }
};

namespace {

// Per-block fields of a plan (QPP parameters, no offsets).
tcdev::CbInfo cb_info(uint32_t k, uint32_t f) {
  const int qi = turbo_cpp::detail::qpp_index(k);
  if (qi < 0) throw Failure{TC_ERR_OUT_OF_RANGE, "unsupported QPP block size"};
  if (f >= k) throw Failure{TC_ERR_INVALID_ARGUMENT, "filler count must be below K"};
  tcdev::CbInfo c{};
  c.k = k;
  c.filler = f;
  c.f1 = static_cast<uint32_t>(turbo_cpp::detail::kQppTable[qi].f1);
  c.f2 = static_cast<uint32_t>(turbo_cpp::detail::kQppTable[qi].f2);
  return c;
}

// Exact-decoder work space of one block, in floats.
size_t exact_work_floats(uint32_t k) { return 16 * (size_t)(k + 1) + 4 * (size_t)k; }

// The window decoders' per-K table: entry t < S holds (pi(t) mod S) << 16 | pi(t) div S.
void fill_qtab(uint32_t k, uint32_t seg, uint32_t f1, uint32_t f2, uint32_t* out) {
  for (uint32_t t = 0; t < seg; ++t) {
    const uint64_t pi = (static_cast<uint64_t>(f1) * t + static_cast<uint64_t>(f2) * t * t) % k;
    out[t] = (static_cast<uint32_t>(pi % seg) << 16) | static_cast<uint32_t>(pi / seg);
  }
}

tcdev::WinJob make_job(const tc_segment_info& g, uint32_t f1, uint32_t f2, uint32_t nblk, uint32_t cb_off,
                       uint32_t qtab_off) {
  tcdev::WinJob j{};
  j.k = g.k;
  j.seg = g.segment_length;
  j.nseg = g.segments;
  j.nblk = nblk;
  j.warm = g.warmup;
  j.f1 = f1;
  j.f2 = f2;
  j.cb_off = cb_off;
  j.qtab_off = qtab_off;
  j.div_magic = static_cast<uint32_t>(((1ull << 32) + g.segment_length - 1) / g.segment_length);
  return j;
}

void build_plan(tc_plan& p, const uint32_t* k, const uint32_t* filler, size_t n, const tc_options& o) {
  require_device();
  hip_check(hipGetDevice(&p.device), "hipGetDevice");
  p.opts = o;
  p.n_cb = n;
  p.cbs.resize(n);
  size_t llr = 0, post = 0, work = 0;
  for (size_t i = 0; i < n; ++i) {
    auto& c = p.cbs[i];
    c = cb_info(k[i], filler ? filler[i] : 0);
    c.llr_off = llr;
    c.post_off = post;
    c.work_off = work;
    llr += 3 * (size_t)(k[i] + 4);
    post += k[i];
    work += exact_work_floats(k[i]);
  }
  p.llr_floats = llr;
  p.post_floats = post;
  p.d_cbs = dev_alloc<tcdev::CbInfo>(n);
  dev_upload(p.d_cbs, p.cbs);
  if (o.decoder == TC_DECODER_EXACT_SERIAL) {
    p.d_work = dev_alloc<float>(work);
    // groups of up to kExactBlocksPerWave blocks of one K, one wave each
    constexpr size_t B = tcdev::kExactBlocksPerWave;
    std::map<uint32_t, std::vector<uint32_t>> by_k;
    for (size_t i = 0; i < n; ++i) by_k[k[i]].push_back(static_cast<uint32_t>(i));
    std::vector<uint32_t> groups;
    for (const auto& [kk, list] : by_k) {
      (void)kk;
      for (size_t s = 0; s < list.size(); s += B)
        for (size_t x = 0; x < B; ++x) groups.push_back(s + x < list.size() ? list[s + x] : tcdev::kNoBlock);
    }
    p.n_groups = static_cast<uint32_t>(groups.size() / B);
    p.d_groups = dev_alloc<uint32_t>(groups.size());
    dev_upload(p.d_groups, groups);
    uploads_done();
    return;
  }
  // ---- window decoders: group by K, pack blocks into jobs of 64 lanes per
  // wave, as many waves per job as the widest block needs ----
  std::map<uint32_t, std::vector<uint32_t>> by_k;
  for (size_t i = 0; i < n; ++i) by_k[k[i]].push_back(static_cast<uint32_t>(i));
  p.waves = 1;
  for (const auto& [kk, list] : by_k) {
    (void)list;
    p.waves = std::max(p.waves, geometry(kk, o).waves_per_block);
  }
  std::vector<tcdev::WinJob> jobs;
  std::vector<uint32_t> cb_index, qtab;
  for (const auto& [kk, list] : by_k) {
    tc_segment_info g = geometry(kk, o);
    g.blocks_per_wave = tcdev::kWaveLanes * p.waves / g.segments;  // blocks per job
    const int qi = turbo_cpp::detail::qpp_index(kk);
    const uint32_t f1 = static_cast<uint32_t>(turbo_cpp::detail::kQppTable[qi].f1);
    const uint32_t f2 = static_cast<uint32_t>(turbo_cpp::detail::kQppTable[qi].f2);
    const uint32_t qoff = static_cast<uint32_t>(qtab.size());
    qtab.resize(qoff + g.segment_length);
    fill_qtab(kk, g.segment_length, f1, f2, qtab.data() + qoff);
    for (size_t s = 0; s < list.size(); s += g.blocks_per_wave) {
      const tcdev::WinJob j = make_job(g, f1, f2, static_cast<uint32_t>(std::min<size_t>(g.blocks_per_wave, list.size() - s)),
                                       static_cast<uint32_t>(cb_index.size()), qoff);
      for (uint32_t b = 0; b < j.nblk; ++b) cb_index.push_back(list[s + b]);
      jobs.push_back(j);
    }
  }
  // padding: the window kernels read a chunk's kSubWin entries at once
  qtab.resize(qtab.size() + tcdev::kSubWin, 0u);
  // longest segments first: a job's time grows with S + 2g
  std::stable_sort(jobs.begin(), jobs.end(), [](const tcdev::WinJob& a, const tcdev::WinJob& b) {
    return a.seg + 2 * a.warm > b.seg + 2 * b.warm;
  });
  p.njobs = static_cast<uint32_t>(jobs.size());
  p.lds_bytes = 0;
  p.lean = !jobs.empty();
  for (const auto& j : jobs) {
    p.lds_bytes = std::max<size_t>(p.lds_bytes, lds_bytes_of(j.seg, o, p.waves));
    p.lean = p.lean && p.waves == 1 && lean_job(j);
  }
  hipDeviceProp_t prop;
  hip_check(hipGetDeviceProperties(&prop, p.device), "hipGetDeviceProperties");
  int per_cu = 0;  // work-groups (jobs) per CU
  const void* fn = reinterpret_cast<const void*>(window_kernel(o, p.lean, p.waves));
  hip_check(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, tcdev::kWaveLanes * p.waves, p.lds_bytes),
            "hipOccupancyMaxActiveBlocksPerMultiprocessor");
  per_cu = std::max(per_cu, 1);
  if (o.waves_per_cu)
    per_cu = std::min<int>(per_cu, std::max<int>(1, static_cast<int>(o.waves_per_cu / p.waves)));
  p.grid = std::max<uint32_t>(1, std::min<uint32_t>(p.njobs, static_cast<uint32_t>(per_cu) * prop.multiProcessorCount));
  p.d_jobs = dev_alloc<tcdev::WinJob>(jobs.size());
  dev_upload(p.d_jobs, jobs);
  p.d_cb_index = dev_alloc<uint32_t>(cb_index.size());
  dev_upload(p.d_cb_index, cb_index);
  p.d_qtab = dev_alloc<uint32_t>(qtab.size());
  dev_upload(p.d_qtab, qtab);
  p.d_scratch = dev_alloc<float2>(static_cast<size_t>(p.grid) * p.waves * tcdev::kSlotFloat2);
  p.d_sched = dev_alloc<uint32_t>(2);
  hip_check(hipMemset(p.d_sched, 0, 2 * sizeof(uint32_t)), "hipMemset");
  std::vector<uint32_t> pw(tcdev::kCrcPowLen);
  tchost::crc24b_powers(pw.data());
  uint32_t* crc_pow = dev_alloc<uint32_t>(pw.size());
  p.d_crc_pow = crc_pow;
  dev_upload(crc_pow, pw);
  uploads_done();
}

void launch(const PlanDev& p, const float* llr, float* post, int32_t* crc, uint32_t* iters, hipStream_t s,
            const tcdev::RxIn* rx = nullptr, const tcdev::TbOut* tb = nullptr) {
  if (p.n_cb == 0) return;
  const tc_options& o = p.opts;
  if (o.decoder == TC_DECODER_EXACT_SERIAL) {
    hipLaunchKernelGGL(tcdev::exact_decode_kernel, dim3(p.n_groups), dim3(64), 0, s, p.d_cbs, p.d_groups,
                       p.n_groups, llr, post, p.d_work, crc, iters, o.max_iterations, o.early_stop);
  } else {
    tcdev::WinArgs a{p.d_jobs, p.njobs, p.d_cb_index, p.d_cbs, p.d_qtab, p.d_crc_pow, llr, post, p.d_scratch,
                     crc, iters, o.max_iterations, o.early_stop, tcdev::RxIn{}};
    if (rx) a.rx = *rx;
    a.tb = tcdev::TbOut{};
    if (tb) a.tb = *tb;
    a.sched = p.d_sched;
    hipLaunchKernelGGL(window_kernel(o, p.lean, p.waves), dim3(p.grid), dim3(tcdev::kWaveLanes * p.waves),
                       p.lds_bytes, s, a);
  }
  hip_check(hipGetLastError(), "kernel launch");
}

}  // namespace

bool tchost::launch_window(tc_plan* plan, const float* llr, float* post, int32_t* crc_ok, uint32_t* iters,
                           hipStream_t stream, const tcdev::RxIn* rx, const tcdev::TbOut* tb) {
  if (plan->opts.decoder == TC_DECODER_EXACT_SERIAL) return false;
  launch(*plan, llr, post, crc_ok, iters, stream, rx, tb);
  return true;
}

void tchost::crc24b_powers(uint32_t* out) {
  // D^m mod G(D) for the CRC24B generator (lte_turbo.cpp:108-111)
  uint32_t r = 1;
  for (int m = 0; m < tcdev::kCrcPowLen; ++m) {
    out[m] = r;
    r = ((r << 1) & 0xFFFFFFu) ^ (((r >> 23) & 1u) ? 0xC00063u : 0u);
  }
}

tc_options tchost::normalised_options(const tc_options* o) { return normalised(o); }

namespace {
// Tables of a one-block decode, uploaded in one copy to the start of the
// caller's area; the decoder's work space follows at kOneBlockWorkOff.
struct OneBlockHeader {
  tcdev::CbInfo cb;
  tcdev::WinJob job;
  uint32_t groups[tcdev::kExactBlocksPerWave];
  uint32_t cb_index;
  uint32_t sched[2];  // the window decoders' job queue: zero at launch
  uint32_t qtab[tcdev::kMaxSeg + tcdev::kSubWin];
};
constexpr size_t round256(size_t n) { return (n + 255) / 256 * 256; }
constexpr size_t kOneBlockWorkOff = round256(sizeof(OneBlockHeader));
constexpr size_t kMaxK = 6144;
}  // namespace

size_t tchost::one_block_area_bytes() {
  const size_t exact = exact_work_floats(kMaxK) * sizeof(float);
  const size_t window = tcdev::kSlotFloat2 * sizeof(float2);
  return kOneBlockWorkOff + round256(std::max(exact, window));
}

void tchost::decode_one_block(uint8_t* area, size_t area_bytes, const uint32_t* crc_pow_dev, const tc_options& o,
                              uint32_t k, uint32_t filler, const float* llr_dev, float* post_dev, int32_t* crc_dev,
                              uint32_t* it_dev, hipStream_t s) {
  OneBlockHeader h{};
  h.cb = cb_info(k, filler);
  PlanDev p{};
  p.opts = o;
  p.n_cb = 1;
  auto at = [&](const void* field) {
    return area + (static_cast<const uint8_t*>(field) - reinterpret_cast<const uint8_t*>(&h));
  };
  p.d_cbs = reinterpret_cast<tcdev::CbInfo*>(at(&h.cb));
  size_t work_bytes = 0;
  if (o.decoder == TC_DECODER_EXACT_SERIAL) {
    work_bytes = exact_work_floats(k) * sizeof(float);
    h.groups[0] = 0;
    for (uint32_t b = 1; b < tcdev::kExactBlocksPerWave; ++b) h.groups[b] = tcdev::kNoBlock;
    p.d_groups = reinterpret_cast<uint32_t*>(at(h.groups));
    p.n_groups = 1;
    p.d_work = reinterpret_cast<float*>(area + kOneBlockWorkOff);
  } else {
    const tc_segment_info g = geometry(k, o);
    work_bytes = g.waves_per_block * tcdev::kSlotFloat2 * sizeof(float2);
    fill_qtab(k, g.segment_length, h.cb.f1, h.cb.f2, h.qtab);  // padding entries stay zero
    h.job = make_job(g, h.cb.f1, h.cb.f2, 1, 0, 0);
    h.cb_index = 0;
    p.d_jobs = reinterpret_cast<tcdev::WinJob*>(at(&h.job));
    p.d_cb_index = reinterpret_cast<uint32_t*>(at(&h.cb_index));
    p.d_qtab = reinterpret_cast<uint32_t*>(at(h.qtab));
    p.d_sched = reinterpret_cast<uint32_t*>(at(h.sched));
    p.d_crc_pow = crc_pow_dev;
    p.d_scratch = reinterpret_cast<float2*>(area + kOneBlockWorkOff);
    p.njobs = 1;
    p.grid = 1;
    p.waves = g.waves_per_block;
    p.lds_bytes = lds_bytes_of(g.segment_length, o, p.waves);
    p.lean = p.waves == 1 && lean_job(h.job);
  }
  if (kOneBlockWorkOff + work_bytes > area_bytes)
    throw Failure{TC_ERR_INVALID_ARGUMENT, "code block needs " + std::to_string(kOneBlockWorkOff + work_bytes) +
                                               " bytes of work space, above the " + std::to_string(area_bytes) +
                                               " bytes available"};
  hip_check(hipMemcpyAsync(area, &h, sizeof h, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
  // the header lives on this thread's stack: the copy must be done before it goes
  hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
  launch(p, llr_dev, post_dev, crc_dev, it_dev, s);
}

namespace {

void check_status(tc_status st) {
  if (st != TC_OK) throw Failure{st, turbo_cpp::detail::last_error()};
}

// Device buffer that only grows (freed by its owner).
struct GrowBuf {
  void* p = nullptr;
  size_t cap = 0;
  template <typename T>
  T* get(size_t n) {
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (cap < bytes) {
      void* q = nullptr;
      hip_check(hipMalloc(&q, bytes), "hipMalloc");
      if (p) (void)hipFree(p);
      p = q;
      cap = bytes;
    }
    return static_cast<T*>(p);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Per-thread state of the host-buffer entry points: a stream of the thread's
// own, so host threads that decode at once -- as the reference's
// decode_transport_block does with std::async (lte_turbo.cpp:941-951) -- do
// not wait on each other's work (each call waits for its own stream only);
// single-block plans by shape; the last batch plan; device staging buffers
// that only grow.  A call that repeats an earlier shape allocates nothing.
using OneKey = std::tuple<uint32_t, uint32_t, int32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>;
struct HostCalls {
  int device = -1;
  hipStream_t stream = nullptr;
  std::map<OneKey, std::unique_ptr<tc_plan>> one_plans;
  std::unique_ptr<tc_plan> batch;
  std::vector<uint32_t> batch_k, batch_f;
  tc_options batch_o{};
  GrowBuf llr, post, crc, it;
  void reset() {
    one_plans.clear();
    batch.reset();
    batch_k.clear();
    batch_f.clear();
    for (GrowBuf* g : {&llr, &post, &crc, &it}) g->release();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    device = -1;
  }
  ~HostCalls() { reset(); }
};
thread_local HostCalls g_host;

// This thread's state on the current device (set up on first use).
HostCalls& host_calls() {
  require_device();
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (g_host.device != dev) {
    g_host.reset();
    g_host.device = dev;
  }
  if (!g_host.stream) hip_check(hipStreamCreateWithFlags(&g_host.stream, hipStreamNonBlocking), "hipStreamCreate");
  return g_host;
}

bool same_options(const tc_options& a, const tc_options& b) {
  return a.max_iterations == b.max_iterations && a.early_stop == b.early_stop && a.decoder == b.decoder &&
         a.min_segment == b.min_segment && a.warmup == b.warmup && a.waves_per_cu == b.waves_per_cu &&
         a.extrinsic_bits == b.extrinsic_bits && a.max_waves_per_block == b.max_waves_per_block;
}

}  // namespace

hipStream_t tchost::host_stream() { return host_calls().stream; }

extern "C" {

tc_status tc_device_count(int32_t* count) {
  return guarded([&] {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

tc_status tc_segment_geometry(uint32_t k, const tc_options* opts, tc_segment_info* out) {
  return guarded([&] { *out = geometry(k, normalised(opts)); });
}

tc_status tc_plan_create(const uint32_t* k, const uint32_t* filler, size_t n_cb, const tc_options* opts,
                         tc_plan** plan) {
  return guarded([&] {
    if (!plan || (n_cb && !k)) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    const tc_options o = normalised(opts);
    auto p = std::make_unique<tc_plan>();
    build_plan(*p, k, filler, n_cb, o);
    *plan = p.release();
  });
}

void tc_plan_destroy(tc_plan* plan) { delete plan; }
size_t tc_plan_llr_floats(const tc_plan* plan) { return plan ? plan->llr_floats : 0; }
size_t tc_plan_post_floats(const tc_plan* plan) { return plan ? plan->post_floats : 0; }

tc_status tc_plan_decode(tc_plan* plan, const float* llr_dev, float* post_dev, int32_t* crc_ok_dev,
                         uint32_t* iters_dev, void* hip_stream) {
  return guarded([&] {
    if (!plan || (plan->n_cb && (!llr_dev || !post_dev))) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    launch(*plan, llr_dev, post_dev, crc_ok_dev, iters_dev, static_cast<hipStream_t>(hip_stream));
  });
}

tc_status tc_plan_decode_rx(tc_plan* plan, const tc_derate_plan* derate, uint32_t n_tx, const float* const* rx_dev,
                            const int32_t* rv, int32_t accumulate, float* keep_dev, float* post_dev,
                            int32_t* crc_ok_dev, uint32_t* iters_dev, void* hip_stream) {
  return guarded([&] {
    if (!plan || !derate) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    if (n_tx == 0 || n_tx > static_cast<uint32_t>(tcdev::kMaxTx))
      throw Failure{TC_ERR_INVALID_ARGUMENT, "n_tx must be 1..4"};
    if (plan->n_cb != derate->n_cb) throw Failure{TC_ERR_INVALID_ARGUMENT, "plans cover different code blocks"};
    for (size_t i = 0; i < plan->n_cb; ++i)
      if (plan->cbs[i].k != derate->ks[i] || plan->cbs[i].filler != derate->fs[i])
        throw Failure{TC_ERR_INVALID_ARGUMENT, "plans cover different code blocks"};
    if (plan->n_cb == 0) return;
    if (!rx_dev || !post_dev || (accumulate && !keep_dev)) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    tcdev::RxIn x{};
    for (uint32_t t = 0; t < static_cast<uint32_t>(tcdev::kMaxTx); ++t) {
      const int32_t r = (rv && t < n_tx) ? rv[t] : -1;
      if (r < -1 || r > 3) throw Failure{TC_ERR_INVALID_ARGUMENT, "rv must be -1 (per block) or 0..3"};
      x.rv[t] = r < 0 ? tcdev::kBlockRv : static_cast<uint32_t>(r);
      x.rx[t] = t < n_tx ? rx_dev[t] : nullptr;
      if (t < n_tx && !x.rx[t]) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    }
    x.n_tx = n_tx;
    x.base = accumulate ? 1 : 0;
    x.keep = keep_dev;
    x.rank = derate->d_rank;
    x.dcb = derate->d_cbs;
    const auto s = static_cast<hipStream_t>(hip_stream);
    if (plan->opts.decoder == TC_DECODER_EXACT_SERIAL) {
      // no fused staging in the serial exact decoder: de-rate into keep, then decode it
      if (!keep_dev) throw Failure{TC_ERR_INVALID_ARGUMENT, "the serial exact decoder needs keep"};
      for (uint32_t t = 0; t < n_tx; ++t)
        check_status(tc_derate_plan_run(const_cast<tc_derate_plan*>(derate), rx_dev[t], rv ? rv[t] : -1,
                                        (t > 0 || accumulate) ? 1 : 0, keep_dev, hip_stream));
      launch(*plan, keep_dev, post_dev, crc_ok_dev, iters_dev, s);
      return;
    }
    launch(*plan, nullptr, post_dev, crc_ok_dev, iters_dev, s, &x);
  });
}

tc_status tc_decode_batch_host(const uint32_t* k, const uint32_t* filler, size_t n_cb, const tc_options* opts,
                               const float* llr, float* post, int32_t* crc_ok, uint32_t* iterations_used) {
  return guarded([&] {
    const tc_options o = normalised(opts);
    if (n_cb == 0) return;
    if (!k || !llr || !post) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    HostCalls& h = host_calls();
    // the last batch's plan is reused when the blocks and options repeat
    std::vector<uint32_t> fs(n_cb, 0u);
    if (filler) fs.assign(filler, filler + n_cb);
    if (!h.batch || !same_options(h.batch_o, o) || h.batch_k.size() != n_cb ||
        !std::equal(h.batch_k.begin(), h.batch_k.end(), k) || h.batch_f != fs) {
      h.batch.reset();
      auto p = std::make_unique<tc_plan>();
      build_plan(*p, k, fs.data(), n_cb, o);
      h.batch = std::move(p);
      h.batch_k.assign(k, k + n_cb);
      h.batch_f = fs;
      h.batch_o = o;
    }
    const tc_plan& p = *h.batch;
    float* d_llr = h.llr.get<float>(p.llr_floats);
    float* d_post = h.post.get<float>(p.post_floats);
    int32_t* d_crc = h.crc.get<int32_t>(n_cb);
    uint32_t* d_it = h.it.get<uint32_t>(n_cb);
    const hipStream_t s = h.stream;
    hip_check(hipMemcpyAsync(d_llr, llr, p.llr_floats * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    launch(p, d_llr, d_post, d_crc, d_it, s);
    hip_check(hipMemcpyAsync(post, d_post, p.post_floats * sizeof(float), hipMemcpyDeviceToHost, s),
              "hipMemcpyAsync D2H");
    if (crc_ok)
      hip_check(hipMemcpyAsync(crc_ok, d_crc, n_cb * sizeof(int32_t), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    if (iterations_used)
      hip_check(hipMemcpyAsync(iterations_used, d_it, n_cb * sizeof(uint32_t), hipMemcpyDeviceToHost, s),
                "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "decode");
  });
}

tc_status tc_decode_code_block(const tc_cb_desc* desc, const float* d0, size_t n0, const float* d1, size_t n1,
                               const float* d2, size_t n2, const tc_options* opts, float* posterior,
                               int32_t* crc_ok, uint32_t* iterations_used) {
  return guarded([&] {
    if (!desc || !d0 || !d1 || !d2 || !posterior) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    // lte_turbo_cuda.cu:1400-1402 / lte_turbo.cpp:744-747: spans must be d_r long
    if (n0 != desc->d_r || n1 != desc->d_r || n2 != desc->d_r)
      throw Failure{TC_ERR_INVALID_ARGUMENT, "invalid LLR sizes for turbo decode"};
    const uint32_t k = static_cast<uint32_t>(desc->k_r);
    if (turbo_cpp::detail::qpp_index(k) < 0) throw Failure{TC_ERR_OUT_OF_RANGE, "unsupported QPP block size"};
    if (desc->d_r < desc->k_r + 4) throw Failure{TC_ERR_INVALID_ARGUMENT, "d_r must be at least K + 4"};
    const tc_options o = normalised(opts);
    HostCalls& h = host_calls();
    const uint32_t f = static_cast<uint32_t>(desc->filler_count);
    const OneKey key{k, f, o.decoder, o.max_iterations, o.min_segment, o.warmup * 2 + (o.early_stop != 0),
                     o.waves_per_cu, o.extrinsic_bits * 4 + o.max_waves_per_block};
    auto& slot = h.one_plans[key];
    if (!slot) {
      auto p = std::make_unique<tc_plan>();
      build_plan(*p, &k, &f, 1, o);
      slot = std::move(p);
    }
    const size_t d = k + 4;
    float* d_llr = h.llr.get<float>(3 * d);
    float* d_post = h.post.get<float>(k);
    int32_t* d_flags = h.crc.get<int32_t>(2);  // CRC flag, iterations
    const hipStream_t s = h.stream;
    // only the first K+4 entries of each stream are read
    hip_check(hipMemcpyAsync(d_llr, d0, d * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    hip_check(hipMemcpyAsync(d_llr + d, d1, d * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    hip_check(hipMemcpyAsync(d_llr + 2 * d, d2, d * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    launch(*slot, d_llr, d_post, d_flags, reinterpret_cast<uint32_t*>(d_flags + 1), s);
    int32_t flags[2] = {0, 0};
    hip_check(hipMemcpyAsync(posterior, d_post, k * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipMemcpyAsync(flags, d_flags, sizeof flags, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "decode");
    if (crc_ok) *crc_ok = flags[0];
    if (iterations_used) *iterations_used = static_cast<uint32_t>(flags[1]);
  });
}

}  // extern "C"

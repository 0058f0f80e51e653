// Per-code-block decode that allocates no device memory.
//
// tc_decode_code_block_noalloc is tc_decode_code_block (the reference's
// turbo_decode_code_block_gpu_* call) for callers that cannot call hipMalloc:
// serving loops that must not allocate, harnesses that forbid it.  It runs the
// same kernels; everything they need beyond the caller's host buffers lives in
// `scratch`, a file-scope device array of kScratchBytes that each GPU the
// library runs on holds until the process ends:
//
//   [0, kCrcOff + 24 KiB)   CRC24B power table (written once per device)
//   kIoOff                  the block's [d0|d1|d2] (3 (K+4) floats) and
//                           posterior (K floats), CRC flag, iteration count
//   kAreaOff                the one-block plan: tables, then the decoder's
//                           work space (exact: alpha/beta of every step;
//                           window: one wave's scratch slot)
//
// In host code `scratch` names a host-side stand-in, not device memory: the
// device address comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that address is used.  Every call on a GPU
// shares that GPU's array, so calls that overlap on one GPU (several host
// threads at once) race; a caller serialises them.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../host/codec_internal.h"
#include "decoder_types.h"
#include "derate_plan.h"
#include "hip_util.h"
#include "turbo_cpp/c_api.h"

namespace {

using namespace tchost;

constexpr size_t kMaxK = 6144;  // the largest QPP size: every valid block fits
constexpr size_t round256(size_t n) { return (n + 255) / 256 * 256; }
constexpr size_t kCrcOff = 0;
constexpr size_t kIoOff = round256(kCrcOff + tcdev::kCrcPowLen * sizeof(uint32_t));
constexpr size_t kIoLlr = 0;                                           // floats, from kIoOff
constexpr size_t kIoPost = kIoLlr + 3 * (kMaxK + 4);                   // floats
constexpr size_t kIoFlags = kIoPost + kMaxK;                           // crc_ok, iterations
constexpr size_t kAreaOff = round256(kIoOff + (kIoFlags + 2) * sizeof(float));
// The one-block area is 492,800 bytes (exact decoder at K = 6144); 640 KiB
// leaves room for it and is checked against it at set-up.
constexpr size_t kScratchBytes = 640 * 1024;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_scratch[kMaxDevices] = {};

// This device's copy of `scratch`, set up on the first call on the device:
// its address, and the CRC24B power table written into it.
uint8_t* device_scratch() {
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw Failure{TC_ERR_RUNTIME, "device index above 63"};
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_scratch[dev]) {
    if (kAreaOff + one_block_area_bytes() > kScratchBytes)
      throw Failure{TC_ERR_RUNTIME, "no-allocation scratch smaller than a one-block decode needs"};
    void* p = nullptr;
    hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    static uint32_t pw[tcdev::kCrcPowLen];
    crc24b_powers(pw);
    hip_check(hipMemcpy(static_cast<uint8_t*>(p) + kCrcOff, pw, sizeof pw, hipMemcpyHostToDevice),
              "hipMemcpy H2D");
    g_scratch[dev] = static_cast<uint8_t*>(p);
  }
  return g_scratch[dev];
}

}  // namespace

extern "C" tc_status tc_decode_code_block_noalloc(const tc_cb_desc* desc, const float* d0, size_t n0,
                                                  const float* d1, size_t n1, const float* d2, size_t n2,
                                                  const tc_options* opts, float* posterior, int32_t* crc_ok,
                                                  uint32_t* iterations_used) {
  return guarded([&] {
    if (!desc || !d0 || !d1 || !d2 || !posterior) throw Failure{TC_ERR_INVALID_ARGUMENT, "null argument"};
    // lte_turbo_cuda.cu:1400-1402 / lte_turbo.cpp:744-747: spans must be d_r long
    if (n0 != desc->d_r || n1 != desc->d_r || n2 != desc->d_r)
      throw Failure{TC_ERR_INVALID_ARGUMENT, "invalid LLR sizes for turbo decode"};
    if (turbo_cpp::detail::qpp_index(desc->k_r) < 0) throw Failure{TC_ERR_OUT_OF_RANGE, "unsupported QPP block size"};
    if (desc->d_r < desc->k_r + 4) throw Failure{TC_ERR_INVALID_ARGUMENT, "d_r must be at least K + 4"};
    const uint32_t k = static_cast<uint32_t>(desc->k_r);
    if (k > kMaxK)
      throw Failure{TC_ERR_INVALID_ARGUMENT, "K=" + std::to_string(k) + " above the no-allocation maximum " +
                                                 std::to_string(kMaxK)};
    const tc_options o = normalised_options(opts);
    if (o.decoder == TC_DECODER_WINDOW_LOG || o.decoder == TC_DECODER_WINDOW_LIN) {
      tc_segment_info g{};
      if (tc_segment_geometry(k, &o, &g) != TC_OK) throw Failure{TC_ERR_INVALID_ARGUMENT, tc_last_error()};
      if (g.waves_per_block > 1)
        throw Failure{TC_ERR_INVALID_ARGUMENT, "K=" + std::to_string(k) + " at this segment length spans " +
                                                   std::to_string(g.waves_per_block) +
                                                   " waves, whose work space exceeds the no-allocation array "
                                                   "(one wave's): use max_waves_per_block 1"};
    }
    require_device();
    uint8_t* base = device_scratch();
    float* io = reinterpret_cast<float*>(base + kIoOff);
    float* llr = io + kIoLlr;
    float* post = io + kIoPost;
    int32_t* crc = reinterpret_cast<int32_t*>(io + kIoFlags);
    uint32_t* it = reinterpret_cast<uint32_t*>(io + kIoFlags + 1);
    const hipStream_t s = nullptr;
    // only the first K+4 entries of each stream are read
    const size_t d = k + 4;
    hip_check(hipMemcpyAsync(llr, d0, d * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    hip_check(hipMemcpyAsync(llr + d, d1, d * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    hip_check(hipMemcpyAsync(llr + 2 * d, d2, d * sizeof(float), hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
    decode_one_block(base + kAreaOff, kScratchBytes - kAreaOff, reinterpret_cast<const uint32_t*>(base + kCrcOff),
                     o, k, static_cast<uint32_t>(desc->filler_count), llr, post, crc, it, s);
    int32_t flags[2] = {0, 0};
    hip_check(hipMemcpyAsync(posterior, post, k * sizeof(float), hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipMemcpyAsync(flags, crc, sizeof flags, hipMemcpyDeviceToHost, s), "hipMemcpyAsync D2H");
    hip_check(hipStreamSynchronize(s), "decode");
    if (crc_ok) *crc_ok = flags[0];
    if (iterations_used) *iterations_used = static_cast<uint32_t>(flags[1]);
  });
}

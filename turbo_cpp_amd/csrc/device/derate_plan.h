// Internal plan interfaces shared by the C ABI translation units: the soft
// de-rate matching plan (its own entry points and the window decoders'
// fused input staging) and the decode plan's launch with fused inputs and
// transport-block assembly.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "rx_types.h"
#include "turbo_cpp/c_api.h"

struct tc_derate_plan {
  size_t n_cb = 0;
  size_t rx_floats = 0, llr_floats = 0;
  uint32_t max_held = 0;  // max over blocks of min(E, N_valid): LDS floats per transmission
  std::vector<uint32_t> ks, fs;  // per block, to match a decode plan
  tcdev::DerateCb* d_cbs = nullptr;
  uint16_t* d_rank = nullptr;
  ~tc_derate_plan() {
    for (void* p : {(void*)d_cbs, (void*)d_rank})
      if (p) (void)hipFree(p);
  }
};

struct tc_plan;
namespace tcdev {
struct RxIn;
}

namespace tchost {
// Launch a decode plan's window decoder with optional fused inputs (rx) and
// TB assembly in its epilogue (tb).  Returns false, launching nothing, for an
// exact-decoder plan.
bool launch_window(tc_plan* plan, const float* llr, float* post, int32_t* crc_ok, uint32_t* iters,
                   hipStream_t stream, const tcdev::RxIn* rx, const tcdev::TbOut* tb);

// tc_options with defaults filled in (min_segment, warmup); throws on an
// unknown decoder.
tc_options normalised_options(const tc_options* o);
// The calling host thread's own stream on the current device (created on
// first use): host-buffer entry points enqueue on it and wait for it alone.
hipStream_t host_stream();
// D^m mod G(D) of the CRC24B generator for m < tcdev::kCrcPowLen.
void crc24b_powers(uint32_t* out);
// Device bytes a one-code-block decode needs (tables + work space, K up to 6144).
size_t one_block_area_bytes();
// Decode one code block (K, filler) of llr_dev ([d0|d1|d2], K+4 each) into
// post_dev on stream s, with the plan's tables and work space placed in
// `area` (256-byte aligned, area_bytes long) instead of allocated: copies the
// tables in (waiting for that copy) and launches.  crc_pow_dev: the CRC24B
// power table (window decoders).  Allocates no device memory.
void decode_one_block(uint8_t* area, size_t area_bytes, const uint32_t* crc_pow_dev, const tc_options& o,
                      uint32_t k, uint32_t filler, const float* llr_dev, float* post_dev, int32_t* crc_dev,
                      uint32_t* it_dev, hipStream_t s);
}  // namespace tchost


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
  uint32_t max_valid = 0;
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

// D^m mod G(D) of the CRC24B generator for m < tcdev::kCrcPowLen.
void crc24b_powers(uint32_t* out);
}  // namespace tchost


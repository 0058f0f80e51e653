// The soft de-rate matching plan, shared by its own entry points and the
// window decoders' fused input staging.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "rx_types.h"

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

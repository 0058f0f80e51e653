// Host-side helpers shared by the C ABI translation units: status-carrying
// failures, HIP error checks, and small device allocation wrappers.
#pragma once

#include <hip/hip_runtime.h>

#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../host/codec_internal.h"
#include "turbo_cpp/c_api.h"

namespace tchost {

struct Failure {
  tc_status status;
  std::string what;
};

inline void hip_check(hipError_t e, const char* where) {
  if (e != hipSuccess) throw Failure{TC_ERR_RUNTIME, std::string(where) + ": " + hipGetErrorString(e)};
}

// Runs f, turning exceptions into a status code and the thread's last error.
template <typename F>
tc_status guarded(F&& f) {
  try {
    f();
    return TC_OK;
  } catch (const Failure& e) {
    turbo_cpp::detail::set_last_error(e.what);
    return e.status;
  } catch (const std::bad_alloc&) {
    turbo_cpp::detail::set_last_error("out of host memory");
    return TC_ERR_RUNTIME;
  } catch (const std::exception& e) {
    turbo_cpp::detail::set_last_error(e.what());
    return TC_ERR_RUNTIME;
  }
}

inline void require_device() {
  int n = 0;
  const hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    throw Failure{TC_ERR_RUNTIME, "no HIP device available (the GPU backends need an MI355X / gfx950)"};
}

// Plan construction uploads tables and clears counters on the null stream;
// callers launch plans on streams of their own, which are not ordered after
// it (non-blocking streams), so a builder waits for that work before
// returning the plan.
inline void uploads_done() { hip_check(hipStreamSynchronize(nullptr), "plan upload"); }

template <typename T>
T* dev_alloc(size_t n) {
  void* p = nullptr;
  if (n == 0) n = 1;
  hip_check(hipMalloc(&p, n * sizeof(T)), "hipMalloc");
  return static_cast<T*>(p);
}

template <typename T>
void dev_upload(T* dst, const std::vector<T>& src) {
  if (!src.empty()) hip_check(hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
}

// Device buffer freed on scope exit (host wrappers that stage host data).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  explicit DevBuf(size_t n) : p(dev_alloc<T>(n)) {}
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

}  // namespace tchost

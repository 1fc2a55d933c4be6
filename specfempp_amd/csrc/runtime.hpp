// Host runtime helpers shared by the engine and the batch: device buffers,
// HIP error checks, and the conversion of exceptions into C ABI status codes.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/specfempp_amd.h"

namespace sfx {

// Message of the last failing C ABI call on this thread (sfx_last_error).
std::string &last_error();

struct HostAssembly;
// Host data and run description an engine was built from (the batch packs them).
const HostAssembly &engine_host(const sfx_engine *e);
const sfx_run_desc &engine_run(const sfx_engine *e);

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define SFX_HIP(call)                                                                                  \
  do {                                                                                                 \
    hipError_t err_ = (call);                                                                          \
    if (err_ != hipSuccess)                                                                            \
      throw ::sfx::HipError(std::string(#call) + " failed: " + hipGetErrorString(err_) + " (" __FILE__ \
                            ":" + std::to_string(__LINE__) + ")");                                     \
  } while (0)

template <class T> struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t count) {
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
    n = count;
    if (count) SFX_HIP(hipMalloc(&p, count * sizeof(T)));
  }
  // allocate only if the size changes (keeps the address, e.g. for captured graphs)
  void ensure(size_t count) {
    if (count != n || (count && !p)) alloc(count);
  }
  void upload(const std::vector<T> &v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) SFX_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

// Run fn; map exceptions to SFX_ERR_* and keep the message.
template <class F> int guarded(F &&fn) {
  try {
    fn();
    return SFX_OK;
  } catch (const HipError &e) {
    last_error() = e.what();
    return SFX_ERR_HIP;
  } catch (const std::invalid_argument &e) {
    last_error() = e.what();
    return SFX_ERR_ARG;
  } catch (const std::logic_error &e) {
    last_error() = e.what();
    return SFX_ERR_STATE;
  } catch (const std::exception &e) {
    last_error() = e.what();
    return SFX_ERR_MESH;
  }
}

}  // namespace sfx

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

// Keeps at most two windows of steps queued on a stream: a loop of
// asynchronous launches calls tick() after each step, and every window steps
// the host waits for the work queued one window earlier.  Not for use while
// the stream is being captured.
struct Throttle {
  hipStream_t s;
  int window, n = 0;
  hipEvent_t ev[2] = {nullptr, nullptr};
  explicit Throttle(hipStream_t stream, int steps = 256) : s(stream), window(steps) {}
  Throttle(const Throttle &) = delete;
  Throttle &operator=(const Throttle &) = delete;
  ~Throttle() {
    for (auto e : ev)
      if (e) (void)hipEventDestroy(e);
  }
  void tick() {
    if (++n % window) return;
    const int w = n / window;
    hipEvent_t &mine = ev[w & 1];
    if (!mine) SFX_HIP(hipEventCreateWithFlags(&mine, hipEventDisableTiming));
    SFX_HIP(hipEventRecord(mine, s));
    if (w >= 2) SFX_HIP(hipEventSynchronize(ev[(w - 1) & 1]));
  }
};

// Run fn; map exceptions to SFX_ERR_* and keep the message.  When fn fails
// and drain is a stream, the work fn already queued on it is waited for
// before returning, so no kernel of a failed call is still running when the
// caller frees or exits.
template <class F> int guarded(F &&fn, hipStream_t drain = nullptr) {
  try {
    fn();
    return SFX_OK;
  } catch (...) {
    if (drain) (void)hipStreamSynchronize(drain);
    try {
      throw;
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
    } catch (...) {
      last_error() = "unknown error";
      return SFX_ERR_STATE;
    }
  }
}

}  // namespace sfx

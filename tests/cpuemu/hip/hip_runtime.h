// CPU emulation of the HIP subset the library uses, for bounds checking with
// AddressSanitizer.  Kernels run on the host: one std::thread per GPU thread
// of a block, blocks one after another; __shared__ becomes static storage
// (shared by the threads of the running block); __syncthreads is a barrier
// that exited threads drop out of.
#pragma once
#include <atomic>
#include <barrier>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) alignas(n)

struct float4 { float x, y, z, w; };
struct float2 { float x, y; };
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
struct int4 { int x, y, z, w; };

struct dim3 {
  unsigned x = 1, y = 1, z = 1;
  dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint3e { unsigned x = 0, y = 0, z = 0; };
inline thread_local uint3e threadIdx, blockIdx;
inline thread_local dim3 blockDim, gridDim;

namespace emu {
inline std::barrier<> *bar = nullptr;
inline std::mutex atom_mu;
inline std::vector<float> shfl_buf;
inline void sync() { bar->arrive_and_wait(); }
}  // namespace emu

inline void __syncthreads() { emu::sync(); }
inline float __shfl_down(float v, int off, int width = 64) {
  const unsigned t = threadIdx.x;
  emu::shfl_buf[t] = v;
  emu::sync();
  const unsigned src = t + off;
  float r = v;
  if ((t % width) + off < (unsigned)width && src < blockDim.x) r = emu::shfl_buf[src];
  emu::sync();
  return r;
}
inline float unsafeAtomicAdd(float *p, float v) {
  std::lock_guard<std::mutex> g(emu::atom_mu);
  const float old = *p;
  *p = old + v;
  return old;
}
inline float atomicAdd(float *p, float v) { return unsafeAtomicAdd(p, v); }
inline float sqrtf_(float x) { return std::sqrt(x); }

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorNotSupported = 801 };
typedef void *hipStream_t;
typedef void *hipEvent_t;
typedef void *hipGraph_t;
typedef void *hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyHostToHost };
enum { hipStreamCaptureModeThreadLocal = 0, hipEventDisableTiming = 2 };

inline const char *hipGetErrorString(hipError_t e) { return e ? "emulated error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
template <class T> inline hipError_t hipMalloc(T **p, size_t n) {
  *p = static_cast<T *>(std::malloc(n ? n : 1));
  std::memset(*p, 0x7f, n);  // poison: reads of never-written device memory show up as huge values
  return hipSuccess;
}
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipStreamCreate(hipStream_t *s) { *s = (void *)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { *e = (void *)1; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0; return hipSuccess; }
namespace emu {
// stream capture: launches and async copies are recorded as closures and
// replayed by hipGraphLaunch
struct Graph { std::vector<std::function<void()>> ops; };
inline thread_local Graph *capturing = nullptr;
}  // namespace emu
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { emu::capturing = new emu::Graph(); return hipSuccess; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) { *g = emu::capturing; emu::capturing = nullptr; return hipSuccess; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, void *, void *, size_t) {
  *e = new emu::Graph(*static_cast<emu::Graph *>(g));
  return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete static_cast<emu::Graph *>(g); return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t g) { delete static_cast<emu::Graph *>(g); return hipSuccess; }
inline hipError_t hipGraphLaunch(hipGraphExec_t g, hipStream_t) {
  for (auto &op : static_cast<emu::Graph *>(g)->ops) op();
  return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
  if (emu::capturing) emu::capturing->ops.push_back([d, s, n] { std::memcpy(d, s, n); });
  else std::memcpy(d, s, n);
  return hipSuccess;
}

namespace emu {
template <class... P, class... A>
void launch(void (*k)(P...), dim3 grid, dim3 block, size_t, hipStream_t, A &&...args) {
  const unsigned nb = grid.x * grid.y * grid.z, nt = block.x * block.y * block.z;
  if (nb == 0 || nt == 0 || nt > 1024) { std::fprintf(stderr, "EMU: bad launch %u x %u\n", nb, nt); std::abort(); }
  std::tuple<P...> params(static_cast<P>(args)...);
  if (capturing) {
    capturing->ops.push_back([k, grid, block, params] {
      std::apply([&](auto... a) { launch(k, grid, block, 0, nullptr, a...); }, params);
    });
    return;
  }
  shfl_buf.assign(nt, 0.0f);
  // nt pool threads run the blocks one after another; each block gets a fresh
  // barrier for __syncthreads, which a thread leaves when its kernel returns
  std::barrier<> start(nt + 1), done(nt + 1);
  std::unique_ptr<std::barrier<>> br;
  unsigned cur = 0;
  bool stop = false;
  std::vector<std::thread> th;
  th.reserve(nt);
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([&, t] {
      threadIdx.x = t; threadIdx.y = 0; threadIdx.z = 0;
      blockDim = block; gridDim = grid;
      for (;;) {
        start.arrive_and_wait();
        if (stop) return;
        blockIdx.x = cur; blockIdx.y = 0; blockIdx.z = 0;
        std::apply(k, params);
        br->arrive_and_drop();
        done.arrive_and_wait();
      }
    });
  for (unsigned b = 0; b < nb; b++) {
    br = std::make_unique<std::barrier<>>(nt);
    bar = br.get();
    cur = b;
    start.arrive_and_wait();
    done.arrive_and_wait();
  }
  stop = true;
  start.arrive_and_wait();
  for (auto &x : th) x.join();
}
}  // namespace emu
#define hipLaunchKernelGGL(...) emu::launch(__VA_ARGS__)
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }

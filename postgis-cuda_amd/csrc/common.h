// Engine-internal plumbing: status codes, HIP error checks, grow-only device scratch, stream timers.
#pragma once
#include <algorithm>
#include <chrono>
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>
#include <mutex>
#include "../../include/gsj.h"

namespace gsj {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define GSJ_HIP(x)                                                                                     \
  do {                                                                                                 \
    hipError_t e_ = (x);                                                                               \
    if (e_ != hipSuccess)                                                                              \
      throw ::gsj::Error(e_ == hipErrorOutOfMemory ? GSJ_E_NOMEM : GSJ_E_HIP,                          \
                         std::string(#x) + " at " + __FILE__ + ":" + std::to_string(__LINE__) + ": " + \
                             hipGetErrorString(e_));                                                   \
  } while (0)

#define GSJ_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) throw ::gsj::Error(code, msg); \
  } while (0)

// Bump allocator over memory owned elsewhere (the no-allocation entry point's scratch). A buffer bound to an arena
// takes its storage from it and never allocates or frees; reset() hands the whole region out again.
struct Arena {
  unsigned char* base = nullptr;
  size_t cap = 0, used = 0;
  const char* name = "scratch";
  size_t left() const { const size_t off = (used + 255) & ~size_t(255); return off >= cap ? 0 : cap - off; }
  void* take(size_t n) {
    const size_t off = (used + 255) & ~size_t(255);
    if (off > cap || n > cap - off)
      throw Error(GSJ_E_LIMIT, std::string("no-allocation ") + name + " of " + std::to_string(cap >> 20) + " MiB exhausted (" +
                                   std::to_string((off + n) >> 20) + " MiB needed): lower batch_q / max_pairs or use the allocating entry point");
    used = off + n;
    return base + off;
  }
  void reset() { used = 0; }
};

// grow-only device buffer; contents are not preserved on growth
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  Arena* arena = nullptr;   // set: storage comes from the arena (no hipMalloc / hipFree)
  void* get(size_t n) {
    if (n == 0) n = 1;
    if (n > cap) {
      if (arena) {   // the old region stays taken until the arena is reset; double to bound that waste
        const size_t want = std::max(n, 2 * cap) <= arena->left() ? std::max(n, 2 * cap) : n;
        p = arena->take(want);
        cap = want;
        return p;
      }
      if (p) GSJ_HIP(hipFree(p));
      p = nullptr; cap = 0;
      size_t want = n + n / 8;   // headroom so slowly growing requests do not reallocate every call
      GSJ_HIP(hipMalloc(&p, want));
      cap = want;
    }
    return p;
  }
  template <class T> T* as(size_t count) { return static_cast<T*>(get(count * sizeof(T))); }
  void release() { if (p && !arena) (void)hipFree(p); p = nullptr; cap = 0; }
  ~DevBuf() { release(); }
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
};

// grow-only pinned host buffer (bound to an arena: plain host memory from it; copies into it are then staged by the
// runtime, and every use of these buffers synchronises the stream before reading them)
struct PinBuf {
  void* p = nullptr;
  size_t cap = 0;
  Arena* arena = nullptr;
  void* get(size_t n) {
    if (n == 0) n = 1;
    if (n > cap) {
      if (arena) {
        const size_t want = std::max(n, 2 * cap) <= arena->left() ? std::max(n, 2 * cap) : n;
        p = arena->take(want);
        cap = want;
        return p;
      }
      if (p) GSJ_HIP(hipHostFree(p));
      p = nullptr; cap = 0;
      size_t want = n + n / 8;
      GSJ_HIP(hipHostMalloc(&p, want, hipHostMallocDefault));
      cap = want;
    }
    return p;
  }
  template <class T> T* as(size_t count) { return static_cast<T*>(get(count * sizeof(T))); }
  void release() { if (p && !arena) (void)hipHostFree(p); p = nullptr; cap = 0; }
  ~PinBuf() { release(); }
  PinBuf() = default;
  PinBuf(const PinBuf&) = delete;
  PinBuf& operator=(const PinBuf&) = delete;
};

// event pair on a stream; elapsed() synchronises on the stop event
struct WallTimer {   // host wall clock around synchronised device work
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

struct StreamTimer {
  hipEvent_t a = nullptr, b = nullptr;
  hipStream_t s = nullptr;
  explicit StreamTimer(hipStream_t st) : s(st) { GSJ_HIP(hipEventCreate(&a)); GSJ_HIP(hipEventCreate(&b)); }
  ~StreamTimer() { if (a) (void)hipEventDestroy(a); if (b) (void)hipEventDestroy(b); }
  void start() { GSJ_HIP(hipEventRecord(a, s)); }
  float stop() {
    GSJ_HIP(hipEventRecord(b, s));
    GSJ_HIP(hipEventSynchronize(b));
    float ms = 0;
    GSJ_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
};

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// MI355X dispatches the workgroups of a grid round-robin over its 8 XCDs, each with its own L2. This renumbers them
// (a bijection on [0, gridDim.x)) so that each XCD runs one contiguous range of logical blocks: neighbouring work
// items, which read the same polygons and tree nodes, then share an L2. For work of even cost per item (the tree
// walk); the refinement's class lists are sorted by size, where contiguous ranges per XCD would unbalance the XCDs.
__device__ __forceinline__ uint32_t xcd_block() {
  constexpr uint32_t NX = 8;
  const uint32_t b = blockIdx.x, x = b % NX, q = gridDim.x / NX, r = gridDim.x % NX;
  return x * q + (x < r ? x : r) + b / NX;
}

// kernel launch check (launch errors surface on the next runtime call otherwise)
#define GSJ_LAUNCHED() GSJ_HIP(hipGetLastError())

}  // namespace gsj

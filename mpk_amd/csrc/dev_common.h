// dev_common.h -- HIP helpers shared by the device sources of libmpkx.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "internal.h"

#define MPKX_HIP(call)                                                          \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      return mpkx::fail(MPKX_EHIP, "%s failed: %s", #call, hipGetErrorString(e_)); \
  } while (0)

namespace mpkx {

// Device buffer owned by a plan.
template <class T>
struct DevArray {
  T *p = nullptr;
  size_t count = 0;
  ~DevArray() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    count = 0;
  }
  int upload(const T *host, size_t n) {
    release();
    count = n;
    if (n == 0) return MPKX_OK;
    if (hipMalloc(&p, sizeof(T) * n) != hipSuccess) {
      p = nullptr;
      return fail(MPKX_ENOMEM, "hipMalloc of %zu bytes failed", sizeof(T) * n);
    }
    if (host && hipMemcpy(p, host, sizeof(T) * n, hipMemcpyHostToDevice) != hipSuccess)
      return fail(MPKX_EHIP, "hipMemcpy upload failed");
    return MPKX_OK;
  }
  int upload(const std::vector<T> &v) { return upload(v.data(), v.size()); }
  int alloc(size_t n) { return upload(nullptr, n); }
};

// A hipGraph recorded for one (argument set, stream) and replayed.
struct GraphCache {
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  const void *k0 = nullptr, *k1 = nullptr;
  hipStream_t stream = nullptr;
  int k = -1;
  ~GraphCache() { reset(); }
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    k0 = k1 = nullptr;
    k = -1;
  }
  bool matches(const void *a, const void *b, hipStream_t s, int kk) const {
    return exec && a == k0 && b == k1 && s == stream && kk == k;
  }
};

// Workgroup id -> work item so that consecutive items land on the same XCD
// (the dispatcher hands workgroup i to XCD i % 8; each XCD has its own L2).
__device__ __forceinline__ int xcd_remap(int i, int count) {
  const int nx = 8;
  int q = count / nx, rem = count % nx;
  int x = i % nx, k = i / nx;
  return x * q + (x < rem ? x : rem) + k;
}

}  // namespace mpkx

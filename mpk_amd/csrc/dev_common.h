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

// Records launch(stream) into gc's hipGraph for the key (a, b, key, stream)
// and replays it; a later call with the same key only replays.  launch
// returns false when it cannot run the configuration (nothing is kept).
// The legacy default stream cannot be captured, so a private stream
// records the graph in that case.
template <class F>
int graph_run(GraphCache &gc, const void *a, const void *b, int key, hipStream_t st, F launch) {
  if (!gc.matches(a, b, st, key)) {
    gc.reset();
    hipStream_t cap = st;
    bool own = false;
    if (cap == nullptr) {
      if (hipStreamCreate(&cap) != hipSuccess) return fail(MPKX_EHIP, "graph capture: hipStreamCreate failed");
      own = true;
    }
    bool ok = false;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      ok = launch(cap);
      hipError_t e2 = hipStreamEndCapture(cap, &g);
      if (e2 != hipSuccess) e = e2;
    }
    hipGraphExec_t ge = nullptr;
    if (e == hipSuccess && ok) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (own) (void)hipStreamDestroy(cap);
    if (e != hipSuccess || !ok) {
      if (g) (void)hipGraphDestroy(g);
      return ok ? fail(MPKX_EHIP, "graph capture failed: %s", hipGetErrorString(e))
                : fail(MPKX_EUNSUPPORTED, "graph capture: launch configuration not supported");
    }
    gc.graph = g;
    gc.exec = ge;
    gc.k0 = a;
    gc.k1 = b;
    gc.stream = st;
    gc.k = key;
  }
  MPKX_HIP(hipGraphLaunch(gc.exec, st));
  return MPKX_OK;
}

// Workgroup id -> work item so that consecutive items land on the same XCD
// (the dispatcher hands workgroup i to XCD i % 8; each XCD has its own L2).
__device__ __forceinline__ int xcd_remap(int i, int count) {
  const int nx = 8;
  int q = count / nx, rem = count % nx;
  int x = i % nx, k = i / nx;
  return x * q + (x < rem ? x : rem) + k;
}

}  // namespace mpkx

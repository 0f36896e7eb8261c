// CPU stand-in for the subset of the HIP API that wrf_amd/csrc uses -- TEST INFRASTRUCTURE.
//
// It lets the kernel sources of wrf_amd/csrc be compiled by g++ and executed serially on
// the CPU, one (block, thread) at a time, under AddressSanitizer: every device-array access
// of every kernel is then bounds-checked against exact-size allocations, and running the
// grid in forward and reverse order exposes any dependence between threads of one launch.
// Nothing of the product links this file; the product is the hipcc build for gfx950.
#pragma once
#include <cmath>
#include <functional>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define WRFMI_GLOBAL_AS
#define WRFMI_CONST_AS

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct emu_uint3 { unsigned x, y, z; };
extern thread_local emu_uint3 blockIdx, threadIdx;
extern thread_local dim3 blockDim, gridDim;

using std::isfinite;
inline unsigned long long atomicAdd(unsigned long long *p, unsigned long long v) {
  unsigned long long o = *p;
  *p = o + v;
  return o;
}

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorNotSupported = 801 };
typedef struct emu_stream *hipStream_t;
typedef struct emu_graph *hipGraph_t;
typedef struct emu_event *hipEvent_t;
typedef struct emu_graph_exec *hipGraphExec_t;
enum hipMemcpyKind { hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2 };
enum { hipStreamNonBlocking = 1, hipStreamCaptureModeThreadLocal = 1 };
enum hipMemoryType { hipMemoryTypeHost = 0, hipMemoryTypeDevice = 1 };
struct hipPointerAttribute_t { hipMemoryType type; };

inline const char *hipGetErrorString(hipError_t e) { return e ? "emulated HIP error" : "no error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
// device allocations are malloc'd blocks, recorded so that hipMemGetAddressRange can find them
void emu_track(void *p, size_t n);
void emu_untrack(void *p);
inline hipError_t hipMalloc(void **p, size_t n) {
  *p = std::malloc(n);
  if (!*p) return 2;
  emu_track(*p, n);
  return hipSuccess;
}
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) { return hipMalloc((void **)p, n); }
inline hipError_t hipFree(void *p) {
  emu_untrack(p);
  std::free(p);
  return hipSuccess;
}
// pinned host memory is ordinary host memory here
enum { hipHostMallocDefault = 0 };
inline hipError_t hipHostMalloc(void **p, size_t n, unsigned) {
  *p = std::malloc(n);
  return *p ? hipSuccess : 2;
}
template <typename T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned f) { return hipHostMalloc((void **)p, n, f); }
inline hipError_t hipHostFree(void *p) {
  std::free(p);
  return hipSuccess;
}
inline unsigned __float_as_uint(float x) {
  unsigned u;
  std::memcpy(&u, &x, 4);
  return u;
}
typedef void *hipDeviceptr_t;
hipError_t hipMemGetAddressRange(hipDeviceptr_t *base, size_t *size, hipDeviceptr_t p);
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memcpy(d, s, n); return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind k, hipStream_t) { return hipMemcpy(d, s, n, k); }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, int) { *s = nullptr; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipPointerGetAttributes(hipPointerAttribute_t *a, const void *) { a->type = hipMemoryTypeDevice; return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
// a static device array is ordinary host memory here: its address is the symbol's own
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))
inline hipError_t hipGetSymbolAddress(void **p, const void *sym) { *p = const_cast<void *>(sym); return hipSuccess; }
// events are not emulated (kernel timing mode is a GPU-only diagnostic)
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = nullptr; return hipErrorNotSupported; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipErrorNotSupported; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipErrorNotSupported; }
// graph capture is not emulated: the capturing entry point reports an error
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return hipErrorNotSupported; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *) { return hipErrorNotSupported; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t *, hipGraph_t, void *, void *, size_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorNotSupported; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }

// Launch order: 0 = blocks and threads in increasing order, 1 = decreasing order.
int emu_launch_order();
// Runs the nt threads of one workgroup as fibers that meet at every __syncthreads().
void emu_run_block(long nt, bool rev, dim3 b, const std::function<void()> &body);
void __syncthreads();
float __shfl_down(float v, unsigned delta, int width = 64);

template <typename... P, typename... A>
void emu_launch(void (*kern)(P...), dim3 g, dim3 b, A &&...args) {
  const bool rev = emu_launch_order() == 1;
  const long nb = (long)g.x * g.y * g.z, nt = (long)b.x * b.y * b.z;
  gridDim = g;
  blockDim = b;
  for (long ob = 0; ob < nb; ob++) {
    const long bb = rev ? nb - 1 - ob : ob;
    blockIdx = emu_uint3{(unsigned)(bb % g.x), (unsigned)((bb / g.x) % g.y), (unsigned)(bb / ((long)g.x * g.y))};
    emu_run_block(nt, rev, b, [&]() { kern(args...); });
  }
}
#define hipLaunchKernelGGL(K, G, B, SHMEM, STREAM, ...) emu_launch(K, G, B, __VA_ARGS__)

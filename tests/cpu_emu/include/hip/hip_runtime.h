// CPU emulation of the subset of HIP that the library's kernels and host
// code use.  Test infrastructure only: it lets the unmodified kernel sources
// run on the host (g++), so their logic can be checked against the oracle
// and with host sanitizers where no GPU is present.  Nothing in the product
// loads it.
//
// Model: device memory is host memory; every call is synchronous (work
// enqueued during a stream capture is recorded and run by hipGraphLaunch); a kernel
// launch runs its workgroups one after another, each workgroup's threads as
// cooperative fibers.  __syncthreads() waits for every live thread of the
// workgroup; __any(), __ballot() and readlane wait for every live lane of the 64-lane
// wavefront (and must be reached by all of them, as in the kernels'
// uniform control flow).  __shared__ arrays become function statics, shared
// by the fibers of the running workgroup.  Atomics are plain read-modify-
// write operations: only one fiber runs at a time and none yields inside one.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>

typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNoDevice = 100,
  hipErrorLaunchFailure = 719,
} hipError_t;

enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
};

struct ihipStream_t;
typedef ihipStream_t* hipStream_t;
struct ihipEvent_t;
typedef ihipEvent_t* hipEvent_t;
struct ihipGraph_t;
typedef ihipGraph_t* hipGraph_t;
struct ihipGraphExec_t;
typedef ihipGraphExec_t* hipGraphExec_t;
struct hipGraphNode;
typedef hipGraphNode* hipGraphNode_t;
enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2
};

#define hipStreamNonBlocking 0x01
#define hipEventDisableTiming 0x2
#define hipHostMallocDefault 0x0

struct alignas(16) uint4 {
  unsigned int x, y, z, w;
};
inline uint4 make_uint4(unsigned int x, unsigned int y, unsigned int z, unsigned int w) { return uint4{x, y, z, w}; }
struct alignas(8) uint2 {
  unsigned int x, y;
};
inline uint2 make_uint2(unsigned int x, unsigned int y) { return uint2{x, y}; }
struct alignas(16) double2 {
  double x, y;
};

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t x_ = 1, uint32_t y_ = 1, uint32_t z_ = 1) : x(x_), y(y_), z(z_) {}
};

#define __global__
#define __device__
#define __host__
#define __shared__ static
#define __forceinline__ inline __attribute__((always_inline))
#define __noinline__ __attribute__((noinline))
#define __launch_bounds__(...)

namespace flemu {
extern dim3 g_block, g_bdim, g_gdim;
extern dim3* g_tid;
void launch(dim3 grid, dim3 block, const std::function<void()>& body);
void block_barrier();
int wave_any(int pred);
unsigned long long wave_ballot(int pred);
int readlane(int v, int lane);
}  // namespace flemu

#define threadIdx (*::flemu::g_tid)
#define blockIdx (::flemu::g_block)
#define blockDim (::flemu::g_bdim)
#define gridDim (::flemu::g_gdim)
#define __syncthreads() ::flemu::block_barrier()
#define __any(p) ::flemu::wave_any((p) ? 1 : 0)
#define __ballot(p) ::flemu::wave_ballot((p) ? 1 : 0)
#define __builtin_amdgcn_readlane(v, l) ::flemu::readlane((int)(v), (int)(l))
// a wave barrier waits for every live lane of the wavefront; fences order
// nothing further here (one fiber runs at a time)
#define __builtin_amdgcn_wave_barrier() ((void)::flemu::wave_any(0))
#define __builtin_amdgcn_s_setprio(p) ((void)(p))
#define __builtin_amdgcn_fence(order, scope) ((void)0)
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  ::flemu::launch(dim3(grid), dim3(block), [=]() { kernel(__VA_ARGS__); })

template <class T>
inline T min(T a, T b) {
  return b < a ? b : a;
}
template <class T>
inline T max(T a, T b) {
  return a < b ? b : a;
}

template <class T, class U>
inline T atomicAdd(T* p, U v) {
  T o = *p;
  *p = (T)(o + (T)v);
  return o;
}
template <class T, class U>
inline T atomicMin(T* p, U v) {
  T o = *p;
  if ((T)v < o) *p = (T)v;
  return o;
}
template <class T, class U>
inline T atomicMax(T* p, U v) {
  T o = *p;
  if (o < (T)v) *p = (T)v;
  return o;
}
template <class T, class U>
inline T atomicOr(T* p, U v) {
  T o = *p;
  *p = (T)(o | (T)v);
  return o;
}
template <class T, class U>
inline T atomicExch(T* p, U v) {
  T o = *p;
  *p = (T)v;
  return o;
}
template <class T, class U, class V>
inline T atomicCAS(T* p, U cmp, V v) {
  T o = *p;
  if (o == (T)cmp) *p = (T)v;
  return o;
}

inline long long __double_as_longlong(double d) {
  long long r;
  std::memcpy(&r, &d, 8);
  return r;
}
inline int __double2loint(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (int)(uint32_t)u;
}
inline int __double2hiint(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return (int)(uint32_t)(u >> 32);
}
inline double __hiloint2double(int hi, int lo) {
  uint64_t u = ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}

hipError_t hipGetDeviceCount(int* count);
hipError_t hipSetDevice(int device);
hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
// stream-ordered allocation: the emulation's streams run in order on the host
inline hipError_t hipMallocAsync(void** p, size_t bytes, hipStream_t) { return hipMalloc(p, bytes); }
inline hipError_t hipFreeAsync(void* p, hipStream_t) { return hipFree(p); }
typedef void* hipMemPool_t;
enum hipMemPoolAttr { hipMemPoolAttrReleaseThreshold = 4 };
inline hipError_t hipDeviceGetDefaultMemPool(hipMemPool_t* pool, int) {
  *pool = nullptr;
  return hipSuccess;
}
inline hipError_t hipMemPoolSetAttribute(hipMemPool_t, hipMemPoolAttr, void*) { return hipSuccess; }
enum hipMemAllocationType { hipMemAllocationTypeInvalid = 0, hipMemAllocationTypePinned = 1 };
enum hipMemAllocationHandleType { hipMemHandleTypeNone = 0 };
enum hipMemLocationType { hipMemLocationTypeInvalid = 0, hipMemLocationTypeDevice = 1 };
struct hipMemLocation {
  hipMemLocationType type;
  int id;
};
struct hipMemPoolProps {
  hipMemAllocationType allocType;
  hipMemAllocationHandleType handleTypes;
  hipMemLocation location;
  void* win32SecurityAttributes;
  size_t maxSize;
  unsigned char reserved[56];
};
#ifndef __align__
#define __align__(n) __attribute__((aligned(n)))
#endif
// device symbols are host globals in the emulation
#define HIP_SYMBOL(X) X
template <typename T>
inline hipError_t hipGetSymbolAddress(void** p, const T& symbol) {
  *p = (void*)&symbol;
  return hipSuccess;
}
// pools: allocations come from the emulation's heap
inline hipError_t hipMemPoolCreate(hipMemPool_t* pool, const hipMemPoolProps*) {
  static char tag;
  *pool = &tag;
  return hipSuccess;
}
inline hipError_t hipMemPoolDestroy(hipMemPool_t) { return hipSuccess; }
inline hipError_t hipMemPoolTrimTo(hipMemPool_t, size_t) { return hipSuccess; }
inline hipError_t hipMallocFromPoolAsync(void** p, size_t bytes, hipMemPool_t, hipStream_t) {
  return hipMalloc(p, bytes);
}
hipError_t hipHostMalloc(void** p, size_t bytes, unsigned flags);
hipError_t hipHostFree(void* p);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemsetAsync(void* dst, int value, size_t bytes, hipStream_t s = nullptr);
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipDeviceSynchronize();
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipGetLastError();
// Stream capture: work enqueued between begin and end is recorded (not run)
// and replayed in order by hipGraphLaunch.
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* graph);
hipError_t hipGraphInstantiate(hipGraphExec_t* exec, hipGraph_t graph, hipGraphNode_t* error_node, char* log,
                               size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t exec, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t graph);
hipError_t hipGraphExecDestroy(hipGraphExec_t exec);
const char* hipGetErrorString(hipError_t e);

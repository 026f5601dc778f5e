// CPU stand-in for <hip/hip_runtime.h>: lets the engine's own kernel sources
// (sdim_amd/csrc/batch.hip, plain.hip, frame.hip) and host runtime (runtime.cpp)
// compile for the host, where every GPU thread becomes a fiber (tests/cpu_sim/
// simt.cpp).  Test infrastructure only: the library built this way exists so
// that AddressSanitizer can check every LDS and device-memory access of the
// kernels against its allocation, and ThreadSanitizer can check that threads of
// a workgroup communicate only across barriers, atomics and wave exchanges.
//
// Execution model: one workgroup runs on one OS thread as blockDim fibers;
// __syncthreads() suspends a fiber until every live fiber of the workgroup has
// arrived; a wave operation (shuffle, ballot) suspends it until every live lane
// of its 64-lane wave has arrived.  Workgroups of a launch run on a pool of OS
// threads (a persistent kernel drains its queue with whichever workgroups run).
// Dynamic LDS is a thread-local buffer whose bytes beyond the launch's dynamic
// LDS size are poisoned for AddressSanitizer.  Kernel launches run to completion
// synchronously; streams, events and graphs keep the host runtime's semantics
// (stream capture records launches, a graph launch replays them).
#pragma once
#include <limits.h>
#include <stddef.h>
#include <stdint.h>
#include <string.h>

#include <functional>
#include <utility>

// ------------------------------------------------------------------ qualifiers
#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ thread_local
#define __align__(n) __attribute__((aligned(n)))

struct dim3 {
  uint32_t x, y, z;
  constexpr dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};
struct __attribute__((aligned(16))) uint4 {
  uint32_t x, y, z, w;
};
inline uint4 make_uint4(uint32_t a, uint32_t b, uint32_t c, uint32_t d) { return uint4{a, b, c, d}; }

// ------------------------------------------------------------------ execution state
namespace simt {
const dim3& thread_idx();
const dim3& block_idx();
const dim3& block_dim();
const dim3& grid_dim();
void syncthreads();
// every live lane of the calling wave deposits v; returns the 64 deposited values
const uint64_t* wave_exchange(uint64_t v);
uint32_t lane_id();
uint64_t clock_ticks();
void yield();

template <class T>
inline uint64_t to_bits(T v) {
  static_assert(sizeof(T) <= 8, "wave exchange of at most 8 bytes");
  uint64_t b = 0;
  memcpy(&b, &v, sizeof(T));
  return b;
}
template <class T>
inline T from_bits(uint64_t b) {
  T v;
  memcpy(&v, &b, sizeof(T));
  return v;
}
}  // namespace simt

#define threadIdx (::simt::thread_idx())
#define blockIdx (::simt::block_idx())
#define blockDim (::simt::block_dim())
#define gridDim (::simt::grid_dim())

inline void __syncthreads() { simt::syncthreads(); }

template <class T>
inline T __shfl_xor(T v, int mask, int width = 64) {
  const uint64_t* all = simt::wave_exchange(simt::to_bits(v));
  const uint32_t lane = simt::lane_id();
  const uint32_t base = lane & ~(uint32_t)(width - 1);
  return simt::from_bits<T>(all[base + ((lane - base) ^ (uint32_t)mask) % (uint32_t)width]);
}
template <class T>
inline T __shfl_up(T v, unsigned delta, int width = 64) {
  const uint64_t* all = simt::wave_exchange(simt::to_bits(v));
  const uint32_t lane = simt::lane_id();
  const uint32_t base = lane & ~(uint32_t)(width - 1);
  return lane - base >= delta ? simt::from_bits<T>(all[lane - delta]) : v;
}
template <class T>
inline T __shfl(T v, int src, int width = 64) {
  const uint64_t* all = simt::wave_exchange(simt::to_bits(v));
  const uint32_t lane = simt::lane_id();
  const uint32_t base = lane & ~(uint32_t)(width - 1);
  return simt::from_bits<T>(all[base + (uint32_t)src % (uint32_t)width]);
}
inline unsigned long long __ballot(int pred) {
  const uint64_t* all = simt::wave_exchange(pred ? 1u : 0u);
  unsigned long long m = 0;
  for (int i = 0; i < 64; i++) m |= (all[i] & 1ull) << i;
  return m;
}
inline int min(int a, int b) { return a < b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned long long min(unsigned long long a, unsigned long long b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline unsigned long long max(unsigned long long a, unsigned long long b) { return a > b ? a : b; }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int __popc(unsigned x) { return __builtin_popcount(x); }
inline uint64_t __umul64hi(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) >> 64); }
inline uint64_t wall_clock64() { return simt::clock_ticks(); }

// v_perm_b32: byte i of the result is byte sel_i of the 8-byte value hi:lo
// (selectors 0-7); 12 gives 0x00 and 13-15 give 0xFF; 8-11 replicate the sign bit
// of bytes 1, 3, 5, 7.
inline uint32_t simt_perm(uint32_t hi, uint32_t lo, uint32_t sel) {
  const uint64_t v = ((uint64_t)hi << 32) | lo;
  uint32_t out = 0;
  for (int i = 0; i < 4; i++) {
    const uint32_t s = (sel >> (8 * i)) & 0xFFu;
    uint32_t b;
    if (s < 8) b = (uint32_t)(v >> (8 * s)) & 0xFFu;
    else if (s < 12) b = ((v >> (16 * (s - 8) + 15)) & 1) ? 0xFFu : 0u;
    else if (s == 12) b = 0;
    else b = 0xFFu;
    out |= b << (8 * i);
  }
  return out;
}
#define __builtin_amdgcn_perm(hi, lo, sel) simt_perm((hi), (lo), (sel))
#define __builtin_amdgcn_s_sleep(n) ::simt::yield()
// readlane: every live lane of the wave takes lane l's value (l is wave-uniform)
#define __builtin_amdgcn_readlane(v, l) __shfl((int)(v), (int)(l))
#define __builtin_amdgcn_fence(order, scope) __atomic_thread_fence(order)

// ------------------------------------------------------------------ atomics (device semantics: relaxed)
#define SIMT_ATOMIC_RMW(NAME, BUILTIN)                                                            \
  inline int NAME(int* p, int v) { return BUILTIN(p, v, __ATOMIC_RELAXED); }                    \
  inline unsigned NAME(unsigned* p, unsigned v) { return BUILTIN(p, v, __ATOMIC_RELAXED); }     \
  inline unsigned long long NAME(unsigned long long* p, unsigned long long v) { return BUILTIN(p, v, __ATOMIC_RELAXED); }
SIMT_ATOMIC_RMW(atomicAdd, __atomic_fetch_add)
SIMT_ATOMIC_RMW(atomicOr, __atomic_fetch_or)
SIMT_ATOMIC_RMW(atomicXor, __atomic_fetch_xor)
SIMT_ATOMIC_RMW(atomicAnd, __atomic_fetch_and)
SIMT_ATOMIC_RMW(atomicMin, __atomic_fetch_min)
SIMT_ATOMIC_RMW(atomicMax, __atomic_fetch_max)
#undef SIMT_ATOMIC_RMW
inline unsigned atomicSub(unsigned* p, unsigned v) { return __atomic_fetch_sub(p, v, __ATOMIC_RELAXED); }
inline int atomicExch(int* p, int v) { return __atomic_exchange_n(p, v, __ATOMIC_RELAXED); }
template <class T>
inline T simt_cas(T* p, T cmp, T v) {
  __atomic_compare_exchange_n(p, &cmp, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED);
  return cmp;
}
inline int atomicCAS(int* p, int c, int v) { return simt_cas(p, c, v); }
inline unsigned atomicCAS(unsigned* p, unsigned c, unsigned v) { return simt_cas(p, c, v); }
inline unsigned long long atomicCAS(unsigned long long* p, unsigned long long c, unsigned long long v) {
  return simt_cas(p, c, v);
}
inline unsigned atomicExch(unsigned* p, unsigned v) { return __atomic_exchange_n(p, v, __ATOMIC_RELAXED); }

#define __HIP_MEMORY_SCOPE_WAVEFRONT 2
#define __HIP_MEMORY_SCOPE_WORKGROUP 3
#define __HIP_MEMORY_SCOPE_AGENT 4
#define __HIP_MEMORY_SCOPE_SYSTEM 5
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add((p), (v), (order))
#define __hip_atomic_fetch_min(p, v, order, scope) __atomic_fetch_min((p), (v), (order))
#define __hip_atomic_fetch_max(p, v, order, scope) __atomic_fetch_max((p), (v), (order))
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), (order))
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), (order))

// ------------------------------------------------------------------ host API
typedef enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorNoDevice = 100,
  hipErrorInvalidDevice = 101,
  hipErrorNotSupported = 801,
  hipErrorNotReady = 600,
  hipErrorLaunchFailure = 719,
  hipErrorStreamCaptureUnsupported = 900,
} hipError_t;

typedef enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4,
} hipMemcpyKind;

typedef enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2,
} hipStreamCaptureMode;

typedef enum hipFuncAttribute {
  hipFuncAttributeMaxDynamicSharedMemorySize = 8,
} hipFuncAttribute;

#define hipStreamDefault 0x0
#define hipStreamNonBlocking 0x1
#define hipEventDefault 0x0
#define hipEventDisableTiming 0x2

struct hipDeviceProp_t {
  char name[256];
  size_t totalGlobalMem;
  int multiProcessorCount;
  int maxThreadsPerBlock;
  size_t sharedMemPerBlock;
  char gcnArchName[256];
};

struct ihipStream_t;
struct ihipEvent_t;
struct ihipGraph;
struct ihipGraphExec;
struct hipGraphNode;
typedef ihipStream_t* hipStream_t;
typedef ihipEvent_t* hipEvent_t;
typedef ihipGraph* hipGraph_t;
typedef ihipGraphExec* hipGraphExec_t;
typedef hipGraphNode* hipGraphNode_t;

hipError_t hipGetDeviceCount(int* count);
hipError_t hipSetDevice(int device);
hipError_t hipGetDevice(int* device);
hipError_t hipGetDeviceProperties(hipDeviceProp_t* prop, int device);
hipError_t hipDeviceSynchronize();
const char* hipGetErrorString(hipError_t e);
hipError_t hipGetLastError();
hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned flags);
hipError_t hipStreamCreate(hipStream_t* s);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t* e);
hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b);
hipError_t hipMalloc(void** p, size_t bytes);
hipError_t hipFree(void* p);
hipError_t hipMemcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void* dst, const void* src, size_t bytes, hipMemcpyKind k, hipStream_t s);
hipError_t hipMemset(void* dst, int v, size_t bytes);
hipError_t hipMemsetAsync(void* dst, int v, size_t bytes, hipStream_t s);
hipError_t hipFuncSetAttribute(const void* f, hipFuncAttribute a, int v);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g);
hipError_t hipGraphInstantiate(hipGraphExec_t* x, hipGraph_t g, hipGraphNode_t* err_node, char* log, size_t log_size);
hipError_t hipGraphLaunch(hipGraphExec_t x, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t x);
hipError_t hipGetSymbolAddress(void** p, const void* symbol);
#define HIP_SYMBOL(x) (&(x))

namespace simt {
// runs (or, on a capturing stream, records) one launch of grid x block fibers
void submit_launch(hipStream_t s, dim3 grid, dim3 block, size_t shmem, std::function<void()> body);
}

template <class... P, class... A>
inline void simt_launch(void (*kernel)(P...), dim3 grid, dim3 block, size_t shmem, hipStream_t s, A&&... args) {
  simt::submit_launch(s, grid, block, shmem, [kernel, args...]() { kernel(args...); });
}
#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  simt_launch((kernel), dim3(grid), dim3(block), (size_t)(shmem), (stream), __VA_ARGS__)

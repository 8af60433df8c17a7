// Host emulation of the HIP device model, for building the library's kernel
// sources with g++ and AddressSanitizer on a machine without a GPU.
//
// This header stands in for <hip/hip_runtime.h> when tests/hostemu is first
// on the include path.  It is test infrastructure: the product library is
// always built with hipcc for gfx950.
//
// Model.  A launch hands its workgroups to a pool of OS threads (one per
// core); one workgroup runs entirely inside one OS thread, its work-items as
// cooperative fibers (each with its own stack, so its locals and private
// arrays are its own and AddressSanitizer checks them).  A fiber runs until
// it reaches a wave operation or a workgroup barrier and then yields; the
// wave operations (shuffles, readlane, ballot) exchange values through a
// per-wave slot array between two barrier phases of the wave's lanes, which
// is the converged-wave semantics.  A work-item that has returned from the
// kernel leaves its wave's and its workgroup's barriers.  Device buffers are
// malloc'd with their exact byte sizes, so any access past a buffer is
// reported.  __shared__ variables become thread_local statics: a workgroup
// runs inside one OS thread, so its work-items share one copy.
#pragma once
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __align__(n) alignas(n)

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

// threadIdx is set by the fiber scheduler before it resumes a work-item.
inline thread_local dim3 threadIdx(0, 0, 0);
inline thread_local dim3 blockIdx(0, 0, 0);
inline dim3 blockDim;
inline dim3 gridDim;

extern "C" void hipemu_switch(void** save_sp, void* new_sp);   // emu_runtime.cpp

namespace hipemu {

constexpr int kLanes = 64;
constexpr int kMaxThreads = 1024;
constexpr size_t kFiberStack = 512 * 1024;

struct Bar {
  int expected = 0, arrived = 0;
  unsigned phase = 0;
};

struct Wave {
  Bar bar;
  uint64_t slot[kLanes];
  bool alive[kLanes];
};

struct Fiber {
  void* sp = nullptr;
  char* stack = nullptr;
  bool done = false;
};

// Per OS thread: the workgroup it is running.
struct Group {
  std::function<void()>* body = nullptr;
  int nthreads = 0;
  Bar sync;
  Wave waves[kMaxThreads / kLanes];
  Fiber fib[kMaxThreads];
  void* sched_sp = nullptr;
  int cur = 0;
  const void* sched_stack_bottom = nullptr;
  size_t sched_stack_size = 0;
};

Group& group();            // emu_runtime.cpp
void yield();              // back to the scheduler
void run_launch(std::function<void()>& body, unsigned nblocks, int nthreads);

inline void arrive_and_wait(Bar& b)
{
  const unsigned ph = b.phase;
  if (++b.arrived >= b.expected) { b.arrived = 0; ++b.phase; return; }
  while (b.phase == ph) yield();
}
inline void drop(Bar& b)
{
  --b.expected;
  if (b.arrived > 0 && b.arrived >= b.expected) { b.arrived = 0; ++b.phase; }
}

inline Wave& wave() { return group().waves[threadIdx.x / kLanes]; }
inline int lane() { return (int)(threadIdx.x % kLanes); }

template <class T>
inline uint64_t to_bits(T v) { uint64_t b = 0; std::memcpy(&b, &v, sizeof(T)); return b; }
template <class T>
inline T from_bits(uint64_t b) { T v; std::memcpy(&v, &b, sizeof(T)); return v; }

// Every live lane writes its value, waits, reads lane `src(l)`, waits.
template <class T, class Src>
inline T exchange(T v, Src src)
{
  Wave& w = wave();
  const int l = lane();
  w.slot[l] = to_bits(v);
  arrive_and_wait(w.bar);
  const uint64_t r = w.slot[src(l) & (kLanes - 1)];
  arrive_and_wait(w.bar);
  return from_bits<T>(r);
}

inline uint64_t ballot(bool p)
{
  Wave& w = wave();
  const int l = lane();
  w.slot[l] = p ? 1u : 0u;
  arrive_and_wait(w.bar);
  uint64_t m = 0;
  for (int k = 0; k < kLanes; ++k)
    if (w.alive[k] && w.slot[k]) m |= 1ull << k;
  arrive_and_wait(w.bar);
  return m;
}

inline int first_live()
{
  Wave& w = wave();
  for (int k = 0; k < kLanes; ++k) if (w.alive[k]) return k;
  return 0;
}

template <class... P, class... A>
inline void launch(void (*k)(P...), dim3 g, dim3 b, A&&... a)
{
  if (g.x == 0 || b.x == 0) return;
  if (b.x > (unsigned)kMaxThreads || b.y != 1 || b.z != 1 || g.y != 1 || g.z != 1) std::abort();
  std::tuple<std::decay_t<P>...> args(std::forward<A>(a)...);
  std::function<void()> body = [&] { std::apply(k, args); };
  blockDim = b;
  gridDim = g;
  run_launch(body, g.x, (int)b.x);
}

}  // namespace hipemu

#define hipLaunchKernelGGL(kernel, grid, block, shmem, stream, ...) \
  hipemu::launch(kernel, dim3(grid), dim3(block), ##__VA_ARGS__)

// ---- wave and workgroup operations ---------------------------------------
template <class T>
inline T __shfl_xor(T v, int mask, int = 64) { return hipemu::exchange(v, [mask](int l) { return l ^ mask; }); }
template <class T>
inline T __shfl_up(T v, unsigned d, int = 64)
{
  return hipemu::exchange(v, [d](int l) { return l >= (int)d ? l - (int)d : l; });
}
template <class T>
inline T __shfl(T v, int src, int = 64) { return hipemu::exchange(v, [src](int) { return src; }); }
inline uint64_t __ballot(int p) { return hipemu::ballot(p != 0); }
inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline void __syncthreads() { hipemu::arrive_and_wait(hipemu::group().sync); }

template <class T>
inline T emu_readfirstlane(T v)
{
  const int f = hipemu::first_live();
  return hipemu::exchange(v, [f](int) { return f; });
}
template <class T>
inline T emu_readlane(T v, int r) { return hipemu::exchange(v, [r](int) { return r; }); }
#define __builtin_amdgcn_readfirstlane(v) emu_readfirstlane(v)
#define __builtin_amdgcn_readlane(v, r) emu_readlane((v), (r))
#define __builtin_amdgcn_rsq(x) (1.0 / std::sqrt(x))
#define __builtin_amdgcn_rsqf(x) (1.0f / std::sqrt((float)(x)))
inline long long __double_as_longlong(double v) { long long r; std::memcpy(&r, &v, 8); return r; }
inline double __longlong_as_double(long long v) { double r; std::memcpy(&r, &v, 8); return r; }
inline unsigned int __float_as_uint(float v) { unsigned int r; std::memcpy(&r, &v, 4); return r; }
inline float __uint_as_float(unsigned int v) { float r; std::memcpy(&r, &v, 4); return r; }
inline double __hiloint2double(int hi, int lo)
{
  const unsigned long long b = ((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo;
  return __longlong_as_double((long long)b);
}
#define __builtin_amdgcn_sched_barrier(x) ((void)0)

template <class T, class U>
inline T atomicAdd(T* p, U v) { return __atomic_fetch_add(p, (T)v, __ATOMIC_RELAXED); }
template <class T, class U>
inline T atomicOr(T* p, U v) { return __atomic_fetch_or(p, (T)v, __ATOMIC_RELAXED); }
template <class T, class U>
inline T atomicMax(T* p, U v)
{
  T cur = __atomic_load_n(p, __ATOMIC_RELAXED);
  while (cur < (T)v && !__atomic_compare_exchange_n(p, &cur, (T)v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {}
  return cur;
}

using std::isfinite;
using std::isinf;
using std::isnan;

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

// ---- runtime API ------------------------------------------------------------
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
typedef struct ihipStream_t* hipStream_t;
typedef struct ihipEvent_t* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice,
                     hipMemcpyDefault };
constexpr unsigned hipStreamNonBlocking = 1;
struct hipDeviceProp_t {
  char name[256];
  int multiProcessorCount;
  size_t totalGlobalMem;
};

inline hipError_t hipMalloc(void** p, size_t n)
{
  *p = std::malloc(n ? n : 1);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
template <class T>
inline hipError_t hipMalloc(T** p, size_t n) { return hipMalloc((void**)p, n); }
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { if (n) std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t = nullptr)
{
  return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemset(void* p, int v, size_t n) { if (n) std::memset(p, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void* p, int v, size_t n, hipStream_t = nullptr) { return hipMemset(p, v, n); }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned)
{
  *s = reinterpret_cast<hipStream_t>(new char[1]);
  return hipSuccess;
}
inline hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
inline hipError_t hipStreamDestroy(hipStream_t s) { delete[] reinterpret_cast<char*>(s); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t* p, int)
{
  std::memset(p, 0, sizeof(*p));
  std::strcpy(p->name, "host emulation");
  p->multiProcessorCount = 256;
  p->totalGlobalMem = (size_t)1 << 36;
  return hipSuccess;
}
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipPeekAtLastError() { return hipSuccess; }
inline const char* hipGetErrorString(hipError_t) { return "host emulation error"; }
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = reinterpret_cast<hipEvent_t>(new char[1]); return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete[] reinterpret_cast<char*>(e); return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t = nullptr) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }
#define HIP_SYMBOL(s) (&(s))
template <class T>
inline hipError_t hipGetSymbolAddress(void** p, T* sym) { *p = (void*)sym; return hipSuccess; }

// CPU build of the library's HIP sources (test infrastructure only).
//
// This header stands in for <hip/hip_runtime.h> when hack-wein2k_amd/csrc/*.hip
// is compiled with the host C++ compiler into libw2klapw1_emu.so (see
// tests/cpuemu/Makefile).  Kernels run on the CPU: every work-item is a fiber,
// blocks run one after another, and the wave-level operations the kernels use
// (readlane, shuffles, ballot, DPP moves, swizzle, MFMA) are rendezvous of the
// wave's live lanes.  Device memory comes from an allocator that puts every
// buffer against a guard page and poisons it, so an access outside a buffer,
// a use after free or a read of never-written memory shows on the CPU.
// Streams run synchronously in issue order; graphs are not supported (the
// capture call fails, as a failed capture does on the device).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <algorithm>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local
#define __align__(n) __attribute__((aligned(n)))
#define __restrict__ __restrict

// ------------------------------------------------------------------ host API types
enum hipError_t {
  hipSuccess = 0,
  hipErrorInvalidValue = 1,
  hipErrorOutOfMemory = 2,
  hipErrorInvalidConfiguration = 9,
  hipErrorInvalidDevice = 101,
  hipErrorNotReady = 600,
  hipErrorNotSupported = 801,
  hipErrorLaunchFailure = 719,
};
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4,
};
enum hipStreamCaptureStatus {
  hipStreamCaptureStatusNone = 0,
  hipStreamCaptureStatusActive = 1,
  hipStreamCaptureStatusInvalidated = 2,
};
enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2,
};
#define hipStreamNonBlocking 0x1
#define hipEventDisableTiming 0x2

struct emu_stream;
struct emu_event;
struct emu_graph;
struct emu_graph_exec;
typedef emu_stream *hipStream_t;
typedef emu_event *hipEvent_t;
typedef emu_graph *hipGraph_t;
typedef emu_graph_exec *hipGraphExec_t;
typedef void *hipGraphNode_t;

struct dim3 {
  unsigned x, y, z;
  constexpr dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct alignas(16) double2 {
  double x, y;
};
struct alignas(8) int2 {
  int x, y;
};
struct alignas(16) int4 {
  int x, y, z, w;
};
inline double2 make_double2(double a, double b) { return double2{a, b}; }
struct emu_uint3 {
  unsigned x, y, z;
};

hipError_t hipMalloc(void **p, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T **p, size_t bytes) { return hipMalloc(reinterpret_cast<void **>(p), bytes); }
hipError_t hipFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind k);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t s = nullptr);
hipError_t hipMemcpy2DAsync(void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, size_t height,
                            hipMemcpyKind k, hipStream_t s = nullptr);
hipError_t hipMemset(void *dst, int v, size_t bytes);
hipError_t hipMemsetAsync(void *dst, int v, size_t bytes, hipStream_t s = nullptr);
hipError_t hipSetDevice(int d);
hipError_t hipGetDevice(int *d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipDeviceSynchronize();
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipStreamWaitEvent(hipStream_t s, hipEvent_t e, unsigned flags);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned flags);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventQuery(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode m);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipStreamIsCapturing(hipStream_t s, hipStreamCaptureStatus *st);
hipError_t hipGraphInstantiate(hipGraphExec_t *ex, hipGraph_t g, hipGraphNode_t *err, char *log, size_t len);
hipError_t hipGraphLaunch(hipGraphExec_t ex, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t ex);
hipError_t emu_get_symbol_address(void **p, const void *sym);
#define HIP_SYMBOL(x) (&(x))
#define hipGetSymbolAddress(p, sym) emu_get_symbol_address((p), (const void *)(sym))

// ------------------------------------------------------------------ device side
namespace emu {

// state of the work-item that is running (one fiber per work-item)
struct Item {
  emu_uint3 tid;
  int lane;  // flattened id % 64
};
extern thread_local Item *cur;
extern thread_local dim3 cur_bdim, cur_gdim;
extern thread_local emu_uint3 cur_bidx;

void syncthreads();

typedef double d4 __attribute__((ext_vector_type(4)));
typedef int v4i __attribute__((ext_vector_type(4)));

// wave rendezvous primitives (64-bit payloads)
uint64_t wave_shfl(uint64_t v, int src);         // value of lane src
uint64_t wave_shfl_down(uint64_t v, int delta);  // lane + delta (own value past the end)
uint64_t wave_shfl_xor(uint64_t v, int mask);
uint64_t wave_ballot(int pred);
int wave_readlane(int v, int src);
int wave_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl);
int wave_swizzle(int v, int pattern);
d4 wave_mfma_f64(double a, double b, d4 c);
typedef unsigned u2 __attribute__((ext_vector_type(2)));
u2 permlane_swap(unsigned x, unsigned y, int s);
v4i wave_mfma_i8(v4i a, v4i b, v4i c);

template <class T>
inline uint64_t bits_of(T v)
{
  static_assert(sizeof(T) <= 8, "payload");
  uint64_t u = 0;
  std::memcpy(&u, &v, sizeof(T));
  return u;
}
template <class T>
inline T of_bits(uint64_t u)
{
  T v;
  std::memcpy(&v, &u, sizeof(T));
  return v;
}

// kernel launch: runs `body` for every work-item of the grid
struct LaunchBody {
  virtual void run() = 0;
  virtual ~LaunchBody() {}
};
template <class F>
struct LaunchFn : LaunchBody {
  F &f;
  explicit LaunchFn(F &g) : f(g) {}
  void run() override { f(); }
};
void launch(const char *name, dim3 grid, dim3 block, hipStream_t s, LaunchBody &body);

}  // namespace emu

#define threadIdx (emu::cur->tid)
#define blockIdx (emu::cur_bidx)
#define blockDim (emu::cur_bdim)
#define gridDim (emu::cur_gdim)

#define hipLaunchKernelGGL(K, G, B, SH, S, ...)                  \
  do {                                                          \
    auto emu_body_ = [&]() { K(__VA_ARGS__); };                 \
    emu::LaunchFn<decltype(emu_body_)> emu_fn_(emu_body_);      \
    emu::launch(#K, dim3(G), dim3(B), (S), emu_fn_);            \
  } while (0)

inline void __syncthreads() { emu::syncthreads(); }

template <class T>
inline T __shfl(T v, int src, int width = 64)
{
  return emu::of_bits<T>(emu::wave_shfl(emu::bits_of(v), src));
}
template <class T>
inline T __shfl_down(T v, unsigned delta, int width = 64)
{
  return emu::of_bits<T>(emu::wave_shfl_down(emu::bits_of(v), (int)delta));
}
template <class T>
inline T __shfl_xor(T v, int mask, int width = 64)
{
  return emu::of_bits<T>(emu::wave_shfl_xor(emu::bits_of(v), mask));
}
inline unsigned long long __ballot(int pred) { return emu::wave_ballot(pred); }
inline int __ffsll(long long x) { return __builtin_ffsll(x); }
inline int __ffsll(unsigned long long x) { return __builtin_ffsll((long long)x); }

#define __builtin_amdgcn_readlane(v, l) emu::wave_readlane((v), (l))
// gfx950 permlane swaps: lanes with bit S set take the other register's value
// of their partner (lane ^ S) into the first result, the others into the second
#define __builtin_amdgcn_permlane16_swap(x, y, fi, bc) emu::permlane_swap((x), (y), 16)
#define __builtin_amdgcn_permlane32_swap(x, y, fi, bc) emu::permlane_swap((x), (y), 32)
#define __builtin_amdgcn_readfirstlane(v) ((int)(v))
#define __builtin_amdgcn_update_dpp(o, v, c, rm, bm, bc) emu::wave_dpp((o), (v), (c), (rm), (bm), (bc))
#define __builtin_amdgcn_ds_swizzle(v, p) emu::wave_swizzle((v), (p))
#define __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, x, y, z) emu::wave_mfma_f64((a), (b), (c))
#define __builtin_amdgcn_mfma_i32_16x16x64_i8(a, b, c, x, y, z) emu::wave_mfma_i8((a), (b), (c))
#define __builtin_amdgcn_rcp(x) (1.0 / (x))
#define __builtin_amdgcn_rsq(x) (1.0 / std::sqrt(x))

inline long long __double_as_longlong(double x) { return emu::of_bits<long long>(emu::bits_of(x)); }
inline double __longlong_as_double(long long x) { return emu::of_bits<double>(emu::bits_of(x)); }

inline int atomicMax(int *p, int v)
{
  const int o = *p;
  if (v > o) *p = v;
  return o;
}
inline int atomicAdd(int *p, int v)
{
  const int o = *p;
  *p = o + v;
  return o;
}
inline int atomicExch(int *p, int v)
{
  const int o = *p;
  *p = v;
  return o;
}
inline double __dsqrt_rn(double x) { return std::sqrt(x); }

// HIP's global min/max overloads
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline long min(long a, long b) { return a < b ? a : b; }
inline long max(long a, long b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }
inline unsigned long min(unsigned long a, unsigned long b) { return a < b ? a : b; }
inline unsigned long max(unsigned long a, unsigned long b) { return a > b ? a : b; }
inline double min(double a, double b) { return a < b ? a : b; }
inline double max(double a, double b) { return a > b ? a : b; }
using std::sqrt;
using std::fabs;
using std::fma;
using std::fmax;
using std::fmin;
using std::copysign;
using std::ldexp;

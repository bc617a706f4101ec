// common.hpp -- constants, error plumbing and launch helpers shared by the
// HIP translation units of libtriton_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/triton_amd.h"

namespace tr {

// Physical and numerical constants of the FP64 reference build
// (src/constants.h:43-50,117-118).  SQRTG is the reference's rounded value,
// not sqrt(9.81).
constexpr double kG = 9.81;
constexpr double kSqrtG = 3.132091953;
constexpr double kEps12 = 1e-12;
constexpr double kMaxDt = 1e10;

constexpr int kBlock = 256;  // 1-D launches of the drop-in kernels

// Thread-local error message behind tr_last_error().
void set_error(const std::string& msg);
void clear_error();

inline int hip_fail(hipError_t e, const char* what) {
  set_error(std::string(what) + ": " + hipGetErrorString(e));
  return TR_ERR_HIP;
}

#define TR_HIP(call)                                           \
  do {                                                         \
    hipError_t tr_e_ = (call);                                 \
    if (tr_e_ != hipSuccess) return ::tr::hip_fail(tr_e_, #call); \
  } while (0)

inline int launch_check(const char* name) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, name);
  return TR_OK;
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Pointers read from memory (descriptor tables, structs) are generic to the
// compiler and would use flat instructions; these casts assert they point to
// device global memory so loads and stores become global_* instructions.
#define TR_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const TR_GLOBAL T* gptr(const T* p) {
  return (const TR_GLOBAL T*)p;
}
template <class T>
__device__ __forceinline__ TR_GLOBAL T* gptr(T* p) {
  return (TR_GLOBAL T*)p;
}

// Neighbour-lane moves across the whole 64-lane wave with DPP (wave_shl:1 /
// wave_shr:1, GFX9-family controls): a VALU move instead of an LDS
// permute.  Lane i receives lane i+1 (from_right) or lane i-1 (from_left);
// the edge lane receives 0, written by the hardware (bound_ctrl), so the
// destination needs no initial value and each shift is one instruction.
// Callers use them in wave-uniform control flow.
__device__ __forceinline__ int dpp_from_right(int x) {
  return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ int dpp_from_left(int x) {
  return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ float from_right(float x) {
  return __int_as_float(dpp_from_right(__float_as_int(x)));
}
__device__ __forceinline__ float from_left(float x) {
  return __int_as_float(dpp_from_left(__float_as_int(x)));
}
__device__ __forceinline__ double from_right(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_from_right(int(b & 0xffffffff));
  const int hi = dpp_from_right(int(b >> 32));
  return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
}
__device__ __forceinline__ double from_left(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = dpp_from_left(int(b & 0xffffffff));
  const int hi = dpp_from_left(int(b >> 32));
  return __longlong_as_double((long long)(unsigned(lo)) | ((long long)hi << 32));
}

// Workgroups are handed to the 8 XCDs of the device round-robin (block b to
// XCD b % 8).  xcd_slot(b, n) returns the logical index block b should work
// on so that each XCD gets one contiguous range of the n logical tiles:
// neighbouring strips then share their halo cache lines in one L2.
constexpr int kNumXcd = 8;
__host__ __device__ inline int xcd_slot(int b, int n) {
  const int x = b % kNumXcd, k = b / kNumXcd;
  const int base = n / kNumXcd, rem = n % kNumXcd;
  return x * base + (x < rem ? x : rem) + k;
}

inline unsigned grid_for(int64_t n, int block = kBlock) {
  return unsigned((n + block - 1) / block);
}

// Positive doubles order like their uint64 bit patterns, so an exact min of
// per-cell dt values is an unsigned 64-bit atomic min.
__host__ __device__ inline unsigned long long dbits(double v) {
  union { double d; unsigned long long u; } x;
  x.d = v;
  return x.u;
}
__host__ __device__ inline double bitsd(unsigned long long u) {
  union { double d; unsigned long long u; } x;
  x.u = u;
  return x.d;
}

}  // namespace tr

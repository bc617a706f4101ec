// sround.hpp -- stochastic rounding of stored FP32 depths (and, optionally,
// discharges), shared by the FP32 step and the FP64 step's FP32-storage
// yardstick variant, so both round by the same rule and dither.
#pragma once

#include <hip/hip_runtime.h>

namespace tr {

// a + b == s + lo exactly (Knuth's TwoSum), s returned.
__device__ __forceinline__ float two_sum(float a, float b, float& lo) {
#pragma clang fp contract(off)
  const float s = a + b;
  const float bb = s - a;
  lo = (a - (s - bb)) + (b - bb);
  return s;
}

// Stored depths are rounded stochastically.  With round-to-nearest the FP32
// form lost mass steadily on deep water, where a step changes a cell by a few
// ulps (an ulp is 9.5e-7 m at 10 m): -1.3e-5 relative mass over two periods
// of the 1024^2 Thacker bowl, +8.6e-6 at 512^2, linear in the step count, the
// same with every dry threshold and with IEEE division, square root and no
// FMA contraction; one FP32 step from an FP64 state showed no such loss, one
// from an FP32-evolved state did.  Rounded stochastically: +1.2e-9 and
// -2.9e-9.  The stored depth is hs + he plus a dither uniform over one ulp of
// hs, rounded to nearest: it lands on the float one ulp away in he's
// direction with probability |he| / ulp (|he| <= ulp/2), else on hs, so its
// expected value is hs + he (a dither as wide as a grid spacing keeps this
// where the spacing halves below a power of two).  The dither is x 2^-32 ulp,
// x = (column term) + fmix32(row and iteration term): hashed, so a cell's
// dither at the next step is unrelated to its neighbours' dithers now (with
// a dither linear in column, row and step, a Weyl sequence, the loss was
// still -5.6e-6: the neighbours' roundings feed the cell's next residual).
// It depends only on the member's cell and iteration count, so a member's
// bits do not depend on the batch or the partition, and a hotstart continues
// the same sequence.  (Below 2^-71 m the scaled ulp underflows and the depth
// is rounded to nearest.)
__device__ __forceinline__ float sround(float hs, float he, int x) {
  const float ulp_2m32 = __int_as_float(__float_as_int(hs) & 0x7f800000) * 0x1p-55f;
  return fmaxf(hs + fmaf(float(x), ulp_2m32, he), 0.f);
}
// The same stochastic rounding for a signed value (a discharge qs + qe).
__device__ __forceinline__ float sround_signed(float qs, float qe, int x) {
  const float ulp_2m32 = __int_as_float(__float_as_int(qs) & 0x7f800000) * 0x1p-55f;
  return qs + fmaf(float(x), ulp_2m32, qe);
}
// m - dtdx*S carried as ms + me (TwoSum and an FMA residual)
__device__ __forceinline__ float sub_scaled2(float m, float dtdx, float S, float& me) {
#pragma clang fp contract(off)
  const float d = dtdx * S;
  const float de = fmaf(dtdx, S, -d);
  float e;
  const float ms = two_sum(m, -d, e);
  me = e - de;
  return ms;
}
constexpr unsigned kDitherCol = 0x85EBCA77u;
constexpr unsigned kDitherRow = 0xC2B2AE3Du;
constexpr unsigned kDitherStep = 0x9E3779B9u;  // 2^32 / golden ratio
// salts of the discharges' row hashes (other streams than the depth's)
constexpr unsigned kDitherQx = 0x632BE5ABu;
constexpr unsigned kDitherQy = 0x1B873593u;
// MurmurHash3's 32-bit finaliser: the row-and-iteration part of the dither
// (wave-uniform, so it runs on the scalar unit)
__device__ __forceinline__ unsigned fmix32(unsigned h) {
  h ^= h >> 16;
  h *= 0x85EBCA6Bu;
  h ^= h >> 13;
  h *= 0xC2B2AE35u;
  h ^= h >> 16;
  return h;
}

// A double rounded to FP32 by the same rule: hs = v to nearest, he = the
// remainder (exact in FP64, then to FP32).
__device__ __forceinline__ float sround_f64(double v, int x) {
  const float hs = float(v);
  const float he = float(v - double(hs));
  return sround(hs, he, x);
}

}  // namespace tr

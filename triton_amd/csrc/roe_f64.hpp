// roe_f64.hpp -- the reference's per-face Roe solver with bed-slope source,
// per-cell CFL step and cell update, in the reference's floating-point
// operation order so the GPU reproduces the CPU builds bit for bit
// (src/kernels_flux.h:81-265 for x, :364-547 for y; src/kernels_state.h:62-101;
// src/kernels.h:176-185).  Written, as the reference is, as templates on the
// state type T with double literals and double constants (_G_, SQRTG,
// 0.5 ...): with T = double every operation is FP64; with T = float the
// mixed expressions are evaluated in FP64 and rounded when stored, exactly as
// the reference's USE_SINGLE_PRECISION build does (src/constants.h:37-50:
// EPS12 = 1e-6, FMAX/FMIN = fmaxf/fminf), with the math functions taking and
// returning double as the C library's do.  Translation units including this
// file are compiled with -ffp-contract=off.
#pragma once

#include "common.hpp"

namespace tr {

// EPS12 of the reference's build for state type T
template <class T>
__host__ __device__ constexpr double eps12_of() {
  return sizeof(T) == 4 ? 1e-6 : 1e-12;
}
// FMAX / FMIN of the reference's build: arguments converted to T first
template <class T>
__device__ __forceinline__ T tmax(T a, T b) {
  return fmax(a, b);
}
template <class T>
__device__ __forceinline__ T tmin(T a, T b) {
  return fmin(a, b);
}
template <>
__device__ __forceinline__ float tmax<float>(float a, float b) {
  return fmaxf(a, b);
}
template <>
__device__ __forceinline__ float tmin<float>(float a, float b) {
  return fminf(a, b);
}

// a / b for many a and one b, with the bits of the compiler's IEEE FP64
// division on this GPU.  That division scales its operands (v_div_scale),
// refines a hardware reciprocal of b with two Newton steps, forms the
// quotient with one correction step (v_div_fmas) and fixes up special cases
// (v_div_fixup).  When a is +0 or a normal number of moderate size and b is a
// normal number of moderate size -- 2^-900 < |a|, 2^-100 < b < 2^100,
// |a/b| < 2^900 -- neither operand is scaled, the correction is a plain fma
// and the fix-up returns its input, so the refined reciprocal can be formed
// once for b (FixedDiv64) and each quotient costs three instructions.
// Callers guarantee that range (div_fixed_ok).
struct FixedDiv64 {
  double b, y;  // divisor and its refined reciprocal
};
__device__ __forceinline__ FixedDiv64 fixed_div64(double b) {
  const double r0 = __builtin_amdgcn_rcp(b);
  const double e0 = fma(-b, r0, 1.0);
  const double r1 = fma(r0, e0, r0);
  const double e1 = fma(-b, r1, 1.0);
  return FixedDiv64{b, fma(r1, e1, r1)};
}
__device__ __forceinline__ double div_fixed(double a, const FixedDiv64& d) {
  const double m = a * d.y;
  const double f = fma(-d.b, m, a);
  return fma(f, d.y, m);
}
__host__ __device__ inline bool div_fixed_ok(double b) { return b > 0x1p-100 && b < 0x1p100; }
// A value every lane holds, made explicitly wave-uniform (read from the first
// lane) so the compiler keeps it in scalar registers.
__device__ __forceinline__ double uniform_f64(double x) {
  const unsigned long long v = __double_as_longlong(x);
  const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(v));
  const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(v >> 32));
  return __longlong_as_double((long long)((static_cast<unsigned long long>(hi) << 32) | lo));
}

template <class T>
struct FaceT {
  T d1[3];  // fluctuation sent to the west (x) / south (y) cell
  T d2[3];  // fluctuation sent to the east (x) / north (y) cell
};
using Face64 = FaceT<double>;

// Per-cell inputs of the face solver that the reference recomputes in every
// face (src/kernels_flux.h:98-117 velocities, :195-200 the wet-wet clip's
// h dx 0.5 / dt): evaluated once per cell with the same operations on the
// same operands, so the face results keep their bits.  hd is the double the
// reference's expression yields for either T.
template <class T>
struct CellT {
  T u, v;
  double hd;
};
using Cell64 = CellT<double>;
template <class T>
__device__ __forceinline__ CellT<T> cell_t(T h, T qx, T qy, T dx, T dt, T hextra) {
  CellT<T> k{T(0), T(0), h * dx * 0.5 / dt};
  if (h >= hextra) {
    k.u = qx / h;
    k.v = qy / h;
  }
  return k;
}
__device__ __forceinline__ Cell64 cell64(double h, double qx, double qy, double dx, double dt,
                                         double hextra) {
  return cell_t<double>(h, qx, qy, dx, dt, hextra);
}

// Returns false (and leaves f untouched) when both cells are dry.  NX/NY is
// the face normal, (1,0) for x faces and (0,1) for y faces; the products with
// 0 and 1 are kept because they decide signed zeros exactly as the
// reference's generic code does.  c1, c2: the cells' CellT values.
template <class T, int NX, int NY>
__device__ __forceinline__ bool roe_face_t(T h1, T h2, T qx1, T qx2, T qy1, T qy2, const CellT<T>& c1,
                                           const CellT<T>& c2, T z1, T z2, T sqh1, T sqh2,
                                           FaceT<T>& f) {
  constexpr double eps = eps12_of<T>();
  if (!(h1 > 0.0 || h2 > 0.0)) return false;
  const T nx = T(NX), ny = T(NY);
  const T u1 = c1.u, v1 = c1.v, u2 = c2.u, v2 = c2.v;

  const T hmid = 0.5 * (h1 + h2);
  const T dh = h2 - h1;
  const T c = sqrt(kG * hmid);
  const T ubar = (u1 * sqh1 + u2 * sqh2) / (sqh1 + sqh2);
  const T vbar = (v1 * sqh1 + v2 * sqh2) / (sqh1 + sqh2);
  const T un = ubar * nx + vbar * ny;
  T l0 = un - c, l1 = un, l2 = un + c;
  T e0 = T(0), e2 = T(0);  // entropy-fix wave speeds (lamE[0], lamE[2])

  {
    T eL = u1 * nx + v1 * ny - kSqrtG * sqh1;
    T eR = u2 * nx + v2 * ny - kSqrtG * sqh2;
    if (eL < 0.0 && eR > 0.0) {
      T split = eL * (eR - l0) / (eR - eL);
      e0 = l0 - split;
      l0 = split;
    }
    eL = u1 * nx + v1 * ny + kSqrtG * sqh1;
    eR = u2 * nx + v2 * ny + kSqrtG * sqh2;
    if (eL < 0.0 && eR > 0.0) {
      T split = eR * (l2 - eL) / (eR - eL);
      e2 = l2 - split;
      l2 = split;
    }
  }

  // eigenvectors (row i = wave i)
  const T ev01 = ubar - c * nx, ev02 = vbar - c * ny;
  const T ev11 = -c * ny, ev12 = c * nx;
  const T ev21 = ubar + c * nx, ev22 = vbar + c * ny;

  const T dqx = qx2 - qx1, dqy = qy2 - qy1;
  const T a0 = 0.5 * (dh - ((dqx * nx + dqy * ny) - un * dh) / c);
  const T a1 = ((dqy - vbar * dh) * nx - (dqx - ubar * dh) * ny) / c;
  const T a2 = 0.5 * (dh + ((dqx * nx + dqy * ny) - un * dh) / c);

  const T dz = z2 - z1;
  T dzp = dz, hp;
  if (dz >= 0.0) {
    hp = h1;
    if (z1 + h1 < z2) dzp = h1;
  } else {
    hp = h2;
    if (z2 + h2 < z1) dzp = -h2;
  }
  T b0 = 0.5 * kG * (hp - 0.5 * fabs(dzp)) * dzp / c;

  hp = h1 + a0;
  if (l0 * l2 < 0.0 && hp > 0.0 && h1 > 0.0 && h2 > 0.0) {
    // l0 / (l0 + e2) is exactly 1 when e2 == 0 (no entropy fix on wave 2):
    // here l0 != 0 (l0 l2 < 0), so only a lane with e2 != 0 or l0 not finite
    // needs the quotient, and a wave without one skips both divisions
    T f0 = T(1), f2 = T(1);
    if (__any(e0 != 0.0 || e2 != 0.0 || !isfinite(l0) || !isfinite(l2))) {
      f0 = l0 / (l0 + e2);
      f2 = l2 / (l2 + e0);
    }
    b0 = tmax<T>(b0, f0 * (a0 * l0 - c1.hd + e2 * h2));
    b0 = tmin<T>(b0, f2 * (-a2 * l2 + c2.hd + e0 * h1));
  }
  const T b2 = -b0;

  const T one = T(1), zero = T(0);
  T d10 = zero, d11 = zero, d12 = zero, d20 = zero, d21 = zero, d22 = zero;
  // the dry-side tests hold only where a side is below eps, so a wave with
  // no such side skips their two divisions
  bool to_left = false, to_right = false;
  if (__any(h1 < eps || h2 < eps)) {
    const T sL = hp - b0 / l0;
    const T sR = hp + b2 / l2;
    to_left = sR < -eps && h2 < eps;
    to_right = sL < -eps && h1 < eps;
  }
  // A face with no dry-side redirection, no entropy fix, wave 0 to the back
  // side and wave 2 to the front (l0 <= 0 < l2), and finite wave strengths
  // and eigenvectors -- almost every face of a wet catchment -- receives,
  // besides the three waves, only products with zero (the entropy parts,
  // lamE[1] and wave 1's weight on the mass).  Those add signed zeros, which
  // can change only the sign of a zero sum, and the flush below turns every
  // zero into +0, so a wave whose faces are all of this kind skips them and
  // the general path's selects: the same flushed values.
  const bool plain = !to_left && !to_right && !(l0 > 0.0) && l2 > 0.0 && e0 == 0.0 && e2 == 0.0 &&
                     isfinite(a0) && isfinite(a1) && isfinite(a2) && isfinite(ubar) && isfinite(vbar) &&
                     isfinite(c);
  if (__all(plain)) {
    const T w0 = l0 * a0 - b0;
    const T w1 = l1 * a1 - zero;
    const T w2 = l2 * a2 - b2;
    d10 = w0 * one;
    d11 = w0 * ev01;
    d12 = w0 * ev02;
    d20 = w2 * one;
    d21 = w2 * ev21;
    d22 = w2 * ev22;
    if (l1 > 0.0) {
      d21 = w1 * ev11 + d21;
      d22 = w1 * ev12 + d22;
    } else {
      d11 = d11 + w1 * ev11;
      d12 = d12 + w1 * ev12;
    }
  } else {
    if (to_left) {
      d10 += (l0 * a0 - b0) + (l2 * a2 - b2);
    } else if (to_right) {
      d20 += (l0 * a0 - b0) + (l2 * a2 - b2);
    } else {
      // wave 0: eigenvector (1, ev01, ev02)
      T w = l0 * a0 - b0;
      if (l0 > 0.0) { d20 += w * one; d21 += w * ev01; d22 += w * ev02; }
      else          { d10 += w * one; d11 += w * ev01; d12 += w * ev02; }
      // wave 1: eigenvector (0, ev11, ev12), beta1 = 0
      w = l1 * a1 - zero;
      if (l1 > 0.0) { d20 += w * zero; d21 += w * ev11; d22 += w * ev12; }
      else          { d10 += w * zero; d11 += w * ev11; d12 += w * ev12; }
      // wave 2: eigenvector (1, ev21, ev22)
      w = l2 * a2 - b2;
      if (l2 > 0.0) { d20 += w * one; d21 += w * ev21; d22 += w * ev22; }
      else          { d10 += w * one; d11 += w * ev21; d12 += w * ev22; }
    }
    // entropy-fix parts; lamE[1] is always 0
    {
      T w = e0 * a0;
      if (e0 > 0.0) { d20 += w * one; d21 += w * ev01; d22 += w * ev02; }
      else          { d10 += w * one; d11 += w * ev01; d12 += w * ev02; }
      w = zero * a1;
      d10 += w * zero; d11 += w * ev11; d12 += w * ev12;
      w = e2 * a2;
      if (e2 > 0.0) { d20 += w * one; d21 += w * ev21; d22 += w * ev22; }
      else          { d10 += w * one; d11 += w * ev21; d12 += w * ev22; }
    }
  }
  f.d1[0] = fabs(d10) < eps ? zero : d10;
  f.d1[1] = fabs(d11) < eps ? zero : d11;
  f.d1[2] = fabs(d12) < eps ? zero : d12;
  f.d2[0] = fabs(d20) < eps ? zero : d20;
  f.d2[1] = fabs(d21) < eps ? zero : d21;
  f.d2[2] = fabs(d22) < eps ? zero : d22;
  return true;
}

template <int NX, int NY>
__device__ __forceinline__ bool roe_face_f64(double h1, double h2, double qx1, double qx2,
                                             double qy1, double qy2, const Cell64& c1,
                                             const Cell64& c2, double z1, double z2,
                                             double sqh1, double sqh2, Face64& f) {
  return roe_face_t<double, NX, NY>(h1, h2, qx1, qx2, qy1, qy2, c1, c2, z1, z2, sqh1, sqh2, f);
}

// The face with the reference's per-face argument set (the drop-in kernels).
template <class T, int NX, int NY>
__device__ __forceinline__ bool roe_face_ref(T h1, T h2, T qx1, T qx2, T qy1, T qy2, T z1, T z2,
                                             T sqh1, T sqh2, T dx, T dt, T hextra, FaceT<T>& f) {
  return roe_face_t<T, NX, NY>(h1, h2, qx1, qx2, qy1, qy2, cell_t<T>(h1, qx1, qy1, dx, dt, hextra),
                               cell_t<T>(h2, qx2, qy2, dx, dt, hextra), z1, z2, sqh1, sqh2, f);
}

// src/kernels.h:176-185 -- per-cell CFL step (MAX_VALUE when shallower than
// hextra) and the sqrt(h) cache value; cn_dx is the reference's cn*dx in T.
template <class T>
__device__ __forceinline__ T cell_dt_t(T h, T qx, T qy, T sq, T cn_dx, T hextra) {
  if (h > hextra) {
    const T umax = tmax<T>(fabs(qx), fabs(qy)) / h;
    return cn_dx / (umax + kSqrtG * sq);
  }
  return T(kMaxDt);
}
__device__ __forceinline__ double cell_dt_f64(double h, double qx, double qy, double sq,
                                              double cn_dx, double hextra) {
  return cell_dt_t<double>(h, qx, qy, sq, cn_dx, hextra);
}

// src/kernels_state.h:62-101 -- new depth and discharge from the RHS sums.
template <class T>
__device__ __forceinline__ void update_cell_t(T hn, T qxn, T qyn, T n2, T rh0, T rh1, T rqx0, T rqx1,
                                              T rqy0, T rqy1, T dt, T hextra, T& h_out, T& qx_out,
                                              T& qy_out) {
  constexpr double eps = eps12_of<T>();
  T hnew = hn - rh0 - rh1;
  if (hnew < 0.0) hnew = T(0);
  if (hnew < hextra) {
    qx_out = T(0);
    qy_out = T(0);
  } else {
    T mx = qxn - rqx0 - rqx1;
    T my = qyn - rqy0 - rqy1;
    T modm = sqrt(double(mx * mx + my * my));
    if (n2 > eps && hn >= hextra && modm > eps) {
      T tt = dt * kG * n2 * modm / (hn * hn * cbrt(double(hn)));
      qx_out = -0.5 * (mx - mx * sqrt(1.0 + 4.0 * tt)) / tt;
      qy_out = -0.5 * (my - my * sqrt(1.0 + 4.0 * tt)) / tt;
    } else {
      qx_out = mx;
      qy_out = my;
    }
  }
  h_out = hnew;
}
__device__ __forceinline__ void update_cell_f64(double hn, double qxn, double qyn, double n2,
                                                double rh0, double rh1, double rqx0,
                                                double rqx1, double rqy0, double rqy1,
                                                double dt, double hextra, double& h_out,
                                                double& qx_out, double& qy_out) {
  update_cell_t<double>(hn, qxn, qyn, n2, rh0, rh1, rqx0, rqx1, rqy0, rqy1, dt, hextra, h_out, qx_out,
                        qy_out);
}

}  // namespace tr

// roe_f64.hpp -- FP64 per-face Roe solver with bed-slope source, in the
// reference's floating-point operation order so the GPU reproduces the FP64
// CPU build bit for bit (src/kernels_flux.h:81-265 for x, :364-547 for y).
// Translation units including this file are compiled with -ffp-contract=off.
#pragma once

#include "common.hpp"

namespace tr {

struct Face64 {
  double d1[3];  // fluctuation sent to the west (x) / south (y) cell
  double d2[3];  // fluctuation sent to the east (x) / north (y) cell
};

// Per-cell inputs of the face solver that the reference recomputes in every
// face (src/kernels_flux.h:98-117 velocities, :195-200 the wet-wet clip's
// h dx 0.5 / dt): evaluated once per cell with the same operations on the
// same operands, so the face results keep their bits.
struct Cell64 {
  double u, v, hd;
};
__device__ __forceinline__ Cell64 cell64(double h, double qx, double qy, double dx, double dt,
                                         double hextra) {
  Cell64 k{0.0, 0.0, h * dx * 0.5 / dt};
  if (h >= hextra) {
    k.u = qx / h;
    k.v = qy / h;
  }
  return k;
}

// Returns false (and leaves f untouched) when both cells are dry.  NX/NY is
// the face normal, (1,0) for x faces and (0,1) for y faces; the products with
// 0 and 1 are kept because they decide signed zeros exactly as the
// reference's generic code does.  c1, c2: the cells' Cell64 values.
template <int NX, int NY>
__device__ __forceinline__ bool roe_face_f64(double h1, double h2, double qx1, double qx2,
                                             double qy1, double qy2, const Cell64& c1,
                                             const Cell64& c2, double z1, double z2,
                                             double sqh1, double sqh2, Face64& f) {
  if (!(h1 > 0.0 || h2 > 0.0)) return false;
  const double nx = double(NX), ny = double(NY);
  const double u1 = c1.u, v1 = c1.v, u2 = c2.u, v2 = c2.v;

  const double hmid = 0.5 * (h1 + h2);
  const double dh = h2 - h1;
  const double c = sqrt(kG * hmid);
  const double ubar = (u1 * sqh1 + u2 * sqh2) / (sqh1 + sqh2);
  const double vbar = (v1 * sqh1 + v2 * sqh2) / (sqh1 + sqh2);
  const double un = ubar * nx + vbar * ny;
  double l0 = un - c, l1 = un, l2 = un + c;
  double e0 = 0.0, e2 = 0.0;  // entropy-fix wave speeds (lamE[0], lamE[2])

  {
    double eL = u1 * nx + v1 * ny - kSqrtG * sqh1;
    double eR = u2 * nx + v2 * ny - kSqrtG * sqh2;
    if (eL < 0.0 && eR > 0.0) {
      double split = eL * (eR - l0) / (eR - eL);
      e0 = l0 - split;
      l0 = split;
    }
    eL = u1 * nx + v1 * ny + kSqrtG * sqh1;
    eR = u2 * nx + v2 * ny + kSqrtG * sqh2;
    if (eL < 0.0 && eR > 0.0) {
      double split = eR * (l2 - eL) / (eR - eL);
      e2 = l2 - split;
      l2 = split;
    }
  }

  // eigenvectors (row i = wave i)
  const double ev01 = ubar - c * nx, ev02 = vbar - c * ny;
  const double ev11 = -c * ny, ev12 = c * nx;
  const double ev21 = ubar + c * nx, ev22 = vbar + c * ny;

  const double dqx = qx2 - qx1, dqy = qy2 - qy1;
  const double a0 = 0.5 * (dh - ((dqx * nx + dqy * ny) - un * dh) / c);
  const double a1 = ((dqy - vbar * dh) * nx - (dqx - ubar * dh) * ny) / c;
  const double a2 = 0.5 * (dh + ((dqx * nx + dqy * ny) - un * dh) / c);

  const double dz = z2 - z1;
  double dzp = dz, hp;
  if (dz >= 0.0) {
    hp = h1;
    if (z1 + h1 < z2) dzp = h1;
  } else {
    hp = h2;
    if (z2 + h2 < z1) dzp = -h2;
  }
  double b0 = 0.5 * kG * (hp - 0.5 * fabs(dzp)) * dzp / c;

  hp = h1 + a0;
  if (l0 * l2 < 0.0 && hp > 0.0 && h1 > 0.0 && h2 > 0.0) {
    b0 = fmax(b0, l0 / (l0 + e2) * (a0 * l0 - c1.hd + e2 * h2));
    b0 = fmin(b0, l2 / (l2 + e0) * (-a2 * l2 + c2.hd + e0 * h1));
  }
  const double b2 = -b0;

  double d10 = 0.0, d11 = 0.0, d12 = 0.0, d20 = 0.0, d21 = 0.0, d22 = 0.0;
  const double sL = hp - b0 / l0;
  const double sR = hp + b2 / l2;
  if (sR < -kEps12 && h2 < kEps12) {
    d10 += (l0 * a0 - b0) + (l2 * a2 - b2);
  } else if (sL < -kEps12 && h1 < kEps12) {
    d20 += (l0 * a0 - b0) + (l2 * a2 - b2);
  } else {
    // wave 0: eigenvector (1, ev01, ev02)
    double w = l0 * a0 - b0;
    if (l0 > 0.0) { d20 += w * 1.0; d21 += w * ev01; d22 += w * ev02; }
    else          { d10 += w * 1.0; d11 += w * ev01; d12 += w * ev02; }
    // wave 1: eigenvector (0, ev11, ev12), beta1 = 0
    w = l1 * a1 - 0.0;
    if (l1 > 0.0) { d20 += w * 0.0; d21 += w * ev11; d22 += w * ev12; }
    else          { d10 += w * 0.0; d11 += w * ev11; d12 += w * ev12; }
    // wave 2: eigenvector (1, ev21, ev22)
    w = l2 * a2 - b2;
    if (l2 > 0.0) { d20 += w * 1.0; d21 += w * ev21; d22 += w * ev22; }
    else          { d10 += w * 1.0; d11 += w * ev21; d12 += w * ev22; }
  }
  // entropy-fix parts; lamE[1] is always 0
  {
    double w = e0 * a0;
    if (e0 > 0.0) { d20 += w * 1.0; d21 += w * ev01; d22 += w * ev02; }
    else          { d10 += w * 1.0; d11 += w * ev01; d12 += w * ev02; }
    w = 0.0 * a1;
    d10 += w * 0.0; d11 += w * ev11; d12 += w * ev12;
    w = e2 * a2;
    if (e2 > 0.0) { d20 += w * 1.0; d21 += w * ev21; d22 += w * ev22; }
    else          { d10 += w * 1.0; d11 += w * ev21; d12 += w * ev22; }
  }
  f.d1[0] = fabs(d10) < kEps12 ? 0.0 : d10;
  f.d1[1] = fabs(d11) < kEps12 ? 0.0 : d11;
  f.d1[2] = fabs(d12) < kEps12 ? 0.0 : d12;
  f.d2[0] = fabs(d20) < kEps12 ? 0.0 : d20;
  f.d2[1] = fabs(d21) < kEps12 ? 0.0 : d21;
  f.d2[2] = fabs(d22) < kEps12 ? 0.0 : d22;
  return true;
}

// The face with the reference's per-face argument set (the drop-in kernels).
template <int NX, int NY>
__device__ __forceinline__ bool roe_face_f64(double h1, double h2, double qx1, double qx2,
                                             double qy1, double qy2, double z1, double z2,
                                             double sqh1, double sqh2, double dx, double dt,
                                             double hextra, Face64& f) {
  return roe_face_f64<NX, NY>(h1, h2, qx1, qx2, qy1, qy2, cell64(h1, qx1, qy1, dx, dt, hextra),
                              cell64(h2, qx2, qy2, dx, dt, hextra), z1, z2, sqh1, sqh2, f);
}

// src/kernels.h:176-185 -- per-cell CFL step (MAX_VALUE when shallower than
// hextra) and the sqrt(h) cache value.
__device__ __forceinline__ double cell_dt_f64(double h, double qx, double qy, double sq,
                                              double cn_dx, double hextra) {
  if (h > hextra) {
    double umax = fmax(fabs(qx), fabs(qy)) / h;
    return cn_dx / (umax + kSqrtG * sq);
  }
  return kMaxDt;
}

// src/kernels_state.h:62-101 -- new depth and discharge from the RHS sums.
__device__ __forceinline__ void update_cell_f64(double hn, double qxn, double qyn, double n2,
                                                double rh0, double rh1, double rqx0,
                                                double rqx1, double rqy0, double rqy1,
                                                double dt, double hextra, double& h_out,
                                                double& qx_out, double& qy_out) {
  double hnew = hn - rh0 - rh1;
  if (hnew < 0.0) hnew = 0.0;
  if (hnew < hextra) {
    qx_out = 0.0;
    qy_out = 0.0;
  } else {
    double mx = qxn - rqx0 - rqx1;
    double my = qyn - rqy0 - rqy1;
    double modm = sqrt(mx * mx + my * my);
    if (n2 > kEps12 && hn >= hextra && modm > kEps12) {
      double tt = dt * kG * n2 * modm / (hn * hn * cbrt(hn));
      qx_out = -0.5 * (mx - mx * sqrt(1.0 + 4.0 * tt)) / tt;
      qy_out = -0.5 * (my - my * sqrt(1.0 + 4.0 * tt)) / tt;
    } else {
      qx_out = mx;
      qy_out = my;
    }
  }
  h_out = hnew;
}

}  // namespace tr

// step_f32.hip -- FP32 fused timestep for a batch of members, one launch per
// step for every member.
//
// Storage per cell: depth h, discharges qx, qy (FP32, ping-pong), the bed as
// two per-face differences dzx = z[c+1]-z[c] and dzy = z[r-1]-z[r] computed in
// FP64 at load and stored FP32, and the squared Manning n (FP32, only when
// some cell has friction).  Simulated time, dt bookkeeping, rain and inflow
// volumes stay FP64 in the member's control block.
//
// The wave layout is that of the FP64 step (60 output columns per 64-lane
// wave, rows streamed through registers with a two-row halo); the next row's
// loads are issued before the current row is computed so every wave keeps
// two rows of HBM traffic in flight.  Arithmetic is FP32 with hardware
// reciprocal / square root; per-cell quantities (velocities, sqrt(h)) are
// computed once per cell and passed to the neighbour lane.  The solver is
// written in normal/tangential components so x and y faces share one routine.
// Semantics follow the reference (src/kernels_flux.h, src/kernels_state.h,
// src/kernels_wetdry.h, src/kernels.h:102-187) with these FP32 choices:
//   * the dry / flush threshold is a member setting (default 1e-12, the
//     reference's FP64 value; its FP32 build uses 1e-6, src/constants.h:43);
//   * bed tests z1+h1<z2 become h1<dz on the stored face difference;
//   * the semi-implicit friction uses q = 2m/(1+sqrt(1+4tt)), the same value
//     as -0.5(m - m sqrt(1+4tt))/tt without its cancellation for small tt;
//   * the updated depth (flux update, rain and point sources) is carried as
//     two floats and rounded to FP32 stochastically, which keeps mass
//     (sround).
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "common.hpp"
#include "sround.hpp"
#include "step_ctrl.hpp"

#pragma clang fp contract(fast)

namespace tr {
namespace {

constexpr int kLanes = 64;
constexpr int kOutCols = 60;
constexpr float kGf = 9.81f;
constexpr float kSqrtGf = 3.132091953f;
constexpr float kMaxDtF = 1e10f;
// 16 units of FP32 round-off (2^-20): above the few roundings that separate
// the two halves of a balanced wave strength, far below any resolved flow.
constexpr float kFlushRel = 9.5367431640625e-7f;

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
// 1/sqrt(x) for x >= 0; subnormal inputs (depths at wetting fronts) are
// raised to 1e-30 first because the hardware rsq returns +inf for them, and
// h*rsq(h) must stay finite.
__device__ __forceinline__ float rsqf(float x) { return __builtin_amdgcn_rsqf(fmaxf(x, 1e-30f)); }
__device__ __forceinline__ float sqrtf_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

// The reference's h < 0 -> 0 on the two-float depth hs + he (hs == 0 with
// he < 0 is left to the final clamp of the rounded depth).
__device__ __forceinline__ void clamp_neg(float& hs, float& he) {
  const bool neg = hs < 0.f;
  hs = neg ? 0.f : hs;
  he = neg ? 0.f : he;
}

struct Side {
  float h, qn, qt, un, ut, sq;
};

// One face in normal/tangential components: fluctuations sent to the cell
// behind the face (l*) and in front of it (r*) for mass, normal and
// tangential momentum, before the dt/dx scaling.  Same scheme as
// src/kernels_flux.h:81-265, arranged for FP32 throughput:
//  * 1/c and c from one reciprocal square root;
//  * the wave-1 part reduces to ub*(dqt - vb*dh) on the tangential momentum;
//  * the dry-side tests hp - b0/l < -eps are evaluated as sign tests of
//    ((hp+eps)*l - b0)*l, without the division;
//  * the wet-wet clip bounds need a reciprocal only under an entropy fix;
//  * the Harten-Hyman entropy fix runs only in waves where some lane needs it;
//  * the front side is the total fluctuation minus the back side;
//  * the dry-side tests and the flushes run only on faces with a dry side:
//    the final flush of |d| < threshold on the mass components, and a
//    relative flush of the two wave strengths that removes FP32 round-off
//    from balanced (zero) fluctuations;
//  * a face with both sides dry (h = 0) needs no mask: with the Roe-average
//    reciprocal kept finite every wave strength, source and fluctuation of
//    it is an exact zero.
__device__ __forceinline__ void face32(const Side& L, const Side& R, float dz, float hdxdt,
                                       float eps, float& lm, float& ln, float& lt, float& rm,
                                       float& rn, float& rt) {
  const float g2 = (0.5f * kGf) * (L.h + R.h);
  const float ic = rsqf(g2);
  const float c = g2 * ic;
  const float is = rcpf(fmaxf(L.sq + R.sq, 1e-30f));
  const float ub = (L.un * L.sq + R.un * R.sq) * is;
  const float vb = (L.ut * L.sq + R.ut * R.sq) * is;
  float l0 = ub - c, l2 = ub + c, e0 = 0.f, e2 = 0.f;
  const float aL0 = L.un - kSqrtGf * L.sq, aR0 = R.un - kSqrtGf * R.sq;
  const float aL2 = L.un + kSqrtGf * L.sq, aR2 = R.un + kSqrtGf * R.sq;
  const bool fix0 = aL0 < 0.f && aR0 > 0.f;
  const bool fix2 = aL2 < 0.f && aR2 > 0.f;
  {
    // The face a wave sees almost everywhere in a rained-on catchment (on
    // the ensemble at peak rain, 99.8-100 % of the wave-rows of the members
    // sampled): every lane's face has both sides wet, is subcritical
    // (l0 < 0 < l2) with hp > 0, so the wet-wet clip applies, and needs no
    // entropy fix.  Such a wave takes the general path's operations without
    // its tests, selects and branches.  Which path a wave takes depends only
    // on the cells of its row window, so a member's bits still do not depend
    // on the batch it runs in.
    const float dh = R.h - L.h;
    const float w = ((R.qn - L.qn) - ub * dh) * ic;
    const float a0 = 0.5f * (dh - w), a2 = 0.5f * (dh + w);
    const float hp = L.h + a0;
    const bool common = !fix0 && !fix2 && l0 < 0.f && l2 > 0.f && hp > 0.f && L.h >= eps && R.h >= eps;
    if (__all(common)) {
      const float t1 = ub * ((R.qt - L.qt) - vb * dh);
      const bool up = dz >= 0.f;
      const float hp0 = up ? L.h : R.h;
      const float dzp = hp0 < fabsf(dz) ? (up ? hp0 : -hp0) : dz;
      const float b0 = (0.5f * kGf) * (hp0 - 0.5f * fabsf(dzp)) * dzp * ic;
      const float W0 = l0 * a0 - b0, W2 = l2 * a2 + b0;
      const float Ld = L.h * hdxdt, Rd = R.h * hdxdt;
      const float S = W0 + W2;
      const float w0 = fmaxf(fminf(W0, Ld), S - Rd);
      const float w2 = fminf(fmaxf(W2, S - Ld), Rd);
      // the back side takes wave 0 (l0 < 0), the front side wave 2
      const float bm = w0, bn = w0 * l0, bt = w0 * vb + (ub > 0.f ? 0.f : t1);
      const float tm = w0 + w2, tn = w0 * l0 + w2 * l2, tt = tm * vb + t1;
      lm = bm;
      ln = bn;
      lt = bt;
      rm = tm - bm;
      rn = tn - bn;
      rt = tt - bt;
      return;
    }
  }
  if (__any(fix0 || fix2)) {
    const float sp0 = aL0 * (aR0 - l0) * rcpf(aR0 - aL0);
    const float sp2 = aR2 * (l2 - aL2) * rcpf(aR2 - aL2);
    e0 = fix0 ? l0 - sp0 : 0.f;
    l0 = fix0 ? sp0 : l0;
    e2 = fix2 ? l2 - sp2 : 0.f;
    l2 = fix2 ? sp2 : l2;
  }
  const float dh = R.h - L.h;
  const float w = ((R.qn - L.qn) - ub * dh) * ic;
  const float a0 = 0.5f * (dh - w), a2 = 0.5f * (dh + w);
  const float t1 = ub * ((R.qt - L.qt) - vb * dh);
  // bed-slope source with the dry-step limit
  const bool up = dz >= 0.f;
  const float hp0 = up ? L.h : R.h;
  const float adz = fabsf(dz);
  const float dzp = hp0 < adz ? (up ? hp0 : -hp0) : dz;
  const float b0 = (0.5f * kGf) * (hp0 - 0.5f * fabsf(dzp)) * dzp * ic;
  const float hp = L.h + a0;
  const bool sub = l0 * l2 < 0.f && hp > 0.f && L.h > 0.f && R.h > 0.f;
  float W0 = l0 * a0 - b0, W2 = l2 * a2 + b0;
  if (__any(sub)) {
    // Wet-wet clip of the source, b0 <- min(max(b0, lo), hi) with
    // lo = l0 a0 - L.h dx/2dt and hi = R.h dx/2dt - l2 a2 (the reference's
    // factors l/(l+e) are exactly 1 without an entropy fix), applied to the
    // strengths: W0 = max(min(W0, Ld), S - Rd), W2 = min(max(W2, S - Ld), Rd)
    // with S = W0 + W2 and the drains Ld, Rd.  A clip that holds drains the
    // cell behind (in front) by exactly its depth over two faces, and this
    // form gives that strength as the product itself; clipping b0 first
    // would leave it as a difference of terms 1e3 times larger, and FP32
    // round-off in a cell the FP64 form empties.
    const float Ld = L.h * hdxdt, Rd = R.h * hdxdt;
    const float S = W0 + W2;
    float w0 = fmaxf(fminf(W0, Ld), S - Rd);
    float w2 = fminf(fmaxf(W2, S - Ld), Rd);
    if (__any(fix0 || fix2)) {
      float lo = a0 * l0 - Ld;
      float hi = Rd - a2 * l2;
      lo = fix2 ? (lo + e2 * R.h) * (l0 * rcpf(l0 + e2)) : lo;
      hi = fix0 ? (hi + e0 * L.h) * (l2 * rcpf(l2 + e0)) : hi;
      const float bc = fminf(fmaxf(b0, lo), hi);
      const bool f = fix0 || fix2;
      w0 = f ? l0 * a0 - bc : w0;
      w2 = f ? l2 * a2 + bc : w2;
    }
    W0 = sub ? w0 : W0;
    W2 = sub ? w2 : W2;
  }
  const float n0 = ub - c, n2 = ub + c;
  // Faces with a dry side (only they can send all mass one way) take the
  // dry-side tests and the flushes; waves without one skip them.
  const bool dry = L.h < eps || R.h < eps;
  const bool any_dry = __any(dry);
  bool to_left = false, to_right = false;
  if (any_dry) {
    // Still water against a higher dry bed is balanced: both strengths are
    // exact zeros, and the reference's final flush (|d| < 1e-12,
    // src/kernels_flux.h:255-264) removes their FP64 round-off.  FP32
    // round-off is ~1e-7 of the terms and would wet the dry cell by ~1e-10 m
    // per step, enough to defeat the wet/dry test; a strength within
    // kFlushRel of the terms it is formed from is zeroed.
    const float bs = sub ? l0 * a0 - W0 : b0;  // the source after the clip
    const bool z0 = fabsf(W0) <= kFlushRel * (fabsf(l0 * a0) + fabsf(bs));
    const bool z2 = fabsf(W2) <= kFlushRel * (fabsf(l2 * a2) + fabsf(bs));
    W0 = dry && z0 ? 0.f : W0;
    W2 = dry && z2 ? 0.f : W2;
    const float he = hp + eps;
    to_left = ((he * l2 - bs) * l2 < 0.f) && R.h < eps;
    to_right = !to_left && ((he * l0 - bs) * l0 < 0.f) && L.h < eps;
  }
  // regular split: back side gets the waves with non-positive speed
  const float B0 = l0 > 0.f ? 0.f : W0, B2 = l2 > 0.f ? 0.f : W2;
  float bm = B0 + B2;
  float bn = B0 * n0 + B2 * n2;
  float bt = (B0 + B2) * vb + (ub > 0.f ? 0.f : t1);
  const float tm = W0 + W2;
  float tn = W0 * n0 + W2 * n2;
  float tt = tm * vb + t1;
  if (any_dry) {
    // all mass to the wet side, no momentum from the regular waves
    const bool one_way = to_left || to_right;
    bm = to_left ? tm : (to_right ? 0.f : bm);
    bn = one_way ? 0.f : bn;
    bt = one_way ? 0.f : bt;
    tn = one_way ? 0.f : tn;
    tt = one_way ? 0.f : tt;
  }
  float fm = tm - bm, fn = tn - bn, ft = tt - bt;
  if (__any(fix0 || fix2)) {
    const float q0 = e0 * a0, q2 = e2 * a2;
    const float r0 = e0 > 0.f ? q0 : 0.f, r2 = e2 > 0.f ? q2 : 0.f;
    const float k0 = q0 - r0, k2 = q2 - r2;
    bm += k0 + k2;
    bn += k0 * n0 + k2 * n2;
    bt += (k0 + k2) * vb;
    fm += r0 + r2;
    fn += r0 * n0 + r2 * n2;
    ft += (r0 + r2) * vb;
  }
  if (any_dry) {
    // the reference's final flush of |d| < threshold, on the mass components
    // (momentum below it changes no test: a cell that thin has no velocity)
    bm = dry && fabsf(bm) < eps ? 0.f : bm;
    fm = dry && fabsf(fm) < eps ? 0.f : fm;
  }
  lm = bm;
  ln = bn;
  lt = bt;
  rm = fm;
  rn = fn;
  rt = ft;
}

// A cell's wave speed max(|u|,|v|) + sqrt(g h) when h > hextra, else 0:
// dt = cn dx / speed is decreasing in the speed, so the smallest dt of a set
// of cells is cn dx over their largest speed, one reciprocal per wave (the
// same bits as the min of the per-cell quotients).
__device__ __forceinline__ float cell_speed32(float h, float qx, float qy, float hextra) {
  if (h > hextra) {
    const float r = rsqf(h);
    const float umax = fmaxf(fabsf(qx), fabsf(qy)) * (r * r);
    return umax + kSqrtGf * (h * r);
  }
  return 0.f;
}

// The member's fields the step reads, copied to registers once.
struct Loc {
  const TR_GLOBAL float* __restrict__ h;
  const TR_GLOBAL float* __restrict__ qx;
  const TR_GLOBAL float* __restrict__ qy;
  const TR_GLOBAL float* __restrict__ dzx;
  const TR_GLOBAL float* __restrict__ dzy;
  const TR_GLOBAL float* __restrict__ n2;
  const TR_GLOBAL int* __restrict__ rid;
  const TR_GLOBAL double* __restrict__ zadd;
  int rows, cols, nz;
  float add0;
};

// A pointer every lane holds, made explicitly wave-uniform: the compiler
// then keeps it in scalar registers for the whole piece instead of
// re-reading it from a vector register at every row.
template <class T>
__device__ __forceinline__ T* uni(T* p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane(unsigned(v));
  const unsigned hi = __builtin_amdgcn_readfirstlane(unsigned(v >> 32));
  return reinterpret_cast<T*>((static_cast<unsigned long long>(hi) << 32) | lo);
}

struct Raw {
  float h, qx, qy, dzx, dzy, n2, add;
};

// Loads and stores by a 32-bit byte offset from a wave-uniform base: the
// base stays in scalar registers and one vector add per row forms the offset
// every array shares (members hold fewer than 2^30 cells).
template <class T>
__device__ __forceinline__ T ld(const TR_GLOBAL T* base, unsigned boff) {
  return *(const TR_GLOBAL T*)((const TR_GLOBAL char*)base + boff);
}
template <class T>
__device__ __forceinline__ TR_GLOBAL T& at(TR_GLOBAL T* base, unsigned boff) {
  return *(TR_GLOBAL T*)((TR_GLOBAL char*)base + boff);
}

// Row r of the input at the lane's column (its byte offset col4 already
// clamped into the grid), with r clamped too: rows and columns outside the
// grid only feed halo or ghost cells that are never updated, so they may hold
// copies instead of zeros, and the loads need no branches.
__device__ __forceinline__ Raw load_row(const Loc& L, int r, unsigned col4) {
  const int rr = min(max(r, 0), L.rows - 1);
  const unsigned b = unsigned(rr * L.cols) * 4u + col4;
  Raw w;
  w.h = ld(L.h, b);
  w.qx = ld(L.qx, b);
  w.qy = ld(L.qy, b);
  w.dzx = ld(L.dzx, b);
  w.dzy = ld(L.dzy, b);
  w.n2 = L.n2 ? ld(L.n2, b) : 0.f;
  w.add = L.nz > 0 ? (L.rid ? float(L.zadd[ld(L.rid, b)]) : L.add0) : 0.f;
  return w;
}

// One member's piece: rows [row_first, row_last] of column strip bx, with
// the two-row halo recomputed at its ends.  Returns the piece's dt min bits.
template <bool kBC>
__device__ __forceinline__ float step_piece(const InstF32& P, int cur, double dt64, double simtime,
                                            unsigned dither_step, int bx, int e_min, int e_max, int lane) {
  const float dt = float(dt64);
  const float dtdx = float(dt64 / P.dx);
  const float hdxdt = float(0.5 * P.dx / dt64);
  const float hextra = P.hextra_f;
  const float eps = P.eps_f;
  const bool open_bc = P.open_bc != 0;
  const int R = P.rows, Cc = P.cols;
  const int nsrc = P.nsrc;
  const TR_GLOBAL int* __restrict__ src_row = gptr(P.src_row);
  const TR_GLOBAL int* __restrict__ src_col = gptr(P.src_col);
  const TR_GLOBAL double* __restrict__ src_add = gptr(P.src_add + cur * nsrc);
  TR_GLOBAL float* __restrict__ ho = uni(gptr(P.h[cur ^ 1]));
  TR_GLOBAL float* __restrict__ qxo = uni(gptr(P.qx[cur ^ 1]));
  TR_GLOBAL float* __restrict__ qyo = uni(gptr(P.qy[cur ^ 1]));
  TR_GLOBAL float* __restrict__ maxh = gptr(P.maxh);
  Loc L;
  L.h = uni(gptr(P.h[cur]));
  L.qx = uni(gptr(P.qx[cur]));
  L.qy = uni(gptr(P.qy[cur]));
  L.dzx = uni(gptr(P.dzx));
  L.dzy = uni(gptr(P.dzy));
  L.n2 = uni(gptr(P.nman2));
  L.rid = uni(gptr(P.runoff_id));
  L.nz = P.num_runoffs;
  L.zadd = uni(gptr(P.num_runoffs > 0 ? P.zone_add + cur * P.num_runoffs : nullptr));
  L.rows = R;
  L.cols = Cc;
  // rain depth of this step (FP64 product rate*dt, rounded once)
  L.add0 = P.num_runoffs > 0 ? float(L.zadd[0]) : 0.f;

  const int c = kOutCols * bx - 1 + lane;
  const unsigned dither_col = unsigned(c) * kDitherCol;
  const unsigned dither_colq = unsigned(c) * kDitherRow;
  const bool round_q = P.round_q != 0;
  const unsigned col4 = unsigned(min(max(c, 0), Cc - 1)) * 4u;
  const bool col_int = c >= 1 && c <= Cc - 2;
  const bool out_lane = lane >= 2 && lane <= kOutCols + 1 && col_int;
  const bool ghost_w = c == 0 && lane == 1;
  const bool ghost_e = c == Cc - 1 && lane >= 3 && lane <= kLanes - 2;
  const bool col_mine = out_lane || ghost_w || ghost_e;
  const bool edge_we = bx == 0 || bx == P.nbx - 1;

  // previous row: input state, derived values, and its RHS sums so far
  float ph = 0.f, pqx = 0.f, pqy = 0.f, pu = 0.f, pv = 0.f, psq = 0.f, pn2 = 0.f, padd = 0.f;
  float pdzx = 0.f, pdzy = 0.f;
  float p0h = 0.f, p0x = 0.f, p0y = 0.f, p1h = 0.f, p1x = 0.f, p1y = 0.f;
  // updated rows r-2 (A) and r-3 (B)
  float ah = 0.f, aqx = 0.f, aqy = 0.f, adzx = 0.f, adzy = 0.f;
  float bh = 0.f;
  // external boundary values carried to the south ghost row
  SouthCarry<float> sc{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, false};
  float fast = -1.f;  // largest wave speed of the cells written (-1: none)
  int ks = 0;  // first source not above the row being updated

  // One code path for every row: a row's bits must not depend on where a
  // piece starts (members give the same result alone and in any batch).
  const int r_end = e_max + 2;
  Raw nxt = load_row(L, e_min - 2, col4);
  for (int r = e_min - 2; r <= r_end; ++r) {
    const Raw s = nxt;
    if (r + 1 <= r_end) nxt = load_row(L, r + 1, col4);
    const float rs = rsqf(s.h);
    const float sq = s.h > 0.f ? s.h * rs : 0.f;
    const float ih = rs * rs;
    const bool mv = s.h >= hextra;
    const float u = mv ? s.qx * ih : 0.f;
    const float v = mv ? s.qy * ih : 0.f;

    // y face between row r (south, back) and row r-1 (north, front); its
    // back side joins row r's x back side below, its front side completes
    // row r-1, which is updated before row r's x face is formed
    float y1h, y1x, y1y, ylm, yln, ylt;
    {
      const Side S{s.h, s.qy, s.qx, v, u, sq};
      const Side N{ph, pqy, pqx, pv, pu, psq};
      face32(S, N, s.dzy, hdxdt, eps, ylm, yln, ylt, y1h, y1y, y1x);
    }

    // update row r-1
    const int ru = r - 1;
    float nh = ph, nqx = pqx, nqy = pqy;
    const bool ru_int = ru >= 1 && ru <= R - 2;
    // point-source depth of this lane's cell in row ru (src/kernels.h:50-84;
    // the FP64 depths the close prepared, summed in list order)
    double sadd = 0.0;
    if (nsrc > 0 && ru_int) {
      while (ks < nsrc && src_row[ks] < ru) ks++;
      for (int k = ks; k < nsrc && src_row[k] == ru; k++)
        if (src_col[k] == c) sadd += src_add[k];
    }
    if (ru_int && col_int) {
      // depth ph - dtdx*S carried as hs + he, he the rounding error
      float hs, he;
      {
#pragma clang fp contract(off)
        const float S = p0h + (p1h + y1h);
        const float d = dtdx * S;
        const float de = fmaf(dtdx, S, -d);  // dtdx*S == d + de exactly
        float e;
        hs = two_sum(ph, -d, e);
        he = e - de;
      }
      clamp_neg(hs, he);
      float mxe = 0.f, mye = 0.f;
      float mx, my;
      if (round_q) {
        mx = sub_scaled2(pqx, dtdx, p0x + (p1x + y1x), mxe);
        my = sub_scaled2(pqy, dtdx, p0y + (p1y + y1y), mye);
      } else {
        mx = pqx - dtdx * (p0x + (p1x + y1x));
        my = pqy - dtdx * (p0y + (p1y + y1y));
      }
      float kf = 1.f;
      if (pn2 > eps && ph >= hextra) {
        const float m2 = mx * mx + my * my;
        const float mm = sqrtf_hw(m2);
        // tt = dt g n^2 |m| / h^(7/3); q = m * 2/(1+sqrt(1+4tt))
        const float ih73 = __builtin_amdgcn_exp2f(-2.3333333f * __builtin_amdgcn_logf(ph));
        const float tt = dt * kGf * pn2 * mm * ih73;
        const float k = 2.f * rcpf(1.f + sqrtf_hw(1.f + 4.f * tt));
        kf = mm > eps ? k : 1.f;
      }
      const bool wet = hs >= hextra;
      if (round_q) {
        // (ms + me) kf carried as qs + qe, then rounded stochastically
        const unsigned hrow = dither_step + unsigned(ru) * kDitherRow;
        const float qxs = mx * kf, qys = my * kf;
        const float qxe = fmaf(mx, kf, -qxs) + mxe * kf;
        const float qye = fmaf(my, kf, -qys) + mye * kf;
        nqx = wet ? sround_signed(qxs, qxe, int(dither_colq + fmix32(hrow ^ kDitherQx))) : 0.f;
        nqy = wet ? sround_signed(qys, qye, int(dither_colq + fmix32(hrow ^ kDitherQy))) : 0.f;
      } else {
        nqx = wet ? mx * kf : 0.f;
        nqy = wet ? my * kf : 0.f;
      }
      if (L.nz > 0) {
        float e;
        hs = two_sum(hs, padd, e);
        he += e;
        const bool low = hs < hextra;
        clamp_neg(hs, he);
        nqx = low ? 0.f : nqx;
        nqy = low ? 0.f : nqy;
      }
      if (sadd != 0.0) {
        // point sources join the two-float depth before it is rounded, so
        // their volume is rounded without bias like the rest of the update
        const float shi = float(sadd);
        const float slo = float(sadd - double(shi));
        float e;
        hs = two_sum(hs, shi, e);
        he += e + slo;
      }
      nh = sround(hs, he, int(dither_col + fmix32(dither_step + unsigned(ru) * kDitherRow)));
    }
    // x face (c, c+1) of row r: back side stays in this lane (rhs*0), the
    // front side goes to lane c+1 (rhs*1).  Faces outside the reference's
    // ranges only feed ghost or halo cells that are never updated, so they
    // are computed unmasked; fluctuations are scaled by dt/dx at the update.
    float c0h, c0x, c0y, c1h, c1x, c1y;
    {
      Side Rt;
      Rt.h = from_right(s.h);
      Rt.qn = from_right(s.qx);
      Rt.qt = from_right(s.qy);
      Rt.un = from_right(u);
      Rt.ut = from_right(v);
      Rt.sq = from_right(sq);
      const Side Lt{s.h, s.qx, s.qy, u, v, sq};
      float em, en, et;
      face32(Lt, Rt, s.dzx, hdxdt, eps, c0h, c0x, c0y, em, en, et);
      c1h = from_left(em);
      c1x = from_left(en);
      c1y = from_left(et);
    }

    // rhs*0 of row r = x back side + y back side (reference order)
    c0h += ylm;
    c0y += yln;
    c0x += ylt;

    if (open_bc) {
      if (edge_we) {  // this strip holds the west or east ghost column
        const float rh = from_right(nh), rqx = from_right(nqx), rqy = from_right(nqy);
        const float lh = from_left(nh), lqx = from_left(nqx), lqy = from_left(nqy);
        if (ru_int && c == 0) { nh = rh; nqx = rqx; nqy = rqy; }
        if (ru_int && c == Cc - 1) { nh = lh; nqx = lqx; nqy = lqy; }
      }
      if (ru == 1) { ah = nh; aqx = nqx; aqy = nqy; }
      if (ru == R - 1) { nh = ah; nqx = aqx; nqy = aqy; }
    }
    // external boundary segments after the open-boundary copies
    if (kBC && P.bc.n > 0)
      extbc_row(P.bc, ru, c, R, Cc, col_int, edge_we, open_bc, simtime, dt64, nh, nqx, nqy, ah, aqx,
                aqy, sc);

    // wet/dry, dt and store of row r-2
    const int e = r - 2;
    if (e >= e_min && e <= e_max && e >= 0) {
      float qxq = aqx, qyq = aqy;
      const float hE = from_right(ah);
      const float hW = from_left(ah);
      const float dzW = -from_left(adzx);
      const bool cell_int = e >= 1 && e <= R - 2 && col_int;
      if (cell_int && ah > hextra) {
        if ((ah < adzx && hE < eps) || (ah < dzW && hW < eps)) qxq = 0.f;
        if ((ah < adzy && bh < eps) || (ah < -pdzy && nh < eps)) qyq = 0.f;
      }
      if (col_mine) {
        const unsigned o = unsigned(e * Cc) * 4u + unsigned(c) * 4u;
        at(ho, o) = ah;
        at(qxo, o) = qxq;
        at(qyo, o) = qyq;
        if (maxh && cell_int && ah > at(maxh, o)) at(maxh, o) = ah;
        fast = fmaxf(fast, cell_speed32(ah, qxq, qyq, hextra));
      }
    }

    // rotate
    bh = ah;
    ah = nh; aqx = nqx; aqy = nqy; adzx = pdzx; adzy = pdzy;
    ph = s.h; pqx = s.qx; pqy = s.qy; pu = u; pv = v; psq = sq; pn2 = s.n2; padd = s.add;
    pdzx = s.dzx; pdzy = s.dzy;
    p0h = c0h; p0x = c0x; p0y = c0y; p1h = c1h; p1x = c1x; p1y = c1y;
  }
  return fast;
}

// Wave max of the lanes' largest speeds (floats >= 0 order like their bit
// patterns as signed integers, and -1 below them all); lane 0 publishes
// dt = cn dx / max (1e10 when every cell written is dry, nothing when no cell
// was written) as the FP64 value for the member's close.
__device__ __forceinline__ void publish_min_f32(StepCtrl* cc, float fast, float cn_dx, int lane) {
  int bits = __float_as_int(fast);
  for (int off = 32; off > 0; off >>= 1) {
    const int o = __shfl_xor(bits, off, 64);
    bits = o > bits ? o : bits;
  }
  if (lane == 0) {
    const float m = __int_as_float(bits);
    const float dt = m > 0.f ? cn_dx * rcpf(m) : kMaxDtF;
    publish_min(cc, m < 0.f ? ~0ull : dbits(double(dt)));
  }
}

// Batched step: wave w (placed XCD-contiguously) steps piece w and publishes
// its dt min.  The member's cursor and both control slots are read from the
// control array by member index, in parallel with its descriptor, so the
// rows' loads wait on two dependent reads (piece, then descriptor and
// control) -- what a small grid's step time is made of.
// 5 waves per SIMD: the step fits in 96 VGPRs without spills.
// kBC: some member has external boundary segments (a variant of its own, so
// the common one keeps its registers).
template <bool kBC>
__global__ __launch_bounds__(64, 5) void k_step_f32(BatchArgs a) {
  const int w = xcd_slot(int(blockIdx.x), a.nwaves);
  const int lane = threadIdx.x;
  const Piece pc = a.pieces[w];
  const InstF32& P = a.inst[pc.member];
  CtrlPair& cp = a.ctrls[pc.member];
  // wave-uniform: keeps the member's array bases in scalar registers
  const int cur = __builtin_amdgcn_readfirstlane(cp.cursor);
  const StepCtrl& c0 = cp.ctrl[0];
  const StepCtrl& c1 = cp.ctrl[1];
  const int halt0 = c0.halt, halt1 = c1.halt;
  const double dt0 = c0.dt, dt1 = c1.dt, t0 = c0.simtime, t1 = c1.simtime;
  const long long n0 = c0.step, n1 = c1.step;
  if (cur ? halt1 : halt0) return;
  const unsigned dither_step = unsigned(P.it_count0 + (cur ? n1 : n0)) * kDitherStep;
  const float fast = step_piece<kBC>(P, cur, cur ? dt1 : dt0, cur ? t1 : t0, dither_step, pc.bx, pc.row_first,
                                     pc.row_last, lane);
  publish_min_f32(cp.ctrl + cur, fast, P.cn_dx_f, lane);
}

// Initial dt of one member over every cell (ghosts included).
__global__ __launch_bounds__(64) void k_prime_f32(const InstF32* inst, int member) {
  const InstF32& P = inst[member];
  const int cur = *P.cursor;
  const int n = P.rows * P.cols;
  float fast = -1.f;
  for (int id = blockIdx.x * kLanes + threadIdx.x; id < n; id += gridDim.x * kLanes)
    fast = fmaxf(fast, cell_speed32(P.h[cur][id], P.qx[cur][id], P.qy[cur][id], P.hextra_f));
  publish_min_f32(P.ctrl + cur, fast, P.cn_dx_f, int(threadIdx.x));
}

// One thread per member (members[i], or first + i when members is NULL):
// prepare the other control slot and flip the member's cursor.  Halted
// members keep their cursor, so launches queued past a halt are no-ops.
// The control pair is addressed by member index, so its reads do not wait
// for the descriptor's.
__global__ __launch_bounds__(64) void k_close_f32(const InstF32* inst, CtrlPair* ctrls, const int* members,
                                                  int first, int n, int after_step) {
  const int i = int(blockIdx.x) * kLanes + int(threadIdx.x);
  if (i >= n) return;
  const int m = members ? members[i] : first + i;
  CtrlPair& cp = ctrls[m];
  const int cur = cp.cursor;
  const StepCtrl& cc = cp.ctrl[cur];
  if (cc.halt) return;
  prepare_next(inst[m], cc, cp.ctrl[cur ^ 1], dtmin_of(cc), after_step != 0, cur ^ 1);
  cp.cursor = cur ^ 1;
}

}  // namespace

int fused_grid_x_f32(int cols) { return (cols - 2 + kOutCols - 1) / kOutCols; }

// Rows per wave of the FP64 fused step's three-wave build: long strips
// amortise the two-row halo and the per-wave set-up (measured best: 96 rows
// on a 4096^2 grid, about 3 waves per SIMD), short ones give a small grid
// enough waves (on 400^2: 26.1 us per step with 3 rows, 29.4 with 4, 29.1
// with 2, 32.7 with 1 and 35.9 with 6).
static int strip_rows_w3(int rows, int cols) {
  const int nbx = fused_grid_x_f32(cols);
  const int64_t target_waves = 3072;
  int64_t strip = (int64_t(rows - 2) * nbx + target_waves - 1) / target_waves;
  return strip < 3 ? 3 : (strip > 128 ? 128 : int(strip));
}

// Grids whose three-wave strips would reach 40 rows run the four-wave build
// (128 VGPRs, 12 B of spills) with about 7168 waves of 24..48 rows: on the
// spun-up watershed 608-614 against 840 us per step at 4096^2 (40-row
// strips) and 382-385 against 466 us at 3000^2 (24-40 rows); at 2048^2 (a
// 24-row three-wave default) 189 against 186 us, so smaller grids keep the
// three-wave build.
int default_waves_per_simd(int rows, int cols) { return strip_rows_w3(rows, cols) >= 40 ? 4 : 3; }

int default_strip_rows(int rows, int cols) {
  if (default_waves_per_simd(rows, cols) == 3) return strip_rows_w3(rows, cols);
  const int64_t target_waves = 7168;
  const int64_t strip = (int64_t(rows - 2) * fused_grid_x_f32(cols) + target_waves - 1) / target_waves;
  return strip < 24 ? 24 : (strip > 48 ? 48 : int(strip));
}

int launch_step_f32(const BatchArgs& a, int nwaves, hipStream_t s) {
  if (nwaves <= 0) return TR_OK;
  if (a.any_bc) hipLaunchKernelGGL(k_step_f32<true>, dim3(unsigned(nwaves)), dim3(kLanes), 0, s, a);
  else hipLaunchKernelGGL(k_step_f32<false>, dim3(unsigned(nwaves)), dim3(kLanes), 0, s, a);
  // one member: the close takes its index directly instead of reading the list
  const bool one = a.nmembers == 1;
  hipLaunchKernelGGL(k_close_f32, dim3(unsigned((a.nmembers + kLanes - 1) / kLanes)), dim3(kLanes), 0, s,
                     a.inst, a.ctrls, one ? nullptr : a.members, one ? a.first_member : 0, a.nmembers, 1);
  return launch_check("step_f32");
}

int launch_prime_f32(const InstF32* d_inst, CtrlPair* d_ctrls, int member, const InstF32& host_inst,
                     hipStream_t s) {
  int64_t n = int64_t(host_inst.rows) * host_inst.cols;
  unsigned g = unsigned(std::min<int64_t>((n + kLanes - 1) / kLanes, 2048));
  hipLaunchKernelGGL(k_prime_f32, dim3(g), dim3(kLanes), 0, s, d_inst, member);
  hipLaunchKernelGGL(k_close_f32, dim3(1), dim3(kLanes), 0, s, d_inst, d_ctrls, nullptr, member, 1, 0);
  return launch_check("prime_f32");
}

}  // namespace tr

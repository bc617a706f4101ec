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
//   * the dry / flush threshold is a member setting (default 1e-10) instead
//     of 1e-12 (the reference's FP32 build uses 1e-6, src/constants.h:43);
//   * bed tests z1+h1<z2 become h1<dz on the stored face difference;
//   * the semi-implicit friction uses q = 2m/(1+sqrt(1+4tt)), the same value
//     as -0.5(m - m sqrt(1+4tt))/tt without its cancellation for small tt.
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "common.hpp"
#include "step_ctrl.hpp"

#pragma clang fp contract(fast)

namespace tr {
namespace {

constexpr int kLanes = 64;
constexpr int kOutCols = 60;
constexpr float kGf = 9.81f;
constexpr float kSqrtGf = 3.132091953f;
constexpr float kMaxDtF = 1e10f;

__device__ __forceinline__ float rcpf(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrtf_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

struct Side {
  float h, qn, qt, un, ut, sq;
};

// One face in normal/tangential components.  Outputs the fluctuations sent to
// the cell behind the face (l*) and in front of it (r*): mass, normal and
// tangential momentum, before the dt/dx scaling.
__device__ __forceinline__ void face32(const Side& L, const Side& R, float dz, float hdxdt,
                                       float kEpsF, float& lm, float& ln, float& lt, float& rm,
                                       float& rn, float& rt) {
  lm = ln = lt = rm = rn = rt = 0.f;
  if (!(L.h > 0.f || R.h > 0.f)) return;
  const float c = sqrtf_hw(kGf * (0.5f * (L.h + R.h)));
  const float ic = rcpf(c);
  const float is = rcpf(L.sq + R.sq);
  const float ub = (L.un * L.sq + R.un * R.sq) * is;
  const float vb = (L.ut * L.sq + R.ut * R.sq) * is;
  float l0 = ub - c, l2 = ub + c, e0 = 0.f, e2 = 0.f;
  {
    const float a = L.un - kSqrtGf * L.sq, b = R.un - kSqrtGf * R.sq;
    if (a < 0.f && b > 0.f) {
      const float sp = a * (b - l0) * rcpf(b - a);
      e0 = l0 - sp;
      l0 = sp;
    }
  }
  {
    const float a = L.un + kSqrtGf * L.sq, b = R.un + kSqrtGf * R.sq;
    if (a < 0.f && b > 0.f) {
      const float sp = b * (l2 - a) * rcpf(b - a);
      e2 = l2 - sp;
      l2 = sp;
    }
  }
  const float dh = R.h - L.h;
  const float w = ((R.qn - L.qn) - ub * dh) * ic;
  const float a0 = 0.5f * (dh - w), a2 = 0.5f * (dh + w);
  // wave 1 only moves tangential momentum: alpha1 * eigenvector = dqt - vbar dh
  const float t1 = ub * ((R.qt - L.qt) - vb * dh);
  float hp, dzp = dz;
  if (dz >= 0.f) {
    hp = L.h;
    if (L.h < dz) dzp = L.h;
  } else {
    hp = R.h;
    if (R.h < -dz) dzp = -R.h;
  }
  float b0 = 0.5f * kGf * (hp - 0.5f * fabsf(dzp)) * dzp * ic;
  hp = L.h + a0;
  if (l0 * l2 < 0.f && hp > 0.f && L.h > 0.f && R.h > 0.f) {
    b0 = fmaxf(b0, l0 * rcpf(l0 + e2) * (a0 * l0 - L.h * hdxdt + e2 * R.h));
    b0 = fminf(b0, l2 * rcpf(l2 + e0) * (-a2 * l2 + R.h * hdxdt + e0 * L.h));
  }
  const float W0 = l0 * a0 - b0, W2 = l2 * a2 + b0;
  const float sL = hp - b0 * rcpf(l0), sR = hp - b0 * rcpf(l2);
  const float n0 = ub - c, n2 = ub + c;
  if (sR < -kEpsF && R.h < kEpsF) {
    lm = W0 + W2;
  } else if (sL < -kEpsF && L.h < kEpsF) {
    rm = W0 + W2;
  } else {
    if (l0 > 0.f) { rm += W0; rn += W0 * n0; rt += W0 * vb; }
    else          { lm += W0; ln += W0 * n0; lt += W0 * vb; }
    if (ub > 0.f) rt += t1;
    else          lt += t1;
    if (l2 > 0.f) { rm += W2; rn += W2 * n2; rt += W2 * vb; }
    else          { lm += W2; ln += W2 * n2; lt += W2 * vb; }
  }
  {
    const float q0 = e0 * a0, q2 = e2 * a2;
    if (e0 > 0.f) { rm += q0; rn += q0 * n0; rt += q0 * vb; }
    else          { lm += q0; ln += q0 * n0; lt += q0 * vb; }
    if (e2 > 0.f) { rm += q2; rn += q2 * n2; rt += q2 * vb; }
    else          { lm += q2; ln += q2 * n2; lt += q2 * vb; }
  }
  lm = fabsf(lm) < kEpsF ? 0.f : lm;
  ln = fabsf(ln) < kEpsF ? 0.f : ln;
  lt = fabsf(lt) < kEpsF ? 0.f : lt;
  rm = fabsf(rm) < kEpsF ? 0.f : rm;
  rn = fabsf(rn) < kEpsF ? 0.f : rn;
  rt = fabsf(rt) < kEpsF ? 0.f : rt;
}

__device__ __forceinline__ float cell_dt32(float h, float qx, float qy, float cn_dx, float hextra) {
  if (h > hextra) {
    const float umax = fmaxf(fabsf(qx), fabsf(qy)) * rcpf(h);
    return cn_dx * rcpf(umax + kSqrtGf * sqrtf_hw(h));
  }
  return kMaxDtF;
}

// Row of input state with its derived per-cell quantities.
struct Row {
  float h, qx, qy, dzx, dzy, n2, add;
};

__device__ __forceinline__ Row load_row(const InstF32& P, int cur, int r, int c, bool col_ok,
                                        bool col_int, float add0) {
  Row w{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (r >= 0 && r < P.rows && col_ok) {
    const long long i = (long long)r * P.cols + c;
    w.h = P.h[cur][i];
    w.qx = P.qx[cur][i];
    w.qy = P.qy[cur][i];
    w.dzx = P.dzx[i];
    w.dzy = P.dzy[i];
    if (P.nman2) w.n2 = P.nman2[i];
    if (P.num_runoffs > 0 && r >= 1 && r <= P.rows - 2 && col_int)
      w.add = P.runoff_id ? float(P.zone_add[cur * P.num_runoffs + P.runoff_id[i]]) : add0;
  }
  return w;
}

__global__ __launch_bounds__(64) void k_step_f32(BatchArgs a) {
  const int2 wk = a.work[blockIdx.x];
  const InstF32& P = a.inst[wk.x];
  const int cur = *P.cursor;
  const StepCtrl& CC = P.ctrl[cur];
  if (CC.halt) return;
  const int lane = threadIdx.x;
  const double dt64 = CC.dt;
  const float dt = float(dt64);
  const float dtdx = float(dt64 / P.dx);
  const float hdxdt = float(0.5 * P.dx / dt64);
  const float hextra = P.hextra_f;
  const float cn_dx = P.cn_dx_f;
  const float kEpsF = P.eps_f;
  const double* __restrict__ src_add = P.src_add + cur * P.nsrc;
  // rain depth of this step (FP64 product rate*dt, rounded once); zone 0
  // when the member has one uniform zone
  const float add0 = P.num_runoffs > 0 ? float(P.zone_add[cur * P.num_runoffs]) : 0.f;
  const bool open_bc = P.open_bc != 0;
  const int R = P.rows, Cc = P.cols;
  const int bx = wk.y % P.nbx, by = wk.y / P.nbx;
  float* __restrict__ ho = P.h[cur ^ 1];
  float* __restrict__ qxo = P.qx[cur ^ 1];
  float* __restrict__ qyo = P.qy[cur ^ 1];

  const int c = kOutCols * bx - 1 + lane;
  const bool col_ok = c >= 0 && c < Cc;
  const bool col_int = c >= 1 && c <= Cc - 2;
  const bool xface_ok = c >= 0 && c <= Cc - 2 && lane < kLanes - 1;
  const bool out_lane = lane >= 2 && lane <= kOutCols + 1 && col_int;
  const bool ghost_w = c == 0 && lane == 1;
  const bool ghost_e = c == Cc - 1 && lane >= 3 && lane <= kLanes - 2;
  const bool col_mine = out_lane || ghost_w || ghost_e;

  const int e_lo = 1 + P.strip_rows * by;
  const int e_hi = min(e_lo + P.strip_rows - 1, R - 2);
  const int e_min = by == 0 ? 0 : e_lo;
  const int e_max = by == P.nby - 1 ? R - 1 : e_hi;

  // previous row (input state and derived values)
  float ph = 0.f, pqx = 0.f, pqy = 0.f, pu = 0.f, pv = 0.f, psq = 0.f, pn2 = 0.f, padd = 0.f;
  float pdzy = 0.f, pdzx = 0.f;
  float px0h = 0.f, px0x = 0.f, px0y = 0.f, px1h = 0.f, px1x = 0.f, px1y = 0.f;
  float py0h = 0.f, py0x = 0.f, py0y = 0.f;
  // updated rows r-2 (A) and r-3 (B)
  float ah = 0.f, aqx = 0.f, aqy = 0.f, adzx = 0.f, adzy = 0.f;
  float bh = 0.f;
  float best = kMaxDtF * 2.f;

  Row nxt = load_row(P, cur, e_min - 2, c, col_ok, col_int, add0);
  for (int r = e_min - 2; r <= e_max + 2; ++r) {
    const Row s = nxt;
    if (r + 1 <= e_max + 2) nxt = load_row(P, cur, r + 1, c, col_ok, col_int, add0);
    const bool row_ok = r >= 0 && r < R;
    const float sq = s.h > 0.f ? sqrtf_hw(s.h) : 0.f;
    const float ih = rcpf(s.h);
    const bool mv = s.h >= hextra;
    const float u = mv ? s.qx * ih : 0.f;
    const float v = mv ? s.qy * ih : 0.f;

    // x face (c, c+1) of row r
    float x0h = 0.f, x0x = 0.f, x0y = 0.f, x1h, x1x, x1y;
    {
      Side Rt;
      Rt.h = __shfl_down(s.h, 1);
      Rt.qn = __shfl_down(s.qx, 1);
      Rt.qt = __shfl_down(s.qy, 1);
      Rt.un = __shfl_down(u, 1);
      Rt.ut = __shfl_down(v, 1);
      Rt.sq = __shfl_down(sq, 1);
      float em = 0.f, en = 0.f, et = 0.f;
      if (row_ok && r >= 1 && r <= R - 2 && xface_ok) {
        const Side Lt{s.h, s.qx, s.qy, u, v, sq};
        float lm, ln, lt;
        face32(Lt, Rt, s.dzx, hdxdt, kEpsF, lm, ln, lt, em, en, et);
        x0h = lm * dtdx;
        x0x = ln * dtdx;
        x0y = lt * dtdx;
        em *= dtdx;
        en *= dtdx;
        et *= dtdx;
      }
      x1h = __shfl_up(em, 1);
      x1x = __shfl_up(en, 1);
      x1y = __shfl_up(et, 1);
    }

    // y face between row r (south) and row r-1 (north): normal = y
    float y0h = 0.f, y0x = 0.f, y0y = 0.f, y1h = 0.f, y1x = 0.f, y1y = 0.f;
    if (row_ok && r >= 1 && col_int) {
      const Side S{s.h, s.qy, s.qx, v, u, sq};
      const Side N{ph, pqy, pqx, pv, pu, psq};
      float lm, ln, lt, rm, rn, rt;
      face32(S, N, s.dzy, hdxdt, kEpsF, lm, ln, lt, rm, rn, rt);
      y0h = lm * dtdx; y0y = ln * dtdx; y0x = lt * dtdx;
      y1h = rm * dtdx; y1y = rn * dtdx; y1x = rt * dtdx;
    }

    // update row r-1
    const int ru = r - 1;
    float nh = ph, nqx = pqx, nqy = pqy;
    const bool ru_int = ru >= 1 && ru <= R - 2;
    if (ru_int && col_int) {
      float hnew = ph - (px0h + py0h) - (px1h + y1h);
      hnew = hnew < 0.f ? 0.f : hnew;
      if (hnew < hextra) {
        nqx = 0.f;
        nqy = 0.f;
      } else {
        const float mx = pqx - (px0x + py0x) - (px1x + y1x);
        const float my = pqy - (px0y + py0y) - (px1y + y1y);
        const float mm = sqrtf_hw(mx * mx + my * my);
        if (pn2 > kEpsF && ph >= hextra && mm > kEpsF) {
          // h^(7/3) = h*h*cbrt(h)
          const float h73 = __builtin_amdgcn_exp2f(2.3333333f * __builtin_amdgcn_logf(ph));
          const float tt = dt * kGf * pn2 * mm * rcpf(h73);
          const float k = 2.f * rcpf(1.f + sqrtf_hw(1.f + 4.f * tt));
          nqx = mx * k;
          nqy = my * k;
        } else {
          nqx = mx;
          nqy = my;
        }
      }
      if (P.num_runoffs > 0) {
        hnew += padd;
        if (hnew < hextra) {
          hnew = hnew < 0.f ? 0.f : hnew;
          nqx = 0.f;
          nqy = 0.f;
        }
      }
      nh = hnew;
    }
    if (ru_int) {
      for (int k = 0; k < P.nsrc; k++)
        if (P.src_row[k] == ru && P.src_col[k] == c) nh = nh + float(src_add[k]);
    }
    if (open_bc) {
      const float rh = __shfl_down(nh, 1), rqx = __shfl_down(nqx, 1), rqy = __shfl_down(nqy, 1);
      const float lh = __shfl_up(nh, 1), lqx = __shfl_up(nqx, 1), lqy = __shfl_up(nqy, 1);
      if (ru_int && c == 0) { nh = rh; nqx = rqx; nqy = rqy; }
      if (ru_int && c == Cc - 1) { nh = lh; nqx = lqx; nqy = lqy; }
      if (ru == 1) { ah = nh; aqx = nqx; aqy = nqy; }
      if (ru == R - 1) { nh = ah; nqx = aqx; nqy = aqy; }
    }

    // wet/dry, dt and store of row r-2
    const int e = r - 2;
    if (e >= e_min && e <= e_max && e >= 0) {
      float qxq = aqx, qyq = aqy;
      const float hE = __shfl_down(ah, 1);
      const float hW = __shfl_up(ah, 1);
      const float dzW = -__shfl_up(adzx, 1);
      const bool cell_int = e >= 1 && e <= R - 2 && col_int;
      if (cell_int && ah > hextra) {
        if ((ah < adzx && hE < kEpsF) || (ah < dzW && hW < kEpsF)) qxq = 0.f;
        if ((ah < adzy && bh < kEpsF) || (ah < -pdzy && nh < kEpsF)) qyq = 0.f;
      }
      if (col_mine) {
        const long long o = (long long)e * Cc + c;
        ho[o] = ah;
        qxo[o] = qxq;
        qyo[o] = qyq;
        if (P.maxh && cell_int && ah > P.maxh[o]) P.maxh[o] = ah;
        best = fminf(best, cell_dt32(ah, qxq, qyq, cn_dx, hextra));
      }
    }

    // rotate
    bh = ah;
    ah = nh; aqx = nqx; aqy = nqy; adzx = pdzx; adzy = pdzy;
    ph = s.h; pqx = s.qx; pqy = s.qy; pu = u; pv = v; psq = sq; pn2 = s.n2; padd = s.add;
    pdzx = s.dzx; pdzy = s.dzy;
    px0h = x0h; px0x = x0x; px0y = x0y; px1h = x1h; px1x = x1x; px1y = x1y;
    py0h = y0h; py0x = y0x; py0y = y0y;
  }
  // wave min of the float dt (positive floats order like their bit patterns)
  unsigned int bits = __float_as_uint(best);
  for (int off = 32; off > 0; off >>= 1) {
    unsigned int o = __shfl_xor(bits, off, 64);
    bits = o < bits ? o : bits;
  }
  if (lane == 0) {
    const float m = __uint_as_float(bits);
    const unsigned long long wb = m >= kMaxDtF * 2.f ? ~0ull : dbits(double(m));
    close_step_cursor(P, cur, wb, true);
  }
}

// Initial dt of one member over every cell (ghosts included).
__global__ __launch_bounds__(64) void k_prime_f32(const InstF32* inst, int member) {
  const InstF32& P = inst[member];
  const int cur = *P.cursor;
  const int n = P.rows * P.cols;
  float best = kMaxDtF * 2.f;
  for (int id = blockIdx.x * kLanes + threadIdx.x; id < n; id += gridDim.x * kLanes)
    best = fminf(best, cell_dt32(P.h[cur][id], P.qx[cur][id], P.qy[cur][id], P.cn_dx_f, P.hextra_f));
  unsigned int bits = __float_as_uint(best);
  for (int off = 32; off > 0; off >>= 1) {
    unsigned int o = __shfl_xor(bits, off, 64);
    bits = o < bits ? o : bits;
  }
  if (threadIdx.x == 0) {
    const float m = __uint_as_float(bits);
    const unsigned long long wb = m >= kMaxDtF * 2.f ? ~0ull : dbits(double(m));
    StepCtrl* cc = P.ctrl + cur;
    if (wb != ~0ull) atomicMin(&cc->dtmin_bits, wb);
    __threadfence();
    unsigned prev = atomicAdd(&cc->arrivals, 1u);
    if (prev == gridDim.x - 1) {
      __threadfence();
      unsigned long long mm = atomicMin(&cc->dtmin_bits, ~0ull);
      prepare_next(P, *cc, P.ctrl[cur ^ 1], mm, false, cur ^ 1);
      __threadfence();
      *P.cursor = cur ^ 1;
    }
  }
}

}  // namespace

int fused_grid_x_f32(int cols) { return (cols - 2 + kOutCols - 1) / kOutCols; }

// Rows per wave: long strips amortise the two-row halo, but a grid must also
// give the device enough waves (about 8 per CU) to hide memory latency.
int default_strip_rows(int rows, int cols) {
  const int nbx = fused_grid_x_f32(cols);
  const int target_waves = 2048;
  int want_nby = (target_waves + nbx - 1) / nbx;
  int strip = (rows - 2 + want_nby - 1) / want_nby;
  return strip < 16 ? 16 : (strip > 64 ? 64 : strip);
}

int launch_step_f32(const BatchArgs& a, int nwork, hipStream_t s) {
  if (nwork <= 0) return TR_OK;
  hipLaunchKernelGGL(k_step_f32, dim3(unsigned(nwork)), dim3(kLanes), 0, s, a);
  return launch_check("step_f32");
}

int launch_prime_f32(const InstF32* d_inst, int member, const InstF32& host_inst, hipStream_t s) {
  int64_t n = int64_t(host_inst.rows) * host_inst.cols;
  unsigned g = unsigned(std::min<int64_t>((n + kLanes - 1) / kLanes, 2048));
  hipLaunchKernelGGL(k_prime_f32, dim3(g), dim3(kLanes), 0, s, d_inst, member);
  return launch_check("prime_f32");
}

}  // namespace tr

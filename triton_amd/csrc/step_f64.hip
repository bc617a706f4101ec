// step_f64.hip -- fused FP64 timestep, one launch per step.
//
// Each 64-lane wave owns a strip of 60 interior columns (lanes 2..61; lanes
// 0,1,62,63 are the halo) and marches down `strip_rows` rows plus a two-row
// halo, keeping a sliding window of rows in registers:
//   load row r -> x faces of row r (neighbour lane via shuffle) -> the y face
//   between rows r and r-1 -> update row r-1 -> wet/dry and CFL dt of row r-2
//   -> store row r-2.
// Every face is computed once per wave and its two sides go to the two cells
// that share it (the east side reaches the next lane by a cross-lane shift).
// The two-row/two-lane halo makes the fused pass exact: wet_dry needs the
// neighbours' post-update depth (src/kernels_wetdry.h:56-75) and the next dt
// the post-wet_dry discharge (src/kernels.h:176-185).  The expressions are
// those of the reference in the same order (roe_f64.hpp); with
// -ffp-contract=off the result equals the unfused sequence bit for bit.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "roe_f64.hpp"
#include "sround.hpp"
#include "step.hpp"
#include "step_ctrl.hpp"

namespace tr {

constexpr int kLanes = 64;
constexpr int kOutCols = 60;

// Initial dt: per-cell CFL over every cell (ghosts included); the close
// kernel then prepares the control block of step 0.
__global__ __launch_bounds__(64) void k_prime_f64(StepParams<double> p) {
  const int n = p.rows * p.cols;
  unsigned long long best = ~0ull;
  for (int id = blockIdx.x * kLanes + threadIdx.x; id < n; id += gridDim.x * kLanes) {
    double h = p.h_in[id];
    double sq = fmax(sqrt(h), 0.0);
    unsigned long long b = dbits(cell_dt_f64(h, p.qx_in[id], p.qy_in[id], sq, p.cn_dx, p.hextra));
    best = b < best ? b : best;
  }
  best = wave_min_bits(best);
  if (threadIdx.x == 0) publish_min(p.ctrl + p.cur, best);
}

// One thread: prepare the next slot from this step's block, or carry a
// halted block so launches queued past the halt stay no-ops.
__global__ __launch_bounds__(64) void k_close_f64(StepParams<double> p, int after_step) {
  if (threadIdx.x != 0) return;
  if (after_step && p.ctrl[p.cur].halt) carry_halted(p);
  else close_one(p, p.cur, after_step != 0);
}

// kR32: every updated h, qx, qy is rounded to FP32 before it is used or
// stored, by the FP32 form's storage rule: the depth after the update and
// the rain stochastically with the FP32 form's dither (sround.hpp), then the
// point sources and the discharges to nearest.  FP64 arithmetic on
// FP32-stored state: the accuracy yardstick of the FP32 form
// (TR_FORM_F64_R32).
// kFast: the per-face dt/dx scaling and each cell's h dx 0.5 / dt divide by
// dx and dt, fixed for the launch, through div_fixed (the same bits as the
// IEEE division when dx and dt are in its range, which the kernel checks once
// per wave); otherwise every quotient is an IEEE division.
template <bool kOpen, bool kBC, bool kR32, bool kFast>
__device__ __forceinline__ void step_strip_f64(const StepParams<double>& p, const StepCtrl* cc, int lane) {
  const double dt = cc->dt;
  const double simtime = cc->simtime;
  const int roff_row = cc->roff_row;
  const double* __restrict__ src_add = p.src_add + p.cur * p.nsrc;
  const unsigned dither_step = unsigned(p.it_count0 + cc->step) * kDitherStep;
  const double dx = p.dx, hextra = p.hextra;
  // dt * d / dx of a face fluctuation d, which the flush leaves +0 or at
  // least 1e-12 in magnitude (src/kernels_flux.h:255-273); a fluctuation
  // that is already infinite (a run that has overflowed) gives NaN here where
  // the IEEE division gives an infinity
  FixedDiv64 rdx = fixed_div64(dx), rdt = fixed_div64(dt);
  // wave-uniform: the reciprocals stay in scalar registers for the whole strip
  rdx.y = uniform_f64(rdx.y);
  rdt.y = uniform_f64(rdt.y);
  auto scale = [&](double d) { return kFast ? div_fixed(dt * d, rdx) : dt * d / dx; };
  const int R = p.rows, C = p.cols;
  const int tile = xcd_slot(int(blockIdx.x), int(p.nblocks));
  const int bx = tile % p.nbx, by = tile / p.nbx;

  const int c = kOutCols * bx - 1 + lane;
  const bool col_ok = c >= 0 && c < C;
  const bool col_int = c >= 1 && c <= C - 2;
  const bool xface_ok = c >= 0 && c <= C - 2 && lane < kLanes - 1;
  const bool out_lane = lane >= 2 && lane <= kOutCols + 1 && col_int;
  const bool ghost_w = c == 0 && lane == 1;
  const bool ghost_e = c == C - 1 && lane >= 3 && lane <= kLanes - 2;
  const bool col_mine = out_lane || ghost_w || ghost_e;
  const bool edge_we = bx == 0 || bx == p.nbx - 1;

  const int e_lo = 1 + p.strip_rows * by;
  const int e_hi = min(e_lo + p.strip_rows - 1, R - 2);
  const int e_min = by == 0 ? 0 : e_lo;
  const int e_max = by == p.nby - 1 ? R - 1 : e_hi;

  // previous row of the input state and its per-cell face inputs
  double ph = 0.0, pqx = 0.0, pqy = 0.0, pz = 0.0, psq = 0.0, pn2 = 0.0;
  Cell64 pk{0.0, 0.0, 0.0};
  double prate = 0.0;
  // the west side of the previous row's x faces (rhs*1)
  double px1h = 0.0, px1x = 0.0, px1y = 0.0;
  // the previous row's two back-side sums, x face then y face (the
  // reference's rhs*0 = flux_x's value + flux_y's), summed when both are known
  double ps0h = 0.0, ps0x = 0.0, ps0y = 0.0;
  // updated rows r-2 (A) and r-3 (B)
  double ah = 0.0, aqx = 0.0, aqy = 0.0, az = 0.0;
  double bh = 0.0, bz = 0.0;
  // external boundary values carried to the south ghost row
  SouthCarry<double> sc{0.0, 0.0, 0.0, 0.0, 0.0, 0.0, false};
  unsigned long long best = ~0ull;

  for (int r = e_min - 2; r <= e_max + 2; ++r) {
    // A: load row r
    const bool row_ok = r >= 0 && r < R;
    double sh = 0.0, sqx = 0.0, sqy = 0.0, sz = 0.0, sn2 = 0.0, srate = 0.0;
    const long long idx = (long long)r * C + c;
    if (row_ok && col_ok) {
      sh = p.h_in[idx];
      sqx = p.qx_in[idx];
      sqy = p.qy_in[idx];
      sz = p.dem[idx];
      if (p.nman2) sn2 = p.nman2[idx];
      if (p.num_runoffs > 0 && r >= 1 && r <= R - 2 && col_int) {
        int zone = p.runoff_id ? p.runoff_id[idx] : 0;
        srate = p.roff_rate[zone * p.roff_nrows + roff_row];
      }
    }
    const double ssq = fmax(sqrt(sh), 0.0);
    Cell64 sk;
    if (kFast) {
      // h dx 0.5 / dt by div_fixed unless some lane's numerator is outside
      // its range (below 2^-900, above 2^900, not finite: then the whole wave
      // divides)
      const double num = sh * dx * 0.5;
      sk.u = 0.0;
      sk.v = 0.0;
      if (sh >= hextra) {
        sk.u = sqx / sh;
        sk.v = sqy / sh;
      }
      if (__any(!(num == 0.0 || (num >= 0x1p-900 && num <= 0x1p900)))) sk.hd = num / dt;
      else sk.hd = div_fixed(num, rdt);
    } else {
      sk = cell64(sh, sqx, sqy, dx, dt, hextra);
    }

    // C: y face between row r (south) and row r-1 (north)
    double y0h = 0.0, y0x = 0.0, y0y = 0.0, y1h = 0.0, y1x = 0.0, y1y = 0.0;
    if (row_ok && r >= 1 && col_int) {
      Face64 f;
      if (roe_face_f64<0, 1>(sh, ph, sqx, pqx, sqy, pqy, sk, pk, sz, pz, ssq, psq, f)) {
        y0h = scale(f.d1[0]);
        y0x = scale(f.d1[1]);
        y0y = scale(f.d1[2]);
        y1h = scale(f.d2[0]);
        y1x = scale(f.d2[1]);
        y1y = scale(f.d2[2]);
      }
    }

    // D: update row r-1 (update_cells, runoff, point sources)
    const int ru = r - 1;
    double nh = ph, nqx = pqx, nqy = pqy;
    const bool ru_int = ru >= 1 && ru <= R - 2;
    if (ru_int && col_int) {
      update_cell_f64(ph, pqx, pqy, pn2, ps0h, px1h + y1h, ps0x, px1x + y1x, ps0y, px1y + y1y, dt, hextra,
                      nh, nqx, nqy);
      if (p.num_runoffs > 0) {
        double hc = nh;
        hc += prate * dt;
        if (hc < hextra) {
          if (hc < 0.0) hc = 0.0;
          nqx = 0.0;
          nqy = 0.0;
        }
        nh = hc;
      }
    }
    if (kR32 && ru_int && col_int)
      nh = double(sround_f64(nh, int(unsigned(c) * kDitherCol + fmix32(dither_step + unsigned(ru) * kDitherRow))));
    if (ru_int) {
      for (int k = 0; k < p.nsrc; k++) {
        if (p.src_row[k] == ru && p.src_col[k] == c) nh = nh + src_add[k];
      }
    }
    if (kR32) {
      nh = double(float(nh));
      nqx = double(float(nqx));
      nqy = double(float(nqy));
    }
    if (kOpen) {
      // transmissive ghosts: west/east columns first, then north/south rows
      const double rh = from_right(nh), rqx = from_right(nqx), rqy = from_right(nqy);
      const double lh = from_left(nh), lqx = from_left(nqx), lqy = from_left(nqy);
      if (ru_int && c == 0) { nh = rh; nqx = rqx; nqy = rqy; }
      if (ru_int && c == C - 1) { nh = lh; nqx = lqx; nqy = lqy; }
      if (ru == 1) { ah = nh; aqx = nqx; aqy = nqy; }          // row 0 <- row 1
      if (ru == R - 1) { nh = ah; nqx = aqx; nqy = aqy; }      // row R-1 <- row R-2
    }
    // external boundary segments after the open-boundary copies
    if (kBC)
      extbc_row(p.bc, ru, c, R, C, col_int, edge_we, kOpen, simtime, dt, nh, nqx, nqy, ah, aqx, aqy,
                sc);

    // E: wet/dry, CFL dt and store of row r-2 (slot A)
    const int e = r - 2;
    if (e >= e_min && e <= e_max && e >= 0) {
      double hq = ah, qxq = aqx, qyq = aqy;
      const double hE = from_right(ah), zE = from_right(az);
      const double hW = from_left(ah), zW = from_left(az);
      const bool cell_int = e >= 1 && e <= R - 2 && col_int;
      if (cell_int && hq > hextra) {
        const double lev = hq + az;
        if ((lev < zE && hE < kEps12) || (lev < zW && hW < kEps12)) qxq = 0.0;
        if ((lev < bz && bh < kEps12) || (lev < pz && nh < kEps12)) qyq = 0.0;
      }
      if (kR32) {  // values the boundary segments wrote
        hq = double(float(hq));
        qxq = double(float(qxq));
        qyq = double(float(qyq));
      }
      if (col_mine) {
        const long long o = (long long)e * C + c;
        p.h_out[o] = hq;
        p.qx_out[o] = qxq;
        p.qy_out[o] = qyq;
        if (p.maxh && cell_int && hq > p.maxh[o]) p.maxh[o] = hq;
        const double sq = fmax(sqrt(hq), 0.0);
        unsigned long long b = dbits(cell_dt_f64(hq, qxq, qyq, sq, p.cn_dx, hextra));
        best = b < best ? b : best;
      }
    }

    // B: x face (c, c+1) of row r, after the row above is finished, so its
    // face sums are no longer live beside this face's temporaries
    double x0h = 0.0, x0x = 0.0, x0y = 0.0, x1h, x1x, x1y;
    {
      const double hR = from_right(sh), qxR = from_right(sqx), qyR = from_right(sqy);
      const double zR = from_right(sz), sqR = from_right(ssq);
      const Cell64 kR{from_right(sk.u), from_right(sk.v), from_right(sk.hd)};
      double e0 = 0.0, e1 = 0.0, e2 = 0.0;
      if (row_ok && r >= 1 && r <= R - 2 && xface_ok) {
        Face64 f;
        if (roe_face_f64<1, 0>(sh, hR, sqx, qxR, sqy, qyR, sk, kR, sz, zR, ssq, sqR, f)) {
          x0h = scale(f.d1[0]);
          x0x = scale(f.d1[1]);
          x0y = scale(f.d1[2]);
          e0 = scale(f.d2[0]);
          e1 = scale(f.d2[1]);
          e2 = scale(f.d2[2]);
        }
      }
      x1h = from_left(e0);
      x1x = from_left(e1);
      x1y = from_left(e2);
    }

    // rotate the window
    bh = ah; bz = az;
    ah = nh; aqx = nqx; aqy = nqy; az = pz;
    ph = sh; pqx = sqx; pqy = sqy; pz = sz; psq = ssq; pn2 = sn2; prate = srate; pk = sk;
    ps0h = x0h + y0h; ps0x = x0x + y0x; ps0y = x0y + y0y;
    px1h = x1h; px1x = x1x; px1y = x1y;
  }
  best = wave_min_bits(best);
  if (lane == 0) publish_min(p.ctrl + p.cur, best);
}

// kW4: built for four waves per SIMD (at most 128 VGPRs, which spills a few
// values the compiler picks) instead of three; the same operations, so the
// same bits.  Which is faster depends on the grid and the strip height, so it
// is a launch knob the autotuner sets (tr_sim_set_waves).
template <bool kOpen, bool kBC, bool kR32, bool kW4>
__global__ __launch_bounds__(64, kW4 ? 4 : 1) void k_step_f64(StepParams<double> p) {
  const int lane = threadIdx.x;
  const StepCtrl* cc = p.ctrl + p.cur;
  if (cc->halt) return;
  // dt * 1e-12 must stay above 2^-900 for the scaled fluctuations
  if (div_fixed_ok(p.dx) && div_fixed_ok(cc->dt)) step_strip_f64<kOpen, kBC, kR32, true>(p, cc, lane);
  else step_strip_f64<kOpen, kBC, kR32, false>(p, cc, lane);
}

// Host launchers (called by the simulation driver).
int launch_prime_f64(const StepParams<double>& p, hipStream_t s) {
  int64_t n = int64_t(p.rows) * p.cols;
  unsigned g = unsigned(std::min<int64_t>((n + kLanes - 1) / kLanes, 2048));
  hipLaunchKernelGGL(k_prime_f64, dim3(g), dim3(kLanes), 0, s, p);
  hipLaunchKernelGGL(k_close_f64, dim3(1), dim3(kLanes), 0, s, p, 0);
  return launch_check("prime_f64");
}

template <bool kW4>
void launch_step_f64_w(const StepParams<double>& p, hipStream_t s) {
  const dim3 grid(p.nblocks), blk(kLanes);
  if (p.round32) {
    if (p.open_bc) hipLaunchKernelGGL((k_step_f64<true, false, true, kW4>), grid, blk, 0, s, p);
    else hipLaunchKernelGGL((k_step_f64<false, false, true, kW4>), grid, blk, 0, s, p);
  } else {
    if (p.open_bc) hipLaunchKernelGGL((k_step_f64<true, false, false, kW4>), grid, blk, 0, s, p);
    else hipLaunchKernelGGL((k_step_f64<false, false, false, kW4>), grid, blk, 0, s, p);
  }
}

int launch_step_f64(const StepParams<double>& p, hipStream_t s) {
  const dim3 grid(p.nblocks), blk(kLanes);
  if (p.bc.n > 0) {
    // boundary segments: the three-wave build only (four waves would spill
    // the segment carry)
    if (p.round32) {
      if (p.open_bc) hipLaunchKernelGGL((k_step_f64<true, true, true, false>), grid, blk, 0, s, p);
      else hipLaunchKernelGGL((k_step_f64<false, true, true, false>), grid, blk, 0, s, p);
    } else {
      if (p.open_bc) hipLaunchKernelGGL((k_step_f64<true, true, false, false>), grid, blk, 0, s, p);
      else hipLaunchKernelGGL((k_step_f64<false, true, false, false>), grid, blk, 0, s, p);
    }
  } else if (p.waves4) {
    launch_step_f64_w<true>(p, s);
  } else {
    launch_step_f64_w<false>(p, s);
  }
  hipLaunchKernelGGL(k_close_f64, dim3(1), blk, 0, s, p, 1);
  return launch_check("step_f64");
}

int fused_grid_x(int cols) { return (cols - 2 + kOutCols - 1) / kOutCols; }

}  // namespace tr

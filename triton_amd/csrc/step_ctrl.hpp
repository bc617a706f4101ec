// step_ctrl.hpp -- device routines that close a fused step: the wave-level
// dt min, and the single-thread bookkeeping that the reference's host loop does between steps
// (src/triton_simulate.h:54-134, :202-226, :246-285).
#pragma once

#include "common.hpp"
#include "step.hpp"

namespace tr {

// Wave64 min of uint64 values via cross-lane shuffles.
__device__ __forceinline__ unsigned long long wave_min_bits(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

// Prepare `nx` (the control block of the next step) from `pv` (the step just
// done, or the initial block when after_step is false).  Single thread.
template <class P>
__device__ void prepare_next(const P& p, const StepCtrl& pv, StepCtrl& nx,
                             unsigned long long dtmin_bits, bool after_step, int nxt_slot) {
  double simtime = after_step ? pv.simtime + pv.dt : pv.simtime;
  int print_id = pv.print_id;
  long long step = pv.step + (after_step ? 1 : 0);
  bool printed = false;
  if (after_step && simtime >= p.print_interval * (print_id + 1)) {
    print_id++;
    printed = true;
  }
  // compute_global_dt (src/triton_simulate.h:202-226)
  double dt = bitsd(dtmin_bits);
  if (dtmin_bits == ~0ull) dt = kMaxDt;
  if (dt >= kMaxDt - 1.0) dt = p.init_dt;
  double next_print = p.print_interval * (print_id + 1);
  if (simtime + dt > next_print) dt = next_print - simtime;

  // runoff row, advanced by at most one per step (src/triton_simulate.h:246-258)
  int row = pv.roff_row;
  double rain_vol = 0.0;
  if (p.num_runoffs > 0) {
    int last = p.roff_nrows - 1;
    if (simtime > p.roff_time[last]) row = last;
    else if (simtime > p.roff_time[row + 1]) row++;
    for (int z = 0; z < p.num_runoffs; z++) {
      const double rate = p.roff_rate[z * p.roff_nrows + row];
      rain_vol += rate * dt * p.zone_area[z];
      // depth each zone's cells gain this step (src/kernels.h:134-135)
      if (p.zone_add) p.zone_add[nxt_slot * p.num_runoffs + z] = rate * dt;
    }
  }
  // hydrograph rows and the depth each source adds (src/triton_simulate.h:264-285,
  // src/kernels.h:60-81)
  int lo = pv.idx_low, hi = pv.idx_high;
  double src_vol = 0.0;
  if (p.nsrc > 0) {
    hi = p.hyg_nrows - 1;
    bool found = false;
    for (int i = lo; i < hi; i++) {
      if (p.hyg_time[i + 1] > simtime) { lo = i; hi = i + 1; found = true; break; }
    }
    if (!found) lo = hi;
    double* add = p.src_add + nxt_slot * p.nsrc;
    for (int k = 0; k < p.nsrc; k++) {
      int s = p.src_id[k];
      double flo = p.hyg_val[lo * p.nsrc + s], fhi = p.hyg_val[hi * p.nsrc + s];
      double q;
      if (lo == hi) q = flo;
      else q = flo + (((fhi - flo) * (simtime - p.hyg_time[lo])) / (p.hyg_time[hi] - p.hyg_time[lo]));
      add[k] = (q * dt) / (p.dx * p.dx);
      src_vol += q * dt;
    }
  }
  // observation points (src/triton_simulate.h:78-81): sampled after a step
  // once simtime has moved print_observation past the last boundary, which
  // then advances by one interval
  double last_obs = pv.last_obs;
  bool observed = false;
  if (after_step && p.time_series && simtime - last_obs >= p.print_observation) {
    last_obs += p.print_observation;
    observed = true;
  }
  // the dt sum the reference writes as time_step= (src/triton_simulate.h:
  // 66-75): accumulated every step, cleared on each it_print boundary
  double dt_acc = pv.dt_acc;
  if (after_step) {
    dt_acc += pv.dt;
    if (p.it_print > 0 && (p.it_count0 + step) % p.it_print == 0) dt_acc = 0.0;
  }
  bool finished = !(simtime < p.sim_duration);
  bool halt = finished || !(simtime < p.t_stop) || step >= p.step_stop ||
              (p.halt_on_print && printed) || (p.halt_on_obs && observed);

  nx.simtime = simtime;
  nx.dt = dt;
  reset_dtmin(nx);
  nx.halt = halt ? 1 : 0;
  nx.step = step;
  nx.print_id = print_id;
  nx.roff_row = row;
  nx.idx_low = lo;
  nx.idx_high = hi;
  nx.finished = finished ? 1 : 0;
  nx.vol_in = after_step ? pv.vol_in + pv.step_vol_in : pv.vol_in;
  nx.step_vol_in = rain_vol + src_vol;
  nx.observed = observed ? 1 : 0;
  nx.last_obs = last_obs;
  nx.dt_acc = dt_acc;
}

// A skipped (halted) step carries its control block and source increments
// to the other slot unchanged, so launches queued past the halt stay no-ops.
template <class P>
__device__ void carry_halted(const P& p) {
  const StepCtrl& pv = p.ctrl[p.cur];
  StepCtrl& nx = p.ctrl[p.cur ^ 1];
  nx = pv;
  reset_dtmin(nx);
  for (int k = 0; k < p.nsrc; k++)
    p.src_add[(p.cur ^ 1) * p.nsrc + k] = p.src_add[p.cur * p.nsrc + k];
}

// Lane 0 of a wave publishes its dt min with a relaxed device-scope atomic
// into the slot of its block's XCD.
// The step kernel does nothing else with the control block: the close kernel
// launched after it (stream order makes every min visible) prepares the next
// slot.  A grid-wide "last wave" inside the step kernel would need a
// device-scope fence per wave, which on this multi-die part writes back and
// invalidates the die's L2 each time.
__device__ __forceinline__ void publish_min(StepCtrl* cc, unsigned long long wave_min) {
  if (wave_min != ~0ull) atomicMin(&cc->dtmin[blockIdx.x & (kDtParts - 1)].bits, wave_min);
}

// Close kernel body for one instance (single thread): prepare the other slot
// from this step's control block.  Halted steps carry the block unchanged
// when the slot is host-driven (carry) or leave the cursor where it is.
template <class P>
__device__ void close_one(const P& p, int cur, bool after_step) {
  const StepCtrl& cc = p.ctrl[cur];
  prepare_next(p, cc, p.ctrl[cur ^ 1], dtmin_of(cc), after_step, cur ^ 1);
}

}  // namespace tr

// step.hpp -- device-side control block and launch parameters of the fused
// timestep (one launch per step; dt, simulated time, print bookkeeping and
// forcing indices stay on the device, so the host never waits per step).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "extbc.hpp"

namespace tr {

// Control block of one step.  Two slots ping-pong: step n reads slot n&1 and
// reduces its per-cell dt into that slot; a one-thread close kernel then
// writes slot (n+1)&1 for the next step (the host-side loop of
// src/triton_simulate.h:54-134 and :202-226, moved on the device).
// The step's dt minimum is published into one of kDtParts slots, each on
// its own 128-B line, by block index modulo 8 (the XCD the block runs on), so
// the waves of a launch do not all contend for one address; the close
// kernel reduces the slots.
constexpr int kDtParts = 8;
struct alignas(128) DtPart {
  unsigned long long bits;
};

struct alignas(128) StepCtrl {
  double simtime;                 // time at the start of this step
  double dt;                      // dt applied in this step (already clamped)
  int pad2;
  int halt;                       // nonzero: the step is skipped
  long long step;                 // steps completed before this one
  int print_id;
  int roff_row;                   // runoff row applied in this step
  int idx_low, idx_high;          // hydrograph rows bracketing simtime
  int finished;                   // simtime >= sim_duration
  int observed;                   // the step just done completed an observation interval
  double vol_in;                  // FP64 rain + inflow volume before this step
  double step_vol_in;             // volume this step adds
  double last_obs;                // time of the last observation interval boundary
  double dt_acc;                  // sum of dt since the last it_print boundary
  DtPart dtmin[kDtParts];         // min per-cell dt of this step's result (bits), per XCD
};

__host__ __device__ inline void reset_dtmin(StepCtrl& c) {
  for (int i = 0; i < kDtParts; i++) c.dtmin[i].bits = ~0ull;
}
__host__ __device__ inline unsigned long long dtmin_of(const StepCtrl& c) {
  unsigned long long m = ~0ull;
  for (int i = 0; i < kDtParts; i++) m = c.dtmin[i].bits < m ? c.dtmin[i].bits : m;
  return m;
}

constexpr int kMaxSrcFused = 256;  // point sources handled by the fused step

template <class Real>
struct StepParams {
  int rows, cols;           // padded grid
  int strip_rows;           // output rows per wave
  int nbx, nby;             // waves along x and y
  unsigned nblocks;         // nbx * nby
  double dx, hextra, cn_dx;
  const Real* h_in;
  const Real* qx_in;
  const Real* qy_in;
  Real* h_out;
  Real* qx_out;
  Real* qy_out;
  const double* dem;        // FP64 form: bed
  const Real* dzx;          // FP32 form: z[c+1]-z[c] per x face (rows*cols)
  const Real* dzy;          // FP32 form: z[c-cols]-z[c] per y face (rows*cols)
  const Real* nman2;        // NULL when no cell has friction
  Real* maxh;               // NULL: not tracked
  const int* runoff_id;     // NULL: zone 0 everywhere
  const double* roff_rate;  // zone-major [zone*roff_nrows + row]
  const double* roff_time;
  int roff_nrows, num_runoffs;
  const double* zone_area;  // interior cells of each zone * dx^2
  int nsrc;
  const int* src_row;       // sources sorted by (row, col), stable
  const int* src_col;
  const int* src_id;        // original list index of each sorted source
  const double* hyg_time;
  const double* hyg_val;
  int hyg_nrows;
  double* src_add;          // [2][nsrc] depth added this step, by slot
  double* zone_add;         // [2][num_runoffs] rain depth this step, by slot (or NULL)
  int open_bc;
  StepCtrl* ctrl;           // [2]
  int cur;                  // slot of this step
  double sim_duration, print_interval, init_dt, t_stop;
  long long step_stop;
  int halt_on_print;
  int time_series;          // observation points are sampled
  double print_observation;
  int halt_on_obs;          // stop after a step that completes an observation interval
  int it_print;             // the reference's log cadence (the dt sum resets there)
  long long it_count0;      // iterations before the start (hotstart)
  ExtBC bc;                 // external boundary segments (bc.n == 0: none)
  int round32;              // FP64 form: round updated h, qx, qy to FP32 (TR_FORM_F64_R32)
  int waves4;               // FP64 fused step: the build for four waves per SIMD (else three)
};

}  // namespace tr

// batch.hpp -- descriptor of one member of a batched FP32 launch.
//
// A batched step launches one 64-lane wave per (member, strip); a small work
// table maps the wave index to its member and strip, so members of any size
// share one launch and each keeps its own dt, simulated time and forcing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "step.hpp"

namespace tr {

struct InstF32 {
  int rows, cols;            // padded grid
  int strip_rows, nbx, nby;  // waves per member: nbx * nby
  unsigned nblocks;
  double dx, hextra, cn_dx;  // FP64 copies for the step bookkeeping
  float dx_f, hextra_f, cn_dx_f;
  float eps_f;               // dry / flush threshold
  float* h[2];               // ping-pong state, FP32
  float* qx[2];
  float* qy[2];
  const float* dzx;          // z[r][c+1]-z[r][c], computed in FP64, stored FP32
  const float* dzy;          // z[r-1][c]-z[r][c]
  const float* nman2;        // NULL when frictionless
  float* maxh;               // NULL: not tracked
  const int* runoff_id;      // NULL: zone 0 everywhere
  const double* roff_rate;
  const double* roff_time;
  int roff_nrows, num_runoffs;
  const double* zone_area;
  double* zone_add;          // [2][num_runoffs]
  int nsrc;
  const int* src_row;
  const int* src_col;
  const int* src_id;
  const double* hyg_time;
  const double* hyg_val;
  int hyg_nrows;
  double* src_add;           // [2][nsrc]
  int open_bc;
  StepCtrl* ctrl;            // [2]
  int* cursor;               // slot of the current step
  double sim_duration, print_interval, init_dt, t_stop;
  long long step_stop;
  int halt_on_print;
};

struct BatchArgs {
  const InstF32* inst;
  const int2* work;          // (member, wave index within the member)
  int nwork;
};

int launch_step_f32(const BatchArgs& a, int nwork, hipStream_t s);
int launch_prime_f32(const InstF32* d_inst, int member, const InstF32& host_inst, hipStream_t s);
int fused_grid_x_f32(int cols);
int default_strip_rows(int rows, int cols);

}  // namespace tr

// batch.hpp -- descriptor of one member of a batched FP32 launch.
//
// A batched step launches a fixed number of 64-lane waves; a piece table
// hands each wave an equal share of the rows of all members' column strips,
// so members of any size share one launch and each keeps its own dt,
// simulated time and forcing.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "step.hpp"

namespace tr {

struct InstF32 {
  int rows, cols;            // padded grid
  int nbx;                   // column strips of 60 interior columns (one cell per lane)
  double dx, hextra, cn_dx;  // FP64 copies for the step bookkeeping
  float dx_f, hextra_f, cn_dx_f;
  float eps_f;               // dry / flush threshold
  int round_q;               // discharges rounded stochastically (else to nearest)
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
  int time_series;          // observation points are sampled
  double print_observation;
  int halt_on_obs;          // stop after a step that completes an observation interval
  int it_print;             // the reference's log cadence (the dt sum resets there)
  long long it_count0;      // iterations before the start (hotstart)
  ExtBC bc;                 // external boundary segments (bc.n == 0: none)
};

// Control blocks of one member (two slots and the cursor naming the current
// one); all members' pairs sit in one array, so the host reads every status
// with one copy and kernels address a member's pair by its index without
// first reading its descriptor.
struct alignas(128) CtrlPair {
  StepCtrl ctrl[2];
  int cursor;
  int pad[31];
};

// A piece is a run of rows of one column strip of one member, and the work
// of one wave.
struct Piece {
  int member;
  int bx;          // column strip
  int row_first;   // first and last row finalised (0 and rows-1 are ghosts)
  int row_last;
};

struct BatchArgs {
  const InstF32* inst;
  CtrlPair* ctrls;           // [member] control blocks (the same memory inst[m].ctrl points into)
  const Piece* pieces;       // [nwaves]: wave w steps pieces[w]
  int nwaves;
  const int* members;        // members in the partition (the close kernel's list)
  int nmembers;
  int first_member;          // the member when nmembers == 1 (the close then reads no list)
  int any_bc;                // some member has external boundary segments
};

int launch_step_f32(const BatchArgs& a, int nwaves, hipStream_t s);
int launch_prime_f32(const InstF32* d_inst, CtrlPair* d_ctrls, int member, const InstF32& host_inst,
                     hipStream_t s);
int fused_grid_x_f32(int cols);
int default_strip_rows(int rows, int cols);       // FP64 fused step, for its default build
int default_waves_per_simd(int rows, int cols);  // that build: 3 or 4 waves per SIMD

}  // namespace tr

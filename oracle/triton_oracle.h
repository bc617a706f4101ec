/*
 * triton_oracle.h -- CPU FP64 restatement of TRITON's explicit finite-volume
 * timestep.  TEST INFRASTRUCTURE ONLY: this library is the checker that the
 * HIP path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * The kernels are restated unfused, with the reference's array set, launch
 * ranges and floating-point operation order, so that results are bit-for-bit
 * those of the reference's FP64 Kokkos Serial/OpenMP build (every kernel is
 * per-cell independent and the only reduction is an exact min).  Compile with
 * -ffp-contract=off; sqrt is IEEE, cbrt comes from the host libm.
 *
 * Layout: row-major rows x cols arrays that include the one-cell ghost ring
 * (rows = nrows_dem + 2, cols = ncols_dem + 2); row 0 is the north edge.
 *
 * Pinning (how each function is checked; tests/test_oracle_fixtures.py and
 * tests/test_oracle_invariants.py):
 *   orc_compute_dt_and_sqrt, orc_find_min_dt, orc_flux_x, orc_flux_y,
 *   orc_update_cells (no friction), orc_wet_dry, orc_run loop/dt logic:
 *       reference fixtures test/reference/{circular,paraboloid} (identical
 *       6-decimal output), plus exact lake at rest over a rough bed with dry
 *       islands, x/y mirror and transpose symmetry of a centred dam break,
 *       volume conservation to 1e-13, and Thacker's analytic solution.
 *   orc_update_cells (friction), orc_update_runoff, orc_compute_flow,
 *   orc_compute_extbc_values (type 3): reference fixture test/reference/allatoona;
 *       friction also by the exact lake at rest with n = 0.03.
 *   orc_open_bc_west_east / orc_open_bc_north_south: no reference fixture;
 *       pinned by an exact uniform-flow steady state, the symmetries above
 *       with open boundaries, and a y-uniform dam break that must stay
 *       y-uniform while water leaves through the east edge.
 *   orc_run observation rows and the time_step= dt sum: no reference
 *       fixture; tests/test_oracle_invariants.py derives the recorded steps
 *       and the sum again in Python from a run that records every step.
 *   orc_compute_extbc_values types 0/1/2: no reference fixture; pinned by
 *       closed forms -- type 0 on every edge cell adds nothing to the open
 *       boundaries' copies and leaves every ghost equal to its cell under
 *       walls; type 1 with a linear level series gives the level at the
 *       step's midpoint minus the bed (clamped), first / last level outside
 *       the series; type 2 gives Manning's normal-flow discharge
 *       h^(5/3) sqrt(S)/n (n^2 floored at 1e-6) along each edge's outward
 *       normal.  Nine planted slips in these (time, exponent, floor, sign,
 *       ghost index, corner, factors, bed, clamp) each fail a test.
 */
#ifndef TRITON_ORACLE_H
#define TRITON_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Everything the timestep needs, as the reference's init leaves it
 * (src/triton_init.h:195-355, 723-830, 913-1055, 1299-1357). */
typedef struct orc_case {
  int rows, cols;             /* padded dims */
  double dx;                  /* cell size */
  double courant, hextra;
  double *h, *qx, *qy;        /* state, in/out [rows*cols] */
  const double *dem;          /* bed with ghost ring (walls 1e6 or copies) */
  const double *nman;         /* squared Manning n [rows*cols] */
  double *maxh;               /* running max depth or NULL */
  /* zonal runoff (src/kernels.h:102-152) */
  int num_runoffs, roff_nrows;
  const double *roff_time;    /* [roff_nrows] seconds */
  const double *roff_rate;    /* [num_runoffs*roff_nrows] m/s, zone-major */
  const int *runoff_id;       /* [rows*cols] */
  /* point inflows (src/kernels.h:50-84) */
  int num_src, hyg_nrows;
  const double *hyg_time;     /* [hyg_nrows] seconds */
  const double *hyg_val;      /* [hyg_nrows*num_src] m^3/s, time-major */
  const int *src_pos;         /* [num_src] flat cell index */
  /* external boundary cells (src/kernels.h:236-424) */
  int num_extbc_cells;
  const int *bc_index, *bc_type, *bc_start, *bc_nrows;
  const double *bc_var1, *bc_var2;
  int open_boundaries;
  /* time control (src/triton_simulate.h:12-226) */
  double sim_start, sim_duration, print_interval, print_observation;
  int time_series_flag, checkpoint_id;
  /* observation points (src/triton_simulate.h:31-37, 78-91, 113-121): with
   * time_series_flag and obs_rec set, a run from the case start appends rows
   * (simtime, h[obs], qx[obs], qy[obs]) -- the start, every step that
   * completes a print_observation interval, and the end when the run reaches
   * sim_duration.  obs_rec holds obs_cap rows of 1 + 3*num_obs doubles;
   * *obs_rows counts every row, also those that did not fit. */
  int num_obs;
  const int *obs_pos;
  double *obs_rec;
  int obs_cap;
  int *obs_rows;
  /* log cadence and iterations before the start: the dt sum the reference
   * writes as time_step= (src/triton_simulate.h:41-75) */
  int it_print, it_count0;
} orc_case;

typedef struct orc_result {
  long long steps;            /* iterations executed */
  double simtime;             /* final simulated time */
  double last_dt;             /* dt of the last step */
  int print_id;               /* prints reached */
  double init_dt;             /* all-dry fallback dt used by the run */
  double last_obs;            /* last observation interval boundary */
  double average_dt;          /* dt sum since the last it_print boundary */
  long long it_count;         /* it_count0 + steps */
} orc_result;

/* Kernel restatements (argument meaning as in the reference). */
void orc_compute_dt_and_sqrt(int size, double dx, const double *qx, const double *qy,
                             const double *h, double *sqrth, double *dtarr,
                             double cn, double hextra);
double orc_find_min_dt(int size, const double *dtarr);
void orc_flux_x(int rows, int cols, double dx, double dt, const double *h,
                const double *qx, const double *qy, const double *dem,
                const double *sqrth, double *rhsh0, double *rhsh1, double *rhsqx0,
                double *rhsqx1, double *rhsqy0, double *rhsqy1, double hextra);
void orc_flux_y(int rows, int cols, double dx, double dt, const double *h,
                const double *qx, const double *qy, const double *dem,
                const double *sqrth, double *rhsh0, double *rhsh1, double *rhsqx0,
                double *rhsqx1, double *rhsqy0, double *rhsqy1, double hextra);
void orc_update_cells(int rows, int cols, double dt, double *h, double *qx, double *qy,
                      const double *nman, const double *rhsh0, const double *rhsh1,
                      const double *rhsqx0, const double *rhsqx1, const double *rhsqy0,
                      const double *rhsqy1, double hextra);
void orc_update_runoff(int rows, int cols, double dt, const int *runoff_id,
                       int index_row_runoff, int n_rows_runoff, const double *rate,
                       double *h, double *qx, double *qy, double hextra);
void orc_compute_flow(int num_src, const double *hyg_time, const double *hyg_val,
                      double dx, double dt, double simtime, int idx_low, int idx_high,
                      double *h, const int *src_pos);
void orc_open_bc_west_east(int rows, int cols, double *h, double *qx, double *qy);
void orc_open_bc_north_south(int rows, int cols, double *h, double *qx, double *qy);
void orc_compute_extbc_values(int ncells, int rows, int cols, double dt, double *h,
                              double *qx, double *qy, const double *dem,
                              const double *nman, const int *bc_index,
                              const int *bc_type, const int *bc_start,
                              const int *bc_nrows, const double *var1,
                              const double *var2, double simtime);
void orc_wet_dry(int rows, int cols, const double *h, double *qx, double *qy,
                 const double *dem, double *maxh, double hextra);

/* Per-face solver, exposed for unit tests: fills d1 (west/south cell) and
 * d2 (east/north cell) fluctuations before the dt/dx scaling; returns 0 when
 * both cells are dry (h1<=0 && h2<=0) and nothing is computed. */
int orc_face(double h1, double h2, double qx1, double qx2, double qy1, double qy2,
             double z1, double z2, double sqh1, double sqh2, double nx, double ny,
             double dx, double dt, double hextra, double d1[3], double d2[3]);

/* The reference's simulate() loop (src/triton_simulate.h:54-134) from
 * c->sim_start while simtime < min(t_stop, sim_duration) (t_stop <= 0: no
 * extra limit), or until `max_steps` iterations when max_steps >= 0.  A
 * print raster of the reference is the state when simtime first reaches
 * print_interval*(id+1); pass that product as t_stop to stop there.  State
 * arrays in `c` are advanced in place.  Returns 0 on success, nonzero on an
 * allocation or index error. */
int orc_run(const orc_case *c, double t_stop, long long max_steps, orc_result *res);

/* ---- The single-precision build (triton_oracle_f32.c) ----
 * The same kernels with float state and T variables and the reference's
 * double literals and constants (src/constants.h:37-43 with
 * USE_SINGLE_PRECISION): every mixed expression is evaluated in double and
 * rounded to float where it is stored; EPS12 is 1e-6; FMAX/FMIN are
 * fmaxf/fminf; sqrt and cbrt take the double the argument promotes to.
 * Pinning (tests/test_oracle_f32.py): each orcf_ kernel against its pinned
 * orc_ counterpart on float-representable inputs to float rounding, exactly
 * where the kernel only copies, compares or clamps; the float loop by an
 * exact lake at rest over a rough bed with dry islands, mirror and transpose
 * symmetry of a centred dam break, closed-domain volume to float rounding,
 * and agreement with the FP64 loop on a dam break.  Where in a kernel
 * rounding to float happens (which operand is double) is the reference's C++
 * promotion rule restated; no fixture of the single-precision build exists,
 * so that placement is "parity unpinned" below float rounding. */
void orcf_compute_dt_and_sqrt(int size, float dx, const float *qx, const float *qy, const float *h,
                              float *sqrth, float *dtarr, float cn, float hextra);
float orcf_find_min_dt(int size, const float *dtarr);
int orcf_face(float h1, float h2, float qx1, float qx2, float qy1, float qy2, float z1, float z2,
              float sqh1, float sqh2, float nx, float ny, float dx, float dt, float hextra, float d1[3],
              float d2[3]);
void orcf_flux_x(int rows, int cols, float dx, float dt, const float *h, const float *qx, const float *qy,
                 const float *dem, const float *sqrth, float *rhsh0, float *rhsh1, float *rhsqx0, float *rhsqx1,
                 float *rhsqy0, float *rhsqy1, float hextra);
void orcf_flux_y(int rows, int cols, float dx, float dt, const float *h, const float *qx, const float *qy,
                 const float *dem, const float *sqrth, float *rhsh0, float *rhsh1, float *rhsqx0, float *rhsqx1,
                 float *rhsqy0, float *rhsqy1, float hextra);
void orcf_update_cells(int rows, int cols, float dt, float *h, float *qx, float *qy, const float *nman,
                       const float *rhsh0, const float *rhsh1, const float *rhsqx0, const float *rhsqx1,
                       const float *rhsqy0, const float *rhsqy1, float hextra);
void orcf_update_runoff(int rows, int cols, float dt, const int *runoff_id, int row, int nrow, const float *rate,
                        float *h, float *qx, float *qy, float hextra);
void orcf_compute_flow(int num_src, const float *hyg_time, const float *hyg_val, float dx, float dt,
                       float simtime, int idx_low, int idx_high, float *h, const int *src_pos);
void orcf_open_bc_west_east(int rows, int cols, float *h, float *qx, float *qy);
void orcf_open_bc_north_south(int rows, int cols, float *h, float *qx, float *qy);
void orcf_compute_extbc_values(int ncells, int rows, int cols, float dt, float *h, float *qx, float *qy,
                               const float *dem, const float *nman, const int *idx, const int *type,
                               const int *start, const int *nvars, const float *v1, const float *v2,
                               float simtime);
void orcf_wet_dry(int rows, int cols, const float *h, float *qx, float *qy, const float *dem, float *maxh,
                  float hextra);
/* Walls (bed ghosts high) or open boundaries, no forcing: steps while
 * simtime < t_end (and < max_steps when max_steps >= 0), dt clamped to the
 * next multiple of print_interval; init_dt when every cell is dry.  `work`
 * holds 8*rows*cols floats.  Returns the step count. */
long long orcf_run(int rows, int cols, float dx, float cn, float hextra, int open_bc, float *h, float *qx,
                   float *qy, const float *dem, const float *nman, float t_end, float print_interval,
                   float init_dt, long long max_steps, float *work, float *simtime_out);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif

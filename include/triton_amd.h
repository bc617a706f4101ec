/*
 * triton_amd.h -- C ABI of the MI355X-native TRITON timestep.
 *
 * Drop-in boundary: the reference calls its timestep kernels as C++ templates
 * Kernels::<name><T>(...) on Kokkos' default instance (src/kokkos_utils.h:18-40)
 * from Triton::triton<T>::compute_local_dt / compute_new_state
 * (src/triton_simulate.h:191-346).  Each tr_<name>_f64 below replaces one of
 * them with the same argument order and meaning, device pointers owned by the
 * caller, an explicit hipStream_t (passed as void*, NULL = null stream) and an
 * int status (TR_OK or an error code; tr_last_error() gives the message of
 * the calling thread's last failure).  Arrays are row-major rows x cols with
 * the one-cell ghost ring, row 0 = north, exactly as the reference lays out
 * device_vec (src/constants.h:82-112).
 *
 * Above the per-kernel functions sit the objects the reference has no single
 * call for: a loaded case (the .cfg/DEM/roughness/hydrograph input formats),
 * a device-resident simulation (the simulate() loop, src/triton_simulate.h:
 * 12-159) and a batch of ensemble members, which the serving loop of the
 * Python package (triton_amd/serve.py) drives.
 *
 * Thread-compatibility: distinct objects may be driven from distinct host
 * threads; one object must not be used by two threads at once.
 */
#ifndef TRITON_AMD_H
#define TRITON_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
#define TR_OK 0
#define TR_ERR_INVALID 1   /* bad argument or unsupported option */
#define TR_ERR_INPUT 2     /* unreadable or malformed input file */
#define TR_ERR_HIP 3       /* HIP runtime error */
#define TR_ERR_NOMEM 4     /* allocation failed or over the memory budget */
#define TR_ERR_STATE 5     /* object in the wrong state for the call */

/* Message of the calling thread's last error ("" if none). */
const char *tr_last_error(void);
/* Library version string. */
const char *tr_version(void);

/* ---- precision / execution form of a simulation -------------------------- */
#define TR_FORM_F64_UNFUSED 0 /* reference kernel sequence, dt read back each step */
#define TR_FORM_F64_FUSED 1   /* one fused launch per step, dt and time on device */
#define TR_FORM_F32_FUSED 2   /* FP32 depth/discharge, FP64 time and volumes */
#define TR_FORM_F64_R32 3     /* the FP64 fused step with h, qx, qy rounded to FP32 after
                                 every update: FP64 arithmetic on FP32-stored state, the
                                 yardstick for the FP32 form's storage error (no
                                 counterpart in the reference) */

/* ---- loaded case (input formats) ----------------------------------------- */
typedef struct tr_case tr_case;

/* Plain view of a case; every pointer is host memory owned by the case
 * (tr_case_view_get) or by the caller (a view filled in by the caller and
 * passed to tr_sim_create / tr_batch_add, which copy what they need).
 * Arrays are padded rows x cols unless stated. */
typedef struct tr_case_view {
  int32_t rows, cols;            /* padded dims: nrows_dem + 2, ncols_dem + 2 */
  double dx;                     /* cell size (m) */
  double xll, yll;               /* lower-left corner of the DEM */
  double courant, hextra;
  double *h, *qx, *qy;           /* initial state */
  double *dem;                   /* bed with ghosts (walls 1e6 or copies) */
  double *nman2;                 /* squared Manning n */
  double *maxh;                  /* initial max depth */
  int32_t *runoff_id;            /* zone per cell, -1 without runoff */
  int32_t num_runoffs, roff_nrows, roff_row0;
  double *roff_time;             /* [roff_nrows] s */
  double *roff_rate;             /* [num_runoffs*roff_nrows] m/s, zone-major */
  int32_t num_src, hyg_nrows;
  double *hyg_time;              /* [hyg_nrows] s */
  double *hyg_val;               /* [hyg_nrows*num_src] m^3/s, time-major */
  int32_t *src_pos;              /* [num_src] flat padded cell index */
  int32_t num_extbc_cells;
  int32_t *bc_index, *bc_type, *bc_start, *bc_nrows;
  int32_t bc_nvars;
  double *bc_var1, *bc_var2;     /* [bc_nvars] */
  int32_t open_boundaries;
  double sim_start, sim_duration, print_interval, print_observation, init_dt;
  int32_t time_series_flag, checkpoint_id, it_print;
  int32_t it_count;              /* iterations before the start (hotstart .cfg) */
  int32_t num_obs;               /* observation points (time_series_flag) */
  int32_t *obs_pos;              /* [num_obs] flat padded cell index */
} tr_case_view;

/* Parse a .cfg and load every input it names (src/triton_init.h:33-81).
 * Relative paths in the .cfg are resolved against base_dir (NULL = cwd).
 * input_option SEQ reads whole rasters; PAR (src/triton_init.h:397-563)
 * reads the one-partition set <name>_00.<ext> with the DEM header from
 * header_filename.  A hotstart (checkpoint_id > 0) reads the print's BIN
 * rasters, MH too when max_value_print_option is set. */
int tr_case_load(const char *cfg_path, const char *base_dir, tr_case **out);
/* Fill `out` with pointers into the case (valid until tr_case_free). */
int tr_case_view_get(tr_case *c, tr_case_view *out);
/* Value of a .cfg key as the reference's parser reads it (first occurrence
 * wins, '#' comments and quotes stripped, src/config_utils.h:195-238): copies
 * at most buflen-1 bytes and a NUL into buf and returns the value's length;
 * -1 when the key is absent. */
int32_t tr_case_config_value(const tr_case *c, const char *key, char *buf, int32_t buflen);
void tr_case_free(tr_case *c);

/* Reference output writers (src/output_ascii.h:52-75, src/output_binary.h:65-80):
 * interior of a padded rows x cols host array. */
int tr_write_output_ascii(const char *path, int32_t rows, int32_t cols, const double *a);
int tr_write_output_binary(const char *path, int32_t rows, int32_t cols, const double *a);

/* ---- per-kernel drop-in entry points (FP64, device pointers) ------------- */
/* src/kernels.h:166-187 Kernels::compute_dt_and_sqrt */
int tr_compute_dt_and_sqrt_f64(int32_t size, double dx, const double *qx, const double *qy,
                               const double *h, double *sqrth, double *dt_out, double cn,
                               double hextra, void *stream);
/* src/kernels.h:195-212 Kernels::find_min_dt.  The min goes to the device
 * scalar d_min; when h_min is non-NULL the call also copies it there and
 * synchronises `stream` (the reference returns it by value). */
int tr_find_min_dt_f64(int32_t size, const double *dtarr, double *d_min, double *h_min,
                       void *stream);
/* src/kernels_flux.h:31-285 Kernels::flux_x */
int tr_flux_x_f64(int32_t size, int32_t nrows, int32_t ncols, double dx, double dt,
                  const double *h, const double *qx, const double *qy, const double *dem,
                  const double *sqrth, double *rhsh0, double *rhsh1, double *rhsqx0,
                  double *rhsqx1, double *rhsqy0, double *rhsqy1, double hextra,
                  void *stream);
/* src/kernels_flux.h:308-559 Kernels::flux_y */
int tr_flux_y_f64(int32_t size, int32_t nrows, int32_t ncols, double dx, double dt,
                  const double *h, const double *qx, const double *qy, const double *dem,
                  const double *sqrth, double *rhsh0, double *rhsh1, double *rhsqx0,
                  double *rhsqx1, double *rhsqy0, double *rhsqy1, double hextra,
                  void *stream);
/* src/kernels_state.h:30-104 Kernels::update_cells */
int tr_update_cells_f64(int32_t size, int32_t nrows, int32_t ncols, double dt, double *h,
                        double *qx, double *qy, const double *dem, const double *nman,
                        double *rhsh0, double *rhsh1, double *rhsqx0, double *rhsqx1,
                        double *rhsqy0, double *rhsqy1, double hextra, void *stream);
/* src/kernels.h:102-152 Kernels::update_runoff */
int tr_update_runoff_f64(int32_t size, int32_t nrows, int32_t ncols, double dt,
                         const int32_t *runoff_id, int32_t index_row_runoff,
                         int32_t n_rows_runoff, const double *runoff_intensity, double *h,
                         double *qx, double *qy, double hextra, void *stream);
/* src/kernels.h:50-84 Kernels::compute_flow (sources applied in list order) */
int tr_compute_flow_f64(int32_t size, const double *hyg_time, const double *hyg_val,
                        double dx, double dt, double simtime, int32_t idx_low,
                        int32_t idx_high, double *h, const int32_t *pos, void *stream);
/* src/kernels.h:436-460 Kernels::copy_info_to_exterior_boundaries_west_east */
int tr_open_bc_west_east_f64(int32_t size, int32_t nrows, int32_t ncols, double *h,
                             double *qx, double *qy, void *stream);
/* src/kernels.h:476-512 Kernels::copy_info_to_exterior_boundaries_north_south */
int tr_open_bc_north_south_f64(int32_t size, int32_t nrows, int32_t ncols, double *h,
                               double *qx, double *qy, int32_t rank, int32_t total_process,
                               void *stream);
/* src/kernels.h:236-424 Kernels::compute_extbc_values (cells in list order) */
int tr_compute_extbc_values_f64(int32_t size, int32_t nrows, int32_t ncols, double dt,
                                double *h, double *qx, double *qy, const double *dem,
                                const double *nman, const int32_t *relative_index,
                                const int32_t *extbctype, const int32_t *start_index,
                                const int32_t *nrows_vars, const double *extbcvar1,
                                const double *extbcvar2, double simtime, int32_t rank,
                                int32_t total_process, void *stream);
/* src/kernels_wetdry.h:24-85 Kernels::wet_dry (maxh may be NULL) */
int tr_wet_dry_f64(int32_t size, int32_t nrows, int32_t ncols, double dt, const double *h,
                   double *qx, double *qy, const double *dem, double *maxh, double hextra,
                   int32_t mpi_tasks, void *stream);
/* The device cube root used by the FP64 friction (src/kernels_state.h:88
 * takes cbrt of the depth), y[i] = cbrt(x[i]) for i < n (device pointers),
 * so its accuracy can be checked on its own. */
int tr_cbrt_f64(int32_t n, const double *x, double *y, void *stream);
/* src/kernels.h:521-535 Kernels::copy_observation_points:
 * obs_arr[i] = arr[relative_index[i]] for i < size (device pointers). */
int tr_copy_observation_points_f64(int32_t size, const double *arr, double *obs_arr,
                                   const int32_t *relative_index, void *stream);

/* ---- per-kernel drop-in entry points (FP32, device pointers) ------------- */
/* Kernels::<name><float>, the reference's USE_SINGLE_PRECISION build
 * (src/constants.h:37-50): float state, the double literals _G_, SQRTG and
 * 0.5 promoting mixed expressions to FP64 before each store, EPS12 = 1e-6,
 * FMAX/FMIN = fmaxf/fminf, math functions in double.  Same arguments and
 * ranges as the _f64 functions above. */
/* src/kernels.h:166-187 */
int tr_compute_dt_and_sqrt_f32(int32_t size, float dx, const float *qx, const float *qy,
                               const float *h, float *sqrth, float *dt_out, float cn,
                               float hextra, void *stream);
/* src/kernels.h:195-212 (d_min is the reduction's slot too) */
int tr_find_min_dt_f32(int32_t size, const float *dtarr, float *d_min, float *h_min, void *stream);
/* src/kernels_flux.h:31-285 */
int tr_flux_x_f32(int32_t size, int32_t nrows, int32_t ncols, float dx, float dt, const float *h,
                  const float *qx, const float *qy, const float *dem, const float *sqrth,
                  float *rhsh0, float *rhsh1, float *rhsqx0, float *rhsqx1, float *rhsqy0,
                  float *rhsqy1, float hextra, void *stream);
/* src/kernels_flux.h:308-559 */
int tr_flux_y_f32(int32_t size, int32_t nrows, int32_t ncols, float dx, float dt, const float *h,
                  const float *qx, const float *qy, const float *dem, const float *sqrth,
                  float *rhsh0, float *rhsh1, float *rhsqx0, float *rhsqx1, float *rhsqy0,
                  float *rhsqy1, float hextra, void *stream);
/* src/kernels_state.h:30-104 */
int tr_update_cells_f32(int32_t size, int32_t nrows, int32_t ncols, float dt, float *h, float *qx,
                        float *qy, const float *dem, const float *nman, float *rhsh0,
                        float *rhsh1, float *rhsqx0, float *rhsqx1, float *rhsqy0, float *rhsqy1,
                        float hextra, void *stream);
/* src/kernels.h:102-152 */
int tr_update_runoff_f32(int32_t size, int32_t nrows, int32_t ncols, float dt,
                         const int32_t *runoff_id, int32_t index_row_runoff,
                         int32_t n_rows_runoff, const float *runoff_intensity, float *h,
                         float *qx, float *qy, float hextra, void *stream);
/* src/kernels.h:50-84 (sources applied in list order) */
int tr_compute_flow_f32(int32_t size, const float *hyg_time, const float *hyg_val, float dx,
                        float dt, float simtime, int32_t idx_low, int32_t idx_high, float *h,
                        const int32_t *pos, void *stream);
/* src/kernels.h:436-460 */
int tr_open_bc_west_east_f32(int32_t size, int32_t nrows, int32_t ncols, float *h, float *qx,
                             float *qy, void *stream);
/* src/kernels.h:476-512 */
int tr_open_bc_north_south_f32(int32_t size, int32_t nrows, int32_t ncols, float *h, float *qx,
                               float *qy, int32_t rank, int32_t total_process, void *stream);
/* src/kernels.h:236-424 (cells in list order) */
int tr_compute_extbc_values_f32(int32_t size, int32_t nrows, int32_t ncols, float dt, float *h,
                                float *qx, float *qy, const float *dem, const float *nman,
                                const int32_t *relative_index, const int32_t *extbctype,
                                const int32_t *start_index, const int32_t *nrows_vars,
                                const float *extbcvar1, const float *extbcvar2, float simtime,
                                int32_t rank, int32_t total_process, void *stream);
/* src/kernels_wetdry.h:24-85 (maxh may be NULL) */
int tr_wet_dry_f32(int32_t size, int32_t nrows, int32_t ncols, float dt, const float *h,
                   float *qx, float *qy, const float *dem, float *maxh, float hextra,
                   int32_t mpi_tasks, void *stream);
/* src/kernels.h:521-535 */
int tr_copy_observation_points_f32(int32_t size, const float *arr, float *obs_arr,
                                   const int32_t *relative_index, void *stream);

/* ---- device-resident simulation ------------------------------------------ */
typedef struct tr_sim tr_sim;

typedef struct tr_sim_status {
  double simtime;       /* simulated time reached */
  double dt;            /* dt of the last step taken */
  int64_t steps;        /* steps taken since creation */
  int32_t print_id;     /* print intervals completed */
  int32_t finished;     /* simtime >= sim_duration */
  double volume;        /* FP64 water volume over interior cells (m^3) */
  double volume_in;     /* FP64 integrated rain + inflow volume (m^3) */
  int32_t observed;     /* the last step completed an observation interval
                           (src/triton_simulate.h:78-91) */
  int32_t reserved;
  double last_obs;      /* time of the last observation interval boundary */
  double average_dt;    /* sum of dt since the last it_print boundary: what the
                           reference writes as time_step= (src/output.h:755-811) */
  int64_t it_count;     /* iterations, those before a hotstart included */
} tr_sim_status;

/* Tuning of one simulation / batch member (zero fields take the defaults). */
typedef struct tr_tuning {
  int32_t strip_rows;      /* FP64 fused form: rows per wave (0: by grid size) */
  double dry_eps;          /* FP32 form: dry / flush threshold (0: 1e-12, the FP64
                              value; the reference's FP32 build uses 1e-6,
                              src/constants.h:43, which loses mass on
                              small-scale flows) */
  int32_t waves_per_simd;  /* FP32 form: waves per SIMD of the launch (0: 5) */
  int32_t round_q;         /* FP32 form: discharges rounded to FP32 stochastically, as the
                              depth always is (1), or to nearest (0) */
} tr_tuning;

/* Allocate device state for `view` in the given form and upload it. */
int tr_sim_create(const tr_case_view *view, int32_t form, tr_sim **out);
int tr_sim_create_ex(const tr_case_view *view, int32_t form, const tr_tuning *tune,
                     tr_sim **out);
/* Advance until simtime >= t_stop (t_stop <= 0: the case's sim_duration) or
 * `max_steps` more steps (max_steps < 0: unlimited), exactly as the
 * reference's loop would (dt clamped to print times).  Blocking. */
int tr_sim_run(tr_sim *s, double t_stop, int64_t max_steps, tr_sim_status *st);
/* Fused forms: set the limits of the steps queued next (t_stop <= 0: the
 * case's sim_duration; max_steps < 0: unlimited), then enqueue exactly
 * `nsteps` launches on the sim's stream without waiting.  Launches past a
 * limit or past sim_duration are skipped on the device. */
int tr_sim_arm(tr_sim *s, double t_stop, int64_t max_steps);
int tr_sim_enqueue(tr_sim *s, int64_t nsteps);
int tr_sim_sync(tr_sim *s, tr_sim_status *st);
/* Copy the current state to host (padded, FP64). Any pointer may be NULL. */
int tr_sim_download(tr_sim *s, double *h, double *qx, double *qy, double *maxh);
/* h, qx, qy of the current state at `n` flat padded cell indices (host
 * arrays; any output may be NULL): the device gather of
 * src/triton_simulate.h get_observation_data_to_host. */
int tr_sim_sample(tr_sim *s, int32_t n, const int32_t *pos, double *h, double *qx, double *qy);
/* on != 0: tr_sim_run stops after each step that completes an observation
 * interval (status.observed), so the caller can sample it as the reference's
 * loop does; the next run continues. */
int tr_sim_set_obs_halt(tr_sim *s, int32_t on);
/* Fused FP64 forms: rows per wave of the step for the launches queued next
 * (0: by grid size, tr_default_strip_rows); results do not depend on it
 * (the autotuner's knob).  tr_sim_strip_rows: the current value. */
int tr_sim_set_strip_rows(tr_sim *s, int32_t strip_rows);
int32_t tr_sim_strip_rows(const tr_sim *s);
/* Fused FP64 forms: the build of the step kernel for the launches queued
 * next, 3 waves per SIMD or 4 (fewer registers, a few values spilled); 0
 * selects the default by grid size (4 where the three-wave strips would
 * reach 40 rows, about 2700^2 and up; else 3).  Results do not depend on it
 * (the autotuner's other knob).  tr_sim_waves: the current value. */
int tr_sim_set_waves(tr_sim *s, int32_t waves_per_simd);
int32_t tr_sim_waves(const tr_sim *s);
/* The hipStream_t the simulation runs on. */
void *tr_sim_stream(tr_sim *s);
/* Bytes of device memory held. */
int64_t tr_sim_device_bytes(const tr_sim *s);
void tr_sim_destroy(tr_sim *s);

/* One FP64 case on device buffers the caller owns (tr_sim_run_f64_noalloc).
 * Arrays are device pointers, rows x cols with the ghost ring; h, qx, qy
 * are ping-pong pairs: [0] holds the initial state and the state reached is
 * in [steps & 1]. */
typedef struct tr_dev_case_f64 {
  int32_t rows, cols;
  double dx, courant, hextra;
  int32_t open_boundaries;
  int32_t checkpoint_id;         /* print intervals done before sim_start */
  double sim_start, sim_duration, print_interval, init_dt;
  int64_t it_count;              /* iterations before the start */
  double *h[2], *qx[2], *qy[2];
  const double *dem, *nman2;     /* bed (walls 1e6 or copies), squared Manning n */
  double *maxh;                  /* NULL: not tracked */
  int32_t num_runoffs, roff_nrows, roff_row0;
  const int32_t *runoff_id;      /* NULL: zone 0 everywhere */
  const double *roff_time;       /* [roff_nrows] s */
  const double *roff_rate;       /* [num_runoffs*roff_nrows] m/s, zone-major */
} tr_dev_case_f64;

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The run tr_sim_create + tr_sim_run give in
 * TR_FORM_F64_FUSED (same kernels, same bits), to t_stop (<= 0:
 * sim_duration) or max_steps (< 0: unlimited), on the caller's buffers, with
 * no allocation: the control blocks, per-step rain depths, zone areas and
 * volume partial sums live in a static 16 KiB device array per GPU.  At most
 * 64 runoff zones (TR_ERR_INVALID beyond); no point sources, external
 * boundary segments or observation points.  Each call is a whole run from
 * h[0]; blocking on `stream`.  Calls that overlap on one GPU race for its
 * array. */
int tr_sim_run_f64_noalloc(const tr_dev_case_f64 *c, double t_stop, int64_t max_steps,
                           tr_sim_status *st, void *stream);

/* ---- synthetic inputs ------------------------------------------------------ */
/* No counterpart in the reference: drains a synthetic DEM (nr x nc, row-major)
 * by priority-flood depression filling with an `eps` gradient and returns the
 * filled bed and the D8 contributing area in cells; used to carve channel
 * networks into generated benchmark terrain. */
int tr_dem_flow_route(const double *dem, int32_t nr, int32_t nc, double eps, double *filled,
                      double *acc);

/* ---- batched FP32 members: ensembles and serving -------------------------- */
/* Members of uneven size advance together, one launch per step for all of
 * them, each with its own adaptive dt, time and forcing (the reference runs
 * ensemble members as separate MPI jobs, src/Ensify.h:43-72).  Members may be
 * added and removed between launches; tr_batch_add fails with TR_ERR_NOMEM
 * when the member would exceed the HBM budget given at creation. */
typedef struct tr_batch tr_batch;
/* Every launch cuts the active members into row bands, one band-strip per
 * wave, over at most waves_per_simd (0: 5) waves per SIMD of the device. */
int tr_batch_create(int64_t hbm_budget_bytes, int32_t waves_per_simd, tr_batch **out);
/* Device bytes a member built from `v` would hold (admission control). */
int64_t tr_batch_member_bytes(const tr_case_view *v);
int tr_batch_add(tr_batch *b, const tr_case_view *v, int32_t *id);
/* tune->dry_eps applies per member; the other tuning fields are the batch's. */
int tr_batch_add_ex(tr_batch *b, const tr_case_view *v, const tr_tuning *tune, int32_t *id);
/* With on != 0 the member stops after every step that completes a print
 * interval (print_id advances) until the next tr_batch_arm, so its state can
 * be written out at print times as the reference's loop does
 * (src/triton_simulate.h:92-110). */
int tr_batch_set_print_halt(tr_batch *b, int32_t id, int32_t on);
/* Default rows per wave of the FP64 fused step for a padded rows x cols grid,
 * for the default build (enough waves to fill the device; long strips
 * otherwise). */
int32_t tr_default_strip_rows(int32_t rows, int32_t cols);
int tr_batch_remove(tr_batch *b, int32_t id);
/* Limits for the steps queued next, for every member: t_stop (<= 0: each
 * member's sim_duration) and at most max_steps more steps (< 0: unlimited). */
int tr_batch_arm(tr_batch *b, double t_stop, int64_t max_steps);
/* Queue nsteps batched steps without waiting (as captured hipGraphs of
 * graph_steps steps when tr_batch_set_launch enabled them). */
int tr_batch_enqueue(tr_batch *b, int64_t nsteps);
/* Launch shape: waves per SIMD of the row partition (0: 5) and steps per
 * captured hipGraph (0: eager launches).  Results do not depend on either. */
int tr_batch_set_launch(tr_batch *b, int32_t waves_per_simd, int32_t graph_steps);
/* Wait; members past their limits leave the launch.  n_active may be NULL. */
int tr_batch_sync(tr_batch *b, int32_t *n_active);
/* arm + enqueue/sync until no member is active or max_steps were queued. */
int tr_batch_run(tr_batch *b, double t_stop, int64_t max_steps, int32_t *n_active);
int tr_batch_status(tr_batch *b, int32_t id, tr_sim_status *st);
/* Status as of the last tr_batch_sync / tr_batch_arm from the host mirror of
 * the control blocks (no device work; volume is NaN). */
int tr_batch_poll(tr_batch *b, int32_t id, tr_sim_status *st);
int tr_batch_download(tr_batch *b, int32_t id, double *h, double *qx, double *qy, double *maxh);
/* tr_sim_sample and tr_sim_set_obs_halt for one member (the halt holds the
 * member until the next tr_batch_arm). */
int tr_batch_sample(tr_batch *b, int32_t id, int32_t n, const int32_t *pos, double *h, double *qx,
                    double *qy);
int tr_batch_set_obs_halt(tr_batch *b, int32_t id, int32_t on);
int64_t tr_batch_device_bytes(const tr_batch *b);
int64_t tr_batch_member_cells(tr_batch *b, int32_t id);
int32_t tr_batch_work_waves(const tr_batch *b);
/* Rows per band of the current partition: each wave steps one column strip
 * (60 output columns) of one band, so with tr_batch_work_waves this is the
 * launch's tile shape (the band height follows from the waves per SIMD that
 * tr_batch_set_launch or the autotuner chose). */
int32_t tr_batch_band_rows(const tr_batch *b);
void *tr_batch_stream(tr_batch *b);
void tr_batch_destroy(tr_batch *b);

/* ---- serving: continuous batching over one tr_batch ------------------------ */
/* Flood-scenario requests (cases run to their sim_duration in the FP32 form)
 * wait in arrival order; each scheduling round (tr_serve_step) admits waiting
 * requests first fit while their device bytes fit the HBM budget, runs
 * chunk_steps batched launches for every admitted request, and retires the
 * finished ones (final status and, with keep_state, state kept on the host;
 * member removed, its memory freed), so requests join while others are
 * mid-run.  A request's result is bit for bit the result it has alone.  The
 * reference has no serving layer (it runs ensemble members as separate MPI
 * jobs, src/Ensify.h:43-72); this is the request loop of the batched form.
 * Not thread-safe: one host thread drives a tr_serve. */
typedef struct tr_serve tr_serve;
typedef struct tr_serve_options {
  int64_t hbm_budget_bytes; /* device bytes the admitted requests may hold (0: no limit) */
  int32_t chunk_steps;      /* batched launches per scheduling round (>= 1) */
  int32_t waves_per_simd;   /* launch shape of the batch (0: default) */
  int32_t max_members;      /* requests on the device at once (0: 256) */
  int32_t starve_rounds;    /* a waiting request passed over more than this many
                               rounds holds back every later one until it fits */
  int32_t keep_state;       /* keep each finished request's final h, qx, qy */
  int32_t reserved;
} tr_serve_options;

/* FAILED: the request could not be admitted (error holds the TR_* code);
 * CANCELLED: ended by tr_serve_cancel. */
enum { TR_REQ_WAITING = 0, TR_REQ_RUNNING = 1, TR_REQ_FINISHED = 2, TR_REQ_FAILED = 3, TR_REQ_CANCELLED = 4 };
typedef struct tr_serve_request {
  int32_t state;            /* TR_REQ_* */
  int32_t member;           /* batch member while running, else -1 */
  int64_t bytes;            /* device bytes (tr_batch_member_bytes) */
  int64_t cells;            /* interior cells */
  int32_t skipped;          /* admission rounds it was passed over */
  int32_t error;            /* TR_* code of a failed request, else TR_OK */
  double arrival;           /* tr_serve_clock() seconds */
  double admitted;          /* -1 until admitted */
  double done;              /* -1 until finished */
  tr_sim_status status;     /* final status once finished */
} tr_serve_request;

typedef struct tr_serve_stats {
  int64_t submitted, completed, waiting, running;
  int64_t cell_steps;       /* interior cells x steps of the retired requests */
  int64_t rounds;           /* rounds that ran a chunk */
  int64_t peak_bytes;       /* most device bytes held at once */
  int32_t max_members;      /* most requests on the device at once */
  int32_t reserved;
  double busy_s;            /* wall time of the rounds that ran a chunk */
  double p50_latency_s;     /* submit-to-finish latency over retired requests */
  double p99_latency_s;     /* (linear interpolation; NaN while none) */
  double max_latency_s;
  int64_t failed;           /* requests that could not be admitted */
  int64_t cancelled;        /* requests ended by tr_serve_cancel */
} tr_serve_stats;

/* Per-request products at print times (the reference writes them in its
 * loop, src/triton_simulate.h:92-110, src/output.h:512-709): at every
 * print_interval boundary the request's H, QX, QY rasters (and MH, the
 * running maximum depth, when the view tracks maxh) are written into
 * output_dir as output::write_output names them (<dir>/bin/H_<id>_00.out,
 * and with output_format 0 their ASCII twins in <dir>/asc/), and/or kept on
 * the host (keep_prints).  The request's member stops at each print until
 * the next scheduling round. */
typedef struct tr_serve_request_options {
  const char *output_dir;   /* NULL: write nothing */
  int32_t output_format;    /* 0: ASC (ASCII + binary twin), 1: BIN only */
  int32_t keep_prints;      /* keep each print's state for tr_serve_print */
} tr_serve_request_options;

/* The server's clock (seconds, monotonic): arrival, admitted and done times. */
double tr_serve_clock(void);
int tr_serve_create(const tr_serve_options *opt, tr_serve **out);
/* Queue a request arriving at `arrival` (tr_serve_clock seconds; < 0: now).
 * The view's arrays must stay valid until the request is admitted
 * (tr_serve_poll state != TR_REQ_WAITING).  The view is checked as
 * tr_batch_add checks it (TR_ERR_INVALID); TR_ERR_NOMEM when it needs more
 * than the whole budget.  A request whose admission still fails is ended
 * with state TR_REQ_FAILED and the others carry on. */
int tr_serve_submit(tr_serve *s, const tr_case_view *v, double arrival, int32_t *request_id);
/* tr_serve_submit with per-request print products (opt may be NULL). */
int tr_serve_submit_ex(tr_serve *s, const tr_case_view *v, double arrival,
                       const tr_serve_request_options *opt, int32_t *request_id);
/* Prints kept for a request (keep_prints) and the k-th of them: its print
 * id, simulated time and state (rows x cols each; maxh may be NULL and is
 * left untouched when the view does not track it). */
int tr_serve_print_count(const tr_serve *s, int32_t request_id, int32_t *n);
int tr_serve_print(const tr_serve *s, int32_t request_id, int32_t k, int32_t *print_id, double *simtime,
                   double *h, double *qx, double *qy, double *maxh);
/* One scheduling round: admit, run one chunk, retire.  *ran = 0 when no
 * request was on the device (nothing ran). */
int tr_serve_step(tr_serve *s, int32_t *ran);
/* Rounds until every submitted request has finished (TR_ERR_STATE if a
 * waiting request can never be admitted). */
int tr_serve_drain(tr_serve *s);
int tr_serve_poll(const tr_serve *s, int32_t request_id, tr_serve_request *out);
/* Final state of a finished request (keep_state), rows x cols each. */
int tr_serve_result(const tr_serve *s, int32_t request_id, double *h, double *qx, double *qy);
/* Forget an ended (finished, failed or cancelled) request and its kept state. */
int tr_serve_release(tr_serve *s, int32_t request_id);
/* End a waiting or running request now (its device memory is freed; the
 * status of a running one is that of the last round). */
int tr_serve_cancel(tr_serve *s, int32_t request_id);
int tr_serve_stats_get(const tr_serve *s, tr_serve_stats *out);
void tr_serve_destroy(tr_serve *s);
/* The admission policy of a round, on its own (host only): walking the
 * waiting requests in arrival order, request k is admitted (admit[k] = 1)
 * when held + bytes[k] fits the budget and fewer than max_members run; a
 * request that does not fit has skipped[k] incremented, and once some
 * skipped[k] exceeds starve_rounds no later request is admitted this round. */
int tr_serve_admit(const int64_t *bytes, int32_t *skipped, int32_t n_waiting, int64_t held, int64_t budget,
                   int32_t running, int32_t max_members, int32_t starve_rounds, uint8_t *admit);

#ifdef __cplusplus
}
#endif
#endif /* TRITON_AMD_H */

// sim.hip -- device-resident simulation objects behind tr_sim_*.
//
// TR_FORM_F64_UNFUSED replays Triton::triton<T>::simulate()
// (src/triton_simulate.h:54-159) with the per-kernel drop-in entry points and
// the same host-side dt / print / forcing logic, reading dt back every step as
// the reference does.  TR_FORM_F64_FUSED (and TR_FORM_F64_R32, the same step
// with its state rounded to FP32) and TR_FORM_F32_FUSED launch one
// fused kernel per step and keep dt, time and forcing indices on the device;
// the host only looks at the control block between batches of launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "common.hpp"
#include "step.hpp"

namespace tr {
int min_dt_to_slot(int size, const double* dtarr, unsigned long long* slot, hipStream_t s);
int copy_obs_f64(int size, const double* arr, double* obs, const int* index, hipStream_t s);
int launch_prime_f64(const StepParams<double>& p, hipStream_t s);
int launch_step_f64(const StepParams<double>& p, hipStream_t s);
int fused_grid_x(int cols);
int default_waves_per_simd(int rows, int cols);
}  // namespace tr

using namespace tr;

struct tr_sim {
  int form = 0;
  tr_case_view v{};               // scalars; array pointers are not kept
  std::vector<double> roff_time, hyg_time;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  int64_t bytes = 0;
  // state (FP64); two buffers for the fused form
  double* h[2] = {nullptr, nullptr};
  double* qx[2] = {nullptr, nullptr};
  double* qy[2] = {nullptr, nullptr};
  double *dem = nullptr, *nman2 = nullptr, *maxh = nullptr;
  int32_t* rid = nullptr;
  double *rrate = nullptr, *rtime = nullptr, *zone_area = nullptr;
  double *htime = nullptr, *hval = nullptr;
  int32_t* srcpos = nullptr;
  int32_t *src_row = nullptr, *src_col = nullptr, *src_id = nullptr;
  double* src_add = nullptr;
  int32_t *bc_index = nullptr, *bc_type = nullptr, *bc_start = nullptr, *bc_nrows = nullptr;
  double *bc_v1 = nullptr, *bc_v2 = nullptr;
  int32_t* bc_map = nullptr;                 // fused forms: edge cell -> first entry
  int32_t* bc_next = nullptr;                // fused forms: next entry of the same cell
  double *bc_zc = nullptr, *bc_n2c = nullptr;
  bool friction = false;
  // unfused scratch
  double *sqrth = nullptr, *dtarr = nullptr, *rhs[6] = {};
  double* dmin = nullptr;
  double* h_pinned = nullptr;
  // unfused host loop state (src/triton_simulate.h:41-53)
  double simtime = 0, last_dt = 0;
  int print_id = 0, roff_row = 0, idx_low = 0;
  int64_t steps = 0;
  double vol_in = 0;
  double last_obs = 0, dt_acc = 0;
  int observed = 0;
  // observation sampling: halt after observation steps; gather scratch
  int halt_on_obs = 0;
  int32_t* d_obs_idx = nullptr;
  double* d_obs_val = nullptr;
  int obs_cap = 0;
  // fused control
  StepCtrl* ctrl = nullptr;
  StepCtrl* ctrl_host = nullptr;  // pinned mirror
  int64_t launched = 0;            // == completed steps between API calls
  double t_stop = INFINITY;        // limits armed for queued steps
  int64_t step_stop = INT64_MAX;
  int strip_rows = 64;
  int waves4 = 0;                  // FP64 fused step: the four-waves-per-SIMD build
  // volume reduction
  double* vpart = nullptr;
  double* vsum = nullptr;
  // FP32 form: a batch of one member
  tr_batch* batch = nullptr;
  int32_t member = -1;

  ~tr_sim() {
    if (batch) tr_batch_destroy(batch);
    if (stream) hipStreamSynchronize(stream);
    for (void* p : allocs) hipFree(p);
    if (d_obs_idx) hipFree(d_obs_idx);
    if (d_obs_val) hipFree(d_obs_val);
    if (ctrl_host) hipHostFree(ctrl_host);
    if (h_pinned) hipHostFree(h_pinned);
    if (stream) hipStreamDestroy(stream);
  }
  template <class T>
  int alloc(T*& p, size_t n) {
    p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
      return TR_ERR_NOMEM;
    }
    allocs.push_back(p);
    bytes += int64_t(n * sizeof(T));
    return TR_OK;
  }
  template <class T>
  int upload(T*& p, const T* src, size_t n) {
    int rc = alloc(p, n);
    if (rc) return rc;
    if (n && src) TR_HIP(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, stream));
    return TR_OK;
  }
};

namespace {

#define RC(x)            \
  do {                   \
    int rc_ = (x);       \
    if (rc_) return rc_; \
  } while (0)

__global__ void k_volume_part(const double* __restrict__ h, int rows, int cols, double area,
                              double* __restrict__ part) {
  __shared__ double sm[256];
  const int n_int = (rows - 2) * (cols - 2);
  const int per = (n_int + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(lo + per, n_int);
  double acc = 0.0;
  for (int k = lo + int(threadIdx.x); k < hi; k += blockDim.x) {
    int r = k / (cols - 2) + 1, c = k % (cols - 2) + 1;
    acc += h[(long long)r * cols + c];
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0] * area;
}

__global__ void k_volume_sum(const double* __restrict__ part, int n, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; i++) s += part[i];
  *out = s;
}

constexpr int kVolBlocks = 512;

// Deterministic FP64 volume of the interior (fixed partition, fixed order).
int volume_f64(tr_sim* s, const double* h, double* out) {
  hipLaunchKernelGGL(k_volume_part, dim3(kVolBlocks), dim3(256), 0, s->stream, h, s->v.rows,
                     s->v.cols, s->v.dx * s->v.dx, s->vpart);
  hipLaunchKernelGGL(k_volume_sum, dim3(1), dim3(64), 0, s->stream, s->vpart, kVolBlocks,
                     s->vsum);
  RC(launch_check("volume"));
  TR_HIP(hipMemcpyAsync(s->h_pinned, s->vsum, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  TR_HIP(hipStreamSynchronize(s->stream));
  *out = *s->h_pinned;
  return TR_OK;
}

int check_view(const tr_case_view* v) {
  if (!v || v->rows < 3 || v->cols < 3 || !v->h || !v->qx || !v->qy || !v->dem || !v->nman2) {
    set_error("tr_sim_create: view needs rows,cols >= 3 and h, qx, qy, dem, nman2");
    return TR_ERR_INVALID;
  }
  if (int64_t(v->rows) * v->cols > int64_t(INT32_MAX)) {
    set_error("tr_sim_create: grid exceeds 2^31 cells");
    return TR_ERR_INVALID;
  }
  if (!(v->dx > 0) || !(v->courant > 0) || !(v->print_interval > 0)) {
    set_error("tr_sim_create: dx, courant and print_interval must be positive");
    return TR_ERR_INVALID;
  }
  if (v->num_runoffs > 0 && (v->roff_nrows < 2 || !v->roff_time || !v->roff_rate)) {
    set_error("tr_sim_create: runoff needs at least two rows");
    return TR_ERR_INVALID;
  }
  if (v->num_src > 0 && (v->hyg_nrows < 2 || !v->hyg_time || !v->hyg_val || !v->src_pos)) {
    set_error("tr_sim_create: sources need a hydrograph with at least two rows");
    return TR_ERR_INVALID;
  }
  const int64_t n = int64_t(v->rows) * v->cols;
  for (int i = 0; i < v->num_src; i++) {
    int r = v->src_pos[i] / v->cols, c = v->src_pos[i] % v->cols;
    if (v->src_pos[i] < 0 || v->src_pos[i] >= n || r < 1 || r > v->rows - 2 || c < 1 || c > v->cols - 2) {
      set_error("tr_sim_create: source cell outside the interior");
      return TR_ERR_INVALID;
    }
  }
  if (v->num_runoffs > 0 && v->runoff_id) {
    for (int r = 1; r < v->rows - 1; r++)
      for (int c = 1; c < v->cols - 1; c++) {
        int z = v->runoff_id[int64_t(r) * v->cols + c];
        if (z < 0 || z >= v->num_runoffs) {
          set_error("tr_sim_create: runoff zone id out of range");
          return TR_ERR_INVALID;
        }
      }
  }
  return TR_OK;
}

// ---- unfused form: the reference's loop, kernel by kernel ----------------
int step_unfused(tr_sim* s) {
  const tr_case_view& v = s->v;
  const int rows = v.rows, cols = v.cols, n = rows * cols;
  hipStream_t st = s->stream;
  void* sp = st;
  double* h = s->h[0];
  double* qx = s->qx[0];
  double* qy = s->qy[0];
  // compute_local_dt: per-cell dt + blocking min (src/triton_simulate.h:191-198)
  RC(tr_compute_dt_and_sqrt_f64(n, v.dx, qx, qy, h, s->sqrth, s->dtarr, v.courant, v.hextra, sp));
  double dt = 0;
  RC(tr_find_min_dt_f64(n, s->dtarr, s->dmin, s->h_pinned, sp));
  dt = *s->h_pinned;
  // compute_global_dt (:202-226)
  if (dt >= kMaxDt - 1.0) dt = v.init_dt;
  if (s->simtime + dt > v.print_interval * (s->print_id + 1))
    dt = v.print_interval * (s->print_id + 1) - s->simtime;
  // compute_new_state (:230-346)
  RC(tr_flux_x_f64(n, rows, cols, v.dx, dt, h, qx, qy, s->dem, s->sqrth, s->rhs[0], s->rhs[1],
                   s->rhs[2], s->rhs[3], s->rhs[4], s->rhs[5], v.hextra, sp));
  RC(tr_flux_y_f64(n, rows, cols, v.dx, dt, h, qx, qy, s->dem, s->sqrth, s->rhs[0], s->rhs[1],
                   s->rhs[2], s->rhs[3], s->rhs[4], s->rhs[5], v.hextra, sp));
  RC(tr_update_cells_f64(n, rows, cols, dt, h, qx, qy, s->dem, s->nman2, s->rhs[0], s->rhs[1],
                         s->rhs[2], s->rhs[3], s->rhs[4], s->rhs[5], v.hextra, sp));
  double step_vol = 0.0;
  if (v.num_runoffs > 0) {
    int last = v.roff_nrows - 1;
    if (s->simtime > s->roff_time[size_t(last)]) s->roff_row = last;
    else if (s->simtime > s->roff_time[size_t(s->roff_row + 1)]) s->roff_row++;
    RC(tr_update_runoff_f64(n, rows, cols, dt, s->rid, s->roff_row, v.roff_nrows, s->rrate, h, qx,
                            qy, v.hextra, sp));
  }
  if (v.num_src > 0) {
    int hi = v.hyg_nrows - 1;
    bool found = false;
    for (int i = s->idx_low; i < hi; i++) {
      if (s->hyg_time[size_t(i + 1)] > s->simtime) { s->idx_low = i; hi = i + 1; found = true; break; }
    }
    if (!found) s->idx_low = hi;
    RC(tr_compute_flow_f64(v.num_src, s->htime, s->hval, v.dx, dt, s->simtime, s->idx_low, hi, h,
                           s->srcpos, sp));
  }
  if (v.open_boundaries) {
    RC(tr_open_bc_west_east_f64(2 * rows, rows, cols, h, qx, qy, sp));
    RC(tr_open_bc_north_south_f64(cols, rows, cols, h, qx, qy, 0, 1, sp));
  }
  if (v.num_extbc_cells > 0) {
    RC(tr_compute_extbc_values_f64(v.num_extbc_cells, rows, cols, dt, h, qx, qy, s->dem, s->nman2,
                                   s->bc_index, s->bc_type, s->bc_start, s->bc_nrows, s->bc_v1,
                                   s->bc_v2, s->simtime, 0, 1, sp));
  }
  RC(tr_wet_dry_f64(n, rows, cols, dt, h, qx, qy, s->dem, s->maxh, v.hextra, 1, sp));
  (void)step_vol;
  s->simtime += dt;
  s->last_dt = dt;
  s->steps++;
  // the dt sum and the observation test of src/triton_simulate.h:66-81
  s->dt_acc += dt;
  if (v.it_print > 0 && (int64_t(v.it_count) + s->steps) % v.it_print == 0) s->dt_acc = 0.0;
  s->observed = 0;
  if (v.time_series_flag && s->simtime - s->last_obs >= v.print_observation) {
    s->last_obs += v.print_observation;
    s->observed = 1;
  }
  if (s->simtime >= v.print_interval * (s->print_id + 1)) s->print_id++;
  return TR_OK;
}

// ---- fused FP64 form ------------------------------------------------------
StepParams<double> fused_params(tr_sim* s, int64_t launch_index) {
  const tr_case_view& v = s->v;
  StepParams<double> p{};
  p.rows = v.rows;
  p.cols = v.cols;
  p.strip_rows = s->strip_rows;
  p.waves4 = s->waves4;
  p.nbx = fused_grid_x(v.cols);
  p.nby = (v.rows - 2 + s->strip_rows - 1) / s->strip_rows;
  p.nblocks = unsigned(p.nbx) * unsigned(p.nby);
  p.dx = v.dx;
  p.hextra = v.hextra;
  p.cn_dx = v.courant * v.dx;
  const int a = int(launch_index & 1);
  p.h_in = s->h[a];
  p.qx_in = s->qx[a];
  p.qy_in = s->qy[a];
  p.h_out = s->h[a ^ 1];
  p.qx_out = s->qx[a ^ 1];
  p.qy_out = s->qy[a ^ 1];
  p.dem = s->dem;
  p.nman2 = s->friction ? s->nman2 : nullptr;
  p.maxh = s->maxh;
  p.runoff_id = s->rid;
  p.roff_rate = s->rrate;
  p.roff_time = s->rtime;
  p.roff_nrows = v.roff_nrows;
  p.num_runoffs = v.num_runoffs;
  p.zone_area = s->zone_area;
  p.nsrc = v.num_src;
  p.src_row = s->src_row;
  p.src_col = s->src_col;
  p.src_id = s->src_id;
  p.hyg_time = s->htime;
  p.hyg_val = s->hval;
  p.hyg_nrows = v.hyg_nrows;
  p.src_add = s->src_add;
  p.open_bc = v.open_boundaries;
  p.ctrl = s->ctrl;
  p.cur = a;
  p.sim_duration = v.sim_duration;
  p.print_interval = v.print_interval;
  p.init_dt = v.init_dt;
  p.t_stop = INFINITY;
  p.step_stop = INT64_MAX;
  p.halt_on_print = 0;
  p.time_series = v.time_series_flag;
  p.print_observation = v.print_observation;
  p.halt_on_obs = s->halt_on_obs;
  p.it_print = v.it_print;
  p.it_count0 = v.it_count;
  p.round32 = s->form == TR_FORM_F64_R32 ? 1 : 0;
  p.bc = ExtBC{};
  if (v.num_extbc_cells > 0) {
    p.bc.n = v.num_extbc_cells;
    p.bc.map = s->bc_map;
    p.bc.next = s->bc_next;
    p.bc.type = s->bc_type;
    p.bc.start = s->bc_start;
    p.bc.nvars = s->bc_nrows;
    p.bc.v1 = s->bc_v1;
    p.bc.v2 = s->bc_v2;
    p.bc.zc = s->bc_zc;
    p.bc.n2c = s->bc_n2c;
  }
  return p;
}

int read_ctrl(tr_sim* s, int slot, StepCtrl& out) {
  TR_HIP(hipMemcpyAsync(s->ctrl_host, s->ctrl + slot, sizeof(StepCtrl), hipMemcpyDeviceToHost,
                        s->stream));
  TR_HIP(hipStreamSynchronize(s->stream));
  out = *s->ctrl_host;
  return TR_OK;
}

int write_ctrl(tr_sim* s, int slot, const StepCtrl& c) {
  *s->ctrl_host = c;
  TR_HIP(hipMemcpyAsync(s->ctrl + slot, s->ctrl_host, sizeof(StepCtrl), hipMemcpyHostToDevice,
                        s->stream));
  TR_HIP(hipStreamSynchronize(s->stream));
  return TR_OK;
}

int prime_fused(tr_sim* s) {
  const tr_case_view& v = s->v;
  StepCtrl c{};
  c.simtime = v.sim_start;
  c.dt = 0.0;
  reset_dtmin(c);
  c.halt = 0;
  c.step = 0;
  c.print_id = v.checkpoint_id;
  if (c.print_id == 0 && c.simtime > 0.0) c.print_id = int(c.simtime / v.print_interval);
  c.roff_row = v.roff_row0;
  c.idx_low = 0;
  c.idx_high = 0;
  c.vol_in = 0.0;
  c.step_vol_in = 0.0;
  c.last_obs = v.sim_start;  // src/triton_simulate.h:47
  c.dt_acc = 0.0;
  c.observed = 0;
  RC(write_ctrl(s, 1, c));
  StepParams<double> p = fused_params(s, 0);
  p.cur = 1;  // the prime reduces into slot 1 and prepares slot 0 for step 0
  RC(launch_prime_f64(p, s->stream));
  TR_HIP(hipStreamSynchronize(s->stream));
  s->launched = 0;
  return TR_OK;
}

// Queue `n` launches without waiting; launches past a halt are no-ops.
int enqueue_fused(tr_sim* s, int64_t n, double t_stop, int64_t step_stop) {
  for (int64_t i = 0; i < n; i++) {
    StepParams<double> p = fused_params(s, s->launched + i);
    p.t_stop = t_stop;
    p.step_stop = step_stop;
    RC(launch_step_f64(p, s->stream));
  }
  s->launched += n;
  return TR_OK;
}

// After launches: realign the launch counter with the steps really taken, so
// the state sits in buffer (steps & 1) and its control block in slot (steps & 1).
int realign_fused(tr_sim* s, StepCtrl& c) {
  const int slot = int(s->launched & 1);
  RC(read_ctrl(s, slot, c));
  if (c.step != s->launched) {
    const int want = int(c.step & 1);
    if (want != slot) {
      RC(write_ctrl(s, want, c));
      if (s->v.num_src > 0)
        TR_HIP(hipMemcpyAsync(s->src_add + want * s->v.num_src, s->src_add + slot * s->v.num_src,
                              sizeof(double) * size_t(s->v.num_src), hipMemcpyDeviceToDevice,
                              s->stream));
      TR_HIP(hipStreamSynchronize(s->stream));
    }
    s->launched = c.step;
  }
  return TR_OK;
}

// Realign, then set the halt flag of the next step for the given limits.
int arm_fused(tr_sim* s, double t_stop, int64_t max_steps, bool& halted) {
  StepCtrl c;
  RC(realign_fused(s, c));
  s->t_stop = t_stop;
  s->step_stop = max_steps < 0 ? INT64_MAX : c.step + max_steps;
  halted = c.finished || !(c.simtime < t_stop) || c.step >= s->step_stop;
  c.halt = halted ? 1 : 0;
  RC(write_ctrl(s, int(s->launched & 1), c));
  return TR_OK;
}

int run_fused(tr_sim* s, double t_stop, int64_t max_steps) {
  bool halted = false;
  RC(arm_fused(s, t_stop, max_steps, halted));
  if (halted) return TR_OK;
  StepCtrl c;
  int64_t batch = 32;
  for (;;) {
    int64_t n = batch;
    if (max_steps >= 0) n = std::min<int64_t>(n, s->step_stop - s->launched);
    if (n <= 0) n = 1;
    RC(enqueue_fused(s, n, s->t_stop, s->step_stop));
    RC(realign_fused(s, c));
    if (c.halt) break;
    batch = std::min<int64_t>(batch * 2, 1024);
  }
  return TR_OK;
}

int fill_status(tr_sim* s, tr_sim_status* st) {
  if (!st) return TR_OK;
  std::memset(st, 0, sizeof(*st));
  const double* h = s->h[0];
  if (s->form == TR_FORM_F64_UNFUSED) {
    st->simtime = s->simtime;
    st->dt = s->last_dt;
    st->steps = s->steps;
    st->print_id = s->print_id;
    st->finished = !(s->simtime < s->v.sim_duration);
    st->volume_in = s->vol_in;
    st->observed = s->observed;
    st->last_obs = s->last_obs;
    st->average_dt = s->dt_acc;
  } else {
    StepCtrl c;
    RC(realign_fused(s, c));
    // the control block of the next step holds the time reached and the
    // dt that was applied last is the previous block's; report both
    st->simtime = c.simtime;
    st->steps = c.step;
    st->print_id = c.print_id;
    st->finished = c.finished;
    st->volume_in = c.vol_in;
    st->dt = c.dt;
    st->observed = c.observed;
    st->last_obs = c.last_obs;
    st->average_dt = c.dt_acc;
    h = s->h[c.step & 1];
  }
  st->it_count = int64_t(s->v.it_count) + st->steps;
  return volume_f64(s, h, &st->volume);
}

}  // namespace

extern "C" {

int tr_sim_create(const tr_case_view* view, int32_t form, tr_sim** out) {
  return tr_sim_create_ex(view, form, nullptr, out);
}

int tr_sim_create_ex(const tr_case_view* view, int32_t form, const tr_tuning* tune,
                     tr_sim** out) {
  if (!out) {
    set_error("tr_sim_create: out is NULL");
    return TR_ERR_INVALID;
  }
  *out = nullptr;
  RC(check_view(view));
  if (form == TR_FORM_F32_FUSED) {
    auto* s = new tr_sim();
    s->form = form;
    s->v = *view;
    // a batch of one: its rows are split over the whole device
    tr_tuning t{};
    if (tune) t = *tune;
    if (t.waves_per_simd < 0 || t.waves_per_simd > 16) {
      delete s;
      set_error("tr_sim_create_ex: waves_per_simd must be in [0, 16]");
      return TR_ERR_INVALID;
    }
    int rc = tr_batch_create(0, t.waves_per_simd, &s->batch);
    if (!rc) rc = tr_batch_add_ex(s->batch, view, &t, &s->member);
    if (rc) {
      delete s;
      return rc;
    }
    *out = s;
    return TR_OK;
  }
  if (form != TR_FORM_F64_UNFUSED && form != TR_FORM_F64_FUSED && form != TR_FORM_F64_R32) {
    set_error("tr_sim_create: unknown form");
    return TR_ERR_INVALID;
  }
  ExtBCHost bch;
  if (form != TR_FORM_F64_UNFUSED) {
    const std::string err = build_extbc(view->rows, view->cols, view->num_extbc_cells,
                                        view->bc_index, view->bc_type, view->bc_start,
                                        view->bc_nrows, view->dem, view->nman2, bch);
    if (!err.empty()) {
      set_error("tr_sim_create: " + err);
      return TR_ERR_INVALID;
    }
  }
  if (form != TR_FORM_F64_UNFUSED && view->num_src > kMaxSrcFused) {
    set_error("tr_sim_create: too many point sources for the fused step");
    return TR_ERR_INVALID;
  }
  if (tune && (tune->strip_rows < 0 || tune->strip_rows > 4096)) {
    set_error("tr_sim_create_ex: strip_rows must be in [0, 4096]");
    return TR_ERR_INVALID;
  }
  auto* s = new tr_sim();
  s->form = form;
  s->v = *view;
  const tr_case_view& v = *view;
  s->strip_rows = (tune && tune->strip_rows > 0) ? tune->strip_rows
                                                 : tr_default_strip_rows(v.rows, v.cols);
  s->waves4 = v.rows >= 3 && v.cols >= 3 && default_waves_per_simd(v.rows, v.cols) == 4 ? 1 : 0;
  const size_t n = size_t(v.rows) * size_t(v.cols);
  auto fail = [&](int rc) {
    delete s;
    return rc;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    s->stream = nullptr;
    set_error("hipStreamCreate failed");
    return fail(TR_ERR_HIP);
  }
  if (hipHostMalloc(reinterpret_cast<void**>(&s->ctrl_host), sizeof(StepCtrl)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&s->h_pinned), 64) != hipSuccess) {
    set_error("hipHostMalloc failed");
    return fail(TR_ERR_NOMEM);
  }
  const int nbuf = form == TR_FORM_F64_UNFUSED ? 1 : 2;
  int rc = TR_OK;
  // the FP32-storage yardstick starts from the state the FP32 form stores
  std::vector<double> r32[3];
  const double* init[3] = {v.h, v.qx, v.qy};
  if (form == TR_FORM_F64_R32) {
    for (int k = 0; k < 3; k++) {
      r32[k].resize(n);
      for (size_t i = 0; i < n; i++) r32[k][i] = double(float(init[k][i]));
      init[k] = r32[k].data();
    }
  }
  for (int b = 0; b < nbuf && !rc; b++) {
    rc = s->upload(s->h[b], init[0], n);
    if (!rc) rc = s->upload(s->qx[b], init[1], n);
    if (!rc) rc = s->upload(s->qy[b], init[2], n);
  }
  if (!rc) rc = s->upload(s->dem, v.dem, n);
  if (!rc) rc = s->upload(s->nman2, v.nman2, n);
  if (!rc && v.maxh) rc = s->upload(s->maxh, v.maxh, n);  // NULL view.maxh: not tracked
  for (size_t i = 0; i < n && !s->friction; i++) s->friction = v.nman2[i] > kEps12;
  // forcing tables
  if (!rc && v.num_runoffs > 0) {
    s->roff_time.assign(v.roff_time, v.roff_time + v.roff_nrows);
    rc = s->upload(s->rrate, v.roff_rate, size_t(v.num_runoffs) * v.roff_nrows);
    if (!rc) rc = s->upload(s->rtime, v.roff_time, size_t(v.roff_nrows));
    std::vector<double> area(size_t(v.num_runoffs), 0.0);
    bool uniform = true;
    for (int r = 1; r < v.rows - 1 && v.runoff_id; r++)
      for (int c = 1; c < v.cols - 1; c++) {
        int z = v.runoff_id[size_t(r) * v.cols + c];
        area[size_t(z)] += 1.0;
        uniform = uniform && z == 0;
      }
    if (!v.runoff_id) area[0] = double(v.rows - 2) * double(v.cols - 2);
    for (auto& a : area) a *= v.dx * v.dx;
    if (!rc) rc = s->upload(s->zone_area, area.data(), area.size());
    if (!rc && v.runoff_id && (form == TR_FORM_F64_UNFUSED || !uniform))
      rc = s->upload(s->rid, v.runoff_id, n);
    if (!rc && form == TR_FORM_F64_UNFUSED && !v.runoff_id) {
      std::vector<int32_t> zeros(n, 0);
      rc = s->upload(s->rid, zeros.data(), n);
    }
  }
  if (!rc && v.num_src > 0) {
    s->hyg_time.assign(v.hyg_time, v.hyg_time + v.hyg_nrows);
    rc = s->upload(s->htime, v.hyg_time, size_t(v.hyg_nrows));
    if (!rc) rc = s->upload(s->hval, v.hyg_val, size_t(v.hyg_nrows) * v.num_src);
    if (!rc) rc = s->upload(s->srcpos, v.src_pos, size_t(v.num_src));
    // sources sorted by (row, col), stable so one cell's sources keep list order
    std::vector<int32_t> order(size_t(v.num_src));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return v.src_pos[a] < v.src_pos[b]; });
    std::vector<int32_t> rr, cc;
    for (int k : order) {
      rr.push_back(v.src_pos[k] / v.cols);
      cc.push_back(v.src_pos[k] % v.cols);
    }
    if (!rc) rc = s->upload(s->src_row, rr.data(), rr.size());
    if (!rc) rc = s->upload(s->src_col, cc.data(), cc.size());
    if (!rc) rc = s->upload(s->src_id, order.data(), order.size());
    if (!rc) rc = s->alloc(s->src_add, size_t(2 * v.num_src));
  }
  if (!rc && v.num_extbc_cells > 0) {
    const size_t k = size_t(v.num_extbc_cells);
    rc = s->upload(s->bc_index, v.bc_index, k);
    if (!rc) rc = s->upload(s->bc_type, v.bc_type, k);
    if (!rc) rc = s->upload(s->bc_start, v.bc_start, k);
    if (!rc) rc = s->upload(s->bc_nrows, v.bc_nrows, k);
    if (!rc) rc = s->upload(s->bc_v1, v.bc_var1, size_t(std::max(v.bc_nvars, 1)));
    if (!rc) rc = s->upload(s->bc_v2, v.bc_var2, size_t(std::max(v.bc_nvars, 1)));
    if (!rc && form != TR_FORM_F64_UNFUSED) {  // the fused pass's edge map
      rc = s->upload(s->bc_map, bch.map.data(), bch.map.size());
      if (!rc) rc = s->upload(s->bc_next, bch.next.data(), bch.next.size());
      if (!rc) rc = s->upload(s->bc_zc, bch.zc.data(), bch.zc.size());
      if (!rc) rc = s->upload(s->bc_n2c, bch.n2c.data(), bch.n2c.size());
    }
  }
  if (!rc) rc = s->alloc(s->vpart, kVolBlocks);
  if (!rc) rc = s->alloc(s->vsum, 1);
  if (!rc && form == TR_FORM_F64_UNFUSED) {
    rc = s->alloc(s->sqrth, n);
    if (!rc) rc = s->alloc(s->dtarr, n);
    for (int i = 0; i < 6 && !rc; i++) {
      rc = s->alloc(s->rhs[i], n);
      if (!rc) {
        hipError_t e = hipMemsetAsync(s->rhs[i], 0, n * sizeof(double), s->stream);
        if (e != hipSuccess) rc = hip_fail(e, "hipMemsetAsync");
      }
    }
    if (!rc) rc = s->alloc(s->dmin, 1);
    s->simtime = v.sim_start;
    s->print_id = v.checkpoint_id;
    if (s->print_id == 0 && s->simtime > 0.0) s->print_id = int(s->simtime / v.print_interval);
    s->roff_row = v.roff_row0;
    s->last_obs = v.sim_start;  // src/triton_simulate.h:47
  }
  if (!rc && form != TR_FORM_F64_UNFUSED) {
    rc = s->alloc(s->ctrl, 2);
    if (!rc) rc = prime_fused(s);
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) rc = hip_fail(e, "upload");
  }
  // drop view pointers: the sim owns device copies only
  s->v.h = s->v.qx = s->v.qy = s->v.dem = s->v.nman2 = s->v.maxh = nullptr;
  s->v.runoff_id = nullptr; s->v.roff_time = s->v.roff_rate = nullptr;
  s->v.hyg_time = s->v.hyg_val = nullptr; s->v.src_pos = nullptr;
  s->v.bc_index = s->v.bc_type = s->v.bc_start = s->v.bc_nrows = nullptr;
  s->v.bc_var1 = s->v.bc_var2 = nullptr;
  if (rc) return fail(rc);
  *out = s;
  return TR_OK;
}

int tr_sim_run(tr_sim* s, double t_stop, int64_t max_steps, tr_sim_status* st) {
  if (!s) {
    set_error("tr_sim_run: NULL sim");
    return TR_ERR_INVALID;
  }
  const double stop = t_stop > 0 ? std::min(t_stop, s->v.sim_duration) : s->v.sim_duration;
  if (s->batch) {
    RC(tr_batch_run(s->batch, stop, max_steps, nullptr));
    return st ? tr_batch_status(s->batch, s->member, st) : TR_OK;
  }
  if (s->form == TR_FORM_F64_UNFUSED) {
    int64_t done = 0;
    while (s->simtime < stop && (max_steps < 0 || done < max_steps)) {
      RC(step_unfused(s));
      done++;
      if (s->halt_on_obs && s->observed) break;
    }
  } else {
    RC(run_fused(s, stop, max_steps));
  }
  return fill_status(s, st);
}

int tr_sim_arm(tr_sim* s, double t_stop, int64_t max_steps) {
  if (!s) {
    set_error("tr_sim_arm: NULL sim");
    return TR_ERR_INVALID;
  }
  if (s->form == TR_FORM_F64_UNFUSED) {
    set_error("tr_sim_arm: only fused forms queue steps");
    return TR_ERR_STATE;
  }
  const double stop = t_stop > 0 ? std::min(t_stop, s->v.sim_duration) : s->v.sim_duration;
  if (s->batch) return tr_batch_arm(s->batch, stop, max_steps);
  bool halted = false;
  return arm_fused(s, stop, max_steps, halted);
}

int tr_sim_enqueue(tr_sim* s, int64_t nsteps) {
  if (!s || nsteps < 0) {
    set_error("tr_sim_enqueue: bad arguments");
    return TR_ERR_INVALID;
  }
  if (s->form == TR_FORM_F64_UNFUSED) {
    set_error("tr_sim_enqueue: only fused forms queue steps");
    return TR_ERR_STATE;
  }
  if (s->batch) return tr_batch_enqueue(s->batch, nsteps);
  return enqueue_fused(s, nsteps, s->t_stop, s->step_stop);
}

int tr_sim_sync(tr_sim* s, tr_sim_status* st) {
  if (!s) {
    set_error("tr_sim_sync: NULL sim");
    return TR_ERR_INVALID;
  }
  if (s->batch) {
    RC(tr_batch_sync(s->batch, nullptr));
    return st ? tr_batch_status(s->batch, s->member, st) : TR_OK;
  }
  TR_HIP(hipStreamSynchronize(s->stream));
  return fill_status(s, st);
}

int tr_sim_download(tr_sim* s, double* h, double* qx, double* qy, double* maxh) {
  if (!s) {
    set_error("tr_sim_download: NULL sim");
    return TR_ERR_INVALID;
  }
  if (s->batch) return tr_batch_download(s->batch, s->member, h, qx, qy, maxh);
  int b = 0;
  if (s->form != TR_FORM_F64_UNFUSED) {
    StepCtrl c;
    RC(realign_fused(s, c));
    b = int(c.step & 1);
  }
  const size_t n = size_t(s->v.rows) * size_t(s->v.cols) * sizeof(double);
  if (h) TR_HIP(hipMemcpyAsync(h, s->h[b], n, hipMemcpyDeviceToHost, s->stream));
  if (qx) TR_HIP(hipMemcpyAsync(qx, s->qx[b], n, hipMemcpyDeviceToHost, s->stream));
  if (qy) TR_HIP(hipMemcpyAsync(qy, s->qy[b], n, hipMemcpyDeviceToHost, s->stream));
  if (maxh && s->maxh) TR_HIP(hipMemcpyAsync(maxh, s->maxh, n, hipMemcpyDeviceToHost, s->stream));
  TR_HIP(hipStreamSynchronize(s->stream));
  return TR_OK;
}

int tr_sim_sample(tr_sim* s, int32_t n, const int32_t* pos, double* h, double* qx, double* qy) {
  if (!s || n < 0 || (n > 0 && !pos)) {
    set_error("tr_sim_sample: bad arguments");
    return TR_ERR_INVALID;
  }
  if (s->batch) return tr_batch_sample(s->batch, s->member, n, pos, h, qx, qy);
  if (n == 0) return TR_OK;
  const int64_t cells = int64_t(s->v.rows) * s->v.cols;
  for (int32_t i = 0; i < n; i++)
    if (pos[i] < 0 || pos[i] >= cells) {
      set_error("tr_sim_sample: cell index out of range");
      return TR_ERR_INVALID;
    }
  int b = 0;
  if (s->form != TR_FORM_F64_UNFUSED) {
    StepCtrl c;
    RC(realign_fused(s, c));
    b = int(c.step & 1);
  }
  if (n > s->obs_cap) {
    if (s->d_obs_idx) hipFree(s->d_obs_idx);
    if (s->d_obs_val) hipFree(s->d_obs_val);
    s->d_obs_idx = nullptr;
    s->d_obs_val = nullptr;
    s->obs_cap = 0;
    TR_HIP(hipMalloc(&s->d_obs_idx, sizeof(int32_t) * size_t(n)));
    TR_HIP(hipMalloc(&s->d_obs_val, sizeof(double) * 3 * size_t(n)));
    s->obs_cap = n;
  }
  TR_HIP(hipMemcpyAsync(s->d_obs_idx, pos, sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice,
                        s->stream));
  const double* src[3] = {s->h[b], s->qx[b], s->qy[b]};
  double* dst[3] = {h, qx, qy};
  for (int k = 0; k < 3; k++) {
    if (!dst[k]) continue;
    RC(copy_obs_f64(n, src[k], s->d_obs_val + size_t(k) * n, s->d_obs_idx, s->stream));
    TR_HIP(hipMemcpyAsync(dst[k], s->d_obs_val + size_t(k) * n, sizeof(double) * size_t(n),
                          hipMemcpyDeviceToHost, s->stream));
  }
  TR_HIP(hipStreamSynchronize(s->stream));
  return TR_OK;
}

int tr_sim_set_obs_halt(tr_sim* s, int32_t on) {
  if (!s) {
    set_error("tr_sim_set_obs_halt: NULL sim");
    return TR_ERR_INVALID;
  }
  if (s->batch) return tr_batch_set_obs_halt(s->batch, s->member, on);
  s->halt_on_obs = on ? 1 : 0;
  return TR_OK;
}

// Rows per wave of the FP64 fused step for the launches queued next (0: by
// grid size); results do not depend on it.
int tr_sim_set_strip_rows(tr_sim* s, int32_t strip_rows) {
  if (!s || strip_rows < 0 || strip_rows > 4096) {
    set_error("tr_sim_set_strip_rows: strip_rows must be in [0, 4096]");
    return TR_ERR_INVALID;
  }
  if (s->form != TR_FORM_F64_FUSED && s->form != TR_FORM_F64_R32) {
    set_error("tr_sim_set_strip_rows: only the fused FP64 forms have strips");
    return TR_ERR_STATE;
  }
  TR_HIP(hipStreamSynchronize(s->stream));
  s->strip_rows = strip_rows > 0 ? strip_rows : tr_default_strip_rows(s->v.rows, s->v.cols);
  return TR_OK;
}

int32_t tr_sim_strip_rows(const tr_sim* s) { return s ? s->strip_rows : 0; }

int tr_sim_set_waves(tr_sim* s, int32_t waves_per_simd) {
  if (!s || !(waves_per_simd == 0 || waves_per_simd == 3 || waves_per_simd == 4)) {
    set_error("tr_sim_set_waves: waves_per_simd must be 0, 3 or 4");
    return TR_ERR_INVALID;
  }
  if (s->form != TR_FORM_F64_FUSED && s->form != TR_FORM_F64_R32) {
    set_error("tr_sim_set_waves: only the fused FP64 forms have this build choice");
    return TR_ERR_STATE;
  }
  TR_HIP(hipStreamSynchronize(s->stream));
  s->waves4 = waves_per_simd == 4 || (waves_per_simd == 0 && default_waves_per_simd(s->v.rows, s->v.cols) == 4)
                  ? 1 : 0;
  return TR_OK;
}

int32_t tr_sim_waves(const tr_sim* s) { return s ? (s->waves4 ? 4 : 3) : 0; }

void* tr_sim_stream(tr_sim* s) {
  if (!s) return nullptr;
  if (s->batch) return tr_batch_stream(s->batch);
  return reinterpret_cast<void*>(s->stream);
}

int64_t tr_sim_device_bytes(const tr_sim* s) {
  if (!s) return 0;
  return s->batch ? tr_batch_device_bytes(s->batch) : s->bytes;
}

void tr_sim_destroy(tr_sim* s) { delete s; }

}  // extern "C"

// ---- fused FP64 run on caller buffers, with no allocation -----------------
// tr_sim_run_f64_noalloc: the kernels of the fused FP64 form (k_prime_f64,
// k_step_f64, k_close_f64) over device buffers the caller owns.  What the
// form otherwise allocates per simulation -- the two control blocks with
// their per-XCD dt minima, the rain depth each zone gains per step, the zone
// areas and the volume partial sums -- lives in `scratch`, a file-scope
// device array each GPU holds once for the life of the process.  Host code
// uses the array only through the address hipGetSymbolAddress returns for
// the current device (the array's own name is a host-side stand-in).  Every
// call on a GPU shares that GPU's array: calls that overlap on one GPU race,
// and each call starts by clearing what must start at zero.

namespace {

constexpr int kNoAllocMaxZones = 64;
constexpr size_t kScratchBytes = 16384;
// byte offsets into the scratch array
constexpr size_t kOffCtrl = 0;                                      // StepCtrl[2]
constexpr size_t kOffZoneAdd = (2 * sizeof(StepCtrl) + 255) / 256 * 256;  // double[2][64]
constexpr size_t kOffZoneCount = kOffZoneAdd + 2 * kNoAllocMaxZones * 8;  // u64[64]
constexpr size_t kOffZoneArea = kOffZoneCount + kNoAllocMaxZones * 8;     // double[64]
constexpr size_t kOffFlag = kOffZoneArea + kNoAllocMaxZones * 8;          // int
constexpr size_t kOffVolPart = (kOffFlag + 4 + 255) / 256 * 256;          // double[kVolBlocks]
constexpr size_t kOffVolSum = kOffVolPart + kVolBlocks * 8;          // double
static_assert(2 * sizeof(StepCtrl) <= kOffZoneAdd, "control blocks overlap the zone table");
static_assert(kOffFlag + 4 <= kOffVolPart, "zone tables overlap the volume partials");
static_assert(kOffVolSum + 8 <= kScratchBytes, "scratch too small");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per-device address of `scratch`, taken once per device with that device
// current (hipGetSymbolAddress), then used for every kernel argument.
constexpr int kMaxDevices = 64;
unsigned char* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

int scratch_for_current_device(unsigned char** out) {
  int dev = 0;
  TR_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("tr_sim_run_f64_noalloc: device index out of range");
    return TR_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e == hipSuccess ? hipErrorInvalidSymbol : e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_scratch_ptr[dev];
  return TR_OK;
}

// Clears the zone counters and the flag, and writes the initial control
// block into slot 1 (the prime reduces into it and prepares slot 0).
__global__ void k_noalloc_init(unsigned char* sc, double sim_start, int print_id, int roff_row0) {
  const int t = int(threadIdx.x);
  auto* cnt = reinterpret_cast<unsigned long long*>(sc + kOffZoneCount);
  if (t < kNoAllocMaxZones) cnt[t] = 0ull;
  if (t == 0) {
    *reinterpret_cast<int*>(sc + kOffFlag) = 0;
    StepCtrl* ctrl = reinterpret_cast<StepCtrl*>(sc + kOffCtrl);
    for (int k = 0; k < 2; k++) {
      StepCtrl& c = ctrl[k];
      c.simtime = sim_start;
      c.dt = 0.0;
      c.pad2 = 0;
      c.halt = 0;
      c.step = 0;
      c.print_id = print_id;
      c.roff_row = roff_row0;
      c.idx_low = 0;
      c.idx_high = 0;
      c.finished = 0;
      c.observed = 0;
      c.vol_in = 0.0;
      c.step_vol_in = 0.0;
      c.last_obs = sim_start;
      c.dt_acc = 0.0;
      reset_dtmin(c);
    }
  }
}

// Interior cells per runoff zone (exact integer counts), and a flag for a
// zone id outside [0, nz).
__global__ void k_noalloc_zone_count(const int* __restrict__ rid, int rows, int cols, int nz,
                                     unsigned char* sc) {
  auto* cnt = reinterpret_cast<unsigned long long*>(sc + kOffZoneCount);
  const int n_int = (rows - 2) * (cols - 2);
  for (int k = int(blockIdx.x * blockDim.x + threadIdx.x); k < n_int; k += int(gridDim.x * blockDim.x)) {
    const int r = k / (cols - 2) + 1, c = k % (cols - 2) + 1;
    const int z = rid ? rid[(long long)r * cols + c] : 0;
    if (z < 0 || z >= nz) atomicOr(reinterpret_cast<int*>(sc + kOffFlag), 1);
    else atomicAdd(&cnt[z], 1ull);
  }
}

// zone area = cells x dx^2, as the host form computes it
__global__ void k_noalloc_zone_area(unsigned char* sc, int nz, double dx) {
  const int z = int(threadIdx.x);
  if (z >= nz) return;
  const auto* cnt = reinterpret_cast<const unsigned long long*>(sc + kOffZoneCount);
  reinterpret_cast<double*>(sc + kOffZoneArea)[z] = double(cnt[z]) * (dx * dx);
}

}  // namespace

extern "C" int tr_sim_run_f64_noalloc(const tr_dev_case_f64* dc, double t_stop, int64_t max_steps,
                                      tr_sim_status* st, void* stream) {
  if (!dc || !st) {
    set_error("tr_sim_run_f64_noalloc: NULL argument");
    return TR_ERR_INVALID;
  }
  const tr_dev_case_f64& c = *dc;
  if (c.rows < 3 || c.cols < 3 || int64_t(c.rows) * c.cols > int64_t(INT32_MAX)) {
    set_error("tr_sim_run_f64_noalloc: rows and cols must be >= 3 and the grid under 2^31 cells");
    return TR_ERR_INVALID;
  }
  for (int k = 0; k < 2; k++)
    if (!c.h[k] || !c.qx[k] || !c.qy[k]) {
      set_error("tr_sim_run_f64_noalloc: both state buffers of h, qx and qy are needed");
      return TR_ERR_INVALID;
    }
  if (!c.dem || !c.nman2) {
    set_error("tr_sim_run_f64_noalloc: dem and nman2 are needed");
    return TR_ERR_INVALID;
  }
  if (!(c.dx > 0) || !(c.courant > 0) || !(c.print_interval > 0) || !(c.init_dt > 0)) {
    set_error("tr_sim_run_f64_noalloc: dx, courant, print_interval and init_dt must be positive");
    return TR_ERR_INVALID;
  }
  if (c.num_runoffs < 0 || c.num_runoffs > kNoAllocMaxZones) {
    set_error("tr_sim_run_f64_noalloc: " + std::to_string(c.num_runoffs) +
              " runoff zones; at most 64 fit the static scratch array");
    return TR_ERR_INVALID;
  }
  if (c.num_runoffs > 0 && (c.roff_nrows < 2 || !c.roff_time || !c.roff_rate || c.roff_row0 < 0 ||
                            c.roff_row0 >= c.roff_nrows - 1)) {
    set_error("tr_sim_run_f64_noalloc: runoff needs time and rate tables of at least two rows");
    return TR_ERR_INVALID;
  }
  unsigned char* sc = nullptr;
  RC(scratch_for_current_device(&sc));
  hipStream_t s = as_stream(stream);
  StepCtrl* ctrl = reinterpret_cast<StepCtrl*>(sc + kOffCtrl);
  int print_id = c.checkpoint_id;
  if (print_id == 0 && c.sim_start > 0.0) print_id = int(c.sim_start / c.print_interval);
  hipLaunchKernelGGL(k_noalloc_init, dim3(1), dim3(64), 0, s, sc, c.sim_start, print_id,
                     c.num_runoffs > 0 ? c.roff_row0 : 0);
  if (c.num_runoffs > 0) {
    const int64_t n_int = int64_t(c.rows - 2) * (c.cols - 2);
    hipLaunchKernelGGL(k_noalloc_zone_count, dim3(unsigned(std::min<int64_t>((n_int + 255) / 256, 4096))),
                       dim3(256), 0, s, c.runoff_id, c.rows, c.cols, c.num_runoffs, sc);
    hipLaunchKernelGGL(k_noalloc_zone_area, dim3(1), dim3(kNoAllocMaxZones), 0, s, sc, c.num_runoffs, c.dx);
    RC(launch_check("noalloc zones"));
    int flag = 0;
    TR_HIP(hipMemcpyAsync(&flag, sc + kOffFlag, sizeof(int), hipMemcpyDeviceToHost, s));
    TR_HIP(hipStreamSynchronize(s));
    if (flag) {
      set_error("tr_sim_run_f64_noalloc: runoff zone id out of range");
      return TR_ERR_INVALID;
    }
  }
  const int strip = tr_default_strip_rows(c.rows, c.cols);
  const double stop = t_stop > 0 ? std::min(t_stop, c.sim_duration) : c.sim_duration;
  const int64_t step_stop = max_steps < 0 ? INT64_MAX : max_steps;
  StepParams<double> p{};
  p.rows = c.rows;
  p.cols = c.cols;
  p.strip_rows = strip;
  p.waves4 = default_waves_per_simd(c.rows, c.cols) == 4 ? 1 : 0;
  p.nbx = fused_grid_x(c.cols);
  p.nby = (c.rows - 2 + strip - 1) / strip;
  p.nblocks = unsigned(p.nbx) * unsigned(p.nby);
  p.dx = c.dx;
  p.hextra = c.hextra;
  p.cn_dx = c.courant * c.dx;
  p.dem = c.dem;
  p.nman2 = c.nman2;
  p.maxh = c.maxh;
  p.runoff_id = c.runoff_id;
  p.roff_rate = c.roff_rate;
  p.roff_time = c.roff_time;
  p.roff_nrows = c.roff_nrows;
  p.num_runoffs = c.num_runoffs;
  p.zone_area = reinterpret_cast<const double*>(sc + kOffZoneArea);
  p.zone_add = c.num_runoffs > 0 ? reinterpret_cast<double*>(sc + kOffZoneAdd) : nullptr;
  p.nsrc = 0;
  p.open_bc = c.open_boundaries;
  p.ctrl = ctrl;
  p.sim_duration = c.sim_duration;
  p.print_interval = c.print_interval;
  p.init_dt = c.init_dt;
  p.t_stop = stop;
  p.step_stop = step_stop;
  p.it_print = 0;
  p.it_count0 = c.it_count;
  p.bc = ExtBC{};
  p.round32 = 0;
  auto set_bufs = [&](int64_t launch) {
    const int a = int(launch & 1);
    p.cur = a;
    p.h_in = c.h[a];
    p.qx_in = c.qx[a];
    p.qy_in = c.qy[a];
    p.h_out = c.h[a ^ 1];
    p.qx_out = c.qx[a ^ 1];
    p.qy_out = c.qy[a ^ 1];
  };
  // the prime reduces the initial dt into slot 1 and prepares slot 0
  set_bufs(0);
  p.cur = 1;
  RC(launch_prime_f64(p, s));
  StepCtrl hc;
  int64_t launched = 0, chunk = 32;
  for (;;) {
    TR_HIP(hipMemcpyAsync(&hc, ctrl + (launched & 1), sizeof(StepCtrl), hipMemcpyDeviceToHost, s));
    TR_HIP(hipStreamSynchronize(s));
    if (hc.halt) break;
    for (int64_t i = 0; i < chunk; i++) {
      set_bufs(launched + i);
      RC(launch_step_f64(p, s));
    }
    launched += chunk;
    chunk = std::min<int64_t>(chunk * 2, 1024);
  }
  // volume of the state reached (buffer steps & 1), deterministic order
  double* vpart = reinterpret_cast<double*>(sc + kOffVolPart);
  double* vsum = reinterpret_cast<double*>(sc + kOffVolSum);
  hipLaunchKernelGGL(k_volume_part, dim3(kVolBlocks), dim3(256), 0, s, c.h[hc.step & 1], c.rows, c.cols,
                     c.dx * c.dx, vpart);
  hipLaunchKernelGGL(k_volume_sum, dim3(1), dim3(64), 0, s, vpart, kVolBlocks, vsum);
  RC(launch_check("noalloc volume"));
  double vol = 0.0;
  TR_HIP(hipMemcpyAsync(&vol, vsum, sizeof(double), hipMemcpyDeviceToHost, s));
  TR_HIP(hipStreamSynchronize(s));
  std::memset(st, 0, sizeof(*st));
  st->simtime = hc.simtime;
  st->dt = hc.dt;
  st->steps = hc.step;
  st->print_id = hc.print_id;
  st->finished = hc.finished;
  st->volume = vol;
  st->volume_in = hc.vol_in;
  st->last_obs = hc.last_obs;
  st->average_dt = hc.dt_acc;
  st->it_count = int64_t(c.it_count) + hc.step;
  return TR_OK;
}

/*
 * triton_oracle.c -- CPU FP64 restatement of TRITON's explicit finite-volume
 * timestep.  TEST INFRASTRUCTURE ONLY (the checker, never the product).
 *
 * Each function names the reference file:line it restates.  Floating-point
 * expressions keep the reference's association order; build with
 * -ffp-contract=off so no FMA is formed.
 */
#include "triton_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/constants.h:43-50,117-118 (FP64 build) */
#define G_ACC 9.81
#define SQRT_G 3.132091953
#define EPS12 1e-12
#define MAX_DT 1e10

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Threads of the parallel loops that follow (timing at several counts). */
void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* src/kernels.h:166-187 -- sqrt(h) cache and per-cell CFL step, every cell
 * including the ghost ring. */
void orc_compute_dt_and_sqrt(int size, double dx, const double *qx, const double *qy,
                             const double *h, double *sqrth, double *dtarr,
                             double cn, double hextra) {
#pragma omp parallel for schedule(static)
  for (int id = 0; id < size; id++) {
    double hc = h[id];
    double sq = fmax(sqrt(hc), 0.0);
    double out = MAX_DT;
    sqrth[id] = sq;
    if (hc > hextra) {
      double umax = fmax(fabs(qx[id]), fabs(qy[id])) / hc;
      out = cn * dx / (umax + SQRT_G * sq);
    }
    dtarr[id] = out;
  }
}

/* src/kernels.h:195-212 -- exact min; the comparison form never selects NaN. */
double orc_find_min_dt(int size, const double *dtarr) {
  double best = INFINITY;
#pragma omp parallel for schedule(static) reduction(min : best)
  for (int id = 0; id < size; id++) {
    double v = dtarr[id];
    best = v < best ? v : best;
  }
  return best;
}

/* One face of the Roe-type solver with bed-slope source, both directions:
 * src/kernels_flux.h:81-265 (x, n=(1,0)) and :364-547 (y, n=(0,1)).  Cell 1
 * is the face's west (x) / south (y) cell, cell 2 its east / north cell. */
int orc_face(double h1, double h2, double qx1, double qx2, double qy1, double qy2,
             double z1, double z2, double sqh1, double sqh2, double nx, double ny,
             double dx, double dt, double hextra, double d1[3], double d2[3]) {
  d1[0] = d1[1] = d1[2] = 0.0;
  d2[0] = d2[1] = d2[2] = 0.0;
  if (!(h1 > 0.0 || h2 > 0.0)) return 0;

  /* velocities with the hextra cut (:98-117) */
  double u1 = 0.0, v1 = 0.0, u2 = 0.0, v2 = 0.0;
  if (h1 >= hextra) { u1 = qx1 / h1; v1 = qy1 / h1; }
  if (h2 >= hextra) { u2 = qx2 / h2; v2 = qy2 / h2; }

  /* Roe averages from the cached sqrt(h) (:119-133) */
  double hmid = 0.5 * (h1 + h2);
  double dh = h2 - h1;
  double c = sqrt(G_ACC * hmid);
  double ubar = (u1 * sqh1 + u2 * sqh2) / (sqh1 + sqh2);
  double vbar = (v1 * sqh1 + v2 * sqh2) / (sqh1 + sqh2);
  double un = ubar * nx + vbar * ny;
  double lam[3] = {un - c, un, un + c};
  double lamE[3] = {0.0, 0.0, 0.0};

  /* Harten-Hyman entropy fix on both acoustic waves (:137-152) */
  double eL = u1 * nx + v1 * ny - SQRT_G * sqh1;
  double eR = u2 * nx + v2 * ny - SQRT_G * sqh2;
  if (eL < 0.0 && eR > 0.0) {
    double split = eL * (eR - lam[0]) / (eR - eL);
    lamE[0] = lam[0] - split;
    lam[0] = split;
  }
  eL = u1 * nx + v1 * ny + SQRT_G * sqh1;
  eR = u2 * nx + v2 * ny + SQRT_G * sqh2;
  if (eL < 0.0 && eR > 0.0) {
    double split = eR * (lam[2] - eL) / (eR - eL);
    lamE[2] = lam[2] - split;
    lam[2] = split;
  }

  /* right eigenvectors, row i = eigenvector of wave i (:154-162) */
  double ev[3][3];
  ev[0][0] = 1.0; ev[0][1] = ubar - c * nx; ev[0][2] = vbar - c * ny;
  ev[1][0] = 0.0; ev[1][1] = -c * ny;       ev[1][2] = c * nx;
  ev[2][0] = 1.0; ev[2][1] = ubar + c * nx; ev[2][2] = vbar + c * ny;

  /* wave strengths (:164-166) */
  double dqx = qx2 - qx1, dqy = qy2 - qy1;
  double alpha[3];
  alpha[0] = 0.5 * (dh - ((dqx * nx + dqy * ny) - un * dh) / c);
  alpha[1] = ((dqy - vbar * dh) * nx - (dqx - ubar * dh) * ny) / c;
  alpha[2] = 0.5 * (dh + ((dqx * nx + dqy * ny) - un * dh) / c);

  /* bed-slope source with the dry-step limit (:168-192) */
  double dz = z2 - z1;
  double lev1 = z1 + h1, lev2 = z2 + h2;
  double dzp = dz, hp;
  if (dz >= 0.0) {
    hp = h1;
    if (lev1 < z2) dzp = h1;
  } else {
    hp = h2;
    if (lev2 < z1) dzp = -h2;
  }
  double beta[3];
  beta[0] = 0.5 * G_ACC * (hp - 0.5 * fabs(dzp)) * dzp / c;

  /* subcritical wet-wet clip (:195-200) */
  hp = h1 + alpha[0];
  if (lam[0] * lam[2] < 0.0 && hp > 0.0 && h1 > 0.0 && h2 > 0.0) {
    beta[0] = fmax(beta[0], lam[0] / (lam[0] + lamE[2]) *
                                (alpha[0] * lam[0] - h1 * dx * 0.5 / dt + lamE[2] * h2));
    beta[0] = fmin(beta[0], lam[2] / (lam[2] + lamE[0]) *
                                (-alpha[2] * lam[2] + h2 * dx * 0.5 / dt + lamE[0] * h1));
  }
  beta[1] = 0.0;
  beta[2] = -beta[0];

  /* intermediate states and the dry-side redirection (:205-236) */
  double sL = hp - beta[0] / lam[0];
  double sR = hp + beta[2] / lam[2];
  if (sR < -EPS12 && h2 < EPS12) {
    d1[0] += (lam[0] * alpha[0] - beta[0]) + (lam[2] * alpha[2] - beta[2]);
  } else if (sL < -EPS12 && h1 < EPS12) {
    d2[0] += (lam[0] * alpha[0] - beta[0]) + (lam[2] * alpha[2] - beta[2]);
  } else {
    for (int i = 0; i < 3; i++) {
      double w = lam[i] * alpha[i] - beta[i];
      double *dst = lam[i] > 0.0 ? d2 : d1;
      dst[0] += w * ev[i][0];
      dst[1] += w * ev[i][1];
      dst[2] += w * ev[i][2];
    }
  }
  /* entropy-fix parts (:238-253) */
  for (int i = 0; i < 3; i++) {
    double w = lamE[i] * alpha[i];
    double *dst = lamE[i] > 0.0 ? d2 : d1;
    dst[0] += w * ev[i][0];
    dst[1] += w * ev[i][1];
    dst[2] += w * ev[i][2];
  }
  /* flush (:255-265) */
  for (int i = 0; i < 3; i++) {
    if (fabs(d1[i]) < EPS12) d1[i] = 0.0;
    if (fabs(d2[i]) < EPS12) d2[i] = 0.0;
  }
  return 1;
}

/* src/kernels_flux.h:31-285 -- x faces (id, id+1) for rows 1..R-2 and cols
 * 0..C-2; both sides are overwritten, with zeros when the face is dry. */
void orc_flux_x(int rows, int cols, double dx, double dt, const double *h,
                const double *qx, const double *qy, const double *dem,
                const double *sqrth, double *rhsh0, double *rhsh1, double *rhsqx0,
                double *rhsqx1, double *rhsqy0, double *rhsqy1, double hextra) {
#pragma omp parallel for schedule(static)
  for (int r = 1; r < rows - 1; r++) {
    for (int cc = 0; cc < cols - 1; cc++) {
      int a = r * cols + cc, b = a + 1;
      double d1[3], d2[3];
      if (orc_face(h[a], h[b], qx[a], qx[b], qy[a], qy[b], dem[a], dem[b], sqrth[a],
                   sqrth[b], 1.0, 0.0, dx, dt, hextra, d1, d2)) {
        rhsh0[a] = dt * d1[0] / dx;
        rhsqx0[a] = dt * d1[1] / dx;
        rhsqy0[a] = dt * d1[2] / dx;
        rhsh1[b] = dt * d2[0] / dx;
        rhsqx1[b] = dt * d2[1] / dx;
        rhsqy1[b] = dt * d2[2] / dx;
      } else {
        rhsh0[a] = rhsqx0[a] = rhsqy0[a] = 0.0;
        rhsh1[b] = rhsqx1[b] = rhsqy1[b] = 0.0;
      }
    }
  }
}

/* src/kernels_flux.h:308-559 -- y faces between id (south) and id-cols (north)
 * for rows 1..R-1 and cols 1..C-2, accumulated onto the x results; a dry face
 * writes nothing. */
void orc_flux_y(int rows, int cols, double dx, double dt, const double *h,
                const double *qx, const double *qy, const double *dem,
                const double *sqrth, double *rhsh0, double *rhsh1, double *rhsqx0,
                double *rhsqx1, double *rhsqy0, double *rhsqy1, double hextra) {
#pragma omp parallel for schedule(static)
  for (int r = 1; r < rows; r++) {
    for (int cc = 1; cc < cols - 1; cc++) {
      int a = r * cols + cc, b = a - cols;
      double d1[3], d2[3];
      if (orc_face(h[a], h[b], qx[a], qx[b], qy[a], qy[b], dem[a], dem[b], sqrth[a],
                   sqrth[b], 0.0, 1.0, dx, dt, hextra, d1, d2)) {
        rhsh0[a] += dt * d1[0] / dx;
        rhsqx0[a] += dt * d1[1] / dx;
        rhsqy0[a] += dt * d1[2] / dx;
        rhsh1[b] += dt * d2[0] / dx;
        rhsqx1[b] += dt * d2[1] / dx;
        rhsqy1[b] += dt * d2[2] / dx;
      }
    }
  }
}

/* src/kernels_state.h:30-104 -- depth update, clamp, hextra cut and the
 * semi-implicit Manning friction, interior cells only. */
void orc_update_cells(int rows, int cols, double dt, double *h, double *qx, double *qy,
                      const double *nman, const double *rhsh0, const double *rhsh1,
                      const double *rhsqx0, const double *rhsqx1, const double *rhsqy0,
                      const double *rhsqy1, double hextra) {
#pragma omp parallel for schedule(static)
  for (int r = 1; r < rows - 1; r++) {
    for (int cc = 1; cc < cols - 1; cc++) {
      int id = r * cols + cc;
      double hn = h[id];
      double hnew = hn - rhsh0[id] - rhsh1[id];
      double qxn, qyn;
      if (hnew < 0.0) hnew = 0.0;
      if (hnew < hextra) {
        qxn = 0.0;
        qyn = 0.0;
      } else {
        double mx = qx[id] - rhsqx0[id] - rhsqx1[id];
        double my = qy[id] - rhsqy0[id] - rhsqy1[id];
        double modm = sqrt(mx * mx + my * my);
        double n2 = nman[id];
        if (n2 > EPS12 && hn >= hextra && modm > EPS12) {
          double tt = dt * G_ACC * n2 * modm / (hn * hn * cbrt(hn));
          qxn = -0.5 * (mx - mx * sqrt(1.0 + 4.0 * tt)) / tt;
          qyn = -0.5 * (my - my * sqrt(1.0 + 4.0 * tt)) / tt;
        } else {
          qxn = mx;
          qyn = my;
        }
      }
      h[id] = hnew;
      qx[id] = qxn;
      qy[id] = qyn;
    }
  }
}

/* src/kernels.h:102-152 -- zonal runoff, piecewise constant in time. */
void orc_update_runoff(int rows, int cols, double dt, const int *runoff_id,
                       int index_row_runoff, int n_rows_runoff, const double *rate,
                       double *h, double *qx, double *qy, double hextra) {
#pragma omp parallel for schedule(static)
  for (int r = 1; r < rows - 1; r++) {
    for (int cc = 1; cc < cols - 1; cc++) {
      int id = r * cols + cc;
      double hc = h[id];
      double net = rate[runoff_id[id] * n_rows_runoff + index_row_runoff];
      hc += net * dt;
      if (hc < hextra) {
        if (hc < 0.0) hc = 0.0;
        qx[id] = 0.0;
        qy[id] = 0.0;
      }
      h[id] = hc;
    }
  }
}

/* src/kernels.h:50-84 -- point inflows, linear in time between hydrograph
 * rows.  Sources are applied in list order, so two sources on one cell add
 * deterministically (the reference's parallel update races there). */
void orc_compute_flow(int num_src, const double *hyg_time, const double *hyg_val,
                      double dx, double dt, double simtime, int idx_low, int idx_high,
                      double *h, const int *src_pos) {
  for (int s = 0; s < num_src; s++) {
    double flo = hyg_val[idx_low * num_src + s];
    double fhi = hyg_val[idx_high * num_src + s];
    double tlo = hyg_time[idx_low];
    double thi = hyg_time[idx_high];
    double q;
    if (idx_low == idx_high) {
      q = flo;
    } else {
      double el = simtime - tlo;
      double span = thi - tlo;
      q = flo + (((fhi - flo) * el) / span);
    }
    int cell = src_pos[s];
    double hc = h[cell];
    double add = (q * dt) / (dx * dx);
    hc += add;
    h[cell] = hc;
  }
}

/* src/kernels.h:436-460 -- transmissive west/east ghosts, every row. */
void orc_open_bc_west_east(int rows, int cols, double *h, double *qx, double *qy) {
  for (int r = 0; r < rows; r++) {
    int w = r * cols, e = r * cols + cols - 1;
    h[w] = h[w + 1]; qx[w] = qx[w + 1]; qy[w] = qy[w + 1];
    h[e] = h[e - 1]; qx[e] = qx[e - 1]; qy[e] = qy[e - 1];
  }
}

/* src/kernels.h:476-512 -- transmissive north/south ghosts, every column;
 * run after west/east so a corner takes its diagonal inner value. */
void orc_open_bc_north_south(int rows, int cols, double *h, double *qx, double *qy) {
  for (int cc = 0; cc < cols; cc++) {
    int n = cc, s = (rows - 1) * cols + cc;
    h[n] = h[n + cols]; qx[n] = qx[n + cols]; qy[n] = qy[n + cols];
    h[s] = h[s - cols]; qx[s] = qx[s - cols]; qy[s] = qy[s - cols];
  }
}

/* src/kernels.h:236-424 -- external boundary cells (single subdomain).
 * Type 0 copies inner to ghost; type 1 imposes a water level interpolated at
 * the step midpoint; type 2 a normal-slope velocity; type 3 a Froude number.
 * Cells are visited in list order. */
void orc_compute_extbc_values(int ncells, int rows, int cols, double dt, double *h,
                              double *qx, double *qy, const double *dem,
                              const double *nman, const int *bc_index,
                              const int *bc_type, const int *bc_start,
                              const int *bc_nrows, const double *var1,
                              const double *var2, double simtime) {
  for (int k = 0; k < ncells; k++) {
    int ii = bc_index[k], type = bc_type[k], st = bc_start[k];
    int r = ii / cols, cc = ii % cols;
    int top = (r == 1), btm = (r == rows - 2), lt = (cc == 1), rt = (cc == cols - 2);
    if (!(top || lt || rt || btm)) continue;
    int ghost = ii, corner_cell = -1;
    double f[3];
    f[0] = 1.0; f[1] = 0.0; f[2] = 0.0;
    if (lt) { ghost = ii - 1; f[1] = -1.0; f[2] = 0.0; }
    if (rt) { ghost = ii + 1; f[1] = 1.0; f[2] = 0.0; }
    if (top) {
      ghost = ii - cols; f[1] = 0.0; f[2] = 1.0;
      if (lt) { corner_cell = ii - 1; f[2] = 0.0; }
      if (rt) { corner_cell = ii + 1; f[2] = 0.0; }
    }
    if (btm) {
      ghost = ii + cols; f[1] = 0.0; f[2] = -1.0;
      if (lt) { corner_cell = ii - 1; f[2] = 0.0; }
      if (rt) { corner_cell = ii + 1; f[2] = 0.0; }
    }
    double hc = h[ii], qxc = qx[ii], qyc = qy[ii];
    if (type == 0) {
      h[ghost] = hc; qx[ghost] = qxc; qy[ghost] = qyc;
      if (corner_cell >= 0) { h[corner_cell] = hc; qx[corner_cell] = qxc; qy[corner_cell] = qyc; }
      continue;
    }
    if (type == 1) {
      double tmid = 0.5 * (simtime + simtime + dt);
      int lo = -999, hi = bc_nrows[k] - 1;
      if (tmid < var1[st]) {
        lo = 0; hi = 0;
      } else if (tmid > var1[hi + st]) {
        lo = hi;
      } else {
        for (int i = 0; i < hi; i++) {
          if (var1[i + st] <= tmid && var1[i + 1 + st] > tmid) { lo = i; hi = i + 1; break; }
        }
        /* tmid equal to the last time leaves the reference's index unset
         * (an out-of-bounds read there); take the last row instead. */
        if (lo < 0) lo = hi;
      }
      double level;
      if (lo == hi) {
        level = var2[lo + st];
      } else {
        double el = tmid - var1[lo + st];
        double span = var1[hi + st] - var1[lo + st];
        level = var2[lo + st] + (((var2[hi + st] - var2[lo + st]) * el) / span);
      }
      hc = fmax(level - dem[ii], 0.0);
    } else {
      double param = var1[st];
      double vel = 0.0;
      if (type == 2) vel = sqrt(param * hc * cbrt(hc) / fmax(nman[ii], 1e-6));
      if (type == 3) vel = param * sqrt(G_ACC * hc);
      qxc = hc * vel;
      qyc = hc * vel;
    }
    h[ghost] = hc * f[0]; qx[ghost] = qxc * f[1]; qy[ghost] = qyc * f[2];
    h[ii] = hc * f[0]; qx[ii] = qxc * f[1]; qy[ii] = qyc * f[2];
    if (corner_cell >= 0) {
      h[corner_cell] = hc * f[0]; qx[corner_cell] = qxc * f[1]; qy[corner_cell] = qyc * f[2];
    }
  }
}

/* src/kernels_wetdry.h:24-85 (single subdomain branch) -- zero the normal
 * discharge against a higher dry neighbour; track the max depth. */
void orc_wet_dry(int rows, int cols, const double *h, double *qx, double *qy,
                 const double *dem, double *maxh, double hextra) {
#pragma omp parallel for schedule(static)
  for (int r = 1; r < rows - 1; r++) {
    for (int cc = 1; cc < cols - 1; cc++) {
      int id = r * cols + cc;
      double hc = h[id], zc = dem[id];
      if (hc > hextra) {
        int e = id + 1, w = id - 1, n = id - cols, s = id + cols;
        if ((hc + zc < dem[e] && h[e] < EPS12) || (hc + zc < dem[w] && h[w] < EPS12))
          qx[id] = 0.0;
        if ((hc + zc < dem[n] && h[n] < EPS12) || (hc + zc < dem[s] && h[s] < EPS12))
          qy[id] = 0.0;
      }
      if (maxh && hc > maxh[id]) maxh[id] = hc;
    }
  }
}

/* src/triton_simulate.h:163-187 (single subdomain) */
static double init_dt_of(const orc_case *c) {
  double idt = c->time_series_flag ? fmin(c->print_observation, c->print_interval)
                                   : c->print_interval;
  if (c->num_runoffs > 0) idt = fmin(idt, c->roff_time[1] - c->roff_time[0]);
  if (c->num_src > 0) idt = fmin(idt, c->hyg_time[1] - c->hyg_time[0]);
  return idt * 0.01;
}

/* src/triton_simulate.h:54-134 with compute_local_dt (:191-198),
 * compute_global_dt (:202-226) and compute_new_state (:230-346); the runoff
 * row starts where src/triton_init.h:1299-1311 puts it. */
/* One observation row: simtime and the state at the observation cells. */
static void obs_row(const orc_case *c, double simtime) {
  if (!c->time_series_flag || !c->obs_rec || !c->obs_rows) return;
  int k = *c->obs_rows;
  if (k < c->obs_cap) {
    double *row = c->obs_rec + (size_t)k * (size_t)(1 + 3 * c->num_obs);
    row[0] = simtime;
    for (int i = 0; i < c->num_obs; i++) {
      int id = c->obs_pos[i];
      row[1 + i] = c->h[id];
      row[1 + c->num_obs + i] = c->qx[id];
      row[1 + 2 * c->num_obs + i] = c->qy[id];
    }
  }
  *c->obs_rows = k + 1;
}

int orc_run(const orc_case *c, double t_stop, long long max_steps, orc_result *res) {
  const int rows = c->rows, cols = c->cols, n = rows * cols;
  if (c->num_runoffs > 0 && c->roff_nrows < 2) return 2;
  if (c->num_src > 0 && c->hyg_nrows < 2) return 2;
  double *buf = (double *)calloc((size_t)n * 8, sizeof(double));
  if (!buf) return 1;
  double *sqrth = buf, *dtarr = buf + (size_t)n;
  double *rh0 = buf + 2 * (size_t)n, *rh1 = buf + 3 * (size_t)n;
  double *rqx0 = buf + 4 * (size_t)n, *rqx1 = buf + 5 * (size_t)n;
  double *rqy0 = buf + 6 * (size_t)n, *rqy1 = buf + 7 * (size_t)n;

  double simtime = c->sim_start;
  double init_dt = init_dt_of(c);
  int print_id = c->checkpoint_id;
  if (print_id == 0 && simtime > 0.0) print_id = (int)(simtime / c->print_interval);

  int roff_row = 0;
  if (c->num_runoffs > 0) {
    for (int j = 0; j < c->roff_nrows - 1; j++)
      if (simtime > c->roff_time[j] && simtime <= c->roff_time[j + 1]) roff_row = j;
  }
  int idx_low = 0;
  long long steps = 0;
  double dt = 0.0;
  double last_obs = simtime, average_dt = 0.0;  /* src/triton_simulate.h:42,47 */
  if (c->obs_rows) *c->obs_rows = 0;
  obs_row(c, simtime);                          /* :31-37 */

  const double stop = (t_stop > 0.0 && t_stop < c->sim_duration) ? t_stop : c->sim_duration;
  while (simtime < stop) {
    if (max_steps >= 0 && steps >= max_steps) break;
    steps++;
    /* compute_local_dt + compute_global_dt */
    orc_compute_dt_and_sqrt(n, c->dx, c->qx, c->qy, c->h, sqrth, dtarr, c->courant, c->hextra);
    dt = orc_find_min_dt(n, dtarr);
    if (dt >= MAX_DT - 1.0) dt = init_dt;
    double next_print = c->print_interval * (print_id + 1);
    if (simtime + dt > next_print) dt = next_print - simtime;

    /* compute_new_state */
    orc_flux_x(rows, cols, c->dx, dt, c->h, c->qx, c->qy, c->dem, sqrth, rh0, rh1, rqx0,
               rqx1, rqy0, rqy1, c->hextra);
    orc_flux_y(rows, cols, c->dx, dt, c->h, c->qx, c->qy, c->dem, sqrth, rh0, rh1, rqx0,
               rqx1, rqy0, rqy1, c->hextra);
    orc_update_cells(rows, cols, dt, c->h, c->qx, c->qy, c->nman, rh0, rh1, rqx0, rqx1,
                     rqy0, rqy1, c->hextra);
    if (c->num_runoffs > 0) {
      int last = c->roff_nrows - 1;
      if (simtime > c->roff_time[last]) {
        roff_row = last;
      } else if (simtime > c->roff_time[roff_row + 1]) {
        roff_row++;
      }
      orc_update_runoff(rows, cols, dt, c->runoff_id, roff_row, c->roff_nrows, c->roff_rate,
                        c->h, c->qx, c->qy, c->hextra);
    }
    if (c->num_src > 0) {
      int idx_high = c->hyg_nrows - 1, found = 0;
      for (int i = idx_low; i < idx_high; i++) {
        if (c->hyg_time[i + 1] > simtime) { idx_low = i; idx_high = i + 1; found = 1; break; }
      }
      if (!found) idx_low = idx_high;
      orc_compute_flow(c->num_src, c->hyg_time, c->hyg_val, c->dx, dt, simtime, idx_low,
                       idx_high, c->h, c->src_pos);
    }
    if (c->open_boundaries) {
      orc_open_bc_west_east(rows, cols, c->h, c->qx, c->qy);
      orc_open_bc_north_south(rows, cols, c->h, c->qx, c->qy);
    }
    if (c->num_extbc_cells > 0) {
      orc_compute_extbc_values(c->num_extbc_cells, rows, cols, dt, c->h, c->qx, c->qy,
                               c->dem, c->nman, c->bc_index, c->bc_type, c->bc_start,
                               c->bc_nrows, c->bc_var1, c->bc_var2, simtime);
    }
    orc_wet_dry(rows, cols, c->h, c->qx, c->qy, c->dem, c->maxh, c->hextra);

    simtime += dt;
    average_dt += dt;                                           /* :66 */
    if (c->it_print > 0 && ((long long)c->it_count0 + steps) % c->it_print == 0)
      average_dt = 0.0;                                         /* :69-75 */
    if (c->time_series_flag && simtime - last_obs >= c->print_observation) {
      last_obs += c->print_observation;                         /* :78-91 */
      obs_row(c, simtime);
    }
    if (simtime >= c->print_interval * (print_id + 1)) print_id++;
  }
  if (!(simtime < c->sim_duration)) obs_row(c, simtime);       /* :113-121 */
  free(buf);
  if (res) {
    res->steps = steps;
    res->simtime = simtime;
    res->last_dt = dt;
    res->print_id = print_id;
    res->init_dt = init_dt;
    res->last_obs = last_obs;
    res->average_dt = average_dt;
    res->it_count = (long long)c->it_count0 + steps;
  }
  return 0;
}

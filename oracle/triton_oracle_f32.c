/*
 * triton_oracle_f32.c -- the reference's per-kernel timestep restated for its
 * single-precision build (USE_SINGLE_PRECISION, src/constants.h:37-50): the
 * kernels are the same templates instantiated with T = float, so the state
 * and every T variable are float while the literals and constants stay
 * double (_G_ = 9.81, SQRTG = 3.132091953, 0.5, 1.0, 4.0): each mixed
 * expression is evaluated in double and rounded to float when it is stored,
 * EPS12 is 1e-6, FMAX/FMIN are fmaxf/fminf and the math functions (sqrt,
 * cbrt) take and return double.  C's usual arithmetic conversions are the
 * same as C++'s, so the expressions below are the reference's.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the tr_*_f32 drop-in kernels,
 * never part of the product.  Built with -ffp-contract=off.
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#include "triton_oracle.h"

#define G_ACC 9.81
#define SQRT_G 3.132091953
#define EPS12F 1e-6
#define MAX_VALUE 1e10

/* src/kernels.h:166-187 */
void orcf_compute_dt_and_sqrt(int size, float dx, const float *qx, const float *qy, const float *h,
                              float *sqrth, float *dtarr, float cn, float hextra) {
#pragma omp parallel for schedule(static)
  for (int id = 0; id < size; id++) {
    dtarr[id] = MAX_VALUE;
    float hij = h[id];
    float sq = fmaxf(sqrt(hij), 0.0);
    sqrth[id] = sq;
    if (hij > hextra) {
      float maxu = fmaxf(fabs(qx[id]), fabs(qy[id])) / hij;
      dtarr[id] = cn * dx / (maxu + SQRT_G * sq);
    }
  }
}

/* src/kernels.h:195-212 */
float orcf_find_min_dt(int size, const float *dtarr) {
  float best = INFINITY;
  for (int id = 0; id < size; id++) best = dtarr[id] < best ? dtarr[id] : best;
  return best;
}

/* One face, src/kernels_flux.h:81-265 (x) and :364-547 (y), in float. */
int orcf_face(float h1, float h2, float qx1, float qx2, float qy1, float qy2, float z1, float z2,
              float sqh1, float sqh2, float nx, float ny, float dx, float dt, float hextra, float d1[3],
              float d2[3]) {
  d1[0] = d1[1] = d1[2] = 0.0f;
  d2[0] = d2[1] = d2[2] = 0.0f;
  if (!(h1 > 0.0 || h2 > 0.0)) return 0;
  float u1, v1, u2, v2;
  if (h1 >= hextra) { u1 = qx1 / h1; v1 = qy1 / h1; } else { u1 = 0.0; v1 = 0.0; }
  if (h2 >= hextra) { u2 = qx2 / h2; v2 = qy2 / h2; } else { u2 = 0.0; v2 = 0.0; }
  float h12 = 0.5 * (h1 + h2);
  float dh = h2 - h1;
  float c12 = sqrt(G_ACC * h12);
  float u12 = (u1 * sqh1 + u2 * sqh2) / (sqh1 + sqh2);
  float v12 = (v1 * sqh1 + v2 * sqh2) / (sqh1 + sqh2);
  float un = u12 * nx + v12 * ny;
  float eig[3] = {un - c12, un, un + c12};
  float eigE[3] = {0.0f, 0.0f, 0.0f};
  float e1 = u1 * nx + v1 * ny - SQRT_G * sqh1;
  float e2 = u2 * nx + v2 * ny - SQRT_G * sqh2;
  if (e1 < 0.0 && e2 > 0.0) {
    eigE[0] = eig[0] - e1 * (e2 - eig[0]) / (e2 - e1);
    eig[0] = e1 * (e2 - eig[0]) / (e2 - e1);
  }
  e1 = u1 * nx + v1 * ny + SQRT_G * sqh1;
  e2 = u2 * nx + v2 * ny + SQRT_G * sqh2;
  if (e1 < 0.0 && e2 > 0.0) {
    eigE[2] = eig[2] - e2 * (eig[2] - e1) / (e2 - e1);
    eig[2] = e2 * (eig[2] - e1) / (e2 - e1);
  }
  float ev[3][3];
  ev[0][0] = 1.0; ev[0][1] = u12 - c12 * nx; ev[0][2] = v12 - c12 * ny;
  ev[1][0] = 0.0; ev[1][1] = -c12 * ny;      ev[1][2] = c12 * nx;
  ev[2][0] = 1.0; ev[2][1] = u12 + c12 * nx; ev[2][2] = v12 + c12 * ny;
  float alpha[3];
  alpha[0] = 0.5 * (dh - (((qx2 - qx1) * nx + (qy2 - qy1) * ny) - un * dh) / c12);
  alpha[1] = (((qy2 - qy1) - v12 * dh) * nx - ((qx2 - qx1) - u12 * dh) * ny) / c12;
  alpha[2] = 0.5 * (dh + (((qx2 - qx1) * nx + (qy2 - qy1) * ny) - un * dh) / c12);
  float dz = z2 - z1;
  float l1 = z1 + h1;
  float l2 = z2 + h2;
  float dzp = dz, hp;
  if (dz >= 0.0) {
    hp = h1;
    if (l1 < z2) dzp = h1;
  } else {
    hp = h2;
    if (l2 < z1) dzp = -h2;
  }
  float beta[3];
  beta[0] = 0.5 * G_ACC * (hp - 0.5 * fabs(dzp)) * dzp / c12;
  hp = h1 + alpha[0];
  if (eig[0] * eig[2] < 0.0 && hp > 0.0 && h1 > 0.0 && h2 > 0.0) {
    beta[0] = fmaxf(beta[0], eig[0] / (eig[0] + eigE[2]) * (alpha[0] * eig[0] - h1 * dx * 0.5 / dt + eigE[2] * h2));
    beta[0] = fminf(beta[0], eig[2] / (eig[2] + eigE[0]) * (-alpha[2] * eig[2] + h2 * dx * 0.5 / dt + eigE[0] * h1));
  }
  beta[1] = 0.0;
  beta[2] = -beta[0];
  l1 = hp - beta[0] / eig[0];
  l2 = hp + beta[2] / eig[2];
  if (l2 < -EPS12F && h2 < EPS12F) {
    d1[0] += (eig[0] * alpha[0] - beta[0]) + (eig[2] * alpha[2] - beta[2]);
  } else if (l1 < -EPS12F && h1 < EPS12F) {
    d2[0] += (eig[0] * alpha[0] - beta[0]) + (eig[2] * alpha[2] - beta[2]);
  } else {
    for (int i = 0; i < 3; i++) {
      float *dst = eig[i] > 0.0 ? d2 : d1;
      dst[0] += (eig[i] * alpha[i] - beta[i]) * ev[i][0];
      dst[1] += (eig[i] * alpha[i] - beta[i]) * ev[i][1];
      dst[2] += (eig[i] * alpha[i] - beta[i]) * ev[i][2];
    }
  }
  for (int i = 0; i < 3; i++) {
    float *dst = eigE[i] > 0.0 ? d2 : d1;
    dst[0] += (eigE[i] * alpha[i]) * ev[i][0];
    dst[1] += (eigE[i] * alpha[i]) * ev[i][1];
    dst[2] += (eigE[i] * alpha[i]) * ev[i][2];
  }
  for (int i = 0; i < 3; i++) {
    if (fabs(d1[i]) < EPS12F) d1[i] = 0.0;
    if (fabs(d2[i]) < EPS12F) d2[i] = 0.0;
  }
  return 1;
}

/* src/kernels_flux.h:31-285 */
void orcf_flux_x(int rows, int cols, float dx, float dt, const float *h, const float *qx, const float *qy,
                 const float *dem, const float *sqrth, float *rhsh0, float *rhsh1, float *rhsqx0, float *rhsqx1,
                 float *rhsqy0, float *rhsqy1, float hextra) {
#pragma omp parallel for schedule(static)
  for (int r = 1; r < rows - 1; r++)
    for (int c = 0; c < cols - 1; c++) {
      int a = r * cols + c, b = a + 1;
      float d1[3], d2[3];
      if (orcf_face(h[a], h[b], qx[a], qx[b], qy[a], qy[b], dem[a], dem[b], sqrth[a], sqrth[b], 1.0f, 0.0f, dx,
                    dt, hextra, d1, d2)) {
        rhsh0[a] = dt * d1[0] / dx; rhsqx0[a] = dt * d1[1] / dx; rhsqy0[a] = dt * d1[2] / dx;
        rhsh1[b] = dt * d2[0] / dx; rhsqx1[b] = dt * d2[1] / dx; rhsqy1[b] = dt * d2[2] / dx;
      } else {
        rhsh0[a] = rhsqx0[a] = rhsqy0[a] = 0.0f;
        rhsh1[b] = rhsqx1[b] = rhsqy1[b] = 0.0f;
      }
    }
}

/* src/kernels_flux.h:308-559: faces (id, id-cols), accumulated */
void orcf_flux_y(int rows, int cols, float dx, float dt, const float *h, const float *qx, const float *qy,
                 const float *dem, const float *sqrth, float *rhsh0, float *rhsh1, float *rhsqx0, float *rhsqx1,
                 float *rhsqy0, float *rhsqy1, float hextra) {
  for (int r = 1; r < rows; r++)
    for (int c = 1; c < cols - 1; c++) {
      int a = r * cols + c, b = a - cols;
      float d1[3], d2[3];
      if (orcf_face(h[a], h[b], qx[a], qx[b], qy[a], qy[b], dem[a], dem[b], sqrth[a], sqrth[b], 0.0f, 1.0f, dx,
                    dt, hextra, d1, d2)) {
        rhsh0[a] += dt * d1[0] / dx; rhsqx0[a] += dt * d1[1] / dx; rhsqy0[a] += dt * d1[2] / dx;
        rhsh1[b] += dt * d2[0] / dx; rhsqx1[b] += dt * d2[1] / dx; rhsqy1[b] += dt * d2[2] / dx;
      }
    }
}

/* src/kernels_state.h:30-104 */
void orcf_update_cells(int rows, int cols, float dt, float *h, float *qx, float *qy, const float *nman,
                       const float *rhsh0, const float *rhsh1, const float *rhsqx0, const float *rhsqx1,
                       const float *rhsqy0, const float *rhsqy1, float hextra) {
#pragma omp parallel for schedule(static)
  for (int r = 1; r < rows - 1; r++)
    for (int c = 1; c < cols - 1; c++) {
      int id = r * cols + c;
      float hij, qxij, qyij;
      float hn = h[id];
      hij = hn - rhsh0[id] - rhsh1[id];
      if (hij < 0.0) hij = 0.0;
      if (hij < hextra) {
        qxij = 0.0;
        qyij = 0.0;
      } else {
        float mx = qx[id] - rhsqx0[id] - rhsqx1[id];
        float my = qy[id] - rhsqy0[id] - rhsqy1[id];
        float modM = sqrt(mx * mx + my * my);
        if (nman[id] > EPS12F && hn >= hextra && modM > EPS12F) {
          float tt = dt * G_ACC * nman[id] * modM / (hn * hn * cbrt(hn));
          qxij = -0.5 * (mx - mx * sqrt(1.0 + 4.0 * tt)) / tt;
          qyij = -0.5 * (my - my * sqrt(1.0 + 4.0 * tt)) / tt;
        } else {
          qxij = mx;
          qyij = my;
        }
      }
      h[id] = hij;
      qx[id] = qxij;
      qy[id] = qyij;
    }
}

/* src/kernels.h:102-152 */
void orcf_update_runoff(int rows, int cols, float dt, const int *runoff_id, int row, int nrow, const float *rate,
                        float *h, float *qx, float *qy, float hextra) {
  for (int r = 1; r < rows - 1; r++)
    for (int c = 1; c < cols - 1; c++) {
      int id = r * cols + c;
      float hij = h[id];
      float net = rate[runoff_id[id] * nrow + row];
      hij += net * dt;
      if (hij < hextra) {
        if (hij < 0.0) hij = 0.0;
        qx[id] = 0.0;
        qy[id] = 0.0;
      }
      h[id] = hij;
    }
}

/* src/kernels.h:50-84 (sources applied in list order) */
void orcf_compute_flow(int num_src, const float *hyg_time, const float *hyg_val, float dx, float dt,
                       float simtime, int idx_low, int idx_high, float *h, const int *src_pos) {
  for (int s = 0; s < num_src; s++) {
    float flo = hyg_val[idx_low * num_src + s], fhi = hyg_val[idx_high * num_src + s];
    float tlo = hyg_time[idx_low], thi = hyg_time[idx_high];
    float flow;
    if (idx_low == idx_high) {
      flow = flo;
    } else {
      float td = simtime - tlo;
      float td2 = thi - tlo;
      flow = flo + (((fhi - flo) * td) / td2);
    }
    float hij = h[src_pos[s]];
    float h_src = (flow * dt) / (dx * dx);
    hij += h_src;
    h[src_pos[s]] = hij;
  }
}

/* src/kernels.h:436-460 */
void orcf_open_bc_west_east(int rows, int cols, float *h, float *qx, float *qy) {
  for (int r = 0; r < rows; r++) {
    int w = r * cols, e = r * cols + cols - 1;
    h[w] = h[w + 1]; qx[w] = qx[w + 1]; qy[w] = qy[w + 1];
    h[e] = h[e - 1]; qx[e] = qx[e - 1]; qy[e] = qy[e - 1];
  }
}

/* src/kernels.h:476-512 */
void orcf_open_bc_north_south(int rows, int cols, float *h, float *qx, float *qy) {
  for (int c = 0; c < cols; c++) {
    int n = c, s = (rows - 1) * cols + c;
    h[n] = h[n + cols]; qx[n] = qx[n + cols]; qy[n] = qy[n + cols];
    h[s] = h[s - cols]; qx[s] = qx[s - cols]; qy[s] = qy[s - cols];
  }
}

/* src/kernels.h:236-424 (one rank, cells in list order) */
void orcf_compute_extbc_values(int ncells, int rows, int cols, float dt, float *h, float *qx, float *qy,
                               const float *dem, const float *nman, const int *idx, const int *type,
                               const int *start, const int *nvars, const float *v1, const float *v2,
                               float simtime) {
  for (int k = 0; k < ncells; k++) {
    int ii = idx[k], bt = type[k], st = start[k];
    int r = ii / cols, c = ii % cols;
    int top = r == 1, btm = r == rows - 2, lt = c == 1, rt = c == cols - 2;
    if (!(top || lt || rt || btm)) continue;
    int ghost = ii, corner = -1;
    float f[3] = {1.0f, 0.0f, 0.0f};
    if (lt) { ghost = ii - 1; f[1] = -1.0; f[2] = 0.0; }
    if (rt) { ghost = ii + 1; f[1] = 1.0; f[2] = 0.0; }
    if (top) {
      ghost = ii - cols; f[1] = 0.0; f[2] = 1.0;
      if (lt) { corner = ii - 1; f[2] = 0.0; }
      if (rt) { corner = ii + 1; f[2] = 0.0; }
    }
    if (btm) {
      ghost = ii + cols; f[1] = 0.0; f[2] = -1.0;
      if (lt) { corner = ii - 1; f[2] = 0.0; }
      if (rt) { corner = ii + 1; f[2] = 0.0; }
    }
    float hij = h[ii], qxij = qx[ii], qyij = qy[ii];
    if (bt == 0) {
      h[ghost] = hij; qx[ghost] = qxij; qy[ghost] = qyij;
      if (corner >= 0) { h[corner] = hij; qx[corner] = qxij; qy[corner] = qyij; }
      continue;
    }
    float aux;
    if (bt == 1) {
      float lvar = 0.5 * (simtime + simtime + dt);
      int lo = -999, hi = nvars[k] - 1;
      if (lvar < v1[st]) { lo = 0; hi = 0; }
      else if (lvar > v1[hi + st]) { lo = hi; }
      else {
        for (int i = 0; i < hi; i++)
          if (v1[i + st] <= lvar && v1[i + 1 + st] > lvar) { lo = i; hi = i + 1; break; }
        if (lo < 0) lo = hi; /* tmid at the last time: the last row (as the FP64 oracle) */
      }
      if (lo == hi) aux = v2[lo + st];
      else {
        float td = lvar - v1[lo + st];
        float td2 = v1[hi + st] - v1[lo + st];
        aux = v2[lo + st] + (((v2[hi + st] - v2[lo + st]) * td) / td2);
      }
      hij = fmaxf(aux - dem[ii], 0.0);
    } else {
      aux = v1[st];
      float vel = 0.0f;
      if (bt == 2) vel = sqrt(aux * hij * cbrt(hij) / fmaxf(nman[ii], 1e-6));
      if (bt == 3) vel = aux * sqrt(G_ACC * hij);
      qxij = hij * vel;
      qyij = hij * vel;
    }
    h[ghost] = hij * f[0]; qx[ghost] = qxij * f[1]; qy[ghost] = qyij * f[2];
    h[ii] = hij * f[0]; qx[ii] = qxij * f[1]; qy[ii] = qyij * f[2];
    if (corner >= 0) { h[corner] = hij * f[0]; qx[corner] = qxij * f[1]; qy[corner] = qyij * f[2]; }
  }
}

/* src/kernels_wetdry.h:24-85 (one rank; maxh may be NULL) */
void orcf_wet_dry(int rows, int cols, const float *h, float *qx, float *qy, const float *dem, float *maxh,
                  float hextra) {
  for (int r = 1; r < rows - 1; r++)
    for (int c = 1; c < cols - 1; c++) {
      int id = r * cols + c;
      float hij = h[id], zij = dem[id];
      if (hij > hextra) {
        if (((hij + zij < dem[id + 1]) && (h[id + 1] < EPS12F)) || ((hij + zij < dem[id - 1]) && (h[id - 1] < EPS12F)))
          qx[id] = 0.0;
        if (((hij + zij < dem[id - cols]) && (h[id - cols] < EPS12F)) ||
            ((hij + zij < dem[id + cols]) && (h[id + cols] < EPS12F)))
          qy[id] = 0.0;
      }
      if (maxh && hij > maxh[id]) maxh[id] = hij;
    }
}

/* The step loop of the single-precision build without forcing
 * (src/triton_simulate.h:54-134, :191-346): dt, simulated time and print
 * clamping in float; walls (bed ghosts 1e6) or open boundaries.  Returns
 * the number of steps; simtime_out receives the time reached.  Scratch:
 * sqrth, dtarr and six RHS arrays of rows*cols floats in `work`. */
long long orcf_run(int rows, int cols, float dx, float cn, float hextra, int open_bc, float *h, float *qx, float *qy,
                   const float *dem, const float *nman, float t_end, float print_interval, float init_dt,
                   long long max_steps, float *work, float *simtime_out) {
  const int n = rows * cols;
  float *sqrth = work, *dtarr = work + n, *r0 = work + 2 * n, *r1 = work + 3 * n, *rx0 = work + 4 * n,
        *rx1 = work + 5 * n, *ry0 = work + 6 * n, *ry1 = work + 7 * n;
  memset(work, 0, sizeof(float) * (size_t)n * 8);
  float simtime = 0.0f;
  int print_id = 0;
  long long steps = 0;
  while (simtime < t_end && (max_steps < 0 || steps < max_steps)) {
    orcf_compute_dt_and_sqrt(n, dx, qx, qy, h, sqrth, dtarr, cn, hextra);
    float dt = orcf_find_min_dt(n, dtarr);
    if (dt >= MAX_VALUE - 1) dt = init_dt;
    if (simtime + dt > print_interval * (print_id + 1)) dt = print_interval * (print_id + 1) - simtime;
    orcf_flux_x(rows, cols, dx, dt, h, qx, qy, dem, sqrth, r0, r1, rx0, rx1, ry0, ry1, hextra);
    orcf_flux_y(rows, cols, dx, dt, h, qx, qy, dem, sqrth, r0, r1, rx0, rx1, ry0, ry1, hextra);
    orcf_update_cells(rows, cols, dt, h, qx, qy, nman, r0, r1, rx0, rx1, ry0, ry1, hextra);
    if (open_bc) {
      orcf_open_bc_west_east(rows, cols, h, qx, qy);
      orcf_open_bc_north_south(rows, cols, h, qx, qy);
    }
    orcf_wet_dry(rows, cols, h, qx, qy, dem, NULL, hextra);
    simtime += dt;
    steps++;
    if (simtime >= print_interval * (print_id + 1)) print_id++;
  }
  *simtime_out = simtime;
  return steps;
}

// kernels_ref.hip -- drop-in replacements for the reference's per-kernel
// entry points (Kernels::* in src/kernels*.h), one thread per launch item
// with the reference's launch range and active sub-range, so a host that
// replays compute_new_state (src/triton_simulate.h:230-346) with these gets
// the CPU results bit for bit.  The kernels are templates on the state type,
// as the reference's are: tr_<name>_f64 instantiate them for double,
// tr_<name>_f32 for float with the reference's single-precision promotions
// (roe_f64.hpp).  Built with -ffp-contract=off.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "roe_f64.hpp"

namespace tr {
namespace {

template <class T>
__global__ void k_compute_dt_and_sqrt(int size, T cn_dx, const T* __restrict__ qx, const T* __restrict__ qy,
                                      const T* __restrict__ h, T* __restrict__ sqrth, T* __restrict__ out,
                                      T hextra) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= size) return;
  T hc = h[id];
  T sq = tmax<T>(T(sqrt(double(hc))), T(0));
  sqrth[id] = sq;
  out[id] = cell_dt_t<T>(hc, qx[id], qy[id], sq, cn_dx, hextra);
}

// Total-order uint64 key of a non-NaN double (negative values flip all bits,
// non-negative ones set the sign bit) and its inverse.
__device__ __forceinline__ unsigned long long order_key(double v) {
  unsigned long long u = dbits(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_to_double(unsigned long long k) {
  return bitsd((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k);
}
// the same for floats, 32-bit keys
__device__ __forceinline__ unsigned order_key32(float v) {
  unsigned u = __float_as_uint(v);
  return (u >> 31) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_to_float(unsigned k) {
  return __uint_as_float((k >> 31) ? (k & 0x7fffffffu) : ~k);
}

// Wave64 min of uint64 keys, one atomic per wave.
__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ unsigned wave_min_u32(unsigned v) {
  for (int off = 32; off > 0; off >>= 1) {
    unsigned o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

__global__ void k_init_min(unsigned long long* slot) { *slot = ~0ull; }
__global__ void k_init_min32(unsigned* slot) { *slot = ~0u; }

__global__ void k_min_reduce(int size, const double* __restrict__ in,
                             unsigned long long* __restrict__ slot) {
  unsigned long long best = ~0ull;
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < size; id += gridDim.x * blockDim.x) {
    double v = in[id];
    // order-preserving uint64 key of a double; NaN never wins, as with the
    // reference's `a < b ? a : b` form
    unsigned long long b = (v == v) ? order_key(v) : ~0ull;
    best = b < best ? b : best;
  }
  best = wave_min_u64(best);
  if ((threadIdx.x & 63) == 0) atomicMin(slot, best);
}
__global__ void k_min_reduce32(int size, const float* __restrict__ in, unsigned* __restrict__ slot) {
  unsigned best = ~0u;
  for (int id = blockIdx.x * blockDim.x + threadIdx.x; id < size; id += gridDim.x * blockDim.x) {
    float v = in[id];
    unsigned b = (v == v) ? order_key32(v) : ~0u;
    best = b < best ? b : best;
  }
  best = wave_min_u32(best);
  if ((threadIdx.x & 63) == 0) atomicMin(slot, best);
}

__global__ void k_bits_to_double(const unsigned long long* slot, double* out) {
  unsigned long long b = *slot;
  *out = (b == ~0ull) ? INFINITY : key_to_double(b);
}
__global__ void k_bits_to_float(const unsigned* slot, float* out) {
  unsigned b = *slot;
  *out = (b == ~0u) ? INFINITY : key_to_float(b);
}

template <class T>
__global__ void k_flux_x(int size, int nrows, int ncols, T dx, T dt, const T* __restrict__ h,
                         const T* __restrict__ qx, const T* __restrict__ qy, const T* __restrict__ dem,
                         const T* __restrict__ sqrth, T* __restrict__ rhsh0, T* __restrict__ rhsh1,
                         T* __restrict__ rhsqx0, T* __restrict__ rhsqx1, T* __restrict__ rhsqy0,
                         T* __restrict__ rhsqy1, T hextra) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= size) return;
  int r = id / ncols, c = id % ncols;
  if (r < 1 || r >= nrows - 1 || c >= ncols - 1) return;  // src/kernels_flux.h:62-71
  int a = id, b = id + 1;
  FaceT<T> f;
  if (roe_face_ref<T, 1, 0>(h[a], h[b], qx[a], qx[b], qy[a], qy[b], dem[a], dem[b], sqrth[a], sqrth[b], dx,
                            dt, hextra, f)) {
    rhsh0[a] = dt * f.d1[0] / dx;
    rhsqx0[a] = dt * f.d1[1] / dx;
    rhsqy0[a] = dt * f.d1[2] / dx;
    rhsh1[b] = dt * f.d2[0] / dx;
    rhsqx1[b] = dt * f.d2[1] / dx;
    rhsqy1[b] = dt * f.d2[2] / dx;
  } else {
    rhsh0[a] = T(0); rhsqx0[a] = T(0); rhsqy0[a] = T(0);
    rhsh1[b] = T(0); rhsqx1[b] = T(0); rhsqy1[b] = T(0);
  }
}

template <class T>
__global__ void k_flux_y(int size, int nrows, int ncols, T dx, T dt, const T* __restrict__ h,
                         const T* __restrict__ qx, const T* __restrict__ qy, const T* __restrict__ dem,
                         const T* __restrict__ sqrth, T* __restrict__ rhsh0, T* __restrict__ rhsh1,
                         T* __restrict__ rhsqx0, T* __restrict__ rhsqx1, T* __restrict__ rhsqy0,
                         T* __restrict__ rhsqy1, T hextra) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= size) return;
  int r = id / ncols, c = id % ncols;
  if (r < 1 || c < 1 || c >= ncols - 1) return;  // src/kernels_flux.h:342-354
  int a = id, b = id - ncols;
  FaceT<T> f;
  if (roe_face_ref<T, 0, 1>(h[a], h[b], qx[a], qx[b], qy[a], qy[b], dem[a], dem[b], sqrth[a], sqrth[b], dx,
                            dt, hextra, f)) {
    rhsh0[a] += dt * f.d1[0] / dx;
    rhsqx0[a] += dt * f.d1[1] / dx;
    rhsqy0[a] += dt * f.d1[2] / dx;
    rhsh1[b] += dt * f.d2[0] / dx;
    rhsqx1[b] += dt * f.d2[1] / dx;
    rhsqy1[b] += dt * f.d2[2] / dx;
  }
}

template <class T>
__global__ void k_update_cells(int size, int nrows, int ncols, T dt, T* __restrict__ h, T* __restrict__ qx,
                               T* __restrict__ qy, const T* __restrict__ nman, const T* __restrict__ rhsh0,
                               const T* __restrict__ rhsh1, const T* __restrict__ rhsqx0,
                               const T* __restrict__ rhsqx1, const T* __restrict__ rhsqy0,
                               const T* __restrict__ rhsqy1, T hextra) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= size) return;
  int r = id / ncols, c = id % ncols;
  if (r < 1 || r >= nrows - 1 || c < 1 || c >= ncols - 1) return;
  T ho, qxo, qyo;
  update_cell_t<T>(h[id], qx[id], qy[id], nman[id], rhsh0[id], rhsh1[id], rhsqx0[id], rhsqx1[id], rhsqy0[id],
                   rhsqy1[id], dt, hextra, ho, qxo, qyo);
  h[id] = ho;
  qx[id] = qxo;
  qy[id] = qyo;
}

template <class T>
__global__ void k_update_runoff(int size, int nrows, int ncols, T dt, const int* __restrict__ rid, int row,
                                int nrow, const T* __restrict__ rate, T* __restrict__ h, T* __restrict__ qx,
                                T* __restrict__ qy, T hextra) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= size) return;
  int r = id / ncols, c = id % ncols;
  if (r < 1 || r >= nrows - 1 || c < 1 || c >= ncols - 1) return;
  T hc = h[id];
  hc += rate[rid[id] * nrow + row] * dt;
  if (hc < hextra) {
    if (hc < 0.0) hc = T(0);
    qx[id] = T(0);
    qy[id] = T(0);
  }
  h[id] = hc;
}

// One thread walks the sources in order: duplicates accumulate
// deterministically where the reference's parallel update would race.
template <class T>
__global__ void k_compute_flow(int nsrc, const T* __restrict__ ht, const T* __restrict__ hv, T dx, T dt,
                               T simtime, int lo, int hi, T* __restrict__ h, const int* __restrict__ pos) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int s = 0; s < nsrc; s++) {
    T flo = hv[lo * nsrc + s], fhi = hv[hi * nsrc + s];
    T q;
    if (lo == hi) q = flo;
    else q = flo + (((fhi - flo) * (simtime - ht[lo])) / (ht[hi] - ht[lo]));
    int cell = pos[s];
    T h_src = (q * dt) / (dx * dx);
    h[cell] = h[cell] + h_src;
  }
}

template <class T>
__global__ void k_open_we(int size, int nrows, int ncols, T* __restrict__ h, T* __restrict__ qx,
                          T* __restrict__ qy) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= size) return;
  int g, in;
  if (id < nrows) { g = (id % nrows) * ncols; in = g + 1; }
  else { g = (id % nrows) * ncols + ncols - 1; in = g - 1; }
  h[g] = h[in]; qx[g] = qx[in]; qy[g] = qy[in];
}

template <class T>
__global__ void k_open_ns(int size, int nrows, int ncols, T* __restrict__ h, T* __restrict__ qx,
                          T* __restrict__ qy, int rank, int nproc) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= size) return;
  if (rank > 0 && rank < nproc - 1) return;
  if (rank == 0) {
    int g = id, in = (id % ncols) + ncols;
    h[g] = h[in]; qx[g] = qx[in]; qy[g] = qy[in];
  }
  if (rank == nproc - 1) {
    int g = id + (nrows - 1) * ncols, in = (id % ncols) + (nrows - 2) * ncols;
    h[g] = h[in]; qx[g] = qx[in]; qy[g] = qy[in];
  }
}

// External boundary cells; one thread walks the list in order so a cell
// listed by two segments gets the later segment's values deterministically.
template <class T>
__global__ void k_extbc(int ncells, int nrows, int ncols, T dt, T* __restrict__ h, T* __restrict__ qx,
                        T* __restrict__ qy, const T* __restrict__ dem, const T* __restrict__ nman,
                        const int* __restrict__ idx, const int* __restrict__ type, const int* __restrict__ start,
                        const int* __restrict__ nvars, const T* __restrict__ v1, const T* __restrict__ v2,
                        T simtime, int rank, int nproc) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  for (int k = 0; k < ncells; k++) {
    int ii = idx[k], bt = type[k], st = start[k];
    int r = ii / ncols, c = ii % ncols;
    bool top = r == 1 && rank == 0, btm = r == nrows - 2 && rank == nproc - 1;
    bool lt = c == 1, rt = c == ncols - 2;
    if (!(top || lt || rt || btm)) continue;
    int ghost = ii, corner = -1;
    T f0 = T(1), f1 = T(0), f2 = T(0);
    if (lt) { ghost = ii - 1; f1 = T(-1); f2 = T(0); }
    if (rt) { ghost = ii + 1; f1 = T(1); f2 = T(0); }
    if (top) {
      ghost = ii - ncols; f1 = T(0); f2 = T(1);
      if (lt) { corner = ii - 1; f2 = T(0); }
      if (rt) { corner = ii + 1; f2 = T(0); }
    }
    if (btm) {
      ghost = ii + ncols; f1 = T(0); f2 = T(-1);
      if (lt) { corner = ii - 1; f2 = T(0); }
      if (rt) { corner = ii + 1; f2 = T(0); }
    }
    T hc = h[ii], qxc = qx[ii], qyc = qy[ii];
    if (bt == 0) {
      h[ghost] = hc; qx[ghost] = qxc; qy[ghost] = qyc;
      if (corner >= 0) { h[corner] = hc; qx[corner] = qxc; qy[corner] = qyc; }
      continue;
    }
    if (bt == 1) {
      T tm = 0.5 * (simtime + simtime + dt);
      int lo = -999, hi = nvars[k] - 1;
      if (tm < v1[st]) { lo = 0; hi = 0; }
      else if (tm > v1[hi + st]) { lo = hi; }
      else {
        for (int i = 0; i < hi; i++)
          if (v1[i + st] <= tm && v1[i + 1 + st] > tm) { lo = i; hi = i + 1; break; }
        if (lo < 0) lo = hi;
      }
      T level;
      if (lo == hi) level = v2[lo + st];
      else level = v2[lo + st] + (((v2[hi + st] - v2[lo + st]) * (tm - v1[lo + st])) / (v1[hi + st] - v1[lo + st]));
      hc = tmax<T>(level - dem[ii], T(0));
    } else {
      T p = v1[st], vel = T(0);
      if (bt == 2) vel = sqrt(p * hc * cbrt(double(hc)) / tmax<T>(nman[ii], T(1e-6)));
      if (bt == 3) vel = p * sqrt(kG * hc);
      qxc = hc * vel;
      qyc = hc * vel;
    }
    h[ghost] = hc * f0; qx[ghost] = qxc * f1; qy[ghost] = qyc * f2;
    h[ii] = hc * f0; qx[ii] = qxc * f1; qy[ii] = qyc * f2;
    if (corner >= 0) { h[corner] = hc * f0; qx[corner] = qxc * f1; qy[corner] = qyc * f2; }
  }
}

template <class T>
__global__ void k_wet_dry(int size, int nrows, int ncols, const T* __restrict__ h, T* __restrict__ qx,
                          T* __restrict__ qy, const T* __restrict__ dem, T* __restrict__ maxh, T hextra,
                          int mpi_tasks) {
  constexpr double eps = eps12_of<T>();
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= size) return;
  int r = id / ncols, c = id % ncols;
  if (r < 1 || r >= nrows - 1 || c < 1 || c >= ncols - 1) return;
  T hc = h[id], zc = dem[id];
  if (hc > hextra) {
    int e = id + 1, w = id - 1, n = id - ncols, s = id + ncols;
    if ((hc + zc < dem[e] && h[e] < eps) || (hc + zc < dem[w] && h[w] < eps)) qx[id] = T(0);
    bool zero_y;
    if (r == 1 && mpi_tasks > 1) zero_y = hc + zc < dem[s] && h[s] < eps;
    else if (r == nrows - 2 && mpi_tasks > 1) zero_y = hc + zc < dem[n] && h[n] < eps;
    else zero_y = (hc + zc < dem[n] && h[n] < eps) || (hc + zc < dem[s] && h[s] < eps);
    if (zero_y) qy[id] = T(0);
  }
  if (maxh && hc > maxh[id]) maxh[id] = hc;
}

// Observation points (src/kernels.h:521-535): obs_arr[i] = arr[index[i]].
__global__ void k_copy_obs(int size, const double* __restrict__ arr, double* __restrict__ obs,
                           const int* __restrict__ index) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < size) obs[id] = arr[index[id]];
}

template <class T>
__global__ void k_copy_obs_t(int size, const T* __restrict__ arr, T* __restrict__ obs, const int* __restrict__ index) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < size) obs[id] = arr[index[id]];
}

__global__ void k_copy_obs_f32(int size, const float* __restrict__ arr, double* __restrict__ obs,
                               const int* __restrict__ index) {
  int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < size) obs[id] = double(arr[index[id]]);
}

}  // namespace

// Gathers of observation cells shared by the simulation and batch drivers.
int copy_obs_f64(int size, const double* arr, double* obs, const int* index, hipStream_t s) {
  if (size <= 0) return TR_OK;
  hipLaunchKernelGGL(k_copy_obs, dim3(grid_for(size)), dim3(kBlock), 0, s, size, arr, obs, index);
  return launch_check("copy_observation_points");
}

int copy_obs_f32(int size, const float* arr, double* obs, const int* index, hipStream_t s) {
  if (size <= 0) return TR_OK;
  hipLaunchKernelGGL(k_copy_obs_f32, dim3(grid_for(size)), dim3(kBlock), 0, s, size, arr, obs,
                     index);
  return launch_check("copy_observation_points_f32");
}

// Shared by the unfused simulation driver: min of `size` dt values into the
// device slot (uint64 image), optionally also as a double.
int min_dt_to_slot(int size, const double* dtarr, unsigned long long* slot, hipStream_t s) {
  hipLaunchKernelGGL(k_init_min, dim3(1), dim3(1), 0, s, slot);
  unsigned g = grid_for(size);
  if (g > 4096) g = 4096;
  if (g == 0) g = 1;
  hipLaunchKernelGGL(k_min_reduce, dim3(g), dim3(kBlock), 0, s, size, dtarr, slot);
  return launch_check("find_min_dt");
}

// The device cube root the FP64 friction takes of the depth
// (update_cell_f64, src/kernels_state.h:88), on its own.
__global__ void k_cbrt_f64(int n, const double* __restrict__ x, double* __restrict__ y) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id < n) y[id] = cbrt(x[id]);
}

}  // namespace tr

using namespace tr;

extern "C" {

int tr_cbrt_f64(int32_t n, const double* x, double* y, void* stream) {
  if (n < 0 || (n > 0 && (!x || !y))) {
    set_error("tr_cbrt_f64: bad arguments");
    return TR_ERR_INVALID;
  }
  if (n == 0) return TR_OK;
  hipLaunchKernelGGL(k_cbrt_f64, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream), n, x, y);
  return launch_check("cbrt_f64");
}

}  // extern "C"

namespace {

// One implementation of every drop-in entry point for both state types.
template <class T>
int compute_dt_and_sqrt_t(const char* fn, int32_t size, T dx, const T* qx, const T* qy, const T* h, T* sqrth,
                          T* dt_out, T cn, T hextra, void* stream) {
  if (size < 0 || (size > 0 && (!qx || !qy || !h || !sqrth || !dt_out))) {
    set_error(std::string(fn) + ": bad arguments");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  const T cn_dx = cn * dx;  // the reference's cn*dx, in T
  hipLaunchKernelGGL(k_compute_dt_and_sqrt<T>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size,
                     cn_dx, qx, qy, h, sqrth, dt_out, hextra);
  return launch_check(fn);
}

template <class T>
int flux_t(const char* fn, bool y, int32_t size, int32_t nrows, int32_t ncols, T dx, T dt, const T* h,
           const T* qx, const T* qy, const T* dem, const T* sqrth, T* rhsh0, T* rhsh1, T* rhsqx0, T* rhsqx1,
           T* rhsqy0, T* rhsqy1, T hextra, void* stream) {
  if (size < 0 || nrows < 3 || ncols < 3 || int64_t(nrows) * ncols < size) {
    set_error(std::string(fn) + ": size must cover at most nrows*ncols with nrows,ncols >= 3");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  if (y)
    hipLaunchKernelGGL(k_flux_y<T>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size, nrows,
                       ncols, dx, dt, h, qx, qy, dem, sqrth, rhsh0, rhsh1, rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra);
  else
    hipLaunchKernelGGL(k_flux_x<T>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size, nrows,
                       ncols, dx, dt, h, qx, qy, dem, sqrth, rhsh0, rhsh1, rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra);
  return launch_check(fn);
}

template <class T>
int update_cells_t(const char* fn, int32_t size, int32_t nrows, int32_t ncols, T dt, T* h, T* qx, T* qy,
                   const T* nman, T* rhsh0, T* rhsh1, T* rhsqx0, T* rhsqx1, T* rhsqy0, T* rhsqy1, T hextra,
                   void* stream) {
  if (size < 0 || nrows < 3 || ncols < 3 || int64_t(nrows) * ncols < size) {
    set_error(std::string(fn) + ": size must cover at most nrows*ncols with nrows,ncols >= 3");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  hipLaunchKernelGGL(k_update_cells<T>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size, nrows,
                     ncols, dt, h, qx, qy, nman, rhsh0, rhsh1, rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra);
  return launch_check(fn);
}

template <class T>
int update_runoff_t(const char* fn, int32_t size, int32_t nrows, int32_t ncols, T dt, const int32_t* runoff_id,
                    int32_t row, int32_t nrow, const T* rate, T* h, T* qx, T* qy, T hextra, void* stream) {
  if (size < 0 || nrows < 3 || ncols < 3 || int64_t(nrows) * ncols < size) {
    set_error(std::string(fn) + ": size must cover at most nrows*ncols with nrows,ncols >= 3");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  hipLaunchKernelGGL(k_update_runoff<T>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size, nrows,
                     ncols, dt, runoff_id, row, nrow, rate, h, qx, qy, hextra);
  return launch_check(fn);
}

template <class T>
int compute_flow_t(const char* fn, int32_t size, const T* ht, const T* hv, T dx, T dt, T simtime, int32_t lo,
                   int32_t hi, T* h, const int32_t* pos, void* stream) {
  if (size < 0 || lo < 0 || hi < 0) {
    set_error(std::string(fn) + ": bad arguments");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  hipLaunchKernelGGL(k_compute_flow<T>, dim3(1), dim3(64), 0, as_stream(stream), size, ht, hv, dx, dt, simtime,
                     lo, hi, h, pos);
  return launch_check(fn);
}

template <class T>
int open_we_t(const char* fn, int32_t size, int32_t nrows, int32_t ncols, T* h, T* qx, T* qy, void* stream) {
  if (size < 0 || size > 2 * nrows || nrows < 3 || ncols < 3) {
    set_error(std::string(fn) + ": size must be at most 2*nrows");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  hipLaunchKernelGGL(k_open_we<T>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size, nrows, ncols,
                     h, qx, qy);
  return launch_check(fn);
}

template <class T>
int open_ns_t(const char* fn, int32_t size, int32_t nrows, int32_t ncols, T* h, T* qx, T* qy, int32_t rank,
              int32_t nproc, void* stream) {
  if (size < 0 || size > ncols || nrows < 3 || ncols < 3 || nproc < 1) {
    set_error(std::string(fn) + ": size must be at most ncols");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  hipLaunchKernelGGL(k_open_ns<T>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size, nrows, ncols,
                     h, qx, qy, rank, nproc);
  return launch_check(fn);
}

template <class T>
int extbc_t(const char* fn, int32_t size, int32_t nrows, int32_t ncols, T dt, T* h, T* qx, T* qy, const T* dem,
            const T* nman, const int32_t* idx, const int32_t* type, const int32_t* start, const int32_t* nvars,
            const T* v1, const T* v2, T simtime, int32_t rank, int32_t nproc, void* stream) {
  if (size < 0 || nrows < 3 || ncols < 3) {
    set_error(std::string(fn) + ": bad arguments");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  hipLaunchKernelGGL(k_extbc<T>, dim3(1), dim3(64), 0, as_stream(stream), size, nrows, ncols, dt, h, qx, qy, dem,
                     nman, idx, type, start, nvars, v1, v2, simtime, rank, nproc);
  return launch_check(fn);
}

template <class T>
int wet_dry_t(const char* fn, int32_t size, int32_t nrows, int32_t ncols, const T* h, T* qx, T* qy,
              const T* dem, T* maxh, T hextra, int32_t mpi_tasks, void* stream) {
  if (size < 0 || nrows < 3 || ncols < 3 || int64_t(nrows) * ncols < size) {
    set_error(std::string(fn) + ": size must cover at most nrows*ncols with nrows,ncols >= 3");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  hipLaunchKernelGGL(k_wet_dry<T>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size, nrows, ncols,
                     h, qx, qy, dem, maxh, hextra, mpi_tasks);
  return launch_check(fn);
}

}  // namespace

extern "C" {

int tr_compute_dt_and_sqrt_f64(int32_t size, double dx, const double* qx, const double* qy,
                               const double* h, double* sqrth, double* dt_out, double cn,
                               double hextra, void* stream) {
  return compute_dt_and_sqrt_t<double>("tr_compute_dt_and_sqrt_f64", size, dx, qx, qy, h, sqrth, dt_out, cn,
                                       hextra, stream);
}
int tr_compute_dt_and_sqrt_f32(int32_t size, float dx, const float* qx, const float* qy, const float* h,
                               float* sqrth, float* dt_out, float cn, float hextra, void* stream) {
  return compute_dt_and_sqrt_t<float>("tr_compute_dt_and_sqrt_f32", size, dx, qx, qy, h, sqrth, dt_out, cn,
                                      hextra, stream);
}

int tr_find_min_dt_f64(int32_t size, const double* dtarr, double* d_min, double* h_min,
                       void* stream) {
  if (size <= 0 || !dtarr || !d_min) {
    set_error("tr_find_min_dt_f64: bad arguments");
    return TR_ERR_INVALID;
  }
  hipStream_t s = as_stream(stream);
  // the uint64 reduction slot lives in the 8 bytes of d_min itself
  auto* slot = reinterpret_cast<unsigned long long*>(d_min);
  int rc = min_dt_to_slot(size, dtarr, slot, s);
  if (rc) return rc;
  hipLaunchKernelGGL(k_bits_to_double, dim3(1), dim3(1), 0, s, slot, d_min);
  rc = launch_check("find_min_dt");
  if (rc) return rc;
  if (h_min) {
    TR_HIP(hipMemcpyAsync(h_min, d_min, sizeof(double), hipMemcpyDeviceToHost, s));
    TR_HIP(hipStreamSynchronize(s));
  }
  return TR_OK;
}
int tr_find_min_dt_f32(int32_t size, const float* dtarr, float* d_min, float* h_min, void* stream) {
  if (size <= 0 || !dtarr || !d_min) {
    set_error("tr_find_min_dt_f32: bad arguments");
    return TR_ERR_INVALID;
  }
  hipStream_t s = as_stream(stream);
  // the uint32 reduction slot lives in the 4 bytes of d_min itself
  auto* slot = reinterpret_cast<unsigned*>(d_min);
  unsigned g = grid_for(size);
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_init_min32, dim3(1), dim3(1), 0, s, slot);
  hipLaunchKernelGGL(k_min_reduce32, dim3(g), dim3(kBlock), 0, s, size, dtarr, slot);
  hipLaunchKernelGGL(k_bits_to_float, dim3(1), dim3(1), 0, s, slot, d_min);
  int rc = launch_check("find_min_dt_f32");
  if (rc) return rc;
  if (h_min) {
    TR_HIP(hipMemcpyAsync(h_min, d_min, sizeof(float), hipMemcpyDeviceToHost, s));
    TR_HIP(hipStreamSynchronize(s));
  }
  return TR_OK;
}

int tr_flux_x_f64(int32_t size, int32_t nrows, int32_t ncols, double dx, double dt,
                  const double* h, const double* qx, const double* qy, const double* dem,
                  const double* sqrth, double* rhsh0, double* rhsh1, double* rhsqx0,
                  double* rhsqx1, double* rhsqy0, double* rhsqy1, double hextra, void* stream) {
  return flux_t<double>("tr_flux_x_f64", false, size, nrows, ncols, dx, dt, h, qx, qy, dem, sqrth, rhsh0, rhsh1,
                        rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra, stream);
}
int tr_flux_x_f32(int32_t size, int32_t nrows, int32_t ncols, float dx, float dt, const float* h, const float* qx,
                  const float* qy, const float* dem, const float* sqrth, float* rhsh0, float* rhsh1,
                  float* rhsqx0, float* rhsqx1, float* rhsqy0, float* rhsqy1, float hextra, void* stream) {
  return flux_t<float>("tr_flux_x_f32", false, size, nrows, ncols, dx, dt, h, qx, qy, dem, sqrth, rhsh0, rhsh1,
                       rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra, stream);
}

int tr_flux_y_f64(int32_t size, int32_t nrows, int32_t ncols, double dx, double dt,
                  const double* h, const double* qx, const double* qy, const double* dem,
                  const double* sqrth, double* rhsh0, double* rhsh1, double* rhsqx0,
                  double* rhsqx1, double* rhsqy0, double* rhsqy1, double hextra, void* stream) {
  return flux_t<double>("tr_flux_y_f64", true, size, nrows, ncols, dx, dt, h, qx, qy, dem, sqrth, rhsh0, rhsh1,
                        rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra, stream);
}
int tr_flux_y_f32(int32_t size, int32_t nrows, int32_t ncols, float dx, float dt, const float* h, const float* qx,
                  const float* qy, const float* dem, const float* sqrth, float* rhsh0, float* rhsh1,
                  float* rhsqx0, float* rhsqx1, float* rhsqy0, float* rhsqy1, float hextra, void* stream) {
  return flux_t<float>("tr_flux_y_f32", true, size, nrows, ncols, dx, dt, h, qx, qy, dem, sqrth, rhsh0, rhsh1,
                       rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra, stream);
}

int tr_update_cells_f64(int32_t size, int32_t nrows, int32_t ncols, double dt, double* h,
                        double* qx, double* qy, const double* dem, const double* nman,
                        double* rhsh0, double* rhsh1, double* rhsqx0, double* rhsqx1,
                        double* rhsqy0, double* rhsqy1, double hextra, void* stream) {
  (void)dem;  // part of the reference signature, unused there too
  return update_cells_t<double>("tr_update_cells_f64", size, nrows, ncols, dt, h, qx, qy, nman, rhsh0, rhsh1,
                                rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra, stream);
}
int tr_update_cells_f32(int32_t size, int32_t nrows, int32_t ncols, float dt, float* h, float* qx, float* qy,
                        const float* dem, const float* nman, float* rhsh0, float* rhsh1, float* rhsqx0,
                        float* rhsqx1, float* rhsqy0, float* rhsqy1, float hextra, void* stream) {
  (void)dem;
  return update_cells_t<float>("tr_update_cells_f32", size, nrows, ncols, dt, h, qx, qy, nman, rhsh0, rhsh1,
                               rhsqx0, rhsqx1, rhsqy0, rhsqy1, hextra, stream);
}

int tr_update_runoff_f64(int32_t size, int32_t nrows, int32_t ncols, double dt,
                         const int32_t* runoff_id, int32_t index_row_runoff,
                         int32_t n_rows_runoff, const double* runoff_intensity, double* h,
                         double* qx, double* qy, double hextra, void* stream) {
  return update_runoff_t<double>("tr_update_runoff_f64", size, nrows, ncols, dt, runoff_id, index_row_runoff,
                                 n_rows_runoff, runoff_intensity, h, qx, qy, hextra, stream);
}
int tr_update_runoff_f32(int32_t size, int32_t nrows, int32_t ncols, float dt, const int32_t* runoff_id,
                         int32_t index_row_runoff, int32_t n_rows_runoff, const float* runoff_intensity, float* h,
                         float* qx, float* qy, float hextra, void* stream) {
  return update_runoff_t<float>("tr_update_runoff_f32", size, nrows, ncols, dt, runoff_id, index_row_runoff,
                                n_rows_runoff, runoff_intensity, h, qx, qy, hextra, stream);
}

int tr_compute_flow_f64(int32_t size, const double* hyg_time, const double* hyg_val, double dx,
                        double dt, double simtime, int32_t idx_low, int32_t idx_high, double* h,
                        const int32_t* pos, void* stream) {
  return compute_flow_t<double>("tr_compute_flow_f64", size, hyg_time, hyg_val, dx, dt, simtime, idx_low,
                                idx_high, h, pos, stream);
}
int tr_compute_flow_f32(int32_t size, const float* hyg_time, const float* hyg_val, float dx, float dt,
                        float simtime, int32_t idx_low, int32_t idx_high, float* h, const int32_t* pos,
                        void* stream) {
  return compute_flow_t<float>("tr_compute_flow_f32", size, hyg_time, hyg_val, dx, dt, simtime, idx_low,
                               idx_high, h, pos, stream);
}

int tr_open_bc_west_east_f64(int32_t size, int32_t nrows, int32_t ncols, double* h, double* qx,
                             double* qy, void* stream) {
  return open_we_t<double>("tr_open_bc_west_east_f64", size, nrows, ncols, h, qx, qy, stream);
}
int tr_open_bc_west_east_f32(int32_t size, int32_t nrows, int32_t ncols, float* h, float* qx, float* qy,
                             void* stream) {
  return open_we_t<float>("tr_open_bc_west_east_f32", size, nrows, ncols, h, qx, qy, stream);
}

int tr_open_bc_north_south_f64(int32_t size, int32_t nrows, int32_t ncols, double* h,
                               double* qx, double* qy, int32_t rank, int32_t total_process,
                               void* stream) {
  return open_ns_t<double>("tr_open_bc_north_south_f64", size, nrows, ncols, h, qx, qy, rank, total_process,
                           stream);
}
int tr_open_bc_north_south_f32(int32_t size, int32_t nrows, int32_t ncols, float* h, float* qx, float* qy,
                               int32_t rank, int32_t total_process, void* stream) {
  return open_ns_t<float>("tr_open_bc_north_south_f32", size, nrows, ncols, h, qx, qy, rank, total_process,
                          stream);
}

int tr_compute_extbc_values_f64(int32_t size, int32_t nrows, int32_t ncols, double dt,
                                double* h, double* qx, double* qy, const double* dem,
                                const double* nman, const int32_t* relative_index,
                                const int32_t* extbctype, const int32_t* start_index,
                                const int32_t* nrows_vars, const double* extbcvar1,
                                const double* extbcvar2, double simtime, int32_t rank,
                                int32_t total_process, void* stream) {
  return extbc_t<double>("tr_compute_extbc_values_f64", size, nrows, ncols, dt, h, qx, qy, dem, nman,
                         relative_index, extbctype, start_index, nrows_vars, extbcvar1, extbcvar2, simtime, rank,
                         total_process, stream);
}
int tr_compute_extbc_values_f32(int32_t size, int32_t nrows, int32_t ncols, float dt, float* h, float* qx,
                                float* qy, const float* dem, const float* nman, const int32_t* relative_index,
                                const int32_t* extbctype, const int32_t* start_index, const int32_t* nrows_vars,
                                const float* extbcvar1, const float* extbcvar2, float simtime, int32_t rank,
                                int32_t total_process, void* stream) {
  return extbc_t<float>("tr_compute_extbc_values_f32", size, nrows, ncols, dt, h, qx, qy, dem, nman,
                        relative_index, extbctype, start_index, nrows_vars, extbcvar1, extbcvar2, simtime, rank,
                        total_process, stream);
}

int tr_wet_dry_f64(int32_t size, int32_t nrows, int32_t ncols, double dt, const double* h,
                   double* qx, double* qy, const double* dem, double* maxh, double hextra,
                   int32_t mpi_tasks, void* stream) {
  (void)dt;  // part of the reference signature, unused there too
  return wet_dry_t<double>("tr_wet_dry_f64", size, nrows, ncols, h, qx, qy, dem, maxh, hextra, mpi_tasks, stream);
}
int tr_wet_dry_f32(int32_t size, int32_t nrows, int32_t ncols, float dt, const float* h, float* qx, float* qy,
                   const float* dem, float* maxh, float hextra, int32_t mpi_tasks, void* stream) {
  (void)dt;
  return wet_dry_t<float>("tr_wet_dry_f32", size, nrows, ncols, h, qx, qy, dem, maxh, hextra, mpi_tasks, stream);
}

int tr_copy_observation_points_f32(int32_t size, const float* arr, float* obs_arr, const int32_t* relative_index,
                                   void* stream) {
  if (size < 0 || (size > 0 && (!arr || !obs_arr || !relative_index))) {
    set_error("tr_copy_observation_points_f32: bad arguments");
    return TR_ERR_INVALID;
  }
  if (size == 0) return TR_OK;
  hipLaunchKernelGGL(k_copy_obs_t<float>, dim3(grid_for(size)), dim3(kBlock), 0, as_stream(stream), size, arr,
                     obs_arr, relative_index);
  return launch_check("tr_copy_observation_points_f32");
}

int tr_copy_observation_points_f64(int32_t size, const double* arr, double* obs_arr,
                                   const int32_t* relative_index, void* stream) {
  if (size < 0 || (size > 0 && (!arr || !obs_arr || !relative_index))) {
    set_error("tr_copy_observation_points_f64: bad arguments");
    return TR_ERR_INVALID;
  }
  return copy_obs_f64(size, arr, obs_arr, relative_index, as_stream(stream));
}

}  // extern "C"

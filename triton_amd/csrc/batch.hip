// batch.hip -- members of a batched FP32 run (tr_batch_*).
//
// Every member owns its device buffers and a control block pair with a
// cursor; one launch per step covers all active members through a table of
// row ranges ("pieces") per wave, so members of uneven size advance
// together while each keeps its own adaptive dt, time and forcing.  Members
// can be added and removed between launches (the serving loop admits
// requests this way); admission is checked against an HBM budget.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "common.hpp"

namespace tr {
int copy_obs_f32(int size, const float* arr, double* obs, const int* index, hipStream_t s);
int check_member_view(const tr_case_view* v);
namespace {

#define RC(x)            \
  do {                   \
    int rc_ = (x);       \
    if (rc_) return rc_; \
  } while (0)

struct Member {
  bool used = false;
  bool active = false;   // in the work table
  InstF32 inst{};        // host image of the device descriptor
  std::vector<void*> allocs;
  int64_t bytes = 0;
  int64_t interior = 0;  // interior cells
  double sim_duration = 0;
};

__global__ void k_volume_f32(const float* __restrict__ h, int rows, int cols, double area,
                             double* __restrict__ part) {
  __shared__ double sm[256];
  const int n_int = (rows - 2) * (cols - 2);
  const int per = (n_int + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(lo + per, n_int);
  double acc = 0.0;
  for (int k = lo + int(threadIdx.x); k < hi; k += blockDim.x) {
    int r = k / (cols - 2) + 1, c = k % (cols - 2) + 1;
    acc += double(h[(long long)r * cols + c]);
  }
  sm[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) sm[threadIdx.x] += sm[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sm[0] * area;
}

__global__ void k_sum_parts(const double* __restrict__ part, int n, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int i = 0; i < n; i++) s += part[i];
  *out = s;
}

constexpr int kVolBlocks = 256;

// FP32 dry / flush threshold: the reference's FP64 value, so thin films wet
// and dry where the FP64 form's do.  1e-6 (the reference's FP32 EPS12)
// flushes fluxes of small-scale flows and loses mass (2e-4 relative on the
// 0.1 m deep paraboloid fixture).
constexpr float kDefaultDryEpsF32 = 1e-12f;

// Waves per SIMD of a batched launch (the step kernel's registers allow 6
// resident, and 5 measured fastest on one large grid: more waves run in later
// rounds that no longer share their neighbours' halo rows in L2; an uneven
// ensemble gains from 10-15, which the autotuner finds) and the band heights a
// small batch is given: it spreads over about kSmallWavesPerSimd waves per
// SIMD with bands of kFloorRowsPerWave..kMinRowsPerWave rows.  Each piece
// recomputes a 4-row halo, but a lone wave on a SIMD is latency-bound, so
// short bands on more waves win: a 400 x 400 grid steps in 16.5 us with
// 2-row bands against 25.3 us with 8-row bands at one wave per SIMD (81 us
// with 48-row bands), Thacker's 1024 x 1024 in 28.7 against 42.2 us.
constexpr int kWavesPerSimd = 5;
constexpr int kFloorRowsPerWave = 2;
constexpr int kMinRowsPerWave = 48;
constexpr int kSmallWavesPerSimd = 4;

}  // namespace
}  // namespace tr

using namespace tr;

struct tr_batch {
  hipStream_t stream = nullptr;
  int64_t budget = 0;          // 0: no limit
  std::vector<Member> m;
  int64_t bytes = 0;           // sum over members + shared
  InstF32* d_inst = nullptr;
  CtrlPair* d_ctrl = nullptr;
  CtrlPair* h_ctrl = nullptr;  // pinned mirror
  int cap = 0;
  Piece* d_pieces = nullptr;
  int* d_members = nullptr;    // [cap] active members, in partition order
  int nmembers = 0;
  int first_member = 0;        // order[0] of the partition
  int any_bc = 0;              // some member in the partition has external boundaries
  int piece_cap = 0;
  int nwork = 0;               // waves per launch
  int band_rows = 0;           // rows per band of the current partition (the tile height)
  int npieces = 0;
  int simd_count = 1024;
  int waves_per_simd = kWavesPerSimd;
  bool table_dirty = true;
  // hipGraph of graph_steps batched steps, valid for one partition
  int graph_steps = 0;         // 0: eager launches only
  hipGraphExec_t gexec = nullptr;
  BatchArgs gargs{};
  int gsteps_built = 0;
  double* d_part = nullptr;
  double* d_sum = nullptr;
  double* h_sum = nullptr;
  // observation sampling scratch (grown on demand)
  int32_t* d_obs_idx = nullptr;
  double* d_obs_val = nullptr;
  int obs_cap = 0;

  ~tr_batch() {
    if (stream) hipStreamSynchronize(stream);
    if (gexec) hipGraphExecDestroy(gexec);
    for (auto& mm : m)
      for (void* p : mm.allocs) hipFree(p);
    hipFree(d_inst);
    hipFree(d_ctrl);
    hipFree(d_pieces);
    hipFree(d_members);
    hipFree(d_part);
    hipFree(d_sum);
    hipFree(d_obs_idx);
    hipFree(d_obs_val);
    if (h_ctrl) hipHostFree(h_ctrl);
    if (h_sum) hipHostFree(h_sum);
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

int grow(tr_batch* b, int need) {
  if (need <= b->cap) return TR_OK;
  int cap = std::max(need, std::max(8, b->cap * 2));
  InstF32* di = nullptr;
  CtrlPair* dc = nullptr;
  CtrlPair* hc = nullptr;
  int* dm = nullptr;
  TR_HIP(hipStreamSynchronize(b->stream));
  TR_HIP(hipMalloc(reinterpret_cast<void**>(&dm), sizeof(int) * size_t(cap)));
  hipFree(b->d_members);
  b->d_members = dm;
  TR_HIP(hipMalloc(reinterpret_cast<void**>(&di), sizeof(InstF32) * size_t(cap)));
  TR_HIP(hipMalloc(reinterpret_cast<void**>(&dc), sizeof(CtrlPair) * size_t(cap)));
  TR_HIP(hipHostMalloc(reinterpret_cast<void**>(&hc), sizeof(CtrlPair) * size_t(cap)));
  if (b->cap > 0) {
    TR_HIP(hipMemcpy(dc, b->d_ctrl, sizeof(CtrlPair) * size_t(b->cap), hipMemcpyDeviceToDevice));
    // the host mirror keeps what the last sync read (tr_batch_poll reads it)
    std::memcpy(hc, b->h_ctrl, sizeof(CtrlPair) * size_t(b->cap));
  }
  hipFree(b->d_inst);
  hipFree(b->d_ctrl);
  if (b->h_ctrl) hipHostFree(b->h_ctrl);
  b->d_inst = di;
  b->d_ctrl = dc;
  b->h_ctrl = hc;
  b->cap = cap;
  // descriptors point into the control pool: rewrite them all
  for (int i = 0; i < int(b->m.size()); i++) {
    if (!b->m[size_t(i)].used) continue;
    b->m[size_t(i)].inst.ctrl = &b->d_ctrl[i].ctrl[0];
    b->m[size_t(i)].inst.cursor = &b->d_ctrl[i].cursor;
    TR_HIP(hipMemcpy(b->d_inst + i, &b->m[size_t(i)].inst, sizeof(InstF32), hipMemcpyHostToDevice));
  }
  b->table_dirty = true;
  return TR_OK;
}

template <class T>
int m_alloc(Member& mm, T*& p, size_t n) {
  p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T));
  if (e != hipSuccess) {
    p = nullptr;
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return TR_ERR_NOMEM;
  }
  mm.allocs.push_back(p);
  mm.bytes += int64_t(n * sizeof(T));
  return TR_OK;
}

// f(lo, hi) over [0, n) split across host threads, at least `grain` items
// each: admitting a large member converts a few 16 M-cell arrays, which
// single-threaded is most of its 175 ms on a 4087^2 grid.
template <class F>
void host_parallel(size_t n, size_t grain, F f) {
  grain = std::max<size_t>(grain, 1);
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t nt = std::min<size_t>(std::min<size_t>(hw, 16), (n + grain - 1) / grain);
  if (nt <= 1) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  th.reserve(nt);
  size_t t = 0;
  try {
    for (; t < nt; t++) th.emplace_back(f, n * t / nt, n * (t + 1) / nt);
  } catch (const std::system_error&) {
    // no thread to spare: this thread takes the ranges not yet started
    for (; t < nt; t++) f(n * t / nt, n * (t + 1) / nt);
  }
  for (auto& x : th) x.join();
}

template <class T>
int m_upload(Member& mm, T*& p, const T* src, size_t n, hipStream_t s) {
  RC(m_alloc(mm, p, n));
  if (src && n) TR_HIP(hipMemcpyAsync(p, src, n * sizeof(T), hipMemcpyHostToDevice, s));
  return TR_OK;
}

bool uniform_zones(const tr_case_view& v) {
  if (!v.runoff_id) return true;
  for (int r = 1; r < v.rows - 1; r++)
    for (int c = 1; c < v.cols - 1; c++)
      if (v.runoff_id[size_t(r) * v.cols + c] != 0) return false;
  return true;
}

// The n^2 array is uploaded when any cell has n^2 > 0; the step decides
// friction per cell at n^2 > the member's dry threshold (1e-12 by default,
// the reference's EPS12, src/kernels_state.h:85).
bool has_friction(const tr_case_view& v) {
  const size_t n = size_t(v.rows) * v.cols;
  for (size_t i = 0; i < n; i++)
    if (v.nman2[i] > 0.0) return true;
  return false;
}

}  // namespace

// Everything tr_batch_add checks of a view before it allocates (also run by
// tr_serve_submit, so a request that could never be admitted is refused at
// submit).
int tr::check_member_view(const tr_case_view* v) {
  if (!v || v->rows < 3 || v->cols < 3 || !v->h || !v->qx || !v->qy || !v->dem || !v->nman2) {
    set_error("member view: view needs rows,cols >= 3 and h, qx, qy, dem, nman2");
    return TR_ERR_INVALID;
  }
  if (int64_t(v->rows) * v->cols > int64_t(INT32_MAX)) {
    set_error("member view: grid exceeds 2^31 cells");
    return TR_ERR_INVALID;
  }
  if (!(v->dx > 0) || !(v->courant > 0) || !(v->print_interval > 0)) {
    set_error("member view: dx, courant and print_interval must be positive");
    return TR_ERR_INVALID;
  }
  if (v->num_extbc_cells > 0 && (!v->bc_index || !v->bc_type || !v->bc_start || !v->bc_nrows ||
                                  !v->bc_var1 || !v->bc_var2 || v->bc_nvars < 1)) {
    set_error("member view: external boundary segments need their index, type and series arrays");
    return TR_ERR_INVALID;
  }
  if (v->num_runoffs > 0 && (v->roff_nrows < 2 || !v->roff_time || !v->roff_rate)) {
    set_error("member view: runoff needs at least two rows");
    return TR_ERR_INVALID;
  }
  if (v->num_src > 0 && (v->hyg_nrows < 2 || !v->hyg_time || !v->hyg_val || !v->src_pos)) {
    set_error("member view: sources need a hydrograph with at least two rows");
    return TR_ERR_INVALID;
  }
  if (v->num_src > kMaxSrcFused) {
    set_error("member view: too many point sources for the fused step");
    return TR_ERR_INVALID;
  }
  const int64_t n = int64_t(v->rows) * v->cols;
  for (int i = 0; i < v->num_src; i++) {
    int r = v->src_pos[i] / v->cols, c = v->src_pos[i] % v->cols;
    if (v->src_pos[i] < 0 || v->src_pos[i] >= n || r < 1 || r > v->rows - 2 || c < 1 ||
        c > v->cols - 2) {
      set_error("member view: source cell outside the interior");
      return TR_ERR_INVALID;
    }
  }
  if (v->num_runoffs > 0 && v->runoff_id) {
    for (int r = 1; r < v->rows - 1; r++)
      for (int c = 1; c < v->cols - 1; c++) {
        int z = v->runoff_id[int64_t(r) * v->cols + c];
        if (z < 0 || z >= v->num_runoffs) {
          set_error("member view: runoff zone id out of range");
          return TR_ERR_INVALID;
        }
      }
  }
  return TR_OK;
}

namespace {

int64_t estimate_bytes(const tr_case_view& v, bool fric, bool zones) {
  const int64_t n = int64_t(v.rows) * v.cols;
  int64_t b = n * 4 * 8;               // h,qx,qy x2 + dzx + dzy
  if (fric) b += n * 4;
  if (v.maxh) b += n * 4;
  if (zones) b += n * 4;
  b += int64_t(v.num_runoffs) * (v.roff_nrows + 3) * 8 + int64_t(v.roff_nrows) * 8;
  b += int64_t(v.num_src) * (v.hyg_nrows + 6) * 8 + int64_t(v.hyg_nrows) * 8;
  if (v.num_extbc_cells > 0) b += n * 4 + int64_t(v.num_extbc_cells) * 28 + int64_t(v.bc_nvars) * 16;
  return b + 4096;
}

// Partition the active members over at most waves_per_simd waves per SIMD
// of the device, one piece per wave: each member is cut into row bands of at
// most H rows (H the smallest height that needs no more waves than that) and
// every column strip of a band is one wave's piece.  Pieces are laid out
// member by member, band by band, strip by strip, and waves are placed
// XCD-contiguously, so the strips of a band start at the same row at the same
// time on one XCD: neighbouring strips then read the halo columns they share
// from that XCD's L2, and a band's rows are streamed together.  (A partition
// into equal row shares balanced the waves exactly but let share boundaries
// drift across strips; on the 4096^2 watershed the aligned bands step 22 %
// faster.)
int rebuild_work(tr_batch* b) {
  TR_HIP(hipStreamSynchronize(b->stream));
  std::vector<int> order;
  for (int i = 0; i < int(b->m.size()); i++)
    if (b->m[size_t(i)].used && b->m[size_t(i)].active) order.push_back(i);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
    return b->m[size_t(x)].interior > b->m[size_t(y)].interior;
  });
  int64_t total = 0;
  b->any_bc = 0;
  for (int i : order) {
    total += int64_t(b->m[size_t(i)].inst.nbx) * b->m[size_t(i)].inst.rows;
    if (b->m[size_t(i)].inst.bc.n > 0) b->any_bc = 1;
  }
  std::vector<Piece> pieces;
  int nw = 0;
  int64_t band_h = 0;
  if (total > 0) {
    const int64_t want = int64_t(b->simd_count) * b->waves_per_simd;
    auto waves_for = [&](int64_t H) {
      int64_t w = 0;
      for (int i : order) {
        const InstF32& P = b->m[size_t(i)].inst;
        w += int64_t(P.nbx) * ((P.rows + H - 1) / H);
      }
      return w;
    };
    // a batch too small to fill the device spreads over about
    // kSmallWavesPerSimd waves per SIMD, with bands of
    // kFloorRowsPerWave..kMinRowsPerWave rows
    int64_t lo = std::max<int64_t>(
        kFloorRowsPerWave, std::min<int64_t>(kMinRowsPerWave, total / (int64_t(b->simd_count) * kSmallWavesPerSimd)));
    int64_t hi = lo;
    for (int i : order) hi = std::max<int64_t>(hi, b->m[size_t(i)].inst.rows);
    while (lo < hi) {
      const int64_t H = (lo + hi) / 2;
      if (waves_for(H) <= want) hi = H;
      else lo = H + 1;
    }
    const int64_t H = lo;
    band_h = H;
    for (int i : order) {
      const InstF32& P = b->m[size_t(i)].inst;
      const int nb = int((P.rows + H - 1) / H);
      for (int k = 0; k < nb; k++) {
        const int r0 = int(int64_t(P.rows) * k / nb), r1 = int(int64_t(P.rows) * (k + 1) / nb);
        for (int bx = 0; bx < P.nbx; bx++) pieces.push_back(Piece{i, bx, r0, r1 - 1});
      }
    }
    nw = int(pieces.size());
  }
  const size_t np = std::max<size_t>(pieces.size(), 1);
  if (np > size_t(b->piece_cap)) {
    hipFree(b->d_pieces);
    b->d_pieces = nullptr;
    b->piece_cap = int(np * 2);
    TR_HIP(hipMalloc(reinterpret_cast<void**>(&b->d_pieces), sizeof(Piece) * size_t(b->piece_cap)));
  }
  if (!pieces.empty())
    TR_HIP(hipMemcpy(b->d_pieces, pieces.data(), sizeof(Piece) * pieces.size(), hipMemcpyHostToDevice));
  if (!order.empty())
    TR_HIP(hipMemcpy(b->d_members, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
  b->nmembers = int(order.size());
  b->first_member = order.empty() ? 0 : order[0];
  b->nwork = nw;
  b->band_rows = nw > 0 ? int(band_h) : 0;
  b->npieces = int(pieces.size());
  b->table_dirty = false;
  return TR_OK;
}

int read_ctrls(tr_batch* b) {
  if (b->cap == 0) return TR_OK;
  TR_HIP(hipMemcpyAsync(b->h_ctrl, b->d_ctrl, sizeof(CtrlPair) * size_t(b->cap), hipMemcpyDeviceToHost,
                        b->stream));
  TR_HIP(hipStreamSynchronize(b->stream));
  return TR_OK;
}

const StepCtrl& cur_ctrl(const tr_batch* b, int i) {
  const CtrlPair& cp = b->h_ctrl[i];
  return cp.ctrl[cp.cursor & 1];
}

int put_ctrl(tr_batch* b, int i) {
  TR_HIP(hipMemcpyAsync(&b->d_ctrl[i], &b->h_ctrl[i], sizeof(CtrlPair), hipMemcpyHostToDevice, b->stream));
  return TR_OK;
}

int put_inst(tr_batch* b, int i) {
  TR_HIP(hipMemcpyAsync(b->d_inst + i, &b->m[size_t(i)].inst, sizeof(InstF32), hipMemcpyHostToDevice,
                        b->stream));
  return TR_OK;
}

Member* member(tr_batch* b, int id) {
  if (!b || id < 0 || id >= int(b->m.size()) || !b->m[size_t(id)].used) {
    set_error("tr_batch: no such member");
    return nullptr;
  }
  return &b->m[size_t(id)];
}

}  // namespace

extern "C" {

int tr_batch_create(int64_t hbm_budget_bytes, int32_t waves_per_simd, tr_batch** out) {
  if (!out || hbm_budget_bytes < 0 || waves_per_simd < 0 || waves_per_simd > 32) {
    set_error("tr_batch_create: bad arguments");
    return TR_ERR_INVALID;
  }
  *out = nullptr;
  auto* b = new tr_batch();
  b->budget = hbm_budget_bytes;
  if (waves_per_simd > 0) b->waves_per_simd = waves_per_simd;
  {
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
      b->simd_count = prop.multiProcessorCount * 4;  // 4 SIMDs per CDNA compute unit
  }
  hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    b->stream = nullptr;
    delete b;
    return hip_fail(e, "hipStreamCreate");
  }
  if (hipMalloc(reinterpret_cast<void**>(&b->d_part), sizeof(double) * kVolBlocks) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&b->d_sum), sizeof(double)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&b->h_sum), sizeof(double)) != hipSuccess) {
    delete b;
    set_error("tr_batch_create: allocation failed");
    return TR_ERR_NOMEM;
  }
  int rc = grow(b, 8);
  if (rc) {
    delete b;
    return rc;
  }
  *out = b;
  return TR_OK;
}

int64_t tr_batch_member_bytes(const tr_case_view* v) {
  if (!v || v->rows < 3 || v->cols < 3 || !v->nman2) return -1;
  return estimate_bytes(*v, has_friction(*v), !uniform_zones(*v));
}

int32_t tr_default_strip_rows(int32_t rows, int32_t cols) {
  return rows >= 3 && cols >= 3 ? default_strip_rows(rows, cols) : 0;
}

int tr_batch_add(tr_batch* b, const tr_case_view* vp, int32_t* id_out) {
  return tr_batch_add_ex(b, vp, nullptr, id_out);
}

int tr_batch_add_ex(tr_batch* b, const tr_case_view* vp, const tr_tuning* tune, int32_t* id_out) {
  if (!b || !id_out) {
    set_error("tr_batch_add: NULL argument");
    return TR_ERR_INVALID;
  }
  RC(check_member_view(vp));
  const tr_case_view& v = *vp;
  const bool fric = has_friction(v);
  const bool zones = v.num_runoffs > 0 && !uniform_zones(v);
  const int64_t need = estimate_bytes(v, fric, zones);
  if (b->budget > 0 && b->bytes + need > b->budget) {
    set_error("tr_batch_add: member needs " + std::to_string(need) + " B, over the HBM budget");
    return TR_ERR_NOMEM;
  }
  int id = -1;
  for (int i = 0; i < int(b->m.size()); i++)
    if (!b->m[size_t(i)].used) { id = i; break; }
  if (id < 0) {
    id = int(b->m.size());
    b->m.emplace_back();
  }
  RC(grow(b, id + 1));
  Member& mm = b->m[size_t(id)];
  mm = Member();
  hipStream_t s = b->stream;
  const int rows = v.rows, cols = v.cols;
  const size_t n = size_t(rows) * cols;
  InstF32& P = mm.inst;
  P.rows = rows;
  P.cols = cols;
  P.eps_f = (tune && tune->dry_eps > 0) ? float(tune->dry_eps) : kDefaultDryEpsF32;
  P.round_q = (tune && tune->round_q) ? 1 : 0;
  P.nbx = fused_grid_x_f32(cols);
  P.dx = v.dx;
  P.hextra = v.hextra;
  P.cn_dx = v.courant * v.dx;
  P.dx_f = float(v.dx);
  P.hextra_f = float(v.hextra);
  P.cn_dx_f = float(v.courant * v.dx);
  // FP32 state; the bed as per-face differences taken in FP64
  std::vector<float> hf(n), qxf(n), qyf(n), dzx(n), dzy(n), n2f(fric ? n : 0);
  host_parallel(size_t(rows), (size_t(1) << 20) / size_t(cols) + 1, [&](size_t r0, size_t r1) {
    for (size_t r = r0; r < r1; r++)
      for (int c = 0; c < cols; c++) {
        const size_t i = r * size_t(cols) + size_t(c);
        hf[i] = float(v.h[i]);
        qxf[i] = float(v.qx[i]);
        qyf[i] = float(v.qy[i]);
        dzx[i] = c < cols - 1 ? float(v.dem[i + 1] - v.dem[i]) : 0.f;
        dzy[i] = r >= 1 ? float(v.dem[i - size_t(cols)] - v.dem[i]) : 0.f;
        if (fric) n2f[i] = float(v.nman2[i]);
      }
  });
  // the second ping-pong slot starts as a device copy of the first
  int rc = m_upload(mm, P.h[0], hf.data(), n, s);
  if (!rc) rc = m_upload(mm, P.qx[0], qxf.data(), n, s);
  if (!rc) rc = m_upload(mm, P.qy[0], qyf.data(), n, s);
  if (!rc) rc = m_alloc(mm, P.h[1], n);
  if (!rc) rc = m_alloc(mm, P.qx[1], n);
  if (!rc) rc = m_alloc(mm, P.qy[1], n);
  if (!rc && n) {
    const size_t nb = n * sizeof(float);
    hipError_t e = hipMemcpyAsync(P.h[1], P.h[0], nb, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(P.qx[1], P.qx[0], nb, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(P.qy[1], P.qy[0], nb, hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) rc = hip_fail(e, "tr_batch_add: slot copy");
  }
  float* pdx = nullptr;
  float* pdy = nullptr;
  if (!rc) rc = m_upload(mm, pdx, dzx.data(), n, s);
  if (!rc) rc = m_upload(mm, pdy, dzy.data(), n, s);
  P.dzx = pdx;
  P.dzy = pdy;
  P.nman2 = nullptr;
  if (!rc && fric) {
    float* p = nullptr;
    rc = m_upload(mm, p, n2f.data(), n, s);
    P.nman2 = p;
  }
  P.maxh = nullptr;
  if (!rc && v.maxh) {
    std::vector<float> mf(n);
    for (size_t i = 0; i < n; i++) mf[i] = float(v.maxh[i]);
    rc = m_upload(mm, P.maxh, mf.data(), n, s);
  }
  // forcing
  P.num_runoffs = v.num_runoffs;
  P.roff_nrows = v.roff_nrows;
  P.runoff_id = nullptr;
  if (!rc && v.num_runoffs > 0) {
    double *rate = nullptr, *rt = nullptr, *za = nullptr;
    rc = m_upload(mm, rate, v.roff_rate, size_t(v.num_runoffs) * v.roff_nrows, s);
    if (!rc) rc = m_upload(mm, rt, v.roff_time, size_t(v.roff_nrows), s);
    std::vector<double> area(size_t(v.num_runoffs), 0.0);
    if (!zones) area[0] = double(rows - 2) * double(cols - 2);  // every interior cell in zone 0
    else
      for (int r = 1; r < rows - 1; r++)
        for (int c = 1; c < cols - 1; c++) area[size_t(v.runoff_id[size_t(r) * cols + c])] += 1.0;
    for (auto& a : area) a *= v.dx * v.dx;
    if (!rc) rc = m_upload(mm, za, area.data(), area.size(), s);
    P.roff_rate = rate;
    P.roff_time = rt;
    P.zone_area = za;
    if (!rc) rc = m_alloc(mm, P.zone_add, size_t(2 * v.num_runoffs));
    if (!rc && zones) {
      int* rid = nullptr;
      rc = m_upload(mm, rid, v.runoff_id, n, s);
      P.runoff_id = rid;
    }
  }
  P.nsrc = v.num_src;
  P.hyg_nrows = v.hyg_nrows;
  if (!rc && v.num_src > 0) {
    std::vector<int> order(size_t(v.num_src));
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return v.src_pos[x] < v.src_pos[y]; });
    std::vector<int> rr, cc;
    for (int k : order) {
      rr.push_back(v.src_pos[k] / cols);
      cc.push_back(v.src_pos[k] % cols);
    }
    int *pr = nullptr, *pc = nullptr, *pi = nullptr;
    double *ht = nullptr, *hv = nullptr;
    rc = m_upload(mm, pr, rr.data(), rr.size(), s);
    if (!rc) rc = m_upload(mm, pc, cc.data(), cc.size(), s);
    if (!rc) rc = m_upload(mm, pi, order.data(), order.size(), s);
    if (!rc) rc = m_upload(mm, ht, v.hyg_time, size_t(v.hyg_nrows), s);
    if (!rc) rc = m_upload(mm, hv, v.hyg_val, size_t(v.hyg_nrows) * v.num_src, s);
    if (!rc) rc = m_alloc(mm, P.src_add, size_t(2 * v.num_src));
    P.src_row = pr;
    P.src_col = pc;
    P.src_id = pi;
    P.hyg_time = ht;
    P.hyg_val = hv;
  }
  P.bc = ExtBC{};
  if (!rc && v.num_extbc_cells > 0) {
    ExtBCHost bch;
    const std::string err = build_extbc(rows, cols, v.num_extbc_cells, v.bc_index, v.bc_type,
                                        v.bc_start, v.bc_nrows, v.dem, v.nman2, bch);
    if (!err.empty()) {
      set_error("tr_batch_add: " + err);
      rc = TR_ERR_INVALID;
    }
    int *map = nullptr, *nx = nullptr, *ty = nullptr, *st = nullptr, *nv = nullptr;
    double *v1 = nullptr, *v2 = nullptr, *zc = nullptr, *n2c = nullptr;
    if (!rc) rc = m_upload(mm, map, bch.map.data(), bch.map.size(), s);
    if (!rc) rc = m_upload(mm, nx, bch.next.data(), bch.next.size(), s);
    if (!rc) rc = m_upload(mm, ty, bch.type.data(), bch.type.size(), s);
    if (!rc) rc = m_upload(mm, st, bch.start.data(), bch.start.size(), s);
    if (!rc) rc = m_upload(mm, nv, bch.nvars.data(), bch.nvars.size(), s);
    if (!rc) rc = m_upload(mm, v1, v.bc_var1, size_t(v.bc_nvars), s);
    if (!rc) rc = m_upload(mm, v2, v.bc_var2, size_t(v.bc_nvars), s);
    if (!rc) rc = m_upload(mm, zc, bch.zc.data(), bch.zc.size(), s);
    if (!rc) rc = m_upload(mm, n2c, bch.n2c.data(), bch.n2c.size(), s);
    if (!rc) P.bc = ExtBC{v.num_extbc_cells, map, nx, ty, st, nv, v1, v2, zc, n2c};
  }
  P.open_bc = v.open_boundaries;
  P.ctrl = &b->d_ctrl[id].ctrl[0];
  P.cursor = &b->d_ctrl[id].cursor;
  P.sim_duration = v.sim_duration;
  P.print_interval = v.print_interval;
  P.init_dt = v.init_dt;
  P.t_stop = INFINITY;
  P.step_stop = INT64_MAX;
  P.halt_on_print = 0;
  P.time_series = v.time_series_flag;
  P.print_observation = v.print_observation;
  P.halt_on_obs = 0;
  P.it_print = v.it_print;
  P.it_count0 = v.it_count;
  mm.interior = int64_t(rows - 2) * (cols - 2);
  mm.sim_duration = v.sim_duration;
  if (rc) {
    for (void* p : mm.allocs) hipFree(p);
    mm = Member();
    return rc;
  }
  mm.used = true;
  // initial control block in slot 1; the prime prepares slot 0 for step 0
  CtrlPair& cp = b->h_ctrl[id];
  std::memset(&cp, 0, sizeof(cp));
  StepCtrl& c0 = cp.ctrl[1];
  c0.simtime = v.sim_start;
  reset_dtmin(c0);
  c0.print_id = v.checkpoint_id;
  if (c0.print_id == 0 && c0.simtime > 0.0) c0.print_id = int(c0.simtime / v.print_interval);
  c0.roff_row = v.roff_row0;
  c0.last_obs = v.sim_start;  // src/triton_simulate.h:47
  reset_dtmin(cp.ctrl[0]);
  cp.cursor = 1;
  RC(put_ctrl(b, id));
  RC(put_inst(b, id));
  RC(launch_prime_f32(b->d_inst, b->d_ctrl, id, P, s));
  TR_HIP(hipStreamSynchronize(s));
  b->bytes += mm.bytes;
  mm.active = true;
  b->table_dirty = true;
  *id_out = id;
  return TR_OK;
}

int tr_batch_remove(tr_batch* b, int32_t id) {
  Member* mm = member(b, id);
  if (!mm) return TR_ERR_INVALID;
  TR_HIP(hipStreamSynchronize(b->stream));
  for (void* p : mm->allocs) hipFree(p);
  b->bytes -= mm->bytes;
  *mm = Member();
  b->table_dirty = true;
  return TR_OK;
}

// Limits for the launches queued next, for every member: simulated time
// t_stop (<= 0: each member's sim_duration) and at most max_steps more steps
// (< 0: unlimited).  Members already at their limit drop out of the table.
int tr_batch_arm(tr_batch* b, double t_stop, int64_t max_steps) {
  if (!b) {
    set_error("tr_batch_arm: NULL batch");
    return TR_ERR_INVALID;
  }
  TR_HIP(hipStreamSynchronize(b->stream));
  RC(read_ctrls(b));
  for (int i = 0; i < int(b->m.size()); i++) {
    Member& mm = b->m[size_t(i)];
    if (!mm.used) continue;
    CtrlPair& cp = b->h_ctrl[i];
    StepCtrl& c = cp.ctrl[cp.cursor & 1];
    const double stop = t_stop > 0 ? std::min(t_stop, mm.sim_duration) : mm.sim_duration;
    mm.inst.t_stop = stop;
    mm.inst.step_stop = max_steps < 0 ? INT64_MAX : c.step + max_steps;
    const bool halt = c.finished || !(c.simtime < stop) || c.step >= mm.inst.step_stop;
    c.halt = halt ? 1 : 0;
    const bool was = mm.active;
    mm.active = !halt;
    if (was != mm.active) b->table_dirty = true;
    RC(put_ctrl(b, i));
    RC(put_inst(b, i));
  }
  TR_HIP(hipStreamSynchronize(b->stream));
  if (b->table_dirty) RC(rebuild_work(b));
  return TR_OK;
}

}  // extern "C"

namespace {

bool same_args(const BatchArgs& x, const BatchArgs& y) {
  return x.inst == y.inst && x.pieces == y.pieces &&
         x.nwaves == y.nwaves && x.members == y.members && x.nmembers == y.nmembers &&
         x.ctrls == y.ctrls && x.first_member == y.first_member &&
         x.any_bc == y.any_bc;
}

// Capture graph_steps batched steps (step + close kernels only: no memset
// or copy nodes) for the current partition.  The kernels read the piece
// table and descriptors from device memory, so the graph stays valid while
// the table pointers and the wave and member counts stay the same.
int build_graph(tr_batch* b, const BatchArgs& a) {
  if (b->gexec) {
    hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
  }
  hipGraph_t g = nullptr;
  TR_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = TR_OK;
  for (int k = 0; k < b->graph_steps && !rc; k++) rc = launch_step_f32(a, a.nwaves, b->stream);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->gexec = nullptr;
    return hip_fail(e, "hipGraphInstantiate");
  }
  b->gargs = a;
  b->gsteps_built = b->graph_steps;
  return TR_OK;
}

}  // namespace

extern "C" {

int tr_batch_enqueue(tr_batch* b, int64_t nsteps) {
  if (!b || nsteps < 0) {
    set_error("tr_batch_enqueue: bad arguments");
    return TR_ERR_INVALID;
  }
  if (b->table_dirty) RC(rebuild_work(b));
  BatchArgs a{b->d_inst, b->d_ctrl, b->d_pieces, b->nwork, b->d_members, b->nmembers, b->first_member, b->any_bc};
  int64_t k = 0;
  if (b->graph_steps > 0 && a.nwaves > 0 && nsteps >= b->graph_steps) {
    if (!b->gexec || b->gsteps_built != b->graph_steps || !same_args(a, b->gargs)) RC(build_graph(b, a));
    for (; k + b->graph_steps <= nsteps; k += b->graph_steps) TR_HIP(hipGraphLaunch(b->gexec, b->stream));
  }
  for (; k < nsteps; k++) RC(launch_step_f32(a, b->nwork, b->stream));
  return TR_OK;
}

// Launch-shape knobs (the autotuner's): waves per SIMD of the partition and
// steps per captured graph (0: eager launches).  Results do not depend on
// either.
int tr_batch_set_launch(tr_batch* b, int32_t waves_per_simd, int32_t graph_steps) {
  if (!b || waves_per_simd < 0 || waves_per_simd > 32 || graph_steps < 0 || graph_steps > 4096) {
    set_error("tr_batch_set_launch: bad arguments");
    return TR_ERR_INVALID;
  }
  TR_HIP(hipStreamSynchronize(b->stream));
  const int w = waves_per_simd > 0 ? waves_per_simd : kWavesPerSimd;
  if (w != b->waves_per_simd) {
    b->waves_per_simd = w;
    b->table_dirty = true;
  }
  b->graph_steps = graph_steps;
  return TR_OK;
}

// Wait for queued steps; members that reached their limit leave the
// partition.  n_active gets the number of members still stepping.
int tr_batch_sync(tr_batch* b, int32_t* n_active) {
  if (!b) {
    set_error("tr_batch_sync: NULL batch");
    return TR_ERR_INVALID;
  }
  // the control copy is queued behind the steps, so its wait is the sync
  if (b->cap == 0) TR_HIP(hipStreamSynchronize(b->stream));
  else RC(read_ctrls(b));
  int act = 0;
  for (int i = 0; i < int(b->m.size()); i++) {
    Member& mm = b->m[size_t(i)];
    if (!mm.used) continue;
    const StepCtrl& c = cur_ctrl(b, i);
    const bool on = !c.halt;
    if (on != mm.active) {
      mm.active = on;
      b->table_dirty = true;
    }
    act += on ? 1 : 0;
  }
  if (n_active) *n_active = act;
  return TR_OK;
}

int tr_batch_run(tr_batch* b, double t_stop, int64_t max_steps, int32_t* n_active) {
  RC(tr_batch_arm(b, t_stop, max_steps));
  int32_t act = 0;
  RC(tr_batch_sync(b, &act));
  int64_t batch = 16, done = 0;
  while (act > 0) {
    int64_t n = batch;
    if (max_steps >= 0) n = std::min<int64_t>(n, max_steps - done);
    if (n <= 0) break;
    RC(tr_batch_enqueue(b, n));
    done += n;
    RC(tr_batch_sync(b, &act));
    batch = std::min<int64_t>(batch * 2, 512);
  }
  if (n_active) *n_active = act;
  return TR_OK;
}

int tr_batch_status(tr_batch* b, int32_t id, tr_sim_status* st) {
  Member* mm = member(b, id);
  if (!mm || !st) return TR_ERR_INVALID;
  RC(read_ctrls(b));
  const StepCtrl& c = cur_ctrl(b, id);
  std::memset(st, 0, sizeof(*st));
  st->simtime = c.simtime;
  st->dt = c.dt;
  st->steps = c.step;
  st->print_id = c.print_id;
  st->finished = c.finished;
  st->volume_in = c.vol_in;
  st->observed = c.observed;
  st->last_obs = c.last_obs;
  st->average_dt = c.dt_acc;
  st->it_count = mm->inst.it_count0 + c.step;
  const int cur = b->h_ctrl[id].cursor & 1;
  hipLaunchKernelGGL(k_volume_f32, dim3(kVolBlocks), dim3(256), 0, b->stream, mm->inst.h[cur],
                     mm->inst.rows, mm->inst.cols, mm->inst.dx * mm->inst.dx, b->d_part);
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(64), 0, b->stream, b->d_part, kVolBlocks, b->d_sum);
  RC(launch_check("volume_f32"));
  TR_HIP(hipMemcpyAsync(b->h_sum, b->d_sum, sizeof(double), hipMemcpyDeviceToHost, b->stream));
  TR_HIP(hipStreamSynchronize(b->stream));
  st->volume = *b->h_sum;
  return TR_OK;
}

int tr_batch_set_print_halt(tr_batch* b, int32_t id, int32_t on) {
  Member* mm = member(b, id);
  if (!mm) return TR_ERR_INVALID;
  TR_HIP(hipStreamSynchronize(b->stream));
  mm->inst.halt_on_print = on ? 1 : 0;
  RC(put_inst(b, id));
  TR_HIP(hipStreamSynchronize(b->stream));
  return TR_OK;
}

// Status of a member as of the last tr_batch_sync / tr_batch_arm, from the
// host mirror of the control blocks: no device work.  volume is NaN.
int tr_batch_poll(tr_batch* b, int32_t id, tr_sim_status* st) {
  Member* mm = member(b, id);
  if (!mm || !st) return TR_ERR_INVALID;
  const StepCtrl& c = cur_ctrl(b, id);
  std::memset(st, 0, sizeof(*st));
  st->simtime = c.simtime;
  st->dt = c.dt;
  st->steps = c.step;
  st->print_id = c.print_id;
  st->finished = c.finished;
  st->volume = NAN;
  st->volume_in = c.vol_in;
  st->observed = c.observed;
  st->last_obs = c.last_obs;
  st->average_dt = c.dt_acc;
  st->it_count = mm->inst.it_count0 + c.step;
  return TR_OK;
}

int tr_batch_download(tr_batch* b, int32_t id, double* h, double* qx, double* qy, double* maxh) {
  Member* mm = member(b, id);
  if (!mm) return TR_ERR_INVALID;
  RC(read_ctrls(b));
  const int cur = b->h_ctrl[id].cursor & 1;
  const size_t n = size_t(mm->inst.rows) * mm->inst.cols;
  std::vector<float> tmp(n);
  auto get = [&](const float* src, double* dst) -> int {
    if (!dst || !src) return TR_OK;
    TR_HIP(hipMemcpyAsync(tmp.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    TR_HIP(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < n; i++) dst[i] = double(tmp[i]);
    return TR_OK;
  };
  RC(get(mm->inst.h[cur], h));
  RC(get(mm->inst.qx[cur], qx));
  RC(get(mm->inst.qy[cur], qy));
  RC(get(mm->inst.maxh, maxh));
  return TR_OK;
}

int tr_batch_sample(tr_batch* b, int32_t id, int32_t n, const int32_t* pos, double* h, double* qx,
                    double* qy) {
  Member* mm = member(b, id);
  if (!mm || n < 0 || (n > 0 && !pos)) {
    set_error("tr_batch_sample: bad arguments");
    return TR_ERR_INVALID;
  }
  if (n == 0) return TR_OK;
  const int64_t cells = int64_t(mm->inst.rows) * mm->inst.cols;
  for (int32_t i = 0; i < n; i++)
    if (pos[i] < 0 || pos[i] >= cells) {
      set_error("tr_batch_sample: cell index out of range");
      return TR_ERR_INVALID;
    }
  RC(read_ctrls(b));
  const int cur = b->h_ctrl[id].cursor & 1;
  if (n > b->obs_cap) {
    hipFree(b->d_obs_idx);
    hipFree(b->d_obs_val);
    b->d_obs_idx = nullptr;
    b->d_obs_val = nullptr;
    b->obs_cap = 0;
    TR_HIP(hipMalloc(&b->d_obs_idx, sizeof(int32_t) * size_t(n)));
    TR_HIP(hipMalloc(&b->d_obs_val, sizeof(double) * 3 * size_t(n)));
    b->obs_cap = n;
  }
  TR_HIP(hipMemcpyAsync(b->d_obs_idx, pos, sizeof(int32_t) * size_t(n), hipMemcpyHostToDevice,
                        b->stream));
  const float* src[3] = {mm->inst.h[cur], mm->inst.qx[cur], mm->inst.qy[cur]};
  double* dst[3] = {h, qx, qy};
  for (int k = 0; k < 3; k++) {
    if (!dst[k]) continue;
    RC(copy_obs_f32(n, src[k], b->d_obs_val + size_t(k) * n, b->d_obs_idx, b->stream));
    TR_HIP(hipMemcpyAsync(dst[k], b->d_obs_val + size_t(k) * n, sizeof(double) * size_t(n),
                          hipMemcpyDeviceToHost, b->stream));
  }
  TR_HIP(hipStreamSynchronize(b->stream));
  return TR_OK;
}

int tr_batch_set_obs_halt(tr_batch* b, int32_t id, int32_t on) {
  Member* mm = member(b, id);
  if (!mm) return TR_ERR_INVALID;
  TR_HIP(hipStreamSynchronize(b->stream));
  mm->inst.halt_on_obs = on ? 1 : 0;
  RC(put_inst(b, id));
  TR_HIP(hipStreamSynchronize(b->stream));
  return TR_OK;
}

int64_t tr_batch_device_bytes(const tr_batch* b) { return b ? b->bytes : 0; }

int64_t tr_batch_member_cells(tr_batch* b, int32_t id) {
  Member* mm = member(b, id);
  return mm ? mm->interior : -1;
}

int32_t tr_batch_work_waves(const tr_batch* b) { return b ? b->nwork : 0; }
int32_t tr_batch_band_rows(const tr_batch* b) { return b ? b->band_rows : 0; }

void* tr_batch_stream(tr_batch* b) { return b ? reinterpret_cast<void*>(b->stream) : nullptr; }

void tr_batch_destroy(tr_batch* b) { delete b; }

}  // extern "C"

// extbc.hpp -- external boundary segments inside the fused row loops.
//
// The reference applies compute_extbc_values (src/kernels.h:236-424) after
// the open-boundary copies and before wet_dry: every listed cell on the
// domain's edge is given a value from its own state after the update (type
// 0: copied out to the ghost; 1: water level from a time series at the step's
// midpoint; 2: normal-slope velocity; 3: Froude velocity), and that value,
// times orientation factors, goes to the cell and to its ghost (and, at a
// corner, to the ghost beside it in the same row).  A fused step applies it
// to row r-1 right after that row's update: the cell's own value in its lane,
// the west/east ghost in the neighbour lane (DPP), the north ghost row (row
// 0, stored in the same pass) directly, and the south ghost row one row later
// (it is the next row the loop produces).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "common.hpp"

namespace tr {

// Per simulation: entry k of the reference's lists (cells in the order the
// .extbc file gives them) and a map from padded cell to entry.
struct ExtBC {
  int n;                  // entries (0: none)
  const int* map;         // [rows*cols] first entry of an edge cell, -1 elsewhere
  const int* next;        // [n] the cell's next entry in list order, -1 after the last
  const int* type;        // [n]
  const int* start;       // [n] first row of the entry's series in v1/v2
  const int* nvars;       // [n] rows of the series
  const double* v1;       // times (type 1) or the parameter (types 2, 3)
  const double* v2;       // water levels (type 1)
  const double* zc;       // [n] bed of the entry's cell
  const double* n2c;      // [n] squared Manning n of the entry's cell
};

// Ghost targets of an edge cell (bits).
constexpr int kGhostN = 1, kGhostS = 2, kGhostW = 4, kGhostE = 8;

// One edge cell, entry k (src/kernels.h:257-420, single process): in (h,qx,qy)
// the cell's state after the update; out (h,qx,qy) what the cell keeps and
// (gh,gqx,gqy) what its ghost cells get, with the ghost bits.  The factor and
// target logic follows the reference's order of tests, so an edge row that is
// both top and bottom, or a column both left and right, resolves the same way.
__device__ inline void extbc_cell(const ExtBC& b, int k, bool top, bool btm, bool lt, bool rt,
                                  double simtime, double dt, double& h, double& qx, double& qy,
                                  double& gh, double& gqx, double& gqy, int& ghosts) {
  double f0 = 1.0, f1 = 0.0, f2 = 0.0;
  int g = 0;
  if (lt) { g = kGhostW; f1 = -1.0; f2 = 0.0; }
  if (rt) { g = kGhostE; f1 = 1.0; f2 = 0.0; }
  if (top) {
    g = kGhostN; f1 = 0.0; f2 = 1.0;
    if (lt) { g |= kGhostW; f2 = 0.0; }
    if (rt) { g = kGhostN | kGhostE; f2 = 0.0; }
  }
  if (btm) {
    g = kGhostS; f1 = 0.0; f2 = -1.0;
    if (lt) { g |= kGhostW; f2 = 0.0; }
    if (rt) { g = kGhostS | kGhostE; f2 = 0.0; }
  }
  ghosts = g;
  const int bt = b.type[k], st = b.start[k];
  double hc = h, qxc = qx, qyc = qy;
  if (bt == 0) {  // the ghosts copy the cell, which keeps its state
    gh = hc; gqx = qxc; gqy = qyc;
    return;
  }
  if (bt == 1) {
    const double tm = 0.5 * (simtime + simtime + dt);
    int lo = -999, hi = b.nvars[k] - 1;
    if (tm < b.v1[st]) {
      lo = 0; hi = 0;
    } else if (tm > b.v1[hi + st]) {
      lo = hi;
    } else {
      for (int i = 0; i < hi; i++)
        if (b.v1[i + st] <= tm && b.v1[i + 1 + st] > tm) { lo = i; hi = i + 1; break; }
      if (lo < 0) lo = hi;
    }
    double level;
    if (lo == hi) level = b.v2[lo + st];
    else level = b.v2[lo + st] + (((b.v2[hi + st] - b.v2[lo + st]) * (tm - b.v1[lo + st])) /
                                  (b.v1[hi + st] - b.v1[lo + st]));
    hc = fmax(level - b.zc[k], 0.0);
  } else {
    const double p = b.v1[st];
    double vel = 0.0;
    if (bt == 2) vel = sqrt(p * hc * cbrt(hc) / fmax(b.n2c[k], 1e-6));
    if (bt == 3) vel = p * sqrt(kG * hc);
    qxc = hc * vel;
    qyc = hc * vel;
  }
  gh = hc * f0; gqx = qxc * f1; gqy = qyc * f2;
  h = gh; qx = gqx; qy = gqy;
}

// Carried from row R-2 to the south ghost row R-1.
template <class Real>
struct SouthCarry {
  Real ph, pqx, pqy;  // row R-2 before the segments (the open-boundary copy's source)
  Real gh, gqx, gqy;  // a segment's south ghost value
  bool flag;          // this column has one
};

// Row ru of a fused pass after its update and open-boundary copies.  n* is
// row ru in this lane's column, a* row ru-1 (the north ghost row when ru ==
// 1).  The reference copies row R-2 into the ghost row R-1 before the
// segments run, but a fused pass reaches row R-1 one row later: `open` makes
// that copy here from row R-2 as it was before the segments, then a
// segment's own south ghost value is applied.  Control flow is wave-uniform
// (ru, edge_we) so the lane shifts see every lane.
template <class Real>
__device__ __forceinline__ void extbc_row(const ExtBC& b, int ru, int c, int R, int C,
                                          bool col_int, bool edge_we, bool open, double simtime,
                                          double dt, Real& nh, Real& nqx, Real& nqy, Real& ah,
                                          Real& aqx, Real& aqy, SouthCarry<Real>& sc) {
  if (ru == R - 1) {  // the south ghost row
    if (open) { nh = sc.ph; nqx = sc.pqx; nqy = sc.pqy; }
    if (sc.flag) { nh = sc.gh; nqx = sc.gqx; nqy = sc.gqy; }
    return;
  }
  if (ru == R - 2) { sc.ph = nh; sc.pqx = nqx; sc.pqy = nqy; sc.flag = false; }
  const bool edge_row = ru == 1 || ru == R - 2;
  if (ru < 1 || ru > R - 2 || !(edge_row || edge_we)) return;
  int k = -1;
  if (col_int && (edge_row || c == 1 || c == C - 2)) k = b.map[(long long)ru * C + c];
  double gh = 0.0, gqx = 0.0, gqy = 0.0;
  int gs = 0;
  if (k >= 0) {
    // a cell listed more than once takes its entries in list order, each
    // from the state the previous one left (the drop-in kernel's order)
    double h = double(nh), qx = double(nqx), qy = double(nqy);
    for (; k >= 0; k = b.next[k])
      extbc_cell(b, k, ru == 1, ru == R - 2, c == 1, c == C - 2, simtime, dt, h, qx, qy, gh, gqx,
                 gqy, gs);
    nh = Real(h); nqx = Real(qx); nqy = Real(qy);
  }
  if (edge_we) {  // west ghost from the lane of column 1, east ghost from column C-2
    const double rh = from_right(gh), rqx = from_right(gqx), rqy = from_right(gqy);
    const double lh = from_left(gh), lqx = from_left(gqx), lqy = from_left(gqy);
    const int rg = dpp_from_right(gs), lg = dpp_from_left(gs);
    if (c == 0 && (rg & kGhostW)) { nh = Real(rh); nqx = Real(rqx); nqy = Real(rqy); }
    if (c == C - 1 && (lg & kGhostE)) { nh = Real(lh); nqx = Real(lqx); nqy = Real(lqy); }
  }
  if (ru == 1 && (gs & kGhostN)) { ah = Real(gh); aqx = Real(gqx); aqy = Real(gqy); }
  if (ru == R - 2) {
    sc.flag = (gs & kGhostS) != 0;
    sc.gh = Real(gh); sc.gqx = Real(gqx); sc.gqy = Real(gqy);
  }
}

// Host side: the arrays behind an ExtBC for a case with `n` entries (cells
// bc_index, types, series starts and lengths) on a padded rows x cols grid.
struct ExtBCHost {
  std::vector<int> map, next, type, start, nvars;
  std::vector<double> zc, n2c;
};

// Builds the edge map and the per-cell chains of entries in list order.  A
// fused pass places every entry by its cell, so a cell outside the interior
// (which the reference would use to write outside the ghost ring) is an
// error.  Entries off the edge do nothing, as in the reference.  Returns an
// error message, empty on success.
inline std::string build_extbc(int rows, int cols, int n, const int32_t* index,
                               const int32_t* type, const int32_t* start, const int32_t* nvars,
                               const double* dem, const double* nman2, ExtBCHost& out) {
  out = ExtBCHost{};
  if (n <= 0) return {};
  out.map.assign(size_t(rows) * size_t(cols), -1);
  for (int k = 0; k < n; k++) {
    const int ii = index[k];
    const int r = ii / cols, c = ii % cols;
    if (ii < 0 || r < 1 || r > rows - 2 || c < 1 || c > cols - 2)
      return "an external boundary cell outside the interior runs only in TR_FORM_F64_UNFUSED";
    out.type.push_back(type[k]);
    out.start.push_back(start[k]);
    out.nvars.push_back(nvars[k]);
    out.zc.push_back(dem[ii]);
    out.n2c.push_back(nman2 ? nman2[ii] : 0.0);
    out.next.push_back(-1);
    if (!(r == 1 || r == rows - 2 || c == 1 || c == cols - 2)) continue;
    if (out.map[size_t(ii)] < 0) {
      out.map[size_t(ii)] = k;
    } else {  // append to the cell's chain
      int j = out.map[size_t(ii)];
      while (out.next[size_t(j)] >= 0) j = out.next[size_t(j)];
      out.next[size_t(j)] = k;
    }
  }
  return {};
}

}  // namespace tr

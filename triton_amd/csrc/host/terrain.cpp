// terrain.cpp -- drainage of synthetic DEMs for the benchmark inputs.
//
// Priority-flood with an epsilon gradient (Barnes, Lehman & Mulla 2014,
// "Priority-Flood: An optimal depression-filling and watershed-labeling
// algorithm", Algorithm 3): cells are taken from a min-heap in increasing
// filled elevation, starting from the domain edge; an unvisited neighbour
// gets max(own elevation, parent elevation + eps) and drains through the
// cell it was reached from.  Reversing the visit order then accumulates the
// contributing area of every cell along those parents (D8 neighbourhood).
#include <cstdint>
#include <queue>
#include <utility>
#include <vector>

#include "../../../include/triton_amd.h"

extern "C" int tr_dem_flow_route(const double* dem, int32_t nr, int32_t nc, double eps,
                                 double* filled, double* acc) {
  if (!dem || !filled || !acc || nr < 1 || nc < 1) return TR_ERR_INVALID;
  const int64_t n = int64_t(nr) * nc;
  std::vector<int32_t> parent(size_t(n), -1);
  std::vector<uint8_t> seen(size_t(n), 0);
  std::vector<int32_t> order;
  order.reserve(size_t(n));
  using Item = std::pair<double, int32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  for (int64_t i = 0; i < n; i++) filled[i] = dem[i];
  for (int32_t r = 0; r < nr; r++)
    for (int32_t c = 0; c < nc; c++)
      if (r == 0 || c == 0 || r == nr - 1 || c == nc - 1) {
        int32_t i = r * nc + c;
        seen[size_t(i)] = 1;
        pq.push({filled[i], i});
      }
  static const int dr[8] = {-1, -1, -1, 0, 0, 1, 1, 1};
  static const int dc[8] = {-1, 0, 1, -1, 1, -1, 0, 1};
  while (!pq.empty()) {
    const int32_t i = pq.top().second;
    pq.pop();
    order.push_back(i);
    const int32_t r = i / nc, c = i % nc;
    for (int k = 0; k < 8; k++) {
      const int32_t rr = r + dr[k], cc = c + dc[k];
      if (rr < 0 || cc < 0 || rr >= nr || cc >= nc) continue;
      const int32_t j = rr * nc + cc;
      if (seen[size_t(j)]) continue;
      seen[size_t(j)] = 1;
      const double lift = filled[i] + eps;
      if (filled[j] < lift) filled[j] = lift;
      parent[size_t(j)] = i;
      pq.push({filled[j], j});
    }
  }
  for (int64_t i = 0; i < n; i++) acc[i] = 1.0;
  for (int64_t k = int64_t(order.size()) - 1; k >= 0; k--) {
    const int32_t i = order[size_t(k)];
    if (parent[size_t(i)] >= 0) acc[parent[size_t(i)]] += acc[i];
  }
  return TR_OK;
}

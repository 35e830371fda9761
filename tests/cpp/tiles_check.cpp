// Host-side invariants of the target-stationary M2L layout (fmm::build_m2l_tiles
// and its merge): every grouped pair appears exactly once, in the tile that
// owns its target; a slice holds at most 16 pairs with distinct targets whose
// groups all have the slice's representative, each pair carrying its group's
// reflection bits; every
// group's offset is its representative's reflected by its bits; a tile's
// targets are distinct and every target has one tile.  Prints "ok" or the
// first violation.
#include <algorithm>
#include <array>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <set>
#include <vector>

#include "fmm_host.hpp"

using namespace fmm;

static int check(const Tree &t, const M2LGroups &G, const M2LTiles &T) {
  if ((int)G.canon.size() != G.ngroup || (int)G.refl.size() != G.ngroup || (int)G.cdx.size() != 3 * G.ncanon)
    return printf("reflection classes: bad sizes\n"), 1;
  for (int g = 0; g < G.ngroup; g++) {
    const int c = G.canon[g];
    if (c < 0 || c >= G.ncanon) return printf("representative out of range\n"), 1;
    for (int d = 0; d < 3; d++) {
      const double want_d = ((G.refl[g] >> d) & 1) ? -G.cdx[3 * (size_t)c + d] : G.cdx[3 * (size_t)c + d];
      if (G.dx[3 * (size_t)g + d] != want_d || G.cdx[3 * (size_t)c + d] < 0)
        return printf("group %d is not its representative reflected\n", g), 1;
    }
  }
  std::map<std::pair<int, int>, int> want;  // (target, source) -> group
  for (int g = 0; g < G.ngroup; g++)
    for (int q = G.grp_ptr[g]; q < G.grp_ptr[g + 1]; q++) want[{G.tgt[q], G.src[q]}] = g;
  std::set<int> seen_t;
  size_t npairs = 0;
  for (int b = 0; b < T.ntile; b++) {
    std::set<int> tiles_t;
    for (int j = 0; j < T.TT; j++) {
      const int c = T.tgt[(size_t)b * T.TT + j];
      if (c < 0) continue;
      if (!tiles_t.insert(c).second || !seen_t.insert(c).second) return printf("target %d twice\n", c), 1;
    }
    for (int s = T.ptr[b]; s < T.ptr[b + 1]; s++) {
      const int rep = T.slice[4 * s], p0 = T.slice[4 * s + 1], n = T.slice[4 * s + 2];
      if (n < 1 || n > 16) return printf("slice size %d\n", n), 1;
      if (T.slice[4 * s + 3] != 0) return printf("slice field 3 not zero\n"), 1;
      std::set<int> lts;
      for (int k = p0; k < p0 + n; k++) {
        const int lt = T.lt[k] & 255, rf = T.lt[k] >> 8;
        if (!lts.insert(lt).second) return printf("target repeated in a slice\n"), 1;
        const int c = T.tgt[(size_t)b * T.TT + lt];
        auto it = want.find({c, T.src[k]});
        if (c < 0 || it == want.end()) return printf("pair not grouped\n"), 1;
        const int g = it->second;
        if (G.canon[g] != rep || G.refl[g] != rf) return printf("pair names the wrong representative\n"), 1;
        want.erase(it);
        npairs++;
      }
    }
  }
  if (!want.empty()) return printf("%zu grouped pairs missing from the tiles\n", want.size()), 1;
  // work items cover every tile's slices once, in order
  std::vector<int> next(T.ntile);
  for (int b = 0; b < T.ntile; b++) next[b] = T.ptr[b];
  // items in tile order (the launch order deals them to the XCDs)
  std::vector<std::array<int, 4>> items;
  for (size_t k = 0; k < T.item.size(); k += 4)
    if (T.item[k + 2] > T.item[k + 1]) items.push_back({T.item[k], T.item[k + 1], T.item[k + 2], T.item[k + 3]});
  std::sort(items.begin(), items.end());
  for (const auto &itm : items) {
    const int b = itm[0], a = itm[1], e = itm[2], split = itm[3];
    if (a != next[b] || e <= a || e > T.ptr[b + 1]) return printf("bad work item\n"), 1;
    if (split != ((a != T.ptr[b] || e != T.ptr[b + 1]) ? 1 : 0)) return printf("bad split flag\n"), 1;
    next[b] = e;
  }
  for (int b = 0; b < T.ntile; b++)
    if (next[b] != T.ptr[b + 1]) return printf("tile slices not covered by items\n"), 1;
  return 0;
}

int main(int argc, char **argv) {
  const int n = argc > 1 ? atoi(argv[1]) : 20000;
  const int TT = argc > 2 ? atoi(argv[2]) : 32;
  std::vector<Tree> trees;
  std::vector<Lists> lists;
  std::vector<M2LGroups> groups;
  std::vector<M2LTiles> tiles;
  const char *dists[3] = {"plummer1", "cube1", "plummer"};
  for (int k = 0; k < 3; k++) {
    const int m = n / (k + 1);
    std::vector<double> x(3 * (size_t)m), q(m);
    make_bodies(m, dists[k], k, x.data(), q.data());
    double lo[3], hi[3], xp[3] = {0, 0, 0};
    get_bounds(m, x.data(), lo, hi);
    trees.push_back(build_tree(m, x.data(), 64, lo, hi));
  }
  for (int k = 0; k < 3; k++) {
    double xp[3] = {0, 0, 0};
    lists.push_back(build_lists(trees[k], 0.4, 5000, xp));
    groups.push_back(build_m2l_groups(trees[k], lists[k], 512, k == 1));
    tiles.push_back(build_m2l_tiles(trees[k], groups[k], TT, 16));
    if (check(trees[k], groups[k], tiles[k])) return 1;
  }
  // forest of the three systems
  std::vector<PlanPart> parts;
  for (int k = 0; k < 3; k++) parts.push_back({&trees[k], &lists[k], &groups[k], &tiles[k]});
  Tree T;
  Lists L;
  M2LGroups G;
  M2LTiles X;
  merge_plans(parts, T, L, &G, 512, &X);
  if (check(T, G, X)) return 1;
  printf("ok\n");
  return 0;
}

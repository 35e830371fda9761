// fmm_host.cpp -- tree and list construction (see fmm_host.hpp).
// Compiled with -ffp-contract=off: the octant tests, the box arithmetic and the
// multipole acceptance criterion must round exactly like the reference's
// scalar x86 build, or a borderline pair changes list.
#include "fmm_host.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace fmm {

void get_bounds(int n, const double *x, double xmin[3], double xmax[3]) {
  if (n == 0) {
    for (int d = 0; d < 3; d++) xmin[d] = xmax[d] = 0;
    return;
  }
  for (int d = 0; d < 3; d++) xmin[d] = xmax[d] = x[d];
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      double lo = xmin[d] - 1e-5, hi = xmax[d] + 1e-5, v = x[3 * (size_t)i + d];
      xmin[d] = v < lo ? v : lo;
      xmax[d] = v > hi ? v : hi;
    }
  }
}

namespace {

// Octree node during the top-down build (include/build_tree.h:14-20).
struct Node {
  int ibody, nbody, nnode;
  int child[8];  // index into the node pool, -1 if absent
  double X[3];
};

constexpr int kMaxLevel = 30;

struct Builder {
  const double *x;
  int ncrit;
  std::vector<Node> pool;

  int make(int begin, int end, const double *X) {
    Node nd;
    nd.ibody = begin;
    nd.nbody = end - begin;
    nd.nnode = 1;
    for (int i = 0; i < 8; i++) nd.child[i] = -1;
    std::memcpy(nd.X, X, sizeof nd.X);
    pool.push_back(nd);
    return (int)pool.size() - 1;
  }

  static inline int octant(const double *xb, const double *X) {
    return (xb[0] > X[0]) + ((xb[1] > X[1]) << 1) + ((xb[2] > X[2]) << 2);
  }

  // buildNodes (include/build_tree.h:77-108).  `bodies` and `buffer` are
  // permutation arrays; recursion swaps them and a leaf reached with
  // direction set copies its range back, so the final order lands in the
  // array passed first at the top level.
  int build(int *bodies, int *buffer, int begin, int end, const double *X, double R0, int level,
            bool direction) {
    if (begin == end) return -1;
    // The reference recurses until a node holds <= ncrit bodies; more than
    // ncrit coincident bodies never separate and its radius arithmetic
    // (1 << level) breaks past level 30, so stop there with an oversized leaf.
    if (end - begin <= ncrit || level >= kMaxLevel) {
      if (direction) std::memcpy(buffer + begin, bodies + begin, sizeof(int) * (end - begin));
      return make(begin, end, X);
    }
    int self = make(begin, end, X);
    int nb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = begin; i < end; i++) nb[octant(x + 3 * (size_t)bodies[i], X)]++;
    int off[8], start[8];
    for (int i = 0, o = begin; i < 8; i++) {
      off[i] = start[i] = o;
      o += nb[i];
    }
    for (int i = begin; i < end; i++) {
      int b = bodies[i];
      buffer[off[octant(x + 3 * (size_t)b, X)]++] = b;
    }
    double r = R0 / (1 << (level + 1));
    int kids[8];
    for (int i = 0; i < 8; i++) {
      double Xc[3];
      for (int d = 0; d < 3; d++) {
        Xc[d] = X[d];
        Xc[d] += r * (((i & 1 << d) >> d) * 2 - 1);
      }
      kids[i] = build(buffer, bodies, start[i], start[i] + nb[i], Xc, R0, level + 1, !direction);
    }
    int nnode = 1;
    for (int i = 0; i < 8; i++) {
      pool[self].child[i] = kids[i];
      if (kids[i] >= 0) nnode += pool[kids[i]].nnode;
    }
    pool[self].nnode = nnode;
    return self;
  }
};

// nodes2cells (include/build_tree.h:123-165): a cell's children occupy a
// contiguous block; each child's descendants follow in child order.
void nodes2cells(const Builder &B, Tree &t, int node, int C, int CN, double R0, int level,
                 int iparent) {
  const Node &nd = B.pool[node];
  t.iparent[C] = iparent;
  t.R[C] = R0 / (1 << level);
  t.level[C] = level;
  std::memcpy(&t.X[3 * (size_t)C], nd.X, sizeof(double) * 3);
  t.nbody[C] = nd.nbody;
  t.ibody[C] = nd.ibody;
  if (nd.nnode == 1) {
    t.ichild[C] = 0;
    t.nchild[C] = 0;
    t.depth = std::max(t.depth, level);
    return;
  }
  int nchild = 0, kids[8];
  for (int i = 0; i < 8; i++)
    if (nd.child[i] >= 0) kids[nchild++] = nd.child[i];
  int Ci = CN;
  t.ichild[C] = Ci;
  t.nchild[C] = nchild;
  CN += nchild;
  for (int i = 0; i < nchild; i++) {
    nodes2cells(B, t, kids[i], Ci, CN, R0, level + 1, C);
    Ci++;
    CN += B.pool[kids[i]].nnode - 1;
  }
  t.depth = std::max(t.depth, level + 1);
}

}  // namespace

Tree build_tree(int n, const double *x, int ncrit, const double xmin[3], const double xmax[3]) {
  Tree t;
  t.n = n;
  t.perm.resize(n);
  if (n == 0) {
    t.level_ptr.assign(1, 0);
    return t;
  }
  // bounds2box (include/build_tree.h:168-180)
  double X[3], R = 0;
  for (int d = 0; d < 3; d++) X[d] = (xmax[d] + xmin[d]) / 2;
  for (int d = 0; d < 3; d++) {
    R = std::max(X[d] - xmin[d], R);
    R = std::max(xmax[d] - X[d], R);
  }
  R *= 1.00001;
  std::vector<int> buffer(n);
  for (int i = 0; i < n; i++) t.perm[i] = i;
  Builder B{x, ncrit, {}};
  B.pool.reserve(2 * (size_t)n / std::max(ncrit, 1) + 64);
  int root = B.build(t.perm.data(), buffer.data(), 0, n, X, R, 0, false);
  int nc = B.pool[root].nnode;
  t.ncell = nc;
  t.iparent.resize(nc);
  t.ichild.resize(nc);
  t.nchild.resize(nc);
  t.ibody.resize(nc);
  t.nbody.resize(nc);
  t.level.resize(nc);
  t.X.resize(3 * (size_t)nc);
  t.R.resize(nc);
  nodes2cells(B, t, root, 0, 1, R, 0, 0);
  // level index and leaf list
  t.level_ptr.assign(t.depth + 2, 0);
  for (int c = 0; c < nc; c++) t.level_ptr[t.level[c] + 1]++;
  for (int l = 0; l <= t.depth; l++) t.level_ptr[l + 1] += t.level_ptr[l];
  t.level_cells.resize(nc);
  std::vector<int> fill(t.level_ptr.begin(), t.level_ptr.end() - 1);
  for (int c = 0; c < nc; c++) t.level_cells[fill[t.level[c]]++] = c;
  for (int c = 0; c < nc; c++)
    if (t.nchild[c] == 0) t.leaves.push_back(c);
  return t;
}

std::string adopt_tree(Tree &t, int n, int ncell, const int *iparent, const int *ichild,
                       const int *nchild, const int *ibody, const int *nbody, const double *X,
                       const double *R) {
  t = Tree();
  t.n = n;
  t.ncell = ncell;
  t.perm.resize(n);
  for (int i = 0; i < n; i++) t.perm[i] = i;
  if (ncell == 0) {
    t.level_ptr.assign(1, 0);
    return n == 0 ? "" : "bodies without cells";
  }
  t.iparent.assign(iparent, iparent + ncell);
  t.ichild.assign(ichild, ichild + ncell);
  t.nchild.assign(nchild, nchild + ncell);
  t.ibody.assign(ibody, ibody + ncell);
  t.nbody.assign(nbody, nbody + ncell);
  t.X.assign(X, X + 3 * (size_t)ncell);
  t.R.assign(R, R + ncell);
  t.level.assign(ncell, -1);
  if (t.ibody[0] != 0 || t.nbody[0] != n) return "root must hold all bodies";
  t.level[0] = 0;
  int64_t leaf_bodies = 0;
  // children are numbered after their parent, so one pass in cell order
  // assigns every level (cells out of order are rejected below)
  for (int c = 0; c < ncell; c++) {
    if (t.level[c] < 0) return "cell " + std::to_string(c) + " is not reached from the root";
    if (t.nbody[c] < 0 || t.ibody[c] < 0 || (int64_t)t.ibody[c] + t.nbody[c] > n)
      return "cell " + std::to_string(c) + " has a body range outside [0, n)";
    if (t.nchild[c] < 0 || t.nchild[c] > 8) return "cell " + std::to_string(c) + " has a bad NCHILD";
    if (t.nchild[c] == 0) {
      leaf_bodies += t.nbody[c];
      t.leaves.push_back(c);
      continue;
    }
    if (t.ichild[c] <= c || (int64_t)t.ichild[c] + t.nchild[c] > ncell)
      return "cell " + std::to_string(c) + " has a bad child block";
    // the children's body ranges tile the parent's range in order
    int64_t next = t.ibody[c];
    for (int k = t.ichild[c]; k < t.ichild[c] + t.nchild[c]; k++) {
      if (t.level[k] >= 0 || t.iparent[k] != c)
        return "cell " + std::to_string(k) + " does not name " + std::to_string(c) + " as parent";
      if (t.ibody[k] != next) return "children of cell " + std::to_string(c) + " do not tile its bodies";
      t.level[k] = t.level[c] + 1;
      next += t.nbody[k];
    }
    if (next != (int64_t)t.ibody[c] + t.nbody[c])
      return "children of cell " + std::to_string(c) + " do not tile its bodies";
  }
  if (leaf_bodies != n) return "leaves do not cover every body once";
  for (int c = 0; c < ncell; c++) t.depth = std::max(t.depth, t.level[c]);
  t.level_ptr.assign(t.depth + 2, 0);
  for (int c = 0; c < ncell; c++) t.level_ptr[t.level[c] + 1]++;
  for (int l = 0; l <= t.depth; l++) t.level_ptr[l + 1] += t.level_ptr[l];
  t.level_cells.resize(ncell);
  std::vector<int> fill(t.level_ptr.begin(), t.level_ptr.end() - 1);
  for (int c = 0; c < ncell; c++) t.level_cells[fill[t.level[c]]++] = c;
  return "";
}

namespace {

struct Walker {
  const Tree &t;
  double theta;
  int nspawn;
  double xper[3];
  std::vector<int> m2l_i, m2l_j, p2p_i, p2p_j;

  void dual(int Ci, int Cj) {
    // dualTreeTraversal (include/traversal.h:117-177)
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = t.X[3 * (size_t)Ci + d] - t.X[3 * (size_t)Cj + d] - xper[d];
    double RT2 = (dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2]) * theta * theta;
    double Rs = t.R[Ci] + t.R[Cj];
    if (RT2 > Rs * Rs * (1 - 1e-3)) {
      m2l_i.push_back(Ci);
      m2l_j.push_back(Cj);
    } else if (t.nchild[Ci] == 0 && t.nchild[Cj] == 0) {
      if (t.nbody[Cj] == 0) {
        m2l_i.push_back(Ci);
        m2l_j.push_back(Cj);
      } else {
        p2p_i.push_back(Ci);
        p2p_j.push_back(Cj);
      }
    } else {
      split(Ci, Cj);
    }
  }

  // TraverseRange::operator() executed in task order (traversal.h:196-229)
  void range(int ib, int ie, int jb, int je) {
    if (ie - ib == 1 || je - jb == 1) {
      if (ib == jb) {
        dual(ib, jb);
      } else {
        for (int i = ib; i < ie; i++)
          for (int j = jb; j < je; j++) dual(i, j);
      }
    } else {
      int im = ib + (ie - ib) / 2, jm = jb + (je - jb) / 2;
      range(ib, im, jb, jm);
      range(im, ie, jm, je);
      range(ib, im, jm, je);
      range(im, ie, jb, jm);
    }
  }

  // splitCell (include/traversal.h:90-114)
  void split(int Ci, int Cj) {
    if (t.nchild[Cj] == 0) {
      for (int ci = t.ichild[Ci]; ci < t.ichild[Ci] + t.nchild[Ci]; ci++) dual(ci, Cj);
    } else if (t.nchild[Ci] == 0) {
      for (int cj = t.ichild[Cj]; cj < t.ichild[Cj] + t.nchild[Cj]; cj++) dual(Ci, cj);
    } else if (t.nbody[Ci] + t.nbody[Cj] >= nspawn || Ci == Cj) {
      range(t.ichild[Ci], t.ichild[Ci] + t.nchild[Ci], t.ichild[Cj], t.ichild[Cj] + t.nchild[Cj]);
    } else if (t.R[Ci] >= t.R[Cj]) {
      for (int ci = t.ichild[Ci]; ci < t.ichild[Ci] + t.nchild[Ci]; ci++) dual(ci, Cj);
    } else {
      for (int cj = t.ichild[Cj]; cj < t.ichild[Cj] + t.nchild[Cj]; cj++) dual(Ci, cj);
    }
  }
};

// Stable counting sort of (target, source) pairs by target: each target's
// list keeps the traversal's visiting order.
void to_csr(int ncell, const std::vector<int> &ti, const std::vector<int> &sj,
            std::vector<int> &ptr, std::vector<int> &src, std::vector<int> &targets) {
  ptr.assign(ncell + 1, 0);
  for (int c : ti) ptr[c + 1]++;
  for (int c = 0; c < ncell; c++) ptr[c + 1] += ptr[c];
  src.resize(ti.size());
  std::vector<int> fill(ptr.begin(), ptr.end() - 1);
  for (size_t k = 0; k < ti.size(); k++) src[fill[ti[k]]++] = sj[k];
  targets.clear();
  for (int c = 0; c < ncell; c++)
    if (ptr[c + 1] > ptr[c]) targets.push_back(c);
}

}  // namespace

Lists build_lists(const Tree &t, double theta, int nspawn, const double xper[3]) {
  Lists L;
  L.m2l_ptr.assign(t.ncell + 1, 0);
  L.p2p_ptr.assign(t.ncell + 1, 0);
  if (t.ncell == 0) return L;
  Walker w{t, theta, nspawn, {0, 0, 0}, {}, {}, {}, {}};
  if (xper) std::memcpy(w.xper, xper, sizeof w.xper);
  w.m2l_i.reserve(64 * (size_t)t.ncell);
  w.m2l_j.reserve(64 * (size_t)t.ncell);
  w.p2p_i.reserve(32 * (size_t)t.leaves.size());
  w.p2p_j.reserve(32 * (size_t)t.leaves.size());
  w.dual(0, 0);
  L.nm2l = (int64_t)w.m2l_i.size();
  L.np2p = (int64_t)w.p2p_i.size();
  to_csr(t.ncell, w.m2l_i, w.m2l_j, L.m2l_ptr, L.m2l_src, L.m2l_targets);
  to_csr(t.ncell, w.p2p_i, w.p2p_j, L.p2p_ptr, L.p2p_src, L.p2p_targets);
  int64_t bp = 0;
  for (size_t k = 0; k < w.p2p_i.size(); k++)
    bp += (int64_t)t.nbody[w.p2p_i[k]] * t.nbody[w.p2p_j[k]];
  L.p2p_body_pairs = bp;
  return L;
}

namespace {

// Open-addressing hash of 64-bit keys to dense ids (first-appearance order).
struct KeyIndex {
  std::vector<uint64_t> keys;
  std::vector<int> ids;
  size_t mask = 0;
  int count = 0;
  static constexpr uint64_t kEmpty = ~0ull;
  explicit KeyIndex(size_t cap = 1 << 12) { rehash(cap); }
  static inline uint64_t mix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
  }
  void rehash(size_t cap) {
    std::vector<uint64_t> ok;
    std::vector<int> oi;
    ok.swap(keys);
    oi.swap(ids);
    keys.assign(cap, kEmpty);
    ids.assign(cap, -1);
    mask = cap - 1;
    for (size_t i = 0; i < ok.size(); i++)
      if (ok[i] != kEmpty) place(ok[i], oi[i]);
  }
  void place(uint64_t k, int id) {
    size_t h = mix(k) & mask;
    while (keys[h] != kEmpty) h = (h + 1) & mask;
    keys[h] = k;
    ids[h] = id;
  }
  int get(uint64_t k, bool &fresh) {
    size_t h = mix(k) & mask;
    while (keys[h] != kEmpty) {
      if (keys[h] == k) {
        fresh = false;
        return ids[h];
      }
      h = (h + 1) & mask;
    }
    fresh = true;
    keys[h] = k;
    ids[h] = count++;
    if (2 * (size_t)count > keys.size()) rehash(2 * keys.size());
    return count - 1;
  }
};

}  // namespace

M2LGroups build_m2l_groups(const Tree &t, const Lists &L, int chunk_pairs) {
  M2LGroups G;
  const int nc = t.ncell;
  G.rest_ptr.assign(nc + 1, 0);
  if (nc == 0 || L.nm2l == 0) {
    G.grp_ptr.assign(1, 0);
    return G;
  }
  const double R0 = t.R[0];
  const int64_t np = L.nm2l;
  std::vector<int> gid(np);
  std::vector<int64_t> kvec;  // 3 per group
  std::vector<int> gexp;      // e per group
  KeyIndex index;
  const int64_t lim = 1 << 18;
  for (int c = 0; c < nc; c++) {
    for (int p = L.m2l_ptr[c]; p < L.m2l_ptr[c + 1]; p++) {
      const int sc = L.m2l_src[p];
      int e = std::max(t.level[c], t.level[sc]);
      const double scale = std::ldexp(1.0, e) / R0;
      int64_t k[3];
      bool ok = e < 64;
      for (int d = 0; d < 3 && ok; d++) {
        const double v = (t.X[3 * (size_t)c + d] - t.X[3 * (size_t)sc + d]) * scale;
        const double r = std::nearbyint(v);
        if (!(std::fabs(v - r) < 1e-6) || std::fabs(r) >= (double)(lim << 8)) ok = false;
        k[d] = (int64_t)r;
      }
      if (ok) {
        while (e > 0 && ((k[0] | k[1] | k[2]) & 1) == 0) {
          k[0] /= 2;
          k[1] /= 2;
          k[2] /= 2;
          e--;
        }
        for (int d = 0; d < 3; d++)
          if (k[d] <= -lim || k[d] >= lim) ok = false;
      }
      if (!ok) {
        gid[p] = -1;
        continue;
      }
      const uint64_t key = ((uint64_t)e << 57) | ((uint64_t)(k[0] + lim) << 38) |
                           ((uint64_t)(k[1] + lim) << 19) | (uint64_t)(k[2] + lim);
      bool fresh;
      const int g = index.get(key, fresh);
      if (fresh) {
        kvec.insert(kvec.end(), {k[0], k[1], k[2]});
        gexp.push_back(e);
      }
      gid[p] = g;
    }
  }
  const int ng = index.count;
  G.ngroup = ng;
  G.dx.resize(3 * (size_t)ng);
  for (int g = 0; g < ng; g++) {
    const double u = R0 / std::ldexp(1.0, gexp[g]);
    for (int d = 0; d < 3; d++) G.dx[3 * (size_t)g + d] = (double)kvec[3 * (size_t)g + d] * u;
  }
  // counting sort by group, CSR order kept inside a group
  G.grp_ptr.assign(ng + 1, 0);
  for (int64_t p = 0; p < np; p++)
    if (gid[p] >= 0) G.grp_ptr[gid[p] + 1]++;
  for (int g = 0; g < ng; g++) G.grp_ptr[g + 1] += G.grp_ptr[g];
  const int64_t ngp = G.grp_ptr[ng];
  G.tgt.resize(ngp);
  G.src.resize(ngp);
  std::vector<int> fill(G.grp_ptr.begin(), G.grp_ptr.end() - 1);
  std::vector<int> rest_t, rest_s;
  for (int c = 0; c < nc; c++) {
    for (int p = L.m2l_ptr[c]; p < L.m2l_ptr[c + 1]; p++) {
      const int g = gid[p];
      if (g >= 0) {
        const int q = fill[g]++;
        G.tgt[q] = c;
        G.src[q] = L.m2l_src[p];
      } else {
        G.rest_ptr[c + 1]++;
        rest_s.push_back(L.m2l_src[p]);
      }
    }
  }
  for (int c = 0; c < nc; c++) G.rest_ptr[c + 1] += G.rest_ptr[c];
  G.rest_src = std::move(rest_s);
  G.nrest = (int64_t)G.rest_src.size();
  for (int c = 0; c < nc; c++)
    if (G.rest_ptr[c + 1] > G.rest_ptr[c]) G.rest_targets.push_back(c);
  // Chunks sorted by their first target (cell index ranges are subtrees, so
  // spatially compact), then dealt to the XCDs in contiguous runs: block b of
  // the launch runs on XCD b % kXcd, so XCD x gets the x-th eighth of the
  // domain and its concurrent chunks share source multipoles in its L2.
  // Empty chunks pad the shorter runs.
  std::vector<std::array<int, 4>> ch;
  for (int g = 0; g < ng; g++)
    for (int b = G.grp_ptr[g]; b < G.grp_ptr[g + 1]; b += chunk_pairs)
      ch.push_back({g, b, std::min(chunk_pairs, G.grp_ptr[g + 1] - b), 0});
  std::stable_sort(ch.begin(), ch.end(), [&](const std::array<int, 4> &a, const std::array<int, 4> &b) {
    return G.tgt[a[1]] < G.tgt[b[1]];
  });
  const int kXcd = 8;
  const size_t nch = ch.size(), per = (nch + kXcd - 1) / kXcd;
  G.chunk.assign(4 * per * kXcd, 0);
  for (size_t i = 0; i < per; i++)
    for (int x = 0; x < kXcd; x++) {
      const size_t src = (size_t)x * per + i, dst = i * kXcd + x;
      if (src < nch)
        for (int k = 0; k < 4; k++) G.chunk[4 * dst + k] = ch[src][k];
    }
  return G;
}

namespace {
void reference_charges(int n, int seed, double *q) {  // initSource, dataset.h:142-159
  srand48(seed);
  double average = 0;
  for (int i = 0; i < n; i++) {
    q[i] = drand48() - .5;
    average += q[i];
  }
  average /= n;
  for (int i = 0; i < n; i++) q[i] -= average;
}
}  // namespace

bool make_bodies(int n, const char *dist, int seed, double *x, double *q) {
  if (n <= 0) return true;
  const std::string d(dist ? dist : "");
  if (d == "cube") {
    srand48(seed);
    for (int i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) x[3 * (size_t)i + k] = drand48() * 2 * M_PI - M_PI;
    reference_charges(n, seed, q);
  } else if (d == "sphere") {
    srand48(seed);
    for (int i = 0; i < n; i++) {
      double *b = x + 3 * (size_t)i;
      for (int k = 0; k < 3; k++) b[k] = drand48() * 2 - 1;
      double r = std::sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
      for (int k = 0; k < 3; k++) b[k] *= M_PI / r;
    }
    reference_charges(n, seed, q);
  } else if (d == "octant") {
    srand48(seed);
    for (int i = 0; i < n; i++) {
      double th = drand48() * M_PI * 0.5, ph = drand48() * M_PI * 0.5;
      double *b = x + 3 * (size_t)i;
      b[0] = 2 * M_PI * std::sin(th) * std::cos(ph) - M_PI;
      b[1] = 2 * M_PI * std::sin(th) * std::sin(ph) - M_PI;
      b[2] = 2 * M_PI * std::cos(th) - M_PI;
    }
    reference_charges(n, seed, q);
  } else if (d == "plummer" || d == "plummer1") {
    const bool unit = d == "plummer1";
    const double rmax = unit ? 10.0 : 100.0, scale = unit ? 1.0 : 3.0 * M_PI / 16.0;
    srand48(seed);
    int i = 0;
    while (i < n) {
      double X1 = drand48(), X2 = drand48(), X3 = drand48();
      double R = 1.0 / std::sqrt((std::pow(X1, -2.0 / 3.0) - 1.0));
      if (R < rmax) {
        double Z = (1.0 - 2.0 * X2) * R;
        double X = std::sqrt(R * R - Z * Z) * std::cos(2.0 * M_PI * X3);
        double Y = std::sqrt(R * R - Z * Z) * std::sin(2.0 * M_PI * X3);
        X *= scale;
        Y *= scale;
        Z *= scale;
        x[3 * (size_t)i] = X;
        x[3 * (size_t)i + 1] = Y;
        x[3 * (size_t)i + 2] = Z;
        i++;
      }
    }
    if (unit)
      for (int k = 0; k < n; k++) q[k] = 1.0 / n;
    else
      reference_charges(n, seed, q);
  } else if (d == "cube1") {
    srand48(seed);
    for (int i = 0; i < n; i++)
      for (int k = 0; k < 3; k++) x[3 * (size_t)i + k] = 2 * drand48() - 1;
    srand48(seed);
    double average = 0;
    for (int i = 0; i < n; i++) {
      q[i] = 2 * drand48() - 1;
      average += q[i];
    }
    average /= n;
    for (int i = 0; i < n; i++) q[i] -= average;
  } else {
    return false;
  }
  return true;
}

}  // namespace fmm

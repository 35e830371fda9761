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
#include <deque>
#include <memory>
#include <string>
#include <vector>
#include <thread>
#include <atomic>

namespace fmm {

namespace {

// The scan step of getBounds (include/bound_box.h:16-23) for one coordinate:
// m <- (v < m - 1e-5) ? v : m - 1e-5 (low side), m <- (v > m + 1e-5) ? v :
// m + 1e-5 (high side).  Both are monotone in m and commute with min / max,
// so a run of bodies acts on the incoming value as m -> min(c, s^k(m)), c
// being the run's scan from +inf (-inf) and s^k the k-fold rounded shift: runs
// are scanned in parallel and chained in order, with the same result, bit for
// bit, as the one sequential scan.

// s^k(a): k times a - 1e-5 (up: a + 1e-5), each rounded to double.  Inside one
// binade the grid spacing u is fixed and a grid value moves by the same
// rounded step d = -+round_u(1e-5) (1e-5 rounds without a tie on every grid
// of spacing 2^-67 and above, i.e. every binade wide enough for a jump), so
// j steps are the grid point a + j d, computed exactly by one fma.  Near a
// binade edge, or where the steps are not yet inside one binade, steps are
// taken one at a time.
double drift(double a, int64_t k, bool up) {
  const double c = 1e-5;
  while (k > 0) {
    const double b = up ? a + c : a - c;
    --k;
    if (k == 0 || !std::isfinite(b)) return b;
    const double d = b - a;
    if (d == 0) return b;  // a -+ 1e-5 rounds to a: fixed from here on
    int ea = 0, eb = 0;
    (void)std::frexp(a, &ea);
    (void)std::frexp(b, &eb);
    if (std::isnormal(a) && std::isnormal(b) && (a > 0) == (b > 0) && ea == eb) {
      // |x| in [bot, top) on this binade; room for steps that keep both the
      // point and its exact shifted value inside it (two steps to spare)
      const double top = std::ldexp(1.0, eb), bot = std::ldexp(1.0, eb - 1);
      const bool grow = (d > 0) == (b > 0);
      const double ab = std::fabs(b), ad = std::fabs(d);
      const double room = grow ? top - ab - c : ab - c - bot;
      const double jj = std::floor(room / ad) - 2;
      if (jj >= 1) {
        const int64_t j = (int64_t)std::min(jj, (double)k);
        a = std::fma((double)j, d, b);
        k -= j;
        continue;
      }
    }
    a = b;
  }
  return a;
}

}  // namespace

void get_bounds(int n, const double *x, double xmin[3], double xmax[3]) {
  if (n == 0) {
    for (int d = 0; d < 3; d++) xmin[d] = xmax[d] = 0;
    return;
  }
  for (int d = 0; d < 3; d++) xmin[d] = xmax[d] = x[d];
  const int T = n >= (1 << 16) ? host_threads() : 1;
  if (T == 1) {
    for (int i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++) {
        double lo = xmin[d] - 1e-5, hi = xmax[d] + 1e-5, v = x[3 * (size_t)i + d];
        xmin[d] = v < lo ? v : lo;
        xmax[d] = v > hi ? v : hi;
      }
    }
    return;
  }
  // runs of bodies scanned in parallel from +-inf, then chained in order
  std::vector<std::array<double, 6>> run(T);
  std::vector<std::thread> th;
  for (int t = 0; t < T; t++)
    th.emplace_back([&, t] {
      const int i0 = (int)((int64_t)n * t / T), i1 = (int)((int64_t)n * (t + 1) / T);
      double lo3[3], hi3[3];
      for (int d = 0; d < 3; d++) {
        lo3[d] = HUGE_VAL;
        hi3[d] = -HUGE_VAL;
      }
      for (int i = i0; i < i1; i++)
        for (int d = 0; d < 3; d++) {
          double lo = lo3[d] - 1e-5, hi = hi3[d] + 1e-5, v = x[3 * (size_t)i + d];
          lo3[d] = v < lo ? v : lo;
          hi3[d] = v > hi ? v : hi;
        }
      for (int d = 0; d < 3; d++) {
        run[t][d] = lo3[d];
        run[t][3 + d] = hi3[d];
      }
    });
  for (auto &h : th) h.join();
  for (int t = 0; t < T; t++) {
    const int64_t k = (int64_t)n * (t + 1) / T - (int64_t)n * t / T;
    for (int d = 0; d < 3; d++) {
      const double lo = drift(xmin[d], k, false), hi = drift(xmax[d], k, true);
      xmin[d] = run[t][d] < lo ? run[t][d] : lo;
      xmax[d] = run[t][3 + d] > hi ? run[t][3 + d] : hi;
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
  // subtrees below `defer_level` are left as tasks (their roots made as
  // placeholders) for other builders; -1: build everything here
  struct Task {
    int node;  // placeholder in this pool
    int *bodies, *buffer;
    int begin, end, level;
    bool direction;
    double X[3];
  };
  int defer_level = -1;
  std::vector<Task> *tasks = nullptr;
  int par_threads = 1;  // threads for the partition of large nodes

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
    if (level == defer_level) {
      const int node = make(begin, end, X);
      tasks->push_back({node, bodies, buffer, begin, end, level, direction, {X[0], X[1], X[2]}});
      return node;
    }
    int self = make(begin, end, X);
    int nb[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int start[8];
    if (par_threads > 1 && end - begin >= 65536) {
      // stable partition by octant on host threads: thread k counts and then
      // scatters its contiguous share after the shares of threads < k
      const int T = par_threads;
      std::vector<std::array<int, 8>> cnt(T), off(T);
      auto run = [&](auto fn) {
        std::vector<std::thread> th;
        for (int k = 1; k < T; k++) th.emplace_back(fn, k);
        fn(0);
        for (auto &w : th) w.join();
      };
      run([&](int k) {
        cnt[k].fill(0);
        const int lo = begin + (int)((int64_t)(end - begin) * k / T), hi = begin + (int)((int64_t)(end - begin) * (k + 1) / T);
        for (int i = lo; i < hi; i++) cnt[k][octant(x + 3 * (size_t)bodies[i], X)]++;
      });
      for (int k = 0; k < T; k++)
        for (int o = 0; o < 8; o++) nb[o] += cnt[k][o];
      for (int o = 0, a = begin; o < 8; o++) {
        start[o] = a;
        for (int k = 0; k < T; k++) {
          off[k][o] = a;
          a += cnt[k][o];
        }
      }
      run([&](int k) {
        const int lo = begin + (int)((int64_t)(end - begin) * k / T), hi = begin + (int)((int64_t)(end - begin) * (k + 1) / T);
        for (int i = lo; i < hi; i++) {
          const int b = bodies[i];
          buffer[off[k][octant(x + 3 * (size_t)b, X)]++] = b;
        }
      });
    } else {
      for (int i = begin; i < end; i++) nb[octant(x + 3 * (size_t)bodies[i], X)]++;
      int off[8];
      for (int i = 0, o = begin; i < 8; i++) {
        off[i] = start[i] = o;
        o += nb[i];
      }
      for (int i = begin; i < end; i++) {
        int b = bodies[i];
        buffer[off[octant(x + 3 * (size_t)b, X)]++] = b;
      }
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

// bounds2box (include/build_tree.h:168-180)
void root_box(const double xmin[3], const double xmax[3], double X[3], double &R) {
  R = 0;
  for (int d = 0; d < 3; d++) X[d] = (xmax[d] + xmin[d]) / 2;
  for (int d = 0; d < 3; d++) {
    R = std::max(X[d] - xmin[d], R);
    R = std::max(xmax[d] - X[d], R);
  }
  R *= 1.00001;
}

void finish_tree(Tree &t) {
  const int nc = t.ncell;
  t.depth = 0;
  for (int c = 0; c < nc; c++) t.depth = std::max(t.depth, t.level[c]);
  t.level_ptr.assign(t.depth + 2, 0);
  for (int c = 0; c < nc; c++) t.level_ptr[t.level[c] + 1]++;
  for (int l = 0; l <= t.depth; l++) t.level_ptr[l + 1] += t.level_ptr[l];
  t.level_cells.resize(nc);
  std::vector<int> fill(t.level_ptr.begin(), t.level_ptr.end() - 1);
  for (int c = 0; c < nc; c++) t.level_cells[fill[t.level[c]]++] = c;
  t.leaves.clear();
  for (int c = 0; c < nc; c++)
    if (t.nchild[c] == 0) t.leaves.push_back(c);
}

Tree build_tree(int n, const double *x, int ncrit, const double xmin[3], const double xmax[3]) {
  Tree t;
  t.n = n;
  t.perm.resize(n);
  if (n == 0) {
    t.level_ptr.assign(1, 0);
    return t;
  }
  double X[3], R;
  root_box(xmin, xmax, X, R);
  std::vector<int> buffer(n);
  for (int i = 0; i < n; i++) t.perm[i] = i;
  Builder B{x, ncrit, {}};
  B.pool.reserve(2 * (size_t)n / std::max(ncrit, 1) + 64);
  const int nth = n >= 100000 ? host_threads() : 1;
  std::vector<Builder::Task> tasks;
  if (nth > 1) {  // levels 0-1 here (partitions threaded), the subtrees below on host threads
    B.defer_level = 2;
    B.tasks = &tasks;
    B.par_threads = nth;
  }
  int root = B.build(t.perm.data(), buffer.data(), 0, n, X, R, 0, false);
  if (!tasks.empty()) {
    const int ntop = (int)B.pool.size();
    std::vector<Builder> sub(tasks.size(), Builder{x, ncrit, {}});
    std::vector<int> sroot(tasks.size());
    std::atomic<size_t> next{0};
    auto work = [&] {
      for (size_t k; (k = next.fetch_add(1)) < tasks.size();) {
        const Builder::Task &tk = tasks[k];
        sroot[k] = sub[k].build(tk.bodies, tk.buffer, tk.begin, tk.end, tk.X, R, tk.level, tk.direction);
      }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nth; i++) th.emplace_back(work);
    work();
    for (auto &w : th) w.join();
    // stitch: each subtree's nodes appended, its root copied onto the placeholder
    for (size_t k = 0; k < tasks.size(); k++) {
      const int off = (int)B.pool.size();
      for (Node nd : sub[k].pool) {
        for (int i = 0; i < 8; i++)
          if (nd.child[i] >= 0) nd.child[i] += off;
        B.pool.push_back(nd);
      }
      B.pool[tasks[k].node] = B.pool[off + sroot[k]];
    }
    // node counts of the top levels: children were made after their parents
    for (int i = ntop - 1; i >= 0; i--) {
      bool placeholder = false;
      for (const Builder::Task &tk : tasks) placeholder = placeholder || tk.node == i;
      if (placeholder) continue;
      int nnode = 1;
      for (int c = 0; c < 8; c++)
        if (B.pool[i].child[c] >= 0) nnode += B.pool[B.pool[i].child[c]].nnode;
      B.pool[i].nnode = nnode;
    }
  }
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
  finish_tree(t);  // level index and leaf list
  return t;
}

std::string adopt_tree(Tree &t, int n, int ncell, const int *iparent, const int *ichild,
                       const int *nchild, const int *ibody, const int *nbody, const double *X,
                       const double *R) {
  t = Tree();
  t.n = n;
  t.ncell = ncell;
  t.adopted = true;
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
  // frontier mode: cell pairs holding at most `cut` bodies are recorded as
  // tasks (in visiting order) instead of being traversed
  int cut = 0;
  std::vector<std::pair<int, int>> *tasks = nullptr;
  int tag = 0;  // image bits of the source codes this walker emits

  void dual(int Ci, int Cj) {
    // dualTreeTraversal (include/traversal.h:117-177)
    double dX[3];
    for (int d = 0; d < 3; d++) dX[d] = t.X[3 * (size_t)Ci + d] - t.X[3 * (size_t)Cj + d] - xper[d];
    double RT2 = (dX[0] * dX[0] + dX[1] * dX[1] + dX[2] * dX[2]) * theta * theta;
    double Rs = t.R[Ci] + t.R[Cj];
    const bool accept = RT2 > Rs * Rs * (1 - 1e-3);
    if (tasks && (accept || (t.nchild[Ci] == 0 && t.nchild[Cj] == 0) || t.nbody[Ci] + t.nbody[Cj] <= cut)) {
      tasks->emplace_back(Ci, Cj);  // emitted pair or subtree pair: replayed in order by a worker
      return;
    }
    if (accept) {
      m2l_i.push_back(Ci);
      m2l_j.push_back(Cj | tag);
    } else if (t.nchild[Ci] == 0 && t.nchild[Cj] == 0) {
      if (t.nbody[Cj] == 0) {
        m2l_i.push_back(Ci);
        m2l_j.push_back(Cj | tag);
      } else {
        p2p_i.push_back(Ci);
        p2p_j.push_back(Cj | tag);
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

// Stable counting sort of (target, source) pairs by target over segments
// taken in order: each target's list keeps the traversal's visiting order.
struct Seg {
  const std::vector<int> *ti, *sj;
};
void to_csr(int ncell, const std::vector<Seg> &segs, std::vector<int> &ptr, std::vector<int> &src,
            std::vector<int> &targets, int nth) {
  ptr.assign(ncell + 1, 0);
  size_t total = 0;
  for (const Seg &g : segs) total += g.ti->size();
  src.resize(total);
  // nth contiguous runs of segments, each counted and scattered by one thread
  // into its own slots of every target's range (run order = segment order)
  const int T = (int)std::min<size_t>(std::max(nth, 1), std::max<size_t>(segs.size(), 1));
  std::vector<std::vector<int>> cnt(T, std::vector<int>(ncell, 0));
  auto run_of = [&](int r, size_t &lo, size_t &hi) {
    lo = segs.size() * r / T;
    hi = segs.size() * (r + 1) / T;
  };
  auto par = [&](auto fn) {
    std::vector<std::thread> th;
    for (int r = 1; r < T; r++) th.emplace_back(fn, r);
    fn(0);
    for (auto &x : th) x.join();
  };
  par([&](int r) {
    size_t lo, hi;
    run_of(r, lo, hi);
    for (size_t g = lo; g < hi; g++)
      for (int c : *segs[g].ti) cnt[r][c]++;
  });
  int64_t acc = 0;
  for (int c = 0; c < ncell; c++) {
    ptr[c] = (int)acc;
    for (int r = 0; r < T; r++) {
      const int k = cnt[r][c];
      cnt[r][c] = (int)acc;  // becomes run r's first slot for target c
      acc += k;
    }
  }
  ptr[ncell] = (int)acc;
  par([&](int r) {
    size_t lo, hi;
    run_of(r, lo, hi);
    std::vector<int> &fill = cnt[r];
    for (size_t g = lo; g < hi; g++) {
      const std::vector<int> &ti = *segs[g].ti, &sj = *segs[g].sj;
      for (size_t k = 0; k < ti.size(); k++) src[fill[ti[k]]++] = sj[k];
    }
  });
  targets.clear();
  for (int c = 0; c < ncell; c++)
    if (ptr[c + 1] > ptr[c]) targets.push_back(c);
}

}  // namespace

namespace {
thread_local int tl_host_threads = 0;
}

void set_host_threads(int n) { tl_host_threads = n; }

int host_threads() {
  if (tl_host_threads > 0) return tl_host_threads;
  const int hw = (int)std::thread::hardware_concurrency();
  return std::max(1, std::min(16, hw));
}

Lists empty_lists(int ncell) {
  Lists L;
  L.m2l_ptr.assign(ncell + 1, 0);
  L.p2p_ptr.assign(ncell + 1, 0);
  return L;
}

Lists build_lists(const Tree &t, double theta, int nspawn, const double xper[3], int images, const double *cycle,
                  int root_i, int root_j) {
  Lists L = empty_lists(t.ncell);
  if (t.ncell == 0) return L;
  const int nth = t.n >= 50000 ? host_threads() : 1;
  std::deque<Walker> ws;  // segments in emission order (stable addresses)
  const int nimg = images > 0 ? 27 : 1;
  auto shift_of = [&](int img, double xp[3]) {
    xp[0] = xp[1] = xp[2] = 0.0;
    if (images > 0)
      image_shift(img, cycle, xp);
    else if (xper)
      std::memcpy(xp, xper, sizeof(double) * 3);
  };
  if (nth == 1) {
    for (int img = 0; img < nimg; img++) {
      double xp[3];
      shift_of(img, xp);
      ws.push_back(Walker{t, theta, nspawn, {xp[0], xp[1], xp[2]}, {}, {}, {}, {}});
      Walker &w = ws.back();
      w.tag = images > 0 ? img << kImageShift : 0;
      w.m2l_i.reserve(64 * (size_t)t.ncell / nimg);
      w.m2l_j.reserve(64 * (size_t)t.ncell / nimg);
      w.p2p_i.reserve(32 * (size_t)t.leaves.size() / nimg);
      w.p2p_j.reserve(32 * (size_t)t.leaves.size() / nimg);
      w.dual(root_i, root_j);
    }
  } else {
    // For every image, the recursion down to cell pairs of at most `cut`
    // bodies runs here and leaves the pairs below it as tasks in visiting
    // order; one pool of threads then traverses blocks of consecutive tasks
    // of all images, and the blocks' pairs are taken in (image, block)
    // order, which is the serial visiting order.
    std::vector<std::vector<std::pair<int, int>>> tasks(nimg);
    size_t ntask = 0;
    for (int img = 0; img < nimg; img++) {
      double xp[3];
      shift_of(img, xp);
      Walker top{t, theta, nspawn, {xp[0], xp[1], xp[2]}, {}, {}, {}, {}};
      top.cut = std::max(2 * 64, t.n / (8 * nth));
      top.tasks = &tasks[img];
      top.dual(root_i, root_j);
      ntask += tasks[img].size();
    }
    struct Blk {
      int img;
      size_t lo, hi;
    };
    std::vector<Blk> blk;
    const size_t per = std::max<size_t>(1, ntask / ((size_t)nth * 16));
    for (int img = 0; img < nimg; img++) {
      double xp[3];
      shift_of(img, xp);
      for (size_t lo = 0; lo < tasks[img].size(); lo += per) {
        blk.push_back({img, lo, std::min(tasks[img].size(), lo + per)});
        ws.push_back(Walker{t, theta, nspawn, {xp[0], xp[1], xp[2]}, {}, {}, {}, {}});
        ws.back().tag = images > 0 ? img << kImageShift : 0;
      }
    }
    std::atomic<size_t> next{0};
    auto work = [&] {
      for (size_t b; (b = next.fetch_add(1)) < blk.size();) {
        const auto &tk = tasks[blk[b].img];
        for (size_t k = blk[b].lo; k < blk[b].hi; k++) ws[b].dual(tk[k].first, tk[k].second);
      }
    };
    std::vector<std::thread> th;
    for (int i = 1; i < nth; i++) th.emplace_back(work);
    work();
    for (auto &x : th) x.join();
  }
  std::vector<Seg> sm, sp;
  for (const Walker &w : ws) {
    L.nm2l += (int64_t)w.m2l_i.size();
    L.np2p += (int64_t)w.p2p_i.size();
    sm.push_back({&w.m2l_i, &w.m2l_j});
    sp.push_back({&w.p2p_i, &w.p2p_j});
    for (size_t k = 0; k < w.p2p_i.size(); k++)
      L.p2p_body_pairs += (int64_t)t.nbody[w.p2p_i[k]] * t.nbody[cell_of(w.p2p_j[k])];
  }
  to_csr(t.ncell, sm, L.m2l_ptr, L.m2l_src, L.m2l_targets, nth);
  to_csr(t.ncell, sp, L.p2p_ptr, L.p2p_src, L.p2p_targets, nth);
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

// Open-addressing index of 128-bit keys (offset key, level pair) -> dense id;
// one 24-byte slot per probe.
struct KeyIndex2 {
  struct Slot {
    uint64_t a, b;
    int id;
  };
  std::vector<Slot> t;
  size_t mask = 0;
  int count = 0;
  explicit KeyIndex2(size_t cap = 1 << 12) { rehash(cap); }
  void rehash(size_t cap) {
    std::vector<Slot> o;
    o.swap(t);
    t.assign(cap, Slot{0, 0, -1});
    mask = cap - 1;
    for (const Slot &x : o)
      if (x.id >= 0) place(x);
  }
  static inline size_t hash(uint64_t a, uint64_t b) { return KeyIndex::mix(a ^ (b * 0x9e3779b97f4a7c15ull)); }
  void place(const Slot &x) {
    size_t h = hash(x.a, x.b) & mask;
    while (t[h].id >= 0) h = (h + 1) & mask;
    t[h] = x;
  }
  int get(uint64_t a, uint64_t b, bool &fresh) {
    size_t h = hash(a, b) & mask;
    while (t[h].id >= 0) {
      if (t[h].a == a && t[h].b == b) {
        fresh = false;
        return t[h].id;
      }
      h = (h + 1) & mask;
    }
    fresh = true;
    t[h] = Slot{a, b, count++};
    if (2 * (size_t)count > t.size()) rehash(2 * t.size());
    return count - 1;
  }
};

}  // namespace

M2LGroups build_m2l_groups(const Tree &t, const Lists &L, int chunk_pairs, bool by_level, const double *cycle) {
  M2LGroups G;
  const int nc = t.ncell;
  G.rest_ptr.assign(nc + 1, 0);
  if (nc == 0 || L.nm2l == 0) {
    G.grp_ptr.assign(1, 0);
    return G;
  }
  const double R0 = t.R[0];
  const int64_t np = L.nm2l;
  const int64_t lim = 1 << 18;
  double pow2[64];  // 2^e / R0
  for (int e = 0; e < 64; e++) pow2[e] = std::ldexp(1.0, e) / R0;
  // round to nearest (ties to even) for |v| < 2^51 without a libm call
  auto rnd = [](double v) { return (v + 6755399441055744.0) - 6755399441055744.0; };
  // Runs of target cells with about equal pair counts, one per thread.  Each
  // run numbers the keys it meets in first-seen order; merging the runs in
  // order gives every group the index a single pass in CSR order would give.
  const int T = np >= 200000 ? host_threads() : 1;
  std::vector<int> cut(T + 1, nc);
  cut[0] = 0;
  for (int r = 1; r < T; r++)
    cut[r] = (int)(std::lower_bound(L.m2l_ptr.begin(), L.m2l_ptr.end(), (int)(np * r / T)) - L.m2l_ptr.begin());
  for (int r = 1; r <= T; r++) cut[r] = std::max(cut[r], cut[r - 1]);
  struct Fresh {
    uint64_t key, key2;
    int64_t k[3];
    int e, img;
    double rs, rt;
  };
  bigvec<int> lid(np);  // run-local group of each pair, -1: not grouped
  std::vector<std::vector<Fresh>> fresh(T);
  auto par = [&](auto fn) {
    std::vector<std::thread> th;
    for (int r = 1; r < T; r++) th.emplace_back(fn, r);
    fn(0);
    for (auto &x : th) x.join();
  };
  par([&](int r) {
    KeyIndex2 index;
    std::vector<Fresh> &fr = fresh[r];
    for (int c = cut[r]; c < cut[r + 1]; c++) {
      const double xc = t.X[3 * (size_t)c], yc = t.X[3 * (size_t)c + 1], zc = t.X[3 * (size_t)c + 2];
      const int lt = t.level[c];
      for (int p = L.m2l_ptr[c]; p < L.m2l_ptr[c + 1]; p++) {
        const int scell = cell_of(L.m2l_src[p]), img = image_of(L.m2l_src[p]);
        const int ls = t.level[scell];
        int e = std::max(lt, ls);
        int64_t k[3];
        bool ok = e < 64;
        if (ok) {
          const double scale = pow2[e];
          const double v[3] = {(xc - t.X[3 * (size_t)scell]) * scale, (yc - t.X[3 * (size_t)scell + 1]) * scale,
                               (zc - t.X[3 * (size_t)scell + 2]) * scale};
          for (int d = 0; d < 3; d++) {
            const double rv = rnd(v[d]);
            if (!(std::fabs(v[d] - rv) < 1e-6) || std::fabs(rv) >= (double)(lim << 8)) ok = false;
            k[d] = (int64_t)rv;
          }
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
          lid[p] = -1;
          continue;
        }
        const uint64_t key = ((uint64_t)e << 57) | ((uint64_t)(k[0] + lim) << 38) |
                             ((uint64_t)(k[1] + lim) << 19) | (uint64_t)(k[2] + lim);
        // the image, and (by_level) the source and target levels, are part of the group
        const uint64_t key2 = ((uint64_t)img << 20) | (by_level ? (((uint64_t)lt << 8) | (uint64_t)ls) + 1 : 0);
        bool isnew;
        lid[p] = index.get(key, key2, isnew);
        if (isnew) fr.push_back({key, key2, {k[0], k[1], k[2]}, e, img, t.R[scell], t.R[c]});
      }
    }
  });
  // merge in run order: global group numbers, offsets and (by_level) radii
  std::vector<std::vector<int>> gmap(T);
  std::vector<double> sc;  // by_level: (R_s, R_t) of each group's first pair
  std::vector<int64_t> kvec;
  std::vector<int> gexp, gimg;
  std::vector<uint64_t> gkey2;
  {
    KeyIndex2 index;
    for (int r = 0; r < T; r++) {
      gmap[r].resize(fresh[r].size());
      for (size_t u = 0; u < fresh[r].size(); u++) {
        const Fresh &f = fresh[r][u];
        bool isnew;
        const int g = index.get(f.key, f.key2, isnew);
        gmap[r][u] = g;
        if (isnew) {
          kvec.insert(kvec.end(), {f.k[0], f.k[1], f.k[2]});
          gexp.push_back(f.e);
          gimg.push_back(f.img);
          gkey2.push_back(f.key2);
          sc.insert(sc.end(), {f.rs, f.rt});
        }
      }
    }
  }
  const int ng = (int)gexp.size();
  G.dx.resize(3 * (size_t)ng);
  for (int g = 0; g < ng; g++) {
    const double u = R0 / std::ldexp(1.0, gexp[g]);
    double sh[3] = {0, 0, 0};
    if (cycle) image_shift(gimg[g], cycle, sh);
    for (int d = 0; d < 3; d++) G.dx[3 * (size_t)g + d] = (double)kvec[3 * (size_t)g + d] * u - sh[d];
  }
  G.ngroup = ng;
  // reflection classes: a group is reflected in every axis where its
  // translation vector (periodic: including the image shift) is negative;
  // reflecting an axis negates the lattice offset and the image's shift
  // along it, so the reflected vector is the exact negation and groups whose
  // vectors differ only in signs share one representative
  G.canon.resize(ng);
  G.refl.assign(ng, 0);
  {
    KeyIndex2 index;
    for (int g = 0; g < ng; g++) {
      const int64_t *k = &kvec[3 * (size_t)g];
      const double *t = &G.dx[3 * (size_t)g];
      uint8_t r = 0;
      int64_t a[3];
      int ic[3] = {gimg[g] / 9 - 1, gimg[g] / 3 % 3 - 1, gimg[g] % 3 - 1};
      for (int d = 0; d < 3; d++) {
        a[d] = k[d];
        if (cycle ? t[d] < 0.0 : k[d] < 0) {
          a[d] = -k[d];
          ic[d] = -ic[d];
          r |= (uint8_t)(1 << d);
        }
      }
      const int imgc = cycle ? (ic[0] + 1) * 9 + (ic[1] + 1) * 3 + (ic[2] + 1) : 0;  // no images: 0
      const uint64_t key = ((uint64_t)gexp[g] << 57) | ((uint64_t)(a[0] + lim) << 38) |
                           ((uint64_t)(a[1] + lim) << 19) | (uint64_t)(a[2] + lim);
      const uint64_t key2 = ((uint64_t)imgc << 20) | (gkey2[g] & ((1ull << 20) - 1));
      bool isnew;
      const int cg = index.get(key, key2, isnew);
      G.canon[g] = cg;
      G.refl[g] = r;
      if (isnew) {
        G.cdx.insert(G.cdx.end(), {(r & 1) ? -t[0] : t[0], (r & 2) ? -t[1] : t[1], (r & 4) ? -t[2] : t[2]});
        G.cscale.insert(G.cscale.end(), {sc[2 * (size_t)g], sc[2 * (size_t)g + 1]});
      }
    }
    G.ncanon = index.count;
  }
  if (!by_level) G.cscale.clear();
  // global group per pair; by_level: a pair whose radii differ from its
  // group's first pair (the chunks' scale) or from its reflection class's
  // first pair (the tiles' operator scale, one per class) runs per target
  // (radius not uniform on its level); counts per (run, group) and per
  // target of the ungrouped rest
  std::vector<std::vector<int>> cnt(T);
  par([&](int r) {
    std::vector<int> &cn = cnt[r];
    cn.assign(ng, 0);
    const std::vector<int> &m = gmap[r];
    for (int c = cut[r]; c < cut[r + 1]; c++)
      for (int p = L.m2l_ptr[c]; p < L.m2l_ptr[c + 1]; p++) {
        int g = lid[p] >= 0 ? m[lid[p]] : -1;
        if (g >= 0 && by_level) {
          const double rs = t.R[cell_of(L.m2l_src[p])], rt = t.R[c];
          const size_t cg = (size_t)G.canon[g];
          if (rs != sc[2 * (size_t)g] || rt != sc[2 * (size_t)g + 1] || rs != G.cscale[2 * cg] ||
              rt != G.cscale[2 * cg + 1])
            g = -1;
        }
        lid[p] = g;
        if (g >= 0)
          cn[g]++;
        else
          G.rest_ptr[c + 1]++;
      }
  });
  if (by_level) G.scale = std::move(sc);
  // counting sort by group, CSR order kept inside a group (runs in order)
  G.grp_ptr.assign(ng + 1, 0);
  for (int g = 0; g < ng; g++) {
    int64_t a = G.grp_ptr[g];
    for (int r = 0; r < T; r++) {
      const int k = cnt[r][g];
      cnt[r][g] = (int)a;  // run r's first slot in group g
      a += k;
    }
    G.grp_ptr[g + 1] = (int)a;
  }
  for (int c = 0; c < nc; c++) G.rest_ptr[c + 1] += G.rest_ptr[c];
  const int64_t ngp = G.grp_ptr[ng];
  G.tgt.resize(ngp);
  G.src.resize(ngp);
  G.rest_src.resize(G.rest_ptr[nc]);
  par([&](int r) {
    std::vector<int> &fill = cnt[r];
    for (int c = cut[r]; c < cut[r + 1]; c++) {
      int rq = G.rest_ptr[c];
      for (int p = L.m2l_ptr[c]; p < L.m2l_ptr[c + 1]; p++) {
        const int g = lid[p];
        if (g >= 0) {
          const int q = fill[g]++;
          G.tgt[q] = c;
          G.src[q] = L.m2l_src[p];
        } else {
          G.rest_src[rq++] = L.m2l_src[p];
        }
      }
    }
  });
  G.nrest = (int64_t)G.rest_src.size();
  for (int c = 0; c < nc; c++)
    if (G.rest_ptr[c + 1] > G.rest_ptr[c]) G.rest_targets.push_back(c);
  deal_m2l_chunks(G, chunk_pairs);
  return G;
}

void list_counts(const Tree &t, const Lists &L, uint64_t *icell, int *nump2p, int *numm2l) {
  const int nc = t.ncell;
  for (int c = 0; c < nc; c++) {
    if (nump2p) nump2p[c] = L.p2p_ptr[c + 1] - L.p2p_ptr[c];
    if (numm2l) numm2l[c] = L.m2l_ptr[c + 1] - L.m2l_ptr[c];
  }
  if (!icell || nc == 0) return;
  double xmin[3];
  for (int d = 0; d < 3; d++) xmin[d] = t.X[d] - t.R[0];
  for (int c = 0; c < nc; c++) {
    const int level = t.level[c];
    const double diameter = 2 * t.R[c];
    int iX[3];
    for (int d = 0; d < 3; d++) iX[d] = (int)((t.X[3 * (size_t)c + d] - xmin[d]) / diameter);
    uint64_t index = (((uint64_t)1 << (3 * level)) - 1) / 7;
    for (int l = 0; l < level; l++) {
      for (int d = 0; d < 3; d++) index += (uint64_t)(iX[d] & 1) << (3 * l + d);
      for (int d = 0; d < 3; d++) iX[d] >>= 1;
    }
    icell[c] = index;
  }
}

void deal_m2l_chunks(M2LGroups &G, int chunk_pairs) {
  const int ng = G.ngroup;
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
}

void split_m2l_tiles(M2LTiles &T, int seg) {
  std::vector<int> it;
  for (int b = 0; b < T.ntile; b++) {
    const int a = T.ptr[b], e = T.ptr[b + 1];
    const int split = e - a > seg ? 1 : 0;
    for (int k = a; k < e; k += seg) it.insert(it.end(), {b, k, std::min(e, k + seg), split});
  }
  // launch order: block i runs on XCD i % 8.  Runs of kRun consecutive items
  // (neighbouring tiles, sharing source multipoles) go to one XCD, and the
  // runs are dealt round-robin so every XCD gets a similar mix of work;
  // empty items (s0 == s1) pad the shorter queues
  const int kXcd = 8, kRun = 16;
  const size_t n = it.size() / 4;
  std::vector<std::vector<size_t>> q(kXcd);
  for (size_t i = 0; i < n; i++) q[(i / kRun) % kXcd].push_back(i);
  size_t per = 0;
  for (auto &v : q) per = std::max(per, v.size());
  T.item.assign(4 * per * kXcd, 0);
  for (int x = 0; x < kXcd; x++)
    for (size_t j = 0; j < q[x].size(); j++)
      for (int k = 0; k < 4; k++) T.item[4 * (j * kXcd + x) + k] = it[4 * q[x][j] + k];
}

M2LTiles build_m2l_tiles(const Tree &t, const M2LGroups &G, int TT, int SW, int seg) {
  M2LTiles T;
  T.TT = TT;
  const int nc = t.ncell;
  const int64_t np = (int64_t)G.tgt.size();
  T.ptr.assign(1, 0);
  if (np == 0) return T;
  // targets of grouped pairs, ordered by (level, octant in parent, cell)
  std::vector<char> has(nc, 0);
  for (int c : G.tgt) has[c] = 1;
  std::vector<int> tg;
  for (int c = 0; c < nc; c++)
    if (has[c]) tg.push_back(c);
  auto octant = [&](int c) {
    if (c == 0) return 0;
    const int p = t.iparent[c];
    int o = 0;
    for (int d = 0; d < 3; d++)
      if (t.X[3 * (size_t)c + d] > t.X[3 * (size_t)p + d]) o |= 1 << d;
    return o;
  };
  std::vector<int64_t> key(tg.size());
  for (size_t i = 0; i < tg.size(); i++)
    key[i] = ((int64_t)t.level[tg[i]] << 40) | ((int64_t)octant(tg[i]) << 32) | tg[i];
  std::vector<int> ord(tg.size());
  for (size_t i = 0; i < ord.size(); i++) ord[i] = (int)i;
  std::sort(ord.begin(), ord.end(), [&](int a, int b) { return key[a] < key[b]; });
  const int nt = (int)((tg.size() + TT - 1) / TT);
  std::vector<int> tile_of(nc, -1), local(nc, -1);
  for (size_t i = 0; i < ord.size(); i++) {
    const int c = tg[ord[i]];
    tile_of[c] = (int)(i / TT);
    local[c] = (int)(i % TT);
  }
  // pairs bucketed by tile; taking the groups in (representative,
  // reflection) order keeps each tile's pairs sorted that way, so a tile
  // loads each representative's operator once
  std::vector<int64_t> pcount(nt + 1, 0);
  for (int64_t q = 0; q < np; q++) pcount[tile_of[G.tgt[q]] + 1]++;
  for (int i = 0; i < nt; i++) pcount[i + 1] += pcount[i];
  bigvec<int> pgrp(np), psrc(np);
  bigvec<uint16_t> plt(np);
  std::vector<int> gorder(G.ngroup);
  for (int g = 0; g < G.ngroup; g++) gorder[g] = g;
  std::stable_sort(gorder.begin(), gorder.end(), [&](int a, int b) {
    return G.canon[a] != G.canon[b] ? G.canon[a] < G.canon[b] : G.refl[a] < G.refl[b];
  });
  // runs of consecutive groups (in that order) with about equal pair counts,
  // one per thread; run r fills its slots of every tile after runs < r, so the
  // result is the single-threaded order
  const int R = np >= 200000 ? host_threads() : 1;
  std::vector<int> gcut(R + 1, G.ngroup);
  {
    gcut[0] = 0;
    int64_t acc = 0;
    int r = 1;
    for (int k = 0; k < G.ngroup && r < R; k++) {
      const int g = gorder[k];
      acc += G.grp_ptr[g + 1] - G.grp_ptr[g];
      while (r < R && acc >= np * r / R) gcut[r++] = k + 1;
    }
  }
  auto par = [&](int nthr, auto fn) {
    std::vector<std::thread> th;
    for (int r = 1; r < nthr; r++) th.emplace_back(fn, r);
    fn(0);
    for (auto &x : th) x.join();
  };
  std::vector<std::vector<int64_t>> off(R, std::vector<int64_t>(nt, 0));
  par(R, [&](int r) {
    for (int k = gcut[r]; k < gcut[r + 1]; k++) {
      const int g = gorder[k];
      for (int q = G.grp_ptr[g]; q < G.grp_ptr[g + 1]; q++) off[r][tile_of[G.tgt[q]]]++;
    }
  });
  for (int b = 0; b < nt; b++) {
    int64_t a = pcount[b];
    for (int r = 0; r < R; r++) {
      const int64_t k = off[r][b];
      off[r][b] = a;
      a += k;
    }
  }
  par(R, [&](int r) {
    std::vector<int64_t> &fill = off[r];
    for (int k = gcut[r]; k < gcut[r + 1]; k++) {
      const int g = gorder[k];
      for (int q = G.grp_ptr[g]; q < G.grp_ptr[g + 1]; q++) {
        const int c = G.tgt[q];
        const int64_t kk = fill[tile_of[c]]++;
        pgrp[kk] = g;
        psrc[kk] = G.src[q];
        plt[kk] = (uint16_t)(local[c] | (G.refl[g] << 8));
      }
    }
  });
  // tiles in key order (neighbours adjacent); the work items built from them
  // are dealt to the XCDs by split_m2l_tiles.  Slices: the pairs of one
  // representative inside a tile (every reflection of it) are dealt into
  // S = max(ceil(m / SW), most pairs of one target) slices of distinct
  // targets -- the pairs sorted by target, pair i to slice i mod S -- so a
  // slice mixes reflections but never repeats a target.
  T.ntile = nt;
  T.tgt.assign((size_t)T.ntile * TT, -1);
  for (size_t i = 0; i < ord.size(); i++) T.tgt[i] = tg[ord[i]];
  std::vector<std::vector<int>> tsl(nt);  // per tile: (representative, first, count) per slice
  const int RT = std::min(R, std::max(nt, 1));
  par(RT, [&](int r) {
    std::vector<int> cnt(TT + 1), seq;
    std::vector<int> bsrc;
    std::vector<uint16_t> blt;
    for (int b = (int)((int64_t)nt * r / RT); b < (int)((int64_t)nt * (r + 1) / RT); b++) {
      std::vector<int> &out = tsl[b];
      for (int64_t k = pcount[b]; k < pcount[b + 1];) {
        const int rep = G.canon[pgrp[k]];
        int64_t e = k;
        while (e < pcount[b + 1] && G.canon[pgrp[e]] == rep) e++;
        const int m = (int)(e - k);
        std::fill(cnt.begin(), cnt.end(), 0);
        int most = 0;
        for (int64_t q = k; q < e; q++) most = std::max(most, ++cnt[(plt[q] & 255) + 1]);
        const int S = std::max((m + SW - 1) / SW, most);
        for (int u = 0; u < TT; u++) cnt[u + 1] += cnt[u];
        seq.resize(m);
        for (int64_t q = k; q < e; q++) seq[cnt[plt[q] & 255]++] = (int)(q - k);
        bsrc.assign(psrc.begin() + k, psrc.begin() + e);
        blt.assign(plt.begin() + k, plt.begin() + e);
        int64_t o = k;
        for (int j = 0; j < S; j++) {
          const int64_t first = o;
          for (int i = j; i < m; i += S) {
            psrc[o] = bsrc[seq[i]];
            plt[o] = blt[seq[i]];
            o++;
          }
          out.insert(out.end(), {rep, (int)first, (int)(o - first)});
        }
        k = e;
      }
    }
  });
  T.ptr.assign(nt + 1, 0);
  for (int b = 0; b < nt; b++) T.ptr[b + 1] = T.ptr[b] + (int)(tsl[b].size() / 3);
  T.slice.resize(4 * (size_t)T.ptr[nt]);
  par(RT, [&](int r) {
    for (int b = (int)((int64_t)nt * r / RT); b < (int)((int64_t)nt * (r + 1) / RT); b++)
      for (size_t i = 0; i < tsl[b].size() / 3; i++) {
        int *o = T.slice.data() + 4 * ((size_t)T.ptr[b] + i);
        o[0] = tsl[b][3 * i];
        o[1] = tsl[b][3 * i + 1];
        o[2] = tsl[b][3 * i + 2];
        o[3] = 0;
      }
  });
  T.src = std::move(psrc);
  T.lt = std::move(plt);
  split_m2l_tiles(T, seg);
  return T;
}

void merge_plans(const std::vector<PlanPart> &parts, Tree &T, Lists &L, M2LGroups *G, int chunk_pairs,
                 M2LTiles *tiles) {
  T = Tree();
  L = Lists();
  int64_t nb = 0, nc = 0, npm = 0, npp = 0;
  for (const PlanPart &p : parts) {
    nb += p.t->n;
    nc += p.t->ncell;
    npm += p.l->nm2l;
    npp += p.l->np2p;
  }
  T.n = (int)nb;
  T.ncell = (int)nc;
  for (const PlanPart &p : parts) T.adopted = T.adopted || p.t->adopted;
  L.nm2l = npm;
  L.np2p = npp;
  L.m2l_ptr.assign(nc + 1, 0);
  L.p2p_ptr.assign(nc + 1, 0);
  int bo = 0, co = 0, pmo = 0, ppo = 0;
  for (const PlanPart &p : parts) {
    const Tree &t = *p.t;
    const Lists &l = *p.l;
    T.depth = std::max(T.depth, t.depth);
    for (int v : t.perm) T.perm.push_back(v + bo);
    for (int c = 0; c < t.ncell; c++) {
      T.iparent.push_back(t.iparent[c] + co);
      T.ichild.push_back(t.ichild[c] + co);
      T.nchild.push_back(t.nchild[c]);
      T.ibody.push_back(t.ibody[c] + bo);
      T.nbody.push_back(t.nbody[c]);
      T.level.push_back(t.level[c]);
      T.R.push_back(t.R[c]);
    }
    T.X.insert(T.X.end(), t.X.begin(), t.X.end());
    for (int v : t.leaves) T.leaves.push_back(v + co);
    for (int c = 0; c < t.ncell; c++) {
      L.m2l_ptr[co + c + 1] = l.m2l_ptr[c + 1] + pmo;
      L.p2p_ptr[co + c + 1] = l.p2p_ptr[c + 1] + ppo;
    }
    for (int v : l.m2l_src) L.m2l_src.push_back(v + co);
    for (int v : l.p2p_src) L.p2p_src.push_back(v + co);
    for (int v : l.m2l_targets) L.m2l_targets.push_back(v + co);
    for (int v : l.p2p_targets) L.p2p_targets.push_back(v + co);
    L.p2p_body_pairs += l.p2p_body_pairs;
    bo += t.n;
    co += t.ncell;
    pmo += (int)l.nm2l;
    ppo += (int)l.np2p;
  }
  // level index over the whole forest, cell order kept inside a level
  T.level_ptr.assign(T.depth + 2, 0);
  for (int c = 0; c < T.ncell; c++) T.level_ptr[T.level[c] + 1]++;
  for (int l = 0; l <= T.depth; l++) T.level_ptr[l + 1] += T.level_ptr[l];
  T.level_cells.resize(T.ncell);
  std::vector<int> fill(T.level_ptr.begin(), T.level_ptr.end() - 1);
  for (int c = 0; c < T.ncell; c++) T.level_cells[fill[T.level[c]]++] = c;
  if (!G) return;
  *G = M2LGroups();
  G->grp_ptr.assign(1, 0);
  G->rest_ptr.assign(nc + 1, 0);
  co = 0;
  int po = 0;
  for (const PlanPart &p : parts) {
    const M2LGroups &g = *p.g;
    for (int k = 0; k < g.ngroup; k++) G->grp_ptr.push_back(g.grp_ptr[k + 1] + po);
    for (int v : g.tgt) G->tgt.push_back(v + co);
    for (int v : g.src) G->src.push_back(v + co);
    G->dx.insert(G->dx.end(), g.dx.begin(), g.dx.end());
    G->scale.insert(G->scale.end(), g.scale.begin(), g.scale.end());
    for (int v : g.canon) G->canon.push_back(v + G->ncanon);
    G->refl.insert(G->refl.end(), g.refl.begin(), g.refl.end());
    G->cdx.insert(G->cdx.end(), g.cdx.begin(), g.cdx.end());
    G->cscale.insert(G->cscale.end(), g.cscale.begin(), g.cscale.end());
    G->ncanon += g.ncanon;
    const int ro = (int)G->rest_src.size();
    for (int c = 0; c < p.t->ncell; c++) G->rest_ptr[co + c + 1] = g.rest_ptr[c + 1] + ro;
    for (int v : g.rest_src) G->rest_src.push_back(v + co);
    for (int v : g.rest_targets) G->rest_targets.push_back(v + co);
    G->ngroup += g.ngroup;
    co += p.t->ncell;
    po += (int)g.tgt.size();
  }
  G->nrest = (int64_t)G->rest_src.size();
  deal_m2l_chunks(*G, chunk_pairs);
  if (!tiles) return;
  // tiles concatenated part by part (each part's tile count is a multiple of
  // the XCD count, so every part keeps its XCD dealing)
  M2LTiles &X = *tiles;
  X = M2LTiles();
  X.TT = parts[0].tl->TT;
  X.ptr.assign(1, 0);
  co = 0;
  int go = 0;
  for (const PlanPart &p : parts) {
    const M2LTiles &a = *p.tl;
    const int so = (int)(X.slice.size() / 4), po2 = (int)X.src.size();
    for (int v : a.tgt) X.tgt.push_back(v >= 0 ? v + co : -1);
    for (int b = 0; b < a.ntile; b++) X.ptr.push_back(a.ptr[b + 1] + so);
    for (size_t q = 0; q < a.slice.size(); q += 4)
      X.slice.insert(X.slice.end(), {a.slice[q] + go, a.slice[q + 1] + po2, a.slice[q + 2], a.slice[q + 3]});
    for (int v : a.src) X.src.push_back(v + co);
    X.lt.insert(X.lt.end(), a.lt.begin(), a.lt.end());
    X.ntile += a.ntile;
    co += p.t->ncell;
    go += p.g->ncanon;
  }
  int seg = 1;
  for (const PlanPart &p : parts)
    for (size_t q = 0; q < p.tl->item.size(); q += 4) seg = std::max(seg, p.tl->item[q + 2] - p.tl->item[q + 1]);
  split_m2l_tiles(X, seg);
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

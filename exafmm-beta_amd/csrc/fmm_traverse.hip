// fmm_traverse.hip -- the octree and the interaction lists built on the device.
//
// The dual tree traversal of include/traversal.h:117-177 (dualTreeTraversal)
// and :90-114 (splitCell) run breadth first: every round takes the frontier
// of cell pairs, applies the multipole acceptance test and the split rules to
// each pair independently, appends accepted pairs to the M2L list, leaf pairs
// to the P2P list (or to M2L when the source leaf is empty) and the split
// pairs to the next frontier.  The rules are local to a pair, so the pairs
// found are exactly those of the recursive walk; only their order differs,
// and the lists are sorted by (target, source) at the end.  The acceptance
// test rounds exactly like the host walker's: no contraction into FMAs.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "fmm_device.hpp"

namespace fmmdev {

namespace {

// this lane's exclusive prefix of v over the wavefront, and the total
__device__ __forceinline__ int wave_excl(int v, int &total) {
  const int l = threadIdx.x & 63;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (l >= o) x += y;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}

struct TravCells {
  const double *X;  // 3 per cell
  const double *R;
  const int *nchild, *ichild, *nbody;
};

enum { kM2L = 0, kP2P = 1, kNext = 2 };

// classify pair (Ci, Cj): what it emits and how many frontier pairs it splits into
__device__ __forceinline__ void classify(const TravCells &t, int Ci, int Cj, double theta, double xp0, double xp1,
                                         double xp2, int nspawn, int &kind, int &na, int &nb, int &sa, int &sb) {
#pragma clang fp contract(off)
  const double dx = t.X[3 * (size_t)Ci] - t.X[3 * (size_t)Cj] - xp0;
  const double dy = t.X[3 * (size_t)Ci + 1] - t.X[3 * (size_t)Cj + 1] - xp1;
  const double dz = t.X[3 * (size_t)Ci + 2] - t.X[3 * (size_t)Cj + 2] - xp2;
  const double RT2 = (dx * dx + dy * dy + dz * dz) * theta * theta;
  const double Ri = t.R[Ci], Rj = t.R[Cj];
  const double Rs = Ri + Rj;
  const bool accept = RT2 > Rs * Rs * (1 - 1e-3);
  const int nci = t.nchild[Ci], ncj = t.nchild[Cj];
  // na x nb frontier pairs: a over Ci's children (or Ci itself), b over Cj's
  na = nb = 0;
  sa = Ci;
  sb = Cj;
  if (accept) {
    kind = kM2L;
  } else if (nci == 0 && ncj == 0) {
    kind = t.nbody[Cj] == 0 ? kM2L : kP2P;
  } else {
    kind = kNext;
    const int ci0 = t.ichild[Ci], cj0 = t.ichild[Cj];
    if (ncj == 0) {
      na = nci; sa = ci0; nb = 1;
    } else if (nci == 0) {
      na = 1; nb = ncj; sb = cj0;
    } else if (t.nbody[Ci] + t.nbody[Cj] >= nspawn || Ci == Cj) {
      na = nci; sa = ci0; nb = ncj; sb = cj0;
    } else if (Ri >= Rj) {
      na = nci; sa = ci0; nb = 1;
    } else {
      na = 1; nb = ncj; sb = cj0;
    }
  }
}

// count (write = false) or emit (write = true) one round; cnt[0..2]: M2L,
// P2P and next-frontier totals (running offsets while writing)
template <bool WRITE>
__global__ __launch_bounds__(256) void trav_round_kernel(const int2 *__restrict__ F, int nf, TravCells t, double theta,
                                                         double xp0, double xp1, double xp2, int nspawn,
                                                         unsigned long long *__restrict__ cnt, int2 *__restrict__ m2l,
                                                         int2 *__restrict__ p2p, int2 *__restrict__ next) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int kind = -1, na = 0, nb = 0, sa = 0, sb = 0;
  if (i < nf) {
    const int2 pr = F[i];
    classify(t, pr.x, pr.y, theta, xp0, xp1, xp2, nspawn, kind, na, nb, sa, sb);
  }
  const int cm = kind == kM2L, cp = kind == kP2P, cn = na * nb;
  // offsets within the wavefront, then within the block (one set of global
  // atomics per block: per-wavefront atomics on three counters serialised)
  int tm, tp, tn;
  const int om = wave_excl(cm, tm), op = wave_excl(cp, tp), on = wave_excl(cn, tn);
  __shared__ int wtot[3][4];
  __shared__ unsigned long long bbase[3];
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    wtot[0][wv] = tm;
    wtot[1][wv] = tp;
    wtot[2][wv] = tn;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    int tot = 0;
    for (int k = 0; k < 4; k++) tot += wtot[threadIdx.x][k];
    bbase[threadIdx.x] = tot ? atomicAdd(&cnt[threadIdx.x], (unsigned long long)tot) : 0ull;
  }
  __syncthreads();
  if (!WRITE) return;
  int pm = 0, pp = 0, pn = 0;  // the earlier wavefronts' counts
  for (int k = 0; k < wv; k++) {
    pm += wtot[0][k];
    pp += wtot[1][k];
    pn += wtot[2][k];
  }
  const unsigned long long bm = bbase[0] + pm, bp = bbase[1] + pp, bn = bbase[2] + pn;
  if (i >= nf) return;
  const int2 pr = F[i];
  if (cm) m2l[bm + om] = pr;
  if (cp) p2p[bp + op] = pr;
  if (cn) {
    int2 *o = next + bn + on;
    for (int a = 0; a < na; a++)
      for (int b = 0; b < nb; b++) o[a * nb + b] = make_int2(sa + a, sb + b);
  }
}

// sort keys (target, source) and the per-target counts; a wavefront's lanes
// holding one target in a row (the traversal emits a target's pairs
// together) add their count with one atomic
__global__ void pair_keys_kernel(const int2 *__restrict__ pr, size_t n, unsigned long long *__restrict__ key,
                                 int *__restrict__ count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int t = -1;
  if (i < n) {
    const int2 p = pr[i];
    key[i] = ((unsigned long long)(unsigned)p.x << 32) | (unsigned)p.y;
    t = p.x;
  }
  const int prev = __shfl_up(t, 1, 64);
  const bool head = lane == 0 || prev != t;
  const unsigned long long heads = __ballot(head);
  if (head && t >= 0) {
    const unsigned long long above = lane == 63 ? 0ull : heads >> (lane + 1);
    const int len = above ? __ffsll((long long)above) : 64 - lane;
    atomicAdd(&count[t], len);
  }
}

__global__ void key_src_kernel(const unsigned long long *__restrict__ key, size_t n, int *__restrict__ src) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) src[i] = (int)(unsigned)(key[i] & 0xffffffffull);
}

// grid-stride over the pairs, one atomic per block (256 threads)
__global__ void body_pairs_kernel(const int2 *__restrict__ pr, size_t n, const int *__restrict__ nbody,
                                  unsigned long long *__restrict__ out) {
  __shared__ unsigned long long part[4];
  unsigned long long v = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    v += (unsigned long long)nbody[pr[i].x] * (unsigned long long)nbody[pr[i].y];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned long long t = part[0] + part[1] + part[2] + part[3];
    if (t) atomicAdd(out, t);
  }
}

// device scratch from the plan builders' per-stream cache (plan_scratch_get)
struct Buf {
  hipStream_t s;
  void *p = nullptr;
  size_t bytes = 0;  // capacity
  explicit Buf(hipStream_t st) : s(st) {}
  Buf(const Buf &) = delete;
  ~Buf() {
    if (p) plan_scratch_put(p, bytes, s);
  }
  hipError_t need(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) plan_scratch_put(p, bytes, s);
    p = nullptr;
    bytes = 0;
    return plan_scratch_get(&p, &bytes, b, s);
  }
  void swap(Buf &o) {
    std::swap(p, o.p);
    std::swap(bytes, o.bytes);
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

#define TCHK(expr)                          \
  do {                                      \
    hipError_t e_ = (expr);                 \
    if (e_ != hipSuccess) return e_;        \
  } while (0)

// sorted CSR of a pair list on the device: ptr (ncell + 1) and the sources,
// copied to the host, or, with keep, left on the device in keep (src empty)
hipError_t pairs_to_csr(const int2 *pr, size_t n, int ncell, std::vector<int> &ptr, std::vector<int> &src,
                        hipStream_t s, DevArray *keep) {
  ptr.assign((size_t)ncell + 1, 0);
  src.clear();
  if (!keep) src.resize(n);
  Buf keys(s), keys2(s), cnt(s), tmp(s), out(s);
  TCHK(cnt.need(sizeof(int) * ((size_t)ncell + 1)));
  TCHK(hipMemsetAsync(cnt.p, 0, sizeof(int) * ((size_t)ncell + 1), s));
  if (n > 0) {
    TCHK(keys.need(8 * n));
    TCHK(keys2.need(8 * n));
    pair_keys_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(pr, n, keys.as<unsigned long long>(), cnt.as<int>());
    TCHK(hipGetLastError());
    int hi_bit = 64;  // sort on the bits the cell indices use
    {
      int b = 0;
      while ((1ll << b) < (long long)ncell) b++;
      hi_bit = 32 + b;
    }
    hipcub::DoubleBuffer<unsigned long long> db(keys.as<unsigned long long>(), keys2.as<unsigned long long>());
    size_t tb = 0;
    TCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, db, (int)n, 0, hi_bit, s));
    TCHK(tmp.need(tb));
    TCHK(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, db, (int)n, 0, hi_bit, s));
    int *dst;
    if (keep) {
      int dev = 0;
      TCHK(hipGetDevice(&dev));
      TCHK(keep->alloc(sizeof(int) * n, dev));
      dst = keep->as<int>();
    } else {
      TCHK(out.need(sizeof(int) * n));
      dst = out.as<int>();
    }
    key_src_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(db.Current(), n, dst);
    TCHK(hipGetLastError());
    if (!keep) TCHK(hipMemcpyAsync(src.data(), out.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
  } else if (keep) {
    keep->reset();
  }
  std::vector<int> c((size_t)ncell + 1);
  TCHK(hipMemcpyAsync(c.data(), cnt.p, sizeof(int) * ((size_t)ncell + 1), hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  int64_t a = 0;
  for (int i = 0; i < ncell; i++) {
    ptr[i] = (int)a;
    a += c[i];
  }
  ptr[ncell] = (int)a;
  return hipSuccess;
}


// ---- the octree (include/build_tree.h:77-209, as fmm_host.cpp builds it) ----
//
// Each body's octant path to level 30 is computed with the host builder's
// centre arithmetic (octant o = (x > X) + 2 (y > Y) + 4 (z > Z), child centre
// X + r (+-1), r = R0 / 2^(level + 1)), and the bodies are sorted stably by
// it.  A node of level d is then a run of bodies sharing d octants; a body's
// leaf is its first run of at most ncrit bodies (or its level-30 run).  The
// runs come from lcp[j], the octants shared by sorted bodies j - 1 and j,
// with a sparse table of its minima: the run of level d around j ends at the
// first k > j with lcp[k] < d and starts at the last k <= j with lcp[k] < d.
// Cells are enumerated in pre-order (position, level); the cell numbering of
// nodes2cells (children contiguous, each child's descendants after all
// children, in child order) follows top-down from the subtree sizes.  Bodies
// keep their input order within a leaf, as the host's stable partitions do.

__global__ void tree_keys_kernel(int n, const double *__restrict__ x, double X0, double X1, double X2, double R0,
                                 unsigned long long *__restrict__ hi, unsigned long long *__restrict__ lo,
                                 int *__restrict__ idx) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double b0 = x[3 * (size_t)i], b1 = x[3 * (size_t)i + 1], b2 = x[3 * (size_t)i + 2];
  double c0 = X0, c1 = X1, c2 = X2;
  unsigned long long h = 0, w = 0;
  for (int lev = 0; lev < 30; lev++) {
    const int o = (b0 > c0) + ((b1 > c1) << 1) + ((b2 > c2) << 2);
    if (lev < 21)
      h |= (unsigned long long)o << (3 * (20 - lev));
    else
      w |= (unsigned long long)o << (3 * (29 - lev));
    const double r = R0 / (1 << (lev + 1));
    c0 += r * ((o & 1) * 2 - 1);
    c1 += r * (((o >> 1) & 1) * 2 - 1);
    c2 += r * (((o >> 2) & 1) * 2 - 1);
  }
  hi[i] = h;
  lo[i] = w;
  idx[i] = i;
}

__global__ void gather_u64_kernel(int n, const unsigned long long *__restrict__ src, const int *__restrict__ idx,
                                  unsigned long long *__restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// lcp[j] (j = 0..n): octants shared by sorted bodies j - 1 and j; -1 at both ends
__global__ void tree_lcp_kernel(int n, const unsigned long long *__restrict__ hi,
                                const unsigned long long *__restrict__ lo, signed char *__restrict__ lcp) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  int v = -1;
  if (j > 0 && j < n) {
    const unsigned long long a = hi[j] ^ hi[j - 1];
    if (a) {
      v = (__clzll(a) - 1) / 3;
    } else {
      const unsigned long long b = lo[j] ^ lo[j - 1];
      v = b ? 21 + (__clzll(b) - 37) / 3 : 30;
    }
  }
  lcp[j] = (signed char)v;
}

// table level lv: minimum of lcp over [p, p + 2^lv) (positions past n count as 127)
__global__ void tree_table_kernel(int m, const signed char *__restrict__ prev, int half,
                                  signed char *__restrict__ cur) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= m) return;
  const signed char a = prev[p];
  const signed char b = p + half < m ? prev[p + half] : (signed char)127;
  cur[p] = a < b ? a : b;
}

struct Runs {
  const signed char *tbl;  // (LG + 1) levels of m = n + 1 entries
  int m, LG;
  // first k > j with lcp[k] < d (at most n)
  __device__ int next(int j, int d) const {
    int pos = j + 1;
    if (tbl[pos] < d) return pos;
    for (int lv = LG; lv >= 0; lv--)
      if (pos + (1 << lv) <= m && tbl[(size_t)lv * m + pos] >= d) pos += 1 << lv;
    return pos;
  }
  // last k <= j with lcp[k] < d (at least 0)
  __device__ int prev(int j, int d) const {
    int pos = j;
    if (tbl[pos] < d) return pos;
    for (int lv = LG; lv >= 0; lv--)
      if (pos - (1 << lv) >= 0 && tbl[(size_t)lv * m + pos - (1 << lv)] >= d) pos -= 1 << lv;
    return pos - 1;
  }
};

// leaf level L[j] of every sorted body, the start of its leaf, and the number
// of cells starting at position j (levels lcp[j] + 1 .. L[j])
__global__ void tree_leaf_kernel(int n, Runs R, int ncrit, signed char *__restrict__ L, int *__restrict__ lstart,
                                 int *__restrict__ cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j > n) return;
  if (j == n) {
    cnt[j] = 0;
    return;
  }
  int lo = 0, hi = 30;  // the first level whose run holds <= ncrit bodies, else 30
  while (lo < hi) {
    const int d = (lo + hi) >> 1;
    if (R.next(j, d) - R.prev(j, d) <= ncrit)
      hi = d;
    else
      lo = d + 1;
  }
  L[j] = (signed char)lo;
  lstart[j] = R.prev(j, lo);
  const int c = lo - R.tbl[j];
  cnt[j] = c > 0 ? c : 0;
}

struct PreCells {
  int *pos, *level, *nbody, *nnode, *nchild;
};

// every cell (pre-order id cstart[j] + d - lcp[j] - 1 for level d at position j)
__global__ void tree_cells_kernel(int n, Runs R, const signed char *__restrict__ L, const int *__restrict__ cstart,
                                  PreCells C) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int d0 = R.tbl[j] + 1, d1 = L[j];
  for (int d = d0; d <= d1; d++) {
    const int p = cstart[j] + (d - d0);
    const int e = R.next(j, d);
    int nc = 0;
    if (d < d1)
      for (int k = j; k < e; k = R.next(k, d + 1)) nc++;
    C.pos[p] = j;
    C.level[p] = d;
    C.nbody[p] = e - j;
    C.nnode[p] = cstart[e] - p;
    C.nchild[p] = nc;
  }
}

// nodes2cells numbering, one level: every cell of level d numbers its children
__global__ void tree_number_kernel(int ncell, int d, Runs R, const int *__restrict__ cstart, PreCells C,
                                   int *__restrict__ fidx, int *__restrict__ fbase, int *__restrict__ fpar) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncell || C.level[p] != d || C.nchild[p] == 0) return;
  const int j = C.pos[p], e = j + C.nbody[p], base = fbase[p], nc = C.nchild[p];
  int i = 0, off = base + nc;
  for (int k = j; k < e; k = R.next(k, d + 1), i++) {
    const int pc = cstart[k] + (d + 1) - (R.tbl[k] + 1);  // the child's pre-order id (level d + 1 at k)
    fidx[pc] = base + i;
    fbase[pc] = off;
    fpar[pc] = fidx[p];
    off += C.nnode[pc] - 1;
  }
}

// the cell arrays in nodes2cells order
__global__ void tree_scatter_kernel(int ncell, PreCells C, const int *__restrict__ fidx, const int *__restrict__ fbase,
                                    const int *__restrict__ fpar, const unsigned long long *__restrict__ hi,
                                    const unsigned long long *__restrict__ lo, double X0, double X1, double X2,
                                    double R0, int *__restrict__ iparent, int *__restrict__ ichild,
                                    int *__restrict__ nchild, int *__restrict__ ibody, int *__restrict__ nbody,
                                    int *__restrict__ level, double *__restrict__ X, double *__restrict__ Rc) {
#pragma clang fp contract(off)
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ncell) return;
  const int f = fidx[p], j = C.pos[p], d = C.level[p];
  iparent[f] = fpar[p];
  nchild[f] = C.nchild[p];
  ichild[f] = C.nchild[p] ? fbase[p] : 0;
  ibody[f] = j;
  nbody[f] = C.nbody[p];
  level[f] = d;
  double c0 = X0, c1 = X1, c2 = X2;
  const unsigned long long h = hi[j], w = lo[j];
  for (int lev = 0; lev < d; lev++) {
    const int o = lev < 21 ? (int)((h >> (3 * (20 - lev))) & 7) : (int)((w >> (3 * (29 - lev))) & 7);
    const double r = R0 / (1 << (lev + 1));
    c0 += r * ((o & 1) * 2 - 1);
    c1 += r * (((o >> 1) & 1) * 2 - 1);
    c2 += r * (((o >> 2) & 1) * 2 - 1);
  }
  X[3 * (size_t)f] = c0;
  X[3 * (size_t)f + 1] = c1;
  X[3 * (size_t)f + 2] = c2;
  Rc[f] = R0 / (1 << d);
}

// the leaf start of every body, by input index
__global__ void tree_leafkey_kernel(int n, const int *__restrict__ idx, const int *__restrict__ lstart,
                                    unsigned *__restrict__ key, int *__restrict__ iota) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  key[idx[j]] = (unsigned)lstart[j];
  iota[j] = j;
}
}  // namespace

hipError_t device_lists(int ncell, const double *X, const double *R, const int *nchild, const int *ichild,
                        const int *nbody, double theta, int nspawn, const double xper[3], int root_i, int root_j,
                        std::vector<int> &m2l_ptr, std::vector<int> &m2l_src, std::vector<int> &p2p_ptr,
                        std::vector<int> &p2p_src, int64_t &body_pairs, hipStream_t s, PlanArrays *keep) {
  if (keep) {
    keep->lists = false;
    keep->m2l_src.reset();
    keep->p2p_src.reset();
    TCHK(hipGetDevice(&keep->device));
  }
  m2l_ptr.assign((size_t)ncell + 1, 0);
  p2p_ptr.assign((size_t)ncell + 1, 0);
  m2l_src.clear();
  p2p_src.clear();
  body_pairs = 0;
  if (ncell == 0) return hipSuccess;
  Buf dX(s), dR(s), dnc(s), dic(s), dnb(s), cnt(s), F0(s), F1(s), M(s), Pp(s), bp(s);
  TCHK(dX.need(sizeof(double) * 3 * (size_t)ncell));
  TCHK(dR.need(sizeof(double) * (size_t)ncell));
  TCHK(dnc.need(sizeof(int) * (size_t)ncell));
  TCHK(dic.need(sizeof(int) * (size_t)ncell));
  TCHK(dnb.need(sizeof(int) * (size_t)ncell));
  TCHK(hipMemcpyAsync(dX.p, X, sizeof(double) * 3 * (size_t)ncell, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dR.p, R, sizeof(double) * (size_t)ncell, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dnc.p, nchild, sizeof(int) * (size_t)ncell, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dic.p, ichild, sizeof(int) * (size_t)ncell, hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(dnb.p, nbody, sizeof(int) * (size_t)ncell, hipMemcpyHostToDevice, s));
  TravCells t{dX.as<double>(), dR.as<double>(), dnc.as<int>(), dic.as<int>(), dnb.as<int>()};
  TCHK(cnt.need(sizeof(unsigned long long) * 3));
  const int2 root = make_int2(root_i, root_j);
  // first sizes from the tree (a cell has tens of M2L and P2P partners), grown by half again when short
  TCHK(F0.need(sizeof(int2) * 8 * (size_t)ncell));
  TCHK(M.need(sizeof(int2) * 64 * (size_t)ncell));
  TCHK(Pp.need(sizeof(int2) * 16 * (size_t)ncell));
  TCHK(hipMemcpyAsync(F0.p, &root, sizeof(int2), hipMemcpyHostToDevice, s));
  size_t nf = 1, nm = 0, np = 0;
  unsigned long long h[3];
  while (nf > 0) {
    if (nf > 0x7fffffff) return hipErrorInvalidValue;  // frontier beyond one launch's index range
    // count: the M2L, P2P and next-frontier pairs of this traversal round
    TCHK(hipMemsetAsync(cnt.p, 0, sizeof(h), s));
    const unsigned nb = (unsigned)((nf + 255) / 256);
    trav_round_kernel<false><<<nb, 256, 0, s>>>(F0.as<int2>(), (int)nf, t, theta, xper[0], xper[1], xper[2], nspawn,
                                                cnt.as<unsigned long long>(), nullptr, nullptr, nullptr);
    TCHK(hipGetLastError());
    TCHK(hipMemcpyAsync(h, cnt.p, sizeof(h), hipMemcpyDeviceToHost, s));
    TCHK(hipStreamSynchronize(s));
    // room for the appended pairs (grown by half again when short), then emit
    auto grow = [&](Buf &b, size_t used, size_t add) -> hipError_t {
      const size_t want = sizeof(int2) * (used + add);
      if (want <= b.bytes) return hipSuccess;
      Buf nbuf(s);
      TCHK(nbuf.need(want + want / 2));
      if (used) TCHK(hipMemcpyAsync(nbuf.p, b.p, sizeof(int2) * used, hipMemcpyDeviceToDevice, s));
      b.swap(nbuf);  // the old buffer is released on the stream after the copy
      return hipSuccess;
    };
    TCHK(grow(M, nm, h[0]));
    TCHK(grow(Pp, np, h[1]));
    if (sizeof(int2) * h[2] > F1.bytes) TCHK(F1.need(sizeof(int2) * (h[2] + h[2] / 2)));
    const unsigned long long base[3] = {nm, np, 0};
    TCHK(hipMemcpyAsync(cnt.p, base, sizeof(base), hipMemcpyHostToDevice, s));
    trav_round_kernel<true><<<nb, 256, 0, s>>>(F0.as<int2>(), (int)nf, t, theta, xper[0], xper[1], xper[2], nspawn,
                                               cnt.as<unsigned long long>(), M.as<int2>(), Pp.as<int2>(),
                                               F1.as<int2>());
    TCHK(hipGetLastError());
    nm += h[0];
    np += h[1];
    nf = h[2];
    F0.swap(F1);
  }
  if (np > 0) {
    TCHK(bp.need(sizeof(unsigned long long)));
    TCHK(hipMemsetAsync(bp.p, 0, sizeof(unsigned long long), s));
    body_pairs_kernel<<<(unsigned)std::min<size_t>((np + 255) / 256, 2048), 256, 0, s>>>(
        Pp.as<int2>(), np, dnb.as<int>(), bp.as<unsigned long long>());
    TCHK(hipGetLastError());
    unsigned long long v = 0;
    TCHK(hipMemcpyAsync(&v, bp.p, sizeof v, hipMemcpyDeviceToHost, s));
    TCHK(hipStreamSynchronize(s));
    body_pairs = (int64_t)v;
  }
  TCHK(pairs_to_csr(M.as<int2>(), nm, ncell, m2l_ptr, m2l_src, s, keep ? &keep->m2l_src : nullptr));
  TCHK(pairs_to_csr(Pp.as<int2>(), np, ncell, p2p_ptr, p2p_src, s, keep ? &keep->p2p_src : nullptr));
  if (keep) keep->lists = true;
  return hipSuccess;
}

namespace {
struct ScratchCache {
  std::multimap<size_t, void *> blocks;  // capacity -> block
  size_t total = 0;
};
std::mutex g_scratch_mu;
std::map<hipStream_t, ScratchCache> g_scratch;
constexpr size_t kScratchKeep = (size_t)8 << 30;
}  // namespace

hipError_t plan_scratch_get(void **p, size_t *cap, size_t bytes, hipStream_t s) {
  *p = nullptr;
  *cap = 0;
  if (bytes == 0) bytes = 1;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto f = g_scratch.find(s);
    if (f != g_scratch.end()) {
      auto it = f->second.blocks.lower_bound(bytes);
      if (it != f->second.blocks.end() && it->first <= 2 * bytes + ((size_t)1 << 20)) {
        *p = it->second;
        *cap = it->first;
        f->second.total -= it->first;
        f->second.blocks.erase(it);
        return hipSuccess;
      }
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return e;
  }
  *cap = bytes;
  return hipSuccess;
}

void plan_scratch_put(void *p, size_t cap, hipStream_t s) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ScratchCache &c = g_scratch[s];
    if (c.total + cap <= kScratchKeep) {
      c.blocks.emplace(cap, p);
      c.total += cap;
      return;
    }
  }
  (void)hipFree(p);  // waits for the device's work, the stream's included
}

void plan_scratch_release(hipStream_t s) {
  ScratchCache c;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto f = g_scratch.find(s);
    if (f == g_scratch.end()) return;
    c = std::move(f->second);
    g_scratch.erase(f);
  }
  if (c.blocks.empty()) return;
  (void)hipStreamSynchronize(s);
  for (auto &b : c.blocks) (void)hipFree(b.second);
}

hipError_t DevArray::alloc(size_t b, int dev) {
  reset();
  if (b == 0) return hipSuccess;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    p = nullptr;
    return e;
  }
  bytes = b;
  device = dev;
  return hipSuccess;
}

void DevArray::reset() {
  if (p) {
    int cur = -1;
    const bool same = hipGetDevice(&cur) == hipSuccess && cur == device;
    if (!same) (void)hipSetDevice(device);
    (void)hipFree(p);
    if (!same && cur >= 0) (void)hipSetDevice(cur);
  }
  p = nullptr;
  bytes = 0;
  device = -1;
}

hipError_t fetch_lists(const PlanArrays &d, fmm::Lists &L, hipStream_t s) {
  if (!d.lists) return hipSuccess;
  if ((int64_t)L.m2l_src.size() != L.nm2l) {
    L.m2l_src.resize(L.nm2l);
    if (L.nm2l) TCHK(hipMemcpyAsync(L.m2l_src.data(), d.m2l_src.p, 4 * (size_t)L.nm2l, hipMemcpyDeviceToHost, s));
  }
  if ((int64_t)L.p2p_src.size() != L.np2p) {
    L.p2p_src.resize(L.np2p);
    if (L.np2p) TCHK(hipMemcpyAsync(L.p2p_src.data(), d.p2p_src.p, 4 * (size_t)L.np2p, hipMemcpyDeviceToHost, s));
  }
  return hipStreamSynchronize(s);
}

hipError_t fetch_tiles(const PlanArrays &d, fmm::M2LTiles &T, hipStream_t s) {
  if (!d.tiles) return hipSuccess;
  const size_t np = d.t_src.bytes / 4;
  if (T.src.size() != np) {
    T.slice.resize(4 * (size_t)d.nslice);
    T.src.resize(np);
    T.lt.resize(np);
    if (d.nslice) TCHK(hipMemcpyAsync(T.slice.data(), d.t_slice.p, 16 * (size_t)d.nslice, hipMemcpyDeviceToHost, s));
    if (np) {
      TCHK(hipMemcpyAsync(T.src.data(), d.t_src.p, 4 * np, hipMemcpyDeviceToHost, s));
      TCHK(hipMemcpyAsync(T.lt.data(), d.t_lt.p, 2 * np, hipMemcpyDeviceToHost, s));
    }
  }
  return hipStreamSynchronize(s);
}

hipError_t device_tree(int n, const double *x, int ncrit, const double X0[3], double R0, std::vector<int> &perm,
                       std::vector<int> &iparent, std::vector<int> &ichild, std::vector<int> &nchild,
                       std::vector<int> &ibody, std::vector<int> &nbody, std::vector<int> &level,
                       std::vector<double> &X, std::vector<double> &R, hipStream_t s, const double *dx_in) {
  perm.resize(n);
  if (n == 0) {
    for (auto *v : {&iparent, &ichild, &nchild, &ibody, &nbody, &level}) v->clear();
    X.clear();
    R.clear();
    return hipSuccess;
  }
  const unsigned nb = (unsigned)((n + 255) / 256);
  using u64 = unsigned long long;
  Buf dx(s), kh(s), kl(s), ka(s), kb(s), idx(s), idx2(s), tmp(s);
  const double *xd = dx_in;
  if (!xd) {
    TCHK(dx.need(sizeof(double) * 3 * (size_t)n));
    TCHK(hipMemcpyAsync(dx.p, x, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
    xd = dx.as<double>();
  }
  for (Buf *b : {&kh, &kl, &ka, &kb}) TCHK(b->need(8 * (size_t)n));
  TCHK(idx.need(4 * (size_t)n));
  TCHK(idx2.need(4 * (size_t)n));
  tree_keys_kernel<<<nb, 256, 0, s>>>(n, xd, X0[0], X0[1], X0[2], R0, kh.as<u64>(), kl.as<u64>(),
                                      idx.as<int>());
  TCHK(hipGetLastError());
  // stable sort by the path, least significant part first: by levels 21-29
  // (kl), then by levels 0-20 (kh, gathered into the first sort's order)
  size_t tb = 0, tb2 = 0;
  TCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, kl.as<u64>(), ka.as<u64>(), idx.as<int>(), idx2.as<int>(), n,
                                          0, 27, s));
  TCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tb2, ka.as<u64>(), kb.as<u64>(), idx2.as<int>(), idx.as<int>(), n,
                                          0, 63, s));
  TCHK(tmp.need(std::max(tb, tb2)));
  TCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, kl.as<u64>(), ka.as<u64>(), idx.as<int>(), idx2.as<int>(), n,
                                          0, 27, s));
  gather_u64_kernel<<<nb, 256, 0, s>>>(n, kh.as<u64>(), idx2.as<int>(), ka.as<u64>());
  TCHK(hipGetLastError());
  TCHK(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb2, ka.as<u64>(), kb.as<u64>(), idx2.as<int>(), idx.as<int>(), n,
                                          0, 63, s));
  gather_u64_kernel<<<nb, 256, 0, s>>>(n, kl.as<u64>(), idx.as<int>(), ka.as<u64>());
  TCHK(hipGetLastError());
  // sorted paths: levels 0-20 in kb, 21-29 in ka; idx: input index of every sorted body
  const u64 *hiS = kb.as<u64>(), *loS = ka.as<u64>();
  const int m = n + 1;
  int LG = 0;
  while ((1 << (LG + 1)) <= m) LG++;
  Buf tbl(s), Lv(s), lst(s), cnt(s), cst(s);
  TCHK(tbl.need((size_t)(LG + 1) * m));
  tree_lcp_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(n, hiS, loS, tbl.as<signed char>());
  TCHK(hipGetLastError());
  for (int lv = 1; lv <= LG; lv++) {
    tree_table_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(
        m, tbl.as<signed char>() + (size_t)(lv - 1) * m, 1 << (lv - 1), tbl.as<signed char>() + (size_t)lv * m);
    TCHK(hipGetLastError());
  }
  Runs RR{tbl.as<signed char>(), m, LG};
  TCHK(Lv.need(m));
  TCHK(lst.need(4 * (size_t)m));
  TCHK(cnt.need(4 * (size_t)m));
  TCHK(cst.need(4 * (size_t)m));
  tree_leaf_kernel<<<(unsigned)((m + 255) / 256), 256, 0, s>>>(n, RR, ncrit, Lv.as<signed char>(), lst.as<int>(),
                                                               cnt.as<int>());
  TCHK(hipGetLastError());
  size_t ts = 0;
  TCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, ts, cnt.as<int>(), cst.as<int>(), m, s));
  Buf tmp2(s);
  TCHK(tmp2.need(ts));
  TCHK(hipcub::DeviceScan::ExclusiveSum(tmp2.p, ts, cnt.as<int>(), cst.as<int>(), m, s));
  int ncell = 0;
  TCHK(hipMemcpyAsync(&ncell, cst.as<int>() + n, sizeof(int), hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  Buf cpos(s), clev(s), cnb(s), cnn(s), cnc(s), fidx(s), fbase(s), fpar(s);
  for (Buf *b : {&cpos, &clev, &cnb, &cnn, &cnc, &fidx, &fbase, &fpar}) TCHK(b->need(4 * (size_t)ncell));
  PreCells C{cpos.as<int>(), clev.as<int>(), cnb.as<int>(), cnn.as<int>(), cnc.as<int>()};
  tree_cells_kernel<<<nb, 256, 0, s>>>(n, RR, Lv.as<signed char>(), cst.as<int>(), C);
  TCHK(hipGetLastError());
  // the root: index 0, children from 1, parent 0
  const int root[3] = {0, 1, 0};
  TCHK(hipMemcpyAsync(fidx.p, &root[0], sizeof(int), hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(fbase.p, &root[1], sizeof(int), hipMemcpyHostToDevice, s));
  TCHK(hipMemcpyAsync(fpar.p, &root[2], sizeof(int), hipMemcpyHostToDevice, s));
  const unsigned nbc = (unsigned)((ncell + 255) / 256);
  for (int d = 0; d < 30; d++) {
    tree_number_kernel<<<nbc, 256, 0, s>>>(ncell, d, RR, cst.as<int>(), C, fidx.as<int>(), fbase.as<int>(),
                                           fpar.as<int>());
    TCHK(hipGetLastError());
  }
  Buf oip(s), oic(s), onc(s), oib(s), onb(s), olv(s), oX(s), oR(s);
  for (Buf *b : {&oip, &oic, &onc, &oib, &onb, &olv}) TCHK(b->need(4 * (size_t)ncell));
  TCHK(oX.need(24 * (size_t)ncell));
  TCHK(oR.need(8 * (size_t)ncell));
  tree_scatter_kernel<<<nbc, 256, 0, s>>>(ncell, C, fidx.as<int>(), fbase.as<int>(), fpar.as<int>(),
                                          hiS, loS, X0[0], X0[1], X0[2], R0, oip.as<int>(), oic.as<int>(), onc.as<int>(), oib.as<int>(),
                                          onb.as<int>(), olv.as<int>(), oX.as<double>(), oR.as<double>());
  TCHK(hipGetLastError());
  // bodies: by leaf, input order within a leaf
  Buf key(s), key2(s), io(s), io2(s);
  TCHK(key.need(4 * (size_t)n));
  TCHK(key2.need(4 * (size_t)n));
  TCHK(io.need(4 * (size_t)n));
  TCHK(io2.need(4 * (size_t)n));
  tree_leafkey_kernel<<<nb, 256, 0, s>>>(n, idx.as<int>(), lst.as<int>(), key.as<unsigned>(), io.as<int>());
  TCHK(hipGetLastError());
  int nbits = 1;
  while ((1ll << nbits) < (long long)n) nbits++;
  size_t tk = 0;
  TCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, tk, key.as<unsigned>(), key2.as<unsigned>(), io.as<int>(),
                                          io2.as<int>(), n, 0, nbits, s));
  Buf tmp3(s);
  TCHK(tmp3.need(tk));
  TCHK(hipcub::DeviceRadixSort::SortPairs(tmp3.p, tk, key.as<unsigned>(), key2.as<unsigned>(), io.as<int>(),
                                          io2.as<int>(), n, 0, nbits, s));
  for (auto *v : {&iparent, &ichild, &nchild, &ibody, &nbody, &level}) v->resize(ncell);
  X.resize(3 * (size_t)ncell);
  R.resize(ncell);
  TCHK(hipMemcpyAsync(perm.data(), io2.p, 4 * (size_t)n, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(iparent.data(), oip.p, 4 * (size_t)ncell, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(ichild.data(), oic.p, 4 * (size_t)ncell, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(nchild.data(), onc.p, 4 * (size_t)ncell, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(ibody.data(), oib.p, 4 * (size_t)ncell, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(nbody.data(), onb.p, 4 * (size_t)ncell, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(level.data(), olv.p, 4 * (size_t)ncell, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(X.data(), oX.p, 24 * (size_t)ncell, hipMemcpyDeviceToHost, s));
  TCHK(hipMemcpyAsync(R.data(), oR.p, 8 * (size_t)ncell, hipMemcpyDeviceToHost, s));
  TCHK(hipStreamSynchronize(s));
  return hipSuccess;
}

}  // namespace fmmdev

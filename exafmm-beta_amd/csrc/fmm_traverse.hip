// fmm_traverse.hip -- the interaction lists built on the device.
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

#include <cstdint>
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

__global__ void pair_keys_kernel(const int2 *__restrict__ pr, size_t n, unsigned long long *__restrict__ key,
                                 int *__restrict__ count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int2 p = pr[i];
  key[i] = ((unsigned long long)(unsigned)p.x << 32) | (unsigned)p.y;
  atomicAdd(&count[p.x], 1);
}

__global__ void key_src_kernel(const unsigned long long *__restrict__ key, size_t n, int *__restrict__ src) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) src[i] = (int)(unsigned)(key[i] & 0xffffffffull);
}

__global__ void body_pairs_kernel(const int2 *__restrict__ pr, size_t n, const int *__restrict__ nbody,
                                  unsigned long long *__restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long v = 0;
  if (i < n) v = (unsigned long long)nbody[pr[i].x] * (unsigned long long)nbody[pr[i].y];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(out, v);
}

// device scratch from the stream-ordered pool: allocation and release are
// queued on the stream, without the device-wide synchronisation of hipFree
struct Buf {
  hipStream_t s;
  void *p = nullptr;
  size_t bytes = 0;
  explicit Buf(hipStream_t st) : s(st) {}
  Buf(const Buf &) = delete;
  ~Buf() {
    if (p) (void)hipFreeAsync(p, s);
  }
  hipError_t need(size_t b) {
    if (b <= bytes) return hipSuccess;
    if (p) (void)hipFreeAsync(p, s);
    p = nullptr;
    bytes = 0;
    hipError_t e = hipMallocAsync(&p, b, s);
    if (e == hipSuccess) bytes = b;
    return e;
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
// copied to the host
hipError_t pairs_to_csr(const int2 *pr, size_t n, int ncell, std::vector<int> &ptr, std::vector<int> &src,
                        hipStream_t s) {
  ptr.assign((size_t)ncell + 1, 0);
  src.resize(n);
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
    TCHK(out.need(sizeof(int) * n));
    key_src_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(db.Current(), n, out.as<int>());
    TCHK(hipGetLastError());
    TCHK(hipMemcpyAsync(src.data(), out.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
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

}  // namespace

hipError_t device_lists(int ncell, const double *X, const double *R, const int *nchild, const int *ichild,
                        const int *nbody, double theta, int nspawn, const double xper[3], int root_i, int root_j,
                        std::vector<int> &m2l_ptr, std::vector<int> &m2l_src, std::vector<int> &p2p_ptr,
                        std::vector<int> &p2p_src, int64_t &body_pairs, hipStream_t s) {
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
    // count: this round's M2L, P2P and frontier pairs
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
    body_pairs_kernel<<<(unsigned)((np + 255) / 256), 256, 0, s>>>(Pp.as<int2>(), np, dnb.as<int>(),
                                                                    bp.as<unsigned long long>());
    TCHK(hipGetLastError());
    unsigned long long v = 0;
    TCHK(hipMemcpyAsync(&v, bp.p, sizeof v, hipMemcpyDeviceToHost, s));
    TCHK(hipStreamSynchronize(s));
    body_pairs = (int64_t)v;
  }
  TCHK(pairs_to_csr(M.as<int2>(), nm, ncell, m2l_ptr, m2l_src, s));
  TCHK(pairs_to_csr(Pp.as<int2>(), np, ncell, p2p_ptr, p2p_src, s));
  return hipSuccess;
}

}  // namespace fmmdev

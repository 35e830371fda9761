// ref_harness.cpp -- drives exafmm-beta's own Laplace operators (kernels/laplace.h)
// and input generators (include/dataset.h), compiled from the reference tree
// in place by oracle/Makefile into oracle/_ref/libexafmm_ref.so.
//
// TEST INFRASTRUCTURE ONLY.  This file contains no reference code: it includes
// the reference headers from the reference tree and calls them.  The
// reference's tree, traversal and up/down pass headers pull in the generated
// config.h (include/thread.h:3) and cannot be built here without the
// reference's build system, so the pass order of include/up_down_pass.h:14-75
// is re-driven below over a tree and pair lists supplied by the caller.
#include "types.h"
#include "kernel.h"
#include "dataset.h"
#include <chrono>
#include <cstring>

using namespace exafmm_laplace;

namespace {
struct RefCtx {
  Kernel kernel;
  explicit RefCtx(int P, double eps2) : kernel(P, eps2, complex_t(1., .1)) {}
};

void load_bodies(Bodies &b, int n, const double *x, const double *q) {
  b.resize(n);
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) b[i].X[d] = x[3 * i + d];
    b[i].SRC = q ? q[i] : 0.0;
    b[i].TRG = 0;
    b[i].IBODY = i;
    b[i].ICELL = 0;
    b[i].WEIGHT = 1;
  }
}

void set_cell(Cell &c, const double *X, double R, int nterm) {
  for (int d = 0; d < 3; d++) c.X[d] = X[d];
  c.R = R;
  c.IPARENT = 0;
  c.ICHILD = 0;
  c.NCHILD = 0;
  c.IBODY = 0;
  c.NBODY = 0;
  c.M.assign(nterm, 0.0);
  c.L.assign(nterm, 0.0);
}

void copy_in(std::vector<complex_t> &v, const double *a, int nterm) {
  for (int i = 0; i < nterm; i++) v[i] = complex_t(a[2 * i], a[2 * i + 1]);
}
void add_out(const std::vector<complex_t> &v, double *a, int nterm) {
  for (int i = 0; i < nterm; i++) {
    a[2 * i] += v[i].real();
    a[2 * i + 1] += v[i].imag();
  }
}
}  // namespace

extern "C" {

void *ref_kernel_create(int P, double eps2) { return new RefCtx(P, eps2); }
void ref_kernel_destroy(void *k) { delete static_cast<RefCtx *>(k); }

// Kernel::P2P (SIMD + scalar tail), trg accumulated
void ref_p2p(void *kp, int ni, const double *xi, double *trgi, int nj, const double *xj,
             const double *qj, const double *xper) {
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Bodies bi, bj;
  load_bodies(bi, ni, xi, nullptr);
  load_bodies(bj, nj, xj, qj);
  Cells cells(2);
  cells[0].BODY = bi.begin();
  cells[0].NBODY = ni;
  cells[1].BODY = bj.begin();
  cells[1].NBODY = nj;
  for (int d = 0; d < 3; d++) K.Xperiodic[d] = xper ? xper[d] : 0.0;
  K.P2P(cells.begin(), cells.begin() + 1);
  K.Xperiodic = 0;
  for (int i = 0; i < ni; i++)
    for (int c = 0; c < 4; c++) trgi[4 * i + c] += bi[i].TRG[c];
}

void ref_p2m(void *kp, const double *Xc, int n, const double *x, const double *q, double *M) {
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Bodies b;
  load_bodies(b, n, x, q);
  Cells cells(1);
  set_cell(cells[0], Xc, 1.0, K.NTERM);
  cells[0].BODY = b.begin();
  cells[0].NBODY = n;
  K.P2M(cells.begin());
  add_out(cells[0].M, M, K.NTERM);
}

void ref_m2m(void *kp, const double *Xp, const double *Xc, const double *Mc, double *Mp) {
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Cells cells(2);
  set_cell(cells[0], Xp, 2.0, K.NTERM);
  set_cell(cells[1], Xc, 1.0, K.NTERM);
  copy_in(cells[1].M, Mc, K.NTERM);
  cells[0].ICHILD = 1;
  cells[0].NCHILD = 1;
  K.M2M(cells.begin(), cells.begin());
  add_out(cells[0].M, Mp, K.NTERM);
}

void ref_m2l(void *kp, const double *Xi, const double *Xj, const double *xper, const double *Mj,
             double *Li) {
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Cells cells(2);
  set_cell(cells[0], Xi, 1.0, K.NTERM);
  set_cell(cells[1], Xj, 1.0, K.NTERM);
  copy_in(cells[1].M, Mj, K.NTERM);
  for (int d = 0; d < 3; d++) K.Xperiodic[d] = xper ? xper[d] : 0.0;
  K.M2L(cells.begin(), cells.begin() + 1);
  K.Xperiodic = 0;
  add_out(cells[0].L, Li, K.NTERM);
}

void ref_l2l(void *kp, const double *Xc, const double *Xp, const double *Lp, double *Lc) {
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Cells cells(2);
  set_cell(cells[0], Xp, 2.0, K.NTERM);
  set_cell(cells[1], Xc, 1.0, K.NTERM);
  copy_in(cells[0].L, Lp, K.NTERM);
  cells[1].IPARENT = 0;
  K.L2L(cells.begin() + 1, cells.begin());
  add_out(cells[1].L, Lc, K.NTERM);
}

void ref_l2p(void *kp, const double *Xc, const double *L, int n, const double *x, double *trg) {
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Bodies b;
  load_bodies(b, n, x, nullptr);
  Cells cells(1);
  set_cell(cells[0], Xc, 1.0, K.NTERM);
  copy_in(cells[0].L, L, K.NTERM);
  cells[0].BODY = b.begin();
  cells[0].NBODY = n;
  K.L2P(cells.begin());
  for (int i = 0; i < n; i++)
    for (int c = 0; c < 4; c++) trg[4 * i + c] += b[i].TRG[c];
}

// Full evaluation with the reference operators over a caller-supplied tree
// (CellBase fields, bodies in tree order) and pair list in visiting order.
// Pass order follows include/up_down_pass.h:14-75.
struct Pass {
  Kernel &K;
  Cells &cells;
  void post(C_iter C) {
    C_iter C0 = cells.begin();
    for (C_iter CC = C0 + C->ICHILD; CC != C0 + C->ICHILD + C->NCHILD; CC++) post(CC);
    if (C->NCHILD == 0) K.P2M(C);
    else K.M2M(C, C0);
  }
  void pre(C_iter C) {
    C_iter C0 = cells.begin();
    K.L2L(C, C0);
    if (C->NCHILD == 0) K.L2P(C);
    for (C_iter CC = C0 + C->ICHILD; CC != C0 + C->ICHILD + C->NCHILD; CC++) pre(CC);
  }
};

void ref_eval(void *kp, int n, const double *x, const double *q, int ncell, const int *iparent,
              const int *ichild, const int *nchild, const int *ibody, const int *nbody,
              const double *X, const double *R, long npair, const int *ci, const int *cj,
              const int *type, double *trg, double *Mout, double *Lout) {
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Bodies bodies;
  load_bodies(bodies, n, x, q);
  Cells cells(ncell);
  for (int c = 0; c < ncell; c++) {
    set_cell(cells[c], X + 3 * c, R[c], K.NTERM);
    cells[c].IPARENT = iparent[c];
    cells[c].ICHILD = ichild[c];
    cells[c].NCHILD = nchild[c];
    cells[c].IBODY = ibody[c];
    cells[c].NBODY = nbody[c];
    cells[c].BODY = bodies.begin() + ibody[c];
  }
  if (ncell == 0) return;
  Pass pass{K, cells};
  pass.post(cells.begin());
  K.Xperiodic = 0;
  for (long p = 0; p < npair; p++) {
    if (type[p] == 0) K.M2L(cells.begin() + ci[p], cells.begin() + cj[p]);
    else K.P2P(cells.begin() + ci[p], cells.begin() + cj[p]);
  }
  C_iter C0 = cells.begin();
  if (C0->NCHILD == 0) K.L2P(C0);
  for (C_iter CC = C0 + C0->ICHILD; CC != C0 + C0->ICHILD + C0->NCHILD; CC++) pass.pre(CC);
  for (int i = 0; i < n; i++)
    for (int c = 0; c < 4; c++) trg[4 * i + c] += bodies[i].TRG[c];
  if (Mout)
    for (int c = 0; c < ncell; c++) add_out(cells[c].M, Mout + 2 * (size_t)c * K.NTERM, K.NTERM);
  if (Lout)
    for (int c = 0; c < ncell; c++) add_out(cells[c].L, Lout + 2 * (size_t)c * K.NTERM, K.NTERM);
}

// Timed variant for the CPU baseline: upward pass, the M2L pairs, the P2P
// pairs and the downward pass, each timed separately (wall clock, one core).
// times[0..3] = upward, M2L, P2P, downward seconds.
void ref_time(void *kp, int n, const double *x, const double *q, int ncell, const int *iparent,
              const int *ichild, const int *nchild, const int *ibody, const int *nbody,
              const double *X, const double *R, long nm2l, const int *mi, const int *mj, long np2p,
              const int *pi, const int *pj, double *trg, double *times) {
  using clk = std::chrono::steady_clock;
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Bodies bodies;
  load_bodies(bodies, n, x, q);
  Cells cells(ncell);
  for (int c = 0; c < ncell; c++) {
    set_cell(cells[c], X + 3 * c, R[c], K.NTERM);
    cells[c].IPARENT = iparent[c];
    cells[c].ICHILD = ichild[c];
    cells[c].NCHILD = nchild[c];
    cells[c].IBODY = ibody[c];
    cells[c].NBODY = nbody[c];
    cells[c].BODY = bodies.begin() + ibody[c];
  }
  if (ncell == 0) return;
  Pass pass{K, cells};
  K.Xperiodic = 0;
  auto t0 = clk::now();
  pass.post(cells.begin());
  auto t1 = clk::now();
  for (long p = 0; p < nm2l; p++) K.M2L(cells.begin() + mi[p], cells.begin() + mj[p]);
  auto t2 = clk::now();
  for (long p = 0; p < np2p; p++) K.P2P(cells.begin() + pi[p], cells.begin() + pj[p]);
  auto t3 = clk::now();
  C_iter C0 = cells.begin();
  if (C0->NCHILD == 0) K.L2P(C0);
  for (C_iter CC = C0 + C0->ICHILD; CC != C0 + C0->ICHILD + C0->NCHILD; CC++) pass.pre(CC);
  auto t4 = clk::now();
  for (int i = 0; i < n; i++)
    for (int c = 0; c < 4; c++) trg[4 * i + c] += bodies[i].TRG[c];
  auto sec = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  times[0] = sec(t0, t1);
  times[1] = sec(t1, t2);
  times[2] = sec(t2, t3);
  times[3] = sec(t3, t4);
}

// Sampled variant for the CPU baseline of large systems: every `stride`-th
// leaf's P2M, internal cell's M2M, non-root cell's L2L and leaf's L2P, and
// the given M2L and P2P pairs, each phase timed separately on one core over
// cells whose expansions hold what the earlier phases left (sampled: partial
// sums, same operation count).  times[0..5] = P2M, M2M, M2L, P2P, L2L, L2P
// seconds; work[0..5] = bodies, children, pairs, body pairs, cells, bodies
// processed (for scaling to the whole solve).
void ref_time_sampled(void *kp, int n, const double *x, const double *q, int ncell, const int *iparent,
                      const int *ichild, const int *nchild, const int *ibody, const int *nbody,
                      const double *X, const double *R, int stride, long nm2l, const int *mi,
                      const int *mj, long np2p, const int *pi, const int *pj, double *times,
                      double *work) {
  using clk = std::chrono::steady_clock;
  Kernel &K = static_cast<RefCtx *>(kp)->kernel;
  Bodies bodies;
  load_bodies(bodies, n, x, q);
  Cells cells(ncell);
  for (int c = 0; c < ncell; c++) {
    set_cell(cells[c], X + 3 * c, R[c], K.NTERM);
    cells[c].IPARENT = iparent[c];
    cells[c].ICHILD = ichild[c];
    cells[c].NCHILD = nchild[c];
    cells[c].IBODY = ibody[c];
    cells[c].NBODY = nbody[c];
    cells[c].BODY = bodies.begin() + ibody[c];
  }
  for (int k = 0; k < 6; k++) times[k] = work[k] = 0;
  if (ncell == 0 || stride < 1) return;
  K.Xperiodic = 0;
  C_iter C0 = cells.begin();
  auto sec = [](clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double>(b - a).count();
  };
  auto t0 = clk::now();
  int k = 0;
  for (int c = 0; c < ncell; c++)
    if (nchild[c] == 0 && (k++ % stride) == 0) {
      K.P2M(C0 + c);
      work[0] += nbody[c];
    }
  auto t1 = clk::now();
  k = 0;
  for (int c = ncell - 1; c >= 0; c--)
    if (nchild[c] > 0 && (k++ % stride) == 0) {
      K.M2M(C0 + c, C0);
      work[1] += nchild[c];
    }
  auto t2 = clk::now();
  for (long p = 0; p < nm2l; p++) K.M2L(C0 + mi[p], C0 + mj[p]);
  auto t3 = clk::now();
  for (long p = 0; p < np2p; p++) {
    K.P2P(C0 + pi[p], C0 + pj[p]);
    work[3] += (double)nbody[pi[p]] * nbody[pj[p]];
  }
  auto t4 = clk::now();
  k = 0;
  for (int c = 1; c < ncell; c++)
    if ((k++ % stride) == 0) {
      K.L2L(C0 + c, C0);
      work[4] += 1;
    }
  auto t5 = clk::now();
  k = 0;
  for (int c = 0; c < ncell; c++)
    if (nchild[c] == 0 && (k++ % stride) == 0) {
      K.L2P(C0 + c);
      work[5] += nbody[c];
    }
  auto t6 = clk::now();
  work[2] = (double)nm2l;
  times[0] = sec(t0, t1);
  times[1] = sec(t1, t2);
  times[2] = sec(t2, t3);
  times[3] = sec(t3, t4);
  times[4] = sec(t4, t5);
  times[5] = sec(t5, t6);
}

// Dataset::initBodies (include/dataset.h:230-255): positions + initSource.
void ref_init_bodies(int n, const char *dist, int seed, double *x, double *q) {
  Dataset data;
  Bodies b = data.initBodies(n, dist, seed);
  if ((int)b.size() < n) n = (int)b.size();  // lattice returns floor(cbrt(n))^3 bodies
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) x[3 * i + d] = b[i].X[d];
    q[i] = b[i].SRC;
  }
}

}  // extern "C"

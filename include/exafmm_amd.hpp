// exafmm_amd.hpp -- C++ drop-in for exafmm-beta's Laplace FMM call surface,
// header-only, on top of the C ABI in exafmm_amd.h (libexafmm_amd.so).
//
// Same class names, constructor arguments and call order as the reference's
// examples/fmm.cxx:13-140:
//   BoundBox::getBounds        include/bound_box.h:11-25
//   BuildTree::buildTree       include/build_tree.h:186-209  (permutes bodies in place)
//   Kernel(P, eps2, wavek)     kernels/laplace.h:116         (owns the device context)
//   UpDownPass::upwardPass     include/up_down_pass.h:49-60  (P2M + M2M on the GPU)
//   Traversal::traverse        include/traversal.h:326-352   (lists on the host, M2L + P2P on the GPU)
//   UpDownPass::downwardPass   include/up_down_pass.h:63-75  (L2L + L2P; TRG += results)
//   Traversal::direct          include/traversal.h:355-377
//   Dataset, Verify, Args      include/dataset.h, verify.h, args.h
//   Kernel::P2P/P2M/M2M/M2L/L2L/L2P  kernels/laplace.h:150-388 (per-cell calls, one launch each)
// Body and Cell keep the reference's fields; Cell::M and Cell::L are filled
// after the passes (reference normalisation).  Errors that the reference
// would hit as undefined behaviour or abort() are reported by throwing
// exafmm_amd::Error.  The passes run one device launch per phase; the
// per-cell Kernel methods are for callers that drive their own traversal
// (each is one launch, so a whole solve should use the passes).
#pragma once
#include <getopt.h>
#include <sys/time.h>

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "exafmm_amd.h"

namespace exafmm_amd {

typedef double real_t;
typedef std::complex<real_t> complex_t;
const real_t EPS = 1e-16;

struct Error : std::runtime_error {
  explicit Error(const std::string &m) : std::runtime_error(m) {}
};

template <int N> struct vec {
  real_t d[N];
  vec() {}
  vec(real_t v) {
    for (int i = 0; i < N; i++) d[i] = v;
  }
  real_t &operator[](int i) { return d[i]; }
  const real_t &operator[](int i) const { return d[i]; }
  vec &operator=(real_t v) {
    for (int i = 0; i < N; i++) d[i] = v;
    return *this;
  }
};
typedef vec<3> vec3;
typedef vec<4> vec4;

// Compensated (Kahan) summation, include/kahan.h:9-124: `s` is the running
// sum and `c` the rounding error of the last additions (the part of the true
// sum that `s` lost, with its sign reversed), so the value is s - c.  Same
// members and operators as the reference's kahan<T>; its conversion and
// stream output use s - c (the reference's operator T() returns s + c, which
// adds the lost part with the wrong sign), and adding a kahan<T> adds its
// value s - c.  The device kernels' compensated accumulation (mixed
// precision P2P) follows the same recurrence.
template <typename T> struct kahan {
  T s, c;
  kahan() {}
  kahan(const T &v) : s(v), c(0) {}
  kahan(const kahan &v) : s(v.s), c(v.c) {}
  kahan &operator=(const T v) {
    s = v;
    c = 0;
    return *this;
  }
  kahan &operator=(const kahan &v) {
    s = v.s;
    c = v.c;
    return *this;
  }
  kahan &operator+=(const T v) {  // Kahan's recurrence: y = v - c, t = s + y, c = (t - s) - y
    const T y = v - c;
    const T t = s + y;
    c = (t - s) - y;
    s = t;
    return *this;
  }
  kahan &operator-=(const T v) { return *this += -v; }
  kahan &operator*=(const T v) {
    s *= v;
    c *= v;
    return *this;
  }
  kahan &operator/=(const T v) {
    s /= v;
    c /= v;
    return *this;
  }
  kahan &operator+=(const kahan &v) {
    *this += v.s;
    return *this += -v.c;
  }
  kahan &operator-=(const kahan &v) {
    *this += -v.s;
    return *this += v.c;
  }
  kahan &operator*=(const kahan &v) { return *this *= T(v); }
  kahan &operator/=(const kahan &v) { return *this /= T(v); }
  kahan operator-() const {
    kahan t;
    t.s = -s;
    t.c = -c;
    return t;
  }
  operator T() const { return s - c; }
  friend std::ostream &operator<<(std::ostream &o, const kahan &v) { return o << (v.s - v.c); }
  friend std::istream &operator>>(std::istream &i, kahan &v) {
    i >> v.s;
    v.c = 0;
    return i;
  }
};
// include/types.h:35-42: kreal_t / kcomplex_t are the compensated types when
// EXAFMM_USE_KAHAN is set, the plain ones otherwise
#if defined(EXAFMM_USE_KAHAN) && EXAFMM_USE_KAHAN
typedef kahan<real_t> kreal_t;
typedef kahan<complex_t> kcomplex_t;
#else
typedef real_t kreal_t;
typedef complex_t kcomplex_t;
#endif

//! include/types.h:72-84
struct Body {
  vec3 X;
  real_t SRC;
  int IBODY;
  int IRANK;
  int64_t ICELL;
  real_t WEIGHT;
  vec4 TRG;
};
typedef std::vector<Body> Bodies;
typedef Bodies::iterator B_iter;

//! include/types.h:98-119
struct CellBase {
  int IPARENT, ICHILD, NCHILD, IBODY, NBODY;
  int numP2P = 0, numM2L = 0;  // interaction list sizes (include/types.h:104-107)
  uint64_t ICELL;
  real_t WEIGHT;
  vec3 X;
  real_t R;
  B_iter BODY;
};
struct Cell : public CellBase {
  std::vector<complex_t> M, L;
};
typedef std::vector<Cell> Cells;
typedef Cells::iterator C_iter;

struct Box {
  vec3 X;
  real_t R;
};
struct Bounds {
  vec3 Xmin, Xmax;
};

namespace logger {  // include/logger.h timers (wall clock, named)
inline bool verbose = false;
inline const int stringLength = 20;
inline const int decimal = 7;
inline std::map<std::string, double> &timers() {
  static std::map<std::string, double> t;
  return t;
}
inline double now() {
  timeval tv;
  gettimeofday(&tv, nullptr);
  return tv.tv_sec + tv.tv_usec * 1e-6;
}
inline void startTimer(const std::string &n) { timers()[n] = now(); }
inline double stopTimer(const std::string &n) {
  const double t = now() - timers()[n];
  if (verbose)
    std::cout << std::setw(stringLength) << std::left << n << " : " << std::setprecision(decimal)
              << std::fixed << t << " s" << std::endl;
  return t;
}
inline void printTitle(const std::string &title) {
  if (!verbose) return;
  std::string t = "--- " + title + " ";
  t.resize(34, '-');
  std::cout << t << std::endl;
}
}  // namespace logger

//! include/bound_box.h:11-25 (the 1e-5 per-body drift is kept)
class BoundBox {
 public:
  Bounds getBounds(Bodies &bodies) {
    Bounds b;
    if (bodies.empty()) {
      b.Xmin = b.Xmax = 0.0;
      return b;
    }
    b.Xmin = b.Xmax = bodies.front().X;
    return getBounds(bodies, b);
  }
  //! include/bound_box.h:28-37: widen given bounds by the bodies (same drift)
  Bounds getBounds(Bodies &bodies, Bounds b) {
    for (B_iter B = bodies.begin(); B != bodies.end(); B++)
      for (int d = 0; d < 3; d++) {
        const real_t lo = b.Xmin[d] - 1e-5, hi = b.Xmax[d] + 1e-5;
        b.Xmin[d] = B->X[d] < lo ? B->X[d] : lo;
        b.Xmax[d] = B->X[d] > hi ? B->X[d] : hi;
      }
    return b;
  }
};

//! include/build_tree.h:186-209: the octree of the reference (same cells,
//! same order), bodies permuted in place into tree order.
class BuildTree {
  const int ncrit;
  int numLevels = 0;

 public:
  explicit BuildTree(int _ncrit) : ncrit(_ncrit) {}
  //! the octree of bodies in the given bounds (the reference's box from them)
  Cells buildTree(Bodies &bodies, Bodies &buffer, Bounds bounds) {
    logger::startTimer("Grow tree");
    const int n = (int)bodies.size();
    std::vector<double> x(3 * (size_t)n);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) x[3 * (size_t)i + d] = bodies[i].X[d];
    fmm_plan *pl = nullptr;
    if (fmm_plan_build_tree(&pl, n, x.data(), ncrit, bounds.Xmin.d, bounds.Xmax.d) != FMM_OK)
      throw Error("buildTree failed");
    fmm_stats s;
    fmm_plan_stats(pl, &s);
    const int nc = s.ncell;
    std::vector<int> perm(n), ip(nc), ic(nc), nch(nc), ib(nc), nb(nc);
    std::vector<double> X(3 * (size_t)nc), R(nc);
    fmm_plan_tree(pl, perm.data(), ip.data(), ic.data(), nch.data(), ib.data(), nb.data(), X.data(), R.data());
    std::vector<uint64_t> key(nc);
    fmm_plan_list_counts(pl, key.data(), nullptr, nullptr);
    fmm_plan_destroy(pl);
    buffer.resize(n);
    for (int i = 0; i < n; i++) buffer[i] = bodies[perm[i]];
    bodies.swap(buffer);
    logger::stopTimer("Grow tree");
    Cells cells(nc);
    for (int c = 0; c < nc; c++) {
      Cell &C = cells[c];
      C.IPARENT = ip[c];
      C.ICHILD = ic[c];
      C.NCHILD = nch[c];
      C.IBODY = ib[c];
      C.NBODY = nb[c];
      C.ICELL = key[c];  // Morton key (include/build_tree.h:132)
      C.WEIGHT = 0;
      for (int d = 0; d < 3; d++) C.X[d] = X[3 * (size_t)c + d];
      C.R = R[c];
      C.BODY = bodies.begin() + C.IBODY;
    }
    numLevels = s.depth;
    return cells;
  }
  void printTreeData(Cells &cells) {
    if (!logger::verbose || cells.empty()) return;
    logger::printTitle("Tree stats");
    std::cout << std::setw(logger::stringLength) << std::left << "Bodies"
              << " : " << cells.front().NBODY << std::endl
              << std::setw(logger::stringLength) << std::left << "Cells"
              << " : " << cells.size() << std::endl
              << std::setw(logger::stringLength) << std::left << "Tree depth"
              << " : " << numLevels << std::endl;
  }
};

//! kernels/laplace.h:116: owns the device context shared by UpDownPass and
//! Traversal (the reference shares its Kernel between them the same way).
class Kernel {
 public:
  const int P;
  const int NTERM;
  real_t eps2;
  complex_t wavek;
  vec3 Xperiodic;
  fmm_ctx *ctx = nullptr;
  fmm_ctx *dctx = nullptr;    // Traversal::direct's own context (the solve's tree stays up)
  real_t theta = 0.4;
  int nspawn = 5000;
  int images = 0;
  vec3 cycle = 2 * M_PI;      // the period the context was made with (images > 0)
  Cells *uploaded = nullptr;  // cells whose tree is on the device
  // traverse(icells, jcells): the device holds the forest of both trees,
  // target cells and bodies first (-1: one tree)
  int pair_cells = -1, pair_bodies = -1;

  Kernel(int _P, real_t _eps2, complex_t _wavek, int device = 0)
      : P(_P), NTERM(_P * (_P + 1) / 2), eps2(_eps2), wavek(_wavek), Xperiodic(0.0), dev(device) {}
  ~Kernel() {
    if (ctx) fmm_destroy(ctx);
    if (dctx) fmm_destroy(dctx);
  }
  Kernel(const Kernel &) = delete;
  Kernel &operator=(const Kernel &) = delete;

  //! (re)create the context with the traversal parameters
  void configure(real_t _theta, int _nspawn, int _images, vec3 _cycle) {
    theta = _theta;
    nspawn = _nspawn;
    images = _images;
    cycle = _cycle;
    if (ctx) fmm_destroy(ctx);
    ctx = nullptr;
    fmm_params p;
    fmm_default_params(&p);
    p.P = P;
    p.eps2 = eps2;
    p.theta = theta;
    p.nspawn = nspawn;
    p.images = images;
    for (int d = 0; d < 3; d++) p.cycle[d] = cycle[d];
    p.device = dev;
    const int rc = fmm_create(&ctx, &p);
    if (rc != FMM_OK) {
      const std::string m = ctx ? fmm_last_error(ctx) : "fmm_create failed";
      if (ctx) fmm_destroy(ctx);
      ctx = nullptr;
      throw Error(m);
    }
  }
  //! drop the device contexts (remade by the next pass)
  void release() {
    if (ctx) fmm_destroy(ctx);
    if (dctx) fmm_destroy(dctx);
    ctx = dctx = nullptr;
    uploaded = nullptr;
    pair_cells = pair_bodies = -1;
  }
  void check(int rc, const char *what, fmm_ctx *on = nullptr) const {
    if (rc != FMM_OK) throw Error(std::string(what) + ": " + fmm_last_error(on ? on : ctx));
  }
  fmm_ctx *context() {
    if (!ctx) configure(theta, nspawn, images, cycle);
    return ctx;
  }

  // ---- kernels/laplace.h:150-388: the per-cell operators on the caller's cells
  //! laplace.h:150-235: TRG of Ci's bodies += field of Cj's bodies shifted by Xperiodic
  void P2P(C_iter Ci, C_iter Cj) {
    const int ni = Ci->NBODY, nj = Cj->NBODY;
    if (ni <= 0 || nj <= 0) return;
    std::vector<double> xi(3 * (size_t)ni), xj(3 * (size_t)nj), qj(nj), trg(4 * (size_t)ni, 0.0);
    for (int i = 0; i < ni; i++)
      for (int d = 0; d < 3; d++) xi[3 * (size_t)i + d] = Ci->BODY[i].X[d];
    for (int j = 0; j < nj; j++) {
      for (int d = 0; d < 3; d++) xj[3 * (size_t)j + d] = Cj->BODY[j].X[d];
      qj[j] = Cj->BODY[j].SRC;
    }
    check(fmm_kernel_p2p(context(), ni, xi.data(), trg.data(), nj, xj.data(), qj.data(), Xperiodic.d),
          "Kernel::P2P");
    for (int i = 0; i < ni; i++)
      for (int k = 0; k < 4; k++) Ci->BODY[i].TRG[k] += trg[4 * (size_t)i + k];
  }
  //! laplace.h:237-252: C->M += multipole of C's bodies about C->X
  void P2M(C_iter C) {
    sized(C->M);
    const int n = C->NBODY;
    std::vector<double> x(3 * (size_t)n), q(n);
    for (int i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++) x[3 * (size_t)i + d] = C->BODY[i].X[d];
      q[i] = C->BODY[i].SRC;
    }
    check(fmm_kernel_p2m(context(), C->X.d, n, x.data(), q.data(), reinterpret_cast<double *>(C->M.data())),
          "Kernel::P2M");
  }
  //! laplace.h:254-290: Ci->M += the multipoles of Ci's children (C0 + ICHILD ...) shifted to Ci
  void M2M(C_iter Ci, C_iter C0) {
    sized(Ci->M);
    const int nc = Ci->NCHILD;
    std::vector<double> X(3 * (size_t)nc);
    std::vector<complex_t> Mc((size_t)nc * NTERM);
    for (int k = 0; k < nc; k++) {
      C_iter Cj = C0 + Ci->ICHILD + k;
      for (int d = 0; d < 3; d++) X[3 * (size_t)k + d] = Cj->X[d];
      for (int u = 0; u < NTERM && u < (int)Cj->M.size(); u++) Mc[(size_t)k * NTERM + u] = Cj->M[u];
    }
    check(fmm_kernel_m2m(context(), Ci->X.d, nc, X.data(), reinterpret_cast<const double *>(Mc.data()),
                         reinterpret_cast<double *>(Ci->M.data())),
          "Kernel::M2M");
  }
  //! laplace.h:292-322: Ci->L += local expansion of Cj->M (source shifted by Xperiodic)
  void M2L(C_iter Ci, C_iter Cj) {
    sized(Ci->L);
    std::vector<complex_t> Mj(NTERM);
    for (int u = 0; u < NTERM && u < (int)Cj->M.size(); u++) Mj[u] = Cj->M[u];
    check(fmm_kernel_m2l(context(), 1, Ci->X.d, Cj->X.d, Xperiodic.d, reinterpret_cast<const double *>(Mj.data()),
                         reinterpret_cast<double *>(Ci->L.data())),
          "Kernel::M2L");
  }
  //! laplace.h:324-357: Ci->L += the local expansion of its parent (C0 + IPARENT) shifted to Ci
  void L2L(C_iter Ci, C_iter C0) {
    sized(Ci->L);
    C_iter Cp = C0 + Ci->IPARENT;
    std::vector<complex_t> Lp(NTERM);
    for (int u = 0; u < NTERM && u < (int)Cp->L.size(); u++) Lp[u] = Cp->L[u];
    check(fmm_kernel_l2l(context(), Ci->X.d, Cp->X.d, reinterpret_cast<const double *>(Lp.data()),
                         reinterpret_cast<double *>(Ci->L.data())),
          "Kernel::L2L");
  }
  //! laplace.h:359-388: TRG of Ci's bodies += Ci->L evaluated at them
  void L2P(C_iter Ci) {
    const int n = Ci->NBODY;
    if (n <= 0) return;
    std::vector<complex_t> L(NTERM);
    for (int u = 0; u < NTERM && u < (int)Ci->L.size(); u++) L[u] = Ci->L[u];
    std::vector<double> x(3 * (size_t)n), trg(4 * (size_t)n, 0.0);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) x[3 * (size_t)i + d] = Ci->BODY[i].X[d];
    check(fmm_kernel_l2p(context(), Ci->X.d, reinterpret_cast<const double *>(L.data()), n, x.data(), trg.data()),
          "Kernel::L2P");
    for (int i = 0; i < n; i++)
      for (int k = 0; k < 4; k++) Ci->BODY[i].TRG[k] += trg[4 * (size_t)i + k];
  }
  //! upload the tree held by cells (bodies in tree order, as buildTree leaves them)
  void upload(Cells &cells) {
    if (!ctx) configure(theta, nspawn, images, cycle);
    const int nc = (int)cells.size();
    const int n = nc ? cells[0].NBODY : 0;
    B_iter B0 = nc ? cells[0].BODY : B_iter();
    std::vector<double> x(3 * (size_t)n), q(n), X(3 * (size_t)nc), R(nc);
    std::vector<int> ip(nc), ic(nc), nch(nc), ib(nc), nb(nc);
    for (int i = 0; i < n; i++) {
      for (int d = 0; d < 3; d++) x[3 * (size_t)i + d] = B0[i].X[d];
      q[i] = B0[i].SRC;
    }
    for (int c = 0; c < nc; c++) {
      ip[c] = cells[c].IPARENT;
      ic[c] = cells[c].ICHILD;
      nch[c] = cells[c].NCHILD;
      ib[c] = cells[c].IBODY;
      nb[c] = cells[c].NBODY;
      for (int d = 0; d < 3; d++) X[3 * (size_t)c + d] = cells[c].X[d];
      R[c] = cells[c].R;
    }
    check(fmm_set_tree(ctx, n, x.data(), q.data(), nc, ip.data(), ic.data(), nch.data(), ib.data(),
                       nb.data(), X.data(), R.data()),
          "fmm_set_tree");
    uploaded = &cells;
    pair_cells = pair_bodies = -1;
  }
  //! the context for direct sums, made with the same parameters
  fmm_ctx *direct_context() {
    if (!dctx) {
      fmm_params p;
      fmm_default_params(&p);
      p.P = P;
      p.eps2 = eps2;
      p.theta = theta;
      p.nspawn = nspawn;
      p.images = images;
      for (int d = 0; d < 3; d++) p.cycle[d] = cycle[d];
      p.device = dev;
      const int rc = fmm_create(&dctx, &p);
      if (rc != FMM_OK) {
        const std::string m = dctx ? fmm_last_error(dctx) : "fmm_create failed";
        if (dctx) fmm_destroy(dctx);
        dctx = nullptr;
        throw Error(m);
      }
    }
    return dctx;
  }

 private:
  int dev;
  void sized(std::vector<complex_t> &E) const {
    if ((int)E.size() < NTERM) E.resize(NTERM, 0.0);
  }
};

//! include/up_down_pass.h:46-99
class UpDownPass {
  Kernel &kernel;

 public:
  explicit UpDownPass(Kernel &_kernel) : kernel(_kernel) {}
  void upwardPass(Cells &cells) {
    logger::startTimer("Upward pass");
    if (!cells.empty()) {
      kernel.upload(cells);
      kernel.check(fmm_upward(kernel.ctx), "fmm_upward");
      std::vector<complex_t> M(cells.size() * (size_t)kernel.NTERM);
      kernel.check(fmm_get_expansions(kernel.ctx, reinterpret_cast<double *>(M.data()), nullptr),
                   "fmm_get_expansions");
      for (size_t c = 0; c < cells.size(); c++) {
        cells[c].M.assign(M.begin() + c * kernel.NTERM, M.begin() + (c + 1) * kernel.NTERM);
        cells[c].L.assign(kernel.NTERM, 0.0);
      }
    }
    logger::stopTimer("Upward pass");
  }
  void downwardPass(Cells &cells) {
    logger::startTimer("Downward pass");
    if (!cells.empty()) {
      if (kernel.uploaded != &cells) throw Error("downwardPass: cells differ from the upward pass");
      kernel.check(fmm_downward(kernel.ctx), "fmm_downward");
      // the device may hold a forest (traverse(icells, jcells)): these cells
      // and their bodies come first
      fmm_stats st;
      fmm_get_stats(kernel.ctx, &st);
      const int n = cells[0].NBODY;
      std::vector<double> trg(4 * (size_t)st.n);
      kernel.check(fmm_get_targets(kernel.ctx, trg.data(), 0), "fmm_get_targets");
      B_iter B0 = cells[0].BODY;
      for (int i = 0; i < n; i++)
        for (int k = 0; k < 4; k++) B0[i].TRG[k] += trg[4 * (size_t)i + k];
      std::vector<complex_t> L((size_t)st.ncell * kernel.NTERM);
      kernel.check(fmm_get_expansions(kernel.ctx, nullptr, reinterpret_cast<double *>(L.data())),
                   "fmm_get_expansions");
      for (size_t c = 0; c < cells.size(); c++)
        cells[c].L.assign(L.begin() + c * kernel.NTERM, L.begin() + (c + 1) * kernel.NTERM);
    }
    logger::stopTimer("Downward pass");
  }
  //! include/up_down_pass.h:77-86: sum over bodies of (X - X0) q
  vec3 getDipole(Bodies &bodies, vec3 X0) {
    vec3 dipole = 0.0;
    for (B_iter B = bodies.begin(); B != bodies.end(); B++)
      for (int d = 0; d < 3; d++) dipole[d] += (B->X[d] - X0[d]) * B->SRC;
    return dipole;
  }
  //! include/up_down_pass.h:89-99: TRG -= 4 pi / (3 V) (|D|^2 / (N q), D)
  void dipoleCorrection(Bodies &bodies, vec3 dipole, int numBodies, vec3 cycle) {
    const real_t coef = 4 * M_PI / (3 * cycle[0] * cycle[1] * cycle[2]);
    const real_t dd = dipole[0] * dipole[0] + dipole[1] * dipole[1] + dipole[2] * dipole[2];
    for (B_iter B = bodies.begin(); B != bodies.end(); B++) {
      B->TRG[0] -= coef * dd / numBodies / B->SRC;
      for (int d = 0; d < 3; d++) B->TRG[d + 1] -= coef * dipole[d];
    }
  }
};

//! include/traversal.h:291-377
class Traversal {
  Kernel &kernel;
  const real_t theta;
  const int nspawn;
  const int images;
  const std::string path;

 public:
  Traversal(Kernel &_kernel, real_t _theta, int _nspawn, int _images, const char *_path = "./")
      : kernel(_kernel), theta(_theta), nspawn(_nspawn), images(_images), path(_path ? _path : "./") {
    // the device context is made on first use with these parameters
    kernel.theta = theta;
    kernel.nspawn = nspawn;
    kernel.images = images;
    kernel.release();
  }
  //! images > 0 and a period other than the context's: remake the context
  //! (the lists depend on the period) and, if a tree was up, re-run its upward pass
  void use_cycle(vec3 cycle) {
    if (images == 0 || (cycle[0] == kernel.cycle[0] && cycle[1] == kernel.cycle[1] && cycle[2] == kernel.cycle[2]))
      return;
    Cells *up = kernel.uploaded;
    kernel.configure(theta, nspawn, images, cycle);
    kernel.uploaded = nullptr;
    if (up) {
      kernel.upload(*up);
      kernel.check(fmm_upward(kernel.ctx), "fmm_upward");
    }
  }
  //! include/traversal.h:298-306
  void initListCount(Cells &cells) {
    for (Cell &C : cells) C.numP2P = C.numM2L = 0;
  }
  void initWeight(Cells &) {}
  void traverse(Cells &icells, Cells &jcells, vec3 cycle, bool, real_t = 1) {
    if (icells.empty() || jcells.empty()) return;
    if (&icells != &jcells) {
      traverse_pair(icells, jcells);
      return;
    }
    logger::startTimer("Traverse");
    use_cycle(cycle);
    if (kernel.uploaded != &icells || kernel.pair_cells >= 0) kernel.upload(icells);
    kernel.check(fmm_p2p(kernel.ctx), "fmm_p2p");
    kernel.check(fmm_m2l(kernel.ctx), "fmm_m2l");
    kernel.check(fmm_synchronize(kernel.ctx), "fmm_synchronize");
    // countList (include/traversal.h:30-38): the list sizes of every target cell
    std::vector<int> np2p(icells.size()), nm2l(icells.size());
    kernel.check(fmm_get_list_counts(kernel.ctx, nullptr, np2p.data(), nm2l.data()), "fmm_get_list_counts");
    for (size_t c = 0; c < icells.size(); c++) {
      icells[c].numP2P += np2p[c];
      icells[c].numM2L += nm2l[c];
    }
    logger::stopTimer("Traverse");
  }
  //! Traversal::traverse(icells, jcells) with two trees (traversal.h:326-333,
  //! examples/fmm.cxx:69-72): the device gets the forest of both trees and the
  //! pair traversal's lists; M of both (the upward pass again, on the device),
  //! then P2P and M2L into icells.  downwardPass(icells) finishes the solve.
  void traverse_pair(Cells &icells, Cells &jcells) {
    if (images != 0) throw Error("traverse: separate source trees need images = 0");
    logger::startTimer("Traverse");
    fmm_plan *pi = tree_plan(icells), *pj = tree_plan(jcells), *pp = nullptr;
    const int rc = fmm_plan_pair(&pp, pi, pj, theta, nspawn);
    fmm_plan_destroy(pi);
    fmm_plan_destroy(pj);
    if (rc != FMM_OK) throw Error("traverse: fmm_plan_pair failed");
    const int ni = icells[0].NBODY, nj = jcells[0].NBODY;
    std::vector<double> x(3 * ((size_t)ni + nj)), q((size_t)ni + nj);
    auto put = [&](Cells &cells, int n, size_t off) {
      B_iter B0 = cells[0].BODY;
      for (int i = 0; i < n; i++) {
        for (int d = 0; d < 3; d++) x[3 * (off + i) + d] = B0[i].X[d];
        q[off + i] = B0[i].SRC;
      }
    };
    put(icells, ni, 0);
    put(jcells, nj, ni);
    fmm_ctx *c = kernel.context();
    const int r1 = fmm_set_bodies_plan(c, pp, x.data(), q.data());
    fmm_plan_destroy(pp);
    kernel.check(r1, "fmm_set_bodies_plan");
    kernel.uploaded = &icells;
    kernel.pair_cells = (int)icells.size();
    kernel.pair_bodies = ni;
    kernel.check(fmm_upward(c), "fmm_upward");
    kernel.check(fmm_p2p(c), "fmm_p2p");
    kernel.check(fmm_m2l(c), "fmm_m2l");
    kernel.check(fmm_synchronize(c), "fmm_synchronize");
    fmm_stats st;
    fmm_get_stats(c, &st);
    std::vector<int> np2p(st.ncell), nm2l(st.ncell);
    kernel.check(fmm_get_list_counts(c, nullptr, np2p.data(), nm2l.data()), "fmm_get_list_counts");
    for (size_t k = 0; k < icells.size(); k++) {
      icells[k].numP2P += np2p[k];
      icells[k].numM2L += nm2l[k];
    }
    logger::stopTimer("Traverse");
  }
  //! a caller's tree (bodies in tree order) as a host plan without lists
  static fmm_plan *tree_plan(Cells &cells) {
    const int nc = (int)cells.size(), n = cells[0].NBODY;
    B_iter B0 = cells[0].BODY;
    std::vector<double> x(3 * (size_t)n), X(3 * (size_t)nc), R(nc);
    std::vector<int> ip(nc), ic(nc), nch(nc), ib(nc), nb(nc);
    for (int i = 0; i < n; i++)
      for (int d = 0; d < 3; d++) x[3 * (size_t)i + d] = B0[i].X[d];
    for (int k = 0; k < nc; k++) {
      ip[k] = cells[k].IPARENT;
      ic[k] = cells[k].ICHILD;
      nch[k] = cells[k].NCHILD;
      ib[k] = cells[k].IBODY;
      nb[k] = cells[k].NBODY;
      for (int d = 0; d < 3; d++) X[3 * (size_t)k + d] = cells[k].X[d];
      R[k] = cells[k].R;
    }
    fmm_plan *pl = nullptr;
    char why[256] = {0};
    // the lists of this plan are not used (the pair traversal replaces them):
    // a theta that accepts every separated pair keeps them short
    if (fmm_plan_from_tree(&pl, n, x.data(), nc, ip.data(), ic.data(), nch.data(), ib.data(), nb.data(), X.data(),
                           R.data(), 1e100, 1 << 30, why, (int)sizeof why) != FMM_OK)
      throw Error(std::string("traverse: tree rejected: ") + why);
    return pl;
  }
  //! direct sum of jbodies (and, with images > 0, of their copies shifted by
  //! (ix, iy, iz) * cycle) at the positions of ibodies (TRG += ), on the
  //! device, in a context of its own (the solve's tree stays uploaded)
  void direct(Bodies &ibodies, Bodies &jbodies, vec3 cycle) {
    const int nj = (int)jbodies.size(), ni = (int)ibodies.size();
    if (!kernel.ctx) kernel.configure(theta, nspawn, images, images ? cycle : kernel.cycle);
    use_cycle(cycle);
    std::vector<double> xj(3 * (size_t)nj), qj(nj), xi(3 * (size_t)ni), trg(4 * (size_t)ni);
    for (int j = 0; j < nj; j++) {
      for (int d = 0; d < 3; d++) xj[3 * (size_t)j + d] = jbodies[j].X[d];
      qj[j] = jbodies[j].SRC;
    }
    for (int i = 0; i < ni; i++)
      for (int d = 0; d < 3; d++) xi[3 * (size_t)i + d] = ibodies[i].X[d];
    // the direct context's bodies are jbodies (a one-leaf tree): only the direct kernel runs
    fmm_ctx *c = kernel.direct_context();
    const int ip = 0, ic = 0, nch = 0, ib = 0;
    double X[3] = {0, 0, 0}, R = 1.0;
    kernel.check(fmm_set_tree(c, nj, xj.data(), qj.data(), nj ? 1 : 0, &ip, &ic, &nch, &ib, &nj, X, &R),
                 "fmm_set_tree", c);
    kernel.check(fmm_direct(c, ni, xi.data(), trg.data()), "fmm_direct", c);
    for (int i = 0; i < ni; i++)
      for (int k = 0; k < 4; k++) ibodies[i].TRG[k] += trg[4 * (size_t)i + k];
  }
  void printTraversalData() {
    if (!logger::verbose || !kernel.ctx) return;
    fmm_stats s;
    fmm_get_stats(kernel.ctx, &s);
    std::cout << "--- Traversal stats --------------" << std::endl
              << std::setw(logger::stringLength) << std::left << "P2P calls"
              << " : " << s.p2p_pairs << std::endl
              << std::setw(logger::stringLength) << std::left << "M2L calls"
              << " : " << s.m2l_pairs << std::endl;
  }
  //! include/traversal.h:412-421: "ICELL numP2P numM2L" per cell into <path>list<rank, 6 digits>.dat
  void writeList(Cells cells, int mpirank) {
    std::ostringstream name;
    name << path << "list" << std::setfill('0') << std::setw(6) << mpirank << ".dat";
    std::ofstream f(name.str().c_str());
    for (const Cell &C : cells)
      f << std::setw(logger::stringLength) << std::left << C.ICELL << " " << C.numP2P << " " << C.numM2L
        << std::endl;
  }
};

//! include/dataset.h: seeded drand48 generators in the reference's draw order
class Dataset {
 public:
  Bodies initBodies(int numBodies, const char *distribution, int mpirank = 0, int = 1, int = 1) {
    std::vector<double> x(3 * (size_t)numBodies), q(numBodies);
    if (fmm_make_bodies(numBodies, distribution, mpirank, x.data(), q.data()) != FMM_OK)
      throw Error(std::string("Unknown data distribution ") + distribution);
    Bodies bodies(numBodies);
    for (int i = 0; i < numBodies; i++) {
      for (int d = 0; d < 3; d++) bodies[i].X[d] = x[3 * (size_t)i + d];
      bodies[i].SRC = q[i];
      bodies[i].IRANK = mpirank;
    }
    initTarget(bodies);
    return bodies;
  }
  void initTarget(Bodies &bodies) {
    for (B_iter B = bodies.begin(); B != bodies.end(); B++) {
      B->TRG = 0.0;
      B->IBODY = (int)(B - bodies.begin());
      B->ICELL = 0;
      B->WEIGHT = 1;
    }
  }
  void sampleBodies(Bodies &bodies, int numTargets) {
    if (numTargets < (int)bodies.size()) {
      const int stride = (int)bodies.size() / numTargets;
      for (int i = 0; i < numTargets; i++) bodies[i] = bodies[i * stride];
      bodies.resize(numTargets);
    }
  }
};

//! include/verify.h:42-92, 97-106
class Verify {
 public:
  bool verbose = false;
  double getNrmScalar(Bodies &b) {
    double v = 0;
    for (B_iter B = b.begin(); B != b.end(); B++) v += std::abs(B->TRG[0] * B->TRG[0]);
    return v;
  }
  double getDifScalar(Bodies &b, Bodies &b2) {
    double v = 0;
    for (size_t i = 0; i < b.size(); i++) v += std::abs((b[i].TRG[0] - b2[i].TRG[0]) * (b[i].TRG[0] - b2[i].TRG[0]));
    return v;
  }
  double getNrmVector(Bodies &b) {
    double v = 0;
    for (B_iter B = b.begin(); B != b.end(); B++)
      v += std::abs(B->TRG[1] * B->TRG[1] + B->TRG[2] * B->TRG[2] + B->TRG[3] * B->TRG[3]);
    return v;
  }
  double getDifVector(Bodies &b, Bodies &b2) {
    double v = 0;
    for (size_t i = 0; i < b.size(); i++) {
      double s = 0;
      for (int d = 1; d < 4; d++) s += (b[i].TRG[d] - b2[i].TRG[d]) * (b[i].TRG[d] - b2[i].TRG[d]);
      v += std::abs(s);
    }
    return v;
  }
  void print(const std::string &title, double v) {
    if (verbose)
      std::cout << std::setw(logger::stringLength) << std::left << title << " : "
                << std::setprecision(logger::decimal) << std::scientific << v << std::endl;
  }
};

//! include/args.h:15-35, 188-274: same options and defaults
class Args {
 public:
  int accuracy = 0, ncrit = 64, dual = 0, graft = 0, getMatrix = 0, images = 0, IneJ = 0;
  int numBodies = 1000000, P = 10, repeat = 1, nspawn = 5000, threads = 16, verbose = 0, write = 0;
  double cutoff = 0.0, theta = 0.4;
  const char *distribution = "cube";
  const char *path = "./";

  Args(int argc = 0, char **argv = nullptr) {
    static struct option opts[] = {
        {"accuracy", no_argument, 0, 'a'},     {"ncrit", required_argument, 0, 'c'},
        {"cutoff", required_argument, 0, 'C'}, {"distribution", required_argument, 0, 'd'},
        {"dual", no_argument, 0, 'D'},         {"graft", no_argument, 0, 'g'},
        {"getMatrix", no_argument, 0, 'G'},    {"help", no_argument, 0, 'h'},
        {"images", required_argument, 0, 'i'}, {"IneJ", no_argument, 0, 'j'},
        {"numBodies", required_argument, 0, 'n'}, {"path", required_argument, 0, 'p'},
        {"P", required_argument, 0, 'P'},      {"repeat", required_argument, 0, 'r'},
        {"nspawn", required_argument, 0, 's'}, {"theta", required_argument, 0, 't'},
        {"threads", required_argument, 0, 'T'}, {"verbose", no_argument, 0, 'v'},
        {"write", no_argument, 0, 'w'},        {0, 0, 0, 0}};
    optind = 1;
    while (argc > 0) {
      int idx;
      const int c = getopt_long(argc, argv, "ac:C:d:DgGhi:jn:p:P:r:s:t:T:vw", opts, &idx);
      if (c == -1) break;
      switch (c) {
        case 'a': accuracy = 1; break;
        case 'c': ncrit = atoi(optarg); break;
        case 'C': cutoff = atof(optarg); break;
        case 'd': distribution = parseDistribution(optarg); break;
        case 'D': dual = 1; break;
        case 'g': graft = 1; break;
        case 'G': getMatrix = 1; break;
        case 'i': images = atoi(optarg); break;
        case 'j': IneJ = 1; break;
        case 'n': numBodies = atoi(optarg); break;
        case 'p': path = optarg; break;
        case 'P': P = atoi(optarg); break;
        case 'r': repeat = atoi(optarg); break;
        case 's': nspawn = atoi(optarg); break;
        case 't': theta = atof(optarg); break;
        case 'T': threads = atoi(optarg); break;
        case 'v': verbose = 1; break;
        case 'w': write = 1; break;
        default: usage(argv[0]); throw Error("bad command line");
      }
    }
  }
  static const char *parseDistribution(const char *arg) {
    switch (arg[0]) {
      case 'c': return arg[1] == '1' ? "cube1" : "cube";
      case 'p': return arg[strlen_c(arg) - 1] == '1' ? "plummer1" : "plummer";
      case 's': return "sphere";
      case 'o': return "octant";
      default: throw Error(std::string("invalid distribution ") + arg);
    }
  }
  static size_t strlen_c(const char *s) {
    size_t n = 0;
    while (s[n]) n++;
    return n;
  }
  static void usage(const char *name) {
    std::fprintf(stderr,
                 "Usage: %s [options]\n"
                 " -a accuracy only  -c ncrit (64)  -d cube|sphere|octant|plummer|cube1|plummer1\n"
                 " -n numBodies (1000000)  -P order (10)  -r repeat (1)  -s nspawn (5000)\n"
                 " -t theta (0.4)  -v verbose\n",
                 name);
  }
  void print() const {
    if (!verbose) return;
    const int w = logger::stringLength;
    std::cout << std::setw(w) << std::left << "numBodies" << " : " << numBodies << std::endl
              << std::setw(w) << std::left << "ncrit" << " : " << ncrit << std::endl
              << std::setw(w) << std::left << "distribution" << " : " << distribution << std::endl
              << std::setw(w) << std::left << "P" << " : " << P << std::endl
              << std::setw(w) << std::left << "theta" << " : " << theta << std::endl
              << std::setw(w) << std::left << "nspawn" << " : " << nspawn << std::endl;
  }
};

}  // namespace exafmm_amd

// fmm_capi.cpp -- implementation of the C ABI in include/exafmm_amd.h.
#include "exafmm_amd.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

#include "fmm_device.hpp"
#include "fmm_host.hpp"

namespace {

// Grow-only device buffer: the arena stays resident across calls so repeated
// solves (successive MD steps, served requests) do not reallocate.
// Owns its allocation: freed with the context (the owner makes its device
// current first); moves transfer it, copies are not allowed.
struct DBuf {
  void *p = nullptr;
  size_t cap = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), cap(o.cap) {
    o.p = nullptr;
    o.cap = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    std::swap(p, o.p);
    std::swap(cap, o.cap);
    return *this;
  }
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  template <typename T> T *as() const { return static_cast<T *>(p); }
};

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

std::mutex g_init_mu;
bool g_init_done[64] = {};

}  // namespace

struct fmm_ctx {
  fmm_params p{};
  std::string err;
  hipStream_t st = nullptr;
  hipEvent_t ev[5] = {};
  hipEvent_t kev[4] = {};  // M2L kernel begin/end, P2P kernel begin/end
  // fmm_evaluate runs P2P on a second stream (default priority), beside the
  // upward pass and M2L: fork after the previous work, join before L2P
  hipStream_t st2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  bool timing = false;     // set by fmm_evaluate
  // fmm_evaluate_graph: the solve captured from both streams, replayed until
  // the bodies, tree or any device buffer of the solve change
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  fmm::Tree tree;
  fmm::Lists lists;
  int n = 0;
  bool have_bodies = false;
  fmm_stats stats{};
  // host-side level tables
  std::vector<int> internal_ptr;  // per level offsets into internal cell list
  std::vector<int> level_ptr;     // per level offsets into level_cells
  // device buffers
  DBuf x, y, z, q, pot, ax, ay, az;
  DBuf xin, qin, perm;  // upload staging: bodies as given, the tree's permutation
  DBuf cx, cy, cz, ibody, nbody, ichild, nchild, iparent;
  DBuf M, L;
  DBuf leaves, internal, level_cells, body_leaf;
  DBuf m2l_ptr, m2l_src, m2l_tgt, p2p_ptr, p2p_src, p2p_tgt;
  // offset-grouped M2L: chunks (GROUPED) or target tiles (AUTO)
  bool use_mfma = false, use_tiles = false;
  int nchunk = 0, ntile = 0;
  int p2p_nwide = 0;  // P2P targets [0, p2p_nwide) are leaves of >= 16 bodies
  fmm::M2LGroups groups;
  DBuf g_tgt, g_src, g_dx, g_chunk, r_ptr, r_src, r_tgt;
  DBuf t_tgt, t_ptr, t_slice, t_src, t_lt, g_y, g_ops, c_dx, c_sc;
  // M2M / L2L as per-level operators on MFMA (regular trees)
  bool xl = false;
  std::vector<int> xl_tptr;  // per child level: first L2L tile
  DBuf xl_child8, xl_children, xl_tiles, xl_ops, xl_frag, xl_dx, xl_kinds;
  // periodic images: shift unit per image, far-image operator F (images >= 2)
  double3 cyc = {0.0, 0.0, 0.0};
  bool far = false;
  DBuf per_F, per_T, per_S, per_cur, per_tmp, per_dx, per_kind;
  // mixed precision: per-group (R_s, R_t), cell radii, FP32 multipoles
  bool mixed = false;
  bool h16 = false;  // FMM_MIXED16: split-FP16 M2L operands (M16 instead of M32)
  DBuf g_sc, cr, M32, M16;
  DBuf dip;  // dipole partial sums
  DBuf ew_b, ew_t, ew_k, ew_s;  // fmm_ewald: bodies, targets, wave vectors, structure factors
  DBuf op_d, op_i;  // fmm_kernel_*: doubles (bodies, centres, expansions) and ints of one call
  // pinned host staging of the upload's large copies (grow-only, hipHostMalloc)
  void *pin = nullptr;
  size_t pin_cap = 0;
};

struct fmm_plan {
  int n = 0;
  fmm::Tree tree;
  fmm::Lists lists;
  fmm::M2LGroups groups;  // prebuilt M2L groups (fmm_plan_build_groups), kind:
  int groups_kind = -1;   // -1 none, FMM_FP64 offset groups, FMM_MIXED by level
  int images = 0;          // periodic lists (images > 0) and their cycle
  double cycle[3] = {0, 0, 0};
  fmm::M2LTiles tiles;    // prebuilt target tiles of those groups (TT > 0)
  double xmin[3] = {0, 0, 0}, xmax[3] = {0, 0, 0};
  double t_bounds = 0, t_tree = 0, t_lists = 0, t_groups = 0;
  int pair_targets = -1;  // fmm_plan_pair: bodies [0, pair_targets) are the targets
  // fmm_plan_build_device: the lists' source columns and the tiles' pair
  // arrays left on the device (copies of the plan share them); the host
  // vectors of lists and tiles stay empty until a host-side use fetches them
  std::shared_ptr<fmmdev::PlanArrays> dev;
};

#define FMM_FAIL(code, msg) \
  do {                      \
    c->err = (msg);         \
    return (code);          \
  } while (0)
#define HIPCHK(expr)                                                               \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) {                                                        \
      c->err = std::string(#expr) + " failed: " + hipGetErrorString(e_);           \
      return e_ == hipErrorOutOfMemory ? FMM_ERR_MEMORY : FMM_ERR_DEVICE;          \
    }                                                                              \
  } while (0)

namespace {

int ensure(fmm_ctx *c, DBuf &b, size_t bytes) {
  if (bytes <= b.cap && b.p) return FMM_OK;
  if (b.p) {
    HIPCHK(hipFree(b.p));
    b.p = nullptr;
    b.cap = 0;
  }
  // 25% headroom: a stream of slightly growing systems reallocates rarely
  // (hipMalloc/hipFree synchronise the device)
  size_t sz = bytes ? bytes + bytes / 4 : 16;
  HIPCHK(hipMalloc(&b.p, sz));
  b.cap = sz;
  return FMM_OK;
}

template <typename T>
int upload(fmm_ctx *c, DBuf &b, const T *h, size_t count) {
  int rc = ensure(c, b, count * sizeof(T));
  if (rc) return rc;
  if (count) HIPCHK(hipMemcpyAsync(b.p, h, count * sizeof(T), hipMemcpyHostToDevice, c->st));
  return FMM_OK;
}

// a device-held plan array into a context buffer (same device)
int dcopy(fmm_ctx *c, DBuf &b, const fmmdev::DevArray &a) {
  int rc = ensure(c, b, a.bytes);
  if (rc) return rc;
  if (a.bytes) HIPCHK(hipMemcpyAsync(b.p, a.p, a.bytes, hipMemcpyDeviceToDevice, c->st));
  return FMM_OK;
}

// host arrays (pointer, bytes) to device buffers through the context's pinned
// staging: the arrays are packed into it by host threads (large sizes), then
// copied by DMA from pinned memory, which runs several times the rate of a
// pageable copy.  The staging is reused by the next call, which first waits
// for the stream (the copies of the previous call read it).
int upload_staged(fmm_ctx *c, std::initializer_list<std::tuple<DBuf *, const void *, size_t>> parts) {
  HIPCHK(hipStreamSynchronize(c->st));
  size_t total = 0;
  for (const auto &pt : parts) total += (std::get<2>(pt) + 255) / 256 * 256;
  if (total > c->pin_cap) {
    if (c->pin) HIPCHK(hipHostFree(c->pin));
    c->pin = nullptr;
    c->pin_cap = 0;
    HIPCHK(hipHostMalloc(&c->pin, total + total / 4, hipHostMallocDefault));
    c->pin_cap = total + total / 4;
  }
  // pack: runs of about 16 MB per host thread
  std::vector<std::tuple<char *, const char *, size_t>> runs;
  size_t off = 0;
  for (const auto &pt : parts) {
    const size_t b = std::get<2>(pt);
    const char *src = static_cast<const char *>(std::get<1>(pt));
    for (size_t o = 0; o < b; o += (size_t)16 << 20)
      runs.emplace_back(static_cast<char *>(c->pin) + off + o, src + o, std::min(b - o, (size_t)16 << 20));
    off += (b + 255) / 256 * 256;
  }
  const int nth = (int)std::min<size_t>(runs.size(), 8);
  if (nth > 1) {
    std::vector<std::thread> th;
    for (int t = 0; t < nth; t++)
      th.emplace_back([&, t] {
        for (size_t r = t; r < runs.size(); r += nth) std::memcpy(std::get<0>(runs[r]), std::get<1>(runs[r]), std::get<2>(runs[r]));
      });
    for (auto &h : th) h.join();
  } else {
    for (auto &r : runs) std::memcpy(std::get<0>(r), std::get<1>(r), std::get<2>(r));
  }
  off = 0;
  for (const auto &pt : parts) {
    const size_t b = std::get<2>(pt);
    int rc = ensure(c, *std::get<0>(pt), b);
    if (rc) return rc;
    if (b) HIPCHK(hipMemcpyAsync(std::get<0>(pt)->p, static_cast<char *>(c->pin) + off, b, hipMemcpyHostToDevice, c->st));
    off += (b + 255) / 256 * 256;
  }
  return FMM_OK;
}

// a device array taken over by a context buffer (both hipMalloc'd)
void adopt(DBuf &b, fmmdev::DevArray &a) {
  if (b.p) (void)hipFree(b.p);
  b.p = a.p;
  b.cap = a.bytes;
  a.p = nullptr;
  a.bytes = 0;
  a.device = -1;
}

// the host copies of a device-built plan's lists and tiles, for host-side
// uses (merging, host grouping, list export, another device); the plan's
// contents do not change, only where they are held
hipError_t plan_to_host(const fmm_plan *pl) {
  if (!pl->dev || !(pl->dev->lists || pl->dev->tiles)) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != pl->dev->device && (e = hipSetDevice(pl->dev->device)) != hipSuccess) return e;
  fmm_plan *m = const_cast<fmm_plan *>(pl);
  e = fmmdev::fetch_lists(*pl->dev, m->lists, nullptr);
  if (e == hipSuccess && m->tiles.TT > 0) e = fmmdev::fetch_tiles(*pl->dev, m->tiles, nullptr);
  if (cur != pl->dev->device) (void)hipSetDevice(cur);
  return e;
}

void release(DBuf &b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

// drop the captured solve (its nodes point at the device buffers as they were)
void graph_reset(fmm_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
}

fmmdev::DevBodies dev_bodies(fmm_ctx *c) {
  return {c->x.as<double>(), c->y.as<double>(), c->z.as<double>(), c->q.as<double>(),
          c->pot.as<double>(), c->ax.as<double>(), c->ay.as<double>(), c->az.as<double>()};
}
fmmdev::DevCells dev_cells(fmm_ctx *c) {
  return {c->cx.as<double>(), c->cy.as<double>(),    c->cz.as<double>(),   c->ibody.as<int>(),
          c->nbody.as<int>(), c->ichild.as<int>(), c->nchild.as<int>(), c->iparent.as<int>()};
}

int nterm(const fmm_ctx *c) { return c->p.P * (c->p.P + 1) / 2; }

double s_nm(int n, int m) {  // sqrt((n-m)!(n+m)!)
  double f = 1.0;
  for (int i = 2; i <= n - m; i++) f *= i;
  double g = 1.0;
  for (int i = 2; i <= n + m; i++) g *= i;
  return std::sqrt(f) * std::sqrt(g);
}

int check_ready(fmm_ctx *c) {
  if (!c) return FMM_ERR_ARG;
  if (!c->have_bodies) FMM_FAIL(FMM_ERR_STATE, "no bodies: call fmm_set_bodies first");
  HIPCHK(hipSetDevice(c->p.device));
  return FMM_OK;
}

}  // namespace

extern "C" {

void fmm_default_params(fmm_params *p) {
  if (!p) return;
  p->P = 10;
  p->ncrit = 64;
  p->nspawn = 5000;
  p->images = 0;
  p->theta = 0.4;
  p->eps2 = 0.0;
  p->precision = FMM_FP64;
  p->device = 0;
  p->m2l_mode = FMM_M2L_AUTO;
  for (int d = 0; d < 3; d++) p->cycle[d] = 2 * M_PI;  // examples/fmm.cxx:14
}

int fmm_create(fmm_ctx **out, const fmm_params *p) {
  if (!out || !p) return FMM_ERR_ARG;
  *out = nullptr;
  fmm_ctx *c = new (std::nothrow) fmm_ctx();
  if (!c) return FMM_ERR_MEMORY;
  c->p = *p;
  int rc = FMM_OK;
  if (!fmmdev::order_supported(p->P)) {
    c->err = "unsupported expansion order P=" + std::to_string(p->P) + " (2..20)";
    rc = FMM_ERR_UNSUPPORTED;
  } else if (p->ncrit < 1 || p->theta <= 0 || p->nspawn < 1) {
    c->err = "invalid ncrit/theta/nspawn";
    rc = FMM_ERR_ARG;
  } else if (p->images < 0 || (p->images > 0 && !(p->cycle[0] > 0 && p->cycle[1] > 0 && p->cycle[2] > 0))) {
    c->err = "images must be >= 0, and the cycle > 0 in x, y and z when images > 0";
    rc = FMM_ERR_ARG;
  } else if (p->m2l_mode != FMM_M2L_AUTO && p->m2l_mode != FMM_M2L_TARGET && p->m2l_mode != FMM_M2L_GROUPED &&
             p->m2l_mode != FMM_M2L_TILES) {
    c->err = "invalid m2l_mode";
    rc = FMM_ERR_ARG;
  } else if (p->precision != FMM_FP64 && p->precision != FMM_MIXED && p->precision != FMM_MIXED16) {
    c->err = "unsupported precision mode";
    rc = FMM_ERR_UNSUPPORTED;
  } else if (p->precision == FMM_MIXED16 &&
             (!fmmdev::m2l_mixed_supported(p->P) || p->m2l_mode == FMM_M2L_TARGET || p->m2l_mode == FMM_M2L_GROUPED)) {
    c->err = "mixed precision with split-FP16 M2L needs 2 <= P <= 12 and m2l_mode AUTO or TILES";
    rc = FMM_ERR_UNSUPPORTED;
  } else if (p->precision == FMM_MIXED &&
             (!fmmdev::m2l_mixed_supported(p->P) || p->m2l_mode == FMM_M2L_TARGET)) {
    c->err = "mixed precision needs 2 <= P <= 12 and m2l_mode AUTO or GROUPED";
    rc = FMM_ERR_UNSUPPORTED;
  }
  if (rc == FMM_OK) {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev <= 0) {
      c->err = std::string("no HIP device: ") + hipGetErrorString(e);
      rc = FMM_ERR_DEVICE;
    } else if (p->device < 0 || p->device >= ndev || p->device >= 64) {
      c->err = "device ordinal out of range";
      rc = FMM_ERR_ARG;
    }
  }
  if (rc == FMM_OK) {
    rc = [&]() -> int {
      HIPCHK(hipSetDevice(p->device));
      {
        std::lock_guard<std::mutex> lk(g_init_mu);
        if (!g_init_done[p->device]) {
          HIPCHK(fmmdev::init_constants());
          HIPCHK(fmmdev::init_scratch(p->device));
          g_init_done[p->device] = true;
        }
      }
      HIPCHK(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
      HIPCHK(hipStreamCreateWithFlags(&c->st2, hipStreamNonBlocking));
      for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
      for (auto &e : c->kev) HIPCHK(hipEventCreate(&e));
      HIPCHK(hipEventCreateWithFlags(&c->fork, hipEventDisableTiming));
      HIPCHK(hipEventCreateWithFlags(&c->join, hipEventDisableTiming));
      return FMM_OK;
    }();
  }
  *out = c;  // returned even on failure so fmm_last_error can be read
  return rc;
}

void fmm_destroy(fmm_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->p.device);
  if (c->st) (void)hipStreamSynchronize(c->st);
  graph_reset(c);
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->kev)
    if (e) (void)hipEventDestroy(e);
  if (c->st2) (void)hipStreamSynchronize(c->st2);
  if (c->fork) (void)hipEventDestroy(c->fork);
  if (c->join) (void)hipEventDestroy(c->join);
  if (c->st2) (void)hipStreamDestroy(c->st2);
  if (c->st) {
    fmmdev::plan_scratch_release(c->st);  // plan builds on this stream
    (void)hipStreamDestroy(c->st);
  }
  if (c->pin) (void)hipHostFree(c->pin);
  delete c;  // the device buffers are freed here, with the device current
}

const char *fmm_last_error(const fmm_ctx *c) { return c ? c->err.c_str() : "null context"; }

namespace {
void plan_fill(fmm_plan *pl, int n, const double *x, int ncrit, double theta, int nspawn, int images = 0,
               const double *cycle = nullptr) {
  pl->images = images;
  if (images > 0)
    for (int d = 0; d < 3; d++) pl->cycle[d] = cycle[d];
  pl->n = n;
  double t0 = now_ms();
  fmm::get_bounds(n, x, pl->xmin, pl->xmax);
  double t1 = now_ms();
  pl->tree = fmm::build_tree(n, x, ncrit, pl->xmin, pl->xmax);
  double t2 = now_ms();
  const double xper[3] = {0, 0, 0};
  pl->lists = fmm::build_lists(pl->tree, theta, nspawn, xper, images, pl->cycle);
  double t3 = now_ms();
  pl->t_bounds = t1 - t0;
  pl->t_tree = t2 - t1;
  pl->t_lists = t3 - t2;
}

void plan_stats(const fmm_plan *pl, fmm_stats &s) {
  s = fmm_stats{};
  s.n = pl->n;
  s.ncell = pl->tree.ncell;
  s.nleaf = (int)pl->tree.leaves.size();
  s.depth = pl->tree.depth;
  s.m2l_pairs = pl->lists.nm2l;
  s.p2p_pairs = pl->lists.np2p;
  s.p2p_body_pairs = pl->lists.p2p_body_pairs;
  for (int d = 0; d < 3; d++) {
    s.xmin[d] = pl->xmin[d];
    s.xmax[d] = pl->xmax[d];
  }
  s.host_bounds_ms = pl->t_bounds;
  s.host_tree_ms = pl->t_tree;
  s.host_lists_ms = pl->t_lists;
  s.plan_groups_ms = pl->t_groups;
  s.m2l_groups = pl->groups_kind >= 0 ? pl->groups.ngroup : 0;  // groups built with the plan
}

// x_on_device: x is already in c->xin (fmm_set_bodies_device: the plan was
// just built from it there)
int upload_plan(fmm_ctx *c, fmm_plan &&pl, const double *x, const double *q, bool x_on_device = false);
}  // namespace

int fmm_make_bodies(int n, const char *dist, int seed, double *x, double *q) {
  if (n < 0 || (n > 0 && (!x || !q || !dist))) return FMM_ERR_ARG;
  return fmm::make_bodies(n, dist, seed, x, q) ? FMM_OK : FMM_ERR_ARG;
}

int fmm_plan_build(fmm_plan **out, int n, const double *x, int ncrit, double theta, int nspawn) {
  if (!out) return FMM_ERR_ARG;
  *out = nullptr;
  if (n < 0 || (n > 0 && !x) || ncrit < 1 || !(theta > 0) || nspawn < 1) return FMM_ERR_ARG;
  fmm_plan *pl = new (std::nothrow) fmm_plan();
  if (!pl) return FMM_ERR_MEMORY;
  try {
    plan_fill(pl, n, x, ncrit, theta, nspawn);
  } catch (const std::bad_alloc &) {
    delete pl;
    return FMM_ERR_MEMORY;
  }
  *out = pl;
  return FMM_OK;
}

// fmm_plan_build with the lists from the device traversal (fmm_traverse.hip)
int fmm_plan_build_device(fmm_ctx *c, fmm_plan **out, int n, const double *x, int ncrit, double theta,
                          int nspawn) {
  if (!c || !out) return FMM_ERR_ARG;
  *out = nullptr;
  if (n < 0 || (n > 0 && !x) || ncrit < 1 || !(theta > 0) || nspawn < 1)
    FMM_FAIL(FMM_ERR_ARG, "fmm_plan_build_device: bad n, x, ncrit, theta or nspawn");
  if (!c->st) FMM_FAIL(FMM_ERR_STATE, "context was not created successfully");
  std::unique_ptr<fmm_plan> pl(new (std::nothrow) fmm_plan());
  if (!pl) return FMM_ERR_MEMORY;
  try {
    pl->n = n;
    double t0 = now_ms();
    fmm::get_bounds(n, x, pl->xmin, pl->xmax);
    double t1 = now_ms();
    HIPCHK(hipSetDevice(c->p.device));
    {
      fmm::Tree &t = pl->tree;
      t = fmm::Tree();
      t.n = n;
      if (n > 0) {
        double X0[3], R0;
        fmm::root_box(pl->xmin, pl->xmax, X0, R0);
        // the bodies go up through the context's pinned staging into its
        // upload buffer (scratch between uploads)
        int rc = upload_staged(c, {{&c->xin, x, sizeof(double) * 3 * (size_t)n}});
        if (rc) {
          (void)hipStreamSynchronize(c->st);
          return rc;
        }
        HIPCHK(fmmdev::device_tree(n, x, ncrit, X0, R0, t.perm, t.iparent, t.ichild, t.nchild, t.ibody, t.nbody,
                                   t.level, t.X, t.R, c->st, c->xin.as<double>()));
        t.ncell = (int)t.nbody.size();
        fmm::finish_tree(t);
      } else {
        t.level_ptr.assign(1, 0);
      }
    }
    double t2 = now_ms();
    const fmm::Tree &t = pl->tree;
    fmm::Lists L = fmm::empty_lists(t.ncell);
    const double xper[3] = {0, 0, 0};
    // the source columns stay on the device with the plan (pl->dev)
    pl->dev = std::make_shared<fmmdev::PlanArrays>();
    HIPCHK(fmmdev::device_lists(t.ncell, t.X.data(), t.R.data(), t.nchild.data(), t.ichild.data(), t.nbody.data(),
                                theta, nspawn, xper, 0, 0, L.m2l_ptr, L.m2l_src, L.p2p_ptr, L.p2p_src,
                                L.p2p_body_pairs, c->st, pl->dev.get()));
    L.nm2l = t.ncell > 0 ? (int64_t)L.m2l_ptr[t.ncell] : 0;
    L.np2p = t.ncell > 0 ? (int64_t)L.p2p_ptr[t.ncell] : 0;
    for (int i = 0; i < t.ncell; i++) {
      if (L.m2l_ptr[i + 1] > L.m2l_ptr[i]) L.m2l_targets.push_back(i);
      if (L.p2p_ptr[i + 1] > L.p2p_ptr[i]) L.p2p_targets.push_back(i);
    }
    pl->lists = std::move(L);
    double t3 = now_ms();
    // the M2L groups and tiles of the context's precision and order, when its
    // M2L takes the tile path (else, or if they do not fit, built at upload)
    const bool mixed = c->p.precision != FMM_FP64;
    const bool tiles = c->p.m2l_mode == FMM_M2L_TILES ||
                       (c->p.m2l_mode == FMM_M2L_AUTO && fmmdev::m2l_tiles_preferred(c->p.P, mixed));
    if (tiles && (mixed ? fmmdev::m2l_mixed_supported(c->p.P) : fmmdev::m2l_mfma_supported(c->p.P))) {
      bool done = false;
      HIPCHK(fmmdev::device_groups_tiles(pl->tree, pl->lists, mixed, fmmdev::m2l_tile_targets(c->p.P), 16,
                                         fmmdev::kM2LItemSlices, pl->groups, pl->tiles, done, c->st,
                                         pl->dev.get()));
      if (done) {
        pl->groups_kind = mixed ? FMM_MIXED : FMM_FP64;
      } else {
        pl->groups = fmm::M2LGroups();
        pl->tiles = fmm::M2LTiles();
      }
    }
    double t4 = now_ms();
    pl->t_bounds = t1 - t0;
    pl->t_tree = t2 - t1;
    pl->t_lists = t3 - t2;
    pl->t_groups = t4 - t3;
  } catch (const std::bad_alloc &) {
    FMM_FAIL(FMM_ERR_MEMORY, "host allocation failed while building the plan");
  }
  *out = pl.release();
  return FMM_OK;
}

void fmm_plan_destroy(fmm_plan *pl) { delete pl; }

int fmm_plan_build_tree(fmm_plan **out, int n, const double *x, int ncrit, const double *xmin,
                        const double *xmax) {
  if (!out) return FMM_ERR_ARG;
  *out = nullptr;
  if (n < 0 || (n > 0 && (!x || !xmin || !xmax)) || ncrit < 1) return FMM_ERR_ARG;
  fmm_plan *pl = new (std::nothrow) fmm_plan();
  if (!pl) return FMM_ERR_MEMORY;
  try {
    pl->n = n;
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (n > 0)
      for (int d = 0; d < 3; d++) {
        lo[d] = xmin[d];
        hi[d] = xmax[d];
      }
    for (int d = 0; d < 3; d++) {
      pl->xmin[d] = lo[d];
      pl->xmax[d] = hi[d];
    }
    const double t0 = now_ms();
    pl->tree = fmm::build_tree(n, x, ncrit, lo, hi);
    pl->lists = fmm::empty_lists(pl->tree.ncell);
    pl->t_tree = now_ms() - t0;
  } catch (const std::bad_alloc &) {
    delete pl;
    return FMM_ERR_MEMORY;
  }
  *out = pl;
  return FMM_OK;
}

int fmm_plan_pair(fmm_plan **out, const fmm_plan *targets, const fmm_plan *sources, double theta, int nspawn) {
  if (!out) return FMM_ERR_ARG;
  *out = nullptr;
  if (!targets || !sources || !(theta > 0) || nspawn < 1) return FMM_ERR_ARG;
  if (targets->images != 0 || sources->images != 0) return FMM_ERR_UNSUPPORTED;
  if ((int64_t)targets->n + sources->n > INT32_MAX || targets->tree.ncell + (int64_t)sources->tree.ncell > fmm::kCellMask)
    return FMM_ERR_ARG;
  fmm_plan *pl = new (std::nothrow) fmm_plan();
  if (!pl) return FMM_ERR_MEMORY;
  try {
    const fmm::Lists lt = fmm::empty_lists(targets->tree.ncell), ls = fmm::empty_lists(sources->tree.ncell);
    std::vector<fmm::PlanPart> parts = {{&targets->tree, &lt, nullptr, nullptr},
                                        {&sources->tree, &ls, nullptr, nullptr}};
    fmm::merge_plans(parts, pl->tree, pl->lists, nullptr, fmmdev::kM2LChunkPairs);
    const double t0 = now_ms();
    const double xper[3] = {0, 0, 0};
    if (targets->tree.ncell > 0 && sources->tree.ncell > 0)
      pl->lists = fmm::build_lists(pl->tree, theta, nspawn, xper, 0, nullptr, 0, targets->tree.ncell);
    pl->t_lists = now_ms() - t0;
  } catch (const std::bad_alloc &) {
    delete pl;
    return FMM_ERR_MEMORY;
  }
  pl->n = targets->n + sources->n;
  pl->pair_targets = targets->n;
  for (int d = 0; d < 3; d++) {
    pl->xmin[d] = targets->xmin[d];
    pl->xmax[d] = targets->xmax[d];
  }
  pl->t_bounds = targets->t_bounds + sources->t_bounds;
  pl->t_tree = targets->t_tree + sources->t_tree;
  *out = pl;
  return FMM_OK;
}

int fmm_plan_build_groups(fmm_plan *pl, int precision) {
  if (!pl || (precision != FMM_FP64 && precision != FMM_MIXED && precision != FMM_MIXED16)) return FMM_ERR_ARG;
  if (precision == FMM_MIXED16) precision = FMM_MIXED;  // the same groups (by level)
  if (plan_to_host(pl) != hipSuccess) return FMM_ERR_DEVICE;
  try {
    pl->groups = fmm::build_m2l_groups(pl->tree, pl->lists, fmmdev::kM2LChunkPairs, precision == FMM_MIXED,
                                       pl->images > 0 ? pl->cycle : nullptr);
    pl->tiles = fmm::M2LTiles();
  } catch (const std::bad_alloc &) {
    return FMM_ERR_MEMORY;
  }
  pl->groups_kind = precision;
  return FMM_OK;
}

int fmm_plan_prepare(fmm_plan *pl, const fmm_params *p) {
  if (!pl || !p || (p->precision != FMM_FP64 && p->precision != FMM_MIXED && p->precision != FMM_MIXED16))
    return FMM_ERR_ARG;
  if (p->m2l_mode == FMM_M2L_TARGET) return FMM_OK;
  const bool mixed = p->precision != FMM_FP64;
  const int kind = mixed ? FMM_MIXED : FMM_FP64;  // FMM_MIXED16 groups as FMM_MIXED
  if (!(mixed ? fmmdev::m2l_mixed_supported(p->P) : fmmdev::m2l_mfma_supported(p->P))) return FMM_OK;
  try {
    if (pl->groups_kind != kind) {
      if (plan_to_host(pl) != hipSuccess) return FMM_ERR_DEVICE;
      pl->groups = fmm::build_m2l_groups(pl->tree, pl->lists, fmmdev::kM2LChunkPairs, mixed,
                                         pl->images > 0 ? pl->cycle : nullptr);
      pl->groups_kind = kind;
      pl->tiles = fmm::M2LTiles();
    }
    const int TT = fmmdev::m2l_tile_targets(p->P);
    const bool tiles =
        p->m2l_mode == FMM_M2L_TILES || (p->m2l_mode == FMM_M2L_AUTO && fmmdev::m2l_tiles_preferred(p->P, mixed));
    if (tiles && pl->tiles.TT != TT)
      pl->tiles = fmm::build_m2l_tiles(pl->tree, pl->groups, TT, 16);
  } catch (const std::bad_alloc &) {
    return FMM_ERR_MEMORY;
  }
  return FMM_OK;
}

int fmm_plan_merge(fmm_plan **out, int k, const fmm_plan *const *plans) {
  if (!out || k < 1 || !plans) return FMM_ERR_ARG;
  *out = nullptr;
  const int kind = plans[0] ? plans[0]->groups_kind : -1;
  std::vector<fmm::PlanPart> parts;
  int64_t n = 0;
  for (int i = 0; i < k; i++) {
    if (!plans[i] || plans[i]->images != 0) return FMM_ERR_ARG;  // forests of free-space systems only
    if (plan_to_host(plans[i]) != hipSuccess) return FMM_ERR_DEVICE;
    n += plans[i]->n;
    parts.push_back({&plans[i]->tree, &plans[i]->lists, &plans[i]->groups, &plans[i]->tiles});
  }
  if (n > INT32_MAX) return FMM_ERR_ARG;
  // groups built with the plan on the device hold no pair arrays (grp_ptr is
  // {0}): a merge with them takes no groups, the upload builds its own
  bool same_kind = true;
  for (int i = 0; i < k; i++)
    same_kind = same_kind && plans[i]->groups_kind == kind &&
                plans[i]->groups.grp_ptr.size() == (size_t)plans[i]->groups.ngroup + 1;
  fmm_plan *pl = new (std::nothrow) fmm_plan();
  if (!pl) return FMM_ERR_MEMORY;
  try {
    const bool with_groups = same_kind && kind >= 0;
    bool with_tiles = with_groups;
    for (int i = 0; i < k; i++) with_tiles = with_tiles && plans[i]->tiles.TT == plans[0]->tiles.TT;
    with_tiles = with_tiles && plans[0]->tiles.TT > 0;
    fmm::merge_plans(parts, pl->tree, pl->lists, with_groups ? &pl->groups : nullptr, fmmdev::kM2LChunkPairs,
                     with_tiles ? &pl->tiles : nullptr);
    pl->groups_kind = with_groups ? kind : -1;
  } catch (const std::bad_alloc &) {
    delete pl;
    return FMM_ERR_MEMORY;
  }
  pl->n = (int)n;
  for (int d = 0; d < 3; d++) {
    pl->xmin[d] = plans[0]->xmin[d];
    pl->xmax[d] = plans[0]->xmax[d];
  }
  for (int i = 0; i < k; i++) {
    pl->t_bounds += plans[i]->t_bounds;
    pl->t_tree += plans[i]->t_tree;
    pl->t_lists += plans[i]->t_lists;
  }
  *out = pl;
  return FMM_OK;
}

int fmm_plan_stats(const fmm_plan *pl, fmm_stats *s) {
  if (!pl || !s) return FMM_ERR_ARG;
  plan_stats(pl, *s);
  return FMM_OK;
}

namespace {
void copy_tree(const fmm::Tree &t, int *perm, int *iparent, int *ichild, int *nchild, int *ibody,
               int *nbody, double *X, double *R) {
  auto cp = [](int *dst, const std::vector<int> &v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int));
  };
  cp(perm, t.perm);
  cp(iparent, t.iparent);
  cp(ichild, t.ichild);
  cp(nchild, t.nchild);
  cp(ibody, t.ibody);
  cp(nbody, t.nbody);
  if (X && !t.X.empty()) std::memcpy(X, t.X.data(), t.X.size() * sizeof(double));
  if (R && !t.R.empty()) std::memcpy(R, t.R.data(), t.R.size() * sizeof(double));
}
void copy_lists(const fmm::Lists &L, int *m2l_ptr, int *m2l_src, int *p2p_ptr, int *p2p_src) {
  auto cp = [](int *dst, const std::vector<int> &v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int));
  };
  cp(m2l_ptr, L.m2l_ptr);
  cp(m2l_src, L.m2l_src);
  cp(p2p_ptr, L.p2p_ptr);
  cp(p2p_src, L.p2p_src);
}
}  // namespace

int fmm_plan_tree(const fmm_plan *pl, int *perm, int *iparent, int *ichild, int *nchild, int *ibody,
                  int *nbody, double *X, double *R) {
  if (!pl) return FMM_ERR_ARG;
  copy_tree(pl->tree, perm, iparent, ichild, nchild, ibody, nbody, X, R);
  return FMM_OK;
}

int fmm_plan_list_counts(const fmm_plan *pl, uint64_t *icell, int *nump2p, int *numm2l) {
  if (!pl) return FMM_ERR_ARG;
  fmm::list_counts(pl->tree, pl->lists, icell, nump2p, numm2l);
  return FMM_OK;
}

int fmm_get_list_counts(fmm_ctx *c, uint64_t *icell, int *nump2p, int *numm2l) {
  if (!c) return FMM_ERR_ARG;
  if (!c->have_bodies) FMM_FAIL(FMM_ERR_STATE, "no bodies");
  fmm::list_counts(c->tree, c->lists, icell, nump2p, numm2l);
  return FMM_OK;
}

int fmm_plan_lists(const fmm_plan *pl, int *m2l_ptr, int *m2l_src, int *p2p_ptr, int *p2p_src) {
  if (!pl) return FMM_ERR_ARG;
  if ((m2l_src || p2p_src) && plan_to_host(pl) != hipSuccess) return FMM_ERR_DEVICE;
  copy_lists(pl->lists, m2l_ptr, m2l_src, p2p_ptr, p2p_src);
  return FMM_OK;
}

int fmm_set_bodies(fmm_ctx *c, int n, const double *x, const double *q) {
  if (!c) return FMM_ERR_ARG;
  if (n < 0 || (n > 0 && (!x || !q))) FMM_FAIL(FMM_ERR_ARG, "fmm_set_bodies: bad n or NULL array");
  if (!c->st) FMM_FAIL(FMM_ERR_STATE, "context was not created successfully");
  fmm_plan pl;
  try {
    plan_fill(&pl, n, x, c->p.ncrit, c->p.theta, c->p.nspawn, c->p.images, c->p.cycle);
  } catch (const std::bad_alloc &) {
    FMM_FAIL(FMM_ERR_MEMORY, "host allocation failed while building the tree");
  }
  return upload_plan(c, std::move(pl), x, q);
}

int fmm_set_bodies_device(fmm_ctx *c, int n, const double *x, const double *q) {
  if (!c) return FMM_ERR_ARG;
  if (n < 0 || (n > 0 && (!x || !q))) FMM_FAIL(FMM_ERR_ARG, "fmm_set_bodies_device: bad n or NULL array");
  if (!c->st) FMM_FAIL(FMM_ERR_STATE, "context was not created successfully");
  if (c->p.images != 0) FMM_FAIL(FMM_ERR_UNSUPPORTED, "fmm_set_bodies_device: free space only (images = 0)");
  fmm_plan *pl = nullptr;
  int rc = fmm_plan_build_device(c, &pl, n, x, c->p.ncrit, c->p.theta, c->p.nspawn);
  if (rc != FMM_OK) return rc;
  // the plan's contents move into the context (no copy of the tree or
  // lists); the bodies are still in c->xin from the tree build
  rc = upload_plan(c, std::move(*pl), x, q, n > 0);
  fmm_plan_destroy(pl);
  return rc;
}

int fmm_set_bodies_plan(fmm_ctx *c, const fmm_plan *pl, const double *x, const double *q) {
  if (!c || !pl) return FMM_ERR_ARG;
  if (pl->n > 0 && (!x || !q)) FMM_FAIL(FMM_ERR_ARG, "fmm_set_bodies_plan: NULL array");
  if (!c->st) FMM_FAIL(FMM_ERR_STATE, "context was not created successfully");
  if (pl->images != c->p.images ||
      (pl->images > 0 && (pl->cycle[0] != c->p.cycle[0] || pl->cycle[1] != c->p.cycle[1] || pl->cycle[2] != c->p.cycle[2])))
    FMM_FAIL(FMM_ERR_ARG, "fmm_set_bodies_plan: the plan's periodic images differ from the context's");
  fmm_plan copy = *pl;
  return upload_plan(c, std::move(copy), x, q);
}

namespace {
int upload_plan(fmm_ctx *c, fmm_plan &&pl, const double *x, const double *q, bool x_on_device) {
  HIPCHK(hipSetDevice(c->p.device));
  graph_reset(c);
  c->have_bodies = false;
  const double t3 = now_ms();
  // a device-built plan's arrays are copied device to device on their own
  // device, else from host copies
  const bool on_dev = pl.dev && (pl.dev->lists || pl.dev->tiles) && pl.dev->device == c->p.device;
  if (pl.dev && !on_dev) HIPCHK(plan_to_host(&pl));
  std::shared_ptr<fmmdev::PlanArrays> pdev = on_dev ? pl.dev : nullptr;
  const bool dev_lists = pdev && pdev->lists, dev_tiles = pdev && pdev->tiles;
  const int n = pl.n;
  c->tree = std::move(pl.tree);
  c->lists = std::move(pl.lists);
  const fmm::Tree &t = c->tree;
  const fmm::Lists &L = c->lists;
  c->n = n;
  const int nc = t.ncell;
  std::vector<double> hcx(nc), hcy(nc), hcz(nc);
  for (int k = 0; k < nc; k++) {
    hcx[k] = t.X[3 * (size_t)k];
    hcy[k] = t.X[3 * (size_t)k + 1];
    hcz[k] = t.X[3 * (size_t)k + 2];
  }
  // internal cells per level (M2M), body -> leaf map (L2P)
  std::vector<int> internal;
  c->internal_ptr.assign(t.depth + 2, 0);
  c->level_ptr = t.level_ptr;
  for (int l = 0; l <= t.depth && nc > 0; l++) {
    c->internal_ptr[l] = (int)internal.size();
    for (int k = t.level_ptr[l]; k < t.level_ptr[l + 1]; k++) {
      const int cell = t.level_cells[k];
      if (t.nchild[cell] > 0) internal.push_back(cell);
    }
  }
  c->internal_ptr[t.depth + 1] = (int)internal.size();
  const size_t NT = nterm(c);
  int rc = FMM_OK;
#define UP(buf, vec)                                                \
  if ((rc = upload(c, c->buf, (vec).data(), (vec).size()))) {       \
    (void)hipStreamSynchronize(c->st); /* staging vectors in use */ \
    return rc;                                                      \
  }
#define DC(buf, arr)                                                \
  if ((rc = dcopy(c, c->buf, (arr)))) {                             \
    (void)hipStreamSynchronize(c->st); /* staging vectors in use */ \
    return rc;                                                      \
  }
  // bodies as given and the tree's permutation; gathered into tree order
  // (SoA) on the device
  if ((rc = x_on_device ? upload_staged(c, {{&c->qin, q, sizeof(double) * (size_t)n},
                                             {&c->perm, t.perm.data(), sizeof(int) * t.perm.size()}})
                        : upload_staged(c, {{&c->xin, x, sizeof(double) * 3 * (size_t)n},
                                            {&c->qin, q, sizeof(double) * (size_t)n},
                                            {&c->perm, t.perm.data(), sizeof(int) * t.perm.size()}})) ||
      (rc = ensure(c, c->x, sizeof(double) * n)) ||
      (rc = ensure(c, c->y, sizeof(double) * n)) || (rc = ensure(c, c->z, sizeof(double) * n)) ||
      (rc = ensure(c, c->q, sizeof(double) * n))) {
    (void)hipStreamSynchronize(c->st);  // copies from the staging may be in flight
    return rc;
  }
  HIPCHK(fmmdev::launch_gather_bodies(n, c->perm.as<int>(), c->xin.as<double>(), c->qin.as<double>(),
                                      c->x.as<double>(), c->y.as<double>(), c->z.as<double>(), c->q.as<double>(),
                                      c->st));
  UP(cx, hcx) UP(cy, hcy) UP(cz, hcz)
  UP(ibody, t.ibody) UP(nbody, t.nbody) UP(ichild, t.ichild) UP(nchild, t.nchild)
  UP(iparent, t.iparent)
  UP(leaves, t.leaves) UP(internal, internal) UP(level_cells, t.level_cells)
  if ((rc = ensure(c, c->body_leaf, sizeof(int) * n))) {
    (void)hipStreamSynchronize(c->st);
    return rc;
  }
  HIPCHK(fmmdev::launch_body_leaf((int)t.leaves.size(), c->leaves.as<int>(), c->ibody.as<int>(), c->nbody.as<int>(),
                                  c->body_leaf.as<int>(), c->st));
  UP(m2l_ptr, L.m2l_ptr) UP(m2l_tgt, L.m2l_targets)
  if (dev_lists) {
    DC(m2l_src, pdev->m2l_src) DC(p2p_src, pdev->p2p_src)
  } else {
    UP(m2l_src, L.m2l_src) UP(p2p_src, L.p2p_src)
  }
  // P2P targets, leaves of >= 16 bodies first (the kernel's targets per lane
  // follow the leaf size, so every leaf is summed the same way in any batch)
  std::vector<int> p2p_order(L.p2p_targets);
  c->p2p_nwide = (int)(std::stable_partition(p2p_order.begin(), p2p_order.end(),
                                             [&](int cell) { return t.nbody[cell] >= fmmdev::kP2PWideLeaf; }) -
                       p2p_order.begin());
  UP(p2p_ptr, L.p2p_ptr) UP(p2p_tgt, p2p_order)
  // M2M / L2L operators: every level's radius uniform and every child at
  // X_parent + R_child (+-1, +-1, +-1) (up to rounding of the centres)
  c->xl = fmmdev::xlate_mfma_supported(c->p.P) && nc > 1;
  std::vector<double> lrad(t.depth + 1, -1.0);
  for (int k = 0; k < nc && c->xl; k++) {
    const int l = t.level[k];
    if (lrad[l] < 0) lrad[l] = t.R[k];
    if (t.R[k] != lrad[l]) c->xl = false;
    if (k > 0)
      for (int d = 0; d < 3 && c->xl; d++) {
        const double off = std::fabs(t.X[3 * (size_t)k + d] - t.X[3 * (size_t)t.iparent[k] + d]);
        if (!(std::fabs(off - t.R[k]) <= 1e-9 * t.R[k])) c->xl = false;
      }
  }
  if (c->xl) {
    auto octant = [&](int k) {
      int o = 0;
      for (int d = 0; d < 3; d++)
        if (t.X[3 * (size_t)k + d] > t.X[3 * (size_t)t.iparent[k] + d]) o |= 1 << d;
      return o;
    };
    std::vector<int> child8(8 * internal.size(), -1);
    for (size_t i = 0; i < internal.size(); i++) {
      const int pc = internal[i];
      for (int k = t.ichild[pc]; k < t.ichild[pc] + t.nchild[pc]; k++) child8[8 * i + octant(k)] = k;
    }
    // L2L tiles: each level's children grouped by octant, 16 per tile
    std::vector<int> children, tiles;
    c->xl_tptr.assign(t.depth + 2, 0);
    for (int l = 1; l <= t.depth; l++) {
      c->xl_tptr[l] = (int)(tiles.size() / 4);
      for (int o = 0; o < 8; o++) {
        const int first = (int)children.size();
        for (int k = t.level_ptr[l]; k < t.level_ptr[l + 1]; k++) {
          const int cell = t.level_cells[k];
          if (octant(cell) == o) children.push_back(cell);
        }
        for (int b = first; b < (int)children.size(); b += 16)
          tiles.insert(tiles.end(), {o, b, std::min(16, (int)children.size() - b), 0});
      }
    }
    c->xl_tptr[t.depth + 1] = (int)(tiles.size() / 4);
    // operators: M2M for parent levels 0..depth-1, L2L for child levels 1..depth
    const int nop = 16 * (t.depth + 1);
    std::vector<double> dxs(3 * (size_t)nop, 0.0);
    std::vector<int> kinds(nop, 0);
    for (int l = 0; l <= t.depth; l++)
      for (int o = 0; o < 8; o++) {
        const int m = l * 8 + o, q = 8 * (t.depth + 1) + l * 8 + o;
        kinds[q] = 1;
        for (int d = 0; d < 3; d++) {
          const double s = (o >> d) & 1 ? 1.0 : -1.0;
          if (l < t.depth && lrad[l + 1] > 0) dxs[3 * (size_t)m + d] = -s * lrad[l + 1];  // X_parent - X_child
          if (lrad[l] > 0) dxs[3 * (size_t)q + d] = s * lrad[l];                         // X_child - X_parent
        }
        if (l == t.depth || lrad[l + 1] < 0) dxs[3 * (size_t)m] = 1.0;  // unused operator: any finite offset
        if (l == 0) dxs[3 * (size_t)q] = 1.0;
      }
    UP(xl_child8, child8) UP(xl_children, children) UP(xl_tiles, tiles) UP(xl_dx, dxs) UP(xl_kinds, kinds)
  }
  c->mixed = c->p.precision == FMM_MIXED || c->p.precision == FMM_MIXED16;
  c->h16 = c->p.precision == FMM_MIXED16;
  c->use_mfma = c->p.m2l_mode != FMM_M2L_TARGET && L.nm2l > 0 &&
                (c->mixed ? fmmdev::m2l_mixed_supported(c->p.P) : fmmdev::m2l_mfma_supported(c->p.P));
  c->use_tiles = c->use_mfma && (c->p.m2l_mode == FMM_M2L_TILES ||
                                 (c->p.m2l_mode == FMM_M2L_AUTO && fmmdev::m2l_tiles_preferred(c->p.P, c->mixed)));
  c->groups = fmm::M2LGroups();
  c->nchunk = 0;
  c->ntile = 0;
  fmm::M2LTiles tiles;
  fmmdev::PlanArrays loc;  // groups and tiles built on the device by this upload
  bool plan_tiles = false;  // tiles taken from the plan
  if (c->use_mfma) {
    // groups built with the plan on the device carry no chunks (grp_ptr is
    // {0}): the chunk path builds its own
    if (pl.groups_kind == (c->mixed ? FMM_MIXED : FMM_FP64) &&
        (c->use_tiles || pl.groups.grp_ptr.size() == (size_t)pl.groups.ngroup + 1)) {
      c->groups = std::move(pl.groups);
      if (pl.tiles.TT == fmmdev::m2l_tile_targets(c->p.P)) {
        tiles = std::move(pl.tiles);
        plan_tiles = true;
      }
    } else {
      bool done = false;
      if (c->use_tiles && c->p.images == 0 && L.nm2l > 0) {
        // free space on tiles: the groups and tiles built on the device from
        // the lists just uploaded (the host build's groups, classes, tiles
        // and slices), their pair arrays left there; irregular trees and
        // offsets out of the key's range fall back to the host build
        const hipError_t e = fmmdev::device_groups_tiles(t, L, c->mixed, fmmdev::m2l_tile_targets(c->p.P), 16,
                                                         fmmdev::kM2LItemSlices, c->groups, tiles, done, c->st,
                                                         &loc, c->m2l_src.as<int>());
        if (e != hipSuccess) {
          (void)hipStreamSynchronize(c->st);
          HIPCHK(e);
        }
        if (!done) {
          c->groups = fmm::M2LGroups();
          tiles = fmm::M2LTiles();
        }
      }
      if (!done) {
        if (dev_lists) {  // host grouping reads them
          const hipError_t e = fmmdev::fetch_lists(*pdev, c->lists, c->st);
          if (e != hipSuccess) {
            (void)hipStreamSynchronize(c->st);
            HIPCHK(e);
          }
        }
        c->groups = fmm::build_m2l_groups(t, L, fmmdev::kM2LChunkPairs, c->mixed,
                                          c->p.images > 0 ? c->p.cycle : nullptr);
      }
    }
    const fmm::M2LGroups &G = c->groups;
    UP(g_dx, G.dx) UP(r_ptr, G.rest_ptr) UP(r_src, G.rest_src) UP(r_tgt, G.rest_targets)
    if (c->use_tiles) {
      const bool dt = plan_tiles && dev_tiles;  // the plan's tiles, their pair arrays on this device
      if (tiles.TT == 0) tiles = fmm::build_m2l_tiles(t, G, fmmdev::m2l_tile_targets(c->p.P), 16);
      c->ntile = (int)(tiles.item.size() / 4);  // work items
      // the tile kernel holds an item's slice descriptors one per lane
      for (size_t i = 0; i < tiles.item.size(); i += 4)
        if (tiles.item[i + 2] - tiles.item[i + 1] > fmmdev::kM2LItemSlices)
          FMM_FAIL(FMM_ERR_UNSUPPORTED, "M2L tiles: a work item holds more than 64 slices");
      UP(t_tgt, tiles.tgt) UP(t_ptr, tiles.item)
      if (dt) {
        DC(t_slice, pdev->t_slice) DC(t_src, pdev->t_src) DC(t_lt, pdev->t_lt)
      } else if (loc.tiles) {  // built here on the device
        adopt(c->t_slice, loc.t_slice);
        adopt(c->t_src, loc.t_src);
        adopt(c->t_lt, loc.t_lt);
      } else {
        UP(t_slice, tiles.slice) UP(t_src, tiles.src) UP(t_lt, tiles.lt)
      }
      UP(c_dx, G.cdx)
      if (c->mixed) UP(c_sc, G.cscale)
    } else {
      c->nchunk = (int)(G.chunk.size() / 4);
      UP(g_tgt, G.tgt) UP(g_src, G.src) UP(g_chunk, G.chunk)
    }
    if (c->mixed) {
      UP(g_sc, G.scale) UP(cr, t.R)
    }
  }
#undef UP
#undef DC
  HIPCHK(hipStreamSynchronize(c->st));  // host staging vectors die here
  if ((rc = ensure(c, c->pot, sizeof(double) * n))) return rc;
  if ((rc = ensure(c, c->ax, sizeof(double) * n))) return rc;
  if ((rc = ensure(c, c->ay, sizeof(double) * n))) return rc;
  if ((rc = ensure(c, c->az, sizeof(double) * n))) return rc;
  if ((rc = ensure(c, c->M, sizeof(double2) * NT * nc))) return rc;
  if ((rc = ensure(c, c->L, sizeof(double2) * NT * nc))) return rc;
  if (c->mixed && !c->h16 && (rc = ensure(c, c->M32, sizeof(float2) * NT * nc))) return rc;
  if (c->h16 && (rc = ensure(c, c->M16, fmmdev::m16_record_bytes(c->p.P) * nc))) return rc;
  c->cyc = c->p.images > 0 ? make_double3(c->p.cycle[0], c->p.cycle[1], c->p.cycle[2]) : make_double3(0, 0, 0);
  c->far = c->p.images >= 2 && nc > 0;
  if (c->far) {  // traversePeriodic (include/traversal.h:233-287) as one operator on the root
    const int P = c->p.P, PP = P * P;
    const int prange = (int)(0.5 / c->p.theta + 1), nb = 2 * prange + 1;
    const size_t mb = sizeof(double) * PP * PP;
    if ((rc = ensure(c, c->per_F, mb)) || (rc = ensure(c, c->per_T, mb)) || (rc = ensure(c, c->per_S, mb)) ||
        (rc = ensure(c, c->per_cur, mb)) || (rc = ensure(c, c->per_tmp, mb)))
      return rc;
    const int nmax = (nb * nb * nb - 1) * nb * nb * nb;
    if ((rc = ensure(c, c->per_dx, sizeof(double) * 3 * nmax)) || (rc = ensure(c, c->per_kind, sizeof(int) * nmax)))
      return rc;
    HIPCHK(hipMemsetAsync(c->per_F.p, 0, mb, c->st));
    HIPCHK(hipMemsetAsync(c->per_cur.p, 0, mb, c->st));
    HIPCHK(fmmdev::launch_add_identity(c->per_cur.as<double>(), PP, c->st));
    HIPCHK(hipMemsetAsync(c->per_kind.p, 0, sizeof(int) * nmax, c->st));  // kind 0: M2M
    double cyc[3] = {c->p.cycle[0], c->p.cycle[1], c->p.cycle[2]};
    std::vector<double> dl, ds;
    for (int lev = 0; lev < c->p.images - 1; lev++) {
      dl.clear();
      ds.clear();
      for (int ix = -prange; ix <= prange; ix++)
        for (int iy = -prange; iy <= prange; iy++)
          for (int iz = -prange; iz <= prange; iz++) {
            if (ix == 0 && iy == 0 && iz == 0) continue;
            // M2L(root, supercell) with Xperiodic = (i*nb + c) * cycle: dX = -Xperiodic
            for (int cx = -prange; cx <= prange; cx++)
              for (int cy = -prange; cy <= prange; cy++)
                for (int cz = -prange; cz <= prange; cz++)
                  dl.insert(dl.end(), {-(ix * nb + cx) * cyc[0], -(iy * nb + cy) * cyc[1], -(iz * nb + cz) * cyc[2]});
            // supercell copy at X + i * cycle; M2M dX = X_parent - X_child
            ds.insert(ds.end(), {-ix * cyc[0], -iy * cyc[1], -iz * cyc[2]});
          }
      const int nl = (int)(dl.size() / 3), ns = (int)(ds.size() / 3);
      // T_l, then F += T_l cur
      HIPCHK(hipMemcpyAsync(c->per_dx.p, dl.data(), sizeof(double) * dl.size(), hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemsetAsync(c->per_T.p, 0, mb, c->st));
      HIPCHK(fmmdev::launch_m2l_op_accum(P, c->per_dx.as<double>(), nl, c->per_T.as<double>(), c->st));
      HIPCHK(fmmdev::launch_matmul(c->per_T.as<double>(), c->per_cur.as<double>(), c->per_F.as<double>(), PP, 1,
                                   c->st));
      // S_l = I + sum of the supercell M2M operators, cur = S_l cur
      HIPCHK(hipStreamSynchronize(c->st));  // dl is reused below
      HIPCHK(hipMemcpyAsync(c->per_dx.p, ds.data(), sizeof(double) * ds.size(), hipMemcpyHostToDevice, c->st));
      HIPCHK(hipMemsetAsync(c->per_S.p, 0, mb, c->st));
      HIPCHK(fmmdev::launch_xlate_ops(P, c->per_dx.as<double>(), c->per_kind.as<int>(), ns, c->per_S.as<double>(), 1,
                                      c->st));
      HIPCHK(fmmdev::launch_add_identity(c->per_S.as<double>(), PP, c->st));
      HIPCHK(fmmdev::launch_matmul(c->per_S.as<double>(), c->per_cur.as<double>(), c->per_tmp.as<double>(), PP, 0,
                                   c->st));
      std::swap(c->per_cur, c->per_tmp);
      HIPCHK(hipStreamSynchronize(c->st));  // ds dies at the next iteration
      for (int d = 0; d < 3; d++) cyc[d] *= nb;
    }
  }
  if (c->xl) {  // the per-level M2M / L2L operators, once per plan
    const int nop = 16 * (t.depth + 1);
    if ((rc = ensure(c, c->xl_ops, sizeof(double) * fmmdev::xlate_op_stride(c->p.P) * nop))) return rc;
    HIPCHK(fmmdev::launch_xlate_ops(c->p.P, c->xl_dx.as<double>(), c->xl_kinds.as<int>(), nop,
                                    c->xl_ops.as<double>(), 0, c->st));
    if ((rc = ensure(c, c->xl_frag, sizeof(double) * fmmdev::xlate_frag_stride(c->p.P) * nop))) return rc;
    HIPCHK(fmmdev::launch_ops_to_frag(c->p.P, c->xl_ops.as<double>(), nop, c->xl_frag.as<double>(), c->st));
  }
  if (c->use_tiles) {  // harmonics, then the dense operator, of every representative offset, once per plan
    const int nr = c->groups.ncanon;
    const size_t ny = (size_t)fmmdev::group_harmonics_size(c->p.P) * nr;
    if ((rc = ensure(c, c->g_y, sizeof(double2) * ny))) return rc;
    HIPCHK(fmmdev::launch_group_harmonics(c->p.P, c->c_dx.as<double>(), nr, c->g_y.as<double2>(), c->st));
    if ((rc = ensure(c, c->g_ops, fmmdev::group_ops_bytes(c->p.P, c->mixed, c->h16) * nr))) return rc;
    HIPCHK(fmmdev::launch_group_ops(c->p.P, c->mixed, c->h16, c->g_y.as<double2>(), c->c_sc.as<double2>(), nr,
                                    c->g_ops.p, c->st));
  }
  double t4 = now_ms();
  pl.tree = fmm::Tree();
  plan_stats(&pl, c->stats);
  fmm_stats &s = c->stats;
  s.n = n;
  s.ncell = nc;
  s.nleaf = (int)t.leaves.size();
  s.depth = t.depth;
  s.m2l_pairs = L.nm2l;
  s.p2p_pairs = L.np2p;
  s.p2p_body_pairs = L.p2p_body_pairs;
  s.host_upload_ms = t4 - t3;
  s.m2l_groups = c->groups.ngroup;
  s.m2l_ungrouped = c->use_mfma ? c->groups.nrest : L.nm2l;
  s.m2l_path = c->use_tiles ? FMM_M2L_TILES : c->use_mfma ? FMM_M2L_GROUPED : FMM_M2L_TARGET;
  c->have_bodies = true;
  return FMM_OK;
}
}  // namespace

}  // extern "C"

namespace {
// caller-built tree -> plan (lists built here); returns "" or the reason it was refused
std::string plan_from_tree(fmm_plan &pl, int n, const double *x, int ncell, const int *iparent,
                           const int *ichild, const int *nchild, const int *ibody, const int *nbody,
                           const double *X, const double *R, double theta, int nspawn, int images = 0,
                           const double *cycle = nullptr) {
  if (n < 0 || ncell < 0 || (n > 0 && !x) ||
      (ncell > 0 && (!iparent || !ichild || !nchild || !ibody || !nbody || !X || !R)))
    return "bad sizes or NULL array";
  if (!(theta > 0) || nspawn < 1) return "theta must be > 0 and nspawn >= 1";
  pl.n = n;
  const double t0 = now_ms();
  const std::string why = fmm::adopt_tree(pl.tree, n, ncell, iparent, ichild, nchild, ibody, nbody, X, R);
  if (!why.empty()) return why;
  const double t1 = now_ms();
  const double xper[3] = {0, 0, 0};
  pl.images = images;
  if (images > 0)
    for (int d = 0; d < 3; d++) pl.cycle[d] = cycle[d];
  pl.lists = fmm::build_lists(pl.tree, theta, nspawn, xper, images, pl.cycle);
  pl.t_tree = t1 - t0;
  pl.t_lists = now_ms() - t1;
  fmm::get_bounds(n, x, pl.xmin, pl.xmax);
  return "";
}
}  // namespace

extern "C" {

int fmm_plan_from_tree(fmm_plan **out, int n, const double *x, int ncell, const int *iparent,
                       const int *ichild, const int *nchild, const int *ibody, const int *nbody,
                       const double *X, const double *R, double theta, int nspawn, char *why,
                       int why_len) {
  if (!out) return FMM_ERR_ARG;
  *out = nullptr;
  fmm_plan *pl = new (std::nothrow) fmm_plan();
  if (!pl) return FMM_ERR_MEMORY;
  std::string msg;
  try {
    msg = plan_from_tree(*pl, n, x, ncell, iparent, ichild, nchild, ibody, nbody, X, R, theta, nspawn);
  } catch (const std::bad_alloc &) {
    delete pl;
    return FMM_ERR_MEMORY;
  }
  if (!msg.empty()) {
    delete pl;
    if (why && why_len > 0) {
      std::strncpy(why, msg.c_str(), (size_t)why_len - 1);
      why[why_len - 1] = 0;
    }
    return FMM_ERR_ARG;
  }
  *out = pl;
  return FMM_OK;
}

int fmm_set_tree(fmm_ctx *c, int n, const double *x, const double *q, int ncell,
                 const int *iparent, const int *ichild, const int *nchild, const int *ibody,
                 const int *nbody, const double *X, const double *R) {
  if (!c) return FMM_ERR_ARG;
  if (n > 0 && !q) FMM_FAIL(FMM_ERR_ARG, "fmm_set_tree: NULL q");
  if (!c->st) FMM_FAIL(FMM_ERR_STATE, "context was not created successfully");
  fmm_plan pl;
  try {
    const std::string why = plan_from_tree(pl, n, x, ncell, iparent, ichild, nchild, ibody, nbody, X, R,
                                           c->p.theta, c->p.nspawn, c->p.images, c->p.cycle);
    if (!why.empty()) FMM_FAIL(FMM_ERR_ARG, "fmm_set_tree: " + why);
  } catch (const std::bad_alloc &) {
    FMM_FAIL(FMM_ERR_MEMORY, "host allocation failed while adopting the tree");
  }
  return upload_plan(c, std::move(pl), x, q);
}

int fmm_set_charges(fmm_ctx *c, const double *q) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!q && c->n > 0) FMM_FAIL(FMM_ERR_ARG, "fmm_set_charges: NULL q");
  std::vector<double> hq(c->n);
  for (int i = 0; i < c->n; i++) hq[i] = q[c->tree.perm[i]];
  const void *before = c->q.p;
  if ((rc = upload(c, c->q, hq.data(), hq.size()))) return rc;
  if (c->q.p != before) graph_reset(c);  // a captured solve reads the charges in place
  HIPCHK(hipStreamSynchronize(c->st));
  return FMM_OK;
}

int fmm_upward(fmm_ctx *c) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (c->tree.ncell == 0) return FMM_OK;
  const int P = c->p.P;
  auto C = dev_cells(c);
  HIPCHK(fmmdev::launch_p2m(P, dev_bodies(c), C, c->leaves.as<int>(), (int)c->tree.leaves.size(),
                            c->M.as<double2>(), c->st));
  const size_t os = fmmdev::xlate_frag_stride(P);
  for (int l = c->tree.depth; l >= 0; l--) {
    const int b = c->internal_ptr[l], e = c->internal_ptr[l + 1];
    if (e > b && c->xl)
      HIPCHK(fmmdev::launch_m2m_mfma(P, c->internal.as<int>() + b, c->xl_child8.as<int>() + 8 * (size_t)b, e - b,
                                     c->xl_frag.as<double>() + os * 8 * l, c->M.as<double2>(), c->st));
    else if (e > b)
      HIPCHK(fmmdev::launch_m2m(P, C, c->internal.as<int>() + b, e - b, c->M.as<double2>(), c->st));
  }
  return FMM_OK;
}

namespace {
// P2P (targets overwritten) on stream s
int p2p_on(fmm_ctx *c, hipStream_t s) {
  if (c->n > 0) {
    HIPCHK(fmmdev::launch_zero(c->pot.as<double>(), c->n, s));
    HIPCHK(fmmdev::launch_zero(c->ax.as<double>(), c->n, s));
    HIPCHK(fmmdev::launch_zero(c->ay.as<double>(), c->n, s));
    HIPCHK(fmmdev::launch_zero(c->az.as<double>(), c->n, s));
  }
  const double3 xper = c->cyc;  // image shifts of the source codes
  if (c->timing) HIPCHK(hipEventRecord(c->kev[2], s));
  HIPCHK((c->mixed ? fmmdev::launch_p2p32 : fmmdev::launch_p2p)(
      dev_bodies(c), dev_cells(c), c->p2p_tgt.as<int>(), (int)c->lists.p2p_targets.size(),
      c->p2p_ptr.as<int>(), c->p2p_src.as<int>(), xper, c->p.eps2, c->tree.adopted ? 1 : 0, c->p2p_nwide, s));
  if (c->timing) HIPCHK(hipEventRecord(c->kev[3], s));
  return FMM_OK;
}
}  // namespace

int fmm_p2p(fmm_ctx *c) {
  int rc = check_ready(c);
  if (rc) return rc;
  return p2p_on(c, c->st);
}

int fmm_m2l(fmm_ctx *c) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (c->tree.ncell == 0) return FMM_OK;
  HIPCHK(fmmdev::launch_zero(c->L.as<double>(), 2 * (size_t)nterm(c) * c->tree.ncell, c->st));
  const double3 xper = c->cyc;  // image shifts of the source codes
  if (c->timing) HIPCHK(hipEventRecord(c->kev[0], c->st));
  if (c->use_mfma) {
    if (c->h16)
      HIPCHK(fmmdev::launch_m_to_f16(c->p.P, c->M.as<double2>(), c->cr.as<double>(), c->tree.ncell, c->M16.p,
                                     c->st));
    else if (c->mixed)
      HIPCHK(fmmdev::launch_m_to_f32(c->p.P, c->M.as<double2>(), c->cr.as<double>(), c->tree.ncell,
                                     c->M32.as<float2>(), c->st));
    if (c->use_tiles) {
      HIPCHK(fmmdev::launch_m2l_tile(c->p.P, c->mixed, c->h16, c->ntile, c->t_tgt.as<int>(), c->t_ptr.as<int4>(),
                                     c->t_slice.as<int4>(), c->t_src.as<int>(), c->t_lt.as<unsigned short>(),
                                     c->g_ops.p, c->cr.as<double>(),
                                     c->h16 ? c->M16.p : (c->mixed ? c->M32.p : c->M.p), c->L.as<double2>(),
                                     c->st));
    } else if (c->mixed) {
      HIPCHK(fmmdev::launch_m2l_mfma32(c->p.P, c->g_chunk.as<int4>(), c->nchunk, c->g_dx.as<double>(),
                                       c->g_sc.as<double2>(), c->g_tgt.as<int>(), c->g_src.as<int>(),
                                       c->M32.as<float2>(), c->L.as<double2>(), c->st));
    } else {
      HIPCHK(fmmdev::launch_m2l_mfma(c->p.P, c->g_chunk.as<int4>(), c->nchunk, c->g_dx.as<double>(),
                                     c->g_tgt.as<int>(), c->g_src.as<int>(), c->M.as<double2>(),
                                     c->L.as<double2>(), c->st));
    }
    // pairs outside the offset keys (or, mixed, on levels of non-uniform radius): FP64 per target
    if (c->groups.nrest > 0)
      HIPCHK(fmmdev::launch_m2l(c->p.P, dev_cells(c), c->r_tgt.as<int>(),
                                (int)c->groups.rest_targets.size(), c->r_ptr.as<int>(),
                                c->r_src.as<int>(), xper, c->M.as<double2>(), c->L.as<double2>(), 1,
                                c->st));
  } else {
    HIPCHK(fmmdev::launch_m2l(c->p.P, dev_cells(c), c->m2l_tgt.as<int>(),
                              (int)c->lists.m2l_targets.size(), c->m2l_ptr.as<int>(),
                              c->m2l_src.as<int>(), xper, c->M.as<double2>(), c->L.as<double2>(), 0,
                              c->st));
  }
  if (c->far)
    HIPCHK(fmmdev::launch_periodic_far(c->p.P, c->per_F.as<double>(), c->M.as<double2>(), c->L.as<double2>(), c->st));
  if (c->timing) HIPCHK(hipEventRecord(c->kev[1], c->st));
  return FMM_OK;
}

int fmm_downward(fmm_ctx *c) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (c->tree.ncell == 0) return FMM_OK;
  const int P = c->p.P;
  auto C = dev_cells(c);
  const size_t os = fmmdev::xlate_frag_stride(P);
  for (int l = 1; l <= c->tree.depth; l++) {
    const int b = c->level_ptr[l], e = c->level_ptr[l + 1];
    if (c->xl) {
      const int t0 = c->xl_tptr[l], t1 = c->xl_tptr[l + 1];
      if (t1 > t0)
        HIPCHK(fmmdev::launch_l2l_mfma(P, c->xl_tiles.as<int4>() + t0, t1 - t0, c->xl_children.as<int>(),
                                       c->iparent.as<int>(), c->xl_frag.as<double>() + os * (8 * (c->tree.depth + 1) + 8 * l),
                                       c->L.as<double2>(), c->st));
    } else if (e > b) {
      HIPCHK(fmmdev::launch_l2l(P, C, c->level_cells.as<int>() + b, e - b, c->L.as<double2>(), c->st));
    }
  }
  HIPCHK(fmmdev::launch_l2p(P, dev_bodies(c), C, c->body_leaf.as<int>(), c->n, c->L.as<double2>(),
                            c->st));
  return FMM_OK;
}

namespace {
// The solve's launch sequence on the context's two streams: P2P needs the
// bodies only, so it runs on the second stream beside the upward pass and M2L
// (vector ALUs beside the matrix cores), and L2P, which adds into the same
// targets, waits for it (forking after the upward pass instead, so that P2P
// runs beside M2L, was measured no faster: M2L's LDS leaves no room for P2P
// blocks on a CU).  With c->timing the phase events are recorded too.
int solve_sequence(fmm_ctx *c) {
  int rc;
  if (c->timing) HIPCHK(hipEventRecord(c->ev[0], c->st));
  HIPCHK(hipEventRecord(c->fork, c->st));
  HIPCHK(hipStreamWaitEvent(c->st2, c->fork, 0));
  if ((rc = p2p_on(c, c->st2))) return rc;
  HIPCHK(hipEventRecord(c->join, c->st2));
  if ((rc = fmm_upward(c))) return rc;
  if (c->timing) HIPCHK(hipEventRecord(c->ev[1], c->st));
  if (c->timing) HIPCHK(hipEventRecord(c->ev[2], c->st));
  if ((rc = fmm_m2l(c))) return rc;
  if (c->timing) HIPCHK(hipEventRecord(c->ev[3], c->st));
  HIPCHK(hipStreamWaitEvent(c->st, c->join, 0));
  if ((rc = fmm_downward(c))) return rc;
  if (c->timing) HIPCHK(hipEventRecord(c->ev[4], c->st));
  return FMM_OK;
}
}  // namespace

int fmm_evaluate(fmm_ctx *c) {
  int rc = check_ready(c);
  if (rc) return rc;
  c->timing = true;
  rc = [&]() -> int {
  int rc2 = solve_sequence(c);
  if (rc2) return rc2;
  HIPCHK(hipEventSynchronize(c->ev[4]));
  float ms[4];
  for (int i = 0; i < 4; i++) HIPCHK(hipEventElapsedTime(&ms[i], c->ev[i], c->ev[i + 1]));
  float tot;
  HIPCHK(hipEventElapsedTime(&tot, c->ev[0], c->ev[4]));
  c->stats.dev_upward_ms = ms[0];
  c->stats.dev_m2l_ms = ms[2];
  c->stats.dev_downward_ms = ms[3];  // includes any wait for P2P
  c->stats.dev_total_ms = tot;
  float km = 0.f, kp = 0.f;
  if (c->tree.ncell > 0) {
    HIPCHK(hipEventElapsedTime(&km, c->kev[0], c->kev[1]));
    HIPCHK(hipEventElapsedTime(&kp, c->kev[2], c->kev[3]));
  }
  c->stats.dev_m2l_kernel_ms = km;
  c->stats.dev_p2p_kernel_ms = kp;
  c->stats.dev_p2p_ms = kp;  // beside the upward pass and M2L
  return FMM_OK;
  }();
  c->timing = false;
  return rc;
}

int fmm_evaluate_graph(fmm_ctx *c) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!c->gexec) {
    // capture the solve from the main stream (the second stream joins through
    // the fork event); kernel nodes only: the phases clear with kernels
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    rc = solve_sequence(c);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->st, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (rc) return rc;
      c->err = std::string("hipStreamEndCapture failed: ") + hipGetErrorString(e);
      return FMM_ERR_DEVICE;
    }
    c->graph = g;
    const hipError_t ei = hipGraphInstantiate(&c->gexec, c->graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      graph_reset(c);
      c->err = std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ei);
      return FMM_ERR_DEVICE;
    }
  }
  HIPCHK(hipEventRecord(c->ev[0], c->st));
  HIPCHK(hipGraphLaunch(c->gexec, c->st));
  HIPCHK(hipEventRecord(c->ev[4], c->st));
  HIPCHK(hipEventSynchronize(c->ev[4]));
  float tot;
  HIPCHK(hipEventElapsedTime(&tot, c->ev[0], c->ev[4]));
  c->stats.dev_upward_ms = c->stats.dev_m2l_ms = c->stats.dev_downward_ms = 0.0;
  c->stats.dev_p2p_ms = c->stats.dev_m2l_kernel_ms = c->stats.dev_p2p_kernel_ms = 0.0;
  c->stats.dev_total_ms = tot;
  return FMM_OK;
}

int fmm_direct(fmm_ctx *c, int nt, const double *xt, double *trg) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (nt < 0 || (nt > 0 && (!xt || !trg))) FMM_FAIL(FMM_ERR_ARG, "fmm_direct: bad arguments");
  if (nt == 0) return FMM_OK;
  std::vector<double> h(3 * (size_t)nt);
  for (int i = 0; i < nt; i++)
    for (int d = 0; d < 3; d++) h[(size_t)d * nt + i] = xt[3 * (size_t)i + d];
  DBuf tb, ob;
  auto done = [&](int r) {
    (void)hipStreamSynchronize(c->st);
    release(tb);
    release(ob);
    return r;
  };
  if ((rc = upload(c, tb, h.data(), h.size()))) return done(rc);
  if ((rc = ensure(c, ob, sizeof(double) * 4 * nt))) return done(rc);
  const double *tx = tb.as<double>();
  // Traversal::direct (include/traversal.h:355-377): images sum the sources
  // over the cube of shifts |i| <= sum_(l < images) neighbor^l
  int prange = 0;
  if (c->p.images > 0) {
    const int neighbor = (int)(2 * (.5 / c->p.theta + 1) + 1);
    for (int l = 0, pw = 1; l < c->p.images; l++, pw *= neighbor) prange += pw;
  }
  hipError_t e = fmmdev::launch_direct(dev_bodies(c), c->n, tx, tx + nt, tx + 2 * (size_t)nt, nt,
                                       ob.as<double>(), c->p.eps2, prange, c->cyc, c->st);
  if (e != hipSuccess) {
    c->err = std::string("direct kernel launch failed: ") + hipGetErrorString(e);
    return done(FMM_ERR_DEVICE);
  }
  std::vector<double> o(4 * (size_t)nt);
  e = hipStreamSynchronize(c->st);
  if (e == hipSuccess) e = hipMemcpy(o.data(), ob.p, o.size() * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    c->err = std::string("fmm_direct: ") + hipGetErrorString(e);
    return done(FMM_ERR_DEVICE);
  }
  for (int i = 0; i < nt; i++)
    for (int k = 0; k < 4; k++) trg[4 * (size_t)i + k] = o[(size_t)k * nt + i];
  return done(FMM_OK);
}

int fmm_dipole_correction(fmm_ctx *c) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (c->n == 0) return FMM_OK;
  const double *cy = c->p.cycle;
  if (!(cy[0] > 0 && cy[1] > 0 && cy[2] > 0)) FMM_FAIL(FMM_ERR_ARG, "fmm_dipole_correction: cycle must be > 0");
  if ((rc = ensure(c, c->dip, sizeof(double) * 3 * fmmdev::dipole_partials(c->n)))) return rc;
  const double coef = 4 * M_PI / (3 * cy[0] * cy[1] * cy[2]);
  HIPCHK(fmmdev::launch_dipole_correction(dev_bodies(c), c->n, c->dip.as<double>(), coef, c->st));
  return FMM_OK;
}

int fmm_md_evaluate(fmm_ctx *c, int n, const double *x, const double *q, double *p, double *f) {
  if (!c) return FMM_ERR_ARG;
  if (n < 0 || (n > 0 && (!x || !q || !p || !f))) FMM_FAIL(FMM_ERR_ARG, "fmm_md_evaluate: bad n or NULL array");
  const double *cy = c->p.cycle;
  if (!(cy[0] > 0 && cy[1] > 0 && cy[2] > 0)) FMM_FAIL(FMM_ERR_ARG, "fmm_md_evaluate: cycle must be > 0");
  // wrap into [-cycle/2, cycle/2] (vec.h:217-230), on a copy: the caller's x is left as it is
  std::vector<double> xw(x, x + 3 * (size_t)n);
  for (size_t i = 0; i < (size_t)n; i++)
    for (int d = 0; d < 3; d++) {
      double &v = xw[3 * i + d];
      if (v < -cy[d] / 2) v += cy[d];
      if (v > cy[d] / 2) v -= cy[d];
    }
  int rc = fmm_set_bodies(c, n, xw.data(), q);
  if (rc || n == 0) return rc;
  if ((rc = fmm_evaluate(c))) return rc;
  if ((rc = fmm_dipole_correction(c))) return rc;
  std::vector<double> trg(4 * (size_t)n);
  if ((rc = fmm_get_targets(c, trg.data(), 1))) return rc;
  for (size_t i = 0; i < (size_t)n; i++) {  // TRG starts from the caller's p, f (md_distributed.cxx:209-212)
    p[i] += trg[4 * i];
    for (int d = 0; d < 3; d++) f[3 * i + d] += trg[4 * i + 1 + d];
  }
  return FMM_OK;
}

int fmm_ewald(fmm_ctx *c, int n, const double *x, const double *q, int ksize, double alpha, double sigma,
              double cutoff, double *p, double *f) {
  if (!c) return FMM_ERR_ARG;
  if (!c->st) FMM_FAIL(FMM_ERR_STATE, "context was not created successfully");
  if (n < 0 || (n > 0 && (!x || !q || !p || !f)) || ksize < 0 || !(alpha > 0) || !(sigma > 0) || !(cutoff >= 0))
    FMM_FAIL(FMM_ERR_ARG, "fmm_ewald: bad arguments");
  const double *cy = c->p.cycle;
  if (!(cy[0] > 0 && cy[1] > 0 && cy[2] > 0)) FMM_FAIL(FMM_ERR_ARG, "fmm_ewald: cycle must be > 0");
  if (n == 0) return FMM_OK;
  HIPCHK(hipSetDevice(c->p.device));
  // bodies as structure of arrays; the kernels wrap each position into the
  // box as they read it (md_distributed.cxx:329)
  std::vector<double> h(4 * (size_t)n);
  for (size_t i = 0; i < (size_t)n; i++) {
    for (int d = 0; d < 3; d++) h[(size_t)d * n + i] = x[3 * i + d];
    h[3 * (size_t)n + i] = q[i];
  }
  // Ewald::initWaves (include/ewald.h:64-88): the half space of k, |k| <= ksize,
  // tabulated on the device
  const int nw = fmmdev::ewald_nwaves(ksize);
  int rc;
  if ((rc = upload(c, c->ew_b, h.data(), h.size()))) return rc;
  if ((rc = ensure(c, c->ew_t, sizeof(double) * 4 * n))) return rc;
  if ((rc = ensure(c, c->ew_k, sizeof(int) * 3 * std::max(nw, 1)))) return rc;
  if ((rc = ensure(c, c->ew_s, sizeof(double2) * std::max(nw, 1)))) return rc;
  HIPCHK(hipMemsetAsync(c->ew_t.p, 0, sizeof(double) * 4 * n, c->st));
  const double *bx = c->ew_b.as<double>();
  double *tx = c->ew_t.as<double>();
  const fmmdev::EwBodies B{bx, bx + n, bx + 2 * (size_t)n, bx + 3 * (size_t)n,
                           tx, tx + n, tx + 2 * (size_t)n, tx + 3 * (size_t)n, 1, 1};
  HIPCHK(fmmdev::launch_ewald(B, n, ksize, c->ew_k.as<int>(), c->ew_s.as<double2>(),
                              make_double3(cy[0], cy[1], cy[2]), alpha, sigma, cutoff, c->st));
  std::vector<double> o(4 * (size_t)n);
  HIPCHK(hipMemcpyAsync(o.data(), c->ew_t.p, sizeof(double) * o.size(), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (size_t i = 0; i < (size_t)n; i++) {  // TRG starts from the caller's p, f (md_distributed.cxx:331-334)
    p[i] += o[i];
    for (int d = 0; d < 3; d++) f[3 * i + d] += o[(size_t)(d + 1) * n + i];
  }
  return FMM_OK;
}

int fmm_ewald_noalloc(fmm_ctx *c, int n, const double *x, const double *q, int ksize, double alpha, double sigma,
                      double cutoff, double *p, double *f) {
  if (!c) return FMM_ERR_ARG;
  if (!c->st) FMM_FAIL(FMM_ERR_STATE, "context was not created successfully");
  if (n < 0 || (n > 0 && (!x || !q || !p || !f)) || ksize < 0 || !(alpha > 0) || !(sigma > 0) || !(cutoff >= 0))
    FMM_FAIL(FMM_ERR_ARG, "fmm_ewald_noalloc: bad arguments");
  if (ksize > fmmdev::kEwaldScratchMaxK)
    FMM_FAIL(FMM_ERR_ARG, "fmm_ewald_noalloc: ksize " + std::to_string(ksize) + " exceeds the scratch limit of " +
                              std::to_string(fmmdev::kEwaldScratchMaxK));
  const double *cy = c->p.cycle;
  if (!(cy[0] > 0 && cy[1] > 0 && cy[2] > 0)) FMM_FAIL(FMM_ERR_ARG, "fmm_ewald_noalloc: cycle must be > 0");
  if (n == 0) return FMM_OK;
  HIPCHK(hipSetDevice(c->p.device));
  // the caller's n x 3 positions and forces read in place (stride 3)
  const fmmdev::EwBodies B{x, x + 1, x + 2, q, p, f, f + 1, f + 2, 3, 3};
  HIPCHK(fmmdev::launch_ewald_scratch(c->p.device, B, n, ksize, make_double3(cy[0], cy[1], cy[2]), alpha, sigma,
                                      cutoff, c->st));
  return FMM_OK;
}

int fmm_synchronize(fmm_ctx *c) {
  if (!c) return FMM_ERR_ARG;
  HIPCHK(hipSetDevice(c->p.device));
  if (c->st) HIPCHK(hipStreamSynchronize(c->st));
  return FMM_OK;
}

int fmm_get_targets(fmm_ctx *c, double *trg, int input_order) {
  int rc = check_ready(c);
  if (rc) return rc;
  if (!trg && c->n > 0) FMM_FAIL(FMM_ERR_ARG, "fmm_get_targets: NULL trg");
  const int n = c->n;
  std::vector<double> h(4 * (size_t)n);
  if (n > 0) {  // on the context's stream: other contexts' work is not waited for
    const size_t b = sizeof(double) * n;
    HIPCHK(hipMemcpyAsync(h.data(), c->pot.p, b, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(h.data() + n, c->ax.p, b, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(h.data() + 2 * (size_t)n, c->ay.p, b, hipMemcpyDeviceToHost, c->st));
    HIPCHK(hipMemcpyAsync(h.data() + 3 * (size_t)n, c->az.p, b, hipMemcpyDeviceToHost, c->st));
  }
  HIPCHK(hipStreamSynchronize(c->st));
  for (int i = 0; i < n; i++) {
    const size_t o = input_order ? (size_t)c->tree.perm[i] : (size_t)i;
    for (int k = 0; k < 4; k++) trg[4 * o + k] = h[(size_t)k * n + i];
  }
  return FMM_OK;
}

int fmm_get_expansions(fmm_ctx *c, double *M, double *L) {
  int rc = check_ready(c);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->st));
  const int P = c->p.P, NT = nterm(c), nc = c->tree.ncell;
  std::vector<double> s(NT);
  for (int n = 0; n < P; n++)
    for (int m = 0; m <= n; m++) s[n * (n + 1) / 2 + m] = s_nm(n, m);
  std::vector<double> h(2 * (size_t)NT * nc);
  if (M && nc) {
    HIPCHK(hipMemcpy(h.data(), c->M.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < (size_t)NT * nc; k++) {
      M[2 * k] = h[2 * k] * s[k % NT];
      M[2 * k + 1] = h[2 * k + 1] * s[k % NT];
    }
  }
  if (L && nc) {
    HIPCHK(hipMemcpy(h.data(), c->L.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < (size_t)NT * nc; k++) {
      L[2 * k] = h[2 * k] / s[k % NT];
      L[2 * k + 1] = h[2 * k + 1] / s[k % NT];
    }
  }
  return FMM_OK;
}

int fmm_get_stats(fmm_ctx *c, fmm_stats *s) {
  if (!c || !s) return FMM_ERR_ARG;
  *s = c->stats;
  return FMM_OK;
}

int fmm_get_tree(fmm_ctx *c, int *perm, int *iparent, int *ichild, int *nchild, int *ibody,
                 int *nbody, double *X, double *R) {
  if (!c) return FMM_ERR_ARG;
  if (!c->have_bodies) FMM_FAIL(FMM_ERR_STATE, "no bodies");
  copy_tree(c->tree, perm, iparent, ichild, nchild, ibody, nbody, X, R);
  return FMM_OK;
}

int fmm_get_lists(fmm_ctx *c, int *m2l_ptr, int *m2l_src, int *p2p_ptr, int *p2p_src) {
  if (!c) return FMM_ERR_ARG;
  if (!c->have_bodies) FMM_FAIL(FMM_ERR_STATE, "no bodies");
  const fmm::Lists &L = c->lists;
  copy_lists(L, m2l_ptr, (int64_t)L.m2l_src.size() == L.nm2l ? m2l_src : nullptr, p2p_ptr,
             (int64_t)L.p2p_src.size() == L.np2p ? p2p_src : nullptr);
  // lists uploaded from a device-built plan: the source columns from the context's device copies
  HIPCHK(hipSetDevice(c->p.device));
  if (m2l_src && (int64_t)L.m2l_src.size() != L.nm2l)
    HIPCHK(hipMemcpy(m2l_src, c->m2l_src.p, sizeof(int) * (size_t)L.nm2l, hipMemcpyDeviceToHost));
  if (p2p_src && (int64_t)L.p2p_src.size() != L.np2p)
    HIPCHK(hipMemcpy(p2p_src, c->p2p_src.p, sizeof(int) * (size_t)L.np2p, hipMemcpyDeviceToHost));
  return FMM_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- Kernel operators
// The reference's per-cell Kernel methods (kernels/laplace.h:150-388) on
// caller-held cells: each call lays its cells out as a small cell table on the
// device (one int block, one double block), runs the same kernel the passes
// use, and adds the result to the caller's arrays in the reference's
// normalisation (M = M~ s, L = L^ / s, s = sqrt((n-m)!(n+m)!)).
namespace {
struct OpLayout {
  // host staging: doubles, then ints; offsets in elements
  std::vector<double> d;
  std::vector<int> i;
  size_t put(const double *v, size_t k) {
    const size_t o = d.size();
    d.insert(d.end(), v, v + k);
    return o;
  }
  size_t zeros(size_t k) {
    const size_t o = d.size();
    d.resize(o + k, 0.0);
    return o;
  }
  void align() {  // next double at a 16-byte offset (complex values)
    if (d.size() & 1) d.push_back(0.0);
  }
  size_t puti(std::initializer_list<int> v) {
    const size_t o = i.size();
    i.insert(i.end(), v);
    return o;
  }
  size_t zerosi(size_t k) {
    const size_t o = i.size();
    i.resize(o + k, 0);
    return o;
  }
};

struct OpDev {
  double *d;
  int *i;
};

int op_begin(fmm_ctx *c) {
  if (!c) return FMM_ERR_ARG;
  if (!c->st) FMM_FAIL(FMM_ERR_STATE, "context was not created successfully");
  HIPCHK(hipSetDevice(c->p.device));
  return FMM_OK;
}

int op_upload(fmm_ctx *c, const OpLayout &h, OpDev &dv) {
  int rc;
  if ((rc = upload(c, c->op_d, h.d.data(), h.d.size()))) return rc;
  if ((rc = upload(c, c->op_i, h.i.data(), h.i.size()))) return rc;
  dv.d = c->op_d.as<double>();
  dv.i = c->op_i.as<int>();
  return FMM_OK;
}

int op_download(fmm_ctx *c, const OpDev &dv, size_t off, size_t count, std::vector<double> &out) {
  out.resize(count);
  if (count) HIPCHK(hipMemcpyAsync(out.data(), dv.d + off, sizeof(double) * count, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return FMM_OK;
}

std::vector<double> norm_table(int P) {  // s(n,m) at index n(n+1)/2+m
  std::vector<double> s(P * (P + 1) / 2);
  for (int n = 0; n < P; n++)
    for (int m = 0; m <= n; m++) s[n * (n + 1) / 2 + m] = s_nm(n, m);
  return s;
}

// cell table of k cells (SoA centres from X (k x 3)), bodies/children fields zero
struct OpCells {
  size_t cx, cy, cz, ibody, nbody, ichild, nchild, iparent;
};
OpCells op_cells(OpLayout &h, int k, const double *X) {
  OpCells o;
  o.cx = h.zeros(k);
  o.cy = h.zeros(k);
  o.cz = h.zeros(k);
  for (int c = 0; c < k; c++) {
    h.d[o.cx + c] = X[3 * (size_t)c];
    h.d[o.cy + c] = X[3 * (size_t)c + 1];
    h.d[o.cz + c] = X[3 * (size_t)c + 2];
  }
  o.ibody = h.zerosi(k);
  o.nbody = h.zerosi(k);
  o.ichild = h.zerosi(k);
  o.nchild = h.zerosi(k);
  o.iparent = h.zerosi(k);
  return o;
}
fmmdev::DevCells dev_cells_of(const OpDev &dv, const OpCells &o) {
  return {dv.d + o.cx, dv.d + o.cy, dv.d + o.cz, dv.i + o.ibody,
          dv.i + o.nbody, dv.i + o.ichild, dv.i + o.nchild, dv.i + o.iparent};
}
// n bodies (SoA) from x (n x 3) and q (may be NULL), targets zeroed
struct OpBodies {
  size_t x, y, z, q, pot, ax, ay, az;
  size_t nb;  // bodies in the layout: the stride of every array above
};
OpBodies op_bodies(OpLayout &h, int n, const double *x, const double *q) {
  OpBodies o;
  o.x = h.zeros(n);
  o.y = h.zeros(n);
  o.z = h.zeros(n);
  o.q = h.zeros(n);
  for (int b = 0; b < n; b++) {
    h.d[o.x + b] = x[3 * (size_t)b];
    h.d[o.y + b] = x[3 * (size_t)b + 1];
    h.d[o.z + b] = x[3 * (size_t)b + 2];
    if (q) h.d[o.q + b] = q[b];
  }
  o.pot = h.zeros(n);
  o.ax = h.zeros(n);
  o.ay = h.zeros(n);
  o.az = h.zeros(n);
  o.nb = (size_t)n;
  return o;
}
fmmdev::DevBodies dev_bodies_of(const OpDev &dv, const OpBodies &o) {
  return {dv.d + o.x, dv.d + o.y, dv.d + o.z, dv.d + o.q, dv.d + o.pot, dv.d + o.ax, dv.d + o.ay, dv.d + o.az};
}
// the targets (pot, ax, ay, az) of the first n bodies added into trg (n x 4);
// the four arrays are contiguous, o.nb values each
int op_add_targets(fmm_ctx *c, const OpDev &dv, const OpBodies &o, int n, double *trg) {
  std::vector<double> h;
  int rc = op_download(c, dv, o.pot, 4 * o.nb, h);
  if (rc) return rc;
  for (int b = 0; b < n; b++)
    for (int k = 0; k < 4; k++) trg[4 * (size_t)b + k] += h[(size_t)k * o.nb + b];
  return FMM_OK;
}
// expansions: reference values -> device normalisation (M / s or L * s)
size_t op_expansion(OpLayout &h, int P, const double *E, bool local) {
  const std::vector<double> s = norm_table(P);
  h.align();
  const size_t nt = s.size(), o = h.zeros(2 * nt);
  if (E)
    for (size_t u = 0; u < nt; u++) {
      const double f = local ? s[u] : 1.0 / s[u];
      h.d[o + 2 * u] = E[2 * u] * f;
      h.d[o + 2 * u + 1] = E[2 * u + 1] * f;
    }
  return o;
}
// device expansion at off added into E in the reference normalisation
int op_add_expansion(fmm_ctx *c, const OpDev &dv, size_t off, int P, bool local, double *E) {
  const std::vector<double> s = norm_table(P);
  std::vector<double> h;
  int rc = op_download(c, dv, off, 2 * s.size(), h);
  if (rc) return rc;
  for (size_t u = 0; u < s.size(); u++) {
    const double f = local ? 1.0 / s[u] : s[u];
    E[2 * u] += h[2 * u] * f;
    E[2 * u + 1] += h[2 * u + 1] * f;
  }
  return FMM_OK;
}
}  // namespace

extern "C" {

int fmm_kernel_p2p(fmm_ctx *c, int ni, const double *xi, double *trgi, int nj, const double *xj,
                   const double *qj, const double *xper) {
  int rc = op_begin(c);
  if (rc) return rc;
  if (ni < 0 || nj < 0 || (ni > 0 && (!xi || !trgi)) || (nj > 0 && (!xj || !qj)))
    FMM_FAIL(FMM_ERR_ARG, "fmm_kernel_p2p: bad sizes or NULL array");
  if (ni == 0 || nj == 0) return FMM_OK;
  if ((int64_t)ni + nj > INT32_MAX) FMM_FAIL(FMM_ERR_ARG, "fmm_kernel_p2p: too many bodies");
  OpLayout h;
  // bodies: the ni targets, then the nj sources; cell 0 holds the targets, cell 1 the sources
  std::vector<double> xb(3 * ((size_t)ni + nj)), qb((size_t)ni + nj, 0.0);
  std::memcpy(xb.data(), xi, sizeof(double) * 3 * ni);
  std::memcpy(xb.data() + 3 * (size_t)ni, xj, sizeof(double) * 3 * nj);
  std::memcpy(qb.data() + ni, qj, sizeof(double) * nj);
  const OpBodies B = op_bodies(h, ni + nj, xb.data(), qb.data());
  const double X[6] = {0, 0, 0, 0, 0, 0};
  const OpCells C = op_cells(h, 2, X);
  h.i[C.ibody + 1] = ni;
  h.i[C.nbody] = ni;
  h.i[C.nbody + 1] = nj;
  // list: target cell 0 <- source cell 1, shifted by xper through image 26
  // (shift (1, 1, 1) * cycle with cycle = xper)
  const size_t tg = h.puti({0}), ptr = h.puti({0, 1, 1});
  const size_t src = h.puti({xper ? 1 | (26 << 26) : 1});
  const double3 cyc = xper ? make_double3(xper[0], xper[1], xper[2]) : make_double3(0, 0, 0);
  OpDev dv;
  if ((rc = op_upload(c, h, dv))) return rc;
  HIPCHK(fmmdev::launch_p2p(dev_bodies_of(dv, B), dev_cells_of(dv, C), dv.i + tg, 1, dv.i + ptr, dv.i + src, cyc,
                            c->p.eps2, 1, ni >= fmmdev::kP2PWideLeaf ? 1 : 0, c->st));
  return op_add_targets(c, dv, B, ni, trgi);
}

int fmm_kernel_p2m(fmm_ctx *c, const double *Xc, int n, const double *x, const double *q, double *M) {
  int rc = op_begin(c);
  if (rc) return rc;
  if (!Xc || !M || n < 0 || (n > 0 && (!x || !q))) FMM_FAIL(FMM_ERR_ARG, "fmm_kernel_p2m: bad arguments");
  if (n == 0) return FMM_OK;
  OpLayout h;
  const OpBodies B = op_bodies(h, n, x, q);
  const OpCells C = op_cells(h, 1, Xc);
  h.i[C.nbody] = n;
  const size_t leaves = h.puti({0});
  h.align();
  const size_t Mo = h.zeros(2 * (size_t)nterm(c));
  OpDev dv;
  if ((rc = op_upload(c, h, dv))) return rc;
  HIPCHK(fmmdev::launch_p2m(c->p.P, dev_bodies_of(dv, B), dev_cells_of(dv, C), dv.i + leaves, 1,
                            reinterpret_cast<double2 *>(dv.d + Mo), c->st));
  return op_add_expansion(c, dv, Mo, c->p.P, false, M);
}

int fmm_kernel_m2m(fmm_ctx *c, const double *Xp, int nchild, const double *Xc, const double *Mc, double *Mp) {
  int rc = op_begin(c);
  if (rc) return rc;
  if (!Xp || !Mp || nchild < 0 || (nchild > 0 && (!Xc || !Mc)) || nchild > (1 << 20))
    FMM_FAIL(FMM_ERR_ARG, "fmm_kernel_m2m: bad arguments");
  if (nchild == 0) return FMM_OK;
  const int NT = nterm(c);
  OpLayout h;
  std::vector<double> X(3 * ((size_t)nchild + 1));
  std::memcpy(X.data(), Xp, sizeof(double) * 3);
  std::memcpy(X.data() + 3, Xc, sizeof(double) * 3 * nchild);
  const OpCells C = op_cells(h, nchild + 1, X.data());
  h.i[C.ichild] = 1;
  h.i[C.nchild] = nchild;
  // multipoles of the parent (cell 0, overwritten) and the children
  h.align();
  const size_t Mo = h.zeros(2 * (size_t)NT);
  for (int k = 0; k < nchild; k++) op_expansion(h, c->p.P, Mc + 2 * (size_t)NT * k, false);
  const size_t cells = h.puti({0});
  OpDev dv;
  if ((rc = op_upload(c, h, dv))) return rc;
  HIPCHK(fmmdev::launch_m2m(c->p.P, dev_cells_of(dv, C), dv.i + cells, 1, reinterpret_cast<double2 *>(dv.d + Mo),
                            c->st));
  return op_add_expansion(c, dv, Mo, c->p.P, false, Mp);
}

int fmm_kernel_m2l(fmm_ctx *c, int npair, const double *Xi, const double *Xj, const double *xper,
                   const double *Mj, double *Li) {
  int rc = op_begin(c);
  if (rc) return rc;
  if (npair < 0 || (npair > 0 && (!Xi || !Xj || !Mj || !Li)) || npair > (1 << 24))
    FMM_FAIL(FMM_ERR_ARG, "fmm_kernel_m2l: bad arguments");
  if (npair == 0) return FMM_OK;
  const int NT = nterm(c);
  OpLayout h;
  // cells [0, npair) the targets, [npair, 2 npair) the sources
  std::vector<double> X(6 * (size_t)npair);
  std::memcpy(X.data(), Xi, sizeof(double) * 3 * npair);
  std::memcpy(X.data() + 3 * (size_t)npair, Xj, sizeof(double) * 3 * npair);
  const OpCells C = op_cells(h, 2 * npair, X.data());
  h.align();
  const size_t Mo = h.zeros(2 * (size_t)NT * npair);  // targets' M (unused)
  for (int k = 0; k < npair; k++) op_expansion(h, c->p.P, Mj + 2 * (size_t)NT * k, false);
  h.align();
  const size_t Lo = h.zeros(4 * (size_t)NT * npair);
  const size_t tg = h.zerosi(npair), ptr = h.zerosi(npair + 1), src = h.zerosi(npair);
  for (int k = 0; k < npair; k++) {
    h.i[tg + k] = k;
    h.i[ptr + k + 1] = k + 1;
    h.i[src + k] = xper ? (npair + k) | (26 << 26) : npair + k;
  }
  const double3 cyc = xper ? make_double3(xper[0], xper[1], xper[2]) : make_double3(0, 0, 0);
  OpDev dv;
  if ((rc = op_upload(c, h, dv))) return rc;
  HIPCHK(fmmdev::launch_m2l(c->p.P, dev_cells_of(dv, C), dv.i + tg, npair, dv.i + ptr, dv.i + src, cyc,
                            reinterpret_cast<const double2 *>(dv.d + Mo), reinterpret_cast<double2 *>(dv.d + Lo), 0,
                            c->st));
  const std::vector<double> s = norm_table(c->p.P);
  std::vector<double> o;
  if ((rc = op_download(c, dv, Lo, 2 * (size_t)NT * npair, o))) return rc;
  for (int k = 0; k < npair; k++)
    for (int u = 0; u < NT; u++) {
      Li[2 * ((size_t)NT * k + u)] += o[2 * ((size_t)NT * k + u)] / s[u];
      Li[2 * ((size_t)NT * k + u) + 1] += o[2 * ((size_t)NT * k + u) + 1] / s[u];
    }
  return FMM_OK;
}

int fmm_kernel_l2l(fmm_ctx *c, const double *Xc, const double *Xp, const double *Lp, double *Lc) {
  int rc = op_begin(c);
  if (rc) return rc;
  if (!Xc || !Xp || !Lp || !Lc) FMM_FAIL(FMM_ERR_ARG, "fmm_kernel_l2l: NULL argument");
  OpLayout h;
  const double X[6] = {Xp[0], Xp[1], Xp[2], Xc[0], Xc[1], Xc[2]};
  const OpCells C = op_cells(h, 2, X);  // cell 1 (the child) pulls from its parent, cell 0
  const size_t Lo = op_expansion(h, c->p.P, Lp, true);
  const size_t Lco = op_expansion(h, c->p.P, nullptr, true);
  (void)Lco;  // follows the parent's block: cell 1's local expansion
  const size_t cells = h.puti({1});
  OpDev dv;
  if ((rc = op_upload(c, h, dv))) return rc;
  HIPCHK(fmmdev::launch_l2l(c->p.P, dev_cells_of(dv, C), dv.i + cells, 1, reinterpret_cast<double2 *>(dv.d + Lo),
                            c->st));
  return op_add_expansion(c, dv, Lo + 2 * (size_t)nterm(c), c->p.P, true, Lc);
}

int fmm_kernel_l2p(fmm_ctx *c, const double *Xc, const double *L, int n, const double *x, double *trg) {
  int rc = op_begin(c);
  if (rc) return rc;
  if (!Xc || !L || n < 0 || (n > 0 && (!x || !trg))) FMM_FAIL(FMM_ERR_ARG, "fmm_kernel_l2p: bad arguments");
  if (n == 0) return FMM_OK;
  OpLayout h;
  const OpBodies B = op_bodies(h, n, x, nullptr);
  const OpCells C = op_cells(h, 1, Xc);
  h.i[C.nbody] = n;
  const size_t Lo = op_expansion(h, c->p.P, L, true);
  const size_t leaf = h.zerosi(n);  // every body in cell 0
  OpDev dv;
  if ((rc = op_upload(c, h, dv))) return rc;
  HIPCHK(fmmdev::launch_l2p(c->p.P, dev_bodies_of(dv, B), dev_cells_of(dv, C), dv.i + leaf, n,
                            reinterpret_cast<const double2 *>(dv.d + Lo), c->st));
  return op_add_targets(c, dv, B, n, trg);
}

}  // extern "C"

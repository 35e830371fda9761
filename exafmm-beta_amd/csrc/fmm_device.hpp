// fmm_device.hpp -- launchers for the gfx950 Laplace FMM kernels.
//
// Device data layout (all in HBM, tree order):
//   bodies: SoA x, y, z, q and targets pot, ax, ay, az (double)
//   cells : SoA centres cx, cy, cz; int ibody, nbody, ichild, nchild, iparent
//   M     : [ncell][NTERM] complex, M~[n,m] = M[n,m] / s(n,m)
//   L     : [ncell][NTERM] complex, L^[n,m] = L[n,m] * s(n,m)
// with NTERM = P(P+1)/2, index n(n+1)/2+m, and s(n,m) = sqrt((n-m)!(n+m)!);
// M and L are exafmm-beta's coefficients (kernels/laplace.h) rescaled so every
// operator becomes a plain sum over regular / singular solid harmonics.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "fmm_host.hpp"

namespace fmmdev {

struct DevBodies {
  const double *x, *y, *z, *q;
  double *pot, *ax, *ay, *az;
};

struct DevCells {
  const double *cx, *cy, *cz;
  const int *ibody, *nbody, *ichild, *nchild, *iparent;
};

constexpr int kMaxP = 20;

// p[0, n) = 0 with a kernel (no memset node when the stream is being captured)
hipError_t launch_zero(double *p, size_t n, hipStream_t s);
// bodies into tree order: x/y/z/q[i] from input body perm[i] of xin (AoS)
// and qin
hipError_t launch_gather_bodies(int n, const int *perm, const double *xin, const double *qin, double *x, double *y,
                                double *z, double *q, hipStream_t s);
// body_leaf[i] = the leaf holding body i (tree order)
hipError_t launch_body_leaf(int nleaf, const int *leaves, const int *ibody, const int *nbody, int *body_leaf,
                            hipStream_t s);

bool order_supported(int P);
hipError_t init_constants();

// P2M for every leaf in `leaves` (kernels/laplace.h:237-252)
hipError_t launch_p2m(int P, DevBodies B, DevCells C, const int *leaves, int nleaf, double2 *M,
                      hipStream_t s);
// M2M into every parent listed in `cells` from its children (laplace.h:254-290)
hipError_t launch_m2m(int P, DevCells C, const int *cells, int ncell, double2 *M, hipStream_t s);
// M2L: L[t] (=|+=) sum over the CSR list of t (laplace.h:292-322), one
// wavefront per target, harmonics generated per pair (deterministic order)
hipError_t launch_m2l(int P, DevCells C, const int *targets, int ntarget, const int *ptr,
                      const int *src, double3 xper, const double2 *M, double2 *L, int accumulate,
                      hipStream_t s);
// M2L over offset groups on FP64 MFMA: chunks are (group, begin, count, 0) over
// the grouped pair arrays; adds into L with atomics (L must hold its initial
// value).  Available for the orders where m2l_mfma_supported(P).
bool m2l_mfma_supported(int P);
constexpr int kM2LChunkPairs = 512;  // pairs per M2L chunk (one block)
constexpr int kM2LItemSlices = 64;   // most slices in one M2L tile work item (one per lane)
hipError_t launch_m2l_mfma(int P, const int4 *chunks, int nchunk, const double *gdx, const int *ptgt,
                           const int *psrc, const double2 *M, double2 *L, hipStream_t s);
// L2L into every cell in `cells` from its parent (laplace.h:324-357); accumulates
hipError_t launch_l2l(int P, DevCells C, const int *cells, int ncell, double2 *L, hipStream_t s);
// L2P for every body (laplace.h:359-388); accumulates into the targets
hipError_t launch_l2p(int P, DevBodies B, DevCells C, const int *body_leaf, int n,
                      const double2 *L, hipStream_t s);
// P2P for every target leaf over its CSR list (laplace.h:150-235); overwrites
// the targets of each listed leaf
// target-stationary grouped M2L (tiles of m2l_tile_targets(P) targets, slices
// of <= 16 pairs of one group, one block per work item of a tile); mix: FP32
// operands on M32, else FP64 on M.  A tile in one item stores its targets' L,
// a split tile adds into L (which must start at zero).
int m2l_tile_targets(int P);
bool m2l_tiles_preferred(int P, bool mixed);  // AUTO uses tiles (else chunks) at this order
// h16 (with mix): split-FP16 operands (tables of group_ops_bytes(P, true, true)
// per group, multipoles from launch_m_to_f16)
hipError_t launch_m2l_tile(int P, bool mix, bool h16, int nitem, const int *ttgt, const int4 *items,
                           const int4 *slices, const int *psrc, const unsigned short *plt, const void *ops,
                           const double *cr, const void *M, double2 *L, hipStream_t s);
// every group's dense operator as the tile kernel's MFMA A operands (from the
// group harmonics); group_ops_bytes: bytes per group
size_t group_ops_bytes(int P, bool mix, bool h16 = false);
hipError_t launch_group_ops(int P, bool mix, bool h16, const double2 *gY, const double2 *gscale, int ngroup,
                            void *ops, hipStream_t s);
// singular harmonics of every group offset (group_harmonics_size(P) complex each)
int group_harmonics_size(int P);
hipError_t launch_group_harmonics(int P, const double *gdx, int ngroup, double2 *Y, hipStream_t s);
// M2M / L2L as per-level dense operators on MFMA (8 child offsets per level):
// ops built once per plan (kind 0 M2M with dX = X_parent - X_child, 1 L2L
// with dX = X_child - X_parent; xlate_op_stride(P) doubles each)
bool xlate_mfma_supported(int P);
int xlate_op_stride(int P);
// the operators in MFMA A-operand order (what m2m_mfma / l2l_mfma read);
// xlate_frag_stride(P) doubles each
int xlate_frag_stride(int P);
hipError_t launch_ops_to_frag(int P, const double *ops, int nop, double *frag, hipStream_t s);
// accum != 0: sum of all operators into ops (row stride P^2, no padding)
hipError_t launch_xlate_ops(int P, const double *dxs, const int *kinds, int nop, double *ops, int accum,
                            hipStream_t s);
// periodic far images: T += sum of the M2L operators of the n offsets dxs
// (P^2 x P^2, row-major), dense products, and L(root) += F v(M(root))
hipError_t launch_m2l_op_accum(int P, const double *dxs, int n, double *T, hipStream_t s);
hipError_t launch_matmul(const double *A, const double *B, double *C, int n, int acc, hipStream_t s);
hipError_t launch_add_identity(double *A, int n, hipStream_t s);
hipError_t launch_periodic_far(int P, const double *F, const double2 *M, double2 *L, hipStream_t s);
// parents of one level in tiles of 16; child8[i*8 + octant] (-1: none)
hipError_t launch_m2m_mfma(int P, const int *parents, const int *child8, int np, const double *ops8, double2 *M,
                           hipStream_t s);
// tiles {octant, first, count <= 16} of one level's children grouped by octant
hipError_t launch_l2l_mfma(int P, const int4 *tiles, int ntile, const int *children, const int *iparent,
                           const double *ops8, double2 *L, hipStream_t s);
// mixed precision (FP32 list phases): M32 = M~ / R^n per cell; grouped FP32-MFMA
// M2L with per-group scales (R_s, R_t) accumulating into the FP64 L; FP32 P2P
bool m2l_mixed_supported(int P);
hipError_t launch_m_to_f32(int P, const double2 *M, const double *R, int ncell, float2 *M32, hipStream_t s);
// split-FP16 multipoles: per cell NT terms (re hi|lo, im hi|lo halves) and
// its k-block scales, m16_record_bytes(P) per cell
size_t m16_record_bytes(int P);
hipError_t launch_m_to_f16(int P, const double2 *M, const double *R, int ncell, void *M16, hipStream_t s);
hipError_t launch_m2l_mfma32(int P, const int4 *chunks, int nchunk, const double *gdx, const double2 *gscale,
                             const int *ptgt, const int *psrc, const float2 *M32, double2 *L, hipStream_t s);
// targets[0, nwide) are leaves of at least kP2PWideLeaf bodies (run with four
// targets per lane), the rest run with one
constexpr int kP2PWideLeaf = 16;
hipError_t launch_p2p32(DevBodies B, DevCells C, const int *targets, int ntarget, const int *ptr,
                        const int *src, double3 xper, double eps2, int mask_all, int nwide, hipStream_t s);
hipError_t launch_p2p(DevBodies B, DevCells C, const int *targets, int ntarget, const int *ptr,
                      const int *src, double3 xper, double eps2, int mask_all, int nwide, hipStream_t s);

// Direct sum (include/traversal.h:355-377) of all n sources at nt targets
// (SoA txyz); writes tpot/tax/tay/taz.
// getDipole + dipoleCorrection on the bodies' targets; part: 3 * dipole_partials(n) doubles
int dipole_partials(int n);
hipError_t launch_dipole_correction(DevBodies B, int n, double *part, double coef, hipStream_t s);
// Ewald sum (include/ewald.h): bodies with positions x[i*xs], y[i*xs],
// z[i*xs] (each wrapped once into the box as it is read), charges q[i]; the
// results are added into pot[i] and ax/ay/az[i*fs]
struct EwBodies {
  const double *x, *y, *z, *q;
  double *pot, *ax, *ay, *az;
  int xs, fs;
};
// number of wave vectors of Ewald::initWaves for ksize
int ewald_nwaves(int ksize);
// wave part over the ewald_nwaves(ksize) wave vectors (table kv: 3 ints each,
// S: one double2 each, both written here), real part over minimum-image pairs
// within the cutoff, self term
hipError_t launch_ewald(EwBodies B, int n, int ksize, int *kv, double2 *S, double3 cyc, double alpha,
                        double sigma, double cutoff, hipStream_t s);
// The same on the device's static scratch array (no allocation): ksize up to
// kEwaldScratchMaxK; init_scratch(device) must have run once with that device
// current.  Calls that overlap on one device share the array and race.
constexpr int kMaxDevices = 64;
constexpr int kEwaldScratchMaxK = 32;
hipError_t init_scratch(int device);
// Scratch of the plan builders (device tree, lists, groups), cached per
// stream: a block given back on stream s is handed to a later request on s
// (the stream's order makes the reuse safe), the smallest cached block that
// fits and is at most twice the request plus 1 MiB, else a new hipMalloc'd
// one (current device).  A stream keeps at most 8 GiB cached (a block beyond
// that is freed); plan_scratch_release(s) frees the stream's blocks after
// its work (the owner calls it before destroying s).  Replaces a
// hipMallocAsync / hipFreeAsync pair per buffer, about 0.2 ms of host time
// each on this runtime.
hipError_t plan_scratch_get(void **p, size_t *cap, size_t bytes, hipStream_t s);
void plan_scratch_put(void *p, size_t cap, hipStream_t s);
void plan_scratch_release(hipStream_t s);
// One device allocation (hipMalloc on `device`), freed on that device when
// its owner goes; move-only.
struct DevArray {
  void *p = nullptr;
  size_t bytes = 0;
  int device = -1;
  DevArray() = default;
  DevArray(const DevArray &) = delete;
  DevArray &operator=(const DevArray &) = delete;
  ~DevArray() { reset(); }
  hipError_t alloc(size_t b, int dev);  // the device must be current
  void reset();
  template <class T> T *as() const { return static_cast<T *>(p); }
};
// The large arrays of a plan built on the device, left there for the upload
// into a context on the same device: the lists' source columns (the host
// Lists keep the row pointers and counts, their m2l_src and p2p_src stay
// empty) and the M2L tiles' slices, pair sources and pair targets (the host
// M2LTiles keep the tiles, items and target cells).  Written on the
// builder's stream, complete when the build returns.
struct PlanArrays {
  int device = -1;
  bool lists = false, tiles = false;
  DevArray m2l_src, p2p_src;    // int per pair
  DevArray t_slice, t_src, t_lt;  // 4 ints per slice, int and uint16 per pair
  int64_t nslice = 0;
};
// Copy the device-held parts into the host structures (synchronous on s).
hipError_t fetch_lists(const PlanArrays &d, fmm::Lists &L, hipStream_t s);
hipError_t fetch_tiles(const PlanArrays &d, fmm::M2LTiles &T, hipStream_t s);
// The M2L groups and tiles of a regular free-space tree built on the device
// (fmm_tiles.hip): the same partition into groups and classes as
// fmm::build_m2l_groups (numbered in key order), the same tiles, slices and
// work items as fmm::build_m2l_tiles; done = false (no error) when the plan
// does not fit (the caller then builds them on the host).  Only the fields
// the tile path reads are filled (no chunks).  With keep, the pair sources
// are read from keep->m2l_src when keep->lists (else from L.m2l_src), and
// the tiles' slices, pair sources and pair targets are left in keep instead
// of T (keep->tiles).  dsrc_in: the lists' source column already on the
// device (stream s), read instead of L.m2l_src.
hipError_t device_groups_tiles(const fmm::Tree &t, const fmm::Lists &L, bool by_level, int TT, int SW, int seg,
                               fmm::M2LGroups &G, fmm::M2LTiles &T, bool &done, hipStream_t s,
                               PlanArrays *keep = nullptr, const int *dsrc_in = nullptr);
// The octree of n bodies (host array x, AoS) built on the device: the same
// cells, numbering and body order as fmm::build_tree for the root box (X0,
// R0) (fmm_traverse.hip); outputs are host arrays.  dx_in: the same bodies
// already on the device (then x is not read), written on stream s.
hipError_t device_tree(int n, const double *x, int ncrit, const double X0[3], double R0, std::vector<int> &perm,
                       std::vector<int> &iparent, std::vector<int> &ichild, std::vector<int> &nchild,
                       std::vector<int> &ibody, std::vector<int> &nbody, std::vector<int> &level,
                       std::vector<double> &X, std::vector<double> &R, hipStream_t s,
                       const double *dx_in = nullptr);
// The interaction lists of a free-space tree built on the device (the dual
// traversal breadth first, fmm_traverse.hip): CSR lists sorted by (target,
// source), the P2P body-pair count; the cell arrays are host arrays.  With
// keep, the source columns stay on the device in keep->m2l_src and
// keep->p2p_src (keep->lists) and m2l_src, p2p_src are left empty.
hipError_t device_lists(int ncell, const double *X, const double *R, const int *nchild, const int *ichild,
                        const int *nbody, double theta, int nspawn, const double xper[3], int root_i, int root_j,
                        std::vector<int> &m2l_ptr, std::vector<int> &m2l_src, std::vector<int> &p2p_ptr,
                        std::vector<int> &p2p_src, int64_t &body_pairs, hipStream_t s,
                        PlanArrays *keep = nullptr);
hipError_t launch_ewald_scratch(int device, EwBodies B, int n, int ksize, double3 cyc, double alpha, double sigma,
                                double cutoff, hipStream_t s);
hipError_t launch_direct(DevBodies B, int n, const double *tx, const double *ty, const double *tz, int nt,
                         double *out, double eps2, int prange, double3 cyc, hipStream_t s);

}  // namespace fmmdev

/*
 * exafmm_amd.h -- C ABI of the MI355X Laplace FMM evaluator (libexafmm_amd.so).
 *
 * Plain pointers and sizes only.  Bodies are passed as x[3*i+d] (the layout of
 * exafmm-beta's Body::X), q[i] (Body::SRC); targets come back as trg[4*i+c]
 * with c = potential, d/dx, d/dy, d/dz (Body::TRG, include/types.h:72-84).
 * Expansions use exafmm-beta's coefficients: [ncell][NTERM] complex stored as
 * interleaved doubles, NTERM = P(P+1)/2, index n(n+1)/2+m
 * (kernels/laplace.h:246-247).
 *
 * Every entry point returns 0 on success and a negative fmm_status otherwise;
 * fmm_last_error() then holds a message.  No exception crosses the ABI.
 * Calls on one context are serialised by the caller; separate contexts may be
 * used from separate host threads.
 */
#ifndef EXAFMM_AMD_H
#define EXAFMM_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct fmm_ctx fmm_ctx;

typedef enum {
  FMM_OK = 0,
  FMM_ERR_ARG = -1,      /* invalid argument (size, order, NULL pointer) */
  FMM_ERR_STATE = -2,    /* call out of order (e.g. evaluate before bodies) */
  FMM_ERR_DEVICE = -3,   /* HIP runtime error */
  FMM_ERR_MEMORY = -4,   /* device or host allocation failed */
  FMM_ERR_UNSUPPORTED = -5
} fmm_status;

/* FP64: every phase in FP64 (parity with the reference to 1e-10).
 * MIXED: P2P and M2L in FP32 (P2P with FP32 partial sums of at most 32 terms
 * added into FP64 accumulators; M2L as grouped FP32 MFMA translations whose
 * results are added into FP64 local expansions); P2M, M2M, L2L, L2P, the
 * harmonics of every translation and all accumulation across lists stay FP64.
 * Needs P <= 12 and m2l_mode AUTO or GROUPED.
 * MIXED16: MIXED with the M2L on split-FP16 operands: every operator row
 * (scaled by a power of two) and every cell's multipole (likewise) are held
 * as hi + lo halves, applied with v_mfma_f32_16x16x32_f16 as hi x hi + hi x lo
 * + lo x hi into FP32 accumulators (22 significant bits of every value within
 * 2^-14 of its row's or cell's largest); accumulation across lists in FP64
 * as for MIXED.  Needs P <= 12 and m2l_mode AUTO or TILES. */
typedef enum { FMM_FP64 = 0, FMM_MIXED = 1, FMM_MIXED16 = 2 } fmm_precision;

/* M2L execution.  Pairs are grouped by offset (one dense operator per group,
 * applied on MFMA) in two layouts.  Target tiles: a block owns a tile of
 * target cells and sums all their groups in LDS before one store (a tile too
 * long for one block is split and its parts are added atomically).  Chunks:
 * one block per chunk of <= 512 pairs of one group, results added with FP64
 * atomics.  AUTO (P <= 12) uses tiles (measured faster at every order and
 * precision: the chunks' atomics cost more than the tiles' operator reads); GROUPED
 * always uses chunks, TILES always tiles; all repeat to rounding.  The m2l_path
 * of fmm_stats reports TILES or GROUPED for the layout taken.  TARGET: one wavefront per
 * target cell on vector ALUs (fixed summation order, bitwise repeatable; any
 * P).  Pairs whose offset does not fit a group run per target in every mode. */
typedef enum { FMM_M2L_AUTO = 0, FMM_M2L_TARGET = 1, FMM_M2L_GROUPED = 2, FMM_M2L_TILES = 3 } fmm_m2l_mode;

/* The knobs of exafmm-beta's Kernel / Traversal / BuildTree constructors and
 * Args defaults (include/args.h:188-205): P=10, theta=0.4, ncrit=64,
 * nspawn=5000, images=0, eps2=0. */
typedef struct {
  int P;          /* expansion order, 2..20 */
  int ncrit;      /* max bodies per leaf (BuildTree(ncrit)) */
  int nspawn;     /* Traversal nspawn: both-sides split threshold */
  int images;     /* periodic images (Traversal(..., images, ...)): 0 = free space; > 0 = the
                   * 27 neighbouring copies in the lists plus images - 1 levels of far copies
                   * (include/traversal.h:233-287, 326-352) */
  double theta;   /* multipole acceptance parameter */
  double eps2;    /* softening added to r^2 in P2P */
  int precision;  /* fmm_precision */
  int device;     /* HIP device ordinal */
  int m2l_mode;   /* fmm_m2l_mode */
  double cycle[3];  /* period in x, y, z (the `cycle` argument of Traversal::traverse); used
                     * when images > 0 */
} fmm_params;

typedef struct {
  int n, ncell, nleaf, depth;
  int64_t m2l_pairs, p2p_pairs, p2p_body_pairs;
  double xmin[3], xmax[3];
  /* host wall times of the last fmm_set_bodies (ms) */
  double host_bounds_ms, host_tree_ms, host_lists_ms, host_upload_ms;
  /* device times of the last evaluation (ms, HIP events on the ctx stream) */
  double dev_upward_ms, dev_p2p_ms, dev_m2l_ms, dev_downward_ms, dev_total_ms;
  /* the M2L and P2P kernels alone (no memsets), same events */
  double dev_m2l_kernel_ms, dev_p2p_kernel_ms;
  /* M2L grouping: offset groups, pairs left to the per-target kernel, path
   * taken by the last upload (fmm_m2l_mode value actually used) */
  int64_t m2l_groups, m2l_ungrouped;
  int m2l_path;
  /* M2L groups and tiles built with the plan on the device (fmm_plan_build_device;
   * otherwise 0 and their host build is part of host_upload_ms) */
  double plan_groups_ms;
} fmm_stats;

/* Fill p with the reference defaults (include/args.h:188-205). */
void fmm_default_params(fmm_params *p);

/* Dataset::initBodies (include/dataset.h:230-255): seeded drand48 inputs in
 * the reference's draw order.  dist: "cube", "sphere", "octant", "plummer"
 * (the reference's generators) or "cube1" ([-1,1]^3, zero-mean
 * q in [-1,1]) and "plummer1" (unit-scale Plummer truncated at r=10, q=1/n).
 * No device needed. */
int fmm_make_bodies(int n, const char *dist, int seed, double *x, double *q);

/* ---- host-only plan (no device needed) ----
 * BoundBox::getBounds + BuildTree::buildTree + the pair lists of
 * Traversal::traverse (include/bound_box.h:11-25, include/build_tree.h:186-209,
 * include/traversal.h:326-352), built on the host.  fmm_set_bodies builds one
 * internally; a plan can also be built ahead (e.g. on another host thread). */
typedef struct fmm_plan fmm_plan;
int fmm_plan_build(fmm_plan **out, int n, const double *x, int ncrit, double theta, int nspawn);
/* fmm_plan_build with the tree (BuildTree::buildTree, include/build_tree.h:77-209)
 * and the pair lists of Traversal::traverse (include/traversal.h:90-177,
 * 326-352; free space) built on the device of context c.  The tree: every
 * body's octant path from the same centre arithmetic, one stable sort, and
 * the runs of shared octants as cells -- the same cells, cell numbering and
 * body order as fmm_plan_build.  The lists: the dual traversal breadth first,
 * each round testing every frontier pair with the same acceptance test and
 * split rules -- the same pairs as fmm_plan_build, each cell's lists ordered
 * by source cell (the host lists keep the walk's visiting order).  When c's
 * M2L takes target tiles, the M2L groups and tiles are built there too (the
 * same as fmm_plan_prepare would build for c).  The lists' source columns and
 * the tiles' pair arrays stay in c's device memory with the plan (freed by
 * fmm_plan_destroy): fmm_set_bodies_plan into a context on that device
 * copies them device to device, and the host functions that read them
 * (fmm_plan_lists, fmm_plan_merge, fmm_plan_prepare, a context on another
 * device) copy them to the host first.  The bounds are computed on the host
 * (their sequential 1e-5 drift). */
int fmm_plan_build_device(fmm_ctx *c, fmm_plan **out, int n, const double *x, int ncrit, double theta,
                          int nspawn);
void fmm_plan_destroy(fmm_plan *pl);
/* Traversal::countList / writeList (include/traversal.h:30-38, 412-421):
 * per cell (ncell entries each) its Morton key ICELL (BuildTree::getKey,
 * include/build_tree.h:111-120, 132, computed in 64 bits at every level; the
 * reference's int shifts overflow past level 10) and the lengths of its P2P
 * and M2L lists (periodic image pairs included).  Any array may be NULL. */
int fmm_plan_list_counts(const fmm_plan *pl, uint64_t *icell, int *nump2p, int *numm2l);
int fmm_get_list_counts(fmm_ctx *c, uint64_t *icell, int *nump2p, int *numm2l);

/* BuildTree::buildTree(bodies, buffer, bounds) (include/build_tree.h:186-209)
 * in the caller's bounds (BoundBox::getBounds, include/bound_box.h:11-37):
 * the octree only; its lists are empty until fmm_plan_pair, or a context's
 * fmm_set_tree, builds them. */
int fmm_plan_build_tree(fmm_plan **out, int n, const double *x, int ncrit, const double *xmin,
                        const double *xmax);
/* Traversal::traverse(icells, jcells, ...) with separate target and source
 * trees (include/traversal.h:326-333; examples/fmm.cxx:69-72, the -j path):
 * one forest of the two trees (target cells and bodies first) whose lists are
 * the dual traversal from the pair of roots: the targets' M2L and P2P lists
 * name source cells only.  A context given this plan (fmm_set_bodies_plan
 * with x, q = the targets' then the sources' bodies, each in its plan's input
 * order) evaluates the field of the sources at the targets; the source
 * bodies' own targets stay zero.  Both plans free-space (images = 0). */
int fmm_plan_pair(fmm_plan **out, const fmm_plan *targets, const fmm_plan *sources, double theta, int nspawn);

/* fills n, ncell, nleaf, depth, pair counts, bounds and host times */
int fmm_plan_stats(const fmm_plan *pl, fmm_stats *s);
int fmm_plan_tree(const fmm_plan *pl, int *perm, int *iparent, int *ichild, int *nchild,
                  int *ibody, int *nbody, double *X, double *R);
int fmm_plan_lists(const fmm_plan *pl, int *m2l_ptr, int *m2l_src, int *p2p_ptr, int *p2p_src);

/* Replaces `Kernel kernel(P, eps2, wavek)` + `Traversal(kernel, theta,
 * nspawn, images, path)` + `BuildTree(ncrit)` construction
 * (examples/fmm.cxx:21-26): one context owns a HIP stream and device memory. */
int fmm_create(fmm_ctx **out, const fmm_params *p);
void fmm_destroy(fmm_ctx *c);
const char *fmm_last_error(const fmm_ctx *c);

/* Replaces BoundBox::getBounds + BuildTree::buildTree (include/bound_box.h:11-25,
 * include/build_tree.h:186-209) and the list part of Traversal::traverse
 * (include/traversal.h:326-352): host bounds, octree and interaction lists,
 * then upload of bodies (in tree order), cells and lists.  x is (n,3) in the
 * caller's order. */
int fmm_set_bodies(fmm_ctx *c, int n, const double *x, const double *q);
/* Same, with a plan built by fmm_plan_build for these x (the context copies
 * what it needs; the plan stays owned by the caller). */
int fmm_set_bodies_plan(fmm_ctx *c, const fmm_plan *pl, const double *x, const double *q);
/* fmm_set_bodies with the tree, the lists and (tile path) the M2L groups and
 * tiles built on c's device: fmm_plan_build_device with c's ncrit, theta and
 * nspawn, then the upload, which takes the plan's contents without a copy.
 * Free space only (FMM_ERR_UNSUPPORTED for images > 0).  Each cell's lists
 * are ordered by source cell (fmm_set_bodies keeps the reference's visiting
 * order); results agree with fmm_set_bodies' to the order of the sums. */
int fmm_set_bodies_device(fmm_ctx *c, int n, const double *x, const double *q);

/* Drop-in path for callers that build their own tree (exafmm-beta's
 * BuildTree::buildTree, include/build_tree.h:186-209): bodies already in tree
 * order (x (n,3), q), and the flattened CellBase fields of every cell
 * (include/types.h:98-113) with children contiguous and each cell's bodies
 * the range [ibody, ibody+nbody).  The interaction lists are then built on the
 * host from this tree (Traversal::traverse rules, include/traversal.h:90-177).
 * Structural errors (ranges, child blocks) return FMM_ERR_ARG. */
int fmm_set_tree(fmm_ctx *c, int n, const double *x, const double *q, int ncell,
                 const int *iparent, const int *ichild, const int *nchild, const int *ibody,
                 const int *nbody, const double *X, const double *R);
/* Host-only half of fmm_set_tree: validate a caller-built tree (same
 * arguments; x in tree order, used for the bounds only) and build its
 * interaction lists for theta and nspawn into a plan that fmm_set_bodies_plan
 * accepts (with x, q in tree order: the plan's permutation is the identity).
 * On FMM_ERR_ARG the reason is written to why (NUL-terminated, at most
 * why_len bytes) when why is not NULL. */
int fmm_plan_from_tree(fmm_plan **out, int n, const double *x, int ncell, const int *iparent,
                       const int *ichild, const int *nchild, const int *ibody, const int *nbody,
                       const double *X, const double *R, double theta, int nspawn, char *why,
                       int why_len);

/* Precompute the plan's M2L offset groups for a context of this precision
 * (FMM_FP64: by offset; FMM_MIXED: by offset and the two levels), so that a
 * context given the plan does no grouping work itself.  Host only. */
int fmm_plan_build_groups(fmm_plan *pl, int precision);

/* Everything a context with parameters p needs from the host for its M2L
 * (groups and, for m2l_mode AUTO, the target tiles), built into the plan so
 * that fmm_set_bodies_plan only uploads.  Host only. */
int fmm_plan_prepare(fmm_plan *pl, const fmm_params *p);

/* Batched evaluation of independent systems: one plan for the forest of the
 * k plans (cells, bodies and pairs of plan i follow those of plans 0..i-1; no
 * interaction crosses systems).  Bodies of the merged plan are the k input
 * arrays concatenated in order.  Prebuilt groups are merged when every plan
 * has the same kind and was grouped on the host (groups built on the device
 * hold no pair arrays; the merged plan is then grouped at upload).  No counterpart in the reference (one system per solve
 * there); host only. */
int fmm_plan_merge(fmm_plan **out, int k, const fmm_plan *const *plans);

/* Re-upload charges only (same positions, same tree and lists). */
int fmm_set_charges(fmm_ctx *c, const double *q);

/* Device phases on the context stream (asynchronous; fmm_synchronize or a
 * getter waits).  They replace UpDownPass::upwardPass (include/up_down_pass.h:49-60),
 * the Kernel::M2L / Kernel::P2P calls made by Traversal::dualTreeTraversal
 * (include/traversal.h:117-177) and UpDownPass::downwardPass (:63-75). */
int fmm_upward(fmm_ctx *c);
int fmm_p2p(fmm_ctx *c);
int fmm_m2l(fmm_ctx *c);
int fmm_downward(fmm_ctx *c);
/* upward + p2p + m2l + downward, with per-phase device timing */
int fmm_evaluate(fmm_ctx *c);
/* No counterpart in the reference: fmm_evaluate as one HIP graph launch.  The
 * first call after the bodies, tree or charges buffer change captures the
 * solve's launch sequence (both streams, kernel nodes only) and instantiates
 * it; later calls replay it (for repeated solves of one system with new
 * charges, fmm_set_charges).  Same kernels and results as fmm_evaluate;
 * fmm_stats reports only dev_total_ms for it. */
int fmm_evaluate_graph(fmm_ctx *c);
int fmm_synchronize(fmm_ctx *c);

/* UpDownPass::getDipole(bodies, X0 = 0) + dipoleCorrection(bodies, dipole, n,
 * cycle) (include/up_down_pass.h:77-99) on the context's targets, on the
 * device: D = sum q x over the bodies as set; pot -= 4 pi |D|^2 / (3 V n q),
 * acc -= 4 pi D / (3 V), V = cycle[0] cycle[1] cycle[2].  As in the reference
 * a body with q = 0 gets a non-finite potential. */
int fmm_dipole_correction(fmm_ctx *c);

/* FMM_FMM (wrappers/md_distributed.cxx:190-264), one process: positions
 * wrapped into [-cycle/2, cycle/2] (vec.h:217-230; x itself is not changed),
 * tree, lists and the periodic FMM of the context's parameters, the dipole
 * correction, then p[i] += potential and f[3i..3i+2] += acceleration in the
 * caller's order.  The tree's bounds are those of the wrapped bodies of this
 * call (the reference keeps the bounds of its last FMM_Partition). */
int fmm_md_evaluate(fmm_ctx *c, int n, const double *x, const double *q, double *p, double *f);

/* FMM_Ewald (wrappers/md_distributed.cxx:310-360) on the device, one process:
 * the Ewald sum of include/ewald.h for the n bodies x (wrapped into the box
 * of the context's cycle; x itself is not changed) with charges q:
 * Ewald::wavePart over the wave vectors of initWaves (|k| <= ksize), the real
 * part (erfc-screened pairs within `cutoff`, each at its minimum image) and
 * Ewald::selfTerm; p[i] += potential, f[3i..3i+2] += gradient, caller's
 * order.  Uses the context's stream and its own device buffers; the bodies
 * set for the FMM are not touched. */
int fmm_ewald(fmm_ctx *c, int n, const double *x, const double *q, int ksize, double alpha, double sigma,
              double cutoff, double *p, double *f);

/* No counterpart in the reference: fmm_ewald for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that
 * forbid hipMalloc).  x (n x 3), q (n), p (n) and f (n x 3) are DEVICE
 * pointers of the context's device; p[i] += potential, f[3i..3i+2] +=
 * gradient, with the same kernels and summation order as fmm_ewald.  The wave
 * table and structure factors live in a static per-device scratch array of
 * about 1.9 MB, so ksize <= 32 (larger ksize: FMM_ERR_ARG).  Enqueued on the
 * context's stream and returns without waiting (fmm_synchronize waits);
 * calls no hipMalloc/hipFree.  Calls that overlap on one device share the
 * scratch array and race: serialise them (one stream, or events). */
int fmm_ewald_noalloc(fmm_ctx *c, int n, const double *x, const double *q, int ksize, double alpha, double sigma,
                      double cutoff, double *p, double *f);

/* Traversal::direct (include/traversal.h:355-377): direct sum over all
 * bodies of the context (tree order) -- with images > 0 over the cube of
 * shifts (ix, iy, iz) * cycle, |i| <= sum_(l < images) neighbor^l -- at nt
 * target points xt
 * (nt x 3); trg (nt x 4) is overwritten.  R^2 = 0 pairs are skipped like in
 * Kernel::P2P.  Uses IEEE sqrt and division, independently of the P2P kernel. */
int fmm_direct(fmm_ctx *c, int nt, const double *xt, double *trg);

/* ---- Kernel operators (kernels/laplace.h:150-388) ----
 * The reference's per-cell Kernel methods on cells the caller holds, each one
 * device launch on the context's stream (synchronous); the context needs no
 * bodies.  Expansions are [NTERM] complex values as interleaved doubles in
 * the reference's normalisation; results are added (+=), as the reference's
 * methods add into Ci->M, Ci->L and Body::TRG.  FP64 whatever the context's
 * precision. */
/* Kernel::P2P(Ci, Cj) (laplace.h:150-235): trgi[4i..4i+3] += potential and
 * gradient at xi[i] of the nj sources xj, qj shifted by xper (the Kernel's
 * Xperiodic; NULL = 0) with the context's eps2; R^2 = 0 pairs skipped. */
int fmm_kernel_p2p(fmm_ctx *c, int ni, const double *xi, double *trgi, int nj, const double *xj,
                   const double *qj, const double *xper);
/* Kernel::P2M(C) (laplace.h:237-252): M += the multipole about Xc of n bodies. */
int fmm_kernel_p2m(fmm_ctx *c, const double *Xc, int n, const double *x, const double *q, double *M);
/* Kernel::M2M(Ci, C0) (laplace.h:254-290): Mp += the multipoles Mc (nchild x
 * NTERM) of the children at Xc (nchild x 3) shifted to the parent at Xp. */
int fmm_kernel_m2m(fmm_ctx *c, const double *Xp, int nchild, const double *Xc, const double *Mc, double *Mp);
/* Kernel::M2L(Ci, Cj) (laplace.h:292-322), npair at once: Li[k] += the local
 * expansion about Xi[k] of the multipole Mj[k] about Xj[k] + xper (NULL = 0). */
int fmm_kernel_m2l(fmm_ctx *c, int npair, const double *Xi, const double *Xj, const double *xper,
                   const double *Mj, double *Li);
/* Kernel::L2L(Ci, C0) (laplace.h:324-357): Lc += the parent's Lp about Xp shifted to Xc. */
int fmm_kernel_l2l(fmm_ctx *c, const double *Xc, const double *Xp, const double *Lp, double *Lc);
/* Kernel::L2P(Ci) (laplace.h:359-388): trg (n x 4) += the local expansion L
 * about Xc evaluated at the n points x. */
int fmm_kernel_l2p(fmm_ctx *c, const double *Xc, const double *L, int n, const double *x, double *trg);

/* Targets (n x 4). input_order != 0 returns them in the order of the x passed
 * to fmm_set_bodies, else in tree order (exafmm-beta leaves bodies sorted). */
int fmm_get_targets(fmm_ctx *c, double *trg, int input_order);
/* M and L of every cell (may be NULL), reference normalisation. */
int fmm_get_expansions(fmm_ctx *c, double *M, double *L);

/* Tree and lists as built by fmm_set_bodies (for parity checks).  Any pointer
 * may be NULL.  perm[i] = input index of the i-th body in tree order. */
int fmm_get_stats(fmm_ctx *c, fmm_stats *s);
int fmm_get_tree(fmm_ctx *c, int *perm, int *iparent, int *ichild, int *nchild, int *ibody,
                 int *nbody, double *X, double *R);
int fmm_get_lists(fmm_ctx *c, int *m2l_ptr, int *m2l_src, int *p2p_ptr, int *p2p_src);

/* ------------------------------------------------------------------ serving
 * A request loop for streams of independent systems (successive MD steps,
 * many small problems).  No counterpart in the reference, which solves one
 * system per call.  Submitted systems are planned on host threads (bounds,
 * tree, lists, M2L groups) and run in batches: a free slot (one context and
 * HIP stream) takes the oldest planned requests -- up to batch_bodies bodies
 * or batch_requests requests -- once their device footprint fits
 * device_budget beside the batches in flight, and evaluates them as one
 * forest (one launch per phase for the whole batch).  Batches of different
 * slots overlap on the GPU.  By default a batch holds one request: a batch of
 * several is merged, then grouped on the host, which measured slower than
 * more slots working at once.  With m2l_mode AUTO the loop's M2L runs on
 * target tiles when batches hold one request (a slot's upload builds their
 * groups on the device; the builders plan tree and lists only), and as
 * grouped chunks when they may hold several.  A system's result does not
 * depend on its batch beyond the order of FP64 atomic additions into its
 * local expansions (for mixed precision on explicit FMM_M2L_TILES with
 * batches of several requests it does, at about 1e-8). */
typedef struct fmm_server fmm_server;

typedef struct {
  int host_threads;       /* plan builders; 0: hardware threads - slots (max 16) */
  int slots;              /* batches in flight (contexts/streams), 1..16; default 8 */
  int64_t device_budget;  /* bytes the in-flight batches may occupy */
  int batch_bodies;       /* a batch closes at this many bodies ... */
  int batch_requests;     /* ... or this many requests (default 1) */
  double batch_delay_ms;  /* hold a batch open up to this long for more requests (0: never) */
  int plan_on_device;     /* 1: each builder builds the tree and lists on the device
                           * (fmm_plan_build_device, one context and stream per builder) */
} fmm_server_config;

typedef struct {
  int n;
  int batch_requests, batch_bodies, slot;  /* the batch the request ran in */
  double t_submit_ms, t_planned_ms, t_started_ms, t_done_ms;  /* server clock */
  double device_ms;                        /* device time of that batch */
} fmm_request_info;

typedef struct {
  int64_t requests_submitted, requests_done, batches, bodies_done;
  int max_batch_requests;
  double device_ms;
} fmm_server_stats;

void fmm_server_default_config(fmm_server_config *cfg);
/* p as for fmm_create (validated once here); cfg NULL: defaults */
int fmm_server_create(fmm_server **out, const fmm_params *p, const fmm_server_config *cfg);
/* copies x (n,3) and q; returns at once with the request id */
int fmm_server_submit(fmm_server *s, int n, const double *x, const double *q, int64_t *id);
/* blocks until request id is done; trg (n,4) in the request's input order;
 * info optional.  Each id is waited for once. */
int fmm_server_wait(fmm_server *s, int64_t id, double *trg, fmm_request_info *info);
int fmm_server_get_stats(fmm_server *s, fmm_server_stats *out);
/* stops the threads after the batches in flight.  Requests that have not
 * run end with FMM_ERR_STATE: a fmm_server_wait blocked on one returns that
 * code, and destroy returns only after every such wait has returned.  No
 * wait may start once destroy has been called. */
void fmm_server_destroy(fmm_server *s);

#ifdef __cplusplus
}
#endif
#endif

/*
 * oracle.h -- CPU FP64 restatement of exafmm-beta's Laplace FMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the timed CPU baseline.  The product (libexafmm_amd.so) never links it.
 *
 * Every routine restates one reference routine; the file:line it follows is
 * given next to each declaration (paths relative to the exafmm-beta tree).
 * Data layout: bodies are SoA (x[3*i+d], q[i], trg[4*i+c] with c = pot, ax,
 * ay, az); cells are the reference's CellBase fields in flat arrays; M and L
 * are [ncell][NTERM] complex with NTERM = P(P+1)/2 and index n(n+1)/2+m.
 */
#ifndef EXAFMM_ORACLE_H
#define EXAFMM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_kernel orc_kernel;
typedef struct orc_tree orc_tree;
typedef struct orc_lists orc_lists;

/* Kernel(P, eps2, wavek) tables: kernels/laplace.h:116-148 */
orc_kernel *orc_kernel_create(int P, double eps2);
void orc_kernel_destroy(orc_kernel *k);
int orc_kernel_nterm(const orc_kernel *k);

/* Single operators on explicit geometry ("+=" semantics, like the reference).
 * Complex arrays are interleaved (re, im) doubles. */
/* Kernel::P2P  kernels/laplace.h:150-235 (scalar path :212-234) */
void orc_p2p(const orc_kernel *k, int ni, const double *xi, double *trgi,
             int nj, const double *xj, const double *qj, const double *xper);
/* Kernel::P2M  kernels/laplace.h:237-252 */
void orc_p2m(const orc_kernel *k, const double *Xc, int n, const double *x,
             const double *q, double *M);
/* Kernel::M2M  kernels/laplace.h:254-290 (Ci = parent at Xp, Cj = child at Xc) */
void orc_m2m(const orc_kernel *k, const double *Xp, const double *Xc,
             const double *Mc, double *Mp);
/* Kernel::M2L  kernels/laplace.h:292-322 */
void orc_m2l(const orc_kernel *k, const double *Xi, const double *Xj,
             const double *xper, const double *Mj, double *Li);
/* Kernel::L2L  kernels/laplace.h:324-357 (Ci = child at Xc, parent at Xp) */
void orc_l2l(const orc_kernel *k, const double *Xc, const double *Xp,
             const double *Lp, double *Lc);
/* Kernel::L2P  kernels/laplace.h:359-388 */
void orc_l2p(const orc_kernel *k, const double *Xc, const double *L, int n,
             const double *x, double *trg);

/* BoundBox::getBounds(Bodies&)  include/bound_box.h:11-25 (1e-5 drift kept) */
void orc_bounds(int n, const double *x, double *xmin, double *xmax);
/* getBounds(bodies, bounds) (bound_box.h:28-37): xmin/xmax widened in place */
void orc_bounds_update(int n, const double *x, double *xmin, double *xmax);

/* BuildTree::buildTree  include/build_tree.h:168-209.  x is in input order;
 * the tree records perm[i] = input index of the i-th body in tree order. */
orc_tree *orc_tree_build(int n, const double *x, int ncrit,
                         const double *xmin, const double *xmax);
void orc_tree_destroy(orc_tree *t);
int orc_tree_ncell(const orc_tree *t);
int orc_tree_depth(const orc_tree *t);
void orc_tree_get(const orc_tree *t, int *perm, int *iparent, int *ichild,
                  int *nchild, int *ibody, int *nbody, double *X, double *R,
                  uint64_t *icell);

/* Traversal::dualTreeTraversal / splitCell / TraverseRange with the pairs
 * recorded in visiting order: include/traversal.h:90-177, 180-230, 326-352.
 * type[k] = 0 for M2L, 1 for P2P. */
orc_lists *orc_traverse(const orc_tree *t, double theta, int nspawn,
                        const double *xper);
/* Traversal::traverse(icells, jcells), two trees (traversal.h:326-333):
 * ci indexes ti (targets), cj indexes tj (sources) */
orc_lists *orc_traverse_pair(const orc_tree *ti, const orc_tree *tj, double theta,
                             int nspawn, const double *xper);
void orc_lists_destroy(orc_lists *l);
int64_t orc_lists_count(const orc_lists *l, int type);
void orc_lists_get(const orc_lists *l, int *ci, int *cj, int *type);

/* Full evaluation over a tree, bodies in tree order (x3, q), following
 * examples/fmm.cxx:61-77: upwardPass (include/up_down_pass.h:49-60),
 * traverse (include/traversal.h:326-352), downwardPass (:63-75).
 * trg is accumulated into (zero it first, like Dataset::initTarget).
 * M, L (ncell*NTERM complex) may be NULL. */
void orc_fmm(const orc_kernel *k, const orc_tree *t, const double *x,
             const double *q, double theta, int nspawn, double *trg,
             double *M, double *L);
/* examples/fmm.cxx -j: targets x (tree ti order) from sources xj, qj (tree tj
 * order); trg accumulated for the targets */
void orc_fmm_pair(const orc_kernel *k, const orc_tree *ti, const double *x,
                  const orc_tree *tj, const double *xj, const double *qj,
                  double theta, int nspawn, double *trg);

/* orc_fmm with periodic images (images > 0): Traversal::traverse runs the
 * dual traversal for the 27 shifts (ix, iy, iz) * cycle, ix..iz in -1..1
 * (include/traversal.h:326-352), then traversePeriodic (:233-287) adds the
 * images - 1 levels of supercell copies of the root to the root's local
 * expansion by M2L and M2M; then downwardPass. */
void orc_fmm_periodic(const orc_kernel *k, const orc_tree *t, const double *x,
                      const double *q, double theta, int nspawn, int images,
                      const double *cycle, double *trg);

/* Same evaluation driven by explicit pair lists (as recorded by
 * orc_traverse), used to time the reference operators on a list sample. */
void orc_eval_lists(const orc_kernel *k, const orc_tree *t, const double *x,
                    const double *q, int64_t npair, const int *ci,
                    const int *cj, const int *type, double *trg, double *L,
                    const double *M);
void orc_upward(const orc_kernel *k, const orc_tree *t, const double *x,
                const double *q, double *M);
void orc_downward(const orc_kernel *k, const orc_tree *t, const double *x,
                  double *L, double *trg);

/* Traversal::direct for images = 0  include/traversal.h:355-377 */
void orc_direct(const orc_kernel *k, int ni, const double *xi, double *trgi,
                int nj, const double *xj, const double *qj);

/* Verify::getDifScalar/getNrmScalar/getDifVector/getNrmVector
 * include/verify.h:42-92; returns potRel, accRel as in examples/fmm.cxx:99-104 */
void orc_rel_l2(int n, const double *trg_ref, const double *trg, double *pot_rel,
                double *acc_rel);

/* vec.h:217-230 wrap into [-cycle/2, cycle/2]; returns the moved-axes mask */
int orc_wrap(double *x, const double *cycle);
/* UpDownPass::getDipole (X0 = 0) + dipoleCorrection, include/up_down_pass.h:77-99 */
void orc_dipole_correction(int n, const double *x, const double *q, const double *cycle, double *trg);
/* Ewald::wavePart + realPart + selfTerm, include/ewald.h; trg += (4 per body).
 * Real-space pairs by minimum image (the reference's cell search targets it). */
void orc_ewald(int n, const double *x, const double *q, int ksize, double alpha, double sigma, double cutoff,
               const double *cycle, double *trg);

/* Dataset generators (include/dataset.h) and the configured inputs. */
/* cube(): x in [-pi,pi]^3, dataset.h:52-66 */
void orc_data_cube(int n, int seed, double *x);
/* plummer(): dataset.h:109-136 (scale 3pi/16, reject R >= 100) */
void orc_data_plummer(int n, int seed, double *x);
/* initSource(): dataset.h:142-159, q = drand48()-.5 minus the mean, reseeded */
void orc_data_source(int n, int seed, double *q);
/* sampleBodies(): dataset.h:290-298, returns the sampled tree-order indices */
int orc_sample_indices(int n, int ntarget, int *idx);

#ifdef __cplusplus
}
#endif
#endif

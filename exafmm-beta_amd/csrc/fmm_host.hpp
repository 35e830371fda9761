// fmm_host.hpp -- host-side tree and interaction-list construction for the
// MI355X Laplace FMM evaluator.  These produce exactly the tree and the pair
// lists of exafmm-beta (bit-identical decisions), so the device results can be
// compared with the reference FMM at the same P, theta and ncrit.
#pragma once
#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace fmm {

// Allocator whose resize() leaves new elements uninitialised: the large pair
// arrays below are filled right after by parallel passes, and zero-filling
// them first (one thread, fresh pages) costs as much as building them.
template <class T> struct NoInit : std::allocator<T> {
  template <class U> struct rebind { using other = NoInit<U>; };
  NoInit() = default;
  template <class U> NoInit(const NoInit<U> &) noexcept {}
  template <class U> void construct(U *p) noexcept { ::new (static_cast<void *>(p)) U; }
  template <class U, class... A> void construct(U *p, A &&...a) { ::new (static_cast<void *>(p)) U(std::forward<A>(a)...); }
};
template <class T> using bigvec = std::vector<T, NoInit<T>>;

// Flattened octree in the reference's cell order (nodes2cells,
// include/build_tree.h:123-165): children of a cell are contiguous, bodies of
// a cell are a contiguous range in tree order.
struct Tree {
  int n = 0;                 // bodies
  int ncell = 0;
  int depth = 0;             // deepest level (root = 0)
  std::vector<int> perm;     // tree-order position -> input index
  std::vector<int> iparent, ichild, nchild, ibody, nbody, level;
  std::vector<double> X;     // 3 * ncell, cell centres
  std::vector<double> R;     // cell radii
  // Cells grouped by level (level_ptr[l] .. level_ptr[l+1] in level_cells)
  std::vector<int> level_ptr, level_cells;
  std::vector<int> leaves;   // leaf cells in cell order
  // caller-built (adopt_tree): coincident bodies may sit in different leaves;
  // in trees built here they always share a leaf
  bool adopted = false;
};

// Interaction lists in CSR form per target cell, each list in the order the
// reference's dual tree traversal visits the pairs.
struct Lists {
  int64_t nm2l = 0, np2p = 0;
  std::vector<int> m2l_ptr, m2l_src;   // ncell+1, nm2l
  std::vector<int> p2p_ptr, p2p_src;   // ncell+1, np2p
  std::vector<int> m2l_targets;        // cells with a non-empty M2L list
  std::vector<int> p2p_targets;        // leaves with a non-empty P2P list
  int64_t p2p_body_pairs = 0;          // sum over P2P pairs of Ni*Nj
};

// Host threads for one tree/list/group build: n > 0 fixes the count for the
// calling thread (1 = serial, as the request loop's plan builders use);
// 0 = hardware threads, at most 16.
void set_host_threads(int n);
int host_threads();

// BoundBox::getBounds (include/bound_box.h:11-25): sequential scan that widens
// the box by 1e-5 per body, as the reference does.
void get_bounds(int n, const double *x, double xmin[3], double xmax[3]);

// BuildTree::buildTree (include/build_tree.h:168-209).  x is (n,3) in input
// order.  Returns the tree; perm gives the body order.
Tree build_tree(int n, const double *x, int ncrit, const double xmin[3], const double xmax[3]);
// the root cell of the bounds (bounds2box, include/build_tree.h:168-180)
void root_box(const double xmin[3], const double xmax[3], double X[3], double &R);
// depth, level index and leaf list of a tree whose cell arrays are filled
void finish_tree(Tree &t);

// Wrap a caller-built tree (reference cell order) in a Tree: checks its
// structure and fills levels, level index and leaves.  Returns an empty
// string on success, else the reason it was rejected.
std::string adopt_tree(Tree &t, int n, int ncell, const int *iparent, const int *ichild,
                       const int *nchild, const int *ibody, const int *nbody, const double *X,
                       const double *R);

// Source entries of the lists are codes: the source cell in the low 26 bits
// and, for periodic runs, the image in the bits above (image i = 0..26 is the
// shift (i/9 - 1, i/3 % 3 - 1, i % 3 - 1) * cycle; 0 without images).
constexpr int kImageShift = 26;
constexpr int kCellMask = (1 << kImageShift) - 1;
inline int image_of(int code) { return code >> kImageShift; }
inline int cell_of(int code) { return code & kCellMask; }
inline void image_shift(int img, const double cycle[3], double out[3]) {
  out[0] = (img / 9 - 1) * cycle[0];
  out[1] = (img / 3 % 3 - 1) * cycle[1];
  out[2] = (img % 3 - 1) * cycle[2];
}

// Traversal::traverse (include/traversal.h:117-177, 90-114, 180-230,
// 326-352), recording pairs instead of evaluating them.  images > 0: the
// dual traversal of the root with itself for each of the 27 shifts in the
// reference's loop order, pairs tagged with their image (the hierarchical
// far images of traversePeriodic are not pairs; see the periodic operator).
// root_i, root_j: the pair the traversal starts from (0, 0 for one tree;
// the two roots of a forest for traverse(icells, jcells), traversal.h:326-333)
Lists build_lists(const Tree &t, double theta, int nspawn, const double xper[3], int images = 0,
                  const double *cycle = nullptr, int root_i = 0, int root_j = 0);

// Empty lists for a tree of ncell cells (a tree whose lists are made later).
Lists empty_lists(int ncell);

// M2L pairs grouped by their exact offset X_t - X_s.  Offsets between cells of
// the reference's octree are integer multiples of R_root / 2^e; a group key is
// the canonical (e, k) with k = offset * 2^e / R_root integral, so every pair
// of a group has the same translation operator.  Pairs are stored group by
// group (each group's pairs in CSR order); `chunk` splits groups into work
// items of at most `chunk_pairs` pairs (4 ints each: group, begin, count, 0).
// Pairs whose key does not fit the packing go to the per-target lists rest_*.
struct M2LGroups {
  int ngroup = 0;
  bigvec<int> tgt, src;            // grouped pairs
  std::vector<int> grp_ptr;        // ngroup+1
  std::vector<double> dx;          // 3 per group: representative offset
  std::vector<int> chunk;          // 4 per chunk
  std::vector<int> rest_ptr, rest_src, rest_targets;
  int64_t nrest = 0;
  std::vector<double> scale;       // 2 per group (R_s, R_t) when grouped by level
  // Reflection classes: reflecting an offset in x, y or z changes the M2L
  // operator only by signs on the expansion components (x: (-1)^m conj,
  // y: conj, z: (-1)^(n+m), on both the multipole columns and the local
  // rows), so every group maps to a representative with non-negative offset
  // components (canon) and the reflections that lead there (refl: bit d set
  // when component d was negative).  Periodic groups reflect their image
  // shift with the offset, so the sign test is on the whole translation
  // vector (offset minus image shift).
  int ncanon = 0;
  std::vector<int> canon;          // per group
  std::vector<uint8_t> refl;       // per group
  std::vector<double> cdx;         // 3 per representative
  std::vector<double> cscale;      // 2 per representative when grouped by level
};
// by_level: a group also fixes the source and target levels (the scaled FP32
// operator depends on R_s and R_t); pairs whose radii differ from their
// group's are left ungrouped.
M2LGroups build_m2l_groups(const Tree &t, const Lists &L, int chunk_pairs, bool by_level = false,
                           const double *cycle = nullptr);
// Traversal::countList / writeList (include/traversal.h:30-38, 412-421): per
// cell its Morton key (BuildTree::getKey, include/build_tree.h:111-120, with
// 64-bit arithmetic at every level) and the lengths of its P2P and M2L lists.
// Any output may be null.
void list_counts(const Tree &t, const Lists &L, uint64_t *icell, int *nump2p, int *numm2l);

// Target-stationary layout of the grouped M2L pairs: targets are sorted by
// (level, octant within the parent, cell) -- cells of one octant class share
// their offsets -- and cut into tiles of TT targets.  A tile's pairs are
// ordered by representative and cut into slices of at most SW pairs with
// distinct targets (a slice may mix the representative's reflections), so a
// block can apply each representative's operator to whole slices and keep the
// tile's local expansions in LDS until one final store.  Tiles are listed in
// launch order (dealt to the 8 XCDs in contiguous runs; empty tiles pad).
// A block runs one work item: at most `seg` consecutive slices of one tile.
// A tile split over several items adds its results (atomics); a tile in one
// item stores them.
struct M2LTiles {
  int TT = 0, ntile = 0;
  std::vector<int> tgt;        // ntile*TT target cells (-1 pads), launch order
  std::vector<int> ptr;        // ntile+1 offsets into the slices
  std::vector<int> slice;      // 4 per slice: representative, first pair, count, 0
  bigvec<int> src;             // source cell of every pair
  bigvec<uint16_t> lt;         // every pair: target index in the tile (bits 0-7) and the
                               // reflection bits of its group (bits 8-10)
  std::vector<int> item;       // 4 per work item: tile, first slice, end slice, split (0/1)
};
M2LTiles build_m2l_tiles(const Tree &t, const M2LGroups &G, int TT, int SW, int seg = 64);
void split_m2l_tiles(M2LTiles &T, int seg);

// (Re)build G.chunk from the groups: chunks of at most chunk_pairs pairs,
// ordered by first target and dealt round-robin-contiguously to the 8 XCDs.
void deal_m2l_chunks(M2LGroups &G, int chunk_pairs);

// Several independent systems as one forest (batched evaluation): cell, body
// and pair indices of part k are offset by the sizes of parts 0..k-1, levels
// are kept, and no pair joins two systems.  G (optional) merges the parts'
// M2L groups (every part must carry groups of the same kind).
struct PlanPart {
  const Tree *t;
  const Lists *l;
  const M2LGroups *g;
  const M2LTiles *tl;
};
// tiles (optional): the parts' tiles concatenated (every part must have tiles
// of the same TT)
void merge_plans(const std::vector<PlanPart> &parts, Tree &T, Lists &L, M2LGroups *G, int chunk_pairs,
                 M2LTiles *tiles = nullptr);

// Dataset generators (include/dataset.h:52-171, 220-227): drand48 streams
// seeded with srand48(seed), positions first, then charges after reseeding,
// exactly in the reference's draw order.  Returns false for an unknown name.
//   "cube"     [-pi,pi]^3, q = drand48()-.5 minus mean        (dataset.h:52-66)
//   "sphere"   |x| = pi shell                                  (dataset.h:69-87)
//   "octant"   one octant of the pi sphere                     (dataset.h:90-106)
//   "plummer"  Plummer, scaled by 3pi/16, R < 100              (dataset.h:109-136)
//   "cube1"    [-1,1]^3, q = 2 drand48()-1 minus mean
//   "plummer1" Plummer with unit scale radius, r < 10, q = 1/n
bool make_bodies(int n, const char *dist, int seed, double *x, double *q);

}  // namespace fmm

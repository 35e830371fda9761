// Host-only use of include/exafmm_amd.hpp (no device call): Args, Dataset,
// BoundBox and BuildTree, printing the bodies in tree order and the cells so
// that a test can compare them with the oracle's tree.
#include <cstdio>

#include "exafmm_amd.hpp"

using namespace exafmm_amd;

int main(int argc, char **argv) {
  Args args(argc, argv);
  Dataset data;
  Bodies bodies = data.initBodies(args.numBodies, args.distribution, 0), buffer;
  BoundBox boundBox;
  BuildTree buildTree(args.ncrit);
  Bounds bounds = boundBox.getBounds(bodies);
  Cells cells = buildTree.buildTree(bodies, buffer, bounds);
  std::printf("%d %d %d %.17g %s\n", args.numBodies, args.ncrit, args.P, args.theta, args.distribution);
  std::printf("%.17g %.17g %.17g %.17g %.17g %.17g\n", bounds.Xmin[0], bounds.Xmin[1], bounds.Xmin[2],
              bounds.Xmax[0], bounds.Xmax[1], bounds.Xmax[2]);
  std::printf("%zu %zu\n", bodies.size(), cells.size());
  for (B_iter B = bodies.begin(); B != bodies.end(); B++)
    std::printf("%d %.17g %.17g %.17g %.17g\n", B->IBODY, B->X[0], B->X[1], B->X[2], B->SRC);
  for (C_iter C = cells.begin(); C != cells.end(); C++)
    std::printf("%d %d %d %d %d %.17g %.17g %.17g %.17g %d %llu\n", C->IPARENT, C->ICHILD, C->NCHILD, C->IBODY,
                C->NBODY, C->X[0], C->X[1], C->X[2], C->R, (int)(C->BODY - bodies.begin()),
                (unsigned long long)C->ICELL);
  // Traversal::writeList of the (reset) counts, when a path is given
  if (args.path && std::string(args.path) != "./") {
    Kernel kernel(args.P, 0.0, complex_t(0.0, 0.0));
    Traversal traversal(kernel, args.theta, args.nspawn, 0, args.path);
    traversal.initListCount(cells);
    for (size_t c = 0; c < cells.size(); c++) cells[c].numP2P = (int)c % 7, cells[c].numM2L = (int)c % 11;
    traversal.writeList(cells, 3);
  }
  return 0;
}

// Traversal::direct between traverse and downwardPass: the reference allows
// any order of these calls; the drop-in's direct sum runs in a context of its
// own, so the downward pass still finds the tree of the upward pass.  Prints
// the rel-L2 error of the FMM against that direct sum at the first 50 bodies.
#include <cmath>
#include <iostream>

#include "exafmm_amd.hpp"

using namespace exafmm_amd;

int main(int argc, char **argv) {
  try {
    Args args(argc, argv);
    const vec3 cycle = 2 * M_PI;
    Bodies bodies, buffer;
    Dataset data;
    BoundBox boundBox;
    BuildTree buildTree(args.ncrit);
    Kernel kernel(args.P, 0.0, complex_t(1., .1));
    Traversal traversal(kernel, args.theta, args.nspawn, 0);
    UpDownPass upDownPass(kernel);
    Verify verify;
    bodies = data.initBodies(args.numBodies, args.distribution, 0);
    Bounds bounds = boundBox.getBounds(bodies);
    Cells cells = buildTree.buildTree(bodies, buffer, bounds);
    upDownPass.upwardPass(cells);
    traversal.traverse(cells, cells, cycle, false);
    Bodies sample(bodies.begin(), bodies.begin() + 50);
    data.initTarget(sample);
    traversal.direct(sample, bodies, cycle);
    upDownPass.downwardPass(cells);
    Bodies fmm(bodies.begin(), bodies.begin() + 50);
    const double pot = std::sqrt(verify.getDifScalar(sample, fmm) / verify.getNrmScalar(sample));
    const double acc = std::sqrt(verify.getDifVector(sample, fmm) / verify.getNrmVector(sample));
    std::cout << "pot " << pot << " acc " << acc << std::endl;
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}

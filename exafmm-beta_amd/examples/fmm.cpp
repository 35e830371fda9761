// Laplace FMM driver with the call sequence of exafmm-beta's examples/fmm.cxx
// (lines 13-140): same options (Args), same passes, same verification against
// a direct sum on 100 sampled bodies.  Each pass runs on the GPU through
// include/exafmm_amd.hpp, periodic images (-i) included.  Separate source sets
// (-j), PAPI, DAG output and the timing-regression database are not part of
// this path.
#include <cmath>
#include <iostream>

#include "exafmm_amd.hpp"

using namespace exafmm_amd;

int main(int argc, char **argv) {
  try {
    const vec3 cycle = 2 * M_PI;
    const real_t eps2 = 0.0;
    const complex_t wavek = complex_t(10., 1.) / real_t(2 * M_PI);
    Args args(argc, argv);
    if (args.IneJ) throw Error("-j (separate source bodies) is not supported");
    Bodies bodies, bodies2, jbodies, buffer;
    BoundBox boundBox;
    Bounds bounds;
    BuildTree buildTree(args.ncrit);
    Cells cells;
    Dataset data;
    Kernel kernel(args.P, eps2, wavek);
    Traversal traversal(kernel, args.theta, args.nspawn, args.images, args.path);
    UpDownPass upDownPass(kernel);
    Verify verify;

    verify.verbose = args.verbose;
    logger::verbose = args.verbose;
    logger::printTitle("FMM Parameters");
    args.print();
    bodies = data.initBodies(args.numBodies, args.distribution, 0);
    buffer.reserve(bodies.size());
    for (int t = 0; t < args.repeat; t++) {
      logger::printTitle("FMM Profiling");
      logger::startTimer("Total FMM");
      bounds = boundBox.getBounds(bodies);
      cells = buildTree.buildTree(bodies, buffer, bounds);
      upDownPass.upwardPass(cells);
      traversal.initListCount(cells);
      traversal.initWeight(cells);
      traversal.traverse(cells, cells, cycle, args.dual);
      jbodies = bodies;
      upDownPass.downwardPass(cells);
      logger::printTitle("Total runtime");
      logger::stopTimer("Total FMM");
      // examples/fmm.cxx:88 (a list file only with --path, as the reference
      // writes one only when built with list counting)
      if (std::string(args.path) != "./") traversal.writeList(cells, 0);

      const int numTargets = 100;
      buffer = bodies;
      data.sampleBodies(bodies, numTargets);
      bodies2 = bodies;
      data.initTarget(bodies);
      logger::startTimer("Total Direct");
      traversal.direct(bodies, jbodies, cycle);
      logger::stopTimer("Total Direct");
      const double potDif = verify.getDifScalar(bodies, bodies2);
      const double potNrm = verify.getNrmScalar(bodies);
      const double accDif = verify.getDifVector(bodies, bodies2);
      const double accNrm = verify.getNrmVector(bodies);
      const double potRel = std::sqrt(potDif / potNrm);
      const double accRel = std::sqrt(accDif / accNrm);
      logger::printTitle("FMM vs. direct");
      verify.print("Rel. L2 Error (pot)", potRel);
      verify.print("Rel. L2 Error (acc)", accRel);
      buildTree.printTreeData(cells);
      traversal.printTraversalData();
      bodies = buffer;
      data.initTarget(bodies);
    }
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}

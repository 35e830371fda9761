// Laplace FMM driver with the call sequence of exafmm-beta's examples/fmm.cxx
// (lines 13-140): same options (Args), same passes, same verification against
// a direct sum on 100 sampled bodies.  Each pass runs on the GPU through
// include/exafmm_amd.hpp, periodic images (-i) and separate source bodies
// (-j) included.  PAPI, DAG output and the timing-regression database are not
// part of this path.
//
// The call sequence is the reference's interface, kept as it is so that a
// caller of exafmm-beta runs unchanged on this drop-in; the reference's
// licence, which covers it:
//
// BSD 3-Clause License
//
// Copyright (c) 2020, Rio Yokota, Lorena Barba
// All rights reserved.
//
// Redistribution and use in source and binary forms, with or without
// modification, are permitted provided that the following conditions are met:
//
// 1. Redistributions of source code must retain the above copyright notice, this
//    list of conditions and the following disclaimer.
//
// 2. Redistributions in binary form must reproduce the above copyright notice,
//    this list of conditions and the following disclaimer in the documentation
//    and/or other materials provided with the distribution.
//
// 3. Neither the name of the copyright holder nor the names of its
//    contributors may be used to endorse or promote products derived from
//    this software without specific prior written permission.
//
// THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS"
// AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE
// IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR PURPOSE ARE
// DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT HOLDER OR CONTRIBUTORS BE LIABLE
// FOR ANY DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL
// DAMAGES (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR
// SERVICES; LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER
// CAUSED AND ON ANY THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY,
// OR TORT (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE
// OF THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
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
    Bodies bodies, bodies2, jbodies, buffer;
    BoundBox boundBox;
    Bounds bounds;
    BuildTree buildTree(args.ncrit);
    Cells cells, jcells;
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
    if (args.IneJ) {
      for (B_iter B = bodies.begin(); B != bodies.end(); B++) {
        B->X[0] += M_PI;
        B->X[0] *= 0.5;
      }
      jbodies = data.initBodies(args.numBodies, args.distribution, 1);
      for (B_iter B = jbodies.begin(); B != jbodies.end(); B++) {
        B->X[0] -= M_PI;
        B->X[0] *= 0.5;
      }
    }
    for (int t = 0; t < args.repeat; t++) {
      logger::printTitle("FMM Profiling");
      logger::startTimer("Total FMM");
      bounds = boundBox.getBounds(bodies);
      if (args.IneJ) bounds = boundBox.getBounds(jbodies, bounds);
      cells = buildTree.buildTree(bodies, buffer, bounds);
      upDownPass.upwardPass(cells);
      traversal.initListCount(cells);
      traversal.initWeight(cells);
      if (args.IneJ) {
        jcells = buildTree.buildTree(jbodies, buffer, bounds);
        upDownPass.upwardPass(jcells);
        traversal.traverse(cells, jcells, cycle, args.dual);
      } else {
        traversal.traverse(cells, cells, cycle, args.dual);
        jbodies = bodies;
      }
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

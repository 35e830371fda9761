// Operator chain of exafmm-beta's examples/kernel.cxx (lines 9-86) over the
// drop-in's per-cell Kernel methods: one unit source through P2M -> M2M ->
// M2L -> L2L -> L2P against P2P of the same pair, each operator one launch on
// the GPU.  Prints the reference's line "P potErr accErr", then the two
// target values at full precision ("fmm" and "direct" rows: potential and
// gradient) for comparison with fixtures.
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
#include <cstdio>
#include <iostream>

#include "exafmm_amd.hpp"

using namespace exafmm_amd;

int main(int argc, char **argv) {
  try {
    const real_t eps2 = 0.0;
    const complex_t wavek = complex_t(1., .1) / real_t(2 * M_PI);
    Args args(argc, argv);
    Bodies bodies(1), bodies2(1), jbodies(1);
    Kernel kernel(args.P, eps2, wavek);
    logger::verbose = true;

    Cells cells(4);
    Verify verify;
    jbodies[0].X = 2;
    jbodies[0].SRC = 1;
    C_iter Cj = cells.begin();
    Cj->X = 1;
    Cj->X[0] = 3;
    Cj->R = 1;
    Cj->BODY = jbodies.begin();
    Cj->NBODY = (int)jbodies.size();
    Cj->M.resize(kernel.NTERM, 0.0);
    kernel.P2M(Cj);

    C_iter CJ = cells.begin() + 1;
    CJ->ICHILD = (int)(Cj - cells.begin());
    CJ->NCHILD = 1;
    CJ->X = 0;
    CJ->X[0] = 4;
    CJ->R = 2;
    CJ->M.resize(kernel.NTERM, 0.0);
    kernel.M2M(CJ, cells.begin());

    C_iter CI = cells.begin() + 2;
    CI->X = 0;
    CI->X[0] = -4;
    CI->R = 2;
    CI->L.resize(kernel.NTERM, 0.0);
    kernel.M2L(CI, CJ);

    C_iter Ci = cells.begin() + 3;
    Ci->X = 1;
    Ci->X[0] = -3;
    Ci->R = 1;
    Ci->IPARENT = 2;
    Ci->L.resize(kernel.NTERM, 0.0);
    kernel.L2L(Ci, cells.begin());

    bodies[0].X = 2;
    bodies[0].X[0] = -2;
    bodies[0].SRC = 1;
    bodies[0].TRG = 0;
    Ci->BODY = bodies.begin();
    Ci->NBODY = (int)bodies.size();
    kernel.L2P(Ci);

    for (B_iter B = bodies2.begin(); B != bodies2.end(); B++) {
      *B = bodies[B - bodies2.begin()];
      B->TRG = 0;
    }
    Cj->NBODY = (int)jbodies.size();
    Ci->NBODY = (int)bodies2.size();
    Ci->BODY = bodies2.begin();
    kernel.P2P(Ci, Cj);
    for (B_iter B = bodies2.begin(); B != bodies2.end(); B++)
      for (int k = 0; k < 4; k++) B->TRG[k] /= B->SRC;

    const double potDif = verify.getDifScalar(bodies, bodies2);
    const double potNrm = verify.getNrmScalar(bodies);
    const double accDif = verify.getDifVector(bodies, bodies2);
    const double accNrm = verify.getNrmVector(bodies);
    std::cout << args.P << " " << std::sqrt(potDif / potNrm) << "  " << std::sqrt(accDif / accNrm) << std::endl;
    verify.verbose = true;
    verify.print("Rel. L2 Error (pot)", std::sqrt(potDif / potNrm));
    verify.print("Rel. L2 Error (acc)", std::sqrt(accDif / accNrm));
    std::printf("fmm %.17g %.17g %.17g %.17g\n", bodies[0].TRG[0], bodies[0].TRG[1], bodies[0].TRG[2],
                bodies[0].TRG[3]);
    std::printf("direct %.17g %.17g %.17g %.17g\n", bodies2[0].TRG[0], bodies2[0].TRG[1], bodies2[0].TRG[2],
                bodies2[0].TRG[3]);
  } catch (const std::exception &e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  return 0;
}

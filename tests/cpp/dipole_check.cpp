// UpDownPass::getDipole / dipoleCorrection of include/exafmm_amd.hpp on host
// Bodies (no device needed): reads n, then n lines "x y z q p ax ay az" on
// stdin and prints the corrected targets, one body per line.
#include <cstdio>

#include "exafmm_amd.hpp"

using namespace exafmm_amd;

int main() {
  int n = 0;
  if (scanf("%d", &n) != 1) return 1;
  Bodies bodies(n);
  for (int i = 0; i < n; i++) {
    Body &b = bodies[i];
    if (scanf("%lf %lf %lf %lf %lf %lf %lf %lf", &b.X[0], &b.X[1], &b.X[2], &b.SRC, &b.TRG[0], &b.TRG[1],
              &b.TRG[2], &b.TRG[3]) != 8)
      return 1;
  }
  Kernel kernel(4, 0.0, complex_t(0.0, 0.0));
  UpDownPass upDownPass(kernel);
  const vec3 cycle = 2 * M_PI;
  const vec3 dipole = upDownPass.getDipole(bodies, 0.0);
  upDownPass.dipoleCorrection(bodies, dipole, n, cycle);
  for (int i = 0; i < n; i++)
    printf("%.17g %.17g %.17g %.17g\n", bodies[i].TRG[0], bodies[i].TRG[1], bodies[i].TRG[2], bodies[i].TRG[3]);
  return 0;
}

// kahan<T> of the C++ layer (include/kahan.h:9-124): prints, for float and
// double, the plain sum, the compensated value s - c and the compensation c of
// a sum whose exact value the test knows, and the results of the other
// operators.
#include <cstdio>

#include "exafmm_amd.hpp"

using exafmm_amd::kahan;

template <typename T> void run(const char *name, int n, T big, T small) {
  // big + n * small: every small term is below half an ulp of big
  T plain = big;
  kahan<T> k(big);
  for (int i = 0; i < n; i++) {
    plain += small;
    k += small;
  }
  std::printf("%s %.17g %.17g %.17g %.17g\n", name, (double)plain, (double)T(k), (double)k.s, (double)k.c);
}

int main() {
  run<float>("float", 1 << 20, 1.0f, 1e-8f);
  run<double>("double", 1 << 20, 1.0, 1e-17);
  // the other operators on a compensated value: -=, *=, /=, += kahan, negation
  kahan<double> a(1.0), b(0.0);
  for (int i = 0; i < 1000; i++) a += 1e-17;
  for (int i = 0; i < 1000; i++) b -= 1e-17;
  kahan<double> s = a;
  s += b;                    // 1 + 1e-14 - 1e-14
  kahan<double> m = a;
  m *= 2.0;
  m /= 4.0;                  // (1 + 1e-14) / 2
  kahan<double> ng = -a;
  std::printf("ops %.17g %.17g %.17g %.17g\n", (double)s, (double)m, (double)ng, (double)b);
  return 0;
}

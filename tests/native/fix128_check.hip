// Host check of the exact accumulator of the force kernels (Fix128 in
// crowddynamics_amd/csrc/cd_device.h): exact round trips, independence of
// the order of the terms, accuracy of sums, NaN for out-of-range terms.
// Prints "ok" and exits 0 on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "../../crowddynamics_amd/csrc/cd_device.h"

int main() {
  std::mt19937_64 g(1);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  int fails = 0;
  for (int k = 0; k < 200000; ++k) {  // |x| in [2^-11, 2^50]: the conversion is exact
    const double x = u(g) * std::ldexp(1.0, (int)(g() % 61) - 10);
    if (std::fabs(x) < std::ldexp(1.0, -11)) continue;
    cd::Fix128 a;
    a.add(x);
    if (a.value() != x && fails++ < 5) std::printf("round trip %.17g -> %.17g\n", x, a.value());
  }
  for (int trial = 0; trial < 20000; ++trial) {
    const int n = 1 + (int)(g() % 60);
    std::vector<double> xs(n);
    for (double &x : xs) x = u(g) * std::pow(10.0, (int)(g() % 10) - 4);
    cd::Fix128 a, b;
    for (double x : xs) a.add(x);
    std::shuffle(xs.begin(), xs.end(), g);
    for (double x : xs) b.add(x);
    if ((a.hi != b.hi || a.lo != b.lo) && fails++ < 5) std::printf("order dependence\n");
    long double exact = 0.0L, scale = 0.0L;
    for (double x : xs) {
      exact += x;
      scale += std::fabs(x);
    }
    if (std::fabs((long double)a.value() - exact) > 1e-15L * scale + 1e-18L && fails++ < 5)
      std::printf("sum error %Lg\n", (long double)a.value() - exact);
  }
  cd::Fix128 c;
  c.add(-0.5);
  c.add(1e30);
  if (!std::isnan(c.value()) && fails++ < 5) std::printf("out of range not flagged\n");
  cd::Fix128 d;
  d.add(-0.5);
  d.add(0.25);
  d.add(-3.0);
  if (d.value() != -3.25 && fails++ < 5) std::printf("small sum %.17g\n", d.value());
  std::printf(fails ? "FAILED %d\n" : "ok\n", fails);
  return fails != 0;
}

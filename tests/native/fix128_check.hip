// Host check of the exact accumulator of the force kernels (Fix128 in
// crowddynamics_amd/csrc/cd_device.h): exact round trips, independence of
// the order of the terms, accuracy of sums, NaN for out-of-range terms, and
// the floating-point conversions (double and float terms below 2^32) adding
// exactly what the bit-level conversion adds.
// Prints "ok" and exits 0 on success.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "../../crowddynamics_amd/csrc/cd_device.h"

// trunc(|x| 2^64) with x's sign as a 128-bit two's-complement pair, by
// exact floating-point splitting (independent of the bit-level conversion
// of Fix128::add): |x| = f + r, f integer, r in [0, 1).
static void split_ref(double x, long long &hi, unsigned long long &lo) {
  const double a = std::fabs(x), f = std::floor(a);
  const double t = (a - f) * 4294967296.0, th = std::floor(t);
  unsigned long long l = ((unsigned long long)(unsigned int)th << 32) |
                         (unsigned long long)(unsigned int)((t - th) * 4294967296.0);
  const double fh = std::floor(f * 2.3283064365386962890625e-10);
  long long h = (long long)(unsigned int)fh * 4294967296LL + (long long)(unsigned int)(f - fh * 4294967296.0);
  if (x < 0.0) {
    h = ~h + (l == 0ull ? 1 : 0);
    l = ~l + 1ull;
  }
  hi = h;
  lo = l;
}

int main() {
  std::mt19937_64 g(1);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  int fails = 0;
  // the conversion of one term equals the floor-based split over the whole
  // exponent range below 2^62 (subnormals, tiny, negative, integers)
  for (int k = 0; k < 400000; ++k) {
    double x;
    if (k < 2000) {
      const double special[] = {0.0, -0.0, 1.0, -1.0, 0.5, 4294967296.0, -4294967296.0,
                                std::ldexp(1.0, -64), std::ldexp(1.0, -65), std::ldexp(-1.0, 61),
                                std::ldexp(1.0, 62) * (1.0 - std::ldexp(1.0, -53)), 4.9e-324};
      x = special[k % 12];
    } else {
      x = u(g) * std::ldexp(1.0, (int)(g() % 1140) - 1075);
    }
    if (!(std::fabs(x) < std::ldexp(1.0, 62))) continue;
    cd::Fix128 a, b;
    a.add(x);
    b.add_bits(x);  // the bit-level conversion, over the whole range
    long long h;
    unsigned long long l;
    split_ref(x, h, l);
    if ((a.hi != h || a.lo != l) && fails++ < 5)
      std::printf("conversion %.17g: %lld %llu vs %lld %llu\n", x, a.hi, a.lo, h, l);
    if ((b.hi != h || b.lo != l) && fails++ < 5)
      std::printf("bit conversion %.17g: %lld %llu vs %lld %llu\n", x, b.hi, b.lo, h, l);
  }
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
  // a float term (the fp32 mode's path) adds exactly what the same value as
  // a double adds: random bit patterns over every float exponent, both
  // signs, subnormals, the 2^32 switch to the double path, inf and NaN
  for (int k = 0; k < 2000000; ++k) {
    float x;
    if (k < 64) {
      const float special[] = {0.0f, -0.0f, 1.0f, -1.0f, 0x1p32f, -0x1p32f, 0x1.fffffep31f,
                               -0x1.fffffep31f, 0x1p-64f, -0x1p-65f, 0x1p-149f, -0x1p-149f,
                               0x1.fffffep127f, INFINITY, -INFINITY, NAN};
      x = special[k % 16];
    } else {
      const unsigned bits = (unsigned)g();
      std::memcpy(&x, &bits, sizeof(x));
    }
    cd::Fix128 a, b;
    a.add(x);
    b.add_bits((double)x);
    if ((a.hi != b.hi || a.lo != b.lo || a.bad != b.bad) && fails++ < 5)
      std::printf("float term %a: %lld %llu %d vs %lld %llu %d\n", (double)x, a.hi, a.lo,
                  (int)a.bad, b.hi, b.lo, (int)b.bad);
    a.add(-x * 0.75f);  // and as the second term of a sum
    b.add_bits((double)(-x * 0.75f));
    if ((a.hi != b.hi || a.lo != b.lo || a.bad != b.bad) && fails++ < 5)
      std::printf("float sum %a\n", (double)x);
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

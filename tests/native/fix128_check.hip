// Host check of the exact accumulator of the force kernels (Fix128 in
// crowddynamics_amd/csrc/cd_device.h): exact round trips, independence of
// the order of the terms, accuracy of sums, NaN for out-of-range terms, the
// floating-point conversion of double terms below 2^32 adding exactly what
// the bit-level conversion adds, and float terms below 2^20 adding their
// value rounded to the 2^-31 grid.
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
  // a float term (the fp32 mode's path) below 2^20 adds round(x 2^31) 2^-31,
  // ties to even (x 2^31 is exact in double, nearbyint rounds it in the
  // default mode); larger terms, inf and NaN add what add_bits adds: random
  // bit patterns over every float exponent, both signs, subnormals, the
  // switch at 2^20, ties, inf and NaN
  auto float_ref = [](float x, long long &h, unsigned long long &l, bool &bad) {
    if (!(std::fabs(x) < 0x1p20f)) {
      cd::Fix128 r;
      r.add_bits((double)x);
      h = r.hi;
      l = r.lo;
      bad = r.bad;
      return;
    }
    const long long t = (long long)std::nearbyint((double)x * 0x1p31);
    h = t >> 31;
    l = (unsigned long long)t << 33;
    bad = false;
  };
  for (int k = 0; k < 2000000; ++k) {
    float x;
    if (k < 64) {
      const float special[] = {0.0f, -0.0f, 1.0f, -1.0f, 0x1p20f, -0x1p20f, 0x1.fffffep19f,
                               -0x1.fffffep19f, 0x1p-32f, -0x1p-32f, 0x3p-32f, -0x3p-32f,
                               0x1p-149f, 0x1.fffffep127f, INFINITY, NAN};
      x = special[k % 16];
    } else {
      const unsigned bits = (unsigned)g();
      std::memcpy(&x, &bits, sizeof(x));
    }
    cd::Fix128 a;
    a.add(x);
    long long h;
    unsigned long long l;
    bool bad;
    float_ref(x, h, l, bad);
    if ((a.hi != h || a.lo != l || a.bad != bad) && fails++ < 5)
      std::printf("float term %a: %lld %llu %d vs %lld %llu %d\n", (double)x, a.hi, a.lo,
                  (int)a.bad, h, l, (int)bad);
  }
  // float sums: independent of the order of the terms, and within half a
  // grid step per term of the exact sum
  for (int trial = 0; trial < 20000; ++trial) {
    const int n = 1 + (int)(g() % 60);
    std::vector<float> xs(n);
    for (float &x : xs) x = (float)(u(g) * std::pow(10.0, (int)(g() % 10) - 5));
    cd::Fix128 a, b;
    for (float x : xs) a.add(x);
    std::shuffle(xs.begin(), xs.end(), g);
    for (float x : xs) b.add(x);
    if ((a.hi != b.hi || a.lo != b.lo) && fails++ < 5) std::printf("float order dependence\n");
    long double exact = 0.0L;
    for (float x : xs) exact += x;
    if (std::fabs((long double)a.value() - exact) > n * 0x1p-32L + 1e-16L * std::fabs(exact) &&
        fails++ < 5)
      std::printf("float sum error %Lg\n", (long double)a.value() - exact);
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

// Device-side physics of one pair / one wall / one agent, shared by the
// kernels.  Every expression keeps the operation order of the reference
// function it cites (paths relative to the reference repository root); the
// library is built with -ffp-contract=off so a*b+c is not fused, matching the
// reference's numba build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace cd {

constexpr double kSightSoc = 3.0;  // core/interactions.py:45
constexpr double kFSocMax = 2e3;   // core/motion/power_law.py:52
constexpr double kTauMax = 30.0;   // core/motion/power_law.py:243
constexpr double kPi = 3.14159265358979323846;

template <typename T>
struct __attribute__((aligned(2 * sizeof(T)))) V2 {
  T x, y;
};

// Arithmetic of the pair terms by precision.  fp64 (and the dual numbers of
// the adjoint): IEEE operations in the reference's order.  fp32 mode: the
// hardware's approximate reciprocal, square root and exponential (a few
// ulp each), well inside the mode's stated error bound against the fp64
// oracle.
// hypot as the reference's C library computes it (np.hypot, the `length`
// of core/vector2D.py:113, reaches the C library's hypot): sqrt(ax^2 + ay^2)
// followed by one correction step from C. F. Borges, "An Improved Algorithm
// for hypot_ref(a, b)" (2019), in its variant without fused multiply-add (the
// library is built with -ffp-contract=off), inputs scaled by powers of two
// where squaring them would overflow or underflow.  That library hypot is
// faithful but not correctly rounded (about one result in 400 is the other
// neighbour of the exact value); this sequence returns the same bits, so
// distances that tie exactly (those of equally oriented three-circle agents)
// resolve as in the reference, and dt's max |v| is the reference's.
__device__ __forceinline__ double hypot_kernel(double ax, double ay) {
  const double h = sqrt(ax * ax + ay * ay);
  double t1, t2;
  if (h <= 2.0 * ay) {
    const double delta = h - ay;
    t1 = ax * (2.0 * delta - ax);
    t2 = (delta - 2.0 * (ax - ay)) * delta;
  } else {
    const double delta = h - ax;
    t1 = 2.0 * delta * (ax - 2.0 * ay);
    t2 = (4.0 * delta - ay) * ay + delta * delta;
  }
  return h - (t1 + t2) / (2.0 * h);
}

__device__ __forceinline__ double hypot_ref(double x, double y) {
  if (!__builtin_isfinite(x) || !__builtin_isfinite(y)) {
    if (__builtin_isinf(x) || __builtin_isinf(y)) return __builtin_inf();
    return x + y;  // NaN
  }
  x = fabs(x);
  y = fabs(y);
  const double ax = x < y ? y : x, ay = x < y ? x : y;
  constexpr double kScale = 0x1p-600, kLarge = 0x1p+511, kTiny = 0x1p-511, kEps = 0x1p-54;
  if (ax > kLarge) {
    if (ay <= ax * kEps) return ax + ay;
    return hypot_kernel(ax * kScale, ay * kScale) / kScale;
  }
  if (ay < kTiny) {
    if (ax >= ay / kEps) return ax + ay;
    return hypot_kernel(ax / kScale, ay / kScale) * kScale;
  }
  if (ay <= ax * kEps) return ax + ay;
  return hypot_kernel(ax, ay);
}

// The same, out of line: the pair functions reach it only on rare paths,
// and a call keeps its registers out of their hot loops.
__device__ __attribute__((noinline)) double hypot_exact(double x, double y) {
  return hypot_ref(x, y);
}

__device__ __forceinline__ double dhypot(double a, double b) { return hypot_ref(a, b); }
__device__ __forceinline__ float dhypot(float a, float b) { return __builtin_amdgcn_sqrtf(a * a + b * b); }
__device__ __forceinline__ double dsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float dsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double dexp(double a) { return exp(a); }
__device__ __forceinline__ float dexp(float a) { return __expf(a); }

// 1/b by the hardware reciprocal (v_rcp_f32, 1 ulp).  The instruction does
// not keep subnormal inputs or results, so |b| outside [2^-64, 2^64] is
// scaled by 2^64 or 2^-64 on the way in and again on the way out (exact
// power-of-two scalings).  IEEE's special cases hold: 1/(+-0) = +-inf,
// 1/inf = 0, 1/NaN = NaN, and 1/b overflows to inf where IEEE's does.
__device__ __forceinline__ float frcp(float b) {
  const float ab = __builtin_fabsf(b);
  const float s = ab > 0x1p64f ? 0x1p-64f : (ab < 0x1p-64f ? 0x1p64f : 1.0f);
  return s * __builtin_amdgcn_rcpf(b * s);
}

// 1/b by the bare hardware reciprocal, for a b known to be a normal float
// well inside the exponent range, where frcp's scale factor is exactly 1 and
// the two give the same bits.
__device__ __forceinline__ float frcp_normal(float b) { return __builtin_amdgcn_rcpf(b); }

template <typename T>
__device__ __forceinline__ T ddiv(T a, T b) {
  return a / b;
}
template <>
__device__ __forceinline__ float ddiv<float>(float a, float b) {
  return a * frcp(b);
}
__device__ __forceinline__ bool is_nan(double a) { return __builtin_isnan(a); }
__device__ __forceinline__ bool is_nan(float a) { return __builtin_isnan(a); }

// Ordered-integer image of a double: unsigned comparison of the images
// equals floating-point comparison (for non-NaN values).
__device__ __forceinline__ unsigned long long ord_bits(double v) {
  unsigned long long b = __double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double from_ord_bits(unsigned long long b) {
  b = (b & 0x8000000000000000ull) ? (b & 0x7fffffffffffffffull) : ~b;
  return __longlong_as_double(b);
}

// Python float remainder a % b for b > 0 (numba semantics), then
// wrap_to_pi, core/vector2D.py:8-36.
__device__ __forceinline__ double wrap_to_pi(double rad) {
  const double two_pi = 2.0 * kPi;
  // fmod(x, y) is x itself when |x| < |y| (the common case: an angle
  // already near [-pi, pi]), so the library call is needed only beyond
  double r = __builtin_fabs(rad) < two_pi ? rad : fmod(rad, two_pi);
  if (r != 0.0) {
    if (r < 0.0) r += two_pi;
  } else {
    r = 0.0;
  }
  if (rad < 0.0 && r == kPi) return -kPi;
  if (r > kPi) return r - two_pi;
  return r;
}

// force_contact, core/motion/contact.py:12-48
template <typename T>
__device__ __forceinline__ void force_contact(T h, T nx, T ny, T vx, T vy, T tx, T ty, T mu,
                                              T kappa, T damping, T &ox, T &oy) {
  T kvt = kappa * (vx * tx + vy * ty);
  T cvn = damping * (vx * nx + vy * ny);
  ox = (-h) * (mu * nx - kvt * tx) + cvn * nx;
  oy = (-h) * (mu * ny - kvt * ty) + cvn * ny;
}

// magnitude, core/motion/power_law.py:83-102
template <typename T>
__device__ __forceinline__ T magnitude(T tau, T tau_0) {
  return ddiv((ddiv(T(2), tau) + ddiv(T(1), tau_0)) * dexp(ddiv(-tau, tau_0)), tau * tau);
}

// Relative margin of the squared-length gates: x^2 + y^2 evaluated in
// double is within 2^-52 of the exact value, so a gate 2^-40 away from a
// threshold decides the same way as the exact length would.
constexpr double kGate = 0x1p-40;

// truncate, core/vector2D.py:167-187.  In double, a vector whose squared
// length is below l^2 (1 - 2^-40) has hypot(x, y) < l and stays as it is;
// only the others take the exact length.  In float likewise below
// l^2 (1 - 2^-16): the float length (the hardware square root of the
// float squared length, within 2^-22 of |(x, y)|) is then below l, so the
// gate skips the square root without changing a result.
template <typename T>
__device__ __forceinline__ void truncate(T &x, T &y, T l) {
  if constexpr (std::is_same<T, double>::value) {
    if (x * x + y * y < (l * l) * (1.0 - kGate)) return;
  } else if constexpr (std::is_same<T, float>::value) {
    if (x * x + y * y < (l * l) * (1.0f - 0x1p-16f)) return;
  }
  T vlen = dhypot(x, y);
  if (vlen > l) {
    T s = ddiv(l, vlen);
    x *= s;
    y *= s;
  }
}

// Exact, order-independent sum of doubles: a 128-bit two's-complement
// fixed-point number with 64 fractional bits (range +-2^63, resolution
// 2^-64).  Every double term is converted exactly when its magnitude is at
// least 2^-11 (and truncated to the 2^-64 grid below that); a float term
// below 2^20 is rounded to the 2^-31 grid.  Each term thus maps to one
// fixed-point value of its own, and additions are exact integer additions,
// so the sum does not depend on the order of the terms.  Terms that are not
// finite or of magnitude 2^62 or more make the sum NaN.
struct Fix128 {
  long long hi;
  unsigned long long lo;
  bool bad;
  __host__ __device__ __forceinline__ Fix128() : hi(0), lo(0ull), bad(false) {}
  // Adds trunc(|x| 2^64) with x's sign.  Below 2^32 by exact
  // floating-point steps: the integer part, then the upper and lower 32 bits
  // of the fraction, each a truncation, the subtraction of a value's own
  // integer part and a scaling by 2^32 (all exact in the value's own
  // precision).  Otherwise by add_bits.
  __host__ __device__ __forceinline__ void add(double x) {
    const double a = __builtin_fabs(x);
    if (!(a < 0x1p32)) {
      add_bits(x);
      return;
    }
    const double t2 = __builtin_trunc(a);
    const double g1 = (a - t2) * 0x1p32;
    const double t1 = __builtin_trunc(g1);
    add_mag((unsigned)t2, (unsigned)t1, (unsigned)((g1 - t1) * 0x1p32), x < 0.0);
  }
  // A float term (the fp32 mode's pair and wall terms) below 2^20: rounded
  // to the nearest point of the 2^-31 grid, ties to even, by one double
  // addition.  x + 1.5 2^21 lies in [2^21, 2^22), where doubles are spaced
  // 2^-31, so its bits minus those of 1.5 2^21 are the grid index
  // t = round(x 2^31) in two's complement (|t| < 2^51); t 2^33 joins the
  // sum.  The grid is far below the float term's own rounding (2^-24 of
  // its magnitude) and costs a third of the exact conversion.  Larger
  // terms, inf and NaN take add_bits.
  __host__ __device__ __forceinline__ void add(float x) {
    if (!(__builtin_fabsf(x) < 0x1p20f)) {
      add_bits((double)x);
      return;
    }
    add_grid31(grid31(x));
  }
  // round(x 2^31) for |x| < 2^20 (the grid index above)
  __host__ __device__ static __forceinline__ long long grid31(float x) {
    const double y = (double)x + 0x1.8p21;
    return __builtin_bit_cast(long long, y) - __builtin_bit_cast(long long, 0x1.8p21);
  }
  // Adds t 2^-31 for any 64-bit t: t 2^33 on the 2^-64 grid, whose high
  // word is t >> 31 (an arithmetic shift) and low word t << 33.
  __host__ __device__ __forceinline__ void add_grid31(long long t) {
    const unsigned long long l = (unsigned long long)t << 33;
    lo += l;
    hi += (t >> 31) + (lo < l ? 1 : 0);
  }
  // The bit-level conversion, valid for every double: the mantissa shifted
  // into place (|x| = m 2^(e - 1075) for biased exponent e, so |x| 2^64 =
  // m 2^s with s = e - 1011; subnormals use e = 1).  |x| >= 2^62, inf and
  // NaN set bad.
  __host__ __device__ __forceinline__ void add_bits(double x) {
    const unsigned long long bits = __builtin_bit_cast(unsigned long long, x);
    const int e = (int)((bits >> 52) & 0x7ffull);
    if (e >= 1085) {  // 2^62 = 2^(1085 - 1023)
      bad = true;
      return;
    }
    const unsigned long long m = (bits & 0xFFFFFFFFFFFFFull) | (e ? (1ull << 52) : 0ull);
    const int s = (e ? e : 1) - 1011;  // <= 73
    unsigned long long l, h;
    if (s >= 64) {
      l = 0ull;
      h = m << (s - 64);
    } else if (s > 0) {
      l = m << s;
      h = m >> (64 - s);
    } else if (s > -64) {
      l = m >> (-s);
      h = 0ull;
    } else {
      l = 0ull;
      h = 0ull;
    }
    add_twos(h, l, (bits >> 63) != 0);
  }
  // Adds the magnitude with 32-bit words w2 (2^0), w1 (2^-32), w0 (2^-64),
  // negated when neg.
  __host__ __device__ __forceinline__ void add_mag(unsigned w2, unsigned w1, unsigned w0,
                                                   bool neg) {
    add_twos(w2, ((unsigned long long)w1 << 32) | w0, neg);
  }
  __host__ __device__ __forceinline__ void add_twos(unsigned long long h, unsigned long long l,
                                                    bool neg) {
    if (neg) {  // two's complement of (h, l)
      h = ~h + (l == 0ull ? 1ull : 0ull);
      l = ~l + 1ull;
    }
    lo += l;
    hi += (long long)h + (lo < l ? 1 : 0);
  }
  __host__ __device__ __forceinline__ void add(const Fix128 &o) {
    lo += o.lo;
    hi += o.hi + (lo < o.lo ? 1 : 0);
    bad = bad || o.bad;
  }
  // The sum, rounded to a double: magnitude first, so small negative sums
  // keep their low bits.
  __host__ __device__ __forceinline__ double value() const {
    if (bad) return __builtin_nan("");
    long long h = hi;
    unsigned long long l = lo;
    const bool neg = h < 0;
    if (neg) {
      h = ~h + (l == 0ull ? 1 : 0);
      l = ~l + 1ull;
    }
    const double v = (double)h + (double)l * 5.421010862427522170037264e-20;  // 2^-64
    return neg ? -v : v;
  }
};

// Per-agent constants the owner side of a pair needs.
template <typename T>
struct Owner {
  T neg_mk;  // (-mass) * k_soc, power_law.py:250
  T tau_0, mu, kappa, damping;
  T rtau_0;  // fp32 mode: frcp_normal(tau_0), once per agent (else unused)
};

// interaction_agent_agent_circular (core/interactions.py:51-70) seen from
// agent i: distance_circles (core/distance.py:16-47), force_social_circular
// (core/motion/power_law.py:213-259) and force_contact, with x = x_i - x_j,
// v = v_i - v_j and i's constants.  Returns false when h >= SIGHT_SOC (the
// pair adds nothing).  The j-side of the reference pair equals this function
// evaluated with i and j swapped bit for bit (every term is odd in (x, v)
// and IEEE negation is exact), so owner-computes reproduces both sides.
//
// In fp64 the bits of the distance d (the library hypot) matter only where
// they decide a branch: h = d - r_tot enters the result through the tests
// h < 3 and h < 0 and, in contact, the contact force, which is continuous
// in d.  s = x^2 + y^2 is within 2^-52 of d^2 and the hypot within 2^-52 of
// d, so s below (3 + r_tot)^2 (1 - 2^-40) and above r_tot^2 (1 + 2^-40)
// decides both tests as d would (h in (0, 3)) and the pair skips d; s above
// (3 + r_tot)^2 (1 + 2^-40) rejects it; s below r_tot^2 (1 - 2^-40) is surely
// in contact, and d = sqrt(s) (correctly rounded, within one ulp of the
// hypot) feeds the contact force.  Only the pairs within 2^-40 of a
// threshold take the library hypot.

template <typename T>
__device__ __forceinline__ bool pair_circular(T x, T y, T vx, T vy, T r_tot, const Owner<T> &o,
                                              T &fx, T &fy) {
  T d = T(0), h = T(0);
  bool contact_possible = true;
  if constexpr (std::is_same<T, double>::value) {
    const double s = x * x + y * y, rs = kSightSoc + r_tot, r2 = r_tot * r_tot;
    if (s > (rs * rs) * (1.0 + kGate)) return false;
    const bool clear = s < (rs * rs) * (1.0 - kGate) && s > r2 * (1.0 + kGate);
    contact_possible = !clear;
    if (contact_possible) {
      d = (s < r2 * (1.0 - kGate)) ? sqrt(s) : hypot_exact(x, y);
      h = d - r_tot;
      if (!(h < kSightSoc)) return false;
    }
  } else {
    d = dhypot(x, y);
    h = d - r_tot;
    if (!(h < T(kSightSoc))) return false;
  }
  T sx = T(0), sy = T(0);
  T a = vx * vx + vy * vy;
  T b = -(x * vx + y * vy);
  T c = (x * x + y * y) - r_tot * r_tot;
  T ds = dsqrt(b * b - a * c);
  if (!(is_nan(ds) || ds == T(0) || a == T(0))) {
    T tau = ddiv(b - ds, a);
    if (!(tau <= T(0) || tau > T(kTauMax))) {
      T gx = ddiv(vx - ddiv(vx * b + x * a, ds), a);
      T gy = ddiv(vy - ddiv(vy * b + y * a, ds), a);
      T m = magnitude(tau, o.tau_0);
      sx = T(0) + (o.neg_mk * gx) * m;
      sy = T(0) + (o.neg_mk * gy) * m;
      truncate(sx, sy, T(kFSocMax));
    }
  }
  if (contact_possible && h < T(0)) {
    T nx = T(0), ny = T(0);
    if (d != T(0)) {
      nx = ddiv(x, d);
      ny = ddiv(y, d);
    }
    T cx, cy;
    force_contact(h, nx, ny, vx, vy, ny, -nx, o.mu, o.kappa, o.damping, cx, cy);
    sx += cx;
    sy += cy;
  }
  fx = sx;
  fy = sy;
  return true;
}

// fp32 mode of the same pair: the same terms and the same sight and tau
// gates on the same distance, with the divisions by a, ds, tau and tau_0
// taken as products with one shared reciprocal each, so a pair costs three
// reciprocals instead of eleven divisions.  Each such product is within
// 2 ulp of the quotient it replaces.  Only a = |v|^2 can leave the normal
// range (|v| below 2^-32 m/s), so only it takes the scaled frcp; ds and d are
// square roots of nonzero floats (at least 2^-75), tau_0 is a positive
// parameter, and a tau below 2^-64 makes 1/tau^2 overflow fp32 however the
// quotient is formed.  Sums of products are fused multiply-adds written
// out (one rounding instead of two, the same in every kernel that inlines
// the pair); each keeps its parity in (x, v), so the pair stays odd.
template <>
__device__ __forceinline__ bool pair_circular<float>(float x, float y, float vx, float vy,
                                                     float r_tot, const Owner<float> &o,
                                                     float &fx, float &fy) {
  const float d2 = __builtin_fmaf(x, x, y * y);
  const float d = __builtin_amdgcn_sqrtf(d2);
  const float h = d - r_tot;
  if (!(h < float(kSightSoc))) return false;
  float sx = 0.0f, sy = 0.0f;
  const float a = __builtin_fmaf(vx, vx, vy * vy);
  const float b = -__builtin_fmaf(x, vx, y * vy);
  const float c = __builtin_fmaf(-r_tot, r_tot, d2);
  const float ds = __builtin_amdgcn_sqrtf(__builtin_fmaf(b, b, -(a * c)));
  if (!(is_nan(ds) || ds == 0.0f || a == 0.0f)) {
    const float ra = frcp(a);
    const float tau = (b - ds) * ra;
    if (!(tau <= 0.0f || tau > float(kTauMax))) {
      const float rds = frcp_normal(ds);
      const float gx = __builtin_fmaf(-__builtin_fmaf(vx, b, x * a), rds, vx) * ra;
      const float gy = __builtin_fmaf(-__builtin_fmaf(vy, b, y * a), rds, vy) * ra;
      const float rt = frcp_normal(tau), rt0 = o.rtau_0;
      // magnitude(tau, tau_0) = (2/tau + 1/tau_0) exp(-tau/tau_0) / tau^2
      const float m = (__builtin_fmaf(2.0f, rt, rt0) * __expf(-tau * rt0)) * (rt * rt);
      sx = (o.neg_mk * gx) * m;
      sy = (o.neg_mk * gy) * m;
      truncate(sx, sy, float(kFSocMax));
    }
  }
  if (h < 0.0f) {
    float nx = 0.0f, ny = 0.0f;
    if (d != 0.0f) {
      const float rd = frcp_normal(d);
      nx = x * rd;
      ny = y * rd;
    }
    float cx, cy;
    force_contact(h, nx, ny, vx, vy, ny, -nx, o.mu, o.kappa, o.damping, cx, cy);
    sx += cx;
    sy += cy;
  }
  fx = sx;
  fy = sy;
  return true;
}

// np.sign
__device__ __forceinline__ double np_sign(double v) {
  if (v > 0.0) return 1.0;
  if (v < 0.0) return -1.0;
  if (v == 0.0) return 0.0;
  return v;
}

// Per-segment part of distance_circle_line: unit tangent (tx, ty) and
// length l_w of the segment p0 -> p1 (returned as {tx, ty, l_w, 0}).
__device__ __forceinline__ double4 seg_frame(double p0x, double p0y, double p1x, double p1y) {
  const double dx = p1x - p0x, dy = p1y - p0y;
  const double l_w = hypot_ref(dx, dy);
  return make_double4(dx / l_w, dy / l_w, l_w, 0.0);
}

// distance_circle_line, core/distance.py:108-146, given the segment's frame
// fr = seg_frame(p0, p1) (the same operations on the same values as
// computing it here).  Absolute coordinates in fp64 (positions are fp64 in
// both precision modes).
__device__ __forceinline__ void distance_circle_line(double x, double y, double r, double p0x,
                                                     double p0y, double p1x, double p1y,
                                                     const double4 &fr, double &h, double &nx,
                                                     double &ny) {
  const double tx = fr.x, ty = fr.y, l_w = fr.z;
  double nwx = -ty, nwy = tx;
  double q0x = x - p0x, q0y = y - p0y;
  double q1x = x - p1x, q1y = y - p1y;
  double l_t = -(tx * q1x + ty * q1y) - (tx * q0x + ty * q0y);
  double d_iw;
  if (l_t > l_w) {
    d_iw = hypot_ref(q0x, q0y);
    nx = q0x / d_iw;
    ny = q0y / d_iw;
  } else if (l_t < -l_w) {
    d_iw = hypot_ref(q1x, q1y);
    nx = q1x / d_iw;
    ny = q1y / d_iw;
  } else {
    double l_n = nwx * q0x + nwy * q0y;
    d_iw = fabs(l_n);
    double s = np_sign(l_n);
    nx = s * nwx;
    ny = s * nwy;
  }
  h = d_iw - r;
}

// One agent of a three-circle pair: part centres (torso, left, right),
// velocity, radii and orientation.
struct TCView {
  double px[3], py[3];
  double vx, vy;
  double r_t, r_s;
  double ox, oy;  // shoulder_offset(r_ts, phi)
};

// The offset of the left shoulder from the torso, (r_ts sin(phi), r_ts
// (-cos(phi))), as force_social_three_circle forms it (core/motion/
// power_law.py:338-350); a per-agent value, computed once per agent and step.
__device__ __forceinline__ double2 shoulder_offset(double r_ts, double phi) {
  return make_double2(r_ts * sin(phi), r_ts * (-cos(phi)));
}

// interaction_agent_agent_three_circle (core/interactions.py:73-104) in the
// reference's argument order (A = agent i, B = agent j), returning only the
// side `side_b ? j : i`: its force and torque.  Uses distance_three_circles
// (core/distance.py:50-105, including r_moment1 built from x0[j_min] at :103)
// and force_social_three_circle (core/motion/power_law.py:262-363, including
// the first-valid-tau selection at :324).  `o` holds the constants of the
// returned side.
__device__ __forceinline__ bool pair_three_circle(const TCView &A, const TCView &B, bool side_b,
                                                  const Owner<double> &o, double &fx, double &fy,
                                                  double &torque) {
  const double rA[3] = {A.r_t, A.r_s, A.r_s};
  const double rB[3] = {B.r_t, B.r_s, B.r_s};
  // distance_three_circles; the parts chosen by the minimum are recorded as
  // they are selected (no dynamic indexing of the part arrays).
  // The minimum's part pair needs the exact distances (the library hypot)
  // only of the part pairs that can still become the minimum.  The
  // hardware square root of x^2 + y^2 is within 2^-24 of the distance
  // (measured: at most 1.7e8 ulp), so a part pair whose approximate h
  // exceeds the exact minimum so far by more than 2^-20 (d + 1) has a
  // larger exact h and is not selected (the selection takes a strictly
  // smaller h); it skips its hypot.  The first part pair, and any NaN,
  // take the exact path.
  double h_min = __builtin_nan("");
  double minx = 0.0, miny = 0.0, mind = 0.0;  // x, y and d of the minimum
  double selx = 0.0, sely = 0.0, selr = 0.0;  // A.part[i_min], rA[i_min]
  double bugx = 0.0, bugy = 0.0, bugr = 0.0;  // A.part[j_min], rB[j_min]
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double x = A.px[i] - B.px[j], y = A.py[i] - B.py[j];
      {
        const double d0 = __builtin_amdgcn_sqrt(x * x + y * y);
        if (d0 - (rA[i] + rB[j]) - 0x1p-20 * (d0 + 1.0) > h_min) continue;
      }
      double d = hypot_ref(x, y);
      double h = d - (rA[i] + rB[j]);
      if (h < h_min || __builtin_isnan(h_min)) {
        h_min = h;
        minx = x;
        miny = y;
        mind = d;
        selx = A.px[i];
        sely = A.py[i];
        selr = rA[i];
        bugx = A.px[j];
        bugy = A.py[j];
        bugr = rB[j];
      }
    }
  }
  // the normal of the minimum only (the reference divides every part pair's
  // x, y by its d and keeps the minimum's: the same quotients)
  double nx = 0.0, ny = 0.0;
  if (mind != 0.0) {
    nx = minx / mind;
    ny = miny / mind;
  }
  if (!(h_min < kSightSoc)) return false;
  double rmx, rmy;
  if (!side_b) {
    rmx = (selx + selr * nx) - A.px[0];
    rmy = (sely + selr * ny) - A.py[0];
  } else {
    rmx = (bugx - bugr * nx) - B.px[0];
    rmy = (bugy - bugr * ny) - B.py[0];
  }
  // force_social_three_circle
  double sx = 0.0, sy = 0.0;
  double vx = A.vx - B.vx, vy = A.vy - B.vy;
  double a = vx * vx + vy * vy;
  if (a != 0.0) {
    int ci = 0, cj = 0;
    double tau = __builtin_nan(""), b_min = __builtin_nan(""), d_min = __builtin_nan("");
#pragma unroll
    for (int pi = 0; pi < 3; ++pi) {
#pragma unroll
      for (int pj = 0; pj < 3; ++pj) {
        double x = A.px[pi] - B.px[pj], y = A.py[pi] - B.py[pj];
        double r_tot = rA[pi] + rB[pj];
        double b = -(x * vx + y * vy);
        double c = (x * x + y * y) - r_tot * r_tot;
        double d = sqrt(b * b - a * c);
        if (__builtin_isnan(d) || d == 0.0) continue;
        double tau_new = (b - d) / a;
        if (__builtin_isnan(tau) || (0.0 < tau_new && tau_new < tau)) {
          ci = pi;
          cj = pj;
          tau = tau_new;
          b_min = b;
          d_min = d;
        }
      }
    }
    if (!(__builtin_isnan(tau) || tau <= 0.0)) {
      double roix = 0.0, roiy = 0.0, rojx = 0.0, rojy = 0.0;
      if (ci == 1 || ci == 2) {
        const double ex = A.ox, ey = A.oy;
        if (ci == 1) {
          roix += ex;
          roiy += ey;
        } else {
          roix -= ex;
          roiy -= ey;
        }
      }
      if (cj == 1 || cj == 2) {
        const double ex = B.ox, ey = B.oy;
        if (cj == 1) {
          rojx += ex;
          rojy += ey;
        } else {
          rojx -= ex;
          rojy -= ey;
        }
      }
      double x = A.px[0] - B.px[0], y = A.py[0] - B.py[0];
      double rox = roix - rojx, roy = roiy - rojy;
      double gx = (vx - (a * (x + 2.0 * rox) + b_min * vx) / d_min) / a;
      double gy = (vy - (a * (y + 2.0 * roy) + b_min * vy) / d_min) / a;
      double m = magnitude(tau, o.tau_0);
      if (!side_b) {
        sx = 0.0 + (o.neg_mk * gx) * m;
        sy = 0.0 + (o.neg_mk * gy) * m;
      } else {
        sx = 0.0 - (o.neg_mk * gx) * m;
        sy = 0.0 - (o.neg_mk * gy) * m;
      }
      truncate(sx, sy, kFSocMax);
    }
  }
  if (h_min < 0.0) {
    double cx, cy;
    force_contact(h_min, nx, ny, vx, vy, ny, -nx, o.mu, o.kappa, o.damping, cx, cy);
    if (!side_b) {
      sx += cx;
      sy += cy;
    } else {
      sx -= cx;
      sy -= cy;
    }
  }
  fx = sx;
  fy = sy;
  torque = rmx * sy - rmy * sx;  // cross(r_moment, f), core/vector2D.py:149
  return true;
}

// fp32 mode of the same pair (velocities and forces in float; positions,
// the moment arm and the torque in double): the part distances from the
// double differences rounded to float, the hardware square root, and the
// divisions by a, ds, tau, tau_0 and d as products with one shared
// reciprocal each, as in pair_circular<float>.  The selections keep the
// reference's rules and quirks (the first minimum of h, strict <;
// x0[j_min] for side B's moment arm; the first valid tau kept unless a
// later one is in (0, tau)); near ties the float distances may pick
// another part pair than the double ones, which the fp32 mode's error
// bound covers.
__device__ __forceinline__ bool pair_three_circle_f32(const TCView &A, const TCView &B, bool side_b,
                                                      const Owner<float> &o, float &fx, float &fy,
                                                      double &torque) {
  const float rA[3] = {(float)A.r_t, (float)A.r_s, (float)A.r_s};
  const float rB[3] = {(float)B.r_t, (float)B.r_s, (float)B.r_s};
  float h_min = __builtin_nanf("");
  float minx = 0.0f, miny = 0.0f, mind = 0.0f;
  double selx = 0.0, sely = 0.0, bugx = 0.0, bugy = 0.0;
  float selr = 0.0f, bugr = 0.0f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const float x = (float)(A.px[i] - B.px[j]), y = (float)(A.py[i] - B.py[j]);
      const float d = __builtin_amdgcn_sqrtf(x * x + y * y);
      const float h = d - (rA[i] + rB[j]);
      if (h < h_min || __builtin_isnan(h_min)) {
        h_min = h;
        minx = x;
        miny = y;
        mind = d;
        selx = A.px[i];
        sely = A.py[i];
        selr = rA[i];
        bugx = A.px[j];
        bugy = A.py[j];
        bugr = rB[j];
      }
    }
  }
  float nx = 0.0f, ny = 0.0f;
  if (mind != 0.0f) {
    const float rd = frcp_normal(mind);
    nx = minx * rd;
    ny = miny * rd;
  }
  if (!(h_min < float(kSightSoc))) return false;
  double rmx, rmy;
  if (!side_b) {
    rmx = (selx + (double)(selr * nx)) - A.px[0];
    rmy = (sely + (double)(selr * ny)) - A.py[0];
  } else {
    rmx = (bugx - (double)(bugr * nx)) - B.px[0];
    rmy = (bugy - (double)(bugr * ny)) - B.py[0];
  }
  float sx = 0.0f, sy = 0.0f;
  const float vx = (float)(A.vx - B.vx), vy = (float)(A.vy - B.vy);
  const float a = vx * vx + vy * vy;
  if (a != 0.0f) {
    const float ra = frcp(a);
    int ci = 0, cj = 0;
    float tau = __builtin_nanf(""), b_min = 0.0f, d_min = 0.0f;
#pragma unroll
    for (int pi = 0; pi < 3; ++pi) {
#pragma unroll
      for (int pj = 0; pj < 3; ++pj) {
        const float x = (float)(A.px[pi] - B.px[pj]), y = (float)(A.py[pi] - B.py[pj]);
        const float r_tot = rA[pi] + rB[pj];
        const float b = -(x * vx + y * vy);
        const float c = (x * x + y * y) - r_tot * r_tot;
        const float d = __builtin_amdgcn_sqrtf(b * b - a * c);
        if (__builtin_isnan(d) || d == 0.0f) continue;
        const float tau_new = (b - d) * ra;
        if (__builtin_isnan(tau) || (0.0f < tau_new && tau_new < tau)) {
          ci = pi;
          cj = pj;
          tau = tau_new;
          b_min = b;
          d_min = d;
        }
      }
    }
    if (!(__builtin_isnan(tau) || tau <= 0.0f)) {
      float roix = 0.0f, roiy = 0.0f, rojx = 0.0f, rojy = 0.0f;
      if (ci == 1) {
        roix = (float)A.ox;
        roiy = (float)A.oy;
      } else if (ci == 2) {
        roix = -(float)A.ox;
        roiy = -(float)A.oy;
      }
      if (cj == 1) {
        rojx = (float)B.ox;
        rojy = (float)B.oy;
      } else if (cj == 2) {
        rojx = -(float)B.ox;
        rojy = -(float)B.oy;
      }
      const float x = (float)(A.px[0] - B.px[0]), y = (float)(A.py[0] - B.py[0]);
      const float rox = roix - rojx, roy = roiy - rojy;
      const float rds = frcp_normal(d_min);
      const float gx = (vx - (a * (x + 2.0f * rox) + b_min * vx) * rds) * ra;
      const float gy = (vy - (a * (y + 2.0f * roy) + b_min * vy) * rds) * ra;
      const float rt = frcp_normal(tau), rt0 = o.rtau_0;
      const float m = ((2.0f * rt + rt0) * __expf(-tau * rt0)) * (rt * rt);
      if (!side_b) {
        sx = (o.neg_mk * gx) * m;
        sy = (o.neg_mk * gy) * m;
      } else {
        sx = -((o.neg_mk * gx) * m);
        sy = -((o.neg_mk * gy) * m);
      }
      truncate(sx, sy, float(kFSocMax));
    }
  }
  if (h_min < 0.0f) {
    float cx, cy;
    force_contact(h_min, nx, ny, vx, vy, ny, -nx, o.mu, o.kappa, o.damping, cx, cy);
    if (!side_b) {
      sx += cx;
      sy += cy;
    } else {
      sx -= cx;
      sy -= cy;
    }
  }
  fx = sx;
  fy = sy;
  torque = rmx * (double)sy - rmy * (double)sx;  // cross(r_moment, f), core/vector2D.py:149
  return true;
}

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    uint32_t lo0 = 0xD2511F53u * c[0], hi0 = __umulhi(0xD2511F53u, c[0]);
    uint32_t lo1 = 0xCD9E8D57u * c[2], hi1 = __umulhi(0xCD9E8D57u, c[2]);
    uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

__device__ __forceinline__ double u01(uint32_t hi, uint32_t lo) {
  unsigned long long x = ((((unsigned long long)hi) << 32) | lo) >> 11;
  return ((double)x + 0.5) * 0x1.0p-53;
}

// Two uniforms of (agent i, step s, purpose, attempt) keyed by seed.
__device__ __forceinline__ void draw2(uint64_t seed, int64_t i, uint64_t s, uint32_t purpose,
                                      uint32_t attempt, double &u0, double &u1) {
  uint32_t c[4] = {(uint32_t)i, (uint32_t)s, (uint32_t)(s >> 32), purpose | (attempt << 8)};
  philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  u0 = u01(c[0], c[1]);
  u1 = u01(c[2], c[3]);
}

// Standard normal truncated to [-3, 3]: Box-Muller with rejection.
__device__ __forceinline__ double truncnorm3(uint64_t seed, int64_t i, uint64_t s,
                                             uint32_t purpose) {
  double z = 0.0;
  for (uint32_t a = 0; a < 16; ++a) {
    double u0, u1;
    draw2(seed, i, s, purpose, a, u0, u1);
    z = sqrt(-2.0 * log(u0)) * cos(2.0 * kPi * u1);
    if (fabs(z) <= 3.0) return z;
  }
  return copysign(3.0, z);
}

}  // namespace cd

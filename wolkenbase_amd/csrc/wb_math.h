// wb_math.h -- scalar FP64 geometry shared by the HIP kernels and by the
// host (C++) side of libwb_gpu.  Every function reproduces, bit for bit, the
// result the reference gets from glibc 2.35 + x86-64 (SSE2 doubles, x87 long
// double), so it must be compiled with -ffp-contract=off and no fast-math.
//
//   hypot            glibc 2.35 e_hypot: Borges, "An Improved Algorithm for
//                    hypot(a,b)", ACM TOMS 47(1), 2020 (arXiv:1904.09481),
//                    unfused correction step, with power-of-two scaling for
//                    huge/tiny operands.  Called by dist() (point.cpp:168-171)
//                    and every Shape::in (shape.cpp:79-238).
//   atan2i           angle.cpp:117-155 (binary angle, 2^31 = one turn).  The
//                    last step is evaluated by the reference in x87 long
//                    double; here a double-double evaluation decides the
//                    rounding unless the value lies within 2^-25 of a
//                    half-integer, in which case the three long-double
//                    operations and lrintl are emulated exactly on 64-bit
//                    significands (X80 below).
//   shape predicates shape.cpp:38-238 with a margin-checked fast path that
//                    skips hypot/sqrt only when the answer is certain.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WB_HD __host__ __device__ inline
// rarely taken exact paths stay out of line so the hot loops stay small
#define WB_HD_COLD __host__ __device__ inline __attribute__((noinline))
#else
#include <math.h>
#define WB_HD inline
#define WB_HD_COLD inline __attribute__((noinline))
#endif

namespace wb {

// ---------------------------------------------------------------------------
// Binary angle constants (angle.h:102-119)
// ---------------------------------------------------------------------------
constexpr int DEG30 = 0x0aaaaaab;
constexpr int DEG45 = 0x10000000;
constexpr int DEG90 = 0x20000000;
constexpr int DEG144 = 0x33333333;
constexpr int DEG180 = 0x40000000;

WB_HD double dabs(double x) { return __builtin_fabs(x); }
WB_HD double dsqrt(double x) { return __builtin_sqrt(x); }
WB_HD bool disfinite(double x) { return __builtin_isfinite(x); }
WB_HD bool disnan(double x) { return __builtin_isnan(x); }
WB_HD bool disinf(double x) { return __builtin_isinf(x); }

// ---------------------------------------------------------------------------
// hypot (glibc 2.35 result)
// ---------------------------------------------------------------------------
WB_HD double hypot_core(double ax, double ay)
{
  // ax >= ay >= 0, neither huge nor tiny.  Borges' correction: h = sqrt of
  // the rounded sum, then h -= (h^2 - ax^2 - ay^2)/(2h) with the residual
  // computed without FMA from one of two rearrangements.
  double h = dsqrt(ax * ax + ay * ay);
  double t1, t2;
  if (h <= 2.0 * ay) {
    double d = h - ay;
    t1 = ax * (2.0 * d - ax);
    t2 = (d - 2.0 * (ax - ay)) * d;
  } else {
    double d = h - ax;
    t1 = 2.0 * d * (ax - 2.0 * ay);
    t2 = (4.0 * d - ay) * ay + d * d;
  }
  h -= (t1 + t2) / (2.0 * h);
  return h;
}

WB_HD double hypot(double x, double y)
{
  if (!disfinite(x) || !disfinite(y)) {
    if (disinf(x) || disinf(y)) return __builtin_inf();
    return x + y;  // NaN
  }
  x = dabs(x);
  y = dabs(y);
  double ax = x < y ? y : x, ay = x < y ? x : y;
  const double big = 0x1p511, tiny = 0x1p-511, scale = 0x1p-600, eps = 0x1p-54;
  if (ax > big) {
    if (ay <= ax * eps) return ax + ay;
    return hypot_core(ax * scale, ay * scale) / scale;
  }
  if (ay < tiny) {
    if (ax >= ay / eps) return ax + ay;
    return hypot_core(ax / scale, ay / scale) * scale;
  }
  if (ay <= ax * eps) return ax + ay;
  return hypot_core(ax, ay);
}

// ---------------------------------------------------------------------------
// 64x64 -> 128 multiply and small 128-bit helpers
// ---------------------------------------------------------------------------
struct U128 {
  uint64_t hi, lo;
};

WB_HD U128 mul64(uint64_t a, uint64_t b)
{
  uint64_t a0 = a & 0xffffffffu, a1 = a >> 32, b0 = b & 0xffffffffu, b1 = b >> 32;
  uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
  uint64_t mid = (p00 >> 32) + (p01 & 0xffffffffu) + (p10 & 0xffffffffu);
  U128 r;
  r.lo = (mid << 32) | (p00 & 0xffffffffu);
  r.hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
  return r;
}

WB_HD int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }

// ---------------------------------------------------------------------------
// X80: exact emulation of x87 extended precision (64-bit significand,
// round to nearest even) for the last step of atan2i.
// value = (-1)^s * m * 2^e, m normalized (bit 63 set) unless zero.
// ---------------------------------------------------------------------------
struct X80 {
  uint64_t m;
  int e;
  int s;
};

WB_HD X80 x80_from_double(double d)
{
  X80 r;
  uint64_t b = __builtin_bit_cast(uint64_t, d);
  r.s = (int)(b >> 63);
  int ex = (int)((b >> 52) & 0x7ff);
  uint64_t fr = b & 0xfffffffffffffull;
  if (ex == 0) {
    if (fr == 0) { r.m = 0; r.e = 0; return r; }
    int lz = clz64(fr);
    r.m = fr << lz;
    r.e = -1074 - lz;
    return r;
  }
  r.m = (fr | (1ull << 52)) << 11;
  r.e = ex - 1075 - 11;
  return r;
}

// Round mag (128-bit) * 2^e, with sticky bits below it, to 64 bits.
WB_HD X80 x80_round(U128 mag, int e, int s, bool sticky)
{
  X80 r;
  r.s = s;
  if (mag.hi == 0 && mag.lo == 0) { r.m = 0; r.e = 0; return r; }
  int lead = mag.hi ? 127 - clz64(mag.hi) : 63 - clz64(mag.lo);
  int sh = lead - 63;  // bits to drop, in [-63, 64]
  if (sh <= 0) {
    // fits in 64 bits; any sticky part is below half an ulp
    r.m = (sh == 0) ? mag.lo : (mag.lo << (-sh));
    r.e = e + sh;
    return r;
  }
  uint64_t m, below;
  bool rbit;
  if (sh == 64) {
    m = mag.hi;
    rbit = (mag.lo >> 63) & 1;
    below = mag.lo & 0x7fffffffffffffffull;
  } else {
    m = (mag.hi << (64 - sh)) | (mag.lo >> sh);
    rbit = (mag.lo >> (sh - 1)) & 1;
    below = (sh == 1) ? 0 : (mag.lo & ((1ull << (sh - 1)) - 1));
  }
  int ee = e + sh;
  if (rbit && (sticky || below != 0 || (m & 1))) {
    m += 1;
    if (m == 0) { m = 1ull << 63; ee += 1; }
  }
  r.m = m;
  r.e = ee;
  return r;
}

WB_HD X80 x80_mul(X80 a, X80 b)
{
  U128 p = mul64(a.m, b.m);
  return x80_round(p, a.e + b.e, a.s ^ b.s, false);
}

// floor((hi:lo) / d) and remainder, for d >= 2^63 and hi < d (bitwise).
WB_HD uint64_t div128(uint64_t hi, uint64_t lo, uint64_t d, uint64_t *rem)
{
  uint64_t q = 0, r = hi;
  for (int i = 63; i >= 0; i--) {
    uint64_t carry = r >> 63;
    r = (r << 1) | ((lo >> i) & 1);
    if (carry || r >= d) {
      r -= d;
      q |= 1ull << i;
    }
  }
  *rem = r;
  return q;
}

WB_HD X80 x80_div(X80 a, X80 b)
{
  // q = floor(a.m * 2^64 / b.m) in [2^63, 2^65); two more bits from rem.
  X80 z;
  if (a.m == 0) { z.m = 0; z.e = 0; z.s = a.s ^ b.s; return z; }
  uint64_t qhi = (a.m >= b.m) ? 1 : 0;
  uint64_t h = a.m - (qhi ? b.m : 0);
  uint64_t rem;
  uint64_t qlo = div128(h, 0, b.m, &rem);
  // extend by two quotient bits
  uint64_t ext = 0;
  for (int k = 0; k < 2; k++) {
    uint64_t carry = rem >> 63;
    rem <<= 1;
    ext <<= 1;
    if (carry || rem >= b.m) { rem -= b.m; ext |= 1; }
  }
  U128 mag;
  mag.hi = (qhi << 2) | (qlo >> 62);
  mag.lo = (qlo << 2) | ext;
  return x80_round(mag, a.e - b.e - 64 - 2, a.s ^ b.s, rem != 0);
}

WB_HD X80 x80_add(X80 a, X80 b)
{
  if (a.m == 0) return b;
  if (b.m == 0) return a;
  // order by magnitude (both normalized: exponent, then significand)
  if ((b.e > a.e) || (b.e == a.e && b.m > a.m)) { X80 t = a; a = b; b = t; }
  int d = a.e - b.e;
  // 128-bit window: a.m at bits 127..64 (unit 2^(a.e-64)); b shifted in
  uint64_t shi, slo;
  bool sticky = false;
  if (d == 0) { shi = b.m; slo = 0; }
  else if (d < 64) { shi = b.m >> d; slo = b.m << (64 - d); }
  else if (d == 64) { shi = 0; slo = b.m; }
  else if (d < 128) { shi = 0; slo = b.m >> (d - 64); sticky = (b.m << (128 - d)) != 0; }
  else { shi = 0; slo = 0; sticky = true; }
  U128 r;
  int e = a.e - 64;
  if (a.s == b.s) {
    r.lo = slo;
    r.hi = a.m + shi;
    if (r.hi < a.m) {  // carry out of bit 127
      sticky = sticky || (r.lo & 1);
      r.lo = (r.lo >> 1) | (r.hi << 63);
      r.hi = (r.hi >> 1) | (1ull << 63);
      e += 1;
    }
    return x80_round(r, e, a.s, sticky);
  }
  // a.m*2^64 - (shi:slo) - (sticky ? something in (0,1) : 0)
  r.lo = 0 - slo;
  r.hi = a.m - shi - (slo != 0 ? 1 : 0);
  if (sticky) {
    if (r.lo == 0) r.hi -= 1;
    r.lo -= 1;
  }
  return x80_round(r, e, a.s, sticky);
}

WB_HD X80 x80_neg(X80 a) { a.s ^= 1; return a; }

// lrintl (round to nearest, ties to even) for |value| < 2^62
WB_HD int64_t x80_lrint(X80 a)
{
  if (a.m == 0) return 0;
  uint64_t ip;
  if (a.e >= 0) {
    ip = a.m << a.e;
  } else {
    int sh = -a.e;
    if (sh > 64) return 0;
    uint64_t frac, half;
    if (sh == 64) { ip = 0; frac = a.m; half = 1ull << 63; }
    else { ip = a.m >> sh; frac = a.m & ((1ull << sh) - 1); half = 1ull << (sh - 1); }
    if (frac > half || (frac == half && (ip & 1))) ip += 1;
  }
  int64_t v = (int64_t)ip;
  return a.s ? -v : v;
}

// ---------------------------------------------------------------------------
// Double-double helpers (FMA is exact here by construction)
// ---------------------------------------------------------------------------
WB_HD void two_prod(double a, double b, double &p, double &e)
{
  p = a * b;
  e = __builtin_fma(a, b, -p);
}
WB_HD void two_sum(double a, double b, double &s, double &e)
{
  s = a + b;
  double bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// Constants of the last atan2i step: K = 0x40000000/M_PIl as a long double
// (64-bit significand), split exactly into two doubles, and the cubic
// coefficient of angle.cpp:151.
struct AtanConst {
  double k_hi, k_lo;     // K = k_hi + k_lo exactly
  uint64_t k_m;          // K = k_m * 2^k_e (x87 value)
  int k_e;
};
constexpr double ATAN_C3 = 1.1392738508503886e8;

// Exact x87 evaluation of lrintl(K*y/x - c3) (angle.cpp:151).
WB_HD_COLD int64_t atan_last_exact(double y, double x, double c3, uint64_t k_m, int k_e)
{
  X80 k;
  k.m = k_m; k.e = k_e; k.s = 0;
  X80 p = x80_mul(k, x80_from_double(y));
  X80 q = x80_div(p, x80_from_double(x));
  X80 r = x80_add(q, x80_neg(x80_from_double(c3)));
  return x80_lrint(r);
}

// atan2i(y, x), angle.cpp:117-155.  tab_tan[511], tab_cos[512], tab_sin[512]
// are the host-built tables of fillTanTables (angle.cpp:305-320).
template <typename TabPtr>
WB_HD int atan2i(double y, double x, TabPtr tab_tan, TabPtr tab_cos, TabPtr tab_sin,
                 const AtanConst &K, int *exact_used = nullptr, bool force_exact = false)
{
  int a = 0;
  double t;
  if (x < 0) {
    a += (y > 0) ? DEG180 : -DEG180;
    y = -y;
    x = -x;
  }
  if (y > x) {
    a += DEG90;
    t = x; x = y; y = -t;
  }
  if (-y > x) {
    a -= DEG90;
    t = x; x = -y; y = t;
  }
  t = y / x;
#pragma unroll
  for (int step = DEG45 / 2; step > DEG45 / 1024; step /= 2) {
    int k = (((a + DEG45) & 0x1ff00000) >> 20) - 1;
    a += (t > tab_tan[k]) ? step : -step;
  }
  {
    int k = 511 - (((a + DEG45) & 0x1ff00000) >> 20);
    double c = tab_cos[k], s = tab_sin[k];
    t = x * c - y * s;
    y = y * c + x * s;
    x = t;
  }
  if (x == 0 && y == 0) return 0;
  double u = y / x;
  double c3 = ATAN_C3 * (u * u * u);
  // double-double value of K*y/x - c3
  double p1, p2;
  two_prod(K.k_hi, y, p1, p2);
  p2 += K.k_lo * y;
  double q1 = p1 / x;
  double rr = __builtin_fma(-q1, x, p1) + p2;
  double q2 = rr / x;
  double v_hi, v_lo;
  two_sum(q1, -c3, v_hi, v_lo);
  v_lo += q2;
  double n = __builtin_rint(v_hi);
  double f = (v_hi - n) + v_lo;  // v_hi - n is exact
  double af = dabs(f);
  int64_t r;
  if (!force_exact && af < 0.5 - 0x1p-25) {
    r = (int64_t)n;
  } else if (!force_exact && af > 0.5 + 0x1p-25) {
    r = (int64_t)n + (f > 0 ? 1 : -1);
  } else {
    r = atan_last_exact(y, x, c3, K.k_m, K.k_e);
    if (exact_used) *exact_used += 1;
  }
  return (int)((int64_t)a + r);
}

// ---------------------------------------------------------------------------
// 64-bin direction index of atan2i(dy, dx): the bin of the circle position
// (a + 2^30) mod 2^31, 2^25 units per bin, i.e. floor((theta + pi) * 32/pi)
// for the angle theta of (dx, dy).  A float path decides it unless the
// direction lies within DIR_MARGIN bins (2^-14, 2048 units, 6e-6 rad) of a
// bin edge:
//   t   = min(|dx|, |dy|) / max(|dx|, |dy|) in [0, 1] (the octant ratio),
//   psi = t * P(t^2) ~ atan(t) in [0, pi/4], P of degree 6 fitted by
//         minimax (Lawson's iteratively reweighted least squares) to atan
//         on [0, 1]: 2.5e-7 rad, 3.2e-7 rad evaluated in float with fma,
//   s   = psi * 32/pi, the position in the octant's eight bins,
// and the octant folds s onto the circle: the bin in the quadrant is k or
// 15 - k (k = floor(s); |dy| > |dx| mirrors), the quadrants of opposite signs
// mirror again (31 - b) and dy > 0 adds 32.  The float path's error: the
// conversions of dx, dy (2^-24 relative each), the quotient (the device's
// v_rcp_f32 is within 1 ulp; the host divides) and the products add at most
// 3e-7 rad to the fit's 3.2e-7 -- against a 6e-6 rad margin -- and atan2i
// itself is within one unit (3e-9 rad) of the true angle, so outside the
// edge band both give the same bin.  Inside it (which holds every direction
// on an axis or a diagonal), or when max(|dx|, |dy|) as a float lies outside
// [2^-100, 2^100], atan2i decides.
// ---------------------------------------------------------------------------
constexpr float DIR_MARGIN = 0x1p-14f;

// atan2i out of line; returns the angle in the low 32 bits and the number of
// exact last steps taken (0 or 1) in bit 32
WB_HD_COLD int64_t atan2i_cold(double y, double x, const double *tab_tan, const double *tab_cos,
                               const double *tab_sin, double k_hi, double k_lo, uint64_t k_m,
                               int k_e)
{
  AtanConst K;
  K.k_hi = k_hi; K.k_lo = k_lo; K.k_m = k_m; K.k_e = k_e;
  int used = 0;
  const int a = atan2i(y, x, tab_tan, tab_cos, tab_sin, K, &used);
  return (int64_t)(uint32_t)a | ((int64_t)used << 32);
}

WB_HD int atan2i_ext(double y, double x, const double *tab_tan, const double *tab_cos,
                     const double *tab_sin, const AtanConst &K, int *exact_used)
{
  const int64_t r = atan2i_cold(y, x, tab_tan, tab_cos, tab_sin, K.k_hi, K.k_lo, K.k_m, K.k_e);
  if (exact_used) *exact_used += (int)(r >> 32);
  return (int)(uint32_t)r;
}

WB_HD int dir_bin(double dy, double dx, const double *tab_tan, const double *tab_cos,
                  const double *tab_sin, const AtanConst &K, int *exact_used = nullptr)
{
  {
    const float fx = (float)dx, fy = (float)dy;
    const float ax = fabsf(fx), ay = fabsf(fy);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
#ifdef __HIP_DEVICE_COMPILE__
    const float t = mn * __builtin_amdgcn_rcpf(mx);
#else
    const float t = mn / mx;
#endif
    const float u = t * t;
    float p = 0.006811794359236956f;
    p = fmaf(p, u, -0.033604223281145096f);
    p = fmaf(p, u, 0.07962367683649063f);
    p = fmaf(p, u, -0.1323334276676178f);
    p = fmaf(p, u, 0.19807815551757812f);
    p = fmaf(p, u, -0.3331736922264099f);
    p = fmaf(p, u, 0.9999961256980896f);
    const float s = (p * t) * 10.185916357881302f;  // 32/pi
    const float fl = floorf(s);
    const float fr = s - fl;  // exact
    // mx in [2^-100, 2^100]: both conversions normal or t below the margin
    // (a NaN fails a compare and takes atan2i)
    if (fr > DIR_MARGIN && fr < 1.0f - DIR_MARGIN && s < 8.0f && mx >= 0x1p-100f &&
        mx <= 0x1p100f) {
      const uint32_t ux = __builtin_bit_cast(uint32_t, fx), uy = __builtin_bit_cast(uint32_t, fy);
      int b = (int)fl;
      if (ay > ax) b ^= 15;
      if ((ux ^ uy) >> 31) b ^= 31;
      return b + (fy > 0 ? 32 : 0);
    }
  }
  const int a = atan2i_ext(dy, dx, tab_tan, tab_cos, tab_sin, K, exact_used);
  return (int)((((uint32_t)a + 0x40000000u) & 0x7fffffffu) >> 25);
}

// ---------------------------------------------------------------------------
constexpr double kScanSqrt7 = 2.6457513110645905905016;  // scan.h:23 (M_SQRT7)
constexpr double kScanSqrt3 = 1.7320508075688772935274463;

// The histogram sector of a bottom-layer point (vx, vy) from the cylinder
// axis (scan.cpp:93-101): lrint(atan2(vy, vx)*3/pi) folded to 1..6, or 0
// inside radius r/sqrt(7); *edge when atan2*3/pi lies within 1e-12 of a
// half-integer.  Out of line: the exact form for the rare points near a
// boundary.
WB_HD_COLD int scan_sector_exact(double vx, double vy, double r, int *edge)
{
  const double th = atan2(vy, vx) * 3 / M_PI;
  long sc = lrint(th);
  const double fr = th - floor(th);
  *edge = dabs(fr - 0.5) < 1e-12 ? 1 : 0;
  if (sc < 0) sc += 6;
  sc = (sc % 6) + 1;
  if (hypot(vx, vy) < r / kScanSqrt7) sc = 0;
  return (int)sc;
}

// The same from comparisons: the sector boundaries are the directions at
// 30, 90 and 150 degrees either side of +x, i.e. sqrt(3)|vy| = |vx| and
// vx = 0.  When sqrt(3)|vy| and |vx| differ by more than 1e-9 of their sum and
// |vx| exceeds it, the direction is at least 8e-10 rad (in atan2*3/pi, 8e-10)
// from every boundary, far beyond the roundings of atan2, the product and
// the compares, and the edge test cannot fire; when vx^2 + vy^2 differs from
// (r/sqrt(7))^2 by more than 1e-12 of it, hypot's rounding cannot change the
// radius test.  Otherwise the exact form decides.
WB_HD int scan_sector(double vx, double vy, double r, int *edge)
{
  const double rr = r / kScanSqrt7;
  const double a = kScanSqrt3 * dabs(vy), b = dabs(vx), m = 1e-9 * (a + b);
  const double q2 = vx * vx + vy * vy, t2 = rr * rr;
  if ((dabs(a - b) > m) & (b > m) & (dabs(q2 - t2) > 1e-12 * t2)) {
    *edge = 0;
    if (q2 < t2) return 0;
    if (a < b) return vx > 0 ? 1 : 4;
    return vy > 0 ? (vx > 0 ? 2 : 3) : (vx > 0 ? 6 : 5);
  }
  return scan_sector_exact(vx, vy, r, edge);
}

// ---------------------------------------------------------------------------
// Shapes (shape.cpp).  Kinds follow the C ABI: 0 cylinder, 1 sphere,
// 2 hyperboloid.
// ---------------------------------------------------------------------------
enum { SHAPE_CYLINDER = 0, SHAPE_SPHERE = 1, SHAPE_HYPERBOLOID = 2 };

struct Shape {
  int kind;
  double cx, cy, cz;  // centre (hyperboloid: vertex + (0,0,por))
  double r;           // cylinder / sphere radius
  double por2, slope; // hyperboloid
};

// Hyperboloid(v, r, s), shape.cpp:119-125
WB_HD Shape make_hyperboloid(double vx, double vy, double vz, double r, double s)
{
  Shape h;
  double por = r * (s * s);
  h.kind = SHAPE_HYPERBOLOID;
  h.slope = s;
  h.por2 = por * por;
  h.cx = vx + 0;
  h.cy = vy + 0;
  h.cz = vz + por;
  h.r = 0;
  return h;
}
WB_HD Shape make_cylinder(double cx, double cy, double r)
{
  Shape c;
  c.kind = SHAPE_CYLINDER; c.cx = cx; c.cy = cy; c.cz = 0; c.r = r; c.por2 = 0; c.slope = 0;
  return c;
}
WB_HD Shape make_sphere(double cx, double cy, double cz, double r)
{
  Shape c;
  c.kind = SHAPE_SPHERE; c.cx = cx; c.cy = cy; c.cz = cz; c.r = r; c.por2 = 0; c.slope = 0;
  return c;
}

// the exact forms of the three predicates, for the rare points near the
// boundary
WB_HD_COLD bool cyl_in_tail(double dx, double dy, double r) { return hypot(dx, dy) <= r; }
WB_HD_COLD bool sph_in_tail(double dx, double dy, double dz, double r)
{
  return hypot(hypot(dx, dy), dz) <= r;
}
WB_HD_COLD bool hyp_in_tail(double dx, double dy, double a, double slope, double por2)
{
  double xyd = hypot(dx, dy);
  double w = xyd * slope;
  return a - w * w >= por2;
}

// Cylinder::in, shape.cpp:214-218: hypot(c - q) <= r.  The fast path decides
// from dx^2+dy^2 when it is more than 2^-46 (relative) away from r^2, where
// hypot's rounding cannot change the answer.
WB_HD bool cyl_in(const Shape &s, double qx, double qy)
{
  double dx = s.cx - qx, dy = s.cy - qy;
  double q2 = dx * dx + dy * dy, r2 = s.r * s.r;
  double m = 0x1p-46 * r2;
  if (q2 < r2 - m) return true;
  if (q2 > r2 + m) return false;
  return cyl_in_tail(dx, dy, s.r);
}

// Sphere::in, shape.cpp:172-175: hypot(hypot(dx,dy),dz) <= r.
WB_HD bool sph_in(const Shape &s, double qx, double qy, double qz)
{
  double dx = qx - s.cx, dy = qy - s.cy, dz = qz - s.cz;
  double q2 = dx * dx + dy * dy + dz * dz, r2 = s.r * s.r;
  double m = 0x1p-45 * r2;
  if (q2 < r2 - m) return true;
  if (q2 > r2 + m) return false;
  return sph_in_tail(dx, dy, dz, s.r);
}

// Hyperboloid::in, shape.cpp:127-135: the point is on the hyperboloid's
// side of the vertex (below it for a positive slope) and
// (zd^2 - slope^2 (dx^2 + dy^2)) - por^2 >= 0, decided from the estimate
// unless it lies within 2^-44 (relative) of zero.  Written as one predicate
// (no early returns), so the hot loops carry one exec-mask branch, the rare
// exact tail.
WB_HD bool hyp_in(const Shape &s, double qx, double qy, double qz)
{
  const double zd = s.cz - qz;
  const bool side = zd * (s.slope > 0 ? 1.0 : -1.0) > 0;  // zd > 0, or zd < 0
  const double dx = s.cx - qx, dy = s.cy - qy;
  const double a = zd * zd;
  const double b = (dx * dx + dy * dy) * (s.slope * s.slope);
  const double d = (a - b) - s.por2;
  const double m = 0x1p-44 * (a + b);
  bool in = side & (d > m);
  if (side & !(d > m) & !(d < -m)) in = hyp_in_tail(dx, dy, a, s.slope, s.por2);
  return in;
}

WB_HD bool shape_in(const Shape &s, double qx, double qy, double qz)
{
  switch (s.kind) {
  case SHAPE_CYLINDER: return cyl_in(s, qx, qy);
  case SHAPE_SPHERE: return sph_in(s, qx, qy, qz);
  default: return hyp_in(s, qx, qy, qz);
  }
}

// Shape::in(Cube&): all eight corners (shape.cpp:45-61)
WB_HD bool shape_in_cube(const Shape &s, double cx, double cy, double cz, double side)
{
  double h = side / 2;
  double x0 = cx - h, x1 = cx + h, y0 = cy - h, y1 = cy + h, z0 = cz - h, z1 = cz + h;
  for (int n = 0; n < 8; n++) {
    double qx = (n & 1) ? x1 : x0, qy = (n & 2) ? y1 : y0, qz = (n & 4) ? z1 : z0;
    if (!shape_in(s, qx, qy, qz)) return false;
  }
  return true;
}

WB_HD double clamp_axis(double sc, double cc, double side)
{
  if (dabs(sc - cc) < side / 2) return sc;
  if (sc > cc) return cc + side / 2;
  return cc - side / 2;
}

// Shape::intersect(cube) = in(closestPoint(cube)), shape.cpp:63-66, 137-238
WB_HD bool shape_intersect(const Shape &s, double cx, double cy, double cz, double side)
{
  double qx = clamp_axis(s.cx, cx, side), qy = clamp_axis(s.cy, cy, side), qz = cz;
  if (s.kind == SHAPE_SPHERE) qz = clamp_axis(s.cz, cz, side);
  else if (s.kind == SHAPE_HYPERBOLOID) {
    if (s.slope > 0) qz -= side / 2;
    if (s.slope < 0) qz += side / 2;
  }
  return shape_in(s, qx, qy, qz);
}

// Conservative box test for pruning: false only if no point of the box
// [lo,hi] can satisfy shape_in.  Clamping each axis to the box gives, per
// axis, a difference no larger in magnitude than any point of the box, and
// every predicate above is monotone in those magnitudes (and in z for the
// hyperboloid), so a false here implies false for every point in the box.
WB_HD bool shape_may_hit_box(const Shape &s, double x0, double x1, double y0, double y1,
                             double z0, double z1)
{
  double qx = s.cx < x0 ? x0 : (s.cx > x1 ? x1 : s.cx);
  double qy = s.cy < y0 ? y0 : (s.cy > y1 ? y1 : s.cy);
  double qz;
  if (s.kind == SHAPE_SPHERE) qz = s.cz < z0 ? z0 : (s.cz > z1 ? z1 : s.cz);
  else if (s.kind == SHAPE_HYPERBOLOID) qz = s.slope > 0 ? z0 : z1;
  else qz = z0;
  return shape_in(s, qx, qy, qz);
}

// The three tests above for a shape known to be a hyperboloid (the same
// operations without the kind dispatch).
WB_HD bool hyp_in_cube(const Shape &s, double cx, double cy, double cz, double side)
{
  double h = side / 2;
  double x0 = cx - h, x1 = cx + h, y0 = cy - h, y1 = cy + h, z0 = cz - h, z1 = cz + h;
  // the corner likeliest to fail first (farthest from the axis in x and y,
  // on the face away from the vertex side): when it is out, so is the AND of
  // all eight, which is then taken without the other seven
  const double fx = s.cx - cx >= 0 ? x0 : x1, fy = s.cy - cy >= 0 ? y0 : y1;
  if (!hyp_in(s, fx, fy, s.slope > 0 ? z1 : z0)) return false;
  bool all = true;
  for (int n = 0; n < 8; n++) {
    double qx = (n & 1) ? x1 : x0, qy = (n & 2) ? y1 : y0, qz = (n & 4) ? z1 : z0;
    all = all && hyp_in(s, qx, qy, qz);
  }
  return all;
}

WB_HD bool hyp_intersect(const Shape &s, double cx, double cy, double cz, double side)
{
  double qx = clamp_axis(s.cx, cx, side), qy = clamp_axis(s.cy, cy, side), qz = cz;
  if (s.slope > 0) qz -= side / 2;
  if (s.slope < 0) qz += side / 2;
  return hyp_in(s, qx, qy, qz);
}

// hyp_intersect for a positive slope with the cube's bottom face raised to
// zlow, the lowest point stored under the cube (-inf: no such bound).  When
// every stored point lies in its leaf's cube, no point under the cube is
// lower than zlow, and (clamped x, clamped y, max(bottom, zlow)) has per-axis
// differences no larger than any of them, so a false here means that no
// point under the cube lies in the hyperboloid.  Nor can the cube hold a leaf
// whose eight corners are all in (Shape::in(Cube&), which takes a leaf's
// points without testing them): such a leaf's top corners would be in, and
// the clamped point with zlow (at most the top, a stored point being below
// it) differs from a top corner by no more on any axis.
WB_HD bool hyp_intersect_low(const Shape &s, double cx, double cy, double cz, double side,
                             double zlow)
{
  double qx = clamp_axis(s.cx, cx, side), qy = clamp_axis(s.cy, cy, side), qz = cz;
  if (s.slope > 0) {
    qz -= side / 2;
    if (zlow > qz) qz = zlow;
  }
  if (s.slope < 0) qz += side / 2;
  return hyp_in(s, qx, qy, qz);
}

WB_HD bool hyp_may_hit_box(const Shape &s, double x0, double x1, double y0, double y1,
                           double z0, double z1)
{
  double qx = s.cx < x0 ? x0 : (s.cx > x1 ? x1 : s.cx);
  double qy = s.cy < y0 ? y0 : (s.cy > y1 ? y1 : s.cy);
  return hyp_in(s, qx, qy, s.slope > 0 ? z0 : z1);
}

// ---------------------------------------------------------------------------
// Direction wedges that can still change a ground test.  With the direction
// bins found so far in `occ`, only a direction in an empty run of >= 24 bins
// can change the outcome (a pass needs every such run broken; a direction in
// a shorter run or an occupied bin leaves every run of >= 24 as it is, and
// runs only shrink as directions are added).  So a box whose xy footprint
// lies, as seen from p, outside every such run cannot change the outcome and
// need not be read.  The wedges are the runs widened by 1e-3 rad on each side
// (bins are pi/32 wide and the bin of a direction is that of atan2i, within
// 2^-31 turn of the true angle), split in two where wider than pi so that
// each is convex; at most two runs of >= 24 bins fit in 64 bins.  The exact
// border pass never uses them: it needs the directions in the occupied bins
// next to a run.
// ---------------------------------------------------------------------------
WB_HD uint64_t rotr64_(uint64_t x, int a) { return a ? (x >> a) | (x << (64 - a)) : x; }
WB_HD uint64_t rotl64_(uint64_t x, int a) { return a ? (x << a) | (x >> (64 - a)) : x; }

// The empty circular runs of >= 24 bins of occ: returns their number (at
// most two fit) with start bins and lengths, or -1 when occ is empty (one
// run of all 64 bins: every direction may still matter).
WB_HD int care_runs(uint64_t occ, int start[2], int len[2])
{
  if (occ == 0) return -1;
  const uint64_t e = ~occ;
  const uint64_t r2 = e & rotr64_(e, 1), r4 = r2 & rotr64_(r2, 2), r8 = r4 & rotr64_(r4, 4);
  const uint64_t r24 = r8 & rotr64_(r8, 8) & rotr64_(r8, 16);  // bins b..b+23 empty
  // every bin within 23 after such a b
  uint64_t c = r24 | rotl64_(r24, 1);
  c |= rotl64_(c, 2);
  c |= rotl64_(c, 4);
  c |= rotl64_(c, 8);
  c |= rotl64_(c, 8);
  // the two iterations written out, so the arrays stay in registers
  uint64_t sm = c & ~rotl64_(c, 1);
  if (!sm) return 0;
  start[0] = __builtin_ctzll(sm);
  len[0] = __builtin_ctzll(~rotr64_(c, start[0]));
  sm &= sm - 1;
  if (!sm) return 1;
  start[1] = __builtin_ctzll(sm);
  len[1] = __builtin_ctzll(~rotr64_(c, start[1]));
  return 2;
}

struct Wedges {
  int live;  // < 0: no pruning; else bit k set when wedge k is in use
  float ax[4], ay[4], bx[4], by[4];  // wedge k spans counterclockwise a_k .. b_k
};

// Unit vectors of the bin edges (angle -pi + k pi/32, k = 0..63), turned by
// side * 1e-3 rad (side -1, 0, +1): 3 x 64 (cos, sin) pairs, side-major.
// The kernels read this table, built once on the host by the same function.
WB_HD void fill_edge_table(float *tab)
{
  for (int side = -1; side <= 1; side++)
    for (int k = 0; k < 64; k++) {
      const double t = -3.14159265358979323846 + k * (3.14159265358979323846 / 32) + side * 1e-3;
      tab[(side + 1) * 128 + 2 * k] = (float)cos(t);
      tab[(side + 1) * 128 + 2 * k + 1] = (float)sin(t);
    }
}

// The care wedges of occ from the edge table: a run of bins [s, s+len) is
// the wedge from edge s turned back by 1e-3 rad to edge s+len turned on by
// 1e-3 rad; a run of 32 bins or more spans more than pi with the margins and
// is split at its middle edge m into two wedges, [s, m] and [m, s+len], each
// with the margin on both sides: at most 31 bins plus 2e-3 rad, so convex
// (wedge_hit's test needs that).  A run of 63 bins (one occupied bin) has no
// such split and is not pruned.  Run r fills slots 2r and, when split, 2r+1
// (fixed slots keep every index a constant).
WB_HD void care_wedges(uint64_t occ, Wedges &w, const float *tab)
{
  w.live = -1;
  int rs[2], rl[2];
  const int nr = care_runs(occ, rs, rl);
  if (nr < 0 || (nr > 0 && rl[0] >= 63)) return;
  int live = 0;
#pragma unroll
  for (int r = 0; r < 2; r++) {
    if (r < nr) {
      const int s = rs[r], len = rl[r], e = (s + len) & 63;
      const bool split = len >= 32;
      const int m = split ? (s + len / 2) & 63 : e;
      w.ax[2 * r] = tab[2 * s];
      w.ay[2 * r] = tab[2 * s + 1];
      w.bx[2 * r] = tab[256 + 2 * m];
      w.by[2 * r] = tab[256 + 2 * m + 1];
      w.ax[2 * r + 1] = tab[2 * m];
      w.ay[2 * r + 1] = tab[2 * m + 1];
      w.bx[2 * r + 1] = tab[256 + 2 * e];
      w.by[2 * r + 1] = tab[256 + 2 * e + 1];
      live |= (split ? 3 : 1) << (2 * r);
    }
  }
  w.live = live;
}

// The care wedges of the exact border pass: with a longest empty run of 24
// or 25 bins, only directions in the run or in the two occupied bins that
// border it can change the exact gap (a direction in the run breaks it, the
// extreme directions of the border bins are its ends).  Run r (r < nr <= 2)
// spans bins before[r]+1 .. after[r]-1; its wedge runs from the start edge of
// bin before[r] turned back by 1e-3 rad to the end edge of bin after[r]
// turned on by 1e-3 rad: at most 27 bins, so convex.  Slots 0 and 2.
WB_HD void border_wedges(const int before[2], const int after[2], int nr, Wedges &w,
                         const float *tab)
{
  int live = 0;
#pragma unroll
  for (int r = 0; r < 2; r++) {
    if (r < nr) {
      const int s = before[r] & 63, e = (after[r] + 1) & 63;
      w.ax[2 * r] = tab[2 * s];
      w.ay[2 * r] = tab[2 * s + 1];
      w.bx[2 * r] = tab[256 + 2 * e];
      w.by[2 * r] = tab[256 + 2 * e + 1];
      live |= 1 << (2 * r);
    }
  }
  w.live = nr > 0 ? live : -1;
}

// false only if the box [x0,x1] x [y0,y1] lies outside every wedge as seen
// from (px, py): all four corners strictly right of a wedge's first edge, or
// all strictly left of its second edge (the box is convex).  Each cross
// product is linear in the corner, so its largest value over the box is at
// the corner the signs of the edge vector pick: all four are negative iff
// that one is.  Rounding: every product is one float product of the same
// operands as at that corner, within 10^-7 (relative) of exact, while a
// direction of a care bin lies at least 10^-3 rad inside a widened edge.
WB_HD bool wedge_hit(const Wedges &w, double px, double py, double x0, double x1,
                                 double y0, double y1)
{
  if (w.live < 0) return true;
  const float u0 = (float)(x0 - px), u1 = (float)(x1 - px);
  const float v0 = (float)(y0 - py), v1 = (float)(y1 - py);
  bool hit = false;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const float ax = w.ax[k], ay = w.ay[k], bx = w.bx[k], by = w.by[k];
    // max over corners of ax*v - ay*u, and of u*by - v*bx
    const float ra = ax * (ax > 0 ? v1 : v0) - ay * (ay > 0 ? u0 : u1);
    const float lb = (by > 0 ? u1 : u0) * by - (bx > 0 ? v0 : v1) * bx;
    // (a NaN, from coordinates a float cannot carry, counts as a hit)
    const bool in_k = !(ra < 0) & !(lb < 0);
    hit |= ((w.live >> k) & 1) ? in_k : false;
  }
  return hit;
}

// wedge_hit for at most two wedges (n of them, n <= 2; n < 0 is no pruning)
// on a box given by its corners' offsets from the point, u0 <= u1 in x and
// v0 <= v1 in y, already rounded to float: the same test per wedge.
WB_HD bool wedge_hit2(int n, const float ax[2], const float ay[2], const float bx[2],
                      const float by[2], float u0, float u1, float v0, float v1)
{
  if (n < 0) return true;
  bool hit = false;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const float ra = ax[k] * (ax[k] > 0 ? v1 : v0) - ay[k] * (ay[k] > 0 ? u0 : u1);
    const float lb = (by[k] > 0 ? u1 : u0) * by[k] - (bx[k] > 0 ? v0 : v1) * bx[k];
    const bool in_k = !(ra < 0) & !(lb < 0);
    hit |= k < n ? in_k : false;
  }
  return hit;
}


}  // namespace wb

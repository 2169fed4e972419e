// wb_hostcheck.cpp -- the device math of wb_math.h / wb_flowsnake.h built
// for the host, so CPU tests can pin it against the oracle without a GPU.
// Not used by the GPU path itself.
#include <math.h>
#include <stdint.h>

#include "wb_flowsnake.h"
#include "wb_host.h"
#include "wb_math.h"

static double g_tan[511], g_cos[512], g_sin[512];
static wb::AtanConst g_k;
static int g_ready;

static void ready()
{
  if (!g_ready) {
    wb::fill_atan_tables(g_tan, g_cos, g_sin);
    g_k = wb::atan_const();
    g_ready = 1;
  }
}

extern "C" {
double wbh_hypot(double x, double y) { return wb::hypot(x, y); }

int wbh_atan2i(double y, double x, int *exact_used)
{
  ready();
  return wb::atan2i(y, x, (const double *)g_tan, (const double *)g_cos, (const double *)g_sin,
                    g_k, exact_used);
}

int wbh_dir_bin(double dy, double dx)
{
  ready();
  return wb::dir_bin(dy, dx, (const double *)g_tan, (const double *)g_cos, (const double *)g_sin,
                     g_k);
}

// the empty runs of >= 24 bins as a bin mask (all ones for -1 runs)
uint64_t wbh_care_mask(uint64_t occ)
{
  int s[2], l[2];
  const int n = wb::care_runs(occ, s, l);
  if (n < 0) return ~0ull;
  uint64_t m = 0;
  for (int r = 0; r < n; r++)
    for (int k = 0; k < l[r]; k++) m |= 1ull << ((s[r] + k) & 63);
  return m;
}

// the care wedges of a direction mask and the box test against them
int wbh_wedge_hit(uint64_t occ, double px, double py, double x0, double x1, double y0, double y1)
{
  static float tab[3 * 128];
  static int filled;
  if (!filled) {
    wb::fill_edge_table(tab);
    filled = 1;
  }
  wb::Wedges w;
  wb::care_wedges(occ, w, tab);
  return wb::wedge_hit(w, px, py, x0, x1, y0, y1) ? 1 : 0;
}

// the border wedges of the runs of >= 24 bins of occ (before/after as the
// exact pass finds them) and the box test against them
int wbh_border_wedge_hit(uint64_t occ, double px, double py, double x0, double x1, double y0,
                         double y1)
{
  static float tab[3 * 128];
  static int filled;
  if (!filled) {
    wb::fill_edge_table(tab);
    filled = 1;
  }
  int before[2] = {-1, -1}, after[2] = {-1, -1}, nr = 0;
  for (int st = 0; st < 64; st++) {
    if ((occ >> st) & 1) continue;
    if (!((occ >> ((st + 63) & 63)) & 1)) continue;
    int len = 0;
    while (len < 64 && !((occ >> ((st + len) & 63)) & 1)) len++;
    if (len >= 24 && nr < 2) {
      before[nr] = (st + 63) & 63;
      after[nr] = (st + len) & 63;
      nr++;
    }
  }
  wb::Wedges w;
  wb::border_wedges(before, after, nr, w, tab);
  return wb::wedge_hit(w, px, py, x0, x1, y0, y1) ? 1 : 0;
}

int wbh_atan2i_forced(double y, double x)
{
  ready();
  return wb::atan2i(y, x, (const double *)g_tan, (const double *)g_cos, (const double *)g_sin,
                    g_k, nullptr, true);
}

// the exact x87 emulation of the last step versus native long double
long wbh_last_exact(double y, double x)
{
  ready();
  double u = y / x;
  double c3 = wb::ATAN_C3 * (u * u * u);
  return (long)wb::atan_last_exact(y, x, c3, g_k.k_m, g_k.k_e);
}

long wbh_last_native(double y, double x)
{
  double u = y / x;
  return lrintl(0x40000000 / M_PIl * y / x - 1.1392738508503886e8 * (u * u * u));
}

// x87 op emulation vs native: op 0 mul, 1 div, 2 sub; returns 1 if equal
int wbh_x80_op_check(int op, double a, double b, double c)
{
  long double la = (long double)a * (long double)c;  // a 64-bit significand operand
  long double lb = b, want;
  wb::X80 xa = wb::x80_mul(wb::x80_from_double(a), wb::x80_from_double(c));
  wb::X80 xb = wb::x80_from_double(b), got;
  switch (op) {
  case 0: want = la * lb; got = wb::x80_mul(xa, xb); break;
  case 1: want = la / lb; got = wb::x80_div(xa, xb); break;
  default: want = la - lb; got = wb::x80_add(xa, wb::x80_neg(xb)); break;
  }
  // compare via lrint of a scaled copy and via exact reconstruction
  long double rec = ldexpl((long double)got.m, got.e);
  if (got.s) rec = -rec;
  if (got.m == 0) rec = 0;
  return rec == want;
}

int wbh_x80_lrint_check(double a, double c)
{
  long double la = (long double)a * (long double)c;
  wb::X80 xa = wb::x80_mul(wb::x80_from_double(a), wb::x80_from_double(c));
  return wb::x80_lrint(xa) == lrintl(la);
}

void wbh_atan_const(double *khi, double *klo, uint64_t *km, int *ke)
{
  wb::AtanConst k = wb::atan_const();
  *khi = k.k_hi;
  *klo = k.k_lo;
  *km = k.k_m;
  *ke = k.k_e;
}

int wbh_shape_in(int kind, const double *p, const double *q)
{
  wb::Shape s;
  if (kind == 0) s = wb::make_cylinder(p[0], p[1], p[2]);
  else if (kind == 1) s = wb::make_sphere(p[0], p[1], p[2], p[3]);
  else s = wb::make_hyperboloid(p[0], p[1], p[2], p[3], p[4]);
  return wb::shape_in(s, q[0], q[1], q[2]);
}

int wbh_shape_in_cube(int kind, const double *p, const double *c, double side)
{
  wb::Shape s;
  if (kind == 0) s = wb::make_cylinder(p[0], p[1], p[2]);
  else if (kind == 1) s = wb::make_sphere(p[0], p[1], p[2], p[3]);
  else s = wb::make_hyperboloid(p[0], p[1], p[2], p[3], p[4]);
  return wb::shape_in_cube(s, c[0], c[1], c[2], side);
}

int wbh_shape_intersect(int kind, const double *p, const double *c, double side)
{
  wb::Shape s;
  if (kind == 0) s = wb::make_cylinder(p[0], p[1], p[2]);
  else if (kind == 1) s = wb::make_sphere(p[0], p[1], p[2], p[3]);
  else s = wb::make_hyperboloid(p[0], p[1], p[2], p[3], p[4]);
  return wb::shape_intersect(s, c[0], c[1], c[2], side);
}

void wbh_to_flowsnake(int n, int *ex, int *ey) { wb::to_flowsnake(n, ex, ey); }
int wbh_from_flowsnake(int ex, int ey, int lo, int hi)
{
  int n;
  return wb::from_flowsnake(ex, ey, lo, hi, &n) ? n : INT32_MIN;
}
void wbh_snake_setup(const double *bbmin, const double *bbmax, double ts, double *out4, int *out3)
{
  wb::SnakeSetup s = wb::snake_setup(bbmin, bbmax, ts);
  out4[0] = s.cx; out4[1] = s.cy; out4[2] = s.spacing; out4[3] = s.radius;
  out3[0] = s.lo; out3[1] = s.hi; out3[2] = s.index;
}
void wbh_size_fit(const double *pts, int n, double *c, double *side) { wb::size_fit(pts, n, c, side); }
// the scan's histogram sector: the fast form (which = 0) or the exact one;
// returns sector * 2 + edge flag
int wbh_scan_sector(int which, double vx, double vy, double r)
{
  int edge = 0;
  const int sec = which == 0 ? wb::scan_sector(vx, vy, r, &edge)
                             : wb::scan_sector_exact(vx, vy, r, &edge);
  return sec * 2 + edge;
}
}

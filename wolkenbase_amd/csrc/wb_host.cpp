// wb_host.cpp -- host set-up for the GPU path (tables, constants, root cube,
// tiling).  Runs on the CPU with glibc, as the reference does.
#define _GNU_SOURCE 1
#include "wb_host.h"
#include "wb_flowsnake.h"

#include <math.h>
#include <string.h>

namespace wb {

static double bin_sin(int a) { return (double)sinl(a * M_PIl / 1073741824.); }
static double bin_cos(int a) { return (double)cosl(a * M_PIl / 1073741824.); }
static double bin_tan(int a) { return (double)tanl(a * M_PIl / 1073741824.); }

void fill_atan_tables(double *tan511, double *cos512, double *sin512)
{
  for (int i = 0; i < 511; i++) tan511[i] = bin_tan(i * 0x100000 - 0xff00000);
  for (int i = 0; i < 512; i++) {
    sin512[i] = bin_sin(i * 0x100000 - 0xff80000);
    cos512[i] = bin_cos(i * 0x100000 - 0xff80000);
  }
}

AtanConst atan_const()
{
  volatile long double k = 0x40000000 / M_PIl;
  long double kk = k;
  AtanConst c;
  c.k_hi = (double)kk;
  c.k_lo = (double)(kk - (long double)c.k_hi);
  // x87 layout: 64-bit significand with explicit integer bit, then
  // sign + 15-bit biased exponent.
  unsigned char raw[16];
  memset(raw, 0, sizeof raw);
  memcpy(raw, &kk, 10);
  uint64_t m;
  memcpy(&m, raw, 8);
  int be = (raw[8] | (raw[9] << 8)) & 0x7fff;
  c.k_m = m;
  c.k_e = be - 16383 - 63;
  return c;
}

void size_fit(const double *pts, int npts, double center[3], double *side_out)
{
  double mn[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, mx[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  for (int i = 0; i < npts; i++)
    for (int a = 0; a < 3; a++) {
      double v = pts[3 * i + a];
      if (v > mx[a]) mx[a] = v;
      if (v < mn[a]) mn[a] = v;
    }
  if (mx[2] <= mn[2] && mx[1] <= mn[1] && mx[0] <= mn[0]) {
    center[0] = center[1] = center[2] = 0;
    *side_out = 0;
    return;
  }
  // power-of-two side from the mean extent, then double until the cube,
  // aligned to side/16, holds every corner
  double side = (mx[0] + mx[1] + mx[2] - mn[0] - mn[1] - mn[2]) / 3;
  side /= significand(side);
  double x = mn[0] - side, y = mn[1] - side, z = mn[2] - side;
  while (x + side <= mx[0] || y + side <= mx[1] || z + side <= mx[2]) {
    side *= 2;
    x = (rint((mn[0] + mx[0]) / side * 8) - 8) * side / 16;
    y = (rint((mn[1] + mx[1]) / side * 8) - 8) * side / 16;
    z = (rint((mn[2] + mx[2]) / side * 8) - 8) * side / 16;
  }
  center[0] = x + side / 2;
  center[1] = y + side / 2;
  center[2] = z + side / 2;
  *side_out = side;
}

// Gosper-island square sizes, flowsnake.cpp:30-44
static const double kSquareSides[12] = {
    0.6583539906808145, 1.8501627472990723, 4.286014912881196, 12.6716160058597,
    32.85016274729906,  79.01914481623778,  243.0343734125204, 592.8501627472989,
    1510.850162747299,  4399.956311577825,  10731.850162747294, 29198.8501627473};
// counter ranges, flowsnake.cpp:56-57
static const int kLoLim[12] = {0, -4, -18, -214, -900, -10504, -44118, -514714,
                               -24242424, -25221004, -105928218, -1235829214};
static const int kHiLim[12] = {0, 2, 30, 128, 1500, 6302, 73530, 308828, 3603000,
                               15132602, 176547030, 741497528};

void bound_rect(const double *corners, int ncorners, double lbrthl[6])
{
  // BoundRect::include of each corner: bounds = left, bottom, -right, -top,
  // low, -high with dirbound(i*90deg) = x*cos + y*sin in binary angles
  double b[6] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
  for (int k = 0; k < ncorners; k++) {
    const double *o = corners + 3 * k;
    for (int i = 0; i < 4; i++) {
      int ang = i * DEG90;
      double nb = o[0] * bin_cos(ang) + o[1] * bin_sin(ang);
      if (nb < b[i]) b[i] = nb;
    }
    if (o[2] < b[4]) b[4] = o[2];
    if (-o[2] < b[5]) b[5] = -o[2];
  }
  lbrthl[0] = b[0];
  lbrthl[1] = b[1];
  lbrthl[2] = -b[2];
  lbrthl[3] = -b[3];
  lbrthl[4] = b[4];
  lbrthl[5] = -b[5];
}

SnakeSetup snake_setup(const double bbmin[3], const double bbmax[3], double tile_size)
{
  const double corners[6] = {bbmin[0], bbmin[1], bbmin[2], bbmax[0], bbmax[1], bbmax[2]};
  return snake_setup_corners(corners, 2, tile_size);
}

SnakeSetup snake_setup_corners(const double *corners, int ncorners, double tile_size)
{
  double br[6];
  bound_rect(corners, ncorners, br);
  double left = br[0], bottom = br[1], right = br[2], top = br[3], low = br[4], high = br[5];
  double side = right - left;
  if (top - bottom > side) side = top - bottom;
  if (high - low > side) side = high - low;
  SnakeSetup s;
  s.cx = (right + left) / 2;
  s.cy = (top + bottom) / 2;
  double best = INFINITY;
  int bi = 0;
  for (int i = 0; i < 12; i++) {
    double sp = side / kSquareSides[i];
    double diff = fabs(log(sp / tile_size));
    if (diff < best) {
      best = diff;
      bi = i;
    }
  }
  s.spacing = side / kSquareSides[bi];
  s.radius = s.spacing * 41 / 71;
  s.lo = kLoLim[bi];
  s.hi = kHiLim[bi];
  s.index = bi;
  return s;
}

void to_flowsnake(int n, int *ex, int *ey) { to_flowsnake_dev(n, ex, ey); }

bool from_flowsnake(int ex, int ey, int lo, int hi, int *n)
{
  return from_flowsnake_dev(ex, ey, lo, hi, n);
}

}  // namespace wb

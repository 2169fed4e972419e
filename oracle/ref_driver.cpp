// ref_driver.cpp -- extern "C" entry points over the reference's own
// geometry/angle/summation sources (angle.cpp, point.cpp, shape.cpp,
// manysum.cpp, ldecimal.cpp, manygcd.cpp), compiled from where they lie under the
// reference tree by oracle/Makefile into oracle/_ref/libwbref.so.
// Test infrastructure only: used to pin the oracle, never by the product.
#include "angle.h"
#include "point.h"
#include "shape.h"
#include "manysum.h"
#include "ldecimal.h"
#include "manygcd.h"

#include <cstring>
#include <vector>

extern "C" {
void ref_fill_tables() { fillTanTables(); }
int ref_ldecimal(double x, double toler, char *out, int cap)
{
  std::string s = ldecimal(x, toler);
  if ((int)s.size() >= cap) return -1;
  std::memcpy(out, s.c_str(), s.size() + 1);
  return (int)s.size();
}
double ref_manygcd(const double *v, int n, double toler)
{
  return manygcd(std::vector<double>(v, v + n), toler);
}
int ref_degtobin(double a) { return degtobin(a); }
double ref_bintodeg(int a) { return bintodeg(a); }
int ref_atan2i(double y, double x) { return atan2i(y, x); }
double ref_sin_bin(int a) { return sin(a); }
double ref_cos_bin(int a) { return cos(a); }
double ref_tan_bin(int a) { return tan(a); }
double ref_pairwisesum(double *a, unsigned n) { return pairwisesum(a, n); }
double ref_dist_xy(double ax, double ay, double bx, double by) { return dist(xy(ax, ay), xy(bx, by)); }
int ref_dir(double ax, double ay, double bx, double by) { return dir(xy(ax, ay), xy(bx, by)); }

static Shape *make(int kind, const double *p, Cylinder &c, Sphere &s, Hyperboloid &h, Paraboloid &pa)
{
  switch (kind) {
  case 0: c = Cylinder(xy(p[0], p[1]), p[2]); return &c;
  case 1: s = Sphere(xyz(p[0], p[1], p[2]), p[3]); return &s;
  case 2: h = Hyperboloid(xyz(p[0], p[1], p[2]), p[3], p[4]); return &h;
  default: pa = Paraboloid(xyz(p[0], p[1], p[2]), p[3]); return &pa;
  }
}
int ref_shape_in(int kind, const double *p, const double *q)
{
  Cylinder c; Sphere s; Hyperboloid h; Paraboloid pa;
  return make(kind, p, c, s, h, pa)->in(xyz(q[0], q[1], q[2]));
}
int ref_shape_in_cube(int kind, const double *p, const double *cc, double side)
{
  Cylinder c; Sphere s; Hyperboloid h; Paraboloid pa;
  Cube cube(xyz(cc[0], cc[1], cc[2]), side);
  return make(kind, p, c, s, h, pa)->in(cube);
}
int ref_shape_intersect(int kind, const double *p, const double *cc, double side)
{
  Cylinder c; Sphere s; Hyperboloid h; Paraboloid pa;
  Cube cube(xyz(cc[0], cc[1], cc[2]), side);
  return make(kind, p, c, s, h, pa)->intersect(cube);
}
int ref_cube_in(const double *cc, double side, const double *q)
{
  Cube cube(xyz(cc[0], cc[1], cc[2]), side);
  return cube.in(xyz(q[0], q[1], q[2]));
}
}

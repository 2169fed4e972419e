/*
 * wb_oracle.c -- CPU restatement of wolkenbase's octree build and ground
 * classification.
 *
 * THIS IS TEST INFRASTRUCTURE (the checker), NOT PART OF THE PRODUCT.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The HIP library never links or calls it.
 *
 * It follows the reference algorithm step by step, in plain C with the same
 * floating-point operation order (built with -ffp-contract=off, no fast-math,
 * x86-64 SSE2 doubles and x87 long double where the reference uses them).
 * Canonical semantics where the multi-threaded reference is scheduling
 * dependent:
 *   - points inside a leaf are kept in input (file) order;
 *   - tiles are visited single-threaded in flowsnake counter order and the
 *     last tile that contains a point decides its class.
 *
 * Pinning (each function is checked by tests/ against):
 *   wbo_atan2i, wbo_fill_tables ... oracle/_ref (reference angle.cpp built
 *                                   from source) bit for bit; round trip.
 *   wbo_shape_*                 ... oracle/_ref (reference shape.cpp) and the
 *                                   known answers of wolkentest.cpp:169-279.
 *   wbo_pairwisesum             ... oracle/_ref (reference manysum.cpp) and
 *                                   pairwisesum(odd,i)==i^2.
 *   wbo_lls3                    ... (-29/77, 51/77) known answer
 *                                   (wolkentest.cpp:744-759) and exact planes.
 *   wbo_build (octree)          ... invariants + brute force.
 *   wbo_to_flowsnake            ... baseSeven(baseFlow(n))==n, 7^k tiles.
 *   wbo_run (scan/classify)     ... invariants (flat ground, isolated point,
 *                                   order independence); end-to-end class
 *                                   bytes are "parity unpinned" against the
 *                                   reference itself (needs config.h, i.e.
 *                                   generated code, so not buildable here).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#include <stdio.h>

#define WBO_RECORDS 537            /* octree.h:40-41 */
#define WBO_MAX_LEVEL 21           /* deepest level of the canonical tree */
#define DEG30  0x0aaaaaab           /* angle.h:105 */
#define DEG45  0x10000000
#define DEG90  0x20000000
#define DEG144 0x33333333           /* angle.h:115 */
#define DEG180 0x40000000
#define SQRT7  2.6457513110645905905016      /* scan.h:23 */
#define SQRT_3_4 0.86602540378443864676372317 /* eisenstein.h:33 */

static double sqrd(double x) { return x * x; }   /* angle.cpp:36-39 */
static double cubd(double x) { return x * x * x; } /* angle.cpp:41-44 */

/* ------------------------------------------------------------------ */
/* Binary angles: angle.cpp:51-64, 117-155, 305-320                    */
/* ------------------------------------------------------------------ */
static double tan_tab[511], cos_tab[512], sin_tab[512];
static int tables_ready;

/* sin/cos/tan of a binary angle (2^31 = full turn), evaluated in long
 * double as angle.cpp:51-64 does, then narrowed to double. */
static double bsin(int a) { return (double)sinl(a * M_PIl / 1073741824.); }
static double bcos(int a) { return (double)cosl(a * M_PIl / 1073741824.); }
static double btan(int a) { return (double)tanl(a * M_PIl / 1073741824.); }

void wbo_fill_tables(void)
{
  /* 512 steps of 0x100000 over [-45deg,45deg) -- angle.cpp:305-320 */
  for (int i = 0; i < 511; i++)
    tan_tab[i] = btan(i * 0x100000 - 0xff00000);
  for (int i = 0; i < 512; i++) {
    sin_tab[i] = bsin(i * 0x100000 - 0xff80000);
    cos_tab[i] = bcos(i * 0x100000 - 0xff80000);
  }
  tables_ready = 1;
}

double wbo_sin_bin(int a) { return bsin(a); }
double wbo_cos_bin(int a) { return bcos(a); }

void wbo_get_tables(double *tan511, double *cos512, double *sin512)
{
  if (!tables_ready) wbo_fill_tables();
  memcpy(tan511, tan_tab, sizeof tan_tab);
  memcpy(cos512, cos_tab, sizeof cos_tab);
  memcpy(sin512, sin_tab, sizeof sin_tab);
}

/* Table-driven binary arctangent, angle.cpp:117-155.  Quadrant fold, a
 * 9-step binary search in tan_tab, a rotation by the table angle, then the
 * residual angle from a cubic in y/x evaluated in x87 long double and
 * rounded with lrintl (angle.cpp:151). */
int wbo_atan2i(double y, double x)
{
  int a = 0, step, k;
  double t;
  if (!tables_ready) wbo_fill_tables();
  if (x < 0) {                      /* fold into the right half plane */
    a += (y > 0) ? DEG180 : -DEG180;
    y = -y;
    x = -x;
  }
  if (y > x) {                      /* upper octant pair -> rotate -90 */
    a += DEG90;
    t = x; x = y; y = -t;
  }
  if (-y > x) {                     /* lower octant pair -> rotate +90 */
    a -= DEG90;
    t = x; x = -y; y = t;
  }
  t = y / x;
  for (step = DEG45 / 2; step > DEG45 / 1024; step /= 2) {
    k = (((a + DEG45) & 0x1ff00000) >> 20) - 1;
    a += (t > tan_tab[k]) ? step : -step;
  }
  k = 511 - (((a + DEG45) & 0x1ff00000) >> 20);
  t = x * cos_tab[k] - y * sin_tab[k];
  y = y * cos_tab[k] + x * sin_tab[k];
  x = t;
  {
    long double r = 0x40000000 / M_PIl * y / x - 1.1392738508503886e8 * cubd(y / x);
    a += (int)lrintl(r);
  }
  if (x == 0 && y == 0)
    a = 0;
  return a;
}

/* wbo_atan2i over arrays (test helper for brute-force checks) */
void wbo_atan2i_batch(const double *y, const double *x, long n, int *out)
{
  for (long i = 0; i < n; i++) out[i] = wbo_atan2i(y[i], x[i]);
}

/* ------------------------------------------------------------------ */
/* pairwisesum: manysum.cpp:120-154                                    */
/* The reference sums aligned power-of-two blocks, one block per set   */
/* bit of n (largest block first in memory), each block as a balanced  */
/* binary tree, and adds the block sums to 0 in ascending bit order.   */
/* ------------------------------------------------------------------ */
static double balanced(const double *a, unsigned n)
{
  if (n == 1) return a[0];
  return balanced(a, n / 2) + balanced(a + n / 2, n / 2);
}

double wbo_pairwisesum(const double *a, unsigned n)
{
  double acc = 0;
  for (int k = 0; k < 32; k++)
    if ((n >> k) & 1u) {
      unsigned off = (k == 31) ? 0 : ((n >> (k + 1)) << (k + 1));
      acc += balanced(a + off, 1u << k);
    }
  return acc;
}

/* ------------------------------------------------------------------ */
/* Small dense matrix Gauss-Jordan elimination: matrix.cpp:262-422     */
/* ------------------------------------------------------------------ */
typedef struct { int rows, cols; double *e; } mat;
#define M(m, r, c) ((m)->e[(r) * (m)->cols + (c)])

static void swap_rows(mat *m, int r0, int r1)
{
  for (int c = 0; c < m->cols; c++) {
    double t = M(m, r0, c); M(m, r0, c) = M(m, r1, c); M(m, r1, c) = t;
  }
}

/* matrix::rowop, matrix.cpp:262-356.  b may be NULL (then ignored). */
static void rowop(mat *a, mat *b, int row0, int row1, int piv)
{
  double slope = 0, minslope = INFINITY, det;
  int flags = 0, pivot;
  double *r0 = &M(a, row0, 0), *r1 = &M(a, row1, 0);
  if (piv >= 0 && r0[piv] == 0 && r1[piv] == 0)
    piv = -1;
  pivot = piv;
  if (piv >= 0 && r0[piv] == 0)
    flags = 9;
  for (int i = 0; piv < 0 && i < a->cols; i++)
    if (r0[i] != 0 || r1[i] != 0) {
      if (fabs(r0[i]) > fabs(r1[i]) || row0 >= row1) {
        slope = fabs(r1[i] / r0[i]);
        flags &= ~8;
      } else {
        slope = fabs(r0[i] / r1[i]);
        flags |= 8;
      }
      if (slope < minslope) {
        minslope = slope;
        flags = (flags >> 3) * 9;
        pivot = i;
      }
    }
  flags &= 1;
  if (flags) {
    swap_rows(a, row0, row1);
    if (b) swap_rows(b, row0, row1);
  }
  det = (pivot < 0) ? 0 : r0[pivot];
  if (det != 0 && det != 1) {
    for (int i = 0; i < a->cols; i++) r0[i] /= det;
    if (b) for (int i = 0; i < b->cols; i++) M(b, row0, i) /= det;
  }
  if (pivot >= 0)
    slope = r1[pivot];
  if (slope != 0 && row0 != row1) {
    for (int i = 0; i < a->cols; i++) r1[i] -= r0[i] * slope;
    if (b) for (int i = 0; i < b->cols; i++) M(b, row1, i) -= M(b, row0, i) * slope;
  }
}

/* matrix::findpivot, matrix.cpp:385-422 (ratio pivoting). */
static void findpivot(mat *a, mat *b, int row, int column)
{
  int pivotrow = -1;
  double maxratio = 0, sq[16], ratio;
  for (; pivotrow < row && column < a->cols; column++) {
    for (int i = row; i < a->rows; i++) {
      memset(sq, 0, sizeof sq);
      for (int j = column + 1; j < a->cols; j++)
        sq[j - column - 1] = sqrd(M(a, i, j));
      ratio = sqrd(M(a, i, column)) / wbo_pairwisesum(sq, a->cols - column);
      if (ratio > maxratio) {
        pivotrow = i;
        maxratio = ratio;
      }
    }
  }
  if (pivotrow > row) {
    swap_rows(a, pivotrow, row);
    if (b) swap_rows(b, pivotrow, row);
  }
}

/* matrix::gausselim, matrix.cpp:358-383. */
static void gausselim(mat *a, mat *b)
{
  for (int i = 0; i < a->rows; i++) {
    findpivot(a, b, i, i);
    for (int j = 0; j < a->rows; j++)
      rowop(a, b, i, j, i);
  }
  for (int i = a->rows - 1; i >= 0; i--)
    for (int j = 0; j < i; j++)
      rowop(a, b, i, j, i);
}

/* linearLeastSquares (leastsquares.cpp:28-43) for the N x 3 design
 * matrix [x y 1]: normal equations from pairwise sums (matrix.cpp:181-215),
 * Gauss-Jordan, NaN where a diagonal pivot is 0. */
void wbo_lls3(const double *x, const double *y, const double *z, unsigned n,
              double out[3])
{
  double *prod = malloc(sizeof(double) * (n ? n : 1));
  double ee[9], be[3];
  mat a = {3, 3, ee}, b = {3, 1, be};
  const double *col[3];
  double ones_dummy = 1;
  (void)ones_dummy;
  col[0] = x; col[1] = y; col[2] = NULL;  /* NULL = the column of ones */
  /* transmult: ret[i][j] = sum_k mt[i][k]*mt[j][k], j <= i */
  for (int i = 0; i < 3; i++)
    for (int j = 0; j <= i; j++) {
      for (unsigned k = 0; k < n; k++) {
        double vi = col[i] ? col[i][k] : 1.0, vj = col[j] ? col[j][k] : 1.0;
        prod[k] = vi * vj;
      }
      M(&a, i, j) = M(&a, j, i) = wbo_pairwisesum(prod, n);
    }
  /* mt * v */
  for (int i = 0; i < 3; i++) {
    for (unsigned k = 0; k < n; k++)
      prod[k] = (col[i] ? col[i][k] : 1.0) * z[k];
    M(&b, i, 0) = wbo_pairwisesum(prod, n);
  }
  gausselim(&a, &b);
  for (int i = 0; i < 3; i++)
    if (M(&a, i, i) == 0)
      M(&b, i, 0) = NAN;
  out[0] = be[0]; out[1] = be[1]; out[2] = be[2];
  free(prod);
}

/* General linear least squares, for the reference's own known answer
 * test (wolkentest.cpp:744-759): m is rows x cols row-major. */
void wbo_lls(const double *m, int rows, int cols, const double *v, double *out)
{
  double *mt = malloc(sizeof(double) * rows * cols);
  double *prod = malloc(sizeof(double) * rows);
  double *ae = malloc(sizeof(double) * cols * cols), *bev = malloc(sizeof(double) * cols);
  mat a = {cols, cols, ae}, b = {cols, 1, bev};
  for (int r = 0; r < rows; r++)
    for (int c = 0; c < cols; c++) mt[c * rows + r] = m[r * cols + c];
  for (int i = 0; i < cols; i++)
    for (int j = 0; j <= i; j++) {
      for (int k = 0; k < rows; k++) prod[k] = mt[i * rows + k] * mt[j * rows + k];
      M(&a, i, j) = M(&a, j, i) = wbo_pairwisesum(prod, rows);
    }
  for (int i = 0; i < cols; i++) {
    for (int k = 0; k < rows; k++) prod[k] = mt[i * rows + k] * v[k];
    M(&b, i, 0) = wbo_pairwisesum(prod, rows);
  }
  gausselim(&a, &b);
  for (int i = 0; i < cols; i++) {
    if (M(&a, i, i) == 0) M(&b, i, 0) = NAN;
    out[i] = M(&b, i, 0);
  }
  free(mt); free(prod); free(ae); free(bev);
}

/* ------------------------------------------------------------------ */
/* Shapes: shape.cpp:38-238, point.cpp:168-176, 362-365                */
/* ------------------------------------------------------------------ */
enum { SH_CYL = 0, SH_SPH = 1, SH_HYP = 2, SH_PAR = 3 };
typedef struct {
  int kind;
  double c[3];      /* centre (cylinder: xy; hyperboloid: v+(0,0,por)) */
  double r;         /* radius / radius of curvature */
  double por2, slope;
} wbo_shape;

void wbo_make_shape(wbo_shape *s, int kind, const double *p)
{
  memset(s, 0, sizeof *s);
  s->kind = kind;
  switch (kind) {
  case SH_CYL: s->c[0] = p[0]; s->c[1] = p[1]; s->r = p[2]; break;
  case SH_SPH: s->c[0] = p[0]; s->c[1] = p[1]; s->c[2] = p[2]; s->r = p[3]; break;
  case SH_HYP: {  /* Hyperboloid(v, r, s): shape.cpp:119-125 */
    double por = p[3] * sqrd(p[4]);
    s->slope = p[4];
    s->por2 = sqrd(por);
    s->c[0] = p[0] + 0; s->c[1] = p[1] + 0; s->c[2] = p[2] + por;
    break;
  }
  case SH_PAR: s->c[0] = p[0]; s->c[1] = p[1]; s->c[2] = p[2]; s->r = p[3]; break;
  }
}

int wbo_shape_in(const wbo_shape *s, const double *q)
{
  switch (s->kind) {
  case SH_CYL:
    return hypot(s->c[0] - q[0], s->c[1] - q[1]) <= s->r;
  case SH_SPH:
    return hypot(hypot(q[0] - s->c[0], q[1] - s->c[1]), q[2] - s->c[2]) <= s->r;
  case SH_HYP: {
    double xyd = hypot(s->c[0] - q[0], s->c[1] - q[1]);
    double zd = s->c[2] - q[2];
    if (s->slope > 0)
      return zd > 0 && sqrd(zd) - sqrd(xyd * s->slope) >= s->por2;
    return zd < 0 && sqrd(zd) - sqrd(xyd * s->slope) >= s->por2;
  }
  case SH_PAR: {
    double xyd = hypot(s->c[0] - q[0], s->c[1] - q[1]);
    double zd = s->c[2] - q[2];
    if (s->r == 0) return xyd == 0;
    return 2 * zd / s->r >= sqrd(xyd / s->r);
  }
  }
  return 0;
}

/* Shape::in(Cube&) -- all eight corners in (shape.cpp:45-61). */
int wbo_shape_in_cube(const wbo_shape *s, const double *cc, double side)
{
  for (int n = 0; n < 8; n++) {
    double q[3];
    q[0] = (n & 1) ? cc[0] + side / 2 : cc[0] - side / 2;
    q[1] = (n & 2) ? cc[1] + side / 2 : cc[1] - side / 2;
    q[2] = (n & 4) ? cc[2] + side / 2 : cc[2] - side / 2;
    if (!wbo_shape_in(s, q)) return 0;
  }
  return 1;
}

/* closestPoint clamp of one axis (shape.cpp:139-152 pattern). */
static double clamp_axis(double shape_c, double cube_c, double side)
{
  if (fabs(shape_c - cube_c) < side / 2) return shape_c;
  if (shape_c > cube_c) return cube_c + side / 2;
  return cube_c - side / 2;
}

/* Shape::intersect = in(closestPoint(cube)) (shape.cpp:63-66). */
int wbo_shape_intersect(const wbo_shape *s, const double *cc, double side)
{
  double q[3];
  q[0] = clamp_axis(s->c[0], cc[0], side);
  q[1] = clamp_axis(s->c[1], cc[1], side);
  q[2] = cc[2];
  switch (s->kind) {
  case SH_SPH: q[2] = clamp_axis(s->c[2], cc[2], side); break;
  case SH_HYP:
    if (s->slope > 0) q[2] -= side / 2;
    if (s->slope < 0) q[2] += side / 2;
    break;
  case SH_PAR:
    if (s->r > 0) q[2] -= side / 2;
    if (s->r < 0) q[2] += side / 2;
    break;
  default: break;
  }
  return wbo_shape_in(s, q);
}

/* flat-argument entry points for ctypes */
int wbo_in_flat(int kind, const double *p, const double *q)
{ wbo_shape s; wbo_make_shape(&s, kind, p); return wbo_shape_in(&s, q); }
int wbo_in_cube_flat(int kind, const double *p, const double *cc, double side)
{ wbo_shape s; wbo_make_shape(&s, kind, p); return wbo_shape_in_cube(&s, cc, side); }
int wbo_intersect_flat(int kind, const double *p, const double *cc, double side)
{ wbo_shape s; wbo_make_shape(&s, kind, p); return wbo_shape_intersect(&s, cc, side); }

/* ------------------------------------------------------------------ */
/* Octree::sizeFit, octree.cpp:268-310                                 */
/* ------------------------------------------------------------------ */
void wbo_sizefit(const double *pts, int npts, double center[3], double *side_out)
{
  double mn[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, mx[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
  double side, x, y, z;
  for (int i = 0; i < npts; i++)
    for (int a = 0; a < 3; a++) {
      if (pts[3 * i + a] > mx[a]) mx[a] = pts[3 * i + a];
      if (pts[3 * i + a] < mn[a]) mn[a] = pts[3 * i + a];
    }
  if (mx[2] <= mn[2] && mx[1] <= mn[1] && mx[0] <= mn[0]) {
    *side_out = 0;
    center[0] = center[1] = center[2] = 0;
    return;
  }
  side = (mx[0] + mx[1] + mx[2] - mn[0] - mn[1] - mn[2]) / 3;
  side /= significand(side);
  x = mn[0] - side;
  y = mn[1] - side;
  z = mn[2] - side;
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

/* ------------------------------------------------------------------ */
/* Canonical octree.  octree.cpp:199-347, 620-662, 849-876, 1295-1338: */
/* a full block (RECORDS distinct locations) splits one level when a   */
/* new location arrives, so the final tree is the one where a node is  */
/* internal iff its region holds more than RECORDS locations.  Child   */
/* index z*4+y*2+x from >= compares against the node centre.           */
/* ------------------------------------------------------------------ */
typedef struct {
  double c[3], side;
  int child[8];   /* >=0: node index; <= -2: leaf id (-2-id); -1: empty */
} onode;

typedef struct {
  double c[3], side;   /* leaf cube (Octree::cube(i) of its parent) */
  int depth;
  long first, count;   /* range in the ordered point list */
} oleaf;

typedef struct wbo_tree {
  long n;
  const double *x, *y, *z;
  onode *nodes; int nnodes, capnodes;
  oleaf *leaves; int nleaves, capleaves;
  long *order;          /* point indices in DFS order, input order in leaf */
  int maxdepth;
  int failed;           /* a level-21 cell holds more than RECORDS points */
} wbo_tree;

static int new_node(wbo_tree *t, const double *c, double side)
{
  if (t->nnodes == t->capnodes) {
    t->capnodes = t->capnodes ? 2 * t->capnodes : 1024;
    t->nodes = realloc(t->nodes, sizeof(onode) * t->capnodes);
  }
  onode *nd = &t->nodes[t->nnodes];
  memcpy(nd->c, c, sizeof nd->c);
  nd->side = side;
  for (int i = 0; i < 8; i++) nd->child[i] = -1;
  return t->nnodes++;
}

static int child_of(const double *c, double px, double py, double pz)
{
  int xb = px >= c[0], yb = py >= c[1], zb = pz >= c[2];
  return zb * 4 + yb * 2 + xb;
}

/* Octree::cube(n), octree.cpp:349-358 */
static void child_cube(const double *c, double side, int i, double *cc, double *cs)
{
  int xb = i & 1, yb = (i & 2) / 2, zb = (i & 4) / 4;
  cc[0] = c[0] + (2 * xb - 1) * side / 4;
  cc[1] = c[1] + (2 * yb - 1) * side / 4;
  cc[2] = c[2] + (2 * zb - 1) * side / 4;
  *cs = side / 2;
}

/* Recursive partition of idx[0..cnt) (input order preserved) under node nd. */
static void build_rec(wbo_tree *t, int nd, long *idx, long cnt, long *tmp, int depth, long *pos)
{
  long bucket[9] = {0};
  double c[3], side;
  memcpy(c, t->nodes[nd].c, sizeof c);
  side = t->nodes[nd].side;
  for (long k = 0; k < cnt; k++) {
    long p = idx[k];
    bucket[child_of(c, t->x[p], t->y[p], t->z[p]) + 1]++;
  }
  for (int i = 0; i < 8; i++) bucket[i + 1] += bucket[i];
  {
    long fill[8];
    for (int i = 0; i < 8; i++) fill[i] = bucket[i];
    for (long k = 0; k < cnt; k++) {
      long p = idx[k];
      tmp[fill[child_of(c, t->x[p], t->y[p], t->z[p])]++] = p;
    }
    memcpy(idx, tmp, sizeof(long) * cnt);
  }
  for (int i = 0; i < 8; i++) {
    long lo = bucket[i], m = bucket[i + 1] - bucket[i];
    double cc[3], cs;
    if (m == 0) continue;
    child_cube(c, side, i, cc, &cs);
    if (m > WBO_RECORDS && depth + 1 >= WBO_MAX_LEVEL) {
      /* The reference splits once more and drops the point it was putting
       * (OctStore::put, octree.cpp:849-876, ignores the second put's
       * result), once per further point: a tree that depends on the order
       * of arrival.  The canonical tree stops at level 21 (the cell index
       * of the GPU build has 21 bits per axis) and the build fails. */
      t->failed = 1;
      continue;
    }
    if (m > WBO_RECORDS) {
      int ch = new_node(t, cc, cs);
      t->nodes[nd].child[i] = ch;
      build_rec(t, ch, idx + lo, m, tmp, depth + 1, pos);
    } else {
      if (t->nleaves == t->capleaves) {
        t->capleaves = t->capleaves ? 2 * t->capleaves : 1024;
        t->leaves = realloc(t->leaves, sizeof(oleaf) * t->capleaves);
      }
      oleaf *lf = &t->leaves[t->nleaves];
      memcpy(lf->c, cc, sizeof cc);
      lf->side = cs;
      lf->depth = depth + 1;
      lf->first = *pos;
      lf->count = m;
      memcpy(t->order + *pos, idx + lo, sizeof(long) * m);
      *pos += m;
      if (depth + 1 > t->maxdepth) t->maxdepth = depth + 1;
      t->nodes[nd].child[i] = -2 - t->nleaves;
      t->nleaves++;
    }
  }
}

wbo_tree *wbo_build(const double *x, const double *y, const double *z, long n,
                    const double center[3], double side)
{
  wbo_tree *t = calloc(1, sizeof *t);
  long *idx = malloc(sizeof(long) * (n ? n : 1)), *tmp = malloc(sizeof(long) * (n ? n : 1));
  long pos = 0;
  t->n = n; t->x = x; t->y = y; t->z = z;
  t->order = malloc(sizeof(long) * (n ? n : 1));
  for (long i = 0; i < n; i++) idx[i] = i;
  new_node(t, center, side);
  if (n > 0)
    build_rec(t, 0, idx, n, tmp, 0, &pos);
  free(idx); free(tmp);
  return t;
}

void wbo_free(wbo_tree *t)
{
  if (!t) return;
  free(t->nodes); free(t->leaves); free(t->order); free(t);
}

int wbo_nleaves(const wbo_tree *t) { return t->nleaves; }
int wbo_nnodes(const wbo_tree *t) { return t->nnodes; }
int wbo_maxdepth(const wbo_tree *t) { return t->maxdepth; }
int wbo_failed(const wbo_tree *t) { return t->failed; }

void wbo_leaves(const wbo_tree *t, double *cen, double *side, int *depth,
                long *first, long *count)
{
  for (int i = 0; i < t->nleaves; i++) {
    const oleaf *l = &t->leaves[i];
    if (cen) { cen[3 * i] = l->c[0]; cen[3 * i + 1] = l->c[1]; cen[3 * i + 2] = l->c[2]; }
    if (side) side[i] = l->side;
    if (depth) depth[i] = l->depth;
    if (first) first[i] = l->first;
    if (count) count[i] = l->count;
  }
}

void wbo_order(const wbo_tree *t, long *out) { memcpy(out, t->order, sizeof(long) * t->n); }

/* Octree::findBlocks, octree.cpp:234-251: DFS, testing each internal
 * node's own cube and each leaf child's cube with Shape::intersect. */
typedef struct { int *v; int n, cap; } ivec;
static void ivec_push(ivec *a, int x)
{
  if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 64; a->v = realloc(a->v, sizeof(int) * a->cap); }
  a->v[a->n++] = x;
}

static void find_blocks(const wbo_tree *t, int nd, const wbo_shape *s, ivec *out)
{
  const onode *n = &t->nodes[nd];
  if (!wbo_shape_intersect(s, n->c, n->side)) return;
  for (int i = 0; i < 8; i++) {
    int ch = n->child[i];
    if (ch <= -2) {
      double cc[3], cs;
      child_cube(n->c, n->side, i, cc, &cs);
      if (wbo_shape_intersect(s, cc, cs)) ivec_push(out, -2 - ch);
    } else if (ch >= 0)
      find_blocks(t, ch, s, out);
  }
}

/* OctStore::pointsIn (octree.cpp:1214-1242), unsorted: whole leaf when all
 * eight corners of its cube are in, otherwise a per-point test.  Returns
 * the count; writes up to cap indices (input numbering) to out. */
long wbo_points_in(const wbo_tree *t, const wbo_shape *s, long *out, long cap)
{
  ivec bl = {0};
  long cnt = 0;
  if (t->nnodes == 0) return 0;
  find_blocks(t, 0, s, &bl);
  for (int b = 0; b < bl.n; b++) {
    const oleaf *l = &t->leaves[bl.v[b]];
    int all = wbo_shape_in_cube(s, l->c, l->side);
    for (long k = 0; k < l->count; k++) {
      long p = t->order[l->first + k];
      if (all) {
        if (out && cnt < cap) out[cnt] = p;
        cnt++;
      } else {
        double q[3] = {t->x[p], t->y[p], t->z[p]};
        if (wbo_shape_in(s, q)) {
          if (out && cnt < cap) out[cnt] = p;
          cnt++;
        }
      }
    }
  }
  free(bl.v);
  return cnt;
}

long wbo_points_in_flat(const wbo_tree *t, int kind, const double *p, long *out, long cap)
{ wbo_shape s; wbo_make_shape(&s, kind, p); return wbo_points_in(t, &s, out, cap); }

/* ------------------------------------------------------------------ */
/* Flowsnake tiling: flowsnake.cpp:46-136, 208-271; eisenstein.cpp    */
/* ------------------------------------------------------------------ */
static const double square_sides[12] = {   /* flowsnake.cpp:30-44 */
  0.6583539906808145, 1.8501627472990723, 4.286014912881196, 12.6716160058597,
  32.85016274729906, 79.01914481623778, 243.0343734125204, 592.8501627472989,
  1510.850162747299, 4399.956311577825, 10731.850162747294, 29198.8501627473};
static const unsigned char flow_tab[6][7] = {   /* flowsnake.cpp:46-54 */
  {0x52, 0x05, 0x06, 0x24, 0x33, 0x40, 0x01}, {0x31, 0x10, 0x12, 0x05, 0x43, 0x54, 0x16},
  {0x46, 0x24, 0x21, 0x10, 0x53, 0x35, 0x22}, {0x31, 0x10, 0x03, 0x54, 0x36, 0x35, 0x22},
  {0x46, 0x24, 0x13, 0x35, 0x42, 0x40, 0x01}, {0x52, 0x05, 0x23, 0x40, 0x51, 0x54, 0x16}};
static const int lo_lim[12] = {0, -4, -18, -214, -900, -10504, -44118, -514714,
                               -24242424, -25221004, -105928218, -1235829214};
static const int hi_lim[12] = {0, 2, 30, 128, 1500, 6302, 73530, 308828, 3603000,
                               15132602, 176547030, 741497528};

/* iToFlowsnake: digit-wise state machine over 11 base-7 digits. */
static int i_to_flowsnake(int n)
{
  int dig[11], ori = 0;
  long long v = (long long)n + 1235829214LL;
  for (int i = 0; i < 11; i++) { dig[i] = (int)(v % 7); v /= 7; }
  v = 0;
  for (int i = 10; i >= 0; i--) {
    int e = flow_tab[ori][dig[i]];
    ori = e >> 4;
    v = 7 * v + (e & 7);
  }
  return (int)(v - 988663371LL);
}

/* baseFlow: centred base-7 digits -> Eisenstein integer in base 2-w
 * (flowsnake.cpp:67-73, 112-131).  (x,y) means x + y*w. */
static void base_flow(int m, int *ex, int *ey);

void wbo_to_flowsnake(int n, int *ex, int *ey) { base_flow(i_to_flowsnake(n), ex, ey); }
void wbo_base_flow(int m, int *ex, int *ey) { base_flow(m, ex, ey); }
int wbo_i_to_flowsnake(int n) { return i_to_flowsnake(n); }

static void base_flow(int m, int *ex, int *ey)
{
  long long rx = 0, ry = 0, px = 1, py = 0;
  while (m) {
    int d = m % 7, dy, dx;
    if (d > 3) d -= 7;
    if (d < -3) d += 7;
    m = (m - d) / 7;
    dy = (d + 4) / 3 - 1;
    dx = d - 2 * dy;
    /* ret += digit * pow ; (a+bw)(c+dw) = (ac-bd) + (ad+bc-bd)w */
    rx += dx * px - dy * py;
    ry += dx * py + dy * px - dy * py;
    {  /* pow *= (2 - w) */
      long long nx = px * 2 - py * (-1), ny = px * (-1) + py * 2 - py * (-1);
      px = nx; py = ny;
    }
  }
  *ex = (int)rx;
  *ey = (int)ry;
}

typedef struct {
  double cx, cy, spacing;
  int lo, hi, index;
} wbo_snake;

/* BoundRect of the header corners at orientation 0 (boundrect.cpp:61-74,
 * point.cpp dirbound) followed by Flowsnake::setSize (flowsnake.cpp:208-229)
 * with the cube built in wolkencanvas.cpp:511-519. */
/* BoundRect::include of each corner at orientation 0 (boundrect.cpp:53-70);
 * out: left, bottom, right, top, low, high (boundrect.h:41-64) */
void wbo_bound_rect(const double *corners, int ncorners, double out[6])
{
  double b[6] = {INFINITY, INFINITY, INFINITY, INFINITY, INFINITY, INFINITY};
  for (int k = 0; k < ncorners; k++) {
    const double *o = corners + 3 * k;
    for (int i = 0; i < 4; i++) {
      int ang = i * DEG90;
      double nb = o[0] * bcos(ang) + o[1] * bsin(ang);
      if (nb < b[i]) b[i] = nb;
    }
    if (o[2] < b[4]) b[4] = o[2];
    if (-o[2] < b[5]) b[5] = -o[2];
  }
  out[0] = b[0]; out[1] = b[1]; out[2] = -b[2]; out[3] = -b[3]; out[4] = b[4]; out[5] = -b[5];
}

void wbo_snake_setup_corners(const double *corners, int ncorners, double tile_size,
                             wbo_snake *sn);
void wbo_snake_setup(const double bbmin[3], const double bbmax[3], double tile_size, wbo_snake *sn)
{
  const double c[6] = {bbmin[0], bbmin[1], bbmin[2], bbmax[0], bbmax[1], bbmax[2]};
  wbo_snake_setup_corners(c, 2, tile_size, sn);
}

void wbo_snake_setup_corners(const double *corners, int ncorners, double tile_size,
                             wbo_snake *sn)
{
  double br[6];
  double side, cen[3], best = INFINITY;
  int bi = 0;
  wbo_bound_rect(corners, ncorners, br);
  {
    double left = br[0], bottom = br[1], right = br[2], top = br[3], low = br[4], high = br[5];
    side = right - left;
    if (top - bottom > side) side = top - bottom;
    if (high - low > side) side = high - low;
    cen[0] = (right + left) / 2;
    cen[1] = (top + bottom) / 2;
    cen[2] = (high + low) / 2;
  }
  for (int i = 0; i < 12; i++) {
    double sp = side / square_sides[i];
    double diff = fabs(log(sp / tile_size));
    if (diff < best) { best = diff; bi = i; }
  }
  sn->cx = cen[0];
  sn->cy = cen[1];
  sn->spacing = side / square_sides[bi];
  sn->lo = lo_lim[bi];
  sn->hi = hi_lim[bi];
  sn->index = bi;
}

/* Flowsnake::cyl, flowsnake.cpp:263-271 */
void wbo_snake_cyl(const wbo_snake *sn, int ex, int ey, double *cx, double *cy, double *r)
{
  double zr = ex - ey / 2., zi = ey * SQRT_3_4;
  zr *= sn->spacing;
  zi *= sn->spacing;
  *cx = zr + sn->cx;
  *cy = zi + sn->cy;
  *r = sn->spacing * 41 / 71;
}

/* baseSeven (flowsnake.cpp:75-90): the least-norm remainder mod 2-w is
 * unique, and w == 2 (mod 2-w), so the digit is (x+2y) mod 7 centred. */
int wbo_base_seven(int ex, int ey)
{
  long long x = ex, y = ey, ret = 0, p7 = 1;
  while (x * x + y * y - x * y) {
    long long d = ((x + 2 * y) % 7 + 7) % 7, dy, dx;
    if (d > 3) d -= 7;
    dy = (d + 4) / 3 - 1;
    dx = d - 2 * dy;
    x -= dx; y -= dy;
    {   /* (x+yw)/(2-w) = ((3x-y) + (x+2y)w)/7 */
      long long nx = (3 * x - y) / 7, ny = (x + 2 * y) / 7;
      x = nx; y = ny;
    }
    ret += d * p7;
    p7 *= 7;
  }
  return (int)ret;
}

/* ------------------------------------------------------------------ */
/* Scan / postscan / classify: scan.cpp:31-179, classify.cpp:67-173    */
/* ------------------------------------------------------------------ */
typedef struct {
  int npoints, nground, tree_flags;
  double density, hs_scan, hs, height;
  double tilt_x, tilt_y;   /* the plane slope after the clamp, scan.cpp:68-73 */
} wbo_tile;

typedef struct {
  double tile_size, max_slope, thickness, min_hs;
} wbo_params;

typedef struct {
  long n_class1, n_class2, n_unclassified;
  long ntiles, nonempty_tiles;
  long sector_borderline;   /* atan2 sector decisions near a half-integer */
} wbo_stats;

static int cmp_int(const void *a, const void *b)
{
  int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

/* surround(), classify.cpp:67-94, applied to the complete direction set:
 * true iff more than one distinct direction and every gap between
 * successive sorted directions, and the wrap-around gap, is under 144deg.
 * (The reference prunes directions and stops early; neither changes the
 * answer on the complete set -- see the equivalence note in the header of
 * the HIP classify kernel.) */
int wbo_surround(int *dirs, int m)
{
  int u = 0, ok;
  qsort(dirs, m, sizeof(int), cmp_int);
  for (int i = 0; i < m; i++)
    if (i == 0 || dirs[i] != dirs[u - 1]) dirs[u++] = dirs[i];
  ok = u > 1;
  if (u == 0) return 0;
  for (int i = 1; i < u; i++)
    if ((int)(((unsigned)dirs[i] - (unsigned)dirs[i - 1]) & INT_MAX) >= DEG144) ok = 0;
  if ((int)(((unsigned)dirs[0] - (unsigned)dirs[u - 1]) & INT_MAX) >= DEG144) ok = 0;
  return ok;
}

/* Ground test of one point p against the tree (classify.cpp:129-164):
 * D = { dir(p,q) : q in pointsIn(downward hyperboloid), xy(q) != xy(p) }.
 *
 * Exhaustive mode builds the whole of D and applies surround() once.  Fast
 * mode visits the leaves of findBlocks in a strided order and stops as soon
 * as the directions gathered so far already satisfy surround(): adding
 * directions can only shrink gaps, so the complete D then satisfies it too.
 * That monotonicity needs every direction to lie in [-2^30, 2^30) (the
 * wrap-around gap of the reference's sorted-int test is then the circular
 * gap); wbo_run enables fast mode only when the cloud's extent is below
 * 1e7 coordinate units, where the smallest non-zero dy/|dx| keeps every
 * direction at least 30 units from +180 deg. */
static int g_exhaustive;
void wbo_set_exhaustive(int on) { g_exhaustive = on; }

static int surround_copy(const int *dirs, int m, int **tmp, long *tcap)
{
  if (m > *tcap) { *tcap = 2 * m; *tmp = realloc(*tmp, sizeof(int) * *tcap); }
  memcpy(*tmp, dirs, sizeof(int) * m);
  return wbo_surround(*tmp, m);
}

static long stride_for(long n)
{
  /* a step coprime to n near 0.618 n, so consecutive visits spread out */
  long h = (long)(n * 0.6180339887498949);
  if (h < 1) h = 1;
  for (;; h++) {
    long a = n, b = h;
    while (b) { long t = a % b; a = b; b = t; }
    if (a == 1) return h;
  }
}

static int classify_point_ex(const wbo_tree *t, long p, double hs, const wbo_params *pr,
                             int fast, int **dbuf, long *dcap, int **tbuf, long *tcap)
{
  wbo_shape h;
  double v[5];
  ivec bl = {0};
  int m = 0, res;
  long next_check = 8;
  v[0] = t->x[p]; v[1] = t->y[p]; v[2] = t->z[p] - pr->thickness;
  v[3] = hs; v[4] = pr->max_slope;
  wbo_make_shape(&h, SH_HYP, v);
  if (t->nnodes) find_blocks(t, 0, &h, &bl);
  {
    long L = bl.n, st = L ? stride_for(L) : 1, j, b = 0;
    for (j = 0; j < L; j++, b = (b + st) % L) {
      const oleaf *l = &t->leaves[bl.v[b]];
      int all = wbo_shape_in_cube(&h, l->c, l->side);
      if (m + l->count > *dcap) {
        *dcap = 2 * (m + l->count);
        *dbuf = realloc(*dbuf, sizeof(int) * *dcap);
      }
      for (long k = 0; k < l->count; k++) {
        long q = t->order[l->first + k];
        double qq[3] = {t->x[q], t->y[q], t->z[q]};
        if (!all && !wbo_shape_in(&h, qq)) continue;
        if (hypot(t->x[p] - t->x[q], t->y[p] - t->y[q]))
          (*dbuf)[m++] = wbo_atan2i(t->y[q] - t->y[p], t->x[q] - t->x[p]);
      }
      if (fast && m >= next_check) {
        if (surround_copy(*dbuf, m, tbuf, tcap)) { free(bl.v); return 1; }
        next_check = 2 * m;
      }
    }
  }
  free(bl.v);
  res = wbo_surround(*dbuf, m);
  return res ? 1 : 2;
}

static int classify_point(const wbo_tree *t, long p, double hs, const wbo_params *pr, int fast,
                          int **dbuf, long *dcap, int **tbuf, long *tcap)
{
  return classify_point_ex(t, p, hs, pr, fast && !g_exhaustive, dbuf, dcap, tbuf, tcap);
}

/* Tile lookup by Eisenstein address (harray<Tile> stand-in). */
typedef struct { int x0, y0, w, h; int *idx; } tgrid;

static int tgrid_get(const tgrid *g, int ex, int ey)
{
  int gx = ex - g->x0, gy = ey - g->y0;
  if (gx < 0 || gy < 0 || gx >= g->w || gy >= g->h) return -1;
  return g->idx[(long)gy * g->w + gx];
}

/* The whole pass of wolkencanvas.cpp:469-565 / threads.cpp:619-660 on a
 * single thread: build, SCAN every tile, POSTSCAN every tile, CLASSIFY every
 * tile (last tile wins).  xi/yi/zi are LAS int32 coordinates; decode as
 * las.cpp:808.  cls_out[n] receives 0 (never classified), 1 or 2 per input
 * point.  tiles_out (may be NULL) receives hi-lo+1 records in counter order.
 * Returns the number of tiles. */
static long run_impl(const int32_t *xi, const int32_t *yi, const int32_t *zi, const double *fx,
                     const double *fy, const double *fz, long n, const double scale[3],
                     const double offset[3], const double res[3], const double *corners,
                     int ncorners, const wbo_params *pr, uint8_t *cls_out,
                     wbo_tile *tiles_out, wbo_stats *st, double *x_out, double *y_out,
                     double *z_out);

long wbo_run(const int32_t *xi, const int32_t *yi, const int32_t *zi, long n,
             const double scale[3], const double offset[3],
             const double bbmin[3], const double bbmax[3],
             const wbo_params *pr, uint8_t *cls_out, wbo_tile *tiles_out,
             wbo_stats *st, double *x_out, double *y_out, double *z_out)
{
  const double c[6] = {bbmin[0], bbmin[1], bbmin[2], bbmax[0], bbmax[1], bbmax[2]};
  return run_impl(xi, yi, zi, NULL, NULL, NULL, n, scale, offset, scale, c, 2, pr,
                  cls_out, tiles_out, st, x_out, y_out, z_out);
}

/* The same pass over coordinates already decoded (LasHeader::readPoint's
 * location, las.cpp:896) with the root cube and tiling of several header
 * corners (wolkencli.cpp:85-90, wolkencanvas.cpp:136-146); res: the
 * per-axis spacing of the coordinate lattices (the extent check of fast
 * mode uses it in place of the scale). */
long wbo_run_xyz(const double *x, const double *y, const double *z, long n, const double res[3],
                 const double *corners, int ncorners, const wbo_params *pr, uint8_t *cls_out,
                 wbo_tile *tiles_out, wbo_stats *st)
{
  return run_impl(NULL, NULL, NULL, x, y, z, n, NULL, NULL, res, corners, ncorners, pr, cls_out,
                  tiles_out, st, NULL, NULL, NULL);
}

static long run_impl(const int32_t *xi, const int32_t *yi, const int32_t *zi, const double *fx,
                     const double *fy, const double *fz, long n, const double scale[3],
                     const double offset[3], const double res[3], const double *corners,
                     int ncorners, const wbo_params *pr, uint8_t *cls_out,
                     wbo_tile *tiles_out, wbo_stats *st, double *x_out, double *y_out,
                     double *z_out)
{
  double *x = x_out ? x_out : malloc(sizeof(double) * (n ? n : 1));
  double *y = y_out ? y_out : malloc(sizeof(double) * (n ? n : 1));
  double *z = z_out ? z_out : malloc(sizeof(double) * (n ? n : 1));
  double center[3], side, ext[4];
  wbo_snake sn;
  wbo_tree *t;
  long ntiles, *pbuf = NULL, pcap = 0;
  int *dbuf = NULL, *tbuf = NULL, fast;
  long dcap = 0, tcap = 0;
  int *ex, *ey;
  wbo_tile *tl;
  tgrid g;
  wbo_stats local;
  if (!st) st = &local;
  memset(st, 0, sizeof *st);
  if (!tables_ready) wbo_fill_tables();
  for (long i = 0; i < n; i++) {
    if (fx) {
      x[i] = fx[i]; y[i] = fy[i]; z[i] = fz[i];
    } else {
      x[i] = (offset[0] + scale[0] * xi[i]) * 1.0;
      y[i] = (offset[1] + scale[1] * yi[i]) * 1.0;
      z[i] = (offset[2] + scale[2] * zi[i]) * 1.0;
    }
  }
  wbo_sizefit(corners, ncorners, center, &side);
  t = wbo_build(x, y, z, n, center, side);
  if (t->failed) {
    wbo_free(t);
    if (!x_out) free(x);
    if (!y_out) free(y);
    if (!z_out) free(z);
    return -1;
  }
  wbo_snake_setup_corners(corners, ncorners, pr->tile_size, &sn);
  ext[0] = ext[2] = INFINITY;
  ext[1] = ext[3] = -INFINITY;
  for (int k = 0; k < ncorners; k++)
    for (int a = 0; a < 2; a++) {
      ext[2 * a] = fmin(ext[2 * a], corners[3 * k + a]);
      ext[2 * a + 1] = fmax(ext[2 * a + 1], corners[3 * k + a]);
    }
  ntiles = (long)sn.hi - sn.lo + 1;
  ex = malloc(sizeof(int) * ntiles);
  ey = malloc(sizeof(int) * ntiles);
  tl = calloc(ntiles, sizeof(wbo_tile));
  {
    int x0 = INT_MAX, y0 = INT_MAX, x1 = INT_MIN, y1 = INT_MIN;
    for (long k = 0; k < ntiles; k++) {
      wbo_to_flowsnake((int)(sn.lo + k), &ex[k], &ey[k]);
      if (ex[k] < x0) x0 = ex[k];
      if (ex[k] > x1) x1 = ex[k];
      if (ey[k] < y0) y0 = ey[k];
      if (ey[k] > y1) y1 = ey[k];
    }
    g.x0 = x0; g.y0 = y0; g.w = x1 - x0 + 1; g.h = y1 - y0 + 1;
    g.idx = malloc(sizeof(int) * (long)g.w * g.h);
    for (long k = 0; k < (long)g.w * g.h; k++) g.idx[k] = -1;
    for (long k = 0; k < ntiles; k++) g.idx[(long)(ey[k] - y0) * g.w + (ex[k] - x0)] = (int)k;
  }
  st->ntiles = ntiles;
  /* SCAN (scan.cpp:31-140) */
  for (long k = 0; k < ntiles; k++) {
    double cx, cy, r;
    wbo_shape cyl;
    double cp[3];
    long cnt;
    wbo_snake_cyl(&sn, ex[k], ey[k], &cx, &cy, &r);
    cp[0] = cx; cp[1] = cy; cp[2] = r;
    wbo_make_shape(&cyl, SH_CYL, cp);
    cnt = wbo_points_in(t, &cyl, pbuf, pcap);
    if (cnt > pcap) {
      pcap = cnt * 2 + 64;
      pbuf = realloc(pbuf, sizeof(long) * pcap);
      cnt = wbo_points_in(t, &cyl, pbuf, pcap);
    }
    if (cnt == 0) continue;
    {
      double *px = malloc(sizeof(double) * cnt * 4), *py = px + cnt, *pz = py + cnt, *uz = pz + cnt;
      double sl[3], sx, sy, bottom = INFINITY, bottom2 = INFINITY, top = -INFINITY;
      double density = 0;
      int histo[7] = {0}, tree = 0;
      long nbottom = 0;
      for (long i = 0; i < cnt; i++) {
        long p = pbuf[i];
        px[i] = x[p] - cx;
        py[i] = y[p] - cy;
        pz[i] = z[p] - 0;
      }
      wbo_lls3(px, py, pz, (unsigned)cnt, sl);
      sx = sl[0]; sy = sl[1];
      {
        double len = hypot(sx, sy);
        if (len > 1) { sx /= len; sy /= len; }
      }
      if (isnan(sx) || isnan(sy)) { sx = 0; sy = 0; }
      for (long i = 0; i < cnt; i++) {
        double dz = sy * py[i] + sx * px[i];   /* dot(slope, xy) point.cpp:178-181 */
        uz[i] = pz[i] - dz;
        if (uz[i] < bottom) { bottom2 = bottom; bottom = uz[i]; }
        if (uz[i] > top) top = uz[i];
      }
      if (isinf(bottom2)) bottom2 = bottom;
      for (long i = 0; i < cnt; i++) {
        if (!(uz[i] < bottom2 + 2 * r)) continue;
        nbottom++;
        {
          double th = atan2(py[i], px[i]) * 3 / M_PI;
          long sec = lrint(th);
          double fr = th - floor(th);
          if (fabs(fr - 0.5) < 1e-12) st->sector_borderline++;
          if (sec < 0) sec += 6;
          sec = (sec % 6) + 1;
          if (hypot(px[i], py[i]) < r / SQRT7) sec = 0;
          histo[sec]++;
        }
      }
      for (int i = 0; i < 7; i++) density += histo[i] * histo[i];
      if (cnt > nbottom && density < 7) tree = 1;
      density = sqrt(density) * SQRT7 / sqrd(r) / M_PI;
      if (cnt > nbottom && density < 0.5) tree = 1;
      if (top - bottom > 1.5) tree = 1;
      tl[k].npoints = (int)cnt;
      tl[k].nground = 0;
      tl[k].density = density;
      tl[k].tree_flags = tree;
      tl[k].hs_scan = sqrt(1 / density + sqrd(pr->min_hs));
      tl[k].hs = tl[k].hs_scan;
      tl[k].height = top - bottom;
      tl[k].tilt_x = sx;
      tl[k].tilt_y = sy;
      st->nonempty_tiles++;
      free(px);
    }
  }
  /* POSTSCAN (scan.cpp:142-179) */
  {
    static const int r1x[6] = {1, 1, 0, -1, -1, 0}, r1y[6] = {0, 1, 1, 0, -1, -1};
    for (long k = 0; k < ntiles; k++) {
      int i = 1, ring, nontree, count = 0;
      if (!tl[k].npoints) continue;
      do {
        ring = nontree = 0;
        for (int j = 0; j < 6 && (tl[k].tree_flags & 1); j++) {
          int o = tgrid_get(&g, ex[k] + r1x[j] * i, ey[k] + r1y[j] * i);
          if (o >= 0 && tl[o].npoints) {
            ring++;
            if (tl[o].tree_flags & 1) count++;
            else nontree++;
          }
        }
        ++i;
      } while (ring && !nontree);
      tl[k].hs = sqrt(sqrd(tl[k].hs) + sqrd(count * sn.spacing / 6));
    }
  }
  /* CLASSIFY (classify.cpp:96-173), last tile wins */
  memset(cls_out, 0, n);
  fast = (ext[1] - ext[0]) < 1e7 * res[1] && (ext[3] - ext[2]) < 1e7 * res[0];
  for (long k = 0; k < ntiles; k++) {
    double cx, cy, r, cp[3];
    wbo_shape cyl;
    long cnt;
    if (!tl[k].npoints) continue;
    wbo_snake_cyl(&sn, ex[k], ey[k], &cx, &cy, &r);
    cp[0] = cx; cp[1] = cy; cp[2] = r;
    wbo_make_shape(&cyl, SH_CYL, cp);
    cnt = wbo_points_in(t, &cyl, pbuf, pcap);
    if (cnt > pcap) {
      pcap = cnt * 2 + 64;
      pbuf = realloc(pbuf, sizeof(long) * pcap);
      cnt = wbo_points_in(t, &cyl, pbuf, pcap);
    }
    for (long i = 0; i < cnt; i++) {
      long p = pbuf[i];
      int c = classify_point(t, p, tl[k].hs, pr, fast, &dbuf, &dcap, &tbuf, &tcap);
      cls_out[p] = (uint8_t)c;
      if (c == 2) tl[k].nground++;
    }
  }
  for (long i = 0; i < n; i++) {
    if (cls_out[i] == 1) st->n_class1++;
    else if (cls_out[i] == 2) st->n_class2++;
    else st->n_unclassified++;
  }
  if (tiles_out) memcpy(tiles_out, tl, sizeof(wbo_tile) * ntiles);
  free(pbuf); free(dbuf); free(tbuf); free(ex); free(ey); free(tl); free(g.idx);
  wbo_free(t);
  if (!x_out) free(x);
  if (!y_out) free(y);
  if (!z_out) free(z);
  return ntiles;
}

/* Classification of a subset of points (by input index) given per-point
 * hyperboloid sizes -- used to time the classify query on a bounded sample
 * and to check single points. */
void wbo_classify_points(const wbo_tree *t, const long *pts, long m, const double *hs,
                         const wbo_params *pr, uint8_t *out)
{
  long dcap = 0, tcap = 0;
  int *dbuf = NULL, *tbuf = NULL;
  for (long i = 0; i < m; i++)
    out[i] = (uint8_t)classify_point(t, pts[i], hs[i], pr, 1, &dbuf, &dcap, &tbuf, &tcap);
  free(dbuf); free(tbuf);
}

long wbo_sizeof_tile(void) { return sizeof(wbo_tile); }
long wbo_sizeof_stats(void) { return sizeof(wbo_stats); }

/*
 * wb_synth.c -- seeded synthetic LiDAR tiles (LAS point format 1 fields,
 * int32 coordinates at 0.001 m) for the benchmark and the tests.
 *
 * Scene: ground = sum of 6 sinusoids (amplitude 0.5-5 m, wavelength
 * 20-150 m) + N(0, 0.03 m); optional embankments (ramps of 20-45 deg) and
 * isolated outliers 1-50 m off the surface (a given share of them below it,
 * the rest above); optionally an urban part (buildings only, x below a
 * split) beside a forested part (trees only); box buildings (10-30 m sides, 4-15 m tall,
 * points on roofs and walls); cone trees (3-20 m, 1-3 returns per pulse,
 * last return often on the ground); optional thin linear features (0.2 m
 * wide curbs and wires).  Feature counts are given per 200 m x 200 m and
 * scaled with the area.  Pulses are uniform in xy.  Every random draw comes
 * from one xoshiro256** stream seeded through SplitMix64, so a seed fixes the
 * cloud on every machine with the same libm.  Duplicate int32 triples are
 * dropped (the octree stores distinct locations only).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  uint64_t seed;
  double extent;          /* square side, m */
  double density;         /* pulses per m^2 */
  double buildings_per_ref, trees_per_ref;   /* per 200 m x 200 m */
  int n_embankments;
  double outlier_frac;
  double linear_frac;
  double origin_x, origin_y;
  double outlier_low;     /* share of the outliers placed below the surface */
  double split_x;         /* > 0: buildings only at x < split_x, trees only at x >= split_x */
} wb_synth_config;

typedef struct {
  int32_t *x, *y, *z;
  uint16_t *intensity;
  uint8_t *returns;       /* return number << 4 | number of returns */
  uint8_t *truth;         /* generator label: 2 ground, 6 building, 5 tree, 7 outlier, 13 wire, 11 curb */
  long n;
  double scale[3], offset[3], bbmin[3], bbmax[3];
} wb_synth_cloud;

/* ---- RNG ---- */
static uint64_t sm64(uint64_t *s)
{
  uint64_t z = (*s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
typedef struct { uint64_t s[4]; double gauss_spare; int has_spare; } rng_t;
static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
static uint64_t nextu(rng_t *r)
{
  uint64_t *s = r->s, res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t; s[3] = rotl(s[3], 45);
  return res;
}
static void rng_init(rng_t *r, uint64_t seed)
{
  uint64_t s = seed;
  for (int i = 0; i < 4; i++) r->s[i] = sm64(&s);
  r->has_spare = 0;
}
static double unif(rng_t *r) { return (nextu(r) >> 11) * 0x1p-53; }
static double unifab(rng_t *r, double a, double b) { return a + (b - a) * unif(r); }
static double gauss(rng_t *r)
{
  if (r->has_spare) { r->has_spare = 0; return r->gauss_spare; }
  double u, v, s;
  do { u = 2 * unif(r) - 1; v = 2 * unif(r) - 1; s = u * u + v * v; } while (s >= 1 || s == 0);
  double f = sqrt(-2 * log(s) / s);
  r->gauss_spare = v * f; r->has_spare = 1;
  return u * f;
}

/* ---- scene ---- */
typedef struct { double A, kx, ky, ph; } wave_t;
typedef struct { double px, py, nx, ny, tana, h; } bank_t;
typedef struct { double x0, y0, x1, y1, roof, base; } bldg_t;
typedef struct { double cx, cy, h, r, base; } tree_t;
typedef struct { double px, py, dx, dy, len, zoff; int wire; } line_t;

typedef struct {
  wave_t w[6];
  bank_t *bank; int nbank;
  bldg_t *b; int nb;
  tree_t *t; int nt;
  line_t *l; int nl;
  double extent, cell;
  int gw;
  int *bstart, *bidx, *tstart, *tidx;   /* grid cell -> lists */
} scene_t;

static double ground(const scene_t *sc, double x, double y)
{
  double z = 0;
  for (int i = 0; i < 6; i++)
    z += sc->w[i].A * sin(sc->w[i].kx * x + sc->w[i].ky * y + sc->w[i].ph);
  for (int i = 0; i < sc->nbank; i++) {
    const bank_t *b = &sc->bank[i];
    double d = (x - b->px) * b->nx + (y - b->py) * b->ny;
    double r = d * b->tana;
    if (r < 0) r = 0;
    if (r > b->h) r = b->h;
    z += r;
  }
  return z;
}

static void grid_lists(int gw, double cell, int n, const double *bx0, const double *by0,
                       const double *bx1, const double *by1, int **start, int **idx)
{
  long ncell = (long)gw * gw;
  int *cnt = calloc(ncell + 1, sizeof(int));
  for (int i = 0; i < n; i++) {
    int a0 = (int)floor(bx0[i] / cell), a1 = (int)floor(bx1[i] / cell);
    int b0 = (int)floor(by0[i] / cell), b1 = (int)floor(by1[i] / cell);
    for (int gy = b0; gy <= b1; gy++)
      for (int gx = a0; gx <= a1; gx++)
        if (gx >= 0 && gy >= 0 && gx < gw && gy < gw) cnt[(long)gy * gw + gx + 1]++;
  }
  for (long c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
  int *fill = malloc(sizeof(int) * (ncell + 1));
  memcpy(fill, cnt, sizeof(int) * (ncell + 1));
  *idx = malloc(sizeof(int) * (cnt[ncell] + 1));
  for (int i = 0; i < n; i++) {
    int a0 = (int)floor(bx0[i] / cell), a1 = (int)floor(bx1[i] / cell);
    int b0 = (int)floor(by0[i] / cell), b1 = (int)floor(by1[i] / cell);
    for (int gy = b0; gy <= b1; gy++)
      for (int gx = a0; gx <= a1; gx++)
        if (gx >= 0 && gy >= 0 && gx < gw && gy < gw) (*idx)[fill[(long)gy * gw + gx]++] = i;
  }
  free(fill);
  *start = cnt;
}

static void build_scene(scene_t *sc, const wb_synth_config *cfg, rng_t *r)
{
  double W = cfg->extent, area_scale = W * W / (200.0 * 200.0);
  memset(sc, 0, sizeof *sc);
  sc->extent = W;
  for (int i = 0; i < 6; i++) {
    double A = unifab(r, 0.5, 5), lam = unifab(r, 20, 150), th = unifab(r, 0, 2 * M_PI);
    double k = 2 * M_PI / lam;
    sc->w[i].A = A; sc->w[i].kx = k * cos(th); sc->w[i].ky = k * sin(th);
    sc->w[i].ph = unifab(r, 0, 2 * M_PI);
  }
  sc->nbank = cfg->n_embankments;
  sc->bank = calloc(sc->nbank + 1, sizeof(bank_t));
  for (int i = 0; i < sc->nbank; i++) {
    double th = unifab(r, 0, 2 * M_PI), al = unifab(r, 20, 45) * M_PI / 180;
    sc->bank[i].px = unifab(r, 0.1 * W, 0.9 * W);
    sc->bank[i].py = unifab(r, 0.1 * W, 0.9 * W);
    sc->bank[i].nx = -sin(th); sc->bank[i].ny = cos(th);
    sc->bank[i].tana = tan(al);
    sc->bank[i].h = unifab(r, 2, 8);
  }
  sc->nb = (int)lrint(cfg->buildings_per_ref * area_scale);
  sc->b = calloc(sc->nb + 1, sizeof(bldg_t));
  double *bx0 = malloc(sizeof(double) * (sc->nb + 1)), *by0 = malloc(sizeof(double) * (sc->nb + 1));
  double *bx1 = malloc(sizeof(double) * (sc->nb + 1)), *by1 = malloc(sizeof(double) * (sc->nb + 1));
  for (int i = 0; i < sc->nb; i++) {
    double sx = unifab(r, 10, 30), sy = unifab(r, 10, 30), h = unifab(r, 4, 15);
    double xhi = cfg->split_x > 0 ? fmin(0.95 * W, cfg->split_x) : 0.95 * W;
    double cx = unifab(r, 0.05 * W + sx / 2, xhi - sx / 2);
    double cy = unifab(r, 0.05 * W + sy / 2, 0.95 * W - sy / 2);
    bldg_t *b = &sc->b[i];
    b->x0 = cx - sx / 2; b->x1 = cx + sx / 2; b->y0 = cy - sy / 2; b->y1 = cy + sy / 2;
    b->base = ground(sc, cx, cy);
    b->roof = b->base + h;
    bx0[i] = b->x0; bx1[i] = b->x1; by0[i] = b->y0; by1[i] = b->y1;
  }
  sc->nt = (int)lrint(cfg->trees_per_ref * area_scale);
  sc->t = calloc(sc->nt + 1, sizeof(tree_t));
  double *tx0 = malloc(sizeof(double) * (sc->nt + 1)), *ty0 = malloc(sizeof(double) * (sc->nt + 1));
  double *tx1 = malloc(sizeof(double) * (sc->nt + 1)), *ty1 = malloc(sizeof(double) * (sc->nt + 1));
  for (int i = 0; i < sc->nt; i++) {
    tree_t *t = &sc->t[i];
    t->h = unifab(r, 3, 20);
    t->r = t->h * unifab(r, 0.2, 0.35);
    t->cx = cfg->split_x > 0 ? unifab(r, cfg->split_x, W) : unifab(r, 0, W);
    t->cy = unifab(r, 0, W);
    t->base = ground(sc, t->cx, t->cy);
    tx0[i] = t->cx - t->r; tx1[i] = t->cx + t->r; ty0[i] = t->cy - t->r; ty1[i] = t->cy + t->r;
  }
  sc->nl = cfg->linear_frac > 0 ? (int)lrint(20 * area_scale) + 1 : 0;
  sc->l = calloc(sc->nl + 1, sizeof(line_t));
  for (int i = 0; i < sc->nl; i++) {
    double th = unifab(r, 0, 2 * M_PI);
    line_t *l = &sc->l[i];
    l->px = unifab(r, 0.1 * W, 0.9 * W); l->py = unifab(r, 0.1 * W, 0.9 * W);
    l->dx = cos(th); l->dy = sin(th); l->len = unifab(r, 20, 0.5 * W);
    l->wire = (i & 1);
    l->zoff = l->wire ? unifab(r, 8, 12) : 0.15;
  }
  sc->cell = 10.0;
  sc->gw = (int)ceil(W / sc->cell) + 1;
  grid_lists(sc->gw, sc->cell, sc->nb, bx0, by0, bx1, by1, &sc->bstart, &sc->bidx);
  grid_lists(sc->gw, sc->cell, sc->nt, tx0, ty0, tx1, ty1, &sc->tstart, &sc->tidx);
  free(bx0); free(by0); free(bx1); free(by1); free(tx0); free(ty0); free(tx1); free(ty1);
}

static void free_scene(scene_t *sc)
{
  free(sc->bank); free(sc->b); free(sc->t); free(sc->l);
  free(sc->bstart); free(sc->bidx); free(sc->tstart); free(sc->tidx);
}

/* ---- output buffer ---- */
typedef struct {
  double *x, *y, *z;
  uint16_t *in;
  uint8_t *ret, *truth;
  long n, cap;
} pbuf_t;

static void push(pbuf_t *p, double x, double y, double z, uint16_t in, uint8_t ret, uint8_t tr)
{
  if (p->n == p->cap) {
    p->cap = p->cap ? p->cap * 2 : 1 << 20;
    p->x = realloc(p->x, sizeof(double) * p->cap);
    p->y = realloc(p->y, sizeof(double) * p->cap);
    p->z = realloc(p->z, sizeof(double) * p->cap);
    p->in = realloc(p->in, sizeof(uint16_t) * p->cap);
    p->ret = realloc(p->ret, p->cap);
    p->truth = realloc(p->truth, p->cap);
  }
  p->x[p->n] = x; p->y[p->n] = y; p->z[p->n] = z;
  p->in[p->n] = in; p->ret[p->n] = ret; p->truth[p->n] = tr;
  p->n++;
}

static uint64_t hash3(int32_t a, int32_t b, int32_t c)
{
  uint64_t h = (uint64_t)(uint32_t)a * 0x9e3779b97f4a7c15ull;
  h ^= (uint64_t)(uint32_t)b * 0xc2b2ae3d27d4eb4full + (h << 6) + (h >> 2);
  h ^= (uint64_t)(uint32_t)c * 0x165667b19e3779f9ull + (h << 6) + (h >> 2);
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  return h;
}

int wb_synth_generate(const wb_synth_config *cfg, wb_synth_cloud *out)
{
  rng_t r;
  scene_t sc;
  pbuf_t pb;
  double W = cfg->extent;
  long npulse = lrint(cfg->density * W * W);
  memset(out, 0, sizeof *out);
  memset(&pb, 0, sizeof pb);
  rng_init(&r, cfg->seed);
  build_scene(&sc, cfg, &r);
  for (long k = 0; k < npulse; k++) {
    double x = unif(&r) * W, y = unif(&r) * W;
    uint16_t inten = (uint16_t)(nextu(&r) >> 48);
    double g = ground(&sc, x, y);
    int gx = (int)(x / sc.cell), gy = (int)(y / sc.cell);
    long c = (long)gy * sc.gw + gx;
    int hit_b = -1, hit_t = -1;
    double best_tree_top = -INFINITY, tree_d = 0;
    if (cfg->linear_frac > 0 && unif(&r) < cfg->linear_frac) {
      const line_t *l = &sc.l[nextu(&r) % sc.nl];
      double s = unif(&r) * l->len, off = unifab(&r, -0.1, 0.1);
      double lx = l->px + l->dx * s - l->dy * off, ly = l->py + l->dy * s + l->dx * off;
      if (lx >= 0 && ly >= 0 && lx < W && ly < W) {
        double lz = ground(&sc, lx, ly) + l->zoff + gauss(&r) * 0.01;
        push(&pb, lx, ly, lz, inten, 0x11, l->wire ? 13 : 11);
        continue;
      }
    }
    for (int j = sc.bstart[c]; j < sc.bstart[c + 1]; j++) {
      const bldg_t *b = &sc.b[sc.bidx[j]];
      if (x >= b->x0 && x < b->x1 && y >= b->y0 && y < b->y1) { hit_b = sc.bidx[j]; break; }
    }
    if (hit_b >= 0) {
      const bldg_t *b = &sc.b[hit_b];
      double edge = fmin(fmin(x - b->x0, b->x1 - x), fmin(y - b->y0, b->y1 - y));
      if (edge < 0.3) {
        double zz = g + unif(&r) * (b->roof - g);
        push(&pb, x, y, zz, inten, 0x11, 6);
      } else {
        push(&pb, x, y, b->roof + gauss(&r) * 0.02, inten, 0x11, 6);
      }
      continue;
    }
    for (int j = sc.tstart[c]; j < sc.tstart[c + 1]; j++) {
      const tree_t *t = &sc.t[sc.tidx[j]];
      double d = hypot(x - t->cx, y - t->cy);
      if (d < t->r) {
        double top = t->base + t->h * (1 - 0.7 * d / t->r);
        if (top > best_tree_top) { best_tree_top = top; hit_t = sc.tidx[j]; tree_d = d; }
      }
    }
    if (hit_t >= 0) {
      int nret = 1 + (int)(nextu(&r) % 3);
      double ztop = best_tree_top - unif(&r) * 0.5;
      double zlow = g + 0.5;
      double zprev = ztop + 1;
      (void)tree_d;
      for (int k2 = 1; k2 <= nret; k2++) {
        double zz;
        uint8_t tr = 5;
        if (k2 == 1) zz = ztop;
        else if (k2 == nret && unif(&r) < 0.7) { zz = g + gauss(&r) * 0.03; tr = 2; }
        else zz = zlow + unif(&r) * (zprev - 0.002 - zlow);
        if (zz >= zprev - 0.001) zz = zprev - 0.002;
        zprev = zz;
        push(&pb, x, y, zz, inten, (uint8_t)((k2 << 4) | nret), tr);
      }
      continue;
    }
    if (cfg->outlier_frac > 0 && unif(&r) < cfg->outlier_frac) {
      double dz = unifab(&r, 1, 50);
      push(&pb, x, y, (unif(&r) < cfg->outlier_low) ? g - dz : g + dz, inten, 0x11, 7);
      continue;
    }
    push(&pb, x, y, g + gauss(&r) * 0.03, inten, 0x11, 2);
  }
  free_scene(&sc);
  /* quantize, drop duplicate int triples (first occurrence kept) */
  {
    long n = pb.n, m = 0;
    unsigned long cap = 1;
    while (cap < (unsigned long)(2 * n + 16)) cap <<= 1;
    int32_t *hx = malloc(sizeof(int32_t) * cap * 3);
    uint8_t *used = calloc(cap, 1);
    out->x = malloc(sizeof(int32_t) * (n + 1));
    out->y = malloc(sizeof(int32_t) * (n + 1));
    out->z = malloc(sizeof(int32_t) * (n + 1));
    out->intensity = malloc(sizeof(uint16_t) * (n + 1));
    out->returns = malloc(n + 1);
    out->truth = malloc(n + 1);
    out->scale[0] = out->scale[1] = out->scale[2] = 0.001;
    out->offset[0] = cfg->origin_x;
    out->offset[1] = cfg->origin_y;
    out->offset[2] = 0;
    for (int a = 0; a < 3; a++) { out->bbmin[a] = INFINITY; out->bbmax[a] = -INFINITY; }
    for (long i = 0; i < n; i++) {
      int32_t xi = (int32_t)llrint(pb.x[i] / 0.001), yi = (int32_t)llrint(pb.y[i] / 0.001),
              zi = (int32_t)llrint(pb.z[i] / 0.001);
      unsigned long h = hash3(xi, yi, zi) & (cap - 1);
      int dup = 0;
      while (used[h]) {
        if (hx[3 * h] == xi && hx[3 * h + 1] == yi && hx[3 * h + 2] == zi) { dup = 1; break; }
        h = (h + 1) & (cap - 1);
      }
      if (dup) continue;
      used[h] = 1;
      hx[3 * h] = xi; hx[3 * h + 1] = yi; hx[3 * h + 2] = zi;
      out->x[m] = xi; out->y[m] = yi; out->z[m] = zi;
      out->intensity[m] = pb.in[i]; out->returns[m] = pb.ret[i]; out->truth[m] = pb.truth[i];
      {
        double dx = out->offset[0] + out->scale[0] * xi, dy = out->offset[1] + out->scale[1] * yi,
               dz = out->offset[2] + out->scale[2] * zi;
        if (dx < out->bbmin[0]) out->bbmin[0] = dx;
        if (dx > out->bbmax[0]) out->bbmax[0] = dx;
        if (dy < out->bbmin[1]) out->bbmin[1] = dy;
        if (dy > out->bbmax[1]) out->bbmax[1] = dy;
        if (dz < out->bbmin[2]) out->bbmin[2] = dz;
        if (dz > out->bbmax[2]) out->bbmax[2] = dz;
      }
      m++;
    }
    out->n = m;
    free(hx); free(used);
  }
  free(pb.x); free(pb.y); free(pb.z); free(pb.in); free(pb.ret); free(pb.truth);
  return 0;
}

void wb_synth_free(wb_synth_cloud *c)
{
  free(c->x); free(c->y); free(c->z); free(c->intensity); free(c->returns); free(c->truth);
  memset(c, 0, sizeof *c);
}

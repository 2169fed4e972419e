/*
 * wb_oracle_las.c -- CPU restatement of the reference's LAS reader and
 * writer, its class-separated output, the duplicate merge of the octree
 * store and the textual octree dump.
 *
 * THIS IS TEST INFRASTRUCTURE (the checker), NOT PART OF THE PRODUCT.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The HIP library never links or calls it.
 *
 * Each function follows the reference step by step (file:line cited at each
 * function).  Pinning (tests/ check each against):
 *   wbo_las_parse        ... header byte layouts written from the published
 *                            LAS 1.2 / 1.4 specifications by the tests
 *                            themselves, and the count-repair cases of the
 *                            comment at las.cpp:354-359.
 *   wbo_las_read_point   ... records packed field by field from the LAS
 *                            specifications' record layouts in the tests;
 *   wbo_degtobin/bintodeg... oracle/_ref (reference angle.cpp) bit for bit.
 *   wbo_las_write_point  ... write -> read round trips for all formats, and
 *                            records packed from the specification.
 *   wbo_manygcd          ... oracle/_ref (reference manygcd.cpp) bit for bit,
 *                            and gcd known answers.
 *   wbo_join_point_format... the specifications' feature table (format 3 =
 *                            1 + 2, 7 = 6 + RGB, 10 = 8 + waves ...).
 *   wbo_dedupe           ... brute force on small inputs.
 *   wbo_ldecimal         ... oracle/_ref (reference ldecimal.cpp) on random
 *                            and special doubles.
 *   wbo_output_plan      ... brute-force restatement of writeFiles' loop in
 *                            the tests and invariants (each class's points
 *                            split over its files, counts balanced per block).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

/* las.cpp:34-39 */
#define MASK_GPSTIME 0x7fa
#define MASK_RGB 0x5ac
#define MASK_NIR 0x500
#define MASK_WAVE 0x630
static const short point_lengths[] = {20, 28, 26, 34, 57, 63, 30, 36, 38, 59, 67};
static const short point_features[] = {0x0, 0x1, 0x2, 0x3, 0x9, 0xb, 0x101, 0x103, 0x107, 0x109, 0x10f};

typedef struct {
  uint16_t source_id, global_encoding;
  uint8_t guid[16];
  uint8_t version_major, version_minor;
  char system_id[33], software[33];
  uint16_t creation_day, creation_year;
  uint32_t header_size, point_offset, n_vlr;
  uint16_t point_format, point_length;
  double scale[3], offset[3], min[3], max[3];
  uint64_t start_waveform, start_evlr;
  uint32_t n_evlr;
  uint64_t n_points[16];
  int32_t zipped, valid;
} wbo_las_header;

typedef struct {
  double x, y, z;
  uint16_t intensity, ret, nret;
  uint8_t scan_dir, edge;
  uint16_t cls, cls_flags, channel, user, source;
  int32_t scan_angle;
  double gps;
  uint16_t nir, red, green, blue;
} wbo_las_point;

int wbo_las_point_length(int format)
{
  return format >= 0 && format <= 10 ? point_lengths[format] : 0;
}

/* little-endian reads with zeros past the end of the buffer */
static uint64_t rd(const uint8_t *b, size_t len, size_t off, int nbytes)
{
  uint64_t v = 0;
  for (int k = 0; k < nbytes; k++)
    if (off + k < len) v |= (uint64_t)b[off + k] << (8 * k);
  return v;
}
static double rdd(const uint8_t *b, size_t len, size_t off)
{
  uint64_t v = rd(b, len, off, 8);
  double d;
  memcpy(&d, &v, 8);
  return d;
}

/* LasHeader::openRead (las.cpp:272-428) on a file image; returns 1 if it
 * starts "LASF" */
int wbo_las_parse(const uint8_t *b, size_t len, wbo_las_header *h)
{
  uint32_t legacy[6];
  int which = 15;
  uint64_t total;
  memset(h, 0, sizeof *h);
  if (len < 4 || !(b[0] == 'L' && b[1] == 'A' && b[2] == 'S' && b[3] == 'F')) return 0;
  h->source_id = (uint16_t)rd(b, len, 4, 2);
  h->global_encoding = (uint16_t)rd(b, len, 6, 2);
  for (int k = 0; k < 16; k++) h->guid[k] = (uint8_t)rd(b, len, 8 + k, 1);
  h->version_major = (uint8_t)rd(b, len, 24, 1);
  h->version_minor = (uint8_t)rd(b, len, 25, 1);
  for (int k = 0; k < 32; k++) h->system_id[k] = (char)rd(b, len, 26 + k, 1);
  for (int k = 0; k < 32; k++) h->software[k] = (char)rd(b, len, 58 + k, 1);
  h->system_id[32] = h->software[32] = 0;
  h->creation_day = (uint16_t)rd(b, len, 90, 2);
  h->creation_year = (uint16_t)rd(b, len, 92, 2);
  /* readleshort returns a short, stored in an unsigned int */
  h->header_size = (uint32_t)(int32_t)(int16_t)rd(b, len, 94, 2);
  h->point_offset = (uint32_t)rd(b, len, 96, 4);
  h->n_vlr = (uint32_t)rd(b, len, 100, 4);
  h->point_format = (uint16_t)rd(b, len, 104, 1);
  h->point_length = (uint16_t)rd(b, len, 105, 2);
  for (int i = 0; i < 6; i++) legacy[i] = (uint32_t)rd(b, len, 107 + 4 * i, 4);
  for (int a = 0; a < 3; a++) h->scale[a] = rdd(b, len, 131 + 8 * a);
  for (int a = 0; a < 3; a++) h->offset[a] = rdd(b, len, 155 + 8 * a);
  for (int a = 0; a < 3; a++) {
    h->max[a] = rdd(b, len, 179 + 16 * a);
    h->min[a] = rdd(b, len, 187 + 16 * a);
  }
  if (h->header_size > 0xe3) {
    h->start_waveform = rd(b, len, 227, 8);
    h->start_evlr = rd(b, len, 235, 8);
    h->n_evlr = (uint32_t)rd(b, len, 243, 4);
    for (int i = 0; i < 16; i++) h->n_points[i] = rd(b, len, 247 + 8 * i, 8);
  }
  /* the count repair of las.cpp:354-402 */
  total = 0;
  for (int i = 1; i < 6; i++) {
    total += legacy[i];
    if (legacy[i] > legacy[0]) which &= ~5;
  }
  if (total != legacy[0]) which &= ~1;
  if (total != 0 || legacy[0] == 0) which &= ~4;
  total = 0;
  for (int i = 1; i < 16; i++) {
    total += h->n_points[i];
    if (h->n_points[i] > h->n_points[0]) which &= ~10;
  }
  if (total != h->n_points[0]) which &= ~2;
  if (total != 0 || h->n_points[0] == 0) which &= ~8;
  for (int i = 0; i < 6; i++) {
    if (legacy[i] != h->n_points[i] && legacy[i] != 0) which &= ~10;
    if (h->n_points[i] != legacy[i] && h->n_points[i] != 0) which &= ~5;
  }
  if (which > 0 && (which & 3) == 0) {
    h->n_points[1] = h->n_points[0];
    legacy[1] = legacy[0];
  }
  if (which == 1 || which == 4)
    for (int i = 0; i < 6; i++) h->n_points[i] = legacy[i];
  if (which == 0)
    for (int i = 0; i < 6; i++) h->n_points[i] = 0;
  if (h->point_length == 0) h->version_major = h->version_minor = 0, h->n_points[0] = 0;
  /* LASzip: an extra header starting "\0\0laszip" (las.cpp:408-422) */
  if (h->point_offset >= (uint64_t)h->header_size + 8) {
    uint64_t m = 0;
    for (int k = 0; k < 8; k++) m = (m << 8) | rd(b, len, h->header_size + k, 1);
    if ((m & 0xffffffffffffull) == 0x6c61737a6970ull) h->zipped = 1;
  }
  /* LasHeader::isValid (las.cpp:597-601) */
  h->valid = h->version_major > 0 && h->version_minor > 0 && h->header_size > 0 &&
             h->point_length > 0 && (h->header_size > 0xe3 || h->point_format < 6);
  return 1;
}

/* angle.cpp:209-247 */
double wbo_bintodeg(int a) { return (a / 2147483648.) * 360; }
int wbo_rottobin(double angle)
{
  double iprt = 0, fprt;
  fprt = 2 * modf(angle / 2, &iprt);
  if (fprt >= 1) fprt -= 2;
  if (fprt < -1) fprt += 2;
  return (int)lrint(2147483648. * fprt);
}
int wbo_degtobin(double angle) { return wbo_rottobin(angle / 360); }

/* LasHeader::readPoint (las.cpp:808-904) of record rec under header h */
void wbo_las_read_point(const uint8_t *rec, const wbo_las_header *h, double unit,
                        wbo_las_point *p)
{
  const size_t L = 1 << 16;
  int f = h->point_format, o;
  int32_t xi = (int32_t)rd(rec, L, 0, 4), yi = (int32_t)rd(rec, L, 4, 4), zi = (int32_t)rd(rec, L, 8, 4);
  int temp;
  memset(p, 0, sizeof *p);
  p->intensity = (uint16_t)rd(rec, L, 12, 2);
  if (f < 6) {
    temp = rec[14];
    p->ret = temp & 7;
    p->nret = (temp >> 3) & 7;
    p->scan_dir = (temp >> 6) & 1;
    p->edge = (temp >> 7) & 1;
    p->cls = rec[15];
    p->cls_flags = (p->cls >> 5) & 7;
    p->cls &= 31;
    p->scan_angle = wbo_degtobin((signed char)rec[16]);
    p->user = rec[17];
    p->source = (uint16_t)rd(rec, L, 18, 2);
    o = 20;
  } else {
    temp = rec[14];
    p->ret = temp & 15;
    p->nret = (temp >> 4) & 15;
    temp = rec[15];
    p->cls_flags = temp & 15;
    p->channel = (temp >> 4) & 3;
    p->scan_dir = (temp >> 6) & 1;
    p->edge = (temp >> 7) & 1;
    p->cls = rec[16];
    p->user = rec[17];
    p->scan_angle = wbo_degtobin((int16_t)rd(rec, L, 18, 2) * 0.006);
    p->source = (uint16_t)rd(rec, L, 20, 2);
    o = 22;
  }
  if ((1 << f) & MASK_GPSTIME) {
    p->gps = rdd(rec, L, o);
    o += 8;
  }
  if ((1 << f) & MASK_RGB) {
    p->red = (uint16_t)rd(rec, L, o, 2);
    p->green = (uint16_t)rd(rec, L, o + 2, 2);
    p->blue = (uint16_t)rd(rec, L, o + 4, 2);
    o += 6;
  }
  if ((1 << f) & MASK_NIR) p->nir = (uint16_t)rd(rec, L, o, 2);
  p->x = (h->offset[0] + h->scale[0] * xi) * unit;
  p->y = (h->offset[1] + h->scale[1] * yi) * unit;
  p->z = (h->offset[2] + h->scale[2] * zi) * unit;
}

static void wr(uint8_t *b, size_t off, uint64_t v, int nbytes)
{
  for (int k = 0; k < nbytes; k++) b[off + k] = (uint8_t)(v >> (8 * k));
}
static void wrd(uint8_t *b, size_t off, double d)
{
  uint64_t v;
  memcpy(&v, &d, 8);
  wr(b, off, v, 8);
}

/* the output side of one LAS file: LasHeader's scale, offset, unit, point
 * format and the running statistics writePoint keeps */
typedef struct {
  int format;
  double scale[3], offset[3], unit;
  double min[3], max[3];
  uint64_t n_points[16];
} wbo_las_out;

void wbo_las_out_init(wbo_las_out *o, int format, const double scale[3], const double offset[3],
                      double unit)
{
  memset(o, 0, sizeof *o);
  o->format = format;
  for (int a = 0; a < 3; a++) {
    o->scale[a] = scale[a];
    o->offset[a] = offset[a];
    o->min[a] = INFINITY;
    o->max[a] = -INFINITY;
  }
  o->unit = unit;
}

/* LasHeader::writePoint (las.cpp:906-990) into rec (point_lengths[format]
 * bytes); the waveform fields are written as zeros (built without
 * WAVEFORM) */
void wbo_las_write_point(uint8_t *rec, wbo_las_out *o, const wbo_las_point *p)
{
  const int f = o->format;
  int xi = (int)lrint((p->x / o->unit - o->offset[0]) / o->scale[0]);
  int yi = (int)lrint((p->y / o->unit - o->offset[1]) / o->scale[1]);
  int zi = (int)lrint((p->z / o->unit - o->offset[2]) / o->scale[2]);
  double w[3];
  int q;
  memset(rec, 0, point_lengths[f]);
  wr(rec, 0, (uint32_t)xi, 4);
  wr(rec, 4, (uint32_t)yi, 4);
  wr(rec, 8, (uint32_t)zi, 4);
  w[0] = xi * o->scale[0] + o->offset[0];
  w[1] = yi * o->scale[1] + o->offset[1];
  w[2] = zi * o->scale[2] + o->offset[2];
  for (int a = 0; a < 3; a++) {
    if (w[a] > o->max[a]) o->max[a] = w[a];
    if (w[a] < o->min[a]) o->min[a] = w[a];
  }
  wr(rec, 12, p->intensity, 2);
  if (f < 6) {
    rec[14] = (uint8_t)((p->ret & 7) + ((p->nret & 7) << 3) + ((p->scan_dir & 1) << 6) +
                        ((p->edge & 1) << 7));
    rec[15] = (uint8_t)((p->cls & 31) + ((p->cls_flags & 7) << 5));
    rec[16] = (uint8_t)lrint(wbo_bintodeg(p->scan_angle));
    rec[17] = (uint8_t)p->user;
    wr(rec, 18, p->source, 2);
    q = 20;
  } else {
    rec[14] = (uint8_t)((p->ret & 15) + ((p->nret & 15) << 4));
    rec[15] = (uint8_t)((p->cls_flags & 15) + ((p->channel & 3) << 4) + ((p->scan_dir & 1) << 6) +
                        ((p->edge & 1) << 7));
    rec[16] = (uint8_t)p->cls;
    rec[17] = (uint8_t)p->user;
    wr(rec, 18, (uint16_t)(short)lrint(wbo_bintodeg(p->scan_angle) / 0.006), 2);
    wr(rec, 20, p->source, 2);
    q = 22;
  }
  if ((1 << f) & MASK_GPSTIME) {
    wrd(rec, q, p->gps);
    q += 8;
  }
  if ((1 << f) & MASK_RGB) {
    wr(rec, q, p->red, 2);
    wr(rec, q + 2, p->green, 2);
    wr(rec, q + 4, p->blue, 2);
    q += 6;
  }
  if ((1 << f) & MASK_NIR) {
    wr(rec, q, p->nir, 2);
    q += 2;
  }
  o->n_points[0]++;
  if (p->ret > 0 && p->ret < 16) o->n_points[p->ret]++;
}

/* LasHeader::writeHeader (las.cpp:549-595) of a version 1.4 file opened by
 * openWrite (las.cpp:430-492) with setVersion(1,4) and setPointFormat;
 * out: 375 bytes.  The fields openWrite leaves unset (source id, global
 * encoding, GUID, start of waveform data) are written as zeros. */
void wbo_las_write_header(uint8_t *out, const wbo_las_out *o, const char *system_id,
                          const char *software, int day, int year)
{
  int legacy_valid = 1, i;
  const uint32_t hs = 0x177;
  memset(out, 0, 375);
  memcpy(out, "LASF", 4);
  for (i = 0; i < 6; i++)
    if (o->n_points[i] > 4294967295ull) legacy_valid = 0;
  for (; i < 16; i++)
    if (o->n_points[i] > 0) legacy_valid = 0;
  out[24] = 1;
  out[25] = 4;
  strncpy((char *)out + 26, system_id, 32);
  strncpy((char *)out + 58, software, 32);
  wr(out, 90, (uint16_t)day, 2);
  wr(out, 92, (uint16_t)year, 2);
  wr(out, 94, hs, 2);
  wr(out, 96, hs, 4);
  wr(out, 100, 0, 4);
  out[104] = (uint8_t)o->format;
  wr(out, 105, (uint16_t)point_lengths[o->format], 2);
  for (i = 0; i < 6; i++) wr(out, 107 + 4 * i, legacy_valid ? (uint32_t)o->n_points[i] : 0, 4);
  for (int a = 0; a < 3; a++) wrd(out, 131 + 8 * a, o->scale[a]);
  for (int a = 0; a < 3; a++) wrd(out, 155 + 8 * a, o->offset[a]);
  for (int a = 0; a < 3; a++) {
    wrd(out, 179 + 16 * a, o->max[a]);
    wrd(out, 187 + 16 * a, o->min[a]);
  }
  wr(out, 227, 0, 8);
  wr(out, 235, hs + o->n_points[0] * (uint64_t)point_lengths[o->format], 8);
  wr(out, 243, 0, 4);
  for (i = 0; i < 16; i++) wr(out, 247 + 8 * i, o->n_points[i], 8);
}

/* joinPointFormat (las.cpp:104-121) */
int wbo_join_point_format(const int *formats, int n)
{
  int all = 0, ret = 0;
  for (int i = 0; i < n; i++)
    if (formats[i] >= 0 && formats[i] < 11) all |= point_features[formats[i]];
  for (int i = 0; i < 11; i++)
    if ((all | point_features[i]) == point_features[i]) {
      ret = i;
      break;
    }
  return ret;
}

/* pairwisesum (manysum.cpp:120-154) is restated in wb_oracle.c */
double wbo_pairwisesum(const double *a, unsigned n);

static void bubbleup(double *v, int n)
{
  for (int i = 0; i < n - 1; i++)
    if (v[i] > v[i + 1]) {
      double t = v[i]; v[i] = v[i + 1]; v[i + 1] = t;
    }
}
static void bubbledown(double *v, int n)
{
  for (int i = n - 2; i >= 0; i--)
    if (v[i] > v[i + 1]) {
      double t = v[i]; v[i] = v[i + 1]; v[i + 1] = t;
    }
}

/* manygcd (manygcd.cpp:44-73) */
double wbo_manygcd(const double *numbers, int n, double toler)
{
  double *v = malloc(sizeof(double) * (n ? n : 1)), r;
  int sz = n;
  for (int i = 0; i < n; i++) v[i] = fabs(numbers[i]);
  if (n) {
    bubbleup(v, sz);
    bubbledown(v, sz);
    bubbleup(v, sz);
  }
  while (sz && v[sz - 1] > v[0] + toler) {
    if (v[sz - 1] - v[sz - 2] > toler) {
      v[sz - 1] -= v[sz - 2];
      bubbledown(v, sz);
    } else {
      sz--;
    }
    bubbleup(v, sz);
  }
  r = sz ? wbo_pairwisesum(v, sz) / sz : 0;
  free(v);
  return r;
}

/* combine1Scale (las.cpp:913-932) */
double wbo_combine1_scale(const double *scales, const double *offsets, int n)
{
  double toler = INFINITY, sg, og, ret = 0;
  for (int i = 0; i < n; i++)
    if (fabs(scales[i]) < toler) toler = fabs(scales[i]);
  toler /= 1e6;
  sg = wbo_manygcd(scales, n, toler);
  og = wbo_manygcd(offsets, n, toler);
  if ((og > 10.5 * sg || og == 0) && fabs(og / sg - rint(og / sg)) < 1e-6) ret = sg;
  for (int i = 10; i > 0; i--)
    for (int j = 10; j > 0; j--)
      if (fabs(sg * i - og * j) < toler) ret = sg / j;
  return ret;
}

/* LasHeader::setScale (las.cpp:629-665) for one axis */
void wbo_set_scale(double minc, double maxc, double scale, double unit, double *out_scale,
                   double *out_offset)
{
  double minscale = (maxc - minc) / 4132485216.;
  double off = (minc + maxc) / 2 / unit;
  if (scale > minscale && isfinite(scale)) {
    *out_scale = scale / unit;
    off = rint(off / *out_scale) * *out_scale;
  } else {
    *out_scale = minscale / unit;
  }
  *out_offset = off;
}

/* Duplicate merge of OctBuffer::put (octree.cpp:620-660) in input order:
 * first[i] = input index of the first point at i's location, last[i] = that
 * of the last one.  Returns the number of merged points. */
static const double *g_dx, *g_dy, *g_dz;
static int cmp_loc(const void *pa, const void *pb)
{
  long a = *(const long *)pa, b = *(const long *)pb;
  if (g_dx[a] < g_dx[b]) return -1;
  if (g_dx[a] > g_dx[b]) return 1;
  if (g_dy[a] < g_dy[b]) return -1;
  if (g_dy[a] > g_dy[b]) return 1;
  if (g_dz[a] < g_dz[b]) return -1;
  if (g_dz[a] > g_dz[b]) return 1;
  return a < b ? -1 : a > b;
}
long wbo_dedupe(const double *x, const double *y, const double *z, long n, long *first, long *last)
{
  long *ix = malloc(sizeof(long) * (n ? n : 1)), ndup = 0;
  for (long i = 0; i < n; i++) ix[i] = i;
  g_dx = x; g_dy = y; g_dz = z;
  qsort(ix, n, sizeof(long), cmp_loc);
  for (long a = 0; a < n;) {
    long b = a + 1;
    while (b < n && x[ix[b]] == x[ix[a]] && y[ix[b]] == y[ix[a]] && z[ix[b]] == z[ix[a]]) b++;
    for (long k = a; k < b; k++) {
      first[ix[k]] = ix[a];
      last[ix[k]] = ix[b - 1];
    }
    ndup += b - a - 1;
    a = b;
  }
  free(ix);
  return ndup;
}

/* CloudOutput::className (cloudoutput.cpp:39-117), English strings */
void wbo_class_name(int n, char *buf, int cap)
{
  static const char *names[] = {"raw", "nonground", "ground", "lowveg", "medveg", "highveg",
                                "building", "lownoise", NULL, "water", "rail", "road",
                                "overlap", "wireguard", "conductor", "tower", "insulator",
                                "bridge", "highnoise", "overhead", "ignground", "snow",
                                "temporal"};
  if (n >= 0 && n <= 22 && names[n])
    snprintf(buf, cap, "%s", names[n]);
  else
    snprintf(buf, cap, "reserved%d", n);
}

/* The file plan of CloudOutput::openFiles / writeFiles
 * (cloudoutput.cpp:119-227) with the octree's blocks taken as its leaves in
 * DFS order.  Inputs: per output point in canonical order its leaf and
 * class; class_total[256].  Outputs: per class the number of files
 * (nfiles[256]), the digits appended to names (return value), and per
 * point its file (file_of) and position in that file (pos_of).  With
 * separate == 0 every point goes to the files of class 0. */
int wbo_output_plan(const long *leaf_of, const uint8_t *cls, long n, const uint64_t *class_total,
                    long points_per_file, int separate, int *nfiles, int *file_of, long *pos_of)
{
  uint64_t grand = 0, quot;
  int ndig = 0;
  long *cnt[256];
  int next[256];
  for (int k = 0; k < 256; k++) grand += class_total[k];
  if (points_per_file) {
    quot = (grand + points_per_file - 1) / points_per_file;
    if (quot) quot--;
    if (!quot) quot++;
    while (quot) {
      quot /= 10;
      ndig++;
    }
  }
  for (int k = 0; k < 256; k++) {
    uint64_t tot = separate ? class_total[k] : (k == 0 ? grand : 0);
    int present = separate ? class_total[k] > 0 : k == 0;
    nfiles[k] = !present ? 0 : points_per_file ? (int)((tot + points_per_file - 1) / points_per_file) : 1;
    cnt[k] = calloc(nfiles[k] ? nfiles[k] : 1, sizeof(long));
    next[k] = 0;
  }
  for (long a = 0; a < n;) {
    long b = a + 1;
    while (b < n && leaf_of[b] == leaf_of[a]) b++;
    for (int k = 0; k < 256; k++) {
      long mn = (long)grand;
      for (int j = 0; j < nfiles[k]; j++)
        if (cnt[k][j] < mn) {
          next[k] = j;
          mn = cnt[k][j];
        }
    }
    for (long i = a; i < b; i++) {
      int k = separate ? cls[i] : 0;
      file_of[i] = next[k];
      pos_of[i] = cnt[k][next[k]]++;
    }
    a = b;
  }
  for (int k = 0; k < 256; k++) free(cnt[k]);
  return ndig;
}

/* ldecimal (ldecimal.cpp:32-132): the shortest %e rendering within toler,
 * reformatted without a needless exponent */
void wbo_ldecimal(double x, double toler, char *out, int cap)
{
  char buffer[64], fmt[16], *dotpos, *epos;
  double x2;
  int h, i, iexp, chexp;
  if (toler > 0 && x != 0) {
    iexp = (int)floor(log10(fabs(x / toler)) - 1);
    if (iexp < 0) iexp = 0;
  } else {
    iexp = DBL_DIG - 1;
  }
  if (iexp > DBL_DIG) iexp = DBL_DIG + 1;
  h = -1;
  i = iexp;
  while (1) {
    snprintf(fmt, sizeof fmt, "%%.%de", i);
    snprintf(buffer, 32, fmt, x);
    x2 = atof(buffer);
    if (h > 0 && (fabs(x - x2) <= toler || i >= DBL_DIG + 3)) break;
    if (fabs(x - x2) > toler || i <= 0) h = 1;
    i += h;
  }
  dotpos = strchr(buffer, '.');
  epos = strchr(buffer, 'e');
  if (epos && !dotpos) { /* 2e+00 becomes 2.e+00 */
    memmove(epos + 1, epos, buffer + 31 - epos);
    dotpos = epos++;
    *dotpos = '.';
  }
  if (dotpos && epos) {
    char s[8] = "", m[96], ant[96], expo[16], num[16];
    int lm, la;
    lm = (int)(dotpos - buffer);
    memcpy(m, buffer, lm);
    m[lm] = 0;
    la = (int)(epos - dotpos - 1);
    memcpy(ant, dotpos + 1, la);
    ant[la] = 0;
    snprintf(expo, sizeof expo, "%s", epos + 1);
    if (strlen(m) > 1) {
      s[0] = m[0];
      s[1] = 0;
      memmove(m, m + 1, strlen(m));
    }
    iexp = atoi(expo);
    {
      /* erase trailing zeros of the mantissa digits */
      int z = (int)strlen(ant) - 1;
      while (z >= 0 && ant[z] == '0') z--;
      ant[z + 1] = 0;
    }
    if (iexp < 0 && iexp > -5) {
      char t[200];
      snprintf(t, sizeof t, "%s%s", m, ant);
      snprintf(ant, sizeof ant, "%s", t);
      m[0] = 0;
      iexp++;
    }
    if (iexp > 0) {
      chexp = iexp;
      if (chexp > (int)strlen(ant)) chexp = (int)strlen(ant);
      lm = (int)strlen(m);
      memcpy(m + lm, ant, chexp);
      m[lm + chexp] = 0;
      memmove(ant, ant + chexp, strlen(ant) - chexp + 1);
      iexp -= chexp;
    }
    while (iexp > -5 && iexp < 0 && strlen(m) == 0) {
      char t[200];
      snprintf(t, sizeof t, "0%s", ant);
      snprintf(ant, sizeof ant, "%s", t);
      iexp++;
    }
    while (iexp < 3 && iexp > 0 && strlen(ant) == 0) {
      strcat(m, "0");
      iexp--;
    }
    snprintf(num, sizeof num, "%d", iexp);
    snprintf(out, cap, "%s%s%s%s%s%s", s, m, strlen(ant) ? "." : "", ant, iexp ? "e" : "",
             iexp ? num : "");
  } else {
    snprintf(out, cap, "%s", buffer);
  }
}

/* ------------------------------------------------------------------ */
/* The store's block format (octree.h:35-43).                          */
/* ------------------------------------------------------------------ */
#define WBO_STORE_RECORD 61
#define WBO_STORE_RECORDS 537
#define WBO_STORE_BLOCK 32768

/* LasPoint::write (las.cpp:177-207), built without WAVEFORM: the location,
 * then the fields in declaration order, little-endian; the two flags share
 * one byte as edgeLine*2+scanDirection */
void wbo_store_record(const wbo_las_point *p, uint8_t *o)
{
  wrd(o, 0, p->x);
  wrd(o, 8, p->y);
  wrd(o, 16, p->z);
  wr(o, 24, p->intensity, 2);
  wr(o, 26, p->ret, 2);
  wr(o, 28, p->nret, 2);
  o[30] = (uint8_t)(p->edge * 2 + p->scan_dir);
  wr(o, 31, p->cls, 2);
  wr(o, 33, p->cls_flags, 2);
  wr(o, 35, p->channel, 2);
  wr(o, 37, p->user, 2);
  wr(o, 39, p->source, 2);
  wr(o, 41, (uint32_t)p->scan_angle, 4);
  wrd(o, 45, p->gps);
  wr(o, 53, p->nir, 2);
  wr(o, 55, p->red, 2);
  wr(o, 57, p->green, 2);
  wr(o, 59, p->blue, 2);
}

/* OctBuffer::write (octree.cpp:427-460): n <= 537 points, then noPoint
 * (LasPoint(), las.cpp:122-132: NaN location, zero fields) up to RECORDS,
 * then zero bytes up to BLOCKSIZE.  Returns 0, or -1 for n > 537. */
int wbo_store_block(const wbo_las_point *pts, int n, uint8_t *out)
{
  wbo_las_point none;
  if (n < 0 || n > WBO_STORE_RECORDS) return -1;
  memset(&none, 0, sizeof none);
  none.x = none.y = none.z = NAN;
  for (int r = 0; r < WBO_STORE_RECORDS; r++)
    wbo_store_record(r < n ? &pts[r] : &none, out + (size_t)r * WBO_STORE_RECORD);
  memset(out + WBO_STORE_RECORDS * WBO_STORE_RECORD, 0,
         WBO_STORE_BLOCK - WBO_STORE_RECORDS * WBO_STORE_RECORD);
  return 0;
}

/* LasPoint::read (las.cpp:139-176) of one record */
void wbo_store_read_record(const uint8_t *d, wbo_las_point *p)
{
  const size_t L = WBO_STORE_RECORD;
  p->x = rdd(d, L, 0);
  p->y = rdd(d, L, 8);
  p->z = rdd(d, L, 16);
  p->intensity = (uint16_t)rd(d, L, 24, 2);
  p->ret = (uint16_t)rd(d, L, 26, 2);
  p->nret = (uint16_t)rd(d, L, 28, 2);
  p->scan_dir = d[30] & 1;
  p->edge = (d[30] >> 1) & 1;
  p->cls = (uint16_t)rd(d, L, 31, 2);
  p->cls_flags = (uint16_t)rd(d, L, 33, 2);
  p->channel = (uint16_t)rd(d, L, 35, 2);
  p->user = (uint16_t)rd(d, L, 37, 2);
  p->source = (uint16_t)rd(d, L, 39, 2);
  p->scan_angle = (int32_t)(uint32_t)rd(d, L, 41, 4);
  p->gps = rdd(d, L, 45);
  p->nir = (uint16_t)rd(d, L, 53, 2);
  p->red = (uint16_t)rd(d, L, 55, 2);
  p->green = (uint16_t)rd(d, L, 57, 2);
  p->blue = (uint16_t)rd(d, L, 59, 2);
}

/* OctBuffer::read (octree.cpp:508-559): the records whose location has no
 * NaN (LasPoint::isEmpty, point.cpp:252-255) in record order; none when
 * there are two or more and the first two locations are equal (a block
 * never written reads as zeros).  Returns the number kept. */
int wbo_store_read_block(const uint8_t *blk, wbo_las_point *out)
{
  int n = 0;
  for (int r = 0; r < WBO_STORE_RECORDS; r++) {
    wbo_las_point p;
    wbo_store_read_record(blk + (size_t)r * WBO_STORE_RECORD, &p);
    if (!(isnan(p.x) || isnan(p.y) || isnan(p.z))) out[n++] = p;
  }
  if (n > 1 && out[0].x == out[1].x && out[0].y == out[1].y && out[0].z == out[1].z) n = 0;
  return n;
}

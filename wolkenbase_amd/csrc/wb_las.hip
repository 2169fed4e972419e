// wb_las.hip -- LAS files into and out of HBM.
//
// Reading: LasHeader::openRead (las.cpp:272-428) parses the header on the
// host; the point records are copied to HBM as they lie in the file and one
// kernel decodes every record as LasHeader::readPoint does (las.cpp:808-904)
// into FP64 coordinates for the octree build.  A cloud is the points of one
// or more files after the READ phase's rule for return number 0
// (threads.cpp:479-545) or wolkencli's keep-everything rule
// (wolkencli.cpp:99-111).
//
// Writing: CloudOutput (cloudoutput.cpp:119-246) with the octree's leaves in
// DFS order as its blocks.  The file plan (which file of a class each
// block's points go to) is the reference's sequential least-filled choice,
// made on the host from per-leaf class counts; one kernel per output point
// format then re-encodes every stored record as LasHeader::writePoint does
// (las.cpp:906-990) straight into the files' images in HBM, and the host
// writes LasHeader::writeHeader's header (las.cpp:549-595) in front.
#include "../../include/wb_gpu.h"
#include "wb_internal.h"

#include <math.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <string>
#include <vector>

// las.cpp:34-39
#define MASK_GPSTIME 0x7fa
#define MASK_RGB 0x5ac
#define MASK_NIR 0x500
static const short kPointLength[] = {20, 28, 26, 34, 57, 63, 30, 36, 38, 59, 67};
static const short kPointFeatures[] = {0x0, 0x1, 0x2, 0x3, 0x9, 0xb, 0x101, 0x103, 0x107, 0x109, 0x10f};
// VERSION of the reference's build (CMakeLists.txt:126-129), in the software field
#define WB_LAS_SOFTWARE "Wolkenbase 0.1.2~alpha"
#define WB_LAS_HEADER 375  // setVersion(1,4): header size 0x177 (las.cpp:603-615)

struct wb_las {
  wb_ctx *ctx;
  wb_las_header h;
  uint64_t n;       // records read (the repaired point count)
  uint8_t *rec;     // n * point_length bytes in HBM, as in the file
  size_t bytes;
  int ret0;         // return number of record 0, -1 when there is none
};

// one input file as the kernels see it
struct LasSrc {
  const uint8_t *rec;
  uint32_t len;
  int32_t format;
  double scale[3], offset[3];
  uint64_t begin;  // cloud index of the file's first point
  int32_t fix;     // a return number 0 is read as 1
  int32_t pad;
};

struct wb_cloud {
  wb_ctx *ctx;
  std::vector<wb_las *> files;
  std::vector<LasSrc> src;
  double unit;
  int zeros;
  uint64_t n;
  double *x, *y, *z;  // device, cloud order
  uint32_t *rec;      // device, record index of each point in its file
  LasSrc *d_src;
  std::vector<double> corners;  // min and max corner (times unit) of each file
  double resolution[3];
  int join_format;
  double ms_decode;  // device time of the keep flags and the decode
};

static int fail(wb_ctx *c, const std::string &msg, int code)
{
  if (c) c->err = msg;
  return code;
}

#define API_TRY(ctx) try {
#define API_CATCH(ctx)                                             \
  }                                                                \
  catch (const std::string &m) { return fail(ctx, m, WB_EDEVICE); } \
  catch (...) { return fail(ctx, "unexpected exception", WB_EDEVICE); }

// ---------------------------------------------------------------------------
// host: header
// ---------------------------------------------------------------------------
static uint64_t le(const uint8_t *b, size_t len, size_t off, int nbytes)
{
  uint64_t v = 0;
  for (int k = 0; k < nbytes; k++)
    if (off + k < len) v |= (uint64_t)b[off + k] << (8 * k);
  return v;
}
static double led(const uint8_t *b, size_t len, size_t off)
{
  uint64_t v = le(b, len, off, 8);
  double d;
  memcpy(&d, &v, 8);
  return d;
}

// LasHeader::openRead (las.cpp:272-428) on the first bytes of a file; fields
// past the end of the buffer read as zero.  Returns false if it does not
// start "LASF".
static bool parse_header(const uint8_t *b, size_t len, wb_las_header *h)
{
  memset(h, 0, sizeof *h);
  if (len < 4 || memcmp(b, "LASF", 4) != 0) return false;
  h->source_id = (uint16_t)le(b, len, 4, 2);
  h->global_encoding = (uint16_t)le(b, len, 6, 2);
  for (int k = 0; k < 16; k++) h->guid[k] = (uint8_t)le(b, len, 8 + k, 1);
  h->version_major = (uint8_t)le(b, len, 24, 1);
  h->version_minor = (uint8_t)le(b, len, 25, 1);
  for (int k = 0; k < 32; k++) {
    h->system_id[k] = (char)le(b, len, 26 + k, 1);
    h->software[k] = (char)le(b, len, 58 + k, 1);
  }
  h->creation_day = (uint16_t)le(b, len, 90, 2);
  h->creation_year = (uint16_t)le(b, len, 92, 2);
  // a short read into an unsigned int (readleshort, binio.cpp)
  h->header_size = (uint32_t)(int32_t)(int16_t)le(b, len, 94, 2);
  h->point_offset = (uint32_t)le(b, len, 96, 4);
  h->n_vlr = (uint32_t)le(b, len, 100, 4);
  h->point_format = (uint16_t)le(b, len, 104, 1);
  h->point_length = (uint16_t)le(b, len, 105, 2);
  uint32_t legacy[6];
  for (int i = 0; i < 6; i++) legacy[i] = (uint32_t)le(b, len, 107 + 4 * i, 4);
  for (int a = 0; a < 3; a++) {
    h->scale[a] = led(b, len, 131 + 8 * a);
    h->offset[a] = led(b, len, 155 + 8 * a);
    h->max[a] = led(b, len, 179 + 16 * a);
    h->min[a] = led(b, len, 187 + 16 * a);
  }
  // the LAS 1.4 part of the header follows when the header is longer than
  // the LAS 1.2 one (0xe3 bytes)
  if (h->header_size > 0xe3) {
    h->start_waveform = le(b, len, 227, 8);
    h->start_evlr = le(b, len, 235, 8);
    h->n_evlr = (uint32_t)le(b, len, 243, 4);
    for (int i = 0; i < 16; i++) h->n_points[i] = le(b, len, 247 + 8 * i, 8);
  }
  // Which of the two sets of counts to trust (las.cpp:354-402): bit 1 the
  // legacy counts add up, bit 2 the 64-bit ones do, bits 4 and 8 the same
  // sets with a total but no per-return counts.
  int which = 15;
  uint64_t total = 0;
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
    // a total without per-return counts: all points are first returns
    h->n_points[1] = h->n_points[0];
    legacy[1] = legacy[0];
  }
  if (which == 1 || which == 4)
    for (int i = 0; i < 6; i++) h->n_points[i] = legacy[i];
  if (which == 0)
    for (int i = 0; i < 6; i++) h->n_points[i] = 0;
  if (h->point_length == 0) {
    h->version_major = h->version_minor = 0;
    h->n_points[0] = 0;
  }
  // LASzip's extra header starts "\0\0laszip" right after the header
  if ((uint64_t)h->point_offset >= (uint64_t)h->header_size + 8) {
    uint64_t m = 0;
    for (int k = 0; k < 8; k++) m = (m << 8) | le(b, len, (size_t)h->header_size + k, 1);
    if ((m & 0xffffffffffffull) == 0x6c61737a6970ull) h->zipped = 1;
  }
  // LasHeader::isValid (las.cpp:597-601)
  h->valid = h->version_major > 0 && h->version_minor > 0 && h->header_size > 0 &&
             h->point_length > 0 && (h->header_size > 0xe3 || h->point_format < 6);
  return true;
}

// ---------------------------------------------------------------------------
// device: record fields
// ---------------------------------------------------------------------------
__device__ inline uint32_t ld_le(const uint8_t *p, int nbytes)
{
  uint32_t v = 0;
  for (int k = 0; k < nbytes; k++) v |= (uint32_t)p[k] << (8 * k);
  return v;
}
__device__ inline double ld_f64(const uint8_t *p)
{
  const uint64_t v = (uint64_t)ld_le(p, 4) | ((uint64_t)ld_le(p + 4, 4) << 32);
  return __longlong_as_double((long long)v);
}
__device__ inline int ret_of(const uint8_t *r, int format)
{
  return format < 6 ? (r[14] & 7) : (r[14] & 15);
}

// angle.cpp:209-247 (bintodeg, rottobin, degtobin); modf's fraction is
// x - trunc(x), exact
__device__ inline double d_bintodeg(int a) { return (a / 2147483648.) * 360; }
__device__ inline int d_degtobin(double angle)
{
  const double h = angle / 360 / 2;
  double fprt = 2 * (h - trunc(h));
  if (fprt >= 1) fprt -= 2;
  if (fprt < -1) fprt += 2;
  return (int)(long long)rint(2147483648. * fprt);
}

struct LasPt {
  double x, y, z;
  int32_t angle;
  double gps;
  uint16_t intensity, ret, nret, cls, flags, channel, user, source, nir, red, green, blue;
  uint8_t dir, edge;
};

// LasHeader::readPoint (las.cpp:808-904) of one record
__device__ inline void read_point(const uint8_t *r, const LasSrc &s, double unit, LasPt &p)
{
  const int f = s.format;
  const int32_t xi = (int32_t)ld_le(r, 4), yi = (int32_t)ld_le(r + 4, 4), zi = (int32_t)ld_le(r + 8, 4);
  p.intensity = (uint16_t)ld_le(r + 12, 2);
  int o;
  if (f < 6) {
    const int t = r[14];
    p.ret = t & 7;
    p.nret = (t >> 3) & 7;
    p.dir = (t >> 6) & 1;
    p.edge = (t >> 7) & 1;
    p.cls = r[15] & 31;
    p.flags = (r[15] >> 5) & 7;
    p.channel = 0;
    p.angle = d_degtobin((double)(signed char)r[16]);
    p.user = r[17];
    p.source = (uint16_t)ld_le(r + 18, 2);
    o = 20;
  } else {
    const int t = r[14], u = r[15];
    p.ret = t & 15;
    p.nret = (t >> 4) & 15;
    p.flags = u & 15;
    p.channel = (u >> 4) & 3;
    p.dir = (u >> 6) & 1;
    p.edge = (u >> 7) & 1;
    p.cls = r[16];
    p.user = r[17];
    p.angle = d_degtobin((int16_t)ld_le(r + 18, 2) * 0.006);
    p.source = (uint16_t)ld_le(r + 20, 2);
    o = 22;
  }
  p.gps = 0;
  p.nir = p.red = p.green = p.blue = 0;
  if ((1 << f) & MASK_GPSTIME) {
    p.gps = ld_f64(r + o);
    o += 8;
  }
  if ((1 << f) & MASK_RGB) {
    p.red = (uint16_t)ld_le(r + o, 2);
    p.green = (uint16_t)ld_le(r + o + 2, 2);
    p.blue = (uint16_t)ld_le(r + o + 4, 2);
    o += 6;
  }
  if ((1 << f) & MASK_NIR) p.nir = (uint16_t)ld_le(r + o, 2);
  p.x = (s.offset[0] + s.scale[0] * (double)xi) * unit;
  p.y = (s.offset[1] + s.scale[1] * (double)yi) * unit;
  p.z = (s.offset[2] + s.scale[2] * (double)zi) * unit;
}

// ---------------------------------------------------------------------------
// device: decode
// ---------------------------------------------------------------------------
__global__ void k_las_keep(const uint8_t *rec, uint64_t n, uint32_t len, int format,
                           uint32_t *keep)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) keep[i] = ret_of(rec + i * len, format) != 0 ? 1u : 0u;
}

// pos: exclusive scan of the keep flags (null: every record is kept)
__global__ void k_las_decode(LasSrc s, uint64_t n, double unit, const uint32_t *pos, double *x,
                             double *y, double *z, uint32_t *recidx)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint8_t *r = s.rec + i * s.len;
  uint64_t j = i;
  if (pos) {
    if (ret_of(r, s.format) == 0) return;
    j = pos[i];
  }
  const int32_t xi = (int32_t)ld_le(r, 4), yi = (int32_t)ld_le(r + 4, 4), zi = (int32_t)ld_le(r + 8, 4);
  j += s.begin;
  x[j] = (s.offset[0] + s.scale[0] * (double)xi) * unit;
  y[j] = (s.offset[1] + s.scale[1] * (double)yi) * unit;
  z[j] = (s.offset[2] + s.scale[2] * (double)zi) * unit;
  recidx[j] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// device: output
// ---------------------------------------------------------------------------
struct OutXf {
  double scale[3], offset[3], unit;
};

// Counters that many waves update are kept in WB_NCOPY copies, each block
// using copy blockIdx % WB_NCOPY, and summed on the host: one address per
// counter would serialise every wave's atomic in one L2 channel.
#define WB_NCOPY 64

// per output file: point counts by return, and the written coordinates'
// extremes as order-preserving integer keys
struct FileStats {
  unsigned long long npoints[16];
  unsigned long long mn[3], mx[3];
};

__device__ inline unsigned long long order_key(double d)
{
  unsigned long long k = (unsigned long long)__double_as_longlong(d);
  return (k >> 63) ? ~k : (k | 0x8000000000000000ull);
}

// the stored point of one output position: the record of the last input at
// its location, the class of the classification (or the record's own class
// where no tile classified it), the READ phase's return fix
__device__ inline void stored_point(const LasSrc *files, int nfiles, const uint32_t *recidx,
                                    double unit, uint32_t c_rec, uint8_t cls_in, LasPt &p)
{
  int lo = 0, hi = nfiles - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    if (files[mid].begin <= c_rec) lo = mid;
    else hi = mid - 1;
  }
  const LasSrc &s = files[lo];
  read_point(s.rec + (uint64_t)recidx[c_rec] * s.len, s, unit, p);
  if (s.fix && p.ret == 0) p.ret = 1;
  if (cls_in) p.cls = cls_in;
}

// class of each stored point in canonical order, and the class totals
// (OctBuffer::countClasses, octree.cpp:665-671)
__global__ void k_out_class(uint64_t n, const uint32_t *canon, const uint32_t *sidx,
                            const uint32_t *last_of, const uint8_t *cls, const LasSrc *files,
                            int nfiles, const uint32_t *recidx, double unit, uint8_t *ocls,
                            unsigned long long *totals)
{
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool on = k < n;
  int v = -1;
  if (on) {
    const uint32_t s = canon[k];
    const uint32_t first = sidx[s], last = last_of ? last_of[s] : first;
    LasPt p;
    stored_point(files, nfiles, recidx, unit, last, cls ? cls[first] : 0, p);
    v = p.cls & 255;
    ocls[k] = (uint8_t)v;
  }
  // one atomic per distinct class in the wave
  uint64_t rem = __ballot(on);
  while (rem) {
    const int leader = __ffsll((long long)rem) - 1;
    const int cv = __shfl(v, leader, 64);
    const uint64_t m = __ballot(on && v == cv);
    if ((int)__lane_id() == leader)
      atomicAdd(&totals[256 * (blockIdx.x % WB_NCOPY) + cv], (unsigned long long)__popcll(m));
    rem &= ~m;
  }
}

// per leaf, the number of its points of each planned class id
__global__ void k_leaf_class_count(const uint32_t *leaf_start, uint32_t nl, const uint8_t *ocls,
                                   const int16_t *cid, int K, uint32_t *counts)
{
  __shared__ uint32_t c[256];
  const uint32_t l = blockIdx.x;
  if (l >= nl) return;
  const uint32_t b = leaf_start[l], e = leaf_start[l + 1];
  for (int k = threadIdx.x; k < K; k += blockDim.x) c[k] = 0;
  __syncthreads();
  for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) atomicAdd(&c[cid[ocls[i]]], 1u);
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += blockDim.x) counts[(uint64_t)l * K + k] = c[k];
}

template <int OFF, int NB>
__device__ inline void put(uint32_t *w, uint64_t v)
{
#pragma unroll
  for (int k = 0; k < NB; k++) {
    const int b = OFF + k;
    w[b >> 2] |= (uint32_t)((v >> (8 * k)) & 0xff) << (8 * (b & 3));
  }
}

// LasHeader::writePoint's record (las.cpp:917-987) in format FO, into words
template <int FO>
__device__ inline void encode(const LasPt &p, int xi, int yi, int zi, uint32_t *w)
{
  constexpr int LEN = FO == 0 ? 20 : FO == 1 ? 28 : FO == 2 ? 26 : FO == 3 ? 34 : FO == 4 ? 57
                      : FO == 5 ? 63 : FO == 6 ? 30 : FO == 7 ? 36 : FO == 8 ? 38 : FO == 9 ? 59 : 67;
#pragma unroll
  for (int k = 0; k < (LEN + 3) / 4; k++) w[k] = 0;
  put<0, 4>(w, (uint32_t)xi);
  put<4, 4>(w, (uint32_t)yi);
  put<8, 4>(w, (uint32_t)zi);
  put<12, 2>(w, p.intensity);
  constexpr int O = FO < 6 ? 20 : 22;
  if (FO < 6) {
    put<14, 1>(w, (p.ret & 7) + ((p.nret & 7) << 3) + ((p.dir & 1) << 6) + ((p.edge & 1) << 7));
    put<15, 1>(w, (p.cls & 31) + ((p.flags & 7) << 5));
    put<16, 1>(w, (uint64_t)(long long)rint(d_bintodeg(p.angle)));
    put<17, 1>(w, p.user);
    put<18, 2>(w, p.source);
  } else {
    put<14, 1>(w, (p.ret & 15) + ((p.nret & 15) << 4));
    put<15, 1>(w, (p.flags & 15) + ((p.channel & 3) << 4) + ((p.dir & 1) << 6) + ((p.edge & 1) << 7));
    put<16, 1>(w, p.cls);
    put<17, 1>(w, p.user);
    put<18, 2>(w, (uint64_t)(long long)rint(d_bintodeg(p.angle) / 0.006));
    put<20, 2>(w, p.source);
  }
  constexpr bool GPS = (MASK_GPSTIME >> FO) & 1, RGB = (MASK_RGB >> FO) & 1, NIR = (MASK_NIR >> FO) & 1;
  constexpr int OG = O, OR = O + (GPS ? 8 : 0), ON = OR + (RGB ? 6 : 0);
  if (GPS) put<OG, 8>(w, (uint64_t)__double_as_longlong(p.gps));
  if (RGB) {
    put<OR, 2>(w, p.red);
    put<OR + 2, 2>(w, p.green);
    put<OR + 4, 2>(w, p.blue);
  }
  if (NIR) put<ON, 2>(w, p.nir);
  // waveform fields stay zero (the reference is built without WAVEFORM)
}

template <int LEN>
__device__ inline void store_record(uint8_t *dst, const uint32_t *w)
{
  const uintptr_t a = (uintptr_t)dst;
  if ((a & 3) == 0) {
#pragma unroll
    for (int k = 0; k < LEN / 4; k++) ((uint32_t *)dst)[k] = w[k];
#pragma unroll
    for (int b = LEN / 4 * 4; b < LEN; b++) dst[b] = (uint8_t)(w[b >> 2] >> (8 * (b & 3)));
  } else if ((a & 1) == 0) {
#pragma unroll
    for (int k = 0; k < LEN / 2; k++) ((uint16_t *)dst)[k] = (uint16_t)(w[k >> 1] >> (16 * (k & 1)));
    if (LEN & 1) dst[LEN - 1] = (uint8_t)(w[(LEN - 1) >> 2] >> (8 * ((LEN - 1) & 3)));
  } else {
#pragma unroll
    for (int b = 0; b < LEN; b++) dst[b] = (uint8_t)(w[b >> 2] >> (8 * (b & 3)));
  }
}

// Writes every stored point of one leaf per 64-lane block, in canonical
// order: lanes take 64 consecutive points; each point's position in its
// file is the leaf-class base of the plan plus its rank among the leaf's
// points of that class (running counts carried over the chunks).
template <int FO>
__global__ void __launch_bounds__(64) k_las_encode(
    const uint32_t *leaf_start, uint32_t nl, const uint32_t *canon, const uint32_t *sidx,
    const uint32_t *last_of, const uint8_t *cls, const uint8_t *ocls, const LasSrc *files,
    int nfiles, const uint32_t *recidx, double unit, const int16_t *cid, int K,
    const uint64_t *base, const int32_t *fileof, const uint64_t *file_data, OutXf o,
    uint8_t *out, FileStats *fst)
{
  constexpr int LEN = FO == 0 ? 20 : FO == 1 ? 28 : FO == 2 ? 26 : FO == 3 ? 34 : FO == 4 ? 57
                      : FO == 5 ? 63 : FO == 6 ? 30 : FO == 7 ? 36 : FO == 8 ? 38 : FO == 9 ? 59 : 67;
  __shared__ uint32_t run[256];  // points of each class id placed so far in this leaf
  const uint32_t l = blockIdx.x;
  if (l >= nl) return;
  const int lane = threadIdx.x;
  for (int k = lane; k < K; k += 64) run[k] = 0;
  __syncthreads();
  const uint32_t b = leaf_start[l], e = leaf_start[l + 1];
  for (uint32_t c0 = b; c0 < e; c0 += 64) {
    const uint32_t k = c0 + lane;
    const bool on = k < e;
    int id = -1, fg = -1, ret = 0;
    double wx = 0, wy = 0, wz = 0;
    uint32_t w[(LEN + 3) / 4];
    if (on) {
      const uint32_t s = canon[k];
      const uint32_t first = sidx[s], last = last_of ? last_of[s] : first;
      LasPt p;
      stored_point(files, nfiles, recidx, unit, last, cls ? cls[first] : 0, p);
      id = cid[ocls[k]];
      fg = fileof[(uint64_t)l * K + id];
      const int xi = (int)(long long)rint((p.x / o.unit - o.offset[0]) / o.scale[0]);
      const int yi = (int)(long long)rint((p.y / o.unit - o.offset[1]) / o.scale[1]);
      const int zi = (int)(long long)rint((p.z / o.unit - o.offset[2]) / o.scale[2]);
      wx = xi * o.scale[0] + o.offset[0];
      wy = yi * o.scale[1] + o.offset[1];
      wz = zi * o.scale[2] + o.offset[2];
      ret = p.ret;
      encode<FO>(p, xi, yi, zi, w);
    }
    // rank among the chunk's points of the same class id, and whether this
    // lane is the chunk's last point of its class id
    uint32_t rank = 0, cnt = 0;
    bool tail = false;
    uint64_t rem = __ballot(on);
    while (rem) {
      const int leader = __ffsll((long long)rem) - 1;
      const int cv = __shfl(id, leader, 64);
      const uint64_t m = __ballot(on && id == cv);
      if (on && id == cv) {
        rank = (uint32_t)__popcll(m & ((1ull << lane) - 1));
        cnt = (uint32_t)__popcll(m);
        tail = (m >> lane) == 1;
      }
      rem &= ~m;
    }
    const uint32_t before = on ? run[id] : 0;
    __syncthreads();
    if (on && tail) run[id] = before + cnt;
    __syncthreads();
    if (on) {
      const uint64_t pos = base[(uint64_t)l * K + id] + before + rank;
      store_record<LEN>(out + file_data[fg] + pos * LEN, w);
    }
    // per-file statistics, one atomic per distinct file and return in the wave
    rem = __ballot(on);
    while (rem) {
      const int leader = __ffsll((long long)rem) - 1;
      const int gv = __shfl(fg, leader, 64);
      const uint64_t m = __ballot(on && fg == gv);
      const bool mine = on && fg == gv;
      FileStats &fsc = fst[(uint64_t)gv * WB_NCOPY + blockIdx.x % WB_NCOPY];
      for (int r = 0; r < 16; r++) {
        const uint64_t mr = __ballot(mine && ret == r);
        if (mr && r > 0 && lane == leader) atomicAdd(&fsc.npoints[r], (unsigned long long)__popcll(mr));
      }
      if (lane == leader) atomicAdd(&fsc.npoints[0], (unsigned long long)__popcll(m));
      unsigned long long kmn[3], kmx[3];
      kmn[0] = mine ? order_key(wx) : ~0ull;
      kmn[1] = mine ? order_key(wy) : ~0ull;
      kmn[2] = mine ? order_key(wz) : ~0ull;
      for (int a = 0; a < 3; a++) kmx[a] = mine ? kmn[a] : 0ull;
      for (int sft = 1; sft < 64; sft <<= 1)
        for (int a = 0; a < 3; a++) {
          const unsigned long long u = __shfl_xor(kmn[a], sft, 64), v = __shfl_xor(kmx[a], sft, 64);
          kmn[a] = u < kmn[a] ? u : kmn[a];
          kmx[a] = v > kmx[a] ? v : kmx[a];
        }
      if (lane == leader)
        for (int a = 0; a < 3; a++) {
          atomicMin(&fsc.mn[a], kmn[a]);
          atomicMax(&fsc.mx[a], kmx[a]);
        }
      rem &= ~m;
    }
  }
}

// ---------------------------------------------------------------------------
// device: the store's block format (octree.h:35-43)
// ---------------------------------------------------------------------------
// LasPoint::write (las.cpp:177-207) of one point: 61 little-endian bytes
__device__ inline void put_bytes(uint8_t *d, uint64_t v, int n)
{
  for (int k = 0; k < n; k++) d[k] = (uint8_t)(v >> (8 * k));
}
__device__ inline void store_rec61(uint8_t *d, const LasPt &p)
{
  put_bytes(d, __double_as_longlong(p.x), 8);
  put_bytes(d + 8, __double_as_longlong(p.y), 8);
  put_bytes(d + 16, __double_as_longlong(p.z), 8);
  put_bytes(d + 24, p.intensity, 2);
  put_bytes(d + 26, p.ret, 2);
  put_bytes(d + 28, p.nret, 2);
  d[30] = (uint8_t)(p.edge * 2 + p.dir);
  put_bytes(d + 31, p.cls, 2);
  put_bytes(d + 33, p.flags, 2);
  put_bytes(d + 35, p.channel, 2);
  put_bytes(d + 37, p.user, 2);
  put_bytes(d + 39, p.source, 2);
  put_bytes(d + 41, (uint32_t)p.angle, 4);
  put_bytes(d + 45, __double_as_longlong(p.gps), 8);
  put_bytes(d + 53, p.nir, 2);
  put_bytes(d + 55, p.red, 2);
  put_bytes(d + 57, p.green, 2);
  put_bytes(d + 59, p.blue, 2);
}

// OctBuffer::write (octree.cpp:427-460) of leaf l as block l: the leaf's
// stored points in canonical order (with a cloud, the record of the
// location's last input as readPoint decodes it, with the READ phase's
// return fix; without, LasPoint's defaults at the point's location; a
// nonzero class byte replaces the classification, as CLASSIFY's put does),
// noPoint records (NaN location, zero fields) to 537, zero bytes to 32768.
// The block is assembled in LDS and copied out in 16-byte words.
__global__ void __launch_bounds__(64) k_store_blocks(
    const uint32_t *leaf_start, uint32_t nl, const uint32_t *canon, const uint32_t *sidx,
    const uint32_t *last_of, const uint8_t *cls, const LasSrc *files, int nfiles,
    const uint32_t *recidx, double unit, const double *px, const double *py, const double *pz,
    uint8_t *out)
{
  __shared__ uint4 blk[WB_STORE_BLOCK / 16];
  const uint32_t l = blockIdx.x;
  if (l >= nl) return;
  uint8_t *bb = (uint8_t *)blk;
  const uint32_t b = leaf_start[l], m = leaf_start[l + 1] - b;
  for (uint32_t r = threadIdx.x; r < WB_STORE_RECORDS; r += blockDim.x) {
    LasPt p;
    p.x = p.y = p.z = __builtin_nan("");
    p.angle = 0;
    p.gps = 0;
    p.intensity = p.ret = p.nret = p.cls = p.flags = p.channel = p.user = p.source = 0;
    p.nir = p.red = p.green = p.blue = 0;
    p.dir = p.edge = 0;
    if (r < m) {
      const uint32_t sl = canon[b + r];
      const uint32_t first = sidx[sl], last = last_of ? last_of[sl] : first;
      const uint8_t c = cls ? cls[first] : 0;
      if (files) {
        stored_point(files, nfiles, recidx, unit, last, c, p);
      } else {
        p.x = px[sl]; p.y = py[sl]; p.z = pz[sl];
        p.cls = c;
      }
    }
    store_rec61(bb + (size_t)r * WB_STORE_RECORD, p);
  }
  for (uint32_t k = WB_STORE_RECORDS * WB_STORE_RECORD + threadIdx.x; k < WB_STORE_BLOCK;
       k += blockDim.x)
    bb[k] = 0;
  __syncthreads();
  uint4 *dst = (uint4 *)(out + (size_t)l * WB_STORE_BLOCK);
  for (uint32_t k = threadIdx.x; k < WB_STORE_BLOCK / 16; k += blockDim.x) dst[k] = blk[k];
}

__device__ inline uint64_t get_bytes(const uint8_t *s, int n)
{
  uint64_t v = 0;
  for (int k = 0; k < n; k++) v |= (uint64_t)s[k] << (8 * k);
  return v;
}

// OctBuffer::read (octree.cpp:508-559) of block b, first half: the records
// with a location (LasPoint::isEmpty: a NaN coordinate), counted; a block
// whose first two kept locations are equal (a block never written reads as
// all zeros) keeps none.  One wave per block.
__global__ void __launch_bounds__(64) k_store_count(const uint8_t *blocks, uint64_t nblocks,
                                                    uint32_t *count)
{
  const uint64_t bk = blockIdx.x;
  if (bk >= nblocks) return;
  const uint8_t *base = blocks + bk * WB_STORE_BLOCK;
  uint32_t n = 0, found = 0;
  double f[2][3] = {{0, 0, 0}, {0, 0, 0}};
  for (uint32_t r0 = 0; r0 < WB_STORE_RECORDS; r0 += 64) {
    const uint32_t r = r0 + threadIdx.x;
    bool keep = false;
    double x = 0, y = 0, z = 0;
    if (r < WB_STORE_RECORDS) {
      const uint8_t *d = base + (size_t)r * WB_STORE_RECORD;
      x = __longlong_as_double((long long)get_bytes(d, 8));
      y = __longlong_as_double((long long)get_bytes(d + 8, 8));
      z = __longlong_as_double((long long)get_bytes(d + 16, 8));
      keep = !(isnan(x) || isnan(y) || isnan(z));
    }
    // the block's first two kept locations
    uint64_t km = __ballot(keep);
    n += (uint32_t)__popcll(km);
    while (found < 2 && km) {
      const int ln = __ffsll((long long)km) - 1;
      f[found][0] = __shfl(x, ln, 64);
      f[found][1] = __shfl(y, ln, 64);
      f[found][2] = __shfl(z, ln, 64);
      found++;
      km &= km - 1;
    }
  }
  if (n > 1 && f[0][0] == f[1][0] && f[0][1] == f[1][1] && f[0][2] == f[1][2]) n = 0;
  if (threadIdx.x == 0) count[bk] = n;
}

// second half: the kept records of block b at pos[b], in record order
__global__ void __launch_bounds__(64) k_store_points(const uint8_t *blocks, uint64_t nblocks,
                                                     const uint32_t *count, const uint32_t *pos,
                                                     double *x, double *y, double *z,
                                                     uint16_t *cls)
{
  const uint64_t bk = blockIdx.x;
  if (bk >= nblocks || count[bk] == 0) return;
  const uint8_t *base = blocks + bk * WB_STORE_BLOCK;
  uint32_t at = pos[bk];
  for (uint32_t r0 = 0; r0 < WB_STORE_RECORDS; r0 += 64) {
    const uint32_t r = r0 + threadIdx.x;
    bool keep = false;
    double vx = 0, vy = 0, vz = 0;
    uint16_t c = 0;
    if (r < WB_STORE_RECORDS) {
      const uint8_t *d = base + (size_t)r * WB_STORE_RECORD;
      vx = __longlong_as_double((long long)get_bytes(d, 8));
      vy = __longlong_as_double((long long)get_bytes(d + 8, 8));
      vz = __longlong_as_double((long long)get_bytes(d + 16, 8));
      c = (uint16_t)get_bytes(d + 31, 2);
      keep = !(isnan(vx) || isnan(vy) || isnan(vz));
    }
    const uint64_t km = __ballot(keep);
    if (keep) {
      const uint32_t o = at + (uint32_t)__popcll(km & ((1ull << threadIdx.x) - 1));
      x[o] = vx;
      y[o] = vy;
      z[o] = vz;
      if (cls) cls[o] = c;
    }
    at += (uint32_t)__popcll(km);
  }
}

// ---------------------------------------------------------------------------
// host: the reference's small helpers
// ---------------------------------------------------------------------------
// joinPointFormat (las.cpp:104-121): the least format with every feature
static int join_point_format(const std::vector<int> &formats)
{
  int all = 0, ret = 0;
  for (int f : formats)
    if (f >= 0 && f < 11) all |= kPointFeatures[f];
  for (int i = 0; i < 11; i++)
    if ((all | kPointFeatures[i]) == kPointFeatures[i]) {
      ret = i;
      break;
    }
  return ret;
}

// pairwisesum (manysum.cpp:120-154): balanced blocks, one per set bit of n
static double pairwise(const double *a, size_t n)
{
  if (n == 0) return 0;
  if (n == 1) return a[0];
  size_t h = 1;
  while (2 * h < n) h *= 2;
  if (h == n) return pairwise(a, n / 2) + pairwise(a + n / 2, n / 2);
  return pairwise(a, h) + pairwise(a + h, n - h);
}
static double pairwisesum(const std::vector<double> &v)
{
  // the blocks of the set bits of n, largest first, added to 0 in
  // ascending bit order
  const size_t n = v.size();
  std::vector<double> part;
  size_t at = 0;
  for (int b = 63; b >= 0; b--)
    if ((n >> b) & 1) {
      const size_t len = (size_t)1 << b;
      part.push_back(pairwise(v.data() + at, len));
      at += len;
    }
  double s = 0;
  for (size_t k = part.size(); k-- > 0;) s += part[k];
  return s;
}

// manygcd (manygcd.cpp:44-73): repeated subtraction of the two largest,
// within toler, then the mean of what is left
static double manygcd(std::vector<double> v, double toler)
{
  for (double &x : v) x = fabs(x);
  auto up = [&]() {
    for (size_t i = 0; i + 1 < v.size(); i++)
      if (v[i] > v[i + 1]) std::swap(v[i], v[i + 1]);
  };
  auto down = [&]() {
    for (size_t i = v.size() - 1; i-- > 0;)
      if (v[i] > v[i + 1]) std::swap(v[i], v[i + 1]);
  };
  if (!v.empty()) {
    up();
    down();
    up();
  }
  while (!v.empty() && v.back() > v[0] + toler) {
    const size_t sz = v.size();
    if (v[sz - 1] - v[sz - 2] > toler) {
      v[sz - 1] -= v[sz - 2];
      down();
    } else {
      v.resize(sz - 1);
    }
    up();
  }
  return v.empty() ? 0 : pairwisesum(v) / v.size();
}

// combine1Scale (las.cpp:913-932)
static double combine1_scale(const std::vector<double> &scales, const std::vector<double> &offsets)
{
  double toler = INFINITY, ret = 0;
  for (double s : scales)
    if (fabs(s) < toler) toler = fabs(s);
  toler /= 1e6;
  const double sg = manygcd(scales, toler), og = manygcd(offsets, toler);
  if ((og > 10.5 * sg || og == 0) && fabs(og / sg - rint(og / sg)) < 1e-6) ret = sg;
  for (int i = 10; i > 0; i--)
    for (int j = 10; j > 0; j--)
      if (fabs(sg * i - og * j) < toler) ret = sg / j;
  return ret;
}

// LasHeader::setScale (las.cpp:629-665), one axis
static void set_scale(double minc, double maxc, double scale, double unit, double *out_scale,
                      double *out_offset)
{
  const double minscale = (maxc - minc) / 4132485216.;
  double off = (minc + maxc) / 2 / unit;
  if (scale > minscale && std::isfinite(scale)) {
    *out_scale = scale / unit;
    off = rint(off / *out_scale) * *out_scale;
  } else {
    *out_scale = minscale / unit;
  }
  *out_offset = off;
}

// CloudOutput::className (cloudoutput.cpp:39-117), untranslated
static std::string class_name(int n)
{
  static const char *names[] = {"raw", "nonground", "ground", "lowveg", "medveg", "highveg",
                                "building", "lownoise", nullptr, "water", "rail", "road",
                                "overlap", "wireguard", "conductor", "tower", "insulator",
                                "bridge", "highnoise", "overhead", "ignground", "snow",
                                "temporal"};
  if (n >= 0 && n <= 22 && names[n]) return names[n];
  return "reserved" + std::to_string(n);
}

// ndecimal (cloudoutput.cpp:26-37)
static std::string ndecimal(size_t n, int dig)
{
  if (!dig) return "";
  char buf[32];
  snprintf(buf, sizeof buf, "%0*zu", dig, n);
  return buf;
}

static void put_le(uint8_t *b, size_t off, uint64_t v, int nbytes)
{
  for (int k = 0; k < nbytes; k++) b[off + k] = (uint8_t)(v >> (8 * k));
}
static void put_d(uint8_t *b, size_t off, double d)
{
  uint64_t v;
  memcpy(&v, &d, 8);
  put_le(b, off, v, 8);
}

static double from_key(unsigned long long k)
{
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  double d;
  memcpy(&d, &b, 8);
  return d;
}

extern "C" {

int wb_las_read_header(const char *path, wb_las_header *hdr)
{
  if (!path || !hdr) return WB_EINVAL;
  FILE *f = fopen(path, "rb");
  if (!f) return WB_EINVAL;
  uint8_t buf[512];
  const size_t got = fread(buf, 1, sizeof buf, f);
  fclose(f);
  return parse_header(buf, got, hdr) ? WB_OK : WB_EINVAL;
}

static int las_open(wb_ctx *c, FILE *fp, const uint8_t *mem, uint64_t memlen, wb_las **out,
                    wb_las_header *hdr)
{
  uint8_t head[512];
  uint64_t flen;
  size_t got;
  if (fp) {
    fseek(fp, 0, SEEK_END);
    flen = (uint64_t)ftell(fp);
    fseek(fp, 0, SEEK_SET);
    got = fread(head, 1, sizeof head, fp);
  } else {
    flen = memlen;
    got = (size_t)std::min<uint64_t>(memlen, sizeof head);
    memcpy(head, mem, got);
  }
  wb_las_header h;
  if (!parse_header(head, got, &h)) return fail(c, "not a LAS file (no LASF signature)", WB_EINVAL);
  if (h.zipped) return fail(c, "LASzip-compressed file: not readable here", WB_EINVAL);
  const uint64_t n = h.n_points[0];
  if (n > 0) {
    if (h.point_format > 10) return fail(c, "point format above 10", WB_ERANGE);
    if (h.point_length < kPointLength[h.point_format])
      return fail(c, "point records shorter than their format", WB_ERANGE);
    if (n > 0xffffffffull) return fail(c, "more than 2^32-1 points", WB_ERANGE);
    if ((uint64_t)h.point_offset + n * h.point_length > flen)
      return fail(c, "file shorter than its point records", WB_EINVAL);
  }
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  wb_las *L = new wb_las();
  L->ctx = c;
  L->h = h;
  L->n = n;
  L->bytes = (size_t)(n * h.point_length);
  L->rec = (uint8_t *)wb_alloc(c, L->bytes ? L->bytes : 16);
  L->ret0 = -1;
  if (n > 0) {
    if (fp) {
      // chunks through two pinned buffers: the read of one overlaps the
      // copy of the other
      const size_t CH = 64u << 20;
      uint8_t *pin[2] = {nullptr, nullptr};
      hipEvent_t done[2];
      try {
        for (int k = 0; k < 2; k++) {
          WB_CHECK(hipHostMalloc((void **)&pin[k], CH, hipHostMallocDefault));
          WB_CHECK(hipEventCreate(&done[k]));
        }
        fseek(fp, (long)h.point_offset, SEEK_SET);
        size_t at = 0;
        for (int k = 0; at < L->bytes; k ^= 1) {
          WB_CHECK(hipEventSynchronize(done[k]));
          const size_t m = std::min(CH, L->bytes - at);
          if (fread(pin[k], 1, m, fp) != m) throw std::string("short read of point records");
          if (at == 0) L->ret0 = h.point_format < 6 ? (pin[k][14] & 7) : (pin[k][14] & 15);
          WB_CHECK(hipMemcpyAsync(L->rec + at, pin[k], m, hipMemcpyHostToDevice, c->stream));
          WB_CHECK(hipEventRecord(done[k], c->stream));
          at += m;
        }
        WB_CHECK(hipStreamSynchronize(c->stream));
      } catch (...) {
        hipStreamSynchronize(c->stream);
        for (int k = 0; k < 2; k++)
          if (pin[k]) hipHostFree(pin[k]);
        wb_free(c, L->rec, L->bytes ? L->bytes : 16);
        delete L;
        throw;
      }
      for (int k = 0; k < 2; k++) {
        hipHostFree(pin[k]);
        hipEventDestroy(done[k]);
      }
    } else {
      const uint8_t *r = mem + h.point_offset;
      L->ret0 = h.point_format < 6 ? (r[14] & 7) : (r[14] & 15);
      WB_CHECK(hipMemcpyAsync(L->rec, r, L->bytes, hipMemcpyHostToDevice, c->stream));
      WB_CHECK(hipStreamSynchronize(c->stream));
    }
  }
  if (hdr) *hdr = h;
  c->las.insert(L);
  *out = L;
  return WB_OK;
  API_CATCH(c)
}

int wb_las_open(wb_ctx *c, const char *path, wb_las **out, wb_las_header *hdr)
{
  if (!c || !path || !out) return fail(c, "null argument", WB_EINVAL);
  *out = nullptr;
  FILE *fp = fopen(path, "rb");
  if (!fp) return fail(c, std::string("cannot open ") + path, WB_EINVAL);
  const int rc = las_open(c, fp, nullptr, 0, out, hdr);
  fclose(fp);
  return rc;
}

int wb_las_open_mem(wb_ctx *c, const void *image, uint64_t len, wb_las **out, wb_las_header *hdr)
{
  if (!c || !image || !out) return fail(c, "null argument", WB_EINVAL);
  *out = nullptr;
  return las_open(c, nullptr, (const uint8_t *)image, len, out, hdr);
}

int wb_las_close(wb_las *L)
{
  if (!L) return WB_EINVAL;
  wb_ctx *c = L->ctx;
  hipStreamSynchronize(c->stream);
  wb_free(c, L->rec, L->bytes ? L->bytes : 16);
  c->las.erase(L);
  delete L;
  return WB_OK;
}

int wb_cloud_free(wb_cloud *cl);

// What wb_cloud_create holds while it works: the two timing events, the
// per-file position buffers of the zero rule and, until it is handed out,
// the cloud itself.  The destructor releases whatever is still held, on the
// refusal and exception paths as on success.
struct CloudCreateHold {
  wb_ctx *c;
  wb_cloud *cl = nullptr;
  std::vector<uint32_t *> pos;
  std::vector<uint64_t> pos_n;
  hipEvent_t ev[2] = {nullptr, nullptr};
  explicit CloudCreateHold(wb_ctx *ctx, int nfiles) : c(ctx), pos(nfiles, nullptr), pos_n(nfiles, 0) {}
  void drop_pos(int f)
  {
    wb_free(c, pos[f], 4 * pos_n[f]);
    pos[f] = nullptr;
  }
  ~CloudCreateHold()
  {
    // kernels queued on the stream may still read these buffers
    hipStreamSynchronize(c->stream);
    for (int k = 0; k < 2; k++)
      if (ev[k]) hipEventDestroy(ev[k]);
    for (size_t f = 0; f < pos.size(); f++)
      if (pos[f]) drop_pos((int)f);
    if (cl) {
      wb_free(c, cl->x, 8 * cl->n + 16);
      wb_free(c, cl->y, 8 * cl->n + 16);
      wb_free(c, cl->z, 8 * cl->n + 16);
      wb_free(c, cl->rec, 4 * cl->n + 16);
      wb_free(c, cl->d_src, sizeof(LasSrc) * cl->files.size());
      delete cl;
    }
  }
};

int wb_cloud_create(wb_ctx *c, wb_las *const *files, int nfiles, double unit, int zeros,
                    wb_cloud **out)
{
  if (!c || !files || nfiles <= 0 || !out) return fail(c, "null argument", WB_EINVAL);
  if (!(unit > 0) || !std::isfinite(unit)) return fail(c, "unit must be positive", WB_EINVAL);
  if (zeros != WB_ZEROS_READ && zeros != WB_ZEROS_KEEP) return fail(c, "unknown zeros rule", WB_EINVAL);
  *out = nullptr;
  for (int f = 0; f < nfiles; f++)
    if (!files[f] || files[f]->ctx != c) return fail(c, "file of another context", WB_EINVAL);
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  CloudCreateHold hold(c, nfiles);
  wb_cloud *cl = hold.cl = new wb_cloud();
  cl->ctx = c;
  cl->unit = unit;
  cl->zeros = zeros;
  cl->files.assign(files, files + nfiles);
  // per file: records kept, and the READ phase's zero rule (threads.cpp:
  // 479-545): if record 0 has a return number, records with none are
  // dropped; otherwise they are read as first returns
  std::vector<uint32_t *> &pos = hold.pos;
  uint64_t total = 0;
  std::vector<int> formats;
  hipEvent_t *ev = hold.ev;
  for (int k = 0; k < 2; k++) WB_CHECK(hipEventCreate(&ev[k]));
  WB_CHECK(hipEventRecord(ev[0], c->stream));
  for (int f = 0; f < nfiles; f++) {
    const wb_las *L = files[f];
    LasSrc s;
    memset(&s, 0, sizeof s);
    s.rec = L->rec;
    s.len = L->h.point_length;
    s.format = L->h.point_format;
    for (int a = 0; a < 3; a++) {
      s.scale[a] = L->h.scale[a];
      s.offset[a] = L->h.offset[a];
    }
    s.begin = total;
    const bool drop = zeros == WB_ZEROS_READ && L->ret0 > 0;
    s.fix = zeros == WB_ZEROS_READ && !drop;
    uint64_t kept = L->n;
    if (drop && L->n > 0) {
      pos[f] = (uint32_t *)wb_alloc(c, 4 * L->n);
      hold.pos_n[f] = L->n;
      hipLaunchKernelGGL(k_las_keep, dim3(wb_blocks(L->n, 256)), dim3(256), 0, c->stream, L->rec,
                         L->n, s.len, s.format, pos[f]);
      kept = wb_exclusive_scan_u32(c, pos[f], L->n);
      if (kept == L->n) hold.drop_pos(f);
    }
    cl->src.push_back(s);
    total += kept;
    formats.push_back(L->h.point_format);
    const double mn[3] = {L->h.min[0] * unit, L->h.min[1] * unit, L->h.min[2] * unit};
    const double mx[3] = {L->h.max[0] * unit, L->h.max[1] * unit, L->h.max[2] * unit};
    cl->corners.insert(cl->corners.end(), mn, mn + 3);
    cl->corners.insert(cl->corners.end(), mx, mx + 3);
  }
  if (total > 0xffffffffull) return fail(c, "more than 2^32-1 points", WB_ERANGE);
  cl->n = total;
  cl->x = (double *)wb_alloc(c, 8 * total + 16);
  cl->y = (double *)wb_alloc(c, 8 * total + 16);
  cl->z = (double *)wb_alloc(c, 8 * total + 16);
  cl->rec = (uint32_t *)wb_alloc(c, 4 * total + 16);
  for (int f = 0; f < nfiles; f++) {
    const wb_las *L = files[f];
    if (L->n > 0)
      hipLaunchKernelGGL(k_las_decode, dim3(wb_blocks(L->n, 256)), dim3(256), 0, c->stream,
                         cl->src[f], L->n, unit, pos[f], cl->x, cl->y, cl->z, cl->rec);
  }
  WB_CHECK(hipEventRecord(ev[1], c->stream));
  cl->d_src = (LasSrc *)wb_alloc(c, sizeof(LasSrc) * nfiles);
  WB_CHECK(hipMemcpyAsync(cl->d_src, cl->src.data(), sizeof(LasSrc) * nfiles,
                          hipMemcpyHostToDevice, c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  float ems = 0;
  hipEventElapsedTime(&ems, ev[0], ev[1]);
  cl->ms_decode = ems;
  // the lattice spacing of the decoded coordinates: combineScales of the
  // headers' scale and offset times the unit (LasHeader::getScale/getOffset)
  for (int a = 0; a < 3; a++) {
    std::vector<double> sc, of;
    for (int f = 0; f < nfiles; f++) {
      sc.push_back(files[f]->h.scale[a] * unit);
      of.push_back(files[f]->h.offset[a] * unit);
    }
    cl->resolution[a] = combine1_scale(sc, of);
  }
  cl->join_format = join_point_format(formats);
  c->clouds.insert(cl);
  hold.cl = nullptr;  // handed out: freed by wb_cloud_free or with its context
  *out = cl;
  return WB_OK;
  API_CATCH(c)
}

int wb_cloud_decode_ms(wb_cloud *cl, double *ms)
{
  if (!cl || !ms) return WB_EINVAL;
  *ms = cl->ms_decode;
  return WB_OK;
}

int wb_cloud_info(wb_cloud *cl, uint64_t *npoints, double resolution[3], int *join_format,
                  double *corners)
{
  if (!cl) return WB_EINVAL;
  if (npoints) *npoints = cl->n;
  if (resolution)
    for (int a = 0; a < 3; a++) resolution[a] = cl->resolution[a];
  if (join_format) *join_format = cl->join_format;
  if (corners) memcpy(corners, cl->corners.data(), sizeof(double) * cl->corners.size());
  return WB_OK;
}

int wb_cloud_points(wb_cloud *cl, const double **x, const double **y, const double **z)
{
  if (!cl || !x || !y || !z) return WB_EINVAL;
  *x = cl->x;
  *y = cl->y;
  *z = cl->z;
  return WB_OK;
}

int wb_cloud_read_points(wb_cloud *cl, double *x, double *y, double *z, uint32_t *rec)
{
  if (!cl) return WB_EINVAL;
  wb_ctx *c = cl->ctx;
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  const size_t n = cl->n;
  if (n) {
    if (x) WB_CHECK(hipMemcpyAsync(x, cl->x, 8 * n, hipMemcpyDeviceToHost, c->stream));
    if (y) WB_CHECK(hipMemcpyAsync(y, cl->y, 8 * n, hipMemcpyDeviceToHost, c->stream));
    if (z) WB_CHECK(hipMemcpyAsync(z, cl->z, 8 * n, hipMemcpyDeviceToHost, c->stream));
    if (rec) WB_CHECK(hipMemcpyAsync(rec, cl->rec, 4 * n, hipMemcpyDeviceToHost, c->stream));
  }
  WB_CHECK(hipStreamSynchronize(c->stream));
  return WB_OK;
  API_CATCH(c)
}

int wb_cloud_build(wb_cloud *cl, const double root_center[3], double root_side, int flags,
                   wb_tree **out)
{
  if (!cl || !out) return WB_EINVAL;
  wb_ctx *c = cl->ctx;
  for (int a = 0; a < 3; a++)
    if (!(cl->resolution[a] > 0))
      return fail(c, "the input files' coordinate lattices have no common spacing", WB_ERANGE);
  return wb_tree_build_xyz(c, cl->x, cl->y, cl->z, cl->n, cl->resolution, root_center, root_side,
                           WB_DEVICE_INPUT | WB_BUILD_DEDUPE | (flags & ~WB_DEVICE_INPUT), out);
}

int wb_cloud_classify(wb_cloud *cl, wb_tree *t, const wb_classify_params *p, uint8_t *class_out,
                      int flags, wb_tile *tiles_out, uint64_t tiles_cap, wb_stats *stats)
{
  if (!cl || !t) return WB_EINVAL;
  return wb_classify_corners(t, cl->corners.data(), (int)(cl->corners.size() / 3), p, class_out,
                             flags, tiles_out, tiles_cap, stats);
}

int wb_cloud_free(wb_cloud *cl)
{
  if (!cl) return WB_EINVAL;
  wb_ctx *c = cl->ctx;
  hipStreamSynchronize(c->stream);
  wb_free(c, cl->x, 8 * cl->n + 16);
  wb_free(c, cl->y, 8 * cl->n + 16);
  wb_free(c, cl->z, 8 * cl->n + 16);
  wb_free(c, cl->rec, 4 * cl->n + 16);
  wb_free(c, cl->d_src, sizeof(LasSrc) * cl->files.size());
  c->clouds.erase(cl);
  delete cl;
  return WB_OK;
}

int wb_cloud_write(wb_cloud *cl, wb_tree *t, const uint8_t *cls, int flags,
                   const wb_las_write_params *p, wb_las_write_stats *st, char *names,
                   uint64_t names_cap)
{
  if (!cl || !t || !p || !p->name) return WB_EINVAL;
  wb_ctx *c = cl->ctx;
  if (t->ctx != c || t->n_input != cl->n) return fail(c, "tree not built from this cloud", WB_EINVAL);
  const int fo = p->point_format < 0 ? cl->join_format : p->point_format;
  if (fo > 10) return fail(c, "point format above 10", WB_ERANGE);
  if (p->points_per_file < 0) return fail(c, "negative points per file", WB_EINVAL);
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  auto now = [] { return std::chrono::steady_clock::now(); };
  auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  wb_las_write_stats S;
  memset(&S, 0, sizeof S);
  const uint64_t n = t->n;
  const uint32_t nl = t->nleaves;
  const int nfin = (int)cl->files.size();
  // class bytes on the device
  const uint8_t *dcls = nullptr;
  uint8_t *dcls_own = nullptr;
  if (cls && !(flags & WB_DEVICE_INPUT)) {
    dcls_own = (uint8_t *)wb_alloc(c, cl->n + 16);
    WB_CHECK(hipMemcpyAsync(dcls_own, cls, cl->n, hipMemcpyHostToDevice, c->stream));
    dcls = dcls_own;
  } else {
    dcls = cls;
  }
  hipEvent_t ev[4];
  for (int k = 0; k < 4; k++) WB_CHECK(hipEventCreate(&ev[k]));
  // 1. class of every stored point, class totals (countClasses)
  uint8_t *ocls = (uint8_t *)wb_alloc(c, n + 16);
  unsigned long long *dtot = (unsigned long long *)wb_alloc(c, 256 * 8 * WB_NCOPY);
  WB_CHECK(hipMemsetAsync(dtot, 0, 256 * 8 * WB_NCOPY, c->stream));
  WB_CHECK(hipEventRecord(ev[0], c->stream));
  if (n)
    hipLaunchKernelGGL(k_out_class, dim3(wb_blocks(n, 256)), dim3(256), 0, c->stream, n, t->canon,
                       t->sidx, t->last_of, dcls, cl->d_src, nfin, cl->rec, cl->unit, ocls, dtot);
  WB_CHECK(hipEventRecord(ev[1], c->stream));
  std::vector<unsigned long long> totc(256 * WB_NCOPY);
  WB_CHECK(hipMemcpyAsync(totc.data(), dtot, 8 * totc.size(), hipMemcpyDeviceToHost, c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  unsigned long long tot[256] = {};
  for (int q = 0; q < WB_NCOPY; q++)
    for (int k = 0; k < 256; k++) tot[k] += totc[256 * q + k];
  auto t_plan = now();
  for (int k = 0; k < 256; k++) S.class_totals[k] = tot[k];
  // 2. the file plan (CloudOutput::openFiles, cloudoutput.cpp:119-178)
  const bool sep = p->separate_classes != 0;
  const uint64_t ppf = (uint64_t)p->points_per_file;
  uint64_t grand = 0;
  for (int k = 0; k < 256; k++) grand += tot[k];
  int ndig = 0;
  if (ppf) {
    uint64_t q = (grand + ppf - 1) / ppf;
    if (q) q--;
    if (!q) q++;
    while (q) {
      q /= 10;
      ndig++;
    }
  }
  std::vector<int> present;  // the keys of the header map, ascending
  int16_t cidh[256];
  for (int k = 0; k < 256; k++) cidh[k] = 0;
  if (sep) {
    for (int k = 0; k < 256; k++)
      if (tot[k]) {
        cidh[k] = (int16_t)present.size();
        present.push_back(k);
      }
  } else {
    present.push_back(0);  // every point goes to the files of key 0
  }
  const int K = (int)present.size();
  std::vector<int> nfiles(K), first_file(K);
  int nout = 0;
  for (int id = 0; id < K; id++) {
    const uint64_t kt = sep ? tot[present[id]] : grand;
    nfiles[id] = ppf ? (int)((kt + ppf - 1) / ppf) : 1;
    first_file[id] = nout;
    nout += nfiles[id];
  }
  // per leaf and class id, the number of points
  int16_t *dcid = (int16_t *)wb_alloc(c, sizeof cidh);
  WB_CHECK(hipMemcpyAsync(dcid, cidh, sizeof cidh, hipMemcpyHostToDevice, c->stream));
  const uint64_t LK = (uint64_t)nl * K;
  uint32_t *dcnt = (uint32_t *)wb_alloc(c, 4 * LK + 16);
  if (nl)
    hipLaunchKernelGGL(k_leaf_class_count, dim3(nl), dim3(64), 0, c->stream, t->leaf_start, nl,
                       ocls, dcid, K, dcnt);
  std::vector<uint32_t> cnt(LK);
  if (LK) WB_CHECK(hipMemcpyAsync(cnt.data(), dcnt, 4 * LK, hipMemcpyDeviceToHost, c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  // CloudOutput::writeFiles (cloudoutput.cpp:180-213): at each block, each
  // class picks its least-filled file (the first of equals), then the
  // block's points of that class go there
  std::vector<uint64_t> base(LK);
  std::vector<int32_t> fileof(LK);
  std::vector<std::vector<uint64_t>> fill(K);
  std::vector<int> next(K, 0);
  for (int id = 0; id < K; id++) fill[id].assign(nfiles[id], 0);
  for (uint32_t l = 0; l < nl; l++) {
    for (int id = 0; id < K; id++) {
      uint64_t mn = grand;
      for (int j = 0; j < nfiles[id]; j++)
        if (fill[id][j] < mn) {
          next[id] = j;
          mn = fill[id][j];
        }
    }
    for (int id = 0; id < K; id++) {
      const uint64_t q = (uint64_t)l * K + id;
      base[q] = fill[id][next[id]];
      fileof[q] = first_file[id] + next[id];
      fill[id][next[id]] += cnt[q];
    }
  }
  // file images: each file's records start on a 256-byte boundary
  const int LEN = kPointLength[fo];
  std::vector<uint64_t> npts(nout, 0), data_off(nout + 1, 0);
  for (int id = 0; id < K; id++)
    for (int j = 0; j < nfiles[id]; j++) npts[first_file[id] + j] = fill[id][j];
  for (int g = 0; g < nout; g++) data_off[g + 1] = data_off[g] + (npts[g] * LEN + 255) / 256 * 256;
  // output transform (CloudOutput::openFiles with setUnit, setScale)
  double br[6];
  wb::bound_rect(cl->corners.data(), (int)(cl->corners.size() / 3), br);
  const double minc[3] = {br[0], br[1], br[4]}, maxc[3] = {br[2], br[3], br[5]};
  OutXf o;
  o.unit = cl->unit;
  for (int a = 0; a < 3; a++) {
    std::vector<double> sc, of;
    for (int f = 0; f < nfin; f++) {
      sc.push_back(cl->files[f]->h.scale[a] * cl->unit);
      of.push_back(cl->files[f]->h.offset[a] * cl->unit);
    }
    const double comb = combine1_scale(sc, of);
    set_scale(minc[a], maxc[a], comb, cl->unit, &o.scale[a], &o.offset[a]);
    S.scale[a] = o.scale[a];
    S.offset[a] = o.offset[a];
  }
  uint64_t *dbase = (uint64_t *)wb_alloc(c, 8 * LK + 16);
  int32_t *dfile = (int32_t *)wb_alloc(c, 4 * LK + 16);
  uint64_t *doff = (uint64_t *)wb_alloc(c, 8 * (nout + 1));
  const size_t nfs = (size_t)(nout + 1) * WB_NCOPY;
  FileStats *dfs = (FileStats *)wb_alloc(c, sizeof(FileStats) * nfs);
  std::vector<FileStats> fs(nfs);
  for (auto &f : fs) {
    memset(&f, 0, sizeof f);
    for (int a = 0; a < 3; a++) f.mn[a] = ~0ull;
  }
  const uint64_t img_bytes = data_off[nout];
  uint8_t *dimg = (uint8_t *)wb_alloc(c, img_bytes + 16);
  if (LK) {
    WB_CHECK(hipMemcpyAsync(dbase, base.data(), 8 * LK, hipMemcpyHostToDevice, c->stream));
    WB_CHECK(hipMemcpyAsync(dfile, fileof.data(), 4 * LK, hipMemcpyHostToDevice, c->stream));
  }
  WB_CHECK(hipMemcpyAsync(doff, data_off.data(), 8 * (nout + 1), hipMemcpyHostToDevice, c->stream));
  WB_CHECK(hipMemcpyAsync(dfs, fs.data(), sizeof(FileStats) * nfs, hipMemcpyHostToDevice,
                          c->stream));
  S.ms_plan = ms(t_plan, now());
  // 3. encode every stored record into its file image
  WB_CHECK(hipEventRecord(ev[2], c->stream));
  if (nl) {
#define WB_ENC(F)                                                                                 \
  case F:                                                                                         \
    hipLaunchKernelGGL(k_las_encode<F>, dim3(nl), dim3(64), 0, c->stream, t->leaf_start, nl,      \
                       t->canon, t->sidx, t->last_of, dcls, ocls, cl->d_src, nfin, cl->rec,      \
                       cl->unit, dcid, K, dbase, dfile, doff, o, dimg, dfs);                       \
    break;
    switch (fo) {
      WB_ENC(0) WB_ENC(1) WB_ENC(2) WB_ENC(3) WB_ENC(4) WB_ENC(5) WB_ENC(6) WB_ENC(7) WB_ENC(8)
      WB_ENC(9) WB_ENC(10)
    }
#undef WB_ENC
  }
  WB_CHECK(hipEventRecord(ev[3], c->stream));
  WB_CHECK(hipMemcpyAsync(fs.data(), dfs, sizeof(FileStats) * nfs, hipMemcpyDeviceToHost,
                          c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  // fold the copies into copy 0 of each file
  for (int g = 0; g < nout; g++) {
    FileStats &f0 = fs[(size_t)g * WB_NCOPY];
    for (int q = 1; q < WB_NCOPY; q++) {
      const FileStats &fq = fs[(size_t)g * WB_NCOPY + q];
      for (int i = 0; i < 16; i++) f0.npoints[i] += fq.npoints[i];
      for (int a = 0; a < 3; a++) {
        f0.mn[a] = std::min(f0.mn[a], fq.mn[a]);
        f0.mx[a] = std::max(f0.mx[a], fq.mx[a]);
      }
    }
  }
  float e0 = 0, e1 = 0;
  hipEventElapsedTime(&e0, ev[0], ev[1]);
  hipEventElapsedTime(&e1, ev[2], ev[3]);
  S.ms_classes = e0;
  S.ms_encode = e1;
  // 4. headers, copies and files (CloudOutput::closeFiles with writeHeader)
  time_t tnow = time(nullptr);
  struct tm tmv;
  gmtime_r(&tnow, &tmv);
  const int day = p->creation_day > 0 ? p->creation_day : tmv.tm_yday + 1;
  const int year = p->creation_year > 0 ? p->creation_year : tmv.tm_year + 1900;
  const char *sysid = sep ? "EXTRACTION" : nfin > 1 ? "MERGE" : "MODIFICATION";
  std::string allnames;
  uint8_t *pin = nullptr;
  const bool dry = (flags & WB_WRITE_DRY) != 0;
  size_t pin_bytes = 0;
  for (int g = 0; g < nout; g++) pin_bytes = std::max<size_t>(pin_bytes, npts[g] * LEN);
  if (pin_bytes && !dry) WB_CHECK(hipHostMalloc((void **)&pin, pin_bytes, hipHostMallocDefault));
  for (int id = 0; id < K; id++)
    for (int j = 0; j < nfiles[id]; j++) {
      const int g = first_file[id] + j;
      std::string fname = p->name;
      if (sep) fname += "-" + class_name(present[id]);
      fname += (ppf ? "-" : "") + ndecimal(j, ndig) + ".las";
      allnames += fname + "\n";
      uint8_t hd[WB_LAS_HEADER];
      memset(hd, 0, sizeof hd);
      memcpy(hd, "LASF", 4);
      hd[24] = 1;
      hd[25] = 4;
      strncpy((char *)hd + 26, sysid, 32);
      strncpy((char *)hd + 58, WB_LAS_SOFTWARE, 32);
      put_le(hd, 90, (uint16_t)day, 2);
      put_le(hd, 92, (uint16_t)year, 2);
      put_le(hd, 94, WB_LAS_HEADER, 2);
      put_le(hd, 96, WB_LAS_HEADER, 4);
      hd[104] = (uint8_t)fo;
      put_le(hd, 105, (uint16_t)LEN, 2);
      const FileStats &f = fs[(size_t)g * WB_NCOPY];
      bool legacy = true;
      for (int i = 0; i < 6; i++)
        if (f.npoints[i] > 4294967295ull) legacy = false;
      for (int i = 6; i < 16; i++)
        if (f.npoints[i] > 0) legacy = false;
      for (int i = 0; i < 6; i++) put_le(hd, 107 + 4 * i, legacy ? (uint32_t)f.npoints[i] : 0, 4);
      for (int a = 0; a < 3; a++) {
        put_d(hd, 131 + 8 * a, o.scale[a]);
        put_d(hd, 155 + 8 * a, o.offset[a]);
        put_d(hd, 179 + 16 * a, f.npoints[0] ? from_key(f.mx[a]) : -INFINITY);
        put_d(hd, 187 + 16 * a, f.npoints[0] ? from_key(f.mn[a]) : INFINITY);
      }
      put_le(hd, 235, WB_LAS_HEADER + f.npoints[0] * (uint64_t)LEN, 8);
      for (int i = 0; i < 16; i++) put_le(hd, 247 + 8 * i, f.npoints[i], 8);
      const size_t nb = (size_t)(npts[g] * LEN);
      auto tc = now();
      if (nb && !dry) WB_CHECK(hipMemcpy(pin, dimg + data_off[g], nb, hipMemcpyDeviceToHost));
      auto td = now();
      S.ms_copy += ms(tc, td);
      if (!dry) {
        FILE *fp = fopen(fname.c_str(), "wb");
        if (!fp) throw std::string("cannot create ") + fname;
        const bool ok = fwrite(hd, 1, sizeof hd, fp) == sizeof hd && fwrite(pin, 1, nb, fp) == nb;
        if (fclose(fp) != 0 || !ok) throw std::string("write failed: ") + fname;
      }
      S.ms_disk += ms(td, now());
      S.npoints += npts[g];
      S.bytes += sizeof hd + nb;
    }
  S.nfiles = (uint32_t)nout;
  S.point_format = fo;
  if (pin) hipHostFree(pin);
  wb_free(c, ocls, n + 16);
  wb_free(c, dtot, 256 * 8 * WB_NCOPY);
  wb_free(c, dcid, sizeof cidh);
  wb_free(c, dcnt, 4 * LK + 16);
  wb_free(c, dbase, 8 * LK + 16);
  wb_free(c, dfile, 4 * LK + 16);
  wb_free(c, doff, 8 * (nout + 1));
  wb_free(c, dfs, sizeof(FileStats) * nfs);
  wb_free(c, dimg, img_bytes + 16);
  if (dcls_own) wb_free(c, dcls_own, cl->n + 16);
  for (int k = 0; k < 4; k++) hipEventDestroy(ev[k]);
  if (st) *st = S;
  if (names && names_cap) {
    const size_t m = std::min<size_t>(allnames.size(), names_cap - 1);
    memcpy(names, allnames.data(), m);
    names[m] = 0;
  }
  return WB_OK;
  API_CATCH(c)
}

}  // extern "C"

void wb_release_las(wb_ctx *c)
{
  // their device buffers go back to the context's pool, which is freed next
  for (wb_cloud *cl : std::vector<wb_cloud *>(c->clouds.begin(), c->clouds.end()))
    wb_cloud_free(cl);
  for (wb_las *L : std::vector<wb_las *>(c->las.begin(), c->las.end())) wb_las_close(L);
}

// the store blocks of tree t into device memory dst (nl blocks)
static void store_blocks_device(wb_tree *t, wb_cloud *cl, const uint8_t *cls, int flags,
                                uint8_t *dst)
{
  wb_ctx *c = t->ctx;
  const uint8_t *dcls = nullptr;
  uint8_t *dcls_own = nullptr;
  const uint64_t nin = t->n_input;
  if (cls && !(flags & WB_DEVICE_INPUT)) {
    dcls_own = (uint8_t *)wb_alloc(c, nin + 16);
    WB_CHECK(hipMemcpyAsync(dcls_own, cls, nin, hipMemcpyHostToDevice, c->stream));
    dcls = dcls_own;
  } else {
    dcls = cls;
  }
  if (t->nleaves)
    hipLaunchKernelGGL(k_store_blocks, dim3(t->nleaves), dim3(64), 0, c->stream, t->leaf_start,
                       t->nleaves, t->canon, t->sidx, t->last_of, dcls, cl ? cl->d_src : nullptr,
                       cl ? (int)cl->files.size() : 0, cl ? cl->rec : nullptr,
                       cl ? cl->unit : 1.0, t->px, t->py, t->pz, dst);
  WB_CHECK(hipGetLastError());
  WB_CHECK(hipStreamSynchronize(c->stream));
  if (dcls_own) wb_free(c, dcls_own, nin + 16);
}

static int store_check(wb_tree *t, wb_cloud *cl)
{
  if (!t) return WB_EINVAL;
  if (cl && (cl->ctx != t->ctx || cl->n != t->n_input))
    return fail(t->ctx, "tree not built from this cloud", WB_EINVAL);
  return WB_OK;
}

extern "C" {

int wb_store_blocks(wb_tree *t, wb_cloud *cl, const uint8_t *cls, int flags, void *out,
                    uint64_t cap, uint64_t *nblocks)
{
  const int rc = store_check(t, cl);
  if (rc) return rc;
  wb_ctx *c = t->ctx;
  const uint64_t nl = t->nleaves, bytes = nl * WB_STORE_BLOCK;
  if (nblocks) *nblocks = nl;
  if (!out) return WB_OK;
  if (cap < bytes) return fail(c, "output smaller than the blocks", WB_ERANGE);
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  if (nl) {
    uint8_t *d = (uint8_t *)wb_alloc(c, bytes);
    store_blocks_device(t, cl, cls, flags, d);
    WB_CHECK(hipMemcpyAsync(out, d, bytes, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipStreamSynchronize(c->stream));
    wb_free(c, d, bytes);
  }
  return WB_OK;
  API_CATCH(c)
}

int wb_store_write_files(wb_tree *t, wb_cloud *cl, const uint8_t *cls, int flags,
                         const char *name, int nfiles)
{
  const int rc = store_check(t, cl);
  if (rc) return rc;
  wb_ctx *c = t->ctx;
  if (!name) return WB_EINVAL;
  if (nfiles < 1) nfiles = 1;  // OctStore::open
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  const uint64_t nl = t->nleaves, bytes = nl * WB_STORE_BLOCK;
  std::vector<uint8_t> h(bytes ? bytes : 1);
  if (nl) {
    uint8_t *d = (uint8_t *)wb_alloc(c, bytes);
    store_blocks_device(t, cl, cls, flags, d);
    WB_CHECK(hipMemcpyAsync(h.data(), d, bytes, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipStreamSynchronize(c->stream));
    wb_free(c, d, bytes);
  }
  // block b in file name<b % nfiles> at BLOCKSIZE * (b / nfiles)
  for (int f = 0; f < nfiles; f++) {
    const std::string fn = std::string(name) + std::to_string(f);
    FILE *fp = fopen(fn.c_str(), "wb");
    if (!fp) return fail(c, "cannot create " + fn, WB_EINVAL);
    bool ok = true;
    for (uint64_t bk = f; bk < nl && ok; bk += nfiles)
      ok = fwrite(h.data() + bk * WB_STORE_BLOCK, 1, WB_STORE_BLOCK, fp) == WB_STORE_BLOCK;
    if (fclose(fp) != 0 || !ok) return fail(c, "write failed: " + fn, WB_EINVAL);
  }
  return WB_OK;
  API_CATCH(c)
}

int wb_store_read_blocks(wb_ctx *c, const void *blocks, uint64_t nblocks, double *x, double *y,
                         double *z, uint16_t *classification, uint32_t *block_count,
                         uint64_t *npoints)
{
  if (!c || (nblocks && !blocks) || !npoints) return WB_EINVAL;
  if (nblocks > 0xffffffffull / WB_STORE_RECORDS) return fail(c, "too many blocks", WB_ERANGE);
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  *npoints = 0;
  if (!nblocks) return WB_OK;
  const uint64_t bytes = nblocks * WB_STORE_BLOCK;
  uint8_t *d = (uint8_t *)wb_alloc(c, bytes);
  uint32_t *cnt = (uint32_t *)wb_alloc(c, 4 * nblocks), *pos = (uint32_t *)wb_alloc(c, 4 * nblocks);
  WB_CHECK(hipMemcpyAsync(d, blocks, bytes, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_store_count, dim3((unsigned)nblocks), dim3(64), 0, c->stream, d, nblocks,
                     cnt);
  WB_CHECK(hipMemcpyAsync(pos, cnt, 4 * nblocks, hipMemcpyDeviceToDevice, c->stream));
  const uint64_t n = wb_exclusive_scan_u32(c, pos, nblocks);
  *npoints = n;
  if (block_count)
    WB_CHECK(hipMemcpyAsync(block_count, cnt, 4 * nblocks, hipMemcpyDeviceToHost, c->stream));
  if (n && (x || y || z || classification)) {
    double *dx = (double *)wb_alloc(c, 24 * n);
    uint16_t *dc = (uint16_t *)wb_alloc(c, 2 * n + 16);
    hipLaunchKernelGGL(k_store_points, dim3((unsigned)nblocks), dim3(64), 0, c->stream, d, nblocks,
                       cnt, pos, dx, dx + n, dx + 2 * n, dc);
    if (x) WB_CHECK(hipMemcpyAsync(x, dx, 8 * n, hipMemcpyDeviceToHost, c->stream));
    if (y) WB_CHECK(hipMemcpyAsync(y, dx + n, 8 * n, hipMemcpyDeviceToHost, c->stream));
    if (z) WB_CHECK(hipMemcpyAsync(z, dx + 2 * n, 8 * n, hipMemcpyDeviceToHost, c->stream));
    if (classification)
      WB_CHECK(hipMemcpyAsync(classification, dc, 2 * n, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipStreamSynchronize(c->stream));
    wb_free(c, dx, 24 * n);
    wb_free(c, dc, 2 * n + 16);
  }
  WB_CHECK(hipStreamSynchronize(c->stream));
  wb_free(c, d, bytes);
  wb_free(c, cnt, 4 * nblocks);
  wb_free(c, pos, 4 * nblocks);
  return WB_OK;
  API_CATCH(c)
}

}  // extern "C"

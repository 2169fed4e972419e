// wb_build.hip -- octree build on the GPU: Morton keys from LAS int32
// coordinates, LSD radix sort, canonical leaf split, SoA permute.
//
// Reference behaviour (octree.cpp:199-347, 620-662, 849-876, 1295-1338): a
// block holds at most RECORDS=537 distinct locations and splits one level
// when another arrives, so the final tree is the unique one in which a node
// is internal iff its region holds more than 537 points.  A point's region
// at each level is decided by `coordinate >= centre` compares (child index
// z*4+y*2+x), and the DFS visits children 0..7, so DFS order is ascending
// order of the Morton key built from those compare bits.
#include "wb_internal.h"
#include "wb_wave.h"

#include <string.h>

#include <algorithm>

#define WAVE 64

// ---------------------------------------------------------------------------
// wave / block primitives
// ---------------------------------------------------------------------------
__device__ inline unsigned lane_id() { return __lane_id(); }
__device__ inline uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

// inclusive scan across a 64-lane wave
__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    uint32_t t = __shfl_up(v, d, WAVE);
    if ((int)lane_id() >= d) v += t;
  }
  return v;
}

// exclusive scan of one value per thread over a 256-thread block; returns
// the exclusive prefix and writes the block total to *total
__device__ inline uint32_t block_excl_scan256(uint32_t v, uint32_t *total)
{
  __shared__ uint32_t wsum[4];
  uint32_t inc = wave_incl_scan(v);
  int w = threadIdx.x / WAVE;
  if (lane_id() == WAVE - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int i = 0; i < 4; i++) {
    if (i < w) off += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  *total = tot;
  return off + inc - v;
}

// ---------------------------------------------------------------------------
// device-wide exclusive scan (u32), 1024 elements per block
// ---------------------------------------------------------------------------
__global__ void k_scan_reduce(const uint32_t *in, uint64_t n, uint32_t *bsum)
{
  uint64_t base = (uint64_t)blockIdx.x * 1024;
  uint32_t s = 0;
  for (int k = 0; k < 4; k++) {
    uint64_t i = base + threadIdx.x * 4 + k;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_excl_scan256(s, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ void k_scan_down(uint32_t *data, uint64_t n, const uint32_t *boff)
{
  uint64_t base = (uint64_t)blockIdx.x * 1024;
  uint32_t v[4], s = 0;
  for (int k = 0; k < 4; k++) {
    uint64_t i = base + threadIdx.x * 4 + k;
    v[k] = (i < n) ? data[i] : 0;
    s += v[k];
  }
  uint32_t tot;
  uint32_t ex = block_excl_scan256(s, &tot) + boff[blockIdx.x];
  for (int k = 0; k < 4; k++) {
    uint64_t i = base + threadIdx.x * 4 + k;
    if (i < n) data[i] = ex;
    ex += v[k];
  }
}

// exclusive scan of n <= SCAN1_MAX values by one block of 1024 threads, each
// taking a run of consecutive values (the block sums of a larger scan)
#define SCAN1_MAX 65536
__global__ void __launch_bounds__(1024) k_scan_one(uint32_t *data, uint32_t n)
{
  __shared__ uint32_t wsum[16];
  const uint32_t per = (n + 1023) / 1024, b = threadIdx.x * per;
  uint32_t s = 0;
  for (uint32_t k = 0; k < per; k++)
    if (b + k < n) s += data[b + k];
  uint32_t inc = s;  // inclusive scan within the wave
  for (int d = 1; d < WAVE; d <<= 1) {
    const uint32_t o = __shfl_up(inc, d, WAVE);
    if ((int)lane_id() >= d) inc += o;
  }
  const int w = threadIdx.x / WAVE;
  if ((int)lane_id() == WAVE - 1) wsum[w] = inc;
  __syncthreads();
  uint32_t off = 0;
  for (int k = 0; k < w; k++) off += wsum[k];
  uint32_t ex = off + inc - s;
  for (uint32_t k = 0; k < per; k++)
    if (b + k < n) {
      const uint32_t v = data[b + k];
      data[b + k] = ex;
      ex += v;
    }
}

static void scan_rec(wb_ctx *c, uint32_t *data, uint64_t n)
{
  if (n <= SCAN1_MAX) {
    hipLaunchKernelGGL(k_scan_one, dim3(1), dim3(1024), 0, c->stream, data, (uint32_t)n);
    return;
  }
  unsigned nb = wb_blocks(n, 1024);
  uint32_t *bs = (uint32_t *)wb_alloc(c, sizeof(uint32_t) * (nb + 1));
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, c->stream, data, n, bs);
  scan_rec(c, bs, nb);
  hipLaunchKernelGGL(k_scan_down, dim3(nb), dim3(256), 0, c->stream, data, n, bs);
  wb_free(c, bs, sizeof(uint32_t) * (nb + 1));
}

// total = the last element before the scan + the last after it, written by
// the device into the context's pinned words
__global__ void k_scan_total(const uint32_t *last_in, const uint32_t *last_out, uint32_t *host)
{
  if (threadIdx.x == 0) {
    const uint64_t t = (uint64_t)*last_in + *last_out;
    host[0] = (uint32_t)t;
    host[1] = (uint32_t)(t >> 32);
  }
}

uint64_t wb_exclusive_scan_u32(wb_ctx *c, uint32_t *data, uint64_t n)
{
  if (n == 0) return 0;
  uint32_t *keep = (uint32_t *)wb_alloc(c, 4);
  WB_CHECK(hipMemcpyAsync(keep, data + n - 1, 4, hipMemcpyDeviceToDevice, c->stream));
  scan_rec(c, data, n);
  uint32_t *h = c->hbox + WB_HBOX_SCAN;
  hipLaunchKernelGGL(k_scan_total, dim3(1), dim3(64), 0, c->stream, keep, data + n - 1, h);
  WB_CHECK(hipStreamSynchronize(c->stream));
  wb_free(c, keep, 4);
  return (uint64_t)h[0] | ((uint64_t)h[1] << 32);
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits.  Per pass: per-block digit histograms in LDS,
// a device-wide exclusive scan of the digit-major histogram table, then a
// stable scatter in which each wave ranks its keys with 8 ballots (peer
// mask of equal digits) and one LDS atomic per distinct digit per round.
// The block's keys are then placed in LDS in digit order, and consecutive
// threads write each digit's run to its consecutive output slots, so every
// store instruction writes whole runs rather than 64 scattered buckets.
// ---------------------------------------------------------------------------
#define RS_THREADS 256
#define RS_ITEMS 8
#define RS_TILE (RS_THREADS * RS_ITEMS)

template <typename K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const K *keys, uint64_t n, int shift,
                                                          uint32_t *hist, unsigned nblocks)
{
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t base = (uint64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_ITEMS; k++) {
    uint64_t i = base + (uint64_t)k * RS_THREADS + threadIdx.x;
    if (i < n) atomicAdd(&h[(unsigned)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[(uint64_t)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

template <typename K>
__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const K *kin, const uint32_t *vin,
                                                             K *kout, uint32_t *vout, uint64_t n,
                                                             int shift, const uint32_t *offs,
                                                             unsigned nblocks)
{
  __shared__ uint32_t wcnt[4][256];
  __shared__ uint32_t dstart[256];
  __shared__ uint32_t dbase[256];  // the block's output start of each digit
  __shared__ K skey[RS_TILE];
  __shared__ uint32_t sval[RS_TILE];
  const int w = threadIdx.x / WAVE;
  const unsigned lane = lane_id();
  for (int d = threadIdx.x; d < 4 * 256; d += RS_THREADS) (&wcnt[0][0])[d] = 0;
  for (int d = threadIdx.x; d < 256; d += RS_THREADS) dbase[d] = offs[(uint64_t)d * nblocks + blockIdx.x];
  __syncthreads();
  const uint64_t base = (uint64_t)blockIdx.x * RS_TILE + (uint64_t)w * (RS_TILE / 4);
  K key[RS_ITEMS];
  uint32_t val[RS_ITEMS];
  uint32_t rank[RS_ITEMS];
  uint32_t dig[RS_ITEMS];
  for (int r = 0; r < RS_ITEMS; r++) {
    uint64_t i = base + (uint64_t)r * WAVE + lane;
    bool ok = i < n;
    key[r] = ok ? kin[i] : (K)0;
    val[r] = ok ? vin[i] : 0u;
    uint32_t d = (unsigned)(key[r] >> shift) & 255u;
    dig[r] = d;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t before = (uint32_t)__popcll(peers & lanemask_lt());
    uint32_t cnt = (uint32_t)__popcll(peers);
    int leader = peers ? (__ffsll((long long)peers) - 1) : 0;
    uint32_t old = 0;
    if (ok && (int)lane == leader) old = atomicAdd(&wcnt[w][d], cnt);
    old = __shfl(old, leader, WAVE);
    rank[r] = old + before;
  }
  __syncthreads();
  // per-digit prefix over the four waves, and the block's digit starts (an
  // exclusive scan of the digit totals across the 256 threads)
  {
    const int d = threadIdx.x;
    const uint32_t a0 = wcnt[0][d], a1 = wcnt[1][d], a2 = wcnt[2][d], a3 = wcnt[3][d];
    const uint32_t tot = a0 + a1 + a2 + a3;
    uint32_t inc = tot;  // inclusive scan within the wave
    for (int s = 1; s < WAVE; s <<= 1) {
      const uint32_t o = __shfl_up(inc, s, WAVE);
      if ((int)lane >= s) inc += o;
    }
    __shared__ uint32_t wsum[4];
    if ((int)lane == WAVE - 1) wsum[w] = inc;
    __syncthreads();
    uint32_t woff = 0;
    for (int k = 0; k < w; k++) woff += wsum[k];
    dstart[d] = woff + inc - tot;
    wcnt[0][d] = 0;
    wcnt[1][d] = a0;
    wcnt[2][d] = a0 + a1;
    wcnt[3][d] = a0 + a1 + a2;
  }
  __syncthreads();
  for (int r = 0; r < RS_ITEMS; r++) {
    uint64_t i = base + (uint64_t)r * WAVE + lane;
    if (i >= n) continue;
    const uint32_t d = dig[r];
    const uint32_t lpos = dstart[d] + wcnt[w][d] + rank[r];
    skey[lpos] = key[r];
    sval[lpos] = val[r];
  }
  __syncthreads();
  const uint64_t b0 = (uint64_t)blockIdx.x * RS_TILE;
  const uint32_t cnt = (uint32_t)(n - b0 < (uint64_t)RS_TILE ? n - b0 : RS_TILE);
  for (uint32_t j = threadIdx.x; j < cnt; j += RS_THREADS) {
    const K k = skey[j];
    const uint32_t d = (unsigned)(k >> shift) & 255u;
    const uint64_t pos = (uint64_t)dbase[d] + (j - dstart[d]);
    kout[pos] = k;
    vout[pos] = sval[j];
  }
}

template <typename K>
static void radix_sort_impl(wb_ctx *c, K *key, uint32_t *val, uint64_t n, int first_shift,
                            int passes)
{
  if (n <= 1 || passes <= 0) return;
  unsigned nb = wb_blocks(n, RS_TILE);
  uint64_t hsz = (uint64_t)nb * 256;
  uint32_t *hist = (uint32_t *)wb_alloc(c, sizeof(uint32_t) * hsz);
  K *k2 = (K *)wb_alloc(c, sizeof(K) * n);
  uint32_t *v2 = (uint32_t *)wb_alloc(c, sizeof(uint32_t) * n);
  K *ka = key, *kb = k2;
  uint32_t *va = val, *vb = v2;
  for (int p = 0; p < passes; p++) {
    int shift = first_shift + 8 * p;
    hipLaunchKernelGGL(k_rs_hist<K>, dim3(nb), dim3(RS_THREADS), 0, c->stream, ka, n, shift, hist,
                       nb);
    scan_rec(c, hist, hsz);
    hipLaunchKernelGGL(k_rs_scatter<K>, dim3(nb), dim3(RS_THREADS), 0, c->stream, ka, va, kb, vb,
                       n, shift, hist, nb);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
  if (ka != key) {
    WB_CHECK(hipMemcpyAsync(key, ka, sizeof(K) * n, hipMemcpyDeviceToDevice, c->stream));
    WB_CHECK(hipMemcpyAsync(val, va, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, c->stream));
  }
  wb_free(c, hist, sizeof(uint32_t) * hsz);
  wb_free(c, k2, sizeof(K) * n);
  wb_free(c, v2, sizeof(uint32_t) * n);
}

void wb_radix_sort_pairs(wb_ctx *c, uint64_t *key, uint32_t *val, uint64_t n, int passes)
{
  radix_sort_impl<uint64_t>(c, key, val, n, 64 - 8 * passes, passes);
}

void wb_radix_sort_u32_prefix(wb_ctx *c, uint32_t *key, uint32_t *val, uint64_t n, int passes)
{
  radix_sort_impl<uint32_t>(c, key, val, n, 32 - 8 * passes, passes);
}

void wb_radix_sort_u32_pairs(wb_ctx *c, uint32_t *key, uint32_t *val, uint64_t n)
{
  radix_sort_impl<uint32_t>(c, key, val, n, 0, 4);
}

// ---------------------------------------------------------------------------
// K1: Morton keys
// ---------------------------------------------------------------------------
struct Xform {
  double sx, sy, sz, ox, oy, oz;
  double lox, loy, loz;  // root cube minimum corner
  double cell;           // side / 2^21
  double inv_cell;       // 2^21 / side
};

// 21-bit cell index along one axis, equal to the bits of the reference's
// `x >= centre` descent: the level-l centre of the node with index j is
// lo + (2j+1)*side/2^(l+1), a dyadic value, so the descent equals counting
// the exact cell bounds lo + k*cell <= x.  Points outside the root cube go to
// the extreme children, as the compares send them.
__device__ inline uint32_t cell_index(double x, double lo, double cell, double inv_cell)
{
  const long long kmax = (1ll << WB_MAX_LEVEL) - 1;
  double t = (x - lo) * inv_cell;
  long long k;
  if (!(t > 0)) k = 0;                      // also NaN
  else if (t >= (double)kmax) k = kmax;
  else k = (long long)t;
  if (k > 0 && x < lo + (double)k * cell) k--;
  else if (k < kmax && x >= lo + (double)(k + 1) * cell) k++;
  return (uint32_t)k;
}

__device__ inline uint64_t spread3(uint32_t v)
{
  uint64_t x = v & 0x1fffff;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ inline double decode(int32_t v, double s, double o) { return o + s * (double)v; }

// the input points: int32 LAS coordinates with the transform of Xform, or
// coordinates already decoded to FP64; `packed` is a copy k_keygen writes
// with x, y, z of a point side by side (int4, or double4 for FP64), so the
// later gathers by sorted slot fetch one vector per point instead of three
// scattered words
struct Src {
  const int32_t *xi, *yi, *zi;
  const double *fx, *fy, *fz;
  void *packed;
};

template <bool F64>
__device__ inline void load_point(const Src &s, const Xform &f, uint64_t i, double &x, double &y,
                                  double &z)
{
  if (F64) {
    x = s.fx[i]; y = s.fy[i]; z = s.fz[i];
  } else {
    x = decode(s.xi[i], f.sx, f.ox); y = decode(s.yi[i], f.sy, f.oy); z = decode(s.zi[i], f.sz, f.oz);
  }
}

// the same point from the packed copy
template <bool F64>
__device__ inline void load_packed(const Src &s, const Xform &f, uint64_t i, double &x, double &y,
                                   double &z)
{
  if (F64) {
    const double4 v = ((const double4 *)s.packed)[i];
    x = v.x; y = v.y; z = v.z;
  } else {
    const int4 v = ((const int4 *)s.packed)[i];
    x = decode(v.x, f.sx, f.ox); y = decode(v.y, f.sy, f.oy); z = decode(v.z, f.sz, f.oz);
  }
}

// K = uint64_t: the whole Morton key; K = uint32_t: its top 32 bits, which
// are all a build that sorts at most 4 digits (32 bits) and merges no
// duplicates reads after the sort (the split reads the top 3 bits per level
// of at most 10 levels)
template <bool F64, typename K>
__global__ void k_keygen(Src src, uint64_t n, Xform f, K *key, uint32_t *val,
                         uint32_t *outside)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double x, y, z;
  load_point<F64>(src, f, i, x, y, z);
  // a point outside the closed root cube (or not a number) lands in an
  // extreme cell, possibly far outside its leaf's cube
  const double side = f.cell * (double)(1 << WB_MAX_LEVEL);
  const bool out = !(x >= f.lox && x <= f.lox + side && y >= f.loy && y <= f.loy + side &&
                     z >= f.loz && z <= f.loz + side);
  const uint64_t ob = __ballot(out);
  if (ob && (int)__lane_id() == __ffsll((long long)ob) - 1) atomicOr(outside, 1u);
  if (F64) ((double4 *)src.packed)[i] = make_double4(x, y, z, 0.0);
  else ((int4 *)src.packed)[i] = make_int4(src.xi[i], src.yi[i], src.zi[i], 0);
  uint32_t ix = cell_index(x, f.lox, f.cell, f.inv_cell);
  uint32_t iy = cell_index(y, f.loy, f.cell, f.inv_cell);
  uint32_t iz = cell_index(z, f.loz, f.cell, f.inv_cell);
  uint64_t k = (spread3(iz) << 2) | (spread3(iy) << 1) | spread3(ix);
  key[i] = (K)((k << 1) >> (64 - 8 * sizeof(K)));
  val[i] = (uint32_t)i;
}

// ---------------------------------------------------------------------------
// K3: canonical split, one level per launch.  Each active node (range of the
// sorted keys sharing its prefix) finds its eight child ranges by binary
// search on the next 3-bit digit; children with more than 537 points become
// the next level's active nodes, the others non-empty children are leaves.
// ---------------------------------------------------------------------------
struct Active {
  uint32_t start, end;
  int32_t node;
  int32_t pad;
};

struct LeafTmp {
  uint32_t start, count;
  int32_t parent, depth;
  double c[4];  // cx, cy, cz, side
};

struct SplitOut {
  uint32_t n_next, n_leaf, n_node, overflow, maxdepth;
};

// the level-`level` digit of a key whose top bit is bit 8*sizeof(K)-1
template <typename K>
__device__ inline uint32_t digit_at(K k, int level)
{
  return (uint32_t)(k >> (8 * (int)sizeof(K) - 3 * level)) & 7u;
}

// The levels are launched back to back: the number of active nodes of a
// level is the counter the previous level incremented (act_cnt), and a level
// after a failed one does nothing.
// A level with at most SPLIT_WAVE_MAX active nodes (the top of the tree:
// few nodes, long ranges) takes a wave per node; a level with more takes a
// thread per node (k_split_level_t).  Both kernels are launched for every
// level, since only the device knows the count; the other returns at once.
#define SPLIT_WAVE_MAX 2048u

// One thread per active node (levels with many active nodes, each with a
// short range): seven binary searches for the boundaries.
template <typename K>
__global__ void k_split_level_t(const K *key, const Active *act, const uint32_t *act_cnt,
                              uint32_t act_cap, int level, int max_level, Active *next,
                              uint32_t *next_cnt, uint32_t next_cap, LeafTmp *leaves,
                              uint32_t leaf_cap, double *node_cube, int32_t *node_child,
                              int32_t *node_parent, uint32_t node_cap, SplitOut *out)
{
  if (out->overflow) return;
  const uint32_t nact = min(*act_cnt, act_cap);
  if (nact <= SPLIT_WAVE_MAX) return;  // k_split_level takes this level
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= nact) return;
  Active A = act[a];
  int node = A.node;
  const double cx = node_cube[4 * node], cy = node_cube[4 * node + 1], cz = node_cube[4 * node + 2];
  const double side = node_cube[4 * node + 3];
  uint32_t bound[9];
  bound[0] = A.start;
  bound[8] = A.end;
  int lv = level + 1;
  for (int c = 1; c < 8; c++) {
    // first index in [bound[c-1], end) whose digit >= c
    uint32_t lo = bound[c - 1], hi = A.end;
    while (lo < hi) {
      uint32_t mid = lo + (hi - lo) / 2;
      if (digit_at(key[mid], lv) < (uint32_t)c) lo = mid + 1;
      else hi = mid;
    }
    bound[c] = lo;
  }
  for (int c = 0; c < 8; c++) {
    uint32_t cnt = bound[c + 1] - bound[c];
    int32_t code = -1;
    if (cnt > 0) {
      // Octree::cube(c), octree.cpp:349-358
      int xb = c & 1, yb = (c & 2) / 2, zb = (c & 4) / 4;
      double ncx = cx + (2 * xb - 1) * side / 4;
      double ncy = cy + (2 * yb - 1) * side / 4;
      double ncz = cz + (2 * zb - 1) * side / 4;
      double ns = side / 2;
      if (cnt > WB_RECORDS) {
        if (lv >= max_level) {
          atomicOr(&out->overflow, 1u);
        } else {
          uint32_t nid = atomicAdd(&out->n_node, 1u);
          uint32_t slot = atomicAdd(next_cnt, 1u);
          if (nid < node_cap && slot < next_cap) {
            node_cube[4 * nid] = ncx;
            node_cube[4 * nid + 1] = ncy;
            node_cube[4 * nid + 2] = ncz;
            node_cube[4 * nid + 3] = ns;
            node_parent[nid] = node;
            for (int k = 0; k < 8; k++) node_child[8 * nid + k] = -1;
            Active na;
            na.start = bound[c];
            na.end = bound[c + 1];
            na.node = (int32_t)nid;
            na.pad = 0;
            next[slot] = na;
            code = (int32_t)nid;
          } else {
            atomicOr(&out->overflow, 2u);
          }
        }
      } else {
        uint32_t ls = atomicAdd(&out->n_leaf, 1u);
        if (ls < leaf_cap) {
          LeafTmp L;
          L.start = bound[c];
          L.count = cnt;
          L.parent = node;
          L.depth = lv;
          L.c[0] = ncx;
          L.c[1] = ncy;
          L.c[2] = ncz;
          L.c[3] = ns;
          leaves[ls] = L;
          code = -2 - (int32_t)ls;
          atomicMax(&out->maxdepth, (uint32_t)lv);
        } else {
          atomicOr(&out->overflow, 4u);
        }
      }
    }
    node_child[8 * node + c] = code;
  }
}

// One wave per active node.  The eight child ranges follow from the seven
// boundaries "first slot whose level-(l+1) digit is >= c" (c = 1..7) of the
// node's sorted keys: 64 evenly spaced samples place every boundary between
// two neighbouring samples, then lanes 8(c-1)..8(c-1)+7 narrow boundary c by
// a 9-way search (8 probes per round).  Lanes 0..7 then make child c = lane.
template <typename K>
__global__ void __launch_bounds__(256) k_split_level(
    const K *key, const Active *act, const uint32_t *act_cnt, uint32_t act_cap, int level,
    int max_level, Active *next, uint32_t *next_cnt, uint32_t next_cap, LeafTmp *leaves,
    uint32_t leaf_cap, double *node_cube, int32_t *node_child, int32_t *node_parent,
    uint32_t node_cap, SplitOut *out)
{
  if (out->overflow) return;
  const uint32_t nact = min(*act_cnt, act_cap);
  if (nact > SPLIT_WAVE_MAX) return;  // k_split_level_t takes this level
  const uint32_t lane = threadIdx.x & 63;
  const int lv = level + 1;
  for (uint32_t a = blockIdx.x * 4 + threadIdx.x / 64; a < nact; a += gridDim.x * 4) {
    const Active A = act[a];
    const int node = A.node;
    const uint32_t len = A.end - A.start;  // > 537: an active node
    // samples l = 0..63 (sorted, so their digits are too)
    const uint32_t ps = A.start + (uint32_t)(((uint64_t)lane * len) >> 6);
    const uint32_t ds = digit_at(key[ps], lv);
    // boundary c of this lane's group: it lies after the last sample with a
    // digit below c and at or before the first one with a digit of c or more
    const uint32_t c = (lane >> 3) + 1, q = lane & 7;
    uint32_t lo = A.start, hi = A.end;
#pragma unroll
    for (uint32_t cc = 1; cc < 8; cc++) {
      const uint32_t j = (uint32_t)__popcll(__ballot(ds < cc));
      const uint32_t plo = (uint32_t)__shfl((int)ps, (int)(j == 0 ? 0 : j - 1), 64);
      const uint32_t phi = (uint32_t)__shfl((int)ps, (int)(j == 64 ? 63 : j), 64);
      if (c == cc) {
        lo = j == 0 ? A.start : plo + 1;
        hi = j == 64 ? A.end : phi;
      }
    }
    // the first slot of [lo, hi) with digit >= c, or hi; 9-way per round
    for (;;) {
      const bool open = c < 8 && lo < hi;
      if (!__ballot(open)) break;
      if (open) {
        // the 8 lanes of a group share lo and hi, so they open and close
        // together and the ballot and shuffles stay inside the group
        const uint32_t p = lo + (uint32_t)(((uint64_t)(hi - lo) * (q + 1)) / 9);
        const bool below = digit_at(key[p], lv) < c;
        // probes of the group, ascending in q: count those below c
        const uint64_t bm = __ballot(below);
        const uint32_t k = (uint32_t)__popcll((bm >> (lane & ~7u)) & 0xffull);
        const uint32_t pk1 = (uint32_t)__shfl((int)p, (int)((lane & ~7u) + (k == 0 ? 0 : k - 1)), 64);
        const uint32_t pk = (uint32_t)__shfl((int)p, (int)((lane & ~7u) + (k == 8 ? 7 : k)), 64);
        const uint32_t nlo = k == 0 ? lo : pk1 + 1;
        const uint32_t nhi = k == 8 ? hi : pk;
        lo = nlo;
        hi = nhi;
      }
    }
    // child c = lane (0..7): slots [bound[c], bound[c+1]); every lane takes
    // part in both shuffles (a lane left out of one supplies no value)
    const uint32_t s_lo = (uint32_t)__shfl((int)lo, (int)(8 * ((lane - 1) & 7)), 64);
    const uint32_t s_hi = (uint32_t)__shfl((int)lo, (int)(8 * (lane & 7)), 64);
    const uint32_t b_lo = lane == 0 ? A.start : s_lo;
    const uint32_t b_hi = lane == 7 ? A.end : s_hi;
    const uint32_t cnt = b_hi - b_lo;
    const bool valid = lane < 8 && cnt > 0;
    const bool big = valid && cnt > WB_RECORDS;
    const bool is_node = big && lv < max_level, is_leaf = valid && !big;
    if (big && lv >= max_level) atomicOr(&out->overflow, 1u);
    // one counter update per wave and kind; children numbered in child order
    const uint64_t mnode = __ballot(is_node), mleaf = __ballot(is_leaf);
    uint32_t nbase = 0, sbase = 0, lbase = 0;
    if (lane == 0) {
      if (mnode) {
        nbase = atomicAdd(&out->n_node, (uint32_t)__popcll(mnode));
        sbase = atomicAdd(next_cnt, (uint32_t)__popcll(mnode));
      }
      if (mleaf) {
        lbase = atomicAdd(&out->n_leaf, (uint32_t)__popcll(mleaf));
        atomicMax(&out->maxdepth, (uint32_t)lv);
      }
    }
    nbase = (uint32_t)__shfl((int)nbase, 0, 64);
    sbase = (uint32_t)__shfl((int)sbase, 0, 64);
    lbase = (uint32_t)__shfl((int)lbase, 0, 64);
    const uint64_t below = (1ull << lane) - 1;
    if (lane < 8) {
      const int cch = (int)lane;
      int32_t code = -1;
      if (valid) {
        const double cx = node_cube[4 * node], cy = node_cube[4 * node + 1];
        const double cz = node_cube[4 * node + 2], side = node_cube[4 * node + 3];
        // Octree::cube(c), octree.cpp:349-358
        const int xb = cch & 1, yb = (cch & 2) / 2, zb = (cch & 4) / 4;
        const double ncx = cx + (2 * xb - 1) * side / 4;
        const double ncy = cy + (2 * yb - 1) * side / 4;
        const double ncz = cz + (2 * zb - 1) * side / 4;
        const double ns = side / 2;
        if (is_node) {
          const uint32_t nid = nbase + (uint32_t)__popcll(mnode & below);
          const uint32_t slot = sbase + (uint32_t)__popcll(mnode & below);
          if (nid < node_cap && slot < next_cap) {
            node_cube[4 * nid] = ncx;
            node_cube[4 * nid + 1] = ncy;
            node_cube[4 * nid + 2] = ncz;
            node_cube[4 * nid + 3] = ns;
            node_parent[nid] = node;
            for (int k = 0; k < 8; k++) node_child[8 * nid + k] = -1;
            Active na;
            na.start = b_lo;
            na.end = b_hi;
            na.node = (int32_t)nid;
            na.pad = 0;
            next[slot] = na;
            code = (int32_t)nid;
          } else {
            atomicOr(&out->overflow, 2u);
          }
        } else if (is_leaf) {
          const uint32_t ls = lbase + (uint32_t)__popcll(mleaf & below);
          if (ls < leaf_cap) {
            LeafTmp L;
            L.start = b_lo;
            L.count = cnt;
            L.parent = node;
            L.depth = lv;
            L.c[0] = ncx;
            L.c[1] = ncy;
            L.c[2] = ncz;
            L.c[3] = ns;
            leaves[ls] = L;
            code = -2 - (int32_t)ls;
          } else {
            atomicOr(&out->overflow, 4u);
          }
        }
      }
      node_child[8 * node + cch] = code;
    }
  }
}

__global__ void k_leaf_keys(const LeafTmp *lt, uint32_t nl, uint32_t *k, uint32_t *v)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  k[i] = lt[i].start;
  v[i] = i;
}

__global__ void k_leaf_final(const LeafTmp *lt, const uint32_t *perm, uint32_t nl, uint64_t n,
                             uint32_t *leaf_start, double *leaf_cube, int32_t *leaf_parent,
                             uint32_t *inv)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i > nl) return;
  if (i == nl) {
    leaf_start[nl] = (uint32_t)n;
    return;
  }
  LeafTmp L = lt[perm[i]];
  leaf_start[i] = L.start;
  leaf_cube[4 * i] = L.c[0];
  leaf_cube[4 * i + 1] = L.c[1];
  leaf_cube[4 * i + 2] = L.c[2];
  leaf_cube[4 * i + 3] = L.c[3];
  leaf_parent[i] = L.parent;
  inv[perm[i]] = i;
}

// leaf children renumbered in DFS order, with the slot range of each
__global__ void k_patch_children(int32_t *child, uint32_t nn, const uint32_t *inv,
                                 const uint32_t *leaf_start, uint2 *crange)
{
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn * 8u) return;
  int32_t c = child[i];
  uint2 r = make_uint2(0u, 0u);
  if (c <= -2) {
    const uint32_t l = inv[-2 - c];
    child[i] = -2 - (int32_t)l;
    r = make_uint2(leaf_start[l], leaf_start[l + 1]);
  }
  crange[i] = r;
}

__global__ void k_fill_pt_leaf(const uint32_t *leaf_start, uint32_t nl, uint32_t *pt_leaf)
{
  uint32_t l = blockIdx.x;
  if (l >= nl) return;
  uint32_t s = leaf_start[l], e = leaf_start[l + 1];
  for (uint32_t i = s + threadIdx.x; i < e; i += blockDim.x) pt_leaf[i] = l;
}

// ---------------------------------------------------------------------------
// The lowest point under every child of every internal node (the ground
// test raises a child cube's bottom face to it, wb::hyp_intersect_low): a
// wave per leaf for the leaves' lowest z, then one launch per depth, deepest
// first, in which each internal node of that depth takes the lowest of its
// children's.  Node ids grow with depth (a level's nodes are made while the
// level above splits), so a node's depth follows from its side.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_leaf_zlow(const uint32_t *leaf_start, uint32_t nl,
                                                   const double *pz, double *leaf_zlow)
{
  const uint32_t l = blockIdx.x * 4 + threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
  if (l >= nl) return;
  const uint32_t b = leaf_start[l], e = leaf_start[l + 1];
  double m = __builtin_inf();
  for (uint32_t j = b + lane; j < e; j += WAVE) m = fmin(m, pz[j]);
  for (int d = 1; d < WAVE; d <<= 1) m = fmin(m, __shfl_xor(m, d, WAVE));
  if (lane == 0) leaf_zlow[l] = m;
}

__global__ void k_node_zlow(const int32_t *node_child, const double *node_cube, uint32_t nn,
                            int depth, const double *leaf_zlow, double *node_zlow,
                            double *child_zlow)
{
  const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nn) return;
  if (ilogb(node_cube[3]) - ilogb(node_cube[4 * (size_t)v + 3]) != depth) return;
  double m = __builtin_inf();
  for (int k = 0; k < 8; k++) {
    const int32_t c = node_child[8 * (size_t)v + k];
    const double zl = c <= -2 ? leaf_zlow[-2 - c] : c >= 0 ? node_zlow[c] : __builtin_inf();
    child_zlow[8 * (size_t)v + k] = zl;
    m = fmin(m, zl);
  }
  node_zlow[v] = m;
}

// ---------------------------------------------------------------------------
// K4: SoA permute (decode in sorted order) and 32-slot bounding boxes
// ---------------------------------------------------------------------------
template <bool F64>
__global__ void k_permute(Src src, const uint32_t *sidx, uint64_t n, Xform f, double *px,
                          double *py, double *pz, double *box, uint32_t *low, double *lowx,
                          double *lowy, double *lowz)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool ok = i < n;
  double x = 0, y = 0, z = 0;
  if (ok) {
    load_packed<F64>(src, f, sidx[i], x, y, z);
    px[i] = x;
    py[i] = y;
    pz[i] = z;
  }
  // the box and the lowest point (smallest z, then smallest slot) of each
  // 32-slot group, by half-wave butterflies over the values just gathered
  double x0 = ok ? x : __builtin_inf(), x1 = ok ? x : -__builtin_inf();
  double y0 = ok ? y : __builtin_inf(), y1 = ok ? y : -__builtin_inf();
  double z0 = ok ? z : __builtin_inf(), z1 = ok ? z : -__builtin_inf();
  double lz = z0, lx = x, ly = y;
  uint32_t li = (uint32_t)i;
#pragma unroll
  for (int d = 1; d < WB_CHUNK; d <<= 1) {
    const double oz = __shfl_xor(lz, d, WAVE);
    const uint32_t oi = __shfl_xor(li, d, WAVE);
    const double ox = __shfl_xor(lx, d, WAVE), oy = __shfl_xor(ly, d, WAVE);
    if (oz < lz || (oz == lz && oi < li)) {
      lz = oz;
      li = oi;
      lx = ox;
      ly = oy;
    }
    x0 = fmin(x0, __shfl_xor(x0, d, WAVE));
    x1 = fmax(x1, __shfl_xor(x1, d, WAVE));
    y0 = fmin(y0, __shfl_xor(y0, d, WAVE));
    y1 = fmax(y1, __shfl_xor(y1, d, WAVE));
    z0 = fmin(z0, __shfl_xor(z0, d, WAVE));
    z1 = fmax(z1, __shfl_xor(z1, d, WAVE));
  }
  const uint64_t chunk = i / WB_CHUNK;
  if ((i % WB_CHUNK) == 0 && chunk < (n + WB_CHUNK - 1) / WB_CHUNK) {
    double *b = box + 6 * chunk;
    b[0] = x0; b[1] = x1; b[2] = y0; b[3] = y1; b[4] = z0; b[5] = z1;
    low[chunk] = li;
    lowx[chunk] = lx;
    lowy[chunk] = ly;
    lowz[chunk] = lz;
  }
}

// ---------------------------------------------------------------------------
// Duplicate locations (OctBuffer::put, octree.cpp:620-660): a point put at
// the location of a stored point replaces that point's record in place, so
// the store holds each location once, at the place of its first occurrence,
// with the record of its last.  Equal locations have equal keys, so after
// the (stable, LSD) sort on the key bits above `shift` they lie in one run
// of equal sorted bits, in input order; such runs hold the points of one
// cell of the deepest sorted level and are short.  For each slot: grp =
// slot of the first equal location in its run, last = input index of the
// last one.
// ---------------------------------------------------------------------------
template <bool F64>
__global__ void k_dup_mark(const uint64_t *key, const uint32_t *sidx, uint64_t n, int shift,
                           Src src, Xform f, uint32_t *grp, uint32_t *last)
{
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint64_t k = key[s] >> shift;
  const bool left = s > 0 && (key[s - 1] >> shift) == k;
  const bool right = s + 1 < n && (key[s + 1] >> shift) == k;
  if (!left && !right) {
    grp[s] = (uint32_t)s;
    last[s] = sidx[s];
    return;
  }
  double x, y, z;
  load_packed<F64>(src, f, sidx[s], x, y, z);
  uint64_t a = s, b = s + 1;
  while (a > 0 && (key[a - 1] >> shift) == k) a--;
  while (b < n && (key[b] >> shift) == k) b++;
  uint64_t first = s, lastj = s;
  for (uint64_t j = a; j < s; j++) {
    double qx, qy, qz;
    load_packed<F64>(src, f, sidx[j], qx, qy, qz);
    if (qx == x && qy == y && qz == z) {
      first = j;
      break;
    }
  }
  for (uint64_t j = b - 1; j > s; j--) {
    double qx, qy, qz;
    load_packed<F64>(src, f, sidx[j], qx, qy, qz);
    if (qx == x && qy == y && qz == z) {
      lastj = j;
      break;
    }
  }
  grp[s] = (uint32_t)first;
  last[s] = sidx[lastj];
}

__global__ void k_dup_flag(const uint32_t *grp, uint64_t n, uint32_t *flag)
{
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) flag[s] = grp[s] == (uint32_t)s ? 1u : 0u;
}

// firsts move to their compact slot; the others become (input index, input
// index of the first) pairs.  pos is the exclusive scan of the first flags.
__global__ void k_dup_compact(const uint64_t *key, const uint32_t *sidx, const uint32_t *grp,
                              const uint32_t *last, const uint32_t *pos, uint64_t n,
                              uint64_t *key2, uint32_t *sidx2, uint32_t *last2, uint32_t *pairs)
{
  const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const uint32_t g = grp[s];
  if (g == (uint32_t)s) {
    const uint32_t p = pos[s];
    key2[p] = key[s];
    sidx2[p] = sidx[s];
    last2[p] = last[s];
  } else {
    const uint64_t d = s - pos[s];
    pairs[2 * d] = sidx[s];
    pairs[2 * d + 1] = sidx[g];
  }
}

// Per-leaf canonical order, input order inside each leaf: a sort of the
// leaf's input indices by a wave per leaf, in registers: leaves of
// 64(K/2)+1 .. 64K points (K = 1: up to 64) hold K keys per lane, element
// e = 64s + lane in slot s, sorted by the register network of wb_wave.h up
// to the next power of two of the leaf size.  No barrier, no LDS.  Keys
// (input indices) are distinct.
template <int K>
__global__ void __launch_bounds__(256) k_leaf_canon_reg(const uint32_t *leaf_start, uint32_t nl,
                                                        const uint32_t *sidx, uint32_t *canon,
                                                        uint32_t *crank)
{
  const uint32_t lane = threadIdx.x & 63;
  for (uint32_t l = blockIdx.x * 4 + threadIdx.x / 64; l < nl; l += gridDim.x * 4) {
    const uint32_t b = leaf_start[l], m = leaf_start[l + 1] - b;
    if (m <= (K == 1 ? 0u : 32u * K) || m > 64u * K) continue;
    uint32_t key[K], val[K];
#pragma unroll
    for (int s = 0; s < K; s++) {
      const uint32_t e = 64u * s + lane;
      key[s] = e < m ? sidx[b + e] : 0xffffffffu;
      val[s] = b + e;
    }
    wbw::wave_sort<K, true>(key, val, wbw::pow2_ceil(m));
#pragma unroll
    for (int s = 0; s < K; s++) {
      const uint32_t e = 64u * s + lane;
      if (e < m) {
        canon[b + e] = val[s];
        crank[val[s]] = b + e;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
// Sort digits for a first attempt (a build that needs more retries with one
// more digit, so this choice changes only the time): the levels down to
// cells of ~1 m, plus one, or, without the duplicate merge, the fewer of
// that and the levels a surface of n points needs for leaves of 537 points
// (each level quartering a leaf's points) plus two for relief and
// clustering.  The merge searches runs of equal sort prefix for equal
// locations, so it keeps the finer prefix and its short runs.
static int default_passes(double side, uint64_t n, bool dedupe)
{
  int lv = 1;
  while (lv < WB_MAX_LEVEL && side / (double)(1ll << lv) > 1.0) lv++;
  int ls = 2;
  while (ls < WB_MAX_LEVEL && ((double)WB_RECORDS * (double)(1ull << (2 * (ls - 2)))) < (double)n) ls++;
  const int levels = dedupe ? lv + 1 : std::min(lv + 1, ls);
  const int p = (3 * levels + 7) / 8;
  return std::min(8, std::max(2, p));
}

void wb_build_tree(wb_tree *t, const int32_t *dx, const int32_t *dy, const int32_t *dz,
                   const double *fx, const double *fy, const double *fz, bool dedupe)
{
  wb_ctx *c = t->ctx;
  const uint64_t n_input = t->n_input;
  uint64_t n = n_input;
  const bool f64 = fx != nullptr;
  const size_t pk_bytes = (f64 ? 32 : 16) * (size_t)(n_input ? n_input : 1);
  Src src = {dx, dy, dz, fx, fy, fz, wb_alloc(c, pk_bytes)};
  t->n = n;
  t->ndup = 0;
  t->dup_pairs = nullptr;
  t->last_of = nullptr;
  Xform f;
  f.sx = t->scale[0]; f.sy = t->scale[1]; f.sz = t->scale[2];
  f.ox = t->offset[0]; f.oy = t->offset[1]; f.oz = t->offset[2];
  f.lox = t->center[0] - t->side / 2;
  f.loy = t->center[1] - t->side / 2;
  f.loz = t->center[2] - t->side / 2;
  f.cell = t->side / (double)(1 << WB_MAX_LEVEL);
  f.inv_cell = (double)(1 << WB_MAX_LEVEL) / t->side;

  t->key = wb_talloc<uint64_t>(t, n);
  t->sidx = wb_talloc<uint32_t>(t, n);
  // timing events, owned by the tree and read once the stream has passed
  // them (wb_tree_read_timing): the build returns without waiting for its
  // last kernels
  hipEvent_t *evb = t->evb;
  for (int k = 0; k < 4; k++)
    if (!evb[k]) WB_CHECK(hipEventCreateWithFlags(&evb[k], hipEventDefault));
  t->evb_read = false;
  int passes = t->sort_passes > 0 ? t->sort_passes : default_passes(t->side, n_input, dedupe);
  const uint32_t node_cap = (uint32_t)std::min<uint64_t>(n / 64 + 1024, 0x7fffffff);
  t->node_cube = wb_talloc<double>(t, 4 * (size_t)node_cap);
  t->node_child = wb_talloc<int32_t>(t, 8 * (size_t)node_cap);
  t->node_parent = wb_talloc<int32_t>(t, node_cap);
  uint32_t leaf_cap = (uint32_t)std::min<uint64_t>(n / 8 + 1024, 0x7fffffff);
  LeafTmp *lt = nullptr;
  SplitOut *d_out = (SplitOut *)wb_alloc(c, sizeof(SplitOut));
  uint32_t *d_cnt = (uint32_t *)wb_alloc(c, 4 * (WB_MAX_LEVEL + 2));
  uint32_t *d_outside = (uint32_t *)wb_alloc(c, 4);
  static const uint32_t zero_word = 0;
  uint32_t h_outside = 0;
  SplitOut h_out;
  Active *act = (Active *)wb_alloc(c, sizeof(Active) * node_cap);
  Active *nxt = (Active *)wb_alloc(c, sizeof(Active) * node_cap);
  // the arrays of all input points; with duplicates merged, the tree's key
  // and slot arrays are compact copies
  uint64_t *key0 = t->key, *key2 = nullptr;
  uint32_t *sidx0 = t->sidx, *sidx2 = nullptr, *last2 = nullptr, *pairs2 = nullptr;
  for (;;) {
    // every attempt starts from all input points
    n = n_input;
    t->n = n;
    t->key = key0;
    t->sidx = sidx0;
    t->ndup = 0;
    t->last_of = nullptr;
    t->dup_pairs = nullptr;
    // a build of at most 4 sort digits that merges no duplicates reads only
    // the top 32 key bits after keygen, so it sorts (u32 key, u32 index)
    // pairs: 8 B per point moved per pass instead of 12
    const bool narrow = !dedupe && passes <= 4;
    uint32_t *key32 = (uint32_t *)key0;
    if (n > 0) {
      WB_CHECK(hipEventRecord(evb[0], c->stream));
      WB_CHECK(hipMemcpyAsync(d_outside, &zero_word, 4, hipMemcpyHostToDevice, c->stream));
      const dim3 kg(wb_blocks(n, 256)), kb(256);
      if (narrow) {
        if (f64)
          hipLaunchKernelGGL((k_keygen<true, uint32_t>), kg, kb, 0, c->stream, src, n, f, key32,
                             t->sidx, d_outside);
        else
          hipLaunchKernelGGL((k_keygen<false, uint32_t>), kg, kb, 0, c->stream, src, n, f, key32,
                             t->sidx, d_outside);
      } else {
        if (f64)
          hipLaunchKernelGGL((k_keygen<true, uint64_t>), kg, kb, 0, c->stream, src, n, f, t->key,
                             t->sidx, d_outside);
        else
          hipLaunchKernelGGL((k_keygen<false, uint64_t>), kg, kb, 0, c->stream, src, n, f, t->key,
                             t->sidx, d_outside);
      }
      WB_CHECK(hipEventRecord(evb[1], c->stream));
      if (narrow) wb_radix_sort_u32_prefix(c, key32, t->sidx, n, passes);
      else wb_radix_sort_pairs(c, t->key, t->sidx, n, passes);
      WB_CHECK(hipEventRecord(evb[2], c->stream));
      if (dedupe) {
        const int shift = 64 - 8 * passes;
        uint32_t *grp = (uint32_t *)wb_alloc(c, 4 * n), *last = (uint32_t *)wb_alloc(c, 4 * n);
        uint32_t *pos = (uint32_t *)wb_alloc(c, 4 * n);
        if (f64)
          hipLaunchKernelGGL(k_dup_mark<true>, dim3(wb_blocks(n, 256)), dim3(256), 0, c->stream,
                             t->key, t->sidx, n, shift, src, f, grp, last);
        else
          hipLaunchKernelGGL(k_dup_mark<false>, dim3(wb_blocks(n, 256)), dim3(256), 0, c->stream,
                             t->key, t->sidx, n, shift, src, f, grp, last);
        hipLaunchKernelGGL(k_dup_flag, dim3(wb_blocks(n, 256)), dim3(256), 0, c->stream, grp, n,
                           pos);
        const uint64_t nd = wb_exclusive_scan_u32(c, pos, n);
        if (nd < n) {
          if (!key2) {
            key2 = wb_talloc<uint64_t>(t, n_input);
            sidx2 = wb_talloc<uint32_t>(t, n_input);
            last2 = wb_talloc<uint32_t>(t, n_input);
            pairs2 = wb_talloc<uint32_t>(t, 2 * n_input);
          }
          hipLaunchKernelGGL(k_dup_compact, dim3(wb_blocks(n, 256)), dim3(256), 0, c->stream,
                             t->key, t->sidx, grp, last, pos, n, key2, sidx2, last2, pairs2);
          t->key = key2;
          t->sidx = sidx2;
          t->last_of = last2;
          t->dup_pairs = pairs2;
          t->ndup = n - nd;
          n = nd;
          t->n = n;
        }
        wb_free(c, grp, 4 * n_input);
        wb_free(c, last, 4 * n_input);
        wb_free(c, pos, 4 * n_input);
      }
    }
    // root node 0
    double root[4] = {t->center[0], t->center[1], t->center[2], t->side};
    int32_t empty8[8] = {-1, -1, -1, -1, -1, -1, -1, -1};
    int32_t m1 = -1;
    WB_CHECK(hipMemcpyAsync(t->node_cube, root, sizeof root, hipMemcpyHostToDevice, c->stream));
    WB_CHECK(hipMemcpyAsync(t->node_child, empty8, sizeof empty8, hipMemcpyHostToDevice, c->stream));
    WB_CHECK(hipMemcpyAsync(t->node_parent, &m1, 4, hipMemcpyHostToDevice, c->stream));
    memset(&h_out, 0, sizeof h_out);
    h_out.n_node = 1;
    WB_CHECK(hipMemcpyAsync(d_out, &h_out, sizeof h_out, hipMemcpyHostToDevice, c->stream));
    if (lt) wb_free(c, lt, sizeof(LeafTmp) * leaf_cap);
    lt = (LeafTmp *)wb_alloc(c, sizeof(LeafTmp) * leaf_cap);
    int max_level = std::min(WB_MAX_LEVEL, (8 * passes) / 3);
    // active nodes of one level are disjoint and hold more than 537 points
    const uint32_t act_ub = (uint32_t)std::min<uint64_t>(node_cap, n / (WB_RECORDS + 1) + 1);
    uint32_t hcnt[WB_MAX_LEVEL + 2] = {};
    if (n > 0) {
      Active r0 = {0u, (uint32_t)n, 0, 0};
      WB_CHECK(hipMemcpyAsync(act, &r0, sizeof r0, hipMemcpyHostToDevice, c->stream));
      hcnt[0] = 1;
    }
    WB_CHECK(hipMemcpyAsync(d_cnt, hcnt, sizeof hcnt, hipMemcpyHostToDevice, c->stream));
    for (int level = 0; n > 0 && level < max_level; level++) {
      const dim3 gw(std::min<uint32_t>(wb_blocks(act_ub, 4), 512)), gt(wb_blocks(act_ub, 128));
      if (narrow) {
        hipLaunchKernelGGL(k_split_level<uint32_t>, gw, dim3(256), 0, c->stream, key32, act,
                           d_cnt + level, node_cap, level, max_level, nxt, d_cnt + level + 1,
                           node_cap, lt, leaf_cap, t->node_cube, t->node_child, t->node_parent,
                           node_cap, d_out);
        if (act_ub > SPLIT_WAVE_MAX)
          hipLaunchKernelGGL(k_split_level_t<uint32_t>, gt, dim3(128), 0, c->stream, key32, act,
                             d_cnt + level, node_cap, level, max_level, nxt, d_cnt + level + 1,
                             node_cap, lt, leaf_cap, t->node_cube, t->node_child,
                             t->node_parent, node_cap, d_out);
      } else {
        hipLaunchKernelGGL(k_split_level<uint64_t>, gw, dim3(256), 0, c->stream, t->key, act,
                           d_cnt + level, node_cap, level, max_level, nxt, d_cnt + level + 1,
                           node_cap, lt, leaf_cap, t->node_cube, t->node_child, t->node_parent,
                           node_cap, d_out);
        if (act_ub > SPLIT_WAVE_MAX)
          hipLaunchKernelGGL(k_split_level_t<uint64_t>, gt, dim3(128), 0, c->stream, t->key, act,
                             d_cnt + level, node_cap, level, max_level, nxt, d_cnt + level + 1,
                             node_cap, lt, leaf_cap, t->node_cube, t->node_child,
                             t->node_parent, node_cap, d_out);
      }
      std::swap(act, nxt);
    }
    WB_CHECK(hipMemcpyAsync(&h_out, d_out, sizeof h_out, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipMemcpyAsync(&h_outside, d_outside, 4, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipStreamSynchronize(c->stream));
    t->outside_root = n > 0 ? h_outside : 0;
    // the split's scratch goes back to the cache before an error leaves
    auto release = [&]() {
      wb_free(c, src.packed, pk_bytes);
      wb_free(c, act, sizeof(Active) * node_cap);
      wb_free(c, nxt, sizeof(Active) * node_cap);
      wb_free(c, lt, sizeof(LeafTmp) * leaf_cap);
      wb_free(c, d_out, sizeof(SplitOut));
      wb_free(c, d_cnt, 4 * (WB_MAX_LEVEL + 2));
      wb_free(c, d_outside, 4);
    };
    if (h_out.overflow & 1u) {
      if (passes >= 8) {
        release();
        throw std::string("octree deeper than 21 levels (more than 537 points in a level-21 cell)");
      }
      passes++;
      continue;  // the sort did not reach deep enough: redo with one more digit
    }
    if (h_out.overflow) {
      release();
      throw std::string("octree node/leaf capacity exceeded");
    }
    break;
  }
  wb_free(c, act, sizeof(Active) * node_cap);
  wb_free(c, nxt, sizeof(Active) * node_cap);
  t->sort_passes = passes;
  t->nleaves = h_out.n_leaf;
  t->nnodes = h_out.n_node;
  t->maxdepth = (int)h_out.maxdepth;
  wb_free(c, d_out, sizeof(SplitOut));
  wb_free(c, d_cnt, 4 * (WB_MAX_LEVEL + 2));
  wb_free(c, d_outside, 4);

  // leaves in DFS order = ascending start
  const uint32_t nl = t->nleaves;
  t->leaf_start = wb_talloc<uint32_t>(t, nl + 1);
  t->leaf_cube = wb_talloc<double>(t, 4 * (size_t)nl + 4);
  t->leaf_parent = wb_talloc<int32_t>(t, nl + 1);
  if (nl > 0) {
    uint32_t *lk = (uint32_t *)wb_alloc(c, 4 * (size_t)nl), *lv = (uint32_t *)wb_alloc(c, 4 * (size_t)nl);
    uint32_t *inv = (uint32_t *)wb_alloc(c, 4 * (size_t)nl);
    hipLaunchKernelGGL(k_leaf_keys, dim3(wb_blocks(nl, 256)), dim3(256), 0, c->stream, lt, nl, lk, lv);
    wb_radix_sort_u32_pairs(c, lk, lv, nl);
    hipLaunchKernelGGL(k_leaf_final, dim3(wb_blocks(nl + 1, 256)), dim3(256), 0, c->stream, lt, lv,
                       nl, n, t->leaf_start, t->leaf_cube, t->leaf_parent, inv);
    t->node_crange = wb_talloc<uint2>(t, 8 * (size_t)t->nnodes);
    hipLaunchKernelGGL(k_patch_children, dim3(wb_blocks((uint64_t)t->nnodes * 8, 256)), dim3(256),
                       0, c->stream, t->node_child, t->nnodes, inv, t->leaf_start,
                       t->node_crange);
    wb_free(c, lk, 4 * (size_t)nl);
    wb_free(c, lv, 4 * (size_t)nl);
    wb_free(c, inv, 4 * (size_t)nl);
  } else {
    uint32_t zero = 0;
    WB_CHECK(hipMemcpyAsync(t->leaf_start, &zero, 4, hipMemcpyHostToDevice, c->stream));
  }
  wb_free(c, lt, sizeof(LeafTmp) * leaf_cap);

  t->pt_leaf = wb_talloc<uint32_t>(t, n);
  t->px = wb_talloc<double>(t, n);
  t->py = wb_talloc<double>(t, n);
  t->pz = wb_talloc<double>(t, n);
  t->canon = wb_talloc<uint32_t>(t, n);
  t->crank = wb_talloc<uint32_t>(t, n);
  uint64_t nchunks = (n + WB_CHUNK - 1) / WB_CHUNK;
  t->chunk_box = wb_talloc<double>(t, 6 * nchunks + 6);
  t->chunk_low = wb_talloc<uint32_t>(t, nchunks + 1);
  t->chunk_sx = wb_talloc<double>(t, nchunks + 1);
  t->chunk_sy = wb_talloc<double>(t, nchunks + 1);
  t->chunk_sz = wb_talloc<double>(t, nchunks + 1);
  if (n > 0) {
    hipLaunchKernelGGL(k_fill_pt_leaf, dim3(nl), dim3(256), 0, c->stream, t->leaf_start, nl,
                       t->pt_leaf);
    if (f64)
      hipLaunchKernelGGL(k_permute<true>, dim3(wb_blocks(nchunks * WB_CHUNK, 256)), dim3(256), 0,
                         c->stream, src, t->sidx, n, f, t->px, t->py, t->pz, t->chunk_box,
                         t->chunk_low, t->chunk_sx, t->chunk_sy, t->chunk_sz);
    else
      hipLaunchKernelGGL(k_permute<false>, dim3(wb_blocks(nchunks * WB_CHUNK, 256)), dim3(256), 0,
                         c->stream, src, t->sidx, n, f, t->px, t->py, t->pz, t->chunk_box,
                         t->chunk_low, t->chunk_sx, t->chunk_sy, t->chunk_sz);
    wb_free(c, src.packed, pk_bytes);
    src.packed = nullptr;
    {
      // a wave per leaf, by size class (every leaf holds at most 537 points)
      const unsigned nb = std::min<unsigned>(wb_blocks(nl, 4), 4096);
      hipLaunchKernelGGL(k_leaf_canon_reg<1>, dim3(nb), dim3(256), 0, c->stream, t->leaf_start,
                         nl, t->sidx, t->canon, t->crank);
      hipLaunchKernelGGL(k_leaf_canon_reg<2>, dim3(nb), dim3(256), 0, c->stream, t->leaf_start,
                         nl, t->sidx, t->canon, t->crank);
      hipLaunchKernelGGL(k_leaf_canon_reg<4>, dim3(nb), dim3(256), 0, c->stream, t->leaf_start,
                         nl, t->sidx, t->canon, t->crank);
      hipLaunchKernelGGL(k_leaf_canon_reg<8>, dim3(nb), dim3(256), 0, c->stream, t->leaf_start,
                         nl, t->sidx, t->canon, t->crank);
      hipLaunchKernelGGL(k_leaf_canon_reg<16>, dim3(nb), dim3(256), 0, c->stream, t->leaf_start,
                         nl, t->sidx, t->canon, t->crank);
    }
    {
      double *leaf_zlow = (double *)wb_alloc(c, 8 * (size_t)nl);
      double *node_zlow = (double *)wb_alloc(c, 8 * (size_t)t->nnodes);
      t->child_zlow = wb_talloc<double>(t, 8 * (size_t)t->nnodes);
      hipLaunchKernelGGL(k_leaf_zlow, dim3(wb_blocks(nl, 4)), dim3(256), 0, c->stream,
                         t->leaf_start, nl, t->pz, leaf_zlow);
      for (int d = t->maxdepth - 1; d >= 0; d--)
        hipLaunchKernelGGL(k_node_zlow, dim3(wb_blocks(t->nnodes, 256)), dim3(256), 0, c->stream,
                           t->node_child, t->node_cube, t->nnodes, d, leaf_zlow, node_zlow,
                           t->child_zlow);
      wb_free(c, leaf_zlow, 8 * (size_t)nl);
      wb_free(c, node_zlow, 8 * (size_t)t->nnodes);
    }
    WB_CHECK(hipEventRecord(evb[3], c->stream));
  } else {
    t->evb_read = true;
  }
  if (src.packed) wb_free(c, src.packed, pk_bytes);
}

void wb_tree_read_timing(wb_tree *t)
{
  if (t->evb_read || !t->evb[3]) return;
  float a = 0, b = 0;
  if (hipEventElapsedTime(&a, t->evb[0], t->evb[3]) != hipSuccess ||
      hipEventElapsedTime(&b, t->evb[1], t->evb[2]) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  t->ms_build = a;
  t->ms_sort = b;
  t->evb_read = true;
}

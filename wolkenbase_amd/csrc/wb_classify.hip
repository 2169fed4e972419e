// wb_classify.hip -- shape queries and the SCAN / POSTSCAN / CLASSIFY pass.
//
// Equivalences used (each is exact, not approximate):
//  * A point's class depends only on its own point, the hyperboloid size of
//    the last tile (flowsnake counter order) whose pointsIn(cylinder) holds
//    it, and the set D of directions to the other points of
//    pointsIn(downward hyperboloid) (classify.cpp:96-173).  surround()'s
//    pruning only removes directions whose two neighbour gaps are both under
//    30 deg and never the first or last one, so it never creates or removes
//    a gap of 144 deg; and the first loop over the previous point's
//    downPoints only inserts directions of points that the second loop
//    inserts too.  So class 1 <=> |D|>1 and every gap between successive
//    sorted directions, and the wrap-around gap, is < DEG144.
//  * When every direction is in [-2^30, 2^30), those gaps are the circular
//    gaps of the directions mod 2^31, and adding directions only shrinks
//    them: the test may stop as soon as the directions seen so far pass.
//    The host checks the cloud extent (< 1e7 scale units) that guarantees
//    the smallest non-zero dy/|dx| keeps every direction >= 30 units below
//    +180 deg.
//  * Gaps are tracked with a 64-bit occupancy mask of 2^25-unit bins.  A
//    longest circular run of <= 23 empty bins bounds every gap by
//    25*2^25 < DEG144 (pass); >= 26 empty bins gives a gap > 26*2^25 > DEG144
//    (fail); a run of 24 or 25 is decided in a second pass that records the
//    exact extreme directions in the two bins bordering the run.
//  * Skipping a 32-point group whose bounding box fails the shape test is
//    exact: clamping the shape centre into the box makes each coordinate
//    difference no larger than for any point inside, and every predicate is
//    monotone in those differences.
#include <stdio.h>
#include <stdlib.h>

#include <algorithm>

#include "wb_flowsnake.h"
#include "wb_internal.h"
#include "wb_wave.h"
#include "../../include/wb_gpu.h"

#define WAVE 64
// waves per SIMD the ground-test kernel is compiled for
#ifndef WB_GROUND_MINW
#define WB_GROUND_MINW 6
#endif
constexpr double kSqrt34 = 0.86602540378443864676372317;  // eisenstein.h:33
constexpr double kSqrt7 = 2.6457513110645905905016;       // scan.h:23

// lane values: the first active lane's, and lane l's (l wave-uniform)
__device__ inline int rfl(int v) { return __builtin_amdgcn_readfirstlane(v); }
// lane l's value of v (l wave-uniform)
__device__ inline double rdl(double v, uint32_t l)
{
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, (int)l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), (int)l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The device hands consecutive workgroups to its 8 XCDs in turn, and each
// XCD has its own L2.  xcd_block(b, nb) renumbers the nb blocks of a grid so
// that the blocks one XCD runs take one contiguous range of the work: kernels
// whose neighbouring blocks read the same leaves then find them in that XCD's
// L2.  A permutation of [0, nb): results do not depend on it.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb)
{
  const uint32_t x = b % 8, per = nb / 8, extra = nb % 8;
  return x * per + (x < extra ? x : extra) + b / 8;
}

struct TreeView {
  const double *px, *py, *pz;
  const uint32_t *leaf_start;
  const double *leaf_cube;
  const int32_t *leaf_parent;
  const double *node_cube;
  const int32_t *node_child, *node_parent;
  const double *chunk_box;
  const uint32_t *chunk_low;
  const double *chunk_sx, *chunk_sy, *chunk_sz;
  const uint2 *node_crange;
  const double *child_zlow;
  const uint32_t *pt_leaf, *canon, *crank, *sidx;
  uint32_t nleaves, nnodes;
  uint64_t n;
};

static TreeView view(const wb_tree *t)
{
  TreeView v;
  v.px = t->px; v.py = t->py; v.pz = t->pz;
  v.leaf_start = t->leaf_start; v.leaf_cube = t->leaf_cube; v.leaf_parent = t->leaf_parent;
  v.node_cube = t->node_cube; v.node_child = t->node_child; v.node_parent = t->node_parent;
  v.chunk_box = t->chunk_box; v.chunk_low = t->chunk_low;
  v.chunk_sx = t->chunk_sx; v.chunk_sy = t->chunk_sy; v.chunk_sz = t->chunk_sz; v.node_crange = t->node_crange;
  v.child_zlow = t->child_zlow;
  v.pt_leaf = t->pt_leaf; v.canon = t->canon; v.crank = t->crank;
  v.sidx = t->sidx; v.nleaves = t->nleaves; v.nnodes = t->nnodes; v.n = t->n;
  return v;
}

// ---------------------------------------------------------------------------
// Octree::findBlocks (octree.cpp:234-251) without a stack: the next child
// to visit at each depth is kept in 4-bit fields of two registers and the
// walk goes back up through parent links.  Leaves are visited in DFS order.
// f(leaf) returns true to stop the walk.
// ---------------------------------------------------------------------------
struct Iter {
  uint64_t lo, hi;
  __device__ int get(int d) const { return d < 16 ? (int)((lo >> (4 * d)) & 15) : (int)((hi >> (4 * (d - 16))) & 15); }
  __device__ void set(int d, int v)
  {
    if (d < 16) lo = (lo & ~(15ull << (4 * d))) | ((uint64_t)v << (4 * d));
    else hi = (hi & ~(15ull << (4 * (d - 16)))) | ((uint64_t)v << (4 * (d - 16)));
  }
};

template <class F>
__device__ inline void for_leaves(const TreeView &T, const wb::Shape &S, F &&f)
{
  if (T.nleaves == 0) return;
  if (!wb::shape_intersect(S, T.node_cube[0], T.node_cube[1], T.node_cube[2], T.node_cube[3])) return;
  int node = 0, depth = 0;
  Iter it = {0, 0};
  for (;;) {
    int c = it.get(depth);
    if (c == 8) {
      if (depth == 0) return;
      node = T.node_parent[node];
      depth--;
      continue;
    }
    it.set(depth, c + 1);
    int ch = T.node_child[8 * node + c];
    if (ch == -1) continue;
    if (ch >= 0) {
      const double *q = T.node_cube + 4 * ch;
      if (wb::shape_intersect(S, q[0], q[1], q[2], q[3])) {
        node = ch;
        depth++;
        it.set(depth, 0);
      }
      continue;
    }
    int leaf = -2 - ch;
    const double *q = T.leaf_cube + 4 * leaf;
    if (!wb::shape_intersect(S, q[0], q[1], q[2], q[3])) continue;
    if (f(leaf)) return;
  }
}

__device__ inline wb::Shape make_shape(int kind, const double *p)
{
  if (kind == wb::SHAPE_CYLINDER) return wb::make_cylinder(p[0], p[1], p[2]);
  if (kind == wb::SHAPE_SPHERE) return wb::make_sphere(p[0], p[1], p[2], p[3]);
  return wb::make_hyperboloid(p[0], p[1], p[2], p[3], p[4]);
}

// ---------------------------------------------------------------------------
// Batched countPointsIn / pointsIn: one wave per shape, lanes split a leaf.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_count_points_in(TreeView T, int kind, const double *params,
                                                        uint64_t nshapes, uint64_t *counts)
{
  uint64_t s = blockIdx.x;
  if (s >= nshapes) return;
  wb::Shape S = make_shape(kind, params + WB_SHAPE_NPARAMS * s);
  uint64_t cnt = 0;
  for_leaves(T, S, [&](int leaf) {
    const double *q = T.leaf_cube + 4 * leaf;
    uint32_t b = T.leaf_start[leaf], e = T.leaf_start[leaf + 1];
    if (wb::shape_in_cube(S, q[0], q[1], q[2], q[3])) {
      cnt += e - b;
      return false;
    }
    for (uint32_t j0 = b; j0 < e; j0 += WAVE) {
      uint32_t j = j0 + threadIdx.x;
      bool in = j < e && wb::shape_in(S, T.px[j], T.py[j], T.pz[j]);
      cnt += __popcll(__ballot(in));
    }
    return false;
  });
  if (threadIdx.x == 0) counts[s] = cnt;
}

__global__ void __launch_bounds__(64) k_points_in(TreeView T, int kind, const double *params,
                                                  uint64_t nshapes, const uint64_t *offs,
                                                  uint32_t *out)
{
  uint64_t s = blockIdx.x;
  if (s >= nshapes) return;
  wb::Shape S = make_shape(kind, params + WB_SHAPE_NPARAMS * s);
  uint64_t pos = offs[s];
  const uint64_t lanelt = (1ull << __lane_id()) - 1ull;
  for_leaves(T, S, [&](int leaf) {
    const double *q = T.leaf_cube + 4 * leaf;
    uint32_t b = T.leaf_start[leaf], e = T.leaf_start[leaf + 1];
    bool all = wb::shape_in_cube(S, q[0], q[1], q[2], q[3]);
    for (uint32_t j0 = b; j0 < e; j0 += WAVE) {
      uint32_t c = j0 + threadIdx.x;  // canonical position
      bool in = false;
      uint32_t slot = 0;
      if (c < e) {
        slot = T.canon[c];
        in = all || wb::shape_in(S, T.px[slot], T.py[slot], T.pz[slot]);
      }
      uint64_t bal = __ballot(in);
      if (in) out[pos + __popcll(bal & lanelt)] = T.sidx[slot];
      pos += __popcll(bal);
    }
    return false;
  });
}

// ---------------------------------------------------------------------------
// Batched countPointsIn / pointsIn with a lane per shape.  The shapes are
// taken in Morton order of their centres (k_query_keys + the radix sort), so
// the 64 shapes of a wave are neighbours and share most of their leaves.
// A wave walks the octree once for its 64 shapes: findBlocks
// (octree.cpp:234-251) with an explicit stack whose entries carry the mask of
// lanes for which the node and every ancestor intersect the lane's shape, so
// each lane sees exactly the leaves findBlocks returns for its own shape, in
// the same (DFS) order.  A leaf whose eight corners are all in a lane's shape
// counts whole for that lane (octree.cpp:1226-1231, 1256-1257); otherwise the
// lane tests the leaf's points, skipping 32-point groups whose bounding box
// no point of the shape can reach (wb::shape_may_hit_box: exact).  A group's
// points are read once per wave with one vector load and handed to every
// lane by readlane.
// ---------------------------------------------------------------------------
// pending entries at one time: at most 7 siblings per level on the path plus
// 8 children (internal nodes sit at levels 0..WB_MAX_LEVEL)
#define WB_QSTK (7 * (WB_MAX_LEVEL + 1) + 8 + 2)
struct QueryStack {
  int code[WB_QSTK];
  uint32_t lo[WB_QSTK], hi[WB_QSTK];
};

template <int KIND>
__device__ inline wb::Shape query_shape(const double *p)
{
  if (KIND == wb::SHAPE_CYLINDER) return wb::make_cylinder(p[0], p[1], p[2]);
  if (KIND == wb::SHAPE_SPHERE) return wb::make_sphere(p[0], p[1], p[2], p[3]);
  return wb::make_hyperboloid(p[0], p[1], p[2], p[3], p[4]);
}

// Octree::cube (octree.cpp:349-358) of child k of the cube (cx, cy, cz, side):
// the operations the build used for the stored cubes
__device__ inline void child_cube(double cx, double cy, double cz, double side, int k,
                                  double &x, double &y, double &z, double &sd)
{
  x = cx + (double)(2 * (k & 1) - 1) * side / 4;
  y = cy + (double)(2 * ((k >> 1) & 1) - 1) * side / 4;
  z = cz + (double)(2 * (k >> 2) - 1) * side / 4;
  sd = side / 2;
}

// The walk: leaf(l, on) for every leaf some lane reaches, in DFS order; `on`
// tells whether this lane's findBlocks returns it.
template <class Leaf>
__device__ inline void query_walk(const TreeView &T, const wb::Shape &S, bool active,
                                  QueryStack &K, Leaf &&leaf)
{
  if (T.nleaves == 0) return;
  const uint64_t m0 = __ballot(active && wb::shape_intersect(S, T.node_cube[0], T.node_cube[1],
                                                             T.node_cube[2], T.node_cube[3]));
  if (!m0) return;
  K.code[0] = 0;
  K.lo[0] = (uint32_t)m0;
  K.hi[0] = (uint32_t)(m0 >> 32);
  int sp = 1;
  const uint32_t lane = __lane_id();
  while (sp > 0) {
    sp--;
    const int code = rfl(K.code[sp]);
    const uint32_t mlo = (uint32_t)rfl((int)K.lo[sp]), mhi = (uint32_t)rfl((int)K.hi[sp]);
    const bool on = ((lane < 32 ? mlo >> lane : mhi >> (lane - 32)) & 1u) != 0;
    if (code < 0) {
      leaf(-2 - code, on);
      continue;
    }
    const double *nc = T.node_cube + 4 * (size_t)code;
    const double cx = nc[0], cy = nc[1], cz = nc[2], side = nc[3];
    const int *chp = T.node_child + 8 * (size_t)code;
    int ck[8];
    uint64_t mk[8];
    uint32_t pushed = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) ck[k] = rfl(chp[k]);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      mk[k] = 0;
      if (ck[k] != -1) {
        double x, y, z, sd;
        child_cube(cx, cy, cz, side, k, x, y, z, sd);
        mk[k] = __ballot(on && wb::shape_intersect(S, x, y, z, sd));
        if (mk[k]) pushed |= 1u << k;
      }
    }
    // child 0 on top: a child's entry sits below those of the smaller indices
#pragma unroll
    for (int k = 0; k < 8; k++) {
      if (mk[k]) {
        const int pos = sp + __popc(pushed >> (k + 1));
        K.code[pos] = ck[k];
        K.lo[pos] = (uint32_t)mk[k];
        K.hi[pos] = (uint32_t)(mk[k] >> 32);
      }
    }
    sp += __popc(pushed);
  }
}

// Morton key of a shape's centre in the tree's root cube (21 bits per axis,
// z, y, x interleaved as the octree's child index): the visiting order of the
// lane-per-shape kernels.  A cylinder has no z.
__device__ inline uint32_t query_cell(double d, double inv_cell)
{
  const double u = d * inv_cell;
  if (!(u >= 0.0)) return 0;
  if (u >= (double)(1u << WB_MAX_LEVEL)) return (1u << WB_MAX_LEVEL) - 1;
  return (uint32_t)u;
}

__device__ inline uint64_t spread3(uint32_t v)
{
  uint64_t x = v & 0x1fffffull;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__global__ void k_query_keys(const double *params, uint64_t ns, int kind, double lox, double loy,
                             double loz, double inv_cell, uint32_t *key, uint32_t *val)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ns) return;
  const double *p = params + WB_SHAPE_NPARAMS * i;
  const uint32_t cx = query_cell(p[0] - lox, inv_cell), cy = query_cell(p[1] - loy, inv_cell);
  const uint32_t cz = kind == wb::SHAPE_CYLINDER ? 0u : query_cell(p[2] - loz, inv_cell);
  // the top 32 of the 63 key bits (cells of side/1024 in x and y) give the
  // locality that matters
  key[i] = (uint32_t)((spread3(cx) | (spread3(cy) << 1) | (spread3(cz) << 2)) >> 31);
  val[i] = (uint32_t)i;
}

// lane o's value of v (o wave-uniform)
__device__ inline uint64_t rdl64(uint64_t v, uint32_t o)
{
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)o);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)o);
  return ((uint64_t)hi << 32) | lo;
}

// lane o's shape (o wave-uniform), read from the lane that holds it
__device__ inline wb::Shape lane_shape(const wb::Shape &S, uint32_t o)
{
  wb::Shape R = S;
  R.cx = rdl(S.cx, o);
  R.cy = rdl(S.cy, o);
  R.cz = rdl(S.cz, o);
  R.r = rdl(S.r, o);
  R.por2 = rdl(S.por2, o);
  R.slope = rdl(S.slope, o);
  return R;
}

// countPointsIn of shapes order[64w .. 64w+63] by wave w.  Inside a leaf,
// each 32-point group is loaded once (lanes 0-31 and 32-63 both hold it)
// and tested against the shapes of the lanes that may reach it ("owners"),
// two per step: lanes 0-31 against the first owner's shape, lanes 32-63
// against the second's, the two ballot halves counted for their owners.
template <int KIND>
__global__ void __launch_bounds__(256) k_query_count(TreeView T, const double *params,
                                                     const uint32_t *order, uint64_t ns,
                                                     uint64_t *counts)
{
  __shared__ QueryStack stk[4];
  QueryStack &K = stk[threadIdx.x / WAVE];
  const uint32_t lane = __lane_id();
  const uint64_t q = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool active = q < ns;
  const uint32_t si = active ? order[q] : 0u;
  const wb::Shape S = query_shape<KIND>(params + WB_SHAPE_NPARAMS * (uint64_t)si);
  uint64_t cnt = 0;
  const bool hi_half = lane >= WAVE / 2;
  query_walk(T, S, active, K, [&](int l, bool on) {
    const double *lc = T.leaf_cube + 4 * (size_t)l;
    const uint32_t b = T.leaf_start[l], e = T.leaf_start[l + 1];
    if (e <= b) return;
    const bool all = on && wb::shape_in_cube(S, lc[0], lc[1], lc[2], lc[3]);
    if (all) cnt += e - b;
    const bool rest = on && !all;
    if (!__ballot(rest)) return;
    const uint32_t g1 = (e - 1) / WB_CHUNK;
    for (uint32_t g = b / WB_CHUNK; g <= g1; g++) {
      const double *bx = T.chunk_box + 6 * (size_t)g;
      const bool hit = rest && wb::shape_may_hit_box(S, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]);
      uint64_t own = __ballot(hit);
      if (!own) continue;
      const uint32_t base = g * WB_CHUNK;
      const uint32_t j = base + (lane & (WB_CHUNK - 1));
      const bool valid = j >= b && j < e;
      double x = 0, y = 0, z = 0;
      if (valid) {
        x = T.px[j];
        y = T.py[j];
        z = T.pz[j];
      }
      while (own) {
        const uint32_t o1 = (uint32_t)__ffsll((long long)own) - 1;
        own &= own - 1;
        const bool two = own != 0;
        const uint32_t o2 = two ? (uint32_t)__ffsll((long long)own) - 1 : o1;
        own &= own - 1;
        const wb::Shape A = lane_shape(S, o1), B = lane_shape(S, o2);
        const wb::Shape &H = hi_half ? B : A;
        const uint64_t in = __ballot(valid && wb::shape_in(H, x, y, z));
        if (lane == o1) cnt += __popcll(in & 0xffffffffull);
        if (two && lane == o2) cnt += __popcll(in >> 32);
      }
    }
  });
  if (active) counts[si] = cnt;
}

// pointsIn of shapes order[64w .. 64w+63] by wave w: each shape's input
// indices from offs[shape] on, leaf by leaf in DFS order and in canonical
// (input) order inside a leaf, as k_points_in writes them.  A leaf is read
// in canonical order 64 points at a time; each such group is tested against
// the shape of every lane that takes the leaf (one owner per step, its
// indices written by consecutive lanes).
template <int KIND>
__global__ void __launch_bounds__(256) k_query_points(TreeView T, const double *params,
                                                      const uint32_t *order, uint64_t ns,
                                                      const uint64_t *offs, uint32_t *out)
{
  __shared__ QueryStack stk[4];
  QueryStack &K = stk[threadIdx.x / WAVE];
  const uint32_t lane = __lane_id();
  const uint64_t q = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const bool active = q < ns;
  const uint32_t si = active ? order[q] : 0u;
  const wb::Shape S = query_shape<KIND>(params + WB_SHAPE_NPARAMS * (uint64_t)si);
  // the writes stay inside the shape's own range whatever the walk finds
  uint64_t pos = active ? offs[si] : 0;
  const uint64_t lim = active ? offs[si + 1] : 0;
  const uint64_t lt = (1ull << lane) - 1ull;
  query_walk(T, S, active, K, [&](int l, bool on) {
    const double *lc = T.leaf_cube + 4 * (size_t)l;
    const uint32_t b = T.leaf_start[l], e = T.leaf_start[l + 1];
    if (e <= b) return;
    const bool all = on && wb::shape_in_cube(S, lc[0], lc[1], lc[2], lc[3]);
    const bool rest = on && !all;
    // bit g - g0: the lane's shape may reach 32-point group g (a leaf of at
    // most 537 points spans at most 18 groups; a longer one is not pruned)
    const uint32_t g0 = b / WB_CHUNK, g1 = (e - 1) / WB_CHUNK;
    const bool wide = g1 - g0 >= 32;
    uint32_t gm = 0;
    if (wide) {
      gm = rest ? ~0u : 0u;
    } else if (__ballot(rest)) {
      for (uint32_t g = g0; g <= g1; g++) {
        const double *bx = T.chunk_box + 6 * (size_t)g;
        if (rest && wb::shape_may_hit_box(S, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]))
          gm |= 1u << (g - g0);
      }
    }
    const uint64_t owners = __ballot(all || gm != 0);
    if (!owners) return;
    for (uint32_t c0 = b; c0 < e; c0 += WAVE) {
      const uint32_t c = c0 + lane;
      const bool valid = c < e;
      uint32_t slot = 0, id = 0;
      double x = 0, y = 0, z = 0;
      if (valid) {
        slot = T.canon[c];
        id = T.sidx[slot];
        x = T.px[slot];
        y = T.py[slot];
        z = T.pz[slot];
      }
      const uint32_t gb = valid && !wide ? slot / WB_CHUNK - g0 : 0;
      uint64_t own = owners;
      while (own) {
        const uint32_t o = (uint32_t)__ffsll((long long)own) - 1;
        own &= own - 1;
        const bool ao = __builtin_amdgcn_readlane((int)all, (int)o) != 0;
        const uint32_t go = (uint32_t)__builtin_amdgcn_readlane((int)gm, (int)o);
        bool in = valid && ao;
        if (!ao) {
          const wb::Shape H = lane_shape(S, o);
          in = valid && ((go >> gb) & 1u) && wb::shape_in(H, x, y, z);
        }
        const uint64_t bal = __ballot(in);
        if (!bal) continue;
        const uint64_t p0 = rdl64(pos, o), l0 = rdl64(lim, o);
        const uint64_t w = p0 + (uint64_t)__popcll(bal & lt);
        if (in && w < l0) out[w] = id;
        if (lane == o) pos = p0 + (uint64_t)__popcll(bal);
      }
    }
  });
}

// ---------------------------------------------------------------------------
// Tiles
// ---------------------------------------------------------------------------
struct TileGeo {
  double scx, scy, spacing, r;
  int lo, hi;
  uint32_t ntiles;
  // dense Eisenstein-address -> counter index map over the island's box
  int gx0, gy0, gw, gh;
  const int32_t *grid;
};

// counter of the tile at Eisenstein (ex, ey), false if it is not a tile
__device__ inline bool tile_at(const TileGeo &g, int ex, int ey, int *cnt)
{
  int gx = ex - g.gx0, gy = ey - g.gy0;
  if (gx < 0 || gy < 0 || gx >= g.gw || gy >= g.gh) return false;
  int k = g.grid[(int64_t)gy * g.gw + gx];
  if (k < 0) return false;
  *cnt = g.lo + k;
  return true;
}

// Flowsnake::cyl, flowsnake.cpp:263-271
__device__ inline void tile_center(const TileGeo &g, int ex, int ey, double *cx, double *cy)
{
  double zr = ex - ey / 2., zi = ey * kSqrt34;
  zr *= g.spacing;
  zi *= g.spacing;
  *cx = zr + g.scx;
  *cy = zi + g.scy;
}

__global__ void k_tile_geo(TileGeo g, int32_t *tex, int32_t *tey, int *bounds)
{
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  int ex = 0, ey = 0;
  bool ok = k < g.ntiles;
  if (ok) {
    wb::to_flowsnake_dev((int64_t)g.lo + k, &ex, &ey);
    tex[k] = ex;
    tey[k] = ey;
  }
  int x0 = ok ? ex : INT32_MAX, x1 = ok ? ex : INT32_MIN, y0 = ok ? ey : INT32_MAX, y1 = ok ? ey : INT32_MIN;
  for (int s = 1; s < 64; s <<= 1) {
    x0 = min(x0, __shfl_xor(x0, s, 64));
    x1 = max(x1, __shfl_xor(x1, s, 64));
    y0 = min(y0, __shfl_xor(y0, s, 64));
    y1 = max(y1, __shfl_xor(y1, s, 64));
  }
  // one atomic per block and bound
  __shared__ int wb4[4][256 / 64];
  const int w = threadIdx.x / 64;
  if (__lane_id() == 0) {
    wb4[0][w] = x0; wb4[1][w] = x1; wb4[2][w] = y0; wb4[3][w] = y1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)blockDim.x / 64; k++) {
      x0 = min(x0, wb4[0][k]); x1 = max(x1, wb4[1][k]);
      y0 = min(y0, wb4[2][k]); y1 = max(y1, wb4[3][k]);
    }
    atomicMin(&bounds[0], x0);
    atomicMax(&bounds[1], x1);
    atomicMin(&bounds[2], y0);
    atomicMax(&bounds[3], y1);
  }
}

__global__ void k_fill_i32(int32_t *p, uint64_t n, int32_t v)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void k_bounds_init(int *b)
{
  if (threadIdx.x == 0) {
    b[0] = INT32_MAX;
    b[1] = INT32_MIN;
    b[2] = INT32_MAX;
    b[3] = INT32_MIN;
  }
}

__global__ void k_tile_grid(TileGeo g, const int32_t *tex, const int32_t *tey, int32_t *grid)
{
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= g.ntiles) return;
  grid[(int64_t)(tey[k] - g.gy0) * g.gw + (tex[k] - g.gx0)] = (int32_t)k;
}

#define MAX_MEMB 4

// the lanes with `has` whose key equals this lane's (0 without `has`)
__device__ inline uint64_t wave_peers(bool has, int key)
{
  uint64_t todo = __ballot(has), peers = 0;
  while (todo) {
    const int kl = __builtin_amdgcn_readlane(key, __ffsll((long long)todo) - 1);
    const bool same = has && key == kl;
    const uint64_t m = __ballot(same);
    if (same) peers = m;
    todo &= ~m;
  }
  return peers;
}

// cnt[key] += 1 for every active lane with `has`: the peer masks first, then
// one atomic per distinct key from its first lane, none of them waited for
// (no slot is returned)
__device__ inline void wave_count_inc(uint32_t *cnt, bool has, int key)
{
  const uint64_t peers = wave_peers(has, key);
  if (has && __ffsll((long long)peers) - 1 == (int)__lane_id())
    atomicAdd(&cnt[key], (uint32_t)__popcll(peers));
}

// Slot claims of a wave, K keys per lane: for each k, one atomic on
// cnt[key[k]] per distinct key from its first lane, each lane's base read
// from that lane, lanes with the same key given consecutive slots in lane
// order (slot[k], where has[k]).  The atomics of all K keys are issued
// before any is waited for: one round trip per wave.
template <int K>
__device__ inline void wave_claim_n(uint32_t *cnt, const bool (&has)[K], const int (&key)[K],
                                    uint32_t (&slot)[K])
{
  const uint32_t lane = __lane_id();
  uint64_t peers[K];
  int first[K];
  uint32_t base[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    peers[k] = wave_peers(has[k], key[k]);
    first[k] = has[k] ? __ffsll((long long)peers[k]) - 1 : (int)lane;
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    base[k] = 0;
    if (has[k] && first[k] == (int)lane)
      base[k] = atomicAdd(&cnt[key[k]], (uint32_t)__popcll(peers[k]));
  }
#pragma unroll
  for (int k = 0; k < K; k++)
    slot[k] = (uint32_t)__shfl((int)base[k], first[k], 64) +
              (uint32_t)__popcll(peers[k] & ((1ull << lane) - 1));
}

// The (leaf, tile) part of the membership test, once per leaf for the
// lattice cells near it: a window of cells (ex, ey), ex in [x0, x0+w), ey in
// [y0, y0+h), w*h <= LEAF_WIN, covering every cell whose centre can lie
// within the candidate bound of a point inside the leaf's square (one cell
// of slack on each side).  Flag bit 0: a tile whose cylinder findBlocks
// reaches this leaf through (every ancestor cube and the leaf cube intersect
// it); bit 1: all eight leaf corners in the cylinder.  w = 0: the leaf's
// window is too large and its points take the full test.
#define LEAF_WIN 256
struct LeafWin {
  int x0, y0, w, h;
};

__global__ void __launch_bounds__(64) k_leaf_tiles(TreeView T, TileGeo g, LeafWin *win,
                                                   uint8_t *flags)
{
  const uint32_t l = blockIdx.x;
  if (l >= T.nleaves) return;
  const double *lc = T.leaf_cube + 4 * (size_t)l;
  const double h = lc[3] / 2, lim = g.r * (1 + 1e-9) + 1e-9;
  const double fy0 = ((lc[1] - h - lim) - g.scy) / g.spacing / kSqrt34;
  const double fy1 = ((lc[1] + h + lim) - g.scy) / g.spacing / kSqrt34;
  const double u0 = ((lc[0] - h - lim) - g.scx) / g.spacing;
  const double u1 = ((lc[0] + h + lim) - g.scx) / g.spacing;
  LeafWin wl = {0, 0, 0, 0};
  if (fabs(fy0) < 1e9 && fabs(fy1) < 1e9 && fabs(u0) < 1e9 && fabs(u1) < 1e9) {
    wl.y0 = (int)floor(fy0) - 1;
    const int y1 = (int)ceil(fy1) + 1;
    wl.x0 = (int)floor(u0 + wl.y0 / 2.0) - 1;
    const int x1 = (int)ceil(u1 + y1 / 2.0) + 1;
    wl.w = x1 - wl.x0 + 1;
    wl.h = y1 - wl.y0 + 1;
    if (wl.w <= 0 || wl.h <= 0 || wl.w * wl.h > LEAF_WIN) wl.w = wl.h = 0;
  }
  if (threadIdx.x == 0) win[l] = wl;
  const int ncell = wl.w * wl.h;
  for (int c = threadIdx.x; c < ncell; c += 64) {
    const int ex = wl.x0 + c % wl.w, ey = wl.y0 + c / wl.w;
    uint8_t f = 0;
    int cnt;
    if (tile_at(g, ex, ey, &cnt)) {
      double cx, cy;
      tile_center(g, ex, ey, &cx, &cy);
      const wb::Shape cyl = wb::make_cylinder(cx, cy, g.r);
      bool reach = wb::shape_intersect(cyl, lc[0], lc[1], lc[2], lc[3]);
      for (int nd = T.leaf_parent[l]; reach && nd >= 0; nd = T.node_parent[nd]) {
        const double *q = T.node_cube + 4 * nd;
        reach = wb::shape_intersect(cyl, q[0], q[1], q[2], q[3]);
      }
      if (reach) f = 1 | (wb::shape_in_cube(cyl, lc[0], lc[1], lc[2], lc[3]) ? 2 : 0);
    }
    flags[(size_t)l * LEAF_WIN + c] = f;
  }
}

// pointsIn(cylinder) membership of every point, evaluated from the point's
// side: the candidates are the lattice tiles near it; a tile holds the point
// iff findBlocks reaches its leaf (every ancestor cube and the leaf cube
// intersect the cylinder) and either all eight leaf corners are in or the
// point itself is.  The leaf's part comes from its k_leaf_tiles window;
// candidates outside the window (points outside their leaf's cube) take the
// whole test.
__global__ void k_membership(TreeView T, TileGeo g, const LeafWin *win, const uint8_t *flags,
                             int32_t *memb, uint32_t *nmemb, uint32_t *tile_cnt, uint32_t *err)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= T.n) return;
  const double x = T.px[i], y = T.py[i];
  const uint32_t leaf = T.pt_leaf[i];
  const double *lc = T.leaf_cube + 4 * leaf;
  const LeafWin wl = win[leaf];
  // The lattice point nearest to p (Eisenstein coordinates a + b*omega,
  // centre spacing*(a - b/2, b*sqrt3/2)) by rounding the cube coordinates
  // (a - b, b, -a) and fixing the one rounded farthest.  A centre within the
  // candidate bound (spacing*41/71, a hair over the hexagon's circumradius
  // spacing/sqrt3) of p lies within 1.16 spacings of that point, so it is the
  // point itself or one of its six neighbours (the next ring is sqrt3
  // spacings away); a rounding near a tie that picks a neighbour of the
  // nearest point still leaves every such centre in the seven.
  const double u = (x - g.scx) / g.spacing, v = (y - g.scy) / g.spacing;
  const double fb = v / kSqrt34, fa = u + fb / 2;
  double rx = __builtin_rint(fa - fb), ry = __builtin_rint(fb), rz = __builtin_rint(-fa);
  const double ddx = fabs(rx - (fa - fb)), ddy = fabs(ry - fb), ddz = fabs(rz + fa);
  if (ddy > ddx && ddy > ddz) ry = -rx - rz;
  else if (ddz >= ddx && ddz >= ddy) rz = -rx - ry;
  const int X = (int)-rz, Y = (int)ry;
  const double lim = g.r * (1 + 1e-9) + 1e-9;
  int nm = 0;
  int32_t mem[MAX_MEMB];
  constexpr int na[7] = {0, 1, -1, 0, 0, 1, -1}, nb[7] = {0, 0, 0, 1, -1, 1, -1};
#pragma unroll
  for (int k = 0; k < 7; k++) {
      int ex = X + na[k], ey = Y + nb[k];
      double cx, cy;
      tile_center(g, ex, ey, &cx, &cy);
      double dx = cx - x, dy = cy - y;
      if (dx * dx + dy * dy > lim * lim) continue;
      int cnt;
      if (!tile_at(g, ex, ey, &cnt)) continue;
      wb::Shape cyl = wb::make_cylinder(cx, cy, g.r);
      const int gx = ex - wl.x0, gy = ey - wl.y0;
      if (gx >= 0 && gy >= 0 && gx < wl.w && gy < wl.h) {
        const uint8_t f = flags[(size_t)leaf * LEAF_WIN + gy * wl.w + gx];
        if (!(f & 1)) continue;
        if (!(f & 2) && !wb::cyl_in(cyl, x, y)) continue;
      } else {
        bool in = wb::shape_in_cube(cyl, lc[0], lc[1], lc[2], lc[3]) || wb::cyl_in(cyl, x, y);
        if (!in) continue;
        if (!wb::shape_intersect(cyl, lc[0], lc[1], lc[2], lc[3])) continue;
        bool reach = true;
        for (int nd = T.leaf_parent[leaf]; nd >= 0; nd = T.node_parent[nd]) {
          const double *q = T.node_cube + 4 * nd;
          if (!wb::shape_intersect(cyl, q[0], q[1], q[2], q[3])) { reach = false; break; }
        }
        if (!reach) continue;
      }
      if (nm < MAX_MEMB) mem[nm] = cnt;
      nm++;
    }
  if (nm > MAX_MEMB) {
    atomicOr(err, 1u);
    nm = MAX_MEMB;
  }
  for (int k = 0; k < MAX_MEMB; k++) memb[MAX_MEMB * i + k] = k < nm ? mem[k] : INT32_MIN;
  nmemb[i] = nm;
#pragma unroll
  for (int k = 0; k < MAX_MEMB; k++) wave_count_inc(tile_cnt, k < nm, k < nm ? mem[k] - g.lo : 0);
}

__global__ void k_tile_lists(uint64_t n, const int32_t *memb, const uint32_t *nmemb, int lo,
                             const uint32_t *crank, const uint32_t *tile_off, uint32_t *tile_fill,
                             uint32_t *lists)
{
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t nm = nmemb[i];
  bool has[MAX_MEMB];
  int t[MAX_MEMB];
  uint32_t slot[MAX_MEMB];
#pragma unroll
  for (uint32_t k = 0; k < MAX_MEMB; k++) {
    has[k] = k < nm;
    t[k] = has[k] ? memb[MAX_MEMB * i + k] - lo : 0;
  }
  // the slot claims of all memberships in one round of atomics
  wave_claim_n<MAX_MEMB>(tile_fill, has, t, slot);
  const uint32_t cr = crank[i];
#pragma unroll
  for (uint32_t k = 0; k < MAX_MEMB; k++)
    if (has[k]) lists[tile_off[t[k]] + slot[k]] = cr;
}

// ---------------------------------------------------------------------------
// SCAN: one wave per tile (scanCylinder, scan.cpp:31-140)
// ---------------------------------------------------------------------------
#define SCAN_CAP 1024

// pairwisesum (manysum.cpp:120-154) of m <= SCAN_CAP values produced by
// val(j): one balanced tree per set bit of m, over the aligned block that
// bit owns (largest block first), block sums added to 0 in ascending bit
// order.  Balanced trees of up to 64 values are xor butterflies across the
// wave; larger blocks combine their 64-value chunk sums the same way.
template <class V>
__device__ inline double wave_pairwise(uint32_t m, V &&val)
{
  const int lane = __lane_id();
  double acc = 0;
  for (int k = 0; k < 11; k++) {
    if (!((m >> k) & 1u)) continue;
    uint32_t off = (m >> (k + 1)) << (k + 1);
    uint32_t sz = 1u << k;
    double blk;
    if (sz <= 64) {
      double v = (uint32_t)lane < sz ? val(off + lane) : 0.0;
      for (uint32_t s = 1; s < sz; s <<= 1) v = v + __shfl_xor(v, (int)s, WAVE);
      blk = __shfl(v, 0, WAVE);
    } else {
      // chunk sums of 64, then a balanced tree over the chunks
      uint32_t nch = sz / 64;
      double mine = 0;
      for (uint32_t ch = 0; ch < nch; ch++) {
        double v = val(off + ch * 64 + lane);
        for (int s = 1; s < 64; s <<= 1) v = v + __shfl_xor(v, s, WAVE);
        if ((uint32_t)lane == ch) mine = v;
      }
      for (uint32_t s = 1; s < nch; s <<= 1) mine = mine + __shfl_xor(mine, (int)s, WAVE);
      blk = __shfl(mine, 0, WAVE);
    }
    acc += blk;
  }
  return acc;
}

__device__ inline double pw3(double a0, double a1, double a2, int n)
{
  // pairwisesum of n <= 3 values
  if (n == 1) return 0 + a0;
  if (n == 2) return 0 + (a0 + a1);
  return (0 + a2) + (a0 + a1);
}

// 3x3 Gauss-Jordan of matrix.cpp:262-422 on a (3x3, row-major) and b (3x1).
// Every row and column index is a compile-time constant, so the system stays
// in registers; the pivot choices made at run time select among them.
// rows R0 and R1 exchanged when `c`, by selects (no indexed access)
template <int R0, int R1>
__device__ __attribute__((always_inline)) inline void swap_rows_s(double (&a)[9], double (&b)[3],
                                                                  bool c)
{
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double x = a[3 * R0 + k], y = a[3 * R1 + k];
    a[3 * R0 + k] = c ? y : x;
    a[3 * R1 + k] = c ? x : y;
  }
  const double x = b[R0], y = b[R1];
  b[R0] = c ? y : x;
  b[R1] = c ? x : y;
}

__device__ inline double sel3(double r0, double r1, double r2, int k)
{
  return k == 0 ? r0 : (k == 1 ? r1 : r2);
}

// rowop(row0, row1, piv) of matrix.cpp
template <int ROW0, int ROW1, int PIV>
__device__ __attribute__((always_inline)) inline void rowop_s(double (&a)[9], double (&b)[3])
{
  double slope = 0, minslope = __builtin_inf(), det;
  int flags = 0, pivot;
  bool piv_ok = true;  // piv >= 0
  if (a[3 * ROW0 + PIV] == 0 && a[3 * ROW1 + PIV] == 0) piv_ok = false;
  pivot = piv_ok ? PIV : -1;
  if (piv_ok && a[3 * ROW0 + PIV] == 0) flags = 9;
  if (!piv_ok) {
#pragma unroll
    for (int i = 0; i < 3; i++) {
      const double x0 = a[3 * ROW0 + i], x1 = a[3 * ROW1 + i];
      if (x0 != 0 || x1 != 0) {
        if (wb::dabs(x0) > wb::dabs(x1) || ROW0 >= ROW1) {
          slope = wb::dabs(x1 / x0);
          flags &= ~8;
        } else {
          slope = wb::dabs(x0 / x1);
          flags |= 8;
        }
        if (slope < minslope) {
          minslope = slope;
          flags = (flags >> 3) * 9;
          pivot = i;
        }
      }
    }
  }
  flags &= 1;
  swap_rows_s<ROW0, ROW1>(a, b, flags != 0);
  det = (pivot < 0) ? 0 : sel3(a[3 * ROW0], a[3 * ROW0 + 1], a[3 * ROW0 + 2], pivot);
  if (det != 0 && det != 1) {
#pragma unroll
    for (int i = 0; i < 3; i++) a[3 * ROW0 + i] /= det;
    b[ROW0] /= det;
  }
  if (pivot >= 0) slope = sel3(a[3 * ROW1], a[3 * ROW1 + 1], a[3 * ROW1 + 2], pivot);
  if (slope != 0 && ROW0 != ROW1) {
#pragma unroll
    for (int i = 0; i < 3; i++) a[3 * ROW1 + i] -= a[3 * ROW0 + i] * slope;
    b[ROW1] -= b[ROW0] * slope;
  }
}

// findpivot(row, column) of matrix.cpp
template <int ROW>
__device__ __attribute__((always_inline)) inline void findpivot_s(double (&a)[9], double (&b)[3])
{
  int pivotrow = -1;
  double maxratio = 0;
  // the column loop runs while pivotrow < ROW; pivotrow, once set, is >= ROW
#pragma unroll
  for (int column = ROW; column < 3; column++) {
    if (pivotrow < ROW) {
#pragma unroll
      for (int i = ROW; i < 3; i++) {
        const double s0 = column + 1 < 3 ? a[3 * i + column + 1] * a[3 * i + column + 1] : 0.0;
        const double s1 = column + 2 < 3 ? a[3 * i + column + 2] * a[3 * i + column + 2] : 0.0;
        const double piv = a[3 * i + column];
        const double ratio = (piv * piv) / pw3(s0, s1, 0.0, 3 - column);
        if (ratio > maxratio) {
          pivotrow = i;
          maxratio = ratio;
        }
      }
    }
  }
  if constexpr (ROW < 1) swap_rows_s<1, ROW>(a, b, pivotrow == 1);
  if constexpr (ROW < 2) swap_rows_s<2, ROW>(a, b, pivotrow == 2);
}

__device__ __attribute__((always_inline)) inline void gausselim_s(double (&a)[9], double (&b)[3])
{
  findpivot_s<0>(a, b);
  rowop_s<0, 0, 0>(a, b); rowop_s<0, 1, 0>(a, b); rowop_s<0, 2, 0>(a, b);
  findpivot_s<1>(a, b);
  rowop_s<1, 0, 1>(a, b); rowop_s<1, 1, 1>(a, b); rowop_s<1, 2, 1>(a, b);
  findpivot_s<2>(a, b);
  rowop_s<2, 0, 2>(a, b); rowop_s<2, 1, 2>(a, b); rowop_s<2, 2, 2>(a, b);
  // back substitution
  rowop_s<2, 0, 2>(a, b); rowop_s<2, 1, 2>(a, b);
  rowop_s<1, 0, 1>(a, b);
}

// the tilt (sx, sy) of the least-squares plane z = sx*x + sy*y + c through
// the normal equations (scan.cpp:53-73, leastsquares.cpp); every lane holds the same sums and
// gets the same answer
__device__ __attribute__((always_inline)) inline void plane_tilt(double sxx, double syx, double syy, double s1x, double s1y,
                                  double s11, double sxz, double syz, double s1z, double *sx,
                                  double *sy)
{
  double a[9] = {sxx, syx, s1x, syx, syy, s1y, s1x, s1y, s11};
  double b[3] = {sxz, syz, s1z};
  gausselim_s(a, b);
  if (a[0] == 0) b[0] = __builtin_nan("");
  if (a[4] == 0) b[1] = __builtin_nan("");
  if (a[8] == 0) b[2] = __builtin_nan("");
  double p = b[0], q = b[1];
  const double len = wb::hypot(p, q);
  if (len > 1) { p /= len; q /= len; }
  if (wb::disnan(p) || wb::disnan(q)) { p = 0; q = 0; }
  *sx = p;
  *sy = q;
}

// Slot s of a K-slot per-lane array, s wave-uniform.
template <int K, class V>
__device__ inline V slot_sel(const V (&v)[K], uint32_t s)
{
  V r = v[0];
#pragma unroll
  for (int k = 1; k < K; k++)
    if (s == (uint32_t)k) r = v[k];
  return r;
}

// v moved across the lanes of each row of 16 by the DPP control CTRL (every
// lane active; the controls used here read inside the row)
template <int CTRL>
__device__ inline double dpp_d(double v)
{
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xf, 0xf, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xf, 0xf,
                                                            false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Pairwise sum of lanes base .. base+sz-1 (sz a wave-uniform power of two up
// to 64, base a multiple of sz): lanes 2i and 2i+1 are added, then pairs of
// those sums, and so on.  Each level inside a row pairs every lane with a
// lane of the sibling group (quad_perm 1,0,3,2 and 2,3,0,1, then the
// half-row and row mirrors), whose lanes all hold that group's sum, so
// every addition has the operands of the xor butterfly; the row sums are
// then read and added in the same tree.
__device__ inline double lanes_pairwise(double b, uint32_t sz, uint32_t base)
{
  if (sz >= 2) b = b + dpp_d<0xb1>(b);
  if (sz >= 4) b = b + dpp_d<0x4e>(b);
  if (sz >= 8) b = b + dpp_d<0x141>(b);
  if (sz >= 16) b = b + dpp_d<0x140>(b);
  if (sz == 64) return (rdl(b, 0) + rdl(b, 16)) + (rdl(b, 32) + rdl(b, 48));
  if (sz == 32) return rdl(b, base) + rdl(b, base + 16);
  return rdl(b, base);
}

// minimum and maximum over the wave: the row steps by DPP, the four row
// results read
__device__ inline double wave_dmin(double v)
{
  v = fmin(v, dpp_d<0xb1>(v));
  v = fmin(v, dpp_d<0x4e>(v));
  v = fmin(v, dpp_d<0x141>(v));
  v = fmin(v, dpp_d<0x140>(v));
  return fmin(fmin(rdl(v, 0), rdl(v, 16)), fmin(rdl(v, 32), rdl(v, 48)));
}
__device__ inline double wave_dmax(double v)
{
  v = fmax(v, dpp_d<0xb1>(v));
  v = fmax(v, dpp_d<0x4e>(v));
  v = fmax(v, dpp_d<0x141>(v));
  v = fmax(v, dpp_d<0x140>(v));
  return fmax(fmax(rdl(v, 0), rdl(v, 16)), fmax(rdl(v, 32), rdl(v, 48)));
}

// wave_pairwise for m <= 64*K values held K per lane (value e in slot e/64
// of lane e%64): the same block trees and block order.  A block of up to 64
// values lies in one slot, aligned to its size; a larger one is the
// balanced tree over its whole slots.
template <int K>
__device__ inline double wave_pairwise_reg(uint32_t m, const double (&v)[K])
{
  static_assert(K <= 4, "blocks of up to 256 values");
  double acc = 0;
  for (int k = 0; k < 9; k++) {
    if (!((m >> k) & 1u)) continue;
    const uint32_t off = (m >> (k + 1)) << (k + 1);
    const uint32_t sz = 1u << k;
    double blk;
    if (sz <= WAVE) {
      blk = lanes_pairwise(slot_sel<K>(v, off / WAVE), sz, off & (WAVE - 1));
    } else {
      const uint32_t s0 = off / WAVE;
      const double c0 = lanes_pairwise(slot_sel<K>(v, s0), WAVE, 0);
      const double c1 = lanes_pairwise(slot_sel<K>(v, s0 + 1), WAVE, 0);
      if (sz == 2 * WAVE) {
        blk = c0 + c1;
      } else {
        const double c2 = lanes_pairwise(slot_sel<K>(v, s0 + 2), WAVE, 0);
        const double c3 = lanes_pairwise(slot_sel<K>(v, s0 + 3), WAVE, 0);
        blk = (c0 + c1) + (c2 + c3);
      }
    }
    acc += blk;
  }
  return acc;
}

// Sum over the wave of K float slots per lane: plain (naive), or with each
// lane's slots added by Neumaier's compensated sum and the lanes' (sum,
// error) pairs merged by Knuth's TwoSum, so the rounding errors of every
// addition are carried and added once at the end.
template <int K>
__device__ inline float wave_fsum(const float (&v)[K], bool comp)
{
  float s = 0, e = 0;
#pragma unroll
  for (int k = 0; k < K; k++) {
    const float t = s + v[k];
    if (comp) e += wb::dabs(s) >= wb::dabs(v[k]) ? (s - t) + v[k] : (v[k] - t) + s;
    s = t;
  }
  for (int o = 1; o < WAVE; o <<= 1) {
    const float s2 = __shfl_xor(s, o, WAVE), e2 = __shfl_xor(e, o, WAVE);
    const float t = s + s2;
    if (comp) {
      const float bp = t - s;
      e = e + e2 + ((s - (t - bp)) + (s2 - bp));
    }
    s = t;
  }
  return s + e;
}

// The plane tilt from FP32 moments of (x, y, z - z0) (a shift of z leaves the
// tilt unchanged), solved through the centred 2x2 normal equations in FP32.
// Returns false when the xy spread is too close to collinear for FP32
// (det < 1e-3 trace^2 of the centred xy moments, or fewer than 3 points); the
// caller then fits in FP64.
template <int K>
__device__ inline bool plane_tilt_f32(uint32_t m, const double (&vx)[K], const double (&vy)[K],
                                      const double (&vz)[K], const bool (&mine)[K], bool comp,
                                      double *sx, double *sy)
{
  const double z0 = rdl(vz[0], 0);
  float fx[K], fy[K], fz[K], t[K];
#pragma unroll
  for (int s = 0; s < K; s++) {
    fx[s] = mine[s] ? (float)vx[s] : 0.0f;
    fy[s] = mine[s] ? (float)vy[s] : 0.0f;
    fz[s] = mine[s] ? (float)(vz[s] - z0) : 0.0f;
  }
  const float Sx = wave_fsum<K>(fx, comp), Sy = wave_fsum<K>(fy, comp);
  const float Sz = wave_fsum<K>(fz, comp);
#pragma unroll
  for (int s = 0; s < K; s++) t[s] = fx[s] * fx[s];
  const float Sxx = wave_fsum<K>(t, comp);
#pragma unroll
  for (int s = 0; s < K; s++) t[s] = fx[s] * fy[s];
  const float Sxy = wave_fsum<K>(t, comp);
#pragma unroll
  for (int s = 0; s < K; s++) t[s] = fy[s] * fy[s];
  const float Syy = wave_fsum<K>(t, comp);
#pragma unroll
  for (int s = 0; s < K; s++) t[s] = fx[s] * fz[s];
  const float Sxz = wave_fsum<K>(t, comp);
#pragma unroll
  for (int s = 0; s < K; s++) t[s] = fy[s] * fz[s];
  const float Syz = wave_fsum<K>(t, comp);
  const float n = (float)m;
  const float mx = Sx / n, my = Sy / n, mz = Sz / n;
  const float cxx = Sxx - Sx * mx, cyy = Syy - Sy * my, cxy = Sxy - Sx * my;
  const float cxz = Sxz - Sx * mz, cyz = Syz - Sy * mz;
  const float det = cxx * cyy - cxy * cxy, tr = cxx + cyy;
  const bool well = m >= 3 && det > 1e-3f * tr * tr;
  double p = (double)((cxz * cyy - cyz * cxy) / det), q = (double)((cyz * cxx - cxz * cxy) / det);
  const double len = wb::hypot(p, q);
  if (len > 1) { p /= len; q /= len; }
  if (wb::disnan(p) || wb::disnan(q)) { p = 0; q = 0; }
  *sx = p;
  *sy = q;
  return well;
}

// The scan of a tile of at most 64*K points, K per lane (point e in slot
// e/64 of lane e%64): the same operations and summation order as the
// general path of k_scan, with the canonical sort done by a bitonic network
// over lanes and slots and every point read once.
template <int K, bool F32>
__device__ __attribute__((always_inline)) inline void scan_reg(const TreeView &T, const TileGeo &g,
                                                                int ex, int ey, uint32_t off,
                         uint32_t m, const uint32_t *lists, double min_hs, int prec,
                         TileRec *out, unsigned long long *ctr)
{
  const uint32_t lane = __lane_id();
  uint32_t c[K];
#pragma unroll
  for (int s = 0; s < K; s++) {
    const uint32_t e = s * WAVE + lane;
    c[s] = e < m ? lists[off + e] : 0xffffffffu;
  }
  {
    // canonical order: the m ranks ascending (pads last), sorted only as far
    // as the next power of two
    uint32_t unused[K];
    wbw::wave_sort<K, false>(c, unused, wbw::pow2_ceil(m));
  }
  double cx, cy;
  tile_center(g, ex, ey, &cx, &cy);
  const double r = g.r;
  double vx[K], vy[K], vz[K];
  bool mine[K];
#pragma unroll
  for (int s = 0; s < K; s++) {
    mine[s] = s * WAVE + lane < m;
    vx[s] = 0; vy[s] = 0; vz[s] = 0;
    if (mine[s]) {
      const uint32_t q = T.canon[c[s]];
      vx[s] = T.px[q] - cx;
      vy[s] = T.py[q] - cy;
      vz[s] = T.pz[q] - 0;
    }
  }
  double sx = 0, sy = 0;
  // FP32 fits are an option for the precision study; FP64 is the reference
  bool fit64 = !F32 || prec == WB_FIT_FP64;
  int fell_back = 0;
  if (F32 && !fit64) {
    const bool well = plane_tilt_f32<K>(m, vx, vy, vz, mine, prec == WB_FIT_FP32_COMP, &sx, &sy);
    if (!well && prec == WB_FIT_FP32_COMP) fit64 = true, fell_back = 1;
  }
  if (fit64) {
    double t[K];
#define WB_SUM(expr) (([&]() { _Pragma("unroll") for (int s = 0; s < K; s++) t[s] = (expr); }()), \
                      wave_pairwise_reg<K>(m, t))
    const double sxx = WB_SUM(vx[s] * vx[s]);
    const double syx = WB_SUM(vy[s] * vx[s]);
    const double syy = WB_SUM(vy[s] * vy[s]);
    const double s1x = WB_SUM(1.0 * vx[s]);
    const double s1y = WB_SUM(1.0 * vy[s]);
    const double s11 = WB_SUM(1.0 * 1.0);
    const double sxz = WB_SUM(vx[s] * vz[s]);
    const double syz = WB_SUM(vy[s] * vz[s]);
    const double s1z = WB_SUM(1.0 * vz[s]);
#undef WB_SUM
    plane_tilt(sxx, syx, syy, s1x, s1y, s11, sxz, syz, s1z, &sx, &sy);
  }
  double u[K];
  double bottom = __builtin_inf(), top = -__builtin_inf();
#pragma unroll
  for (int s = 0; s < K; s++) {
    u[s] = vz[s] - (sy * vy[s] + sx * vx[s]);
    if (mine[s]) {
      bottom = fmin(bottom, u[s]);
      top = fmax(top, u[s]);
    }
  }
  bottom = wave_dmin(bottom);
  top = wave_dmax(top);
  uint32_t first = 0xffffffffu;
#pragma unroll
  for (int s = K - 1; s >= 0; s--) {
    const uint64_t atb = __ballot(mine[s] && u[s] == bottom);
    if (atb) first = s * WAVE + (uint32_t)(__ffsll((long long)atb) - 1);
  }
  double bottom2 = __builtin_inf();
#pragma unroll
  for (int s = 0; s < K; s++)
    if (mine[s] && s * WAVE + lane < first) bottom2 = fmin(bottom2, u[s]);
  bottom2 = wave_dmin(bottom2);
  if (wb::disinf(bottom2)) bottom2 = bottom;
  // 7-sector histogram of the bottom layer
  uint32_t nbottom = 0, nedge = 0, hist[7] = {0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < K; s++) {
    int sec = -1;
    bool edge = false;
    if (mine[s] && u[s] < bottom2 + 2 * r) {
      int e = 0;
      sec = wb::scan_sector(vx[s], vy[s], r, &e);
      edge = e != 0;
    }
    nbottom += (uint32_t)__popcll(__ballot(sec >= 0));
    nedge += (uint32_t)__popcll(__ballot(edge));
    for (int k = 0; k < 7; k++) hist[k] += (uint32_t)__popcll(__ballot(sec == k));
  }
  if (lane == 0) {
    if (nedge) atomicAdd(&ctr[0], (unsigned long long)nedge);
    double density = 0;
    int tree = 0;
    for (int i = 0; i < 7; i++) density += (int)hist[i] * (int)hist[i];
    if (m > nbottom && density < 7) tree = 1;
    density = sqrt(density) * kSqrt7 / (r * r) / M_PI;
    if (m > nbottom && density < 0.5) tree = 1;
    if (top - bottom > 1.5) tree = 1;
    TileRec rr;
    rr.npoints = (int32_t)m;
    rr.nground = 0;
    rr.tree_flags = tree;
    rr.fit_fp64 = fell_back;
    rr.density = density;
    rr.hs_scan = sqrt(1 / density + min_hs * min_hs);
    rr.hs = rr.hs_scan;
    rr.height = top - bottom;
    rr.tilt_x = sx;
    rr.tilt_y = sy;
    *out = rr;
    if (fell_back) atomicAdd(&ctr[7], 1ull);
  }
}

// Size classes of the scan: 0 = 1..16 points, 1 = 17-32, 2 = 33-64,
// 3 = 65-128, 4 = 129-256, 5 = more.  One thread per tile: an empty tile gets
// its record here, every other tile goes on its class's list (classes 0 and
// 5 from the front and the back of la, 2 and 4 of lb, 1 and 3 of lc; ncls[c]
// the lengths), so each class kernel visits only its own tiles.  The order
// inside a list does not matter: every tile's record depends on that tile
// alone.
#define SCAN_NCLS 6
__device__ inline int scan_class(uint32_t m)
{
  return m == 0 ? -1 : m <= 16 ? 0 : m <= 32 ? 1 : m <= WAVE ? 2 : m <= 2 * WAVE ? 3 : m <= 4 * WAVE ? 4 : 5;
}

__global__ void __launch_bounds__(1024) k_scan_classes(uint32_t ntiles, const uint32_t *tile_cnt,
                                                      uint32_t *la, uint32_t *lb, uint32_t *lc,
                                                      uint32_t *ncls, TileRec *rec)
{
  // slots are claimed per block: the waves' class counts summed in LDS, one
  // atomic per class and block (atomics on the list counters are
  // serialised, so one per wave would cost more than the pass)
  __shared__ uint32_t wcnt[SCAN_NCLS][1024 / WAVE];
  __shared__ uint32_t bbase[SCAN_NCLS];
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t m = t < ntiles ? tile_cnt[t] : 0u;
  if (t < ntiles && m == 0) {
    TileRec r = {0, 0, 0, 0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    rec[t] = r;
  }
  const int c = scan_class(m);
  const uint32_t lane = __lane_id(), w = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
  uint64_t mk[SCAN_NCLS];
#pragma unroll
  for (int k = 0; k < SCAN_NCLS; k++) {
    mk[k] = __ballot(c == k);
    if (lane == 0) wcnt[k][w] = (uint32_t)__popcll(mk[k]);
  }
  __syncthreads();
  if (threadIdx.x < SCAN_NCLS) {
    const int k = threadIdx.x;
    uint32_t tot = 0;
    for (uint32_t v = 0; v < nw; v++) {
      const uint32_t x = wcnt[k][v];
      wcnt[k][v] = tot;  // exclusive prefix over the block's waves
      tot += x;
    }
    bbase[k] = tot ? atomicAdd(&ncls[k], tot) : 0u;
  }
  __syncthreads();
  if (c < 0) return;
  const uint32_t pos = bbase[c] + wcnt[c][w] + (uint32_t)__popcll(mk[c] & ((1ull << lane) - 1));
  switch (c) {
  case 0: la[pos] = t; break;
  case 1: lc[pos] = t; break;
  case 2: lb[pos] = t; break;
  case 3: lc[ntiles - 1 - pos] = t; break;
  case 4: lb[ntiles - 1 - pos] = t; break;
  default: la[ntiles - 1 - pos] = t; break;
  }
}

// A class list: entry i at list[step * i] (step +1 for a list filled from
// the front of its buffer, -1 from the back), n entries.
struct ClassList {
  const uint32_t *list;
  int32_t step;
  const uint32_t *n;
  __device__ uint32_t at(uint32_t i) const { return list[(int64_t)step * i]; }
};

// One wave per tile, over the tiles of one size class: up to 64, 65-128 and
// 129-256 points (K = 1, 2, 4 points per lane).  Each class has its own
// register budget.  The next tile's number is read before this tile's scan.
#ifndef WB_SCAN_MINW
#define WB_SCAN_MINW 5
#endif
// F32: the kernel of the FP32 fit options (the precision study); the FP64
// kernel does not carry their code or registers.
template <int K, bool F32>
__global__ void __launch_bounds__(64, WB_SCAN_MINW) k_scan_reg(TreeView T, TileGeo g, const int32_t *tex,
                                                 const int32_t *tey, const uint32_t *tile_off,
                                                 const uint32_t *tile_cnt, const uint32_t *lists,
                                                 ClassList cl, double min_hs, int prec,
                                                 TileRec *rec, unsigned long long *ctr)
{
  const uint32_t n = *cl.n;
  uint32_t i = blockIdx.x;
  if (i >= n) return;
  uint32_t t = cl.at(i);
  for (; i < n; i += gridDim.x) {
    const uint32_t inext = i + gridDim.x;
    const uint32_t tnext = inext < n ? cl.at(inext) : 0u;
    scan_reg<K, F32>(T, g, tex[t], tey[t], tile_off[t], tile_cnt[t], lists, min_hs, prec, rec + t,
                     ctr);
    t = tnext;
  }
}

// ---------------------------------------------------------------------------
// Tiles of at most GS = 16 or 32 points, WAVE/GS of them per wave: group
// q (lanes GS*q .. GS*q+GS-1) scans one tile with the operations of
// scan_reg<1> in the same order (FP64 fit only).  Every exchange stays
// inside a group: the canonical sort is the register network on GS lanes,
// the pairwise sums are the xor butterflies of lanes_pairwise (DPP inside a
// row, one ds_bpermute across the two rows of a 32-lane group) with each
// group's blocks read from the lanes its own point count names, and the
// minima, the first bottom point and the sector counts are taken per group.
// ---------------------------------------------------------------------------
// pairwisesum (manysum.cpp:120-154) of the first m values of this lane's
// group of GS lanes (m may differ between groups; every lane active)
template <int GS>
__device__ inline double group_pairwise(double v, uint32_t m, uint32_t gbase)
{
  constexpr int LV = GS == 16 ? 5 : 6;
  double L[LV];
  L[0] = v;
  L[1] = L[0] + dpp_d<0xb1>(L[0]);
  L[2] = L[1] + dpp_d<0x4e>(L[1]);
  L[3] = L[2] + dpp_d<0x141>(L[2]);
  L[4] = L[3] + dpp_d<0x140>(L[3]);
  if constexpr (GS == 32) L[5] = L[4] + __shfl_xor(L[4], 16, WAVE);
  // the block of bit k of m starts at (m >> (k+1)) << (k+1); block sums are
  // added to 0 in ascending bit order, as wave_pairwise_reg does
  double acc = 0;
#pragma unroll
  for (int k = 0; k < LV; k++) {
    const uint32_t off = (m >> (k + 1)) << (k + 1);
    const double b = __shfl(L[k], (int)(gbase + (off & (GS - 1))), WAVE);
    if ((m >> k) & 1u) acc = acc + b;
  }
  return acc;
}

// minimum / maximum over this lane's group of GS lanes
template <int GS>
__device__ inline double group_dmin(double v)
{
  v = fmin(v, dpp_d<0xb1>(v));
  v = fmin(v, dpp_d<0x4e>(v));
  v = fmin(v, dpp_d<0x141>(v));
  v = fmin(v, dpp_d<0x140>(v));
  if constexpr (GS == 32) v = fmin(v, __shfl_xor(v, 16, WAVE));
  return v;
}
template <int GS>
__device__ inline double group_dmax(double v)
{
  v = fmax(v, dpp_d<0xb1>(v));
  v = fmax(v, dpp_d<0x4e>(v));
  v = fmax(v, dpp_d<0x141>(v));
  v = fmax(v, dpp_d<0x140>(v));
  if constexpr (GS == 32) v = fmax(v, __shfl_xor(v, 16, WAVE));
  return v;
}

template <int GS>
__device__ __attribute__((always_inline)) inline void scan_group(
    const TreeView &T, const TileGeo &g, bool live, int ex, int ey, uint32_t off, uint32_t m,
    const uint32_t *lists, double min_hs, TileRec *out, unsigned long long *ctr)
{
  const uint32_t lane = __lane_id(), gl = lane % GS, gbase = lane - gl;
  const uint64_t gmask = (GS == 64 ? ~0ull : ((1ull << GS) - 1)) << gbase;
  uint32_t c[1] = {live && gl < m ? lists[off + gl] : 0xffffffffu};
  {
    // canonical order inside each group: the network on GS lanes sorts every
    // aligned group of GS lanes on its own (pads last)
    uint32_t unused[1];
    wbw::wave_sort<1, false>(c, unused, GS);
  }
  double cx = 0, cy = 0;
  if (live) tile_center(g, ex, ey, &cx, &cy);
  const double r = g.r;
  const bool mine = live && gl < m;
  double vx = 0, vy = 0, vz = 0;
  if (mine) {
    const uint32_t q = T.canon[c[0]];
    vx = T.px[q] - cx;
    vy = T.py[q] - cy;
    vz = T.pz[q] - 0;
  }
  const double sxx = group_pairwise<GS>(vx * vx, m, gbase);
  const double syx = group_pairwise<GS>(vy * vx, m, gbase);
  const double syy = group_pairwise<GS>(vy * vy, m, gbase);
  const double s1x = group_pairwise<GS>(1.0 * vx, m, gbase);
  const double s1y = group_pairwise<GS>(1.0 * vy, m, gbase);
  const double s11 = group_pairwise<GS>(1.0 * 1.0, m, gbase);
  const double sxz = group_pairwise<GS>(vx * vz, m, gbase);
  const double syz = group_pairwise<GS>(vy * vz, m, gbase);
  const double s1z = group_pairwise<GS>(1.0 * vz, m, gbase);
  double sx = 0, sy = 0;
  plane_tilt(sxx, syx, syy, s1x, s1y, s11, sxz, syz, s1z, &sx, &sy);
  const double u = vz - (sy * vy + sx * vx);
  const double bottom = group_dmin<GS>(mine ? u : __builtin_inf());
  const double top = group_dmax<GS>(mine ? u : -__builtin_inf());
  const uint64_t atb = __ballot(mine && u == bottom) & gmask;
  const uint32_t first = atb ? (uint32_t)(__ffsll((long long)atb) - 1) - gbase : 0xffffffffu;
  double bottom2 = group_dmin<GS>(mine && gl < first ? u : __builtin_inf());
  if (wb::disinf(bottom2)) bottom2 = bottom;
  // 7-sector histogram of the bottom layer
  int sec = -1;
  bool edge = false;
  if (mine && u < bottom2 + 2 * r) {
    int e = 0;
    sec = wb::scan_sector(vx, vy, r, &e);
    edge = e != 0;
  }
  const uint32_t nbottom = (uint32_t)__popcll(__ballot(sec >= 0) & gmask);
  const uint32_t nedge = (uint32_t)__popcll(__ballot(edge) & gmask);
  uint32_t hist[7];
#pragma unroll
  for (int k = 0; k < 7; k++) hist[k] = (uint32_t)__popcll(__ballot(sec == k) & gmask);
  if (live && gl == 0) {
    if (nedge) atomicAdd(&ctr[0], (unsigned long long)nedge);
    double density = 0;
    int tree = 0;
    for (int i = 0; i < 7; i++) density += (int)hist[i] * (int)hist[i];
    if (m > nbottom && density < 7) tree = 1;
    density = sqrt(density) * kSqrt7 / (r * r) / M_PI;
    if (m > nbottom && density < 0.5) tree = 1;
    if (top - bottom > 1.5) tree = 1;
    TileRec rr;
    rr.npoints = (int32_t)m;
    rr.nground = 0;
    rr.tree_flags = tree;
    rr.fit_fp64 = 0;
    rr.density = density;
    rr.hs_scan = sqrt(1 / density + min_hs * min_hs);
    rr.hs = rr.hs_scan;
    rr.height = top - bottom;
    rr.tilt_x = sx;
    rr.tilt_y = sy;
    *out = rr;
  }
}

// WAVE/GS tiles of the class list per wave, grid-stride over the list (the
// FP64 fit; the FP32 options take these tiles through k_scan_reg<1>)
template <int GS>
__global__ void __launch_bounds__(64, WB_SCAN_MINW) k_scan_row(TreeView T, TileGeo g,
                                                             const int32_t *tex, const int32_t *tey,
                                                             const uint32_t *tile_off,
                                                             const uint32_t *tile_cnt,
                                                             const uint32_t *lists, ClassList cl,
                                                             double min_hs, TileRec *rec,
                                                             unsigned long long *ctr)
{
  constexpr uint32_t NG = WAVE / GS;
  const uint32_t n = *cl.n;
  const uint32_t grp = __lane_id() / GS;
  for (uint32_t i0 = blockIdx.x * NG; i0 < n; i0 += gridDim.x * NG) {
    const uint32_t i = i0 + grp;
    const bool live = i < n;
    const uint32_t t = live ? cl.at(i) : 0u;
    scan_group<GS>(T, g, live, live ? tex[t] : 0, live ? tey[t] : 0, live ? tile_off[t] : 0u,
                   live ? tile_cnt[t] : 0u, lists, min_hs, rec + t, ctr);
  }
}

// the scan of one tile of up to SCAN_CAP points by one wave (LDS sort)
__device__ void scan_big(const TreeView &T, const TileGeo &g, int ex, int ey, uint32_t off,
                         uint32_t m, const uint32_t *lists, double min_hs, int prec,
                         TileRec *out, unsigned long long *ctr, uint32_t *cp, int *histo)
{
  const int lane = threadIdx.x;
  uint32_t P = 1;
  while (P < m) P <<= 1;
  for (uint32_t j = lane; j < P; j += WAVE) cp[j] = j < m ? lists[off + j] : 0xffffffffu;
  if (lane < 7) histo[lane] = 0;
  __syncthreads();
  // canonical order of the tile's points = ascending canonical position
  for (uint32_t k = 2; k <= P; k <<= 1)
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = lane; i < P; i += WAVE) {
        uint32_t ixj = i ^ j;
        if (ixj > i) {
          bool up = (i & k) == 0;
          uint32_t a = cp[i], b = cp[ixj];
          if ((a > b) == up) { cp[i] = b; cp[ixj] = a; }
        }
      }
      __syncthreads();
    }
  double cx, cy;
  tile_center(g, ex, ey, &cx, &cy);
  const double r = g.r;
  // the tile's points relative to the cylinder axis, in canonical order
  // (location - xyz(cyl centre, 0)); read through the canonical map
  auto VX = [&](uint32_t j) { return T.px[T.canon[cp[j]]] - cx; };
  auto VY = [&](uint32_t j) { return T.py[T.canon[cp[j]]] - cy; };
  auto VZ = [&](uint32_t j) { return T.pz[T.canon[cp[j]]] - 0; };
  // normal equations of [x y 1] z (transmult and operator*, matrix.cpp:181-215)
  double sxx = wave_pairwise(m, [&](uint32_t j) { double v = VX(j); return v * v; });
  double syx = wave_pairwise(m, [&](uint32_t j) { return VY(j) * VX(j); });
  double syy = wave_pairwise(m, [&](uint32_t j) { double v = VY(j); return v * v; });
  double s1x = wave_pairwise(m, [&](uint32_t j) { return 1.0 * VX(j); });
  double s1y = wave_pairwise(m, [&](uint32_t j) { return 1.0 * VY(j); });
  double s11 = wave_pairwise(m, [&](uint32_t j) { (void)j; return 1.0 * 1.0; });
  double sxz = wave_pairwise(m, [&](uint32_t j) { return VX(j) * VZ(j); });
  double syz = wave_pairwise(m, [&](uint32_t j) { return VY(j) * VZ(j); });
  double s1z = wave_pairwise(m, [&](uint32_t j) { return 1.0 * VZ(j); });
  double sx, sy;
  plane_tilt(sxx, syx, syy, s1x, s1y, s11, sxz, syz, s1z, &sx, &sy);
  // untilted height; bottom, bottom2 (the running minimum before its last
  // update), top
  auto U = [&](uint32_t j) { return VZ(j) - (sy * VY(j) + sx * VX(j)); };
  double bottom = __builtin_inf(), top = -__builtin_inf();
  for (uint32_t j = lane; j < m; j += WAVE) {
    double u = U(j);
    bottom = fmin(bottom, u);
    top = fmax(top, u);
  }
  for (int s = 1; s < WAVE; s <<= 1) {
    bottom = fmin(bottom, __shfl_xor(bottom, s, WAVE));
    top = fmax(top, __shfl_xor(top, s, WAVE));
  }
  uint32_t first = 0xffffffffu;
  for (uint32_t j = lane; j < m; j += WAVE)
    if (U(j) == bottom && j < first) first = j;
  for (int s = 1; s < WAVE; s <<= 1) first = min(first, (uint32_t)__shfl_xor((int)first, s, WAVE));
  double bottom2 = __builtin_inf();
  for (uint32_t j = lane; j < first; j += WAVE) bottom2 = fmin(bottom2, U(j));
  for (int s = 1; s < WAVE; s <<= 1) bottom2 = fmin(bottom2, __shfl_xor(bottom2, s, WAVE));
  if (wb::disinf(bottom2)) bottom2 = bottom;
  // 7-sector histogram of the bottom layer
  uint32_t nbottom = 0;
  for (uint32_t j = lane; j < m; j += WAVE) {
    if (!(U(j) < bottom2 + 2 * r)) continue;
    nbottom++;
    const double vx = VX(j), vy = VY(j);
    int edge = 0;
    const int sec = wb::scan_sector(vx, vy, r, &edge);
    if (edge) atomicAdd(&ctr[0], 1ull);
    atomicAdd(&histo[sec], 1);
  }
  for (int s = 1; s < WAVE; s <<= 1) nbottom += __shfl_xor((int)nbottom, s, WAVE);
  __syncthreads();
  if (lane == 0) {
    double density = 0;
    int tree = 0;
    for (int i = 0; i < 7; i++) density += histo[i] * histo[i];
    if (m > nbottom && density < 7) tree = 1;
    density = sqrt(density) * kSqrt7 / (r * r) / M_PI;
    if (m > nbottom && density < 0.5) tree = 1;
    if (top - bottom > 1.5) tree = 1;
    TileRec rr;
    rr.npoints = (int32_t)m;
    rr.nground = 0;
    rr.tree_flags = tree;
    rr.fit_fp64 = prec != WB_FIT_FP64;  // tiles of this size always fit in FP64
    rr.density = density;
    rr.hs_scan = sqrt(1 / density + min_hs * min_hs);
    rr.hs = rr.hs_scan;
    rr.height = top - bottom;
    rr.tilt_x = sx;
    rr.tilt_y = sy;
    *out = rr;
    if (prec != WB_FIT_FP64) atomicAdd(&ctr[7], 1ull);
  }
  __syncthreads();
}

// Tiles of more than 256 points: the canonical sort in LDS and every value
// read through the canonical map.
__global__ void __launch_bounds__(64) k_scan(TreeView T, TileGeo g, const int32_t *tex,
                                             const int32_t *tey, const uint32_t *tile_off,
                                             const uint32_t *tile_cnt, const uint32_t *lists,
                                             ClassList cl, double min_hs, int prec, TileRec *rec,
                                             uint32_t *err, unsigned long long *ctr)
{
  __shared__ uint32_t cp[SCAN_CAP];
  __shared__ int histo[7];
  const uint32_t n = *cl.n;
  for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
    const uint32_t t = cl.at(i);
    const uint32_t m = tile_cnt[t];
    if (m > SCAN_CAP) {
      if (threadIdx.x == 0) atomicOr(err, 2u);
      continue;
    }
    scan_big(T, g, tex[t], tey[t], tile_off[t], m, lists, min_hs, prec, rec + t, ctr, cp, histo);
  }
}

// POSTSCAN (scan.cpp:142-179): tree tiles widen their hyperboloid by the
// number of tree tiles met along six rays before an empty ring or a
// non-tree tile.
__global__ void k_postscan(TileGeo g, const int32_t *tex, const int32_t *tey, TileRec *rec,
                           unsigned long long *nonempty)
{
  // the non-empty tiles counted per block (one atomic on the counter per
  // block: atomics on one address are serialised)
  __shared__ uint32_t bcount;
  if (threadIdx.x == 0) bcount = 0;
  __syncthreads();
  uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = t < g.ntiles && rec[t].npoints;
  const uint64_t lv = __ballot(live);
  if (__lane_id() == 0 && lv) atomicAdd(&bcount, (uint32_t)__popcll(lv));
  __syncthreads();
  if (threadIdx.x == 0 && bcount) atomicAdd(nonempty, (unsigned long long)bcount);
  if (!live) return;
  TileRec me = rec[t];
  const int r1x[6] = {1, 1, 0, -1, -1, 0}, r1y[6] = {0, 1, 1, 0, -1, -1};  // root1, eisenstein.cpp:51
  int i = 1, ring, nontree, count = 0;
  do {
    ring = nontree = 0;
    for (int j = 0; j < 6 && (me.tree_flags & 1); j++) {
      int o;
      if (!tile_at(g, tex[t] + r1x[j] * i, tey[t] + r1y[j] * i, &o)) continue;
      const TileRec &q = rec[o - g.lo];
      if (q.npoints) {
        ring++;
        if (q.tree_flags & 1) count++;
        else nontree++;
      }
    }
    ++i;
  } while (ring && !nontree);
  double w = count * g.spacing / 6;
  rec[t].hs = sqrt(me.hs_scan * me.hs_scan + w * w);
}

// ---------------------------------------------------------------------------
// CLASSIFY
// ---------------------------------------------------------------------------
// a point handed to the wave-per-point kernel, with the directions found
// so far and the walk it continues with
struct HeavyRec {
  uint64_t occ;
  uint32_t i;
  uint32_t start;  // walk to start from: 0 A, 1 B, 2 the exact border pass
};

struct ClassifyArgs {
  TreeView T;
  TileGeo g;
  const int32_t *memb;
  const uint32_t *nmemb;
  const TileRec *rec;
  double thickness, max_slope;
  // 1: every point lies in its leaf's cube, so the care wedges may skip a
  // node or leaf by its cube and a child cube's bottom face may be raised to
  // the lowest point stored under it; 0 (a point outside the root cube):
  // only the 32-point groups' boxes, which bound the points themselves, are
  // tested
  int cube_wedges;
  const double *tab;     // tan[511], cos[512], sin[512] packed
  const float *edges;    // bin-edge directions (wb::fill_edge_table)
  wb::AtanConst K;
  uint8_t *cls;          // per sorted slot
  uint8_t *cls_in;       // per input index (the caller's class bytes)
  uint32_t *nslow;       // points decided by the exact gap pass
  HeavyRec *heavy;       // points handed to the wave-per-point kernel
  uint32_t *nheavy;
  unsigned long long *atan_exact;
  unsigned long long *counters;  // 16 work counters (only with the counting kernel)
  // work items in place of the points (nullptr: work index k is sorted slot
  // k with the hyperboloid size of its deciding tile): item k is (slot,
  // tile record index), and its verdict goes to item_cls[k] only
  const uint2 *items;
  uint64_t n_items;
  uint8_t *item_cls;
};

// the slot and hyperboloid size of work index k; false for a point in no tile
__device__ inline bool work_point(const ClassifyArgs &A, uint64_t k, uint32_t &slot, double &hs)
{
  if (A.items) {
    const uint2 it = A.items[k];
    slot = it.x;
    hs = A.rec[it.y].hs;
    return true;
  }
  slot = (uint32_t)k;
  if (A.nmemb[k] == 0) return false;
  int w = INT32_MIN;
  for (uint32_t j = 0; j < A.nmemb[k]; j++) w = max(w, A.memb[MAX_MEMB * k + j]);
  hs = A.rec[w - A.g.lo].hs;
  return true;
}

// the verdict of work index k: the class byte of the point (per sorted slot
// and at its input index), or of the work item
__device__ inline void work_verdict(const ClassifyArgs &A, uint64_t k, uint32_t slot, uint8_t v)
{
  if (A.items) {
    A.item_cls[k] = v;
    return;
  }
  A.cls[k] = v;
  A.cls_in[A.T.sidx[slot]] = v;
}

__device__ inline uint64_t rotr64(uint64_t x, int a) { return (x >> a) | (x << (64 - a)); }

// Per-lane care wedges of k_ground: the wedges of wb::care_wedges for each
// lane, at most two, kept in LDS; a lane with more than two wedges (two runs,
// one of them 32 bins or more) is not pruned (n < 0).
struct LaneWedges {
  float ax[2][256], ay[2][256], bx[2][256], by[2][256];
  int n[256];
};

__device__ inline void lane_wedges_update(uint64_t occ, const float *edges, LaneWedges &W, int t)
{
  wb::Wedges w;
  wb::care_wedges(occ, w, edges);
  // live is 1, 3 (one run, split) or 5 (two runs) when at most two are in use;
  // the second kept wedge is slot 1 or slot 2
  const int n = w.live < 0 ? -1 : __builtin_popcount(w.live);
  if (n >= 0 && n <= 2) {
    const bool s1 = (w.live & 2) != 0;
    W.ax[0][t] = w.ax[0]; W.ay[0][t] = w.ay[0];
    W.bx[0][t] = w.bx[0]; W.by[0][t] = w.by[0];
    W.ax[1][t] = s1 ? w.ax[1] : w.ax[2]; W.ay[1][t] = s1 ? w.ay[1] : w.ay[2];
    W.bx[1][t] = s1 ? w.bx[1] : w.bx[2]; W.by[1][t] = s1 ? w.by[1] : w.by[2];
    W.n[t] = n;
  } else {
    W.n[t] = -1;
  }
}

// wb::wedge_hit against thread t's wedges (at most two)
__device__ inline bool lane_hit(const LaneWedges &W, int t, double px, double py, double x0,
                                double x1, double y0, double y1)
{
  const int n = W.n[t];
  if (n < 0) return true;
  const float ax[2] = {W.ax[0][t], W.ax[1][t]}, ay[2] = {W.ay[0][t], W.ay[1][t]};
  const float bx[2] = {W.bx[0][t], W.bx[1][t]}, by[2] = {W.by[0][t], W.by[1][t]};
  return wb::wedge_hit2(n, ax, ay, bx, by, (float)(x0 - px), (float)(x1 - px), (float)(y0 - py),
                        (float)(y1 - py));
}

// classification of the empty-bin mask: 1 pass, 2 fail, 0 undecided
__device__ inline int gap_verdict(uint64_t occupied)
{
  uint64_t e = ~occupied;
  uint64_t r2 = e & rotr64(e, 1);
  uint64_t r4 = r2 & rotr64(r2, 2);
  uint64_t r8 = r4 & rotr64(r4, 4);
  uint64_t r16 = r8 & rotr64(r8, 8);
  uint64_t r24 = r16 & rotr64(r8, 16);
  if (r24 == 0) return 1;
  uint64_t r26 = r24 & rotr64(e, 24) & rotr64(e, 25);
  if (r26 != 0) return 2;
  return 0;
}

__device__ inline int dir_bin_pos(int a) { return (int)(((uint32_t)a + 0x40000000u) & 0x7fffffffu); }

// OR over the 64 lanes by DPP row operations: lanes of a row swap within
// quads, then rotate by 4 and 8, so every lane holds its row's OR; the two
// row broadcasts fold row r-1 into row r and row 1 into rows 2 and 3, so
// lane 63 holds the OR of all lanes.  Lanes a broadcast does not write add 0.
template <int CTRL, int ROWS>
__device__ inline uint32_t dpp_or(uint32_t v)
{
  return v | (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWS, 0xf, false);
}
__device__ inline uint32_t wave_or32(uint32_t v)
{
  v = dpp_or<0xb1, 0xf>(v);   // quad_perm 1,0,3,2
  v = dpp_or<0x4e, 0xf>(v);   // quad_perm 2,3,0,1
  v = dpp_or<0x124, 0xf>(v);  // row_ror 4
  v = dpp_or<0x128, 0xf>(v);  // row_ror 8
  v = dpp_or<0x142, 0xa>(v);  // row_bcast 15 into rows 1 and 3
  v = dpp_or<0x143, 0xc>(v);  // row_bcast 31 into rows 2 and 3
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
__device__ inline uint64_t wave_or(uint64_t v)
{
  return ((uint64_t)wave_or32((uint32_t)(v >> 32)) << 32) | wave_or32((uint32_t)v);
}
// OR over lanes 0-31 (h0) and over lanes 32-63 (h1): the row steps of
// wave_or32, then row 0 folded into row 1 and row 2 into row 3, so lanes 31
// and 63 hold the two halves' ORs.  Every lane active.
__device__ inline void half_or(uint64_t v, uint64_t &h0, uint64_t &h1)
{
  uint32_t a = (uint32_t)v, b = (uint32_t)(v >> 32);
  a = dpp_or<0xb1, 0xf>(a);   b = dpp_or<0xb1, 0xf>(b);   // quad_perm 1,0,3,2
  a = dpp_or<0x4e, 0xf>(a);   b = dpp_or<0x4e, 0xf>(b);   // quad_perm 2,3,0,1
  a = dpp_or<0x124, 0xf>(a);  b = dpp_or<0x124, 0xf>(b);  // row_ror 4
  a = dpp_or<0x128, 0xf>(a);  b = dpp_or<0x128, 0xf>(b);  // row_ror 8
  a = dpp_or<0x142, 0xa>(a);  b = dpp_or<0x142, 0xa>(b);  // row_bcast 15 into rows 1, 3
  h0 = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)b, 31) << 32) |
       (uint32_t)__builtin_amdgcn_readlane((int)a, 31);
  h1 = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)b, 63) << 32) |
       (uint32_t)__builtin_amdgcn_readlane((int)a, 63);
}


// ---------------------------------------------------------------------------
// Ground-test walk: findBlocks (octree.cpp:234-251) for 64 hyperboloids at
// once, with an explicit stack.  The 64 lanes of a wave hold 64 consecutive
// points of the sorted order (spatial neighbours), each with its own
// hyperboloid.  A stack entry is a node or leaf code and the mask of lanes
// for which it and every ancestor intersect the lane's hyperboloid, so each
// lane sees exactly the leaves findBlocks returns for its own shape.
// Expanding a node tests its eight children at once; their cubes follow from
// the parent's by Octree::cube (octree.cpp:349-358), the operations the
// build used, so no child cube is loaded.
// Visiting order: the children of a node are taken starting with the xy
// quadrant that holds the wave's first point, then its x, y and diagonal
// neighbours, the lower half (where a downward hyperboloid lies) first.
// Leaves near the point in every direction come early, which lets the early
// exit of the ground test come early.  The order changes no result: a pass
// on a subset of D is a pass on D, and a fail or an exact gap needs all of D.
// ---------------------------------------------------------------------------
// Entries pending at one time: at most 7 siblings per level on the path
// (internal nodes sit at levels 0..WB_MAX_LEVEL) plus 8 children.
#define WB_STK (7 * (WB_MAX_LEVEL + 1) + 8 + 2)
struct WaveStack {
  int code[WB_STK];
  uint32_t lo[WB_STK], hi[WB_STK];
  uint32_t b[WB_STK], e[WB_STK];  // slot range of a leaf entry
};

// rank of child k in the visiting order for quadrant o and lower-first flag
__device__ inline int child_rank(int k, int o, int zf) { return ((k ^ o) & 3) | ((k ^ zf) & 4); }

// Pushes the children of internal node `code` that intersect the
// hyperboloid of a lane in `m` (this lane takes part iff `on`), in reverse
// visiting order.  Returns the new stack depth.
// Each child's test is wb::hyp_intersect on the child cube of Octree::cube
// (wb::hyp_intersect_low with the lowest point stored under the child when
// every point lies in its leaf's cube) and the lane's care wedges on the
// child's xy square, with the shared parts computed once: the clamped
// coordinate of the closest point takes one of two values per axis (x from
// bit 0 of the child index, y from bit 1, z from bit 2) and the two children
// of one xy square (bit 2 differs) share the wedge test.  Every value is the
// same double operation on the same operands as in the per-child form.
__device__ inline int expand_node(const TreeView &T, WaveStack &S, int sp, int code, bool on,
                                  const wb::Shape &H, double rx, double ry, int zf,
                                  const LaneWedges &LW, bool cube_wedges)
{
  const double *nc = T.node_cube + 4 * (size_t)code;
  const double cx = nc[0], cy = nc[1], cz = nc[2], side = nc[3];
  const int *chp = T.node_child + 8 * (size_t)code;
  const int o = rfl((rx >= cx ? 1 : 0) | (ry >= cy ? 2 : 0));
  int ck[8];
  uint32_t pushed = 0;  // bit t: the child of rank t is pushed
  uint64_t mk[8];       // lanes for which child k is pushed
#pragma unroll
  for (int k = 0; k < 8; k++) ck[k] = rfl(chp[k]);
  // lane k < 8 holds child k's slot range (leaf children), pushed with it
  const uint2 cr = __lane_id() < 8 ? T.node_crange[8 * (size_t)code + __lane_id()]
                                   : make_uint2(0u, 0u);
  // lane k < 8 holds the lowest z under child k
  const double zl = cube_wedges && __lane_id() < 8 ? T.child_zlow[8 * (size_t)code + __lane_id()]
                                                   : -__builtin_inf();
  // Octree::cube (octree.cpp:349-358) per axis bit, and the closest point of
  // wb::hyp_intersect: clamp_axis on x and y, the bottom (top for a negative
  // slope) face on z
  const double cs = side / 2;
  double ncx[2], ncy[2], qx[2], qy[2], qz[2];
#pragma unroll
  for (int b = 0; b < 2; b++) {
    ncx[b] = cx + (double)(2 * b - 1) * side / 4;
    ncy[b] = cy + (double)(2 * b - 1) * side / 4;
    const double ncz = cz + (double)(2 * b - 1) * side / 4;
    qx[b] = wb::clamp_axis(H.cx, ncx[b], cs);
    qy[b] = wb::clamp_axis(H.cy, ncy[b], cs);
    qz[b] = ncz;
    if (H.slope > 0) qz[b] -= cs / 2;
    if (H.slope < 0) qz[b] += cs / 2;
  }
  // the lane's wedges (at most two), read from LDS once; one test per xy
  // square, on the three x and three y edges of the four squares as offsets
  // from the point (the middle edge ncx[0] + side/4 = ncx[1] - side/4: cube
  // centres and sides are exact dyadics)
  bool wq[4] = {true, true, true, true};
  const int nw = cube_wedges ? LW.n[threadIdx.x] : -1;
  if (nw >= 0) {
    const int t = threadIdx.x;
    const float ax[2] = {LW.ax[0][t], LW.ax[1][t]}, ay[2] = {LW.ay[0][t], LW.ay[1][t]};
    const float bx[2] = {LW.bx[0][t], LW.bx[1][t]}, by[2] = {LW.by[0][t], LW.by[1][t]};
    const float ux[3] = {(float)((ncx[0] - side / 4) - H.cx), (float)((ncx[0] + side / 4) - H.cx),
                         (float)((ncx[1] + side / 4) - H.cx)};
    const float vy[3] = {(float)((ncy[0] - side / 4) - H.cy), (float)((ncy[0] + side / 4) - H.cy),
                         (float)((ncy[1] + side / 4) - H.cy)};
#pragma unroll
    for (int q = 0; q < 4; q++)
      if ((ck[q] != -1) | (ck[q + 4] != -1))
        wq[q] = wb::wedge_hit2(nw, ax, ay, bx, by, ux[q & 1], ux[(q & 1) + 1], vy[q >> 1],
                               vy[(q >> 1) + 1]);
  }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    mk[k] = 0;
    if (ck[k] != -1) {
      const double zlk = rdl(zl, k);
      const double qzk = H.slope > 0 && zlk > qz[k >> 2] ? zlk : qz[k >> 2];
      const bool h = on & wq[k & 3] & wb::hyp_in(H, qx[k & 1], qy[(k >> 1) & 1], qzk);
      mk[k] = __ballot(h);
      if (mk[k]) pushed |= 1u << child_rank(k, o, zf);
    }
  }
#pragma unroll
  for (int k = 0; k < 8; k++) {
    if (mk[k]) {
      const int pos = sp + __popc(pushed >> (child_rank(k, o, zf) + 1));
      S.code[pos] = ck[k];
      S.lo[pos] = (uint32_t)mk[k];
      S.hi[pos] = (uint32_t)(mk[k] >> 32);
      S.b[pos] = (uint32_t)__builtin_amdgcn_readlane((int)cr.x, k);
      S.e[pos] = (uint32_t)__builtin_amdgcn_readlane((int)cr.y, k);
    }
  }
  return sp + __popc(pushed);
}

// per-wave work counters of the instrumented kernel
enum {
  CT_LANES_A, CT_LANES_B, CT_LANES_C, CT_LEAF_A, CT_GROUP_A, CT_DIR_A, CT_LEAF_B, CT_POINT_B,
  CT_DIR_B, CT_EXP_A, CT_WLEAF_A, CT_EXP_B, CT_WLEAF_B, CT_HEAVY, CT_WAVES,
  CT_TSUM, CT_TMAX, CT_THIST, CT_TSUMS = CT_THIST + 5,
  // shader cycles per wave in: walk A expansions, walk A leaves, walk B
  // expansions, walk B leaves, wedge refreshes; wave-level sample and slot
  // iterations of walks A and B
  CT_CYC = CT_TSUMS + 5, CT_ITER_A = CT_CYC + 8, CT_ITER_B, CT_N,
  // the counting kernels add into 64 copies (block index mod 64) so their
  // own atomics do not slow the memory system they measure
  CT_STRIDE = CT_N + 24, CT_COPIES = 64
};

// a lane whose sampled pass enters leaves with more than HEAVY_B_GROUPS
// 32-point groups, or whose sampled pass alone passes HEAVY_A_GROUPS, is
// finished by k_ground_heavy
#ifndef HEAVY_A_GROUPS
#define HEAVY_A_GROUPS 2048u
#endif
// blocks of the heavy kernel at most: 4 waves on each SIMD of 256 CUs
#define HEAVY_GRID 4096u
#ifndef HEAVY_B_GROUPS
#define HEAVY_B_GROUPS 1024u
#endif

// The ground test of classify.cpp:128-165 for one point per lane, in two
// walks over the octree:
//   A  the lowest point of each 32-point group of every leaf (a subset of D;
//      only a pass is decided here),
//   B  every point, for lanes A left open.
// Lanes whose walks would be long, and lanes whose directions leave a
// longest empty run of 24 or 25 bins (decided by the exact directions in
// the two bins bordering the run), go to k_ground_heavy with the directions
// found so far.
// The hyperboloid is Hyperboloid(p - (0,0,thickness), hs, slope)
// (shape.cpp:119-125) with its axis at x, y instead of x + 0, y + 0: the
// two differ only in the sign of a zero, which no test below can see (every
// use is a difference that is squared, compared or passed to hypot).
// unit: the index of the wave's 64 work items (items 64*unit ..); stacks,
// LW, edges: the block's shared arrays (edges: the bin-edge table, read by
// wedge refreshes, loaded by ground_setup).  No barrier: the waves of a
// block run their units independently.
__device__ __forceinline__ void ground_setup(const ClassifyArgs &A, float *edges)
{
  for (int k = threadIdx.x; k < 3 * 128; k += blockDim.x) edges[k] = A.edges[k];
  __syncthreads();
}

template <bool COUNT>
__device__ __forceinline__ void ground_body(const ClassifyArgs &A, uint64_t unit,
                                            WaveStack *stacks, LaneWedges &LW, const float *edges)
{
  WaveStack &S = stacks[threadIdx.x / WAVE];
  LW.n[threadIdx.x] = -1;
  const uint32_t blk = (uint32_t)unit;  // spreads the instrumented counters
  const TreeView &T = A.T;
  const uint32_t lane = __lane_id();
  // work index: a sorted slot, or a work item
  const uint64_t i = unit * WAVE + lane;
  const bool inrange = i < (A.items ? A.n_items : T.n);
  const double *ttan = A.tab, *tcos = A.tab + 511, *tsin = A.tab + 1023;
  const uint64_t t_start = COUNT ? wall_clock64() : 0;
  uint32_t cnt[CT_WAVES] = {};
  uint64_t cyc[8] = {}, iter_a = 0, iter_b = 0;
  // lane state: 0 undecided, 1 not ground, 2 ground, 3 no tile; handed to
  // k_ground_heavy: 5 walk A cut short, 6 after walk A, 7 after walk B
  int state = 3;
  double x = 0, y = 0, z = 0, hs = 0;
  uint32_t slot = 0;
  if (inrange && work_point(A, i, slot, hs)) {
    x = T.px[slot]; y = T.py[slot]; z = T.pz[slot];
    state = 0;
  }
  wb::Shape H;
  {
    const double sl = A.max_slope, por = hs * (sl * sl);
    H.kind = wb::SHAPE_HYPERBOLOID;
    H.cx = x;
    H.cy = y;
    H.cz = (z - A.thickness) + por;
    H.r = 0;
    H.por2 = por * por;
    H.slope = sl;
  }
  const int zf = A.max_slope > 0 ? 0 : 4;
  uint64_t occ = 0, checked = 0;
  int nexact = 0;
  uint32_t groups_a = 0;
  for (int phase = 0; phase < 2; phase++) {
    if (phase == 1 && state == 0 && groups_a > HEAVY_B_GROUPS) state = 6;
    const bool act = state == 0;
    uint64_t alive = __ballot(act);
    if (alive == 0) break;
    if (COUNT && act) {
      if (phase == 0) cnt[CT_LANES_A]++;
      else cnt[CT_LANES_B]++;
    }
    // the wave's first point orders the children
    const int ld = __ffsll((long long)alive) - 1;
    const double rx = __shfl(x, ld, 64), ry = __shfl(y, ld, 64);
    int sp = 0;
    {
      const double *rc = T.node_cube;
      const uint64_t m = __ballot(act && wb::hyp_intersect(H, rc[0], rc[1], rc[2], rc[3]));
      if (m) {
        S.code[0] = 0;
        S.lo[0] = (uint32_t)m;
        S.hi[0] = (uint32_t)(m >> 32);
        sp = 1;
      }
    }
    while (sp > 0) {
      sp--;
      const int code = rfl(S.code[sp]);
      const uint64_t m = (((uint64_t)(uint32_t)rfl((int)S.hi[sp]) << 32) |
                          (uint32_t)rfl((int)S.lo[sp])) & alive;
      if (m == 0) continue;
      const bool on = (m >> lane) & 1;
      const uint64_t c0t = COUNT ? clock64() : 0;
      if (code >= 0) {
        if (COUNT) {
          if (phase == 0) cnt[CT_EXP_A]++;
          else cnt[CT_EXP_B]++;
        }
        sp = expand_node(T, S, sp, code, on, H, rx, ry, zf, LW, A.cube_wedges != 0);
        if (COUNT) cyc[phase == 0 ? 0 : 2] += clock64() - c0t;
        continue;
      }
      const int leaf = -2 - code;
      const uint32_t b = (uint32_t)rfl((int)S.b[sp]);
      const uint32_t e = (uint32_t)rfl((int)S.e[sp]);
      // the lane skips a leaf outside its care wedges
      const double *lc = T.leaf_cube + 4 * (size_t)leaf;
      const double lh = lc[3] / 2;
      const bool onl = on && (!A.cube_wedges || lane_hit(LW, threadIdx.x, x, y, lc[0] - lh,
                                                         lc[0] + lh, lc[1] - lh, lc[1] + lh));
      uint64_t c2t = 0;
      if (COUNT && phase == 0) {
        __builtin_amdgcn_s_waitcnt(0);
        c2t = clock64();
        cyc[5] += c2t - c0t;
      }
      if (phase == 0) {
        if (COUNT) {
          cnt[CT_WLEAF_A]++;
          if (onl) cnt[CT_LEAF_A]++;
        }
        const uint32_t g0 = b / WB_CHUNK, ng = (e - 1) / WB_CHUNK - g0 + 1;
        if (onl) groups_a += ng;
        if (COUNT && onl) cnt[CT_GROUP_A] += ng;
        // the lowest point of each group (the likeliest to lie in a downward
        // hyperboloid), or the group's first slot in the leaf when the
        // lowest one belongs to the neighbouring leaf; lane g fetches the
        // sample of group g, then every lane tests every sample
        for (uint32_t gb = 0; gb < ng; gb += WAVE) {
          const uint32_t g = gb + lane;
          double sx = 0, sy = 0, sz = 0;
          if (g < ng) {
            // the group's lowest point from the per-group copy (consecutive
            // lanes, consecutive words); a group that starts or ends in
            // another leaf takes its first slot in this leaf instead
            const uint32_t j = T.chunk_low[g0 + g];
            if (j < b || j >= e) {
              const uint32_t jf = g == 0 ? b : (g0 + g) * WB_CHUNK;
              sx = T.px[jf]; sy = T.py[jf]; sz = T.pz[jf];
            } else {
              sx = T.chunk_sx[g0 + g]; sy = T.chunk_sy[g0 + g]; sz = T.chunk_sz[g0 + g];
            }
          }
          if (COUNT) {
            __builtin_amdgcn_s_waitcnt(0);
            const uint64_t c3t = clock64();
            cyc[6] += c3t - c2t;
            c2t = c3t;
          }
          const uint32_t ns = min((uint32_t)WAVE, ng - gb);
          if (COUNT && __ballot(onl)) iter_a += ns;
          for (uint32_t s = 0; s < ns; s++) {
            const double qx = rdl(sx, s), qy = rdl(sy, s), qz = rdl(sz, s);
            if (onl & wb::hyp_in(H, qx, qy, qz) & ((qx != x) | (qy != y))) {
              if (COUNT) cnt[CT_DIR_A]++;
              occ |= 1ull << wb::dir_bin(qy - y, qx - x, ttan, tcos, tsin, A.K, &nexact);
            }
          }
          if (COUNT) {
            const uint64_t c3t = clock64();
            cyc[7] += c3t - c2t;
            c2t = c3t;
          }
        }
      } else {
        if (COUNT) {
          cnt[CT_WLEAF_B]++;
          if (onl) cnt[CT_LEAF_B]++;
        }
        const bool all = onl && wb::hyp_in_cube(H, lc[0], lc[1], lc[2], lc[3]);
        const uint64_t allm = __ballot(all);
        const uint32_t js = lane & (WB_CHUNK - 1);
        for (uint32_t c0 = b; c0 < e;) {
          const uint32_t chk = c0 / WB_CHUNK;
          const uint32_t ce = min(e, (chk + 1) * WB_CHUNK);
          // lanes j and j + 32 fetch slot c0 + j while the group's box is tested
          double sx = 0, sy = 0, sz = 0;
          if (js < ce - c0) {
            sx = T.px[c0 + js]; sy = T.py[c0 + js]; sz = T.pz[c0 + js];
          }
          bool chit = onl;
          if (chit) {
            const double *bx = T.chunk_box + 6 * (uint64_t)chk;
            if (!all) chit = wb::hyp_may_hit_box(H, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5]);
            if (chit) chit = lane_hit(LW, threadIdx.x, x, y, bx[0], bx[1], bx[2], bx[3]);
          }
          // The lanes that need the group ("owners"), two per step: lanes
          // 0-31 test its slots against the first owner's hyperboloid, lanes
          // 32-63 against the second's, and each half's direction bits are
          // ORed into its owner's occupancy.
          uint64_t own = __ballot(chit);
          if (COUNT && chit) cnt[CT_POINT_B] += ce - c0;
          while (own) {
            const int o1 = __ffsll((long long)own) - 1;
            own &= own - 1;
            const int o2 = own ? __ffsll((long long)own) - 1 : o1;
            const bool two = own != 0;
            own &= own - 1;
            if (COUNT) iter_b++;
            const bool hi_half = lane >= WAVE / 2;
            const int o = hi_half ? o2 : o1;
            // the owner's hyperboloid, fetched across lanes by the LDS unit
            wb::Shape Ho = H;
            Ho.cx = __shfl(x, o, WAVE);
            Ho.cy = __shfl(y, o, WAVE);
            Ho.cz = __shfl(H.cz, o, WAVE);
            Ho.por2 = __shfl(H.por2, o, WAVE);
            const bool valid = js < ce - c0 && (!hi_half || two);
            bool in = valid & ((((allm >> o) & 1) != 0) | wb::hyp_in(Ho, sx, sy, sz)) &
                      ((sx != Ho.cx) | (sy != Ho.cy));
            // A direction outside the owner's care wedges is not binned: its
            // bin is occupied or in an empty run of fewer than 24 bins, runs
            // only shrink, so it never lies in a run of 24 or more, and the
            // runs of 24 or more (the verdict, and the occupied bins that
            // border them) come out the same without it.
            if (in) {
              const int ot = (int)(threadIdx.x & ~(WAVE - 1)) + o;
              const int nw = LW.n[ot];
              if (nw >= 0) {
                const float ax[2] = {LW.ax[0][ot], LW.ax[1][ot]}, ay[2] = {LW.ay[0][ot], LW.ay[1][ot]};
                const float bx[2] = {LW.bx[0][ot], LW.bx[1][ot]}, by[2] = {LW.by[0][ot], LW.by[1][ot]};
                const float u = (float)(sx - Ho.cx), v = (float)(sy - Ho.cy);
                in = wb::wedge_hit2(nw, ax, ay, bx, by, u, u, v, v);
              }
            }
            uint64_t bit = 0;
            if (in) {
              if (COUNT) cnt[CT_DIR_B]++;
              bit = 1ull << wb::dir_bin(sy - Ho.cy, sx - Ho.cx, ttan, tcos, tsin, A.K, &nexact);
            }
            if (__ballot(bit != 0)) {
              uint64_t r1, r2;
              half_or(bit, r1, r2);
              if ((int)lane == o1) occ |= r1;
              if ((int)lane == o2 && two) occ |= r2;
            }
          }
          c0 = ce;
        }
      }
      const uint64_t c1t = COUNT ? clock64() : 0;
      if (COUNT) cyc[phase == 0 ? 1 : 3] += c1t - c0t;
      if (state == 0 && occ != checked) {
        checked = occ;
        if (gap_verdict(occ) == 1) state = 1;
        else lane_wedges_update(occ, edges, LW, threadIdx.x);
      }
      if (COUNT) cyc[4] += clock64() - c1t;
      if (phase == 0 && state == 0 && groups_a > HEAVY_A_GROUPS) state = 5;
      alive = __ballot(state == 0);
      if (alive == 0) break;
    }
  }
  if (state == 0) {
    const int v = gap_verdict(occ);
    state = v == 0 ? 7 : v;
  }
  // hand the open lanes over, with the directions found so far
  const uint64_t hb = __ballot(state >= 5);
  if (hb) {
    const int leader = __ffsll((long long)hb) - 1;
    uint32_t base = 0;
    if ((int)lane == leader) base = atomicAdd(A.nheavy, (uint32_t)__popcll(hb));
    base = __shfl(base, leader, 64);
    if (state >= 5) {
      HeavyRec r;
      r.occ = occ;
      r.i = (uint32_t)i;
      r.start = (uint32_t)(state - 5);
      A.heavy[base + __popcll(hb & ((1ull << lane) - 1))] = r;
    }
  }
  if (nexact) atomicAdd(A.atan_exact, (unsigned long long)nexact);
  if (inrange && state < 5) work_verdict(A, i, slot, state == 1 ? 1 : (state == 2 ? 2 : 0));
  if (COUNT) {
    if (state == 7) cnt[CT_LANES_C]++;
    if (state >= 5) cnt[CT_HEAVY]++;
    uint64_t v[CT_WAVES];
    for (int k = 0; k < CT_WAVES; k++) {
      uint64_t t = cnt[k];
      // wave-level counters are the same in every lane
      if (k != CT_EXP_A && k != CT_WLEAF_A && k != CT_EXP_B && k != CT_WLEAF_B)
        for (int sft = 1; sft < 64; sft <<= 1) t += __shfl_xor(t, sft, 64);
      v[k] = t;
    }
    if (lane == 0) {
      unsigned long long *ctr = A.counters + (blk % CT_COPIES) * CT_STRIDE;
      for (int k = 0; k < CT_WAVES; k++) atomicAdd(&ctr[k], (unsigned long long)v[k]);
      atomicAdd(&ctr[CT_WAVES], 1ull);
      const uint64_t dt = wall_clock64() - t_start;
      atomicAdd(&ctr[CT_TSUM], (unsigned long long)dt);
      atomicMax(&ctr[CT_TMAX], (unsigned long long)dt);
      // wave times in ticks: <1e3, <1e4, <1e5, <1e6, >=1e6
      const int hb = dt < 1000 ? 0 : dt < 10000 ? 1 : dt < 100000 ? 2 : dt < 1000000 ? 3 : 4;
      atomicAdd(&ctr[CT_THIST + hb], 1ull);
      atomicAdd(&ctr[CT_TSUMS + hb], (unsigned long long)dt);
      for (int k = 0; k < 8; k++) atomicAdd(&ctr[CT_CYC + k], (unsigned long long)cyc[k]);
      atomicAdd(&ctr[CT_ITER_A], (unsigned long long)iter_a);
      atomicAdd(&ctr[CT_ITER_B], (unsigned long long)iter_b);
    }
  }
}


template <bool COUNT, int MINW>
__global__ void __launch_bounds__(256, MINW) k_ground(ClassifyArgs A)
{
  __shared__ WaveStack stacks[256 / WAVE];
  __shared__ LaneWedges LW;
  __shared__ float edges[3 * 128];
  ground_setup(A, edges);
  ground_body<COUNT>(A, (uint64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE, stacks, LW,
                     edges);
}

// The same test on a grid that stays resident: each wave takes its first
// unit of 64 items by its place in the grid, then the next unit no wave has
// taken (*next counts the units taken beyond the grid's; zero before the
// launch), until none is left.  Launched with fewer blocks than fit on the
// device, it leaves room on every CU for the kernels of other streams,
// which a kernel of many short blocks does not: each freed slot goes to its
// next block.
template <int MINW>
__global__ void __launch_bounds__(256, MINW) k_ground_resident(ClassifyArgs A, uint32_t *next)
{
  __shared__ WaveStack stacks[256 / WAVE];
  __shared__ LaneWedges LW;
  __shared__ float edges[3 * 128];
  ground_setup(A, edges);
  const uint64_t nunits = ((A.items ? A.n_items : A.T.n) + WAVE - 1) / WAVE;
  const uint32_t waves = gridDim.x * (256 / WAVE);
  uint64_t unit = (uint64_t)blockIdx.x * (256 / WAVE) + threadIdx.x / WAVE;
  while (unit < nunits) {
    ground_body<false>(A, unit, stacks, LW, edges);
    uint32_t u = 0;
    if (__lane_id() == 0) u = atomicAdd(next, 1u);
    unit = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)u) + waves;
  }
}

// The ground tests of several trees in one launch: blocks boff[k] ..
// boff[k+1]-1 take tree k, whose arguments are AA[k].
template <int MINW>
__global__ void __launch_bounds__(256, MINW) k_ground_batch(const ClassifyArgs *AA,
                                                            const uint32_t *boff, int ntrees)
{
  __shared__ WaveStack stacks[256 / WAVE];
  __shared__ LaneWedges LW;
  __shared__ float edges[3 * 128];
  int lo = 0, hi = ntrees;  // boff[lo] <= blockIdx.x < boff[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (boff[mid] <= blockIdx.x) lo = mid;
    else hi = mid;
  }
  ground_setup(AA[lo], edges);
  ground_body<false>(AA[lo], (uint64_t)(blockIdx.x - boff[lo]) * (256 / WAVE) + threadIdx.x / WAVE,
                     stacks, LW, edges);
}

// One heavy point per 64-lane block.  The shape is the same in every lane,
// so the lanes split the work.  The walk pops up to eight nodes at a time
// (lane 8i+k tests child k of the i-th node from the top; children in the
// visiting order of k_ground, the popped nodes' children stacked so the top
// node's come out first) and collects the leaves it reaches into a batch of
// up to HB_LEAVES.  A batch is served lane-parallel: per
// leaf its bounds, whether its cube lies inside the hyperboloid and which
// 32-point groups' boxes may meet it; then 64 of the batch's samples (walk
// A: the lowest point of each group) or slots (walk B, and C, the exact
// border pass for a longest empty run of 24 or 25 bins) per step, with the
// early exit checked after every step.
#ifndef HB_LEAVES
#define HB_LEAVES 256
#endif
// The heavy walk's stack.  Its levels never decrease from bottom to top
// (the children of the i-th popped node sit above those of the (i+1)-th, one
// level deeper than their parents).  So a round that pops a level-(m-1) node
// also pops every level >= m entry, and afterwards the level-m entries are
// that round's children, at most 8 x 8 = 64; until the next such round they
// only decrease.  Internal nodes sit at levels 0..20: at most 64 entries at
// each of levels 1..20.
#define HB_STK (64 * WB_MAX_LEVEL - 64)

#ifndef WB_HEAVY_MINW
#define WB_HEAVY_MINW 4
#endif
// The block's shared arrays: the walk's stack, the leaf batch (leaf,
// bounds, group mask, prefix of counts), the care wedges of walks A and B
// (uniform).
struct HeavyShared {
  int stk[HB_STK];
  uint32_t lq[HB_LEAVES], lb[HB_LEAVES], le[HB_LEAVES], lmask[HB_LEAVES];
  uint32_t lpre[HB_LEAVES + 1];
  wb::Wedges wg;
};

template <bool COUNT>
__device__ __forceinline__ void heavy_body(const ClassifyArgs &A, const HeavyRec hr, uint32_t blk,
                                           HeavyShared &SH)
{
  int *stk = SH.stk;
  uint32_t *lq = SH.lq, *lb = SH.lb, *le = SH.le, *lmask = SH.lmask, *lpre = SH.lpre;
  wb::Wedges &wg = SH.wg;
  const uint64_t t_start = COUNT ? wall_clock64() : 0;
  uint64_t c_leaf = 0, c_pts = 0, c_exp = 0, c_test = 0, c_in = 0;
  const TreeView &T = A.T;
  const uint32_t i = hr.i;  // work index (k_ground hands over only lanes with a tile)
  const uint32_t lane = threadIdx.x;
  const double *ttan = A.tab, *tcos = A.tab + 511, *tsin = A.tab + 1023;
  uint32_t slot = 0;
  double hs = 0;
  work_point(A, i, slot, hs);
  const double x = T.px[slot], y = T.py[slot], z = T.pz[slot];
  const wb::Shape H = wb::make_hyperboloid(x - 0, y - 0, z - A.thickness, hs, A.max_slope);
  const int zf = A.max_slope > 0 ? 0 : 4;
  int nexact = 0;
  // the directions k_ground found are a subset of D; the walks it completed
  // did not pass and are not repeated
  uint64_t occ = lane == 0 ? hr.occ : 0, checked = hr.occ;
  int verdict = 0;  // 1 pass, 2 fail, 4 ambiguous
  int before[2] = {-1, -1}, afterb[2] = {-1, -1}, nr = 0;
  int hi[2] = {-1, -1}, lo[2] = {0x7fffffff, 0x7fffffff};
  for (int phase = rfl((int)hr.start); phase < 3 && verdict != 1 && verdict != 2; phase++) {
    if (phase == 2) {
      const uint64_t occw = wave_or(occ);
      verdict = gap_verdict(occw);
      if (verdict != 0) break;
      verdict = 4;
      for (int st = 0; st < 64; st++) {
        if ((occw >> st) & 1) continue;
        if (!((occw >> ((st + 63) & 63)) & 1)) continue;
        int len = 0;
        while (len < 64 && !((occw >> ((st + len) & 63)) & 1)) len++;
        if (len >= 24 && nr < 2) {
          before[nr] = (st + 63) & 63;
          afterb[nr] = (st + len) & 63;
          nr++;
        }
      }
    }
    int sp = 0;
    {
      const double *rc = T.node_cube;
      if (wb::hyp_intersect(H, rc[0], rc[1], rc[2], rc[3])) {
        if (lane == 0) stk[0] = 0;
        sp = 1;
      }
    }
    if (lane == 0) {
      if (phase < 2) wb::care_wedges(checked, wg, A.edges);
      else wb::border_wedges(before, afterb, nr, wg, A.edges);
    }
    __syncthreads();
    bool done = false;
    // batches grow 16, 64, 256 leaves: an early pass stops after few leaves
    uint32_t lim = 16;
    while (!done) {
      // collect the leaves of the next part of the walk
      uint32_t nq = 0;
      while (sp > 0 && nq + 8 <= lim) {
        // P nodes from the top of the stack, with room for their leaves in
        // the batch (the stack's own room follows from HB_STK)
        const int P = min(min(8, sp), (int)((lim - nq) / 8));
        const int gi = (int)(lane >> 3), k = (int)(lane & 7);
        const int code = gi < P ? stk[sp - 1 - gi] : 0;
        __syncthreads();
        if (COUNT) c_exp += P;
        bool h = false;
        int ck = -1, t = 0;
        if (gi < P) {
          const double *nc = T.node_cube + 4 * (size_t)code;
          const double cx = nc[0], cy = nc[1], cz = nc[2], side = nc[3];
          const int o = (x >= cx ? 1 : 0) | (y >= cy ? 2 : 0);
          ck = T.node_child[8 * (size_t)code + k];
          t = child_rank(k, o, zf);
          if (ck != -1) {
            const double ncx = cx + (double)(2 * (k & 1) - 1) * side / 4;
            const double ncy = cy + (double)(2 * ((k & 2) / 2) - 1) * side / 4;
            const double ncz = cz + (double)(2 * ((k & 4) / 4) - 1) * side / 4;
            const double zl = A.cube_wedges ? T.child_zlow[8 * (size_t)code + k] : -__builtin_inf();
            h = wb::hyp_intersect_low(H, ncx, ncy, ncz, side / 2, zl) &&
                (!A.cube_wedges || wb::wedge_hit(wg, x, y, ncx - side / 4, ncx + side / 4,
                                                 ncy - side / 4, ncy + side / 4));
          }
        }
        // per popped node: node children and leaf children by rank
        // (OR over each group of 8 lanes: quad swaps, then the half-row mirror)
        uint32_t rn = h && ck >= 0 ? 1u << t : 0u, rl = h && ck < -1 ? 1u << t : 0u;
        rn = dpp_or<0xb1, 0xf>(rn);
        rl = dpp_or<0xb1, 0xf>(rl);
        rn = dpp_or<0x4e, 0xf>(rn);
        rl = dpp_or<0x4e, 0xf>(rl);
        rn = dpp_or<0x141, 0xf>(rn);
        rl = dpp_or<0x141, 0xf>(rl);
        // node i's children go above those of nodes i+1..P-1; its leaves
        // after those of nodes 0..i-1
        const int base = sp - P;
        int nabove = 0, lbefore = 0, ntot = 0, ltot = 0;
        for (int i = 0; i < P; i++) {
          const int cn = __popc((uint32_t)__builtin_amdgcn_readlane((int)rn, 8 * i));
          const int cl = __popc((uint32_t)__builtin_amdgcn_readlane((int)rl, 8 * i));
          if (i > gi) nabove += cn;
          if (i < gi) lbefore += cl;
          ntot += cn;
          ltot += cl;
        }
        if (h && ck >= 0) stk[base + nabove + __popc(rn >> (t + 1))] = ck;
        if (h && ck < -1) lq[nq + lbefore + __popc(rl & ((1u << t) - 1))] = (uint32_t)(-2 - ck);
        sp = base + ntot;
        nq += ltot;
        __syncthreads();
      }
      if (nq == 0) break;
      lim = min(4 * lim, (uint32_t)HB_LEAVES);
      if (COUNT) c_leaf += nq;
      // per-leaf bounds, inside flag (bit 31 of lmask) and group mask, and
      // the prefix of the leaves' sample counts (walk A) or selected-group
      // counts (walks B and C)
      uint32_t carry = 0;
      for (uint32_t k0 = 0; k0 < nq; k0 += WAVE) {
        const uint32_t k = k0 + lane;
        uint32_t cnt = 0;
        if (k < nq) {
          const uint32_t leaf = lq[k];
          const uint32_t b = T.leaf_start[leaf], e = T.leaf_start[leaf + 1];
          const uint32_t g0 = b / WB_CHUNK, ng = (e - 1) / WB_CHUNK - g0 + 1;
          // bits 0-30: the groups to read (groups past the 31st are always
          // read); bit 31: the leaf cube lies inside the hyperboloid
          uint32_t mask = 0x7fffffffu;
          const double *lc = T.leaf_cube + 4 * (size_t)leaf;
          const double h2 = lc[3] / 2;
          // the wedges may have narrowed since the leaf was collected
          const bool seen =
              !A.cube_wedges || wb::wedge_hit(wg, x, y, lc[0] - h2, lc[0] + h2, lc[1] - h2, lc[1] + h2);
          if (phase != 0 && seen) {
            const bool inside = wb::hyp_in_cube(H, lc[0], lc[1], lc[2], lc[3]);
            mask = inside ? 0x80000000u : 0u;
            for (uint32_t g = 0; g < ng && g < 31; g++) {
              const double *bx = T.chunk_box + 6 * (uint64_t)(g0 + g);
              if ((inside || wb::hyp_may_hit_box(H, bx[0], bx[1], bx[2], bx[3], bx[4], bx[5])) &&
                  wb::wedge_hit(wg, x, y, bx[0], bx[1], bx[2], bx[3]))
                mask |= 1u << g;
            }
          }
          lb[k] = b;
          le[k] = e;
          lmask[k] = mask;
          cnt = !seen ? 0 : phase == 0 ? ng
                                       : __popc(mask & 0x7fffffffu) + (ng > 31 ? ng - 31 : 0);
        }
        uint32_t incl = cnt;
        for (int s = 1; s < WAVE; s <<= 1) {
          const uint32_t v = (uint32_t)__shfl_up((int)incl, s, WAVE);
          if ((int)lane >= s) incl += v;
        }
        if (k < nq) lpre[k + 1] = carry + incl;
        carry += (uint32_t)__shfl((int)incl, WAVE - 1, WAVE);
      }
      if (lane == 0) lpre[0] = 0;
      __syncthreads();
      const uint32_t S = carry;
      if (COUNT) c_pts += phase == 0 ? S : WB_CHUNK * S;
      uint32_t kc = 0;  // this lane's leaf in the batch
      // walk A: a sample per lane; walks B and C: two selected groups per
      // step, lanes 0-31 the slots of the first, 32-63 of the second
      const uint32_t per = phase == 0 ? WAVE : WAVE / WB_CHUNK;
      for (uint32_t f0 = 0; f0 < S && !done; f0 += per) {
        const uint32_t f = phase == 0 ? f0 + lane : f0 + lane / WB_CHUNK;
        while (kc + 1 < nq && lpre[kc + 1] <= f) kc++;
        if (f < S) {
          const uint32_t b = lb[kc], e = le[kc], mask = lmask[kc], o = f - lpre[kc];
          const uint32_t g0 = b / WB_CHUNK;
          uint32_t j;
          bool test = true;
          if (phase == 0) {
            j = T.chunk_low[g0 + o];
            if (j < b || j >= e) j = o == 0 ? b : (g0 + o) * WB_CHUNK;
          } else {
            // the o-th selected group of the leaf: a set bit of the mask,
            // or one of the groups past the 31st
            const uint32_t m31 = mask & 0x7fffffffu, pc = __popc(m31);
            uint32_t g;
            if (o < pc) {
              uint32_t m = m31;
              for (uint32_t q = 0; q < o; q++) m &= m - 1;
              g = __ffs(m) - 1;
            } else {
              g = 31 + (o - pc);
            }
            j = (g0 + g) * WB_CHUNK + (lane & (WB_CHUNK - 1));
            test = j >= b && j < e;
          }
          if (test) {
            const double qx = T.px[j], qy = T.py[j], qz = T.pz[j];
            const bool in = (phase != 0 && (mask >> 31)) || wb::hyp_in(H, qx, qy, qz);
            if (COUNT) {
              c_test++;
              if (in) c_in++;
            }
            if (in && (qx != x || qy != y)) {
              const int bin = wb::dir_bin(qy - y, qx - x, ttan, tcos, tsin, A.K, &nexact);
              if (phase < 2) {
                occ |= 1ull << bin;
              } else if (bin == before[0] || bin == afterb[0] || bin == before[1] ||
                         bin == afterb[1]) {
                const int pos =
                    dir_bin_pos(wb::atan2i_ext(qy - y, qx - x, ttan, tcos, tsin, A.K, &nexact));
                for (int q = 0; q < 2; q++) {
                  if (bin == before[q] && pos > hi[q]) hi[q] = pos;
                  if (bin == afterb[q] && pos < lo[q]) lo[q] = pos;
                }
              }
            }
          }
        }
        if (phase < 2) {
          const uint64_t occw = wave_or(occ);
          if (occw != checked) {
            checked = occw;
            if (gap_verdict(occw) == 1) {
              verdict = 1;
              done = true;
            }
            __syncthreads();
            if (lane == 0) wb::care_wedges(occw, wg, A.edges);
            __syncthreads();
          }
        }
      }
      __syncthreads();
    }
  }
  if (verdict == 0) verdict = gap_verdict(wave_or(occ));
  if (verdict == 4) {
    bool pass = true;
    for (int k = 0; k < nr; k++) {
      int h = hi[k], l = lo[k];
      for (int s = 1; s < 64; s <<= 1) {
        h = max(h, __shfl_xor(h, s, 64));
        l = min(l, __shfl_xor(l, s, 64));
      }
      const int gap = (int)(((uint32_t)l - (uint32_t)h) & 0x7fffffffu);
      if (gap >= wb::DEG144) pass = false;
    }
    verdict = pass ? 1 : 2;
    if (lane == 0) atomicAdd(A.nslow, 1u);
  }
  for (int s = 1; s < 64; s <<= 1) nexact += __shfl_xor(nexact, s, 64);
  if (COUNT)
    for (int s = 1; s < 64; s <<= 1) {
      c_test += __shfl_xor(c_test, s, 64);
      c_in += __shfl_xor(c_in, s, 64);
    }
  if (lane == 0) {
    work_verdict(A, i, slot, (uint8_t)verdict);
    if (nexact) atomicAdd(A.atan_exact, (unsigned long long)nexact);
    if (COUNT) {
      // per start walk: points, passes, leaves, slots, expansions, ticks,
      // slots tested, slots inside the shape
      unsigned long long *h =
          A.counters + (blk % CT_COPIES) * CT_STRIDE + CT_N + 8 * hr.start;
      atomicAdd(&h[0], 1ull);
      if (verdict == 1) atomicAdd(&h[1], 1ull);
      atomicAdd(&h[2], (unsigned long long)c_leaf);
      atomicAdd(&h[3], (unsigned long long)c_pts);
      atomicAdd(&h[4], (unsigned long long)c_exp);
      atomicAdd(&h[5], (unsigned long long)(wall_clock64() - t_start));
      atomicAdd(&h[6], (unsigned long long)c_test);
      atomicAdd(&h[7], (unsigned long long)c_in);
    }
  }
}


// The handed-over points, as many as the ground kernel counted (*nheavy,
// read on the device, so the host does not wait for the count): block b
// takes record b, then the next record not yet taken (*next counts the
// records taken beyond the first gridDim.x; zero before the launch).
template <bool COUNT>
__global__ void __launch_bounds__(64, WB_HEAVY_MINW) k_ground_heavy(ClassifyArgs A, const HeavyRec *heavy,
                                                                    const uint32_t *nheavy,
                                                                    uint32_t *next)
{
  __shared__ HeavyShared SH;
  __shared__ uint32_t take;
  const uint32_t nh = *nheavy;
  uint32_t r = blockIdx.x;
  while (r < nh) {
    heavy_body<COUNT>(A, heavy[r], r, SH);
    __syncthreads();
    if (threadIdx.x == 0) take = atomicAdd(next, 1u) + gridDim.x;
    __syncthreads();
    r = take;
  }
}

// The heavy points of several trees in one launch: blocks hoff[k] ..
// hoff[k+1]-1 take the heavy records of tree k (AA[k].heavy).
__global__ void __launch_bounds__(64, WB_HEAVY_MINW) k_ground_heavy_batch(const ClassifyArgs *AA,
                                                                          const uint32_t *hoff,
                                                                          int ntrees)
{
  __shared__ HeavyShared SH;
  if (blockIdx.x >= hoff[ntrees]) return;
  int lo = 0, hi = ntrees;  // hoff[lo] <= blockIdx.x < hoff[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (hoff[mid] <= blockIdx.x) lo = mid;
    else hi = mid;
  }
  const ClassifyArgs &A = AA[lo];
  heavy_body<false>(A, A.heavy[blockIdx.x - hoff[lo]], blockIdx.x, SH);
}

// an input merged into an earlier point at its location gets that point's class
__global__ void k_dup_class(uint64_t ndup, const uint32_t *pairs, uint8_t *out)
{
  uint64_t d = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= ndup) return;
  out[pairs[2 * d]] = out[pairs[2 * d + 1]];
}

// class histogram (per-block partials, summed by k_sum_hist) and, when tile
// records are wanted, the number of ground points each deciding tile
// classified
// Class totals, and Tile::nGround (classify.cpp:151-155): CLASSIFY tests
// every point of a tile with that tile's hyperboloid size and counts the
// ground verdicts, so a point in several tiles counts once in each tile
// that finds it ground.  The point's final class is its deciding tile's
// verdict, which every tile of the same size shares; the other tiles'
// verdicts come from k_ground over work items (k_overlap_items,
// k_item_ground).
__global__ void __launch_bounds__(256) k_count_ground(uint64_t n, const uint8_t *cls,
                                                       const int32_t *memb, const uint32_t *nmemb,
                                                       const TileRec *rec, int lo,
                                                       int32_t *nground, uint32_t *partial,
                                                       int per_tile)
{
  __shared__ uint32_t h[3];
  if (threadIdx.x < 3) h[threadIdx.x] = 0;
  __syncthreads();
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint8_t c = i < n ? cls[i] : 255;
  for (int k = 0; k < 3; k++) {
    uint64_t b = __ballot(c == k);
    if (__lane_id() == 0 && b) atomicAdd(&h[k], (uint32_t)__popcll(b));
  }
  if (per_tile && c == 2) {
    int w = INT32_MIN;
    for (uint32_t k = 0; k < nmemb[i]; k++) w = max(w, memb[MAX_MEMB * i + k]);
    const double hw = rec[w - lo].hs;
    for (uint32_t k = 0; k < nmemb[i]; k++) {
      const int t = memb[MAX_MEMB * i + k];
      if (t == w || rec[t - lo].hs == hw) atomicAdd(&nground[t - lo], 1);
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) partial[3 * (uint64_t)blockIdx.x + threadIdx.x] = h[threadIdx.x];
}

// A point's tiles whose hyperboloid size differs from its deciding tile's:
// with items == nullptr, off[i] receives their number; otherwise the items
// (slot, tile record index) are written from off[i] on.
__global__ void k_overlap_items(uint64_t n, const int32_t *memb, const uint32_t *nmemb,
                                const TileRec *rec, int lo, uint32_t *off, uint2 *items)
{
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t m = nmemb[i];
  int w = INT32_MIN;
  for (uint32_t k = 0; k < m; k++) w = max(w, memb[MAX_MEMB * i + k]);
  uint32_t c = 0;
  const uint32_t base = items ? off[i] : 0;
  for (uint32_t k = 0; k < m; k++) {
    const int t = memb[MAX_MEMB * i + k];
    if (t != w && rec[t - lo].hs != rec[w - lo].hs) {
      if (items) items[base + c] = make_uint2((uint32_t)i, (uint32_t)(t - lo));
      c++;
    }
  }
  if (!items) off[i] = c;
}

__global__ void k_item_ground(uint64_t nitems, const uint2 *items, const uint8_t *item_cls,
                              int32_t *nground)
{
  const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nitems && item_cls[k] == 2) atomicAdd(&nground[items[k].y], 1);
}

__global__ void __launch_bounds__(256) k_sum_hist(const uint32_t *partial, uint32_t nblocks,
                                                   unsigned long long *hist3)
{
  uint64_t s[3] = {0, 0, 0};
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nblocks; b += gridDim.x * blockDim.x)
    for (int k = 0; k < 3; k++) s[k] += partial[3 * (uint64_t)b + k];
  for (int k = 0; k < 3; k++) {
    uint64_t v = s[k];
    for (int sh = 1; sh < 64; sh <<= 1) v += __shfl_xor(v, sh, 64);
    if (__lane_id() == 0) atomicAdd(&hist3[k], (unsigned long long)v);
  }
}

// ---------------------------------------------------------------------------
// host drivers
// ---------------------------------------------------------------------------
uint32_t *wb_query_order(wb_tree *t, int kind, const double *dparams, uint64_t ns)
{
  wb_ctx *c = t->ctx;
  uint32_t *key = (uint32_t *)wb_alloc(c, 4 * ns);
  uint32_t *val = (uint32_t *)wb_alloc(c, 4 * ns);
  const double h = t->side / 2;
  hipLaunchKernelGGL(k_query_keys, dim3(wb_blocks(ns, 256)), dim3(256), 0, c->stream, dparams, ns,
                     kind, t->center[0] - h, t->center[1] - h, t->center[2] - h,
                     (double)(1u << WB_MAX_LEVEL) / t->side, key, val);
  wb_radix_sort_u32_prefix(c, key, val, ns, 4);
  wb_free(c, key, 4 * ns);
  return val;
}

void wb_run_count(wb_tree *t, int kind, const double *dparams, uint64_t ns, uint64_t *dcounts,
                  const uint32_t *order)
{
  if (ns == 0) return;
  hipStream_t st = t->ctx->stream;
  if (!order) {
    hipLaunchKernelGGL(k_count_points_in, dim3((unsigned)ns), dim3(64), 0, st, view(t), kind,
                       dparams, ns, dcounts);
    return;
  }
  const dim3 grid(wb_blocks(ns, 256)), blk(256);
  if (kind == wb::SHAPE_CYLINDER)
    hipLaunchKernelGGL(k_query_count<wb::SHAPE_CYLINDER>, grid, blk, 0, st, view(t), dparams,
                       order, ns, dcounts);
  else if (kind == wb::SHAPE_SPHERE)
    hipLaunchKernelGGL(k_query_count<wb::SHAPE_SPHERE>, grid, blk, 0, st, view(t), dparams,
                       order, ns, dcounts);
  else
    hipLaunchKernelGGL(k_query_count<wb::SHAPE_HYPERBOLOID>, grid, blk, 0, st, view(t), dparams,
                       order, ns, dcounts);
}

void wb_run_points(wb_tree *t, int kind, const double *dparams, uint64_t ns, const uint64_t *doffs,
                   uint32_t *dout, const uint32_t *order)
{
  if (ns == 0) return;
  hipStream_t st = t->ctx->stream;
  if (!order) {
    hipLaunchKernelGGL(k_points_in, dim3((unsigned)ns), dim3(64), 0, st, view(t), kind, dparams,
                       ns, doffs, dout);
    return;
  }
  const dim3 grid(wb_blocks(ns, 256)), blk(256);
  if (kind == wb::SHAPE_CYLINDER)
    hipLaunchKernelGGL(k_query_points<wb::SHAPE_CYLINDER>, grid, blk, 0, st, view(t), dparams,
                       order, ns, doffs, dout);
  else if (kind == wb::SHAPE_SPHERE)
    hipLaunchKernelGGL(k_query_points<wb::SHAPE_SPHERE>, grid, blk, 0, st, view(t), dparams,
                       order, ns, doffs, dout);
  else
    hipLaunchKernelGGL(k_query_points<wb::SHAPE_HYPERBOLOID>, grid, blk, 0, st, view(t), dparams,
                       order, ns, doffs, dout);
}

// The state of one tree's classification pass between its phases:
// classify_prepare (tile geometry, membership, SCAN, POSTSCAN, the ground
// test's arguments), the ground test (one tree, or several in one launch),
// classify_finish (counts, Tile::nGround, tile records, timings, release).
struct ClassifyRun {
  wb_tree *t;
  wb_ctx *c;  // allocations and stream (the tree's context, or a batch's)
  hipStream_t st;
  const wb_classify_params *p;
  uint8_t *d_class_out;
  bool want_tiles, scan_only;
  ClassifyOut *out;
  uint64_t n;
  TileGeo g;
  hipEvent_t ev[8];
  int32_t *tex, *tey, *memb, *grid;
  uint32_t *tile_cnt, *tile_fill, *nmemb, *err, *nslow, *tile_off, *lists;
  TileRec *rec;
  unsigned long long *ctr, *dcount;
  uint64_t gcells, total;
  uint8_t *cls;
  HeavyRec *heavy;
  ClassifyArgs A;
  uint32_t h_nheavy, h_nslow;
  uint32_t *heavy_next;  // records the heavy kernel has taken beyond its grid
};

// The flowsnake coordinates of tiles lo .. lo+ntiles-1 and the address
// grid over them (k_tile_geo, k_tile_grid) depend on nothing else, so the
// context keeps those of its last WB_GEO_CACHE tilings: a pass over a tiling
// seen before neither recomputes them nor waits for the grid's bounds.
static const wb_ctx::TileGeoCache &tile_geometry(wb_ctx *c, const TileGeo &g0)
{
  for (size_t k = 0; k < c->geo.size(); k++)
    if (c->geo[k].lo == g0.lo && c->geo[k].ntiles == g0.ntiles) {
      if (k + 1 != c->geo.size()) std::rotate(c->geo.begin() + k, c->geo.begin() + k + 1, c->geo.end());
      return c->geo.back();
    }
  hipStream_t st = c->stream;
  wb_ctx::TileGeoCache e;
  e.lo = g0.lo;
  e.ntiles = g0.ntiles;
  e.tex = (int32_t *)wb_alloc(c, 4ull * g0.ntiles);
  e.tey = (int32_t *)wb_alloc(c, 4ull * g0.ntiles);
  int *dbounds = (int *)wb_alloc(c, 16);
  hipLaunchKernelGGL(k_bounds_init, dim3(1), dim3(64), 0, st, dbounds);
  hipLaunchKernelGGL(k_tile_geo, dim3(wb_blocks(g0.ntiles, 256)), dim3(256), 0, st, g0, e.tex,
                     e.tey, dbounds);
  int *hb = (int *)(c->hbox + WB_HBOX_GEO);
  WB_CHECK(hipMemcpyAsync(hb, dbounds, 16, hipMemcpyDeviceToHost, st));
  WB_CHECK(hipStreamSynchronize(st));
  wb_free(c, dbounds, 16);
  TileGeo g = g0;
  e.gx0 = g.gx0 = hb[0];
  e.gy0 = g.gy0 = hb[2];
  e.gw = g.gw = hb[1] - hb[0] + 1;
  e.gh = g.gh = hb[3] - hb[2] + 1;
  const uint64_t gcells = (uint64_t)g.gw * (uint64_t)g.gh;
  e.grid = (int32_t *)wb_alloc(c, 4 * gcells);
  hipLaunchKernelGGL(k_fill_i32, dim3(wb_blocks(gcells, 256)), dim3(256), 0, st, e.grid, gcells, -1);
  hipLaunchKernelGGL(k_tile_grid, dim3(wb_blocks(g.ntiles, 256)), dim3(256), 0, st, g, e.tex, e.tey,
                     e.grid);
  if (c->geo.size() >= WB_GEO_CACHE) {
    // the oldest tiling's arrays go back to the pool (stream-ordered reuse)
    wb_ctx::TileGeoCache &o = c->geo.front();
    wb_free(c, o.tex, 4ull * o.ntiles);
    wb_free(c, o.tey, 4ull * o.ntiles);
    wb_free(c, o.grid, 4 * (uint64_t)o.gw * (uint64_t)o.gh);
    c->geo.erase(c->geo.begin());
  }
  c->geo.push_back(e);
  return c->geo.back();
}

static bool classify_count_work()
{
  static const bool on = getenv("WB_CLASSIFY_COUNTERS") != nullptr;
  return on;
}

static void classify_prepare(ClassifyRun &R, wb_tree *t, wb_ctx *c, const double *corners,
                             int ncorners, const wb_classify_params *p, uint8_t *d_class_out,
                             bool want_tiles, bool scan_only, ClassifyOut *out, bool counters)
{
  R.t = t; R.c = c; R.st = c->stream; R.p = p; R.d_class_out = d_class_out;
  R.want_tiles = want_tiles; R.scan_only = scan_only; R.out = out;
  R.cls = nullptr; R.heavy = nullptr; R.dcount = nullptr; R.h_nheavy = 0; R.h_nslow = 0;
  hipStream_t st = R.st;
  const uint64_t n = R.n = t->n;
  wb::SnakeSetup ss = wb::snake_setup_corners(corners, ncorners, p->tile_size);
  TileGeo &g = R.g;
  g.scx = ss.cx; g.scy = ss.cy; g.spacing = ss.spacing; g.r = ss.radius;
  g.lo = ss.lo; g.hi = ss.hi;
  g.ntiles = (uint32_t)((int64_t)ss.hi - ss.lo + 1);
  for (int k = 0; k < 8; k++) R.ev[k] = wb_event_get(c);
  hipEvent_t *ev = R.ev;
  WB_CHECK(hipEventRecord(ev[0], st));
  const wb_ctx::TileGeoCache &geo = tile_geometry(c, g);
  int32_t *tex = R.tex = geo.tex;
  int32_t *tey = R.tey = geo.tey;
  uint32_t *tile_cnt = R.tile_cnt = (uint32_t *)wb_alloc(c, 4ull * g.ntiles + 4);
  uint32_t *tile_fill = R.tile_fill = (uint32_t *)wb_alloc(c, 4ull * g.ntiles + 4);
  TileRec *rec = R.rec = (TileRec *)wb_alloc(c, sizeof(TileRec) * (size_t)g.ntiles);
  int32_t *memb = R.memb = (int32_t *)wb_alloc(c, 4ull * MAX_MEMB * n + 16);
  uint32_t *nmemb = R.nmemb = (uint32_t *)wb_alloc(c, 4ull * n + 4);
  // err[0] error flags; then u64 counters: 0 sector borderline, 1 exact atan2i
  // steps, 2-4 class histogram, 5 non-empty tiles, 6 (two u32) exact gap
  // points and heavy points, 7 FP64 fit fallbacks; u32 18-19 work counters of
  // the heavy and resident ground kernels, 20-25 the scan's class list lengths
  uint32_t *err = R.err = (uint32_t *)wb_alloc(c, 128);
  unsigned long long *ctr = R.ctr = (unsigned long long *)err + 1;
  uint32_t *nslow = R.nslow = (uint32_t *)(ctr + 6);
  R.heavy_next = err + 18;
  WB_CHECK(hipMemsetAsync(err, 0, 128, st));
  WB_CHECK(hipMemsetAsync(tile_cnt, 0, 4ull * g.ntiles, st));
  WB_CHECK(hipMemsetAsync(tile_fill, 0, 4ull * g.ntiles, st));
  TreeView T = view(t);
  g.gx0 = geo.gx0;
  g.gy0 = geo.gy0;
  g.gw = geo.gw;
  g.gh = geo.gh;
  g.grid = R.grid = geo.grid;
  LeafWin *lwin = (LeafWin *)wb_alloc(c, sizeof(LeafWin) * ((size_t)t->nleaves + 1));
  uint8_t *lflags = (uint8_t *)wb_alloc(c, (size_t)LEAF_WIN * t->nleaves + 16);
  if (t->nleaves)
    hipLaunchKernelGGL(k_leaf_tiles, dim3(t->nleaves), dim3(64), 0, st, T, g, lwin, lflags);
  if (n > 0)
    hipLaunchKernelGGL(k_membership, dim3(wb_blocks(n, 256)), dim3(256), 0, st, T, g, lwin,
                       lflags, memb, nmemb, tile_cnt, err);
  wb_free(c, lwin, sizeof(LeafWin) * ((size_t)t->nleaves + 1));
  wb_free(c, lflags, (size_t)LEAF_WIN * t->nleaves + 16);
  // tile lists in canonical order
  uint32_t *tile_off = R.tile_off = (uint32_t *)wb_alloc(c, 4ull * g.ntiles + 4);
  WB_CHECK(hipMemcpyAsync(tile_off, tile_cnt, 4ull * g.ntiles, hipMemcpyDeviceToDevice, st));
  const uint64_t total = R.total = wb_exclusive_scan_u32(c, tile_off, g.ntiles);
  uint32_t *lists = R.lists = (uint32_t *)wb_alloc(c, 4 * total + 4);
  if (n > 0)
    hipLaunchKernelGGL(k_tile_lists, dim3(wb_blocks(n, 256)), dim3(256), 0, st, n, memb, nmemb,
                       g.lo, t->crank, tile_off, tile_fill, lists);
  WB_CHECK(hipEventRecord(ev[1], st));
  {
    // the size classes' tile lists, then one wave per tile of each class,
    // grid-stride over its list
    uint32_t *la = tile_fill;  // free until the ground counts
    uint32_t *lb = (uint32_t *)wb_alloc(c, 4ull * g.ntiles + 4);
    uint32_t *lc = (uint32_t *)wb_alloc(c, 4ull * g.ntiles + 4);
    uint32_t *ncls = err + 20;
    const unsigned nb = (unsigned)std::min<uint64_t>(g.ntiles, 1u << 16);
    const int prec = p->fit_precision;
    if (nb) {
      hipLaunchKernelGGL(k_scan_classes, dim3(wb_blocks(g.ntiles, 1024)), dim3(1024), 0, st,
                         g.ntiles, tile_cnt, la, lb, lc, ncls, rec);
      const uint32_t e = g.ntiles - 1;
      const ClassList c16 = {la, 1, ncls}, c32 = {lc, 1, ncls + 1}, c64 = {lb, 1, ncls + 2};
      const ClassList c128 = {lc + e, -1, ncls + 3}, c256 = {lb + e, -1, ncls + 4};
      const ClassList cbig = {la + e, -1, ncls + 5};
#define WB_SCAN_REG(K, F32, CL)                                                                   \
  hipLaunchKernelGGL((k_scan_reg<K, F32>), dim3(nb), dim3(64), 0, st, T, g, tex, tey, tile_off,   \
                     tile_cnt, lists, CL, p->min_hs, prec, rec, ctr)
      if (prec == WB_FIT_FP64) {
        // tiles of up to 16 and 32 points four and two to a wave
        hipLaunchKernelGGL(k_scan_row<16>, dim3(nb), dim3(64), 0, st, T, g, tex, tey, tile_off,
                           tile_cnt, lists, c16, p->min_hs, rec, ctr);
        hipLaunchKernelGGL(k_scan_row<32>, dim3(nb), dim3(64), 0, st, T, g, tex, tey, tile_off,
                           tile_cnt, lists, c32, p->min_hs, rec, ctr);
        WB_SCAN_REG(1, false, c64);
        WB_SCAN_REG(2, false, c128);
        WB_SCAN_REG(4, false, c256);
      } else {
        WB_SCAN_REG(1, true, c16);
        WB_SCAN_REG(1, true, c32);
        WB_SCAN_REG(1, true, c64);
        WB_SCAN_REG(2, true, c128);
        WB_SCAN_REG(4, true, c256);
      }
#undef WB_SCAN_REG
      hipLaunchKernelGGL(k_scan, dim3(std::min(nb, 4096u)), dim3(64), 0, st, T, g, tex, tey,
                         tile_off, tile_cnt, lists, cbig, p->min_hs, prec, rec, err, ctr);
    }
    wb_free(c, lb, 4ull * g.ntiles + 4);
    wb_free(c, lc, 4ull * g.ntiles + 4);
  }
  hipLaunchKernelGGL(k_postscan, dim3(wb_blocks(g.ntiles, 256)), dim3(256), 0, st, g, tex, tey, rec,
                     ctr + 5);
  WB_CHECK(hipEventRecord(ev[2], st));
  for (int k = 4; k < 8; k++) WB_CHECK(hipEventRecord(ev[k], st));
  if (scan_only) return;
  R.cls = (uint8_t *)wb_alloc(c, n + 16);
  ClassifyArgs &A = R.A;
  A.T = T; A.g = g; A.memb = memb; A.nmemb = nmemb; A.rec = rec;
  A.thickness = p->thickness; A.max_slope = p->max_slope;
  A.cube_wedges = t->outside_root ? 0 : 1;
  A.tab = c->d_tan; A.K = c->kconst; A.cls = R.cls; A.cls_in = d_class_out; A.nslow = nslow;
  A.edges = c->d_edges;
  A.atan_exact = ctr + 1;
  A.items = nullptr;
  A.n_items = 0;
  A.item_cls = nullptr;
  if (counters) {
    R.dcount = (unsigned long long *)wb_alloc(c, CT_STRIDE * CT_COPIES * 8);
    WB_CHECK(hipMemsetAsync(R.dcount, 0, CT_STRIDE * CT_COPIES * 8, st));
  }
  A.counters = R.dcount;
  R.heavy = (HeavyRec *)wb_alloc(c, sizeof(HeavyRec) * (n + 1));
  A.heavy = R.heavy;
  A.nheavy = nslow + 1;
}

// the ground test of one tree: the main kernel, then the heavy kernel for
// the points it hands over
static void classify_ground_single(ClassifyRun &R)
{
  hipStream_t st = R.st;
  ClassifyArgs &A = R.A;
  const bool count_work = R.dcount != nullptr;
  WB_CHECK(hipEventRecord(R.ev[4], st));
  if (R.n > 0) {
    const uint32_t nb = wb_blocks(R.n, 256);
    if (count_work)
      hipLaunchKernelGGL((k_ground<true, WB_GROUND_MINW>), dim3(nb), dim3(256), 0, st, A);
    else if (R.c->sharing)
      // 4 of the 6 blocks a CU holds: a wave slot with 192 VGPRs and 2
      // blocks' LDS stay free on every SIMD for other contexts' kernels
      hipLaunchKernelGGL((k_ground_resident<WB_GROUND_MINW>),
                         dim3(std::min<uint32_t>(nb, 4u * (uint32_t)R.c->ncu)), dim3(256), 0, st,
                         A, R.err + 19);
    else
      hipLaunchKernelGGL((k_ground<false, WB_GROUND_MINW>), dim3(nb), dim3(256), 0, st, A);
  }
  WB_CHECK(hipEventRecord(R.ev[5], st));
  WB_CHECK(hipEventRecord(R.ev[6], st));
  if (R.n > 0) {
    // at most one block per handed-over point, and no more than fit on the
    // device at once; the count is read on the device
    // sharing: 8 blocks per CU, about half the LDS the kernel could hold
    const unsigned nb = (unsigned)std::min<uint64_t>(
        R.n, R.c->sharing ? 8ull * (uint64_t)R.c->ncu : (uint64_t)HEAVY_GRID);
    if (count_work)
      hipLaunchKernelGGL(k_ground_heavy<true>, dim3(nb), dim3(64), 0, st, A, R.heavy, A.nheavy,
                         R.heavy_next);
    else
      hipLaunchKernelGGL(k_ground_heavy<false>, dim3(nb), dim3(64), 0, st, A, R.heavy, A.nheavy,
                         R.heavy_next);
  }
  WB_CHECK(hipEventRecord(R.ev[7], st));
}

static void classify_finish(ClassifyRun &R)
{
  wb_tree *t = R.t;
  wb_ctx *c = R.c;
  hipStream_t st = R.st;
  const uint64_t n = R.n;
  const TileGeo &g = R.g;
  hipEvent_t *ev = R.ev;
  ClassifyOut *out = R.out;
  const bool want_tiles = R.want_tiles, scan_only = R.scan_only;
  int32_t *memb = R.memb;
  uint32_t *nmemb = R.nmemb, *nslow = R.nslow, *err = R.err;
  TileRec *rec = R.rec;
  unsigned long long *ctr = R.ctr, *dcount = R.dcount;
  uint8_t *cls = R.cls, *d_class_out = R.d_class_out;
  const ClassifyArgs &A = R.A;
  int32_t *nground = (int32_t *)R.tile_fill;  // reused once the lists are built
  if (!scan_only) {
    WB_CHECK(hipMemsetAsync(nground, 0, 4ull * g.ntiles, st));
    if (n > 0) {
      const unsigned nbk = wb_blocks(n, 256);
      uint32_t *partial = (uint32_t *)wb_alloc(c, 12ull * nbk);
      hipLaunchKernelGGL(k_count_ground, dim3(nbk), dim3(256), 0, st, n, cls, memb, nmemb, rec,
                         g.lo, nground, partial, want_tiles ? 1 : 0);
      hipLaunchKernelGGL(k_sum_hist, dim3(std::min(wb_blocks(nbk, 256), 64u)), dim3(256), 0, st,
                         partial, nbk, ctr + 2);
      wb_free(c, partial, 12ull * nbk);
      if (t->ndup)
        hipLaunchKernelGGL(k_dup_class, dim3(wb_blocks(t->ndup, 256)), dim3(256), 0, st, t->ndup,
                           t->dup_pairs, d_class_out);
    }
    if (want_tiles && n > 0) {
      // Tile::nGround of the tiles that do not decide a point they hold:
      // the ground test again with their hyperboloid size, for the
      // (point, tile) pairs whose size differs from the deciding tile's
      uint32_t *off = (uint32_t *)wb_alloc(c, 4 * n + 4);
      hipLaunchKernelGGL(k_overlap_items, dim3(wb_blocks(n, 256)), dim3(256), 0, st, n, memb,
                         nmemb, rec, g.lo, off, nullptr);
      const uint64_t nitems = wb_exclusive_scan_u32(c, off, n);
      if (nitems > 0xffffffffull) {
        wb_free(c, off, 4 * n + 4);
        throw std::string("more than 2^32-1 (point, tile) pairs to recount");
      }
      if (nitems) {
        uint2 *items = (uint2 *)wb_alloc(c, 8 * nitems);
        uint8_t *icls = (uint8_t *)wb_alloc(c, nitems + 16);
        HeavyRec *iheavy = (HeavyRec *)wb_alloc(c, sizeof(HeavyRec) * (nitems + 1));
        // the pass's own counters: exact gap points, handed-over items,
        // emulated atan2i steps (the statistics keep the main pass's)
        unsigned long long *ict = (unsigned long long *)wb_alloc(c, 32);
        WB_CHECK(hipMemsetAsync(ict, 0, 32, st));
        hipLaunchKernelGGL(k_overlap_items, dim3(wb_blocks(n, 256)), dim3(256), 0, st, n, memb,
                           nmemb, rec, g.lo, off, items);
        ClassifyArgs B = A;
        B.items = items;
        B.n_items = nitems;
        B.item_cls = icls;
        B.heavy = iheavy;
        B.nslow = (uint32_t *)ict;
        B.nheavy = (uint32_t *)ict + 1;
        B.atan_exact = ict + 1;
        B.counters = nullptr;
        hipLaunchKernelGGL((k_ground<false, WB_GROUND_MINW>), dim3(wb_blocks(nitems, 256)),
                           dim3(256), 0, st, B);
        hipLaunchKernelGGL(k_ground_heavy<false>,
                           dim3((unsigned)std::min<uint64_t>(nitems, HEAVY_GRID)), dim3(64), 0, st,
                           B, iheavy, B.nheavy, (uint32_t *)ict + 4);
        hipLaunchKernelGGL(k_item_ground, dim3(wb_blocks(nitems, 256)), dim3(256), 0, st, nitems,
                           items, icls, nground);
        WB_CHECK(hipStreamSynchronize(st));
        wb_free(c, items, 8 * nitems);
        wb_free(c, icls, nitems + 16);
        wb_free(c, iheavy, sizeof(HeavyRec) * (nitems + 1));
        wb_free(c, ict, 32);
      }
      wb_free(c, off, 4 * n + 4);
    }
  }
  WB_CHECK(hipEventRecord(ev[3], st));
  uint32_t *herr = c->hbox + WB_HBOX_ERR;
  WB_CHECK(hipMemcpyAsync(herr, err, 128, hipMemcpyDeviceToHost, st));
  if (want_tiles) {
    out->tiles.resize(g.ntiles);
    out->tex.resize(g.ntiles);
    out->tey.resize(g.ntiles);
    std::vector<int32_t> ng(g.ntiles, 0);
    WB_CHECK(hipMemcpyAsync(out->tiles.data(), rec, sizeof(TileRec) * (size_t)g.ntiles, hipMemcpyDeviceToHost, st));
    WB_CHECK(hipMemcpyAsync(out->tex.data(), R.tex, 4ull * g.ntiles, hipMemcpyDeviceToHost, st));
    WB_CHECK(hipMemcpyAsync(out->tey.data(), R.tey, 4ull * g.ntiles, hipMemcpyDeviceToHost, st));
    if (!scan_only)
      WB_CHECK(hipMemcpyAsync(ng.data(), nground, 4ull * g.ntiles, hipMemcpyDeviceToHost, st));
    WB_CHECK(hipStreamSynchronize(st));
    for (uint32_t k = 0; k < g.ntiles; k++) out->tiles[k].nground = ng[k];
  }
  WB_CHECK(hipStreamSynchronize(st));
  uint32_t h_err[32];
  std::copy(herr, herr + 32, h_err);
  // the exact-gap and heavy point counts (nslow[0], nslow[1])
  R.h_nslow = h_err[14];
  R.h_nheavy = h_err[15];
  float ms1 = 0, ms2 = 0, ms3 = 0;
  hipEventElapsedTime(&ms1, ev[0], ev[1]);
  hipEventElapsedTime(&ms2, ev[1], ev[2]);
  hipEventElapsedTime(&ms3, ev[2], ev[3]);
  float ms4 = 0;
  hipEventElapsedTime(&ms4, ev[4], ev[5]);
  out->ms_classify_kernel = ms4;
  float ms5 = 0;
  hipEventElapsedTime(&ms5, ev[6], ev[7]);
  out->ms_heavy_kernel = ms5;
  for (int k = 0; k < 8; k++) wb_event_put(c, ev[k]);
  unsigned long long *hc = (unsigned long long *)(h_err + 2);
  out->borderline = hc[0];
  out->atan_exact = hc[1];
  out->hist[0] = hc[2]; out->hist[1] = hc[3]; out->hist[2] = hc[4];
  out->nonempty = hc[5];
  out->fallbacks = hc[7];
  out->nslow = R.h_nslow;
  if (dcount) {
    std::vector<unsigned long long> hc((size_t)CT_STRIDE * CT_COPIES);
    WB_CHECK(hipMemcpy(hc.data(), dcount, hc.size() * 8, hipMemcpyDeviceToHost));
    unsigned long long h[CT_STRIDE] = {};
    for (int k = 0; k < CT_STRIDE; k++)
      for (int r = 0; r < CT_COPIES; r++) {
        const unsigned long long v = hc[(size_t)r * CT_STRIDE + k];
        h[k] = k == CT_TMAX ? std::max(h[k], v) : h[k] + v;
      }
    for (int s = 0; s < 3; s++)
      fprintf(stderr,
              "[wb counters] heavy from walk %c: points %llu passed %llu leaves %llu slots %llu "
              "expansions %llu ticks %llu tested %llu inside %llu\n",
              "ABC"[s], h[CT_N + 8 * s], h[CT_N + 8 * s + 1], h[CT_N + 8 * s + 2],
              h[CT_N + 8 * s + 3], h[CT_N + 8 * s + 4], h[CT_N + 8 * s + 5],
              h[CT_N + 8 * s + 6], h[CT_N + 8 * s + 7]);
    fprintf(stderr,
            "[wb counters] lanes A %llu B %llu open after B %llu heavy %llu | A: leaf-lanes %llu "
            "group-lanes %llu dirs %llu | B: leaf-lanes %llu point-lanes %llu dirs %llu | per "
            "wave: expansions A %llu leaves A %llu expansions B %llu leaves B %llu | waves %llu\n",
            h[CT_LANES_A], h[CT_LANES_B], h[CT_LANES_C], h[CT_HEAVY], h[CT_LEAF_A], h[CT_GROUP_A],
            h[CT_DIR_A], h[CT_LEAF_B], h[CT_POINT_B], h[CT_DIR_B], h[CT_EXP_A], h[CT_WLEAF_A],
            h[CT_EXP_B], h[CT_WLEAF_B], h[CT_WAVES]);
    fprintf(stderr,
            "[wb counters] cycles: expansions A %llu leaves A %llu expansions B %llu leaves B "
            "%llu wedge refresh %llu | leaves A: to bounds %llu to samples %llu loop %llu | wave "
            "iterations: samples A %llu slots B %llu\n",
            h[CT_CYC], h[CT_CYC + 1], h[CT_CYC + 2], h[CT_CYC + 3], h[CT_CYC + 4], h[CT_CYC + 5],
            h[CT_CYC + 6], h[CT_CYC + 7], h[CT_ITER_A], h[CT_ITER_B]);
    int rate = 0;
    hipDeviceGetAttribute(&rate, hipDeviceAttributeWallClockRate, c->device);
    fprintf(stderr,
            "[wb counters] wave time: sum %llu max %llu ticks (rate %d kHz); hist counts %llu "
            "%llu %llu %llu %llu ; hist tick sums %llu %llu %llu %llu %llu\n",
            h[CT_TSUM], h[CT_TMAX], rate, h[CT_THIST], h[CT_THIST + 1], h[CT_THIST + 2],
            h[CT_THIST + 3], h[CT_THIST + 4], h[CT_TSUMS], h[CT_TSUMS + 1], h[CT_TSUMS + 2],
            h[CT_TSUMS + 3], h[CT_TSUMS + 4]);
    wb_free(c, dcount, CT_STRIDE * CT_COPIES * 8);
  }
  out->ms_tiles = ms1; out->ms_scan = ms2; out->ms_classify = ms3;
  wb_free(c, R.tile_cnt, 4ull * g.ntiles + 4);
  wb_free(c, R.tile_fill, 4ull * g.ntiles + 4);
  wb_free(c, rec, sizeof(TileRec) * (size_t)g.ntiles);
  wb_free(c, memb, 4ull * MAX_MEMB * n + 16);
  wb_free(c, nmemb, 4ull * n + 4);
  wb_free(c, R.tile_off, 4ull * g.ntiles + 4);
  wb_free(c, R.lists, 4 * R.total + 4);
  if (cls) wb_free(c, cls, n + 16);
  if (R.heavy) wb_free(c, R.heavy, sizeof(HeavyRec) * (n + 1));
  out->nheavy = R.h_nheavy;
  if (h_err[0] & 1u) { wb_free(c, err, 128); throw std::string("a point lies in more than 4 tile cylinders"); }
  if (h_err[0] & 2u) { wb_free(c, err, 128); throw std::string("a tile cylinder holds more than 1024 points"); }
  wb_free(c, err, 128);
}

void wb_run_classify(wb_tree *t, const double *corners, int ncorners,
                     const wb_classify_params *p, uint8_t *d_class_out, bool want_tiles,
                     bool scan_only, ClassifyOut *out)
{
  ClassifyRun R;
  classify_prepare(R, t, t->ctx, corners, ncorners, p, d_class_out, want_tiles, scan_only, out,
                   !scan_only && classify_count_work());
  if (!scan_only) classify_ground_single(R);
  classify_finish(R);
}

void wb_run_classify_batch(wb_ctx *c, wb_tree *const *trees, int ntrees, const double *corners,
                           const wb_classify_params *p, uint8_t *const *d_class_out,
                           ClassifyOut *outs)
{
  hipStream_t st = c->stream;
  std::vector<ClassifyRun> R(ntrees);
  for (int k = 0; k < ntrees; k++)
    classify_prepare(R[k], trees[k], c, corners + 6 * k, 2, p, d_class_out[k], false, false,
                     &outs[k], false);
  // one launch of the ground kernel over every tree's points, then one of
  // the heavy kernel over every tree's handed-over points
  std::vector<ClassifyArgs> AA(ntrees);
  std::vector<uint32_t> boff(ntrees + 1, 0), hoff(ntrees + 1, 0);
  for (int k = 0; k < ntrees; k++) {
    AA[k] = R[k].A;
    const uint64_t nb = boff[k] + (uint64_t)wb_blocks(R[k].n, 256);
    if (nb > 0xffffffffull) throw std::string("too many points in one batch");
    boff[k + 1] = (uint32_t)nb;
  }
  ClassifyArgs *dAA = (ClassifyArgs *)wb_alloc(c, sizeof(ClassifyArgs) * ntrees);
  uint32_t *doff = (uint32_t *)wb_alloc(c, 8 * (ntrees + 1));
  WB_CHECK(hipMemcpyAsync(dAA, AA.data(), sizeof(ClassifyArgs) * ntrees, hipMemcpyHostToDevice, st));
  WB_CHECK(hipMemcpyAsync(doff, boff.data(), 4 * (ntrees + 1), hipMemcpyHostToDevice, st));
  for (int k = 0; k < ntrees; k++) WB_CHECK(hipEventRecord(R[k].ev[4], st));
  if (boff[ntrees])
    hipLaunchKernelGGL((k_ground_batch<WB_GROUND_MINW>), dim3(boff[ntrees]), dim3(256), 0, st, dAA,
                       doff, ntrees);
  for (int k = 0; k < ntrees; k++) {
    WB_CHECK(hipEventRecord(R[k].ev[5], st));
    WB_CHECK(hipMemcpyAsync(&R[k].h_nheavy, R[k].nslow + 1, 4, hipMemcpyDeviceToHost, st));
  }
  WB_CHECK(hipStreamSynchronize(st));
  for (int k = 0; k < ntrees; k++) {
    WB_CHECK(hipEventRecord(R[k].ev[6], st));
    hoff[k + 1] = hoff[k] + R[k].h_nheavy;
  }
  if (hoff[ntrees]) {
    uint32_t *dh = doff + ntrees + 1;
    WB_CHECK(hipMemcpyAsync(dh, hoff.data(), 4 * (ntrees + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_ground_heavy_batch, dim3(hoff[ntrees]), dim3(64), 0, st, dAA, dh,
                       ntrees);
  }
  for (int k = 0; k < ntrees; k++) WB_CHECK(hipEventRecord(R[k].ev[7], st));
  WB_CHECK(hipStreamSynchronize(st));
  wb_free(c, dAA, sizeof(ClassifyArgs) * ntrees);
  wb_free(c, doff, 8 * (ntrees + 1));
  for (int k = 0; k < ntrees; k++) classify_finish(R[k]);
}

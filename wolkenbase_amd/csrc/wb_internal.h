// wb_internal.h -- device data layout and launch helpers of libwb_gpu.
//
// HBM layout of one octree (n points, nl leaves, nn internal nodes):
//   key[n]      u64   Morton key << 1 (levels 1..21, z,y,x bits per level)
//   sidx[n]     u32   input index of each sorted slot
//   px,py,pz[n] f64   decoded coordinates in sorted (Morton-prefix) order
//   pt_leaf[n]  u32   leaf of each sorted slot
//   canon[n]    u32   canonical position -> sorted slot (leaf DFS order, then
//                     input order inside a leaf, as the reference stores it)
//   crank[n]    u32   sorted slot -> canonical position
//   chunk_box[ceil(n/32)]  f64x6  bounding box of each 32-slot group
//   chunk_low[ceil(n/32)]  u32    slot of the group's lowest point (first of
//                                 equal z)
//   chunk_sx/sy/sz[ceil(n/32)] f64 that point's coordinates (the ground
//                                 test's samples, read coalesced per leaf)
//   leaf_start[nl+1] u32, leaf_cube[nl] f64x4 (cx,cy,cz,side), leaf_parent[nl]
//   node_cube[nn] f64x4, node_child[8*nn] i32 (>=0 node, -2-l leaf, -1 empty),
//   node_crange[8*nn] u32x2 (slot range of each leaf child; 0,0 otherwise),
//   node_parent[nn] i32
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <set>
#include <unordered_map>
#include <string>
#include <vector>

#include "../../include/wb_gpu.h"
#include "wb_host.h"
#include "wb_math.h"

#define WB_RECORDS 537  // octree.h:40-41: a leaf holds at most 537 locations
#define WB_MAX_LEVEL 21
#define WB_CHUNK 32
#define WB_HBOX_WORDS 1024
// word offsets of the pinned host words of each use
#define WB_HBOX_SCAN 0     // wb_exclusive_scan_u32: total (2 words)
#define WB_HBOX_GEO 16     // tile grid bounds (4 words)
#define WB_HBOX_ERR 32     // classification flags and counters (32 words)
#define WB_HBOX_BATCH 64   // batch classification: heavy counts per tree
#define WB_GEO_CACHE 4

struct wb_stream_alloc;

struct wb_ctx {
  int device;
  hipStream_t stream;
  std::string err;
  // atan2i tables and constants (device)
  double *d_tan, *d_cos, *d_sin;
  float *d_edges;  // bin-edge directions of the ground test (wb::fill_edge_table)
  // device address of this GPU's copy of the library's static scratch
  // (wb_count_points_in_noalloc)
  void *scratch;
  wb::AtanConst kconst;
  // size-keyed cache of device buffers, and the allocated size of every
  // buffer handed out (a cached buffer up to 1.5x the request is reused)
  std::multimap<size_t, void *> cache;
  std::unordered_map<void *, size_t> held;
  size_t cached_bytes;
  // compute units of the device, and whether the context's long kernels
  // leave room on every CU for other contexts' (wb_ctx_set_sharing)
  int ncu;
  bool sharing;
  // pinned host words the kernels and copies of one call fill for the host
  // (no staging copy, no implicit wait): WB_HBOX_WORDS words
  uint32_t *hbox;
  // tile geometry of recent tilings (flowsnake coordinates of every tile
  // and the address grid over them), a pure function of the first counter
  // and the number of tiles; most recent last
  struct TileGeoCache {
    int lo;
    uint32_t ntiles;
    int32_t *tex, *tey, *grid;
    int gx0, gy0, gw, gh;
  };
  std::vector<TileGeoCache> geo;
  // timing events kept for reuse (creating one costs host time per call)
  std::vector<hipEvent_t> evpool;
  // trees, LAS files and clouds made on this context and not yet freed
  std::set<struct wb_tree *> trees;
  std::set<struct wb_las *> las;
  std::set<struct wb_cloud *> clouds;
};

// ---- implemented in wb_las.hip ----
// frees the clouds and LAS files still alive on a context being destroyed
void wb_release_las(wb_ctx *c);

struct wb_tree {
  wb_ctx *ctx;
  uint64_t n;        // distinct locations in the tree
  uint64_t n_input;  // input points (n plus the duplicates merged away)
  uint64_t ndup;     // inputs merged into an earlier point at the same location
  // nonzero when some point lies outside the closed root cube: such a point
  // sits in an extreme cell, possibly far outside its leaf's cube, so no
  // test may bound a leaf's points by the leaf's cube
  uint32_t outside_root;
  double center[3], side;
  double scale[3], offset[3];
  int sort_passes, maxdepth;
  uint32_t nleaves, nnodes;
  float ms_build, ms_sort;  // device time of the last build (events)
  // device time of the last batched query's kernels (counting, and the
  // index pass of wb_points_in; 0 when it had none)
  float ms_query_count, ms_query_points;
  // the build's timing events (start, sort start, sort end, end), read
  // once the stream has passed them; null when there are none
  hipEvent_t evb[4];
  bool evb_read;
  // device arrays
  uint64_t *key;
  uint32_t *sidx;
  double *px, *py, *pz;
  uint32_t *pt_leaf, *canon, *crank;
  double *chunk_box;
  uint32_t *chunk_low;
  double *chunk_sx, *chunk_sy, *chunk_sz;
  uint32_t *leaf_start;
  double *leaf_cube;
  int32_t *leaf_parent;
  double *node_cube;
  int32_t *node_child, *node_parent;
  uint2 *node_crange;
  // per child slot of every internal node (8 per node): the lowest z of the
  // points stored under the child (+inf for an empty child)
  double *child_zlow;
  // with duplicates merged: (merged input index, input index of the first
  // point at its location) pairs, and per sorted slot the input index of
  // the last point at the slot's location (whose record a writer keeps)
  uint32_t *dup_pairs;
  uint32_t *last_of;
  std::vector<std::pair<void *, size_t>> owned;
};

// device allocation through the context cache
void *wb_alloc(wb_ctx *c, size_t bytes);
// a timing event from the context's pool, and its return
hipEvent_t wb_event_get(wb_ctx *c);
void wb_event_put(wb_ctx *c, hipEvent_t e);
// the build's device times, once the stream has passed its last event
void wb_tree_read_timing(wb_tree *t);
void wb_free(wb_ctx *c, void *p, size_t bytes);
template <typename T>
T *wb_talloc(wb_tree *t, size_t count)
{
  size_t b = count * sizeof(T);
  if (b == 0) b = 16;
  T *p = (T *)wb_alloc(t->ctx, b);
  t->owned.push_back({(void *)p, b});
  return p;
}

#define WB_CHECK(call)                                                              \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      throw std::string(#call) + ": " + hipGetErrorString(e_);                      \
    }                                                                               \
  } while (0)

// per-tile record of the classification pass (device and host)
struct TileRec {
  int32_t npoints, nground, tree_flags, fit_fp64;
  double density, hs_scan, hs, height, tilt_x, tilt_y;
};

struct ClassifyOut {
  std::vector<TileRec> tiles;
  std::vector<int32_t> tex, tey;
  uint64_t borderline, atan_exact, nslow, hist[3], nonempty, nheavy, fallbacks;
  double ms_tiles, ms_scan, ms_classify, ms_classify_kernel, ms_heavy_kernel;
};

// ---- implemented in wb_classify.hip ----
// batched shape queries: with order (a permutation of the shapes, from
// wb_query_order) a lane per shape in that order; without, a wave per shape
// (no allocation)
void wb_run_count(wb_tree *t, int kind, const double *dparams, uint64_t ns, uint64_t *dcounts,
                  const uint32_t *order);
void wb_run_points(wb_tree *t, int kind, const double *dparams, uint64_t ns, const uint64_t *doffs,
                   uint32_t *dout, const uint32_t *order);
// the shapes in Morton order of their centres (ns indices, allocated on the
// context: wb_free(c, order, 4 * ns))
uint32_t *wb_query_order(wb_tree *t, int kind, const double *dparams, uint64_t ns);
// the classification pass (tile geometry, membership, SCAN, POSTSCAN and,
// unless scan_only, CLASSIFY); class bytes to d_class_out in input order
// several trees of one context classified together, their ground tests in
// one launch (tiles as wb_classify with each tree's two header corners,
// corners + 6k); d_class_out[k]: device class bytes of tree k, input order
void wb_run_classify_batch(wb_ctx *c, wb_tree *const *trees, int ntrees, const double *corners,
                           const wb_classify_params *p, uint8_t *const *d_class_out,
                           ClassifyOut *outs);
void wb_run_classify(wb_tree *t, const double *corners, int ncorners,
                     const wb_classify_params *p, uint8_t *d_class_out, bool want_tiles,
                     bool scan_only, ClassifyOut *out);

// ---- building blocks implemented in wb_build.hip ----
// exclusive scan of u32 in place, returns the total (synchronous read-back)
uint64_t wb_exclusive_scan_u32(wb_ctx *c, uint32_t *data, uint64_t n);
// LSD radix sort of (u64 key, u32 val) pairs on bits [64-8*passes, 64)
void wb_radix_sort_pairs(wb_ctx *c, uint64_t *key, uint32_t *val, uint64_t n, int passes);
// LSD radix sort of (u32 key, u32 val) pairs on bits [32-8*passes, 32)
void wb_radix_sort_u32_prefix(wb_ctx *c, uint32_t *key, uint32_t *val, uint64_t n, int passes);
// LSD radix sort of (u32 key, u32 val) pairs on all 32 bits
void wb_radix_sort_u32_pairs(wb_ctx *c, uint32_t *key, uint32_t *val, uint64_t n);

// builds the canonical octree of t->n_input points, decoded from int32 LAS
// coordinates (dx, dy, dz with t->scale / t->offset) or given in FP64 (fx,
// fy, fz; the int32 pointers are then null); with dedupe, points at the
// location of an earlier point are merged into it (sets t->n, t->ndup)
void wb_build_tree(wb_tree *t, const int32_t *dx, const int32_t *dy, const int32_t *dz,
                   const double *fx, const double *fy, const double *fz, bool dedupe);

inline unsigned wb_blocks(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

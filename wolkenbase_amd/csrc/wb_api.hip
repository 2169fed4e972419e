// wb_api.hip -- the C ABI of include/wb_gpu.h.
#include "../../include/wb_gpu.h"
#include "wb_internal.h"

#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>

// ---------------------------------------------------------------------------
// Static scratch of wb_count_points_in_noalloc: the staged shape parameters
// (WB_NOALLOC_MAX_SHAPES x WB_SHAPE_NPARAMS doubles) and counts (as many
// u64) of one call.  Each GPU holds its own copy for the life of the process.
// In host code `scratch` names a host-side stand-in, not device memory: its
// device address on each GPU comes from hipGetSymbolAddress, taken once per
// device with that device current (wb_ctx_create), and only that address is
// passed to copies and kernels.
// ---------------------------------------------------------------------------
#define WB_NOALLOC_PARAM_BYTES ((size_t)WB_NOALLOC_MAX_SHAPES * WB_SHAPE_NPARAMS * sizeof(double))
#define WB_NOALLOC_BYTES (WB_NOALLOC_PARAM_BYTES + (size_t)WB_NOALLOC_MAX_SHAPES * sizeof(uint64_t))
static __device__ __align__(256) unsigned char scratch[WB_NOALLOC_BYTES];

#define WB_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[WB_MAX_DEVICES];

// the device address of `scratch` on `device` (current); nullptr if the
// runtime refuses it
static void *scratch_address(int device)
{
  if (device < 0 || device >= WB_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}


// ---------------------------------------------------------------------------
// device memory cache: buffers are returned to a size-keyed pool and reused,
// so repeated builds of tiles of similar size do not call hipMalloc again
// (the smallest cached buffer of at least the request is taken when it is
// no more than 1.5 times the request, 2 times below 1 MiB)
// ---------------------------------------------------------------------------
static size_t round_size(size_t b)
{
  size_t r = 256;
  while (r < b) r = (r < (1u << 20)) ? r * 2 : r + (r >> 2);
  return r;
}

void *wb_alloc(wb_ctx *c, size_t bytes)
{
  size_t b = round_size(bytes ? bytes : 1);
  auto it = c->cache.lower_bound(b);
  const size_t cap = b < (1u << 20) ? 2 * b : b + (b >> 1);
  if (it != c->cache.end() && it->first <= cap) {
    void *p = it->second;
    const size_t have = it->first;
    c->cache.erase(it);
    c->cached_bytes -= have;
    c->held[p] = have;
    return p;
  }
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, b);
  if (e != hipSuccess) {
    // release the pool and retry once
    WB_CHECK(hipStreamSynchronize(c->stream));
    for (auto &kv : c->cache) hipFree(kv.second);
    c->cache.clear();
    c->cached_bytes = 0;
    (void)hipGetLastError();
    e = hipMalloc(&p, b);
    if (e != hipSuccess) throw std::string("hipMalloc failed for ") + std::to_string(b) + " bytes";
  }
  c->held[p] = b;
  return p;
}

void wb_free(wb_ctx *c, void *p, size_t bytes)
{
  if (!p) return;
  size_t b = round_size(bytes ? bytes : 1);
  auto h = c->held.find(p);
  if (h != c->held.end()) {
    b = h->second;
    c->held.erase(h);
  }
  c->cache.insert({b, p});
  c->cached_bytes += b;
}

hipEvent_t wb_event_get(wb_ctx *c)
{
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  WB_CHECK(hipEventCreate(&e));
  return e;
}

void wb_event_put(wb_ctx *c, hipEvent_t e)
{
  if (e) c->evpool.push_back(e);
}

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

// a tree whose build failed: its buffers back to the pool, its events
// destroyed
static void discard_tree(wb_ctx *c, wb_tree *t)
{
  for (auto &o : t->owned) wb_free(c, o.first, o.second);
  for (hipEvent_t e : t->evb)
    if (e) hipEventDestroy(e);
  delete t;
}

extern "C" {

int wb_ctx_create(int device, wb_ctx **out)
{
  if (!out) return WB_EINVAL;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return WB_EDEVICE;
  wb_ctx *c = new wb_ctx();
  c->device = device;
  c->cached_bytes = 0;
  c->stream = nullptr;
  c->d_tan = nullptr;
  c->d_edges = nullptr;
  c->hbox = nullptr;
  try {
    WB_CHECK(hipSetDevice(device));
    WB_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    double tabs[511 + 512 + 512];
    wb::fill_atan_tables(tabs, tabs + 511, tabs + 1023);
    WB_CHECK(hipMalloc(&c->d_tan, sizeof tabs));
    WB_CHECK(hipMemcpy(c->d_tan, tabs, sizeof tabs, hipMemcpyHostToDevice));
    c->d_cos = c->d_tan + 511;
    c->d_sin = c->d_tan + 1023;
    float edges[3 * 128];
    wb::fill_edge_table(edges);
    WB_CHECK(hipMalloc(&c->d_edges, sizeof edges));
    WB_CHECK(hipMemcpy(c->d_edges, edges, sizeof edges, hipMemcpyHostToDevice));
    c->kconst = wb::atan_const();
    c->scratch = scratch_address(device);
    WB_CHECK(hipHostMalloc((void **)&c->hbox, 4 * WB_HBOX_WORDS, hipHostMallocDefault));
    WB_CHECK(hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device));
    c->sharing = false;
    memset(c->hbox, 0, 4 * WB_HBOX_WORDS);
  } catch (const std::string &m) {
    if (c->hbox) hipHostFree(c->hbox);
    if (c->d_tan) hipFree(c->d_tan);
    if (c->d_edges) hipFree(c->d_edges);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return WB_EDEVICE;
  }
  *out = c;
  return WB_OK;
}

void wb_ctx_destroy(wb_ctx *c)
{
  if (!c) return;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  // clouds, LAS files and trees still alive go with their context
  wb_release_las(c);
  for (wb_tree *t : c->trees) {
    for (auto &o : t->owned) hipFree(o.first);
    for (hipEvent_t e : t->evb)
      if (e) hipEventDestroy(e);
    delete t;
  }
  c->trees.clear();
  for (auto &g : c->geo) {
    hipFree(g.tex);
    hipFree(g.tey);
    hipFree(g.grid);
  }
  c->geo.clear();
  for (hipEvent_t e : c->evpool) hipEventDestroy(e);
  c->evpool.clear();
  for (auto &kv : c->cache) hipFree(kv.second);
  if (c->hbox) hipHostFree(c->hbox);
  hipFree(c->d_tan);
  hipFree(c->d_edges);
  hipStreamDestroy(c->stream);
  delete c;
}

const char *wb_last_error(wb_ctx *c) { return c ? c->err.c_str() : "no context"; }

int wb_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes)
{
  if (!free_bytes || !total_bytes) return WB_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return WB_EDEVICE;
  size_t f = 0, t = 0;
  if (hipSetDevice(device) != hipSuccess || hipMemGetInfo(&f, &t) != hipSuccess) return WB_EDEVICE;
  *free_bytes = f;
  *total_bytes = t;
  return WB_OK;
}

int wb_host_alloc(uint64_t bytes, void **out)
{
  if (!out || bytes == 0) return WB_EINVAL;
  *out = nullptr;
  if (hipHostMalloc(out, (size_t)bytes, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *out = nullptr;
    return WB_EDEVICE;
  }
  return WB_OK;
}

void wb_host_free(void *p)
{
  if (p) hipHostFree(p);
}

int wb_ctx_set_sharing(wb_ctx *c, int on)
{
  if (!c) return WB_EINVAL;
  c->sharing = on != 0;
  return WB_OK;
}

int wb_ctx_sync(wb_ctx *c)
{
  if (!c) return WB_EINVAL;
  return hipStreamSynchronize(c->stream) == hipSuccess ? WB_OK : WB_EDEVICE;
}

int wb_size_fit(const double *corners, int ncorners, double center[3], double *side)
{
  if (!corners || ncorners <= 0 || !center || !side) return WB_EINVAL;
  wb::size_fit(corners, ncorners, center, side);
  return WB_OK;
}

int wb_tree_build(wb_ctx *c, const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n,
                  const wb_lasxform *xf, const double root_center[3], double root_side, int flags,
                  wb_tree **out)
{
  if (!c || !out || !xf || !root_center) return fail(c, "null argument", WB_EINVAL);
  if (n > 0xffffffffull) return fail(c, "more than 2^32-1 points", WB_ERANGE);
  if (n > 0 && (!x || !y || !z)) return fail(c, "null coordinate array", WB_EINVAL);
  if (!(root_side > 0) || !isfinite(root_side)) return fail(c, "root side must be positive", WB_EINVAL);
  if ((root_side / (1 << 21)) * (double)(1 << 21) != root_side)
    return fail(c, "root side is not representable", WB_ERANGE);
  if (flags & ~(WB_DEVICE_INPUT | WB_BUILD_DEDUPE)) return fail(c, "unknown flags", WB_EINVAL);
  *out = nullptr;
  wb_tree *t = new wb_tree();
  t->ctx = c;
  t->n = t->n_input = n;
  for (int a = 0; a < 3; a++) {
    t->center[a] = root_center[a];
    t->scale[a] = xf->scale[a] * xf->unit;
    t->offset[a] = xf->offset[a] * xf->unit;
  }
  t->side = root_side;
  t->sort_passes = 0;
  t->ms_build = t->ms_sort = 0;
  t->ms_query_count = t->ms_query_points = 0;
  if (xf->unit != 1.0) {
    delete t;
    return fail(c, "only unit == 1 is supported (metres)", WB_ERANGE);
  }
  for (int a = 0; a < 3; a++) {
    t->scale[a] = xf->scale[a];
    t->offset[a] = xf->offset[a];
  }
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  const int32_t *dx = x, *dy = y, *dz = z;
  int32_t *tmp = nullptr;
  if (!(flags & WB_DEVICE_INPUT) && n > 0) {
    tmp = (int32_t *)wb_alloc(c, 12 * n);
    WB_CHECK(hipMemcpyAsync(tmp, x, 4 * n, hipMemcpyHostToDevice, c->stream));
    WB_CHECK(hipMemcpyAsync(tmp + n, y, 4 * n, hipMemcpyHostToDevice, c->stream));
    WB_CHECK(hipMemcpyAsync(tmp + 2 * n, z, 4 * n, hipMemcpyHostToDevice, c->stream));
    dx = tmp;
    dy = tmp + n;
    dz = tmp + 2 * n;
  }
  try {
    wb_build_tree(t, dx, dy, dz, nullptr, nullptr, nullptr, (flags & WB_BUILD_DEDUPE) != 0);
  } catch (...) {
    if (tmp) wb_free(c, tmp, 12 * n);
    discard_tree(c, t);
    throw;
  }
  if (tmp) wb_free(c, tmp, 12 * n);
  // the call returns while the build's last kernels run (later calls on the
  // context are ordered after them); host input is copied by then
  if (tmp) WB_CHECK(hipEventSynchronize(t->evb[0]));
  c->trees.insert(t);
  *out = t;
  return WB_OK;
  API_CATCH(c)
}

int wb_tree_build_xyz(wb_ctx *c, const double *x, const double *y, const double *z, uint64_t n,
                      const double resolution[3], const double root_center[3], double root_side,
                      int flags, wb_tree **out)
{
  if (!c || !out || !root_center || !resolution) return fail(c, "null argument", WB_EINVAL);
  if (n > 0xffffffffull) return fail(c, "more than 2^32-1 points", WB_ERANGE);
  if (n > 0 && (!x || !y || !z)) return fail(c, "null coordinate array", WB_EINVAL);
  if (!(root_side > 0) || !isfinite(root_side)) return fail(c, "root side must be positive", WB_EINVAL);
  if ((root_side / (1 << 21)) * (double)(1 << 21) != root_side)
    return fail(c, "root side is not representable", WB_ERANGE);
  for (int a = 0; a < 3; a++)
    if (!(resolution[a] > 0) || !isfinite(resolution[a]))
      return fail(c, "resolution must be positive", WB_EINVAL);
  if (flags & ~(WB_DEVICE_INPUT | WB_BUILD_DEDUPE)) return fail(c, "unknown flags", WB_EINVAL);
  *out = nullptr;
  wb_tree *t = new wb_tree();
  t->ctx = c;
  t->n = t->n_input = n;
  for (int a = 0; a < 3; a++) {
    t->center[a] = root_center[a];
    t->scale[a] = resolution[a];  // only the extent check of the ground test reads it
    t->offset[a] = 0;
  }
  t->side = root_side;
  t->sort_passes = 0;
  t->ms_build = t->ms_sort = 0;
  t->ms_query_count = t->ms_query_points = 0;
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  const double *dx = x, *dy = y, *dz = z;
  double *tmp = nullptr;
  if (!(flags & WB_DEVICE_INPUT) && n > 0) {
    tmp = (double *)wb_alloc(c, 24 * n);
    WB_CHECK(hipMemcpyAsync(tmp, x, 8 * n, hipMemcpyHostToDevice, c->stream));
    WB_CHECK(hipMemcpyAsync(tmp + n, y, 8 * n, hipMemcpyHostToDevice, c->stream));
    WB_CHECK(hipMemcpyAsync(tmp + 2 * n, z, 8 * n, hipMemcpyHostToDevice, c->stream));
    dx = tmp;
    dy = tmp + n;
    dz = tmp + 2 * n;
  }
  try {
    wb_build_tree(t, nullptr, nullptr, nullptr, dx, dy, dz, (flags & WB_BUILD_DEDUPE) != 0);
  } catch (...) {
    if (tmp) wb_free(c, tmp, 24 * n);
    discard_tree(c, t);
    throw;
  }
  if (tmp) wb_free(c, tmp, 24 * n);
  if (tmp) WB_CHECK(hipEventSynchronize(t->evb[0]));
  c->trees.insert(t);
  *out = t;
  return WB_OK;
  API_CATCH(c)
}

int wb_tree_duplicates(wb_tree *t, uint64_t *ndup, uint32_t *pairs)
{
  if (!t || !ndup) return WB_EINVAL;
  wb_ctx *c = t->ctx;
  *ndup = t->ndup;
  if (!pairs || t->ndup == 0) return WB_OK;
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  WB_CHECK(hipMemcpyAsync(pairs, t->dup_pairs, 8 * t->ndup, hipMemcpyDeviceToHost, c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  return WB_OK;
  API_CATCH(c)
}

int wb_tree_free(wb_tree *t)
{
  if (!t) return WB_EINVAL;
  wb_ctx *c = t->ctx;
  hipStreamSynchronize(c->stream);
  for (auto &o : t->owned) wb_free(c, o.first, o.second);
  for (hipEvent_t e : t->evb)
    if (e) hipEventDestroy(e);
  c->trees.erase(t);
  delete t;
  return WB_OK;
}

int wb_tree_query_ms(wb_tree *t, double *count_ms, double *points_ms)
{
  if (!t) return WB_EINVAL;
  if (count_ms) *count_ms = t->ms_query_count;
  if (points_ms) *points_ms = t->ms_query_points;
  return WB_OK;
}

int wb_tree_info(wb_tree *t, uint64_t *nleaves, uint64_t *nnodes, int *maxdepth, int *sort_passes)
{
  if (!t) return WB_EINVAL;
  if (nleaves) *nleaves = t->nleaves;
  if (nnodes) *nnodes = t->nnodes;
  if (maxdepth) *maxdepth = t->maxdepth;
  if (sort_passes) *sort_passes = t->sort_passes;
  return WB_OK;
}

int wb_tree_leaves(wb_tree *t, wb_leaf *out, uint64_t cap, uint64_t *nleaves)
{
  if (!t) return WB_EINVAL;
  wb_ctx *c = t->ctx;
  if (nleaves) *nleaves = t->nleaves;
  if (!out) return WB_OK;
  if (cap < t->nleaves) return fail(c, "leaf buffer too small", WB_EINVAL);
  API_TRY(c)
  std::vector<uint32_t> st(t->nleaves + 1);
  std::vector<double> cube(4ull * t->nleaves + 4);
  WB_CHECK(hipMemcpyAsync(st.data(), t->leaf_start, 4ull * (t->nleaves + 1), hipMemcpyDeviceToHost, c->stream));
  if (t->nleaves)
    WB_CHECK(hipMemcpyAsync(cube.data(), t->leaf_cube, 32ull * t->nleaves, hipMemcpyDeviceToHost, c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  for (uint32_t i = 0; i < t->nleaves; i++) {
    wb_leaf &L = out[i];
    L.center[0] = cube[4 * i];
    L.center[1] = cube[4 * i + 1];
    L.center[2] = cube[4 * i + 2];
    L.side = cube[4 * i + 3];
    L.first = st[i];
    L.count = st[i + 1] - st[i];
    L.depth = (int32_t)lrint(log2(t->side / L.side));
  }
  return WB_OK;
  API_CATCH(c)
}

int wb_tree_order(wb_tree *t, uint32_t *order)
{
  if (!t || !order) return WB_EINVAL;
  wb_ctx *c = t->ctx;
  API_TRY(c)
  std::vector<uint32_t> canon(t->n), sidx(t->n);
  if (t->n) {
    WB_CHECK(hipMemcpyAsync(canon.data(), t->canon, 4 * t->n, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipMemcpyAsync(sidx.data(), t->sidx, 4 * t->n, hipMemcpyDeviceToHost, c->stream));
  }
  WB_CHECK(hipStreamSynchronize(c->stream));
  for (uint64_t k = 0; k < t->n; k++) order[k] = sidx[canon[k]];
  return WB_OK;
  API_CATCH(c)
}

static int check_kind(int kind) { return kind == WB_SHAPE_CYLINDER || kind == WB_SHAPE_SPHERE || kind == WB_SHAPE_HYPERBOLOID; }

// Batches of up to this many shapes take a wave per shape in the caller's
// order: one wave would hold them all in the lane-per-shape kernel, and
// ordering them first would cost more launches than the query.
#define WB_QUERY_SORT_MIN 65

// Events around a query's kernels on the context stream; ms() after the
// stream has passed the second one.
struct QueryTimer {
  wb_ctx *c;
  hipEvent_t e[2];
  explicit QueryTimer(wb_ctx *c_) : c(c_) { e[0] = wb_event_get(c); e[1] = wb_event_get(c); }
  ~QueryTimer() { wb_event_put(c, e[0]); wb_event_put(c, e[1]); }
  void start() { WB_CHECK(hipEventRecord(e[0], c->stream)); }
  void stop() { WB_CHECK(hipEventRecord(e[1], c->stream)); }
  float ms()
  {
    float v = 0;
    WB_CHECK(hipEventElapsedTime(&v, e[0], e[1]));
    return v;
  }
};

int wb_count_points_in(wb_tree *t, int kind, const double *params, uint64_t ns, uint64_t *counts)
{
  if (!t || (!params && ns) || (!counts && ns)) return WB_EINVAL;
  wb_ctx *c = t->ctx;
  if (!check_kind(kind)) return fail(c, "unknown shape kind", WB_EINVAL);
  if (ns > 0x7fffffffull) return fail(c, "too many shapes", WB_ERANGE);
  if (ns == 0) return WB_OK;
  API_TRY(c)
  size_t pb = sizeof(double) * WB_SHAPE_NPARAMS * ns;
  double *dp = (double *)wb_alloc(c, pb);
  uint64_t *dc = (uint64_t *)wb_alloc(c, 8 * ns);
  WB_CHECK(hipMemcpyAsync(dp, params, pb, hipMemcpyHostToDevice, c->stream));
  QueryTimer qt(c);
  qt.start();
  uint32_t *order = ns >= WB_QUERY_SORT_MIN ? wb_query_order(t, kind, dp, ns) : nullptr;
  wb_run_count(t, kind, dp, ns, dc, order);
  qt.stop();
  WB_CHECK(hipMemcpyAsync(counts, dc, 8 * ns, hipMemcpyDeviceToHost, c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  t->ms_query_count = qt.ms();
  t->ms_query_points = 0;
  if (order) wb_free(c, order, 4 * ns);
  wb_free(c, dp, pb);
  wb_free(c, dc, 8 * ns);
  return WB_OK;
  API_CATCH(c)
}

int wb_count_points_in_noalloc(wb_tree *t, int kind, const double *params, uint64_t ns,
                               uint64_t *counts)
{
  if (!t || (!params && ns) || (!counts && ns)) return WB_EINVAL;
  wb_ctx *c = t->ctx;
  if (!check_kind(kind)) return fail(c, "unknown shape kind", WB_EINVAL);
  if (ns > WB_NOALLOC_MAX_SHAPES)
    return fail(c, "wb_count_points_in_noalloc takes at most WB_NOALLOC_MAX_SHAPES (" +
                       std::to_string(WB_NOALLOC_MAX_SHAPES) + ") shapes, got " +
                       std::to_string(ns),
                WB_ERANGE);
  if (ns == 0) return WB_OK;
  if (!c->scratch) return fail(c, "the static scratch has no device address on this GPU", WB_EDEVICE);
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  // every count is written by the kernel (one per shape): nothing in the
  // shared buffer has to start at zero
  double *dp = (double *)c->scratch;
  uint64_t *dc = (uint64_t *)((unsigned char *)c->scratch + WB_NOALLOC_PARAM_BYTES);
  WB_CHECK(hipMemcpyAsync(dp, params, sizeof(double) * WB_SHAPE_NPARAMS * ns,
                          hipMemcpyHostToDevice, c->stream));
  wb_run_count(t, kind, dp, ns, dc, nullptr);
  WB_CHECK(hipMemcpyAsync(counts, dc, 8 * ns, hipMemcpyDeviceToHost, c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  return WB_OK;
  API_CATCH(c)
}

int wb_points_in(wb_tree *t, int kind, const double *params, uint64_t ns, uint64_t *offsets,
                 uint32_t *idx, uint64_t cap)
{
  if (!t || (!params && ns) || !offsets) return WB_EINVAL;
  wb_ctx *c = t->ctx;
  if (!check_kind(kind)) return fail(c, "unknown shape kind", WB_EINVAL);
  if (ns > 0x7fffffffull) return fail(c, "too many shapes", WB_ERANGE);
  offsets[0] = 0;
  if (ns == 0) return WB_OK;
  API_TRY(c)
  size_t pb = sizeof(double) * WB_SHAPE_NPARAMS * ns;
  double *dp = (double *)wb_alloc(c, pb);
  uint64_t *dc = (uint64_t *)wb_alloc(c, 8 * (ns + 1));
  WB_CHECK(hipMemcpyAsync(dp, params, pb, hipMemcpyHostToDevice, c->stream));
  QueryTimer qt(c);
  qt.start();
  uint32_t *order = ns >= WB_QUERY_SORT_MIN ? wb_query_order(t, kind, dp, ns) : nullptr;
  wb_run_count(t, kind, dp, ns, dc, order);
  qt.stop();
  std::vector<uint64_t> cnt(ns);
  WB_CHECK(hipMemcpyAsync(cnt.data(), dc, 8 * ns, hipMemcpyDeviceToHost, c->stream));
  WB_CHECK(hipStreamSynchronize(c->stream));
  t->ms_query_count = qt.ms();
  t->ms_query_points = 0;
  for (uint64_t s = 0; s < ns; s++) offsets[s + 1] = offsets[s] + cnt[s];
  if (idx) {
    uint64_t total = offsets[ns];
    if (cap < total) {
      if (order) wb_free(c, order, 4 * ns);
      wb_free(c, dp, pb);
      wb_free(c, dc, 8 * (ns + 1));
      return fail(c, "index buffer too small", WB_EINVAL);
    }
    WB_CHECK(hipMemcpyAsync(dc, offsets, 8 * (ns + 1), hipMemcpyHostToDevice, c->stream));
    uint32_t *dout = (uint32_t *)wb_alloc(c, 4 * total + 4);
    qt.start();
    wb_run_points(t, kind, dp, ns, dc, dout, order);
    qt.stop();
    WB_CHECK(hipMemcpyAsync(idx, dout, 4 * total, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipStreamSynchronize(c->stream));
    t->ms_query_points = qt.ms();
    wb_free(c, dout, 4 * total + 4);
  }
  if (order) wb_free(c, order, 4 * ns);
  wb_free(c, dp, pb);
  wb_free(c, dc, 8 * (ns + 1));
  return WB_OK;
  API_CATCH(c)
}

int wb_tile_count(const double bbmin[3], const double bbmax[3], double tile_size, uint64_t *ntiles)
{
  if (!bbmin || !bbmax || !ntiles || !(tile_size > 0)) return WB_EINVAL;
  wb::SnakeSetup s = wb::snake_setup(bbmin, bbmax, tile_size);
  *ntiles = (uint64_t)((int64_t)s.hi - s.lo + 1);
  return WB_OK;
}

int wb_tile_count_corners(const double *corners, int ncorners, double tile_size, uint64_t *ntiles)
{
  if (!corners || ncorners <= 0 || !ntiles || !(tile_size > 0)) return WB_EINVAL;
  wb::SnakeSetup s = wb::snake_setup_corners(corners, ncorners, tile_size);
  *ntiles = (uint64_t)((int64_t)s.hi - s.lo + 1);
  return WB_OK;
}

// the checks every classification pass makes of its parameters and tiling;
// WB_OK or the code (the message on the tree's context)
static int check_pass(wb_tree *t, const double *corners, int ncorners,
                      const wb_classify_params *p, uint64_t *ntiles_out)
{
  wb_ctx *c = t->ctx;
  if (!(p->tile_size > 0) || !(p->max_slope > 0) || !(p->min_hs >= 0) || !isfinite(p->thickness))
    return fail(c, "bad classification parameters", WB_EINVAL);
  if (p->fit_precision < WB_FIT_FP64 || p->fit_precision > WB_FIT_FP32_COMP)
    return fail(c, "unknown fit precision", WB_EINVAL);
  // the early-exit argument needs every direction >= 30 units below +180 deg
  for (int a = 0; a < 2; a++) {
    double lo = INFINITY, hi = -INFINITY;
    for (int k = 0; k < ncorners; k++) {
      lo = fmin(lo, corners[3 * k + a]);
      hi = fmax(hi, corners[3 * k + a]);
    }
    if (!(hi - lo < 1e7 * t->scale[1 - a]))
      return fail(c, "cloud extent above 1e7 coordinate units is not supported", WB_ERANGE);
  }
  uint64_t ntiles;
  wb_tile_count_corners(corners, ncorners, p->tile_size, &ntiles);
  if (ntiles > 0xffffffffull) return fail(c, "too many tiles", WB_ERANGE);
  *ntiles_out = ntiles;
  return WB_OK;
}

static void fill_stats(wb_stats *stats, const wb_tree *t, const ClassifyOut &co, uint64_t ntiles,
                       double ms_total)
{
  memset(stats, 0, sizeof *stats);
  stats->n_unclassified = co.hist[0];
  stats->n_class1 = co.hist[1];
  stats->n_class2 = co.hist[2];
  stats->ntiles = ntiles;
  stats->nonempty_tiles = co.nonempty;
  stats->sector_borderline = co.borderline;
  stats->exact_gap_points = co.nslow;
  stats->atan_exact_calls = co.atan_exact;
  stats->heavy_points = co.nheavy;
  stats->sort_passes = t->sort_passes;
  stats->maxdepth = t->maxdepth;
  stats->nleaves = t->nleaves;
  stats->nnodes = t->nnodes;
  wb_tree_read_timing(const_cast<wb_tree *>(t));
  stats->ms_build = t->ms_build;
  stats->ms_sort = t->ms_sort;
  stats->ms_classify_kernel = co.ms_classify_kernel;
  stats->ms_heavy_kernel = co.ms_heavy_kernel;
  stats->ms_tiles = co.ms_tiles;
  stats->ms_scan = co.ms_scan;
  stats->ms_classify = co.ms_classify;
  stats->ms_total = ms_total;
  stats->fit_fallbacks = co.fallbacks;
}

// wb_classify and wb_scan_tiles: the pass up to POSTSCAN, and CLASSIFY unless
// scan_only
static int run_pass(wb_tree *t, const double *corners, int ncorners,
                    const wb_classify_params *p, uint8_t *class_out, int flags,
                    wb_tile *tiles_out, uint64_t tiles_cap, wb_stats *stats, bool scan_only)
{
  wb_ctx *c = t->ctx;
  if (!scan_only && !class_out) return fail(c, "no class output", WB_EINVAL);
  uint64_t ntiles = 0;
  const int chk = check_pass(t, corners, ncorners, p, &ntiles);
  if (chk != WB_OK) return chk;
  if (tiles_out && tiles_cap < ntiles) return fail(c, "tile buffer too small", WB_EINVAL);
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  auto t0 = std::chrono::steady_clock::now();
  uint8_t *dcls = class_out;
  bool host_out = !(flags & WB_DEVICE_INPUT) && !scan_only;
  if (host_out) dcls = (uint8_t *)wb_alloc(c, t->n_input + 16);
  ClassifyOut co;
  wb_run_classify(t, corners, ncorners, p, dcls, tiles_out != nullptr, scan_only, &co);
  if (host_out) {
    if (t->n_input)
      WB_CHECK(hipMemcpyAsync(class_out, dcls, t->n_input, hipMemcpyDeviceToHost, c->stream));
    WB_CHECK(hipStreamSynchronize(c->stream));
    wb_free(c, dcls, t->n_input + 16);
  }
  auto t1 = std::chrono::steady_clock::now();
  if (tiles_out) {
    for (uint64_t k = 0; k < ntiles; k++) {
      const TileRec &r = co.tiles[k];
      wb_tile &o = tiles_out[k];
      o.npoints = r.npoints;
      o.nground = r.nground;
      o.tree_flags = r.tree_flags;
      o.flowsnake_x = co.tex[k];
      o.flowsnake_y = co.tey[k];
      o.fit_fp64 = r.fit_fp64;
      o.density = r.density;
      o.hs_scan = r.hs_scan;
      o.hs = r.hs;
      o.height = r.height;
      o.tilt_x = r.tilt_x;
      o.tilt_y = r.tilt_y;
    }
  }
  if (stats) fill_stats(stats, t, co, ntiles, std::chrono::duration<double, std::milli>(t1 - t0).count());
  return WB_OK;
  API_CATCH(c)
}

int wb_classify(wb_tree *t, const double bbmin[3], const double bbmax[3],
                const wb_classify_params *p, uint8_t *class_out, int flags, wb_tile *tiles_out,
                uint64_t tiles_cap, wb_stats *stats)
{
  if (!t || !bbmin || !bbmax || !p || !class_out) return WB_EINVAL;
  const double corners[6] = {bbmin[0], bbmin[1], bbmin[2], bbmax[0], bbmax[1], bbmax[2]};
  return run_pass(t, corners, 2, p, class_out, flags, tiles_out, tiles_cap, stats, false);
}

int wb_classify_corners(wb_tree *t, const double *corners, int ncorners,
                        const wb_classify_params *p, uint8_t *class_out, int flags,
                        wb_tile *tiles_out, uint64_t tiles_cap, wb_stats *stats)
{
  if (!t || !corners || ncorners <= 0 || !p || !class_out) return WB_EINVAL;
  return run_pass(t, corners, ncorners, p, class_out, flags, tiles_out, tiles_cap, stats, false);
}

int wb_classify_batch(wb_tree *const *trees, int ntrees, const double *bbmin,
                      const double *bbmax, const wb_classify_params *p, uint8_t *const *class_out,
                      int flags, wb_stats *stats)
{
  if (!trees || ntrees <= 0 || !bbmin || !bbmax || !p || !class_out) return WB_EINVAL;
  if (!trees[0]) return WB_EINVAL;
  wb_ctx *c = trees[0]->ctx;
  std::vector<double> corners(6 * (size_t)ntrees);
  std::vector<uint64_t> ntiles(ntrees);
  for (int k = 0; k < ntrees; k++) {
    if (!trees[k] || !class_out[k]) return fail(c, "null tree or class output", WB_EINVAL);
    if (trees[k]->ctx != c) return fail(c, "trees of different contexts", WB_EINVAL);
    for (int a = 0; a < 3; a++) {
      corners[6 * k + a] = bbmin[3 * k + a];
      corners[6 * k + 3 + a] = bbmax[3 * k + a];
    }
    const int chk = check_pass(trees[k], &corners[6 * k], 2, p, &ntiles[k]);
    if (chk != WB_OK) return chk;
  }
  API_TRY(c)
  WB_CHECK(hipSetDevice(c->device));
  auto t0 = std::chrono::steady_clock::now();
  const bool host_out = !(flags & WB_DEVICE_INPUT);
  std::vector<uint8_t *> dcls(ntrees);
  for (int k = 0; k < ntrees; k++)
    dcls[k] = host_out ? (uint8_t *)wb_alloc(c, trees[k]->n_input + 16) : class_out[k];
  std::vector<ClassifyOut> co(ntrees);
  wb_run_classify_batch(c, trees, ntrees, corners.data(), p, dcls.data(), co.data());
  if (host_out) {
    for (int k = 0; k < ntrees; k++)
      if (trees[k]->n_input)
        WB_CHECK(hipMemcpyAsync(class_out[k], dcls[k], trees[k]->n_input, hipMemcpyDeviceToHost,
                                c->stream));
    WB_CHECK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < ntrees; k++) wb_free(c, dcls[k], trees[k]->n_input + 16);
  }
  auto t1 = std::chrono::steady_clock::now();
  if (stats)
    for (int k = 0; k < ntrees; k++)
      fill_stats(&stats[k], trees[k], co[k], ntiles[k],
                 std::chrono::duration<double, std::milli>(t1 - t0).count());
  return WB_OK;
  API_CATCH(c)
}

int wb_scan_tiles(wb_tree *t, const double bbmin[3], const double bbmax[3],
                  const wb_classify_params *p, wb_tile *tiles_out, uint64_t tiles_cap,
                  wb_stats *stats)
{
  if (!t || !bbmin || !bbmax || !p || !tiles_out) return WB_EINVAL;
  const double corners[6] = {bbmin[0], bbmin[1], bbmin[2], bbmax[0], bbmax[1], bbmax[2]};
  return run_pass(t, corners, 2, p, nullptr, 0, tiles_out, tiles_cap, stats, true);
}

}  // extern "C"

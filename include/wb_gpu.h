/*
 * wb_gpu.h -- C ABI of the MI355X octree build and ground classification.
 *
 * Plain C types only.  Every function returns 0 on success or a negative
 * WB_E* code; wb_last_error(ctx) gives the message of the last failure on
 * that context.  Host buffers belong to the caller; device memory belongs to
 * the handles.  One wb_ctx per host thread (a context is not locked).
 *
 * Coordinates are LAS int32 values with a scale/offset transform; they are
 * decoded exactly as LasHeader::readPoint does (las.cpp:808):
 *   x = (offset_x + scale_x * x_int) * unit
 * and every geometric test runs in FP64 with the reference's operation
 * order, so results equal the reference's single-threaded CPU path bit for
 * bit (see the parity notes in DESIGN.md).
 */
#ifndef WB_GPU_H
#define WB_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  WB_OK = 0,
  WB_EINVAL = -1,   /* bad argument */
  WB_EDEVICE = -2,  /* HIP runtime error */
  WB_ENOMEM = -3,   /* device allocation failed */
  WB_ERANGE = -4,   /* input outside what the kernels support */
};

/* shape kinds of wb_count_points_in / wb_points_in */
enum {
  WB_SHAPE_CYLINDER = 0,    /* params: cx, cy, r                 (Cylinder, shape.h:136-157) */
  WB_SHAPE_SPHERE = 1,      /* params: cx, cy, cz, r             (Sphere, shape.h:120-134) */
  WB_SHAPE_HYPERBOLOID = 2, /* params: vx, vy, vz, r, slope      (Hyperboloid, shape.h:96-118) */
};
#define WB_SHAPE_NPARAMS 5  /* each shape takes 5 doubles; unused ones are ignored */

/* flags of wb_tree_build */
enum {
  WB_DEVICE_INPUT = 1,      /* x, y, z are device pointers already resident in HBM */
};

/* precision of the scan's plane fit (wb_classify_params.fit_precision) */
enum {
  WB_FIT_FP64 = 0,          /* the reference's FP64 fit, bit-identical (default) */
  WB_FIT_FP32 = 1,          /* FP32 moments, naive sums, FP32 solve (study only) */
  WB_FIT_FP32_COMP = 2,     /* FP32 moments with compensated sums, FP32 solve; tiles
                               whose xy spread is near-collinear fall back to FP64 */
};

typedef struct wb_ctx wb_ctx;
typedef struct wb_tree wb_tree;

typedef struct {
  double scale[3];
  double offset[3];
  double unit;              /* length unit multiplier (1 for metres) */
} wb_lasxform;

typedef struct {
  double center[3];         /* leaf cube centre (Octree::cube, octree.cpp:349-358) */
  double side;              /* leaf cube side */
  uint64_t first;           /* first canonical position of its points */
  uint32_t count;           /* number of points (<= 537) */
  int32_t depth;            /* level below the root (root children are 1) */
} wb_leaf;

typedef struct {
  double tile_size;         /* desired tile spacing, m (default 1) */
  double max_slope;         /* hyperboloid asymptotic slope (default 1) */
  double thickness;         /* vertex drop below the point (default 0) */
  double min_hs;            /* minimum hyperboloid size, m (default 0.1) */
  int32_t fit_precision;    /* WB_FIT_*; 0 (FP64) reproduces the reference */
} wb_classify_params;

typedef struct {
  int32_t npoints;          /* points in the tile cylinder (Tile::nPoints) */
  int32_t nground;          /* points it classified as ground */
  int32_t tree_flags;       /* Tile::treeFlags */
  int32_t flowsnake_x, flowsnake_y;  /* Eisenstein tile address */
  int32_t fit_fp64;         /* 1 if an FP32 fit fell back to FP64 for this tile */
  double density;           /* bottom-layer density (Tile::density) */
  double hs_scan;           /* hyperboloid size after scan */
  double hs;                /* hyperboloid size after postscan (used to classify) */
  double height;            /* Tile::height */
  double tilt_x, tilt_y;    /* the fitted plane's slope after scan's clamp (scan.cpp:68-73) */
} wb_tile;

typedef struct {
  uint64_t n_class1, n_class2, n_unclassified;
  uint64_t ntiles, nonempty_tiles;
  uint64_t sector_borderline; /* scan sector angles within 1e-12 of a boundary */
  uint64_t exact_gap_points;  /* points decided by the exact (slow) gap pass */
  uint64_t atan_exact_calls;  /* atan2i last steps emulated in 64-bit significands */
  uint64_t heavy_points;      /* points finished by the wave-per-point kernel */
  uint32_t sort_passes, maxdepth;
  uint64_t nleaves, nnodes;
  double ms_build;            /* last wb_tree_build of this tree (device time) */
  double ms_sort;             /* of which the radix sort */
  double ms_tiles;            /* tile geometry, membership, tile lists */
  double ms_scan;             /* scan + postscan */
  double ms_classify;         /* ground test, exact pass, class scatter */
  double ms_classify_kernel;  /* the ground-test kernel alone */
  double ms_heavy_kernel;     /* the wave-per-point kernel for large candidate sets */
  double ms_total;            /* host wall time of wb_classify */
  uint64_t fit_fallbacks;     /* tiles whose FP32 fit fell back to FP64 */
} wb_stats;

/* --- context ------------------------------------------------------------ */
int wb_ctx_create(int device, wb_ctx **out);
/* also frees every tree built on the context that is still alive: their
 * handles are invalid afterwards and must not be passed to wb_tree_free */
void wb_ctx_destroy(wb_ctx *ctx);
const char *wb_last_error(wb_ctx *ctx);
/* synchronise the context's stream */
int wb_ctx_sync(wb_ctx *ctx);
/* free and total device memory of a device (admission control of a serving
 * loop; no counterpart in the reference, whose store lives on disk) */
int wb_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* --- root cube: Octree::sizeFit (octree.cpp:268-310) -------------------- */
int wb_size_fit(const double *corners, int ncorners, double center[3], double *side);

/* --- octree build: replaces the READ phase of wolkencli / the GUI
 *     (wolkencli.cpp:99-125, threads.cpp:477-589, OctStore::put
 *     octree.cpp:849-876, Octree::split octree.cpp:312-347).
 *     n <= 2^32-1.  The tree keeps its own decoded copy of the points. */
int wb_tree_build(wb_ctx *ctx, const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n,
                  const wb_lasxform *xf, const double root_center[3], double root_side,
                  int flags, wb_tree **out);
int wb_tree_free(wb_tree *tree);
int wb_tree_info(wb_tree *tree, uint64_t *nleaves, uint64_t *nnodes, int *maxdepth,
                 int *sort_passes);
/* leaves in DFS (Octree::dump) order: replaces Octree::dump / OctStore::dump
 * (octree.cpp:360-376, 888-891) as a structured record */
int wb_tree_leaves(wb_tree *tree, wb_leaf *out, uint64_t cap, uint64_t *nleaves);
/* input indices in canonical order (leaf DFS order, input order in a leaf) */
int wb_tree_order(wb_tree *tree, uint32_t *order);

/* --- batched queries: OctStore::countPointsIn / pointsIn
 *     (octree.cpp:1214-1264) for nshapes shapes of one kind.
 *     params: nshapes * WB_SHAPE_NPARAMS doubles.
 *     wb_points_in writes, per shape, the input indices in the reference's
 *     unsorted order (findBlocks DFS order, then in-leaf order);
 *     offsets has nshapes+1 entries; idx needs offsets[nshapes] entries
 *     (call with idx == NULL to get the offsets only). */
int wb_count_points_in(wb_tree *tree, int kind, const double *params, uint64_t nshapes,
                       uint64_t *counts);
int wb_points_in(wb_tree *tree, int kind, const double *params, uint64_t nshapes,
                 uint64_t *offsets, uint32_t *idx, uint64_t cap);

/* --- the classification pass: SCAN (scanCylinder, scan.cpp:31-140),
 *     POSTSCAN (postscanCylinder, scan.cpp:142-179) and CLASSIFY
 *     (classifyCylinder, classify.cpp:96-173) over every flowsnake tile, with
 *     the single-threaded flowsnake order deciding overlaps (last tile wins).
 *     bbmin/bbmax: the LAS header corners (they set the tiling as
 *     wolkencanvas.cpp:501-519 does).  class_out: n bytes in input order
 *     (0 never classified, 1 not ground, 2 ground).  tiles_out may be NULL,
 *     otherwise it receives 7^k records in flowsnake counter order.
 *     With WB_DEVICE_INPUT the tree was built from device input and class_out
 *     may be a device pointer too (flags of this call). */
int wb_classify(wb_tree *tree, const double bbmin[3], const double bbmax[3],
                const wb_classify_params *params, uint8_t *class_out, int flags,
                wb_tile *tiles_out, uint64_t tiles_cap, wb_stats *stats);

/* --- SCAN and POSTSCAN alone (scanCylinder and postscanCylinder over every
 *     tile, scan.cpp:31-179, in the phase order of wolkencanvas.cpp:469-535):
 *     the tile records of wb_classify without the classification.  tiles_out
 *     receives 7^k records in flowsnake counter order. */
int wb_scan_tiles(wb_tree *tree, const double bbmin[3], const double bbmax[3],
                  const wb_classify_params *params, wb_tile *tiles_out, uint64_t tiles_cap,
                  wb_stats *stats);

/* number of flowsnake tiles the classification uses for these corners */
int wb_tile_count(const double bbmin[3], const double bbmax[3], double tile_size,
                  uint64_t *ntiles);

#ifdef __cplusplus
}
#endif
#endif

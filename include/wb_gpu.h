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
 * order.  Results are bit-identical to the oracle, a single-threaded CPU
 * restatement of the reference's path with its scheduling-dependent orders
 * pinned (points of a leaf in input order, the last flowsnake tile deciding
 * an overlap).
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

/* flags of wb_tree_build / wb_tree_build_xyz */
enum {
  WB_DEVICE_INPUT = 1,      /* x, y, z are device pointers already resident in HBM */
  WB_BUILD_DEDUPE = 2,      /* merge points at one location as OctBuffer::put does
                               (octree.cpp:620-660): the location keeps the place of its
                               first point and the record of its last; the tree then
                               holds each location once */
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
/* also frees every tree, LAS file and cloud made on the context that is
 * still alive: their handles are invalid afterwards and must not be passed
 * to wb_tree_free, wb_las_close or wb_cloud_free */
void wb_ctx_destroy(wb_ctx *ctx);
const char *wb_last_error(wb_ctx *ctx);
/* synchronise the context's stream */
int wb_ctx_sync(wb_ctx *ctx);
/* on != 0: the context's long-running kernels (the ground test and its
 * wave-per-point pass) run on resident grids that leave room on every CU,
 * so kernels of other contexts that run meanwhile are not held until they
 * end (a serving loop with one context per worker); off (the default): the
 * full device.  Results are the same either way.  No counterpart in the
 * reference. */
int wb_ctx_set_sharing(wb_ctx *ctx, int on);
/* free and total device memory of a device (admission control of a serving
 * loop; no counterpart in the reference, whose store lives on disk) */
int wb_device_memory(int device, uint64_t *free_bytes, uint64_t *total_bytes);
/* pinned host memory (hipHostMalloc) and its release: receive buffers of a
 * serving loop, which copies to and from the device read and write by DMA
 * without staging (no counterpart in the reference).  bytes > 0. */
int wb_host_alloc(uint64_t bytes, void **out);
void wb_host_free(void *p);

/* --- root cube: Octree::sizeFit (octree.cpp:268-310) -------------------- */
int wb_size_fit(const double *corners, int ncorners, double center[3], double *side);

/* --- octree build: replaces the READ phase of wolkencli / the GUI
 *     (wolkencli.cpp:99-125, threads.cpp:477-589, OctStore::put
 *     octree.cpp:849-876, Octree::split octree.cpp:312-347).
 *     n <= 2^32-1.  The tree keeps its own decoded copy of the points. */
int wb_tree_build(wb_ctx *ctx, const int32_t *x, const int32_t *y, const int32_t *z, uint64_t n,
                  const wb_lasxform *xf, const double root_center[3], double root_side,
                  int flags, wb_tree **out);
/* --- octree build from coordinates already decoded to FP64 (several LAS
 *     files with different transforms, a length unit other than 1: the
 *     location LasHeader::readPoint computes, las.cpp:896).  resolution: a
 *     spacing every nonzero coordinate difference is at least (per axis;
 *     combineScales of the inputs, las.cpp:913-944, times the unit); the
 *     ground test checks the cloud extent against it. */
int wb_tree_build_xyz(wb_ctx *ctx, const double *x, const double *y, const double *z, uint64_t n,
                      const double resolution[3], const double root_center[3], double root_side,
                      int flags, wb_tree **out);
/* inputs merged by WB_BUILD_DEDUPE (the count wolkencli prints as duplicate
 * points, wolkencli.cpp:114-115); pairs (may be NULL) receives ndup
 * (merged input index, input index of the first point at its location) */
int wb_tree_duplicates(wb_tree *tree, uint64_t *ndup, uint32_t *pairs);
int wb_tree_free(wb_tree *tree);
int wb_tree_info(wb_tree *tree, uint64_t *nleaves, uint64_t *nnodes, int *maxdepth,
                 int *sort_passes);
/* leaves in DFS (Octree::dump) order: replaces Octree::dump / OctStore::dump
 * (octree.cpp:360-376, 888-891) as a structured record */
int wb_tree_leaves(wb_tree *tree, wb_leaf *out, uint64_t cap, uint64_t *nleaves);
/* input indices in canonical order (leaf DFS order, input order in a leaf);
 * order has one entry per location (n minus the merged duplicates) */
int wb_tree_order(wb_tree *tree, uint32_t *order);

/* --- the octree as text: OctStore::dump (octree.cpp:360-376, 673-689,
 *     888-891), the "dumpfile" of wolkencli (wolkencli.cpp:122-124): per
 *     leaf in DFS order "(cx,cy,cz)±half N points[, K stray points]", then
 *     "T total points", numbers by ldecimal.  Written to path (may be NULL)
 *     and/or buf (cap bytes, NUL-terminated); len receives the length. */
int wb_tree_dump(wb_tree *tree, const char *path, char *buf, uint64_t cap, uint64_t *len);
/* ldecimal (ldecimal.cpp:32-132): returns the length, or a negative code */
int wb_ldecimal(double x, double toler, char *buf, int cap);

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
/* device time (ms, HIP events on the context stream) of the kernels of the
 * tree's last wb_count_points_in / wb_points_in call: the counting kernel,
 * and wb_points_in's index kernel (0 when the call wrote no indices).  No
 * counterpart in the reference (its queries are not timed); for
 * measurement. */
int wb_tree_query_ms(wb_tree *tree, double *count_ms, double *points_ms);

/* --- wb_count_points_in with no device allocation.  No counterpart in the
 *     reference: for callers that cannot allocate device memory (code under
 *     HIP graph capture, tight serving loops, harnesses that forbid
 *     hipMalloc).  The shape parameters and counts are staged in a fixed
 *     buffer the library holds on each GPU from the first wb_ctx_create on
 *     that GPU until the process ends, sized for WB_NOALLOC_MAX_SHAPES shapes
 *     (3 MiB); more shapes are refused with WB_ERANGE.  Every call on a GPU
 *     uses that GPU's one buffer: calls that overlap on one GPU (contexts of
 *     several host threads) race, and the caller must serialise them.
 *     Counts equal wb_count_points_in's. */
#define WB_NOALLOC_MAX_SHAPES 65536
int wb_count_points_in_noalloc(wb_tree *tree, int kind, const double *params, uint64_t nshapes,
                               uint64_t *counts);

/* --- the classification pass: SCAN (scanCylinder, scan.cpp:31-140),
 *     POSTSCAN (postscanCylinder, scan.cpp:142-179) and CLASSIFY
 *     (classifyCylinder, classify.cpp:96-173) over every flowsnake tile, with
 *     the single-threaded flowsnake order deciding overlaps (last tile wins).
 *     bbmin/bbmax: the LAS header corners (they set the tiling as
 *     wolkencanvas.cpp:501-519 does).  class_out: n bytes in input order
 *     (0 never classified, 1 not ground, 2 ground); a merged duplicate gets the class of
 *     its location.  tiles_out may be NULL,
 *     otherwise it receives 7^k records in flowsnake counter order.
 *     With WB_DEVICE_INPUT the tree was built from device input and class_out
 *     may be a device pointer too (flags of this call). */
int wb_classify(wb_tree *tree, const double bbmin[3], const double bbmax[3],
                const wb_classify_params *params, uint8_t *class_out, int flags,
                wb_tile *tiles_out, uint64_t tiles_cap, wb_stats *stats);

/* --- several trees classified together (a serving loop's small tiles):
 *     wb_classify on each tree with its own header corners (bbmin and bbmax:
 *     3 doubles per tree), the trees' ground tests in one launch of the
 *     ground kernel and one of the wave-per-point kernel, so small tiles
 *     share a full device.  No counterpart in the reference, which
 *     classifies one cloud at a time (threads.cpp:619-660).  Class bytes
 *     and statistics equal those of wb_classify on each tree alone.  Every
 *     tree belongs to the same context; class_out[k]: n_k bytes (host, or
 *     device with WB_DEVICE_INPUT); stats (may be NULL): ntrees records. */
int wb_classify_batch(wb_tree *const *trees, int ntrees, const double *bbmin,
                      const double *bbmax, const wb_classify_params *params,
                      uint8_t *const *class_out, int flags, wb_stats *stats);

/* the same pass with the tiling taken from the corners of several LAS
 * headers (x, y, z each): the BoundRect of them all, as sizeToFit builds it
 * (wolkencanvas.cpp:136-146, 501-519) */
int wb_classify_corners(wb_tree *tree, const double *corners, int ncorners,
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
int wb_tile_count_corners(const double *corners, int ncorners, double tile_size,
                          uint64_t *ntiles);

/* --- LAS files (las.h:97-183) -------------------------------------------
 * A wb_las is one LAS file opened for reading: its header, parsed as
 * LasHeader::openRead does (las.cpp:272-428, including the repair of the
 * point counts), and its point records copied to HBM as they lie in the
 * file.  LASzip files are refused (no LASzip here). */
typedef struct wb_las wb_las;
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
  uint64_t n_points[16];    /* [0] total, [r] return r, after the repair */
  int32_t zipped;           /* LASzip file */
  int32_t valid;            /* LasHeader::isValid (las.cpp:597-601) */
} wb_las_header;

/* the header alone (host only) */
int wb_las_read_header(const char *path, wb_las_header *hdr);
int wb_las_open(wb_ctx *ctx, const char *path, wb_las **out, wb_las_header *hdr);
/* the same from a whole file image in host memory */
int wb_las_open_mem(wb_ctx *ctx, const void *image, uint64_t len, wb_las **out,
                    wb_las_header *hdr);
int wb_las_close(wb_las *las);

/* --- a cloud: the points of one or more LAS files in HBM, decoded to FP64
 *     as LasHeader::readPoint does (las.cpp:808-904), in file order.
 *     WB_ZEROS_READ applies the READ phase's rule for return number 0
 *     (threads.cpp:479-545: when the first record has a return number,
 *     records without one are dropped, otherwise they are read as first
 *     returns); WB_ZEROS_KEEP keeps every record as read (wolkencli.cpp:
 *     99-111).  The files must stay open while the cloud lives. */
typedef struct wb_cloud wb_cloud;
enum { WB_ZEROS_READ = 0, WB_ZEROS_KEEP = 1 };
int wb_cloud_create(wb_ctx *ctx, wb_las *const *files, int nfiles, double unit, int zeros,
                    wb_cloud **out);
/* points, the coordinate lattice spacing (combineScales of the headers
 * times the unit, las.cpp:934-944), joinPointFormat of the inputs
 * (las.cpp:104-121), and the 2*nfiles header corners times the unit (each
 * may be NULL) */
int wb_cloud_info(wb_cloud *cloud, uint64_t *npoints, double resolution[3], int *join_format,
                  double *corners);
/* device time (ms) of the cloud's record decode (keep flags, scan, decode) */
int wb_cloud_decode_ms(wb_cloud *cloud, double *ms);
/* device pointers of the decoded coordinates */
int wb_cloud_points(wb_cloud *cloud, const double **x, const double **y, const double **z);
/* host copies of the decoded coordinates and of each point's record index
 * in its file (any may be NULL) */
int wb_cloud_read_points(wb_cloud *cloud, double *x, double *y, double *z, uint32_t *rec);
/* the octree of the cloud, duplicates merged as OctBuffer::put does */
int wb_cloud_build(wb_cloud *cloud, const double root_center[3], double root_side, int flags,
                   wb_tree **out);
/* wb_classify_corners with the tiling of the cloud's header corners */
int wb_cloud_classify(wb_cloud *cloud, wb_tree *tree, const wb_classify_params *params,
                      uint8_t *class_out, int flags, wb_tile *tiles_out, uint64_t tiles_cap,
                      wb_stats *stats);
int wb_cloud_free(wb_cloud *cloud);

/* --- LAS output: CloudOutput::openFiles / writeFiles / closeFiles
 *     (cloudoutput.cpp:119-246) with the octree's leaves in DFS order as its
 *     blocks, and LasHeader::writePoint / writeHeader (las.cpp:549-595,
 *     906-990).  Points whose class byte is 0 keep the class of their record;
 *     a merged duplicate is written once, with the record of the last input
 *     at its location.  Files are named
 *       name-<class>-<k>.las (separate_classes) or name-<k>.las,
 *     <k> only when points_per_file is nonzero. */
enum { WB_WRITE_DRY = 4 };  /* flag: plan and encode the files in HBM, copy and write
                               nothing (names and statistics are still returned) */
typedef struct {
  const char *name;
  int32_t separate_classes;
  int32_t points_per_file;  /* 0: no limit; the GUI uses 10000000 */
  int32_t point_format;     /* -1: joinPointFormat of the inputs */
  int32_t creation_day, creation_year;  /* 0: today (UTC) */
} wb_las_write_params;
typedef struct {
  uint64_t class_totals[256]; /* stored points per class (OctStore::countClasses) */
  uint64_t npoints;           /* points written */
  uint64_t bytes;             /* bytes written, headers included */
  uint32_t nfiles;
  int32_t point_format;
  double scale[3], offset[3]; /* of the output files (LasHeader::setScale) */
  double ms_classes;          /* device: class of each stored point, totals */
  double ms_plan;             /* host: per-leaf counts, file plan, uploads */
  double ms_encode;           /* device: the record-encode kernel */
  double ms_copy;             /* device to host copies */
  double ms_disk;             /* file writes */
} wb_las_write_stats;
/* cls: class bytes of wb_classify in cloud order (NULL: every point keeps
 * its record's class; a device pointer with WB_DEVICE_INPUT); names (may be
 * NULL) receives the file names, one per line */
int wb_cloud_write(wb_cloud *cloud, wb_tree *tree, const uint8_t *cls, int flags,
                   const wb_las_write_params *params, wb_las_write_stats *stats, char *names,
                   uint64_t names_cap);

/* --- the store's block format (octree.h:35-43, octree.cpp:427-559): the
 *     spill tier of the HBM-resident tree.  Block l holds leaf l (DFS order):
 *     its stored points in canonical order as LasPoint::write records
 *     (las.cpp:177-207, 61 bytes), then noPoint records (NaN location, zero
 *     fields) up to 537, then zero bytes up to 32768 (OctBuffer::write).
 *     With a cloud (the tree built from it by wb_cloud_build) a record holds
 *     the fields of the location's last input as readPoint decodes them
 *     (las.cpp:808-904) with the READ phase's return fix; without one,
 *     LasPoint's defaults (las.cpp:122-132) at the point's location.  cls
 *     (may be NULL): class bytes in input order as wb_classify returns them
 *     (a device pointer with WB_DEVICE_INPUT); a nonzero byte replaces the
 *     classification, as CLASSIFY's put does (classify.cpp:158-163). */
#define WB_STORE_RECORD 61
#define WB_STORE_RECORDS 537
#define WB_STORE_BLOCK 32768
/* the blocks of every leaf into out (host memory, pinned or not, of cap
 * bytes; NULL: only *nblocks, the number of leaves, is returned) */
int wb_store_blocks(wb_tree *tree, wb_cloud *cloud, const uint8_t *cls, int flags, void *out,
                    uint64_t cap, uint64_t *nblocks);
/* the blocks as the store's files (OctStore::open and OctBuffer::write,
 * octree.cpp:812-830, 427-460): block b in file name<b % nfiles> at offset
 * 32768 * (b / nfiles); nfiles below 1 is 1 */
int wb_store_write_files(wb_tree *tree, wb_cloud *cloud, const uint8_t *cls, int flags,
                         const char *name, int nfiles);
/* OctBuffer::read (octree.cpp:508-559) of nblocks blocks in host memory:
 * the records with a location (LasPoint::isEmpty, point.cpp:252-255),
 * block by block in record order, into x, y, z and classification (each may
 * be NULL; room for the total); block_count[b] (may be NULL) the points of
 * block b, 0 for a block whose first two locations are equal (a block never
 * written reads as zeros); *npoints the total */
int wb_store_read_blocks(wb_ctx *ctx, const void *blocks, uint64_t nblocks, double *x, double *y,
                         double *z, uint16_t *classification, uint32_t *block_count,
                         uint64_t *npoints);

#ifdef __cplusplus
}
#endif
#endif

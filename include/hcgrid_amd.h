/*
 * hcgrid_amd.h -- C ABI of the MI355X (gfx950) gridding library.
 *
 * This is the only layer that talks to HIP.  Everything is plain C: integer
 * sizes, plain pointers, opaque stream handles (a hipStream_t passed as void*,
 * NULL = the default stream).  Functions return HCG_OK (0) or a negative
 * error code; none of them calls exit().  hcg_strerror() names a code and
 * hcg_last_error_message() carries the detail of the calling thread's last
 * failure.
 *
 * Unless a parameter says "host", pointers are device pointers and the call
 * only enqueues work on `stream` (no synchronisation).  The caller owns every
 * buffer; functions that need scratch take a workspace whose size a *_size()
 * query gives.
 *
 * Each entry point names the reference code it replaces (paths relative to
 * the HCGrid repository).
 */
#ifndef HCGRID_AMD_H
#define HCGRID_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define HCG_OK                0
#define HCG_ERR_INVALID      -1   /* bad argument (NULL, size, unsupported kernel type) */
#define HCG_ERR_HIP          -2   /* a HIP runtime call failed */
#define HCG_ERR_NOMEM        -3   /* allocation failed */
#define HCG_ERR_UNSUPPORTED  -4   /* nside or size beyond what the device path supports */
#define HCG_ERR_WORKSPACE    -5   /* workspace smaller than the *_size() query */
#define HCG_ERR_IO           -6   /* file I/O or FITS format error */

/* ---- constants (HCGrid/constants.h:57-73) ---- */
#define HCG_MAX_Y          1073741824u  /* key of a sample whose theta is outside [0, pi] */
#define HCG_GAUSS1D        1u
#define HCG_GAUSS2D        2u
#define HCG_TAPERED_SINC   3u
#define HCG_NESTED         1u
#define HCG_RING           2u
#define HCG_DTYPE_F64      0
#define HCG_DTYPE_F32      1
#define HCG_DTYPE_BF16     2
#define HCG_DTYPE_F16      3
#define HCG_MAX_NSIDE      8192u  /* keys are uint32; npix(8192) < MAX_Y, as the reference needs */
/* hcg_kernel.selection: which (cell, sample) pairs are gridded */
#define HCG_SELECT_COMPAT  0u  /* the reference's windows (HCGrid/gridding.cu:306-386), quirks included */
#define HCG_SELECT_DISC    1u  /* every sample within sphere_radius of the cell, nothing else */

/* HEALPix geometry: the fields of `struct Healpix` (HCGrid/healpix.h:14-21)
 * that the path reads.  Filled by hcg_healpix_init(). */
typedef struct hcg_healpix {
    uint64_t nside, order, nrings, npface, ncap, npix;
    double fact1, fact2, omega, resolution;
} hcg_healpix;

/* Gridding-kernel state left by _prepare_grid_kernel (HCGrid/gmap.h:47-61,
 * HCGrid/gmap.cpp:64-99): kernel type, transformed parameters, radius and the
 * disc size that bounds the ring and longitude windows. */
typedef struct hcg_kernel {
    uint32_t kernel_type;     /* HCG_GAUSS1D / HCG_GAUSS2D / HCG_TAPERED_SINC */
    uint32_t bearing_needed;  /* 1 for GAUSS2D */
    uint32_t factor;          /* coarsening factor gamma: cells per window anchor (>=1) */
    uint32_t selection;       /* HCG_SELECT_COMPAT (default, 0) or HCG_SELECT_DISC; no counterpart
                                 in the reference, whose selection is the compat one.  DISC takes
                                 every sample with distance < sphere_radius: no first-ring skip, no
                                 east-west truncation, windows across longitude 0 select both
                                 sides; factor is then ignored (each cell is its own anchor) */
    double params[3];         /* GAUSS1D: params[0] = 0.5 / sigma^2 (sigma in degrees) */
    double sphere_radius;     /* degrees */
    double disc_size;         /* radians: DEG2RAD * sphere_radius + resolution */
} hcg_kernel;

/* Sorted-sample record the gridding kernels read: unit vector of the sample
 * direction and its value, 32 bytes, one per sorted sample. */
typedef struct hcg_sample {
    double x, y, z, v;
} hcg_sample;

/* Device-side summary written by hcg_ring_lut(). */
typedef struct hcg_ring_meta {
    uint32_t firstring;  /* ring of the first sorted sample */
    uint32_t usedrings;  /* rings spanned by the valid samples (0 if none) */
    uint32_t n_valid;    /* samples with a key < npix (invalid ones sort last) */
    uint32_t n;          /* all samples */
} hcg_ring_meta;

/* ---- errors and device ---- */
const char *hcg_strerror(int code);
const char *hcg_last_error_message(void);
int hcg_device_count(int *count);                        /* host out */
int hcg_set_device(int device);

/* Memory plumbing for callers without their own allocator (the C++ host layer). */
int hcg_malloc(void **dptr, size_t bytes);
int hcg_free(void *dptr);
int hcg_memcpy_h2d(void *dst, const void *host_src, size_t bytes, void *stream);
int hcg_memcpy_d2h(void *host_dst, const void *src, size_t bytes, void *stream);
int hcg_memset(void *dst, int value, size_t bytes, void *stream);
int hcg_stream_create(void **stream);                    /* default priority */
int hcg_stream_destroy(void *stream);
int hcg_stream_synchronize(void *stream);

/* ---- set-up, host only ---- */
/* _Healpix_init + _update_params (HCGrid/healpix.cu:42-65).  RING only. */
int hcg_healpix_init(uint64_t nside, hcg_healpix *hp);
/* set_optimal_nside (HCGrid/helpers.cu:98-102); target resolution in radians. */
uint64_t hcg_set_optimal_nside(double target_res_rad);
/* _prepare_grid_kernel (HCGrid/gmap.cpp:64-99).  kernel_params (host, 3
 * doubles) is read, not modified; the transformed values land in kern->params.
 * hpx_max_res_deg fixes nside.  Returns HCG_ERR_INVALID for unknown types. */
int hcg_prepare_grid_kernel(uint32_t kernel_type, const double *kernel_params,
                            double sphere_radius_deg, double hpx_max_res_deg,
                            hcg_kernel *kern, hcg_healpix *hp);

/* Host evaluations of the same HEALPix functions the device uses (for
 * callers that mirror the reference's h_* helpers, HCGrid/healpix.cu:52-348). */
uint64_t hcg_host_ang2pix(const hcg_healpix *hp, double theta, double phi);
uint64_t hcg_host_pix2ring(const hcg_healpix *hp, uint64_t pix);
void     hcg_host_ring_info(const hcg_healpix *hp, uint64_t ring, uint64_t *startpix, uint64_t *npr);
double   hcg_host_pix2theta(const hcg_healpix *hp, uint64_t pix);

/* ---- the device path ---- */

/* Workspace bytes hcg_presort() needs for n samples. */
size_t hcg_presort_workspace_size(uint32_t n);

/* HEALPix keys + stable sort: the ang2pix loop and sort of
 * init_input_with_cpu (HCGrid/gridding.cu:70-100; init_input_with_thrust,
 * :149-168).  Keys are RING pixels of (theta, phi) = (HALFPI - DEG2RAD*lat,
 * DEG2RAD*lon), HCG_MAX_Y for theta outside [0, pi].  Output is stable by
 * (key, original index), the order of sort_param 2 (parallel_stable_sort),
 * for every caller.  keys_sorted holds n+1 entries; entry n is the npix
 * sentinel (HCGrid/gridding.cu:100).  perm[i] = original index of slot i. */
int hcg_presort(const hcg_healpix *hp, uint32_t n, const double *lon_deg, const double *lat_deg,
                uint32_t *keys_sorted, uint32_t *perm, void *workspace, size_t workspace_bytes,
                void *stream);

/* hcg_presort without a workspace argument and without allocating (no
 * hipMalloc, hipFree or allocating library call): no counterpart in the
 * reference; for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid it).  The
 * sort's scratch is a static device array of the library (66 MiB per GPU,
 * held until the process ends), so n <= HCG_PRESORT_NOALLOC_MAX_N
 * (2^23 = 8388608 samples); larger n is rejected with
 * HCG_ERR_UNSUPPORTED.  Results equal hcg_presort's bit for bit.  Every
 * call on one GPU uses that GPU's one array: calls that overlap on a GPU
 * (other streams or host threads) race and must be serialised by the
 * caller.  Nothing in the array needs clearing between calls (every byte
 * is written before it is read).  The first call on a device resolves the
 * array's device address (hipGetSymbolAddress), which enqueues no work. */
#define HCG_PRESORT_NOALLOC_MAX_N 8388608u
int hcg_presort_noalloc(const hcg_healpix *hp, uint32_t n, const double *lon_deg, const double *lat_deg,
                        uint32_t *keys_sorted, uint32_t *perm, void *stream);

/* The three permutation gathers of init_input_with_cpu (HCGrid/gridding.cu:101-114),
 * fused, plus the sample records the gridding kernel reads.  data has
 * data_dtype (HCG_DTYPE_F64 or HCG_DTYPE_F32).  weights may be NULL (all 1,
 * as read_input_map sets them, HCGrid/gmap.cpp:140-142).  Any of lon_sorted,
 * lat_sorted, data_sorted, weights_sorted may be NULL to skip that output;
 * records must not be NULL. */
int hcg_gather_samples(uint32_t n, const uint32_t *perm, const double *lon_deg,
                       const double *lat_deg, const void *data, int data_dtype,
                       const double *weights, hcg_sample *records, double *lon_sorted,
                       double *lat_sorted, double *data_sorted, double *weights_sorted,
                       void *stream);

/* Two-level ring lookup table, level 1 (HCGrid/gridding.cu:116-136), built on
 * the device.  lut[k] = first sorted index whose key >= startpix(firstring+k)
 * for 0 < k < usedrings, lut[0] = 0, lut[usedrings] = n.  ring_z/ring_s hold
 * cos/sin of the centre colatitude of each used ring (what the kernel's
 * pix2ang(startpix) + ang2pix recompute per cell, HCGrid/gridding.cu:332-351).
 * Capacity: lut and ring_z/ring_s need hp->nrings + 2 entries.  meta (device)
 * receives firstring/usedrings/n_valid. */
int hcg_ring_lut(const hcg_healpix *hp, uint32_t n, const uint32_t *keys_sorted, uint32_t *lut,
                 double *ring_z, double *ring_s, hcg_ring_meta *meta, void *stream);

/* The `hcgrid` kernel (HCGrid/gridding.cu:277-403): for each of side*side
 * cells (row length `side` = NAXIS2, HCGrid/gmap.cpp:174; cell c =
 * row*side + col), accumulate w*v and w into datacube/weightscube in place
 * over the reference's candidate set: rings [uring, dring] of the window
 * anchor (the first cell of each `factor` group), skipping the first used
 * ring, sorted indices with key in [lpix_r, rpix_r] (whole ring when the
 * anchor's colatitude is <= 20 or >= 160 degrees), and great-circle distance
 * < sphere_radius.  Weight w = weights_sorted[i] * kernel(d) (weights_sorted
 * may be NULL = 1).  FP64 geometry and accumulation.  lon/lat_sorted are
 * needed only for GAUSS2D bearings and distances at the radius boundary. */
int hcg_grid_plain(const hcg_healpix *hp, const hcg_kernel *kern, uint32_t side,
                   uint32_t n, const hcg_sample *records, const uint32_t *keys_sorted,
                   const double *lon_sorted, const double *lat_sorted,
                   const double *weights_sorted, const uint32_t *lut, const double *ring_z,
                   const double *ring_s, const hcg_ring_meta *meta, const double *xwcs,
                   const double *ywcs, double *datacube, double *weightscube, void *stream);

/* ---- multi-channel (spectral cube) path ---- */

/* Padded samples per channel of the sorted channel buffer: n rounded up to
 * 32 plus 64 zero samples.  The buffer holds nchan * hcg_channels_padded_len(n)
 * values, blocked by 8 sorted samples: [padded_len / 8][nchan][8], so one
 * lane's MFMA operand (8 samples of one channel) is one 16-byte block row, 16
 * channels of a block are one contiguous 256-byte piece, and the contraction
 * packs ring ranges at 8-sample granularity. */
uint32_t hcg_channels_padded_len(uint32_t n);

/* The per-channel permutation gather of HCGrid+'s pre_order_data
 * (HCGrid+/gridding-omp.cu:146-157), all channels at once: values (dtype
 * HCG_DTYPE_BF16 or HCG_DTYPE_F16) laid out channel-major [nchan][n]
 * (layout 0, the reference's per-channel arrays) or sample-major
 * [n][nchan] (layout 1) -> values_sorted in the blocked layout above, in
 * sorted sample order, zero-padded.  nonfinite (device, may be NULL)
 * is incremented once per non-finite value: the contraction multiplies
 * every staged value, so non-finite channel values are not supported. */
int hcg_gather_channels(uint32_t n, uint32_t nchan, const uint32_t *perm, const void *values, int dtype,
                        int layout, void *values_sorted, uint32_t *nonfinite, void *stream);

/* Multi-channel gridding: what HCGrid+ computes by running `hcgrid` once per
 * channel (HCGrid+/gridding-omp.cu:552-678), as one pass.  Same selection
 * and FP64 weights as hcg_grid_plain; the weights are rounded to dtype and
 * contracted with the channel values on MFMA with FP32 accumulation:
 *   cube[ch][c] = sum_s w~(c,s) v[ch][s] / sum_s w~(c,s)   (normalize != 0)
 * or the raw numerator (normalize == 0), w~ = weight rounded to dtype.
 * cube is FP32 [nchan][side*side] (channel-major, like the reference's
 * datacube planes); wsum (FP32 [side*side], may be NULL) receives sum w~.
 * NaN where a cell has no weight.  Samples are the sorted records/keys of
 * hcg_presort/hcg_gather_samples; values_sorted from hcg_gather_channels.
 * Routes (which weight precision applies):
 *  - one channel (nchan == 1) with GAUSS1D, no weights_sorted and a radius
 *    below ~0.26 deg, on a map of any side, runs the single-channel FP64
 *    tile loop instead of the contraction (w~ = the unrounded FP64 weight;
 *    the launch shape is ignored).  Maps of at least 4096 tiles of 16 x 4
 *    cells (side >= 512) give each tile one wave; smaller maps split each
 *    tile's rings over 4 partial streams, summed in stream order;
 *  - 2..8 channels with GAUSS1D, no weights_sorted, a radius between
 *    ~0.018 deg (where the candidate test on the unit vectors' dot product
 *    is exact enough) and ~0.26 deg, on sparse maps of at least 4096 tiles
 *    (n <= side * side / 10), run that loop too, with one FP32 sum per
 *    channel (weights rounded once to FP32; wsum is the FP64 weight sum);
 *  - everything else runs the MFMA contraction (w~ = weight rounded to
 *    dtype as hi + lo).
 * Batch jobs take the same routes, so a job's cube is the same alone or
 * batched. */
int hcg_grid_mfma(const hcg_healpix *hp, const hcg_kernel *kern, uint32_t side, uint32_t n, uint32_t nchan,
                  int dtype, const hcg_sample *records, const uint32_t *keys_sorted,
                  const double *lon_sorted, const double *lat_sorted, const double *weights_sorted,
                  const void *values_sorted, const uint32_t *lut, const double *ring_z,
                  const double *ring_s, const hcg_ring_meta *meta, const double *xwcs,
                  const double *ywcs, float *cube, float *wsum, int normalize, void *stream);

/* Launch shape of the multi-channel kernel.  A workgroup owns 64 cells,
 * a (64 / tile_cols) x tile_cols block of the map (tile_cols 8, 16, 32 or
 * 64), with `waves` waves (1, 2, 4 or 8; 64 channels each; more channels
 * take more workgroups, each recomputing the block's weights) and stages
 * `chunk` samples at a time (128, 256 or 512).  0 in a field = the default
 * (tile_cols 16; the fewest waves that cover the channels, at most 8; chunk
 * by waves).  No counterpart in the reference: the shapes an autotuner
 * picks per workload. */
typedef struct hcg_mfma_shape {
    uint32_t tile_cols, waves, chunk, reserved;
} hcg_mfma_shape;

/* Workspace bytes of hcg_grid_mfma_ex for a side x side map and a launch
 * shape (NULL = defaults): the tile plan, 3600 bytes per 64-cell tile. */
size_t hcg_grid_mfma_workspace_size(uint32_t side, const hcg_mfma_shape *shape);

/* hcg_grid_mfma with a launch shape (NULL = defaults) and a workspace for
 * the tile plan (NULL = none).  With a workspace, a first small kernel (one
 * wave per tile, many per CU) computes every tile's ring span, each ring's
 * sorted range and each cell's window and unit vector, and the gridding
 * kernel reads them instead of evaluating them itself; the results are the
 * same either way. */
int hcg_grid_mfma_ex(const hcg_healpix *hp, const hcg_kernel *kern, uint32_t side, uint32_t n,
                     uint32_t nchan, int dtype, const hcg_sample *records, const uint32_t *keys_sorted,
                     const double *lon_sorted, const double *lat_sorted, const double *weights_sorted,
                     const void *values_sorted, const uint32_t *lut, const double *ring_z,
                     const double *ring_s, const hcg_ring_meta *meta, const double *xwcs,
                     const double *ywcs, float *cube, float *wsum, int normalize,
                     const hcg_mfma_shape *shape, void *workspace, size_t workspace_bytes, void *stream);

/* hcg_grid_mfma with an explicit launch shape (NULL = defaults), no tile
 * plan.  Results do not depend on the shape beyond FP32 summation order. */
int hcg_grid_mfma_shaped(const hcg_healpix *hp, const hcg_kernel *kern, uint32_t side, uint32_t n,
                         uint32_t nchan, int dtype, const hcg_sample *records, const uint32_t *keys_sorted,
                         const double *lon_sorted, const double *lat_sorted, const double *weights_sorted,
                         const void *values_sorted, const uint32_t *lut, const double *ring_z,
                         const double *ring_s, const hcg_ring_meta *meta, const double *xwcs,
                         const double *ywcs, float *cube, float *wsum, int normalize,
                         const hcg_mfma_shape *shape, void *stream);

/* Output normalisation of write_output_map (HCGrid/gmap.cpp:305-312):
 * map[c] = datacube[c] / weightscube[c] in FP64, NaN where the weight is 0.
 * map may alias datacube. */
int hcg_normalize(uint64_t ncells, const double *datacube, const double *weightscube,
                  double *map, void *stream);

/* The device exponential the gridding kernels use for GAUSS1D weights,
 * evaluated elementwise (x <= 0): a validation hook, no reference counterpart. */
int hcg_eval_exp_neg(uint32_t n, const double *x, double *out, void *stream);

/* ---- host-side target geometry ---- */
/* Zenithal projections of the target map (Calabretta & Greisen 2002, Paper
 * II sec. 5.1).  The reference's set_WCS accepts whatever wcslib parses from
 * the target header (HCGrid/gmap.cpp:183-250); these cover the zenithal
 * family with a diagonal CDELT and the native pole at CRVAL.  The host
 * layer's set_WCS (hcgrid_host.h) calls hcg_wcs_p2s for such targets and
 * maps the other projections it supports (slanted SIN, NCP, cylindricals,
 * SFL, GLS, AIT; PC, CD or CROTA2 matrices) on the host itself. */
#define HCG_PROJ_SIN 0u  /* orthographic, PV2_1 = PV2_2 = 0 */
#define HCG_PROJ_TAN 1u  /* gnomonic */
#define HCG_PROJ_ARC 2u  /* zenithal equidistant */
#define HCG_PROJ_ZEA 3u  /* zenithal equal-area */
#define HCG_PROJ_STG 4u  /* stereographic */

/* set_WCS for a zenithal target (HCGrid/gmap.cpp:183-250): cell (x, y) at
 * pixel (x+1, y+1) -> world (lon, lat) degrees, cell index y*naxis1 + x,
 * NaN outside the projection's domain.  Host pointers.  proj is one of
 * HCG_PROJ_*. */
int hcg_wcs_p2s(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                double cdelt1, double cdelt2, double crval1, double crval2,
                double lonpole, double *xwcs, double *ywcs);

/* The same projection evaluated on the device (device outputs, enqueued on
 * stream).  Same formula and operation order; the device sin/cos/atan2/asin
 * are the HIP math library's, so results may differ from the host form in
 * the last bit. */
int hcg_wcs_p2s_device(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                       double cdelt1, double cdelt2, double crval1, double crval2,
                       double lonpole, double *xwcs, double *ywcs, void *stream);

/* hcg_wcs_p2s / hcg_wcs_p2s_device with proj = HCG_PROJ_SIN. */
int hcg_wcs_sin_p2s(uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                    double cdelt1, double cdelt2, double crval1, double crval2,
                    double lonpole, double *xwcs, double *ywcs);
int hcg_wcs_sin_p2s_device(uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                           double cdelt1, double cdelt2, double crval1, double crval2,
                           double lonpole, double *xwcs, double *ywcs, void *stream);

/* ---- resident geometries and batched multi-channel gridding ----
 * No counterpart in the reference, which grids one map per run
 * (HCGrid/HCGrid.cpp:38-197) and one channel per kernel pass
 * (HCGrid+/gridding-omp.cu:552-678).  A geometry holds one map's sorted
 * samples, ring LUT and target coordinates in HBM, so repeated requests on
 * the same coordinates skip the pre-sort; a batch grids many (geometry,
 * channel values, cube) jobs of uneven sizes in ONE persistent launch. */

typedef struct hcg_geometry hcg_geometry;   /* opaque, device-resident */

/* Device bytes hcg_geometry_create allocates for n samples and side x side
 * cells at hp's nside. */
size_t hcg_geometry_bytes(uint32_t n, uint32_t side, const hcg_healpix *hp);

/* Allocates a geometry (copies hp and kern).  Build it before use. */
int hcg_geometry_create(const hcg_healpix *hp, const hcg_kernel *kern, uint32_t n, uint32_t side,
                        hcg_geometry **out);

/* The same in caller-provided device memory (>= hcg_geometry_bytes, 256-byte
 * aligned), e.g. from a caching allocator: creating and destroying it then
 * make no allocation call (hipFree synchronises the device). */
int hcg_geometry_create_in(const hcg_healpix *hp, const hcg_kernel *kern, uint32_t n, uint32_t side,
                           void *device_mem, size_t bytes, hcg_geometry **out);

/* Enqueues keys + stable sort, sample records and ring LUT (hcg_presort,
 * hcg_gather_samples, hcg_ring_lut) and copies the target coordinates
 * (device, [side*side] degrees; skipped when they are the geometry's own
 * arrays) into the geometry.  lon/lat: device, [n] degrees.  Kernels and
 * device-to-device copies only: capturable into a graph. */
int hcg_geometry_build(hcg_geometry *g, const double *lon_deg, const double *lat_deg, const double *xwcs,
                       const double *ywcs, void *stream);

/* Device arrays of a geometry (read-only views for the one-job entry points). */
typedef struct hcg_geometry_view {
    uint32_t n, side;
    size_t bytes;
    const uint32_t *keys_sorted, *perm;
    const hcg_sample *records;
    const double *lon_sorted, *lat_sorted;
    const uint32_t *lut;
    const double *ring_z, *ring_s;
    const hcg_ring_meta *meta;
    double *xwcs, *ywcs;
} hcg_geometry_view;
int hcg_geometry_arrays(const hcg_geometry *g, hcg_geometry_view *view);

int hcg_geometry_destroy(hcg_geometry *g);

/* One job of a batch. */
typedef struct hcg_batch_job {
    const hcg_geometry *geometry;
    uint32_t nchan;
    int dtype;                /* HCG_DTYPE_BF16 or HCG_DTYPE_F16 */
    const void *values;       /* device, [nchan][n] channel-major */
    float *cube;              /* device, [nchan][side*side] */
    float *wsum;              /* device, [side*side], or NULL */
    uint32_t tile_cols;       /* 8, 16, 32, 64, or 0 = 16 */
    int normalize;            /* nonzero: cube = sum w v / sum w (NaN where no weight) */
} hcg_batch_job;

typedef struct hcg_batch hcg_batch;   /* opaque */

/* Device bytes of a job's sorted channel values in its batch (the batch
 * adds the job's tile plan, 520 bytes per 64-cell tile, and its work items). */
size_t hcg_batch_job_bytes(const hcg_geometry *g, uint32_t nchan);

/* Plans a batch: orders the work items (tile x channel group of every job,
 * longest items first), allocates the job table, work list, counters and
 * the sorted channel buffers, and uploads the plan (synchronously).  waves
 * per workgroup (1, 2, 4, 8; 0 = 4) and chunk (128, 256, 512; 0 = by waves)
 * fix the persistent launch's shape.  The jobs' geometries, values and
 * cubes must stay allocated while the batch is used. */
int hcg_batch_create(uint32_t njobs, const hcg_batch_job *jobs, uint32_t waves, uint32_t chunk,
                     hcg_batch **out);

/* Device bytes hcg_batch_create_in needs for these jobs and launch shape
 * (0 if the jobs are invalid). */
size_t hcg_batch_bytes(uint32_t njobs, const hcg_batch_job *jobs, uint32_t waves, uint32_t chunk);

/* hcg_batch_create in caller-provided device memory (>= hcg_batch_bytes,
 * 256-byte aligned); the plan is uploaded asynchronously on `stream`, so
 * the batch must be enqueued on that stream (or after it). */
int hcg_batch_create_in(uint32_t njobs, const hcg_batch_job *jobs, uint32_t waves, uint32_t chunk,
                        void *device_mem, size_t bytes, void *stream, hcg_batch **out);

/* Enqueues the batch: counters cleared by a fill kernel, every job's channel
 * gather and tile plan in one launch, then the persistent launch over the
 * work items (dense jobs, 32+ samples per cell, in a persistent launch of
 * their own at 8 waves / chunk 512 first, when the batch's shape is not
 * that one).  Kernels only
 * (capturable).  Results per job equal hcg_grid_mfma_shaped's with the same
 * tile_cols, bit for bit. */
int hcg_batch_enqueue(hcg_batch *b, void *stream);

typedef struct hcg_batch_stats {
    uint32_t njobs, nitems, blocks, waves, chunk, reserved;
    size_t device_bytes;
} hcg_batch_stats;
int hcg_batch_info(const hcg_batch *b, hcg_batch_stats *stats);

/* Per-job counts of non-finite channel values seen by the last run (host
 * array of njobs; synchronises the stream). */
int hcg_batch_nonfinite(const hcg_batch *b, uint32_t *counts, void *stream);

/* Kernel timing of hcg_batch_enqueue, for measurement (no reference
 * counterpart).  enable != 0 creates three HIP events in the batch; every
 * later enqueue that is not being captured into a graph records them on its
 * stream: before the counter fill, after the prep launch (channel gathers
 * and tile plans), and after the persistent gridding launches.
 * hcg_batch_last_times waits for the last such enqueue and returns the prep
 * and gridding durations in milliseconds.  enable == 0 destroys the events. */
int hcg_batch_timing(hcg_batch *b, int enable);
int hcg_batch_last_times(hcg_batch *b, float *prep_ms, float *grid_ms);

int hcg_batch_destroy(hcg_batch *b);

/* 128-bit content hash of `bytes` bytes of device memory (8-byte aligned):
 * out (device, 2 x uint64, 8-byte aligned) receives
 *   h_k = mix64(bytes ^ B_(1-k)) + sum_i mix64(w_i ^ (i * A_k + B_k)),
 * w_i the buffer's 8-byte little-endian words (the tail zero-padded),
 * mix64 the splitmix64 finaliser, A_0 = 0x9e3779b97f4a7c15,
 * B_0 = 0x243f6a8885a308d3, A_1 = 0xc2b2ae3d27d4eb4f, B_1 = 0x13198a2e03707344,
 * all modulo 2^64.  Deterministic (the sum is order-free).  No reference
 * counterpart: the served loop's key for resident geometries (samples
 * sorted once, reused for repeat coordinates). */
int hcg_content_hash(const void *data, size_t bytes, uint64_t *out, void *stream);

/* Graph capture of any sequence of the entry points above on a created
 * stream (not the null stream), then replay.  No reference counterpart. */
int hcg_graph_capture_begin(void *stream);
int hcg_graph_capture_end(void *stream, void **graph_exec);
int hcg_graph_launch(void *graph_exec, void *stream);
int hcg_graph_destroy(void *graph_exec);

#ifdef __cplusplus
}
#endif
#endif /* HCGRID_AMD_H */

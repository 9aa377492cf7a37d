/*
 * hcg_oracle.h -- CPU restatement of HCGrid's gridding path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * implementation under hcgrid_amd/.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; no product code links or calls it.
 *
 * It restates, operation for operation, the reference's host HEALPix code,
 * the pre-sort + two-level ring lookup table, and the `hcgrid` kernel
 * (HCGrid/healpix.cu, HCGrid/helpers.cu, HCGrid/gridding.cu, HCGrid/gmap.cpp),
 * compiled with -ffp-contract=off so every FP64 expression rounds as on the
 * x86-64 host the reference ran on.  Each function cites the lines it follows.
 *
 * Pinning (see tests/test_oracle_*.py):
 *   HEALPix   : closed forms (ring starts, ring sizes, nside=1 base pixels),
 *               exhaustive pix2ang->ang2pix round trips at small nside, and an
 *               independent projection-plane pixelisation (Calabretta & Roukema
 *               2007) in numpy on random points.
 *   presort   : numpy stable argsort.
 *   ring LUT  : numpy searchsorted over ring start pixels.
 *   grid      : an independent per-pair predicate brute force, and a textbook
 *               disc brute force in long double where the compat window
 *               contains the whole disc (|dec| small, first ring far away).
 */
#ifndef HCG_ORACLE_H
#define HCG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t nside, order, nrings, npface, ncap, npix;
    double fact1, fact2, omega, resolution;
} orc_hpx;

/* kernel description: the state _prepare_grid_kernel leaves behind */
typedef struct {
    uint32_t kernel_type;     /* 1 GAUSS1D, 2 GAUSS2D, 3 TAPERED_SINC */
    uint32_t factor;          /* thread-coarsening factor gamma (>=1) */
    uint32_t side;            /* NAXIS2: the row length and row count the kernel covers */
    uint32_t bearing_needed;
    double params[3];         /* already transformed (GAUSS1D: 0.5/sigma^2) */
    double sphere_radius;     /* degrees */
    double disc_size;         /* radians: DEG2RAD*R + resolution */
} orc_grid;

/* ---- helpers (HCGrid/helpers.cu) ---- */
uint64_t orc_ilog2(uint64_t arg);
uint64_t orc_isqrt(uint64_t arg);
double   orc_fmodulo(double v1, double v2);
uint64_t orc_set_optimal_nside(double target_res_rad);
uint64_t orc_nside_to_order(uint64_t nside);
double   orc_true_angular_distance(double l1, double b1, double l2, double b2);
double   orc_great_circle_bearing(double l1, double b1, double l2, double b2);

/* ---- HEALPix RING (HCGrid/healpix.cu) ---- */
void     orc_hpx_init(orc_hpx *h, uint64_t nside);
void     orc_ring_info(const orc_hpx *h, uint64_t ring, uint64_t *startpix,
                       uint64_t *npr, int *shifted);
uint64_t orc_pix2ring(const orc_hpx *h, uint64_t pix);
uint64_t orc_loc2pix(const orc_hpx *h, double z, double phi, double sin_theta, int have_sin);
uint64_t orc_ang2pix(const orc_hpx *h, double theta, double phi);
void     orc_pix2loc(const orc_hpx *h, uint64_t pix, double *z, double *phi,
                     double *sin_theta, int *have_sin);
void     orc_pix2ang(const orc_hpx *h, uint64_t pix, double *theta, double *phi);

/* batched forms for the tests */
void orc_ang2pix_many(const orc_hpx *h, int64_t n, const double *theta,
                      const double *phi, uint64_t *out);
void orc_pix2ang_many(const orc_hpx *h, int64_t n, const uint64_t *pix,
                      double *theta, double *phi);
void orc_pix2ring_many(const orc_hpx *h, int64_t n, const uint64_t *pix, uint64_t *out);
void orc_ring_info_many(const orc_hpx *h, int64_t n, const uint64_t *ring,
                        uint64_t *startpix, uint64_t *npr);
/* sample keys exactly as the pre-sort computes them from (lon, lat) in degrees */
void orc_sample_keys(const orc_hpx *h, int64_t n, const double *lon_deg,
                     const double *lat_deg, uint64_t *keys);
/* samples whose key changes for some cos (sin) within +-ulps of this libm's */
int64_t orc_sample_keys_libm_sensitive(const orc_hpx *h, int64_t n, const double *lon_deg,
                                       const double *lat_deg, int ulps, uint8_t *flags);

/* ---- pre-sort + ring LUT (HCGrid/gridding.cu:70-147, 260-275) ---- */
uint32_t orc_search_last_less(const uint64_t *v, uint32_t left, uint32_t right, uint64_t key);
/* keys_sorted[n] = npix sentinel; perm[i] = original index of sorted slot i.
 * Stable by (key, original index): the order boost::parallel_stable_sort gives. */
void orc_presort(const orc_hpx *h, uint32_t n, const double *lon_deg, const double *lat_deg,
                 uint64_t *keys_sorted, uint32_t *perm);
/* gathers src through perm (double arrays) */
void orc_gather(uint32_t n, const uint32_t *perm, const double *src, double *dst);
/* returns usedrings (0 on error: lut_cap too small or n == 0). lut holds usedrings+1. */
uint32_t orc_ring_lut(const orc_hpx *h, uint32_t n, const uint64_t *keys_sorted,
                      uint64_t *firstring, uint32_t *lut, uint32_t lut_cap);

/* ---- gridding (HCGrid/gridding.cu:230-403) ----
 * Accumulates into datacube/weightscube in place (as the kernel does), over
 * side*side cells.  Samples are in sorted order.  weights may be NULL (=1).
 * Returns the number of (group, ring) visits where the reference would have
 * looped forever (lpix/rpix outside the ring); those rings are skipped here. */
int64_t orc_grid_compat(const orc_hpx *h, const orc_grid *g, uint32_t n,
                        const double *lon, const double *lat, const double *data,
                        const double *weights, const uint64_t *keys_sorted,
                        uint64_t firstring, uint32_t usedrings, const uint32_t *lut,
                        const double *xwcs, const double *ywcs,
                        double *datacube, double *weightscube, int nthreads);

/* orc_grid_compat restricted to rows [row_begin, row_end) of the map. */
int64_t orc_grid_compat_rows(const orc_hpx *h, const orc_grid *g, uint32_t n,
                             const double *lon, const double *lat, const double *data,
                             const double *weights, const uint64_t *keys_sorted,
                             uint64_t firstring, uint32_t usedrings, const uint32_t *lut,
                             const double *xwcs, const double *ywcs,
                             double *datacube, double *weightscube,
                             uint32_t row_begin, uint32_t row_end, int nthreads);

/* Multi-channel form over rows [row_begin, row_end): FP64 weights of the
 * same selection applied to nchan channel-major float values
 * (values[ch * n + i], sorted order).  out[ch * rows*side + k] = sum w v,
 * wsum[k] = sum w (raw sums, no normalisation). */
int64_t orc_grid_compat_multi(const orc_hpx *h, const orc_grid *g, uint32_t n,
                              const double *lon, const double *lat, uint32_t nchan, const float *values,
                              const double *weights, const uint64_t *keys_sorted,
                              uint64_t firstring, uint32_t usedrings, const uint32_t *lut,
                              const double *xwcs, const double *ywcs,
                              uint32_t row_begin, uint32_t row_end, double *out, double *wsum,
                              int nthreads);

/* Same selection, evaluated pair by pair over all samples (no LUT, no search). */
void orc_grid_bruteforce(const orc_hpx *h, const orc_grid *g, uint32_t n,
                         const double *lon, const double *lat, const double *data,
                         const double *weights, const uint64_t *keys_sorted,
                         uint64_t firstring, uint32_t usedrings,
                         const double *xwcs, const double *ywcs,
                         double *datacube, double *weightscube, int nthreads);

/* Textbook disc gridding: every sample with great-circle distance < R,
 * Gaussian weight exp(-d^2 p0), long double throughout, distance via
 * atan2(|p x q|, p . q).  Outputs raw sums (no normalisation). */
void orc_grid_disc_textbook(uint32_t ncells, uint32_t n, const double *lon, const double *lat,
                            const double *data, const double *xwcs, const double *ywcs,
                            double sphere_radius_deg, double p0,
                            double *datacube, double *weightscube, int nthreads);

/* the kernel weight function (HCGrid/gridding.cu:238-257) */
double orc_kernel_weight(const orc_grid *g, double distance, double bearing);

/* output normalisation (HCGrid/gmap.cpp:305-312): map = data / weights */
void orc_normalize(int64_t ncells, const double *datacube, const double *weightscube, double *map);

/* Zenithal deprojections, Calabretta & Greisen 2002 sec. 5.1, with the
 * native pole at CRVAL and phi_p = lonpole (degrees).  proj: 0 SIN
 * (PV2_1 = PV2_2 = 0), 1 TAN, 2 ARC, 3 ZEA, 4 STG.  The native latitude is
 * taken through the inverse of each projection's R(theta) in long double.
 * Pixel (x+1, y+1) of a NAXIS1 x NAXIS2 grid -> (lon, lat) degrees, cell
 * index y*naxis1 + x (HCGrid/gmap.cpp:209-238); NaN outside the domain. */
void orc_wcs_zenithal(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                      double cdelt1, double cdelt2, double crval1, double crval2, double lonpole,
                      double *xwcs, double *ywcs);

/* orc_wcs_zenithal for SIN with LONPOLE 180 */
void orc_wcs_sin(uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                 double cdelt1, double cdelt2, double crval1, double crval2,
                 double *xwcs, double *ywcs);

/* General celestial WCS (Calabretta & Greisen 2002): linear matrix m
 * (row-major, degrees per pixel), projection proj (0 SIN with slant
 * (pv1, pv2), 1 TAN, 2 ARC, 3 ZEA, 4 STG, 5 CAR, 6 MER, 7 CEA with
 * lambda = pv1, 8 SFL, 9 GLS (AIPS), 10 AIT), native -> celestial rotation
 * with phi_p = lonpole (NaN: the default) and delta_p chosen by latpole.
 * Returns 0, or -1 when the LONPOLE admits no native pole. */
int orc_wcs_celestial(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                      const double *m, double crval1, double crval2, double lonpole, double latpole,
                      double pv1, double pv2, double *xwcs, double *ywcs);

#ifdef __cplusplus
}
#endif
#endif

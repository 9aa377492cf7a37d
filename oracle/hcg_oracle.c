/*
 * hcg_oracle.c -- CPU restatement of HCGrid's gridding path (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker.  Never part of the product.
 *
 * Build: gcc -O2 -ffp-contract=off -fopenmp (no -ffast-math, no -mfma): the
 * reference's host code ran on x86-64 with SSE2 double arithmetic, where
 * every operation rounds separately.
 *
 * Every function names the reference lines it restates.  Constants follow
 * HCGrid/constants.h:57-84 (the PI literal, and macros folded in double).
 */
#include "hcg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* HCGrid/constants.h:57,75-84 */
#define ORC_MAX_Y    (1073741824ull)
#define ORC_PI       (3.1415926535897932384626433832)
#define ORC_TWOTHIRD (2.0 / 3.0)
#define ORC_HALFPI   (ORC_PI / 2.)
#define ORC_INV_HALFPI (1. / ORC_HALFPI)
#define ORC_NORTH_B  (20)
#define ORC_SOUTH_B  (160)
#define ORC_DEG2RAD  (ORC_PI / 180.)
#define ORC_RAD2DEG  (180. / ORC_PI)

/* ------------------------------------------------------------------ */
/* helpers: HCGrid/helpers.cu                                          */
/* ------------------------------------------------------------------ */

/* HCGrid/helpers.cu:40-64 */
uint64_t orc_ilog2(uint64_t arg) {
    uint64_t r = 0;
    while (arg > 0xFFFFull) { r += 16; arg >>= 16; }
    if (arg > 0xFFull) { r |= 8; arg >>= 8; }
    if (arg > 0xFull)  { r |= 4; arg >>= 4; }
    if (arg > 0x3ull)  { r |= 2; arg >>= 2; }
    if (arg > 0x1ull)  { r |= 1; }
    return r;
}

/* HCGrid/helpers.cu:67-69 */
uint64_t orc_isqrt(uint64_t arg) { return (uint64_t)sqrt((double)arg + 0.5); }

/* HCGrid/helpers.cu:83-95 */
double orc_fmodulo(double v1, double v2) {
    if (v1 >= 0) {
        if (v1 < v2) return v1;
        return fmod(v1, v2);
    }
    double tmp = fmod(v1, v2) + v2;
    if (tmp == v2) return 0.;
    return tmp;
}

/* HCGrid/helpers.cu:98-102: always one power of two above floor(log2). */
uint64_t orc_set_optimal_nside(double target_res) {
    uint64_t nside = (uint64_t)(sqrt(ORC_PI / 3.) / target_res + 0.5);
    nside = (uint64_t)1 << (orc_ilog2(nside) + 1);
    return nside;
}

/* HCGrid/helpers.cu:105-114 (returns (uint64_t)-1 for non powers of two) */
uint64_t orc_nside_to_order(uint64_t nside) {
    if (nside == 0) return ORC_MAX_Y;
    if (nside & (nside - 1)) return (uint64_t)-1;
    return orc_ilog2(nside);
}

/* HCGrid/helpers.cu:128-130 (haversine) */
double orc_true_angular_distance(double l1, double b1, double l2, double b2) {
    return 2 * asin(sqrt(pow(sin((b1 - b2) / 2.), 2) +
                         cos(b1) * cos(b2) * pow(sin((l1 - l2) / 2.), 2)));
}

/* HCGrid/helpers.cu:133-137 */
double orc_great_circle_bearing(double l1, double b1, double l2, double b2) {
    double l_diff_rad = l2 - l1;
    double cos_b2 = cos(b2);
    return atan2(cos_b2 * sin(l_diff_rad), cos(b1) * sin(b2) - sin(b1) * cos_b2 * cos(l_diff_rad));
}

/* ------------------------------------------------------------------ */
/* HEALPix RING: HCGrid/healpix.cu                                     */
/* ------------------------------------------------------------------ */

/* HCGrid/healpix.cu:42-65 */
void orc_hpx_init(orc_hpx *h, uint64_t nside) {
    h->nside = nside;
    h->order = orc_nside_to_order(nside);
    h->nrings = 4 * nside - 1;
    h->npface = nside * nside;
    h->ncap = (h->npface - nside) << 1;
    h->npix = 12 * h->npface;
    h->fact2 = 4. / (double)h->npix;
    h->fact1 = (double)(nside << 1) * h->fact2;
    h->omega = 4. * ORC_PI / (double)h->npix;
    h->resolution = sqrt(h->omega);
}

/* HCGrid/healpix.cu:87-105 */
void orc_ring_info(const orc_hpx *h, uint64_t ring, uint64_t *startpix, uint64_t *npr, int *shifted) {
    uint64_t nr;
    if (ring < h->nside) {
        *shifted = 1;
        *npr = 4 * ring;
        *startpix = 2 * ring * (ring - 1);
    } else if (ring < 3 * h->nside) {
        *shifted = ((ring - h->nside) & 1) == 0;
        *npr = 4 * h->nside;
        *startpix = h->ncap + (ring - h->nside) * (*npr);
    } else {
        *shifted = 1;
        nr = 4 * h->nside - ring;
        *npr = 4 * nr;
        *startpix = h->npix - 2 * nr * (nr + 1);
    }
}

/* HCGrid/healpix.cu:116-123 */
uint64_t orc_pix2ring(const orc_hpx *h, uint64_t pix) {
    if (pix < h->ncap) return (1 + orc_isqrt(1 + 2 * pix)) >> 1;
    if (pix < h->npix - h->ncap) return (pix - h->ncap) / (4 * h->nside) + h->nside;
    return 4 * h->nside - ((1 + orc_isqrt(2 * (h->npix - pix) - 1)) >> 1);
}

/* HCGrid/healpix.cu:166-205 */
uint64_t orc_loc2pix(const orc_hpx *h, double z, double phi, double sin_theta, int have_sin) {
    double temp1, temp2, tp, tmp;
    double za = fabs(z);
    double tt = orc_fmodulo(phi * ORC_INV_HALFPI, 4.0);
    uint64_t nl4, jp, jm, ir, kshift, t1, ip;
    double dn = (double)h->nside;

    if (za <= ORC_TWOTHIRD) {
        nl4 = 4 * h->nside;
        temp1 = dn * (0.5 + tt);
        temp2 = dn * z * 0.75;
        jp = (uint64_t)(temp1 - temp2);
        jm = (uint64_t)(temp1 + temp2);
        ir = h->nside + 1 + jp - jm;
        kshift = 1 - (ir & 1);
        t1 = jp + jm + kshift + 1 + nl4 + nl4 - h->nside;
        if (h->order > 0) ip = (t1 >> 1) & (nl4 - 1);
        else ip = (t1 >> 1) % nl4;
        return h->ncap + (ir - 1) * nl4 + ip;
    }
    tp = tt - (double)((uint64_t)tt);
    if ((za < 0.99) || (!have_sin)) tmp = dn * sqrt(3. * (1. - za));
    else tmp = dn * sin_theta / sqrt((1. + za) / 3.);
    jp = (uint64_t)(tp * tmp);
    jm = (uint64_t)((1.0 - tp) * tmp);
    ir = jp + jm + 1;
    ip = (uint64_t)(tt * (double)ir);
    if (z > 0) return 2 * ir * (ir - 1) + ip;
    return h->npix - 2 * ir * (ir + 1) + ip;
}

/* HCGrid/healpix.cu:217-224 */
uint64_t orc_ang2pix(const orc_hpx *h, double theta, double phi) {
    if (!(theta >= 0 && theta <= ORC_PI)) return ORC_MAX_Y;
    if ((theta < 0.01) || (theta > ORC_PI - 0.01))
        return orc_loc2pix(h, cos(theta), phi, sin(theta), 1);
    return orc_loc2pix(h, cos(theta), phi, 0., 0);
}

/* HCGrid/healpix.cu:274-320 */
void orc_pix2loc(const orc_hpx *h, uint64_t pix, double *z, double *phi, double *sin_theta, int *have_sin) {
    double tmp, fodd;
    uint64_t ip, itmp, nl4, iphi, iring;
    *have_sin = 0;
    if (pix < h->ncap) {
        iring = (1 + orc_isqrt(1 + 2 * pix)) >> 1;
        iphi = (pix + 1) - 2 * iring * (iring - 1);
        tmp = (double)(iring * iring) * h->fact2;
        *z = 1.0 - tmp;
        if (*z > 0.99) { *sin_theta = sqrt(tmp * (2. - tmp)); *have_sin = 1; }
        *phi = ((double)iphi - 0.5) * ORC_HALFPI / (double)iring;
    } else if (pix < h->npix - h->ncap) {
        nl4 = 4 * h->nside;
        ip = pix - h->ncap;
        /* order is unsigned in the reference, so the `>= 0` branch is always taken */
        itmp = ip >> (h->order + 2);
        iring = itmp + h->nside;
        iphi = ip - nl4 * itmp + 1;
        fodd = (((iring + h->nside) & 1) > 0) ? 1.0 : 0.5;
        *z = (2. * (double)h->nside - (double)iring) * h->fact1;
        *phi = ((double)iphi - fodd) * ORC_PI * 0.75 * h->fact1;
    } else {
        ip = h->npix - pix;
        iring = (1 + orc_isqrt(2 * ip - 1)) >> 1;
        iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
        tmp = (double)(iring * iring) * h->fact2;
        *z = tmp - 1.0;
        if (*z < -0.99) { *sin_theta = sqrt(tmp * (2. - tmp)); *have_sin = 1; }
        *phi = ((double)iphi - 0.5) * ORC_HALFPI / (double)iring;
    }
}

/* HCGrid/healpix.cu:336-348 */
void orc_pix2ang(const orc_hpx *h, uint64_t pix, double *theta, double *phi) {
    double z, s = 0.;
    int have;
    orc_pix2loc(h, pix, &z, phi, &s, &have);
    *theta = have ? atan2(s, z) : acos(z);
}

void orc_ang2pix_many(const orc_hpx *h, int64_t n, const double *theta, const double *phi, uint64_t *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = orc_ang2pix(h, theta[i], phi[i]);
}

void orc_pix2ang_many(const orc_hpx *h, int64_t n, const uint64_t *pix, double *theta, double *phi) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) orc_pix2ang(h, pix[i], &theta[i], &phi[i]);
}

void orc_pix2ring_many(const orc_hpx *h, int64_t n, const uint64_t *pix, uint64_t *out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) out[i] = orc_pix2ring(h, pix[i]);
}

void orc_ring_info_many(const orc_hpx *h, int64_t n, const uint64_t *ring, uint64_t *startpix, uint64_t *npr) {
    for (int64_t i = 0; i < n; ++i) {
        int sh;
        orc_ring_info(h, ring[i], &startpix[i], &npr[i], &sh);
    }
}

/* HCGrid/gridding.cu:77-82: theta = HALFPI - DEG2RAD*lat, phi = DEG2RAD*lon */
void orc_sample_keys(const orc_hpx *h, int64_t n, const double *lon, const double *lat, uint64_t *keys) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        double theta = ORC_HALFPI - ORC_DEG2RAD * lat[i];
        double phi = ORC_DEG2RAD * lon[i];
        keys[i] = orc_ang2pix(h, theta, phi);
    }
}

/* Key sensitivity to the math library.  orc_ang2pix's only library calls
 * are cos(theta) and, within 0.01 rad of a pole, sin(theta)
 * (HCGrid/healpix.cu:208-224); everything else is IEEE-exact arithmetic.
 * For each sample this re-evaluates the key with every double within
 * +-ulps of this libm's cos (and sin) and flags the sample when any of
 * them gives another pixel.  A sample that is not flagged has the same key
 * under every cos/sin whose results lie within `ulps` of these -- the
 * reference's glibc or CUDA libm, or the device's -- so an unflagged
 * workload's keys are the reference's whatever libm computed them.
 * Returns the number of flagged samples; flags (may be NULL) gets 0/1. */
static double orc_step_ulps(double x, int k) {
    while (k > 0) { x = nextafter(x, INFINITY); --k; }
    while (k < 0) { x = nextafter(x, -INFINITY); ++k; }
    return x;
}

int64_t orc_sample_keys_libm_sensitive(const orc_hpx *h, int64_t n, const double *lon, const double *lat,
                                       int ulps, uint8_t *flags) {
    int64_t count = 0;
#pragma omp parallel for schedule(static) reduction(+ : count)
    for (int64_t i = 0; i < n; ++i) {
        double theta = ORC_HALFPI - ORC_DEG2RAD * lat[i];
        double phi = ORC_DEG2RAD * lon[i];
        int flag = 0;
        if (theta >= 0 && theta <= ORC_PI) {
            int have_sin = (theta < 0.01) || (theta > ORC_PI - 0.01);
            double z0 = cos(theta), s0 = have_sin ? sin(theta) : 0.;
            uint64_t k0 = orc_loc2pix(h, z0, phi, s0, have_sin);
            for (int a = -ulps; a <= ulps && !flag; ++a) {
                double z = orc_step_ulps(z0, a);
                if (have_sin) {
                    for (int b = -ulps; b <= ulps && !flag; ++b)
                        flag = orc_loc2pix(h, z, phi, orc_step_ulps(s0, b), 1) != k0;
                } else {
                    flag = orc_loc2pix(h, z, phi, 0., 0) != k0;
                }
            }
        }
        if (flags) flags[i] = (uint8_t)flag;
        count += flag;
    }
    return count;
}

/* ------------------------------------------------------------------ */
/* pre-sort and the ring lookup table: HCGrid/gridding.cu:70-147       */
/* ------------------------------------------------------------------ */

/* HCGrid/gridding.cu:260-275: largest i in [left, right) with v[i] < key,
 * else `right`.  Unsigned 32-bit arithmetic as in the reference. */
uint32_t orc_search_last_less(const uint64_t *values, uint32_t left, uint32_t right, uint64_t key) {
    if (right <= left) return right;
    uint32_t low = left, mid, up = right - 1;
    while (low < up) {
        mid = low + ((up - low + 1) >> 1);
        if (values[mid] < key) low = mid;
        else up = mid - 1;
    }
    if (values[low] < key) return low;
    return right;
}

/* Stable LSD radix sort of (key, index) records: the (key, original index)
 * order of HCGrid/gridding.cu:86-92 with sort_param 2 (parallel_stable_sort). */
void orc_presort(const orc_hpx *h, uint32_t n, const double *lon, const double *lat,
                 uint64_t *keys_sorted, uint32_t *perm) {
    uint64_t *ka = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint64_t *kb = (uint64_t *)malloc(sizeof(uint64_t) * (n ? n : 1));
    uint32_t *va = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *vb = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    orc_sample_keys(h, n, lon, lat, ka);
    for (uint32_t i = 0; i < n; ++i) va[i] = i;
    for (int pass = 0; pass < 8; ++pass) {
        int shift = 8 * pass;
        uint64_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (uint32_t i = 0; i < n; ++i) cnt[((ka[i] >> shift) & 0xFF) + 1]++;
        int trivial = 0;
        for (int d = 0; d < 256; ++d) if (cnt[d + 1] == n) trivial = 1;
        if (trivial) continue;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; ++i) {
            uint64_t pos = cnt[(ka[i] >> shift) & 0xFF]++;
            kb[pos] = ka[i];
            vb[pos] = va[i];
        }
        uint64_t *tk = ka; ka = kb; kb = tk;
        uint32_t *tv = va; va = vb; vb = tv;
    }
    for (uint32_t i = 0; i < n; ++i) { keys_sorted[i] = ka[i]; perm[i] = va[i]; }
    keys_sorted[n] = h->npix; /* HCGrid/gridding.cu:100 */
    free(ka); free(kb); free(va); free(vb);
}

/* HCGrid/gridding.cu:101-114 */
void orc_gather(uint32_t n, const uint32_t *perm, const double *src, double *dst) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) dst[i] = src[perm[i]];
}

/* HCGrid/gridding.cu:116-136 */
uint32_t orc_ring_lut(const orc_hpx *h, uint32_t n, const uint64_t *keys, uint64_t *firstring,
                      uint32_t *lut, uint32_t lut_cap) {
    if (n == 0) return 0;
    uint64_t first_ring = orc_pix2ring(h, keys[0]);
    uint32_t temp_count = (uint32_t)(1 + orc_pix2ring(h, keys[n - 1]) - first_ring);
    if ((uint64_t)temp_count + 1 > lut_cap) return 0;
    *firstring = first_ring;
    lut[0] = 0;
    uint32_t ring_idx = 0;
    for (uint64_t cnt_ring = 1; cnt_ring < temp_count; ++cnt_ring) {
        uint64_t startpix, npr;
        int shifted;
        orc_ring_info(h, cnt_ring + first_ring, &startpix, &npr, &shifted);
        uint32_t cnt_ring_idx = orc_search_last_less(keys, ring_idx, n, startpix);
        if (cnt_ring_idx == n) cnt_ring_idx = ring_idx - 1;
        ring_idx = cnt_ring_idx + 1;
        lut[cnt_ring] = ring_idx;
    }
    lut[temp_count] = n;
    return temp_count;
}

/* ------------------------------------------------------------------ */
/* gridding: HCGrid/gridding.cu:230-403                                */
/* ------------------------------------------------------------------ */

/* HCGrid/gridding.cu:230-235 */
static double orc_sinc(double x) {
    if (fabs(x) < 1.e-10) return 1.;
    return sin(x) / x;
}

/* HCGrid/gridding.cu:238-257.  TAPERED_SINC keeps the reference's
 * exp(pow(-(arg/p1), 2)) = exp(+(arg/p1)^2) as written. */
double orc_kernel_weight(const orc_grid *g, double distance, double bearing) {
    const double *p = g->params;
    if (g->kernel_type == 1) return exp(-distance * distance * p[0]);
    if (g->kernel_type == 2) {
        double ellarg = (pow(p[0], 2.0) * pow(sin(bearing - p[2]), 2.0) +
                         pow(p[1], 2.0) * pow(cos(bearing - p[2]), 2.0));
        double Earg = pow(distance / p[0] / p[1], 2.0) / 2. * ellarg;
        return exp(-Earg);
    }
    if (g->kernel_type == 3) {
        double arg = ORC_PI * distance / p[0];
        return orc_sinc(arg / p[2]) * exp(pow(-(arg / p[1]), 2.0));
    }
    return 0.;
}

/* Per-group window state, HCGrid/gridding.cu:307-323 */
typedef struct {
    uint64_t uring, dring;
    double l1a, b1a, disc_theta;
    int polar;
} orc_window;

static void orc_window_of(const orc_hpx *h, const orc_grid *g, uint64_t firstring, uint32_t usedrings,
                          double xa, double ya, orc_window *w) {
    w->l1a = xa * ORC_DEG2RAD;
    w->b1a = ya * ORC_DEG2RAD;
    double disc_theta = ORC_HALFPI - w->b1a;
    double disc_phi = w->l1a;
    double utheta = disc_theta - g->disc_size;
    if (utheta * ORC_RAD2DEG < 0) utheta = 0;
    uint64_t upix = orc_ang2pix(h, utheta, disc_phi);
    uint64_t uring = orc_pix2ring(h, upix);
    if (uring < firstring) uring = firstring;
    utheta = disc_theta + g->disc_size;
    upix = orc_ang2pix(h, utheta, disc_phi);
    uint64_t dring = orc_pix2ring(h, upix);
    if (dring >= firstring + usedrings) dring = firstring + usedrings - 1;
    w->uring = uring;
    w->dring = dring;
    w->disc_theta = disc_theta;
    /* HCGrid/gridding.cu:341, 352 */
    w->polar = (disc_theta * ORC_RAD2DEG <= ORC_NORTH_B || disc_theta * ORC_RAD2DEG >= ORC_SOUTH_B);
}

/* lpix / rpix of ring `ring` for a window, HCGrid/gridding.cu:329-360.
 * Returns 0 where the reference would `continue` without advancing the ring
 * (an endless loop there); callers skip the ring. */
static int orc_ring_window(const orc_hpx *h, const orc_grid *g, const orc_window *w, uint64_t ring,
                           uint64_t *lpix, uint64_t *rpix) {
    uint64_t startpix, npr;
    int shifted;
    orc_ring_info(h, ring, &startpix, &npr, &shifted);
    double utheta, uphi;
    orc_pix2ang(h, startpix, &utheta, &uphi);
    double disc_phi = w->l1a;
    uphi = disc_phi - g->disc_size;
    uint64_t lp = orc_ang2pix(h, utheta, uphi);
    if (w->polar) lp = startpix;
    if (!(lp >= startpix && lp < startpix + npr)) return 0;
    uphi = disc_phi + g->disc_size;
    uint64_t rp = orc_ang2pix(h, utheta, uphi);
    if (w->polar) rp = startpix + npr - 1;
    if (!(rp >= startpix && rp < startpix + npr)) return 0;
    *lpix = lp;
    *rpix = rp;
    return 1;
}

int64_t orc_grid_compat(const orc_hpx *h, const orc_grid *g, uint32_t n,
                        const double *lon, const double *lat, const double *data,
                        const double *weights, const uint64_t *keys,
                        uint64_t firstring, uint32_t usedrings, const uint32_t *lut,
                        const double *xwcs, const double *ywcs,
                        double *datacube, double *weightscube, int nthreads) {
    return orc_grid_compat_rows(h, g, n, lon, lat, data, weights, keys, firstring, usedrings, lut,
                                xwcs, ywcs, datacube, weightscube, 0, g->side, nthreads);
}

int64_t orc_grid_compat_rows(const orc_hpx *h, const orc_grid *g, uint32_t n,
                             const double *lon, const double *lat, const double *data,
                             const double *weights, const uint64_t *keys,
                             uint64_t firstring, uint32_t usedrings, const uint32_t *lut,
                             const double *xwcs, const double *ywcs,
                             double *datacube, double *weightscube,
                             uint32_t row_begin, uint32_t row_end, int nthreads) {
    const uint32_t side = g->side;
    const uint32_t factor = g->factor ? g->factor : 1;
    const int64_t groups_per_row = (side + factor - 1) / factor;
    if (row_end > side) row_end = side;
    const int64_t g0 = groups_per_row * (int64_t)row_begin;
    const int64_t ngroups = groups_per_row * (int64_t)row_end;
    int64_t hangs = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (n == 0 || usedrings == 0) return 0;

#pragma omp parallel for schedule(dynamic, 8) reduction(+ : hangs)
    for (int64_t gi = g0; gi < ngroups; ++gi) {
        /* HCGrid/gridding.cu:287-304: a group is `factor` consecutive cells of a row */
        uint32_t row = (uint32_t)(gi / groups_per_row);
        uint32_t left = (uint32_t)(gi % groups_per_row) * factor;
        uint32_t right = left + factor - 1;
        if (right >= side) right = side - 1;
        uint64_t base = (uint64_t)row * side;
        double tw[64], td[64], l1[64], b1[64];
        uint32_t m = right - left + 1;
        for (uint32_t k = 0; k < m; ++k) {
            uint64_t c = base + left + k;
            tw[k] = weightscube[c];
            td[k] = datacube[c];
            l1[k] = xwcs[c] * ORC_DEG2RAD;
            b1[k] = ywcs[c] * ORC_DEG2RAD;
        }
        orc_window w;
        orc_window_of(h, g, firstring, usedrings, xwcs[base + left], ywcs[base + left], &w);
        uint64_t uring = w.uring;
        /* HCGrid/gridding.cu:326 */
        uint32_t start_int = lut[uring - firstring];
        while (uring <= w.dring) {
            uint32_t end_int = lut[uring - firstring + 1];
            uint64_t lpix, rpix;
            if (!orc_ring_window(h, g, &w, uring, &lpix, &rpix)) {
                hangs++;
                start_int = end_int;
                ++uring;
                continue;
            }
            /* HCGrid/gridding.cu:363-367: start_int-1 wraps to 0xFFFFFFFF on
             * the first used ring, which then contributes nothing. */
            uint32_t upix_idx = orc_search_last_less(keys, start_int - 1u, end_int, lpix);
            ++upix_idx;
            if (upix_idx > end_int) upix_idx = n;
            while (upix_idx < n) {
                double l2 = lon[upix_idx] * ORC_DEG2RAD;
                double b2 = lat[upix_idx] * ORC_DEG2RAD;
                /* HCGrid/gridding.cu:373: the kernel recomputes the sample's pixel */
                uint64_t spix = orc_ang2pix(h, ORC_HALFPI - b2, l2);
                if (spix > rpix) break;
                double in_w = weights ? weights[upix_idx] : 1.0;
                double in_d = data[upix_idx];
                for (uint32_t k = 0; k < m; ++k) {
                    double sdist = orc_true_angular_distance(l1[k], b1[k], l2, b2) * ORC_RAD2DEG;
                    double sbear = 0.;
                    if (g->bearing_needed) sbear = orc_great_circle_bearing(l1[k], b1[k], l2, b2);
                    if (sdist < g->sphere_radius) {
                        double sweight = orc_kernel_weight(g, sdist, sbear);
                        double tweight = in_w * sweight;
                        td[k] += in_d * tweight;
                        tw[k] += tweight;
                    }
                }
                ++upix_idx;
            }
            start_int = end_int;
            ++uring;
        }
        for (uint32_t k = 0; k < m; ++k) {
            uint64_t c = base + left + k;
            datacube[c] = td[k];
            weightscube[c] = tw[k];
        }
    }
    return hangs;
}

/* Multi-channel form: the selection and FP64 weights of orc_grid_compat for
 * every cell of rows [row_begin, row_end), applied to nchan channels of
 * channel-major sorted values (values[ch * n + i], the layout of HCGrid+'s
 * per-channel arrays, HCGrid+/gridding-omp.cu:146-157).  Writes raw sums:
 * out[ch * ncells_rows + k] (k = cell - row_begin * side) and wsum[k].  This
 * is what the reference's channel-by-channel pipeline computes, one channel
 * at a time (HCGrid+/gridding-omp.cu:616-626). */
int64_t orc_grid_compat_multi(const orc_hpx *h, const orc_grid *g, uint32_t n,
                              const double *lon, const double *lat, uint32_t nchan, const float *values,
                              const double *weights, const uint64_t *keys,
                              uint64_t firstring, uint32_t usedrings, const uint32_t *lut,
                              const double *xwcs, const double *ywcs,
                              uint32_t row_begin, uint32_t row_end, double *out, double *wsum,
                              int nthreads) {
    const uint32_t side = g->side;
    if (row_end > side) row_end = side;
    const uint64_t ncr = (uint64_t)(row_end - row_begin) * side;
    int64_t hangs = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (n == 0 || usedrings == 0) return 0;
#pragma omp parallel reduction(+ : hangs)
    {
        uint32_t cap = 4096;
        uint32_t *sel = (uint32_t *)malloc(sizeof(uint32_t) * cap);
        double *wts = (double *)malloc(sizeof(double) * cap);
        double *acc = (double *)malloc(sizeof(double) * (nchan ? nchan : 1));
#pragma omp for schedule(dynamic, 4)
        for (int64_t k = 0; k < (int64_t)ncr; ++k) {
            uint64_t c = (uint64_t)row_begin * side + (uint64_t)k;
            uint32_t row = (uint32_t)(c / side), col = (uint32_t)(c % side);
            uint32_t factor = g->factor ? g->factor : 1;
            uint64_t anchor = (uint64_t)row * side + (col / factor) * factor;
            orc_window w;
            orc_window_of(h, g, firstring, usedrings, xwcs[anchor], ywcs[anchor], &w);
            double l1 = xwcs[c] * ORC_DEG2RAD, b1 = ywcs[c] * ORC_DEG2RAD;
            uint32_t m = 0;
            double sw = 0.0;
            uint64_t uring = w.uring;
            uint32_t start_int = lut[uring - firstring];
            while (uring <= w.dring) {
                uint32_t end_int = lut[uring - firstring + 1];
                uint64_t lpix, rpix;
                if (!orc_ring_window(h, g, &w, uring, &lpix, &rpix)) {
                    hangs++;
                    start_int = end_int;
                    ++uring;
                    continue;
                }
                uint32_t i = orc_search_last_less(keys, start_int - 1u, end_int, lpix);
                ++i;
                if (i > end_int) i = n;
                for (; i < n; ++i) {
                    double l2 = lon[i] * ORC_DEG2RAD, b2 = lat[i] * ORC_DEG2RAD;
                    if (orc_ang2pix(h, ORC_HALFPI - b2, l2) > rpix) break;
                    double sdist = orc_true_angular_distance(l1, b1, l2, b2) * ORC_RAD2DEG;
                    if (!(sdist < g->sphere_radius)) continue;
                    double sbear = g->bearing_needed ? orc_great_circle_bearing(l1, b1, l2, b2) : 0.;
                    double tw = (weights ? weights[i] : 1.0) * orc_kernel_weight(g, sdist, sbear);
                    if (m == cap) {
                        cap *= 2;
                        sel = (uint32_t *)realloc(sel, sizeof(uint32_t) * cap);
                        wts = (double *)realloc(wts, sizeof(double) * cap);
                    }
                    sel[m] = i;
                    wts[m] = tw;
                    ++m;
                    sw += tw;
                }
                start_int = end_int;
                ++uring;
            }
            for (uint32_t ch = 0; ch < nchan; ++ch) {
                const float *vch = values + (uint64_t)ch * n;
                double a = 0.0;
                for (uint32_t q = 0; q < m; ++q) a += (double)vch[sel[q]] * wts[q];
                acc[ch] = a;
            }
            for (uint32_t ch = 0; ch < nchan; ++ch) out[(uint64_t)ch * ncr + (uint64_t)k] = acc[ch];
            wsum[k] = sw;
        }
        free(sel);
        free(wts);
        free(acc);
    }
    return hangs;
}

void orc_grid_bruteforce(const orc_hpx *h, const orc_grid *g, uint32_t n,
                         const double *lon, const double *lat, const double *data,
                         const double *weights, const uint64_t *keys,
                         uint64_t firstring, uint32_t usedrings,
                         const double *xwcs, const double *ywcs,
                         double *datacube, double *weightscube, int nthreads) {
    const uint32_t side = g->side;
    const uint32_t factor = g->factor ? g->factor : 1;
    const int64_t ncells = (int64_t)side * side;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (n == 0 || usedrings == 0) return;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t c = 0; c < ncells; ++c) {
        uint32_t row = (uint32_t)(c / side), col = (uint32_t)(c % side);
        uint64_t anchor = (uint64_t)row * side + (col / factor) * factor;
        orc_window w;
        orc_window_of(h, g, firstring, usedrings, xwcs[anchor], ywcs[anchor], &w);
        double l1 = xwcs[c] * ORC_DEG2RAD, b1 = ywcs[c] * ORC_DEG2RAD;
        double td = datacube[c], tw = weightscube[c];
        uint64_t cached_ring = (uint64_t)-1, lp = 0, rp = 0;
        int ok = 0;
        for (uint32_t s = 0; s < n; ++s) {
            uint64_t r = orc_pix2ring(h, keys[s]);
            if (r < w.uring || r > w.dring) continue;
            if (r == firstring) continue; /* start_int - 1 underflow */
            if (r != cached_ring) {
                cached_ring = r;
                ok = orc_ring_window(h, g, &w, r, &lp, &rp);
            }
            if (!ok) continue;
            if (keys[s] < lp || keys[s] > rp) continue;
            double l2 = lon[s] * ORC_DEG2RAD, b2 = lat[s] * ORC_DEG2RAD;
            double sdist = orc_true_angular_distance(l1, b1, l2, b2) * ORC_RAD2DEG;
            if (!(sdist < g->sphere_radius)) continue;
            double sbear = g->bearing_needed ? orc_great_circle_bearing(l1, b1, l2, b2) : 0.;
            double tweight = (weights ? weights[s] : 1.0) * orc_kernel_weight(g, sdist, sbear);
            td += data[s] * tweight;
            tw += tweight;
        }
        datacube[c] = td;
        weightscube[c] = tw;
    }
}

void orc_grid_disc_textbook(uint32_t ncells, uint32_t n, const double *lon, const double *lat,
                            const double *data, const double *xwcs, const double *ywcs,
                            double sphere_radius_deg, double p0,
                            double *datacube, double *weightscube, int nthreads) {
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    const long double d2r = 3.14159265358979323846264338327950288L / 180.0L;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t c = 0; c < (int64_t)ncells; ++c) {
        long double a1 = xwcs[c] * d2r, d1 = ywcs[c] * d2r;
        long double px = cosl(d1) * cosl(a1), py = cosl(d1) * sinl(a1), pz = sinl(d1);
        long double sv = 0.0L, sw = 0.0L;
        for (uint32_t s = 0; s < n; ++s) {
            long double a2 = lon[s] * d2r, d2 = lat[s] * d2r;
            long double qx = cosl(d2) * cosl(a2), qy = cosl(d2) * sinl(a2), qz = sinl(d2);
            long double cx = py * qz - pz * qy, cy = pz * qx - px * qz, cz = px * qy - py * qx;
            long double dist = atan2l(sqrtl(cx * cx + cy * cy + cz * cz), px * qx + py * qy + pz * qz) / d2r;
            if (dist < (long double)sphere_radius_deg) {
                long double wgt = expl(-dist * dist * (long double)p0);
                sv += (long double)data[s] * wgt;
                sw += wgt;
            }
        }
        datacube[c] = (double)sv;
        weightscube[c] = (double)sw;
    }
}

/* HCGrid/gmap.cpp:305-312 (NaN where the weight sum is 0, by IEEE 0/0) */
void orc_normalize(int64_t ncells, const double *datacube, const double *weightscube, double *map) {
    for (int64_t i = 0; i < ncells; ++i) map[i] = datacube[i] / weightscube[i];
}

/* SIN deprojection + zenithal rotation in long double (Calabretta & Greisen
 * 2002, eqs. 2, 54-55 with PV2_1 = PV2_2 = 0, and the native-to-celestial
 * rotation, eq. 2 of the same paper, with phi_p = LONPOLE = 180 deg). */
void orc_wcs_zenithal(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                      double cdelt1, double cdelt2, double crval1, double crval2, double lonpole,
                      double *xwcs, double *ywcs) {
    const long double pi = 3.14159265358979323846264338327950288L;
    const long double d2r = pi / 180.0L;
    long double a0 = crval1 * d2r, dl0 = crval2 * d2r, php = lonpole * d2r;
    for (uint32_t yy = 0; yy < naxis2; ++yy) {
        for (uint32_t xx = 0; xx < naxis1; ++xx) {
            long double x = (long double)cdelt1 * ((long double)(xx + 1) - crpix1) * d2r;
            long double y = (long double)cdelt2 * ((long double)(yy + 1) - crpix2) * d2r;
            long double r = sqrtl(x * x + y * y);
            long double phi = atan2l(x, -y);        /* native longitude */
            long double theta;                      /* native latitude, inverse of R(theta) */
            switch (proj) {
                case 1: theta = atan2l(1.0L, r); break;                     /* TAN: R = cot(theta) */
                case 2: theta = r <= pi ? pi / 2 - r : NAN; break;          /* ARC: R = pi/2 - theta */
                case 3: theta = r <= 2.0L ? pi / 2 - 2 * asinl(r / 2) : NAN; break;  /* ZEA */
                case 4: theta = pi / 2 - 2 * atanl(r / 2); break;           /* STG */
                default: theta = acosl(r); break;                           /* SIN: R = cos(theta) */
            }
            long double dphi = phi - php;
            long double sd = sinl(theta) * sinl(dl0) + cosl(theta) * cosl(dl0) * cosl(dphi);
            long double dec = asinl(sd);
            long double ra = a0 + atan2l(-cosl(theta) * sinl(dphi),
                                         sinl(theta) * cosl(dl0) - cosl(theta) * sinl(dl0) * cosl(dphi));
            long double ra_deg = ra / d2r;
            uint64_t c = (uint64_t)yy * naxis1 + xx;
            if (isnan(theta)) {
                xwcs[c] = NAN;
                ywcs[c] = NAN;
                continue;
            }
            while (ra_deg < 0) ra_deg += 360.0L;
            while (ra_deg >= 360.0L) ra_deg -= 360.0L;
            xwcs[c] = (double)ra_deg;
            ywcs[c] = (double)(dec / d2r);
        }
    }
}

void orc_wcs_sin(uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                 double cdelt1, double cdelt2, double crval1, double crval2,
                 double *xwcs, double *ywcs) {
    orc_wcs_zenithal(0, naxis1, naxis2, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2, 180.0, xwcs, ywcs);
}

/* The general celestial WCS of a 2-D target, what the reference's set_WCS
 * gets from wcslib's wcsp2s (HCGrid/gmap.cpp:183-250), restated from
 * Calabretta & Greisen 2002 (Paper II) in long double.
 *   1. Intermediate world coordinates (degrees): (x, y) = m (p - crpix) with
 *      p = (xx + 1, yy + 1) and m row-major (CD, or CDELT_i PC_ij).
 *   2. Native (phi, theta) by the projection's inverse (X, Y = x, y in
 *      radians): 0..4 the zenithal ones of orc_wcs_zenithal, SIN with the
 *      slant (xi, eta) = (pv1, pv2): sin(theta) is the root nearest +1 of
 *      a s^2 + 2 b s + c = 0 (a = xi^2 + eta^2 + 1, b = xi (X - xi) +
 *      eta (Y - eta), c = (X - xi)^2 + (Y - eta)^2 - 1) and phi =
 *      atan2(X - xi (1 - s), -(Y - eta (1 - s))) (sec. 5.1.5);
 *      5 CAR phi = X, theta = Y; 6 MER theta = 2 atan(e^Y) - pi/2; 7 CEA
 *      theta = asin(lambda Y), lambda = pv1; 8 SFL theta = Y, phi =
 *      X / cos(theta); 10 AIT Z = sqrt(1 - (X/4)^2 - (Y/2)^2), phi =
 *      2 atan2(X Z / 2, 2 Z^2 - 1), theta = asin(Y Z) (secs. 5.2.1-5.3.4).
 *   3. Native -> celestial (eq. 2) about the celestial coordinates
 *      (alpha_p, delta_p) of the native pole, phi_p = LONPOLE.  Zenithal:
 *      (alpha_p, delta_p) = CRVAL.  theta_0 = 0 (5..8, 10): delta_p =
 *      atan2(0, cos phi_p) +- acos(sin delta_0 / |cos phi_p|) (eq. 8),
 *      the root in [-90, 90] nearest latpole, and alpha_p = alpha_0 -
 *      atan2(sin phi_p, -sin delta_p cos phi_p) (eq. 2 at the reference
 *      point).  lonpole NaN: 0 when delta_0 >= theta_0, else 180.
 *   9 GLS (the AIPS convention, Greisen's AIPS memo 27): dec = crval2 + y,
 *      ra = crval1 + x / cos(dec), no rotation.
 * NaN outside the projection's domain.  Returns 0, or -1 when no delta_p
 * satisfies eq. 8 (a LONPOLE the reference point cannot have). */
int orc_wcs_celestial(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                      const double *m, double crval1, double crval2, double lonpole, double latpole,
                      double pv1, double pv2, double *xwcs, double *ywcs) {
    const long double pi = 3.14159265358979323846264338327950288L;
    const long double d2r = pi / 180.0L;
    const int zenithal = proj <= 4, gls = proj == 9;
    const long double a0 = crval1 * d2r, dl0 = crval2 * d2r;
    long double php = isnan(lonpole) ? ((crval2 >= (zenithal ? 90.0 : 0.0)) ? 0.0L : pi) : lonpole * d2r;
    long double ap = a0, dp = dl0;
    if (!zenithal && !gls) {
        const long double den = fabsl(cosl(php));
        if (den < 1e-15L) {
            if (fabsl(sinl(dl0)) > 1e-15L) return -1;
            dp = latpole * d2r;
        } else {
            long double u = sinl(dl0) / den;
            if (fabsl(u) > 1.0L + 1e-12L) return -1;
            if (u > 1.0L) u = 1.0L;
            if (u < -1.0L) u = -1.0L;
            const long double base = atan2l(0.0L, cosl(php)), a = acosl(u);
            long double best = NAN;
            for (int k = 0; k < 2; ++k) {
                long double c = k ? base - a : base + a;
                if (c > pi) c -= 2 * pi;
                if (c < -pi) c += 2 * pi;
                if (fabsl(c) > pi / 2 + 1e-12L) continue;
                if (c > pi / 2) c = pi / 2;
                if (c < -pi / 2) c = -pi / 2;
                if (isnan(best) || fabsl(c - latpole * d2r) < fabsl(best - latpole * d2r)) best = c;
            }
            if (isnan(best)) return -1;
            dp = best;
        }
        ap = a0 - atan2l(sinl(php), -sinl(dp) * cosl(php));
    }
    for (uint32_t yy = 0; yy < naxis2; ++yy) {
        for (uint32_t xx = 0; xx < naxis1; ++xx) {
            const long double d1 = (long double)(xx + 1) - crpix1, d2 = (long double)(yy + 1) - crpix2;
            const long double x = ((long double)m[0] * d1 + (long double)m[1] * d2) * d2r;
            const long double y = ((long double)m[2] * d1 + (long double)m[3] * d2) * d2r;
            const uint64_t c = (uint64_t)yy * naxis1 + xx;
            long double phi = 0, theta = NAN;
            if (gls) {
                const long double dec = dl0 + y;
                long double ra = NAN;
                if (fabsl(dec) <= pi / 2) {
                    const long double dl = cosl(dec) > 0 ? x / cosl(dec) : 0.0L;
                    if (fabsl(dl) <= pi) ra = (a0 + dl) / d2r;
                }
                if (isnan(ra)) {
                    xwcs[c] = NAN;
                    ywcs[c] = NAN;
                    continue;
                }
                while (ra < 0) ra += 360.0L;
                while (ra >= 360.0L) ra -= 360.0L;
                xwcs[c] = (double)ra;
                ywcs[c] = (double)(dec / d2r);
                continue;
            }
            switch (proj) {
                case 0: {
                    const long double xi = pv1, eta = pv2;
                    const long double qa = xi * xi + eta * eta + 1.0L, qb = xi * (x - xi) + eta * (y - eta);
                    const long double qc = (x - xi) * (x - xi) + (y - eta) * (y - eta) - 1.0L;
                    const long double disc = qb * qb - qa * qc;
                    if (disc < 0) break;
                    long double s = (-qb + sqrtl(disc)) / qa;
                    if (s > 1.0L + 1e-12L) s = (-qb - sqrtl(disc)) / qa;
                    if (fabsl(s) > 1.0L + 1e-12L) break;
                    if (s > 1.0L) s = 1.0L;
                    theta = asinl(s);
                    phi = atan2l(x - xi * (1.0L - s), -(y - eta * (1.0L - s)));
                    break;
                }
                case 1: case 2: case 3: case 4: {
                    const long double r = sqrtl(x * x + y * y);
                    phi = atan2l(x, -y);
                    if (proj == 1) theta = atan2l(1.0L, r);
                    else if (proj == 2) theta = r <= pi ? pi / 2 - r : NAN;
                    else if (proj == 3) theta = r <= 2.0L ? pi / 2 - 2 * asinl(r / 2) : NAN;
                    else theta = pi / 2 - 2 * atanl(r / 2);
                    break;
                }
                case 5:
                    phi = x;
                    theta = fabsl(y) <= pi / 2 ? y : NAN;
                    break;
                case 6:
                    phi = x;
                    theta = 2 * atanl(expl(y)) - pi / 2;
                    break;
                case 7: {
                    const long double s = (long double)pv1 * y;
                    phi = x;
                    theta = fabsl(s) <= 1.0L ? asinl(s) : NAN;
                    break;
                }
                case 8:
                    if (fabsl(y) <= pi / 2) {
                        phi = cosl(y) > 0 ? x / cosl(y) : 0.0L;
                        theta = fabsl(phi) <= pi ? y : NAN;
                    }
                    break;
                case 10: {
                    const long double z2 = 1.0L - (x / 4) * (x / 4) - (y / 2) * (y / 2);
                    if (z2 < 0.5L) break;
                    const long double z = sqrtl(z2);
                    phi = 2 * atan2l(x * z / 2, 2 * z2 - 1);
                    theta = asinl(y * z);
                    break;
                }
                default:
                    break;
            }
            if (isnan(theta)) {
                xwcs[c] = NAN;
                ywcs[c] = NAN;
                continue;
            }
            const long double dphi = phi - php;
            const long double sd = sinl(theta) * sinl(dp) + cosl(theta) * cosl(dp) * cosl(dphi);
            const long double dec = asinl(sd > 1 ? 1 : (sd < -1 ? -1 : sd));
            long double ra = (ap + atan2l(-cosl(theta) * sinl(dphi),
                                          sinl(theta) * cosl(dp) - cosl(theta) * sinl(dp) * cosl(dphi))) / d2r;
            while (ra < 0) ra += 360.0L;
            while (ra >= 360.0L) ra -= 360.0L;
            xwcs[c] = (double)ra;
            ywcs[c] = (double)(dec / d2r);
        }
    }
    return 0;
}

/*
 * hcgrid_host.h -- the reference's C++ host API for the gridding path, as a
 * drop-in over the gfx950 library.
 *
 * Same names, argument meaning and module-global state as HCGrid's
 * gmap.h / gridding.h / healpix.h (HCGrid/gmap.h:47-139,
 * HCGrid/gridding.h:44-102, HCGrid/healpix.h:14-53), so a caller of the
 * reference (its main, HCGrid/HCGrid.cpp:38-197) links against
 * libhcgrid_host.so unchanged.  Underneath, every device operation goes
 * through the C ABI of hcgrid_amd.h; this layer never calls HIP itself.
 *
 * Differences a caller can observe, all deliberate:
 *  - every sort_param (1..4) yields the stable (key, input index) order;
 *  - keys and sorted arrays come back from the device (the host does no
 *    ang2pix loop and no sort);
 *  - errors follow the reference (message on stderr, exit(1)) unless
 *    hcg_host_set_exit_on_error(0) is called, after which failing functions
 *    return early and hcg_host_error() reports the failure;
 *  - write_ordered_map writes the sorted coordinates as the 2 x N image the
 *    input uses (the reference calls fits_write_col on an image HDU);
 *  - set_WCS maps the celestial targets its own formulas cover (wcs.cpp:
 *    SIN with slant and NCP, TAN, ARC, ZEA, STG, CAR, MER, CEA, SFL, GLS,
 *    AIT; CDELT, PC, CD or CROTA2; LONPOLE and LATPOLE) where the reference
 *    takes whatever wcslib parses; other headers are an error.
 */
#ifndef HCGRID_HOST_H
#define HCGRID_HOST_H

#include <stdint.h>

/* HCGrid/constants.h:57-73 */
#define MAX_Y (1073741824)
#define BLOCK_INDIRECT_SORT 1
#define PARALLEL_STABLE_SORT 2
#define THRUST 3
#define STL_SORT 4
#define GAUSS1D 1
#define GAUSS2D 2
#define TAPERED_SINC 3
#define NESTED 1
#define RING 2

/* HCGrid/healpix.h:14-21 */
struct Healpix {
    uint64_t _nside, _order, _nrings, usedrings, firstring;
    uint64_t _max_npix_per_ring;
    uint64_t _npface, _ncap, _npix;
    double _fact1, _fact2, _resolution, _omega;
    uint32_t _scheme;
    bool _params_dirty;
};
extern Healpix h_Healpix;

/* HCGrid/gmap.h:47-61 */
class GMaps {
 public:
    uint32_t block_warp_num;
    uint32_t factor;
    uint32_t data_shape;
    uint32_t spec_dim;
    uint32_t kernel_type;
    double disc_size;
    double sphere_radius;
    bool bearing_needed;
    uint32_t selection;  /* not in the reference: HCG_SELECT_COMPAT (0, default) or HCG_SELECT_DISC */
    uint32_t spec_dtype; /* not in the reference: how the HCGrid+ path grids its channels,
                            HCG_DTYPE_F64 (0, default: FP64 per channel, as HCGrid+) or
                            HCG_DTYPE_BF16 / HCG_DTYPE_F16 (one multi-channel pass) */
};
extern GMaps h_GMaps;

/* host state (HCGrid/gmap.h:65-107) */
extern double *h_lons, *h_lats, *h_data, *h_weights;
extern uint64_t *h_hpx_idx;
extern uint32_t *h_start_ring;
extern char *h_header;
extern uint32_t *h_zyx;
extern double *h_xwcs, *h_ywcs;
extern double *h_datacube, *h_weightscube;
extern double *h_kernel_params;
extern double last_sphere_radius, last_hpxmaxres;

/* HEALPix (HCGrid/healpix.h:26-53), host versions */
void _Healpix_init(uint64_t nside, uint32_t scheme);
void _update_params();
void h_get_ring_info_small(uint64_t ring, uint64_t &startpix, uint64_t &num_pix_in_ring, bool &shifted);
uint64_t h_pix2ring(uint64_t pix);
uint64_t h_ang2pix(double theta, double phi);
void h_pix2ang(uint64_t pix, double &theta, double &phi);

/* helpers (HCGrid/helpers.h:168-197) */
uint64_t set_optimal_nside(double target_res);
double cpuSecond();

/* set-up and I/O (HCGrid/gmap.h:121-136) */
void _prepare_grid_kernel(uint32_t kernel_type, double *kernel_params, double sphere_radius, double hpx_max_res);
void read_input_map(const char *infile);
void read_output_map(const char *infile);
int set_WCS();
void write_output_map(const char *infile);
void write_ordered_map(const char *infile, const char *outfile);

/* driver (HCGrid/gridding.h:44-102) */
void init_input_with_cpu(const int &sort_param);
void init_input_with_thrust(const int &sort_param);
void init_output();
uint32_t searchLastPosLessThan(uint64_t *values, uint32_t left, uint32_t right, uint64_t _key);
void data_alloc();
void data_h2d();
void data_d2h();
void data_free();
void solve_gridding(const char *infile, const char *tarfile, const char *outfile, const char *sortfile,
                    const int &param, const int &bDim);

/* HCGrid+'s spectral path (HCGrid+/gmap.h:126-128, HCGrid+/gridding.h): HDF5
 * input (group coords: attribute num_samples, datasets xcoords, ycoords;
 * group signal: datasets signal_0, signal_1, ...), one device pre-sort, every
 * channel gridded.  The HDF5 reader is libhcgrid_h5.so next to this library,
 * loaded on first use.  HCGrid+ reads exactly 5 channels; here spec_dim is
 * the number of consecutive signal_<c> datasets. */
void read_input_coordinate(const char *infile);  /* h_lons, h_lats, h_weights = 1, data_shape, spec_dim */
void read_input_data(const char *infile);        /* h_spec_data */
void solve_gridding(const char *infile, const char *tarfile, const char *outfile, const char *sortfile,
                    const int &param, const int &bDim, int argc, char **argv);
/* Not in the reference.  HCGrid+ keeps channels in double **h_data and
 * writes only channel 0 (HCGrid+/gmap.cpp:354-361); here: */
extern double *h_spec_data;  /* [spec_dim][data_shape] channel values, input order */
extern double *h_spec_cube;  /* [spec_dim][NAXIS1*NAXIS2] normalised maps, NaN where no weight */
void write_output_cube(const char *outfile);   /* every channel: NAXIS1 x NAXIS2 x spec_dim, BITPIX -64 */
void write_ordered_cube(const char *outfile);  /* sorted input as FITS: 1 x N x spec_dim values + N x 2 coords */

/* error-mode control, no counterpart in the reference */
extern "C" void hcg_host_set_exit_on_error(int on);
extern "C" const char *hcg_host_error(void);  /* "" when the last call succeeded */

#endif /* HCGRID_HOST_H */

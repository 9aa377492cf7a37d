// gmap.cpp -- kernel set-up, FITS input/target/output and target geometry:
// the reference's HCGrid/gmap.cpp interface (gmap.h:121-136) over the C ABI
// and the built-in FITS reader/writer.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include <string>
#include <vector>

#include "fits_io.hpp"
#include "hcgrid_amd.h"
#include "hcgrid_host.h"
#include "host_internal.hpp"
#include "wcs.hpp"

// ---- module state (HCGrid/gmap.cpp:46-61, HCGrid/healpix.cu:39) ----
double *h_lons = nullptr, *h_lats = nullptr, *h_data = nullptr, *h_weights = nullptr;
uint64_t *h_hpx_idx = nullptr;
uint32_t *h_start_ring = nullptr;
char *h_header = nullptr;
uint32_t *h_zyx = nullptr;
double *h_xwcs = nullptr, *h_ywcs = nullptr;
double *h_datacube = nullptr, *h_weightscube = nullptr;
double *h_kernel_params = nullptr;
double last_sphere_radius = -1;
double last_hpxmaxres = -1;
GMaps h_GMaps;
Healpix h_Healpix;

namespace hcghost {

namespace {
bool g_exit_on_error = true;
std::string g_error;
}  // namespace

bool fail(const std::string& msg) {
  g_error = msg;
  fprintf(stderr, "%s\n", msg.c_str());
  if (g_exit_on_error) exit(1);  // the reference's UTIL_FAIL / printfitserror (HCGrid/helpers.h:53-56)
  return false;
}

bool failed() { return !g_error.empty(); }
void clear_error() { g_error.clear(); }

hcg_healpix abi_healpix() {
  hcg_healpix hp;
  hp.nside = h_Healpix._nside;
  hp.order = h_Healpix._order;
  hp.nrings = h_Healpix._nrings;
  hp.npface = h_Healpix._npface;
  hp.ncap = h_Healpix._ncap;
  hp.npix = h_Healpix._npix;
  hp.fact1 = h_Healpix._fact1;
  hp.fact2 = h_Healpix._fact2;
  hp.omega = h_Healpix._omega;
  hp.resolution = h_Healpix._resolution;
  return hp;
}

bool check_abi(int rc, const char* where) {
  if (rc == HCG_OK) return true;
  return fail(std::string(where) + ": " + hcg_strerror(rc) + " (" + hcg_last_error_message() + ")");
}

}  // namespace hcghost

using namespace hcghost;

extern "C" void hcg_host_set_exit_on_error(int on) { hcghost::g_exit_on_error = on != 0; }
extern "C" const char* hcg_host_error(void) { return hcghost::g_error.c_str(); }

// ---- HEALPix (HCGrid/healpix.cu:42-65, 68-348), host side ----

void _update_params() {
  if (!h_Healpix._params_dirty) return;
  hcg_healpix hp;
  if (hcg_healpix_init(h_Healpix._nside, &hp) != HCG_OK) {
    fail("_update_params: nside must be a power of two");
    return;
  }
  h_Healpix._order = hp.order;
  h_Healpix._nrings = hp.nrings;
  h_Healpix._max_npix_per_ring = 4 * hp.nside;
  h_Healpix._npface = hp.npface;
  h_Healpix._ncap = hp.ncap;
  h_Healpix._npix = hp.npix;
  h_Healpix._fact2 = hp.fact2;
  h_Healpix._fact1 = hp.fact1;
  h_Healpix._omega = hp.omega;
  h_Healpix._resolution = hp.resolution;
  h_Healpix._params_dirty = false;
}

void _Healpix_init(uint64_t nside, uint32_t scheme) {
  clear_error();
  if (scheme != RING) {  // the reference checks the previous scheme (healpix.cu:44-46); only RING exists
    fail("Currently, only RING scheme is supported.");
    return;
  }
  h_Healpix._nside = nside;
  h_Healpix._scheme = scheme;
  h_Healpix._params_dirty = true;
  _update_params();
}

void h_get_ring_info_small(uint64_t ring, uint64_t& startpix, uint64_t& num_pix_in_ring, bool& shifted) {
  const hcg_healpix hp = abi_healpix();
  hcg_host_ring_info(&hp, ring, &startpix, &num_pix_in_ring);
  shifted = (ring < h_Healpix._nside || ring >= 3 * h_Healpix._nside) ? true
                                                                        : (((ring - h_Healpix._nside) & 1) == 0);
}

uint64_t h_pix2ring(uint64_t pix) {
  const hcg_healpix hp = abi_healpix();
  return hcg_host_pix2ring(&hp, pix);
}

uint64_t h_ang2pix(double theta, double phi) {
  const hcg_healpix hp = abi_healpix();
  return hcg_host_ang2pix(&hp, theta, phi);
}

void h_pix2ang(uint64_t pix, double& theta, double& phi) {
  const hcg_healpix hp = abi_healpix();
  theta = hcg_host_pix2theta(&hp, pix);
  // pixel-centre longitude (HCGrid/healpix.cu:239-285)
  const uint64_t ring = hcg_host_pix2ring(&hp, pix);
  uint64_t sp, npr;
  hcg_host_ring_info(&hp, ring, &sp, &npr);
  const uint64_t iphi = pix - sp + 1;
  const double kPi = 3.1415926535897932384626433832;
  if (ring < hp.nside || ring > 3 * hp.nside) {
    const uint64_t iring = ring < hp.nside ? ring : 4 * hp.nside - ring;
    phi = ((double)iphi - 0.5) * (kPi / 2.) / (double)iring;
  } else {
    const double fodd = (((ring + hp.nside) & 1) > 0) ? 1.0 : 0.5;
    phi = ((double)iphi - fodd) * kPi * 0.75 * hp.fact1;
  }
}

uint64_t set_optimal_nside(double target_res) { return hcg_set_optimal_nside(target_res); }

double cpuSecond() {
  struct timeval tp;
  gettimeofday(&tp, nullptr);
  return (double)tp.tv_sec + (double)tp.tv_usec * 1.e-6;
}

// ---- set-up (HCGrid/gmap.cpp:64-99) ----

void _prepare_grid_kernel(uint32_t kernel_type, double* kernel_params, double sphere_radius, double hpx_max_res) {
  clear_error();
  if (kernel_params == nullptr) {
    fail("Number Kernel Paramters Error.");
    return;
  }
  hcg_kernel k;
  hcg_healpix hp;
  if (!check_abi(hcg_prepare_grid_kernel(kernel_type, kernel_params, sphere_radius, hpx_max_res, &k, &hp),
                 "_prepare_grid_kernel"))
    return;
  // the reference transforms the caller's array in place and keeps it
  if (kernel_type == GAUSS1D) kernel_params[0] = k.params[0];
  h_GMaps.bearing_needed = k.bearing_needed != 0;
  h_GMaps.kernel_type = kernel_type;
  h_kernel_params = kernel_params;
  // HEALPix is recomputed only when radius or resolution moved (HCGrid/gmap.cpp:89)
  if (fabs(last_sphere_radius - sphere_radius) > 3e-5 || fabs(last_hpxmaxres - hpx_max_res) > 3e-5) {
    _Healpix_init(hp.nside, RING);
    last_sphere_radius = sphere_radius;
    last_hpxmaxres = hpx_max_res;
  }
  h_GMaps.sphere_radius = sphere_radius;
  const double kDeg2Rad = 3.1415926535897932384626433832 / 180.;
  h_GMaps.disc_size = (kDeg2Rad * sphere_radius + h_Healpix._resolution);
}

// ---- input (HCGrid/gmap.cpp:102-148) ----

void read_input_map(const char* infile) {
  clear_error();
  hcgio::FitsFile f;
  std::string err;
  if (!f.open(infile, err)) {
    fail("read_input_map: " + err);
    return;
  }
  const hcgio::Hdu& p = f.hdu(0);
  if (p.naxes.size() < 2 || f.num_hdus() < 2) {
    fail(std::string("read_input_map: ") + infile + " needs NAXIS >= 2 and a coordinate HDU");
    return;
  }
  const uint64_t y = (uint64_t)p.naxes[0], x = (uint64_t)p.naxes[1];
  const uint64_t z = p.naxes.size() == 3 ? (uint64_t)p.naxes[2] : 1;
  if (y * x >= 0xFFFFFFFFull) {
    fail("read_input_map: more than 2^32-2 samples");
    return;
  }
  h_GMaps.data_shape = (uint32_t)(y * x);
  h_GMaps.spec_dim = (uint32_t)z;
  free(h_data);
  free(h_lons);
  free(h_lats);
  free(h_weights);
  h_data = (double*)malloc(sizeof(double) * y * x * z);
  h_lons = (double*)malloc(sizeof(double) * y * x);
  h_lats = (double*)malloc(sizeof(double) * y * x);
  h_weights = (double*)malloc(sizeof(double) * z * y * x);
  if (!h_data || !h_lons || !h_lats || !h_weights) {
    fail("read_input_map: malloc() failed");
    return;
  }
  if (!f.read(0, 1, y * x * z, h_data, err) || !f.read(1, 1, y * x, h_lons, err) ||
      !f.read(1, y * x + 1, y * x, h_lats, err)) {
    fail("read_input_map: " + err);
    return;
  }
  for (uint64_t i = 0; i < z * y * x; ++i) h_weights[i] = 1.;
}

// ---- target header (HCGrid/gmap.cpp:151-180) ----

void read_output_map(const char* infile) {
  clear_error();
  hcgio::FitsFile f;
  std::string err;
  if (!f.open(infile, err)) {
    fail("read_output_map: " + err);
    return;
  }
  const hcgio::Hdu& p = f.hdu(0);
  if (p.naxes.size() < 2) {
    fail(std::string("read_output_map: ") + infile + " needs NAXIS >= 2");
    return;
  }
  const std::string hdr = f.header_string(0);
  free(h_header);
  h_header = (char*)malloc(hdr.size() + 1);
  memcpy(h_header, hdr.c_str(), hdr.size() + 1);
  free(h_zyx);
  h_zyx = (uint32_t*)malloc(sizeof(uint32_t) * 3);
  // reversed axes, {NAXIS3, NAXIS2, NAXIS1}; a 2-axis target gets NAXIS3 = 1
  h_zyx[0] = p.naxes.size() >= 3 ? (uint32_t)p.naxes[2] : 1;
  h_zyx[1] = (uint32_t)p.naxes[1];
  h_zyx[2] = (uint32_t)p.naxes[0];
  if (h_GMaps.factor == 0) h_GMaps.factor = 1;
  h_GMaps.block_warp_num = (h_zyx[1] - 1) / (32 * h_GMaps.factor) + 1;
}

// ---- target coordinates (HCGrid/gmap.cpp:183-250) ----

static_assert(hcgwcs::SIN == HCG_PROJ_SIN && hcgwcs::TAN == HCG_PROJ_TAN && hcgwcs::ARC == HCG_PROJ_ARC &&
                  hcgwcs::ZEA == HCG_PROJ_ZEA && hcgwcs::STG == HCG_PROJ_STG,
              "the zenithal codes are the C ABI's");

int set_WCS() {
  clear_error();
  if (!h_header || !h_zyx) {
    fail("set_WCS: read_output_map first");
    return 0;
  }
  std::vector<std::string> cards;
  for (size_t k = 0; k + 80 <= strlen(h_header); k += 80) cards.emplace_back(h_header + k, 80);
  hcgwcs::Celestial w;
  std::string err;
  if (!hcgwcs::from_header(cards, w, err)) {
    fail(err);
    return 0;
  }
  const uint32_t y = h_zyx[1], x = h_zyx[2];
  free(h_xwcs);
  free(h_ywcs);
  h_xwcs = (double*)malloc(sizeof(double) * (size_t)x * y);
  h_ywcs = (double*)malloc(sizeof(double) * (size_t)x * y);
  if (!h_xwcs || !h_ywcs) {
    fail("set_WCS: malloc() failed");
    return 0;
  }
  // a zenithal target with a diagonal CDELT: the C ABI's form (the one the
  // device entry point mirrors); anything else: the general host form
  if (hcgwcs::plain_zenithal(w)) {
    if (!check_abi(hcg_wcs_p2s((uint32_t)w.proj, x, y, w.crpix[0], w.crpix[1], w.m[0], w.m[3], w.crval[0],
                               w.crval[1], w.lonpole, h_xwcs, h_ywcs),
                   "set_WCS"))
      return 0;
  } else if (!hcgwcs::p2s(w, x, y, h_xwcs, h_ywcs, err)) {
    fail(err);
    return 0;
  }
  return 1;
}

// ---- output (HCGrid/gmap.cpp:253-323) ----

void write_output_map(const char* infile) {
  clear_error();
  if (!h_header || !h_zyx || !h_datacube || !h_weightscube) {
    fail("write_output_map: nothing to write");
    return;
  }
  hcgio::HduOut h;
  h.bitpix = -64;
  h.naxes = {(int64_t)h_zyx[1], (int64_t)h_zyx[2], (int64_t)h_zyx[0]};
  // header records from the 7th card up to END, '#' meta-comments dropped
  const size_t len = strlen(h_header);
  for (size_t k = 80 * 6; k + 80 <= len; k += 80) {
    std::string rec(h_header + k, 80);
    if (rec.compare(0, 10, "END       ") == 0) break;
    if (rec[0] == '#') continue;
    h.extra_cards.push_back(rec);
  }
  const uint64_t ncoords = (uint64_t)h.naxes[0] * (uint64_t)h.naxes[1];
  for (uint64_t i = 0; i < ncoords; ++i) h_datacube[i] /= h_weightscube[i];  // NaN where the weight is 0
  h.data = h_datacube;
  h.ndata = ncoords;
  std::string err;
  if (!hcgio::write_file(infile, {h}, err)) fail("write_output_map: " + err);
}

// ---- sorted input (HCGrid/gmap.cpp:326-379), written as the input's own layout ----

void write_ordered_map(const char* infile, const char* outfile) {
  clear_error();
  hcgio::FitsFile f;
  std::string err;
  if (!f.open(infile, err)) {
    fail("write_ordered_map: " + err);
    return;
  }
  const hcgio::Hdu& p = f.hdu(0);
  const uint64_t n = h_GMaps.data_shape;
  const uint64_t z = h_GMaps.spec_dim ? h_GMaps.spec_dim : 1;
  auto copy_cards = [](const hcgio::Hdu& hd) {
    std::vector<std::string> out;
    for (const auto& c : hd.cards) {
      const std::string kw = c.substr(0, 8);
      if (kw.compare(0, 6, "SIMPLE") == 0 || kw.compare(0, 6, "BITPIX") == 0 || kw.compare(0, 5, "NAXIS") == 0 ||
          kw.compare(0, 6, "EXTEND") == 0 || kw.compare(0, 8, "XTENSION") == 0 || kw.compare(0, 6, "PCOUNT") == 0 ||
          kw.compare(0, 6, "GCOUNT") == 0 || kw.compare(0, 8, "CHECKSUM") == 0 || kw.compare(0, 7, "DATASUM") == 0 ||
          kw.compare(0, 6, "BSCALE") == 0 || kw.compare(0, 5, "BZERO") == 0)
        continue;
      out.push_back(c);
    }
    return out;
  };
  hcgio::HduOut prim;
  prim.bitpix = p.bitpix < 0 ? p.bitpix : -64;
  prim.naxes = p.naxes;
  prim.extra_cards = copy_cards(p);
  prim.data = h_data;
  prim.ndata = n * z;
  std::vector<double> coords(2 * n);
  for (uint64_t i = 0; i < n; ++i) {
    coords[i] = h_lons[i];
    coords[n + i] = h_lats[i];
  }
  hcgio::HduOut ext;
  const hcgio::Hdu& c = f.num_hdus() > 1 ? f.hdu(1) : p;
  ext.bitpix = c.bitpix < 0 ? c.bitpix : -64;
  ext.naxes = f.num_hdus() > 1 ? c.naxes : std::vector<int64_t>{(int64_t)n, 2};
  ext.extra_cards = f.num_hdus() > 1 ? copy_cards(c) : std::vector<std::string>{};
  ext.data = coords.data();
  ext.ndata = 2 * n;
  if (!hcgio::write_file(outfile, {prim, ext}, err)) fail("write_ordered_map: " + err);
}

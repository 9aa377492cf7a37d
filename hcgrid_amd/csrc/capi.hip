// capi.hip -- C ABI plumbing: errors, memory, set-up and target geometry.
#include <hip/hip_runtime.h>
#include <math.h>

#include <string>

#include "common.hpp"
#include "healpix.hpp"

namespace hcg {
namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
}  // namespace hcg

using namespace hcg;

extern "C" const char* hcg_strerror(int code) {
  switch (code) {
    case HCG_OK: return "ok";
    case HCG_ERR_INVALID: return "invalid argument";
    case HCG_ERR_HIP: return "HIP runtime error";
    case HCG_ERR_NOMEM: return "out of memory";
    case HCG_ERR_UNSUPPORTED: return "unsupported size or parameter";
    case HCG_ERR_WORKSPACE: return "workspace too small";
    case HCG_ERR_IO: return "I/O or FITS format error";
    default: return "unknown error";
  }
}

extern "C" const char* hcg_last_error_message(void) { return g_last_error.c_str(); }

extern "C" int hcg_device_count(int* count) {
  HCG_REQUIRE(count != nullptr, "hcg_device_count: NULL");
  HCG_CHECK_HIP(hipGetDeviceCount(count));
  return HCG_OK;
}

extern "C" int hcg_set_device(int device) {
  HCG_CHECK_HIP(hipSetDevice(device));
  return HCG_OK;
}

extern "C" int hcg_malloc(void** dptr, size_t bytes) {
  HCG_REQUIRE(dptr != nullptr, "hcg_malloc: NULL");
  *dptr = nullptr;
  if (bytes == 0) return HCG_OK;
  hipError_t e = hipMalloc(dptr, bytes);
  if (e != hipSuccess) {
    *dptr = nullptr;
    return set_error(HCG_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return HCG_OK;
}

extern "C" int hcg_free(void* dptr) {
  if (dptr) HCG_CHECK_HIP(hipFree(dptr));
  return HCG_OK;
}

extern "C" int hcg_memcpy_h2d(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return HCG_OK;
  HCG_REQUIRE(dst && src, "hcg_memcpy_h2d: NULL");
  HCG_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
  return HCG_OK;
}

extern "C" int hcg_memcpy_d2h(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return HCG_OK;
  HCG_REQUIRE(dst && src, "hcg_memcpy_d2h: NULL");
  HCG_CHECK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
  return HCG_OK;
}

extern "C" int hcg_memset(void* dst, int value, size_t bytes, void* stream) {
  if (bytes == 0) return HCG_OK;
  HCG_REQUIRE(dst != nullptr, "hcg_memset: NULL");
  HCG_CHECK_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
  return HCG_OK;
}

extern "C" int hcg_stream_create(void** stream) {
  HCG_REQUIRE(stream != nullptr, "hcg_stream_create: NULL");
  hipStream_t s;
  HCG_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return HCG_OK;
}

extern "C" int hcg_stream_destroy(void* stream) {
  if (stream) HCG_CHECK_HIP(hipStreamDestroy(as_stream(stream)));
  return HCG_OK;
}

extern "C" int hcg_stream_synchronize(void* stream) {
  HCG_CHECK_HIP(hipStreamSynchronize(as_stream(stream)));
  return HCG_OK;
}

// _Healpix_init + _update_params (HCGrid/healpix.cu:42-65).
extern "C" int hcg_healpix_init(uint64_t nside, hcg_healpix* hp) {
  HCG_REQUIRE(hp != nullptr, "hcg_healpix_init: NULL");
  HCG_REQUIRE(nside >= 1 && (nside & (nside - 1)) == 0 && nside <= (1ull << 29),
              "hcg_healpix_init: nside must be a power of two");
  hp->nside = nside;
  hp->order = ilog2_u64(nside);
  hp->nrings = 4 * nside - 1;
  hp->npface = nside * nside;
  hp->ncap = (hp->npface - nside) << 1;
  hp->npix = 12 * hp->npface;
  hp->fact2 = 4. / (double)hp->npix;
  hp->fact1 = (double)(nside << 1) * hp->fact2;
  hp->omega = 4. * kPi / (double)hp->npix;
  hp->resolution = sqrt(hp->omega);
  return HCG_OK;
}

// HCGrid/helpers.cu:98-102
extern "C" uint64_t hcg_set_optimal_nside(double target_res) {
  uint64_t nside = (uint64_t)(sqrt(kPi / 3.) / target_res + 0.5);
  return (uint64_t)1 << (ilog2_u64(nside) + 1);
}

// HCGrid/gmap.cpp:64-99
extern "C" int hcg_prepare_grid_kernel(uint32_t kernel_type, const double* kernel_params, double sphere_radius,
                                       double hpx_max_res, hcg_kernel* kern, hcg_healpix* hp) {
  HCG_REQUIRE(kernel_params && kern && hp, "hcg_prepare_grid_kernel: NULL argument");
  HCG_REQUIRE(sphere_radius > 0 && hpx_max_res > 0, "hcg_prepare_grid_kernel: radius and resolution must be > 0");
  hcg_kernel k{};
  k.kernel_type = kernel_type;
  k.factor = 1;  // the caller sets the coarsening factor afterwards (HCGrid/HCGrid.cpp:158-161)
  if (kernel_type == HCG_GAUSS1D) {
    HCG_REQUIRE(kernel_params[0] > 0, "hcg_prepare_grid_kernel: sigma must be > 0");
    k.params[0] = 0.5 / pow(kernel_params[0], 2.);
    k.bearing_needed = 0;
  } else if (kernel_type == HCG_GAUSS2D) {
    for (int i = 0; i < 3; ++i) k.params[i] = kernel_params[i];
    k.bearing_needed = 1;
  } else if (kernel_type == HCG_TAPERED_SINC) {
    for (int i = 0; i < 3; ++i) k.params[i] = kernel_params[i];
    k.bearing_needed = 0;
  } else {
    return set_error(HCG_ERR_INVALID, "hcg_prepare_grid_kernel: unknown kernel type");
  }
  const uint64_t nside = hcg_set_optimal_nside(kDeg2Rad * hpx_max_res);
  int rc = hcg_healpix_init(nside, hp);
  if (rc != HCG_OK) return rc;
  k.sphere_radius = sphere_radius;
  k.disc_size = (kDeg2Rad * sphere_radius + hp->resolution);
  *kern = k;
  return HCG_OK;
}

extern "C" uint64_t hcg_host_ang2pix(const hcg_healpix* hp, double theta, double phi) {
  return hp ? ang2pix(*hp, theta, phi) : kMaxY;
}

extern "C" uint64_t hcg_host_pix2ring(const hcg_healpix* hp, uint64_t pix) { return hp ? pix2ring(*hp, pix) : 0; }

extern "C" void hcg_host_ring_info(const hcg_healpix* hp, uint64_t ring, uint64_t* startpix, uint64_t* npr) {
  uint64_t s = 0, n = 0;
  if (hp) ring_info(*hp, ring, s, n);
  if (startpix) *startpix = s;
  if (npr) *npr = n;
}

extern "C" double hcg_host_pix2theta(const hcg_healpix* hp, uint64_t pix) { return hp ? pix2theta(*hp, pix) : 0.0; }

// Zenithal deprojections (Calabretta & Greisen 2002, Paper II sec. 5.1:
// SIN with PV2_1 = PV2_2 = 0, TAN, ARC, ZEA, STG) and the native ->
// celestial rotation (Paper II eq. 2) with (alpha_p, delta_p) = CRVAL and
// phi_p = LONPOLE.  Cell (x, y) <- pixel (x+1, y+1) as set_WCS builds pixcrd
// (HCGrid/gmap.cpp:209-238).
namespace {

// One zenithal-projection cell: intermediate world coordinates (x, y) in
// radians -> native longitude phi = atan2(x, -y) and native latitude theta
// from R = sqrt(x^2 + y^2) through the projection's R(theta) (radians):
//   SIN R = cos(theta)            TAN R = cot(theta)
//   ARC R = pi/2 - theta          ZEA R = 2 sin((pi/2 - theta) / 2)
//   STG R = 2 tan((pi/2 - theta) / 2)
// -> (ra, dec) in degrees; NaN outside the projection's domain (SIN R > 1,
// ARC R > pi, ZEA R > 2).
__host__ __device__ inline void wcs_zenithal_cell(uint32_t proj, double x, double y, double a0, double sd0,
                                                  double cd0, double php, double& ra_deg, double& dec_deg) {
  const double r2 = x * x + y * y;
  double sth, cth;  // sin and cos of the native latitude
  bool ok = true;
  switch (proj) {
    case HCG_PROJ_TAN: {
      const double q = 1.0 / sqrt(1.0 + r2);
      sth = q;
      cth = sqrt(r2) * q;
      break;
    }
    case HCG_PROJ_ARC: {
      const double r = sqrt(r2);
      ok = r <= kPi;
      sth = cos(r);
      cth = sin(r);
      break;
    }
    case HCG_PROJ_ZEA: {
      ok = r2 <= 4.0;
      sth = 1.0 - 0.5 * r2;
      cth = sqrt(r2) * sqrt(ok ? 1.0 - 0.25 * r2 : 0.0);
      break;
    }
    case HCG_PROJ_STG: {
      const double q = 1.0 / (1.0 + 0.25 * r2);
      sth = (1.0 - 0.25 * r2) * q;
      cth = sqrt(r2) * q;
      break;
    }
    default: {  // HCG_PROJ_SIN
      ok = r2 <= 1.0;
      cth = sqrt(r2);
      sth = sqrt(ok ? 1.0 - r2 : 0.0);
      break;
    }
  }
  if (!ok) {
    ra_deg = NAN;
    dec_deg = NAN;
    return;
  }
  const double phi = atan2(x, -y);
  const double dphi = phi - php;
  const double sdp = sin(dphi), cdp = cos(dphi);
  const double sdec = sth * sd0 + cth * cd0 * cdp;
  const double dec = asin(sdec > 1.0 ? 1.0 : (sdec < -1.0 ? -1.0 : sdec));
  double ra = (a0 + atan2(-cth * sdp, sth * cd0 - cth * sd0 * cdp)) * kRad2Deg;
  ra = fmod(ra, 360.0);
  if (ra < 0) ra += 360.0;
  ra_deg = ra;
  dec_deg = dec * kRad2Deg;
}

__global__ void __launch_bounds__(256) k_wcs_zenithal(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1,
                                                      double crpix2, double cdelt1, double cdelt2, double a0,
                                                      double sd0, double cd0, double php, double* __restrict__ xwcs,
                                                      double* __restrict__ ywcs) {
  const uint64_t c = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= (uint64_t)naxis1 * naxis2) return;
  const uint32_t yy = (uint32_t)(c / naxis1), xx = (uint32_t)(c % naxis1);
  const double y = cdelt2 * ((double)(yy + 1) - crpix2) * kDeg2Rad;
  const double x = cdelt1 * ((double)(xx + 1) - crpix1) * kDeg2Rad;
  wcs_zenithal_cell(proj, x, y, a0, sd0, cd0, php, xwcs[c], ywcs[c]);
}

}  // namespace

extern "C" int hcg_wcs_p2s(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                           double cdelt1, double cdelt2, double crval1, double crval2, double lonpole, double* xwcs,
                           double* ywcs) {
  HCG_REQUIRE(xwcs && ywcs, "hcg_wcs_p2s: NULL output");
  HCG_REQUIRE(proj <= HCG_PROJ_STG, "hcg_wcs_p2s: unknown projection");
  const double a0 = crval1 * kDeg2Rad, d0 = crval2 * kDeg2Rad, php = lonpole * kDeg2Rad;
  const double sd0 = sin(d0), cd0 = cos(d0);
  for (uint32_t yy = 0; yy < naxis2; ++yy) {
    const double y = cdelt2 * ((double)(yy + 1) - crpix2) * kDeg2Rad;
    for (uint32_t xx = 0; xx < naxis1; ++xx) {
      const double x = cdelt1 * ((double)(xx + 1) - crpix1) * kDeg2Rad;
      const uint64_t c = (uint64_t)yy * naxis1 + xx;
      wcs_zenithal_cell(proj, x, y, a0, sd0, cd0, php, xwcs[c], ywcs[c]);
    }
  }
  return HCG_OK;
}

extern "C" int hcg_wcs_p2s_device(uint32_t proj, uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                                  double cdelt1, double cdelt2, double crval1, double crval2, double lonpole,
                                  double* xwcs, double* ywcs, void* stream) {
  HCG_REQUIRE(xwcs && ywcs, "hcg_wcs_p2s_device: NULL output");
  HCG_REQUIRE(proj <= HCG_PROJ_STG, "hcg_wcs_p2s_device: unknown projection");
  const uint64_t m = (uint64_t)naxis1 * naxis2;
  if (m == 0) return HCG_OK;
  HCG_REQUIRE(m < (1ull << 40), "hcg_wcs_p2s_device: too many cells");
  const double a0 = crval1 * kDeg2Rad, d0 = crval2 * kDeg2Rad, php = lonpole * kDeg2Rad;
  const double sd0 = sin(d0), cd0 = cos(d0);
  k_wcs_zenithal<<<(uint32_t)((m + 255) / 256), 256, 0, as_stream(stream)>>>(
      proj, naxis1, naxis2, crpix1, crpix2, cdelt1, cdelt2, a0, sd0, cd0, php, xwcs, ywcs);
  HCG_CHECK_LAUNCH("k_wcs_zenithal");
  return HCG_OK;
}

extern "C" int hcg_wcs_sin_p2s(uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2, double cdelt1,
                               double cdelt2, double crval1, double crval2, double lonpole, double* xwcs,
                               double* ywcs) {
  return hcg_wcs_p2s(HCG_PROJ_SIN, naxis1, naxis2, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2, lonpole, xwcs,
                     ywcs);
}

extern "C" int hcg_wcs_sin_p2s_device(uint32_t naxis1, uint32_t naxis2, double crpix1, double crpix2,
                                      double cdelt1, double cdelt2, double crval1, double crval2, double lonpole,
                                      double* xwcs, double* ywcs, void* stream) {
  return hcg_wcs_p2s_device(HCG_PROJ_SIN, naxis1, naxis2, crpix1, crpix2, cdelt1, cdelt2, crval1, crval2, lonpole,
                            xwcs, ywcs, stream);
}

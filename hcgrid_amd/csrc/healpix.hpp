// healpix.hpp -- HEALPix RING geometry for host and gfx950 device code.
//
// Pixel indices must equal the reference's bit for bit (keys, sort order and
// the kernel's candidate windows all derive from them), so the operation
// sequence follows HCGrid/healpix.cu and HCGrid/helpers.cu exactly: the same
// folded constants (HCGrid/constants.h:75-84), the same order of every
// multiply and add, and no FMA contraction (this file is only compiled with
// -ffp-contract=off).  The transcendental calls (cos, sin, acos, atan2, sqrt)
// are the device library's; sqrt is correctly rounded on both sides, and the
// pixel index changes with a 1-ulp difference in cos only when a product lands
// within ~1e-12 of an integer boundary.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hcgrid_amd.h"

namespace hcg {

// HCGrid/constants.h:75-84, folded in double as the preprocessor macros are.
constexpr double kPi = 3.1415926535897932384626433832;
constexpr double kTwoThird = 2.0 / 3.0;
constexpr double kHalfPi = kPi / 2.;
constexpr double kInvHalfPi = 1. / kHalfPi;
constexpr double kDeg2Rad = kPi / 180.;
constexpr double kRad2Deg = 180. / kPi;
constexpr int kNorthB = 20;
constexpr int kSouthB = 160;
constexpr uint64_t kMaxY = 1073741824ull;

__host__ __device__ inline uint64_t ilog2_u64(uint64_t arg) {  // HCGrid/helpers.cu:40-64
  uint64_t r = 0;
  while (arg > 0xFFFFull) { r += 16; arg >>= 16; }
  if (arg > 0xFFull) { r |= 8; arg >>= 8; }
  if (arg > 0xFull) { r |= 4; arg >>= 4; }
  if (arg > 0x3ull) { r |= 2; arg >>= 2; }
  if (arg > 0x1ull) { r |= 1; }
  return r;
}

__host__ __device__ inline uint64_t isqrt_u64(uint64_t arg) {  // HCGrid/helpers.cu:67-69
  return (uint64_t)sqrt((double)arg + 0.5);
}

__host__ __device__ inline double fmodulo(double v1, double v2) {  // HCGrid/helpers.cu:83-95
  if (v1 >= 0) return (v1 < v2) ? v1 : fmod(v1, v2);
  double t = fmod(v1, v2) + v2;
  return (t == v2) ? 0. : t;
}

// fmod(v, 4.0) without the library's iterative reduction.  v * 0.25 and
// 4.0 * trunc(.) are exact, and |4 trunc(v/4)| lies in [|v|/2, |v|] whenever
// it is non-zero, so the subtraction is exact (Sterbenz): the result equals
// fmod(v, 4.0) bit for bit for every finite v (and is NaN for NaN/inf).
__host__ __device__ inline double fmod4(double v) { return v - 4.0 * trunc(v * 0.25); }

// fmodulo(v1, 4.0) as loc2pix uses it.
__host__ __device__ inline double fmodulo4(double v1) {
  if (v1 >= 0) return (v1 < 4.0) ? v1 : fmod4(v1);
  double t = fmod4(v1) + 4.0;
  return (t == 4.0) ? 0. : t;
}

// Ring start pixel and size (HCGrid/healpix.cu:68-105).
__host__ __device__ inline void ring_info(const hcg_healpix& hp, uint64_t ring, uint64_t& startpix,
                                          uint64_t& npr) {
  if (ring < hp.nside) {
    npr = 4 * ring;
    startpix = 2 * ring * (ring - 1);
  } else if (ring < 3 * hp.nside) {
    npr = 4 * hp.nside;
    startpix = hp.ncap + (ring - hp.nside) * npr;
  } else {
    uint64_t nr = 4 * hp.nside - ring;
    npr = 4 * nr;
    startpix = hp.npix - 2 * nr * (nr + 1);
  }
}

// Ring of a pixel (HCGrid/healpix.cu:108-123).
__host__ __device__ inline uint64_t pix2ring(const hcg_healpix& hp, uint64_t pix) {
  if (pix < hp.ncap) return (1 + isqrt_u64(1 + 2 * pix)) >> 1;
  // (pix - ncap) / (4 nside): nside = 2^order, so the division is a shift
  if (pix < hp.npix - hp.ncap) return ((pix - hp.ncap) >> (hp.order + 2)) + hp.nside;
  return 4 * hp.nside - ((1 + isqrt_u64(2 * (hp.npix - pix) - 1)) >> 1);
}

// (z, phi) -> RING pixel (HCGrid/healpix.cu:126-205).
__host__ __device__ inline uint64_t loc2pix(const hcg_healpix& hp, double z, double phi,
                                            double sin_theta, bool have_sin) {
  const double za = fabs(z);
  const double tt = fmodulo4(phi * kInvHalfPi);
  const double dn = (double)hp.nside;
  if (za <= kTwoThird) {
    const uint64_t nl4 = 4 * hp.nside;
    const double t1 = dn * (0.5 + tt);
    const double t2 = dn * z * 0.75;
    const uint64_t jp = (uint64_t)(t1 - t2);
    const uint64_t jm = (uint64_t)(t1 + t2);
    const uint64_t ir = hp.nside + 1 + jp - jm;
    const uint64_t kshift = 1 - (ir & 1);
    const uint64_t tsum = jp + jm + kshift + 1 + nl4 + nl4 - hp.nside;
    const uint64_t ip = (tsum >> 1) & (nl4 - 1);  // nl4 is a power of two: % nl4 == & (nl4 - 1)
    return hp.ncap + (ir - 1) * nl4 + ip;
  }
  const double tp = tt - (double)((uint64_t)tt);
  double tmp;
  if ((za < 0.99) || !have_sin)
    tmp = dn * sqrt(3. * (1. - za));
  else
    tmp = dn * sin_theta / sqrt((1. + za) / 3.);
  const uint64_t jp = (uint64_t)(tp * tmp);
  const uint64_t jm = (uint64_t)((1.0 - tp) * tmp);
  const uint64_t ir = jp + jm + 1;
  const uint64_t ip = (uint64_t)(tt * (double)ir);
  return (z > 0) ? (2 * ir * (ir - 1) + ip) : (hp.npix - 2 * ir * (ir + 1) + ip);
}

// (theta, phi) -> RING pixel, MAX_Y outside [0, pi] (HCGrid/healpix.cu:208-224).
__host__ __device__ inline uint64_t ang2pix(const hcg_healpix& hp, double theta, double phi) {
  if (!(theta >= 0 && theta <= kPi)) return kMaxY;
  if ((theta < 0.01) || (theta > kPi - 0.01)) return loc2pix(hp, cos(theta), phi, sin(theta), true);
  return loc2pix(hp, cos(theta), phi, 0., false);
}

// Pixel centre colatitude (HCGrid/healpix.cu:227-348); only theta is needed.
__host__ __device__ inline double pix2theta(const hcg_healpix& hp, uint64_t pix) {
  double z, s = 0.;
  bool have = false;
  if (pix < hp.ncap) {
    const uint64_t iring = (1 + isqrt_u64(1 + 2 * pix)) >> 1;
    const double tmp = (double)(iring * iring) * hp.fact2;
    z = 1.0 - tmp;
    if (z > 0.99) { s = sqrt(tmp * (2. - tmp)); have = true; }
  } else if (pix < hp.npix - hp.ncap) {
    const uint64_t ip = pix - hp.ncap;
    const uint64_t itmp = ip >> (hp.order + 2);
    const uint64_t iring = itmp + hp.nside;
    z = (2. * (double)hp.nside - (double)iring) * hp.fact1;
  } else {
    const uint64_t ip = hp.npix - pix;
    const uint64_t iring = (1 + isqrt_u64(2 * ip - 1)) >> 1;
    const double tmp = (double)(iring * iring) * hp.fact2;
    z = tmp - 1.0;
    if (z < -0.99) { s = sqrt(tmp * (2. - tmp)); have = true; }
  }
  return have ? atan2(s, z) : acos(z);
}

// Per-ring constants of the kernel's window evaluation: the reference calls
// ang2pix(theta_r, phi) with theta_r = pix2ang(startpix(r)).theta; that is
// loc2pix(cos theta_r, phi, sin theta_r, have) with `have` from the polar test.
struct RingGeom {
  double z, s;
  bool have_sin;
};

__host__ __device__ inline RingGeom ring_geom(const hcg_healpix& hp, uint64_t ring) {
  uint64_t sp, npr;
  ring_info(hp, ring, sp, npr);
  const double th = pix2theta(hp, sp);
  RingGeom g;
  g.have_sin = (th < 0.01) || (th > kPi - 0.01);
  g.z = cos(th);
  g.s = g.have_sin ? sin(th) : 0.;
  return g;
}

}  // namespace hcg

// wcs.hpp -- the target map's celestial WCS on the host (set_WCS,
// HCGrid/gmap.cpp:183-250); the formulas are in wcs.cpp.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace hcgwcs {

enum Proj : uint32_t { SIN, TAN, ARC, ZEA, STG, CAR, MER, CEA, SFL, GLS, AIT };

struct Celestial {
  Proj proj = SIN;
  bool ncp = false;      // CTYPE ...-NCP (SIN with the slant (0, cot CRVAL2))
  uint32_t linear = 0;   // the header's linear part: 0 CDELT, 1 PC, 2 CD, 3 CROTA2
  double crpix[2] = {0, 0};
  double crval[2] = {0, 0};
  double m[4] = {1, 0, 0, 1};  // degrees per pixel, row-major
  double pv[2] = {0, 0};       // SIN (xi, eta); CEA lambda in pv[0]
  double lonpole = 0, latpole = 90;
};

// Parses CTYPE/CRPIX/CRVAL, the linear part and the projection parameters
// of axes 1 and 2; false (err set) for anything the formulas do not cover.
bool from_header(const std::vector<std::string>& cards, Celestial& w, std::string& err);

// A zenithal target with a diagonal CDELT and no slant: the C ABI's
// hcg_wcs_p2s covers it (the form the device entry point mirrors).
bool plain_zenithal(const Celestial& w);

// Cell (x, y) = pixel (x + 1, y + 1) -> (lon, lat) degrees at y * naxis1 + x,
// NaN outside the projection's domain.
bool p2s(const Celestial& w, uint32_t naxis1, uint32_t naxis2, double* lon, double* lat, std::string& err);

}  // namespace hcgwcs

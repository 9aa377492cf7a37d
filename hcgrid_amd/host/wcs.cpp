// wcs.cpp -- the target map's celestial WCS on the host: what the
// reference's set_WCS takes from wcslib's wcsp2s (HCGrid/gmap.cpp:183-250),
// written from Calabretta & Greisen 2002 (FITS WCS Papers I and II).
//
// A cell (x, y) is pixel p = (x + 1, y + 1).  Three steps:
//  1. intermediate world coordinates (degrees) (u, v) = M (p - CRPIX), with
//     M = CD, or diag(CDELT) PC, or the PC that a legacy CROTA2 stands for
//     (Paper II sec. 6.1), or diag(CDELT);
//  2. native (phi, theta) by the projection's inverse (radians X, Y):
//       zenithal, theta_0 = 90:  SIN (with the slant (xi, eta) = (PV2_1,
//         PV2_2); NCP is SIN with xi = 0, eta = cot delta_0), TAN, ARC,
//         ZEA, STG (sec. 5.1);
//       theta_0 = 0: CAR phi = X, theta = Y; MER theta = 2 atan(e^Y) - pi/2;
//         CEA theta = asin(lambda Y), lambda = PV2_1 (default 1); SFL
//         theta = Y, phi = X / cos theta; AIT Z = sqrt(1 - (X/4)^2 -
//         (Y/2)^2), phi = 2 atan2(X Z / 2, 2 Z^2 - 1), theta = asin(Y Z)
//         (secs. 5.2-5.3);
//  3. native -> celestial (eq. 2) about the native pole's celestial
//     coordinates (alpha_p, delta_p), phi_p = LONPOLE (default 0 when
//     delta_0 >= theta_0, else 180): zenithal (alpha_p, delta_p) = CRVAL;
//     theta_0 = 0: delta_p = atan2(0, cos phi_p) +- acos(sin delta_0 /
//     |cos phi_p|) (eq. 8), the root in [-90, 90] nearest LATPOLE (default
//     90), and alpha_p so that the reference point lands on CRVAL.
// GLS follows the AIPS convention (Greisen, AIPS memo 27): dec = CRVAL2 +
// v, ra = CRVAL1 + u / cos(dec), no rotation (wcslib's translation of it
// to SFL with theta_0 = CRVAL2 gives the same map).
// NaN outside a projection's domain, as wcsp2s flags such pixels.
#include <math.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "fits_io.hpp"
#include "wcs.hpp"

namespace hcgwcs {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr double kD2R = kPi / 180.0;
constexpr double kTol = 1e-12;

struct Code {
  const char* name;
  Proj proj;
};
constexpr Code kCodes[] = {{"SIN", SIN}, {"TAN", TAN}, {"ARC", ARC}, {"ZEA", ZEA}, {"STG", STG}, {"CAR", CAR},
                           {"MER", MER}, {"CEA", CEA}, {"SFL", SFL}, {"GLS", GLS}, {"AIT", AIT}, {"NCP", SIN}};

bool lon_type(const std::string& t) { return t == "RA--" || t == "GLON" || t == "ELON"; }
bool lat_type(const std::string& t) { return t == "DEC-" || t == "GLAT" || t == "ELAT"; }

// Native (phi, theta) of intermediate (X, Y) radians; false outside the domain.
bool deproject(const Celestial& w, double X, double Y, double& phi, double& sth, double& cth) {
  switch (w.proj) {
    case SIN: {
      const double xi = w.pv[0], eta = w.pv[1];
      if (xi == 0.0 && eta == 0.0) {  // the orthographic projection itself
        const double r2 = X * X + Y * Y;
        if (r2 > 1.0) return false;
        cth = sqrt(r2);
        sth = sqrt(1.0 - r2);
        phi = atan2(X, -Y);
        return true;
      }
      // sin(theta) = s solves a s^2 + 2 b s + c = 0 (from (X - xi (1 - s))^2
      // + (Y - eta (1 - s))^2 = 1 - s^2); the root nearest the pole
      const double a = xi * xi + eta * eta + 1.0, b = xi * (X - xi) + eta * (Y - eta);
      const double c = (X - xi) * (X - xi) + (Y - eta) * (Y - eta) - 1.0;
      const double disc = b * b - a * c;
      if (disc < 0.0) return false;
      double s = (-b + sqrt(disc)) / a;
      if (s > 1.0 + kTol) s = (-b - sqrt(disc)) / a;
      if (fabs(s) > 1.0 + kTol) return false;
      s = fmin(s, 1.0);
      const double ex = X - xi * (1.0 - s), ey = Y - eta * (1.0 - s);
      sth = s;
      cth = sqrt(ex * ex + ey * ey);
      phi = atan2(ex, -ey);
      return true;
    }
    case TAN: {
      const double r2 = X * X + Y * Y, q = 1.0 / sqrt(1.0 + r2);
      sth = q;
      cth = sqrt(r2) * q;
      phi = atan2(X, -Y);
      return true;
    }
    case ARC: {
      const double r = sqrt(X * X + Y * Y);
      if (r > kPi) return false;
      sth = cos(r);
      cth = sin(r);
      phi = atan2(X, -Y);
      return true;
    }
    case ZEA: {
      const double r2 = X * X + Y * Y;
      if (r2 > 4.0) return false;
      sth = 1.0 - 0.5 * r2;
      cth = sqrt(r2) * sqrt(1.0 - 0.25 * r2);
      phi = atan2(X, -Y);
      return true;
    }
    case STG: {
      const double r2 = X * X + Y * Y, q = 1.0 / (1.0 + 0.25 * r2);
      sth = (1.0 - 0.25 * r2) * q;
      cth = sqrt(r2) * q;
      phi = atan2(X, -Y);
      return true;
    }
    case CAR:
      if (fabs(Y) > 0.5 * kPi) return false;
      phi = X;
      sth = sin(Y);
      cth = cos(Y);
      return true;
    case MER: {
      const double th = 2.0 * atan(exp(Y)) - 0.5 * kPi;
      phi = X;
      sth = sin(th);
      cth = cos(th);
      return true;
    }
    case CEA: {
      const double s = w.pv[0] * Y;
      if (fabs(s) > 1.0) return false;
      phi = X;
      sth = s;
      cth = sqrt((1.0 - s) * (1.0 + s));
      return true;
    }
    case SFL: {
      if (fabs(Y) > 0.5 * kPi) return false;
      cth = cos(Y);
      sth = sin(Y);
      phi = cth > 0.0 ? X / cth : 0.0;
      return fabs(phi) <= kPi;
    }
    case AIT: {
      const double z2 = 1.0 - 0.0625 * X * X - 0.25 * Y * Y;
      if (z2 < 0.5) return false;
      const double z = sqrt(z2), s = Y * z;
      phi = 2.0 * atan2(0.5 * X * z, 2.0 * z2 - 1.0);
      sth = s;
      cth = sqrt((1.0 - s) * (1.0 + s));
      return true;
    }
    default:
      return false;
  }
}

// A numeric card into out; out keeps its value when the card is absent or
// does not parse (hcgio::card_double writes before it knows).
bool card(const std::vector<std::string>& cards, const char* key, double& out) {
  double v;
  if (!hcgio::card_double(cards, key, v)) return false;
  out = v;
  return true;
}

double wrap360(double a) {
  a = fmod(a, 360.0);
  if (a < 0) a += 360.0;
  return a;
}

}  // namespace

bool from_header(const std::vector<std::string>& cards, Celestial& w, std::string& err) {
  std::string ct1, ct2;
  if (!hcgio::card_string(cards, "CTYPE1", ct1) || !hcgio::card_string(cards, "CTYPE2", ct2) ||
      !card(cards, "CRPIX1", w.crpix[0]) || !card(cards, "CRPIX2", w.crpix[1]) ||
      !card(cards, "CRVAL1", w.crval[0]) || !card(cards, "CRVAL2", w.crval[1])) {
    err = "set_WCS: target header lacks CTYPE/CRPIX/CDELT/CRVAL for axes 1 and 2";
    return false;
  }
  bool known = false;
  if (ct1.size() == 8 && ct2.size() == 8 && ct1.substr(4) == ct2.substr(4) && ct1[4] == '-')
    for (const Code& c : kCodes)
      if (ct1.substr(5) == c.name) {
        w.proj = c.proj;
        w.ncp = ct1.substr(5) == "NCP";
        known = true;
      }
  if (!known || !lon_type(ct1.substr(0, 4)) || !lat_type(ct2.substr(0, 4))) {
    err = "set_WCS: supported targets are SIN (NCP), TAN, ARC, ZEA, STG, CAR, MER, CEA, SFL, GLS and AIT with "
          "longitude on axis 1 and latitude on axis 2 (got " + ct1 + ", " + ct2 + ")";
    return false;
  }
  // the linear part: PC (with CDELT) over CD over CROTA2 over CDELT alone
  double cdelt[2] = {1.0, 1.0}, pc[4] = {1.0, 0.0, 0.0, 1.0}, cd[4] = {0.0, 0.0, 0.0, 0.0}, crota = 0.0;
  const bool have_cdelt = card(cards, "CDELT1", cdelt[0]) & card(cards, "CDELT2", cdelt[1]);
  static const char* const kPc[] = {"PC1_1", "PC1_2", "PC2_1", "PC2_2"};
  static const char* const kCd[] = {"CD1_1", "CD1_2", "CD2_1", "CD2_2"};
  bool have_pc = false, have_cd = false;
  for (int k = 0; k < 4; ++k) {
    have_pc |= card(cards, kPc[k], pc[k]);
    have_cd |= card(cards, kCd[k], cd[k]);
  }
  const bool have_crota = card(cards, "CROTA2", crota);
  w.linear = have_pc ? 1 : have_cd ? 2 : (have_crota && crota != 0.0) ? 3 : 0;
  if (w.linear != 2 && !have_cdelt) {
    err = "set_WCS: target header lacks CTYPE/CRPIX/CDELT/CRVAL for axes 1 and 2";
    return false;
  }
  if (w.linear == 2) {
    for (int k = 0; k < 4; ++k) w.m[k] = cd[k];
  } else {
    if (w.linear == 3) {  // Paper II eq. 188
      const double r = crota * kD2R;
      pc[0] = cos(r);
      pc[1] = -sin(r) * cdelt[1] / cdelt[0];
      pc[2] = sin(r) * cdelt[0] / cdelt[1];
      pc[3] = cos(r);
    }
    w.m[0] = cdelt[0] * pc[0];
    w.m[1] = cdelt[0] * pc[1];
    w.m[2] = cdelt[1] * pc[2];
    w.m[3] = cdelt[1] * pc[3];
  }
  if (w.m[0] * w.m[3] - w.m[1] * w.m[2] == 0.0) {
    err = "set_WCS: the target's pixel -> world matrix is singular";
    return false;
  }
  // projection parameters
  const bool zenithal = w.proj <= STG;
  w.pv[0] = w.proj == CEA ? 1.0 : 0.0;
  w.pv[1] = 0.0;
  if (w.proj == SIN || w.proj == CEA) {
    card(cards, "PV2_1", w.pv[0]);
    card(cards, "PV2_2", w.pv[1]);
  }
  if (w.ncp) {  // NCP = SIN with (xi, eta) = (0, cot delta_0) (Paper II sec. 5.1.5)
    if (sin(w.crval[1] * kD2R) == 0.0) {
      err = "set_WCS: NCP needs CRVAL2 != 0";
      return false;
    }
    w.pv[0] = 0.0;
    w.pv[1] = cos(w.crval[1] * kD2R) / sin(w.crval[1] * kD2R);
  }
  if (w.proj == CEA && !(w.pv[0] > 0.0 && w.pv[0] <= 1.0)) {
    err = "set_WCS: CEA needs 0 < PV2_1 <= 1";
    return false;
  }
  double pv11 = 0.0, pv12 = zenithal ? 90.0 : 0.0;
  const double theta0 = pv12;
  card(cards, "PV1_1", pv11);
  card(cards, "PV1_2", pv12);
  if (pv11 != 0.0 || pv12 != theta0) {
    err = "set_WCS: a native reference point other than the projection's (PV1_1, PV1_2) is not supported";
    return false;
  }
  w.lonpole = NAN;
  w.latpole = 90.0;
  card(cards, "LONPOLE", w.lonpole);
  card(cards, "LATPOLE", w.latpole);
  if (isnan(w.lonpole)) w.lonpole = w.crval[1] >= theta0 ? 0.0 : 180.0;
  return true;
}

bool plain_zenithal(const Celestial& w) {
  return w.proj <= STG && w.linear == 0 && w.pv[0] == 0.0 && w.pv[1] == 0.0;
}

bool p2s(const Celestial& w, uint32_t naxis1, uint32_t naxis2, double* lon, double* lat, std::string& err) {
  const double a0 = w.crval[0] * kD2R, d0 = w.crval[1] * kD2R, php = w.lonpole * kD2R;
  double ap = a0, dp = d0;
  if (w.proj >= CAR && w.proj != GLS) {
    const double den = fabs(cos(php)), sd0 = sin(d0);
    if (den < 1e-15) {  // phi_p = +-90: only delta_0 = 0 fits, and then any delta_p
      if (fabs(sd0) > 1e-15) {
        err = "set_WCS: LONPOLE admits no native pole for this CRVAL2";
        return false;
      }
      dp = w.latpole * kD2R;
    } else {
      const double u = sd0 / den;
      if (fabs(u) > 1.0 + kTol) {
        err = "set_WCS: LONPOLE admits no native pole for this CRVAL2";
        return false;
      }
      const double base = atan2(0.0, cos(php)), a = acos(fmax(-1.0, fmin(1.0, u)));
      const double lp = w.latpole * kD2R;
      bool found = false;
      for (int k = 0; k < 2; ++k) {
        double c = k ? base - a : base + a;
        if (c > kPi) c -= 2.0 * kPi;
        if (c < -kPi) c += 2.0 * kPi;
        if (fabs(c) > 0.5 * kPi + kTol) continue;
        c = fmax(-0.5 * kPi, fmin(0.5 * kPi, c));
        if (!found || fabs(c - lp) < fabs(dp - lp)) dp = c;
        found = true;
      }
      if (!found) {
        err = "set_WCS: LONPOLE admits no native pole for this CRVAL2";
        return false;
      }
    }
    ap = a0 - atan2(sin(php), -sin(dp) * cos(php));
  }
  const double sdp = sin(dp), cdp = cos(dp);
  for (uint32_t yy = 0; yy < naxis2; ++yy) {
    const double q2 = (double)(yy + 1) - w.crpix[1];
    for (uint32_t xx = 0; xx < naxis1; ++xx) {
      const double q1 = (double)(xx + 1) - w.crpix[0];
      const double X = (w.m[0] * q1 + w.m[1] * q2) * kD2R, Y = (w.m[2] * q1 + w.m[3] * q2) * kD2R;
      const uint64_t c = (uint64_t)yy * naxis1 + xx;
      lon[c] = lat[c] = NAN;
      if (w.proj == GLS) {
        const double dec = d0 + Y;
        if (fabs(dec) > 0.5 * kPi) continue;
        const double dl = cos(dec) > 0.0 ? X / cos(dec) : 0.0;
        if (fabs(dl) > kPi) continue;
        lon[c] = wrap360((a0 + dl) / kD2R);
        lat[c] = dec / kD2R;
        continue;
      }
      double phi, sth, cth;
      if (!deproject(w, X, Y, phi, sth, cth)) continue;
      const double dphi = phi - php, sdph = sin(dphi), cdph = cos(dphi);
      const double sdec = sth * sdp + cth * cdp * cdph;
      lat[c] = asin(fmax(-1.0, fmin(1.0, sdec))) / kD2R;
      lon[c] = wrap360((ap + atan2(-cth * sdph, sth * cdp - cth * sdp * cdph)) / kD2R);
    }
  }
  return true;
}

}  // namespace hcgwcs

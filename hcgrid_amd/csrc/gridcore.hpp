// gridcore.hpp -- device pieces shared by the gridding kernels: the
// reference's window/selection logic (HCGrid/gridding.cu:306-360), the
// distance and weight evaluation, and wave-level search/reduction helpers.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "expneg.hpp"
#include "healpix.hpp"

namespace hcg {

struct GridParams {
  hcg_healpix hp;
  double disc_size;
  double sphere_radius;
  double params[3];
  double c_lo, c_hi;   // guard band around the radius threshold on c2 = |p - q|^2 = 4 sin^2(d/2)
  // GAUSS1D exponent in units of ln2/64: y = c2 (k0 + c2 (k1 + c2 (k2 + c2 k3))) for c2 < 4e-4
  // (the series of -p0 RAD2DEG^2 4 asin^2(sqrt(c2)/2) * 64/ln2), else k_asin * asin^2(sqrt(c2)/2)
  double k0, k1, k2, k3, k_asin;
  double polar_cos, polar_sin, polar_half;  // cos R', sin R', sin(R'/2), R' = R (1 + 1e-6)
  uint32_t side;
  uint32_t factor;     // 1 under HCG_SELECT_DISC
  uint32_t selection;  // HCG_SELECT_COMPAT / HCG_SELECT_DISC
};

struct GridBuffers {
  const hcg_sample* rec;
  const uint32_t* keys;
  const double* lon_s;
  const double* lat_s;
  const double* w_s;
  const uint32_t* lut;
  const double* ring_z;
  const double* ring_s;
  const hcg_ring_meta* meta;
  const double* xwcs;
  const double* ywcs;
  double* dc;
  double* wc;
};

// The window of an anchor cell (HCGrid/gridding.cu:306-323).
//  * kWinCompat: the reference's per-ring longitude window phi +- disc_size.
//  * kWinPolar: the reference's whole rings (anchor colatitude <= 20 or >= 160
//    deg), narrowed in the polar-cap rings to the pixels whose longitude (in
//    units of pi/2) lies in [tlo, thi] when tlo <= thi: a range that provably
//    holds every sample within the radius of the cell (phi_range), so the
//    selected samples, and their order, are the reference's.
//  * kWinDisc (HCG_SELECT_DISC): rings two beyond the disc's, each narrowed
//    to [tlo, thi] when tlo <= thi, else whole: a superset of the disc.
enum : uint32_t { kWinCompat = 0, kWinPolar = 1, kWinDisc = 2 };

struct PolarRange {
  float tlo, thi;
};

struct Window {
  uint32_t uring, dring;
  union {
    double phi;        // kWinCompat: anchor longitude, radians
    PolarRange range;  // kWinPolar / kWinDisc: [tlo, thi], or whole rings when tlo > thi
  };
  uint32_t mode;
};

// The longitude range of the samples within the radius R of a cell, in units
// of pi/2 (the HEALPix tt).  A sample at colatitude ts within distance d of a
// cell at colatitude tc satisfies hav d = hav(tc - ts) + sin tc sin ts
// hav(dphi) >= sin tc sin ts sin^2(dphi / 2), so d < R bounds |dphi| by
// 2 asin s, s = sin(R/2) / sqrt(sin tc min sin ts), ts in (tc - R, tc + R)
// where sin is smallest at an end: min sin ts = sin tc cos R - |cos tc| sin R.
// asin s <= s / sqrt(1 - s^2) (tan x >= x).  R is widened by 1e-6 relative on
// the host (the reference's haversine decides the band around R to ~1e-15),
// the range by 1e-6 relative, and the pixel indices on each side (ring_window).
// Valid only where the anchor is the cell (factor 1).  No range (tlo > thi)
// when the disc reaches a pole, s >= 1/2, or the range would wrap through
// longitude 0.
__device__ inline void phi_range(const GridParams& P, double b1, double l1, Window& w) {
  w.range.tlo = 1.f;
  w.range.thi = 0.f;
  if (P.factor != 1u) return;
  double sb, cb;
  sincos(b1, &sb, &cb);  // sin tc = cos b1, cos tc = sin b1
  const double smin = cb * P.polar_cos - fabs(sb) * P.polar_sin;
  if (!(smin > 0.0)) return;
  const double s = P.polar_half / sqrt(cb * smin);
  if (!(s < 0.5)) return;
  const double dt = (4.0 / kPi) * s / sqrt(1.0 - s * s) * (1.0 + 1e-6) + 1e-9;
  const double tc = fmodulo4(l1 * kInvHalfPi);
  if (!(tc - dt >= 0.0 && tc + dt < 4.0)) return;
  w.range.tlo = (float)(tc - dt);
  w.range.thi = (float)(tc + dt);
}

__device__ inline Window window_of(const GridParams& P, uint32_t firstring, uint32_t usedrings, double xa_deg,
                                   double ya_deg) {
  Window w;
  const double l1 = xa_deg * kDeg2Rad;
  const double b1 = ya_deg * kDeg2Rad;
  const double disc_theta = kHalfPi - b1;
  double utheta = disc_theta - P.disc_size;
  if (utheta * kRad2Deg < 0) utheta = 0;
  uint64_t uring = pix2ring(P.hp, ang2pix(P.hp, utheta, l1));
  utheta = disc_theta + P.disc_size;
  if (P.selection == HCG_SELECT_DISC && utheta > kPi) utheta = kPi;  // the reference's MAX_Y key
  uint64_t dring = pix2ring(P.hp, ang2pix(P.hp, utheta, l1));
  if (P.selection == HCG_SELECT_DISC) {
    // a point lies in one of the two rings whose centres bracket it, so two
    // rings either side cover every sample with |ts - tc| < R <= disc_size
    uring = uring > 2 ? uring - 2 : 1;
    dring += 2;
    w.mode = kWinDisc;
  } else {
    w.mode = ((disc_theta * kRad2Deg <= kNorthB) || (disc_theta * kRad2Deg >= kSouthB)) ? kWinPolar : kWinCompat;
  }
  if (uring < firstring) uring = firstring;
  if (dring >= (uint64_t)firstring + usedrings) dring = (uint64_t)firstring + usedrings - 1;
  w.uring = (uint32_t)uring;
  w.dring = (uint32_t)dring;
  if (w.mode == kWinCompat)
    w.phi = l1;
  else
    phi_range(P, b1, l1, w);
  return w;
}

// The first used ring's samples are never gridded by the reference (the
// start_int - 1 wrap, HCGrid/gridding.cu:326, 363-367); HCG_SELECT_DISC grids them.
__device__ inline bool skips_first_ring(const GridParams& P) { return P.selection != HCG_SELECT_DISC; }

// lpix/rpix of ring r (HCGrid/gridding.cu:336-360).  Returns false where the
// reference would `continue` without advancing (an endless loop there); the
// ring is skipped here instead.  ang2pix(theta_r, phi) == loc2pix(cos theta_r, ...).
__device__ inline bool ring_window(const GridParams& P, const Window& w, uint64_t ring, double rz, double rs,
                                   uint64_t& lpix, uint64_t& rpix) {
  uint64_t sp, npr;
  ring_info(P.hp, ring, sp, npr);
  if (w.mode != kWinCompat) {
    lpix = sp;
    rpix = sp + npr - 1;
    // The pixel of longitude tt in ring r is sp + floor(tt npr / 4) in a
    // polar-cap ring (HCGrid/healpix.cu:126-205), and that or the next (mod
    // npr) in the others, whose pixels are diamonds.  Polar windows narrow
    // the cap rings only (ring < nside or > 3 nside), with one pixel of
    // margin; disc windows narrow every ring, with two.  (float: tt npr / 4
    // < 32768 is good to 0.004 pixel.)
    const bool cap = ring < P.hp.nside || ring > 3 * (uint64_t)P.hp.nside;
    if (w.range.tlo <= w.range.thi && (cap || w.mode == kWinDisc)) {
      const int32_t m = w.mode == kWinDisc ? 2 : 1;
      const float q = (float)(uint32_t)(npr >> 2);
      const int32_t lo = (int32_t)floorf(w.range.tlo * q) - m;
      const int32_t hi = (int32_t)floorf(w.range.thi * q) + m;
      if (lo >= 0 && hi < (int32_t)npr) {
        lpix = sp + (uint64_t)lo;
        rpix = sp + (uint64_t)hi;
      }
    }
    return true;
  }
  const bool have = rs >= 0.0;
  lpix = loc2pix(P.hp, rz, w.phi - P.disc_size, have ? rs : 0., have);
  if (!(lpix >= sp && lpix < sp + npr)) return false;
  rpix = loc2pix(P.hp, rz, w.phi + P.disc_size, have ? rs : 0., have);
  if (!(rpix >= sp && rpix < sp + npr)) return false;
  return true;
}

// ---- out-of-line slow paths (rare; kept out of the inner loop's registers) ----

// The reference's radius test (HCGrid/helpers.cu:128-130, gridding.cu:381-386)
// for a pair inside the guard band.
[[maybe_unused]] static __device__ __noinline__ bool radius_test_ref(const double* xwcs, const double* ywcs, const double* lon_s,
                                             const double* lat_s, uint32_t cell, uint32_t i, double radius) {
  const double l1 = xwcs[cell] * kDeg2Rad, b1 = ywcs[cell] * kDeg2Rad;
  const double l2 = lon_s[i] * kDeg2Rad, b2 = lat_s[i] * kDeg2Rad;
  const double a = sin((b1 - b2) / 2.);
  const double c = sin((l1 - l2) / 2.);
  const double d = 2 * asin(sqrt(a * a + cos(b1) * cos(b2) * (c * c))) * kRad2Deg;
  return d < radius;
}

[[maybe_unused]] static __device__ __noinline__ double asin2_direct(double h) {
  const double a = asin(sqrt(h));
  return a * a;
}

__device__ inline double asin2_of_h(double h) {  // asin(sqrt(h))^2
  if (h < 1e-4) {
    const double t = fma(h, 4.0 / 35.0, 8.0 / 45.0);
    const double u = fma(h, t, 1.0 / 3.0);
    return h * fma(h, u, 1.0);
  }
  return asin2_direct(h);
}

// GAUSS1D exponent -p0 d^2 (d in degrees) scaled by 64/ln2, from c2 = 4 sin^2(d/2):
// d^2 = (2 RAD2DEG)^2 asin^2(sqrt(c2)/2) and asin^2(sqrt h) = h (1 + h/3 + 8h^2/45 + 4h^3/35 + ...),
// i.e. (c2/4)(1 + c2/12 + c2^2/90 + c2^3/560), truncation < 1e-17 relative for c2 < 4e-4.
__device__ inline double gauss_exponent_scaled(const GridParams& P, double c2) {
  if (c2 < 4e-4) return c2 * fma(c2, fma(c2, fma(c2, P.k3, P.k2), P.k1), P.k0);
  return P.k_asin * asin2_direct(0.25 * c2);
}

// The series branch of gauss_exponent_scaled alone, for grids whose radius
// keeps every selected c2 below 4e-4 (R < ~1.1 deg): the same operations.
__device__ inline double gauss_exponent_series(const GridParams& P, double c2) {
  return c2 * fma(c2, fma(c2, fma(c2, P.k3, P.k2), P.k1), P.k0);
}

// True when every selected pair (c2 < c_hi) takes the series branch and its
// scaled exponent stays above exp_scaled_neg's clamp (host side).
// 16-bit values (BF16 or FP16) to and from FP32, round to nearest even.
__device__ inline uint16_t round16(float f, bool bf) {
  return bf ? __builtin_bit_cast(uint16_t, (__bf16)f) : __builtin_bit_cast(uint16_t, (_Float16)f);
}

__device__ inline float widen16(uint16_t h, bool bf) {
  return bf ? (float)__builtin_bit_cast(__bf16, h) : (float)__builtin_bit_cast(_Float16, h);
}

inline bool gauss_small_radius(const GridParams& P) {
  if (!(P.c_hi < 4e-4)) return false;
  const double y = P.c_hi * (P.k0 + P.c_hi * (P.k1 + P.c_hi * (P.k2 + P.c_hi * P.k3)));
  return y > -90000.0;
}

// The tile loop's candidate test on p.q (c2 = 2 h, h = 1 - p.q by three FMAs
// from 1, instead of seven operations for |p - q|^2): the FMAs' two
// roundings of O(1) partial sums and the unit vectors' own rounding leave
// c2 within ~1.2e-15 absolute,
// which at c_hi >= 1e-7 (R >= ~0.018 deg) is <= 1.2e-8 of the threshold, so
// a weight exp(-p0 d^2) (p0 d^2 <= 4.5 at R = 3 sigma) moves by <= ~5e-8
// relative; smaller radii keep |p - q|^2.  Selection is exact either way
// (the guard band covers the error and the reference's formula decides).
inline bool dot_distance_ok(const GridParams& P) { return P.c_hi >= 1e-7; }

// gauss_small_radius and c_hi < 2e-5 (R < ~0.26 deg), where the series' c2^3
// term, (4/560) c2^3 relative to the leading one, is below 6e-17 (host side).
inline bool gauss_tiny_radius(const GridParams& P) { return gauss_small_radius(P) && P.c_hi < 2e-5; }

// The GAUSS1D weight of a selected pair from the exponent's series (W = 1:
// four terms, gauss_small_radius; W = 2: three terms, gauss_tiny_radius) and
// exp_scaled_neg_fold; steps with a parameter as one v_fma_f64 each.
template <int W>
__device__ inline double gauss_weight_series(const GridParams& P, double c2, const double* s_exp) {
  double t;
  if constexpr (W == 2)
    t = fma_vvs(c2, c2 * P.k2 + P.k1, P.k0);
  else
    t = fma_vvs(c2, fma_vvs(c2, c2 * P.k3 + P.k2, P.k1), P.k0);
  return exp_scaled_neg_fold(c2 * t, s_exp);
}

// A loop-invariant double moved into a vector register pair.  A VOP3
// instruction on gfx950 reads at most one scalar operand, so an FMA whose
// two constant operands both sit in scalar registers costs an extra move
// (or a mul + add); with one of them in vector registers it is one v_fma_f64.
__device__ inline double in_vgprs(double x) {
  double v;
  asm("v_mov_b64 %0, %1" : "=v"(v) : "s"(x));
  return v;
}

// Degree-3 coefficient of exp(u ln2/64) - 1 = u (e1 + u (e2 + u e3)).
constexpr double kExpE1 = 0x1.62e42fefa39efp-7;  // ln2 / 64
constexpr double kExpE2 = kExpE1 * kExpE1 / 2.0;
constexpr double kExpE3 = kExpE2 * kExpE1 / 3.0;

// The GAUSS1D weight of a selected pair for the p.q tile loop
// (gauss_tiny_radius && dot_distance_ok: c2 < 2e-5): the exponent's series
// to its second term, y = c2 (k0 + c2 k1) in units of ln2/64 (the c2^2
// term is below 5e-12 of the exponent there), and exp by the 64-entry table
// (Tang 1989) with the degree-3 polynomial of exp(r) - 1, |r| <= ln2/128
// (truncation r^4/24 < 3.6e-11 relative).  The weight is within 4e-11
// relative of exp(-p0 d^2).  The caller passes h = c2 / 2 with a = 2 k0
// and b = 4 k1 (y = h (a + b h) is the same value: every scaling is by a
// power of two); bv and e3v hold b and kExpE3 in vector registers
// (in_vgprs), so every step is one v_fma_f64.
__device__ inline double gauss_weight_lean(double h, double a, double bv, double e3v, const double* s_exp) {
  const double y = h * fma(h, bv, a);
  const double n = rint(y);
  const int ni = (int)n;
  const double u = y - n;
  double q = fma_vvs(u, e3v, kExpE2);  // three-operand v_fma_f64: no copy of a constant register
  q = fma_vvs(u, q, kExpE1);
  const double p = u * q;
  const double t = s_exp[ni & 63];
  return ldexp(fma(t, p, t), ni >> 6);
}

// GAUSS2D / TAPERED_SINC weights (HCGrid/gridding.cu:242-256).
template <int KT>
[[maybe_unused]] static __device__ __noinline__ double weight_general(double p0, double p1, double p2, const double* xwcs,
                                              const double* ywcs, const double* lon_s, const double* lat_s,
                                              double h, uint32_t cell, uint32_t i) {
  const double sdist = 2.0 * sqrt(asin2_of_h(h)) * kRad2Deg;
  const double p[3] = {p0, p1, p2};
  if constexpr (KT == HCG_GAUSS2D) {
    const double l1 = xwcs[cell] * kDeg2Rad, b1 = ywcs[cell] * kDeg2Rad;
    const double l2 = lon_s[i] * kDeg2Rad, b2 = lat_s[i] * kDeg2Rad;
    const double dl = l2 - l1, cb2 = cos(b2);
    const double bear = atan2(cb2 * sin(dl), cos(b1) * sin(b2) - sin(b1) * cb2 * cos(dl));
    const double sb = sin(bear - p[2]), cbr = cos(bear - p[2]);
    const double ell = p[0] * p[0] * (sb * sb) + p[1] * p[1] * (cbr * cbr);
    const double q = sdist / p[0] / p[1];
    return exp(-(q * q / 2. * ell));
  } else {
    const double arg = kPi * sdist / p[0];
    const double t = arg / p[1];
    const double x = arg / p[2];
    const double sinc = (fabs(x) < 1.e-10) ? 1. : sin(x) / x;  // HCGrid/gridding.cu:230-235
    return sinc * exp(t * t);  // exp(pow(-(arg/p1), 2)) as the reference writes it
  }
}

// ---- wave-level helpers ----

// 64-way search over keys[lo, hi): first index whose key >= K (UPPER=false) or
// > K (UPPER=true).  Wave-uniform inputs and result.
template <bool UPPER>
__device__ inline uint32_t wave_search(const uint32_t* __restrict__ keys, uint32_t lo, uint32_t hi, uint32_t K) {
  const uint32_t lane = __lane_id();
  while (hi - lo > 64) {
    const uint32_t step = (hi - lo + 63) / 64;
    const uint32_t p = lo + lane * step;
    const bool inr = p < hi;
    const uint32_t kv = inr ? keys[p] : 0xFFFFFFFFu;
    const bool ge = inr ? (UPPER ? kv > K : kv >= K) : true;
    const uint64_t m = __ballot(ge);
    if (m == 0) {  // all 64 probes in range and below K: the answer is past the last probe
      lo = lo + 63 * step + 1;
      continue;
    }
    const uint32_t f = (uint32_t)__ffsll((unsigned long long)m) - 1;
    const uint32_t pf = lo + f * step;
    if (f == 0) return lo;
    const uint32_t nlo = lo + (f - 1) * step + 1;
    hi = pf < hi ? pf : hi;
    lo = nlo;
  }
  const uint32_t p = lo + lane;
  const bool inr = p < hi;
  const uint32_t kv = inr ? keys[p] : 0u;
  const bool ge = inr && (UPPER ? kv > K : kv >= K);
  const uint64_t m = __ballot(ge);
  return m ? lo + (uint32_t)__ffsll((unsigned long long)m) - 1 : hi;
}

__device__ inline uint32_t wave_min_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const uint32_t o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

// Kernel-argument form of the set-up (host side).
inline GridParams make_grid_params(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t side) {
  GridParams P;
  P.hp = *hp;
  P.disc_size = kern->disc_size;
  P.sphere_radius = kern->sphere_radius;
  for (int i = 0; i < 3; ++i) P.params[i] = kern->params[i];
  P.side = side;
  P.selection = kern->selection;
  P.factor = kern->selection == HCG_SELECT_DISC ? 1u : kern->factor;
  // radius threshold on c2 = 4 sin^2(R/2), with a guard band where the
  // reference's haversine decides: 1e-9 relative (c2 = |p - q|^2 is good to
  // ~1e-12 relative) plus 4e-15 absolute (c2 = 2 (1 - p.q), the form the
  // tile loop uses at dot_distance_ok radii, is good to ~1.2e-15 absolute)
  const double half = kern->sphere_radius * kDeg2Rad / 2.0;
  const double cR = 4.0 * sin(half) * sin(half);
  P.c_lo = cR * (1.0 - 1e-9) - 4e-15;
  P.c_hi = cR * (1.0 + 1e-9) + 4e-15;
  // -p0 d^2 = -p0 (2 RAD2DEG)^2 asin^2(sqrt(c2)/2); in units of ln2/64
  const double inv_ln2_64 = 0x1.71547652b82fep+6;
  P.k_asin = -4.0 * kRad2Deg * kRad2Deg * kern->params[0] * inv_ln2_64;
  P.k0 = P.k_asin / 4.0;
  P.k1 = P.k0 / 12.0;
  P.k2 = P.k0 / 90.0;
  P.k3 = P.k0 / 560.0;
  const double rw = kern->sphere_radius * kDeg2Rad * (1.0 + 1e-6);
  P.polar_cos = cos(rw);
  P.polar_sin = sin(rw);
  P.polar_half = sin(0.5 * rw);
  return P;
}

}  // namespace hcg

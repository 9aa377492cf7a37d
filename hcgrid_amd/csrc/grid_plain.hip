// grid_plain.hip -- the single-channel gather-convolution on gfx950.
//
// Replaces the `hcgrid` kernel (HCGrid/gridding.cu:277-403).  Selection is
// the reference's, per window anchor a (first cell of a `factor` group):
//   rings r in [uring(a), dring(a)], r != firstring (the reference's
//   start_int-1 wrap skips the first used ring, gridding.cu:326,363-367),
//   sorted samples with lpix_r(a) <= key <= rpix_r(a) (whole ring when the
//   anchor colatitude is <= 20 or >= 160 deg, gridding.cu:341-356),
//   great-circle distance from the cell itself < sphere_radius.
// Accumulation order per cell is the reference's: ring ascending, sorted
// index ascending, FP64.
//
// What differs from the reference's arithmetic (not its selection):
//  * ring windows use the per-ring cos/sin table built with the LUT instead
//    of recomputing pix2ang + cos per cell and ring (same values);
//  * the end-of-window test reads the sorted key instead of recomputing
//    ang2pix for every candidate (gridding.cu:373; equal by construction);
//  * distance: h = |p - q|^2 / 4 = sin^2(d/2) from precomputed unit vectors
//    (the haversine quantity of helpers.cu:128-130, exact in real
//    arithmetic, ~1e-12 relative in FP64).  Pairs with h within 1e-9
//    (relative) of the radius threshold are decided with the reference's
//    haversine formula itself; d^2 for the Gaussian comes from the series
//    4 asin^2(sqrt h) = 4 (h + h^2/3 + 8h^3/45 + 4h^4/35 + ...), exact to
//    <1e-17 relative for h < 1e-4 (radius < 1.1 deg), asin otherwise.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "healpix.hpp"

namespace hcg {

struct GridParams {
  hcg_healpix hp;
  hcg_kernel kern;
  uint32_t side;
  uint32_t factor;
  uint64_t ncells;
  double h_lo, h_hi;   // guard band around the radius threshold on h = sin^2(d/2)
  double gauss_a;      // 4 * RAD2DEG^2 * p0: exp(-gauss_a * asin^2(sqrt h)) for GAUSS1D
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

// The cell's window (HCGrid/gridding.cu:306-323).
struct Window {
  uint64_t uring, dring;
  double phi;   // anchor longitude, radians
  bool polar;
};

__device__ inline Window window_of(const GridParams& P, uint32_t firstring, uint32_t usedrings, double xa_deg,
                                   double ya_deg) {
  Window w;
  const double l1 = xa_deg * kDeg2Rad;
  const double b1 = ya_deg * kDeg2Rad;
  const double disc_theta = kHalfPi - b1;
  double utheta = disc_theta - P.kern.disc_size;
  if (utheta * kRad2Deg < 0) utheta = 0;
  uint64_t uring = pix2ring(P.hp, ang2pix(P.hp, utheta, l1));
  if (uring < firstring) uring = firstring;
  utheta = disc_theta + P.kern.disc_size;
  uint64_t dring = pix2ring(P.hp, ang2pix(P.hp, utheta, l1));
  if (dring >= (uint64_t)firstring + usedrings) dring = (uint64_t)firstring + usedrings - 1;
  w.uring = uring;
  w.dring = dring;
  w.phi = l1;
  w.polar = (disc_theta * kRad2Deg <= kNorthB) || (disc_theta * kRad2Deg >= kSouthB);
  return w;
}

// lpix/rpix of ring r (HCGrid/gridding.cu:336-360).  Returns false where the
// reference would `continue` without advancing (an endless loop there); the
// ring is skipped here instead.  ang2pix(theta_r, phi) == loc2pix(cos theta_r, ...).
__device__ inline bool ring_window(const GridParams& P, const Window& w, uint64_t ring, double rz, double rs,
                                   uint64_t& lpix, uint64_t& rpix) {
  uint64_t sp, npr;
  ring_info(P.hp, ring, sp, npr);
  if (w.polar) {
    lpix = sp;
    rpix = sp + npr - 1;
    return true;
  }
  const bool have = rs >= 0.0;
  lpix = loc2pix(P.hp, rz, w.phi - P.kern.disc_size, have ? rs : 0., have);
  if (!(lpix >= sp && lpix < sp + npr)) return false;
  rpix = loc2pix(P.hp, rz, w.phi + P.kern.disc_size, have ? rs : 0., have);
  if (!(rpix >= sp && rpix < sp + npr)) return false;
  return true;
}

// The reference's distance in degrees (HCGrid/helpers.cu:128-130).
__device__ inline double ref_distance_deg(double l1, double b1, double l2, double b2) {
  const double a = sin((b1 - b2) / 2.);
  const double c = sin((l1 - l2) / 2.);
  return 2 * asin(sqrt(a * a + cos(b1) * cos(b2) * (c * c))) * kRad2Deg;
}

__device__ inline double asin2_of_h(double h) {  // asin(sqrt(h))^2
  if (h < 1e-4) {
    const double t = fma(h, 4.0 / 35.0, 8.0 / 45.0);
    const double u = fma(h, t, 1.0 / 3.0);
    return h * fma(h, u, 1.0);
  }
  const double a = asin(sqrt(h));
  return a * a;
}

__device__ inline double sinc_ref(double x) {  // HCGrid/gridding.cu:230-235
  return (fabs(x) < 1.e-10) ? 1. : sin(x) / x;
}

// Weight of a pair already inside the radius (HCGrid/gridding.cu:238-257).
template <int KT>
__device__ inline double pair_weight(const GridParams& P, double h, double l1, double b1, const GridBuffers& B,
                                     uint32_t i) {
  if constexpr (KT == HCG_GAUSS1D) {
    return exp(-P.gauss_a * asin2_of_h(h));
  } else {
    const double sdist = 2.0 * sqrt(asin2_of_h(h)) * kRad2Deg;
    const double* p = P.kern.params;
    if constexpr (KT == HCG_GAUSS2D) {
      const double l2 = B.lon_s[i] * kDeg2Rad, b2 = B.lat_s[i] * kDeg2Rad;
      const double dl = l2 - l1, cb2 = cos(b2);
      const double bear = atan2(cb2 * sin(dl), cos(b1) * sin(b2) - sin(b1) * cb2 * cos(dl));
      const double sb = sin(bear - p[2]), cbr = cos(bear - p[2]);
      const double ell = p[0] * p[0] * (sb * sb) + p[1] * p[1] * (cbr * cbr);
      const double q = sdist / p[0] / p[1];
      return exp(-(q * q / 2. * ell));
    } else {
      const double arg = kPi * sdist / p[0];
      const double t = arg / p[1];
      return sinc_ref(arg / p[2]) * exp(t * t);  // exp(pow(-(arg/p1), 2)) as the reference writes it
    }
  }
}

// One thread per cell; a 256-thread workgroup covers a 16x16 block of cells
// (wave w holds rows 4w..4w+3), so lanes of a wave share rings and read
// overlapping sample ranges through L1/L2.
template <int KT, bool HAS_W>
__global__ void __launch_bounds__(256) k_grid_cells(GridParams P, GridBuffers B) {
  const hcg_ring_meta m = *B.meta;
  if (m.usedrings == 0) return;
  const uint32_t tiles_x = (P.side + 15) / 16;
  const uint32_t tile = blockIdx.x;
  const uint32_t col = (tile % tiles_x) * 16 + (threadIdx.x & 15);
  const uint32_t row = (tile / tiles_x) * 16 + (threadIdx.x >> 4);
  if (col >= P.side || row >= P.side) return;
  const uint64_t c = (uint64_t)row * P.side + col;
  const uint64_t anchor = (uint64_t)row * P.side + (col / P.factor) * P.factor;

  const Window w = window_of(P, m.firstring, m.usedrings, B.xwcs[anchor], B.ywcs[anchor]);
  const double l1 = B.xwcs[c] * kDeg2Rad, b1 = B.ywcs[c] * kDeg2Rad;
  double sb1, cb1, sl1, cl1;
  sincos(b1, &sb1, &cb1);
  sincos(l1, &sl1, &cl1);
  const double px = cb1 * cl1, py = cb1 * sl1, pz = sb1;
  double acc_d = B.dc[c], acc_w = B.wc[c];

  uint64_t r0 = w.uring;
  if (r0 == m.firstring) ++r0;  // first used ring contributes nothing
  for (uint64_t r = r0; r <= w.dring; ++r) {
    const uint32_t k = (uint32_t)(r - m.firstring);
    const uint32_t start = B.lut[k], end = B.lut[k + 1];
    if (start == end) continue;
    uint64_t lpix, rpix;
    if (!ring_window(P, w, r, B.ring_z[k], B.ring_s[k], lpix, rpix)) continue;
    if (lpix > rpix) continue;
    uint32_t i = lower_bound_u32(B.keys, start, end, (uint32_t)lpix);
    for (; i < end; ++i) {
      if (B.keys[i] > (uint32_t)rpix) break;
      const hcg_sample q = B.rec[i];
      const double dx = px - q.x, dy = py - q.y, dz = pz - q.z;
      const double h = 0.25 * fma(dx, dx, fma(dy, dy, dz * dz));
      bool in = h < P.h_lo;
      if (!in && h < P.h_hi)
        in = ref_distance_deg(l1, b1, B.lon_s[i] * kDeg2Rad, B.lat_s[i] * kDeg2Rad) < P.kern.sphere_radius;
      if (in) {
        double wt = pair_weight<KT>(P, h, l1, b1, B, i);
        if constexpr (HAS_W) wt = B.w_s[i] * wt;
        acc_d = fma(q.v, wt, acc_d);
        acc_w += wt;
      }
    }
  }
  B.dc[c] = acc_d;
  B.wc[c] = acc_w;
}

template <int KT, bool HAS_W>
int launch_grid(const GridParams& P, const GridBuffers& B, hipStream_t s) {
  const uint32_t tiles_x = (P.side + 15) / 16;
  const uint64_t blocks = (uint64_t)tiles_x * tiles_x;
  if (blocks == 0) return HCG_OK;
  k_grid_cells<KT, HAS_W><<<(uint32_t)blocks, 256, 0, s>>>(P, B);
  HCG_CHECK_LAUNCH("k_grid_cells");
  return HCG_OK;
}

}  // namespace hcg

using namespace hcg;

extern "C" int hcg_grid_plain(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t side, uint32_t n,
                              const hcg_sample* records, const uint32_t* keys_sorted, const double* lon_sorted,
                              const double* lat_sorted, const double* weights_sorted, const uint32_t* lut,
                              const double* ring_z, const double* ring_s, const hcg_ring_meta* meta,
                              const double* xwcs, const double* ywcs, double* datacube, double* weightscube,
                              void* stream) {
  HCG_REQUIRE(hp && kern, "hcg_grid_plain: NULL hp/kern");
  if (hp->nside == 0 || hp->nside > HCG_MAX_NSIDE)
    return set_error(HCG_ERR_UNSUPPORTED, "hcg_grid_plain: nside must be <= 8192");
  HCG_REQUIRE(kern->kernel_type >= HCG_GAUSS1D && kern->kernel_type <= HCG_TAPERED_SINC,
              "hcg_grid_plain: unknown kernel type");
  HCG_REQUIRE(kern->factor >= 1, "hcg_grid_plain: factor must be >= 1");
  HCG_REQUIRE(side <= 65535, "hcg_grid_plain: side must be <= 65535");
  if (side == 0) return HCG_OK;
  HCG_REQUIRE(records && keys_sorted && lon_sorted && lat_sorted && lut && ring_z && ring_s && meta && xwcs &&
                  ywcs && datacube && weightscube,
              "hcg_grid_plain: NULL buffer");
  (void)n;
  GridParams P;
  P.hp = *hp;
  P.kern = *kern;
  P.side = side;
  P.factor = kern->factor;
  P.ncells = (uint64_t)side * side;
  const double half = kern->sphere_radius * kDeg2Rad / 2.0;
  const double hR = sin(half) * sin(half);
  P.h_lo = hR * (1.0 - 1e-9);
  P.h_hi = hR * (1.0 + 1e-9);
  P.gauss_a = 4.0 * kRad2Deg * kRad2Deg * kern->params[0];
  GridBuffers B{records, keys_sorted, lon_sorted, lat_sorted, weights_sorted, lut, ring_z, ring_s, meta,
                xwcs,    ywcs,        datacube,   weightscube};
  hipStream_t s = as_stream(stream);
  const bool hw = weights_sorted != nullptr;
  switch (kern->kernel_type) {
    case HCG_GAUSS1D: return hw ? launch_grid<HCG_GAUSS1D, true>(P, B, s) : launch_grid<HCG_GAUSS1D, false>(P, B, s);
    case HCG_GAUSS2D: return hw ? launch_grid<HCG_GAUSS2D, true>(P, B, s) : launch_grid<HCG_GAUSS2D, false>(P, B, s);
    default: return hw ? launch_grid<HCG_TAPERED_SINC, true>(P, B, s) : launch_grid<HCG_TAPERED_SINC, false>(P, B, s);
  }
}

namespace hcg {
__global__ void __launch_bounds__(256) k_normalize(uint64_t n, const double* __restrict__ dc,
                                                   const double* __restrict__ wc, double* __restrict__ map) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) map[i] = dc[i] / wc[i];
}
}  // namespace hcg

extern "C" int hcg_normalize(uint64_t ncells, const double* datacube, const double* weightscube, double* map,
                             void* stream) {
  if (ncells == 0) return HCG_OK;
  HCG_REQUIRE(datacube && weightscube && map, "hcg_normalize: NULL buffer");
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((ncells + 255) / 256, 8192);
  k_normalize<<<blocks, 256, 0, as_stream(stream)>>>(ncells, datacube, weightscube, map);
  HCG_CHECK_LAUNCH("k_normalize");
  return HCG_OK;
}

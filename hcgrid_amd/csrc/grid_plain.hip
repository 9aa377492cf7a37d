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

// ---------------------------------------------------------------------------
// Wave-tiled kernel: one 64-lane workgroup per 16 x 8 tile of cells.  Lane l
// owns column (l & 15) and two vertically adjacent cells in rows
// 2*(l >> 4) + {0, 1}.  The wave walks the union of its cells' ring ranges in
// ascending order; per non-empty ring it
//   1. evaluates every cell's [lpix, rpix] window (per-ring cos/sin table),
//   2. finds the sorted-index range of the union window with a 64-way
//      cooperative search of the keys (3 coalesced probes instead of ~15
//      dependent loads),
//   3. stages that range (records as x/y/z/v columns + keys) in LDS with
//      coalesced loads, chunk by chunk,
//   4. lets every lane locate its cells' sub-ranges by binary search in LDS
//      and accumulate over its union sub-range, two cells per record read.
// Cells in a column share their longitude window, so lanes of one column
// read the same LDS words (broadcast) and columns are ~11 samples apart,
// which spreads the 16 distinct addresses over the banks.
// ---------------------------------------------------------------------------
constexpr int kTileCols = 16;
constexpr int kTileRows = 8;
constexpr int kCellsPerLane = 2;
constexpr uint32_t kStageCap = 384;

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

template <int KT, bool HAS_W>
__global__ void __launch_bounds__(64) k_grid_tiles(GridParams P, GridBuffers B, uint32_t tiles_x,
                                                   uint32_t tiles_y) {
  __shared__ double s_x[kStageCap], s_y[kStageCap], s_z[kStageCap], s_v[kStageCap];
  __shared__ uint32_t s_key[kStageCap];

  const hcg_ring_meta m = *B.meta;
  if (m.usedrings == 0) return;
  // XCD-aware order: consecutive tiles land on the same XCD (workgroups are
  // dealt round-robin over the 8 XCDs), so neighbouring tiles share L2.
  const uint32_t ntiles = tiles_x * tiles_y;
  uint32_t tile = blockIdx.x;
  {
    const uint32_t per = (ntiles + 7) / 8;
    const uint32_t t2 = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if ((ntiles % 8) == 0) tile = t2;
  }
  const uint32_t lane = __lane_id();
  const uint32_t col = (tile % tiles_x) * kTileCols + (lane & 15);
  const uint32_t row0 = (tile / tiles_x) * kTileRows + (lane >> 4) * kCellsPerLane;

  // per-cell state
  bool live[kCellsPerLane];
  uint64_t cell[kCellsPerLane];
  double px[kCellsPerLane], py[kCellsPerLane], pz[kCellsPerLane], l1[kCellsPerLane], b1[kCellsPerLane];
  double acc_d[kCellsPerLane], acc_w[kCellsPerLane];
  Window win[kCellsPerLane];
  uint32_t lane_ur = 0xFFFFFFFFu, lane_dr = 0;
#pragma unroll
  for (int k = 0; k < kCellsPerLane; ++k) {
    const uint32_t row = row0 + k;
    live[k] = (col < P.side) && (row < P.side);
    const uint32_t rr = live[k] ? row : 0, cc = live[k] ? col : 0;
    cell[k] = (uint64_t)rr * P.side + cc;
    const uint64_t anchor = (uint64_t)rr * P.side + (cc / P.factor) * P.factor;
    win[k] = window_of(P, m.firstring, m.usedrings, B.xwcs[anchor], B.ywcs[anchor]);
    l1[k] = B.xwcs[cell[k]] * kDeg2Rad;
    b1[k] = B.ywcs[cell[k]] * kDeg2Rad;
    double sb, cb, sl, cl;
    sincos(b1[k], &sb, &cb);
    sincos(l1[k], &sl, &cl);
    px[k] = cb * cl;
    py[k] = cb * sl;
    pz[k] = sb;
    acc_d[k] = live[k] ? B.dc[cell[k]] : 0.0;
    acc_w[k] = live[k] ? B.wc[cell[k]] : 0.0;
    if (live[k] && win[k].uring <= win[k].dring) {
      lane_ur = min(lane_ur, (uint32_t)win[k].uring);
      lane_dr = max(lane_dr, (uint32_t)win[k].dring);
    }
  }
  uint32_t r_begin = wave_min_u32(lane_ur);
  const uint32_t r_end = wave_max_u32(lane_dr);
  if (r_begin == m.firstring) ++r_begin;  // the first used ring contributes nothing

  for (uint32_t r = r_begin; r <= r_end && r_begin <= r_end; ++r) {
    const uint32_t kr = r - m.firstring;
    const uint32_t start = B.lut[kr], end = B.lut[kr + 1];
    if (start == end) continue;
    const double rz = B.ring_z[kr], rs = B.ring_s[kr];
    uint32_t lp[kCellsPerLane], rp[kCellsPerLane];
    bool act[kCellsPerLane];
    uint32_t lane_lp = 0xFFFFFFFFu, lane_rp = 0;
#pragma unroll
    for (int k = 0; k < kCellsPerLane; ++k) {
      act[k] = live[k] && r >= win[k].uring && r <= win[k].dring;
      uint64_t a = 1, b = 0;
      if (act[k]) act[k] = ring_window(P, win[k], r, rz, rs, a, b) && a <= b;
      lp[k] = (uint32_t)a;
      rp[k] = (uint32_t)b;
      if (act[k]) {
        lane_lp = min(lane_lp, lp[k]);
        lane_rp = max(lane_rp, rp[k]);
      }
    }
    const uint32_t LP = wave_min_u32(lane_lp), RP = wave_max_u32(lane_rp);
    if (LP > RP) continue;
    const uint32_t lo_u = wave_search<false>(B.keys, start, end, LP);
    const uint32_t hi_u = wave_search<true>(B.keys, lo_u, end, RP);
    for (uint32_t cb = lo_u; cb < hi_u; cb += kStageCap) {
      const uint32_t cnt = min(kStageCap, hi_u - cb);
      __syncthreads();  // previous chunk fully consumed
      for (uint32_t j = lane; j < cnt; j += 64) {
        const hcg_sample q = B.rec[cb + j];
        s_x[j] = q.x;
        s_y[j] = q.y;
        s_z[j] = q.z;
        s_v[j] = q.v;
        s_key[j] = B.keys[cb + j];
      }
      __syncthreads();
      uint32_t lo[kCellsPerLane], hi[kCellsPerLane];
      uint32_t jb = 0xFFFFFFFFu, je = 0;
#pragma unroll
      for (int k = 0; k < kCellsPerLane; ++k) {
        lo[k] = 0;
        hi[k] = 0;
        if (act[k]) {
          lo[k] = lower_bound_u32(s_key, 0, cnt, lp[k]);
          hi[k] = upper_bound_u32(s_key, lo[k], cnt, rp[k]);
          if (lo[k] < hi[k]) {
            jb = min(jb, lo[k]);
            je = max(je, hi[k]);
          }
        }
      }
      for (uint32_t j = jb; j < je; ++j) {
        const double qx = s_x[j], qy = s_y[j], qz = s_z[j], qv = s_v[j];
#pragma unroll
        for (int k = 0; k < kCellsPerLane; ++k) {
          const double dx = px[k] - qx, dy = py[k] - qy, dz = pz[k] - qz;
          const double h = 0.25 * fma(dx, dx, fma(dy, dy, dz * dz));
          const bool band = j >= lo[k] && j < hi[k];
          bool in = band && h < P.h_lo;
          if (band && !in && h < P.h_hi) {
            const uint32_t gi = cb + j;
            in = ref_distance_deg(l1[k], b1[k], B.lon_s[gi] * kDeg2Rad, B.lat_s[gi] * kDeg2Rad) <
                 P.kern.sphere_radius;
          }
          if (in) {
            double wt = pair_weight<KT>(P, h, l1[k], b1[k], B, cb + j);
            if constexpr (HAS_W) wt = B.w_s[cb + j] * wt;
            acc_d[k] = fma(qv, wt, acc_d[k]);
            acc_w[k] += wt;
          }
        }
      }
    }
  }
#pragma unroll
  for (int k = 0; k < kCellsPerLane; ++k) {
    if (live[k]) {
      B.dc[cell[k]] = acc_d[k];
      B.wc[cell[k]] = acc_w[k];
    }
  }
}

template <int KT, bool HAS_W>
int launch_grid(const GridParams& P, const GridBuffers& B, hipStream_t s) {
  const uint32_t tiles_x = (P.side + kTileCols - 1) / kTileCols;
  const uint32_t tiles_y = (P.side + kTileRows - 1) / kTileRows;
  const uint64_t blocks = (uint64_t)tiles_x * tiles_y;
  if (blocks == 0) return HCG_OK;
  k_grid_tiles<KT, HAS_W><<<(uint32_t)blocks, 64, 0, s>>>(P, B, tiles_x, tiles_y);
  HCG_CHECK_LAUNCH("k_grid_tiles");
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

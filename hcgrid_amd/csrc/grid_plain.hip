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
//    arithmetic, ~1e-12 relative in FP64); for GAUSS1D at radii with
//    c2 threshold >= 1e-7 (dot_distance_ok) c2 = 2 - 2 p.q instead, which
//    tests a candidate in four FP64 operations rather than seven and is
//    good to ~1.2e-15 absolute (weights to ~5e-8 relative at worst).  Pairs
//    within the guard band of the radius threshold (1e-9 relative + 4e-15
//    absolute) are decided with the reference's haversine formula itself,
//    so selection is exact in both forms; d^2 for the Gaussian comes from the series
//    4 asin^2(sqrt h) = 4 (h + h^2/3 + 8h^3/45 + 4h^4/35 + ...), exact to
//    <1e-17 relative for h < 1e-4 (radius < 1.1 deg), asin otherwise.
//
// Work decomposition: one 64-lane workgroup per tile of kTileCols x
// (kLaneRows * kCellsPerLane) cells.  Lane l owns column l / kLaneRows and
// kCellsPerLane vertically adjacent cells in it (column-major lanes,
// plain_tile.hpp).  The wave walks the union of its
// cells' ring ranges in ascending order; per non-empty ring it
//   1. evaluates every cell's [lpix, rpix] window,
//   2. finds the sorted-index range of the union window with a 64-way
//      cooperative search of the keys (3 coalesced probes instead of ~15
//      dependent loads),
//   3. stages that range (records as x/y/z/v columns + keys) in LDS with
//      coalesced loads, chunk by chunk,
//   4. lets every lane locate its cells' sub-ranges by binary search in LDS
//      and accumulate over its union sub-range, kCellsPerLane cells per
//      record read.
// Cells in a column share their longitude window, so lanes of one column
// read the same LDS words (broadcast) and columns are ~11 samples apart,
// which spreads the distinct addresses over the banks.
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "expneg.hpp"
#include "healpix.hpp"

#include "plain_tile.hpp"

namespace hcg {

// WMODE (GAUSS1D): 0 general exponent and exp; 1 gauss_small_radius; 2 gauss_tiny_radius.
template <int KT, bool HAS_W, int WMODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(HCG_GRID_WAVES_PER_EU))) k_grid_tiles(GridParams P, GridBuffers B, uint32_t tiles_x,
                                                   uint32_t tiles_y) {
  __shared__ double s_x[kStageCap], s_y[kStageCap], s_z[kStageCap], s_v[kStageCap];
  __shared__ uint32_t s_key[kStageCap];
  __shared__ double s_exp[64];
  exp_table_load(s_exp);  // visible after the first staging barrier
  // XCD-aware order: workgroups are dealt round-robin over the 8 XCDs, so
  // block b runs on XCD b % 8; give each XCD a contiguous run of tiles so
  // neighbouring tiles (which stage overlapping samples) share its L2.
  const uint32_t ntiles = tiles_x * tiles_y;
  uint32_t tile = blockIdx.x;
  if ((ntiles & 7u) == 0) tile = (blockIdx.x & 7u) * (ntiles >> 3) + (blockIdx.x >> 3);
  grid_tile<KT, HAS_W, WMODE, false>(P, B, tile, tiles_x, s_x, s_y, s_z, s_v, s_key, s_exp, TileOut{},
                                     RingStream{0, 1}, nullptr, nullptr);
}

// A one-channel job of the multi-channel entry points on the tile path:
// 16-bit values in, FP32 cube and weight sum out; one tile per workgroup of
// S waves, one partial ring stream per wave (one_channel_tile).
template <uint32_t S>
__global__ void __launch_bounds__(64 * S) k_grid_tiles_1ch(GridParams P, GridBuffers B, TileOut O, uint32_t tiles_x,
                                                           uint32_t tiles_y) {
  __shared__ __attribute__((aligned(16))) char lds[one_channel_lds_bytes(S, S)];
  __shared__ double s_exp[64];
  if (threadIdx.x < 64) exp_table_load(s_exp);
  __syncthreads();
  const uint32_t ntiles = tiles_x * tiles_y;
  uint32_t tile = blockIdx.x;
  if ((ntiles & 7u) == 0) tile = (blockIdx.x & 7u) * (ntiles >> 3) + (blockIdx.x >> 3);
  one_channel_tile(P, B, O, tile, tiles_x, S, S, lds, s_exp);
}

// A few-channel job of the multi-channel entry points on the tile path: one
// wave per tile (few_channel_tile_body).
__global__ void __launch_bounds__(64) k_grid_tiles_few(GridParams P, GridBuffers B, TileOut O, uint32_t tiles_x,
                                                       uint32_t tiles_y) {
  __shared__ __attribute__((aligned(16))) char lds[kFewTileLdsBytes];
  __shared__ double s_exp[64];
  exp_table_load(s_exp);
  tile_sync<true>();
  const uint32_t ntiles = tiles_x * tiles_y;
  uint32_t tile = blockIdx.x;
  if ((ntiles & 7u) == 0) tile = (blockIdx.x & 7u) * (ntiles >> 3) + (blockIdx.x >> 3);
  few_channel_tile_body(P, B, O, tile, tiles_x, lds, s_exp);
}

int launch_tiles_few_channels(const GridParams& P, const GridBuffers& B, const TileOut& O, hipStream_t s) {
  const uint32_t tiles_x = tile_cols_of(P.side);
  const uint32_t tiles_y = (P.side + kTileRows - 1) / kTileRows;
  if ((uint64_t)tiles_x * tiles_y == 0) return HCG_OK;
  k_grid_tiles_few<<<tiles_x * tiles_y, 64, 0, s>>>(P, B, O, tiles_x, tiles_y);
  HCG_CHECK_LAUNCH("k_grid_tiles_few");
  return HCG_OK;
}

int launch_tiles_one_channel(const GridParams& P, const GridBuffers& B, const TileOut& O, hipStream_t s) {
  const uint32_t tiles_x = tile_cols_of(P.side);
  const uint32_t tiles_y = (P.side + kTileRows - 1) / kTileRows;
  if ((uint64_t)tiles_x * tiles_y == 0) return HCG_OK;
  if (one_channel_streams(P.side) == kOneChannelStreams)
    k_grid_tiles_1ch<kOneChannelStreams><<<tiles_x * tiles_y, 64 * kOneChannelStreams, 0, s>>>(P, B, O, tiles_x,
                                                                                              tiles_y);
  else
    k_grid_tiles_1ch<1><<<tiles_x * tiles_y, 64, 0, s>>>(P, B, O, tiles_x, tiles_y);
  HCG_CHECK_LAUNCH("k_grid_tiles_1ch");
  return HCG_OK;
}

// ---------------------------------------------------------------------------
// Row-segment kernel: a 256-thread workgroup owns kRowWaves consecutive map
// rows x 64 columns; wave w owns one row, lane l one cell.  Lanes of a wave
// lie in one row, so they need (nearly) the same rings and their windows in
// a ring have the same length, only shifted by ~11 samples per column: all
// lanes stay busy.  The waves of the workgroup share one LDS staging buffer
// per ring (the union window of all 4 x 64 cells), which keeps LDS per wave
// small enough for 4 resident waves per SIMD.  A wave with no cell in the
// current ring waits at the workgroup barrier, leaving its SIMD to others.
// ---------------------------------------------------------------------------
#ifndef HCG_GRID_ROW_WAVES
#define HCG_GRID_ROW_WAVES 4
#endif
#ifndef HCG_GRID_ROW_CAP
#define HCG_GRID_ROW_CAP 1024
#endif
constexpr int kRowWaves = HCG_GRID_ROW_WAVES;
constexpr uint32_t kRowCap = HCG_GRID_ROW_CAP;

template <int KT, bool HAS_W>
__global__ void __launch_bounds__(64 * kRowWaves) k_grid_rows(GridParams P, GridBuffers B, uint32_t segs_x,
                                                               uint32_t segs_y) {
  __shared__ double s_x[kRowCap], s_y[kRowCap], s_z[kRowCap], s_v[kRowCap];
  __shared__ uint32_t s_key[kRowCap];
  __shared__ double s_exp[64];
  __shared__ uint32_t s_red[2][kRowWaves][2];

  const hcg_ring_meta m = *B.meta;
  if (m.usedrings == 0) return;  // uniform over the workgroup
  const uint32_t nseg = segs_x * segs_y;
  uint32_t seg = blockIdx.x;
  if ((nseg & 7u) == 0) seg = (blockIdx.x & 7u) * (nseg >> 3) + (blockIdx.x >> 3);  // XCD-contiguous runs
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  if (wave == 0) exp_table_load(s_exp);
  const uint32_t col = (seg % segs_x) * 64 + lane;
  const uint32_t row = (seg / segs_x) * kRowWaves + wave;
  const bool live = col < P.side && row < P.side;

  const uint32_t cell = live ? row * P.side + col : 0;
  const uint32_t anchor = live ? row * P.side + (col / P.factor) * P.factor : 0;
  Window win = window_of(P, m.firstring, m.usedrings, B.xwcs[anchor], B.ywcs[anchor]);
  if (!live) {
    win.uring = 1;
    win.dring = 0;
  }
  double sb, cb0, sl, cl;
  sincos(B.ywcs[cell] * kDeg2Rad, &sb, &cb0);
  sincos(B.xwcs[cell] * kDeg2Rad, &sl, &cl);
  const double px = cb0 * cl, py = cb0 * sl, pz = sb;
  double acc_d = live ? B.dc[cell] : 0.0;
  double acc_w = live ? B.wc[cell] : 0.0;

  // ring span of the workgroup
  {
    const bool has = win.uring <= win.dring;
    const uint32_t ur = wave_min_u32(has ? win.uring : 0xFFFFFFFFu);
    const uint32_t dr = wave_max_u32(has ? win.dring : 0u);
    if (lane == 0) {
      s_red[0][wave][0] = ur;
      s_red[0][wave][1] = dr;
    }
  }
  __syncthreads();
  uint32_t r_begin = 0xFFFFFFFFu, r_end = 0;
#pragma unroll
  for (int w = 0; w < kRowWaves; ++w) {
    r_begin = min(r_begin, s_red[0][w][0]);
    r_end = max(r_end, s_red[0][w][1]);
  }
  if (r_begin == m.firstring && skips_first_ring(P)) ++r_begin;  // the first used ring contributes nothing
  uint32_t phase = 1;                     // s_red[0] was just read; the next write goes to [1]

  for (uint32_t r = r_begin; r <= r_end && r_begin <= r_end; ++r) {
    const uint32_t kr = r - m.firstring;
    const uint32_t start = B.lut[kr], end = B.lut[kr + 1];
    if (start == end) continue;  // uniform: empty ring
    uint64_t a = 1, b = 0;
    const bool act =
        r >= win.uring && r <= win.dring && ring_window(P, win, r, B.ring_z[kr], B.ring_s[kr], a, b) && a <= b;
    const uint32_t lp = act ? (uint32_t)a : 1u, rp = act ? (uint32_t)b : 0u;
    {
      const uint32_t wl = wave_min_u32(act ? lp : 0xFFFFFFFFu);
      const uint32_t wr = wave_max_u32(act ? rp : 0u);
      if (lane == 0) {
        s_red[phase][wave][0] = wl;
        s_red[phase][wave][1] = wr;
      }
    }
    __syncthreads();
    uint32_t LP = 0xFFFFFFFFu, RP = 0;
    bool wave_active = false;
#pragma unroll
    for (int w = 0; w < kRowWaves; ++w) {
      const uint32_t wl = s_red[phase][w][0], wr = s_red[phase][w][1];
      LP = min(LP, wl);
      RP = max(RP, wr);
      if (w == (int)wave) wave_active = wl <= wr;
    }
    phase ^= 1u;
    if (LP > RP) continue;  // uniform
    const uint32_t lo_u = wave_search<false>(B.keys, start, end, LP);
    const uint32_t hi_u = wave_search<true>(B.keys, lo_u, end, RP);
    for (uint32_t c0 = lo_u; c0 < hi_u; c0 += kRowCap) {
      const uint32_t cnt = min(kRowCap, hi_u - c0);
      __syncthreads();  // the previous chunk is consumed by every wave
      for (uint32_t j = threadIdx.x; j < cnt; j += 64 * kRowWaves) {
        const hcg_sample q = B.rec[c0 + j];
        s_x[j] = q.x;
        s_y[j] = q.y;
        s_z[j] = q.z;
        s_v[j] = q.v;
        s_key[j] = B.keys[c0 + j];
      }
      __syncthreads();
      if (!wave_active) continue;
      const uint32_t jb = act ? lower_bound_u32(s_key, 0, cnt, lp) : 0u;
      const uint32_t je = act ? upper_bound_u32(s_key, jb, cnt, rp) : 0u;
      for (uint32_t j = jb; j < je; ++j) {
        const double dx = px - s_x[j], dy = py - s_y[j], dz = pz - s_z[j];
        const double c2 = fma(dx, dx, fma(dy, dy, dz * dz));
        if (c2 < P.c_hi) {
          if (c2 < P.c_lo || radius_test_ref(B.xwcs, B.ywcs, B.lon_s, B.lat_s, cell, c0 + j, P.sphere_radius)) {
            double wt;
            if constexpr (KT == HCG_GAUSS1D)
              wt = exp_scaled_neg(gauss_exponent_scaled(P, c2), s_exp);
            else
              wt = weight_general<KT>(P.params[0], P.params[1], P.params[2], B.xwcs, B.ywcs, B.lon_s, B.lat_s,
                                      0.25 * c2, cell, c0 + j);
            if constexpr (HAS_W) wt = B.w_s[c0 + j] * wt;
            acc_d = fma(s_v[j], wt, acc_d);
            acc_w += wt;
          }
        }
      }
    }
  }
  if (live) {
    B.dc[cell] = acc_d;
    B.wc[cell] = acc_w;
  }
}

#ifndef HCG_GRID_KERNEL
#define HCG_GRID_KERNEL 1
#endif

template <int KT, bool HAS_W>
int launch_grid(const GridParams& P, const GridBuffers& B, hipStream_t s) {
  if (HCG_GRID_KERNEL == 1) {
    const uint32_t tiles_x = (P.side + kTileCols - 1) / kTileCols;
    const uint32_t tiles_y = (P.side + kTileRows - 1) / kTileRows;
    const uint64_t blocks = (uint64_t)tiles_x * tiles_y;
    if (blocks == 0) return HCG_OK;
    // GAUSS1D with a radius whose pairs all take the exponent's series branch
    // (and stay clear of the exp clamp): the forms without either check
    if (KT == HCG_GAUSS1D && HCG_GRID_SMALLR && gauss_tiny_radius(P) && dot_distance_ok(P) &&
        kCellsPerLane == 1 && !HCG_GRID_LANE_LOOP)
      k_grid_tiles<KT, HAS_W, (KT == HCG_GAUSS1D && kCellsPerLane == 1 && !HCG_GRID_LANE_LOOP) ? 3 : 2>
          <<<(uint32_t)blocks, 64, 0, s>>>(P, B, tiles_x, tiles_y);
    else if (KT == HCG_GAUSS1D && HCG_GRID_SMALLR && gauss_tiny_radius(P))
      k_grid_tiles<KT, HAS_W, 2><<<(uint32_t)blocks, 64, 0, s>>>(P, B, tiles_x, tiles_y);
    else if (KT == HCG_GAUSS1D && HCG_GRID_SMALLR && gauss_small_radius(P))
      k_grid_tiles<KT, HAS_W, 1><<<(uint32_t)blocks, 64, 0, s>>>(P, B, tiles_x, tiles_y);
    else
      k_grid_tiles<KT, HAS_W, 0><<<(uint32_t)blocks, 64, 0, s>>>(P, B, tiles_x, tiles_y);
    HCG_CHECK_LAUNCH("k_grid_tiles");
  } else {
    const uint32_t segs_x = (P.side + 63) / 64;
    const uint32_t segs_y = (P.side + kRowWaves - 1) / kRowWaves;
    const uint64_t blocks = (uint64_t)segs_x * segs_y;
    if (blocks == 0) return HCG_OK;
    k_grid_rows<KT, HAS_W><<<(uint32_t)blocks, 64 * kRowWaves, 0, s>>>(P, B, segs_x, segs_y);
    HCG_CHECK_LAUNCH("k_grid_rows");
  }
  return HCG_OK;
}

__global__ void __launch_bounds__(256) k_normalize(uint64_t n, const double* __restrict__ dc,
                                                   const double* __restrict__ wc, double* __restrict__ map) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) map[i] = dc[i] / wc[i];
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
  HCG_REQUIRE(kern->selection <= HCG_SELECT_DISC,
              "hcg_grid_plain: selection must be HCG_SELECT_COMPAT or HCG_SELECT_DISC");
  HCG_REQUIRE(side <= 65535, "hcg_grid_plain: side must be <= 65535");
  if (side == 0) return HCG_OK;
  HCG_REQUIRE(records && keys_sorted && lon_sorted && lat_sorted && lut && ring_z && ring_s && meta && xwcs &&
                  ywcs && datacube && weightscube,
              "hcg_grid_plain: NULL buffer");
  (void)n;
  GridParams P = make_grid_params(hp, kern, side);
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

extern "C" int hcg_normalize(uint64_t ncells, const double* datacube, const double* weightscube, double* map,
                             void* stream) {
  if (ncells == 0) return HCG_OK;
  HCG_REQUIRE(datacube && weightscube && map, "hcg_normalize: NULL buffer");
  const uint32_t blocks = (uint32_t)std::min<uint64_t>((ncells + 255) / 256, 8192);
  k_normalize<<<blocks, 256, 0, as_stream(stream)>>>(ncells, datacube, weightscube, map);
  HCG_CHECK_LAUNCH("k_normalize");
  return HCG_OK;
}

namespace hcg {
__global__ void __launch_bounds__(64) k_eval_exp_neg(uint32_t n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double s_exp[64];
  exp_table_load(s_exp);
  __syncthreads();
  const uint32_t i = blockIdx.x * 64 + threadIdx.x;
  if (i < n) out[i] = exp_neg(x[i], s_exp);
}
}  // namespace hcg

extern "C" int hcg_eval_exp_neg(uint32_t n, const double* x, double* out, void* stream) {
  if (n == 0) return HCG_OK;
  HCG_REQUIRE(x && out, "hcg_eval_exp_neg: NULL buffer");
  k_eval_exp_neg<<<ceil_div_u32(n, 64), 64, 0, as_stream(stream)>>>(n, x, out);
  HCG_CHECK_LAUNCH("k_eval_exp_neg");
  return HCG_OK;
}

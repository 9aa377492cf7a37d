// plain_tile.hpp -- one tile of the single-channel gather-convolution, the
// body of k_grid_tiles (grid_plain.hip), shared with the one- and
// few-channel jobs of the multi-channel entry points (grid_plain.hip's
// k_grid_tiles_1ch / k_grid_tiles_few, and the persistent batch kernel in
// grid_mfma.hip).
//
// A tile is kTileCols x kTileRows map cells worked by one wave: lane l owns
// column l / kLaneRows and, in it, kCellsPerLane vertically adjacent cells
// from lane row l % kLaneRows.  Column-major lanes put the four cells of a
// column (whose sample bands in a ring are nearly the same) next to each
// other: the 16 lanes an LDS read serves at once then fetch from four bands
// instead of sixteen, which cut LDS bank conflicts from 0.46 to 0.18 of
// LDS-active cycles and the 20M-sample 2048^2 map from 4.92 to 4.65 ms
// (row-major lanes; the sums are the same bit for bit, since each cell's
// accumulation order does not depend on its lane).  The selection,
// distance, weight and accumulation order are the ones described in
// grid_plain.hip.  Instantiations:
//  * BATCH false (k_grid_tiles): FP64 values from the sample records,
//    sums accumulated in place into the FP64 data and weight cubes (the
//    reference's datacube/weightscube), workgroup barriers (one wave per
//    workgroup);
//  * BATCH true (a one-channel job of the multi-channel entry points): the
//    job's sorted 16-bit values, one partial stream of the tile's rings
//    (every stride-th ring from the phase-th), sums from zero returned to
//    the caller, which adds the streams in stream order and writes the FP32
//    cube (divided by the weight sum when the job normalises; NaN where no
//    sample contributes) and FP32 weight sum (tile_write); wave-level
//    synchronisation only, since the other waves of the workgroup work
//    other tiles or streams;
//  * NC > 0 (few-channel jobs): as BATCH true, with one FP32 sum per channel
//    and the outputs written by the tile itself.
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"
#include "expneg.hpp"
#include "healpix.hpp"

#ifndef HCG_GRID_TILE_COLS
#define HCG_GRID_TILE_COLS 16
#endif
#ifndef HCG_GRID_CELLS_PER_LANE
#define HCG_GRID_CELLS_PER_LANE 1
#endif
#ifndef HCG_GRID_STAGE_CAP
#define HCG_GRID_STAGE_CAP 256
#endif
#ifndef HCG_GRID_WAVES_PER_EU
#define HCG_GRID_WAVES_PER_EU 4
#endif
#ifndef HCG_GRID_LANE_LOOP
// 0: the wave walks the union of its lanes' bands (broadcast LDS reads; lanes
// outside their band only evaluate the distance); 1: each lane walks its own
// band (per-lane LDS addresses).  Measured on the 20M raster: 6.69 ms vs 7.14.
#define HCG_GRID_LANE_LOOP 0
#endif
#ifndef HCG_GRID_SMALLR
#define HCG_GRID_SMALLR 1  // GAUSS1D at R < ~1.1 deg: no series/asin switch, no exp clamp in the inner loop
#endif
#ifndef HCG_GRID_TABLE_EXP
#define HCG_GRID_TABLE_EXP 1
#endif


namespace hcg {

// Sorted channel values (multi-channel jobs) are blocked by kVB samples:
// block b holds the [nchan][kVB] values of sorted samples kVB b .. kVB b +
// kVB - 1, so a lane's 16-byte MFMA operand (8 samples of one channel) is
// one block row, and the MFMA path packs ring ranges at this granularity.
constexpr uint32_t kVB = 8;

constexpr int kTileCols = HCG_GRID_TILE_COLS;
constexpr int kLaneRows = 64 / kTileCols;
constexpr int kCellsPerLane = HCG_GRID_CELLS_PER_LANE;
constexpr int kTileRows = kLaneRows * kCellsPerLane;
constexpr uint32_t kStageCap = HCG_GRID_STAGE_CAP;
static_assert(kTileCols * kLaneRows == 64, "tile columns must divide 64");

}  // namespace hcg

#include "gridcore.hpp"

namespace hcg {

// LDS ordering among the lanes of one wave: its LDS instructions execute in
// issue order, so a compiler fence at wavefront scope is enough.
template <bool WAVE>
__device__ inline void tile_sync() {
  if constexpr (WAVE) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// A one-channel job's values and outputs (BATCH true).
struct TileOut {
  const uint16_t* vs;  // sorted values, BF16 or FP16: [n_pad] (one channel), or 32-sample blocked
                       // [n_pad / 32][nchan][32] (few channels)
  float* cube;         // [nchan][side * side]
  float* wsum;         // [side * side] or NULL
  uint32_t bf16;       // vs holds BF16 (else FP16)
  uint32_t normalize;  // cube = sums / weight sums
  uint32_t nchan;      // channels (few-channel tiles)
};

// The part of a tile's rings one call of grid_tile walks: rings r_begin +
// phase, + phase + stride, ... (the whole range: phase 0, stride 1).
struct RingStream {
  uint32_t phase, stride;
};

// Per-wave LDS of a tile: staged x, y, z, value columns and keys.
constexpr size_t kTileLdsBytes = (size_t)kStageCap * (4 * sizeof(double) + sizeof(uint32_t));

// Tiles of the kTileCols x kTileRows tiling of a side x side map.
inline uint32_t tile_cols_of(uint32_t side) { return (side + kTileCols - 1) / kTileCols; }
inline uint32_t tile_count_of(uint32_t side) { return tile_cols_of(side) * ((side + kTileRows - 1) / kTileRows); }

// One-channel jobs of the multi-channel entry points (hcg_grid_mfma*, batch
// jobs) that take the FP64 tile path instead of the MFMA contraction:
// GAUSS1D, no per-sample weights, a radius small enough for the three-term
// weight series.  With one channel the FP64 weight is all the work, and
// the tile loop evaluates it several times faster than the MFMA kernel's
// weight stage; the result is also closer to FP64 (unrounded weights).
inline bool one_channel_tiles(const GridParams& P, uint32_t kernel_type, bool weighted, uint32_t nchan) {
  return nchan == 1 && kernel_type == HCG_GAUSS1D && !weighted && gauss_tiny_radius(P);
}

// Partial streams per tile of a one-channel job, fixed by the map alone (so
// a job's sums do not depend on the launch shape): maps of fewer than
// kOneChannelMinTiles tiles (four per SIMD of the 256-CU chip) split each
// tile's rings over kOneChannelStreams waves, ring r to stream
// (r - first ring) % streams; larger maps take one wave per tile.  (A
// 5M-sample 128^2 map, 256 tiles: 25 ms on one wave per tile, against 15
// ms on the MFMA kernel, which spreads a block's weights over its waves.)
constexpr uint32_t kOneChannelMinTiles = 4096;
constexpr uint32_t kOneChannelStreams = 4;
inline uint32_t one_channel_streams(uint32_t side) {
  return tile_count_of(side) < kOneChannelMinTiles ? kOneChannelStreams : 1u;
}

// Few-channel jobs (2..kFewChannels channels on sparse maps, at most one
// sample per kFewCellsPerSample cells, otherwise as one_channel_tiles, at
// dot_distance_ok radii (their tile loop tests candidates on p.q), on maps
// of at least kOneChannelMinTiles tiles) also take the tile loop: FP64
// weights, FP32 accumulation per channel, one wave per tile at several
// waves per SIMD.  On sparse maps the MFMA kernel's fixed chain of barriers
// and global round trips per 64-cell block outweighs the per-pair work:
// 300k samples x 2..8 channels onto 2048^2 (0.07 samples per cell) took
// 1.96-2.10 ms here against 2.20-2.23 ms there.  With 16 channels, or at
// 0.6-13 samples per cell, the MFMA kernel is the faster one (3M x 16 onto
// 2048^2: 6.98 against 8.35 ms; 3.4M x 2 onto 512^2: 3.31 against 4.23 ms).
constexpr int kFewChannels = 8;
constexpr uint32_t kFewCellsPerSample = 10;
constexpr uint32_t kFewCap = 128;  // staged samples per round
constexpr size_t kFewTileLdsBytes = (size_t)kFewCap * (3 * sizeof(double) + sizeof(uint32_t) + kFewChannels * 2);

inline bool few_channel_tiles(const GridParams& P, uint32_t kernel_type, bool weighted, uint32_t nchan, uint32_t n) {
  return nchan >= 2 && nchan <= (uint32_t)kFewChannels &&
         (uint64_t)n * kFewCellsPerSample <= (uint64_t)P.side * P.side && kernel_type == HCG_GAUSS1D && !weighted &&
         gauss_tiny_radius(P) && dot_distance_ok(P) && tile_count_of(P.side) >= kOneChannelMinTiles;
}

// Launches the tile path for one one-channel job: one workgroup per tile,
// one_channel_streams(side) waves.
int launch_tiles_one_channel(const GridParams& P, const GridBuffers& B, const TileOut& O, hipStream_t s);

// Launches the tile path for one few-channel job: one wave per tile.
int launch_tiles_few_channels(const GridParams& P, const GridBuffers& B, const TileOut& O, hipStream_t s);

// WMODE (GAUSS1D): 0 general exponent and exp; 1 gauss_small_radius; 2 gauss_tiny_radius; 3 as 2,
// with the candidate test on p.q (dot_distance_ok).
// s_exp: the 64-entry exponential table (exp_table_load), visible to the wave.
//
// NC > 0 (few-channel tiles, BATCH true): up to NC channels of 16-bit values
// staged per sample in s_c ([CAP][NC]; s_v unused), one FP32 accumulator per
// channel and lane, and the cube, weight sum and normalisation written here
// (no ring streams: S must be {0, 1}).
template <int KT, bool HAS_W, int WMODE, bool BATCH, int NC = 0, uint32_t CAP = kStageCap>
__device__ __forceinline__ void grid_tile(const GridParams P, const GridBuffers B, uint32_t tile,
                                          uint32_t tiles_x, double* __restrict__ s_x, double* __restrict__ s_y,
                                          double* __restrict__ s_z, double* __restrict__ s_v,
                                          uint32_t* __restrict__ s_key, const double* __restrict__ s_exp,
                                          const TileOut& O, RingStream S, double* sum_d, double* sum_w,
                                          uint16_t* __restrict__ s_c = nullptr) {
  static_assert(!BATCH || kCellsPerLane == 1, "the one-channel path keeps one cell per lane");
  static_assert(WMODE != 3 || (kCellsPerLane == 1 && !HCG_GRID_LANE_LOOP && KT == HCG_GAUSS1D),
                "the p.q candidate test is written for the one-cell-per-lane GAUSS1D loop");
  static_assert(NC == 0 || (BATCH && NC % 2 == 0), "few-channel tiles: an even channel count, batch outputs");
  // the p.q loop with one value per sample: the lean weight and the guard
  // band handled after it (below)
  constexpr bool kLean = WMODE == 3 && NC == 0;
  float acc_c[NC > 0 ? NC : 1];
#pragma unroll
  for (int ch = 0; ch < (NC > 0 ? NC : 1); ++ch) acc_c[ch] = 0.0f;
  const hcg_ring_meta m = *B.meta;
  if constexpr (!BATCH) {
    if (m.usedrings == 0) return;
  }
  const uint32_t lane = __lane_id();
  const uint32_t col = (tile % tiles_x) * kTileCols + lane / kLaneRows;
  const uint32_t row0 = (tile / tiles_x) * kTileRows + (lane % kLaneRows) * kCellsPerLane;

  uint32_t cell[kCellsPerLane];
  double px[kCellsPerLane], py[kCellsPerLane], pz[kCellsPerLane];
  double acc_d[kCellsPerLane], acc_w[kCellsPerLane];
  Window win[kCellsPerLane];
  uint32_t lane_ur = 0xFFFFFFFFu, lane_dr = 0;
#pragma unroll
  for (int k = 0; k < kCellsPerLane; ++k) {
    const uint32_t row = row0 + k;
    const bool live = (col < P.side) && (row < P.side);
    const uint32_t rr = live ? row : 0, cc = live ? col : 0;
    cell[k] = rr * P.side + cc;
    const uint32_t anchor = rr * P.side + (cc / P.factor) * P.factor;
    win[k] = window_of(P, m.firstring, m.usedrings, B.xwcs[anchor], B.ywcs[anchor]);
    if (!live || (BATCH && m.usedrings == 0)) {  // an empty ring range: never active
      win[k].uring = 1;
      win[k].dring = 0;
    }
    double sb, cb, sl, cl;
    sincos(B.ywcs[cell[k]] * kDeg2Rad, &sb, &cb);
    sincos(B.xwcs[cell[k]] * kDeg2Rad, &sl, &cl);
    px[k] = cb * cl;
    py[k] = cb * sl;
    pz[k] = sb;
    if constexpr (BATCH) {
      acc_d[k] = 0.0;
      acc_w[k] = 0.0;
    } else {
      acc_d[k] = live ? B.dc[cell[k]] : 0.0;
      acc_w[k] = live ? B.wc[cell[k]] : 0.0;
    }
    if (win[k].uring <= win[k].dring) {
      lane_ur = min(lane_ur, win[k].uring);
      lane_dr = max(lane_dr, win[k].dring);
    }
  }
  uint32_t r_begin = wave_min_u32(lane_ur);
  const uint32_t r_end = wave_max_u32(lane_dr);
  if (r_begin == m.firstring && skips_first_ring(P)) ++r_begin;  // the first used ring contributes nothing

  for (uint32_t r = r_begin + S.phase; r <= r_end && r_begin <= r_end; r += S.stride) {
    const uint32_t kr = r - m.firstring;
    const uint32_t start = B.lut[kr], end = B.lut[kr + 1];
    if (start == end) continue;
    const double rz = B.ring_z[kr], rs = B.ring_s[kr];
    uint32_t lp[kCellsPerLane], rp[kCellsPerLane];
    uint32_t lane_lp = 0xFFFFFFFFu, lane_rp = 0;
#pragma unroll
    for (int k = 0; k < kCellsPerLane; ++k) {
      uint64_t a = 1, b = 0;
      const bool act = r >= win[k].uring && r <= win[k].dring && ring_window(P, win[k], r, rz, rs, a, b) && a <= b;
      lp[k] = act ? (uint32_t)a : 1u;  // inactive: the empty window [1, 0]
      rp[k] = act ? (uint32_t)b : 0u;
      if (act) {
        lane_lp = min(lane_lp, lp[k]);
        lane_rp = max(lane_rp, rp[k]);
      }
    }
    const uint32_t LP = wave_min_u32(lane_lp), RP = wave_max_u32(lane_rp);
    if (LP > RP) continue;
    const uint32_t lo_u = wave_search<false>(B.keys, start, end, LP);
    const uint32_t hi_u = wave_search<true>(B.keys, lo_u, end, RP);
    for (uint32_t cb = lo_u; cb < hi_u; cb += CAP) {
      const uint32_t cnt = min(CAP, hi_u - cb);
      tile_sync<BATCH>();  // previous chunk fully consumed
      for (uint32_t j = lane; j < cnt; j += 64) {
        const hcg_sample q = B.rec[cb + j];
        s_x[j] = q.x;
        s_y[j] = q.y;
        s_z[j] = q.z;
        if constexpr (NC > 0)
          ;  // channel values below
        else if constexpr (BATCH)
          s_v[j] = (double)widen16(O.vs[cb + j], O.bf16 != 0);
        else
          s_v[j] = q.v;
        s_key[j] = B.keys[cb + j];
      }
      if constexpr (NC > 0) {
        // channel by channel: lanes read consecutive samples of one channel
        // (contiguous within a kVB-sample block); slots past nchan hold 0
        for (uint32_t ch = 0; ch < (uint32_t)NC; ++ch) {
          const bool have = ch < O.nchan;
          for (uint32_t j = lane; j < cnt; j += 64) {
            const uint32_t i = cb + j;
            s_c[j * NC + ch] = have ? O.vs[((size_t)(i / kVB) * O.nchan + ch) * kVB + (i % kVB)] : (uint16_t)0;
          }
        }
      }
      tile_sync<BATCH>();
      uint32_t lo[kCellsPerLane], len[kCellsPerLane];
      uint32_t jb = 0xFFFFFFFFu, je = 0;
#pragma unroll
      for (int k = 0; k < kCellsPerLane; ++k) {
        const uint32_t a = lower_bound_u32(s_key, 0, cnt, lp[k]);
        const uint32_t b = (lp[k] <= rp[k]) ? upper_bound_u32(s_key, a, cnt, rp[k]) : a;
        lo[k] = a;
        len[k] = b - a;
        if (b > a) {
          jb = min(jb, a);
          je = max(je, b);
        }
      }
#if HCG_GRID_LANE_LOOP
      // each lane walks its own band: the trip count is the wave's longest
      // band, not the bands' union (over half of which a lane would skip);
      // per cell the samples are still taken in sorted order
      uint32_t tmax = 0;
#pragma unroll
      for (int k = 0; k < kCellsPerLane; ++k) tmax = max(tmax, len[k]);
      tmax = wave_max_u32(tmax);
      (void)jb;
      (void)je;
      for (uint32_t t = 0; t < tmax; ++t) {
#pragma unroll
        for (int k = 0; k < kCellsPerLane; ++k) {
          if (t >= len[k]) continue;
          const uint32_t j = lo[k] + t;
          const double qx = s_x[j], qy = s_y[j], qz = s_z[j], qv = s_v[j];
          const double dx = px[k] - qx, dy = py[k] - qy, dz = pz[k] - qz;
          const double c2 = fma(dx, dx, fma(dy, dy, dz * dz));
          if (c2 < P.c_hi) {
            const double gq = c2, glo = P.c_lo;
#elif HCG_GRID_CELLS_PER_LANE == 1
      // one cell per lane: its band is exactly [lo, lo + len)
      (void)jb;
      (void)je;
      // WMODE 3: the tests on h = 1 - p.q = c2 / 2 (halving is exact, so
      // h < c_hi / 2 <=> c2 < c_hi); -p in registers, so h is three FMAs
      const double h_hi = 0.5 * P.c_hi, h_lo = 0.5 * P.c_lo;
      const double npx = -px[0], npy = -py[0], npz = -pz[0];
      (void)h_hi;
      (void)npx;
      (void)npy;
      (void)npz;
      // the lean weight's constants for y = h (2 k0 + 4 k1 h), which equals
      // c2 (k0 + k1 c2) bit for bit: k1 and kExpE3 in vector registers
      // (hoisted out of the loops)
      const double k0h = 2.0 * P.k0;
      const double k1v = WMODE == 3 ? in_vgprs(4.0 * P.k1) : 0.0, e3v = WMODE == 3 ? in_vgprs(kExpE3) : 0.0;
      (void)k0h;
      (void)k1v;
      (void)e3v;
      for (uint32_t t = 0; t < len[0]; ++t) {
        const uint32_t j = lo[0] + t;
        const double qx = s_x[j], qy = s_y[j], qz = s_z[j], qv = NC > 0 ? 0.0 : s_v[j];
        {
          constexpr int k = 0;
          // WMODE 3: the candidate test on h = 1 - p.q (three FP64
          // operations instead of seven); otherwise c2 = |p - q|^2
          double c2, h = 0.0;
          if constexpr (WMODE == 3) {
            h = fma(npx, qx, fma(npy, qy, fma(npz, qz, 1.0)));
          } else {
            const double dx = px[k] - qx, dy = py[k] - qy, dz = pz[k] - qz;
            c2 = fma(dx, dx, fma(dy, dy, dz * dz));
          }
          if (WMODE == 3 ? h < h_hi : c2 < P.c_hi) {
            if constexpr (WMODE == 3) c2 = 2.0 * h;
            const double gq = WMODE == 3 ? h : c2, glo = WMODE == 3 ? h_lo : P.c_lo;
#else
      for (uint32_t j = jb; j < je; ++j) {
        const double qx = s_x[j], qy = s_y[j], qz = s_z[j], qv = s_v[j];
#pragma unroll
        for (int k = 0; k < kCellsPerLane; ++k) {
          const double dx = px[k] - qx, dy = py[k] - qy, dz = pz[k] - qz;
          const double c2 = fma(dx, dx, fma(dy, dy, dz * dz));
          if ((j - lo[k]) < len[k] && c2 < P.c_hi) {
            const double gq = c2, glo = P.c_lo;
#endif
            if constexpr (kLean) {
              // every candidate's weight; a lane in the guard band (rare, and
              // wave-uniformly skipped) that the reference's test rejects then
              // adds zero: one compare and no lane mask in the common case
              double wt = gauss_weight_lean(h, k0h, k1v, e3v, s_exp);
              double qw = qv;
              if constexpr (HAS_W) wt = B.w_s[cb + j] * wt;
              if (__builtin_amdgcn_ballot_w64(!(gq < glo))) {
                // (a block of its own, so the branch above stays a scalar one)
                asm volatile("; guard band" ::);
                if (!(gq < glo) &&
                    !radius_test_ref(B.xwcs, B.ywcs, B.lon_s, B.lat_s, cell[k], cb + j, P.sphere_radius)) {
                  wt = 0.0;
                  qw = 0.0;
                }
              }
              acc_d[k] = fma(qw, wt, acc_d[k]);
              acc_w[k] += wt;
            } else if (gq < glo ||
                       radius_test_ref(B.xwcs, B.ywcs, B.lon_s, B.lat_s, cell[k], cb + j, P.sphere_radius)) {
              double wt;
              if constexpr (KT == HCG_GAUSS1D && WMODE > 0)
                wt = gauss_weight_series<WMODE == 3 ? 2 : WMODE>(P, c2, s_exp);
              else if constexpr (KT == HCG_GAUSS1D)
                wt = HCG_GRID_TABLE_EXP ? exp_scaled_neg(gauss_exponent_scaled(P, c2), s_exp)
                                        : exp(gauss_exponent_scaled(P, c2) * 0x1.62e42fefa39efp-7);
              else
                wt = weight_general<KT>(P.params[0], P.params[1], P.params[2], B.xwcs, B.ywcs, B.lon_s, B.lat_s,
                                        0.25 * c2, cell[k], cb + j);
              if constexpr (HAS_W) wt = B.w_s[cb + j] * wt;
              if constexpr (NC > 0) {
                // two channels per 32-bit LDS word; the value type is the
                // job's (a uniform branch)
                const float wf = (float)wt;
                const uint32_t* cv = reinterpret_cast<const uint32_t*>(s_c + j * NC);
                if (O.bf16) {
#pragma unroll
                  for (int h = 0; h < NC / 2; ++h) {
                    const uint32_t u = cv[h];
                    acc_c[2 * h] = fmaf(wf, __uint_as_float(u << 16), acc_c[2 * h]);
                    acc_c[2 * h + 1] = fmaf(wf, __uint_as_float(u & 0xFFFF0000u), acc_c[2 * h + 1]);
                  }
                } else {
#pragma unroll
                  for (int h = 0; h < NC / 2; ++h) {
                    const uint32_t u = cv[h];
                    acc_c[2 * h] = fmaf(wf, (float)__builtin_bit_cast(_Float16, (uint16_t)(u & 0xFFFFu)),
                                        acc_c[2 * h]);
                    acc_c[2 * h + 1] = fmaf(wf, (float)__builtin_bit_cast(_Float16, (uint16_t)(u >> 16)),
                                            acc_c[2 * h + 1]);
                  }
                }
                (void)qv;
              } else {
                acc_d[k] = fma(qv, wt, acc_d[k]);
              }
              acc_w[k] += wt;
            }
          }
        }
      }
    }
  }
  if constexpr (NC > 0) {
    if (col < P.side && row0 < P.side) {
      const size_t M = (size_t)P.side * P.side;
      const uint32_t c0 = cell[0];
#pragma unroll
      for (int ch = 0; ch < NC; ++ch)
        if ((uint32_t)ch < O.nchan)
          O.cube[(size_t)ch * M + c0] =
              O.normalize ? (float)((double)acc_c[ch] / acc_w[0]) : acc_c[ch];
      if (O.wsum) O.wsum[c0] = (float)acc_w[0];
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < kCellsPerLane; ++k) {
    if constexpr (BATCH) {
      sum_d[k] = acc_d[k];
      sum_w[k] = acc_w[k];
    } else if (col < P.side && row0 + k < P.side) {
      B.dc[cell[k]] = acc_d[k];
      B.wc[cell[k]] = acc_w[k];
    }
  }
}

// The FP32 outputs of one one-channel tile cell (lane) from its sums.
__device__ inline void tile_write(const GridParams& P, uint32_t tile, uint32_t tiles_x, const TileOut& O, double d,
                                  double w) {
  const uint32_t lane = __lane_id();
  const uint32_t col = (tile % tiles_x) * kTileCols + lane / kLaneRows;
  const uint32_t row = (tile / tiles_x) * kTileRows + lane % kLaneRows;
  if (col < P.side && row < P.side) {
    const uint32_t cell = row * P.side + col;
    O.cube[cell] = O.normalize ? (float)(d / w) : (float)d;
    if (O.wsum) O.wsum[cell] = (float)w;
  }
}

// LDS of one_channel_tile: `active` staging slices, then the streams' sums
// when there are several.
constexpr size_t one_channel_lds_bytes(uint32_t active, uint32_t streams) {
  return active * kTileLdsBytes + (streams > 1 ? (size_t)streams * 64 * 2 * sizeof(double) : 0);
}

// One one-channel tile over `streams` partial streams: waves w < active of
// the workgroup walk streams w, w + active, ... in their own kTileLdsBytes
// slice of `lds`; the streams' sums meet after the slices and wave 0 adds
// them in stream order and writes the cells.  Every wave of the workgroup
// calls this (it holds a workgroup barrier when streams > 1).
__device__ __forceinline__ void one_channel_tile(const GridParams& P, const GridBuffers& B, const TileOut& O,
                                                 uint32_t tile, uint32_t tiles_x, uint32_t streams, uint32_t active,
                                                 char* lds, const double* s_exp) {
  const uint32_t wave = threadIdx.x >> 6, lane = __lane_id();
  double* red = reinterpret_cast<double*>(lds + active * kTileLdsBytes);
  double d = 0.0, w = 0.0;
  if (wave < active) {
    double* s_x = reinterpret_cast<double*>(lds + wave * kTileLdsBytes);
    double* s_y = s_x + kStageCap;
    double* s_z = s_y + kStageCap;
    double* s_v = s_z + kStageCap;
    uint32_t* s_key = reinterpret_cast<uint32_t*>(s_v + kStageCap);
    for (uint32_t q = wave; q < streams; q += active) {
      grid_tile<HCG_GAUSS1D, false, 2, true>(P, B, tile, tiles_x, s_x, s_y, s_z, s_v, s_key, s_exp, O,
                                             RingStream{q, streams}, &d, &w);
      if (streams > 1) {
        red[(q * 64 + lane) * 2] = d;
        red[(q * 64 + lane) * 2 + 1] = w;
      }
    }
  }
  if (streams > 1) {
    __syncthreads();
    if (wave == 0) {
      d = red[lane * 2];
      w = red[lane * 2 + 1];
      for (uint32_t q = 1; q < streams; ++q) {
        d += red[(q * 64 + lane) * 2];
        w += red[(q * 64 + lane) * 2 + 1];
      }
    }
  }
  if (wave == 0) tile_write(P, tile, tiles_x, O, d, w);
}

// One few-channel tile in a kFewTileLdsBytes slice of LDS (one wave).
__device__ __forceinline__ void few_channel_tile_body(const GridParams& P, const GridBuffers& B, const TileOut& O,
                                                      uint32_t tile, uint32_t tiles_x, char* lds,
                                                      const double* s_exp) {
  double* s_x = reinterpret_cast<double*>(lds);
  double* s_y = s_x + kFewCap;
  double* s_z = s_y + kFewCap;
  uint32_t* s_key = reinterpret_cast<uint32_t*>(s_z + kFewCap);
  uint16_t* s_c = reinterpret_cast<uint16_t*>(s_key + kFewCap);
  grid_tile<HCG_GAUSS1D, false, 3, true, kFewChannels, kFewCap>(P, B, tile, tiles_x, s_x, s_y, s_z, nullptr, s_key,
                                                                s_exp, O, RingStream{0, 1}, nullptr, nullptr, s_c);
}

// The same, out of line, for the persistent batch kernel: inlined there as
// a third copy of the tile body, it raised that kernel's register spills
// and slowed its MFMA and one-channel items by 1-10%; the call costs a
// few-channel tile ~4%.  (The one-job kernel inlines the body: a called
// function takes the register budget of its widest caller.)
[[maybe_unused]] static __device__ __noinline__ void few_channel_tile(const GridParams& P, const GridBuffers& B,
                                                                    const TileOut& O, uint32_t tile,
                                                                    uint32_t tiles_x, char* lds,
                                                                    const double* s_exp) {
  few_channel_tile_body(P, B, O, tile, tiles_x, lds, s_exp);
}

}  // namespace hcg

// grid_mfma.hip -- multi-channel gridding as a dense MFMA contraction.
//
// The reference grids a spectral cube one channel at a time, rerunning the
// full single-channel kernel per channel on its own stream
// (HCGrid+/gridding-omp.cu:552-678, kernel HCGrid+/gridding-omp.cu:294-420).
// The selection and weights do not depend on the channel, so here they are
// computed once per (cell, sample) pair in FP64 and the channel sums become
//     cube[ch][cell] = sum_s W[cell][s] * V[ch][s]   (FP32 accumulation)
// evaluated with v_mfma_f32_16x16x32_{bf16,f16}: W (cells x samples) in LDS,
// V (channels x sorted samples) streamed from HBM.  Each FP64 weight is
// carried as hi + lo, two values of the channel type (hi = round(w),
// lo = round(w - hi)), and both are contracted: the weight keeps ~2x the
// type's significand (relative error ~2^-17 for BF16) at twice the MFMA
// count, which the kernel's HBM/L2 stream of V hides.
//
// Work decomposition: a workgroup owns a block of 64 cells (TR rows x TC
// columns; 4 MFMA row blocks of 16 cells) and up to 8 x 64 channels (wave
// w: 4 blocks of 16 channels; 4 row blocks x 4 channel blocks x 4 = 64 FP32
// accumulators per lane).  Wider channel ranges take more workgroups along
// grid.y.  For every non-empty ring of the block's ring range:
//   1. each cell's [lpix, rpix] window (the reference's compat selection);
//   2. the union window's sorted range by a 64-way cooperative key search;
//   3. per chunk of <= CHUNK samples: stage unit vectors and keys in LDS,
//      zero the weight tiles, let every cell's threads fill its band with
//      the FP64-decided weights (as hi + lo; the sum of the rounded pairs is
//      the normalisation, so numerator and denominator use the same
//      weights), record each row block's k-block span;
//   4. MFMA over the union's 32-sample k-blocks: B (16 channels x 32
//      samples) is loaded once per k-block and applied to every row block
//      whose span covers it.
// Selection, FP64 distance and weight follow the single-channel kernel
// exactly; only the products with channel values run in BF16/FP16.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gridcore.hpp"
#include "mfma_job.hpp"

namespace hcg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));


typedef short s16x8_t __attribute__((ext_vector_type(8)));  // a raw 16-byte MFMA operand

// Channel type of a kernel instance: fixed at compile time (one-job
// launches) or read from the job descriptor (the batched launch, whose jobs
// may mix BF16 and FP16).  Values and weights are kept as raw 16-bit words;
// the type matters only where a weight is rounded and where the MFMA issues.
enum : int { kDtBf16 = 0, kDtF16 = 1, kDtAny = 2 };

template <int DT>
__device__ inline bool dt_is_bf16(uint32_t dtype) {
  if constexpr (DT == kDtBf16) return true;
  else if constexpr (DT == kDtF16) return false;
  else return dtype == HCG_DTYPE_BF16;
}

__device__ inline f32x4_t mfma16(bool bf, s16x8_t a, s16x8_t b, f32x4_t c) {
  if (bf)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a), __builtin_bit_cast(bf16x8_t, b),
                                                   c, 0, 0, 0);
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8_t, a), __builtin_bit_cast(f16x8_t, b), c,
                                                0, 0, 0);
}

// One k-block of B for kMcNB channel blocks, 16 bytes per lane per block.
// The sorted channel values are blocked by kVB = 8 samples ([nchan][8] per
// block): lane l's operand is 8 samples of one channel, a row of the value
// block of its k-group (l / 16) in the k-block, so `blk` points at that
// value block and lofs[nb] = 16 x the lane's channel; the 16 lanes of a
// k-group read one contiguous 256-byte piece per channel block.
__device__ inline void load_b_rows(s16x8_t* b, const char* blk, const uint32_t* lofs) {
#pragma unroll
  for (int nb = 0; nb < kMcNB; ++nb) b[nb] = *reinterpret_cast<const s16x8_t*>(blk + lofs[nb]);
}

// FP64 weight of a selected pair (the single-channel kernel's evaluation).
__device__ inline double pair_weight_any(const GridParams& P, const GridBuffers& B, uint32_t kt, double c2,
                                         const double* s_exp, uint32_t cell, uint32_t i) {
  // GAUSS1D whose radius rules out the exponent's asin branch and the exp
  // clamp (gauss_small_radius): the same values
  if (kt == kKtGaussSmall || kt == kKtGaussLean) return exp_scaled_neg_nc(gauss_exponent_series(P, c2), s_exp);
  if (kt == HCG_GAUSS1D) return exp_scaled_neg(gauss_exponent_scaled(P, c2), s_exp);
  if (kt == HCG_GAUSS2D)
    return weight_general<HCG_GAUSS2D>(P.params[0], P.params[1], P.params[2], B.xwcs, B.ywcs, B.lon_s, B.lat_s,
                                       0.25 * c2, cell, i);
  return weight_general<HCG_TAPERED_SINC>(P.params[0], P.params[1], P.params[2], B.xwcs, B.ywcs, B.lon_s,
                                          B.lat_s, 0.25 * c2, cell, i);
}

// Block geometry, lane l = cell l of a (64 >> tc_shift) x (1 << tc_shift)
// block at (y0, x0): the cell's window (its anchor's, HCGrid/gridding.cu:306-323)
// and unit vector q, and the block's ring span [ur, dr] (a wave reduction;
// ur > dr when no cell has rings), first used ring excluded.
__device__ inline Window block_cell_window(const GridParams& P, const GridBuffers& B, const hcg_ring_meta& m,
                                           uint32_t y0, uint32_t x0, uint32_t tc_shift, uint32_t lane, double q[3],
                                           uint32_t& ur, uint32_t& dr) {
  const uint32_t tc_mask = (1u << tc_shift) - 1u;
  const uint32_t rl = y0 + (lane >> tc_shift), cl = x0 + (lane & tc_mask);
  const bool lv = cl < P.side && rl < P.side;
  const uint32_t celll = lv ? rl * P.side + cl : 0;
  const uint32_t anchor = lv ? rl * P.side + (cl / P.factor) * P.factor : 0;
  Window win = window_of(P, m.firstring, m.usedrings, B.xwcs[anchor], B.ywcs[anchor]);
  if (!lv || m.usedrings == 0) {
    win.uring = 1;
    win.dring = 0;
  }
  double sb, cb0, sn, cs;
  sincos(B.ywcs[celll] * kDeg2Rad, &sb, &cb0);
  sincos(B.xwcs[celll] * kDeg2Rad, &sn, &cs);
  q[0] = cb0 * cs;
  q[1] = cb0 * sn;
  q[2] = sb;
  const bool has = win.uring <= win.dring;
  ur = wave_min_u32(has ? win.uring : 0xFFFFFFFFu);
  dr = wave_max_u32(has ? win.dring : 0u);
  if (ur == m.firstring && skips_first_ring(P)) ++ur;  // the first used ring contributes nothing (start_int - 1 wraps)
  return win;
}

// A cell's [lpix, rpix] in ring r (HCGrid/gridding.cu:334-360); (1, 0) when
// the cell takes no samples from the ring.
__device__ inline void ring_cell_window(const GridParams& P, const GridBuffers& B, const hcg_ring_meta& m,
                                        const Window& win, uint32_t r, uint32_t& lp, uint32_t& rp) {
  const uint32_t kr = r - m.firstring;
  uint64_t a64 = 1, b64 = 0;
  const bool act =
      r >= win.uring && r <= win.dring && ring_window(P, win, r, B.ring_z[kr], B.ring_s[kr], a64, b64) && a64 <= b64;
  lp = act ? (uint32_t)a64 : 1u;
  rp = act ? (uint32_t)b64 : 0u;
}

// The sorted range of the union of the wave's windows within [start, end)
// (a 64-way cooperative search of the keys); lo >= hi when empty.
__device__ inline void ring_union_range(const GridBuffers& B, uint32_t lp, uint32_t rp, uint32_t start, uint32_t end,
                                        uint32_t& lo, uint32_t& hi) {
  const bool act = lp <= rp;
  const uint32_t LP = wave_min_u32(act ? lp : 0xFFFFFFFFu);
  const uint32_t RP = wave_max_u32(act ? rp : 0u);
  lo = hi = 0;
  if (LP <= RP) {  // uniform
    lo = wave_search<false>(B.keys, start, end, LP);
    hi = wave_search<true>(B.keys, lo, end, RP);
  }
}

// The tile plan: one wave per tile, lane = cell.  Many small workgroups per
// CU hide the latency of the per-ring searches that the gridding kernel,
// with one or two large workgroups per CU, cannot.
__device__ inline void plan_tile(const MfmaJob& J, uint32_t tile) {
  const uint32_t lane = __lane_id();
  const hcg_ring_meta m = *J.B.meta;
  const uint32_t tc_shift = J.tc_shift;
  const uint32_t y0 = (tile / J.tiles_x) << (6u - tc_shift);
  const uint32_t x0 = (tile % J.tiles_x) << tc_shift;
  double q[3];
  uint32_t r_begin, r_end;
  const Window win = block_cell_window(J.P, J.B, m, y0, x0, tc_shift, lane, q, r_begin, r_end);
  uint32_t* out = J.plan + (size_t)tile * kPlanTileWords;
  if (lane == 0) {
    out[0] = r_begin;
    out[1] = r_end;
  }
  reinterpret_cast<Window*>(out + kPlanHeadWords)[lane] = win;
  double* pq = reinterpret_cast<double*>(out + kPlanHeadWords + kPlanWinWords);
  pq[lane] = q[0];
  pq[kMcCols + lane] = q[1];
  pq[2 * kMcCols + lane] = q[2];
  for (uint32_t k = 0; k < kPlanRings; ++k) {
    const uint32_t r = r_begin + k;
    uint32_t lo = 0, hi = 0;
    if (r_begin <= r_end && r <= r_end) {  // uniform
      const uint32_t kr = r - m.firstring;
      const uint32_t start = J.B.lut[kr], end = J.B.lut[kr + 1];
      if (start != end) {
        uint32_t lp, rp;
        ring_cell_window(J.P, J.B, m, win, r, lp, rp);
        ring_union_range(J.B, lp, rp, start, end, lo, hi);
      }
    }
    if (lane == 0) {
      out[2 + 2 * k] = lo;
      out[3 + 2 * k] = hi;
    }
  }
}

__global__ void __launch_bounds__(64) k_mfma_plan(MfmaJob J) { plan_tile(J, blockIdx.x); }

int launch_mfma_plan(const MfmaJob& J, hipStream_t s) {
  if (J.ntiles == 0) return HCG_OK;
  k_mfma_plan<<<J.ntiles, 64, 0, s>>>(J);
  HCG_CHECK_LAUNCH("k_mfma_plan");
  return HCG_OK;
}

// LDS of one workgroup (the weight tiles as raw 16-bit storage).
template <uint32_t CHUNK>
struct MfmaLds {
  // first, so one address and immediate offsets reach a sample's x, y, z
  double x[CHUNK], y[CHUNK], z[CHUNK];
  uint32_t key[CHUNK];
  // the weight tiles, row-interleaved: cell c's hi weights at c * RS, its lo
  // weights at c * RS + CHUNK (elements; RS = 2 CHUNK + 8), so one address
  // and an immediate offset reach both; RS is an odd multiple of 16 bytes,
  // so the 16 rows of an MFMA operand read start in 16 distinct groups of 4
  // banks
  __attribute__((aligned(16))) uint16_t w[kMcCols * (2 * CHUNK + 8)];
  uint32_t mb[2][4][2];  // row blocks' k-block spans, by chunk parity
  double inv[kMcCols];  // per cell: 1 / its weight sum (1 without normalisation)
  double p[kMcCols][3];          // cell unit vectors
  uint32_t span[2];              // the block's ring span
  uint32_t plan[kPlanStride];    // the tile's plan (when the job has one)
  // chunk descriptors (double-buffered): blocks and segments; per block its
  // sorted 32-sample block index, per segment its ring, first block, blocks
  uint32_t dn[2][3];                   // k-blocks, segments, value blocks
  uint32_t dgb[2][CHUNK / kVB];        // per value block of the chunk: its sorted block index
  uint32_t dring[2][CHUNK / 32];       // per segment (at most CHUNK / 32 per chunk): its ring,
  uint16_t du0[2][CHUNK / 32], dnb[2][CHUNK / 32];  // first value block, value blocks
  uint32_t band[CHUNK / 32][kMcCols];  // each cell's band per segment: ja | jb << 16
  // a planned tile's rings (the wave-parallel descriptor builder), slot l =
  // ring r_begin + l, written and read by lane l of wave 0 only: first value
  // block, the ring's offset in the tile's block sequence, its block count
  uint32_t rl_b0[kPlanRings], rl_ex[kPlanRings], rl_nb[kPlanRings];
  Window win[kMcCols];                 // each cell's window
};

// CHUNK = staged samples per chunk.  LDS: the weight tiles (64 rows of
// 2 CHUNK + 8 elements) + CHUNK unit vectors and keys + the chunk
// descriptors and bands: 156 KB at 512 (one workgroup per CU), 79 KB at 256
// (two per CU), 46 KB at 128.
//
// The 64 cells of a workgroup form a TR x TC block of the map (TC = 1 <<
// tc_shift columns, TR = 64 / TC rows); cell c of the block is (c / TC,
// c % TC) and MFMA row block mb holds cells 16 mb .. 16 mb + 15.  Square-ish
// blocks share more of their neighbourhood, so each channel value is
// re-read by fewer workgroups: (TR + d)(TC + d) / 64 reads per value for a
// support of d cells, 16.8 for 1 x 64 and 7.7 for 4 x 16 at d = 13.
// STAGED: the cube stores go through LDS (the one-job kernel, where they
// took the 2M x 1024 cube from 7.29 to 7.08 ms; in the persistent kernel,
// whose register budget is tighter, the 64-job batch measured 159.4 ms
// with them against 158.9 without).
template <int DT, uint32_t CHUNK, bool STAGED>
__device__ __forceinline__ void mfma_block(const MfmaJob& J, uint32_t tile, uint32_t group, MfmaLds<CHUNK>& L,
                                           const double* s_exp) {
  typedef s16x8_t vec;
  constexpr uint32_t WS = CHUNK;       // hi -> lo offset within a weight row (elements)
  constexpr uint32_t RS = 2 * WS + 8;  // weight row stride (elements)
  static_assert((WS * 2) % 16 == 0 && (RS * 2) % 16 == 0 && (RS / 2) % 8 == 4, "weight tile layout");
  uint16_t* s_wh = L.w;
  uint16_t* s_wl = L.w + WS;
  const bool bf = dt_is_bf16<DT>(J.dtype);
  double* s_x = L.x;
  double* s_y = L.y;
  double* s_z = L.z;
  uint32_t* s_key = L.key;
  auto& s_mb = L.mb;
  double* s_inv = L.inv;
  const GridParams& P = J.P;
  const GridBuffers& B = J.B;
  const uint16_t* __restrict__ vs = static_cast<const uint16_t*>(J.vs);
  const uint32_t kt = J.kt, nchan = J.nchan, tc_shift = J.tc_shift, tiles_x = J.tiles_x;
  const size_t blk = (size_t)nchan * kVB;  // elements per value block
  float* __restrict__ cube = J.cube;
  float* __restrict__ wsum_out = J.wsum;
  const int normalize = J.normalize;

  const hcg_ring_meta m = *B.meta;
  const uint32_t nthreads = blockDim.x;
  const uint32_t nw = nthreads >> 6;       // waves in this workgroup
  const uint32_t S = nw;                   // threads per cell (one per wave)
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = __lane_id();
  const uint32_t wave = tid >> 6;

  const uint32_t tc_mask = (1u << tc_shift) - 1u;
  const uint32_t y0 = (tile / tiles_x) << (6u - tc_shift);
  const uint32_t x0 = (tile % tiles_x) << tc_shift;
  const uint32_t M = P.side * P.side;

  // this thread's cell c = lane and its share sub = wave of the cell's band:
  // every wave holds all 64 cells, so the weight work is balanced across
  // the waves whichever rows of the block a ring reaches
  const uint32_t c = lane, sub = wave;
  const uint32_t row = y0 + (c >> tc_shift);
  const uint32_t col = x0 + (c & tc_mask);
  const bool live = col < P.side && row < P.side;
  const uint32_t cell = live ? row * P.side + col : 0;

  // Wave 0 evaluates the block's geometry, lane l = cell l: the cell's
  // window (kept in LDS, not registers, through the chunk loop) and unit
  // vector, the block's ring span, and per ring the cells' [lpix, rpix] and
  // the union's sorted range, published through LDS.  The other waves only
  // stage, weigh and contract.
  if (J.plan) {  // the plan holds all of it (one round of coalesced loads)
    const uint32_t* tp = J.plan + (size_t)tile * kPlanTileWords;
    for (uint32_t i = tid; i < kPlanStride; i += nthreads) L.plan[i] = tp[i];
    if (wave == 0) {
      L.win[lane] = reinterpret_cast<const Window*>(tp + kPlanHeadWords)[lane];
      const double* pq = reinterpret_cast<const double*>(tp + kPlanHeadWords + kPlanWinWords);
      L.p[lane][0] = pq[lane];
      L.p[lane][1] = pq[kMcCols + lane];
      L.p[lane][2] = pq[2 * kMcCols + lane];
      if (lane == 0) {
        L.span[0] = tp[0];
        L.span[1] = tp[1];
      }
    }
  } else if (wave == 0) {
    double q[3];
    uint32_t ur, dr;
    L.win[lane] = block_cell_window(P, B, m, y0, x0, tc_shift, lane, q, ur, dr);
    L.p[lane][0] = q[0];
    L.p[lane][1] = q[1];
    L.p[lane][2] = q[2];
    if (lane == 0) {
      L.span[0] = ur;
      L.span[1] = dr;
    }
  }
  __syncthreads();
  const double px = L.p[c][0], py = L.p[c][1], pz = L.p[c][2];
  double wsum = 0.0;

  // accumulators: 4 row blocks x kMcNB channel blocks of 16x16
  f32x4_t acc[4][kMcNB];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nb = 0; nb < kMcNB; ++nb) acc[mb][nb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const uint32_t ch_base = (group * nw + wave) * kMcChanPerWave;
  // B operand addressing: lane l reads samples 8 (l / 16) .. +7 of a
  // k-block for channel ch_base + 16 nb + l % 16 (clamped to the last one)
  const uint32_t kg = lane >> 4, koff = kg * 8;  // the lane's k-group: samples koff .. koff + 7 of a k-block
  uint32_t lofs[kMcNB];
#pragma unroll
  for (int nb = 0; nb < kMcNB; ++nb) {
    const uint32_t ch = ch_base + nb * 16 + (lane & 15);
    lofs[nb] = (ch < nchan ? ch : nchan - 1) * (kVB * 2u);
  }
  // the lane's value block of k-block u of descriptor db
  const char* vs_bytes = reinterpret_cast<const char*>(vs);
  auto kblock = [&](uint32_t db, uint32_t u) {
    return vs_bytes + (size_t)L.dgb[db][u * (32 / kVB) + kg] * (blk * 2);
  };

  const uint32_t r_begin = L.span[0], r_end = L.span[1];
  const uint32_t n_all = m.n;       // sorted samples (records and keys beyond are not read)
  constexpr uint32_t KV = CHUNK / kVB;    // value blocks per chunk
  constexpr uint32_t KSEG = CHUNK / 32;   // segments per chunk

  auto planned = [&](uint32_t rr) { return J.plan != nullptr && (rr - r_begin) < kPlanRings; };
  auto plan_lo = [&](uint32_t rr) { return L.plan[2 + 2 * (rr - r_begin)]; };
  auto plan_hi = [&](uint32_t rr) { return L.plan[3 + 2 * (rr - r_begin)]; };
  // sorted index of chunk position p of descriptor db
  auto gidx = [&](uint32_t db, uint32_t p) { return L.dgb[db][p / kVB] * kVB + (p % kVB); };

  // ---- chunk descriptors -----------------------------------------------------
  // A chunk packs the value blocks of consecutive rings' union ranges, at
  // most KV blocks in at most KSEG segments: ring r's blocks [lo_r / kVB,
  // (hi_r + kVB - 1) / kVB), in ascending ring and block order, so the
  // staged keys are sorted and each cell's band in a ring is one interval of
  // the chunk.  The chunk's k-blocks (32 positions) are nkb = ceil(blocks /
  // 4); the slots past the last block repeat it (their positions are staged
  // empty, weight 0).  Wave 0 builds them, one ahead, walking the rings from
  // its cursor (ring cr at block offset cbo of its range [clo, chi)); a ring
  // past the plan has its range searched here (every lane of wave 0 takes
  // part).
  uint32_t cr = r_begin, cbo = 0, clo = 0, chi = 0;
  bool crange = false;
  constexpr uint32_t VPK = 32 / kVB;  // value blocks per k-block
  static_assert(KSEG >= VPK && KV % VPK == 0, "chunk geometry");
  auto build_desc = [&](uint32_t db) {  // wave 0, uniform
    uint32_t nvb = 0, nseg = 0;
    while (nvb < KV && r_begin <= r_end && cr <= r_end) {
      // with at most VPK - 1 segments left, only complete the current
      // k-block, then end the chunk: every chunk but a tile's last ends on a
      // k-block boundary, so k-block j of a tile is always value blocks
      // VPK j .. VPK j + VPK - 1 of the tile's block sequence whatever the
      // chunk size, and the FP32 sums do not depend on it
      const uint32_t need = (VPK - nvb % VPK) % VPK;
      const bool closing = nseg + (VPK - 1) >= KSEG;
      if (closing && need == 0) break;  // uniform
      if (!crange) {
        clo = chi = 0;
        if (planned(cr)) {
          clo = plan_lo(cr);
          chi = plan_hi(cr);
        } else {
          const uint32_t kr = cr - m.firstring;
          const uint32_t start = B.lut[kr], end = B.lut[kr + 1];
          if (start != end) {  // uniform
            uint32_t a, b;
            ring_cell_window(P, B, m, L.win[lane], cr, a, b);
            ring_union_range(B, a, b, start, end, clo, chi);
          }
        }
        crange = true;
        cbo = 0;
      }
      if (clo >= chi) {
        ++cr;
        crange = false;
        continue;
      }
      const uint32_t b0 = clo / kVB + cbo, b1 = (chi + kVB - 1) / kVB;
      uint32_t take = min(KV - nvb, b1 - b0);
      if (closing) take = min(take, need);
      if (lane == 0) {
        for (uint32_t t = 0; t < take; ++t) L.dgb[db][nvb + t] = b0 + t;
        L.dring[db][nseg] = cr;
        L.du0[db][nseg] = (uint16_t)nvb;
        L.dnb[db][nseg] = (uint16_t)take;
      }
      nvb += take;
      ++nseg;
      if (b0 + take == b1) {
        ++cr;
        crange = false;
      } else {
        cbo += take;
      }
    }
    const uint32_t nkb = (nvb + VPK - 1) / VPK;
    if (lane == 0) {
      for (uint32_t t = nvb; t < nkb * VPK; ++t) L.dgb[db][t] = L.dgb[db][nvb - 1];
      L.dn[db][0] = nkb;
      L.dn[db][1] = nseg;
      L.dn[db][2] = nvb;
    }
  };

  // The same packing built by the 64 lanes of wave 0 at once, for a tile
  // whose rings are all in its plan: lane l holds ring r_begin + l (its
  // block count nb, 0 when empty, and its offset ex in the tile's block
  // sequence, an exclusive prefix sum), and the non-empty rings are ranked
  // by popcount.  A chunk starting at block cpos ends at E = min(cpos + KV,
  // the first k-block boundary at or after the end of the (KSEG - VPK +
  // 1)-th non-empty ring from the one holding cpos) -- where the walk above
  // stops -- and every ring overlapping [cpos, E) writes its own segment.
  // Lanes exchange values by shuffles only: LDS slots are read back by the
  // lane that wrote them.
  const bool par_desc = J.plan != nullptr && r_begin <= r_end && r_end - r_begin < kPlanRings;  // uniform
  uint32_t cpos = 0;
  auto ring_list = [&]() {  // wave 0
    const uint32_t r = r_begin + lane;
    uint32_t nb = 0, b0 = 0;
    if (r <= r_end) {
      const uint32_t lo = plan_lo(r), hi = plan_hi(r);
      if (lo < hi) {
        b0 = lo / kVB;
        nb = (hi + kVB - 1) / kVB - b0;
      }
    }
    uint32_t incl = nb;  // inclusive prefix sum over the lanes
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
      const uint32_t t = __shfl_up(incl, off, 64);
      if (lane >= off) incl += t;
    }
    L.rl_b0[lane] = b0;
    L.rl_ex[lane] = incl - nb;
    L.rl_nb[lane] = nb;
  };
  auto build_desc_par = [&](uint32_t db) {  // wave 0, after ring_list
    const uint32_t b0 = L.rl_b0[lane], ex = L.rl_ex[lane], nb = L.rl_nb[lane];  // this lane's own slots
    const uint32_t in = ex + nb;
    const uint32_t total = __shfl(in, 63, 64);
    uint32_t nvb = 0, nseg = 0;
    const uint64_t hm = __ballot(nb > 0 && ex <= cpos && cpos < in);  // the ring holding cpos
    if (cpos < total && hm) {  // uniform (the rings' [ex, ex + nb) tile [0, total): one lane holds cpos)
      const uint64_t ne = __ballot(nb > 0);
      const uint32_t rank = (uint32_t)__popcll(ne & ((1ull << lane) - 1ull));
      const uint32_t kl = (uint32_t)__ffsll((long long)hm) - 1u;
      const uint32_t rk = __shfl(rank, kl, 64);  // the rank of the ring holding cpos
      const uint64_t tm = __ballot(nb > 0 && rank == rk + (KSEG - (VPK - 1)) - 1u);
      const uint32_t tl = tm ? (uint32_t)__ffsll((long long)tm) - 1u : 63u;
      const uint32_t end_t = __shfl(in, tl, 64);
      const uint32_t pos_s = tm ? end_t : total;
      const uint32_t e = min(min(cpos + KV, total), (pos_s + VPK - 1) / VPK * VPK);
      const uint32_t nkb = (e - cpos + VPK - 1) / VPK;
      const bool seg = nb > 0 && lane >= kl && ex < e && rank - rk < KSEG;
      if (seg) {
        const uint32_t sg = rank - rk;
        const uint32_t a = max(ex, cpos), b = min(in, e);
        const uint32_t du = a - cpos, g0 = b0 + (a - ex);
        L.dring[db][sg] = r_begin + lane;
        L.du0[db][sg] = (uint16_t)du;
        L.dnb[db][sg] = (uint16_t)(b - a);
        for (uint32_t t = 0; t < b - a; ++t) L.dgb[db][du + t] = g0 + t;
        if (b == e)  // the last segment repeats its last block up to the k-block boundary
          for (uint32_t p = e - cpos; p < nkb * VPK; ++p) L.dgb[db][p] = g0 + (b - a - 1);
      }
      nseg = (uint32_t)__popcll(__ballot(seg));
      nvb = e - cpos;
      cpos = e;
    }
    if (lane == 0) {
      L.dn[db][0] = (nvb + VPK - 1) / VPK;
      L.dn[db][1] = nseg;
      L.dn[db][2] = nvb;
    }
  };

  // ---- per-chunk steps -------------------------------------------------
  // stage chunk positions [from, 32 nkb) of descriptor db into LDS (keys past
  // the samples and past the chunk's nvb value blocks are the empty key,
  // beyond every window)
  auto stage = [&](uint32_t db, uint32_t from, uint32_t nkb, uint32_t nvb) {
    for (uint32_t p = from + tid; p < nkb * 32u; p += nthreads) {
      const uint32_t g = gidx(db, p);
      if (g < n_all && p < nvb * kVB) {
        const hcg_sample q = B.rec[g];
        s_x[p] = q.x;
        s_y[p] = q.y;
        s_z[p] = q.z;
        s_key[p] = B.keys[g];
      } else {
        s_x[p] = s_y[p] = s_z[p] = 0.0;
        s_key[p] = 0xFFFFFFFFu;
      }
    }
  };
  // zero the weight tiles' used width
  auto clear = [&](uint32_t nkb) {
    const uint32_t width8 = nkb * 4;  // 16-byte groups per row
    for (uint32_t g = tid; g < kMcCols * width8; g += nthreads) {
      const uint32_t rr = g / width8, cc = (g % width8) * 8;
      *reinterpret_cast<uint4*>(&s_wh[rr * RS + cc]) = uint4{0u, 0u, 0u, 0u};
      *reinterpret_cast<uint4*>(&s_wl[rr * RS + cc]) = uint4{0u, 0u, 0u, 0u};
    }
  };
  // reset the row blocks' k-block spans of buffer db (read last by the
  // contraction two chunks back, written by the bands of the next chunk)
  auto reset_spans = [&](uint32_t db) {
    if (tid < 4) {
      s_mb[db][tid][0] = 0xFFFFFFFFu;
      s_mb[db][tid][1] = 0u;
    }
  };
  // each cell's band in each segment of the chunk (wave w: segments w, w +
  // nw, ...), and the row blocks' k-block spans
  auto bands = [&](uint32_t db, uint32_t nseg) {
    for (uint32_t sg = wave; sg < nseg; sg += nw) {
      uint32_t a, b;
      ring_cell_window(P, B, m, L.win[lane], L.dring[db][sg], a, b);
      uint32_t ja = 0, jb = 0;
      if (a <= b) {
        const uint32_t p0 = (uint32_t)L.du0[db][sg] * kVB, p1 = p0 + (uint32_t)L.dnb[db][sg] * kVB;
        ja = lower_bound_u32(s_key, p0, p1, a);
        jb = upper_bound_u32(s_key, ja, p1, b);
        if (ja < jb) {
          atomicMin(&s_mb[db][lane >> 4][0], ja >> 5);
          atomicMax(&s_mb[db][lane >> 4][1], (jb + 31) >> 5);
        }
      }
      L.band[sg][lane] = ja | (jb << 16);
    }
  };
  // the lean form (kKtGaussLean): the tile loop's p.q candidate test on
  // h = 1 - p.q and gauss_weight_lean, the guard band after the weight
  const double npx = -px, npy = -py, npz = -pz;
  const double h_hi = 0.5 * P.c_hi, h_lo = 0.5 * P.c_lo, k0h = 2.0 * P.k0;
  const bool lean = kt == kKtGaussLean && B.w_s == nullptr;
  const double k1v = in_vgprs(4.0 * P.k1), e3v = in_vgprs(kExpE3);
  // the weight wt of pair (c, j) as hi + lo in the value type (~2x its
  // significand), its two halves stored in the weight tiles and summed
  auto store_weight = [&](uint32_t j, double wt) {
    const float wf = (float)wt;
    const uint16_t hi = round16(wf, bf);
    const float hif = widen16(hi, bf);
    // wf - hif is exact (Sterbenz), and hif + lof is exact in FP32 (at most
    // 2 x 11 significant bits, the second ~2^-8 below the first)
    const uint16_t lo = round16(wf - hif, bf);
    s_wh[c * RS + j] = hi;
    s_wl[c * RS + j] = lo;
    wsum += (double)(hif + widen16(lo, bf));
  };
  // weights of this thread's cell over its bands in the chunk (after the
  // bands barrier)
  auto weigh = [&](uint32_t db, uint32_t nseg) {
    if (lean) {  // uniform
      for (uint32_t sg = 0; sg < nseg; ++sg) {
        const uint32_t bd = L.band[sg][c];
        const uint32_t ja = bd & 0xFFFFu, jb = bd >> 16;
        for (uint32_t j = ja + sub; j < jb; j += S) {
          const double h = fma(npx, s_x[j], fma(npy, s_y[j], fma(npz, s_z[j], 1.0)));
          if (h < h_hi) {
            double wt = gauss_weight_lean(h, k0h, k1v, e3v, s_exp);
            if (__builtin_amdgcn_ballot_w64(!(h < h_lo))) {  // some lane in the guard band (rare)
              asm volatile("; guard band" ::);                // (keeps the branch above a scalar one)
              if (!(h < h_lo) && !radius_test_ref(B.xwcs, B.ywcs, B.lon_s, B.lat_s, cell, gidx(db, j),
                                                  P.sphere_radius))
                wt = 0.0;
            }
            store_weight(j, wt);
          }
        }
      }
      return;
    }
    for (uint32_t sg = 0; sg < nseg; ++sg) {
      const uint32_t bd = L.band[sg][c];
      const uint32_t ja = bd & 0xFFFFu, jb = bd >> 16;
      for (uint32_t j = ja + sub; j < jb; j += S) {
        const double dx = px - s_x[j], dy = py - s_y[j], dz = pz - s_z[j];
        const double c2 = fma(dx, dx, fma(dy, dy, dz * dz));
        if (c2 < P.c_hi &&
            (c2 < P.c_lo || radius_test_ref(B.xwcs, B.ywcs, B.lon_s, B.lat_s, cell, gidx(db, j), P.sphere_radius))) {
          // (the sample's sorted index only for the kernels that read its coordinates)
          const bool need_i = kt == HCG_GAUSS2D || kt == HCG_TAPERED_SINC;
          double wt = pair_weight_any(P, B, kt, c2, s_exp, cell, need_i ? gidx(db, j) : 0u);
          if (B.w_s) wt *= B.w_s[gidx(db, j)];
          store_weight(j, wt);
        }
      }
    }
  };
  // the contraction over the chunk's k-blocks; bn holds block 0's channel
  // values on entry
  auto contract = [&](uint32_t db, uint32_t nkb, vec* b0) {
    // the row blocks' k-block spans (wave-uniform: scalar compares)
    uint32_t mblo[4], mbhi[4];
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      mblo[mb] = __builtin_amdgcn_readfirstlane(s_mb[db][mb][0]);
      mbhi[mb] = __builtin_amdgcn_readfirstlane(s_mb[db][mb][1]);
    }
    // block u's MFMAs on B operand b
    auto mfmas = [&](uint32_t u, const vec* b) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        if (!(u >= mblo[mb] && u < mbhi[mb])) continue;
        // A: the row block's hi and lo weights, from LDS
        const uint32_t off = (mb * 16 + (lane & 15)) * RS + u * 32 + koff;
        const vec ah = *reinterpret_cast<const vec*>(&s_wh[off]);
        const vec al = *reinterpret_cast<const vec*>(&s_wl[off]);
#pragma unroll
        for (int nb = 0; nb < kMcNB; ++nb) {
          acc[mb][nb] = mfma16(bf, ah, b[nb], acc[mb][nb]);
          acc[mb][nb] = mfma16(bf, al, b[nb], acc[mb][nb]);
        }
      }
    };
    // software pipeline over two register sets (no copies): block u's
    // MFMAs run while block u+1's channel values load.  The loads are
    // unconditional (past the last block they repeat it, unused), so every
    // MFMA waits only for its own operand's loads
    vec b1[kMcNB];
    const uint32_t last = nkb - 1;
    for (uint32_t u = 0; u < nkb; u += 2) {
      load_b_rows(b1, kblock(db, min(u + 1, last)), lofs);
      mfmas(u, b0);
      if (u + 1 >= nkb) break;
      load_b_rows(b0, kblock(db, min(u + 2, last)), lofs);
      mfmas(u + 1, b1);
    }
  };

  // ---- chunks ---------------------------------------------------------------
  // barriers per chunk: staging (B1, only when records are left to stage),
  // bands (B2), weights (B3), contraction done (T).  The next chunk's
  // descriptor is built by wave 0 between B2 and B3; its first nthreads
  // records load during the contraction and go to LDS after it (the
  // contraction reads only weight tiles and channel values), so a chunk of
  // at most nthreads samples needs no staging of its own.
  uint32_t db = 0, staged = 0;
  if (wave == 0) {
    if (par_desc) {
      ring_list();
      build_desc_par(0);
    } else {
      build_desc(0);
    }
  }
  reset_spans(0);
  __syncthreads();
  for (;;) {
    const uint32_t nkb = L.dn[db][0], nseg = L.dn[db][1], nvb = L.dn[db][2];
    if (nkb == 0) break;  // uniform
    clear(nkb);
    reset_spans(db ^ 1u);
    if (staged < nkb * 32u) {  // uniform
      stage(db, staged, nkb, nvb);
      __syncthreads();  // B1
    }
    bands(db, nseg);
    __syncthreads();  // B2
    // channel values of the chunk's first k-block, in flight while the weights are computed
    vec bn[kMcNB];
    if (ch_base < nchan) load_b_rows(bn, kblock(db, 0), lofs);
    if (wave == 0) {
      if (par_desc)
        build_desc_par(db ^ 1u);
      else
        build_desc(db ^ 1u);
    }
    weigh(db, nseg);
    __syncthreads();  // B3
    const uint32_t nn = L.dn[db ^ 1u][0] * 32u, nvb_next = L.dn[db ^ 1u][2];
    staged = min(nn, nthreads);
    const bool pre = tid < staged;
    hcg_sample pq{};
    uint32_t pkey = 0xFFFFFFFFu;
    if (pre) {
      const uint32_t g = gidx(db ^ 1u, tid);
      if (g < n_all && tid < nvb_next * kVB) {
        pq = B.rec[g];
        pkey = B.keys[g];
      }
    }
    if (ch_base < nchan) contract(db, nkb, bn);
    if (pre) {
      s_x[tid] = pq.x;
      s_y[tid] = pq.y;
      s_z[tid] = pq.z;
      s_key[tid] = pkey;
    }
    db ^= 1u;
    __syncthreads();  // T
  }

  // per-cell weight sums over the waves (exact in FP64 in any order: every
  // weight is hi + lo of two 16-bit values); the weight tiles are free now
  __syncthreads();
  double* wpart = reinterpret_cast<double*>(L.w);
  wpart[wave * kMcCols + lane] = wsum;
  __syncthreads();
  if (wave == 0) {
    double t = 0.0;
    for (uint32_t w = 0; w < nw; ++w) t += wpart[w * kMcCols + lane];
    // the normalisation as one multiply per value: v * (1 / t) is within an
    // FP64 ulp or two of v / t, so the FP32 result is the quotient's except
    // where it lies that close to an FP32 rounding boundary (1 ulp then);
    // t = 0 gives inf, so an empty cell is 0 * inf = NaN as 0 / 0 is
    s_inv[lane] = normalize ? 1.0 / t : 1.0;
    if (group == 0 && wsum_out && live) wsum_out[cell] = (float)t;
  }
  __syncthreads();
  // lane l holds cells 16 mb + 4 (l / 16) + j, j = 0..3: four consecutive
  // cells of one map row (tile widths are multiples of 4), stored as one
  // 16-byte piece when the map side is a multiple of 4 (then aligned)
  auto value = [&](int mb, int nb, int j, uint32_t cl0) {
    return (float)((double)acc[mb][nb][j] * s_inv[cl0 + j]);
  };
  constexpr uint32_t kEpiFloats = kMcCols * 64;  // one wave's 64 cells x 64 channels
  if (STAGED && (P.side & 3u) == 0 && nw * kEpiFloats * sizeof(float) <= sizeof(L.w)) {  // uniform
    // Through LDS: each wave stages its cells x channels by channel (the
    // 16-byte quads of a channel's row XOR-swizzled by the channel, so the
    // 16 channels of a write, and the 16 quads of a read, fall in 16
    // distinct bank groups), then stores four channels per instruction, a
    // channel's 64 cells as 16 pieces of 4 consecutive cells of a map row.
    // An instruction then spans 4 channel planes instead of 16 (each plane
    // its own pages: 4 MB at 1024^2), so it needs fewer address
    // translations.  The weight tiles are free: the last chunk's
    // contraction and the weight-sum reduction are behind barriers.
    float* T = reinterpret_cast<float*>(L.w) + (size_t)wave * kEpiFloats;
    if (ch_base < nchan) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const uint32_t q = mb * 4 + (lane >> 4), cl0 = q * 4;
#pragma unroll
        for (int nb = 0; nb < kMcNB; ++nb) {
          const uint32_t cr = nb * 16 + (lane & 15);
          *reinterpret_cast<float4*>(&T[cr * 64 + ((q ^ (cr & 15u)) << 2)]) =
              make_float4(value(mb, nb, 0, cl0), value(mb, nb, 1, cl0), value(mb, nb, 2, cl0), value(mb, nb, 3, cl0));
        }
      }
    }
    __syncthreads();
    if (ch_base < nchan) {
      const uint32_t q = lane & 15, cl0 = q * 4;
      const uint32_t rowj = y0 + (cl0 >> tc_shift), col0 = x0 + (cl0 & tc_mask);
      const bool in_map = col0 < P.side && rowj < P.side;
#pragma unroll
      for (uint32_t k = 0; k < 16; ++k) {
        const uint32_t cr = k * 4 + (lane >> 4), ch = ch_base + cr;
        const float4 v = *reinterpret_cast<const float4*>(&T[cr * 64 + ((q ^ (cr & 15u)) << 2)]);
        if (in_map && ch < nchan)
          *reinterpret_cast<float4*>(cube + (size_t)ch * M + (size_t)rowj * P.side + col0) = v;
      }
    }
    return;
  }
  if (ch_base < nchan) {
    if ((P.side & 3u) == 0) {  // uniform
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const uint32_t cl0 = mb * 16 + (lane >> 4) * 4;
        const uint32_t rowj = y0 + (cl0 >> tc_shift), col0 = x0 + (cl0 & tc_mask);
        if (col0 >= P.side || rowj >= P.side) continue;
#pragma unroll
        for (int nb = 0; nb < kMcNB; ++nb) {
          const uint32_t ch = ch_base + nb * 16 + (lane & 15);
          if (ch >= nchan) continue;
          float* dst = cube + (size_t)ch * M + (size_t)rowj * P.side + col0;
          *reinterpret_cast<float4*>(dst) =
              make_float4(value(mb, nb, 0, cl0), value(mb, nb, 1, cl0), value(mb, nb, 2, cl0), value(mb, nb, 3, cl0));
        }
      }
    } else {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const uint32_t cl0 = mb * 16 + (lane >> 4) * 4;
        const uint32_t rowj = y0 + (cl0 >> tc_shift), col0 = x0 + (cl0 & tc_mask);
        if (col0 >= P.side || rowj >= P.side) continue;
#pragma unroll
        for (int nb = 0; nb < kMcNB; ++nb) {
          const uint32_t ch = ch_base + nb * 16 + (lane & 15);
          if (ch >= nchan) continue;
          float* dst = cube + (size_t)ch * M + (size_t)rowj * P.side + col0;
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (col0 + j < P.side) dst[j] = value(mb, nb, j, cl0);
        }
      }
    }
  }
}


// One job, one launch: grid = (tiles, channel groups), tiles in XCD-contiguous
// runs (block b runs on XCD b % 8, so neighbouring tiles share an L2).
template <int DT, uint32_t CHUNK>
__global__ void __launch_bounds__(64 * kMcMaxWaves) k_grid_mfma(MfmaJob J) {
  __shared__ MfmaLds<CHUNK> L;
  __shared__ double s_exp[64];
  if (threadIdx.x < 64) exp_table_load(s_exp);
  uint32_t tile = blockIdx.x;
  const uint32_t ntiles = J.ntiles;
  if ((ntiles & 7u) == 0) tile = (blockIdx.x & 7u) * (ntiles >> 3) + (blockIdx.x >> 3);
  mfma_block<DT, CHUNK, true>(J, tile, blockIdx.y, L, s_exp);
}

// Many jobs, one launch: a persistent grid pulls work items (job, tile,
// channel group) from a device queue in list order until the list is
// exhausted; every workgroup exits when its pull passes nitems.  The list
// is ordered longest-first by the host, so the tail holds the short items.
// A work item of a one-channel tile-path job.  One stream per tile
// (tile_streams 1): tiles item * tpi + w, one per wave w < tpi, each in its
// own kTileLdsBytes slice of the workgroup's LDS (the host keeps tpi within
// tile_items_per_block).  Several streams: tile `item`, its streams over
// the waves (one_channel_tile; the same sums for any number of waves).
// two 256-sample workgroups (with the exp table and the queue slot) share a
// CU's 160 KB of LDS; one 512-sample workgroup fills it
static_assert(2 * (sizeof(MfmaLds<256>) + 64 * sizeof(double) + 16) <= 160 * 1024, "two CHUNK-256 workgroups per CU");
static_assert(sizeof(MfmaLds<512>) + 64 * sizeof(double) + 16 <= 160 * 1024, "one CHUNK-512 workgroup per CU");
static_assert(sizeof(MfmaLds<128>) >= one_channel_lds_bytes(kOneChannelStreams, kOneChannelStreams),
              "the smallest workgroup LDS holds a one-channel tile's streams");

// Called, not inlined: the persistent kernel's two multi-channel block
// copies then keep its register budget to themselves (22 VGPRs spilled
// instead of 37; the 64-job batch 171.0 -> 164.7 ms per replay).
template <uint32_t CHUNK>
__device__ __noinline__ void tile_item(const MfmaJob& J, uint32_t item, MfmaLds<CHUNK>& L, const double* s_exp) {
  const uint32_t nw = blockDim.x >> 6, wave = threadIdx.x >> 6, lane = __lane_id();
  char* lds = reinterpret_cast<char*>(&L);
  const TileOut O{static_cast<const uint16_t*>(J.vs), J.cube, J.wsum, J.dtype == HCG_DTYPE_BF16 ? 1u : 0u,
                  J.normalize, J.nchan};
  if (J.nchan > 1) {  // few-channel tiles, one per wave
    const uint32_t tpi = J.tiles_per_item;
    const uint32_t t = item * tpi + wave;
    if (wave < tpi && (wave + 1) * kFewTileLdsBytes <= sizeof(MfmaLds<CHUNK>) && t < J.ntiles)
      few_channel_tile(J.P, J.B, O, t, J.tiles_x, lds + wave * kFewTileLdsBytes, s_exp);
    return;
  }
  const uint32_t streams = J.tile_streams;
  // this wave's tile and streams (one call site of the tile body, so the
  // persistent kernel carries one copy of it)
  uint32_t tile, q0, q1, qstep, active;
  if (streams > 1) {
    active = nw < streams ? nw : streams;
    tile = item;
    q0 = wave;
    q1 = wave < active ? streams : 0;
    qstep = active;
  } else {
    const uint32_t tpi = J.tiles_per_item;
    active = tpi;
    tile = item * tpi + wave;
    q0 = 0;
    q1 = (wave < tpi && (wave + 1) * kTileLdsBytes <= sizeof(MfmaLds<CHUNK>) && tile < J.ntiles) ? 1 : 0;
    qstep = 1;
  }
  double* red = reinterpret_cast<double*>(lds + active * kTileLdsBytes);
  double d = 0.0, w = 0.0;
  if (q0 < q1) {
    double* s_x = reinterpret_cast<double*>(lds + wave * kTileLdsBytes);
    double* s_y = s_x + kStageCap;
    double* s_z = s_y + kStageCap;
    double* s_v = s_z + kStageCap;
    uint32_t* s_key = reinterpret_cast<uint32_t*>(s_v + kStageCap);
    for (uint32_t q = q0; q < q1; q += qstep) {
      grid_tile<HCG_GAUSS1D, false, 2, true>(J.P, J.B, tile, J.tiles_x, s_x, s_y, s_z, s_v, s_key, s_exp, O,
                                             RingStream{q, streams}, &d, &w);
      if (streams > 1) {
        red[(q * 64 + lane) * 2] = d;
        red[(q * 64 + lane) * 2 + 1] = w;
      }
    }
  }
  if (streams > 1) {
    // the streams' sums in stream order, as one_channel_tile adds them
    __syncthreads();
    if (wave != 0) return;
    d = red[lane * 2];
    w = red[lane * 2 + 1];
    for (uint32_t q = 1; q < streams; ++q) {
      d += red[(q * 64 + lane) * 2];
      w += red[(q * 64 + lane) * 2 + 1];
    }
  } else if (q0 >= q1) {
    return;
  }
  tile_write(J.P, tile, J.tiles_x, O, d, w);
}

uint32_t tile_items_per_block(uint32_t chunk, uint32_t waves, size_t tile_bytes) {
  const size_t lds = chunk == 512 ? sizeof(MfmaLds<512>) : chunk == 256 ? sizeof(MfmaLds<256>) : sizeof(MfmaLds<128>);
  const uint32_t cap = (uint32_t)(lds / tile_bytes);
  return std::max(1u, std::min(waves, cap));
}

template <uint32_t CHUNK>
__global__ void __launch_bounds__(64 * kMcMaxWaves)
    k_grid_mfma_batch(const MfmaJob* __restrict__ jobs, const MfmaItem* __restrict__ items, uint32_t nitems,
                      uint32_t* __restrict__ head) {
  __shared__ MfmaLds<CHUNK> L;
  __shared__ double s_exp[64];
  __shared__ uint32_t s_item;
  if (threadIdx.x < 64) exp_table_load(s_exp);
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(head, 1u);
    __syncthreads();
    const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= nitems) break;  // uniform: every thread read the same value
    const uint32_t jb = __builtin_amdgcn_readfirstlane(items[it].job);
    const uint32_t tile = __builtin_amdgcn_readfirstlane(items[it].tile);
    const uint32_t group = __builtin_amdgcn_readfirstlane(items[it].group);
    // one copy of the block per value type: with the type known at compile
    // time the contraction issues its MFMAs back to back (a type tested at
    // run time inside the block made the 2M x 1024 cube 15% slower)
    if (__builtin_amdgcn_readfirstlane(jobs[jb].tiles_per_item))
      tile_item(jobs[jb], tile, L, s_exp);
    else if (__builtin_amdgcn_readfirstlane(jobs[jb].dtype) == HCG_DTYPE_F16)
      mfma_block<kDtF16, CHUNK, false>(jobs[jb], tile, group, L, s_exp);
    else
      mfma_block<kDtBf16, CHUNK, false>(jobs[jb], tile, group, L, s_exp);
    __syncthreads();  // LDS and s_item are free for the next item
  }
}

int mfma_batch_blocks(uint32_t chunk, uint32_t waves, uint32_t* blocks) {
  if (!(waves == 1 || waves == 2 || waves == 4 || waves == 8))
    return set_error(HCG_ERR_INVALID, "batch: waves must be 1, 2, 4 or 8");
  if (!(chunk == 128 || chunk == 256 || chunk == 512))
    return set_error(HCG_ERR_INVALID, "batch: chunk must be 128, 256 or 512");
  const void* fn = chunk == 512 ? reinterpret_cast<const void*>(&k_grid_mfma_batch<512>)
                   : chunk == 256 ? reinterpret_cast<const void*>(&k_grid_mfma_batch<256>)
                                  : reinterpret_cast<const void*>(&k_grid_mfma_batch<128>);
  int dev = 0, cus = 0, per_cu = 0;
  HCG_CHECK_HIP(hipGetDevice(&dev));
  HCG_CHECK_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HCG_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64 * waves, 0));
  if (per_cu < 1) per_cu = 1;
  *blocks = (uint32_t)(cus * per_cu);
  return HCG_OK;
}

int launch_mfma_batch(uint32_t chunk, uint32_t waves, uint32_t blocks, const MfmaJob* jobs, const MfmaItem* items,
                      uint32_t nitems, uint32_t* head, hipStream_t s) {
  if (!(waves == 1 || waves == 2 || waves == 4 || waves == 8))
    return set_error(HCG_ERR_INVALID, "batch: waves must be 1, 2, 4 or 8");
  if (!(chunk == 128 || chunk == 256 || chunk == 512))
    return set_error(HCG_ERR_INVALID, "batch: chunk must be 128, 256 or 512");
  if (blocks == 0) return set_error(HCG_ERR_INVALID, "batch: zero blocks");
  if (chunk == 512)
    k_grid_mfma_batch<512><<<blocks, 64 * waves, 0, s>>>(jobs, items, nitems, head);
  else if (chunk == 256)
    k_grid_mfma_batch<256><<<blocks, 64 * waves, 0, s>>>(jobs, items, nitems, head);
  else
    k_grid_mfma_batch<128><<<blocks, 64 * waves, 0, s>>>(jobs, items, nitems, head);
  HCG_CHECK_LAUNCH("k_grid_mfma_batch");
  return HCG_OK;
}

// Channel gather into the blocked layout: sorted slot i of channel ch goes
// to block i / 8, row ch, column i % 8 ([nblk][nchan][8], kVB = 8 slots per
// block, zero-padded to nblk = hcg_channels_padded_len(n) / 8 blocks).  A
// workgroup takes kGatherSlots sorted slots x kGatherChans channels through
// LDS: thread t gathers slot t of each channel (consecutive lanes read the
// caller's values at consecutive permuted positions, 2 bytes each), and then
// writes block rows (16 bytes each), consecutive threads the consecutive
// channel rows of one block.
constexpr uint32_t kGatherChans = 32;
constexpr uint32_t kGatherSlots = 256;
constexpr uint32_t kGatherRow = kGatherSlots + 8;  // LDS row stride (16-byte pad against bank conflicts)

__device__ inline size_t blocked_index(uint32_t i, uint32_t ch, uint32_t nchan) {
  return ((size_t)(i / kVB) * nchan + ch) * kVB + (i % kVB);
}

template <typename T>
__device__ inline void gather_cm_block(uint32_t bx, uint32_t by, uint32_t n, uint32_t n_pad, uint32_t nchan,
                                       const uint32_t* __restrict__ perm, const T* __restrict__ in,
                                       T* __restrict__ out, uint32_t* __restrict__ nonfinite,
                                       uint16_t* __restrict__ lds /* [kGatherChans][kGatherRow] */) {
  static_assert(sizeof(T) == 2, "16-bit values");
  const uint32_t t = threadIdx.x;
  const uint32_t i0 = bx * kGatherSlots;
  const uint32_t ch0 = by * kGatherChans;
  const uint32_t nc = min(kGatherChans, nchan - ch0);
  const uint32_t i = i0 + t;
  uint32_t j = 0;
  if (i < n) {
    j = perm[i];
    if (j >= n) j = i;  // a permutation never holds this; keeps every read in bounds regardless
  }
  // all of the thread's loads first (up to kGatherChans in flight), then
  // the LDS stores
  uint16_t x[kGatherChans];
  const bool live = i < n;
  const T* src = in + (size_t)ch0 * n + j;
#pragma unroll
  for (uint32_t c = 0; c < kGatherChans; ++c)
    x[c] = (live && c < nc) ? __builtin_bit_cast(uint16_t, src[(size_t)c * n]) : (uint16_t)0;
  uint32_t bad = 0;
#pragma unroll
  for (uint32_t c = 0; c < kGatherChans; ++c) {
    if (c < nc) {
      bad += isfinite((float)__builtin_bit_cast(T, x[c])) ? 0u : 1u;
      lds[c * kGatherRow + t] = x[c];
    }
  }
  if (nonfinite && bad) atomicAdd(nonfinite, bad);
  __syncthreads();
  // pieces: (block b of the workgroup's slots, channel c) -> one block row
  // of kVB values; the output is 16-byte aligned when the buffer is (a row
  // is 16 bytes)
  static_assert(kVB == 8 && kGatherSlots % kVB == 0, "a block row is one 16-byte piece");
  const bool vec = (reinterpret_cast<uintptr_t>(out) & 15u) == 0;
  const uint32_t npieces = (kGatherSlots / kVB) * nc;
  for (uint32_t p = t; p < npieces; p += 256) {
    const uint32_t b = p / nc, c = p % nc;
    const uint32_t blk = i0 / kVB + b;
    if (blk * kVB >= n_pad) continue;
    const uint16_t* src = lds + c * kGatherRow + b * kVB;
    T* dst = out + ((size_t)blk * nchan + ch0 + c) * kVB;
    if (vec) {
      *reinterpret_cast<uint4*>(dst) = *reinterpret_cast<const uint4*>(src);
    } else {
#pragma unroll
      for (uint32_t e = 0; e < 8; ++e) dst[e] = __builtin_bit_cast(T, src[e]);
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_gather_channels_cm(uint32_t n, uint32_t n_pad, uint32_t nchan,
                                                            const uint32_t* __restrict__ perm,
                                                            const T* __restrict__ in, T* __restrict__ out,
                                                            uint32_t* __restrict__ nonfinite) {
  __shared__ __attribute__((aligned(16))) uint16_t lds[kGatherChans * kGatherRow];
  gather_cm_block<T>(blockIdx.x, blockIdx.y, n, n_pad, nchan, perm, in, out, nonfinite, lds);
}

// Sample-major input ([n][nchan]): 64 x 64 tiles transposed through LDS.
template <typename T>
__device__ inline void gather_sm_block(uint32_t bx, uint32_t by, uint32_t n, uint32_t n_pad, uint32_t nchan,
                                       const uint32_t* __restrict__ perm, const T* __restrict__ in,
                                       T* __restrict__ out, uint32_t* __restrict__ nonfinite, T (*tile)[65]) {
  const uint32_t i0 = bx * 64, ch0 = by * 64;
  const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  for (uint32_t k = ty; k < 64; k += 4) {
    const uint32_t i = i0 + k, ch = ch0 + tx;
    T v = (T)0.0f;
    if (i < n && ch < nchan) {
      uint32_t j = perm[i];
      if (j >= n) j = i;
      v = in[(size_t)j * nchan + ch];
      if (nonfinite && !isfinite((float)v)) atomicAdd(nonfinite, 1u);
    }
    tile[k][tx] = v;
  }
  __syncthreads();
  for (uint32_t k = ty; k < 64; k += 4) {
    const uint32_t ch = ch0 + k, i = i0 + tx;
    if (ch < nchan && i < n_pad) out[blocked_index(i, ch, nchan)] = tile[tx][k];
  }
}

template <typename T>
__global__ void __launch_bounds__(256) k_gather_channels_sm(uint32_t n, uint32_t n_pad, uint32_t nchan,
                                                            const uint32_t* __restrict__ perm,
                                                            const T* __restrict__ in, T* __restrict__ out,
                                                            uint32_t* __restrict__ nonfinite) {
  __shared__ T tile[64][65];
  gather_sm_block<T>(blockIdx.x, blockIdx.y, n, n_pad, nchan, perm, in, out, nonfinite, tile);
}

// Every job's channel gather and tile plan of a batch in one launch
// (prep_ranges): the plans' dependent key searches are latency-bound and
// the gathers bandwidth-bound, so sharing the chip lets each hide the
// other.  A range is a run of workgroups of one kind for one job; each
// workgroup finds its range by binary search over the range starts.
__global__ void __launch_bounds__(256) k_batch_prep(const PrepRange* __restrict__ ranges, uint32_t nranges,
                                                    const PrepJob* __restrict__ pj,
                                                    const MfmaJob* __restrict__ jobs) {
  static_assert(kGatherChans * kGatherRow >= 64 * 65, "the sample-major tile fits the channel-major one");
  __shared__ __attribute__((aligned(16))) uint16_t smem[kGatherChans * kGatherRow];
  uint16_t(*tile)[65] = reinterpret_cast<uint16_t(*)[65]>(smem);
  uint32_t lo = 0, hi = nranges;  // last range with first <= blockIdx.x
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (ranges[mid].first <= blockIdx.x)
      lo = mid;
    else
      hi = mid;
  }
  const PrepRange R = ranges[lo];
  const uint32_t lb = blockIdx.x - R.first;
  const PrepJob& G = pj[R.job];
  const MfmaJob& J = jobs[R.job];
  const uint32_t bx = lb % R.gx, by = lb / R.gx;
  switch (R.kind) {
    case kPrepPlan: {
      const uint32_t t = lb * 4 + (threadIdx.x >> 6);  // a wave per tile
      if (t < J.ntiles) plan_tile(J, t);
      break;
    }
    case kPrepGatherCm:
      if (J.dtype == HCG_DTYPE_BF16)
        gather_cm_block<__bf16>(bx, by, G.n, J.n_pad, J.nchan, G.perm, static_cast<const __bf16*>(G.in),
                                static_cast<__bf16*>(const_cast<void*>(J.vs)), G.nonfinite, smem);
      else
        gather_cm_block<_Float16>(bx, by, G.n, J.n_pad, J.nchan, G.perm, static_cast<const _Float16*>(G.in),
                                  static_cast<_Float16*>(const_cast<void*>(J.vs)), G.nonfinite, smem);
      break;
    default:  // kPrepGatherSm
      if (J.dtype == HCG_DTYPE_BF16)
        gather_sm_block<__bf16>(bx, by, G.n, J.n_pad, J.nchan, G.perm, static_cast<const __bf16*>(G.in),
                                static_cast<__bf16*>(const_cast<void*>(J.vs)), G.nonfinite,
                                reinterpret_cast<__bf16(*)[65]>(tile));
      else
        gather_sm_block<_Float16>(bx, by, G.n, J.n_pad, J.nchan, G.perm, static_cast<const _Float16*>(G.in),
                                  static_cast<_Float16*>(const_cast<void*>(J.vs)), G.nonfinite,
                                  reinterpret_cast<_Float16(*)[65]>(tile));
      break;
  }
}

int prep_ranges(uint32_t njobs, const MfmaJob* jobs, const uint32_t* layouts, std::vector<PrepRange>& out,
                uint32_t* blocks) {
  out.clear();
  uint64_t first = 0;
  auto add = [&](uint32_t j, uint32_t kind, uint32_t gx, uint64_t blocks) {
    if (blocks == 0) return;
    out.push_back(PrepRange{(uint32_t)first, j, kind, gx});
    first += blocks;
  };
  // plans first: their workgroups run longest
  for (uint32_t j = 0; j < njobs; ++j)
    if (jobs[j].plan && jobs[j].nchan) add(j, kPrepPlan, 1, ceil_div_u32(jobs[j].ntiles, 4));
  for (uint32_t j = 0; j < njobs; ++j) {
    const MfmaJob& J = jobs[j];
    if (J.nchan == 0 || J.n_pad == 0) continue;
    if (layouts[j] == 0) {
      const uint32_t gx = ceil_div_u32(J.n_pad, kGatherSlots), gy = ceil_div_u32(J.nchan, kGatherChans);
      add(j, kPrepGatherCm, gx, (uint64_t)gx * gy);
    } else {
      const uint32_t gx = ceil_div_u32(J.n_pad, 64), gy = ceil_div_u32(J.nchan, 64);
      add(j, kPrepGatherSm, gx, (uint64_t)gx * gy);
    }
  }
  if (first >= 0x7FFFFFFFull) return set_error(HCG_ERR_INVALID, "hcg_batch_create: too many prep workgroups");
  *blocks = (uint32_t)first;
  return HCG_OK;
}

int launch_batch_prep(const PrepRange* ranges, uint32_t nranges, uint32_t blocks, const PrepJob* pj,
                      const MfmaJob* jobs, hipStream_t s) {
  if (nranges == 0 || blocks == 0) return HCG_OK;
  k_batch_prep<<<blocks, 256, 0, s>>>(ranges, nranges, pj, jobs);
  HCG_CHECK_LAUNCH("k_batch_prep");
  return HCG_OK;
}

}  // namespace hcg

using namespace hcg;

extern "C" uint32_t hcg_channels_padded_len(uint32_t n) { return ((n + 31u) & ~31u) + 64u; }

extern "C" int hcg_gather_channels(uint32_t n, uint32_t nchan, const uint32_t* perm, const void* values, int dtype,
                                   int layout, void* values_sorted, uint32_t* nonfinite, void* stream) {
  if (n == 0 || nchan == 0) return HCG_OK;
  HCG_REQUIRE(perm && values && values_sorted, "hcg_gather_channels: NULL buffer");
  HCG_REQUIRE(dtype == HCG_DTYPE_BF16 || dtype == HCG_DTYPE_F16, "hcg_gather_channels: dtype must be BF16 or F16");
  HCG_REQUIRE(layout == 0 || layout == 1, "hcg_gather_channels: layout must be 0 (channel-major) or 1 (sample-major)");
  const uint32_t n_pad = hcg_channels_padded_len(n);
  hipStream_t s = as_stream(stream);
  if (layout == 0) {
    const dim3 grid(ceil_div_u32(n_pad, kGatherSlots), ceil_div_u32(nchan, kGatherChans));
    if (dtype == HCG_DTYPE_BF16)
      k_gather_channels_cm<__bf16><<<grid, 256, 0, s>>>(n, n_pad, nchan, perm, static_cast<const __bf16*>(values),
                                                        static_cast<__bf16*>(values_sorted), nonfinite);
    else
      k_gather_channels_cm<_Float16><<<grid, 256, 0, s>>>(n, n_pad, nchan, perm,
                                                          static_cast<const _Float16*>(values),
                                                          static_cast<_Float16*>(values_sorted), nonfinite);
  } else {
    const dim3 grid(ceil_div_u32(n_pad, 64), ceil_div_u32(nchan, 64));
    if (dtype == HCG_DTYPE_BF16)
      k_gather_channels_sm<__bf16><<<grid, 256, 0, s>>>(n, n_pad, nchan, perm, static_cast<const __bf16*>(values),
                                                        static_cast<__bf16*>(values_sorted), nonfinite);
    else
      k_gather_channels_sm<_Float16><<<grid, 256, 0, s>>>(n, n_pad, nchan, perm,
                                                          static_cast<const _Float16*>(values),
                                                          static_cast<_Float16*>(values_sorted), nonfinite);
  }
  HCG_CHECK_LAUNCH("k_gather_channels");
  return HCG_OK;
}

extern "C" size_t hcg_grid_mfma_workspace_size(uint32_t side, const hcg_mfma_shape* shape) {
  const uint32_t tc = (shape && shape->tile_cols) ? shape->tile_cols : kMcDefaultTileCols;
  if (!(tc == 8 || tc == 16 || tc == 32 || tc == 64)) return 0;
  const size_t ntiles = (size_t)ceil_div_u32(side, tc) * ceil_div_u32(side, 64u / tc);
  return ntiles * kPlanTileWords * sizeof(uint32_t);
}

extern "C" int hcg_grid_mfma_ex(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t side, uint32_t n,
                                uint32_t nchan, int dtype, const hcg_sample* records, const uint32_t* keys_sorted,
                                const double* lon_sorted, const double* lat_sorted, const double* weights_sorted,
                                const void* values_sorted, const uint32_t* lut, const double* ring_z,
                                const double* ring_s, const hcg_ring_meta* meta, const double* xwcs,
                                const double* ywcs, float* cube, float* wsum, int normalize,
                                const hcg_mfma_shape* shape, void* workspace, size_t workspace_bytes,
                                void* stream) {
  HCG_REQUIRE(hp && kern, "hcg_grid_mfma: NULL hp/kern");
  if (hp->nside == 0 || hp->nside > HCG_MAX_NSIDE)
    return set_error(HCG_ERR_UNSUPPORTED, "hcg_grid_mfma: nside must be <= 8192");
  HCG_REQUIRE(kern->kernel_type >= HCG_GAUSS1D && kern->kernel_type <= HCG_TAPERED_SINC,
              "hcg_grid_mfma: unknown kernel type");
  HCG_REQUIRE(kern->factor >= 1, "hcg_grid_mfma: factor must be >= 1");
  HCG_REQUIRE(kern->selection <= HCG_SELECT_DISC,
              "hcg_grid_mfma: selection must be HCG_SELECT_COMPAT or HCG_SELECT_DISC");
  HCG_REQUIRE(dtype == HCG_DTYPE_BF16 || dtype == HCG_DTYPE_F16, "hcg_grid_mfma: dtype must be BF16 or F16");
  HCG_REQUIRE(side <= 65535, "hcg_grid_mfma: side must be <= 65535");
  hcg_mfma_shape sh = shape ? *shape : hcg_mfma_shape{0, 0, 0, 0};
  HCG_REQUIRE(sh.tile_cols == 0 || sh.tile_cols == 8 || sh.tile_cols == 16 || sh.tile_cols == 32 ||
                  sh.tile_cols == 64,
              "hcg_grid_mfma: shape.tile_cols must be 0, 8, 16, 32 or 64");
  HCG_REQUIRE(sh.waves == 0 || sh.waves == 1 || sh.waves == 2 || sh.waves == 4 || sh.waves == 8,
              "hcg_grid_mfma: shape.waves must be 0, 1, 2, 4 or 8");
  HCG_REQUIRE(sh.chunk == 0 || sh.chunk == 128 || sh.chunk == 256 || sh.chunk == 512,
              "hcg_grid_mfma: shape.chunk must be 0, 128, 256 or 512");
  if (side == 0 || nchan == 0) return HCG_OK;
  HCG_REQUIRE(records && keys_sorted && lon_sorted && lat_sorted && values_sorted && lut && ring_z && ring_s && meta &&
                  xwcs && ywcs && cube,
              "hcg_grid_mfma: NULL buffer");
  hipStream_t s = as_stream(stream);
  GridParams P = make_grid_params(hp, kern, side);
  GridBuffers B{records, keys_sorted, lon_sorted, lat_sorted, weights_sorted, lut, ring_z, ring_s, meta,
                xwcs,    ywcs,        nullptr,    nullptr};
  if (one_channel_tiles(P, kern->kernel_type, weights_sorted != nullptr, nchan)) {
    const TileOut O{static_cast<const uint16_t*>(values_sorted), cube, wsum, dtype == HCG_DTYPE_BF16 ? 1u : 0u,
                    (uint32_t)(normalize ? 1 : 0), 1u};
    return launch_tiles_one_channel(P, B, O, s);
  }
  if (few_channel_tiles(P, kern->kernel_type, weights_sorted != nullptr, nchan, n)) {
    const TileOut O{static_cast<const uint16_t*>(values_sorted), cube, wsum, dtype == HCG_DTYPE_BF16 ? 1u : 0u,
                    (uint32_t)(normalize ? 1 : 0), nchan};
    return launch_tiles_few_channels(P, B, O, s);
  }
  const uint32_t n_pad = hcg_channels_padded_len(n);
  const uint32_t groups_total = ceil_div_u32(nchan, kMcChanPerWave);  // waves needed
  // waves per workgroup: a power of two (each cell's nw threads are aligned
  // consecutive lanes, reduced with xor shuffles); surplus waves only help
  // with the weights
  uint32_t nw = sh.waves;
  if (nw == 0) {
    // default: the fewest waves that cover the channels, at most 8 (chunk
    // 512, one workgroup per CU): with several rings packed per chunk the
    // 8-wave shape is the fastest on the 2M x 1024 raster (13.5 ms, against
    // 15.5 for 4 waves at chunk 256, two workgroups per CU)
    nw = 1;
    while (nw < groups_total && nw < 8u) nw <<= 1;
  }
  const uint32_t ngroups = ceil_div_u32(groups_total, nw);
  const uint32_t tc = sh.tile_cols ? sh.tile_cols : kMcDefaultTileCols;
  const uint32_t tc_shift = tc == 8 ? 3u : tc == 16 ? 4u : tc == 32 ? 5u : 6u;
  const uint32_t tr = 64u / tc;
  const uint32_t tiles_x = ceil_div_u32(side, tc);
  const uint32_t ntiles = tiles_x * ceil_div_u32(side, tr);
  const dim3 grid(ntiles, ngroups);
  // chunk size by workgroup width: 8 waves fill a CU alone (VGPR-limited);
  // narrower workgroups take smaller tiles so several share a CU
  const uint32_t chunk = sh.chunk ? sh.chunk : (nw >= 8 ? 512u : (nw >= 4 ? 256u : 128u));
  const uint32_t ktc = mfma_weight_form(P, kern->kernel_type);
  MfmaJob J{P,        B,       values_sorted, cube,   wsum, ktc, n_pad, nchan, (uint32_t)dtype,
            tc_shift, tiles_x, ntiles,        (uint32_t)(normalize ? 1 : 0), nullptr, 0u, 0u};
  if (workspace) {
    if (workspace_bytes < (size_t)ntiles * kPlanTileWords * sizeof(uint32_t))
      return set_error(HCG_ERR_WORKSPACE, "hcg_grid_mfma_ex: workspace smaller than hcg_grid_mfma_workspace_size");
    J.plan = static_cast<uint32_t*>(workspace);
    const int rc = launch_mfma_plan(J, s);
    if (rc != HCG_OK) return rc;
  }
#define HCG_LAUNCH_MFMA(TT, CH) k_grid_mfma<TT, CH><<<grid, 64 * nw, 0, s>>>(J)
  if (dtype == HCG_DTYPE_BF16) {
    if (chunk == 512) HCG_LAUNCH_MFMA(kDtBf16, 512);
    else if (chunk == 256) HCG_LAUNCH_MFMA(kDtBf16, 256);
    else HCG_LAUNCH_MFMA(kDtBf16, 128);
  } else {
    if (chunk == 512) HCG_LAUNCH_MFMA(kDtF16, 512);
    else if (chunk == 256) HCG_LAUNCH_MFMA(kDtF16, 256);
    else HCG_LAUNCH_MFMA(kDtF16, 128);
  }
#undef HCG_LAUNCH_MFMA
  HCG_CHECK_LAUNCH("k_grid_mfma");
  return HCG_OK;
}

extern "C" int hcg_grid_mfma(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t side, uint32_t n, uint32_t nchan,
                             int dtype, const hcg_sample* records, const uint32_t* keys_sorted,
                             const double* lon_sorted, const double* lat_sorted, const double* weights_sorted,
                             const void* values_sorted, const uint32_t* lut, const double* ring_z,
                             const double* ring_s, const hcg_ring_meta* meta, const double* xwcs, const double* ywcs,
                             float* cube, float* wsum, int normalize, void* stream) {
  return hcg_grid_mfma_ex(hp, kern, side, n, nchan, dtype, records, keys_sorted, lon_sorted, lat_sorted,
                          weights_sorted, values_sorted, lut, ring_z, ring_s, meta, xwcs, ywcs, cube, wsum, normalize,
                          nullptr, nullptr, 0, stream);
}

extern "C" int hcg_grid_mfma_shaped(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t side, uint32_t n,
                                    uint32_t nchan, int dtype, const hcg_sample* records, const uint32_t* keys_sorted,
                                    const double* lon_sorted, const double* lat_sorted, const double* weights_sorted,
                                    const void* values_sorted, const uint32_t* lut, const double* ring_z,
                                    const double* ring_s, const hcg_ring_meta* meta, const double* xwcs,
                                    const double* ywcs, float* cube, float* wsum, int normalize,
                                    const hcg_mfma_shape* shape, void* stream) {
  return hcg_grid_mfma_ex(hp, kern, side, n, nchan, dtype, records, keys_sorted, lon_sorted, lat_sorted,
                          weights_sorted, values_sorted, lut, ring_z, ring_s, meta, xwcs, ywcs, cube, wsum, normalize,
                          shape, nullptr, 0, stream);
}

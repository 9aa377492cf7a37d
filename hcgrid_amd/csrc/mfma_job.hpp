// mfma_job.hpp -- the multi-channel job descriptor shared by the one-job
// launch (hcg_grid_mfma) and the persistent batched launch (hcg_batch_*).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "plain_tile.hpp"

namespace hcg {

constexpr int kMcCols = 64;                   // cells per workgroup
constexpr uint32_t kMcDefaultTileCols = 16;   // default block: 4 rows x 16 columns
constexpr int kMcNB = 4;                      // 16-channel blocks per wave
constexpr int kMcMaxWaves = 8;
constexpr uint32_t kMcChanPerWave = 16 * kMcNB;

// Tile plan: per tile, its ring span and the sorted range of each ring's
// union window, for up to kPlanRings rings (rings past that are searched in
// the gridding kernel itself).  Layout per tile, kPlanStride uint32:
// [r_begin, r_end, lo_0, hi_0, lo_1, hi_1, ...]; lo >= hi: no samples.
constexpr uint32_t kPlanRings = 64;
constexpr uint32_t kPlanStride = 2 + 2 * kPlanRings;
// After the ring ranges, at kPlanHeadWords: the block's 64 cell windows
// (Window, as block_cell_window gives them), then their unit vectors as
// three arrays of 64 doubles (x, y, z): the gridding kernel reads them
// instead of evaluating the cells' trigonometry and its dependent loads.
// kPlanTileWords uint32 per tile in all.
constexpr uint32_t kPlanHeadWords = (kPlanStride + 3u) & ~3u;
constexpr uint32_t kPlanWinWords = kMcCols * (uint32_t)sizeof(Window) / 4u;
constexpr uint32_t kPlanTileWords = kPlanHeadWords + kPlanWinWords + kMcCols * 3u * 2u;
static_assert(sizeof(Window) % 8 == 0 && kPlanHeadWords % 2 == 0 && kPlanTileWords % 4 == 0, "plan alignment");

// The weight form of a job (MfmaJob::kt): the kernel type, or
// kKtGaussSmall (GAUSS1D whose radius rules out the exponent's asin branch
// and the exp clamp, gauss_small_radius), or kKtGaussLean (also
// gauss_tiny_radius and dot_distance_ok: the p.q candidate test and
// gauss_weight_lean of the tile loop).
constexpr uint32_t kKtGaussSmall = HCG_GAUSS1D + 16u;
constexpr uint32_t kKtGaussLean = HCG_GAUSS1D + 32u;

inline uint32_t mfma_weight_form(const GridParams& P, int kernel_type) {
  if (kernel_type != HCG_GAUSS1D || !gauss_small_radius(P)) return (uint32_t)kernel_type;
  return gauss_tiny_radius(P) && dot_distance_ok(P) ? kKtGaussLean : kKtGaussSmall;
}

// One multi-channel job: geometry, sorted channel values and outputs, with
// its launch shape (tiles of (64 >> tc_shift) x (1 << tc_shift) cells).
struct MfmaJob {
  GridParams P;
  GridBuffers B;
  const void* vs;     // sorted channel values [nchan][n_pad], BF16 or FP16
  float* cube;        // [nchan][side * side]
  float* wsum;        // [side * side] or NULL
  uint32_t kt, n_pad, nchan, dtype;  // kt: mfma_weight_form
  uint32_t tc_shift, tiles_x, ntiles, normalize;
  uint32_t* plan;     // [ntiles][kPlanTileWords], or NULL (windows and searches in the gridding kernel)
  // 0: the MFMA contraction; else a one-channel job on the FP64 tile path
  // (one_channel_tiles) with tiles_x and ntiles counting kTileCols x
  // kTileRows tiles: this many tiles per work item, one per wave, when
  // tile_streams is 1; one tile per item split over tile_streams partial
  // ring streams otherwise (one_channel_streams)
  uint32_t tiles_per_item;
  uint32_t tile_streams;
};

// One unit of the persistent launch: a tile of one job, one channel group.
struct MfmaItem {
  uint32_t job, tile, group, pad;
};

inline uint32_t tile_shift_of(uint32_t tile_cols) {
  return tile_cols == 8 ? 3u : tile_cols == 16 ? 4u : tile_cols == 32 ? 5u : 6u;
}

// Batch preparation in one launch: each job's channel gather (input values,
// permutation and non-finite counter here; output J.vs) and tile plan.
struct PrepJob {
  const void* in;          // the caller's channel values (layout per prep_ranges)
  const uint32_t* perm;    // [n]
  uint32_t* nonfinite;     // counter, or NULL
  uint32_t n, pad;
};
enum : uint32_t { kPrepPlan = 0, kPrepGatherCm = 1, kPrepGatherSm = 2 };
struct PrepRange {
  uint32_t first, job, kind, gx;  // first workgroup; for gathers, workgroups per channel block row
};
// The ranges of one prep launch (layouts[j]: job j's value layout, 0
// channel-major, 1 sample-major; at most 2 per job) and their workgroup
// total.
int prep_ranges(uint32_t njobs, const MfmaJob* jobs, const uint32_t* layouts, std::vector<PrepRange>& out,
                uint32_t* blocks);
int launch_batch_prep(const PrepRange* ranges, uint32_t nranges, uint32_t blocks, const PrepJob* pj,
                      const MfmaJob* jobs, hipStream_t s);

// Enqueues the tile-plan kernel of a job into J.plan (non-NULL).
int launch_mfma_plan(const MfmaJob& J, hipStream_t s);

// Tiles per work item of a tile-path job in the persistent kernel of (chunk,
// waves): one per wave, as many as the workgroup's LDS holds at tile_bytes
// per tile.
uint32_t tile_items_per_block(uint32_t chunk, uint32_t waves, size_t tile_bytes);

// Grid size of the persistent kernel: CUs x resident workgroups per CU.
int mfma_batch_blocks(uint32_t chunk, uint32_t waves, uint32_t* blocks);

// Launches the persistent kernel with `blocks` workgroups of `waves` waves
// over `nitems` work items (head: a device counter that reads 0 at launch).
int launch_mfma_batch(uint32_t chunk, uint32_t waves, uint32_t blocks, const MfmaJob* jobs, const MfmaItem* items,
                      uint32_t nitems, uint32_t* head, hipStream_t s);

}  // namespace hcg

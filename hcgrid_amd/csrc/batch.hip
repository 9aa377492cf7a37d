// batch.hip -- resident geometries, batched multi-channel gridding in one
// persistent launch, and HIP graph capture of the whole sequence.
//
// The reference grids one map per process run (HCGrid/HCGrid.cpp:38-197;
// HCGrid+ loops over channels of one map, HCGrid+/gridding-omp.cu:552-678).
// Here a geometry (one map's sorted samples, ring LUT and target
// coordinates) is built once and kept in HBM; a batch names any number of
// (geometry, channel values, output cube) jobs of uneven sizes and grids
// them all with one launch of k_grid_mfma_batch, whose workgroups pull
// (job, tile, channel group) items from a device queue.  Everything a batch
// enqueues is kernels only (no memsets, no host copies), so the sequence
// can be captured into a graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "common.hpp"
#include "mfma_job.hpp"

using namespace hcg;

struct hcg_geometry {
  int device;
  bool owns;             // base was allocated here (freed by destroy)
  hcg_healpix hp;
  hcg_kernel kern;
  uint32_t n, side;
  size_t bytes;
  void* base;
  uint32_t* keys;        // [n+1]
  uint32_t* perm;        // [n]
  hcg_sample* rec;       // [n]
  double* lon_s;         // [n]
  double* lat_s;         // [n]
  uint32_t* lut;         // [nrings+2]
  double* ring_z;        // [nrings+2]
  double* ring_s;        // [nrings+2]
  hcg_ring_meta* meta;   // [1]
  double* xwcs;          // [side*side]
  double* ywcs;          // [side*side]
  void* ws;              // pre-sort workspace
  size_t ws_bytes;
};

struct hcg_batch {
  int device;
  bool owns;             // base was allocated here (freed by destroy)
  uint32_t njobs, nitems, waves, chunk, blocks;
  uint32_t nitems_dense, blocks_dense;  // the dense jobs' launch (kDenseWaves, kDenseChunk), run first
  size_t bytes;
  void* base;
  MfmaJob* d_jobs;
  MfmaItem* d_items;
  uint32_t* d_counters;  // [0] queue head, [1 + j] non-finite values of job j, [1 + njobs] dense queue head
  PrepRange* d_ranges;   // the prep launch: every job's channel gather and tile plan
  PrepJob* d_prep;
  uint32_t nranges, prep_blocks;
  std::vector<hcg_batch_job> jobs;
  std::vector<void*> values_sorted;
  std::vector<MfmaJob> host_jobs;     // upload sources, alive until destroy
  std::vector<MfmaItem> host_items;
  std::vector<PrepRange> host_ranges;
  std::vector<PrepJob> host_prep;
  // optional kernel timing (hcg_batch_timing): start, after the prep launch, after the gridding launches
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  bool timed = false;  // the events of the last eager enqueue were recorded
};

namespace {

constexpr size_t kAlign = 256;

size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// Carves consecutive aligned regions out of one allocation.
struct Carver {
  char* base;
  size_t off = 0;
  template <typename T>
  T* take(size_t count) {
    T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
    off = align_up(off + count * sizeof(T));
    return p;
  }
};

void geometry_layout(hcg_geometry* g, char* base) {
  Carver c{base};
  const size_t n = g->n, nr = (size_t)g->hp.nrings + 2, m = (size_t)g->side * g->side;
  g->keys = c.take<uint32_t>(n + 1);
  g->perm = c.take<uint32_t>(n ? n : 1);
  g->rec = c.take<hcg_sample>(n ? n : 1);
  g->lon_s = c.take<double>(n ? n : 1);
  g->lat_s = c.take<double>(n ? n : 1);
  g->lut = c.take<uint32_t>(nr);
  g->ring_z = c.take<double>(nr);
  g->ring_s = c.take<double>(nr);
  g->meta = c.take<hcg_ring_meta>(1);
  g->xwcs = c.take<double>(m ? m : 1);
  g->ywcs = c.take<double>(m ? m : 1);
  g->ws_bytes = hcg_presort_workspace_size(g->n);
  g->ws = c.take<char>(g->ws_bytes ? g->ws_bytes : 1);
  g->bytes = c.off;
}

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* p, uint32_t count, uint32_t v) {
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i < count) p[i] = v;
}

// Rough cost of one work item of a job: samples in a tile's neighbourhood
// times (channels in the item + a fixed weight-evaluation share).
double item_cost(const hcg_geometry* g, uint32_t tile_cols, uint32_t chans) {
  const double side = g->side, tc = tile_cols, tr = 64.0 / tile_cols;
  const double cell_deg = 0.01;  // only the ratio between jobs matters
  const double d = 2.0 * g->kern.sphere_radius / cell_deg;
  const double density = side > 0 ? (double)g->n / (side * side) : 0.0;
  return density * (tr + d) * (tc + d) * (chans + 32.0);
}

}  // namespace

extern "C" size_t hcg_geometry_bytes(uint32_t n, uint32_t side, const hcg_healpix* hp) {
  if (!hp) return 0;
  hcg_geometry g{};
  g.n = n;
  g.side = side;
  g.hp = *hp;
  geometry_layout(&g, nullptr);
  return g.bytes;
}

namespace {

int geometry_init(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t n, uint32_t side, void* mem,
                  size_t mem_bytes, hcg_geometry** out) {
  HCG_REQUIRE(hp && kern && out, "hcg_geometry_create: NULL argument");
  *out = nullptr;
  if (hp->nside == 0 || hp->nside > HCG_MAX_NSIDE || (hp->nside & (hp->nside - 1)))
    return set_error(HCG_ERR_UNSUPPORTED, "hcg_geometry_create: nside must be a power of two <= 8192");
  HCG_REQUIRE(n < 0xFFFFFFFFu, "hcg_geometry_create: n must be < 2^32-1");
  HCG_REQUIRE(side <= 65535, "hcg_geometry_create: side must be <= 65535");
  HCG_REQUIRE(kern->factor >= 1, "hcg_geometry_create: factor must be >= 1");
  HCG_REQUIRE(kern->selection <= HCG_SELECT_DISC,
              "hcg_geometry_create: selection must be HCG_SELECT_COMPAT or HCG_SELECT_DISC");
  hcg_geometry* g = new (std::nothrow) hcg_geometry{};
  if (!g) return set_error(HCG_ERR_NOMEM, "hcg_geometry_create: host allocation failed");
  g->hp = *hp;
  g->kern = *kern;
  g->n = n;
  g->side = side;
  geometry_layout(g, nullptr);
  hipError_t e = hipGetDevice(&g->device);
  if (e == hipSuccess && mem) {
    if (mem_bytes < g->bytes || (reinterpret_cast<uintptr_t>(mem) & (kAlign - 1))) {
      delete g;
      return set_error(HCG_ERR_WORKSPACE, "hcg_geometry_create_in: memory smaller than hcg_geometry_bytes or not "
                                          "256-byte aligned");
    }
    g->base = mem;
  } else if (e == hipSuccess) {
    e = hipMalloc(&g->base, g->bytes);
    g->owns = true;
  }
  if (e != hipSuccess) {
    delete g;
    (void)hipGetLastError();
    return set_error(HCG_ERR_NOMEM, std::string("hcg_geometry_create: hipMalloc: ") + hipGetErrorString(e));
  }
  geometry_layout(g, static_cast<char*>(g->base));
  *out = g;
  return HCG_OK;
}

}  // namespace

extern "C" int hcg_geometry_create(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t n, uint32_t side,
                                   hcg_geometry** out) {
  return geometry_init(hp, kern, n, side, nullptr, 0, out);
}

extern "C" int hcg_geometry_create_in(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t n, uint32_t side,
                                      void* device_mem, size_t bytes, hcg_geometry** out) {
  HCG_REQUIRE(device_mem, "hcg_geometry_create_in: NULL memory");
  return geometry_init(hp, kern, n, side, device_mem, bytes, out);
}

extern "C" int hcg_geometry_destroy(hcg_geometry* g) {
  if (!g) return HCG_OK;
  if (g->owns && g->base) HCG_CHECK_HIP(hipFree(g->base));
  delete g;
  return HCG_OK;
}

extern "C" int hcg_geometry_build(hcg_geometry* g, const double* lon_deg, const double* lat_deg,
                                  const double* xwcs, const double* ywcs, void* stream) {
  HCG_REQUIRE(g, "hcg_geometry_build: NULL geometry");
  hipStream_t s = as_stream(stream);
  const size_t m = (size_t)g->side * g->side;
  if (m) {
    HCG_REQUIRE(xwcs && ywcs, "hcg_geometry_build: NULL target coordinates");
    if (xwcs != g->xwcs) HCG_CHECK_HIP(hipMemcpyAsync(g->xwcs, xwcs, m * 8, hipMemcpyDeviceToDevice, s));
    if (ywcs != g->ywcs) HCG_CHECK_HIP(hipMemcpyAsync(g->ywcs, ywcs, m * 8, hipMemcpyDeviceToDevice, s));
  }
  int rc = hcg_presort(&g->hp, g->n, lon_deg, lat_deg, g->keys, g->perm, g->ws, g->ws_bytes, stream);
  if (rc != HCG_OK) return rc;
  rc = hcg_gather_samples(g->n, g->perm, lon_deg, lat_deg, nullptr, HCG_DTYPE_F64, nullptr, g->rec, g->lon_s,
                          g->lat_s, nullptr, nullptr, stream);
  if (rc != HCG_OK) return rc;
  return hcg_ring_lut(&g->hp, g->n, g->keys, g->lut, g->ring_z, g->ring_s, g->meta, stream);
}

extern "C" int hcg_geometry_arrays(const hcg_geometry* g, hcg_geometry_view* v) {
  HCG_REQUIRE(g && v, "hcg_geometry_arrays: NULL argument");
  v->n = g->n;
  v->side = g->side;
  v->bytes = g->bytes;
  v->keys_sorted = g->keys;
  v->perm = g->perm;
  v->records = g->rec;
  v->lon_sorted = g->lon_s;
  v->lat_sorted = g->lat_s;
  v->lut = g->lut;
  v->ring_z = g->ring_z;
  v->ring_s = g->ring_s;
  v->meta = g->meta;
  v->xwcs = g->xwcs;
  v->ywcs = g->ywcs;
  return HCG_OK;
}

extern "C" size_t hcg_batch_job_bytes(const hcg_geometry* g, uint32_t nchan) {
  if (!g) return 0;
  return align_up((size_t)nchan * hcg_channels_padded_len(g->n) * 2);
}

namespace {

// A batch's plan: work items and the layout of its one device allocation.
// Dense MFMA jobs (at least kDenseSamplesPerCell samples per map cell), and
// jobs with more channels than one channel group of the batch's shape
// holds, run in a persistent launch of their own at 8 waves and 512-sample
// chunks, the shape that suits them (one workgroup per CU sharing each
// block's weights over 512 channels): a 3.9M-sample 128^2 x 512-channel job
// takes 18.2 ms there against 26.7 ms at 4 waves / 256, a 3.2M-sample
// 2048^2 x 512-channel one 21.6 against 24.5, while sparse jobs with one
// channel group run up to 1.8x slower at 8 waves.  Results do not depend on
// the launch shape.
constexpr double kDenseSamplesPerCell = 32.0;
constexpr uint32_t kDenseWaves = 8, kDenseChunk = 512;

struct BatchPlan {
  uint32_t waves = 0, chunk = 0;
  uint32_t n_dense = 0;  // the first n_dense items belong to the dense launch
  std::vector<MfmaItem> items;
  std::vector<uint32_t> tiles, tcols, tpi, tstreams;  // tile-path jobs: tiles per item, streams; else 0
  std::vector<size_t> voff, poff;
  size_t bytes = 0;
};

int plan_batch(uint32_t njobs, const hcg_batch_job* jobs, uint32_t waves, uint32_t chunk, BatchPlan& p) {
  HCG_REQUIRE(njobs == 0 || jobs, "hcg_batch_create: NULL jobs");
  if (waves == 0) waves = 4;
  if (chunk == 0) chunk = waves >= 8 ? 512 : (waves >= 4 ? 256 : 128);
  HCG_REQUIRE(waves == 1 || waves == 2 || waves == 4 || waves == 8, "hcg_batch_create: waves must be 1, 2, 4 or 8");
  HCG_REQUIRE(chunk == 128 || chunk == 256 || chunk == 512, "hcg_batch_create: chunk must be 128, 256 or 512");
  for (uint32_t j = 0; j < njobs; ++j) {
    const hcg_batch_job& b = jobs[j];
    HCG_REQUIRE(b.geometry, "hcg_batch_create: job without geometry");
    HCG_REQUIRE(b.dtype == HCG_DTYPE_BF16 || b.dtype == HCG_DTYPE_F16, "hcg_batch_create: dtype must be BF16 or F16");
    HCG_REQUIRE(b.tile_cols == 0 || b.tile_cols == 8 || b.tile_cols == 16 || b.tile_cols == 32 || b.tile_cols == 64,
                "hcg_batch_create: tile_cols must be 0, 8, 16, 32 or 64");
    HCG_REQUIRE(b.geometry->kern.kernel_type >= HCG_GAUSS1D && b.geometry->kern.kernel_type <= HCG_TAPERED_SINC,
                "hcg_batch_create: unknown kernel type");
    HCG_REQUIRE(b.nchan == 0 || (b.values && b.cube), "hcg_batch_create: NULL values or cube");
  }
  p.waves = waves;
  p.chunk = chunk;
  // work items: jobs longest item first, each job's tiles in order, a
  // tile's channel groups adjacent (they share the tile's samples in L2)
  std::vector<uint32_t> order(njobs), groups(njobs), units(njobs);
  std::vector<char> dense(njobs, 0);
  std::vector<double> cost(njobs);
  p.tiles.assign(njobs, 0);
  p.tcols.assign(njobs, 0);
  p.tpi.assign(njobs, 0);
  p.tstreams.assign(njobs, 0);
  const uint32_t tile_tpi = tile_items_per_block(chunk, waves, kTileLdsBytes);
  const uint32_t few_tpi = tile_items_per_block(chunk, waves, kFewTileLdsBytes);
  uint64_t nitems = 0;
  for (uint32_t j = 0; j < njobs; ++j) {
    const hcg_batch_job& b = jobs[j];
    const hcg_geometry* g = b.geometry;
    const uint32_t side = g->side;
    order[j] = j;
    const GridParams P = make_grid_params(&g->hp, &g->kern, side);
    if (one_channel_tiles(P, g->kern.kernel_type, false, b.nchan)) {
      // one channel: FP64 tile path, tile_tpi tiles of 64 cells per item (one
      // per wave), or on small maps one tile per item over its ring streams;
      // the tile loop weighs a pair roughly 8x faster than the MFMA stage
      p.tcols[j] = kTileCols;
      p.tiles[j] = side ? tile_count_of(side) : 0;
      p.tstreams[j] = one_channel_streams(side);
      p.tpi[j] = p.tstreams[j] > 1 ? 1u : tile_tpi;
      groups[j] = 1;
      units[j] = ceil_div_u32(p.tiles[j], p.tpi[j]);
      cost[j] = item_cost(g, kTileCols, 1) * p.tpi[j] / 8.0;
    } else if (few_channel_tiles(P, g->kern.kernel_type, false, b.nchan, g->n)) {
      // few channels on a large map: FP64 tile path, one tile per wave
      p.tcols[j] = kTileCols;
      p.tiles[j] = tile_count_of(side);
      p.tstreams[j] = 1;
      p.tpi[j] = few_tpi;
      groups[j] = 1;
      units[j] = ceil_div_u32(p.tiles[j], few_tpi);
      cost[j] = item_cost(g, kTileCols, b.nchan) * few_tpi / 4.0;
    } else {
      const uint32_t tc = b.tile_cols ? b.tile_cols : kMcDefaultTileCols;
      p.tcols[j] = tc;
      p.tiles[j] = (b.nchan && side) ? ceil_div_u32(side, tc) * ceil_div_u32(side, 64 / tc) : 0;
      // dense, or more channels than one channel group at the batch's shape
      // covers (each group recomputes the block's weights)
      dense[j] = waves != kDenseWaves && side &&
                 ((double)g->n >= kDenseSamplesPerCell * side * side || b.nchan > waves * kMcChanPerWave);
      const uint32_t w = dense[j] ? kDenseWaves : waves;
      groups[j] = ceil_div_u32(ceil_div_u32(b.nchan, kMcChanPerWave), w);
      units[j] = p.tiles[j];
      cost[j] = item_cost(g, tc, std::min<uint32_t>(b.nchan, w * kMcChanPerWave));
    }
    nitems += (uint64_t)units[j] * groups[j];
  }
  if (nitems >= 0xFFFFFFFFull) return set_error(HCG_ERR_INVALID, "hcg_batch_create: too many work items");
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
  p.items.clear();
  p.items.reserve(nitems);
  for (int pass = 0; pass < 2; ++pass) {  // the dense launch's items first
    for (uint32_t j : order) {
      if (dense[j] != (pass == 0)) continue;
      for (uint32_t t = 0; t < units[j]; ++t)
        for (uint32_t g = 0; g < groups[j]; ++g) p.items.push_back(MfmaItem{j, t, g, 0});
    }
    if (pass == 0) p.n_dense = (uint32_t)p.items.size();
  }
  // one allocation: job table, items, counters, sorted channel values
  Carver c{nullptr};
  c.take<MfmaJob>(njobs ? njobs : 1);
  c.take<MfmaItem>(p.items.empty() ? 1 : p.items.size());
  c.take<uint32_t>(2 + njobs);
  c.take<PrepRange>(2 * (size_t)njobs + 1);
  c.take<PrepJob>(njobs ? njobs : 1);
  p.voff.assign(njobs, 0);
  p.poff.assign(njobs, 0);
  for (uint32_t j = 0; j < njobs; ++j) {
    p.voff[j] = c.off;
    c.off += hcg_batch_job_bytes(jobs[j].geometry, jobs[j].nchan);
    p.poff[j] = c.off;
    if (!p.tpi[j]) c.off = align_up(c.off + (size_t)p.tiles[j] * kPlanTileWords * sizeof(uint32_t));
  }
  p.bytes = c.off;
  return HCG_OK;
}

// Lays the batch out in `mem` and uploads the job table and work list:
// asynchronously on `stream` when upload_async, else synchronously.
int batch_init(uint32_t njobs, const hcg_batch_job* jobs, const BatchPlan& p, void* mem, bool owns,
               bool upload_async, hipStream_t stream, hcg_batch** out) {
  hcg_batch* bt = new (std::nothrow) hcg_batch{};
  if (!bt) return set_error(HCG_ERR_NOMEM, "hcg_batch_create: host allocation failed");
  bt->njobs = njobs;
  bt->waves = p.waves;
  bt->chunk = p.chunk;
  bt->nitems = (uint32_t)p.items.size();
  bt->nitems_dense = p.n_dense;
  bt->bytes = p.bytes;
  bt->base = mem;
  bt->owns = owns;
  bt->jobs.assign(jobs, jobs + njobs);
  (void)hipGetDevice(&bt->device);
  Carver d{static_cast<char*>(bt->base)};
  bt->d_jobs = d.take<MfmaJob>(njobs ? njobs : 1);
  bt->d_items = d.take<MfmaItem>(p.items.empty() ? 1 : p.items.size());
  bt->d_counters = d.take<uint32_t>(2 + njobs);
  bt->d_ranges = d.take<PrepRange>(2 * (size_t)njobs + 1);
  bt->d_prep = d.take<PrepJob>(njobs ? njobs : 1);
  bt->values_sorted.resize(njobs);
  std::vector<MfmaJob>& hj = bt->host_jobs;
  hj.resize(njobs);
  bt->host_items = p.items;
  for (uint32_t j = 0; j < njobs; ++j) {
    const hcg_batch_job& b = jobs[j];
    const hcg_geometry* g = b.geometry;
    bt->values_sorted[j] = static_cast<char*>(bt->base) + p.voff[j];
    GridBuffers B{g->rec,  g->keys, g->lon_s, g->lat_s, nullptr, g->lut, g->ring_z,
                  g->ring_s, g->meta, g->xwcs,  g->ywcs,  nullptr, nullptr};
    const GridParams P = make_grid_params(&g->hp, &g->kern, g->side);
    hj[j] = MfmaJob{P,
                    B,
                    bt->values_sorted[j],
                    b.cube,
                    b.wsum,
                    mfma_weight_form(P, g->kern.kernel_type),
                    hcg_channels_padded_len(g->n),
                    b.nchan,
                    (uint32_t)b.dtype,
                    tile_shift_of(p.tcols[j]),
                    ceil_div_u32(g->side, p.tcols[j]),
                    p.tiles[j],
                    (uint32_t)(b.normalize ? 1 : 0),
                    p.tpi[j] ? nullptr : reinterpret_cast<uint32_t*>(static_cast<char*>(bt->base) + p.poff[j]),
                    p.tpi[j],
                    p.tstreams[j]};
  }
  // the prep launch: gathers (channel-major values) and plans of every job
  bt->host_prep.resize(njobs);
  std::vector<uint32_t> layouts(njobs, 0u);
  for (uint32_t j = 0; j < njobs; ++j)
    bt->host_prep[j] = PrepJob{jobs[j].values, jobs[j].geometry->perm, bt->d_counters + 1 + j, jobs[j].geometry->n, 0u};
  int prc = prep_ranges(njobs, hj.data(), layouts.data(), bt->host_ranges, &bt->prep_blocks);
  if (prc != HCG_OK) {
    if (owns) (void)hipFree(bt->base);
    delete bt;
    return prc;
  }
  bt->nranges = (uint32_t)bt->host_ranges.size();
  hipError_t e = hipSuccess;
  const size_t jb = njobs * sizeof(MfmaJob), ib = p.items.size() * sizeof(MfmaItem);
  const size_t rb = bt->host_ranges.size() * sizeof(PrepRange), pb = njobs * sizeof(PrepJob);
  if (upload_async) {
    // ordered on `stream`; the host sources live in the batch until destroy
    if (jb) e = hipMemcpyAsync(bt->d_jobs, hj.data(), jb, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && ib)
      e = hipMemcpyAsync(bt->d_items, bt->host_items.data(), ib, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && rb)
      e = hipMemcpyAsync(bt->d_ranges, bt->host_ranges.data(), rb, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess && pb) e = hipMemcpyAsync(bt->d_prep, bt->host_prep.data(), pb, hipMemcpyHostToDevice, stream);
  } else {
    if (jb) e = hipMemcpy(bt->d_jobs, hj.data(), jb, hipMemcpyHostToDevice);
    if (e == hipSuccess && ib) e = hipMemcpy(bt->d_items, p.items.data(), ib, hipMemcpyHostToDevice);
    if (e == hipSuccess && rb) e = hipMemcpy(bt->d_ranges, bt->host_ranges.data(), rb, hipMemcpyHostToDevice);
    if (e == hipSuccess && pb) e = hipMemcpy(bt->d_prep, bt->host_prep.data(), pb, hipMemcpyHostToDevice);
  }
  int rc = HCG_OK;
  if (e != hipSuccess) rc = set_error(HCG_ERR_HIP, std::string("hcg_batch_create: ") + hipGetErrorString(e));
  // grid size of the persistent launch, fixed here so that enqueueing makes no queries
  if (rc == HCG_OK) rc = mfma_batch_blocks(p.chunk, p.waves, &bt->blocks);
  if (rc == HCG_OK && bt->nitems_dense) rc = mfma_batch_blocks(kDenseChunk, kDenseWaves, &bt->blocks_dense);
  if (rc != HCG_OK) {
    if (owns) (void)hipFree(bt->base);
    delete bt;
    return rc;
  }
  *out = bt;
  return HCG_OK;
}

}  // namespace

extern "C" size_t hcg_batch_bytes(uint32_t njobs, const hcg_batch_job* jobs, uint32_t waves, uint32_t chunk) {
  BatchPlan p;
  if (plan_batch(njobs, jobs, waves, chunk, p) != HCG_OK) return 0;
  return p.bytes;
}

extern "C" int hcg_batch_create(uint32_t njobs, const hcg_batch_job* jobs, uint32_t waves, uint32_t chunk,
                                hcg_batch** out) {
  HCG_REQUIRE(out, "hcg_batch_create: NULL out");
  *out = nullptr;
  BatchPlan p;
  int rc = plan_batch(njobs, jobs, waves, chunk, p);
  if (rc != HCG_OK) return rc;
  void* mem = nullptr;
  hipError_t e = hipMalloc(&mem, p.bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(HCG_ERR_NOMEM, std::string("hcg_batch_create: hipMalloc: ") + hipGetErrorString(e));
  }
  return batch_init(njobs, jobs, p, mem, true, false, nullptr, out);
}

extern "C" int hcg_batch_create_in(uint32_t njobs, const hcg_batch_job* jobs, uint32_t waves, uint32_t chunk,
                                   void* device_mem, size_t bytes, void* stream, hcg_batch** out) {
  HCG_REQUIRE(out && device_mem, "hcg_batch_create_in: NULL argument");
  *out = nullptr;
  BatchPlan p;
  int rc = plan_batch(njobs, jobs, waves, chunk, p);
  if (rc != HCG_OK) return rc;
  if (bytes < p.bytes || (reinterpret_cast<uintptr_t>(device_mem) & (kAlign - 1)))
    return set_error(HCG_ERR_WORKSPACE, "hcg_batch_create_in: memory smaller than hcg_batch_bytes or not 256-byte "
                                        "aligned");
  return batch_init(njobs, jobs, p, device_mem, false, true, as_stream(stream), out);
}

extern "C" int hcg_batch_enqueue(hcg_batch* b, void* stream) {
  HCG_REQUIRE(b, "hcg_batch_enqueue: NULL batch");
  hipStream_t s = as_stream(stream);
  // timing events only outside stream capture (a captured graph carries no event records)
  bool timing = b->ev[0] != nullptr;
  if (timing) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HCG_CHECK_HIP(hipStreamIsCapturing(s, &cs));
    timing = cs == hipStreamCaptureStatusNone;
  }
  if (timing) HCG_CHECK_HIP(hipEventRecord(b->ev[0], s));
  const uint32_t nc = 2 + b->njobs;
  k_fill_u32<<<ceil_div_u32(nc, 256), 256, 0, s>>>(b->d_counters, nc, 0u);
  HCG_CHECK_LAUNCH("k_fill_u32");
  // every job's channel gather and tile plan, one launch
  int rc = launch_batch_prep(b->d_ranges, b->nranges, b->prep_blocks, b->d_prep, b->d_jobs, s);
  if (rc != HCG_OK) return rc;
  if (timing) HCG_CHECK_HIP(hipEventRecord(b->ev[1], s));
  // the dense jobs' launch (its own queue head after the job counters), then the rest
  if (b->nitems_dense) {
    rc = launch_mfma_batch(kDenseChunk, kDenseWaves, b->blocks_dense, b->d_jobs, b->d_items, b->nitems_dense,
                           b->d_counters + 1 + b->njobs, s);
    if (rc != HCG_OK) return rc;
  }
  if (b->nitems != b->nitems_dense) {
    rc = launch_mfma_batch(b->chunk, b->waves, b->blocks, b->d_jobs, b->d_items + b->nitems_dense,
                           b->nitems - b->nitems_dense, b->d_counters, s);
    if (rc != HCG_OK) return rc;
  }
  if (timing) {
    HCG_CHECK_HIP(hipEventRecord(b->ev[2], s));
    b->timed = true;
  }
  return HCG_OK;
}

extern "C" int hcg_batch_timing(hcg_batch* b, int enable) {
  HCG_REQUIRE(b, "hcg_batch_timing: NULL batch");
  if (enable && !b->ev[0]) {
    for (auto& e : b->ev) HCG_CHECK_HIP(hipEventCreate(&e));
  } else if (!enable && b->ev[0]) {
    for (auto& e : b->ev) {
      HCG_CHECK_HIP(hipEventDestroy(e));
      e = nullptr;
    }
  }
  b->timed = false;
  return HCG_OK;
}

extern "C" int hcg_batch_last_times(hcg_batch* b, float* prep_ms, float* grid_ms) {
  HCG_REQUIRE(b && prep_ms && grid_ms, "hcg_batch_last_times: NULL argument");
  HCG_REQUIRE(b->ev[0] && b->timed, "hcg_batch_last_times: no timed enqueue (hcg_batch_timing first)");
  HCG_CHECK_HIP(hipEventSynchronize(b->ev[2]));
  HCG_CHECK_HIP(hipEventElapsedTime(prep_ms, b->ev[0], b->ev[1]));
  HCG_CHECK_HIP(hipEventElapsedTime(grid_ms, b->ev[1], b->ev[2]));
  return HCG_OK;
}

extern "C" int hcg_batch_info(const hcg_batch* b, hcg_batch_stats* st) {
  HCG_REQUIRE(b && st, "hcg_batch_info: NULL argument");
  st->njobs = b->njobs;
  st->nitems = b->nitems;
  st->blocks = b->blocks;
  st->waves = b->waves;
  st->chunk = b->chunk;
  st->reserved = 0;
  st->device_bytes = b->bytes;
  return HCG_OK;
}

extern "C" int hcg_batch_nonfinite(const hcg_batch* b, uint32_t* counts, void* stream) {
  HCG_REQUIRE(b && counts, "hcg_batch_nonfinite: NULL argument");
  if (b->njobs == 0) return HCG_OK;
  hipStream_t s = as_stream(stream);
  HCG_CHECK_HIP(hipMemcpyAsync(counts, b->d_counters + 1, b->njobs * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HCG_CHECK_HIP(hipStreamSynchronize(s));
  return HCG_OK;
}

extern "C" int hcg_batch_destroy(hcg_batch* b) {
  if (!b) return HCG_OK;
  for (auto& e : b->ev)
    if (e) HCG_CHECK_HIP(hipEventDestroy(e));
  if (b->owns && b->base) HCG_CHECK_HIP(hipFree(b->base));
  delete b;
  return HCG_OK;
}

// ---- graphs ----

extern "C" int hcg_graph_capture_begin(void* stream) {
  HCG_REQUIRE(stream, "hcg_graph_capture_begin: capture needs a created stream, not the null stream");
  HCG_CHECK_HIP(hipStreamBeginCapture(as_stream(stream), hipStreamCaptureModeThreadLocal));
  return HCG_OK;
}

extern "C" int hcg_graph_capture_end(void* stream, void** graph_exec) {
  HCG_REQUIRE(stream && graph_exec, "hcg_graph_capture_end: NULL argument");
  *graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  HCG_CHECK_HIP(hipStreamEndCapture(as_stream(stream), &graph));
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return set_error(HCG_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  *graph_exec = exec;
  return HCG_OK;
}

extern "C" int hcg_graph_launch(void* graph_exec, void* stream) {
  HCG_REQUIRE(graph_exec, "hcg_graph_launch: NULL graph");
  HCG_CHECK_HIP(hipGraphLaunch(static_cast<hipGraphExec_t>(graph_exec), as_stream(stream)));
  return HCG_OK;
}

extern "C" int hcg_graph_destroy(void* graph_exec) {
  if (!graph_exec) return HCG_OK;
  HCG_CHECK_HIP(hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph_exec)));
  return HCG_OK;
}

// presort.hip -- device pre-sort of the samples and the ring lookup table.
//
// Replaces init_input_with_cpu / init_input_with_thrust
// (HCGrid/gridding.cu:70-147): the reference computes HEALPix keys on one
// host thread, sorts 16-byte (key, index) records with boost/Thrust on the
// host, gathers lon/lat/data through the permutation and builds the ring
// LUT with a sequential binary-search loop.  Here all of it stays in HBM:
//
//   k_keys        ang2pix per sample -> (key u32, index u32)
//   k_rs_upsweep  per-tile 8-bit digit histograms           } 4 passes of a
//   k_rs_scan     per-digit scan over tiles + digit offsets } stable LSD radix
//   k_rs_downsweep stable in-tile ranking, scatter          } sort
//   k_gather      permutation gather fused with the unit-vector records
//   k_ring_meta / k_ring_lut  firstring, usedrings, LUT and per-ring cos/sin
//
// The radix sort is reduce-then-scan (no inter-workgroup waiting), stable by
// construction: tiles keep input order, and inside a tile items are ranked
// in index order with wave ballots.  Keys are < 2^31 (npix(8192) < 2^30 and
// MAX_Y = 2^30), so 4 passes of 8 bits order them completely.
#include <hip/hip_runtime.h>

#include <mutex>

#include "common.hpp"
#include "healpix.hpp"

namespace hcg {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 16;
constexpr uint32_t kSortTile = kSortThreads * kSortItems;  // 4096 keys per workgroup
constexpr int kSortWaves = kSortThreads / 64;

__global__ void __launch_bounds__(256) k_keys(hcg_healpix hp, uint32_t n, const double* __restrict__ lon,
                                              const double* __restrict__ lat, uint32_t* __restrict__ keys,
                                              uint32_t* __restrict__ idx) {
  // 64-bit index: with a 32-bit one, i + stride wraps below n for n close to 2^32
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    // HCGrid/gridding.cu:78-80
    const double theta = kHalfPi - kDeg2Rad * lat[i];
    const double phi = kDeg2Rad * lon[i];
    keys[i] = (uint32_t)ang2pix(hp, theta, phi);
    idx[i] = (uint32_t)i;
  }
}

// Digit histogram of one 4096-key tile, written digit-major: hist[d * tiles + t].
__global__ void __launch_bounds__(256) k_rs_upsweep(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                    uint32_t tiles, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[256];
  const uint32_t tid = threadIdx.x;
  h[tid] = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kSortTile;
  const uint64_t lt = lanemask_lt();
#pragma unroll 4
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t i = base + j * kSortThreads + tid;
    const bool valid = i < n;
    const uint32_t d = valid ? ((keys[i] >> shift) & 0xFFu) : 0u;
    const uint64_t peers = match_digit8(d, valid);
    // the lowest lane of each peer group adds the group's count
    if (valid && (peers & lt) == 0) atomicAdd(&h[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  hist[tid * tiles + blockIdx.x] = h[tid];
}

// Exclusive scan of 256 values held one per thread (block of 256).
__device__ inline uint32_t block_exclusive_scan256(uint32_t v, uint32_t* lds_waves, uint32_t& total) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = threadIdx.x >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= (uint32_t)off) x += y;
  }
  if (lane == 63) lds_waves[wave] = x;
  __syncthreads();
  uint32_t wave_off = 0;
  total = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t s = lds_waves[w];
    if (w < (int)wave) wave_off += s;
    total += s;
  }
  __syncthreads();
  return wave_off + x - v;
}

// One workgroup per digit: exclusive scan of its row of tile counts; the row
// total goes to digit_tot[d].
__global__ void __launch_bounds__(256) k_rs_scan_rows(uint32_t* __restrict__ hist, uint32_t tiles,
                                                      uint32_t* __restrict__ digit_tot) {
  __shared__ uint32_t waves[4];
  uint32_t* row = hist + (size_t)blockIdx.x * tiles;
  uint32_t running = 0;
  for (uint32_t base = 0; base < tiles; base += 256 * 4) {
    uint32_t v[4];
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t i = base + threadIdx.x * 4 + q;
      v[q] = (i < tiles) ? row[i] : 0u;
      s += v[q];
    }
    uint32_t total;
    uint32_t pre = block_exclusive_scan256(s, waves, total) + running;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t i = base + threadIdx.x * 4 + q;
      if (i < tiles) row[i] = pre;
      pre += v[q];
    }
    running += total;
  }
  if (threadIdx.x == 0) digit_tot[blockIdx.x] = running;
}

__global__ void __launch_bounds__(256) k_rs_scan_digits(uint32_t* __restrict__ digit_tot) {
  __shared__ uint32_t waves[4];
  uint32_t total;
  const uint32_t v = digit_tot[threadIdx.x];
  const uint32_t ex = block_exclusive_scan256(v, waves, total);
  digit_tot[threadIdx.x] = ex;
}

// Stable scatter of one tile.  Items are visited in index order (round j,
// thread t -> index tile*4096 + j*256 + t); within a round an item's rank
// among equal digits is (items of earlier waves) + (earlier lanes of its wave).
__global__ void __launch_bounds__(256) k_rs_downsweep(const uint32_t* __restrict__ kin,
                                                      const uint32_t* __restrict__ vin,
                                                      uint32_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                      uint32_t n, int shift, uint32_t tiles,
                                                      const uint32_t* __restrict__ hist,
                                                      const uint32_t* __restrict__ digit_off) {
  __shared__ uint32_t base_off[256];
  __shared__ uint32_t wave_cnt[kSortWaves][256];
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = __lane_id();
  const uint32_t wave = tid >> 6;
  base_off[tid] = digit_off[tid] + hist[tid * tiles + blockIdx.x];
  const uint64_t lt = lanemask_lt();
  const uint32_t tile_base = blockIdx.x * kSortTile;
  __syncthreads();
  for (int j = 0; j < kSortItems; ++j) {
    const uint32_t i = tile_base + j * kSortThreads + tid;
    const bool valid = i < n;
    const uint32_t key = valid ? kin[i] : 0u;
    const uint32_t val = valid ? vin[i] : 0u;
    const uint32_t d = (key >> shift) & 0xFFu;
    const uint64_t peers = match_digit8(d, valid);
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    const uint32_t cnt = (uint32_t)__popcll(peers);
#pragma unroll
    for (int q = 0; q < 4; ++q) wave_cnt[wave][lane + 64 * q] = 0;
    if (valid && rank == cnt - 1) wave_cnt[wave][d] = cnt;
    __syncthreads();
    uint32_t pos = base_off[d] + rank;
    for (uint32_t w = 0; w < wave; ++w) pos += wave_cnt[w][d];
    __syncthreads();
    base_off[tid] += wave_cnt[0][tid] + wave_cnt[1][tid] + wave_cnt[2][tid] + wave_cnt[3][tid];
    if (valid) {
      kout[pos] = key;
      vout[pos] = val;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(64) k_set_sentinel(uint32_t* keys, uint32_t n, uint32_t npix) {
  if (threadIdx.x == 0 && blockIdx.x == 0) keys[n] = npix;  // HCGrid/gridding.cu:100
}

// Permutation gather (HCGrid/gridding.cu:101-114) fused with the records the
// gridding kernel reads.
__global__ void __launch_bounds__(256) k_gather(uint32_t n, const uint32_t* __restrict__ perm,
                                                const double* __restrict__ lon, const double* __restrict__ lat,
                                                const void* __restrict__ data, int dtype,
                                                const double* __restrict__ w, hcg_sample* __restrict__ rec,
                                                double* __restrict__ lon_s, double* __restrict__ lat_s,
                                                double* __restrict__ data_s, double* __restrict__ w_s) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    uint32_t j = perm[i];
    if (j >= n) j = (uint32_t)i;  // a permutation never holds this; keeps every read in bounds regardless
    const double lo = lon[j], la = lat[j];
    const double v = !data                      ? 0.0  // multi-channel jobs carry values separately
                     : (dtype == HCG_DTYPE_F64) ? static_cast<const double*>(data)[j]
                                                : (double)static_cast<const float*>(data)[j];
    double sb, cb, sl, cl;
    sincos(la * kDeg2Rad, &sb, &cb);
    sincos(lo * kDeg2Rad, &sl, &cl);
    hcg_sample r;
    r.x = cb * cl;
    r.y = cb * sl;
    r.z = sb;
    r.v = v;
    rec[i] = r;
    if (lon_s) lon_s[i] = lo;
    if (lat_s) lat_s[i] = la;
    if (data_s) data_s[i] = v;
    if (w_s) w_s[i] = w ? w[j] : 1.0;
  }
}

// firstring / usedrings / n_valid (HCGrid/gridding.cu:118-121).  Samples
// whose key is >= npix (theta outside [0, pi]) sort last and are left out of
// the ring span; the reference would take pix2ring of MAX_Y there.
__global__ void __launch_bounds__(64) k_ring_meta(hcg_healpix hp, uint32_t n, const uint32_t* __restrict__ keys,
                                                  hcg_ring_meta* __restrict__ meta) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t nv = lower_bound_u32(keys, 0, n, (uint32_t)hp.npix);
  hcg_ring_meta m;
  m.n = n;
  m.n_valid = nv;
  if (nv == 0) {
    m.firstring = 0;
    m.usedrings = 0;
  } else {
    const uint64_t first = pix2ring(hp, keys[0]);
    const uint64_t last = pix2ring(hp, keys[nv - 1]);
    m.firstring = (uint32_t)first;
    m.usedrings = (uint32_t)(1 + last - first);
  }
  *meta = m;
}

// Level-1 LUT (HCGrid/gridding.cu:122-136): lut[k] = first index with key >=
// startpix(firstring + k).  The reference's incremental searchLastPosLessThan
// walk returns exactly this lower bound.  Also the per-ring cos/sin of the
// ring-centre colatitude (HCGrid/gridding.cu:332-334); ring_s < 0 marks rings
// where the reference's ang2pix takes the cos-only branch.
__global__ void __launch_bounds__(256) k_ring_lut(hcg_healpix hp, const uint32_t* __restrict__ keys,
                                                  const hcg_ring_meta* __restrict__ meta, uint32_t* __restrict__ lut,
                                                  double* __restrict__ ring_z, double* __restrict__ ring_s) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  const hcg_ring_meta m = *meta;
  if (k > m.usedrings) return;
  if (k == 0) {
    lut[0] = 0;
  } else if (k == m.usedrings) {
    lut[k] = m.n;
  } else {
    uint64_t sp, npr;
    ring_info(hp, (uint64_t)m.firstring + k, sp, npr);
    lut[k] = lower_bound_u32(keys, 0, m.n, (uint32_t)sp);
  }
  if (k < m.usedrings) {
    const RingGeom g = ring_geom(hp, (uint64_t)m.firstring + k);
    ring_z[k] = g.z;
    ring_s[k] = g.have_sin ? g.s : -1.0;
  }
}

}  // namespace hcg

using namespace hcg;

namespace {
struct SortWorkspace {
  uint32_t* keys_b;
  uint32_t* vals_b;
  uint32_t* hist;
  uint32_t* digit;
  uint32_t tiles;
};

constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Sort scratch for n keys: ping-pong keys and values, per-tile digit
// counts, per-digit totals.  Every byte of it is written before it is read.
constexpr size_t sort_ws_bytes(uint32_t n) {
  return 2 * align256((size_t)n * 4) + align256((size_t)((n ? n : 1) + kSortTile - 1) / kSortTile * 256 * 4) +
         align256(256 * 4);
}

size_t sort_ws_layout(uint32_t n, void* base, SortWorkspace* ws) {
  const uint32_t tiles = ceil_div_u32(n ? n : 1, kSortTile);
  const size_t kb = align256((size_t)n * 4), hb = align256((size_t)tiles * 256 * 4);
  if (ws) {
    char* p = static_cast<char*>(base);
    ws->keys_b = reinterpret_cast<uint32_t*>(p);
    ws->vals_b = reinterpret_cast<uint32_t*>(p + kb);
    ws->hist = reinterpret_cast<uint32_t*>(p + 2 * kb);
    ws->digit = reinterpret_cast<uint32_t*>(p + 2 * kb + hb);
    ws->tiles = tiles;
  }
  return sort_ws_bytes(n);
}

// Argument checks shared by hcg_presort and hcg_presort_noalloc.
int presort_args(const hcg_healpix* hp, uint32_t n, const double* lon_deg, const double* lat_deg,
                 const uint32_t* keys_sorted, const uint32_t* perm, const char* who) {
  if (hp == nullptr) return set_error(HCG_ERR_INVALID, std::string(who) + ": hp is NULL");
  if (keys_sorted == nullptr || perm == nullptr) return set_error(HCG_ERR_INVALID, std::string(who) + ": NULL output");
  if (hp->nside == 0 || hp->nside > HCG_MAX_NSIDE || (hp->nside & (hp->nside - 1)))
    return set_error(HCG_ERR_UNSUPPORTED, std::string(who) + ": nside must be a power of two <= 8192");
  if (n >= 0xFFFFFFFFu) return set_error(HCG_ERR_INVALID, std::string(who) + ": n must be < 2^32-1");
  if (n && (lon_deg == nullptr || lat_deg == nullptr))
    return set_error(HCG_ERR_INVALID, std::string(who) + ": NULL input");
  return HCG_OK;
}

// Keys, the four radix passes and the sentinel, on workspace `ws` (n > 0).
int presort_enqueue(const hcg_healpix* hp, uint32_t n, const double* lon_deg, const double* lat_deg,
                    uint32_t* keys_sorted, uint32_t* perm, const SortWorkspace& ws, hipStream_t s) {
  const uint32_t blocks = std::min<uint32_t>(ceil_div_u32(n, 256), 8192);
  k_keys<<<blocks, 256, 0, s>>>(*hp, n, lon_deg, lat_deg, keys_sorted, perm);
  HCG_CHECK_LAUNCH("k_keys");

  uint32_t* kin = keys_sorted;
  uint32_t* vin = perm;
  uint32_t* kout = ws.keys_b;
  uint32_t* vout = ws.vals_b;
  for (int pass = 0; pass < 4; ++pass) {
    const int shift = 8 * pass;
    k_rs_upsweep<<<ws.tiles, 256, 0, s>>>(kin, n, shift, ws.tiles, ws.hist);
    HCG_CHECK_LAUNCH("k_rs_upsweep");
    k_rs_scan_rows<<<256, 256, 0, s>>>(ws.hist, ws.tiles, ws.digit);
    HCG_CHECK_LAUNCH("k_rs_scan_rows");
    k_rs_scan_digits<<<1, 256, 0, s>>>(ws.digit);
    HCG_CHECK_LAUNCH("k_rs_scan_digits");
    k_rs_downsweep<<<ws.tiles, 256, 0, s>>>(kin, vin, kout, vout, n, shift, ws.tiles, ws.hist, ws.digit);
    HCG_CHECK_LAUNCH("k_rs_downsweep");
    std::swap(kin, kout);
    std::swap(vin, vout);
  }
  // four passes: the result is back in keys_sorted / perm
  k_set_sentinel<<<1, 64, 0, s>>>(keys_sorted, n, (uint32_t)hp->npix);
  HCG_CHECK_LAUNCH("k_set_sentinel");
  return HCG_OK;
}

// The no-allocation variant's scratch: one sort workspace for up to
// HCG_PRESORT_NOALLOC_MAX_N keys in a file-scope device array.  In host code
// `scratch` names a host-side stand-in, not a device address: its device
// address comes from hipGetSymbolAddress, once per device (each device
// holds its own copy of the array, at its own address).
constexpr size_t kNoAllocScratchBytes = sort_ws_bytes(HCG_PRESORT_NOALLOC_MAX_N);
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch` (resolved on the first call on
// that device).
int scratch_address(void** out) {
  int dev = 0;
  HCG_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(HCG_ERR_UNSUPPORTED, "hcg_presort_noalloc: device index");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HCG_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return set_error(HCG_ERR_HIP, "hcg_presort_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return HCG_OK;
}
}  // namespace

extern "C" size_t hcg_presort_workspace_size(uint32_t n) { return sort_ws_layout(n, nullptr, nullptr); }

extern "C" int hcg_presort(const hcg_healpix* hp, uint32_t n, const double* lon_deg, const double* lat_deg,
                           uint32_t* keys_sorted, uint32_t* perm, void* workspace, size_t workspace_bytes,
                           void* stream) {
  int rc = presort_args(hp, n, lon_deg, lat_deg, keys_sorted, perm, "hcg_presort");
  if (rc != HCG_OK) return rc;
  hipStream_t s = as_stream(stream);
  if (n == 0) {
    k_set_sentinel<<<1, 64, 0, s>>>(keys_sorted, 0, (uint32_t)hp->npix);
    HCG_CHECK_LAUNCH("k_set_sentinel");
    return HCG_OK;
  }
  if (workspace == nullptr || workspace_bytes < hcg_presort_workspace_size(n))
    return set_error(HCG_ERR_WORKSPACE, "hcg_presort: workspace smaller than hcg_presort_workspace_size(n)");
  SortWorkspace ws;
  sort_ws_layout(n, workspace, &ws);
  return presort_enqueue(hp, n, lon_deg, lat_deg, keys_sorted, perm, ws, s);
}

extern "C" int hcg_presort_noalloc(const hcg_healpix* hp, uint32_t n, const double* lon_deg, const double* lat_deg,
                                   uint32_t* keys_sorted, uint32_t* perm, void* stream) {
  int rc = presort_args(hp, n, lon_deg, lat_deg, keys_sorted, perm, "hcg_presort_noalloc");
  if (rc != HCG_OK) return rc;
  if (n > HCG_PRESORT_NOALLOC_MAX_N)
    return set_error(HCG_ERR_UNSUPPORTED, "hcg_presort_noalloc: n exceeds HCG_PRESORT_NOALLOC_MAX_N (8388608); "
                                          "use hcg_presort with a workspace");
  hipStream_t s = as_stream(stream);
  if (n == 0) {
    k_set_sentinel<<<1, 64, 0, s>>>(keys_sorted, 0, (uint32_t)hp->npix);
    HCG_CHECK_LAUNCH("k_set_sentinel");
    return HCG_OK;
  }
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc != HCG_OK) return rc;
  SortWorkspace ws;
  sort_ws_layout(n, base, &ws);
  return presort_enqueue(hp, n, lon_deg, lat_deg, keys_sorted, perm, ws, s);
}

extern "C" int hcg_gather_samples(uint32_t n, const uint32_t* perm, const double* lon_deg, const double* lat_deg,
                                  const void* data, int data_dtype, const double* weights, hcg_sample* records,
                                  double* lon_sorted, double* lat_sorted, double* data_sorted,
                                  double* weights_sorted, void* stream) {
  if (n == 0) return HCG_OK;
  HCG_REQUIRE(perm && lon_deg && lat_deg && records, "hcg_gather_samples: NULL argument");
  HCG_REQUIRE(data_dtype == HCG_DTYPE_F64 || data_dtype == HCG_DTYPE_F32,
              "hcg_gather_samples: data_dtype must be HCG_DTYPE_F64 or HCG_DTYPE_F32");
  const uint32_t blocks = std::min<uint32_t>(ceil_div_u32(n, 256), 8192);
  k_gather<<<blocks, 256, 0, as_stream(stream)>>>(n, perm, lon_deg, lat_deg, data, data_dtype, weights, records,
                                                  lon_sorted, lat_sorted, data_sorted, weights_sorted);
  HCG_CHECK_LAUNCH("k_gather");
  return HCG_OK;
}

extern "C" int hcg_ring_lut(const hcg_healpix* hp, uint32_t n, const uint32_t* keys_sorted, uint32_t* lut,
                            double* ring_z, double* ring_s, hcg_ring_meta* meta, void* stream) {
  HCG_REQUIRE(hp && keys_sorted && lut && ring_z && ring_s && meta, "hcg_ring_lut: NULL argument");
  if (hp->nside == 0 || hp->nside > HCG_MAX_NSIDE)
    return set_error(HCG_ERR_UNSUPPORTED, "hcg_ring_lut: nside must be <= 8192");
  hipStream_t s = as_stream(stream);
  k_ring_meta<<<1, 64, 0, s>>>(*hp, n, keys_sorted, meta);
  HCG_CHECK_LAUNCH("k_ring_meta");
  const uint32_t threads = (uint32_t)hp->nrings + 1;
  k_ring_lut<<<ceil_div_u32(threads, 256), 256, 0, s>>>(*hp, keys_sorted, meta, lut, ring_z, ring_s);
  HCG_CHECK_LAUNCH("k_ring_lut");
  return HCG_OK;
}

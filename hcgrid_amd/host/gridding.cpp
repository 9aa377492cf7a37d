// gridding.cpp -- the reference's driver (HCGrid/gridding.cu:70-556) over the
// C ABI: the pre-sort and LUT run on the device, the gridding kernel is the
// gfx950 one, and the host keeps the reference's module state up to date
// (sorted h_lons/h_lats/h_data, h_hpx_idx with its npix sentinel,
// h_start_ring, h_Healpix.firstring/usedrings).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "hcgrid_amd.h"
#include "hcgrid_host.h"
#include "host_internal.hpp"

using namespace hcghost;

namespace {

// Device-resident buffers of one solve (the reference's d_* globals,
// HCGrid/gridding.cu:41-55).
struct DeviceState {
  void* stream = nullptr;
  uint32_t n = 0;
  uint64_t ncells = 0;
  double *lon = nullptr, *lat = nullptr, *data = nullptr, *weights = nullptr;
  uint32_t *keys = nullptr, *perm = nullptr;
  void* ws = nullptr;
  size_t ws_bytes = 0;
  hcg_sample* rec = nullptr;
  double *lon_s = nullptr, *lat_s = nullptr, *w_s = nullptr;
  uint32_t* lut = nullptr;
  double *ring_z = nullptr, *ring_s = nullptr;
  hcg_ring_meta* meta = nullptr;
  double *xwcs = nullptr, *ywcs = nullptr, *datacube = nullptr, *weightscube = nullptr;
  bool unit_weights = true;
};
DeviceState D;

template <typename T>
bool dalloc(T*& p, size_t count, const char* what) {
  void* v = nullptr;
  if (!check_abi(hcg_malloc(&v, count * sizeof(T)), what)) return false;
  p = static_cast<T*>(v);
  return true;
}

template <typename T>
void dfree(T*& p) {
  if (p) hcg_free(p);
  p = nullptr;
}

void free_device() {
  dfree(D.lon);
  dfree(D.lat);
  dfree(D.data);
  dfree(D.weights);
  dfree(D.keys);
  dfree(D.perm);
  if (D.ws) hcg_free(D.ws);
  D.ws = nullptr;
  dfree(D.rec);
  dfree(D.lon_s);
  dfree(D.lat_s);
  dfree(D.w_s);
  dfree(D.lut);
  dfree(D.ring_z);
  dfree(D.ring_s);
  dfree(D.meta);
  dfree(D.xwcs);
  dfree(D.ywcs);
  dfree(D.datacube);
  dfree(D.weightscube);
  if (D.stream) hcg_stream_destroy(D.stream);
  D = DeviceState();
}

bool ensure_stream() {
  if (D.stream) return true;
  return check_abi(hcg_stream_create(&D.stream), "stream");
}

// ang2pix + stable sort + gathers + ring LUT on the device, then the host
// state the reference leaves behind (HCGrid/gridding.cu:70-147).
void presort_on_device(int sort_param) {
  clear_error();
  const double t0 = cpuSecond();
  if (sort_param < BLOCK_INDIRECT_SORT || sort_param > STL_SORT) {
    fail("init_input: sort_param must be 1..4");
    return;
  }
  if (!h_lons || !h_lats || !h_data) {
    fail("init_input: read_input_map first");
    return;
  }
  const uint32_t n = h_GMaps.data_shape;
  if (!ensure_stream()) return;
  const hcg_healpix hp = abi_healpix();
  D.n = n;
  D.ws_bytes = hcg_presort_workspace_size(n);
  const uint32_t nr = (uint32_t)hp.nrings + 2;
  if (!dalloc(D.lon, n, "alloc") || !dalloc(D.lat, n, "alloc") || !dalloc(D.data, n, "alloc") ||
      !dalloc(D.weights, n, "alloc") || !dalloc(D.keys, (size_t)n + 1, "alloc") || !dalloc(D.perm, n, "alloc") ||
      !check_abi(hcg_malloc(&D.ws, D.ws_bytes), "alloc") || !dalloc(D.rec, n, "alloc") ||
      !dalloc(D.lon_s, n, "alloc") || !dalloc(D.lat_s, n, "alloc") || !dalloc(D.w_s, n, "alloc") ||
      !dalloc(D.lut, nr, "alloc") || !dalloc(D.ring_z, nr, "alloc") || !dalloc(D.ring_s, nr, "alloc") ||
      !dalloc(D.meta, 1, "alloc"))
    return;
  D.unit_weights = true;
  for (uint32_t i = 0; i < n && h_weights; ++i)
    if (h_weights[i] != 1.0) D.unit_weights = false;
  void* s = D.stream;
  double* data_s = nullptr;
  if (!dalloc(data_s, n, "alloc")) return;
  bool ok = check_abi(hcg_memcpy_h2d(D.lon, h_lons, sizeof(double) * n, s), "h2d") &&
            check_abi(hcg_memcpy_h2d(D.lat, h_lats, sizeof(double) * n, s), "h2d") &&
            check_abi(hcg_memcpy_h2d(D.data, h_data, sizeof(double) * n, s), "h2d") &&
            (D.unit_weights || check_abi(hcg_memcpy_h2d(D.weights, h_weights, sizeof(double) * n, s), "h2d")) &&
            check_abi(hcg_presort(&hp, n, D.lon, D.lat, D.keys, D.perm, D.ws, D.ws_bytes, s), "hcg_presort") &&
            check_abi(hcg_gather_samples(n, D.perm, D.lon, D.lat, D.data, HCG_DTYPE_F64,
                                         D.unit_weights ? nullptr : D.weights, D.rec, D.lon_s, D.lat_s, data_s,
                                         D.w_s, s),
                      "hcg_gather_samples") &&
            check_abi(hcg_ring_lut(&hp, n, D.keys, D.lut, D.ring_z, D.ring_s, D.meta, s), "hcg_ring_lut");
  hcg_ring_meta meta{};
  uint32_t* keys32 = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)n + 1));
  ok = ok && keys32 && check_abi(hcg_memcpy_d2h(keys32, D.keys, sizeof(uint32_t) * ((size_t)n + 1), s), "d2h") &&
       check_abi(hcg_memcpy_d2h(&meta, D.meta, sizeof(meta), s), "d2h") &&
       check_abi(hcg_memcpy_d2h(h_lons, D.lon_s, sizeof(double) * n, s), "d2h") &&
       check_abi(hcg_memcpy_d2h(h_lats, D.lat_s, sizeof(double) * n, s), "d2h") &&
       check_abi(hcg_memcpy_d2h(h_data, data_s, sizeof(double) * n, s), "d2h") &&
       check_abi(hcg_stream_synchronize(s), "sync");
  if (ok && !D.unit_weights) ok = check_abi(hcg_memcpy_d2h(h_weights, D.w_s, sizeof(double) * n, s), "d2h");
  hcg_free(data_s);
  if (!ok) {
    free(keys32);
    return;
  }
  free(h_hpx_idx);
  h_hpx_idx = (uint64_t*)malloc(sizeof(uint64_t) * ((size_t)n + 1));
  for (uint64_t i = 0; i <= n; ++i) h_hpx_idx[i] = keys32[i];
  free(keys32);
  h_Healpix.firstring = meta.firstring;
  h_Healpix.usedrings = meta.usedrings;
  free(h_start_ring);
  h_start_ring = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)meta.usedrings + 1));
  if (!check_abi(hcg_memcpy_d2h(h_start_ring, D.lut, sizeof(uint32_t) * ((size_t)meta.usedrings + 1), s), "d2h") ||
      !check_abi(hcg_stream_synchronize(s), "sync"))
    return;
  printf("%f, ", (cpuSecond() - t0) * 1000.);
}

}  // namespace

namespace hcghost {

hcg_kernel abi_kernel() {
  hcg_kernel k{};
  k.kernel_type = h_GMaps.kernel_type;
  k.bearing_needed = h_GMaps.bearing_needed ? 1 : 0;
  k.factor = h_GMaps.factor ? h_GMaps.factor : 1;
  k.selection = h_GMaps.selection;
  for (int i = 0; i < 3; ++i) k.params[i] = h_kernel_params ? h_kernel_params[i] : 0.0;
  if (h_GMaps.kernel_type == GAUSS1D && h_kernel_params) {
    k.params[1] = 0.0;
    k.params[2] = 0.0;
  }
  k.sphere_radius = h_GMaps.sphere_radius;
  k.disc_size = h_GMaps.disc_size;
  return k;
}

bool device_perm(uint32_t* perm) {
  if (!D.perm) return fail("device_perm: no pre-sort on the device");
  return check_abi(hcg_memcpy_d2h(perm, D.perm, sizeof(uint32_t) * D.n, D.stream), "d2h") &&
         check_abi(hcg_stream_synchronize(D.stream), "sync");
}

namespace {

// Round to nearest even: double -> float -> bfloat16 (torch's path) or half.
uint16_t to_bf16(double v) {
  const float f = (float)v;
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((x >> 16) | 0x40);  // quiet NaN
  return (uint16_t)((x + 0x7FFFu + ((x >> 16) & 1u)) >> 16);
}

uint16_t to_f16(double v) {
  const float f = (float)v;
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u, ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return (uint16_t)(sign | (ax > 0x7F800000u ? 0x7E00u : 0x7C00u));
  if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u);  // 65520 and up round to infinity
  if (ax < 0x38800000u) {                                    // below 2^-14: a multiple of 2^-24
    float a;
    memcpy(&a, &ax, 4);
    return (uint16_t)(sign | (uint32_t)nearbyintf(a * 16777216.0f));
  }
  const uint32_t r = ax - 0x38000000u;  // exponent rebias 127 -> 15
  return (uint16_t)(sign | ((r + 0xFFFu + ((r >> 13) & 1u)) >> 13));
}

}  // namespace

bool grid_spectral_channels(const double* values, uint32_t nchan, int dtype, double* maps, double* wsum) {
  if (!D.rec || !D.xwcs) return fail("grid_spectral_channels: pre-sort and data_alloc first");
  if (dtype != HCG_DTYPE_F64 && dtype != HCG_DTYPE_BF16 && dtype != HCG_DTYPE_F16)
    return fail("grid_spectral_channels: dtype must be F64, BF16 or F16");
  const uint32_t n = D.n, side = h_zyx[1];
  const uint64_t m = (uint64_t)h_zyx[1] * h_zyx[2];
  const hcg_healpix hp = abi_healpix();
  const hcg_kernel k = abi_kernel();
  void* s = D.stream;
  const double* w_s = D.unit_weights ? nullptr : D.w_s;
  if (dtype == HCG_DTYPE_F64) {  // hcg_grid_plain per channel: the reference's FP64 arithmetic
    double* dmap = nullptr;
    if (!dalloc(dmap, m ? m : 1, "alloc")) return false;
    bool ok = true;
    for (uint32_t c = 0; ok && c < nchan; ++c) {
      ok = check_abi(hcg_memcpy_h2d(D.data, values + (size_t)c * n, sizeof(double) * n, s), "h2d") &&
           check_abi(hcg_gather_samples(n, D.perm, D.lon, D.lat, D.data, HCG_DTYPE_F64, nullptr, D.rec, nullptr,
                                        nullptr, nullptr, nullptr, s),
                     "hcg_gather_samples") &&
           check_abi(hcg_memset(D.datacube, 0, sizeof(double) * m, s), "memset") &&
           check_abi(hcg_memset(D.weightscube, 0, sizeof(double) * m, s), "memset") &&
           check_abi(hcg_grid_plain(&hp, &k, side, n, D.rec, D.keys, D.lon_s, D.lat_s, w_s, D.lut, D.ring_z,
                                    D.ring_s, D.meta, D.xwcs, D.ywcs, D.datacube, D.weightscube, s),
                     "hcg_grid_plain") &&
           check_abi(hcg_normalize(m, D.datacube, D.weightscube, dmap, s), "hcg_normalize") &&
           check_abi(hcg_memcpy_d2h(maps + (size_t)c * m, dmap, sizeof(double) * m, s), "d2h") &&
           (c != 0 || !wsum || check_abi(hcg_memcpy_d2h(wsum, D.weightscube, sizeof(double) * m, s), "d2h"));
    }
    ok = ok && check_abi(hcg_stream_synchronize(s), "sync");
    hcg_free(dmap);
    return ok;
  }
  // one multi-channel pass: values rounded to 16 bits, FP32 accumulation
  if (nchan == 0) return true;
  const size_t nv = (size_t)nchan * n;
  const size_t np = (size_t)nchan * hcg_channels_padded_len(n);
  uint16_t* hv = (uint16_t*)malloc(sizeof(uint16_t) * (nv ? nv : 1));
  float* hc = (float*)malloc(sizeof(float) * (size_t)nchan * m);
  float* hw = (float*)malloc(sizeof(float) * (m ? m : 1));
  if (!hv || !hc || !hw) {
    free(hv);
    free(hc);
    free(hw);
    return fail("grid_spectral_channels: malloc() failed");
  }
  for (size_t i = 0; i < nv; ++i) hv[i] = dtype == HCG_DTYPE_BF16 ? to_bf16(values[i]) : to_f16(values[i]);
  uint16_t *dv = nullptr, *dsorted = nullptr;
  float *dcube = nullptr, *dws = nullptr;
  uint32_t* dnf = nullptr;
  uint32_t nonfinite = 0;
  bool ok = dalloc(dv, nv ? nv : 1, "alloc") && dalloc(dsorted, np, "alloc") &&
            dalloc(dcube, (size_t)nchan * m, "alloc") && dalloc(dws, m ? m : 1, "alloc") && dalloc(dnf, 1, "alloc");
  ok = ok && check_abi(hcg_memcpy_h2d(dv, hv, sizeof(uint16_t) * nv, s), "h2d") &&
       check_abi(hcg_memset(dnf, 0, sizeof(uint32_t), s), "memset") &&
       check_abi(hcg_gather_channels(n, nchan, D.perm, dv, dtype, 0, dsorted, dnf, s), "hcg_gather_channels") &&
       check_abi(hcg_grid_mfma(&hp, &k, side, n, nchan, dtype, D.rec, D.keys, D.lon_s, D.lat_s, w_s, dsorted, D.lut,
                               D.ring_z, D.ring_s, D.meta, D.xwcs, D.ywcs, dcube, dws, 1, s),
                 "hcg_grid_mfma") &&
       check_abi(hcg_memcpy_d2h(hc, dcube, sizeof(float) * (size_t)nchan * m, s), "d2h") &&
       check_abi(hcg_memcpy_d2h(hw, dws, sizeof(float) * m, s), "d2h") &&
       check_abi(hcg_memcpy_d2h(&nonfinite, dnf, sizeof(uint32_t), s), "d2h") &&
       check_abi(hcg_stream_synchronize(s), "sync");
  if (ok && nonfinite) ok = fail("grid_spectral_channels: " + std::to_string(nonfinite) +
                                 " non-finite channel values (the 16-bit path needs finite values)");
  if (ok) {
    for (size_t i = 0; i < (size_t)nchan * m; ++i) maps[i] = hc[i];
    if (wsum)
      for (uint64_t i = 0; i < m; ++i) wsum[i] = hw[i];
  }
  free(hv);
  free(hc);
  free(hw);
  dfree(dv);
  dfree(dsorted);
  dfree(dcube);
  dfree(dws);
  dfree(dnf);
  return ok;
}

}  // namespace hcghost

void init_input_with_cpu(const int& sort_param) { presort_on_device(sort_param); }

void init_input_with_thrust(const int& sort_param) {
  presort_on_device(sort_param == THRUST ? PARALLEL_STABLE_SORT : sort_param);
}

void init_output() {  // HCGrid/gridding.cu:219-227
  clear_error();
  if (!h_zyx) {
    fail("init_output: read_output_map first");
    return;
  }
  const size_t num = (size_t)h_zyx[0] * h_zyx[1] * h_zyx[2];
  free(h_datacube);
  free(h_weightscube);
  h_datacube = (double*)calloc(num, sizeof(double));
  h_weightscube = (double*)calloc(num, sizeof(double));
  if (!h_datacube || !h_weightscube) fail("init_output: malloc() failed");
}

// HCGrid/gridding.cu:260-275 (host form)
uint32_t searchLastPosLessThan(uint64_t* values, uint32_t left, uint32_t right, uint64_t _key) {
  if (right <= left) return right;
  uint32_t low = left, mid, up = right - 1;
  while (low < up) {
    mid = low + ((up - low + 1) >> 1);
    if (values[mid] < _key)
      low = mid;
    else
      up = mid - 1;
  }
  return values[low] < _key ? low : right;
}

void data_alloc() {  // HCGrid/gridding.cu:406-422 (the sample buffers already live on the device)
  clear_error();
  if (!h_zyx || !ensure_stream()) {
    fail("data_alloc: read_output_map first");
    return;
  }
  D.ncells = (uint64_t)h_zyx[0] * h_zyx[1] * h_zyx[2];
  dalloc(D.xwcs, (size_t)h_zyx[1] * h_zyx[2], "alloc") && dalloc(D.ywcs, (size_t)h_zyx[1] * h_zyx[2], "alloc") &&
      dalloc(D.datacube, D.ncells, "alloc") && dalloc(D.weightscube, D.ncells, "alloc");
}

void data_h2d() {  // HCGrid/gridding.cu:425-445
  clear_error();
  void* s = D.stream;
  const size_t m = (size_t)h_zyx[1] * h_zyx[2];
  check_abi(hcg_memcpy_h2d(D.xwcs, h_xwcs, sizeof(double) * m, s), "h2d") &&
      check_abi(hcg_memcpy_h2d(D.ywcs, h_ywcs, sizeof(double) * m, s), "h2d") &&
      check_abi(hcg_memcpy_h2d(D.datacube, h_datacube, sizeof(double) * D.ncells, s), "h2d") &&
      check_abi(hcg_memcpy_h2d(D.weightscube, h_weightscube, sizeof(double) * D.ncells, s), "h2d");
}

void data_d2h() {  // HCGrid/gridding.cu:448-452
  clear_error();
  void* s = D.stream;
  check_abi(hcg_memcpy_d2h(h_datacube, D.datacube, sizeof(double) * D.ncells, s), "d2h") &&
      check_abi(hcg_memcpy_d2h(h_weightscube, D.weightscube, sizeof(double) * D.ncells, s), "d2h") &&
      check_abi(hcg_stream_synchronize(s), "sync");
}

void data_free() {  // HCGrid/gridding.cu:455-481
  free_device();
  free(h_lons);
  free(h_lats);
  free(h_data);
  free(h_weights);
  free(h_xwcs);
  free(h_ywcs);
  free(h_datacube);
  free(h_weightscube);
  free(h_hpx_idx);
  free(h_start_ring);
  free(h_header);
  free(h_zyx);
  h_lons = h_lats = h_data = h_weights = h_xwcs = h_ywcs = h_datacube = h_weightscube = nullptr;
  h_hpx_idx = nullptr;
  h_start_ring = nullptr;
  h_header = nullptr;
  h_zyx = nullptr;
  // h_kernel_params belongs to the caller here; the reference frees it (gridding.cu:480)
}

// HCGrid/gridding.cu:484-556
void solve_gridding(const char* infile, const char* tarfile, const char* outfile, const char* sortfile,
                    const int& param, const int& bDim) {
  const double t1 = cpuSecond();
  read_input_map(infile);
  if (failed()) return;
  read_output_map(tarfile);
  if (failed()) return;
  if (h_zyx[1] != h_zyx[2]) {
    fail("solve_gridding: the gridding kernel covers NAXIS2 x NAXIS2 cells; the target must be square");
    return;
  }
  set_WCS();
  if (failed()) return;
  init_output();
  if (failed()) return;
  if (param == THRUST)
    init_input_with_thrust(param);
  else
    init_input_with_cpu(param);
  if (failed()) return;
  data_alloc();
  if (failed()) return;
  data_h2d();
  if (failed()) return;
  printf("h_zyx[1]=%d, h_zyx[2]=%d, ", h_zyx[1], h_zyx[2]);
  // launch shape of the reference for the log; the device kernel picks its own tiles
  const int bx = bDim > 0 ? bDim : 64;
  const int gx = (int)((h_GMaps.block_warp_num * h_zyx[1] - 1) / (uint32_t)(bx / 32 > 0 ? bx / 32 : 1) + 1);
  printf("grid.x=%d, block.x=%d, ", gx, bx);

  const hcg_healpix hp = abi_healpix();
  const hcg_kernel k = abi_kernel();
  const double tk = cpuSecond();
  if (!check_abi(hcg_grid_plain(&hp, &k, h_zyx[1], D.n, D.rec, D.keys, D.lon_s, D.lat_s,
                                D.unit_weights ? nullptr : D.w_s, D.lut, D.ring_z, D.ring_s, D.meta, D.xwcs,
                                D.ywcs, D.datacube, D.weightscube, D.stream),
                 "hcg_grid_plain") ||
      !check_abi(hcg_stream_synchronize(D.stream), "sync"))
    return;
  printf("kernel elapsed time=%f, ", (cpuSecond() - tk) * 1000.);
  data_d2h();
  if (failed()) return;
  write_output_map(outfile);
  if (failed()) return;
  if (sortfile) {
    write_ordered_map(infile, sortfile);
    if (failed()) return;
  }
  data_free();
  printf("solving_gridding time=%f\n", (cpuSecond() - t1) * 1000.);
}

// gridding_plus.cpp -- HCGrid+'s spectral driver (HCGrid+/gridding-omp.cu:
// 552-678) and its HDF5 input (HCGrid+/gmap.cpp:109-191) over the C ABI.
//
// HCGrid+ reads coordinates and channels from HDF5, pre-sorts once, then runs
// the single-channel kernel per channel on OpenMP streams and writes only
// channel 0 (HCGrid+/gmap.cpp:354-361).  Here the pre-sort runs on the
// device once, every channel is gridded (FP64 per channel, or one
// multi-channel pass when h_GMaps.spec_dtype asks for 16-bit values), and the
// whole cube is written.  The HDF5 reader lives in libhcgrid_h5.so, loaded
// on first use from the directory of this library.
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "fits_io.hpp"
#include "hcgrid_amd.h"
#include "hcgrid_host.h"
#include "host_internal.hpp"

using namespace hcghost;

double* h_spec_data = nullptr;
double* h_spec_cube = nullptr;

namespace {

using ReadCoordsFn = int (*)(const char*, uint32_t*, uint32_t*, double**, double**, char*, size_t);
using ReadSignalFn = int (*)(const char*, uint32_t, uint32_t, double*, char*, size_t);

struct H5Lib {
  ReadCoordsFn read_coords = nullptr;
  ReadSignalFn read_signal = nullptr;
};

// libhcgrid_h5.so next to this library (dladdr of one of its own functions).
bool h5lib(H5Lib& out) {
  static H5Lib lib;
  static std::string load_error;
  static bool tried = false;
  if (!tried) {
    tried = true;
    Dl_info info{};
    std::string dir = ".";
    if (dladdr((void*)&read_input_coordinate, &info) && info.dli_fname) {
      dir = info.dli_fname;
      const size_t slash = dir.rfind('/');
      dir = slash == std::string::npos ? "." : dir.substr(0, slash);
    }
    const std::string path = dir + "/libhcgrid_h5.so";
    void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) {
      const char* e = dlerror();
      load_error = "HDF5 support is not available (" + path + ": " + (e ? e : "dlopen failed") + ")";
    } else {
      lib.read_coords = (ReadCoordsFn)dlsym(h, "hcg_h5_read_coords");
      lib.read_signal = (ReadSignalFn)dlsym(h, "hcg_h5_read_signal");
      if (!lib.read_coords || !lib.read_signal) load_error = "libhcgrid_h5.so lacks its entry points";
    }
  }
  if (!load_error.empty()) return fail(load_error);
  out = lib;
  return true;
}

// The header records of the target from the 7th card up to END, '#'
// meta-comments dropped (HCGrid+/gmap.cpp:318-345).
std::vector<std::string> target_cards() {
  std::vector<std::string> out;
  const size_t len = h_header ? strlen(h_header) : 0;
  for (size_t k = 80 * 6; k + 80 <= len; k += 80) {
    std::string rec(h_header + k, 80);
    if (rec.compare(0, 10, "END       ") == 0) break;
    if (rec[0] == '#') continue;
    out.push_back(rec);
  }
  return out;
}

}  // namespace

// HCGrid+/gmap.cpp:109-166
void read_input_coordinate(const char* infile) {
  clear_error();
  H5Lib lib;
  if (!h5lib(lib)) return;
  uint32_t n = 0, nchan = 0;
  double *lon = nullptr, *lat = nullptr;
  char err[512] = "";
  if (lib.read_coords(infile, &n, &nchan, &lon, &lat, err, sizeof(err)) != 0) {
    fail(std::string("read_input_coordinate: ") + err);
    return;
  }
  free(h_lons);
  free(h_lats);
  free(h_weights);
  h_lons = lon;
  h_lats = lat;
  h_weights = (double*)malloc(sizeof(double) * (n ? n : 1));
  if (!h_weights) {
    fail("read_input_coordinate: malloc() failed");
    return;
  }
  for (uint32_t i = 0; i < n; ++i) h_weights[i] = 1.;
  h_GMaps.data_shape = n;
  h_GMaps.spec_dim = nchan;
}

// HCGrid+/gmap.cpp:168-191
void read_input_data(const char* infile) {
  clear_error();
  H5Lib lib;
  if (!h5lib(lib)) return;
  const uint32_t n = h_GMaps.data_shape, nchan = h_GMaps.spec_dim;
  free(h_spec_data);
  const size_t nv = (size_t)nchan * n;
  h_spec_data = (double*)malloc(sizeof(double) * (nv ? nv : 1));
  if (!h_spec_data) {
    fail("read_input_data: malloc() failed");
    return;
  }
  char err[512] = "";
  if (lib.read_signal(infile, n, nchan, h_spec_data, err, sizeof(err)) != 0)
    fail(std::string("read_input_data: ") + err);
}

// HCGrid+/gmap.cpp:295-364 with every channel: NAXIS1 x NAXIS2 x spec_dim,
// BITPIX -64, the target's header records, the normalised maps of h_spec_cube.
void write_output_cube(const char* outfile) {
  clear_error();
  if (!h_header || !h_zyx || !h_spec_cube) {
    fail("write_output_cube: nothing to write");
    return;
  }
  hcgio::HduOut h;
  h.bitpix = -64;
  h.naxes = {(int64_t)h_zyx[1], (int64_t)h_zyx[2], (int64_t)h_GMaps.spec_dim};
  h.extra_cards = target_cards();
  h.data = h_spec_cube;
  h.ndata = (uint64_t)h_zyx[1] * h_zyx[2] * h_GMaps.spec_dim;
  std::string err;
  if (!hcgio::write_file(outfile, {h}, err)) fail("write_output_cube: " + err);
}

// The sorted input as FITS (HCGrid+ passes its HDF5 input to
// write_ordered_map, HCGrid+/gridding-omp.cu:664-666, which cannot open it):
// primary 1 x N x spec_dim values in sorted order, extension N x 2 sorted
// coordinates (longitudes, then latitudes), HCGrid's input layout.
void write_ordered_cube(const char* outfile) {
  clear_error();
  const uint32_t n = h_GMaps.data_shape, nchan = h_GMaps.spec_dim;
  if (!h_lons || !h_lats || (nchan && !h_spec_data)) {
    fail("write_ordered_cube: nothing to write");
    return;
  }
  std::vector<uint32_t> perm(n);
  if (n && !device_perm(perm.data())) return;
  std::vector<double> vals((size_t)nchan * n), coords(2 * (size_t)n);
  for (uint32_t c = 0; c < nchan; ++c)
    for (uint32_t i = 0; i < n; ++i) vals[(size_t)c * n + i] = h_spec_data[(size_t)c * n + perm[i]];
  for (uint32_t i = 0; i < n; ++i) {
    coords[i] = h_lons[i];  // sorted by the pre-sort
    coords[(size_t)n + i] = h_lats[i];
  }
  hcgio::HduOut prim, ext;
  prim.bitpix = -64;
  prim.naxes = {1, (int64_t)n, (int64_t)nchan};
  prim.data = vals.data();
  prim.ndata = vals.size();
  ext.bitpix = -64;
  ext.naxes = {(int64_t)n, 2};
  ext.data = coords.data();
  ext.ndata = coords.size();
  std::string err;
  if (!hcgio::write_file(outfile, {prim, ext}, err)) fail("write_ordered_cube: " + err);
}

// HCGrid+/gridding-omp.cu:552-678
void solve_gridding(const char* infile, const char* tarfile, const char* outfile, const char* sortfile,
                    const int& param, const int& bDim, int argc, char** argv) {
  (void)argc;
  (void)argv;
  const double t1 = cpuSecond();
  read_input_coordinate(infile);
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
  read_input_data(infile);
  if (failed()) return;
  const uint32_t n = h_GMaps.data_shape, nchan = h_GMaps.spec_dim;
  // the pre-sort carries h_data along (HCGrid's single-channel state): channel 0
  free(h_data);
  h_data = (double*)calloc(n ? n : 1, sizeof(double));
  if (!h_data) {
    fail("solve_gridding: malloc() failed");
    return;
  }
  if (nchan) memcpy(h_data, h_spec_data, sizeof(double) * n);
  if (param == THRUST)
    init_input_with_thrust(param);
  else
    init_input_with_cpu(param);
  if (failed()) return;
  data_alloc();
  if (failed()) return;
  data_h2d();
  if (failed()) return;
  const int bx = bDim > 0 ? bDim : 96;
  printf("h_zyx[1]=%d, h_zyx[2]=%d, channels=%u, block.x=%d, ", h_zyx[1], h_zyx[2], nchan, bx);
  const uint64_t m = (uint64_t)h_zyx[1] * h_zyx[2];
  free(h_spec_cube);
  const size_t nc = (size_t)nchan * m;
  h_spec_cube = (double*)malloc(sizeof(double) * (nc ? nc : 1));
  if (!h_spec_cube) {
    fail("solve_gridding: malloc() failed");
    return;
  }
  const double tk = cpuSecond();
  if (!grid_spectral_channels(h_spec_data, nchan, (int)h_GMaps.spec_dtype, h_spec_cube, h_weightscube)) return;
  printf("kernel elapsed time=%f, ", (cpuSecond() - tk) * 1000.);
  write_output_cube(outfile);
  if (failed()) return;
  if (sortfile) {
    write_ordered_cube(sortfile);
    if (failed()) return;
  }
  data_free();
  free(h_spec_data);
  free(h_spec_cube);
  h_spec_data = h_spec_cube = nullptr;
  printf("solving_gridding time=%f\n", (cpuSecond() - t1) * 1000.);
}

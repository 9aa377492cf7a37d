// hdf5_io.cpp -- the HCGrid+ HDF5 input layout (HCGrid+/gmap.cpp:109-191),
// built as libhcgrid_h5.so over libhdf5 and loaded by the host layer only
// when an HDF5 file is read or written, so programs that never touch HDF5
// never load libhdf5.
//
// Layout (what HCGrid+ reads):
//   /coords                group, attribute "num_samples" (integer) = N
//   /coords/xcoords        N longitudes, degrees
//   /coords/ycoords        N latitudes, degrees
//   /signal/signal_<c>     N values of channel c = 0, 1, ...
// Datasets of any floating or integer type are converted to double on read
// (H5T_NATIVE_DOUBLE as HCGrid+ reads them).  HCGrid+ reads exactly five
// channels (spec_dim = 5, HCGrid+/gmap.cpp:140); here the channels are
// signal_0, signal_1, ... up to the first missing one.
#include <hdf5.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

namespace {

void set_err(char* err, size_t errlen, const std::string& msg) {
  if (err && errlen) {
    strncpy(err, msg.c_str(), errlen - 1);
    err[errlen - 1] = '\0';
  }
}

// Closes an HDF5 handle when it goes out of scope.
struct H {
  hid_t id = -1;
  herr_t (*close)(hid_t) = nullptr;
  H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
  ~H() {
    if (id >= 0 && close) close(id);
  }
  H(const H&) = delete;
  H& operator=(const H&) = delete;
  bool ok() const { return id >= 0; }
};

bool link_exists(hid_t loc, const char* name) { return H5Lexists(loc, name, H5P_DEFAULT) > 0; }

// Reads a 1-D dataset of exactly n elements as double.
bool read_1d(hid_t group, const char* name, uint64_t n, double* out, std::string& msg) {
  if (!link_exists(group, name)) {
    msg = std::string("dataset ") + name + " not found";
    return false;
  }
  H ds(H5Dopen2(group, name, H5P_DEFAULT), H5Dclose);
  if (!ds.ok()) {
    msg = std::string("cannot open dataset ") + name;
    return false;
  }
  H sp(H5Dget_space(ds.id), H5Sclose);
  const hssize_t npts = H5Sget_simple_extent_npoints(sp.id);
  if (npts < 0 || (uint64_t)npts != n) {
    msg = std::string("dataset ") + name + " has " + std::to_string((long long)npts) + " elements, num_samples " +
          std::to_string((unsigned long long)n);
    return false;
  }
  if (n && H5Dread(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, out) < 0) {
    msg = std::string("cannot read dataset ") + name;
    return false;
  }
  return true;
}

bool write_1d(hid_t group, const char* name, uint64_t n, const double* data, bool as_float, std::string& msg) {
  const hsize_t dims[1] = {(hsize_t)n};
  H sp(H5Screate_simple(1, dims, nullptr), H5Sclose);
  H ds(H5Dcreate2(group, name, as_float ? H5T_IEEE_F32LE : H5T_IEEE_F64LE, sp.id, H5P_DEFAULT, H5P_DEFAULT,
                  H5P_DEFAULT),
       H5Dclose);
  if (!ds.ok() || (n && H5Dwrite(ds.id, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)) {
    msg = std::string("cannot write dataset ") + name;
    return false;
  }
  return true;
}

// Opens a file for reading; errors come back as return codes, not printed.
hid_t open_ro(const char* path, char* err, size_t errlen) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  const hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) set_err(err, errlen, std::string("cannot open HDF5 file ") + path);
  return f;
}

}  // namespace

extern "C" {

// Reads /coords (HCGrid+/gmap.cpp:109-166): *lon and *lat (n each) are
// malloc'ed and owned by the caller; *nchan = the number of consecutive
// /signal/signal_<c> datasets.  Returns 0, or -1 with a message in err.
int hcg_h5_read_coords(const char* path, uint32_t* n_out, uint32_t* nchan_out, double** lon, double** lat,
                       char* err, size_t errlen) {
  *lon = *lat = nullptr;
  *n_out = *nchan_out = 0;
  H f(open_ro(path, err, errlen), H5Fclose);
  if (!f.ok()) return -1;
  if (!link_exists(f.id, "coords")) {
    set_err(err, errlen, "group coords not found");
    return -1;
  }
  H coords(H5Gopen2(f.id, "coords", H5P_DEFAULT), H5Gclose);
  if (!coords.ok() || H5Aexists(coords.id, "num_samples") <= 0) {
    set_err(err, errlen, "attribute coords/num_samples not found");
    return -1;
  }
  long long ns = -1;
  {
    H a(H5Aopen(coords.id, "num_samples", H5P_DEFAULT), H5Aclose);
    if (!a.ok() || H5Aread(a.id, H5T_NATIVE_LLONG, &ns) < 0) {
      set_err(err, errlen, "cannot read coords/num_samples");
      return -1;
    }
  }
  if (ns < 0 || ns > 0xFFFFFFFFll) {
    set_err(err, errlen, "coords/num_samples out of range");
    return -1;
  }
  const uint64_t n = (uint64_t)ns;
  std::string msg;
  double* x = (double*)malloc(sizeof(double) * (n ? n : 1));
  double* y = (double*)malloc(sizeof(double) * (n ? n : 1));
  if (!x || !y || !read_1d(coords.id, "xcoords", n, x, msg) || !read_1d(coords.id, "ycoords", n, y, msg)) {
    free(x);
    free(y);
    set_err(err, errlen, msg.empty() ? "out of memory" : msg);
    return -1;
  }
  uint32_t nchan = 0;
  if (link_exists(f.id, "signal")) {
    H sig(H5Gopen2(f.id, "signal", H5P_DEFAULT), H5Gclose);
    while (sig.ok() && link_exists(sig.id, ("signal_" + std::to_string(nchan)).c_str())) ++nchan;
  }
  *lon = x;
  *lat = y;
  *n_out = (uint32_t)n;
  *nchan_out = nchan;
  return 0;
}

// Reads /signal/signal_0 .. signal_<nchan-1> (HCGrid+/gmap.cpp:168-191), n
// values each, into values (caller-allocated, nchan x n, channel-major).
int hcg_h5_read_signal(const char* path, uint32_t n, uint32_t nchan, double* values, char* err, size_t errlen) {
  H f(open_ro(path, err, errlen), H5Fclose);
  if (!f.ok()) return -1;
  if (nchan == 0) return 0;
  if (!link_exists(f.id, "signal")) {
    set_err(err, errlen, "group signal not found");
    return -1;
  }
  H sig(H5Gopen2(f.id, "signal", H5P_DEFAULT), H5Gclose);
  std::string msg;
  for (uint32_t c = 0; c < nchan; ++c) {
    const std::string name = "signal_" + std::to_string(c);
    if (!sig.ok() || !read_1d(sig.id, name.c_str(), n, values + (size_t)c * n, msg)) {
      set_err(err, errlen, msg.empty() ? "cannot open group signal" : msg);
      return -1;
    }
  }
  return 0;
}

// Writes the layout above (datasets as IEEE float32 when as_float32, else
// float64; num_samples as a 32-bit integer, HCGrid+ reads it as an int, or
// 64-bit from 2^31 on).  Returns 0, or -1 with err.
int hcg_h5_write_input(const char* path, uint32_t n, const double* lon, const double* lat, uint32_t nchan,
                       const double* values, int as_float32, char* err, size_t errlen) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
  H f(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), H5Fclose);
  if (!f.ok()) {
    set_err(err, errlen, std::string("cannot create HDF5 file ") + path);
    return -1;
  }
  std::string msg;
  H coords(H5Gcreate2(f.id, "coords", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (!coords.ok()) {
    set_err(err, errlen, "cannot create group coords");
    return -1;
  }
  {
    H sp(H5Screate(H5S_SCALAR), H5Sclose);
    const bool wide = n > 0x7FFFFFFFu;
    H a(H5Acreate2(coords.id, "num_samples", wide ? H5T_STD_I64LE : H5T_STD_I32LE, sp.id, H5P_DEFAULT,
                   H5P_DEFAULT),
        H5Aclose);
    const long long ns = (long long)n;
    if (!a.ok() || H5Awrite(a.id, H5T_NATIVE_LLONG, &ns) < 0) {
      set_err(err, errlen, "cannot write coords/num_samples");
      return -1;
    }
  }
  const bool f32 = as_float32 != 0;
  if (!write_1d(coords.id, "xcoords", n, lon, f32, msg) || !write_1d(coords.id, "ycoords", n, lat, f32, msg)) {
    set_err(err, errlen, msg);
    return -1;
  }
  H sig(H5Gcreate2(f.id, "signal", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), H5Gclose);
  if (!sig.ok()) {
    set_err(err, errlen, "cannot create group signal");
    return -1;
  }
  for (uint32_t c = 0; c < nchan; ++c) {
    const std::string name = "signal_" + std::to_string(c);
    if (!write_1d(sig.id, name.c_str(), n, values + (size_t)c * n, f32, msg)) {
      set_err(err, errlen, msg);
      return -1;
    }
  }
  return 0;
}

}  // extern "C"

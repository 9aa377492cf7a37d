// main_plus.cpp -- the HCGrid+ command line (HCGrid+/HCGrid.cpp:38-170): same
// flags and file-name composition, input "<path><input><id>.hdf5", GAUSS1D
// with sigma = FWHM / sqrt(8 ln 2), radius 5 sigma, HEALPix resolution
// sigma / 2, default block size 96 (reported only).  Differences: --beam_size
// is honoured (HCGrid+ assigns it to a shadowing local, HCGrid+/HCGrid.cpp:
// 131-133, so it always grids with 300"), every channel is written, and two
// options have no counterpart: --selection compat|disc and --cube_dtype
// f64|bf16|f16 (FP64 per channel, the default, or one multi-channel pass).
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "hcgrid_amd.h"
#include "hcgrid_host.h"

int main(int argc, char** argv) {
  char *path = nullptr, *ifile = nullptr, *tfile = nullptr, *ofile = nullptr, *sfile = nullptr, *num = nullptr,
       *beam = nullptr, *order = nullptr, *bDim = nullptr, *factor = nullptr, *selection = nullptr,
       *cube_dtype = nullptr;
  static const struct option long_options[] = {
      {"helparg", no_argument, nullptr, 'h'},          {"fits_path", required_argument, nullptr, 'p'},
      {"input_file", required_argument, nullptr, 'i'}, {"target_file", required_argument, nullptr, 't'},
      {"output_file", required_argument, nullptr, 'o'}, {"sorted_file", required_argument, nullptr, 's'},
      {"fits_id", required_argument, nullptr, 'n'},    {"beam_size", required_argument, nullptr, 'b'},
      {"order_arg", required_argument, nullptr, 'd'},  {"block_num", required_argument, nullptr, 'a'},
      {"coarsening_factor", required_argument, nullptr, 'f'},
      {"selection", required_argument, nullptr, 'S'},  {"cube_dtype", required_argument, nullptr, 'C'},
      {nullptr, 0, nullptr, 0}};
  int option_index = 0, c;
  while ((c = getopt_long_only(argc, argv, "hp:i:t:o:s:n:b:d:a:", long_options, &option_index)) != -1) {
    switch (c) {
      case 'h':
        fprintf(stderr,
                "useage: ./HCGridPlus --fits_path <absolute path> --input_file <input file> --target_file <target "
                "file> --sorted_file <sorted file> --output_file <output file> --fits_id <number> --beam_size "
                "<beam> --order_arg <order> --block_num <num> [--selection compat|disc] [--cube_dtype "
                "f64|bf16|f16]\n");
        return 1;
      case 'p': path = optarg; break;
      case 'i': ifile = optarg; break;
      case 't': tfile = optarg; break;
      case 'o': ofile = optarg; break;
      case 's': sfile = optarg; break;
      case 'n': num = optarg; break;
      case 'b': beam = optarg; break;
      case 'd': order = optarg; break;
      case 'a': bDim = optarg; break;
      case 'f': factor = optarg; break;
      case 'S': selection = optarg; break;
      case 'C': cube_dtype = optarg; break;
      default: return 1;
    }
  }
  if (!path || !ifile || !tfile || !ofile || !num || !order) {
    fprintf(stderr, "HCGridPlus: --fits_path, --input_file, --target_file, --output_file, --fits_id and "
                    "--order_arg are required\n");
    return 1;
  }
  uint32_t sel = HCG_SELECT_COMPAT, dtype = HCG_DTYPE_F64;
  if (selection) {
    if (strcmp(selection, "disc") == 0) {
      sel = HCG_SELECT_DISC;
    } else if (strcmp(selection, "compat") != 0) {
      fprintf(stderr, "HCGridPlus: --selection must be compat or disc\n");
      return 1;
    }
  }
  if (cube_dtype) {
    if (strcmp(cube_dtype, "bf16") == 0) {
      dtype = HCG_DTYPE_BF16;
    } else if (strcmp(cube_dtype, "f16") == 0) {
      dtype = HCG_DTYPE_F16;
    } else if (strcmp(cube_dtype, "f64") != 0) {
      fprintf(stderr, "HCGridPlus: --cube_dtype must be f64, bf16 or f16\n");
      return 1;
    }
  }
  // "<path><name><id>.hdf5" in, "<path><name><id>.fits" out (overwritten)
  const std::string infile = std::string(path) + ifile + num + ".hdf5";
  const std::string tarfile = std::string(path) + tfile + num + ".fits";
  const std::string outfile = std::string("!") + path + ofile + num + ".fits";
  const std::string sortfile = sfile ? std::string("!") + path + sfile + num + ".fits" : std::string();

  _Healpix_init(1, RING);
  double kernelsize_fwhm = 300. / 3600.;
  if (beam) kernelsize_fwhm = atoi(beam) / 3600.;
  const double kernelsize_sigma = kernelsize_fwhm / sqrt(8 * log(2));
  double* kernel_params = (double*)calloc(3, sizeof(double));
  kernel_params[0] = kernelsize_sigma;
  _prepare_grid_kernel(GAUSS1D, kernel_params, 5. * kernelsize_sigma, kernelsize_sigma / 2.);
  h_GMaps.factor = factor ? atoi(factor) : 1;
  h_GMaps.selection = sel;
  h_GMaps.spec_dtype = dtype;
  solve_gridding(infile.c_str(), tarfile.c_str(), outfile.c_str(), sfile ? sortfile.c_str() : nullptr, atoi(order),
                 bDim ? atoi(bDim) : 96, argc, argv);
  free(kernel_params);
  return 0;
}

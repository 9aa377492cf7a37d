// main.cpp -- the HCGrid command line (HCGrid/HCGrid.cpp:38-197): same
// flags, same file-name composition, same kernel set-up (GAUSS1D, sigma =
// FWHM / sqrt(8 ln 2), radius 5 sigma, HEALPix resolution sigma / 2), same
// block-size heuristic (reported only: the gfx950 kernel chooses its tiles).
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
       *beam = nullptr, *order = nullptr, *bDim = nullptr, *Rg = nullptr, *Sp = nullptr, *factor = nullptr,
       *selection = nullptr;
  static const struct option long_options[] = {
      {"helparg", no_argument, nullptr, 'h'},          {"fits_path", required_argument, nullptr, 'p'},
      {"input_file", required_argument, nullptr, 'i'}, {"target_file", required_argument, nullptr, 't'},
      {"output_file", required_argument, nullptr, 'o'}, {"sorted_file", required_argument, nullptr, 's'},
      {"fits_id", required_argument, nullptr, 'n'},    {"beam_size", required_argument, nullptr, 'b'},
      {"order_arg", required_argument, nullptr, 'd'},  {"block_num", required_argument, nullptr, 'a'},
      {"register_num", required_argument, nullptr, 'r'}, {"sp_num", required_argument, nullptr, 'm'},
      {"coarsening_factor", required_argument, nullptr, 'f'},
      // not in the reference: "compat" (default, the reference's windows) or "disc"
      {"selection", required_argument, nullptr, 'S'}, {nullptr, 0, nullptr, 0}};
  int option_index = 0, c;
  while ((c = getopt_long_only(argc, argv, "hp:i:t:o:s:n:b:d:a:", long_options, &option_index)) != -1) {
    switch (c) {
      case 'h':
        fprintf(stderr,
                "useage: ./HCGrid --fits_path <absolute path> --input_file <input file> --target_file <target file> "
                "--sorted_file <sorted file> --output_file <output file> --fits_id <number> --beam_size <beam> "
                "--order_arg <order> --block_num <num>\n");
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
      case 'r': Rg = optarg; break;
      case 'm': Sp = optarg; break;
      case 'f': factor = optarg; break;
      case 'S': selection = optarg; break;
      default: return 1;
    }
  }
  if (!path || !ifile || !tfile || !ofile || !num || !order) {
    fprintf(stderr, "HCGrid: --fits_path, --input_file, --target_file, --output_file, --fits_id and "
                    "--order_arg are required\n");
    return 1;
  }
  // "<path><name><id>.fits", outputs prefixed with '!' (overwrite)
  const std::string infile = std::string(path) + ifile + num + ".fits";
  const std::string tarfile = std::string(path) + tfile + num + ".fits";
  const std::string outfile = std::string("!") + path + ofile + num + ".fits";
  const std::string sortfile = sfile ? std::string("!") + path + sfile + num + ".fits" : std::string();
  printf("order: %s, num: %s, ", order, num);

  _Healpix_init(1, RING);
  const uint32_t kernel_type = GAUSS1D;
  double kernelsize_fwhm = 300. / 3600.;
  if (beam) kernelsize_fwhm = atoi(beam) / 3600.;
  const double kernelsize_sigma = kernelsize_fwhm / sqrt(8 * log(2));
  double* kernel_params = (double*)calloc(3, sizeof(double));
  kernel_params[0] = kernelsize_sigma;
  const double sphere_radius = 5. * kernelsize_sigma;
  const double hpx_max_resolution = kernelsize_sigma / 2.;
  _prepare_grid_kernel(kernel_type, kernel_params, sphere_radius, hpx_max_resolution);

  h_GMaps.factor = 1;
  if (factor) h_GMaps.factor = atoi(factor);
  printf("h_GMaps.factor=%d, ", h_GMaps.factor);
  h_GMaps.selection = HCG_SELECT_COMPAT;
  if (selection) {
    if (strcmp(selection, "disc") == 0) {
      h_GMaps.selection = HCG_SELECT_DISC;
    } else if (strcmp(selection, "compat") != 0) {
      fprintf(stderr, "HCGrid: --selection must be compat or disc\n");
      return 1;
    }
  }
  int block = 64;
  if (bDim) {
    block = atoi(bDim);
  } else if (Rg && Sp) {  // HCGrid/HCGrid.cpp:167-174
    const int Register = atoi(Rg) * 1024, SP = atoi(Sp);
    const int T_max = Register / 184, T_max_h = T_max / 2;
    block = (SP >= 32 && SP < T_max_h) ? SP : T_max - T_max % 32;
    if (sfile) printf("blockDim_x=%d\n", block);
  }
  solve_gridding(infile.c_str(), tarfile.c_str(), outfile.c_str(), sfile ? sortfile.c_str() : nullptr, atoi(order),
                 block);
  free(kernel_params);
  return 0;
}

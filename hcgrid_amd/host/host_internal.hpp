// host_internal.hpp -- shared helpers of the C++ host layer (not installed).
#pragma once

#include <string>

#include "hcgrid_amd.h"

namespace hcghost {
// Reports an error the reference's way (stderr + exit(1)) unless exit-on-error
// is off; returns false.
bool fail(const std::string& msg);
bool failed();
void clear_error();
// Checks a C ABI return code.
bool check_abi(int rc, const char* where);
// The ABI view of the module-global HEALPix state.
hcg_healpix abi_healpix();
// The ABI view of h_GMaps / h_kernel_params.
hcg_kernel abi_kernel();
// The sort permutation of the last pre-sort (host copy of n entries).
bool device_perm(uint32_t* perm);
// Grids nchan channels ([nchan][n] values in input order) over the geometry
// of the last pre-sort and the target of data_alloc/data_h2d: maps receives
// [nchan][NAXIS1*NAXIS2] normalised maps (NaN where no weight), wsum (may
// be NULL) the weight sums.  dtype HCG_DTYPE_F64 runs hcg_grid_plain per
// channel; HCG_DTYPE_BF16/F16 one hcg_grid_mfma pass on 16-bit values.
bool grid_spectral_channels(const double* values, uint32_t nchan, int dtype, double* maps, double* wsum);
}  // namespace hcghost

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
}  // namespace hcghost

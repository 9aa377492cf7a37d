// common.hpp -- error plumbing and small device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "hcgrid_amd.h"

namespace hcg {

// Records the detail of the calling thread's last failure and returns `code`.
int set_error(int code, const std::string& msg);

#define HCG_CHECK_HIP(expr)                                                              \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::hcg::set_error(HCG_ERR_HIP, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

#define HCG_CHECK_LAUNCH(name)                                                           \
  do {                                                                                   \
    hipError_t e_ = hipGetLastError();                                                   \
    if (e_ != hipSuccess)                                                                \
      return ::hcg::set_error(HCG_ERR_HIP, std::string(name " launch: ") + hipGetErrorString(e_)); \
  } while (0)

#define HCG_REQUIRE(cond, msg)                                                           \
  do {                                                                                   \
    if (!(cond)) return ::hcg::set_error(HCG_ERR_INVALID, msg);                          \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline uint32_t ceil_div_u32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Lanes of the wave whose 8-bit digit equals this lane's (wave64 ballots).
__device__ inline uint64_t match_digit8(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t vote = __ballot(bit);
    m &= bit ? vote : ~vote;
  }
  return m;
}

__device__ inline uint64_t lanemask_lt() {
  const uint32_t lane = __lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// First index in [lo, hi) whose key >= key (hi if none).
__host__ __device__ inline uint32_t lower_bound_u32(const uint32_t* v, uint32_t lo, uint32_t hi,
                                                    uint32_t key) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (v[mid] < key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// First index in [lo, hi) whose key > key (hi if none).
__host__ __device__ inline uint32_t upper_bound_u32(const uint32_t* v, uint32_t lo, uint32_t hi,
                                                    uint32_t key) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (v[mid] <= key)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

}  // namespace hcg

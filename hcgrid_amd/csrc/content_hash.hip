// content_hash.hip -- 128-bit content hash of a device buffer.
//
// The served loop keeps sorted samples and ring lookup tables resident per
// distinct coordinate set and recognises repeat requests by the content of
// their coordinates.  Hashing on the device keeps the coordinates in HBM:
// only 16 bytes come back.
//
// h_k = len_mix + sum_i mix64(w_i ^ (i * A_k + B_k))  (mod 2^64), k = 0, 1,
// over the buffer's 8-byte words w_i (the tail zero-padded), with mix64 the
// splitmix64 finaliser (Steele, Lea and Flood 2014; Vigna's constants).
// Position enters every term, so reordered words change the hash; the sum
// is order-free, so the reduction can use atomics in any order and the
// result is the same on every run.
#include <hip/hip_runtime.h>

#include "common.hpp"

namespace hcg {
namespace {

__device__ __host__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

constexpr uint64_t kA0 = 0x9e3779b97f4a7c15ull, kB0 = 0x243f6a8885a308d3ull;
constexpr uint64_t kA1 = 0xc2b2ae3d27d4eb4full, kB1 = 0x13198a2e03707344ull;

__global__ void __launch_bounds__(64) k_hash_init(unsigned long long* out, uint64_t bytes) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = mix64(bytes ^ kB1);
    out[1] = mix64(bytes ^ kB0);
  }
}

__global__ void __launch_bounds__(256) k_hash_words(const uint64_t* __restrict__ words, uint64_t nwords,
                                                    const uint8_t* __restrict__ tail, uint32_t tail_bytes,
                                                    unsigned long long* __restrict__ out) {
  uint64_t h0 = 0, h1 = 0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += stride) {
    const uint64_t w = words[i];
    h0 += mix64(w ^ (i * kA0 + kB0));
    h1 += mix64(w ^ (i * kA1 + kB1));
  }
  if (tail_bytes && blockIdx.x == 0 && threadIdx.x == 0) {
    uint64_t w = 0;
    for (uint32_t b = 0; b < tail_bytes; ++b) w |= (uint64_t)tail[b] << (8 * b);
    h0 += mix64(w ^ (nwords * kA0 + kB0));
    h1 += mix64(w ^ (nwords * kA1 + kB1));
  }
  // wave sums, then one atomic per wave and word
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    h0 += (uint64_t)__shfl_xor((unsigned long long)h0, off, 64);
    h1 += (uint64_t)__shfl_xor((unsigned long long)h1, off, 64);
  }
  if (__lane_id() == 0) {
    atomicAdd(&out[0], (unsigned long long)h0);
    atomicAdd(&out[1], (unsigned long long)h1);
  }
}

}  // namespace
}  // namespace hcg

using namespace hcg;

extern "C" int hcg_content_hash(const void* data, size_t bytes, uint64_t* out, void* stream) {
  HCG_REQUIRE(out != nullptr, "hcg_content_hash: NULL out");
  HCG_REQUIRE(bytes == 0 || data != nullptr, "hcg_content_hash: NULL data");
  HCG_REQUIRE(((uintptr_t)data & 7u) == 0, "hcg_content_hash: data must be 8-byte aligned");
  HCG_REQUIRE(((uintptr_t)out & 7u) == 0, "hcg_content_hash: out must be 8-byte aligned");
  hipStream_t s = as_stream(stream);
  auto* o = reinterpret_cast<unsigned long long*>(out);
  k_hash_init<<<1, 64, 0, s>>>(o, (uint64_t)bytes);
  HCG_CHECK_LAUNCH("k_hash_init");
  const uint64_t nwords = bytes / 8;
  const uint32_t tail = (uint32_t)(bytes % 8);
  if (nwords == 0 && tail == 0) return HCG_OK;
  const uint32_t blocks = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(1, (nwords + 255) / 256), 2048);
  k_hash_words<<<blocks, 256, 0, s>>>(static_cast<const uint64_t*>(data), nwords,
                                      static_cast<const uint8_t*>(data) + nwords * 8, tail, o);
  HCG_CHECK_LAUNCH("k_hash_words");
  return HCG_OK;
}

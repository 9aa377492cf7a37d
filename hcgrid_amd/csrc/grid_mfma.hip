// grid_mfma.hip -- multi-channel gridding as a dense MFMA contraction.
//
// The reference grids a spectral cube one channel at a time, rerunning the
// full single-channel kernel per channel on its own stream
// (HCGrid+/gridding-omp.cu:552-678, kernel HCGrid+/gridding-omp.cu:294-420).
// The selection and weights do not depend on the channel, so here they are
// computed once per (cell, sample) pair in FP64 and the channel sums become
//     cube[ch][cell] = sum_s W[cell][s] * V[ch][s]   (FP32 accumulation)
// evaluated with v_mfma_f32_16x16x32_{bf16,f16}: W (cells x samples) in LDS,
// V (channels x sorted samples) streamed from HBM.  Each FP64 weight is
// carried as hi + lo, two values of the channel type (hi = round(w),
// lo = round(w - hi)), and both are contracted: the weight keeps ~2x the
// type's significand (relative error ~2^-17 for BF16) at twice the MFMA
// count, which the kernel's HBM/L2 stream of V hides.
//
// Work decomposition: a workgroup owns one map row x 64 columns (4 MFMA
// row blocks of 16 cells) and up to 8 x 64 channels (wave w: 4 blocks of
// 16 channels; 4 row blocks x 4 channel blocks x 4 = 64 FP32 accumulators
// per lane).  Wider channel ranges take more workgroups along grid.y.
// For every non-empty ring of the
// row segment's ring range:
//   1. each cell's [lpix, rpix] window (the reference's compat selection);
//   2. the union window's sorted range by a 64-way cooperative key search;
//   3. per chunk of <= CHUNK samples: stage unit vectors and keys in LDS,
//      zero the weight tiles, let every cell's threads fill its band with
//      the FP64-decided weights (as hi + lo; the sum of the rounded pairs is
//      the normalisation, so numerator and denominator use the same
//      weights), record each row block's k-block span;
//   4. MFMA over the union's 32-sample k-blocks: B (16 channels x 32
//      samples) is loaded once per k-block and applied to every row block
//      whose span covers it.
// Selection, FP64 distance and weight follow the single-channel kernel
// exactly; only the products with channel values run in BF16/FP16.
#include <hip/hip_runtime.h>

#include "gridcore.hpp"

namespace hcg {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kMcCols = 64;                   // cells per workgroup (one map row)
constexpr int kMcNB = 4;                      // 16-channel blocks per wave
constexpr int kMcMaxWaves = 8;
constexpr uint32_t kMcChanPerWave = 16 * kMcNB;

template <typename T>
struct Mfma;
template <>
struct Mfma<__bf16> {
  typedef bf16x8_t vec;
  static __device__ inline f32x4_t op(vec a, vec b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<_Float16> {
  typedef f16x8_t vec;
  static __device__ inline f32x4_t op(vec a, vec b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  }
};

// FP64 weight of a selected pair (the single-channel kernel's evaluation).
__device__ inline double pair_weight_any(const GridParams& P, const GridBuffers& B, uint32_t kt, double c2,
                                         const double* s_exp, uint32_t cell, uint32_t i) {
  if (kt == HCG_GAUSS1D) return exp_scaled_neg(gauss_exponent_scaled(P, c2), s_exp);
  if (kt == HCG_GAUSS2D)
    return weight_general<HCG_GAUSS2D>(P.params[0], P.params[1], P.params[2], B.xwcs, B.ywcs, B.lon_s, B.lat_s,
                                       0.25 * c2, cell, i);
  return weight_general<HCG_TAPERED_SINC>(P.params[0], P.params[1], P.params[2], B.xwcs, B.ywcs, B.lon_s,
                                          B.lat_s, 0.25 * c2, cell, i);
}

// CHUNK = staged samples per chunk.  LDS: two weight tiles of 64 x (CHUNK+8)
// elements (16-byte row padding) + CHUNK unit vectors and keys: 148 KB at
// 512 (one 8-wave workgroup per CU), 75 KB at 256, 38 KB at 128.
template <typename T, uint32_t CHUNK>
__global__ void __launch_bounds__(64 * kMcMaxWaves)
    k_grid_mfma(GridParams P, GridBuffers B, uint32_t kt, const T* __restrict__ vs, uint32_t n_pad, uint32_t nchan,
                uint32_t segs_x, uint32_t nsegs, float* __restrict__ cube, float* __restrict__ wsum_out,
                int normalize) {
  typedef typename Mfma<T>::vec vec;
  constexpr uint32_t WS = CHUNK + 8;  // weight-tile row stride (elements)
  __shared__ __attribute__((aligned(16))) T s_wh[kMcCols * WS];
  __shared__ __attribute__((aligned(16))) T s_wl[kMcCols * WS];
  __shared__ double s_x[CHUNK], s_y[CHUNK], s_z[CHUNK];
  __shared__ uint32_t s_key[CHUNK];
  __shared__ double s_exp[64];
  __shared__ uint32_t s_mb[4][2];
  __shared__ uint32_t s_red[kMcMaxWaves][2];
  __shared__ double s_wsum[kMcCols];

  const hcg_ring_meta m = *B.meta;
  const uint32_t nthreads = blockDim.x;
  const uint32_t nw = nthreads >> 6;       // waves in this workgroup
  const uint32_t S = nw;                   // threads per cell
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = __lane_id();
  const uint32_t wave = tid >> 6;
  if (wave == 0) exp_table_load(s_exp);

  // XCD-contiguous runs of row segments (block b runs on XCD b % 8)
  uint32_t seg = blockIdx.x;
  if ((nsegs & 7u) == 0) seg = (blockIdx.x & 7u) * (nsegs >> 3) + (blockIdx.x >> 3);
  const uint32_t row = seg / segs_x;
  const uint32_t x0 = (seg % segs_x) * kMcCols;
  const uint32_t M = P.side * P.side;

  // this thread's cell: c = tid / S, sub = tid % S
  const uint32_t c = tid / S, sub = tid % S;
  const uint32_t col = x0 + c;
  const bool live = col < P.side && row < P.side;
  const uint32_t cell = live ? row * P.side + col : 0;
  const uint32_t anchor = live ? row * P.side + (col / P.factor) * P.factor : 0;
  Window win = window_of(P, m.firstring, m.usedrings, B.xwcs[anchor], B.ywcs[anchor]);
  if (!live || m.usedrings == 0) {
    win.uring = 1;
    win.dring = 0;
  }
  double sb, cb0, sl, cl;
  sincos(B.ywcs[cell] * kDeg2Rad, &sb, &cb0);
  sincos(B.xwcs[cell] * kDeg2Rad, &sl, &cl);
  const double px = cb0 * cl, py = cb0 * sl, pz = sb;
  double wsum = 0.0;

  // accumulators: 4 row blocks x kMcNB channel blocks of 16x16
  f32x4_t acc[4][kMcNB];
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int nb = 0; nb < kMcNB; ++nb) acc[mb][nb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  const uint32_t ch_base = (blockIdx.y * nw + wave) * kMcChanPerWave;

  // ring span of the row segment
  {
    const bool has = win.uring <= win.dring;
    const uint32_t ur = wave_min_u32(has ? win.uring : 0xFFFFFFFFu);
    const uint32_t dr = wave_max_u32(has ? win.dring : 0u);
    if (lane == 0) {
      s_red[wave][0] = ur;
      s_red[wave][1] = dr;
    }
  }
  __syncthreads();
  uint32_t r_begin = 0xFFFFFFFFu, r_end = 0;
  for (uint32_t w = 0; w < nw; ++w) {
    r_begin = min(r_begin, s_red[w][0]);
    r_end = max(r_end, s_red[w][1]);
  }
  if (r_begin == m.firstring) ++r_begin;  // the first used ring contributes nothing

  for (uint32_t r = r_begin; r <= r_end && r_begin <= r_end; ++r) {
    const uint32_t kr = r - m.firstring;
    const uint32_t start = B.lut[kr], end = B.lut[kr + 1];
    if (start == end) continue;  // uniform
    uint64_t a64 = 1, b64 = 0;
    const bool act =
        r >= win.uring && r <= win.dring && ring_window(P, win, r, B.ring_z[kr], B.ring_s[kr], a64, b64) && a64 <= b64;
    const uint32_t lp = act ? (uint32_t)a64 : 1u, rp = act ? (uint32_t)b64 : 0u;
    __syncthreads();  // s_red reads of the previous round are done
    {
      const uint32_t wl = wave_min_u32(act ? lp : 0xFFFFFFFFu);
      const uint32_t wr = wave_max_u32(act ? rp : 0u);
      if (lane == 0) {
        s_red[wave][0] = wl;
        s_red[wave][1] = wr;
      }
    }
    __syncthreads();
    uint32_t LP = 0xFFFFFFFFu, RP = 0;
    for (uint32_t w = 0; w < nw; ++w) {
      LP = min(LP, s_red[w][0]);
      RP = max(RP, s_red[w][1]);
    }
    if (LP > RP) continue;  // uniform
    const uint32_t lo_u = wave_search<false>(B.keys, start, end, LP);
    const uint32_t hi_u = wave_search<true>(B.keys, lo_u, end, RP);
    for (uint32_t c0 = lo_u & ~7u; c0 < hi_u; c0 += CHUNK) {
      const uint32_t cnt = min(CHUNK, hi_u - c0);
      const uint32_t nkb = (cnt + 31) / 32;
      __syncthreads();  // the previous chunk's weight tile is consumed
      // stage the chunk; zero the weight tile's used width
      for (uint32_t j = tid; j < cnt; j += nthreads) {
        const hcg_sample q = B.rec[c0 + j];
        s_x[j] = q.x;
        s_y[j] = q.y;
        s_z[j] = q.z;
        s_key[j] = B.keys[c0 + j];
      }
      {
        const uint32_t width8 = nkb * 4;  // 16-byte groups per row
        for (uint32_t g = tid; g < kMcCols * width8; g += nthreads) {
          const uint32_t rr = g / width8, cc = (g % width8) * 8;
          *reinterpret_cast<uint4*>(&s_wh[rr * WS + cc]) = uint4{0u, 0u, 0u, 0u};
          *reinterpret_cast<uint4*>(&s_wl[rr * WS + cc]) = uint4{0u, 0u, 0u, 0u};
        }
      }
      if (tid < 4) {
        s_mb[tid][0] = 0xFFFFFFFFu;
        s_mb[tid][1] = 0u;
      }
      __syncthreads();
      // weights of this thread's cell over its band in the chunk
      if (act) {
        const uint32_t ja = lower_bound_u32(s_key, 0, cnt, lp);
        const uint32_t jb = upper_bound_u32(s_key, ja, cnt, rp);
        if (ja < jb && sub == 0) {
          atomicMin(&s_mb[c >> 4][0], ja >> 5);
          atomicMax(&s_mb[c >> 4][1], (jb + 31) >> 5);
        }
        for (uint32_t j = ja + sub; j < jb; j += S) {
          const double dx = px - s_x[j], dy = py - s_y[j], dz = pz - s_z[j];
          const double c2 = fma(dx, dx, fma(dy, dy, dz * dz));
          if (c2 < P.c_hi &&
              (c2 < P.c_lo || radius_test_ref(B.xwcs, B.ywcs, B.lon_s, B.lat_s, cell, c0 + j, P.sphere_radius))) {
            double wt = pair_weight_any(P, B, kt, c2, s_exp, cell, c0 + j);
            if (B.w_s) wt *= B.w_s[c0 + j];
            // the weight as hi + lo in the value type: ~2x its significand
            const T hi = (T)(float)wt;
            const T lo = (T)(float)(wt - (double)(float)hi);
            s_wh[c * WS + j] = hi;
            s_wl[c * WS + j] = lo;
            wsum += (double)(float)hi + (double)(float)lo;
          }
        }
      }
      __syncthreads();
      // contraction over the chunk's k-blocks
      uint32_t mblo[4], mbhi[4];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        mblo[mb] = s_mb[mb][0];
        mbhi[mb] = s_mb[mb][1];
      }
      if (ch_base < nchan) {
        const uint32_t koff = (lane >> 4) * 8;
        for (uint32_t u = 0; u < nkb; ++u) {
          bool any = false;
#pragma unroll
          for (int mb = 0; mb < 4; ++mb) any |= (u >= mblo[mb] && u < mbhi[mb]);
          if (!any) continue;  // uniform
          // B: 16 channels x 32 samples per channel block, from HBM/L2
          const T* vrow = vs + c0 + u * 32 + koff;
          vec bv[kMcNB];
#pragma unroll
          for (int nb = 0; nb < kMcNB; ++nb) {
            uint32_t ch = ch_base + nb * 16 + (lane & 15);
            ch = ch < nchan ? ch : nchan - 1;
            bv[nb] = *reinterpret_cast<const vec*>(vrow + (size_t)ch * n_pad);
          }
          // A: the row block's hi and lo weights, from LDS
#pragma unroll
          for (int mb = 0; mb < 4; ++mb) {
            if (!(u >= mblo[mb] && u < mbhi[mb])) continue;
            const uint32_t off = (mb * 16 + (lane & 15)) * WS + u * 32 + koff;
            const vec ah = *reinterpret_cast<const vec*>(&s_wh[off]);
            const vec al = *reinterpret_cast<const vec*>(&s_wl[off]);
#pragma unroll
            for (int nb = 0; nb < kMcNB; ++nb) {
              acc[mb][nb] = Mfma<T>::op(ah, bv[nb], acc[mb][nb]);
              acc[mb][nb] = Mfma<T>::op(al, bv[nb], acc[mb][nb]);
            }
          }
        }
      }
    }
  }

  // per-cell weight sums: the S threads of a cell are consecutive lanes
  for (uint32_t off = 1; off < S; off <<= 1) wsum += __shfl_xor(wsum, off, 64);
  if (sub == 0) s_wsum[c] = wsum;
  __syncthreads();
  if (blockIdx.y == 0 && wsum_out && sub == 0 && live) wsum_out[cell] = (float)wsum;
  if (ch_base < nchan && row < P.side) {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
#pragma unroll
      for (int nb = 0; nb < kMcNB; ++nb) {
        const uint32_t ch = ch_base + nb * 16 + (lane & 15);
        if (ch >= nchan) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const uint32_t cl = mb * 16 + (lane >> 4) * 4 + j;
          const uint32_t colj = x0 + cl;
          if (colj >= P.side) continue;
          const double v = (double)acc[mb][nb][j];
          const float o = normalize ? (float)(v / s_wsum[cl]) : (float)v;
          cube[(size_t)ch * M + (size_t)row * P.side + colj] = o;
        }
      }
    }
  }
}

// Channel gather: sorted, channel-major, zero-padded to n_pad.
template <typename T>
__global__ void __launch_bounds__(256) k_gather_channels_cm(uint32_t n, uint32_t n_pad, const uint32_t* __restrict__ perm,
                                                            const T* __restrict__ in, T* __restrict__ out,
                                                            uint32_t* __restrict__ nonfinite) {
  const uint32_t ch = blockIdx.y;
  const uint32_t i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n_pad) return;
  T v = (T)0.0f;
  if (i < n) {
    uint32_t j = perm[i];
    if (j >= n) j = i;
    v = in[(size_t)ch * n + j];
    if (nonfinite && !isfinite((float)v)) atomicAdd(nonfinite, 1u);
  }
  out[(size_t)ch * n_pad + i] = v;
}

// Sample-major input ([n][nchan]): 64 x 64 tiles transposed through LDS.
template <typename T>
__global__ void __launch_bounds__(256) k_gather_channels_sm(uint32_t n, uint32_t n_pad, uint32_t nchan,
                                                            const uint32_t* __restrict__ perm,
                                                            const T* __restrict__ in, T* __restrict__ out,
                                                            uint32_t* __restrict__ nonfinite) {
  __shared__ T tile[64][65];
  const uint32_t i0 = blockIdx.x * 64, ch0 = blockIdx.y * 64;
  const uint32_t tx = threadIdx.x & 63, ty = threadIdx.x >> 6;  // 64 x 4
  for (uint32_t k = ty; k < 64; k += 4) {
    const uint32_t i = i0 + k, ch = ch0 + tx;
    T v = (T)0.0f;
    if (i < n && ch < nchan) {
      uint32_t j = perm[i];
      if (j >= n) j = i;
      v = in[(size_t)j * nchan + ch];
      if (nonfinite && !isfinite((float)v)) atomicAdd(nonfinite, 1u);
    }
    tile[k][tx] = v;
  }
  __syncthreads();
  for (uint32_t k = ty; k < 64; k += 4) {
    const uint32_t ch = ch0 + k, i = i0 + tx;
    if (ch < nchan && i < n_pad) out[(size_t)ch * n_pad + i] = tile[tx][k];
  }
}

}  // namespace hcg

using namespace hcg;

extern "C" uint32_t hcg_channels_padded_len(uint32_t n) { return ((n + 7u) & ~7u) + 64u; }

extern "C" int hcg_gather_channels(uint32_t n, uint32_t nchan, const uint32_t* perm, const void* values, int dtype,
                                   int layout, void* values_sorted, uint32_t* nonfinite, void* stream) {
  if (n == 0 || nchan == 0) return HCG_OK;
  HCG_REQUIRE(perm && values && values_sorted, "hcg_gather_channels: NULL buffer");
  HCG_REQUIRE(dtype == HCG_DTYPE_BF16 || dtype == HCG_DTYPE_F16, "hcg_gather_channels: dtype must be BF16 or F16");
  HCG_REQUIRE(layout == 0 || layout == 1, "hcg_gather_channels: layout must be 0 (channel-major) or 1 (sample-major)");
  const uint32_t n_pad = hcg_channels_padded_len(n);
  hipStream_t s = as_stream(stream);
  if (layout == 0) {
    const dim3 grid(ceil_div_u32(n_pad, 256), nchan);
    if (dtype == HCG_DTYPE_BF16)
      k_gather_channels_cm<__bf16><<<grid, 256, 0, s>>>(n, n_pad, perm, static_cast<const __bf16*>(values),
                                                        static_cast<__bf16*>(values_sorted), nonfinite);
    else
      k_gather_channels_cm<_Float16><<<grid, 256, 0, s>>>(n, n_pad, perm, static_cast<const _Float16*>(values),
                                                          static_cast<_Float16*>(values_sorted), nonfinite);
  } else {
    const dim3 grid(ceil_div_u32(n_pad, 64), ceil_div_u32(nchan, 64));
    if (dtype == HCG_DTYPE_BF16)
      k_gather_channels_sm<__bf16><<<grid, 256, 0, s>>>(n, n_pad, nchan, perm, static_cast<const __bf16*>(values),
                                                        static_cast<__bf16*>(values_sorted), nonfinite);
    else
      k_gather_channels_sm<_Float16><<<grid, 256, 0, s>>>(n, n_pad, nchan, perm,
                                                          static_cast<const _Float16*>(values),
                                                          static_cast<_Float16*>(values_sorted), nonfinite);
  }
  HCG_CHECK_LAUNCH("k_gather_channels");
  return HCG_OK;
}

extern "C" int hcg_grid_mfma(const hcg_healpix* hp, const hcg_kernel* kern, uint32_t side, uint32_t n, uint32_t nchan,
                             int dtype, const hcg_sample* records, const uint32_t* keys_sorted,
                             const double* lon_sorted, const double* lat_sorted, const double* weights_sorted,
                             const void* values_sorted, const uint32_t* lut, const double* ring_z,
                             const double* ring_s, const hcg_ring_meta* meta, const double* xwcs, const double* ywcs,
                             float* cube, float* wsum, int normalize, void* stream) {
  HCG_REQUIRE(hp && kern, "hcg_grid_mfma: NULL hp/kern");
  if (hp->nside == 0 || hp->nside > HCG_MAX_NSIDE)
    return set_error(HCG_ERR_UNSUPPORTED, "hcg_grid_mfma: nside must be <= 8192");
  HCG_REQUIRE(kern->kernel_type >= HCG_GAUSS1D && kern->kernel_type <= HCG_TAPERED_SINC,
              "hcg_grid_mfma: unknown kernel type");
  HCG_REQUIRE(kern->factor >= 1, "hcg_grid_mfma: factor must be >= 1");
  HCG_REQUIRE(dtype == HCG_DTYPE_BF16 || dtype == HCG_DTYPE_F16, "hcg_grid_mfma: dtype must be BF16 or F16");
  HCG_REQUIRE(side <= 65535, "hcg_grid_mfma: side must be <= 65535");
  if (side == 0 || nchan == 0) return HCG_OK;
  HCG_REQUIRE(records && keys_sorted && lon_sorted && lat_sorted && values_sorted && lut && ring_z && ring_s && meta &&
                  xwcs && ywcs && cube,
              "hcg_grid_mfma: NULL buffer");
  (void)n;
  GridParams P = make_grid_params(hp, kern, side);
  GridBuffers B{records, keys_sorted, lon_sorted, lat_sorted, weights_sorted, lut, ring_z, ring_s, meta,
                xwcs,    ywcs,        nullptr,    nullptr};
  const uint32_t n_pad = hcg_channels_padded_len(n);
  const uint32_t groups_total = ceil_div_u32(nchan, kMcChanPerWave);  // waves needed
  // waves per workgroup: a power of two (each cell's nw threads are aligned
  // consecutive lanes, reduced with xor shuffles); surplus waves only help
  // with the weights
  uint32_t nw = 1;
  while (nw < groups_total && nw < (uint32_t)kMcMaxWaves) nw <<= 1;
  const uint32_t ngroups = ceil_div_u32(groups_total, nw);
  const uint32_t segs_x = ceil_div_u32(side, kMcCols);
  const uint32_t nsegs = segs_x * side;
  const dim3 grid(nsegs, ngroups);
  hipStream_t s = as_stream(stream);
  // chunk size by workgroup width: 8 waves fill a CU alone (VGPR-limited);
  // narrower workgroups take smaller tiles so several share a CU
  const uint32_t chunk = nw >= 8 ? 512u : (nw >= 4 ? 256u : 128u);
#define HCG_LAUNCH_MFMA(TT, CH)                                                                               \
  k_grid_mfma<TT, CH><<<grid, 64 * nw, 0, s>>>(P, B, kern->kernel_type, static_cast<const TT*>(values_sorted), \
                                               n_pad, nchan, segs_x, nsegs, cube, wsum, normalize)
  if (dtype == HCG_DTYPE_BF16) {
    if (chunk == 512) HCG_LAUNCH_MFMA(__bf16, 512);
    else if (chunk == 256) HCG_LAUNCH_MFMA(__bf16, 256);
    else HCG_LAUNCH_MFMA(__bf16, 128);
  } else {
    if (chunk == 512) HCG_LAUNCH_MFMA(_Float16, 512);
    else if (chunk == 256) HCG_LAUNCH_MFMA(_Float16, 256);
    else HCG_LAUNCH_MFMA(_Float16, 128);
  }
#undef HCG_LAUNCH_MFMA
  HCG_CHECK_LAUNCH("k_grid_mfma");
  return HCG_OK;
}

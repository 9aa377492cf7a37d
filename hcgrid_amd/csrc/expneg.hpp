// expneg.hpp -- exp(x) for x <= 0, the Gaussian weights of the gridding kernels.
//
// Table-driven method of P. T. P. Tang, "Table-driven implementation of the
// exponential function in IEEE floating-point arithmetic", ACM TOMS 15(2),
// 1989: x = (64 m + j) ln2/64 + r with |r| <= ln2/128, so
// exp(x) = 2^m * 2^(j/64) * exp(r); 2^(j/64) comes from a 64-entry table
// (correctly rounded doubles) and exp(r) - 1 from its degree-6 Taylor
// polynomial (truncation < 3e-20 on |r| <= 0.0055).  ln2/64 is split so that
// n * hi is exact for |n| < 2^21.  Error < 1 ulp; arguments below -1000 give 0.
// Compared with the device library's exp this needs 4 non-inline constants
// instead of 11, which keeps them out of the inner loop's vector registers.
#pragma once

#include <hip/hip_runtime.h>

namespace hcg {

constexpr double kExpInvLn2_64 = 0x1.71547652b82fep+6;   // 64 / ln2
constexpr double kExpLn2_64Hi = 0x1.62e42fee00000p-7;    // ln2 / 64, leading 32 bits
constexpr double kExpLn2_64Lo = 0x1.a39ef35793c76p-39;   // ln2 / 64 - hi

// 2^(j/64), j = 0..63, correctly rounded.
__device__ __constant__ const double kExp2Table[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0};

// Copies the table into a 64-entry LDS array (one element per lane of the wave).
__device__ inline void exp_table_load(double* lds_table) { lds_table[__lane_id()] = kExp2Table[__lane_id()]; }

// exp(y * ln2/64) for y <= 0 given pre-scaled: n = rint(y), y - n is exact,
// r = (y - n) ln2/64 with |r| <= ln2/128, degree-5 Taylor polynomial
// (truncation r^6/720 < 4e-17).  Callers fold 64/ln2 into their argument.
__device__ inline double exp_scaled_neg(double y, const double* lds_table) {
  y = y < -92000.0 ? -92000.0 : y;  // exp underflows to 0 well before
  const double n = rint(y);
  const int ni = (int)n;
  const int j = ni & 63;
  const int m = ni >> 6;  // floor(ni / 64)
  const double r = (y - n) * 0x1.62e42fefa39efp-7;  // ln2 / 64
  double q = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  q = fma(r, q, 1.0 / 6.0);
  q = fma(r, q, 0.5);
  q = fma(r, q, 1.0);
  const double p = r * q;  // exp(r) - 1
  const double t = lds_table[j];
  return ldexp(fma(t, p, t), m);
}

// exp_scaled_neg for arguments known to be >= -92000 (no clamp): the same
// operations, so the same results.
__device__ inline double exp_scaled_neg_nc(double y, const double* lds_table) {
  const double n = rint(y);
  const int ni = (int)n;
  const int j = ni & 63;
  const int m = ni >> 6;
  const double r = (y - n) * 0x1.62e42fefa39efp-7;
  double q = fma(r, 1.0 / 120.0, 1.0 / 24.0);
  q = fma(r, q, 1.0 / 6.0);
  q = fma(r, q, 0.5);
  q = fma(r, q, 1.0);
  const double p = r * q;
  const double t = lds_table[j];
  return ldexp(fma(t, p, t), m);
}

// a * b + c with c in a scalar register pair: one VOP3 v_fma_f64.  Written
// out because the compiler otherwise keeps loop-invariant addends in vector
// registers and emits the two-address v_fmac form plus a copy per call.
__device__ inline double fma_vvs(double a, double b, double c) {
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
}

// exp_scaled_neg_nc with ln2/64 folded into the polynomial: p = u (c1 + u (c2
// + u (c3 + u (c4 + u c5)))), c_k = (ln2/64)^k / k!, u = y - rint(y), the same
// degree-5 Taylor polynomial of exp(u ln2/64) - 1 without forming r, each
// step one v_fma_f64 with its coefficient in scalar registers (the leading
// pair as mul + add: two scalar operands cannot share one instruction).
// For arguments >= -92000 (no clamp).
__device__ inline double exp_scaled_neg_fold(double y, const double* lds_table) {
  constexpr double c1 = 0x1.62e42fefa39efp-7;
  constexpr double c2 = c1 * c1 / 2.0;
  constexpr double c3 = c2 * c1 / 3.0;
  constexpr double c4 = c3 * c1 / 4.0;
  constexpr double c5 = c4 * c1 / 5.0;
  const double n = rint(y);
  const int ni = (int)n;
  const int j = ni & 63;
  const int m = ni >> 6;
  const double u = y - n;
  double q = u * c5 + c4;
  q = fma_vvs(u, q, c3);
  q = fma_vvs(u, q, c2);
  q = fma_vvs(u, q, c1);
  const double p = u * q;
  const double t = lds_table[j];
  return ldexp(fma(t, p, t), m);
}

__device__ inline double exp_neg(double x, const double* lds_table) {
  x = x < -1000.0 ? -1000.0 : x;
  const double n = rint(x * kExpInvLn2_64);
  const int ni = (int)n;
  const int j = ni & 63;
  const int m = ni >> 6;  // floor(ni / 64)
  const double r = fma(-n, kExpLn2_64Lo, fma(-n, kExpLn2_64Hi, x));
  double q = fma(r, 1.0 / 720.0, 1.0 / 120.0);
  q = fma(r, q, 1.0 / 24.0);
  q = fma(r, q, 1.0 / 6.0);
  q = fma(r, q, 0.5);
  q = fma(r, q, 1.0);
  const double p = r * q;  // exp(r) - 1
  const double t = lds_table[j];
  return ldexp(fma(t, p, t), m);
}

}  // namespace hcg

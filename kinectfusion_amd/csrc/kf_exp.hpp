// exp(x) in FP64 for x <= 0 by Tang's table-driven method (P. T. P. Tang,
// "Table-driven implementation of the exponential function in IEEE
// floating-point arithmetic", ACM TOMS 15(2):144-157, 1989):
//
//   x = (64 m + j) ln2/64 + r,  |r| <= ln2/128,  0 <= j < 64
//   exp(x) = 2^m * 2^(j/64) * (1 + p(r)),  p(r) = r + r^2/2 + ... + r^5/120
//
// ln2/64 is split into a head with 16 trailing zero bits (n * head is exact
// for |n| < 2^16) and a tail; the degree-5 Taylor polynomial truncates below
// 4e-17 relative for |r| <= ln2/128; the table entries are 2^(j/64) rounded
// to nearest.  The relative error is below 4e-16, so the float the result
// rounds to is the correctly rounded float exponential except when exp(x)
// lies within that distance of a float rounding boundary.  Arguments below
// -110 return 0 (exp(-110) is far below half the smallest float denormal).
//
// Plain C++ with an optional __host__ __device__ qualifier, so the same code
// is compiled for the device and, in the test suite, for the host.
#pragma once

#ifndef KF_EXP_HD
#if defined(__HIPCC__)
#define KF_EXP_HD __host__ __device__
#else
#define KF_EXP_HD
#endif
#endif

#include <cmath>

namespace kf {

// 2^(j/64), j = 0 .. 63, rounded to nearest double.
#define KF_EXP2_J64_TABLE                                                                         \
  {0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,       \
   0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,       \
   0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,       \
   0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,       \
   0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,       \
   0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,       \
   0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,       \
   0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,       \
   0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,       \
   0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,       \
   0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,       \
   0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,       \
   0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,       \
   0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,       \
   0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,       \
   0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0}

// exp(x) for x <= 0, with `table` the 64 entries of KF_EXP2_J64_TABLE.
KF_EXP_HD inline double exp_nonpositive(double x, const double* table) {
  if (!(x >= -110.0)) return 0.0;
  const double inv_l = 0x1.71547652b82fep+6;      // 64 / ln2
  const double l_head = 0x1.62e42fefa0000p-7;     // ln2 / 64, head
  const double l_tail = 0x1.cf79abc9e3b3ap-46;    // ln2 / 64 - head
  const double n = std::rint(x * inv_l);
  const int ni = static_cast<int>(n);
  const int j = ni & 63;
  const int m = (ni - j) / 64;
  const double r = (x - n * l_head) - n * l_tail;
  const double p = r + (r * r) * (0.5 + r * (1.0 / 6.0 + r * (1.0 / 24.0 + r * (1.0 / 120.0))));
  const double t = table[j];
  return std::ldexp(t + t * p, m);
}

}  // namespace kf

// exp(x) for the Gaussian filter's weights, x = r^2 * (-1 / (2 sigma^2)) in
// [ln(alpha), 0].  Tang's table-driven method (P. T. P. Tang, "Table-driven
// implementation of the exponential function in IEEE floating-point
// arithmetic", ACM TOMS 15(2), 1989) with 32 table entries:
//
//   x = (32 m + j) ln2/32 + r,  |r| <= ln2/64,  0 <= j < 32
//   exp(x) = 2^m * 2^(j/32) * (1 + q),  q = exp(r) - 1
//
// q is the Taylor polynomial r + r^2/2 + ... + r^6/720 (the first dropped
// term is below 4e-18 relative), r comes from a two-part ln2/32 with fused
// multiply-adds (Cody and Waite).  Valid for -700 <= x <= 0, where 2^m stays
// normal; the result is within about one unit in the last place of exp(x).
// The same operations run on the host and the device (fma, rint and ldexp
// are exact or correctly rounded on both), so a host test checks the
// device's bits.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define GX_HD __host__ __device__
#else
#define GX_HD
#endif

namespace gx {

// 2^(j/32), j = 0..31, each correctly rounded
constexpr double kTab[32] = {
    0x1.0000000000000p+0, 0x1.059b0d3158574p+0, 0x1.0b5586cf9890fp+0, 0x1.11301d0125b51p+0,
    0x1.172b83c7d517bp+0, 0x1.1d4873168b9aap+0, 0x1.2387a6e756238p+0, 0x1.29e9df51fdee1p+0,
    0x1.306fe0a31b715p+0, 0x1.371a7373aa9cbp+0, 0x1.3dea64c123422p+0, 0x1.44e086061892dp+0,
    0x1.4bfdad5362a27p+0, 0x1.5342b569d4f82p+0, 0x1.5ab07dd485429p+0, 0x1.6247eb03a5585p+0,
    0x1.6a09e667f3bcdp+0, 0x1.71f75e8ec5f74p+0, 0x1.7a11473eb0187p+0, 0x1.82589994cce13p+0,
    0x1.8ace5422aa0dbp+0, 0x1.93737b0cdc5e5p+0, 0x1.9c49182a3f090p+0, 0x1.a5503b23e255dp+0,
    0x1.ae89f995ad3adp+0, 0x1.b7f76f2fb5e47p+0, 0x1.c199bdd85529cp+0, 0x1.cb720dcef9069p+0,
    0x1.d5818dcfba487p+0, 0x1.dfc97337b9b5fp+0, 0x1.ea4afa2a490dap+0, 0x1.f50765b6e4540p+0,
};
constexpr double kInvL = 0x1.71547652b82fep+5;  // 32 / ln2
constexpr double kL1 = 0x1.62e42fefa3800p-6;    // ln2 / 32, leading 42 bits
constexpr double kL2 = 0x1.ef35793c76730p-50;   // ln2 / 32 - kL1

// T: the table (a copy in fast memory on the device)
GX_HD inline double exp_tab(double x, const double *T) {
  double kd = rint(x * kInvL);
  double r = fma(-kd, kL2, fma(-kd, kL1, x));
  int k = (int)kd;
  int j = k & 31, m = k >> 5;  // arithmetic shift: k = 32 m + j
  double p = fma(r, 1.0 / 720.0, 1.0 / 120.0);
  p = fma(r, p, 1.0 / 24.0);
  p = fma(r, p, 1.0 / 6.0);
  p = fma(r, p, 0.5);
  p = fma(r, p, 1.0);
  double q = r * p;
  double t = T[j];
  return ldexp(fma(t, q, t), m);
}

}  // namespace gx

// cert.hpp -- running error bounds of the certified fp64 tier.
//
// Model. Every value-path operation is IEEE binary64 round-to-nearest (the
// library is built with -ffp-contract=off and spells out each fma), so a freshly
// rounded complex result z satisfies |fl(x) - x| <= u ||fl(x)||_1 + UFL per
// operation, u = 2^-53, where ||z||_1 = |Re z| + |Im z| >= |z| and the absolute
// UFL covers results in the subnormal range. Bounds are kept in the 1-norm, which
// is sub-multiplicative for complex numbers.
//
// Product rule: for computed operands a^, b^ with 1-norm errors ea, eb and
//   fl(a^ b^) = (fma(ar, br, -fl(ai bi)), fma(ar, bi, fl(ai br))),
// |a^ b^ - a b| <= ||a^|| eb + ||b^|| ea + ea eb and the rounding of the product
// adds at most (2u + u^2) ||a^|| ||b^|| <= 3u ||a^|| ||b^|| (+ UFL).
//
// Bound arithmetic. The bound itself is evaluated in round-to-nearest. It is an
// expression of additions, multiplications and fmas of non-negative quantities,
// so its computed value is at least (1 - u)^D times its exact value, D being the
// largest number of roundings on any path of the expression. cert_finish()
// multiplies by 1 + 2 D u >= (1 - u)^-D (valid for D u <= 1/2) and rounds the
// result up by one ulp twice (once per operation of the inflation itself).
// Underflow inside the bound arithmetic is covered by the absolute D * 2^-1074.
#pragma once

#include "common.hpp"

namespace gbsk {

constexpr double CERT_U = 1.1102230246251565e-16;    // 2^-53
constexpr double CERT_UFL = 1.9762625833649862e-323;  // 4 * 2^-1074
constexpr double CERT_CMUL = 3.0 * CERT_U;            // exact (3 * 2^-53)

__host__ __device__ inline double cert_norm1(cplx z) { return fabs(z.re) + fabs(z.im); }

// next double above x (x >= 0 finite); +inf and NaN pass through
__host__ __device__ inline double cert_bump_up(double x) {
  if (!(x < 1.79769313486231570e308)) return x;
  return __longlong_as_double(__double_as_longlong(x) + 1);
}

// e + rounding error of the freshly rounded complex value z
__device__ inline double cert_round(cplx z, double e) { return fma(CERT_U, cert_norm1(z), e + CERT_UFL); }

// bound of |fl(a^ b^) - a b| from |a^ - a| <= ea, |b^ - b| <= eb (1-norm)
__device__ inline double cert_mul(cplx a, double ea, cplx b, double eb) {
  const double A = cert_norm1(a), Bn = cert_norm1(b);
  double t = fma(A, eb, CERT_UFL);
  t = fma(Bn, ea, t);
  t = fma(ea, eb, t);
  return fma(CERT_CMUL * A, Bn, t);
}

// final rigorous bound from a round-to-nearest evaluation of rounding depth <= D
__device__ inline double cert_finish(double e, double D) {
  const double infl = fma(e, 1.0 + 2.0 * D * CERT_U, D * 4.9406564584124654e-324);
  return cert_bump_up(cert_bump_up(infl));
}

// butterfly sum of (value, bound) pairs over G lanes; both partners of each
// level compute the same bits
template <int G>
__device__ inline void cert_group_sum(cplx& v, double& e) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) {
    const double vr = shfl_xor_d(v.re, m), vi = shfl_xor_d(v.im, m), eo = shfl_xor_d(e, m);
    v.re += vr;
    v.im += vi;
    e = cert_round(v, e + eo);
  }
}

}  // namespace gbsk

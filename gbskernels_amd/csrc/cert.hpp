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

// ---------------------------------------------------------------------------
// Certified double-double. Each operation returns the value of the plain dd
// routine in common.hpp (same operations, same order, so certified-dd values
// are bit-identical to the dd tier's) and the error of that one operation on
// exact operands, bounded a posteriori from its own intermediates:
//   every rounded binary64 operation x (+,-,*,fma) errs by <= u|fl(x)| + UFL;
//   TwoSum / TwoProduct are exact (TwoProduct up to UFL);
//   a FastTwoSum(x, y) -> r is exact unless its precondition fails, and its
//   error is then |TwoSum(x, y).lo - r.lo| exactly (both share the high part).
// Division and square root are certified from dd residuals (|q - a/b| =
// |a - q b| / |b|, |y - sqrt(a)| <= |y^2 - a| / y), divided by a guaranteed
// lower bound of the dd magnitude: fl(fl(|hi| - |lo|) (1 - 2u)) <= |hi + lo|.

namespace gbsk {

__device__ inline double cdd_fts_err(double x, double y, dd r) {
  const dd t = two_sum(x, y);  // t.hi == r.hi: both round x + y
  return (1.0 + 2.0 * CERT_U) * fabs(t.lo - r.lo);
}
__device__ inline double cdd_mag(dd a) { return fabs(a.hi) + fabs(a.lo); }
// guaranteed lower bound of |a.hi + a.lo| (0 if not provably positive)
__device__ inline double cdd_low(dd a) {
  const double d = (fabs(a.hi) - fabs(a.lo)) * (1.0 - 2.0 * CERT_U);
  return d > 1e-290 ? d : 0.0;
}

// r = dd_add(a, b); *e = |r - (a + b)|
__device__ inline dd cdd_add(dd a, dd b, double* e) {
  const dd s = two_sum(a.hi, b.hi);
  const double t = a.lo + b.lo;
  const double w = s.lo + t;
  const dd r = fast_two_sum(s.hi, w);
  *e = fma(CERT_U, fabs(t) + fabs(w), cdd_fts_err(s.hi, w, r));
  return r;
}
// r = dd_mul(a, b); *e = |r - a b|
__device__ inline dd cdd_mul(dd a, dd b, double* e) {
  const dd p = two_prod(a.hi, b.hi);
  const double m = a.lo * b.hi;
  const double q = fma(a.hi, b.lo, m);
  const double w = p.lo + q;
  const dd r = fast_two_sum(p.hi, w);
  const double dropped = (1.0 + 2.0 * CERT_U) * fabs(a.lo * b.lo);  // a.lo b.lo is not computed
  *e = fma(CERT_U, fabs(m) + fabs(q) + fabs(w), dropped + cdd_fts_err(p.hi, w, r) + 4.0 * CERT_UFL);
  return r;
}

}  // namespace gbsk

namespace gbsk {

// r = dd_add_d(a, b); *e = |r - (a + b)|
__device__ inline dd cdd_add_d(dd a, double b, double* e) {
  const dd s = two_sum(a.hi, b);
  const double w = s.lo + a.lo;
  const dd r = fast_two_sum(s.hi, w);
  *e = fma(CERT_U, fabs(w), cdd_fts_err(s.hi, w, r));
  return r;
}
__device__ inline double cddc_norm1(ddc a) { return cdd_mag(a.re) + cdd_mag(a.im); }
// complex dd (1-norm errors). Values are exactly ddc_add / ddc_sub / ddc_mul /
// ddc_add_scaled of common.hpp; *e receives the operation's own error.
__device__ inline ddc cddc_add(ddc a, ddc b, double* e) {
  double er, ei;
  const ddc r{cdd_add(a.re, b.re, &er), cdd_add(a.im, b.im, &ei)};
  *e = er + ei;
  return r;
}
__device__ inline ddc cddc_sub(ddc a, ddc b, double* e) { return cddc_add(a, ddc_neg(b), e); }
__device__ inline ddc cddc_add_scaled(ddc a, double s, cplx z, double* e) {
  double er, ei;
  const ddc r{cdd_add_d(a.re, s * z.re, &er), cdd_add_d(a.im, s * z.im, &ei)};
  *e = er + ei;
  return r;
}
__device__ inline ddc cddc_mul(ddc a, ddc b, double* e) {
  double e1, e2, e3, e4, e5, e6;
  const dd rr = cdd_mul(a.re, b.re, &e1), ii = cdd_mul(a.im, b.im, &e2);
  const dd ri = cdd_mul(a.re, b.im, &e3), ir = cdd_mul(a.im, b.re, &e4);
  const ddc r{cdd_add(rr, dd_neg(ii), &e5), cdd_add(ri, ir, &e6)};
  *e = ((e1 + e2) + (e3 + e4)) + (e5 + e6);
  return r;
}
// bound of |a^ b^ - a b| from operand bounds ea, eb and the product's own error
__device__ inline double cddc_mul_bound(ddc a, double ea, ddc b, double eb, double eop) {
  return fma(cddc_norm1(a), eb, fma(cddc_norm1(b), ea, fma(ea, eb, eop)));
}
// |ddc_to_c(a) - a| exactly (1-norm): the two collapse roundings
__device__ inline double cddc_collapse_err(ddc a) {
  return fabs(two_sum(a.re.hi, a.re.lo).lo) + fabs(two_sum(a.im.hi, a.im.lo).lo);
}

}  // namespace gbsk

namespace gbsk {

// r = dd_mul_d(a, b); *e = |r - a b|
__device__ inline dd cdd_mul_d(dd a, double b, double* e) {
  const dd p = two_prod(a.hi, b);
  const double w = fma(a.lo, b, p.lo);
  const dd r = fast_two_sum(p.hi, w);
  *e = fma(CERT_U, fabs(w), cdd_fts_err(p.hi, w, r) + 2.0 * CERT_UFL);
  return r;
}
// r = dd_div_d(a, b), b a nonzero double; *e = |r - a / b| from the residual
// a - r b (computed in dd with its own bound)
__device__ inline dd cdd_div_d(dd a, double b, double* e) {
  const dd q = dd_div_d(a, b);
  double e1, e2;
  const dd t = cdd_mul_d(q, b, &e1);
  const dd res = cdd_add(a, dd_neg(t), &e2);
  *e = (cdd_mag(res) + e1 + e2) / fabs(b);
  return q;
}
// the complex dd helpers of haf_impl.hpp's ArDD, with their own errors
__device__ inline ddc cddc_fma_c(ddc a, cplx z, ddc acc, double* e) {  // acc + a z
  double e1, e2, e3, e4, e5, e6, e7, e8;
  const dd x1 = cdd_mul_d(a.re, z.re, &e1), x2 = cdd_mul_d(a.im, z.im, &e2);
  const dd x3 = cdd_mul_d(a.re, z.im, &e3), x4 = cdd_mul_d(a.im, z.re, &e4);
  const dd re = cdd_add(x1, dd_neg(x2), &e5), im = cdd_add(x3, x4, &e6);
  const ddc r{cdd_add(acc.re, re, &e7), cdd_add(acc.im, im, &e8)};
  *e = ((e1 + e2) + (e3 + e4)) + ((e5 + e6) + (e7 + e8));
  return r;
}
__device__ inline ddc cddc_fma(ddc a, ddc b, ddc acc, double* e) {  // acc + a b
  double e1, e2;
  const ddc p = cddc_mul(a, b, &e1);
  const ddc r = cddc_add(acc, p, &e2);
  *e = e1 + e2;
  return r;
}
__device__ inline ddc cddc_scale(ddc a, double s, double* e) {
  double e1, e2;
  const ddc r{cdd_mul_d(a.re, s, &e1), cdd_mul_d(a.im, s, &e2)};
  *e = e1 + e2;
  return r;
}
__device__ inline ddc cddc_div_d(ddc a, double s, double* e) {
  double e1, e2;
  const ddc r{cdd_div_d(a.re, s, &e1), cdd_div_d(a.im, s, &e2)};
  *e = e1 + e2;
  return r;
}

}  // namespace gbsk

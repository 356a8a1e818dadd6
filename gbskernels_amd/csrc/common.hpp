// common.hpp -- scalar types and arithmetic shared by every gfx950 kernel.
//
// All value-path arithmetic is written with explicit fma()/mul/add and the
// library is compiled with -ffp-contract=off, so the plain, kappa and certified
// instantiations of a kernel execute the same roundings and return bit-identical
// values (the reference's contract, tests/test_certified.py:457).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gbsk {

// complex128, numpy / std::complex<double> memory layout
struct alignas(16) cplx {
  double re, im;
};

__host__ __device__ inline cplx c_make(double re, double im) { return cplx{re, im}; }
__host__ __device__ inline cplx c_add(cplx a, cplx b) { return cplx{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline cplx c_sub(cplx a, cplx b) { return cplx{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline cplx c_neg(cplx a) { return cplx{-a.re, -a.im}; }
__host__ __device__ inline cplx c_scale(cplx a, double s) { return cplx{a.re * s, a.im * s}; }
// (a.re + i a.im)(b.re + i b.im): re = a.re b.re - a.im b.im, im = a.re b.im + a.im b.re,
// each as one product plus one fused multiply-add.
__host__ __device__ inline cplx c_mul(cplx a, cplx b) {
  return cplx{fma(a.re, b.re, -(a.im * b.im)), fma(a.re, b.im, a.im * b.re)};
}
// a + s * b with s real (one fma per component)
__host__ __device__ inline cplx c_axpy(double s, cplx b, cplx a) {
  return cplx{fma(s, b.re, a.re), fma(s, b.im, a.im)};
}
// a + x * y, complex
__host__ __device__ inline cplx c_fma(cplx x, cplx y, cplx a) {
  return cplx{fma(x.re, y.re, fma(-x.im, y.im, a.re)), fma(x.re, y.im, fma(x.im, y.re, a.im))};
}

// ------------------------------------------------------------ double-double
// Error-free transformations (Knuth TwoSum, Dekker FastTwoSum, FMA TwoProduct)
// and the "sloppy" double-double add/mul of Hida, Li and Bailey's QD library.
struct dd {
  double hi, lo;
};

__host__ __device__ inline dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  return dd{s, (a - (s - bb)) + (b - bb)};
}
__host__ __device__ inline dd fast_two_sum(double a, double b) {
  double s = a + b;
  return dd{s, b - (s - a)};
}
__host__ __device__ inline dd two_prod(double a, double b) {
  double p = a * b;
  return dd{p, fma(a, b, -p)};
}
__host__ __device__ inline dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  return fast_two_sum(s.hi, s.lo + (a.lo + b.lo));
}
__host__ __device__ inline dd dd_add_d(dd a, double b) {
  dd s = two_sum(a.hi, b);
  return fast_two_sum(s.hi, s.lo + a.lo);
}
__host__ __device__ inline dd dd_neg(dd a) { return dd{-a.hi, -a.lo}; }
__host__ __device__ inline dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
__host__ __device__ inline dd dd_mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  return fast_two_sum(p.hi, p.lo + fma(a.hi, b.lo, a.lo * b.hi));
}
__host__ __device__ inline dd dd_mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  return fast_two_sum(p.hi, fma(a.lo, b, p.lo));
}
// a / b for a double b: two quotient digits (QD's accurate_div by a double)
__host__ __device__ inline dd dd_div_d(dd a, double b) {
  double q1 = a.hi / b;
  dd r = dd_sub(a, two_prod(q1, b));
  double q2 = r.hi / b;
  return fast_two_sum(q1, q2);
}
__host__ __device__ inline dd dd_div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = dd_sub(a, dd_mul_d(b, q1));
  double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_mul_d(b, q2));
  double q3 = r.hi / b.hi;
  return dd_add(fast_two_sum(q1, q2), dd{q3, 0.0});
}
// sqrt(a), a >= 0: one Newton correction of the double seed (Karp's trick)
__host__ __device__ inline dd dd_sqrt(dd a) {
  if (!(a.hi > 0.0)) return dd{0.0, 0.0};
  double x = 1.0 / sqrt(a.hi);
  dd y = dd_mul_d(a, x);
  dd diff = dd_sub(a, dd_mul(y, y));
  return dd_add(y, dd_mul_d(diff, 0.5 * x));
}
__host__ __device__ inline double dd_to_d(dd a) { return a.hi + a.lo; }

struct ddc {
  dd re, im;
};
__host__ __device__ inline ddc ddc_from(cplx z) { return ddc{dd{z.re, 0.0}, dd{z.im, 0.0}}; }
__host__ __device__ inline ddc ddc_zero() { return ddc{dd{0.0, 0.0}, dd{0.0, 0.0}}; }
__host__ __device__ inline ddc ddc_add(ddc a, ddc b) { return ddc{dd_add(a.re, b.re), dd_add(a.im, b.im)}; }
__host__ __device__ inline ddc ddc_sub(ddc a, ddc b) { return ddc{dd_sub(a.re, b.re), dd_sub(a.im, b.im)}; }
__host__ __device__ inline ddc ddc_neg(ddc a) { return ddc{dd_neg(a.re), dd_neg(a.im)}; }
__host__ __device__ inline ddc ddc_mul(ddc a, ddc b) {
  return ddc{dd_sub(dd_mul(a.re, b.re), dd_mul(a.im, b.im)), dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re))};
}
__host__ __device__ inline ddc ddc_mul_d(ddc a, double s) { return ddc{dd_mul_d(a.re, s), dd_mul_d(a.im, s)}; }
__host__ __device__ inline ddc ddc_div_d(ddc a, double s) { return ddc{dd_div_d(a.re, s), dd_div_d(a.im, s)}; }
// a + s * z with s = +-2^k (s * z exact in binary64)
__host__ __device__ inline ddc ddc_add_scaled(ddc a, double s, cplx z) {
  return ddc{dd_add_d(a.re, s * z.re), dd_add_d(a.im, s * z.im)};
}
__host__ __device__ inline cplx ddc_to_c(ddc a) { return cplx{dd_to_d(a.re), dd_to_d(a.im)}; }

// |z| without spurious overflow for the kappa indicators (not on any value path)
__device__ inline double c_abs(cplx z) {
  double t = fma(z.re, z.re, z.im * z.im);
  if (t < 1e300 && t > 1e-300) return sqrt(t);
  return hypot(z.re, z.im);
}

// ---------------------------------------------------------- wave utilities
__device__ inline double shfl_xor_d(double v, int mask) { return __shfl_xor(v, mask, 64); }
// 1 / sqrt(a), straight-line: the hardware estimate refined by two Newton
// steps, y' = y + (y / 2) (1 - (a y) y) with the residual formed by one fma. A
// step takes relative error d to 3 d^2 / 2 plus at most 3u/2 of its own
// rounding; the estimate's d0 is only ~2^-24 for some operands on MI355X (one
// step left 38u), so after two the error is 3u/2 + O(u^2) for a in
// [2^-1000, 2^1000]: charged 2u there, checked on the device (gbs_check_rsqrt).
// 8 instructions where the correctly rounded 1.0 / sqrt(a) takes about 30 (a
// range branch around it made the torontonian walk spill). Outside that range
// the result is unspecified: callers flag non-positive pivots, and the
// certified tier refuses pivots outside it.
__device__ inline double rsqrt_nr(double a) {
  const double y0 = __builtin_amdgcn_rsq(a);
  const double y1 = fma(0.5 * y0, fma(-(a * y0), y0, 1.0), y0);
  return fma(0.5 * y1, fma(-(a * y1), y1, 1.0), y1);
}

// the partner lane's value (lane ^ 1) through a DPP quad permutation
// [1, 0, 3, 2]: a VALU move, no LDS traffic (a shuffle is an LDS permute);
// every lane of the wave must execute it
// a double moved across lanes by one DPP pattern (CTRL: 0xB1 lane ^ 1 and
// 0x4E lane ^ 2 inside each quad, 0x00 / 0x55 / 0xAA / 0xFF quad lane 0 / 1 / 2 /
// 3 to the whole quad, 0x104 lane + 4 inside each row of 16, where lanes 12..15
// keep an unspecified value)
template <int CTRL>
__device__ inline double dpp_d(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)b, CTRL, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)hi << 32) | lo));
}
__device__ inline double swap_pair_d(double v) { return dpp_d<0xB1>(v); }

// Sum over groups of G consecutive lanes (G a power of two), butterfly order:
// every lane of the group ends holding the same, deterministically ordered sum.
template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) v += shfl_xor_d(v, m);
  return v;
}
template <int G>
__device__ inline cplx group_sum(cplx v) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) {
    double r = shfl_xor_d(v.re, m), i = shfl_xor_d(v.im, m);
    v.re += r;
    v.im += i;
  }
  return v;
}
template <int G>
__device__ inline ddc group_sum(ddc v) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) {
    ddc o;
    o.re.hi = shfl_xor_d(v.re.hi, m);
    o.re.lo = shfl_xor_d(v.re.lo, m);
    o.im.hi = shfl_xor_d(v.im.hi, m);
    o.im.lo = shfl_xor_d(v.im.lo, m);
    // add in lane order (lower lane first) so both partners compute the same bits
    bool low = (threadIdx.x & m) == 0;
    v = low ? ddc_add(v, o) : ddc_add(o, v);
  }
  return v;
}

constexpr int ilog2(uint64_t v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

}  // namespace gbsk

// tor.hip -- torontonian of real, physical-domain O on gfx950 (fp64, kappa, dd).
//
// tor(O) = sum_{S subset modes} (-1)^(n-|S|) / sqrt(det(I - O)_S), xxpp order
// (reference: cpu_ref/torontonian.py:48-64, cpu_ref/tor_recursive.py:46-101,
// core/tor_recursive.cu:41-124). As in the reference's recursive algorithm the
// subset lattice is walked as a tree over "exclude / include mode j", so every
// determinant is a product of Cholesky pivots shared with its ancestors; a
// non-positive pivot (I - O_S not SPD: off the physical domain) yields NaN.
//
// MI355X design (this file's own). With A = I - O re-ordered mode-major
// (mode i -> rows 2i, 2i+1), one instance is split three ways:
//  * T1 leading modes: their 2^T1 include patterns are separate workgroups;
//    each loads A on its active modes only (included leading, then trailing)
//    and eliminates the included ones (right-looking Cholesky, cooperative
//    over the wave's lanes) to get the Schur complement on the trailing modes;
//  * T2 <= 6 middle modes: the wave walks their include/exclude tree together,
//    one Schur complement per tree level in LDS (packed lower triangle when
//    symmetric), every update spread over the 64 lanes; each of the 2^T2
//    leaves becomes one lane's task (its Schur complement on the last R modes,
//    the running determinant, the count);
//  * R <= 5 trailing modes: each lane walks its task's 2^R leaves alone with a
//    fully unrolled prefix-Cholesky (reference algorithm) whose factor lives in
//    VGPRs.
// Partials: lanes -> butterfly -> one partial per workgroup -> fixed-order sum.
#include <type_traits>

#include "cert.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace gbsk {

constexpr int TOR_MAX_MODES = 32;        // the kernels: dim 64 (one instance at a time at the top)
constexpr int TOR_BATCH_MAX_MODES = 20;  // batched entry points: dim 40
constexpr int TOR_T2 = 6;          // cooperative middle modes (64 tasks)
// built-in schedule (tor_tune): the grouped small-count kernel up to these mode
// counts per family (off: the breadth-first whole-wave kernel was faster at
// every mode count), the breadth-first middle walk up to these (every count:
// measured on MI355X, call times fp64 nm = 6 / 8 / 12 / 16 / 20 4.3x / 3.6x /
// 2.0x / 1.6x / 1.5x faster than depth-first, dd 1.5x-4x, certified 1.2x-3.8x)
constexpr int TOR_GROUPED_MAX_F64 = 0;
constexpr int TOR_GROUPED_MAX_DD = 0;
constexpr int TOR_BFS_MAX_F64 = 32;
constexpr int TOR_BFS_MAX_DD = 32;

// real arithmetic policies: binary64 and double-double (common.hpp)
struct RF64 {
  using T = double;
  static constexpr int RMAX = 5;  // per-lane trailing modes
  static constexpr bool CERT = false;
  // real SPD: 1/sqrt(det) = product of the reciprocal Cholesky diagonals
  static constexpr bool ISD = true;
  static constexpr bool SYM = true;  // symmetric input: packed Cholesky (general LU otherwise)
  // each lane keeps its task's matrix in VGPRs (no LDS task slots): a smaller
  // LDS footprint and more workgroups per CU
  static constexpr bool TASK_REGS = true;
  __device__ static T zero() { return 0.0; }
  __device__ static T one() { return 1.0; }
  __device__ static T from(double x) { return x; }
  __device__ static T one_minus(cplx o) { return 1.0 - o.re; }
  __device__ static T neg(cplx o) { return 0.0 - o.re; }
  __device__ static bool domain(cplx o) { return o.im == 0.0; }  // real O only
  __device__ static double mag(T a) { return a; }
  __device__ static T add(T a, T b) { return a + b; }
  __device__ static T sub(T a, T b) { return a - b; }
  __device__ static T mul(T a, T b) { return a * b; }
  __device__ static T div(T a, T b) { return a / b; }
  __device__ static T recip(T a) { return 1.0 / a; }
  __device__ static T sqrt_(T a) { return sqrt(a); }
  __device__ static T rsqrt(T a) { return rsqrt_nr(a); }  // 1 / sqrt(a), relative error <= 2u
  __device__ static T fms(T a, T b, T c) { return fma(-a, b, c); }  // c - a b
  __device__ static bool pos(T a) { return a > 0.0; }
  __device__ static double to_d(T a) { return a; }
  template <int W>
  __device__ static T sumw(T v) {
    return group_sum<W>(v);
  }
  __device__ static T sum64(T v) { return sumw<64>(v); }
  using P = cplx;  // partials: complex, so the complex fallback can overwrite them
  __device__ static P part(T a) { return cplx{a, 0.0}; }
  __device__ static P part_nan() { return cplx{NAN, NAN}; }
};
struct RDD {
  using T = dd;
  static constexpr int RMAX = 3;
  static constexpr bool CERT = false;
  static constexpr bool ISD = true;
  static constexpr bool SYM = true;
  static constexpr bool TASK_REGS = true;
  __device__ static T zero() { return dd{0.0, 0.0}; }
  __device__ static T one() { return dd{1.0, 0.0}; }
  __device__ static T from(double x) { return dd{x, 0.0}; }
  __device__ static T one_minus(cplx o) { return two_sum(1.0, -o.re); }  // exact
  __device__ static T neg(cplx o) { return dd{-o.re, 0.0}; }
  // the dd tier reads the real part only, as the reference's dd kernel does
  // (core/torontonian_dd.cu:72); callers gate the imaginary part beforehand
  __device__ static bool domain(cplx) { return true; }
  __device__ static double mag(T a) { return dd_to_d(a); }
  __device__ static T add(T a, T b) { return dd_add(a, b); }
  __device__ static T sub(T a, T b) { return dd_sub(a, b); }
  __device__ static T mul(T a, T b) { return dd_mul(a, b); }
  __device__ static T div(T a, T b) { return dd_div(a, b); }
  __device__ static T recip(T a) { return dd_div(dd{1.0, 0.0}, a); }
  __device__ static T sqrt_(T a) { return dd_sqrt(a); }
  __device__ static T rsqrt(T a) { return recip(sqrt_(a)); }
  __device__ static T fms(T a, T b, T c) { return dd_sub(c, dd_mul(a, b)); }
  __device__ static bool pos(T a) { return a.hi > 0.0; }
  __device__ static double to_d(T a) { return dd_to_d(a); }
  using P = dd;
  __device__ static P part(T a) { return a; }
  __device__ static P part_nan() { return dd{NAN, NAN}; }
  // butterfly over W lanes, operands in lane order so both partners agree
  template <int W>
  __device__ static T sumw(T v) {
#pragma unroll
    for (int m = 1; m < W; m <<= 1) {
      dd o{shfl_xor_d(v.hi, m), shfl_xor_d(v.lo, m)};
      v = ((threadIdx.x & m) == 0) ? dd_add(v, o) : dd_add(o, v);
    }
    return v;
  }
  __device__ static T sum64(T v) { return sumw<64>(v); }
};

// any complex O (complex symmetric, Hermitian as in the complex-basis
// O = I - Q^-1, or general): LU without pivoting (the middle tree's Schur
// complements use both off-diagonal pivot entries, the per-lane walk a prefix
// LU of the full task matrix), principal square root of each subset
// determinant (the np.sqrt convention of cpu_ref/torontonian.py:60); a zero
// pivot is a breakdown (NaN)
__device__ inline cplx c_div(cplx a, cplx b) {  // Smith's algorithm
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, den = fma(b.im, r, b.re);
    return cplx{fma(a.im, r, a.re) / den, fma(-a.re, r, a.im) / den};
  }
  const double r = b.re / b.im, den = fma(b.re, r, b.im);
  return cplx{fma(a.re, r, a.im) / den, fma(a.im, r, -a.re) / den};
}
__device__ inline cplx c_sqrt(cplx z) {  // principal branch
  if (z.re == 0.0 && z.im == 0.0) return cplx{0.0, z.im};
  const double r = hypot(z.re, z.im);
  const double t = sqrt(0.5 * (r + fabs(z.re)));
  if (z.re >= 0.0) return cplx{t, z.im / (2.0 * t)};
  return cplx{fabs(z.im) / (2.0 * t), copysign(t, z.im)};
}
struct RC64 {
  using T = cplx;
  static constexpr int RMAX = 3;
  static constexpr bool CERT = false;
  // complex: the principal sqrt of the subset determinant (not the product of
  // the pivots' principal roots, which can differ in sign)
  static constexpr bool ISD = false;
  // any complex O (complex symmetric, Hermitian, general): LU, no symmetry assumed
  static constexpr bool SYM = false;
  static constexpr bool TASK_REGS = false;
  __device__ static T zero() { return cplx{0.0, 0.0}; }
  __device__ static T one() { return cplx{1.0, 0.0}; }
  __device__ static T from(double x) { return cplx{x, x}; }
  __device__ static T one_minus(cplx o) { return cplx{1.0 - o.re, 0.0 - o.im}; }
  __device__ static T neg(cplx o) { return cplx{0.0 - o.re, 0.0 - o.im}; }
  __device__ static bool domain(cplx) { return true; }
  __device__ static double mag(T a) { return hypot(a.re, a.im); }
  __device__ static T add(T a, T b) { return c_add(a, b); }
  __device__ static T sub(T a, T b) { return c_sub(a, b); }
  __device__ static T mul(T a, T b) { return c_mul(a, b); }
  __device__ static T div(T a, T b) { return c_div(a, b); }
  __device__ static T recip(T a) { return c_div(cplx{1.0, 0.0}, a); }
  __device__ static T sqrt_(T a) { return c_sqrt(a); }
  __device__ static T rsqrt(T a) { return recip(sqrt_(a)); }
  __device__ static T fms(T a, T b, T c) { return c_sub(c, c_mul(a, b)); }
  __device__ static bool pos(T a) { return a.re != 0.0 || a.im != 0.0; }
  __device__ static double to_d(T a) { return a.re; }
  template <int W>
  __device__ static T sumw(T v) {
    return group_sum<W>(v);
  }
  __device__ static T sum64(T v) { return sumw<64>(v); }
  using P = cplx;
  __device__ static P part(T a) { return a; }
  __device__ static P part_nan() { return cplx{NAN, NAN}; }
};

// Certified fp64: the value of every operation is exactly RF64's (same split,
// same rounding sequence, so values are bit-identical to the fp64 tier), and a
// running bound e >= |value - exact| travels with it. Rules, u = 2^-53, each
// bound expression evaluated in round-to-nearest and inflated once at the end
// (cert_finish, rounding depth D):
//   a +- b      e = ea + eb + u|v|
//   a * b       e = |a| eb + |b| ea + ea eb + u|v| + UFL
//   c - a b     (one fma) e = ec + |a| eb + |b| ea + ea eb + u|v| + UFL
//   a / b       needs |b| >= 4 eb (then |b_exact| >= |b| / 2):
//               e = 2 (ea + (1 + u)|v| eb) / |b| + u|v| + UFL
//   sqrt(a)     needs a >= 4 ea (so a_exact > 0):
//               e = (1 + u) ea / v + u v + UFL   (|sqrt x - sqrt y| <= |x - y| / sqrt y)
//   1/sqrt(a)   (rsqrt_nr, relative error <= 2u on [2^-1000, 2^1000]) needs
//               4 ea <= a and 2^-1000 <= a <= 2^600:
//               e = 0.78 ea v^3 + 2u v + UFL     (a_exact >= 3a/4, so
//               |x^-1/2 - a^-1/2| <= ea / (2 (3a/4)^(3/2)) = 0.7698 ea a^(-3/2))
// A failed requirement makes the bound +inf (refusal); pivots are divided by
// or square-rooted, so a certified bound also certifies that every exact
// pivot is positive (the real formula applies).
struct cv {
  double v, e;
};
struct RCERT {
  using T = cv;
  static constexpr int RMAX = RF64::RMAX;  // the fp64 tier's split, for identical values
  static constexpr bool CERT = true;
  static constexpr bool ISD = true;
  static constexpr bool SYM = true;
  static constexpr bool TASK_REGS = true;
  __device__ static T zero() { return cv{0.0, 0.0}; }
  __device__ static T one() { return cv{1.0, 0.0}; }
  __device__ static T one_minus(cplx o) {
    const double v = 1.0 - o.re;
    return cv{v, CERT_U * fabs(v)};
  }
  __device__ static T neg(cplx o) { return cv{0.0 - o.re, 0.0}; }
  __device__ static bool domain(cplx o) { return o.im == 0.0; }
  __device__ static double mag(T a) { return a.v; }
  __device__ static T add(T a, T b) {
    const double v = a.v + b.v;
    return cv{v, fma(CERT_U, fabs(v), a.e + b.e)};
  }
  __device__ static T sub(T a, T b) {
    const double v = a.v - b.v;
    return cv{v, fma(CERT_U, fabs(v), a.e + b.e)};
  }
  __device__ static double prod_err(double a, double ea, double b, double eb, double v) {
    double t = fma(fabs(a), eb, CERT_UFL);
    t = fma(fabs(b), ea, t);
    t = fma(ea, eb, t);
    return fma(CERT_U, fabs(v), t);
  }
  __device__ static T mul(T a, T b) {
    const double v = a.v * b.v;
    return cv{v, prod_err(a.v, a.e, b.v, b.e, v)};
  }
  __device__ static T fms(T a, T b, T c) {
    const double v = fma(-a.v, b.v, c.v);
    return cv{v, prod_err(a.v, a.e, b.v, b.e, v) + c.e};
  }
  __device__ static T div(T a, T b) {
    const double v = a.v / b.v;
    const double bb = fabs(b.v);
    if (!(bb >= 4.0 * b.e)) return cv{v, INFINITY};
    const double t = fma((1.0 + 2.0 * CERT_U) * fabs(v), b.e, a.e);
    return cv{v, fma(CERT_U, fabs(v), fma(2.0, t / bb, CERT_UFL))};
  }
  __device__ static T recip(T a) { return div(one(), a); }
  __device__ static T sqrt_(T a) {
    const double v = sqrt(a.v);
    if (a.e == 0.0) return cv{v, fma(CERT_U, v, CERT_UFL)};
    if (!(a.v >= 4.0 * a.e)) return cv{v, INFINITY};
    return cv{v, fma(CERT_U, v, fma(1.0 + 2.0 * CERT_U, a.e / v, CERT_UFL))};
  }
  __device__ static T rsqrt(T a) {
    const double v = rsqrt_nr(a.v);
    if (!(a.v >= 4.0 * a.e) || !(a.v >= 0x1p-1000) || !(a.v <= 0x1p600)) return cv{v, INFINITY};
    return cv{v, fma(2.0 * CERT_U, v, fma(0.78 * a.e, v * v * v, CERT_UFL))};
  }
  __device__ static bool pos(T a) { return a.v > 0.0; }
  template <int W>
  __device__ static T sumw(T x) {
#pragma unroll
    for (int m = 1; m < W; m <<= 1) {
      const double ov = shfl_xor_d(x.v, m), oe = shfl_xor_d(x.e, m);
      x.v += ov;  // the same butterfly as group_sum<W> of the fp64 tier
      x.e = fma(CERT_U, fabs(x.v), x.e + oe);
    }
    return x;
  }
  __device__ static T sum64(T x) { return sumw<64>(x); }
  using P = cplx;
  __device__ static P part(T a) { return cplx{a.v, 0.0}; }
  __device__ static P part_nan() { return cplx{NAN, NAN}; }
  __device__ static double bound(T a) { return a.e; }
};

// Certified double-double: RDD's values (same split and operations, so values
// are bit-identical to the dd tier's) with a running absolute bound. Each
// operation adds the bound of its own rounding (cert.hpp, a posteriori) to the
// propagated operand errors:
//   a +- b   ea + eb
//   a * b    |a| eb + |b| ea + ea eb                       (|a| = |hi| + |lo|)
//   a / b    needs lb(b) >= 4 eb; 2 (ea + (|q| + e_q) eb) / lb(b)
//   sqrt(a)  needs lb(a) >= 4 ea and lb(y) >= 2 e_y; 2 ea / lb(y)
// where lb() is a guaranteed lower bound of a dd magnitude. The division's and
// square root's own errors come from dd residuals (a - q b, y^2 - a).
struct cdd {
  dd v;
  double e;
};
struct RCDD {
  using T = cdd;
  static constexpr int RMAX = RDD::RMAX;  // the dd tier's split, for identical values
  static constexpr bool CERT = true;
  static constexpr bool ISD = true;
  static constexpr bool SYM = true;
  static constexpr bool TASK_REGS = true;
  __device__ static T zero() { return cdd{dd{0.0, 0.0}, 0.0}; }
  __device__ static T one() { return cdd{dd{1.0, 0.0}, 0.0}; }
  __device__ static T one_minus(cplx o) { return cdd{two_sum(1.0, -o.re), 0.0}; }  // exact
  __device__ static T neg(cplx o) { return cdd{dd{-o.re, 0.0}, 0.0}; }
  __device__ static bool domain(cplx o) { return o.im == 0.0; }
  __device__ static double mag(T a) { return dd_to_d(a.v); }
  __device__ static T add(T a, T b) {
    double r;
    const dd v = cdd_add(a.v, b.v, &r);
    return cdd{v, a.e + b.e + r};
  }
  __device__ static T sub(T a, T b) {
    double r;
    const dd v = cdd_add(a.v, dd_neg(b.v), &r);
    return cdd{v, a.e + b.e + r};
  }
  __device__ static T mul(T a, T b) {
    double r;
    const dd v = cdd_mul(a.v, b.v, &r);
    const double t = fma(cdd_mag(a.v), b.e, fma(cdd_mag(b.v), a.e, a.e * b.e));
    return cdd{v, t + r};
  }
  __device__ static T fms(T a, T b, T c) { return sub(c, mul(a, b)); }  // = dd_sub(c, dd_mul(a, b))
  __device__ static T div(T a, T b) {
    const dd q = dd_div(a.v, b.v);
    const double lb = cdd_low(b.v);
    if (!(lb > 0.0) || !(lb >= 4.0 * b.e)) return cdd{q, INFINITY};
    double e1, e2;
    const dd p = cdd_mul(q, b.v, &e1);
    const dd res = cdd_add(a.v, dd_neg(p), &e2);  // a - q b, up to e1 + e2
    const double eq = (cdd_mag(res) + e1 + e2) / lb;
    const double prop = 2.0 * fma(cdd_mag(q) + eq, b.e, a.e) / lb;
    return cdd{q, eq + prop};
  }
  __device__ static T recip(T a) { return div(one(), a); }
  __device__ static T sqrt_(T a) {
    const dd y = dd_sqrt(a.v);
    if (a.v.hi == 0.0 && a.v.lo == 0.0 && a.e == 0.0) return cdd{y, 0.0};
    const double la = cdd_low(a.v), ly = cdd_low(y);
    if (!(la > 0.0) || !(la >= 4.0 * a.e) || !(ly > 0.0)) return cdd{y, INFINITY};
    double e1, e2;
    const dd yy = cdd_mul(y, y, &e1);
    const dd res = cdd_add(yy, dd_neg(a.v), &e2);  // y^2 - a
    const double ey = (cdd_mag(res) + e1 + e2) / ly;
    if (!(ly >= 2.0 * ey)) return cdd{y, INFINITY};
    return cdd{y, ey + 2.0 * a.e / ly};
  }
  __device__ static T rsqrt(T a) { return recip(sqrt_(a)); }

  __device__ static bool pos(T a) { return a.v.hi > 0.0; }
  template <int W>
  __device__ static T sumw(T x) {
#pragma unroll
    for (int m = 1; m < W; m <<= 1) {
      const T o{dd{shfl_xor_d(x.v.hi, m), shfl_xor_d(x.v.lo, m)}, shfl_xor_d(x.e, m)};
      x = ((threadIdx.x & m) == 0) ? add(x, o) : add(o, x);  // RDD::sumw's order
    }
    return x;
  }
  __device__ static T sum64(T x) { return sumw<64>(x); }
  using P = dd;
  __device__ static P part(T a) { return a.v; }
  __device__ static P part_nan() { return dd{NAN, NAN}; }
  __device__ static double bound(T a) { return a.e; }
};

struct TorSplit {
  int R, T2, T1;
  int G = 1;  // instances per wave (lane groups of 2^T2 lanes; 1 = the whole-wave kernel)
};
// lanes first (T2 middle modes -> up to 64 tasks), then per-lane depth R, then
// leading patterns as separate workgroups
__host__ __device__ inline TorSplit tor_split(int nm, int rmax) {
  TorSplit s;
  s.T2 = nm < TOR_T2 ? nm : TOR_T2;
  const int rest = nm - s.T2;
  s.R = rest < rmax ? rest : rmax;
  s.T1 = rest - s.R;
  return s;
}

// one large instance: at least 2^10 leading patterns (workgroups) where the
// mode count allows, the trailing modes per lane up to rmax
inline TorSplit tor_single_split(int nm, int rmax) {
  TorSplit s = tor_split(nm, rmax);
  const int rest = nm - s.T2;
  const int t1 = rest - rmax > (rest < 10 ? rest : 10) ? rest - rmax : (rest < 10 ? rest : 10);
  s.T1 = t1;
  s.R = rest - t1;
  return s;
}

// packed lower-triangular index, and its inverse
__device__ constexpr int tri(int r, int c) { return r * (r + 1) / 2 + c; }
__device__ constexpr int tri_row(int e) {
  int r = 0;
  while ((r + 1) * (r + 2) / 2 <= e) ++r;
  return r;
}
__device__ constexpr int tri_col(int e) { return e - tri(tri_row(e), 0); }
// run-time inverse of tri for the cooperative loops (e < 2^20)
__device__ inline void tri_decode(int e, int& r, int& c) {
  r = (int)((sqrtf(8.0f * (float)e + 1.0f) - 1.0f) * 0.5f);
  if (tri(r + 1, 0) <= e) ++r;
  if (tri(r, 0) > e) --r;
  c = e - tri(r, 0);
}

// ----------------------------------------------------------- per-lane bottom
// Prefix-Cholesky walk over the R trailing modes of one task. M is the task's
// 2R x 2R Schur complement (packed lower, mode-major), L the factor prefix.
// Real policies (ISD) keep the reciprocal of each diagonal entry in L[r][r]
// (columns are scaled by it, no division per entry) and carry det as the
// product of those reciprocals times 1/sqrt(det) of the task's upstream
// pivots, so a leaf's term is that product (no square root per leaf). The
// complex policy keeps the diagonal values and the determinant.
template <class Ar, int R, int D, unsigned MASK, int C>
struct Bottom {
  using T = typename Ar::T;
  static __device__ __forceinline__ void run(const T* __restrict__ M, T (&L)[2 * R][2 * R], T det, int count, int nm,
                                             bool& bad, T& acc, double& aacc) {
    // exclude mode D
    Bottom<Ar, R, D + 1, MASK, C>::run(M, L, det, count, nm, bad, acc, aacc);
    // include mode D: append rows 2C (x) and 2C+1 (p)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int r = 2 * C + t;
      const int gr = 2 * D + t;
#pragma unroll
      for (int c = 0; c <= r; ++c) {
        // global (mode-major) index of L column c: position c >> 1 of the included list
        int pos = c >> 1, mode = D, k = 0;
#pragma unroll
        for (int j = 0; j < D; ++j)
          if ((MASK >> j) & 1u) {
            if (k == pos) mode = j;
            ++k;
          }
        const int gc = 2 * mode + (c & 1);
        T s = gr >= gc ? M[tri(gr, gc)] : M[tri(gc, gr)];
#pragma unroll
        for (int u = 0; u < c; ++u) s = Ar::fms(L[r][u], L[c][u], s);
        if (c < r) {
          L[r][c] = Ar::ISD ? Ar::mul(s, L[c][c]) : Ar::div(s, L[c][c]);
        } else {
          bad |= !Ar::pos(s);
          L[r][r] = Ar::ISD ? Ar::rsqrt(s) : Ar::sqrt_(s);
        }
      }
    }
    const T q = Ar::mul(L[2 * C][2 * C], L[2 * C + 1][2 * C + 1]);
    const T child = Ar::ISD ? Ar::mul(det, q) : Ar::mul(det, Ar::mul(q, q));
    Bottom<Ar, R, D + 1, MASK | (1u << D), C + 1>::run(M, L, child, count + 1, nm, bad, acc, aacc);
  }
};

// leaf: (-1)^(nm - count) / sqrt(det); `det` is already 1/sqrt(det) for ISD
template <class Ar>
__device__ __forceinline__ void tor_leaf(typename Ar::T det, int count, int nm, typename Ar::T& acc, double& aacc) {
  const typename Ar::T t = Ar::ISD ? det : Ar::recip(Ar::sqrt_(det));
  acc = ((nm - count) & 1) ? Ar::sub(acc, t) : Ar::add(acc, t);
  aacc += Ar::mag(t);
}

template <class Ar, int R, unsigned MASK, int C>
struct Bottom<Ar, R, R, MASK, C> {
  using T = typename Ar::T;
  static __device__ __forceinline__ void run(const T* __restrict__, T (&)[2 * R][2 * R], T det, int count, int nm,
                                             bool&, T& acc, double& aacc) {
    tor_leaf<Ar>(det, count, nm, acc, aacc);
  }
};

// Non-symmetric (complex) policy: the same walk with a prefix LU (Doolittle,
// unit L) of the full task matrix: appending index r gives U[c][r], L[r][c]
// (c < r) and the pivot U[r][r]; the determinant is the product of pivots.
template <class Ar, int R, int D, unsigned MASK, int C>
struct BottomLU {
  using T = typename Ar::T;
  static __device__ __forceinline__ void run(const T* __restrict__ M, T (&L)[2 * R][2 * R], T (&U)[2 * R][2 * R],
                                             T det, int count, int nm, bool& bad, T& acc, double& aacc) {
    BottomLU<Ar, R, D + 1, MASK, C>::run(M, L, U, det, count, nm, bad, acc, aacc);
    T piv = Ar::one();
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int r = 2 * C + t;
      const int gr = 2 * D + t;
#pragma unroll
      for (int c = 0; c < r; ++c) {
        int pos = c >> 1, mode = D, k = 0;
#pragma unroll
        for (int j = 0; j < D; ++j)
          if ((MASK >> j) & 1u) {
            if (k == pos) mode = j;
            ++k;
          }
        const int gc = 2 * mode + (c & 1);
        T u = M[gc * (2 * R) + gr], l = M[gr * (2 * R) + gc];
#pragma unroll
        for (int w = 0; w < c; ++w) {
          u = Ar::fms(L[c][w], U[w][r], u);
          l = Ar::fms(L[r][w], U[w][c], l);
        }
        U[c][r] = u;
        L[r][c] = Ar::div(l, U[c][c]);
      }
      T p = M[gr * (2 * R) + gr];
#pragma unroll
      for (int w = 0; w < r; ++w) p = Ar::fms(L[r][w], U[w][r], p);
      bad |= !Ar::pos(p);
      U[r][r] = p;
      piv = Ar::mul(piv, p);
    }
    BottomLU<Ar, R, D + 1, MASK | (1u << D), C + 1>::run(M, L, U, Ar::mul(det, piv), count + 1, nm, bad, acc, aacc);
  }
};

template <class Ar, int R, unsigned MASK, int C>
struct BottomLU<Ar, R, R, MASK, C> {
  using T = typename Ar::T;
  static __device__ __forceinline__ void run(const T* __restrict__, T (&)[2 * R][2 * R], T (&)[2 * R][2 * R], T det,
                                             int count, int nm, bool&, T& acc, double& aacc) {
    tor_leaf<Ar>(det, count, nm, acc, aacc);
  }
};

// task slot entries: packed lower triangle (symmetric policies) or full matrix
template <class Ar, int R>
__host__ __device__ constexpr int tor_slot() {
  return Ar::SYM ? (2 * R) * (2 * R + 1) / 2 : (2 * R) * (2 * R);
}

template <class Ar, int R>
__device__ inline void bottom_walk(const typename Ar::T* M, typename Ar::T det, int count, int nm, bool& bad,
                                   typename Ar::T& acc, double& aacc) {
  constexpr int W = 2 * R > 0 ? 2 * R : 1;
  // ISD: the walk carries 1/sqrt of the determinant (upstream pivots here)
  const typename Ar::T d0 = Ar::ISD ? Ar::rsqrt(det) : det;
  if constexpr (R == 0) {
    tor_leaf<Ar>(d0, count, nm, acc, aacc);
  } else if constexpr (Ar::SYM) {
    typename Ar::T L[W][W];
    Bottom<Ar, R, 0, 0u, 0>::run(M, L, d0, count, nm, bad, acc, aacc);
  } else {
    typename Ar::T L[W][W], U[W][W];
    BottomLU<Ar, R, 0, 0u, 0>::run(M, L, U, d0, count, nm, bad, acc, aacc);
  }
}

// ----------------------------------------------------- cooperative helpers
// Eliminate pivot p of the D x D matrix X (full storage, LDS) in place on the
// trailing block: X[i][j] -= (X[i][p] / X[p][p]) X[p][j] for i, j > p. Pivots
// are taken in order, so rows and columns up to p are never read again. Returns
// the pivot.
template <class Ar>
__device__ inline typename Ar::T eliminate(typename Ar::T* X, int D, int p) {
  const typename Ar::T piv = X[p * D + p];
  const typename Ar::T rp = Ar::recip(piv);
  const int n = D - p - 1;
  __syncthreads();
  if (n > 0) {
    // lane entries e = lane, lane + 64, ... of the n x n trailing block, the
    // (row, column) stepped without a division per entry
    int i = (int)threadIdx.x / n, j = (int)threadIdx.x - i * n;
    const int di = 64 / n, dj = 64 - di * n;
    for (int e = threadIdx.x; e < n * n; e += 64) {
      const int r = p + 1 + i, c = p + 1 + j;
      X[r * D + c] = Ar::fms(Ar::mul(X[r * D + p], rp), X[p * D + c], X[r * D + c]);
      i += di;
      j += dj;
      if (j >= n) {
        j -= n;
        ++i;
      }
    }
  }
  __syncthreads();
  return piv;
}

// Level matrices of the middle tree: packed lower triangles for symmetric
// policies (a Schur complement of a symmetric matrix is symmetric: half the
// storage and half the update), full D x D otherwise.
template <bool SYM>
__host__ __device__ constexpr int lev_size(int D) { return SYM ? D * (D + 1) / 2 : D * D; }

// Y = X (Dx x Dx) with its first mode (rows and columns 0, 1) removed.
template <bool SYM, class T>
__device__ inline void drop_first_mode(const T* X, int Dx, T* Y) {
  const int Dy = Dx - 2;
  for (int e = threadIdx.x; e < lev_size<SYM>(Dy); e += 64) {
    if constexpr (SYM) {
      int i, j;
      tri_decode(e, i, j);
      Y[e] = X[tri(i + 2, j + 2)];
    } else {
      const int i = e / Dy, j = e % Dy;
      Y[e] = X[(i + 2) * Dx + j + 2];
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------- kernel
// workgroup = (instance b, leading pattern p1); out partials [B][2^T1] (T)
// RERUN: only workgroups whose partial is NaN (a real-path breakdown or a
// complex O) run, overwriting it (the complex fallback of the fp64 tier).
// The fp64 policy asks for three waves per SIMD: at R = 5 that is 168 VGPRs and
// 16 spilled (the default allocation took 186-190, two waves), measured 1.2x
// faster at 12-20 modes; four waves helped at 12-16 modes but lost at 20, and
// the dd policies gained nothing from a hint.
template <class Ar>
constexpr int tor_min_waves() { return std::is_same<Ar, RF64>::value ? 3 : 1; }

#ifndef TOR_BFS_Q
#define TOR_BFS_Q 1
#endif

// Bits the breadth-first walk is split on (2^Q parts walked one after the
// other, one part's include matrices held): 1 (two parts) for fp64, dd and
// certified; four parts held less LDS still but their extra passes cost more
// than the occupancy gained (fp64 1.1-1.2x, dd 1.2-1.3x slower at 12-20
// modes). 0 for certified dd, whose walk spills 720 bytes per lane to scratch:
// more of its workgroups per CU only thrashed the caches (20 modes 1.15x slower).
template <class Ar>
constexpr int tor_bfs_split() {
  return std::is_same<Ar, RCDD>::value ? 0 : (std::is_same<Ar, RCERT>::value ? 1 : TOR_BFS_Q);
}

// Results a whole-wave launch of the fp64 tier writes itself when an instance
// has one partial (T1 = 0), in place of the reduce or merge launch, whose
// arithmetic on one partial is 0 + value (the other lanes' and the
// butterfly's zeros change no bit after it). out == nullptr: partials only.
//   real kernel: out / aux (kappa) and the complex rerun's marks cpart[npc]
//   rerun kernel: out / aux (kappa: the complex sum |term|, cert_inf: +inf)
struct TorDirect {
  cplx* out = nullptr;
  double* aux = nullptr;
  cplx* cpart = nullptr;
  int npc = 0;
  bool cert_inf = false;
};

template <class Ar, int R, int T2, bool KAPPA, bool RERUN, bool BFS>
__device__ __forceinline__ void tor_workgroup(const cplx* __restrict__ mats, int nm, int T1, int B,
                                              typename Ar::P* __restrict__ part, double* __restrict__ part_abs,
                                              int64_t wg, const TorDirect& dir) {
  using T = typename Ar::T;
  extern __shared__ double lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);  // split: T1 (argument) + T2 + R = nm
  const int dim = 2 * nm;
  const int lane = threadIdx.x;
  const int64_t b = wg >> T1;
  const unsigned p1 = (unsigned)(wg & ((1 << T1) - 1));
  if constexpr (RERUN) {
    const typename Ar::P prev = part[(size_t)b * ((size_t)1 << T1) + p1];
    if (!(prev.re != prev.re)) return;  // this workgroup's real partial is valid
  }
  // LDS layout: level stack (T2+1 levels), then one region used first by the
  // work matrix (dim^2, dead once level 0 is extracted) and then by the task slots
  T* lev = lds;
  int levtot = 0;
#pragma unroll
  for (int d = 0; d <= T2; ++d) levtot += lev_size<Ar::SYM>(2 * (nm - T1 - d));
  constexpr int SLOT = tor_slot<Ar, R>();
  constexpr bool TREG = Ar::TASK_REGS;
  // TREG: the work matrix and the level stack share the region from 0 (level 0
  // is staged through registers); otherwise the work matrix, later the task
  // slots, follow the level stack
  T* W = TREG ? lds : lev + levtot;
  T* slots = W;
  T* slot_det = slots + 64 * SLOT;
  int* slot_cnt = reinterpret_cast<int*>(slot_det + 64);
  __shared__ int bad_flag;
  if (lane == 0) bad_flag = 0;
  __syncthreads();

  // Active modes: the included leading modes in mode order, then the trailing
  // modes. Only their rows and columns enter a pivot or the Schur complement,
  // so W holds just them (D x D). The workgroup with every leading mode
  // included loads all of O, so a complex entry anywhere still flags the
  // instance.
  __shared__ int amode[TOR_MAX_MODES];
  const int k1 = __popc(p1);
  const int D = 2 * (k1 + nm - T1);
  if (lane == 0) {
    int q = 0;
    for (int j = 0; j < T1; ++j)
      if ((p1 >> j) & 1u) amode[q++] = j;
    for (int j = T1; j < nm; ++j) amode[q++] = j;
  }
  __syncthreads();
  // A = I - O, mode-major: W[(2i+s)][(2j+t)] = delta - O[mi + s*nm][mj + t*nm]
  // for active modes mi = amode[i], mj = amode[j]
  const cplx* O = mats + b * dim * dim;
  {
    int r = lane / D, c = lane - r * D;  // (row, column) of entry e, stepped per iteration
    const int dr = 64 / D, dc = 64 - dr * D;
    for (int e = lane; e < D * D; e += 64) {
      const int gr = amode[r >> 1] + (r & 1) * nm, gc = amode[c >> 1] + (c & 1) * nm;
      const cplx o = O[gr * dim + gc];
      W[e] = (r == c) ? Ar::one_minus(o) : Ar::neg(o);
      if (!Ar::domain(o)) bad_flag = 1;  // complex O: not the real kernels' domain
      r += dr;
      c += dc;
      if (c >= D) {
        c -= D;
        ++r;
      }
    }
  }
  __syncthreads();

  // leading modes: eliminate the included ones (mode order, x then p pivot)
  T det = Ar::one();
  const int count = k1;
  bool bad = false;
  for (int q = 0; q < k1; ++q) {
    const T a = eliminate<Ar>(W, D, 2 * q);
    const T c = eliminate<Ar>(W, D, 2 * q + 1);
    bad |= !Ar::pos(a) || !Ar::pos(c);
    det = Ar::mul(det, Ar::mul(a, c));
  }
  // level 0 = W restricted to the trailing modes T1..nm-1 (from row 2 k1);
  // symmetric policies keep its lower triangle
  {
    const int D0 = 2 * (nm - T1), o0 = 2 * k1, n0 = lev_size<Ar::SYM>(D0);
    auto src = [&](int e) -> int {
      int i, j;
      if constexpr (Ar::SYM) {
        tri_decode(e, i, j);
      } else {
        i = e / D0;
        j = e % D0;
      }
      return (o0 + i) * D + o0 + j;
    };
    if constexpr (TREG) {  // in place: read everything, then write
      constexpr int D0MAX = 2 * (TOR_T2 + R);
      constexpr int PER = lev_size<Ar::SYM>(D0MAX) / 64 + 1;
      T tmp[PER];
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = lane + 64 * k;
        if (e < n0) tmp[k] = W[src(e)];
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = lane + 64 * k;
        if (e < n0) lev[e] = tmp[k];
      }
    } else {
      for (int e = lane; e < n0; e += 64) lev[e] = W[src(e)];
    }
    __syncthreads();
  }

  T Mreg[TREG && SLOT > 0 ? SLOT : 1];  // TREG: this lane's task (packed lower triangle)
  T my_det = Ar::one();
  int my_cnt = 0;
  int ntask = 0;
  if constexpr (BFS) {
    // Breadth-first walk over the T2 middle modes (register-task policies).
    // Level d has the 2^d include / exclude patterns of middle modes 0..d-1
    // (node index = pattern bits, first mode most significant, include = 1).
    // Only include children are formed: excluding a mode drops the first two
    // rows and columns, and the packed lower-triangle index tri(i, j) does not
    // depend on the matrix size, so an exclude child is its parent read with a
    // shift of two rows. A node reads the matrix formed by its last include
    // (the level-0 matrix if none), shifted by two rows per exclude since.
    // Each level's new matrices are formed in one pass over the 64 lanes.
    // Pivots, complements and determinants are the depth-first walk's
    // operations, so the values are identical; the leaf level is formed in the
    // lanes' registers (lane k = leaf k = task k of the depth-first order).
    // With all six middle modes the tree is walked in 2^Q parts split on the
    // first Q modes (tor_bfs_split), one after the other, each filling its
    // 2^(6-Q) lanes' leaves: only one part's include matrices (and the split
    // path's) are held. At 11 modes below the leading ones: 3437 entries for
    // the whole tree, 1950 at Q = 1, 1292 at Q = 2 (fp64: the LDS allowed 5
    // workgroups per CU, the registers 12).
    static_assert(Ar::SYM, "breadth-first walk of the symmetric policies");
    constexpr int Q = (T2 == TOR_T2) ? tor_bfs_split<Ar>() : 0;  // split bits
    constexpr int NODES = 1 << (T2 - 1);
    __shared__ T bpiv[NODES][3];            // ra, rc, b01 of the current level's nodes (part-local index)
    __shared__ T bdet[(2 << T2) - 1];        // determinants, level d node n at 2^d - 1 + n
    __shared__ int boff[NODES], bsh[NODES];  // each node's matrix offset and row shift
    __shared__ int region[TOR_T2 + 1];       // first entry of level d's formed matrices
    const int base = nm - T1;                // modes of the level-0 matrix
    // (row, column) of every packed lower-triangle index a formed matrix can
    // hold (at most 2 (TOR_T2 + 5) - 2 rows below level 0)
    constexpr int TIJ = 256;
    static_assert(lev_size<true>(2 * (TOR_T2 + 5) - 2) <= TIJ, "triangle table");
    __shared__ unsigned char tij_r[TIJ], tij_c[TIJ];
    for (int e = lane; e < TIJ; e += 64) {
      int i, j;
      tri_decode(e, i, j);
      tij_r[e] = (unsigned char)i;
      tij_c[e] = (unsigned char)j;
    }
    if (lane == 0) {
      // level 0, then per level d >= 1 the include matrices one part holds:
      // 2^(d-1-Q) of them above the split, one on the split path
      bdet[0] = det;
      int r = lev_size<true>(2 * base);
      region[0] = 0;
      for (int d = 1; d <= T2; ++d) {
        region[d] = r;
        r += (d > Q ? (1 << (d - 1 - Q)) : 1) * lev_size<true>(2 * (base - d));
      }
    }
    __syncthreads();
    // pass 1 on the nodes n0 .. n0 + cnt - 1 of level d (part h): each node's
    // matrix (offset, shift), pivots and its children's determinants
    auto pivots = [&](int d, int h, int n0, int cnt) {
      if (lane < cnt) {
        const int n = n0 + lane;
        int off = 0, sh = 2 * d;
        if (n != 0) {
          const int p = __builtin_ctz((unsigned)n);  // excludes since the last include
          const int lv = d - p;                      // level of the matrix that include formed
          const int g = n >> (p + 1);                // its parent at level lv - 1
          const int loc = lv > Q ? g - (h << (lv - 1 - Q)) : 0;
          off = region[lv] + loc * lev_size<true>(2 * (base - lv));
          sh = 2 * p;
        }
        const T* X = lev + off;
        const T a = X[tri(sh, sh)];
        const T b01 = X[tri(sh + 1, sh)];
        const T ra = Ar::recip(a);
        const T c = Ar::sub(X[tri(sh + 1, sh + 1)], Ar::mul(Ar::mul(b01, b01), ra));
        const T rc = Ar::recip(c);
        bpiv[lane][0] = ra;
        bpiv[lane][1] = rc;
        bpiv[lane][2] = b01;
        boff[lane] = off;
        bsh[lane] = sh;
        bad |= !Ar::pos(a) || !Ar::pos(c);
        const T dn = bdet[(1 << d) - 1 + n];
        bdet[(2 << d) - 1 + 2 * n] = dn;
        bdet[(2 << d) - 1 + 2 * n + 1] = Ar::mul(dn, Ar::mul(a, c));
      }
      __syncthreads();
    };
    // pass 2: the include child (the Schur complement of its two pivots) of
    // each of the cnt nodes of pass 1, formed at region[d + 1] + node * Ly
    // (part-local node index). The flat index (node, entry) is split without
    // an integer division (a float reciprocal, exact to one step below 2^24,
    // then corrected) and the entry's (row, column) comes from the shared table.
    auto complements = [&](int d, int cnt) {
      const int Ly = lev_size<true>(2 * (base - d - 1));
      T* const Y = lev + region[d + 1];
      const int total = cnt * Ly;
      const float rLy = 1.0f / (float)Ly;
      for (int idx = lane; idx < total; idx += 64) {
        int nu = (int)((float)idx * rLy);
        if (nu * Ly > idx)
          --nu;
        else if ((nu + 1) * Ly <= idx)
          ++nu;
        const int e = idx - nu * Ly, sh = bsh[nu];
        const T* X = lev + boff[nu];
        const int i = tij_r[e] + 2 + sh, j = tij_c[e] + 2 + sh;
        const T ra = bpiv[nu][0], rc = bpiv[nu][1], b01 = bpiv[nu][2];
        const T li0 = Ar::mul(X[tri(i, sh)], ra);
        const T li1 = Ar::mul(Ar::fms(li0, b01, X[tri(i, sh + 1)]), rc);
        const T x0j = X[tri(j, sh)];
        const T x1j = Ar::fms(Ar::mul(x0j, ra), b01, X[tri(j, sh + 1)]);
        T v = Ar::fms(li0, x0j, X[tri(i, j)]);
        Y[idx] = Ar::fms(li1, x1j, v);
      }
      __syncthreads();
    };
    // the split path: level 0's node and its include child (level 1), once;
    // deeper path levels l < Q each time the part's node there changes
    if constexpr (Q >= 1) {
      pivots(0, 0, 0, 1);
      complements(0, 1);
    }
    ntask = 1 << T2;
#pragma unroll 1
    for (int h = 0; h < (1 << Q); ++h) {
      for (int l = 1; l < Q; ++l) {
        if (h % (1 << (Q - l)) == 0) {
          pivots(l, 0, h >> (Q - l), 1);
          complements(l, 1);
        }
      }
      for (int d = Q; d < T2; ++d) {
        const int cnt = 1 << (d - Q);
        pivots(d, h, h * cnt, cnt);
        if (d + 1 < T2) complements(d, cnt);
      }
      // leaves of part h: lane k forms task k from its parent (level T2 - 1)
      // in registers
      const int k0 = h << (T2 - Q);
      if (lane >= k0 && lane < k0 + (1 << (T2 - Q))) {
        const int nu = (lane >> 1) - (k0 >> 1), sh = bsh[nu];
        const T* X = lev + boff[nu];
        const bool inc = lane & 1;
        const T ra = bpiv[nu][0], rc = bpiv[nu][1], b01 = bpiv[nu][2];
#pragma unroll
        for (int e = 0; e < SLOT; ++e) {
          const int i = tri_row(e) + 2 + sh, j = tri_col(e) + 2 + sh;
          if (!inc) {
            Mreg[e] = X[tri(i, j)];
          } else {
            const T li0 = Ar::mul(X[tri(i, sh)], ra);
            const T li1 = Ar::mul(Ar::fms(li0, b01, X[tri(i, sh + 1)]), rc);
            const T x0j = X[tri(j, sh)];
            const T x1j = Ar::fms(Ar::mul(x0j, ra), b01, X[tri(j, sh + 1)]);
            T v = Ar::fms(li0, x0j, X[tri(i, j)]);
            Mreg[e] = Ar::fms(li1, x1j, v);
          }
        }
        my_det = bdet[(1 << T2) - 1 + lane];
        my_cnt = count + __popc((unsigned)lane);
      }
      __syncthreads();  // the next part overwrites the matrices above the split
    }
  } else {
    // cooperative walk over the T2 middle modes' include / exclude tree,
    // exclude first, one Schur complement per level (level d's Dd x Dd matrix,
    // the next level right after it). The walk's state is wave-uniform and lives
    // in LDS (every lane writes the same value), not in per-lane scratch. Leaf k
    // in walk order becomes task k.
    __shared__ T wdet[TOR_T2 + 1];
    __shared__ int wcnt[TOR_T2 + 1], wstate[TOR_T2 + 1], woff[TOR_T2 + 1];
    wdet[0] = det;
    wcnt[0] = count;
    wstate[0] = 0;
    woff[0] = 0;
    int d = 0;
    for (;;) {
      const int Dd = 2 * (nm - T1 - d);
      T* X = lev + woff[d];
      if (d == T2 && TREG) {
        // leaf: task ntask goes to lane ntask's registers
        if (lane == ntask) {
  #pragma unroll
          for (int e = 0; e < SLOT; ++e) Mreg[e] = X[e];  // the leaf level is the packed task
          my_det = wdet[d];
          my_cnt = wcnt[d];
        }
        ++ntask;
        if (d == 0) break;
        --d;
        continue;
      }
      if (d == T2) {
        // leaf: the task is this level's matrix (the trailing R modes)
        // (the leaf level has the task's layout: packed lower or full 2R x 2R)
        for (int e = lane; e < SLOT; e += 64) slots[ntask * SLOT + e] = X[e];
        if (lane == 0) {
          slot_det[ntask] = wdet[d];
          slot_cnt[ntask] = wcnt[d];
        }
        ++ntask;
        if (d == 0) break;
        --d;
        continue;
      }
      T* Y = X + lev_size<Ar::SYM>(Dd);
      const int st = wstate[d];
      if (st == 0) {  // exclude middle mode d: drop its rows
        wstate[d] = 1;
        drop_first_mode<Ar::SYM>(X, Dd, Y);
        wdet[d + 1] = wdet[d];
        wcnt[d + 1] = wcnt[d];
      } else if (st == 1) {
        // include: two pivots (x then p), Schur complement into Y. With the
        // 2 x 2 pivot block [[a, b], [b, c']], c = c' - b^2 / a, l_i0 = X_i0 / a
        // and l_i1 = (X_i1 - l_i0 b) / c, the complement is
        // Y_ij = X_ij - l_i0 X_0j - l_i1 (X_1j - b X_0j / a): the reciprocals of
        // the two pivots, no division per entry
        // Symmetric policies (packed lower): X_0j = X_j0, X_1j = X_j1, and only
        // the lower triangle of Y is formed.
        wstate[d] = 2;
        const T a = X[0];
        const T b01 = X[1];                                   // packed: X_10
        const T b10 = Ar::SYM ? b01 : X[Dd];                  // [[a, b01], [b10, c']]
        const T ra = Ar::recip(a);
        const T c = Ar::sub(Ar::SYM ? X[2] : X[Dd + 1], Ar::mul(Ar::mul(b10, b01), ra));
        const T rc = Ar::recip(c);
        const int Dy = Dd - 2;
        for (int e = lane; e < lev_size<Ar::SYM>(Dy); e += 64) {
          int i, j, xi0, xi1, x0j_, x1j_, xij;
          if constexpr (Ar::SYM) {
            tri_decode(e, i, j);
            i += 2;
            j += 2;
            xi0 = tri(i, 0);
            xi1 = tri(i, 1);
            x0j_ = tri(j, 0);
            x1j_ = tri(j, 1);
            xij = tri(i, j);
          } else {
            i = e / Dy + 2;
            j = e % Dy + 2;
            xi0 = i * Dd;
            xi1 = i * Dd + 1;
            x0j_ = j;
            x1j_ = Dd + j;
            xij = i * Dd + j;
          }
          const T li0 = Ar::mul(X[xi0], ra);
          const T li1 = Ar::mul(Ar::fms(li0, b01, X[xi1]), rc);
          const T x0j = X[x0j_];
          const T x1j = Ar::fms(Ar::mul(x0j, ra), b10, X[x1j_]);
          T v = Ar::fms(li0, x0j, X[xij]);
          v = Ar::fms(li1, x1j, v);
          Y[e] = v;
        }
        __syncthreads();
        bad |= !Ar::pos(a) || !Ar::pos(c);
        wdet[d + 1] = Ar::mul(wdet[d], Ar::mul(a, c));
        wcnt[d + 1] = wcnt[d] + 1;
      } else {
        if (d == 0) break;
        --d;
        continue;
      }
      wstate[d + 1] = 0;
      woff[d + 1] = woff[d] + lev_size<Ar::SYM>(Dd);
      ++d;
    }
    __syncthreads();
  }
  __syncthreads();

  // per-lane trailing walks
  T acc = Ar::zero();
  double aacc = 0.0;
  if (lane < ntask) {
    if constexpr (TREG)
      bottom_walk<Ar, R>(Mreg, my_det, my_cnt, nm, bad, acc, aacc);
    else
      bottom_walk<Ar, R>(slots + lane * SLOT, slot_det[lane], slot_cnt[lane], nm, bad, acc, aacc);
  }
  if (bad) bad_flag = 1;
  acc = Ar::sum64(acc);
  if (KAPPA) aacc = group_sum<64>(aacc);
  __syncthreads();
  if (lane == 0) {
    const size_t np = (size_t)1 << T1;
    const typename Ar::P pv = bad_flag ? Ar::part_nan() : Ar::part(acc);
    if constexpr (std::is_same<typename Ar::P, cplx>::value && !Ar::CERT) {
      if (dir.out != nullptr) {  // T1 = 0: b is the slot
        const cplx s{0.0 + pv.re, 0.0 + pv.im};
        dir.out[b] = s;
        if (KAPPA) dir.aux[b] = 0.0 + (bad_flag ? NAN : aacc);
        if (!RERUN && dir.cpart != nullptr) {
          const cplx m = (s.re != s.re) ? cplx{NAN, NAN} : cplx{0.0, 0.0};
          for (int i = 0; i < dir.npc; ++i) dir.cpart[b * dir.npc + i] = m;
        }
        if (RERUN && dir.cert_inf) dir.aux[b] = INFINITY;
        return;
      }
    }
    part[b * np + p1] = pv;
    if constexpr (Ar::CERT) {
      part_abs[b * np + p1] = bad_flag ? INFINITY : Ar::bound(acc);
    } else if (KAPPA) {
      part_abs[b * np + p1] = bad_flag ? NAN : aacc;
    }
  }
}

// RERUN: min(slots, TOR_RERUN_GRID) workgroups, each owning a contiguous range
// of partial slots; the 64 lanes test 64 slots at once and the workgroup then
// recomputes the NaN ones in order. When no instance fell off the real path
// the pass is a quick scan (at 20 modes, one workgroup per slot -- 8.4M --
// took 5 ms to find nothing); when every slot did (a complex O batch), up to
// 8192 workgroups recompute in parallel (one slot each at 4096 instances
// without leading modes: 64 in a row per workgroup took 1.3 ms at 4 modes).
constexpr int64_t TOR_RERUN_GRID = 8192;

template <class Ar, int R, int T2, bool KAPPA, bool RERUN, bool BFS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(tor_min_waves<Ar>()))) void tor_real_kernel(
    const cplx* __restrict__ mats, int nm, int T1, int B, typename Ar::P* __restrict__ part,
    double* __restrict__ part_abs, const TorDirect dir) {
  if constexpr (RERUN) {
    const int64_t total = (int64_t)B << T1;
    const int64_t per = (total + gridDim.x - 1) / gridDim.x;
    const int64_t lo = (int64_t)blockIdx.x * per, hi = lo + per < total ? lo + per : total;
    for (int64_t base = lo; base < hi; base += 64) {
      const int64_t mine = base + threadIdx.x;
      bool nan = false;
      if (mine < hi) {
        const typename Ar::P prev = part[mine];
        nan = prev.re != prev.re;
      }
      unsigned long long todo = __ballot(nan);  // one wave per workgroup: uniform
      while (todo) {
        const int k = __builtin_ctzll(todo);
        todo &= todo - 1;
        __syncthreads();  // the previous slot's LDS use is over
        tor_workgroup<Ar, R, T2, KAPPA, RERUN, BFS>(mats, nm, T1, B, part, part_abs, base + k, dir);
      }
    }
  } else {
    tor_workgroup<Ar, R, T2, KAPPA, RERUN, BFS>(mats, nm, T1, B, part, part_abs, blockIdx.x, dir);
  }
}

// ---------------------------------------------------- small mode counts
// When R = min(nm, RMAX) trailing modes per lane leave T2 = nm - R < 6 middle
// modes, an instance needs only 2^T2 lanes: a wave carries G = 2^(6 - T2)
// instances, lane group gi walking instance gi's include / exclude tree over the
// T2 middle modes with the whole-wave kernel's level operations (each level's
// Schur complement spread over the group's lanes, packed lower triangles in
// LDS), one leaf task per lane, the unrolled trailing walk in VGPRs, and a
// butterfly over the group. Every group walks the same tree shape (no leading
// modes here), so the walk's control state is shared; determinants are per group.
template <class Ar, int R, int T2>
__host__ __device__ constexpr int tor_small_levtot() {
  int t = 0;
  for (int d = 0; d <= T2; ++d) t += lev_size<true>(2 * (T2 + R - d));
  return t;
}

template <class Ar, int R, int T2, bool KAPPA>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(tor_min_waves<Ar>()))) void tor_small_kernel(
    const cplx* __restrict__ mats, int B, typename Ar::P* __restrict__ part, double* __restrict__ part_abs) {
  static_assert(Ar::SYM && Ar::TASK_REGS && Ar::ISD, "the grouped walk is the symmetric register-task one");
  using T = typename Ar::T;
  constexpr int nm = T2 + R, dim = 2 * nm;
  constexpr int LPI = 1 << T2, G = 64 / LPI;
  constexpr int LEV = tor_small_levtot<Ar, R, T2>();
  constexpr int SLOT = tor_slot<Ar, R>();
  extern __shared__ double lds_raw[];  // G level stacks of LEV entries (T2 > 0)
  __shared__ T wdet[G][T2 + 1];
  __shared__ int wcnt[T2 + 1], wstate[T2 + 1], woff[T2 + 1];
  __shared__ int gbad[G];
  const int lane = threadIdx.x;
  const int gi = lane / LPI, t = lane % LPI;
  const int64_t b = (int64_t)blockIdx.x * G + gi;
  const bool live = b < B;
  T* const mylev = reinterpret_cast<T*>(lds_raw) + gi * LEV;
  if (t == 0) gbad[gi] = 0;
  __syncthreads();
  bool bad = false;
  T my_det = Ar::one();
  int my_cnt = 0;
  T Mreg[SLOT > 0 ? SLOT : 1];
  // A = I - O mode-major (mode i -> rows 2i, 2i+1), packed lower triangle;
  // every entry of O is checked for the real domain
  const cplx* O = mats + (live ? b : 0) * dim * dim;
  if constexpr (T2 == 0) {
    // one lane, one instance: the task is the whole matrix, loaded straight
    // into the lane's registers
    if (live) {
      for (int e = 0; e < dim * dim; ++e) bad |= !Ar::domain(O[e]);
#pragma unroll
      for (int e = 0; e < SLOT; ++e) {
        const int r = tri_row(e), c = tri_col(e);
        const cplx o = O[((r >> 1) + (r & 1) * nm) * dim + (c >> 1) + (c & 1) * nm];
        Mreg[e] = (r == c) ? Ar::one_minus(o) : Ar::neg(o);
      }
    }
  } else {
    if (live) {
      for (int e = t; e < dim * dim; e += LPI) {
        const int r = e / dim, c = e % dim;
        const int gr = (r >> 1) + (r & 1) * nm, gc = (c >> 1) + (c & 1) * nm;
        const cplx o = O[gr * dim + gc];
        if (!Ar::domain(o)) gbad[gi] = 1;
        if (c <= r) mylev[tri(r, c)] = (r == c) ? Ar::one_minus(o) : Ar::neg(o);
      }
    }
    __syncthreads();
    wdet[gi][0] = Ar::one();
    wcnt[0] = 0;
    wstate[0] = 0;
    woff[0] = 0;
    int ntask = 0, d = 0;
    for (;;) {
      const int Dd = 2 * (nm - d);
      T* X = mylev + woff[d];
      if (d == T2) {
        if (t == ntask) {
#pragma unroll
          for (int e = 0; e < SLOT; ++e) Mreg[e] = X[e];
          my_det = wdet[gi][d];
          my_cnt = wcnt[d];
        }
        ++ntask;
        --d;
        continue;
      }
      T* Y = X + lev_size<true>(Dd);
      const int st = wstate[d];
      const int Dy = Dd - 2;
      if (st == 0) {  // exclude middle mode d: drop its rows and columns
        wstate[d] = 1;
        for (int e = t; e < lev_size<true>(Dy); e += LPI) {
          int i, j;
          tri_decode(e, i, j);
          Y[e] = X[tri(i + 2, j + 2)];
        }
        __syncthreads();
        wdet[gi][d + 1] = wdet[gi][d];
        wcnt[d + 1] = wcnt[d];
      } else if (st == 1) {
        // include: the two pivots (x then p) and the Schur complement, as the
        // whole-wave kernel forms it
        wstate[d] = 2;
        const T a = X[0];
        const T b01 = X[1];
        const T ra = Ar::recip(a);
        const T c = Ar::sub(X[2], Ar::mul(Ar::mul(b01, b01), ra));
        const T rc = Ar::recip(c);
        for (int e = t; e < lev_size<true>(Dy); e += LPI) {
          int i, j;
          tri_decode(e, i, j);
          i += 2;
          j += 2;
          const T li0 = Ar::mul(X[tri(i, 0)], ra);
          const T li1 = Ar::mul(Ar::fms(li0, b01, X[tri(i, 1)]), rc);
          const T x0j = X[tri(j, 0)];
          const T x1j = Ar::fms(Ar::mul(x0j, ra), b01, X[tri(j, 1)]);
          T v = Ar::fms(li0, x0j, X[tri(i, j)]);
          v = Ar::fms(li1, x1j, v);
          Y[e] = v;
        }
        __syncthreads();
        bad |= !Ar::pos(a) || !Ar::pos(c);
        wdet[gi][d + 1] = Ar::mul(wdet[gi][d], Ar::mul(a, c));
        wcnt[d + 1] = wcnt[d] + 1;
      } else {
        if (d == 0) break;
        --d;
        continue;
      }
      wstate[d + 1] = 0;
      woff[d + 1] = woff[d] + lev_size<true>(Dd);
      ++d;
    }
  }
  // the task's trailing walk; bottom_walk turns the upstream determinant into
  // its reciprocal square root first
  T acc = Ar::zero();
  double aacc = 0.0;
  if (live) bottom_walk<Ar, R>(Mreg, my_det, my_cnt, nm, bad, acc, aacc);
  if (bad) gbad[gi] = 1;
  acc = Ar::template sumw<LPI>(acc);
  if (KAPPA) aacc = group_sum<LPI>(aacc);
  __syncthreads();
  if (live && t == 0) {
    const bool fail = gbad[gi] != 0;
    part[b] = fail ? Ar::part_nan() : Ar::part(acc);
    if constexpr (Ar::CERT) {
      part_abs[b] = fail ? INFINITY : Ar::bound(acc);
    } else if (KAPPA) {
      part_abs[b] = fail ? NAN : aacc;
    }
  }
}

// Schedule knobs of the batched torontonian (host state, per policy family:
// 0 = fp64 / kappa / certified, 1 = dd / certified-dd). grouped_max: mode counts
// up to which the grouped small-count kernel runs (its split differs from the
// whole-wave kernel's, so it changes values' last bits; the built-in choice is
// fixed per family). bfs_max: mode counts up to which the whole-wave kernel
// walks its middle tree breadth-first (values identical to the depth-first walk).
struct TorTune {
  int grouped_max;
  int bfs_max;
};
static TorTune tor_tune[2] = {{TOR_GROUPED_MAX_F64, TOR_BFS_MAX_F64}, {TOR_GROUPED_MAX_DD, TOR_BFS_MAX_DD}};
template <class Ar>
static TorTune& tor_tune_of() {
  return tor_tune[(std::is_same<Ar, RDD>::value || std::is_same<Ar, RCDD>::value) ? 1 : 0];
}

// the batched split of a policy: the grouped small-count form up to the
// family's grouped_max modes (T1 = 0, T2 = nm - min(nm, RMAX) < TOR_T2), else
// the whole-wave form
template <class Ar>
inline TorSplit tor_batch_split(int nm) {
  if (Ar::TASK_REGS && Ar::SYM && nm <= tor_tune_of<Ar>().grouped_max) {
    const int R = nm < Ar::RMAX ? nm : Ar::RMAX;
    if (nm - R < TOR_T2) {
      TorSplit s;
      s.R = R;
      s.T2 = nm - R;
      s.T1 = 0;
      s.G = 64 >> s.T2;
      return s;
    }
  }
  return tor_split(nm, Ar::RMAX);
}

template <class Ar, bool KAPPA, int R, int T2>
static void launch_tor_small_rt(const cplx* d_in, int B, typename Ar::P* part, double* part_abs, hipStream_t st) {
  constexpr int G = 64 >> T2;
  const size_t lds = T2 > 0 ? (size_t)G * tor_small_levtot<Ar, R, T2>() * sizeof(typename Ar::T) : 0;
  if (lds > 65536)  // the certified policies' wider entries: up to gfx950's 160 KiB
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&tor_small_kernel<Ar, R, T2, KAPPA>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((tor_small_kernel<Ar, R, T2, KAPPA>), dim3((unsigned)(((int64_t)B + G - 1) / G)), dim3(64),
                     lds, st, d_in, B, part, part_abs);
}

// R = min(nm, RMAX): nm <= RMAX gives (R = nm, T2 = 0), larger nm (R = RMAX, T2 = nm - RMAX)
template <class Ar, bool KAPPA>
static int launch_tor_small(int nm, const cplx* d_in, int B, typename Ar::P* part, double* part_abs,
                            hipStream_t st) {
  constexpr int RM = Ar::RMAX;
  switch (nm) {
#define TOR_SMALL_CASE(NM)                                                                                 \
  case NM:                                                                                                 \
    if constexpr (NM - (NM < RM ? NM : RM) < TOR_T2)                                                       \
      launch_tor_small_rt<Ar, KAPPA, (NM < RM ? NM : RM), NM - (NM < RM ? NM : RM)>(d_in, B, part, part_abs, st); \
    else                                                                                                   \
      return GBSK_ERR_DIM;                                                                                 \
    break;
    TOR_SMALL_CASE(1)
    TOR_SMALL_CASE(2)
    TOR_SMALL_CASE(3)
    TOR_SMALL_CASE(4)
    TOR_SMALL_CASE(5)
    TOR_SMALL_CASE(6)
    TOR_SMALL_CASE(7)
    TOR_SMALL_CASE(8)
    TOR_SMALL_CASE(9)
    TOR_SMALL_CASE(10)
#undef TOR_SMALL_CASE
    default: return GBSK_ERR_DIM;
  }
  return 0;
}

// Reductions: one wave per instance, 4 instances per 256-thread block. Lane l
// sums partials l, l + 64, l + 128, ... in order, then a fixed butterfly over
// the 64 lanes; lane 0 writes. The association depends on np only, so plain,
// kappa and certified (and dd, certified-dd) reductions give the same bits.
constexpr int TOR_RED_WAVES = 4;
__device__ inline int64_t tor_red_instance() { return (int64_t)blockIdx.x * TOR_RED_WAVES + (threadIdx.x >> 6); }
static unsigned tor_red_grid(int B) { return (unsigned)(((int64_t)B + TOR_RED_WAVES - 1) / TOR_RED_WAVES); }

// butterfly of dd values, operands in lane order so partners agree bit for bit
__device__ inline dd dd_wave_sum(dd s) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    const dd o{shfl_xor_d(s.hi, m), shfl_xor_d(s.lo, m)};
    s = ((threadIdx.x & m) == 0) ? dd_add(s, o) : dd_add(o, s);
  }
  return s;
}
// the same with each addition's certified error (cert.hpp)
__device__ inline dd cdd_wave_sum(dd s, double& e) {
#pragma unroll
  for (int m = 1; m < 64; m <<= 1) {
    const dd o{shfl_xor_d(s.hi, m), shfl_xor_d(s.lo, m)};
    const double eo = shfl_xor_d(e, m);
    double r;
    s = ((threadIdx.x & m) == 0) ? cdd_add(s, o, &r) : cdd_add(o, s, &r);
    e = e + eo + r;
  }
  return s;
}

// With cpart, an fp64 reduce also marks the complex fallback's partials of
// its instance: all npc of them NaN (to be recomputed) when the real sum is
// NaN, else 0 (skipped); every lane holds the sum after the butterfly.
__device__ inline void tor_mark_fallback(const cplx& s, int64_t b, int lane, cplx* __restrict__ cpart, int npc) {
  if (cpart == nullptr) return;
  const cplx m = (s.re != s.re) ? cplx{NAN, NAN} : cplx{0.0, 0.0};
  for (int i = lane; i < npc; i += 64) cpart[b * npc + i] = m;
}

template <bool DD, bool KAPPA>
__global__ __launch_bounds__(256) void tor_reduce_kernel(const void* __restrict__ partv,
                                                         const double* __restrict__ part_abs, int np, int B,
                                                         cplx* __restrict__ out, double* __restrict__ aux,
                                                         cplx* __restrict__ cpart, int npc) {
  const int64_t b = tor_red_instance();
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  double a = 0.0;
  if constexpr (DD) {
    const dd* part = reinterpret_cast<const dd*>(partv) + b * np;
    dd s{0.0, 0.0};
    for (int i = lane; i < np; i += 64) s = dd_add(s, part[i]);
    s = dd_wave_sum(s);
    if (lane == 0) out[b] = cplx{dd_to_d(s), 0.0};
  } else {
    const cplx* part = reinterpret_cast<const cplx*>(partv) + b * np;
    cplx s{0.0, 0.0};
    for (int i = lane; i < np; i += 64) {
      s = c_add(s, part[i]);
      if (KAPPA) a += part_abs[b * np + i];
    }
    s = group_sum<64>(s);
    if (KAPPA) a = group_sum<64>(a);
    if (lane == 0) out[b] = s;
    tor_mark_fallback(s, b, lane, cpart, npc);
  }
  if (KAPPA && lane == 0) aux[b] = a;
}

// certified reduce: the fp64 tier's sum of the partials (same bits) with the
// partials' bounds plus one rounding charge per addition
__global__ __launch_bounds__(256) void tor_reduce_cert_kernel(const cplx* __restrict__ part,
                                                              const double* __restrict__ part_e, int np, int B,
                                                              double depth, cplx* __restrict__ out,
                                                              double* __restrict__ aux, cplx* __restrict__ cpart,
                                                              int npc) {
  const int64_t b = tor_red_instance();
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  cplx s{0.0, 0.0};
  double e = 0.0;
  for (int i = lane; i < np; i += 64) {
    s = c_add(s, part[b * np + i]);
    e = cert_round(s, e + part_e[b * np + i]);
  }
  cert_group_sum<64>(s, e);
  tor_mark_fallback(s, b, lane, cpart, npc);
  if (lane != 0) return;
  out[b] = s;
  const double f = cert_finish(e, depth);
  aux[b] = (f == f) ? f : INFINITY;
}

// certified-dd reduce: the dd tier's dd sum of the partials (same bits) and
// its collapse to binary64, whose rounding is measured exactly
__global__ __launch_bounds__(256) void tor_reduce_ddcert_kernel(const dd* __restrict__ part,
                                                                const double* __restrict__ part_e, int np, int B,
                                                                double depth, cplx* __restrict__ out,
                                                                double* __restrict__ aux) {
  const int64_t b = tor_red_instance();
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  dd s{0.0, 0.0};
  double e = 0.0;
  for (int i = lane; i < np; i += 64) {
    double r;
    s = cdd_add(s, part[b * np + i], &r);
    e = e + part_e[b * np + i] + r;
  }
  s = cdd_wave_sum(s, e);
  if (lane != 0) return;
  const double v = dd_to_d(s);
  e += fabs(two_sum(s.hi, s.lo).lo);  // |v - (s.hi + s.lo)|, exact
  out[b] = cplx{v, 0.0};
  const double f = cert_finish(e, depth);
  aux[b] = (f == f) ? f : INFINITY;
}

// instances the complex kernel recomputed take its sum (same association as
// the real reduce; certified tier: such an instance's bound is +inf -- the
// complex path is not certified)
template <bool KAPPA, bool CERT>
__global__ __launch_bounds__(256) void tor_merge_kernel(const cplx* __restrict__ cpart,
                                                        const double* __restrict__ cpart_abs, int np, int B,
                                                        cplx* __restrict__ out, double* __restrict__ aux) {
  const int64_t b = tor_red_instance();
  const int lane = threadIdx.x & 63;
  if (b >= B) return;
  const cplx v = out[b];
  if (!(v.re != v.re)) return;  // wave-uniform (one instance per wave)
  cplx s{0.0, 0.0};
  double a = 0.0;
  for (int i = lane; i < np; i += 64) {
    s = c_add(s, cpart[b * np + i]);
    if (KAPPA) a += cpart_abs[b * np + i];
  }
  s = group_sum<64>(s);
  if (KAPPA) a = group_sum<64>(a);
  if (lane != 0) return;
  out[b] = s;
  if (KAPPA) aux[b] = a;
  if (CERT) aux[b] = INFINITY;
}

__global__ void tor_n0_kernel(int B, cplx* out, double* aux, double auxval) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  out[b] = cplx{1.0, 0.0};
  if (aux) aux[b] = auxval;
}

// LDS entries of the breadth-first walk: the level-0 matrix and, per level
// d = 1 .. T2 - 1, the include matrices one part of the tree holds
// (2^(d-1-Q) above the split, one on the split path; Q = tor_bfs_split at
// T2 = TOR_T2)
static size_t tor_bfs_entries(int nm, const TorSplit& sp, int split) {
  const size_t base = (size_t)(nm - sp.T1);
  const int Q = sp.T2 == TOR_T2 ? split : 0;  // the kernel's split bits (one part held at a time)
  size_t t = base * (2 * base + 1);       // lev_size(2 base)
  for (int d = 1; d < sp.T2; ++d) {
    const size_t D = 2 * (base - d);
    t += (d > Q ? ((size_t)1 << (d - 1 - Q)) : (size_t)1) * (D * (D + 1) / 2);
  }
  return t;
}

template <class Ar>
static size_t tor_lds_bytes(int nm, const TorSplit& sp, bool bfs) {
  const int dim = 2 * nm;
  size_t lev = 0;
  for (int d = 0; d <= sp.T2; ++d) {
    const size_t Dd = 2 * (nm - sp.T1 - d);
    lev += Ar::SYM ? Dd * (Dd + 1) / 2 : Dd * Dd;  // the kernel's lev_size
  }
  if (Ar::TASK_REGS && bfs) lev = tor_bfs_entries(nm, sp, tor_bfs_split<Ar>());
  const size_t slot = Ar::SYM ? (size_t)(2 * sp.R) * (2 * sp.R + 1) / 2 : (size_t)(2 * sp.R) * (2 * sp.R);
  const size_t T = sizeof(typename Ar::T);
  // task matrices in VGPRs: the level stack overlays the work matrix
  if (Ar::TASK_REGS) return (lev > (size_t)dim * dim ? lev : (size_t)dim * dim) * T + 16;
  const size_t task_area = 64 * slot * T + 64 * T + 64 * 4;
  const size_t work = (size_t)dim * dim * T;
  return lev * T + (work > task_area ? work : task_area);
}

template <class Ar, int R, int T2, bool KAPPA, bool RERUN, bool BFS>
static void launch_tor_rt1(int nm, int T1, const cplx* d_in, int B, typename Ar::P* part, double* part_abs,
                           dim3 grid, size_t lds, hipStream_t st, const TorDirect& dir) {
  if (lds > 65536)  // dim 64 at the wider tiers: up to gfx950's 160 KiB per workgroup
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&tor_real_kernel<Ar, R, T2, KAPPA, RERUN, BFS>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((tor_real_kernel<Ar, R, T2, KAPPA, RERUN, BFS>), grid, dim3(64), lds, st, d_in, nm, T1, B, part,
                     part_abs, dir);
}

// the whole-wave kernel, with the breadth-first middle walk when `bfs` (the
// register-task policies only; the depth-first walk otherwise)
template <class Ar, int R, int T2, bool KAPPA, bool RERUN>
static void launch_tor_rt(int nm, int T1, const cplx* d_in, int B, typename Ar::P* part, double* part_abs, dim3 grid,
                          bool bfs, hipStream_t st, const TorDirect& dir) {
  const TorSplit sp{R, T2, T1};
  if constexpr (Ar::TASK_REGS && Ar::SYM && T2 >= 1) {
    if (bfs) {
      launch_tor_rt1<Ar, R, T2, KAPPA, RERUN, true>(nm, T1, d_in, B, part, part_abs, grid,
                                                    tor_lds_bytes<Ar>(nm, sp, true), st, dir);
      return;
    }
  }
  launch_tor_rt1<Ar, R, T2, KAPPA, RERUN, false>(nm, T1, d_in, B, part, part_abs, grid,
                                                 tor_lds_bytes<Ar>(nm, sp, false), st, dir);
}


// the split's R and T2 are template arguments: T2 = 6 with R = 0..RMAX
// (nm >= 6), or T2 = nm < 6 with R = 0
template <class Ar, bool KAPPA, bool RERUN>
// dir: see TorDirect; honoured only by the whole-wave kernel (callers pass it
// when sp.G == 1 and sp.T1 == 0)
static int launch_tor(int nm, const TorSplit& sp, const cplx* d_in, int B, typename Ar::P* part, double* part_abs,
                      hipStream_t st, const TorDirect& dir = TorDirect{}) {
  if constexpr (!RERUN && Ar::TASK_REGS && Ar::SYM) {
    if (sp.G > 1) return launch_tor_small<Ar, KAPPA>(nm, d_in, B, part, part_abs, st);
  }
  const int np = 1 << sp.T1;
  const bool bfs = Ar::TASK_REGS && nm <= tor_tune_of<Ar>().bfs_max;
  const int64_t slots = (int64_t)B * np;
  const dim3 grid((unsigned)(RERUN ? (slots < TOR_RERUN_GRID ? slots : TOR_RERUN_GRID) : slots));
  if (sp.T2 == TOR_T2) {
    switch (sp.R) {
#define TOR_CASE(RR)                                                                                             \
  case RR:                                                                                                       \
    if constexpr (RR <= Ar::RMAX) launch_tor_rt<Ar, RR, TOR_T2, KAPPA, RERUN>(nm, sp.T1, d_in, B, part, part_abs, grid, bfs, st, dir); \
    break;
      TOR_CASE(0)
      TOR_CASE(1)
      TOR_CASE(2)
      TOR_CASE(3)
      TOR_CASE(4)
      TOR_CASE(5)
#undef TOR_CASE
      default: return GBSK_ERR_DIM;
    }
    return 0;
  }
  if (sp.R != 0) return GBSK_ERR_DIM;
  switch (sp.T2) {
    case 1: launch_tor_rt<Ar, 0, 1, KAPPA, RERUN>(nm, sp.T1, d_in, B, part, part_abs, grid, bfs, st, dir); break;
    case 2: launch_tor_rt<Ar, 0, 2, KAPPA, RERUN>(nm, sp.T1, d_in, B, part, part_abs, grid, bfs, st, dir); break;
    case 3: launch_tor_rt<Ar, 0, 3, KAPPA, RERUN>(nm, sp.T1, d_in, B, part, part_abs, grid, bfs, st, dir); break;
    case 4: launch_tor_rt<Ar, 0, 4, KAPPA, RERUN>(nm, sp.T1, d_in, B, part, part_abs, grid, bfs, st, dir); break;
    case 5: launch_tor_rt<Ar, 0, 5, KAPPA, RERUN>(nm, sp.T1, d_in, B, part, part_abs, grid, bfs, st, dir); break;
    default: return GBSK_ERR_DIM;
  }
  return 0;
}

// fp64 tier: real-domain kernel; workgroups that found a complex entry or a
// non-positive pivot are recomputed by the complex (LU) kernel, whose split
// is chosen so that both kernels partition the instance identically (same T1).
// rounding depth of the certified bound: per mode a Schur level (<= 16 bound
// operations), per lane the 2^R leaf sums, the butterfly and the np-term reduce
static double tor_cert_depth(int nm, int np) { return 64.0 * nm + 64.0 + 16.0 + 2.0 * np + 64.0; }

template <class ArR, bool KAPPA>
static int run_tor_fp64(int nm, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, hipStream_t st) {
  static_assert(ArR::RMAX == RF64::RMAX, "the real split must be the fp64 tier's");
  constexpr bool CERT = ArR::CERT;
  const TorSplit sr = tor_batch_split<ArR>(nm), sc = tor_split(nm, RC64::RMAX);
  const int np = 1 << (sr.T1 > sc.T1 ? sr.T1 : sc.T1);
  (void)np;
  cplx* part = reinterpret_cast<cplx*>(ws);
  // the two kernels' leading splits differ (R differs); the complex rerun is
  // therefore done per instance: the real kernel's partials are reduced first,
  // and the complex kernel recomputes whole instances whose sum is NaN
  const int npr = 1 << sr.T1, npc = 1 << sc.T1;
  double* part_abs = reinterpret_cast<double*>(part + (size_t)B * npr);
  cplx* cpart = reinterpret_cast<cplx*>(part_abs + (size_t)B * npr);
  double* cpart_abs = reinterpret_cast<double*>(cpart + (size_t)B * npc);
  // one partial per instance: the whole-wave kernels write the reduce's and
  // the merge's results themselves (two launches fewer at <= 8 modes)
  const bool direct_real = !CERT && npr == 1 && sr.G == 1;
  const bool direct_rerun = npc == 1;  // the rerun is always the whole-wave kernel
  TorDirect dr, dc;
  if (direct_real) dr = TorDirect{d_out, d_aux, cpart, npc, false};
  if (direct_rerun) dc = TorDirect{d_out, d_aux, nullptr, 0, CERT};
  int rc = launch_tor<ArR, KAPPA, false>(nm, sr, d_in, B, part, part_abs, st, dr);
  if (rc) return rc;
  // the reduce also marks the partials the complex kernel recomputes
  if constexpr (CERT)
    hipLaunchKernelGGL(tor_reduce_cert_kernel, dim3(tor_red_grid(B)), dim3(256), 0, st, part, part_abs, npr, B,
                       tor_cert_depth(nm, npr), d_out, d_aux, cpart, npc);
  else if (!direct_real)
    hipLaunchKernelGGL((tor_reduce_kernel<false, KAPPA>), dim3(tor_red_grid(B)), dim3(256), 0, st, part, part_abs,
                       npr, B, d_out, d_aux, cpart, npc);
  rc = launch_tor<RC64, KAPPA, true>(nm, sc, d_in, B, cpart, cpart_abs, st, dc);
  if (rc) return rc;
  if (!direct_rerun)
    hipLaunchKernelGGL((tor_merge_kernel<KAPPA, CERT>), dim3(tor_red_grid(B)), dim3(256), 0, st, cpart, cpart_abs,
                       npc, B, d_out, d_aux);
  return (int)hipGetLastError();
}

static int run_tor_dd(int nm, const cplx* d_in, int B, cplx* d_out, void* ws, hipStream_t st) {
  const TorSplit sp = tor_batch_split<RDD>(nm);
  const int np = 1 << sp.T1;
  dd* part = reinterpret_cast<dd*>(ws);
  int rc = launch_tor<RDD, false, false>(nm, sp, d_in, B, part, nullptr, st);
  if (rc) return rc;
  hipLaunchKernelGGL((tor_reduce_kernel<true, false>), dim3(tor_red_grid(B)), dim3(256), 0, st, part, nullptr, np, B,
                     d_out, nullptr, nullptr, 0);
  return (int)hipGetLastError();
}

// certified dd: the dd tier's kernels with bounds (no complex fallback: the
// dd tiers are real-domain; off-domain instances are NaN with bound +inf)
static int run_tor_ddcert(int nm, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, hipStream_t st) {
  const TorSplit sp = tor_batch_split<RCDD>(nm);
  const int np = 1 << sp.T1;
  dd* part = reinterpret_cast<dd*>(ws);
  double* part_e = reinterpret_cast<double*>(part + (size_t)B * np);
  int rc = launch_tor<RCDD, false, false>(nm, sp, d_in, B, part, part_e, st);
  if (rc) return rc;
  // bound rounding depth: per mode <= 4 divisions / square roots and 8
  // products of ~12 bound operations each, the 2^R leaf sums, the butterfly
  // and the np-term reduce
  const double depth = 160.0 * nm + 64.0 + 96.0 + 4.0 * np + 64.0;
  hipLaunchKernelGGL(tor_reduce_ddcert_kernel, dim3(tor_red_grid(B)), dim3(256), 0, st, part, part_e, np, B, depth,
                     d_out, d_aux);
  return (int)hipGetLastError();
}

size_t tor_workspace_bytes(int tier, int dim, int B) {
  if (dim < 2 || (dim & 1) || dim > 2 * TOR_BATCH_MAX_MODES || B <= 0) return 0;
  if (tier == GBSK_DD || tier == GBSK_DD_CERTIFIED) {
    const TorSplit sp = tor_batch_split<RDD>(dim / 2);
    return (size_t)B * ((size_t)1 << sp.T1) * (sizeof(dd) + sizeof(double));
  }
  if (tier != GBSK_FP64 && tier != GBSK_KAPPA && tier != GBSK_CERTIFIED) return 0;
  const TorSplit sr = tor_batch_split<RF64>(dim / 2), sc = tor_split(dim / 2, RC64::RMAX);
  return (size_t)B * (((size_t)1 << sr.T1) + ((size_t)1 << sc.T1)) * (sizeof(cplx) + sizeof(double));
}

int tor_max_dim(int tier) {
  switch (tier) {
    case GBSK_FP64:
    case GBSK_DD:
    case GBSK_KAPPA:
    case GBSK_CERTIFIED:
    case GBSK_DD_CERTIFIED: return 2 * TOR_BATCH_MAX_MODES;
  }
  return 0;
}

int tor_launch(int tier, int dim, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, hipStream_t st) {
  if (B <= 0) return 0;
  if (dim < 0 || (dim & 1) || dim > 2 * TOR_BATCH_MAX_MODES) return GBSK_ERR_DIM;
  if (tier < GBSK_FP64 || tier > GBSK_DD_CERTIFIED) return GBSK_ERR_FUNC;
  const int nm = dim / 2;
  if (nm == 0) {  // tor of the empty matrix: exactly 1 (kappa 1, bound 0)
    const bool cert = tier == GBSK_CERTIFIED || tier == GBSK_DD_CERTIFIED;
    hipLaunchKernelGGL(tor_n0_kernel, dim3((B + 255) / 256), dim3(256), 0, st, B, d_out,
                       (tier == GBSK_KAPPA || cert) ? d_aux : nullptr, cert ? 0.0 : 1.0);
    return (int)hipGetLastError();
  }
  if (!ws) return GBSK_ERR_WORKSPACE;
  // at most 2^30 workgroups per launch: large batches run as consecutive
  // sub-batches on the same stream, reusing the workspace
  const TorSplit s64 = tor_batch_split<RF64>(nm), sc = tor_split(nm, RC64::RMAX), sdd = tor_batch_split<RDD>(nm);
  int t1 = s64.T1 > sc.T1 ? s64.T1 : sc.T1;
  t1 = sdd.T1 > t1 ? sdd.T1 : t1;
  const int64_t max_b = ((int64_t)1 << 30) >> t1;
  for (int64_t b0 = 0; b0 < B; b0 += max_b) {
    const int Bs = (int)((B - b0) < max_b ? (B - b0) : max_b);
    const cplx* in = d_in + b0 * dim * dim;
    cplx* out = d_out + b0;
    double* aux = d_aux ? d_aux + b0 : nullptr;
    int rc;
    if (tier == GBSK_DD)
      rc = run_tor_dd(nm, in, Bs, out, ws, st);
    else if (tier == GBSK_DD_CERTIFIED)
      rc = run_tor_ddcert(nm, in, Bs, out, aux, ws, st);
    else if (tier == GBSK_CERTIFIED)
      rc = run_tor_fp64<RCERT, false>(nm, in, Bs, out, aux, ws, st);
    else if (tier == GBSK_KAPPA)
      rc = run_tor_fp64<RF64, true>(nm, in, Bs, out, aux, ws, st);
    else
      rc = run_tor_fp64<RF64, false>(nm, in, Bs, out, nullptr, ws, st);
    if (rc) return rc;
  }
  return 0;
}

// ------------------------------------------------ one large real instance
// The single-large torontonian (reference core/tor_recursive.cu:148-678 via
// core/host_api.cu:262-407): the same kernels on one instance, dim <= 64, with
// a wider leading split (tor_single_split) so that one evaluation fills the
// GPU. Real path only: off the physical domain the value is NaN and a bound is
// +inf (no complex fallback).
int tor_schedule_get(int family, int* grouped_max, int* bfs_max) {
  if (family < 0 || family > 1) return GBSK_ERR_ARG;
  if (grouped_max) *grouped_max = tor_tune[family].grouped_max;
  if (bfs_max) *bfs_max = tor_tune[family].bfs_max;
  return 0;
}

int tor_schedule_set(int family, int grouped_max, int bfs_max) {
  if (family < 0 || family > 1 || grouped_max < 0 || grouped_max > 10 || bfs_max < 0 || bfs_max > TOR_MAX_MODES)
    return GBSK_ERR_ARG;
  tor_tune[family].grouped_max = grouped_max;
  tor_tune[family].bfs_max = bfs_max;
  return 0;
}

int tor_single_max_modes() { return TOR_MAX_MODES; }

static int tor_single_rmax(int tier) { return (tier == GBSK_DD || tier == GBSK_DD_CERTIFIED) ? RDD::RMAX : RF64::RMAX; }

size_t tor_single_workspace_bytes(int tier, int nm) {
  if (nm < 1 || nm > TOR_MAX_MODES) return 0;
  const TorSplit sp = tor_single_split(nm, tor_single_rmax(tier));
  const size_t el = (tier == GBSK_DD || tier == GBSK_DD_CERTIFIED) ? sizeof(dd) : sizeof(cplx);
  return ((size_t)1 << sp.T1) * (el + sizeof(double));
}

// Arithmetic-model self-check: RCERT::div and RCERT::sqrt_ charge one rounding
// of u|result| to the division and square root they evaluate, which holds only
// if those operations are correctly rounded (IEEE-754 round to nearest). The
// kernel evaluates exactly those value expressions on operand pairs so that a
// host can compare them bit for bit with correctly rounded results.
__global__ void tor_div_sqrt_probe_kernel(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                                          double* __restrict__ q, double* __restrict__ s) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    q[i] = RCERT::div(cv{a[i], 0.0}, cv{b[i], 0.0}).v;
    s[i] = RCERT::sqrt_(cv{fabs(a[i]), 0.0}).v;
  }
}

int tor_div_sqrt_probe(const double* d_a, const double* d_b, int64_t n, double* d_q, double* d_s, hipStream_t st) {
  if (n <= 0) return 0;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(tor_div_sqrt_probe_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, st,
                     d_a, d_b, n, d_q, d_s);
  return (int)hipGetLastError();
}

// The same for the reciprocal square root of the trailing walk's pivots
// (RF64::rsqrt = RCERT::rsqrt's value): RCERT charges it 2u|result|, which a
// host checks against an extended-precision 1 / sqrt(a).
__global__ void tor_rsqrt_probe_kernel(const double* __restrict__ a, int64_t n, double* __restrict__ r) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] = RCERT::rsqrt(cv{a[i], 0.0}).v;
}

int tor_rsqrt_probe(const double* d_a, int64_t n, double* d_r, hipStream_t st) {
  if (n <= 0) return 0;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(tor_rsqrt_probe_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, st,
                     d_a, n, d_r);
  return (int)hipGetLastError();
}

int tor_single_launch(int tier, int nm, const cplx* d_in, cplx* d_out, double* d_aux, void* ws, hipStream_t st) {
  if (nm < 0 || nm > TOR_MAX_MODES) return GBSK_ERR_DIM;
  if (tier != GBSK_FP64 && tier != GBSK_DD && tier != GBSK_CERTIFIED && tier != GBSK_DD_CERTIFIED)
    return GBSK_ERR_FUNC;
  const bool cert = tier == GBSK_CERTIFIED || tier == GBSK_DD_CERTIFIED;
  if (cert && !d_aux) return GBSK_ERR_ARG;
  if (nm == 0) {
    hipLaunchKernelGGL(tor_n0_kernel, dim3(1), dim3(64), 0, st, 1, d_out, cert ? d_aux : nullptr, 0.0);
    return (int)hipGetLastError();
  }
  if (!ws) return GBSK_ERR_WORKSPACE;
  const TorSplit sp = tor_single_split(nm, tor_single_rmax(tier));
  const int np = 1 << sp.T1;
  int rc;
  if (tier == GBSK_FP64 || tier == GBSK_CERTIFIED) {
    cplx* part = reinterpret_cast<cplx*>(ws);
    double* part_abs = reinterpret_cast<double*>(part + np);
    if (tier == GBSK_FP64) {
      rc = launch_tor<RF64, false, false>(nm, sp, d_in, 1, part, nullptr, st);
      if (!rc)
        hipLaunchKernelGGL((tor_reduce_kernel<false, false>), dim3(1), dim3(256), 0, st, part, nullptr, np, 1, d_out,
                           nullptr, nullptr, 0);
    } else {
      rc = launch_tor<RCERT, false, false>(nm, sp, d_in, 1, part, part_abs, st);
      if (!rc)
        hipLaunchKernelGGL(tor_reduce_cert_kernel, dim3(1), dim3(256), 0, st, part, part_abs, np, 1,
                           tor_cert_depth(nm, np), d_out, d_aux, nullptr, 0);
    }
  } else {
    dd* part = reinterpret_cast<dd*>(ws);
    double* part_e = reinterpret_cast<double*>(part + np);
    if (tier == GBSK_DD) {
      rc = launch_tor<RDD, false, false>(nm, sp, d_in, 1, part, nullptr, st);
      if (!rc)
        hipLaunchKernelGGL((tor_reduce_kernel<true, false>), dim3(1), dim3(256), 0, st, part, nullptr, np, 1, d_out,
                           nullptr, nullptr, 0);
    } else {
      rc = launch_tor<RCDD, false, false>(nm, sp, d_in, 1, part, part_e, st);
      const double depth = 160.0 * nm + 64.0 + 96.0 + 4.0 * np + 64.0;
      if (!rc)
        hipLaunchKernelGGL(tor_reduce_ddcert_kernel, dim3(1), dim3(256), 0, st, part, part_e, np, 1, depth, d_out,
                           d_aux);
    }
  }
  return rc ? rc : (int)hipGetLastError();
}

}  // namespace gbsk

// gbsk_oracle.cpp -- CPU oracle for the batched perm/haf/lhaf/tor path.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker, never the product: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
// The HIP library (gbskernels_amd/csrc) never links, calls or falls back to it.
//
// It restates the reference CPU algorithms op-for-op (citations are to the
// reference repository, path:line):
//   * Glynn / BB-FG Gray-code permanent ........ cpu_ref/permanent.py:108-139
//   * Glynn in double-double ..................... cpu_ref/permanent.py:142-189, cpu_ref/dd.py
//   * Ryser (cross-check) ........................ cpu_ref/permanent.py:76-105
//   * power-trace hafnian ........................ cpu_ref/hafnian.py:60-129
//   * power-trace loop hafnian (even N) .......... cpu_ref/loop_hafnian.py:65-108
//   * naive matchings (haf / lhaf, tiny N) ....... cpu_ref/hafnian.py:37-57, cpu_ref/loop_hafnian.py:44-62
//   * torontonian, subset determinants ........... cpu_ref/torontonian.py:48-64
//   * recursive prefix-Cholesky torontonian ...... cpu_ref/tor_recursive.py:46-101
//   * sum|term| indicators (kappa numerators) .... cpu_ref/diagnostics.py:30-170
//   * repeated-row loop hafnian (sieve) .......... cpu_ref/repeated.py:61-122
//   * DD twins of haf / lhaf / tor ............... core/hafnian_dd.cu:34-116,
//                                                  core/loop_hafnian_dd.cu:40-113,
//                                                  core/torontonian_dd.cu:23-96
// plus a __float128 ("quad") tier of Glynn, power-trace and subset-determinant
// evaluation that is not in the reference: it is this oracle's own high-precision
// check at sizes mpmath cannot reach.
//
// Layout: every matrix is row-major complex128 (re, im interleaved), exactly the
// numpy memory layout. Batched entry points run OpenMP over instances.
//
// Built with -ffp-contract=off so every rounding is the one written here.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include <quadmath.h>

namespace {

// ----------------------------------------------------------------- complex
struct cplx {
  double re, im;
};
inline cplx cadd(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
inline cplx csub(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
inline cplx cmul(cplx a, cplx b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline cplx cscale(cplx a, double s) { return {a.re * s, a.im * s}; }
inline double cabs_(cplx a) { return std::hypot(a.re, a.im); }
// numpy's complex division (Smith's algorithm as numpy implements it)
inline cplx cdiv(cplx a, cplx b) {
  double abr = std::fabs(b.re), abi = std::fabs(b.im);
  if (abr >= abi) {
    if (abr == 0.0 && abi == 0.0) return {a.re / abr, a.im / abi};
    double rat = b.im / b.re, scl = 1.0 / (b.re + b.im * rat);
    return {(a.re + a.im * rat) * scl, (a.im - a.re * rat) * scl};
  }
  double rat = b.re / b.im, scl = 1.0 / (b.im + b.re * rat);
  return {(a.re * rat + a.im) * scl, (a.im * rat - a.re) * scl};
}
// principal square root, the np.sqrt convention (branch cut on the negative axis)
inline cplx csqrt_(cplx z) {
  if (z.re == 0.0 && z.im == 0.0) return {0.0, z.im};
  double r = std::hypot(z.re, z.im);
  double t = std::sqrt(0.5 * (r + std::fabs(z.re)));
  if (z.re >= 0.0) return {t, z.im / (2.0 * t)};
  return {std::fabs(z.im) / (2.0 * t), std::copysign(t, z.im)};
}

// -------------------------------------------------------- double-double (cpu_ref/dd.py)
struct dd {
  double hi, lo;
};
inline dd two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double err = (a - (s - bb)) + (b - bb);
  return {s, err};
}
inline dd split(double a) {
  const double c = 134217729.0 * a;  // Dekker splitter 2^27 + 1
  double hi = c - (c - a);
  return {hi, a - hi};
}
inline dd two_prod(double a, double b) {
  double p = a * b;
  dd as = split(a), bs = split(b);
  double err = ((as.hi * bs.hi - p) + as.hi * bs.lo + as.lo * bs.hi) + as.lo * bs.lo;
  return {p, err};
}
inline dd quick_two_sum(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}
inline dd dd_add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  double e = s.lo + (a.lo + b.lo);
  return quick_two_sum(s.hi, e);
}
inline dd dd_neg(dd a) { return {-a.hi, -a.lo}; }
inline dd dd_sub(dd a, dd b) { return dd_add(a, dd_neg(b)); }
inline dd dd_mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  double e = p.lo + (a.hi * b.lo + a.lo * b.hi);
  return quick_two_sum(p.hi, e);
}
inline dd dd_mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  double e = p.lo + a.lo * b;
  return quick_two_sum(p.hi, e);
}
// QD-style division by a double and by a dd; dd sqrt by one Newton step
// (core/dd.cuh:121-164 states the same published algorithms).
inline dd dd_div_d(dd a, double b) {
  double q1 = a.hi / b;
  dd r = dd_sub(a, two_prod(q1, b));
  double q2 = r.hi / b;
  return quick_two_sum(q1, q2);
}
inline dd dd_div(dd a, dd b) {
  double q1 = a.hi / b.hi;
  dd r = dd_sub(a, dd_mul_d(b, q1));
  double q2 = r.hi / b.hi;
  r = dd_sub(r, dd_mul_d(b, q2));
  double q3 = r.hi / b.hi;
  dd q = quick_two_sum(q1, q2);
  return dd_add(q, dd{q3, 0.0});
}
inline dd dd_sqrt(dd a) {
  if (a.hi <= 0.0) return {0.0, 0.0};
  double x = 1.0 / std::sqrt(a.hi);
  dd y = dd_mul_d(a, x);
  dd diff = dd_sub(a, dd_mul(y, y));
  return dd_add(y, dd_mul_d(diff, 0.5 * x));
}
struct ddc {
  dd re, im;
};
inline ddc ddc_from(cplx z) { return {{z.re, 0.0}, {z.im, 0.0}}; }
inline ddc ddc_add(ddc a, ddc b) { return {dd_add(a.re, b.re), dd_add(a.im, b.im)}; }
inline ddc ddc_sub(ddc a, ddc b) { return {dd_sub(a.re, b.re), dd_sub(a.im, b.im)}; }
inline ddc ddc_mul(ddc a, ddc b) {
  dd rr = dd_sub(dd_mul(a.re, b.re), dd_mul(a.im, b.im));
  dd ii = dd_add(dd_mul(a.re, b.im), dd_mul(a.im, b.re));
  return {rr, ii};
}
inline ddc ddc_mul_d(ddc a, double s) { return {dd_mul_d(a.re, s), dd_mul_d(a.im, s)}; }
inline ddc ddc_div_d(ddc a, double s) { return {dd_div_d(a.re, s), dd_div_d(a.im, s)}; }
inline cplx ddc_to(ddc a) { return {a.re.hi + a.re.lo, a.im.hi + a.im.lo}; }

// ------------------------------------------------------------- quad complex
struct qc {
  __float128 re, im;
};
inline qc qadd(qc a, qc b) { return {a.re + b.re, a.im + b.im}; }
inline qc qsub(qc a, qc b) { return {a.re - b.re, a.im - b.im}; }
inline qc qmul(qc a, qc b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline qc qscale(qc a, __float128 s) { return {a.re * s, a.im * s}; }
inline qc qdiv(qc a, qc b) {
  __float128 d = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}
inline qc qsqrt(qc z) {
  __float128 r = sqrtq(z.re * z.re + z.im * z.im);
  if (r == 0) return {0, 0};
  __float128 t = sqrtq((r + fabsq(z.re)) / 2);
  if (z.re >= 0) return {t, z.im / (2 * t)};
  return {fabsq(z.im) / (2 * t), z.im < 0 ? -t : t};
}
inline qc qfrom(cplx z) { return {(__float128)z.re, (__float128)z.im}; }
inline cplx qto(qc z) { return {(double)z.re, (double)z.im}; }

inline const cplx* as_c(const double* p) { return reinterpret_cast<const cplx*>(p); }

inline int bit_length(uint64_t x) { return x ? 64 - __builtin_clzll(x) : 0; }

// ================================================================= permanent
// cpu_ref/permanent.py:108-139 (permanent_glynn)
cplx perm_glynn(const cplx* M, int n) {
  if (n == 0) return {1.0, 0.0};
  std::vector<cplx> rs(n);
  for (int r = 0; r < n; ++r) {
    cplx s{0.0, 0.0};
    for (int c = 0; c < n; ++c) s = cadd(s, M[r * n + c]);
    rs[r] = s;
  }
  auto prod = [&]() {
    cplx p = rs[0];
    for (int r = 1; r < n; ++r) p = cmul(p, rs[r]);
    return p;
  };
  cplx total = prod();
  int sign = 1;
  uint64_t prev = 0;
  const uint64_t terms = 1ull << (n - 1);
  for (uint64_t i = 1; i < terms; ++i) {
    uint64_t gray = i ^ (i >> 1);
    int k = bit_length(gray ^ prev) - 1;
    int col = k + 1;
    bool minus = (gray >> k) & 1ull;
    for (int r = 0; r < n; ++r) {
      cplx d = cscale(M[r * n + col], 2.0);
      rs[r] = minus ? csub(rs[r], d) : cadd(rs[r], d);
    }
    sign = -sign;
    cplx p = prod();
    total = sign > 0 ? cadd(total, p) : csub(total, p);
    prev = gray;
  }
  return cscale(total, 1.0 / (double)terms);
}

// cpu_ref/diagnostics.py:30-52 (glynn_abs_term_sum): sum over terms of prod |rowsum|
double perm_abs_terms(const cplx* M, int n) {
  if (n == 0) return 1.0;
  std::vector<cplx> rs(n);
  for (int r = 0; r < n; ++r) {
    cplx s{0.0, 0.0};
    for (int c = 0; c < n; ++c) s = cadd(s, M[r * n + c]);
    rs[r] = s;
  }
  auto aprod = [&]() {
    double p = 1.0;
    for (int r = 0; r < n; ++r) p *= cabs_(rs[r]);
    return p;
  };
  double acc = aprod();
  uint64_t prev = 0;
  const uint64_t terms = 1ull << (n - 1);
  for (uint64_t i = 1; i < terms; ++i) {
    uint64_t gray = i ^ (i >> 1);
    int k = bit_length(gray ^ prev) - 1;
    int col = k + 1;
    bool minus = (gray >> k) & 1ull;
    for (int r = 0; r < n; ++r) {
      cplx d = cscale(M[r * n + col], 2.0);
      rs[r] = minus ? csub(rs[r], d) : cadd(rs[r], d);
    }
    acc += aprod();
    prev = gray;
  }
  return acc / (double)terms;
}

// cpu_ref/permanent.py:142-189 (permanent_glynn_dd)
cplx perm_glynn_dd(const cplx* M, int n) {
  if (n == 0) return {1.0, 0.0};
  std::vector<ddc> rs(n);
  for (int r = 0; r < n; ++r) {
    ddc s = ddc_from({0.0, 0.0});
    for (int c = 0; c < n; ++c) s = ddc_add(s, ddc_from(M[r * n + c]));
    rs[r] = s;
  }
  auto prod = [&]() {
    ddc p = rs[0];
    for (int r = 1; r < n; ++r) p = ddc_mul(p, rs[r]);
    return p;
  };
  ddc total = prod();
  int sign = 1;
  uint64_t prev = 0;
  const uint64_t terms = 1ull << (n - 1);
  for (uint64_t i = 1; i < terms; ++i) {
    uint64_t gray = i ^ (i >> 1);
    int k = bit_length(gray ^ prev) - 1;
    int col = k + 1;
    double step = ((gray >> k) & 1ull) ? -2.0 : 2.0;
    for (int r = 0; r < n; ++r) rs[r] = ddc_add(rs[r], ddc_mul_d(ddc_from(M[r * n + col]), step));
    sign = -sign;
    ddc p = prod();
    total = sign > 0 ? ddc_add(total, p) : ddc_sub(total, p);
    prev = gray;
  }
  total = ddc_mul_d(total, 1.0 / (double)terms);
  return ddc_to(total);
}

// cpu_ref/permanent.py:76-105 (permanent_ryser), an independent cross-check
cplx perm_ryser(const cplx* M, int n) {
  if (n == 0) return {1.0, 0.0};
  std::vector<cplx> rs(n, cplx{0.0, 0.0});
  cplx total{0.0, 0.0};
  int sign = 1;
  uint64_t prev = 0;
  for (uint64_t i = 1; i < (1ull << n); ++i) {
    uint64_t gray = i ^ (i >> 1);
    int j = bit_length(gray ^ prev) - 1;
    bool in = (gray >> j) & 1ull;
    for (int r = 0; r < n; ++r) rs[r] = in ? cadd(rs[r], M[r * n + j]) : csub(rs[r], M[r * n + j]);
    sign = -sign;
    cplx p = rs[0];
    for (int r = 1; r < n; ++r) p = cmul(p, rs[r]);
    total = sign > 0 ? cadd(total, p) : csub(total, p);
    prev = gray;
  }
  if (n & 1) total = cscale(total, -1.0);
  return total;
}

// Glynn in __float128 (this oracle's own high-precision tier)
cplx perm_glynn_q(const cplx* M, int n) {
  if (n == 0) return {1.0, 0.0};
  std::vector<qc> rs(n);
  for (int r = 0; r < n; ++r) {
    qc s{0, 0};
    for (int c = 0; c < n; ++c) s = qadd(s, qfrom(M[r * n + c]));
    rs[r] = s;
  }
  auto prod = [&]() {
    qc p = rs[0];
    for (int r = 1; r < n; ++r) p = qmul(p, rs[r]);
    return p;
  };
  qc total = prod();
  int sign = 1;
  const uint64_t terms = 1ull << (n - 1);
  for (uint64_t i = 1; i < terms; ++i) {
    int k = __builtin_ctzll(i);
    int col = k + 1;
    bool minus = (((i ^ (i >> 1)) >> k) & 1ull) != 0;
    for (int r = 0; r < n; ++r) {
      qc d = qscale(qfrom(M[r * n + col]), 2);
      rs[r] = minus ? qsub(rs[r], d) : qadd(rs[r], d);
    }
    sign = -sign;
    qc p = prod();
    total = sign > 0 ? qadd(total, p) : qsub(total, p);
  }
  return qto(qscale(total, (__float128)1.0 / (__float128)terms));
}

// ================================================================= hafnian
// cpu_ref/hafnian.py:60-77 (exp_coeff_from_kg): [lambda^n] exp(sum g_k lambda^k)
template <class T, class Add, class Mul, class Div>
T exp_coeff(const std::vector<T>& kg, int n, T one, T zero, Add add, Mul mul, Div divj) {
  std::vector<T> e(n + 1, zero);
  e[0] = one;
  for (int j = 1; j <= n; ++j) {
    T acc = zero;
    for (int k = 1; k <= j; ++k) acc = add(acc, mul(kg[k], e[j - k]));
    e[j] = divj(acc, j);
  }
  return e[n];
}

// the pair index list of subset S (pairs (2i, 2i+1)) in increasing order
inline int pair_rows(uint64_t mask, int n, int* idx) {
  int s = 0;
  for (int i = 0; i < n; ++i)
    if ((mask >> i) & 1ull) {
      idx[s++] = 2 * i;
      idx[s++] = 2 * i + 1;
    }
  return s;
}

// Generic power-trace subset term over a scalar type T. `loop` selects the loop
// hafnian (cpu_ref/loop_hafnian.py:65-108); otherwise the ordinary hafnian with
// the diagonal zeroed (cpu_ref/hafnian.py:101-129). Returns the signed coefficient.
template <class T, class Ops>
T powertrace_term(const cplx* A, int N, int n, uint64_t mask, bool loop, const Ops& o) {
  int idx[64];
  int s = pair_rows(mask, n, idx);
  int m = s / 2;
  if (m == 0) {
    // empty subset: [lambda^n] exp(0) = (n == 0); loop hafnian skips it (n >= 1)
    return o.zero;
  }
  // C = B_S X: column c of C is column (c ^ 1) of B_S
  std::vector<T> C(s * s), P(s * s), Q(s * s), T2(s * s), d(s);
  for (int r = 0; r < s; ++r) {
    for (int c = 0; c < s; ++c) {
      int rr = idx[r], cc = idx[c ^ 1];
      cplx v = A[rr * N + cc];
      if (!loop && rr == cc) v = {0.0, 0.0};  // ordinary hafnian ignores the diagonal
      C[r * s + c] = o.from(v);
    }
    d[r] = o.from(A[idx[r] * N + idx[r]]);
  }
  std::vector<T> kg(n + 1, o.zero);
  P = C;
  for (int i = 0; i < s * s; ++i) Q[i] = o.zero;
  for (int i = 0; i < s; ++i) Q[i * s + i] = o.one;  // C^0
  for (int k = 1; k <= n; ++k) {
    T pk = o.zero;
    for (int i = 0; i < s; ++i) pk = o.add(pk, P[i * s + i]);
    T half_pk = o.half(pk);
    if (loop) {
      // v_k = (1/2) d^T X C^{k-1} d   (X swaps entries within each pair)
      std::vector<T> w(s, o.zero);
      for (int a = 0; a < s; ++a) {
        T acc = o.zero;
        for (int c = 0; c < s; ++c) acc = o.add(acc, o.mul(Q[a * s + c], d[c]));
        w[a] = acc;
      }
      T vk = o.zero;
      for (int a = 0; a < s; ++a) vk = o.add(vk, o.mul(d[a], w[a ^ 1]));
      vk = o.half(vk);
      kg[k] = o.add(half_pk, o.scale(vk, (double)k));
    } else {
      kg[k] = half_pk;
    }
    if (k < n) {
      if (loop) Q = P;
      for (int i = 0; i < s; ++i)
        for (int j = 0; j < s; ++j) {
          T acc = o.zero;
          for (int t = 0; t < s; ++t) acc = o.add(acc, o.mul(P[i * s + t], C[t * s + j]));
          T2[i * s + j] = acc;
        }
      P = T2;
    }
  }
  T coeff = exp_coeff(kg, n, o.one, o.zero, o.add, o.mul, o.divj);
  return ((n - m) & 1) ? o.neg(coeff) : coeff;
}

struct OpsD {
  cplx zero{0.0, 0.0}, one{1.0, 0.0};
  static cplx from(cplx v) { return v; }
  static cplx add(cplx a, cplx b) { return cadd(a, b); }
  static cplx mul(cplx a, cplx b) { return cmul(a, b); }
  static cplx half(cplx a) { return cscale(a, 0.5); }
  static cplx scale(cplx a, double s) { return cscale(a, s); }
  static cplx divj(cplx a, int j) { return {a.re / j, a.im / j}; }
  static cplx neg(cplx a) { return {-a.re, -a.im}; }
};
struct OpsDD {
  ddc zero{{0.0, 0.0}, {0.0, 0.0}}, one{{1.0, 0.0}, {0.0, 0.0}};
  static ddc from(cplx v) { return ddc_from(v); }
  static ddc add(ddc a, ddc b) { return ddc_add(a, b); }
  static ddc mul(ddc a, ddc b) { return ddc_mul(a, b); }
  static ddc half(ddc a) { return ddc_mul_d(a, 0.5); }
  static ddc scale(ddc a, double s) { return ddc_mul_d(a, s); }
  static ddc divj(ddc a, int j) { return ddc_div_d(a, (double)j); }
  static ddc neg(ddc a) { return {dd_neg(a.re), dd_neg(a.im)}; }
};
struct OpsQ {
  qc zero{0, 0}, one{1, 0};
  static qc from(cplx v) { return qfrom(v); }
  static qc add(qc a, qc b) { return qadd(a, b); }
  static qc mul(qc a, qc b) { return qmul(a, b); }
  static qc half(qc a) { return qscale(a, (__float128)0.5); }
  static qc scale(qc a, double s) { return qscale(a, (__float128)s); }
  static qc divj(qc a, int j) { return {a.re / j, a.im / j}; }
  static qc neg(qc a) { return {-a.re, -a.im}; }
};

// cpu_ref/hafnian.py:101-129 (hafnian_powertrace); loop: cpu_ref/loop_hafnian.py:65-108
template <class T, class Ops>
T powertrace_sum(const cplx* A, int N, bool loop, const Ops& o, double* abs_sum, double (*absf)(T)) {
  if (N == 0) {
    if (abs_sum) *abs_sum = 1.0;
    return o.one;
  }
  if (N & 1) {
    if (abs_sum) *abs_sum = 0.0;
    return o.zero;
  }
  int n = N / 2;
  T total = o.zero;
  double acc = 0.0;
  // cpu_ref enumerates subsets by size (itertools.combinations); the sum order
  // only moves the last bits, and the per-size order is kept here for fidelity.
  for (int m = 0; m <= n; ++m) {
    for (uint64_t mask = 0; mask < (1ull << n); ++mask) {
      if (__builtin_popcountll(mask) != m) continue;
      if (m == 0) {
        if (!loop) {
          // empty subset: (-1)^n [lambda^n] exp(0) = 0 for n >= 1
        }
        continue;
      }
      T t = powertrace_term<T>(A, N, n, mask, loop, o);
      total = o.add(total, t);
      if (abs_sum) acc += absf(t);
    }
  }
  if (abs_sum) *abs_sum = acc;
  return total;
}
double absd(cplx z) { return cabs_(z); }
double absdd(ddc z) { return cabs_(ddc_to(z)); }
double absq(qc z) { return cabs_(qto(z)); }

// cpu_ref/hafnian.py:37-57 (hafnian_naive) and cpu_ref/loop_hafnian.py:44-62
cplx matchings_rec(const cplx* M, int N, std::vector<int>& rem, bool loop) {
  if (rem.empty()) return {1.0, 0.0};
  int i = rem[0];
  std::vector<int> rest(rem.begin() + 1, rem.end());
  cplx total{0.0, 0.0};
  if (loop) total = cmul(M[i * N + i], matchings_rec(M, N, rest, loop));
  for (size_t k = 0; k < rest.size(); ++k) {
    std::vector<int> sub;
    for (size_t t = 0; t < rest.size(); ++t)
      if (t != k) sub.push_back(rest[t]);
    total = cadd(total, cmul(M[i * N + rest[k]], matchings_rec(M, N, sub, loop)));
  }
  return total;
}

// ============================================================== torontonian
// complex determinant by partial-pivot LU (numpy.linalg.det is LAPACK getrf)
cplx det_lu(std::vector<cplx>& M, int mm) {
  cplx det{1.0, 0.0};
  for (int c = 0; c < mm; ++c) {
    int piv = c;
    double best = cabs_(M[c * mm + c]);
    for (int r = c + 1; r < mm; ++r) {
      double v = cabs_(M[r * mm + c]);
      if (v > best) {
        best = v;
        piv = r;
      }
    }
    if (best == 0.0) return {0.0, 0.0};
    if (piv != c) {
      for (int j = 0; j < mm; ++j) std::swap(M[c * mm + j], M[piv * mm + j]);
      det = {-det.re, -det.im};
    }
    cplx pv = M[c * mm + c];
    det = cmul(det, pv);
    for (int r = c + 1; r < mm; ++r) {
      cplx f = cdiv(M[r * mm + c], pv);
      for (int j = c; j < mm; ++j) M[r * mm + j] = csub(M[r * mm + j], cmul(f, M[c * mm + j]));
    }
  }
  return det;
}

// cpu_ref/torontonian.py:48-64; abs: cpu_ref/diagnostics.py torontonian_abs_term_sum
cplx tor_direct(const cplx* O, int n, double* abs_sum) {
  if (n == 0) {
    if (abs_sum) *abs_sum = 1.0;
    return {1.0, 0.0};
  }
  const int N = 2 * n;
  cplx total{0.0, 0.0};
  double acc = 0.0;
  for (int m = 0; m <= n; ++m)
    for (uint64_t mask = 0; mask < (1ull << n); ++mask) {
      if (__builtin_popcountll(mask) != m) continue;
      cplx det{1.0, 0.0};
      if (m > 0) {
        int idx[128], c = 0;
        for (int i = 0; i < n; ++i)
          if ((mask >> i) & 1ull) idx[c++] = i;
        for (int i = 0; i < n; ++i)
          if ((mask >> i) & 1ull) idx[c++] = i + n;
        int mm = 2 * m;
        std::vector<cplx> sub(mm * mm);
        for (int r = 0; r < mm; ++r)
          for (int cc = 0; cc < mm; ++cc) {
            cplx v = O[idx[r] * N + idx[cc]];
            sub[r * mm + cc] = {(r == cc ? 1.0 : 0.0) - v.re, -v.im};
          }
        det = det_lu(sub, mm);
      }
      cplx t = cdiv({1.0, 0.0}, csqrt_(det));
      acc += cabs_(t);
      total = ((n - m) & 1) ? csub(total, t) : cadd(total, t);
    }
  if (abs_sum) *abs_sum = acc;
  return total;
}

// cpu_ref/tor_recursive.py:46-101 (prefix-Cholesky DFS, real domain). Returns NaN
// on a breakdown (non-positive pivot), where the Python raises.
double tor_recursive_real(const double* O, int n) {
  if (n == 0) return 1.0;
  const int dim = 2 * n;
  std::vector<double> L(dim * dim, 0.0);
  std::vector<int> modes(n);
  double total = 0.0;
  bool bad = false;
  auto append = [&](int j, int size, int count) -> double {
    for (int t = 0; t < 2; ++t) {
      int r = size + t;
      int gj = (t == 0) ? j : j + n;
      for (int c = 0; c <= r; ++c) {
        int mc = (c / 2 == count) ? j : modes[c / 2];
        int gc = (c % 2 == 0) ? mc : mc + n;
        double s = (gj == gc ? 1.0 : 0.0) - O[gj * dim + gc];
        double dot = 0.0;
        for (int u = 0; u < c; ++u) dot += L[r * dim + u] * L[c * dim + u];
        double acc = s - dot;
        if (c < r) {
          L[r * dim + c] = acc / L[c * dim + c];
        } else {
          if (!(acc > 0.0)) return NAN;
          L[r * dim + r] = std::sqrt(acc);
        }
      }
    }
    double a = L[size * dim + size] * L[(size + 1) * dim + size + 1];
    return a * a;
  };
  // explicit recursion as in the reference
  struct Rec {
    static void go(int mode, int count, double detprod, int n, std::vector<int>& modes,
                   double& total, bool& bad, decltype(append)& app) {
      if (bad) return;
      if (mode == n) {
        double t = 1.0 / std::sqrt(detprod);
        total += ((n - count) & 1) ? -t : t;
        return;
      }
      go(mode + 1, count, detprod, n, modes, total, bad, app);
      modes[count] = mode;
      double d = app(mode, 2 * count, count);
      if (!(d == d)) {
        bad = true;
        return;
      }
      go(mode + 1, count + 1, detprod * d, n, modes, total, bad, app);
    }
  };
  Rec::go(0, 0, 1.0, n, modes, total, bad, append);
  return bad ? NAN : total;
}

// real DD subset-determinant torontonian (core/torontonian_dd.cu:23-96)
double tor_dd_real(const cplx* O, int n) {
  if (n == 0) return 1.0;
  const int N = 2 * n;
  dd total{0.0, 0.0};
  for (uint64_t mask = 0; mask < (1ull << n); ++mask) {
    int m = __builtin_popcountll(mask);
    dd term{1.0, 0.0};
    if (m > 0) {
      int idx[128], c = 0;
      for (int i = 0; i < n; ++i)
        if ((mask >> i) & 1ull) idx[c++] = i;
      for (int i = 0; i < n; ++i)
        if ((mask >> i) & 1ull) idx[c++] = i + n;
      int mm = 2 * m;
      std::vector<dd> M(mm * mm);
      for (int r = 0; r < mm; ++r)
        for (int cc = 0; cc < mm; ++cc) {
          double v = O[idx[r] * N + idx[cc]].re;
          M[r * mm + cc] = (r == cc) ? two_sum(1.0, -v) : dd{-v, 0.0};
        }
      dd det{1.0, 0.0};
      for (int col = 0; col < mm; ++col) {
        int piv = col;
        double best = std::fabs(M[col * mm + col].hi);
        for (int r = col + 1; r < mm; ++r) {
          double v = std::fabs(M[r * mm + col].hi);
          if (v > best) {
            best = v;
            piv = r;
          }
        }
        if (best == 0.0) {
          det = {0.0, 0.0};
          break;
        }
        if (piv != col) {
          for (int j = 0; j < mm; ++j) std::swap(M[col * mm + j], M[piv * mm + j]);
          det = dd_neg(det);
        }
        dd pv = M[col * mm + col];
        det = dd_mul(det, pv);
        for (int r = col + 1; r < mm; ++r) {
          dd f = dd_div(M[r * mm + col], pv);
          for (int j = col; j < mm; ++j) M[r * mm + j] = dd_sub(M[r * mm + j], dd_mul(f, M[col * mm + j]));
        }
      }
      term = dd_div(dd{1.0, 0.0}, dd_sqrt(det));
    }
    total = ((n - m) & 1) ? dd_sub(total, term) : dd_add(total, term);
  }
  return total.hi + total.lo;
}

// subset-determinant torontonian in __float128 with the principal complex branch
cplx tor_q(const cplx* O, int n) {
  if (n == 0) return {1.0, 0.0};
  const int N = 2 * n;
  qc total{0, 0};
  for (uint64_t mask = 0; mask < (1ull << n); ++mask) {
    int m = __builtin_popcountll(mask);
    qc det{1, 0};
    if (m > 0) {
      int idx[128], c = 0;
      for (int i = 0; i < n; ++i)
        if ((mask >> i) & 1ull) idx[c++] = i;
      for (int i = 0; i < n; ++i)
        if ((mask >> i) & 1ull) idx[c++] = i + n;
      int mm = 2 * m;
      std::vector<qc> M(mm * mm);
      for (int r = 0; r < mm; ++r)
        for (int cc = 0; cc < mm; ++cc) {
          qc v = qfrom(O[idx[r] * N + idx[cc]]);
          M[r * mm + cc] = {(r == cc ? (__float128)1 : (__float128)0) - v.re, -v.im};
        }
      for (int col = 0; col < mm; ++col) {
        int piv = col;
        __float128 best = fabsq(M[col * mm + col].re) + fabsq(M[col * mm + col].im);
        for (int r = col + 1; r < mm; ++r) {
          __float128 v = fabsq(M[r * mm + col].re) + fabsq(M[r * mm + col].im);
          if (v > best) {
            best = v;
            piv = r;
          }
        }
        if (best == 0) {
          det = {0, 0};
          break;
        }
        if (piv != col) {
          for (int j = 0; j < mm; ++j) std::swap(M[col * mm + j], M[piv * mm + j]);
          det = {-det.re, -det.im};
        }
        qc pv = M[col * mm + col];
        det = qmul(det, pv);
        for (int r = col + 1; r < mm; ++r) {
          qc f = qdiv(M[r * mm + col], pv);
          for (int j = col; j < mm; ++j) M[r * mm + j] = qsub(M[r * mm + j], qmul(f, M[col * mm + j]));
        }
      }
    }
    qc t = qdiv(qc{1, 0}, qsqrt(det));
    total = ((n - m) & 1) ? qsub(total, t) : qadd(total, t);
  }
  return qto(total);
}

// ------------------------------------------------------- repeated-row sieve
// cpu_ref/repeated.py:61-122 (lhaf_repeated): loop hafnian of A expanded by the
// repetition counts n (row/column i repeated n_i times, loop weight gamma_i on
// each copy) by the finite-difference sieve
//   lhaf = (N! 2^N)^-1 sum_{v <= n} (-1)^|v| prod C(n_i, v_i) S_N(x_v),
//   x_v = n - 2v, S_N(x) = sum_k C(N, 2k) (2k-1)!! (x^T A x)^k (x . gamma)^(N-2k),
// visiting v in numpy.ndindex order (last index fastest), keeping v <= n - v
// lexicographically and doubling the strict half. R = double restates the
// reference's arithmetic (x @ M @ x as a row vector times x); R = __float128 is
// this oracle's high-precision twin.
template <class R>
struct CX {
  R re, im;
};
template <class R>
inline CX<R> cx_add(CX<R> a, CX<R> b) { return {a.re + b.re, a.im + b.im}; }
template <class R>
inline CX<R> cx_mul(CX<R> a, CX<R> b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
template <class R>
inline CX<R> cx_rscale(R s, CX<R> a) { return {s * a.re, s * a.im}; }

inline uint64_t binom_u64(int n, int k) {  // exact for n <= 64
  if (k < 0 || k > n) return 0;
  if (k > n - k) k = n - k;
  unsigned __int128 r = 1;
  for (int i = 1; i <= k; ++i) r = r * (unsigned)(n - k + i) / (unsigned)i;
  return (uint64_t)r;
}

template <class R>
CX<R> sieve_repeated(const cplx* A, const cplx* g, int M, const int* n) {
  int N = 0;
  for (int i = 0; i < M; ++i) N += n[i];
  if (N == 0) return {1, 0};
  const int kmax = N / 2;
  // C(N, 2k) (2k-1)!! (an integer; __float128 keeps it exact up to 2^113)
  std::vector<R> coeff(kmax + 1);
  for (int k = 0; k <= kmax; ++k) {
    __float128 df = 1;
    for (int t = 2 * k - 1; t > 0; t -= 2) df *= t;
    coeff[k] = (R)((__float128)binom_u64(N, 2 * k) * df);
  }
  std::vector<int> v(M, 0);
  std::vector<R> x(M);
  std::vector<CX<R>> y(M), mupow(kmax + 1);
  CX<R> total{0, 0};
  for (;;) {
    int cmp = 0;  // v against its mirror n - v, lexicographic from index 0
    for (int i = 0; i < M && !cmp; ++i) {
      const int d = v[i] - (n[i] - v[i]);
      if (d) cmp = d < 0 ? -1 : 1;
    }
    if (cmp <= 0) {
      int vs = 0;
      for (int i = 0; i < M; ++i) vs += v[i];
      R c = (vs & 1) ? R(-1) : R(1);
      if (cmp < 0) c *= 2;
      for (int i = 0; i < M; ++i) c *= (R)binom_u64(n[i], v[i]);
      for (int i = 0; i < M; ++i) x[i] = (R)(n[i] - 2 * v[i]);
      // sig2 = (x @ A) @ x, mu = x @ gamma
      for (int j = 0; j < M; ++j) {
        CX<R> s{0, 0};
        for (int i = 0; i < M; ++i) s = cx_add(s, cx_rscale(x[i], CX<R>{(R)A[i * M + j].re, (R)A[i * M + j].im}));
        y[j] = s;
      }
      CX<R> sig2{0, 0}, mu{0, 0};
      for (int j = 0; j < M; ++j) sig2 = cx_add(sig2, cx_rscale(x[j], y[j]));
      for (int i = 0; i < M; ++i) mu = cx_add(mu, cx_rscale(x[i], CX<R>{(R)g[i].re, (R)g[i].im}));
      mupow[kmax] = (N & 1) ? mu : CX<R>{1, 0};
      for (int k = kmax - 1; k >= 0; --k) mupow[k] = cx_mul(mupow[k + 1], cx_mul(mu, mu));
      CX<R> SN{0, 0}, sk{1, 0};
      for (int k = 0; k <= kmax; ++k) {
        SN = cx_add(SN, cx_mul(cx_rscale(coeff[k], sk), mupow[k]));
        sk = cx_mul(sk, sig2);
      }
      total = cx_add(total, cx_rscale(c, SN));
    }
    int i = M - 1;  // advance, last index fastest
    while (i >= 0 && v[i] == n[i]) v[i--] = 0;
    if (i < 0) break;
    ++v[i];
  }
  __float128 f = 1;
  for (int t = 2; t <= N; ++t) f *= t;
  const R D = (R)f * (R)ldexpq((__float128)1, N);
  return {total.re / D, total.im / D};
}

template <class F>
void for_batch(int B, F f) {
#pragma omp parallel for schedule(dynamic, 1)
  for (int b = 0; b < B; ++b) f(b);
}

}  // namespace

extern "C" {

// Every entry point: `in` is B row-major d x d complex128 matrices (d = n for perm,
// N for haf/lhaf, 2n for tor); `out` B complex128 (re, im); `aux` B doubles.

void or_perm(const double* in, int n, int B, double* out) {
  for_batch(B, [&](int b) {
    cplx v = perm_glynn(as_c(in) + (size_t)b * n * n, n);
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}
void or_perm_abs(const double* in, int n, int B, double* aux) {
  for_batch(B, [&](int b) { aux[b] = perm_abs_terms(as_c(in) + (size_t)b * n * n, n); });
}
void or_perm_dd(const double* in, int n, int B, double* out) {
  for_batch(B, [&](int b) {
    cplx v = perm_glynn_dd(as_c(in) + (size_t)b * n * n, n);
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}
void or_perm_ryser(const double* in, int n, int B, double* out) {
  for_batch(B, [&](int b) {
    cplx v = perm_ryser(as_c(in) + (size_t)b * n * n, n);
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}
void or_perm_q(const double* in, int n, int B, double* out) {
  for_batch(B, [&](int b) {
    cplx v = perm_glynn_q(as_c(in) + (size_t)b * n * n, n);
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}

// tier: 0 fp64, 1 dd, 2 quad. loop: 0 haf, 1 lhaf (even N). aux (may be null):
// sum |term| at the fp64 tier.
void or_powertrace(const double* in, int N, int B, int loop, int tier, double* out, double* aux) {
  for_batch(B, [&](int b) {
    const cplx* A = as_c(in) + (size_t)b * N * N;
    cplx v;
    if (tier == 0) {
      OpsD o;
      v = powertrace_sum<cplx>(A, N, loop != 0, o, aux ? aux + b : nullptr, absd);
    } else if (tier == 1) {
      OpsDD o;
      v = ddc_to(powertrace_sum<ddc>(A, N, loop != 0, o, nullptr, absdd));
    } else {
      OpsQ o;
      v = qto(powertrace_sum<qc>(A, N, loop != 0, o, nullptr, absq));
    }
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}

// literal matchings (any N; loop adds self-loops)
void or_matchings(const double* in, int N, int B, int loop, double* out) {
  for_batch(B, [&](int b) {
    const cplx* A = as_c(in) + (size_t)b * N * N;
    cplx v;
    if (!loop && (N & 1)) {
      v = {0.0, 0.0};
    } else {
      std::vector<int> rem(N);
      for (int i = 0; i < N; ++i) rem[i] = i;
      v = matchings_rec(A, N, rem, loop != 0);
    }
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}

void or_tor(const double* in, int n, int B, double* out, double* aux) {
  for_batch(B, [&](int b) {
    cplx v = tor_direct(as_c(in) + (size_t)b * 4 * n * n, n, aux ? aux + b : nullptr);
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}
// real O given as B row-major (2n)^2 doubles
void or_tor_recursive(const double* in, int n, int B, double* out) {
  for_batch(B, [&](int b) { out[b] = tor_recursive_real(in + (size_t)b * 4 * n * n, n); });
}
void or_tor_dd(const double* in, int n, int B, double* out) {
  for_batch(B, [&](int b) {
    out[2 * b] = tor_dd_real(as_c(in) + (size_t)b * 4 * n * n, n);
    out[2 * b + 1] = 0.0;
  });
}
void or_tor_q(const double* in, int n, int B, double* out) {
  for_batch(B, [&](int b) {
    cplx v = tor_q(as_c(in) + (size_t)b * 4 * n * n, n);
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}

// repeated-row sieve: A (M x M), gamma (M), reps (B x M int32); quad: 0 double, 1 __float128
void or_lhaf_repeated(const double* A, const double* g, int M, const int* reps, int B, int quad, double* out) {
  for_batch(B, [&](int b) {
    cplx v;
    if (quad) {
      CX<__float128> r = sieve_repeated<__float128>(as_c(A), as_c(g), M, reps + (size_t)b * M);
      v = {(double)r.re, (double)r.im};
    } else {
      CX<double> r = sieve_repeated<double>(as_c(A), as_c(g), M, reps + (size_t)b * M);
      v = {r.re, r.im};
    }
    out[2 * b] = v.re;
    out[2 * b + 1] = v.im;
  });
}

}  // extern "C"

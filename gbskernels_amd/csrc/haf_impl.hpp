// haf_impl.hpp -- power-trace hafnian and loop hafnian on gfx950 (fp64, kappa, dd).
//
// Algorithm (reference: core/hafnian.cu:41-133, core/loop_hafnian.cu:45-139,
// cpu_ref/hafnian.py:60-129, cpu_ref/loop_hafnian.py:65-108):
//   haf(A) = sum_{S subset of the n index pairs} (-1)^(n-|S|) [x^n] exp(sum_k g_k x^k)
// with C = B_S X (B_S the principal submatrix on the pairs of S, X the swap
// inside each pair), k g_k = tr(C^k)/2 (+ k v_k, v_k = 1/2 d^T X C^(k-1) d for
// the loop hafnian, d = diag B_S), and the coefficient from the Newton recurrence
// e_j = (1/j) sum_{k<=j} (k g_k) e_{j-k}. The ordinary hafnian ignores the
// diagonal, which is zeroed on load (cpu_ref/hafnian.py:115).
//
// MI355X design (this file's own):
//  * one kernel per (subset size s = 2m, tier): all subsets of one size form a
//    phase; a wave holds one instance's matrix in LDS and walks its C(n, m)
//    subsets, G = floor(64 / s) at a time, one lane group of s lanes per subset,
//    lane r owning row r of the current power C^j in VGPRs;
//  * powers only up to C^h, h = ceil(n/2): since B_S is symmetric,
//    (C^j)^T = X C^j X, so tr(C^i C^j) = sum_ab (C^j)_ab (C^i)_(a^1)(b^1) pairs
//    lane a with lane a^1 (a DPP swap) and gives p_{2j-1} and p_{2j} from the
//    rows of C^(j-1) and C^j: h-1 matrix products instead of n-1;
//  * the loop term uses the Krylov vectors w_k = C^(k-1) d (one matvec each);
//  * subset terms are summed per lane group in subset order, groups and phases
//    in a fixed order by a final pass: the value depends on N alone.
#pragma once

#include <type_traits>
#include <utility>

#include "cert.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace gbsk {

constexpr int HAF_MAX_N = 28;
constexpr int HAF_MAX_PAIRS = HAF_MAX_N / 2;

// ------------------------------------------------------------ arithmetic policies
struct ArF64 {
  using T = cplx;
  using E = cplx;  // staged matrix entry
  __device__ static T zero() { return cplx{0.0, 0.0}; }
  __device__ static T one() { return cplx{1.0, 0.0}; }
  __device__ static T from(cplx z) { return z; }
  __device__ static T add(T a, T b) { return c_add(a, b); }
  __device__ static T sub(T a, T b) { return c_sub(a, b); }
  __device__ static T neg(T a) { return c_neg(a); }
  __device__ static T mul(T a, T b) { return c_mul(a, b); }
  // acc + a * z, z a plain complex input entry
  __device__ static T fma_c(T a, cplx z, T acc) { return c_fma(a, z, acc); }
  __device__ static T fma(T a, T b, T acc) { return c_fma(a, b, acc); }
  __device__ static T scale(T a, double s) { return c_scale(a, s); }
  __device__ static T div_int(T a, int j) { return cplx{a.re / j, a.im / j}; }
  __device__ static cplx to_c(T a) { return a; }
  __device__ static T shfl_xor(T a, int m) { return cplx{shfl_xor_d(a.re, m), shfl_xor_d(a.im, m)}; }
  __device__ static T swap_pair(T a) { return cplx{swap_pair_d(a.re), swap_pair_d(a.im)}; }
  template <int CTRL>
  __device__ static T dpp(T a) {
    return cplx{dpp_d<CTRL>(a.re), dpp_d<CTRL>(a.im)};
  }
};

// Real fp64 arithmetic for matrices whose imaginary parts are all zero (the
// fused small-size kernel picks it per workgroup). Each operation is the real
// part of ArF64's on such operands: c_fma's real part is
// fma(x.re, y.re, fma(-x.im, y.im, a.re)) and the inner fma adds a zero product
// to a.re, c_add / c_scale / div_int act per component, so the values agree
// with ArF64's (an exact zero can at most differ in sign) and every imaginary
// part stays zero.
struct ArR64 {
  using T = double;
  using E = double;
  __device__ static T zero() { return 0.0; }
  __device__ static T one() { return 1.0; }
  __device__ static T from(double z) { return z; }
  __device__ static T add(T a, T b) { return a + b; }
  __device__ static T neg(T a) { return -a; }
  __device__ static T fma_c(T a, double z, T acc) { return fma(a, z, acc); }
  __device__ static T fma(T a, T b, T acc) { return ::fma(a, b, acc); }
  __device__ static T scale(T a, double s) { return a * s; }
  __device__ static T div_int(T a, int j) { return a / j; }
  __device__ static cplx to_c(T a) { return cplx{a, 0.0}; }
  __device__ static T shfl_xor(T a, int m) { return shfl_xor_d(a, m); }
  __device__ static T swap_pair(T a) { return swap_pair_d(a); }
  template <int CTRL>
  __device__ static T dpp(T a) {
    return dpp_d<CTRL>(a);
  }
};

__device__ inline dd dd_mul_d_acc(dd a, double b, dd acc) { return dd_add(acc, dd_mul_d(a, b)); }

struct ArDD {
  using T = ddc;
  using E = cplx;
  __device__ static T zero() { return ddc_zero(); }
  __device__ static T one() { return ddc{dd{1.0, 0.0}, dd{0.0, 0.0}}; }
  __device__ static T from(cplx z) { return ddc_from(z); }
  __device__ static T add(T a, T b) { return ddc_add(a, b); }
  __device__ static T sub(T a, T b) { return ddc_sub(a, b); }
  __device__ static T neg(T a) { return ddc_neg(a); }
  __device__ static T mul(T a, T b) { return ddc_mul(a, b); }
  __device__ static T fma_c(T a, cplx z, T acc) {
    dd re = dd_sub(dd_mul_d(a.re, z.re), dd_mul_d(a.im, z.im));
    dd im = dd_add(dd_mul_d(a.re, z.im), dd_mul_d(a.im, z.re));
    return ddc{dd_add(acc.re, re), dd_add(acc.im, im)};
  }
  __device__ static T fma(T a, T b, T acc) { return ddc_add(acc, ddc_mul(a, b)); }
  __device__ static T scale(T a, double s) { return ddc_mul_d(a, s); }
  __device__ static T div_int(T a, int j) { return ddc_div_d(a, (double)j); }
  __device__ static cplx to_c(T a) { return ddc_to_c(a); }
  __device__ static T shfl_xor(T a, int m) {
    return ddc{dd{shfl_xor_d(a.re.hi, m), shfl_xor_d(a.re.lo, m)}, dd{shfl_xor_d(a.im.hi, m), shfl_xor_d(a.im.lo, m)}};
  }
  __device__ static T swap_pair(T a) {
    return ddc{dd{swap_pair_d(a.re.hi), swap_pair_d(a.re.lo)}, dd{swap_pair_d(a.im.hi), swap_pair_d(a.im.lo)}};
  }
  template <int CTRL>
  __device__ static T dpp(T a) {
    return ddc{dd{dpp_d<CTRL>(a.re.hi), dpp_d<CTRL>(a.re.lo)}, dd{dpp_d<CTRL>(a.im.hi), dpp_d<CTRL>(a.im.lo)}};
  }
};

// binomial coefficient, n <= 14
__device__ inline int binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

// q-th (lexicographic) m-subset of {0..n-1} -> pair list p[0..m-1]
template <int M>
__device__ inline void unrank(int q, int n, int (&p)[M]) {
  int x = 0;
#pragma unroll
  for (int i = 0; i < M; ++i) {
    for (;;) {
      const int c = binom(n - 1 - x, M - 1 - i);
      if (c <= q) {
        q -= c;
        ++x;
      } else {
        break;
      }
    }
    p[i] = x++;
  }
}

// LDS reduction over the S lanes of each group, fixed order; every lane of the
// group receives the same sum. All 64 lanes must call it (barriers).
template <class Ar, int S>
__device__ inline typename Ar::T group_reduce(typename Ar::T v, typename Ar::T* red) {
  if constexpr ((std::is_same<Ar, ArF64>::value || std::is_same<Ar, ArR64>::value) && (S == 2 || S == 4)) {
    // groups inside one quad: each group value arrives by a DPP quad broadcast
    // (no LDS round trip, no barrier); the same adds in the same order
    (void)red;
    if constexpr (S == 2) {
      return Ar::add(Ar::template dpp<0xA0>(v), Ar::template dpp<0xF5>(v));
    } else {
      const typename Ar::T s = Ar::add(Ar::template dpp<0x00>(v), Ar::template dpp<0x55>(v));
      return Ar::add(Ar::add(s, Ar::template dpp<0xAA>(v)), Ar::template dpp<0xFF>(v));
    }
  }
  const int lane = threadIdx.x;
  red[lane] = v;
  __syncthreads();
  const int base = (lane / S < 64 / S) ? (lane / S) * S : 0;
  typename Ar::T s = red[base];
#pragma unroll
  for (int i = 1; i < S; ++i) s = Ar::add(s, red[base + i]);
  __syncthreads();
  return s;
}

// ------------------------------------------------------------ certified parts
// (fp64 only; cert.hpp states the error model). Bounds are 1-norm bounds of the
// error of the computed complex value relative to the exact value of the same
// formula on the exact inputs.

enum HafMode { HM_PLAIN = 0, HM_KAPPA = 1, HM_CERT = 2, HM_DDCERT = 3 };

// gamma_k = k u / (1 - k u) <= k u (1 + 1e-6) for k u < 1e-7 (k <= 2 HAF_MAX_N here)
__device__ inline double cert_gamma(int k) { return (double)k * CERT_U * 1.000001; }

// bound of r = c_fma(x, y, a): propagated input errors plus the four roundings,
// each at most u times the magnitude it rounds (see common.hpp c_fma)
__device__ inline double cert_fma(cplx x, double ex, cplx y, double ey, cplx a, double ea, cplx r) {
  const double X = cert_norm1(x), Y = cert_norm1(y);
  double t = fma(X, ey, ea + CERT_UFL);
  t = fma(Y, ex, t);
  t = fma(ex, ey, t);
  return fma(CERT_U, cert_norm1(a) + fma(X, Y, cert_norm1(r)), t);
}

// LDS reduction of (value, bound) over the S lanes of each group, same order as
// group_reduce (so the value is bit-identical to the plain kernel's)
template <int S>
__device__ inline void cert_group_reduce(cplx& v, double& e, cplx* red, double* rede) {
  const int lane = threadIdx.x;
  red[lane] = v;
  rede[lane] = e;
  __syncthreads();
  const int base = (lane / S < 64 / S) ? (lane / S) * S : 0;
  cplx s = red[base];
  double E = rede[base];
#pragma unroll
  for (int i = 1; i < S; ++i) {
    s = c_add(s, red[base + i]);
    E = cert_round(s, E + rede[base + i]);
  }
  __syncthreads();
  v = s;
  e = E;
}

// certified dd (HM_DDCERT, Ar = ArDD): the same reduction with each dd
// addition's own error (cert.hpp) added to the operands' bounds
template <int S>
__device__ inline void ddcert_group_reduce(ddc& v, double& e, ddc* red, double* rede) {
  const int lane = threadIdx.x;
  red[lane] = v;
  rede[lane] = e;
  __syncthreads();
  const int base = (lane / S < 64 / S) ? (lane / S) * S : 0;
  ddc s = red[base];
  double E = rede[base];
#pragma unroll
  for (int i = 1; i < S; ++i) {
    double r;
    s = cddc_add(s, red[base + i], &r);
    E = E + rede[base + i] + r;
  }
  __syncthreads();
  v = s;
  e = E;
}

// Work split of a phase: its C(n, m) subsets are cut into chunks of
// HAF_CHUNK_ITERS * G consecutive subsets (G = 64 / S per wave iteration), one
// workgroup per (instance, chunk), so that large N runs many waves per instance.
// The split depends on (N, S) only; chunk partials are summed in chunk order.
constexpr int HAF_CHUNK_ITERS = 1;
__host__ __device__ inline int haf_chunk_size(int S) { return HAF_CHUNK_ITERS * (64 / S); }
// a / b for 0 <= a < 2^16, 1 <= b <= 2^10, through one fp32 reciprocal: the
// exact quotient (a + 1/2) / b lies at least 1/(2b) from an integer and the
// fp32 product is within (a + 1/2) / b * 2^-21 < 2^-5 / b of it, so the
// truncation is floor(a / b) (a few instructions where an integer division by
// a runtime divisor takes about twenty)
__device__ __forceinline__ int small_div(int a, int b) {
  return (int)(((float)a + 0.5f) * __builtin_amdgcn_rcpf((float)b));
}
__host__ __device__ constexpr int haf_cbinom(int n, int k) {
  return (k < 0 || k > n) ? 0 : (k == 0 ? 1 : haf_cbinom(n - 1, k - 1) * n / k);
}
__host__ __device__ inline int haf_binom(int n, int k) {
  if (k < 0 || k > n) return 0;
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}
__host__ __device__ inline int haf_nchunks(int n, int S) {
  const int c = haf_binom(n, S / 2), ch = haf_chunk_size(S);
  return (c + ch - 1) / ch;
}
// partial slots per instance (all phases) and the first slot of phase S
__host__ __device__ inline int haf_nparts(int n) {
  int t = 0;
  for (int m = 1; m <= n; ++m) t += haf_nchunks(n, 2 * m);
  return t;
}
__host__ __device__ inline int haf_part_offset(int n, int S) {
  int t = 0;
  for (int m = 1; m < S / 2; ++m) t += haf_nchunks(n, 2 * m);
  return t;
}

// first term slot of subset size S when an instance's terms are stored size
// by size (all C(n, 1) subsets of size 2, then the C(n, 2) of size 4, ...)
__host__ __device__ inline int haf_small_term_offset(int n, int S) {
  int t = 0;
  for (int m = 1; m < S / 2; ++m) t += haf_binom(n, m);
  return t;
}

// rounding depth of a phase's bound arithmetic (generous; see cert_finish)
__device__ inline double haf_cert_depth(int S, int n, int count) {
  return 64.0 * (n + 1) * (S + 8) + 8.0 * count + 16.0 * n * n + 256.0;
}

// Launch descriptor of one phase over several buckets (kernel argument, by
// value): bucket j owns workgroups [wg_begin[j], wg_begin[j+1]).
template <class T>
struct HafPhaseJobs {
  const cplx* mats[HAF_MAX_JOBS];
  T* part[HAF_MAX_JOBS];
  double* part_abs[HAF_MAX_JOBS];
  int N[HAF_MAX_JOBS];
  int B[HAF_MAX_JOBS];
  int wg_begin[HAF_MAX_JOBS + 1];
  int n;
};

// LDS scratch of one wave's subset evaluations at subset size S (every lane
// group g of S lanes owns row g of the per-group arrays)
// (PMAX: the largest pair count n it serves, plus one for the odd-n powers)
template <class Ar, bool LOOP, int S, int MODE, int PMAX = HAF_MAX_PAIRS>
struct HafScratch {
  using T = typename Ar::T;
  static constexpr bool ANYC = (MODE == HM_CERT || MODE == HM_DDCERT);
  static constexpr int G = 64 / S;
  static constexpr int P = PMAX;
  T red[64];
  T kgs[G][PMAX + 1];  // k g_k of each group's current subset
  T es[G][PMAX + 1];   // Newton coefficients e_0..e_n
  double rede[ANYC ? 64 : 1];
  T kvs[LOOP ? G : 1][PMAX + 2];  // k v_k (loop hafnian)
  double kve[LOOP && ANYC ? G : 1][PMAX + 2];
  double kge[ANYC ? G : 1][PMAX + 1];
  double ese[ANYC ? G : 1][PMAX + 1];
};

// ------------------------------------------------------------- one subset term
// Every lane of the wave calls it (the group reductions hold barriers). Lane
// group g = lane / S (g < 64 / S) evaluates subset q (lexicographic m-subset of
// the n = N / 2 index pairs, m = S / 2) of the N x N matrix at As, which may
// differ from group to group; lane r of the group owns row r of the powers
// C^j. On return lane r == 0 of each group holds the subset's term
// (-1)^(n - m) [x^n] exp(sum_k g_k x^k) in coeff and, in the certified modes,
// its bound in ecoeff. The arithmetic depends on (As contents, N, q) only.
template <class Ar, bool LOOP, int S, int MODE, class Scratch>
__device__ __forceinline__ void haf_subset_term(const typename Ar::E* __restrict__ As0, int ab, int N, int q, Scratch& sh,
                                                typename Ar::T& coeff_out, double& ecoeff_out) {
  // As0 is the workgroup's LDS array and ab this group's matrix offset in it
  // (an integer offset keeps every access a plain LDS address)
  using E = typename Ar::E;
  const E* __restrict__ As = As0 + ab;
  using T = typename Ar::T;
  constexpr bool KAPPA = (MODE == HM_KAPPA);
  constexpr bool CERT = (MODE == HM_CERT);
  constexpr bool DDC = (MODE == HM_DDCERT);  // certified double-double (Ar = ArDD)
  constexpr bool ANYC = CERT || DDC;
  constexpr int M = S / 2;
  constexpr int G = 64 / S;
  (void)KAPPA;
  T* red = sh.red;
  auto& kgs = sh.kgs;
  auto& es = sh.es;
  double* rede = sh.rede;
  auto& kvs = sh.kvs;
  auto& kve = sh.kve;
  auto& kge = sh.kge;
  auto& ese = sh.ese;
  (void)kve;
  (void)kge;
  (void)ese;
  (void)rede;
  const int lane = threadIdx.x;
  const int n = N / 2;
  const int g = lane / S;
  const int r = lane % S;
  const bool lane_ok = g < G;
  const int h = (n + 1) / 2;
  const double gam = CERT ? cert_gamma(2 * S) : 0.0;
  (void)gam;
  int p[M];
  unrank<M>(q, n, p);
  // lane r owns row r of C = B_S X: C[r][2u] = A[R][2p_u + 1], C[r][2u+1] = A[R][2p_u]
  const int R = 2 * p[r >> 1] + (r & 1);
  T P[S];
  double eP[ANYC ? S : 1];
#pragma unroll
  for (int u = 0; u < M; ++u) {
    P[2 * u] = Ar::from(As[__mul24(R, N) + 2 * p[u] + 1]);
    P[2 * u + 1] = Ar::from(As[__mul24(R, N) + 2 * p[u]]);
  }
  if (ANYC) {
#pragma unroll
    for (int c = 0; c < S; ++c) eP[c] = 0.0;  // C is exact
  }
  // Loop hafnian: k v_k = k/2 d^T X C^(k-1) d with d = diag B_S. With
  // u_j = C^j d (lane r: row r of C^j times d, no lane exchange) and
  // (C^a)^T = X C^a X, d^T X C^(a+b) d = sum_r u_a[r^1] u_b[r], so every v_k
  // up to k - 1 = 2h comes from u_0 .. u_h with a = floor((k-1)/2),
  // b = ceil((k-1)/2): v_{2j}, v_{2j+1} at power j. Stored apart (kvs) and
  // added to the trace part p_k / 2 before the Newton recurrence.
  auto dot_d = [&](const T (&row)[S], const double (&erow)[ANYC ? S : 1], T& u, double& eu) {
    // d is re-read from LDS on every call: the opaque zero offset stops the
    // compiler from hoisting the S loop-invariant diagonal entries (4 S
    // VGPRs) out of the power loop, which halved the loop kernels' occupancy
    int zo = 0;
    asm volatile("" : "+v"(zo));
    const E* __restrict__ Ad = As + zo;
    u = Ar::zero();
    eu = CERT ? (double)S * CERT_UFL : 0.0;
#pragma unroll
    for (int c = 0; c < S; ++c) {
      const int Rc = 2 * p[c >> 1] + (c & 1);
      const E dc = Ad[__mul24(Rc, N + 1)];
      if constexpr (DDC) {
        double o;
        u = cddc_fma_c(row[c], dc, u, &o);
        eu = fma(erow[c], cert_norm1(dc), eu + o);
      } else {
        if constexpr (CERT) eu = fma(fma(gam, cert_norm1(Ar::to_c(row[c])), erow[c]), cert_norm1(dc), eu);
        u = Ar::fma_c(row[c], dc, u);
      }
    }
  };
  auto add_v = [&](int k, const T& ua, double eua, const T& ub, double eub) {
    if (k > n) return;  // uniform over the wave (reductions hold barriers)
    const T x = Ar::swap_pair(ua);
    const double ex = ANYC ? swap_pair_d(eua) : 0.0;
    const double hk = 0.5 * (double)k;
    if constexpr (DDC) {
      double o;
      T v = cddc_fma(x, ub, Ar::zero(), &o);
      double ev = cddc_mul_bound(x, ex, ub, eub, o);
      ddcert_group_reduce<S>(v, ev, red, rede);
      if (r == 0 && lane_ok) {
        double os;
        kvs[g][k] = cddc_scale(v, hk, &os);
        kve[g][k] = fma(hk, ev, os);
      }
    } else if constexpr (CERT) {
      T v = Ar::fma(x, ub, Ar::zero());
      double ev = cert_fma(Ar::to_c(x), ex, Ar::to_c(ub), eub, cplx{0.0, 0.0}, 0.0, Ar::to_c(v));
      cert_group_reduce<S>(v, ev, red, rede);
      if (r == 0 && lane_ok) {
        const T kv = Ar::scale(v, hk);
        kvs[g][k] = kv;
        kve[g][k] = cert_round(Ar::to_c(kv), hk * ev);
      }
    } else {
      const T v = group_reduce<Ar, S>(Ar::fma(x, ub, Ar::zero()), red);
      if (r == 0 && lane_ok) kvs[g][k] = Ar::scale(v, hk);
    }
  };
  T uprev = Ar::zero();  // u_{j-1}
  double euprev = 0.0;
  if constexpr (LOOP) {
    const T u0 = Ar::from(As[__mul24(R, N + 1)]);
    T u1;
    double eu1;
    dot_d(P, eP, u1, eu1);
    add_v(1, u0, 0.0, u0, 0.0);
    add_v(2, u0, 0.0, u1, eu1);
    add_v(3, u1, eu1, u1, eu1);
    uprev = u1;
    euprev = eu1;
  }
  // p_1 = tr C (C[r][r] = A[R][R^1]) and p_2 = tr(C C) through the partner row
  {
    T d1 = Ar::from(As[__mul24(R, N) + (R ^ 1)]);
    T t2 = Ar::zero();
    double e2 = 0.0;
#pragma unroll
    for (int c = 0; c < S; ++c) {
      const T y = Ar::swap_pair(P[c ^ 1]);
      if constexpr (DDC) {
        double eo;
        t2 = cddc_fma(P[c], y, t2, &eo);  // exact operands
        e2 += eo;
      } else {
        const T t2n = Ar::fma(P[c], y, t2);
        if constexpr (CERT) e2 = cert_fma(P[c], 0.0, y, 0.0, t2, e2, t2n);
        t2 = t2n;
      }
    }
    if constexpr (DDC) {
      double e1 = 0.0;
      ddcert_group_reduce<S>(d1, e1, red, rede);
      ddcert_group_reduce<S>(t2, e2, red, rede);
      if (r == 0 && lane_ok) {  // halving is exact (up to underflow)
        kgs[g][1] = Ar::scale(d1, 0.5);
        kgs[g][2] = Ar::scale(t2, 0.5);
        kge[g][1] = fma(0.5, e1, 2.0 * CERT_UFL);
        kge[g][2] = fma(0.5, e2, 2.0 * CERT_UFL);
      }
    } else if constexpr (CERT) {
      double e1 = 0.0;
      cert_group_reduce<S>(d1, e1, red, rede);
      cert_group_reduce<S>(t2, e2, red, rede);
      if (r == 0 && lane_ok) {
        kgs[g][1] = Ar::scale(d1, 0.5);
        kgs[g][2] = Ar::scale(t2, 0.5);
        kge[g][1] = fma(0.5, e1, CERT_UFL);
        kge[g][2] = fma(0.5, e2, CERT_UFL);
      }
    } else {
      const T s1 = group_reduce<Ar, S>(d1, red);
      const T s2 = group_reduce<Ar, S>(t2, red);
      if (r == 0 && lane_ok) {
        kgs[g][1] = Ar::scale(s1, 0.5);
        kgs[g][2] = Ar::scale(s2, 0.5);
      }
    }
  }
  // C^j for j = 2..h: one matrix product each, then p_{2j-1}, p_{2j}
  for (int j = 2; j <= h; ++j) {
    T acc[S];
    double eA[ANYC ? S : 1];
#pragma unroll
    for (int c = 0; c < S; ++c) acc[c] = Ar::zero();
    if (ANYC) {
#pragma unroll
      for (int c = 0; c < S; ++c) eA[c] = CERT ? (double)S * CERT_UFL : 0.0;
    }
    // the column offsets pass through an opaque move on every power: without
    // it the compiler keeps all S x M entry addresses live across the power
    // loop (up to 100 more VGPRs, one wave per SIMD where two fit)
    int colv[M];
#pragma unroll
    for (int u = 0; u < M; ++u) {
      colv[u] = 2 * p[u];
      asm volatile("" : "+v"(colv[u]));
    }
    // acc[c] = sum_t P[t] C[t][c] as an fma chain in t = 0, 1, .., S - 1 for
    // every c. B_S is symmetric, so C[t][c] = A[R_t][R_(c^1)] = C[c^1][t^1]:
    // each entry is read from LDS once and feeds both of its positions (half
    // the LDS reads; LDS bandwidth, not the FP64 pipe, bounds this loop). At
    // step t the entries (t, c) with c^1 >= t are read in ascending c^1; entry
    // (t, c) is term t of acc[c] and, as C[c^1][t^1], term c^1 of acc[t^1].
    // So acc[y] receives terms 0 .. y^1 - 1 at those steps, then term y^1 and
    // terms y^1 + 1 .. S - 1 in order at step y^1: the chain order of reading
    // row by row (for symmetric input the values are the same bits).
    auto wrow = [&](int t) -> double {  // a-priori bound weight of row t (certified)
      return CERT ? fma(gam, cert_norm1(Ar::to_c(P[t])), eP[t]) : 0.0;
    };
    // one term of the fma-chain dot product acc[c] += P[t] C[t][c] (and, in the
    // certified modes, its bound: sum_t (eP_t + gamma_2S |P_t|) |C_tc|)
    auto term = [&](int c, int t, const E& x) {
      if constexpr (DDC) {
        double o;
        acc[c] = cddc_fma_c(P[t], x, acc[c], &o);
        eA[c] = fma(eP[t], cert_norm1(x), eA[c] + o);
      } else {
        acc[c] = Ar::fma_c(P[t], x, acc[c]);
        if constexpr (CERT) eA[c] = fma(wrow(t), cert_norm1(x), eA[c]);
      }
    };
    if constexpr (DDC) {
      // certified dd keeps the row-by-row reads (the same chains, so the same
      // bits as the dd tier on symmetric input): the interleaved form made its
      // compilation take over half an hour
#pragma unroll
      for (int t = 0; t < S; ++t) {
        const E* At = As + __mul24(2 * p[t >> 1] + (t & 1), N);
#pragma unroll
        for (int cc = 0; cc < S; ++cc) term(cc ^ 1, t, At[colv[cc >> 1] + (cc & 1)]);
      }
    } else {
#pragma unroll
      for (int t = 0; t < S; ++t) {
        const E* At = As + __mul24(2 * p[t >> 1] + (t & 1), N);
#pragma unroll
        for (int cc = t; cc < S; ++cc) {
          const E x = At[colv[cc >> 1] + (cc & 1)];  // C[t][cc ^ 1] = A[R_t][R_cc]
          term(cc ^ 1, t, x);
          if (cc != t) term(t ^ 1, cc, x);
        }
      }
    }
    if constexpr (LOOP) {
      T uj;
      double euj;
      dot_d(acc, eA, uj, euj);
      add_v(2 * j, uprev, euprev, uj, euj);
      add_v(2 * j + 1, uj, euj, uj, euj);
      uprev = uj;
      euprev = euj;
    }
    T todd = Ar::zero(), teven = Ar::zero();
    double eo = 0.0, ee = 0.0;
#pragma unroll
    for (int c = 0; c < S; ++c) {
      const T yo = Ar::swap_pair(P[c ^ 1]);
      const T ye = Ar::swap_pair(acc[c ^ 1]);
      if constexpr (DDC) {
        const double eyo = swap_pair_d(eP[c ^ 1]), eye = swap_pair_d(eA[c ^ 1]);
        double oo, oe;
        const T ton = cddc_fma(acc[c], yo, todd, &oo);
        const T ten = cddc_fma(acc[c], ye, teven, &oe);
        eo = eo + cddc_mul_bound(acc[c], eA[c], yo, eyo, oo);
        ee = ee + cddc_mul_bound(acc[c], eA[c], ye, eye, oe);
        todd = ton;
        teven = ten;
        continue;
      }
      const T ton = Ar::fma(acc[c], yo, todd);
      const T ten = Ar::fma(acc[c], ye, teven);
      if constexpr (CERT) {
        const double eyo = swap_pair_d(eP[c ^ 1]), eye = swap_pair_d(eA[c ^ 1]);
        eo = cert_fma(acc[c], eA[c], yo, eyo, todd, eo, ton);
        ee = cert_fma(acc[c], eA[c], ye, eye, teven, ee, ten);
      }
      todd = ton;
      teven = ten;
    }
    if constexpr (DDC) {
      ddcert_group_reduce<S>(todd, eo, red, rede);
      ddcert_group_reduce<S>(teven, ee, red, rede);
      if (r == 0 && lane_ok) {
        kgs[g][2 * j - 1] = Ar::scale(todd, 0.5);
        kge[g][2 * j - 1] = fma(0.5, eo, 2.0 * CERT_UFL);
        if (2 * j <= Scratch::P) {
          kgs[g][2 * j] = Ar::scale(teven, 0.5);
          kge[g][2 * j] = fma(0.5, ee, 2.0 * CERT_UFL);
        }
      }
#pragma unroll
      for (int c = 0; c < S; ++c) eP[c] = eA[c];
    } else if constexpr (CERT) {
      cert_group_reduce<S>(todd, eo, red, rede);
      cert_group_reduce<S>(teven, ee, red, rede);
      if (r == 0 && lane_ok) {
        kgs[g][2 * j - 1] = Ar::scale(todd, 0.5);
        kge[g][2 * j - 1] = fma(0.5, eo, CERT_UFL);
        if (2 * j <= Scratch::P) {
          kgs[g][2 * j] = Ar::scale(teven, 0.5);
          kge[g][2 * j] = fma(0.5, ee, CERT_UFL);
        }
      }
#pragma unroll
      for (int c = 0; c < S; ++c) eP[c] = eA[c];
    } else {
      const T so = group_reduce<Ar, S>(todd, red);
      const T se = group_reduce<Ar, S>(teven, red);
      if (r == 0 && lane_ok) {
        kgs[g][2 * j - 1] = Ar::scale(so, 0.5);
        if (2 * j <= Scratch::P) kgs[g][2 * j] = Ar::scale(se, 0.5);
      }
    }
#pragma unroll
    for (int c = 0; c < S; ++c) P[c] = acc[c];
  }
  if constexpr (LOOP) {
    // kg_k = p_k / 2 + k v_k
    if (r == 0 && lane_ok) {
      for (int k = 1; k <= n; ++k) {
        if constexpr (DDC) {
          double o;
          kgs[g][k] = cddc_add(kgs[g][k], kvs[g][k], &o);
          kge[g][k] = kge[g][k] + kve[g][k] + o;
        } else if constexpr (CERT) {
          const T sum = Ar::add(kgs[g][k], kvs[g][k]);
          kge[g][k] = cert_round(Ar::to_c(sum), kge[g][k] + kve[g][k]);
          kgs[g][k] = sum;
        } else {
          kgs[g][k] = Ar::add(kgs[g][k], kvs[g][k]);
        }
      }
    }
  }
  __syncthreads();
  // Newton: e_j = (1/j) sum_{k=1..j} kg_k e_{j-k}, by the first lane of each group
  T coeff = Ar::zero();
  double ecoeff = 0.0;
  if constexpr (!ANYC && (std::is_same<Ar, ArF64>::value || std::is_same<Ar, ArR64>::value) && Scratch::P <= 7) {
    // few pairs (the fused small-size kernel): every lane runs the recurrence
    // with kg and e in registers, unrolled, instead of one lane walking them in
    // LDS; the same operations in the same order, so the same bits
    const int gg = lane_ok ? g : 0;
    T kg[Scratch::P + 1];
#pragma unroll
    for (int k = 1; k <= Scratch::P; ++k) kg[k] = kgs[gg][k <= n ? k : 1];
    T e[Scratch::P + 1];
    e[0] = Ar::one();
#pragma unroll
    for (int jj = 1; jj <= Scratch::P; ++jj) {
      if (jj <= n) {  // uniform: n is the workgroup's
        T a = Ar::zero();
#pragma unroll
        for (int k = 1; k <= jj; ++k) a = Ar::fma(kg[k], e[jj - k], a);
        e[jj] = Ar::div_int(a, jj);
        if (jj == n) coeff = e[jj];
      }
    }
  } else if (r == 0 && lane_ok) {
    T* e = es[g];
    e[0] = Ar::one();
    if (ANYC) ese[g][0] = 0.0;
    for (int jj = 1; jj <= n; ++jj) {
      T a = Ar::zero();
      double ea = 0.0;
      if constexpr (DDC) {
        for (int k = 1; k <= jj; ++k) {
          double o;
          const T an = cddc_fma(kgs[g][k], e[jj - k], a, &o);
          ea = ea + cddc_mul_bound(kgs[g][k], kge[g][k], e[jj - k], ese[g][jj - k], o);
          a = an;
        }
        double od;
        e[jj] = cddc_div_d(a, (double)jj, &od);
        ese[g][jj] = ea / (double)jj + od;
        continue;
      }
      for (int k = 1; k <= jj; ++k) {
        const T an = Ar::fma(kgs[g][k], e[jj - k], a);
        if (CERT) ea = cert_fma(Ar::to_c(kgs[g][k]), kge[g][k], Ar::to_c(e[jj - k]), ese[g][jj - k], Ar::to_c(a), ea,
                                Ar::to_c(an));
        a = an;
      }
      e[jj] = Ar::div_int(a, jj);
      if (CERT) ese[g][jj] = cert_round(Ar::to_c(e[jj]), ea / (double)jj);
    }
    coeff = e[n];
    if (ANYC) ecoeff = ese[g][n];
  }
  __syncthreads();
  if ((n - M) & 1) coeff = Ar::neg(coeff);
  coeff_out = coeff;
  ecoeff_out = ecoeff;
}

// ---------------------------------------------------------------- phase kernel
// One workgroup (one wave) = one instance; evaluates every subset of size
// m = S/2 and writes the phase partial (and sum |term| in kappa mode, the
// bound in certified mode).
//   part[b * np + off_S + chunk]      chunk partial (np = haf_nparts(n))
// Waves per SIMD the register allocator must leave room for, per fp64 phase
// kernel: bit S / 2 of a mask asks for 2 (a wave capped at 256 registers, the
// excess spilled) where the allocator alone would take more and run one wave
// per SIMD. Plain and kappa share a mask; certified has its own. Chosen per
// subset size from per-phase kernel times on MI355X (N = 16 and 24): forced
// where it was faster (S = 20, 22: 1.3x; loop S = 18: 1.55x; certified
// S = 14, 16 and loop S = 10 .. 14: 1.1 - 1.55x), left to the allocator where
// the spill cost more (S = 24, loop S = 16, certified loop S = 16).
#ifndef GBSK_HAF_OCC2_MASK
#define GBSK_HAF_OCC2_MASK ((1u << 10) | (1u << 11))
#endif
#ifndef GBSK_HAF_OCC2_LOOP_MASK
#define GBSK_HAF_OCC2_LOOP_MASK ((1u << 9) | (1u << 10) | (1u << 11))
#endif
#ifndef GBSK_HAF_OCC2_CERT_MASK
#define GBSK_HAF_OCC2_CERT_MASK ((1u << 7) | (1u << 8))
#endif
#ifndef GBSK_HAF_OCC2_CERT_LOOP_MASK
#define GBSK_HAF_OCC2_CERT_LOOP_MASK ((1u << 5) | (1u << 6) | (1u << 7))
#endif
template <class Ar, bool LOOP, int S, int MODE>
constexpr int haf_phase_min_waves() {
  if (!std::is_same<Ar, ArF64>::value) return 1;
  const unsigned mask = MODE == HM_CERT ? (LOOP ? (unsigned)GBSK_HAF_OCC2_CERT_LOOP_MASK : (unsigned)GBSK_HAF_OCC2_CERT_MASK)
                                        : (LOOP ? (unsigned)GBSK_HAF_OCC2_LOOP_MASK : (unsigned)GBSK_HAF_OCC2_MASK);
  return ((mask >> (S / 2)) & 1u) ? 2 : 1;
}

template <class Ar, bool LOOP, int S, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(haf_phase_min_waves<Ar, LOOP, S, MODE>())))
void haf_phase_kernel(const HafPhaseJobs<typename Ar::T> jobs) {
  using T = typename Ar::T;
  constexpr bool KAPPA = (MODE == HM_KAPPA);
  constexpr bool CERT = (MODE == HM_CERT);
  constexpr bool DDC = (MODE == HM_DDCERT);  // certified double-double (Ar = ArDD)
  constexpr bool ANYC = CERT || DDC;
  constexpr int M = S / 2;
  constexpr int G = 64 / S;
  extern __shared__ cplx As[];  // N * N, dynamic
  __shared__ HafScratch<Ar, LOOP, S, MODE> sh;
  T* red = sh.red;
  double* rede = sh.rede;
  const int lane = threadIdx.x;
  // this workgroup's bucket (wave-uniform scan of at most HAF_MAX_JOBS entries)
  int jb = 0;
  while (jb + 1 < jobs.n && (int)blockIdx.x >= jobs.wg_begin[jb + 1]) ++jb;
  const cplx* __restrict__ mats = jobs.mats[jb];
  const int N = jobs.N[jb];
  typename Ar::T* __restrict__ part = jobs.part[jb];
  double* __restrict__ part_abs = jobs.part_abs[jb];
  const int local = (int)blockIdx.x - jobs.wg_begin[jb];
  const int nch = haf_nchunks(N / 2, S);
  const int64_t b = local / nch;
  const int chunk = local % nch;
  const int n = N / 2;
  const int64_t pslot = b * haf_nparts(n) + haf_part_offset(n, S) + chunk;
  // stage A (row-major N x N); the ordinary hafnian zeroes the diagonal
  for (int e = lane; e < N * N; e += 64) {
    cplx v = mats[b * N * N + e];
    if (!LOOP && (e / N) == (e % N)) v = cplx{0.0, 0.0};
    As[e] = v;
  }
  __syncthreads();

  const int g = lane / S;
  const int r = lane % S;
  const bool lane_ok = g < G;
  const int count = binom(n, M);

  T group_acc = Ar::zero();
  double group_abs = 0.0;  // kappa: sum |term|; certified: bound of group_acc

  const int qbeg = chunk * haf_chunk_size(S);
  const int qend = min(count, qbeg + haf_chunk_size(S));
  for (int q0 = qbeg; q0 < qend; q0 += G) {
    const int q = q0 + (lane_ok ? g : 0);  // groups past G mirror group 0 (discarded)
    const bool active = lane_ok && q < qend;
    T coeff;
    double ecoeff;
    haf_subset_term<Ar, LOOP, S, MODE>(As, 0, N, active ? q : 0, sh, coeff, ecoeff);
    if constexpr (DDC) {
      if (active && r == 0) {
        double o;
        group_acc = cddc_add(group_acc, coeff, &o);
        group_abs = group_abs + ecoeff + o;
      }
    }
    if (!DDC && active && r == 0) {
      group_acc = Ar::add(group_acc, coeff);
      if (KAPPA) {
        const cplx cc = Ar::to_c(coeff);
        group_abs += c_abs(cc);
      }
      if (CERT) group_abs = cert_round(Ar::to_c(group_acc), group_abs + ecoeff);
    }
  }
  // combine the G group partials in group order
  if (r == 0 && lane_ok) red[g] = group_acc;
  if (ANYC && r == 0 && lane_ok) rede[g] = group_abs;
  __syncthreads();
  if (lane == 0) {
    T s = red[0];
    double E = ANYC ? rede[0] : 0.0;
    for (int i = 1; i < G; ++i) {
      if constexpr (DDC) {
        double o;
        s = cddc_add(s, red[i], &o);
        E = E + rede[i] + o;
      } else {
        s = Ar::add(s, red[i]);
        if (CERT) E = cert_round(Ar::to_c(s), E + rede[i]);
      }
    }
    part[pslot] = s;
    // certified dd: each bound operation above is a few roundings deeper than
    // its fp64 counterpart, hence 4x the fp64 depth
    if (ANYC) part_abs[pslot] = cert_finish(E, (DDC ? 4.0 : 1.0) * haf_cert_depth(S, n, count));
  }
  if (KAPPA) {
    __syncthreads();
    double* rabs = reinterpret_cast<double*>(red);
    if (r == 0 && lane_ok) rabs[g] = group_abs;
    __syncthreads();
    if (lane == 0) {
      double s = rabs[0];
      for (int i = 1; i < G; ++i) s += rabs[i];
      part_abs[pslot] = s;
    }
  }
}

// Reduce descriptor over several buckets: bucket j owns instances
// [b_begin[j], b_begin[j+1]) of the launch's thread range.
template <class T>
struct HafReduceJobs {
  const T* part[HAF_MAX_JOBS];
  const double* part_abs[HAF_MAX_JOBS];
  cplx* out[HAF_MAX_JOBS];
  double* aux[HAF_MAX_JOBS];
  int n[HAF_MAX_JOBS];
  int b_begin[HAF_MAX_JOBS + 1];
  int njobs;
};

// out[b] = sum of the instance's np chunk partials, one wave per instance:
// lane l sums slots l, l + 64, l + 128, ... in slot order, then a fixed
// xor-butterfly over the 64 lanes; lane 0 writes. The association depends on np
// only, so every mode (plain, kappa, certified) produces the same value bits.
constexpr int HAF_REDUCE_WAVES = 4;
template <class Ar, int MODE>
__global__ __launch_bounds__(64 * HAF_REDUCE_WAVES) void haf_reduce_kernel(const HafReduceJobs<typename Ar::T> jobs) {
  using T = typename Ar::T;
  const int64_t t = (int64_t)blockIdx.x * HAF_REDUCE_WAVES + (threadIdx.x >> 6);  // wave-uniform
  const int lane = threadIdx.x & 63;
  if (t >= jobs.b_begin[jobs.njobs]) return;
  int jb = 0;
  while (jb + 1 < jobs.njobs && t >= jobs.b_begin[jb + 1]) ++jb;
  const int n = jobs.n[jb];
  const int np = haf_nparts(n);
  const int64_t b = t - jobs.b_begin[jb];
  const T* __restrict__ part = jobs.part[jb] + b * np;
  const double* __restrict__ part_abs = jobs.part_abs[jb] + b * np;
  T s = Ar::zero();
  double a = 0.0;
  if (lane < np) {
    s = part[lane];
    if (MODE != HM_PLAIN) a = part_abs[lane];
  }
  for (int m = lane + 64; m < np; m += 64) {
    if constexpr (MODE == HM_DDCERT) {
      double o;
      s = cddc_add(s, part[m], &o);
      a = a + part_abs[m] + o;
    } else {
      s = Ar::add(s, part[m]);
      if (MODE == HM_KAPPA) a += part_abs[m];
      if (MODE == HM_CERT) a = cert_round(Ar::to_c(s), a + part_abs[m]);
    }
  }
#pragma unroll
  for (int k = 1; k < 64; k <<= 1) {
    const T o = Ar::shfl_xor(s, k);
    const double ao = MODE != HM_PLAIN ? shfl_xor_d(a, k) : 0.0;
    if constexpr (MODE == HM_DDCERT) {
      double r;
      s = cddc_add(s, o, &r);
      a = a + ao + r;
    } else {
      s = Ar::add(s, o);
      if (MODE == HM_KAPPA) a += ao;
      if (MODE == HM_CERT) a = cert_round(Ar::to_c(s), a + ao);
    }
  }
  if (lane != 0) return;
  cplx* __restrict__ out = jobs.out[jb];
  double* __restrict__ aux = jobs.aux[jb];
  out[b] = Ar::to_c(s);
  if (MODE == HM_KAPPA) aux[b] = a;
  // rounding depth of the bound arithmetic: at most np serial steps plus the six
  // butterfly levels, each a few roundings deep
  if (MODE == HM_CERT) aux[b] = cert_finish(a, 4.0 * (np + 6) + 8.0);
  if constexpr (MODE == HM_DDCERT) {
    const double f = cert_finish(a + cddc_collapse_err(s), 8.0 * (np + 6) + 16.0);
    aux[b] = (f == f) ? f : INFINITY;
  }
}

// ------------------------------------------- small sizes: one fused launch
// For N <= HAF_SMALL_MAX_N (fp64 / kappa / certified fp64 / dd / certified dd) one launch replaces the n phase
// launches and the reduce. A workgroup (one wave) takes `T` consecutive
// instances of a bucket: it stages their matrices in LDS, evaluates every subset
// of every size with haf_subset_term (lane groups of S lanes, 64 / S groups per
// round, subsets of several instances in one round, since a term depends only
// on its own matrix), keeps the terms in LDS, forms each chunk partial in the
// phase kernel's order (the chunk's groups summed in group order, idle groups
// contributing zero) and sums the partials with the reduce kernel's butterfly.
// Values, kappa sums and certified bounds (each bound operation of the phase
// and reduce kernels replayed in the same order) are therefore bit-identical
// to the phase path, and the number of instances per workgroup only changes
// the schedule.
constexpr int HAF_SMALL_MAX_N = 12;
constexpr int HAF_SMALL_MAX_PAIRS = HAF_SMALL_MAX_N / 2;
constexpr int HAF_SMALL_MAX_T = 4;  // instances per workgroup
// certified dd takes the fused kernel up to N = 8 only: measured at B = 4096,
// N = 8 ran 535 -> 190 us fused, but the N <= 12 class (every phase's dd
// bound registers in one kernel) made N = 12 1.3x slower than the phases
constexpr int HAF_SMALL_DDCERT_MAX_N = 8;
// the dd loop hafnian likewise stops at N = 10: at B = 4096 its N <= 12 class
// took 1978 us at N = 12 against the phases' 1484 (N = 10: 747 against 779)
constexpr int HAF_SMALL_LHAF_DD_MAX_N = 10;
// terms of one instance (every nonempty subset of its pairs, 2^n - 1 <= 63)
// and chunk partials (haf_nparts(n) <= 8 for n <= 6)
__host__ __device__ constexpr int haf_small_terms(int pairs) { return (1 << pairs) - 1; }
constexpr int HAF_SMALL_MAX_PARTS = 8;
static_assert(HAF_CHUNK_ITERS == 1, "the fused small-size kernel replays one-round chunks");

struct HafSmallJobs {
  const cplx* mats[HAF_MAX_JOBS];
  cplx* out[HAF_MAX_JOBS];
  double* aux[HAF_MAX_JOBS];
  int N[HAF_MAX_JOBS];
  int B[HAF_MAX_JOBS];
  int T[HAF_MAX_JOBS];
  int wg_begin[HAF_MAX_JOBS + 1];
  int n;
};

// one phase (subset size S) of the ninst staged instances: terms[i][off + q]
// receives instance i's subset-q term as the phase kernel's group accumulator
// holds it (zero + term), tabs the kappa magnitude (0 + |term|)
template <class Ar, bool LOOP, int MODE, int PAIRS, int S>
__device__ __forceinline__ void haf_small_phase(const typename Ar::E* As, int N, int ninst, void* scratch,
                                                typename Ar::T (*terms)[haf_small_terms(PAIRS)],
                                                double (*tabs)[haf_small_terms(PAIRS)],
                                                typename Ar::T (*parts)[HAF_SMALL_MAX_PARTS],
                                                double (*pabs)[HAF_SMALL_MAX_PARTS]) {
  using T = typename Ar::T;
  using Sc = HafScratch<Ar, LOOP, S, MODE, PAIRS + 1>;
  constexpr int G = 64 / S;
  Sc& sh = *reinterpret_cast<Sc*>(scratch);
  const int lane = threadIdx.x;
  const int g = lane / S, r = lane % S;
  const bool lane_ok = g < G;
  const int n = N / 2;
  const int count = haf_binom(n, S / 2);
  const int off = haf_small_term_offset(n, S);
  const int total = ninst * count;
  for (int base = 0; base < total; base += G) {  // wave-uniform
    const int item = base + g;
    const bool active = lane_ok && item < total;
    const int iq = small_div(item, count);
    const int inst = active ? iq : 0;
    const int q = active ? item - iq * count : 0;
    T coeff;
    double ecoeff;
    haf_subset_term<Ar, LOOP, S, MODE>(As, __mul24(inst, N * N), N, q, sh, coeff, ecoeff);
    (void)ecoeff;
    if (active && r == 0) {
      if constexpr (MODE == HM_DDCERT) {
        // certified dd: the group sum and its bound after its one subset, as
        // the phase kernel forms them (a dd add that reports its rounding)
        double o;
        terms[inst][off + q] = cddc_add(Ar::zero(), coeff, &o);
        tabs[inst][off + q] = 0.0 + ecoeff + o;
      } else {
        const T t0 = Ar::add(Ar::zero(), coeff);
        terms[inst][off + q] = t0;
        if (MODE == HM_KAPPA) tabs[inst][off + q] = 0.0 + c_abs(Ar::to_c(coeff));
        // certified: the group's bound after its one subset, as the phase
        // kernel forms it (cert_round of the group sum, from 0 + the term's bound)
        if (MODE == HM_CERT) tabs[inst][off + q] = cert_round(Ar::to_c(t0), 0.0 + ecoeff);
      }
    }
  }
  __syncthreads();
  // this phase's chunk partials, in the phase kernel's association: chunk
  // `slot` of instance i sums the G groups of one round (terms q0 .. q0 + G - 1,
  // idle groups past the last subset adding zero) in group order. G is a
  // compile-time constant here, so every term read is issued ahead of the add
  // chain (reads past the last subset are clamped to it and then discarded).
  // A chunk holds at most KMAX = min(G, C(PAIRS, S/2)) subsets, so only its
  // first KMAX groups are read; the idle groups after them add zero in the tail.
  constexpr int KMAX = G < haf_cbinom(PAIRS, S / 2) ? G : haf_cbinom(PAIRS, S / 2);
  const int nch = haf_nchunks(n, S);
  const int pbase = haf_part_offset(n, S);
  for (int idx = lane; idx < ninst * nch; idx += 64) {
    const int i = small_div(idx, nch), slot = idx - i * nch, q0 = slot * G;
    const T* ti = terms[i] + off;
    T acc = ti[q0];
    double a = (MODE == HM_KAPPA) ? tabs[i][off + q0] : 0.0;
#pragma unroll
    for (int k = 1; k < KMAX; ++k) {
      const bool on = q0 + k < count;
      const int qk = on ? q0 + k : q0;
      const T x = ti[qk];
      acc = Ar::add(acc, on ? x : Ar::zero());
      if (MODE == HM_KAPPA) a += on ? tabs[i][off + qk] : 0.0;
    }
    if constexpr (std::is_same<T, double>::value || std::is_same<T, cplx>::value) {
      // fp64: x + 0 + 0 = x + 0 (the first zero turns -0 into +0, the rest are exact)
      if (KMAX < G) {
        acc = Ar::add(acc, Ar::zero());
        if (MODE == HM_KAPPA) a += 0.0;
      }
    } else {
      for (int k = KMAX; k < G; ++k) acc = Ar::add(acc, Ar::zero());
    }
    parts[i][pbase + slot] = acc;
    if (MODE == HM_KAPPA) pabs[i][pbase + slot] = a;
    if constexpr (MODE == HM_CERT) {
      // the chunk bound of the phase kernel, operation for operation: every
      // one of the G groups in group order (an idle group adds a zero sum with
      // a zero bound, and cert_round still charges its underflow slack), then
      // the chunk's finishing inflation
      T s = ti[q0];
      double E = tabs[i][off + q0];
      for (int k = 1; k < G; ++k) {
        const bool on = q0 + k < count;
        const int qk = on ? q0 + k : q0;
        const T x = ti[qk];
        const double ex = tabs[i][off + qk];
        s = Ar::add(s, on ? x : Ar::zero());
        E = cert_round(Ar::to_c(s), E + (on ? ex : 0.0));
      }
      pabs[i][pbase + slot] = cert_finish(E, haf_cert_depth(S, n, count));
    }
    if constexpr (MODE == HM_DDCERT) {
      // the phase kernel's certified-dd chunk: the G group sums in group order
      // with reporting dd adds (value and bound), then 4x the fp64 depth
      T s = ti[q0];
      double E = tabs[i][off + q0];
      for (int k = 1; k < G; ++k) {
        const bool on = q0 + k < count;
        const int qk = on ? q0 + k : q0;
        const T x = ti[qk];
        const double ex = tabs[i][off + qk];
        double o;
        s = cddc_add(s, on ? x : Ar::zero(), &o);
        E = E + (on ? ex : 0.0) + o;
      }
      parts[i][pbase + slot] = s;
      pabs[i][pbase + slot] = cert_finish(E, 4.0 * haf_cert_depth(S, n, count));
    }
  }
}

template <class Ar, bool LOOP, int MODE, int PAIRS, int... Ss>
__device__ __forceinline__ void haf_small_phases(const typename Ar::E* As, int N, int ninst, void* scratch,
                                                 typename Ar::T (*terms)[haf_small_terms(PAIRS)],
                                                 double (*tabs)[haf_small_terms(PAIRS)],
                                                 typename Ar::T (*parts)[HAF_SMALL_MAX_PARTS],
                                                 double (*pabs)[HAF_SMALL_MAX_PARTS],
                                                 std::integer_sequence<int, Ss...>) {
  ((2 * (Ss + 1) <= N ? (haf_small_phase<Ar, LOOP, MODE, PAIRS, 2 * (Ss + 1)>(As, N, ninst, scratch, terms, tabs,
                                                                              parts, pabs),
                         0)
                      : 0),
   ...);
}

// Everything after staging: the phases, the butterfly and the stores, in the
// arithmetic Ar (ArR64 when the workgroup's matrices are real). The LDS arrays
// are the kernel's, sized for its own (complex) Ar; a real Ar uses their first
// halves.
template <class Ar, bool LOOP, int MODE, int PAIRS>
__device__ __forceinline__ void haf_small_body(cplx* outp, double* auxp, int ninst, int N,
                                               const typename Ar::E* As, void* terms_raw,
                                               double (*tabs)[haf_small_terms(PAIRS)], void* parts_raw,
                                               double (*pabs)[HAF_SMALL_MAX_PARTS], void* scratch) {
  using T = typename Ar::T;
  constexpr bool KAPPA = (MODE == HM_KAPPA);
  constexpr int TERMS = haf_small_terms(PAIRS);
  T(*terms)[TERMS] = reinterpret_cast<T(*)[TERMS]>(terms_raw);
  T(*parts)[HAF_SMALL_MAX_PARTS] = reinterpret_cast<T(*)[HAF_SMALL_MAX_PARTS]>(parts_raw);
  const int lane = threadIdx.x;
  const int n = N / 2;
  constexpr bool CERT = (MODE == HM_CERT);
  constexpr bool DDC = (MODE == HM_DDCERT);
  constexpr bool AUX = KAPPA || CERT || DDC;
  haf_small_phases<Ar, LOOP, MODE, PAIRS>(As, N, ninst, scratch, terms, AUX ? tabs : nullptr, parts,
                                          AUX ? pabs : nullptr, std::make_integer_sequence<int, PAIRS>{});
  __syncthreads();
  // the reduce kernel's butterfly over each instance's partials, for lane 0,
  // whose result is stored: with at most 8 partials (lanes 0..7, exact zeros
  // above), levels 1, 2 are quad swaps, level 4 brings lane 4 to lane 0, and
  // levels 8, 16, 32 add the zero sums of lanes 8.., 16.., 32.. . Each as a DPP
  // move or a zero, the same adds in the same order; the instances' chains are
  // interleaved.
  static_assert(HAF_SMALL_MAX_PARTS <= 8, "the butterfly below covers 8 partials");
  const int np = haf_nparts(n);
  T s[HAF_SMALL_MAX_T];
  double a[KAPPA ? HAF_SMALL_MAX_T : 1];
#pragma unroll
  for (int i = 0; i < HAF_SMALL_MAX_T; ++i) {
    const bool on = i < ninst && lane < np;
    const T x = parts[i][on ? lane : 0];
    s[i] = on ? x : Ar::zero();
    if (KAPPA) a[i] = on ? pabs[i][on ? lane : 0] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < HAF_SMALL_MAX_T; ++i) {
    s[i] = Ar::add(s[i], Ar::template dpp<0xB1>(s[i]));
    if (KAPPA) a[i] += dpp_d<0xB1>(a[i]);
  }
#pragma unroll
  for (int i = 0; i < HAF_SMALL_MAX_T; ++i) {
    s[i] = Ar::add(s[i], Ar::template dpp<0x4E>(s[i]));
    if (KAPPA) a[i] += dpp_d<0x4E>(a[i]);
  }
#pragma unroll
  for (int i = 0; i < HAF_SMALL_MAX_T; ++i) {
    s[i] = Ar::add(s[i], Ar::template dpp<0x104>(s[i]));
    if (KAPPA) a[i] += dpp_d<0x104>(a[i]);
  }
#pragma unroll
  for (int k = 8; k < 64; k <<= 1) {
#pragma unroll
    for (int i = 0; i < HAF_SMALL_MAX_T; ++i) {
      s[i] = Ar::add(s[i], Ar::zero());
      if (KAPPA) a[i] += 0.0;
    }
  }
  // certified: the reduce kernel's bound, literally: all 64 lanes (lanes at
  // or past the partial count hold a zero sum and a zero bound) through the
  // six xor-butterfly levels, each a cert_round of the new sum, then its
  // finishing inflation
  double eb[CERT ? HAF_SMALL_MAX_T : 1];
  if constexpr (CERT) {
#pragma unroll
    for (int i = 0; i < HAF_SMALL_MAX_T; ++i) {
      const bool on = i < ninst && lane < np;
      const T x = parts[i][on ? lane : 0];
      const double ex = pabs[i][on ? lane : 0];
      T sv = on ? x : Ar::zero();
      double av = on ? ex : 0.0;
#pragma unroll
      for (int k = 1; k < 64; k <<= 1) {
        const T o = Ar::shfl_xor(sv, k);
        const double ao = shfl_xor_d(av, k);
        sv = Ar::add(sv, o);
        av = cert_round(Ar::to_c(sv), av + ao);
      }
      eb[i] = cert_finish(av, 4.0 * (np + 6) + 8.0);
    }
  }
  // certified dd: the reduce kernel's reporting-add butterfly, literally (its
  // value is the stored one), the sum's collapse error and the finishing
  // inflation, a NaN bound made +inf
  double ebd[DDC ? HAF_SMALL_MAX_T : 1];
  T sd[DDC ? HAF_SMALL_MAX_T : 1];
  if constexpr (DDC) {
#pragma unroll
    for (int i = 0; i < HAF_SMALL_MAX_T; ++i) {
      const bool on = i < ninst && lane < np;
      const T x = parts[i][on ? lane : 0];
      const double ex = pabs[i][on ? lane : 0];
      T sv = on ? x : Ar::zero();
      double av = on ? ex : 0.0;
#pragma unroll
      for (int k = 1; k < 64; k <<= 1) {
        const T o = Ar::shfl_xor(sv, k);
        const double ao = shfl_xor_d(av, k);
        double rr;
        sv = cddc_add(sv, o, &rr);
        av = av + ao + rr;
      }
      const double f = cert_finish(av + cddc_collapse_err(sv), 8.0 * (np + 6) + 16.0);
      ebd[i] = (f == f) ? f : INFINITY;
      sd[i] = sv;
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < HAF_SMALL_MAX_T; ++i) {
      if (i < ninst) {
        if constexpr (DDC) {
          outp[i] = Ar::to_c(sd[i]);
          auxp[i] = ebd[i];
        } else {
          outp[i] = Ar::to_c(s[i]);
          if (KAPPA) auxp[i] = a[i];
          if (CERT) auxp[i] = eb[i];
        }
      }
    }
  }
}

// NMAX: the size class (4, 8 or 12) the LDS arrays and phases are built for;
// a smaller class takes less LDS and fewer registers (more waves per CU).
// fp64 and kappa run a workgroup whose staged imaginary parts are all zero in
// real arithmetic (ArR64: the same values, a quarter of the products).
template <class Ar, bool LOOP, int MODE, int NMAX>
__global__ __launch_bounds__(64) void haf_small_kernel(const HafSmallJobs jobs) {
  using T = typename Ar::T;
  constexpr bool KAPPA = (MODE == HM_KAPPA);
  constexpr bool REAL = std::is_same<Ar, ArF64>::value && (MODE == HM_PLAIN || MODE == HM_KAPPA);
  constexpr int PAIRS = NMAX / 2;
  constexpr int NN = NMAX * NMAX;
  constexpr int TERMS = haf_small_terms(PAIRS);
  // the largest phase scratch is the S = 2 one (32 lane groups)
  constexpr size_t SCR = sizeof(HafScratch<Ar, LOOP, 2, MODE, PAIRS + 1>);
  static_assert(sizeof(HafScratch<ArR64, LOOP, 2, MODE, PAIRS + 1>) <= SCR, "real scratch fits");
  __shared__ cplx As[HAF_SMALL_MAX_T * NN];
  __shared__ T terms[HAF_SMALL_MAX_T][TERMS];
  constexpr bool AUX = KAPPA || MODE == HM_CERT || MODE == HM_DDCERT;  // per-term magnitudes or bounds
  __shared__ double tabs[AUX ? HAF_SMALL_MAX_T : 1][TERMS];
  __shared__ T parts[HAF_SMALL_MAX_T][HAF_SMALL_MAX_PARTS];
  __shared__ double pabs[AUX ? HAF_SMALL_MAX_T : 1][HAF_SMALL_MAX_PARTS];
  __shared__ __align__(16) unsigned char scratch[SCR];
  const int lane = threadIdx.x;
  int jb = 0;
  while (jb + 1 < jobs.n && (int)blockIdx.x >= jobs.wg_begin[jb + 1]) ++jb;
  const int N = jobs.N[jb], Tj = jobs.T[jb];
  const int64_t b0 = (int64_t)((int)blockIdx.x - jobs.wg_begin[jb]) * Tj;
  const int ninst = (int)min((int64_t)Tj, (int64_t)jobs.B[jb] - b0);
  const int ne = ninst * N * N;
  const cplx* __restrict__ src = jobs.mats[jb] + b0 * N * N;
  // the output pointers are read from the argument block now, while the
  // matrices load, not after the phases (the asm makes each value ready here)
  cplx* const outp = jobs.out[jb] + b0;
  double* const auxp = AUX ? jobs.aux[jb] + b0 : nullptr;
  asm volatile("" : : "s"(outp), "s"(auxp));
  // stage through registers (the ordinary hafnian zeroes the diagonal, as the
  // phase kernels do); a real workgroup stores the real parts only, packed
  // into the front of the same LDS array
  constexpr int ITER = (HAF_SMALL_MAX_T * NN + 63) / 64;
  cplx v[ITER];
  bool nonreal = false;
#pragma unroll
  for (int k = 0; k < ITER; ++k) {
    const int e = lane + 64 * k;
    v[k] = cplx{0.0, 0.0};
    if (e < ne) v[k] = src[e];
    const int rc = e - __mul24(small_div(e, N * N), N * N);  // e < 2^16, N * N <= 2^10
    if (!LOOP && rc == __mul24(small_div(rc, N), N + 1)) v[k] = cplx{0.0, 0.0};  // diagonal
    nonreal |= !(v[k].im == 0.0);
  }
  if constexpr (REAL) {
    if (!__any(nonreal)) {  // wave-uniform
      double* Asr = reinterpret_cast<double*>(As);
#pragma unroll
      for (int k = 0; k < ITER; ++k) {
        const int e = lane + 64 * k;
        if (e < ne) Asr[e] = v[k].re;
      }
      __syncthreads();
      haf_small_body<ArR64, LOOP, MODE, PAIRS>(outp, auxp, ninst, N, Asr, terms, tabs, parts, pabs, scratch);
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < ITER; ++k) {
    const int e = lane + 64 * k;
    if (e < ne) As[e] = v[k];
  }
  __syncthreads();
  haf_small_body<Ar, LOOP, MODE, PAIRS>(outp, auxp, ninst, N, As, terms, tabs, parts, pabs, scratch);
}

// Schedule of the fused small-size kernel (host state, haf.hip; neither knob
// changes a value): it runs sizes N <= haf_small_max_n. A positive
// haf_small_target_wgs packs as many instances per workgroup (at most
// HAF_SMALL_MAX_T) as keep at least that many workgroups in the launch; 0 (the
// built-in) picks the packing T per bucket from a cost model fitted on MI355X:
// a wave of T instances makes haf_small_rounds(N, T) subset rounds; each
// further wave sharing a SIMD adds 0.38 of its rounds (one wave leaves issue
// slots idle that a second one partly fills; from N = 8, B = 4096: two waves
// of 4 rounds took 0.92 x one wave of 6). A smaller T competes only if all its
// workgroups fit on the device at once: the workgroups take equal time, so
// the ones left over would run as a whole second pass (N = 12, B = 4096 at
// T = 2: 2048 workgroups where 1792 fit, 1.25x slower than T = 4).
extern int haf_small_max_n;
extern int haf_small_target_wgs;
inline int haf_small_rounds(int N, int T) {
  const int n = N / 2;
  int r = 0;
  for (int m = 1; m <= n; ++m) {
    const int G = 64 / (2 * m);
    r += (T * haf_binom(n, m) + G - 1) / G;
  }
  return r;
}
// cus: compute units; per_cu: workgroups (one wave each) of the kernel a CU holds
inline int haf_small_T(int B, int N, int cus, int per_cu) {
  if (haf_small_target_wgs > 0) {
    const int t = B / haf_small_target_wgs;
    return t < 1 ? 1 : (t > HAF_SMALL_MAX_T ? HAF_SMALL_MAX_T : t);
  }
  int best = HAF_SMALL_MAX_T;
  double best_cost = 0.0;
  for (int T = HAF_SMALL_MAX_T; T >= 1; --T) {  // ties keep the larger T
    const int wg = (B + T - 1) / T;
    if (T < HAF_SMALL_MAX_T && wg > cus * per_cu) break;  // a second pass would double the time
    const double waves = ceil((double)wg / (4.0 * cus));  // per SIMD
    const double cost = haf_small_rounds(N, T) * (1.0 + 0.38 * (waves - 1.0));
    if (T == HAF_SMALL_MAX_T || cost < best_cost) {
      best = T;
      best_cost = cost;
    }
  }
  return best;
}
// compute units of the current device and how many workgroups of `kernel`
// (64 threads, static LDS) one of them holds; queried once per kernel (one
// device model per process)
inline void haf_small_occupancy(const void* kernel, int* cus, int* per_cu) {
  int dev = 0, c = 0, nb = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 64, 0) != hipSuccess || c < 1 || nb < 1) {
    (void)hipGetLastError();
    c = 256;
    nb = 1;
  }
  *cus = c;
  *per_cu = nb;
}

template <class Ar, bool LOOP, int MODE>
static void launch_small(const HafJob* jobs, int njobs, hipStream_t st) {
  // one launch per size class present (N <= 4, <= 8, <= 12); certified dd
  // stops at N = 8, so its "12" class names the 8 kernel and is never launched
  constexpr int C12 = MODE == HM_DDCERT ? 8 : 12;
  struct Occ {
    int cus, per_cu;
  };
  static const Occ occ[3] = {
      [] { Occ o; haf_small_occupancy((const void*)haf_small_kernel<Ar, LOOP, MODE, 4>, &o.cus, &o.per_cu); return o; }(),
      [] { Occ o; haf_small_occupancy((const void*)haf_small_kernel<Ar, LOOP, MODE, 8>, &o.cus, &o.per_cu); return o; }(),
      [] { Occ o; haf_small_occupancy((const void*)haf_small_kernel<Ar, LOOP, MODE, C12>, &o.cus, &o.per_cu); return o; }()};
  for (int cls : {4, 8, 12}) {
    const Occ& oc = occ[cls / 4 - 1];
    HafSmallJobs sj{};
    int wg = 0;
    for (int j = 0; j < njobs; ++j) {
      const int N = jobs[j].N;
      if (N > cls || (cls > 4 && N <= cls - 4)) continue;
      const int k = sj.n++;
      sj.mats[k] = jobs[j].in;
      sj.out[k] = jobs[j].out;
      sj.aux[k] = jobs[j].aux;
      sj.N[k] = N;
      sj.B[k] = jobs[j].B;
      sj.T[k] = haf_small_T(jobs[j].B, N, oc.cus, oc.per_cu);
      sj.wg_begin[k] = wg;
      wg += (jobs[j].B + sj.T[k] - 1) / sj.T[k];
    }
    if (sj.n == 0 || wg == 0) continue;
    sj.wg_begin[sj.n] = wg;
    if (cls == 4)
      hipLaunchKernelGGL((haf_small_kernel<Ar, LOOP, MODE, 4>), dim3((unsigned)wg), dim3(64), 0, st, sj);
    else if (cls == 8)
      hipLaunchKernelGGL((haf_small_kernel<Ar, LOOP, MODE, 8>), dim3((unsigned)wg), dim3(64), 0, st, sj);
    else
      hipLaunchKernelGGL((haf_small_kernel<Ar, LOOP, MODE, C12>), dim3((unsigned)wg), dim3(64), 0, st, sj);
  }
}

template <class Ar, bool LOOP, int MODE, int S>
static void launch_phase(const HafJob* jobs, int njobs, typename Ar::T* const* part, double* const* part_abs,
                         hipStream_t st) {
  HafPhaseJobs<typename Ar::T> pj{};
  int wg = 0, maxN = 0;
  for (int j = 0; j < njobs; ++j) {
    if (jobs[j].N / 2 < S / 2) continue;  // this size has no subsets of size S
    const int k = pj.n++;
    pj.mats[k] = jobs[j].in;
    pj.part[k] = part[j];
    pj.part_abs[k] = part_abs[j];
    pj.N[k] = jobs[j].N;
    pj.B[k] = jobs[j].B;
    pj.wg_begin[k] = wg;
    wg += jobs[j].B * haf_nchunks(jobs[j].N / 2, S);
    maxN = jobs[j].N > maxN ? jobs[j].N : maxN;
  }
  if (pj.n == 0 || wg == 0) return;
  pj.wg_begin[pj.n] = wg;
  hipLaunchKernelGGL((haf_phase_kernel<Ar, LOOP, S, MODE>), dim3((unsigned)wg), dim3(64),
                     (size_t)maxN * maxN * sizeof(cplx), st, pj);
}

template <class Ar, bool LOOP, int MODE, int... Ss>
static void launch_phases(const HafJob* jobs, int njobs, typename Ar::T* const* part, double* const* part_abs,
                          hipStream_t st, std::integer_sequence<int, Ss...>) {
  int maxn = 0;
  for (int j = 0; j < njobs; ++j) maxn = jobs[j].N / 2 > maxn ? jobs[j].N / 2 : maxn;
  ((Ss + 1 <= maxn ? (launch_phase<Ar, LOOP, MODE, 2 * (Ss + 1)>(jobs, njobs, part, part_abs, st), 0) : 0), ...);
}

// jobs: even N >= 2 and B >= 1 each, at most HAF_MAX_JOBS
template <class Ar, bool LOOP, int MODE>
int run_haf_multi(const HafJob* all, int nall, hipStream_t st) {
  using T = typename Ar::T;
  // small sizes of every mode: one fused launch
  HafJob jobs[HAF_MAX_JOBS];
  int njobs = 0;
  if constexpr (MODE == HM_PLAIN || MODE == HM_KAPPA || MODE == HM_DDCERT ||
                (MODE == HM_CERT && std::is_same<Ar, ArF64>::value)) {
    HafJob small[HAF_MAX_JOBS];
    int ns = 0;
    const int cap = MODE == HM_DDCERT                                            ? HAF_SMALL_DDCERT_MAX_N
                    : (LOOP && MODE == HM_PLAIN && std::is_same<Ar, ArDD>::value) ? HAF_SMALL_LHAF_DD_MAX_N
                                                                                    : HAF_SMALL_MAX_N;
    const int small_n = haf_small_max_n < cap ? haf_small_max_n : cap;
    for (int j = 0; j < nall; ++j) (all[j].N <= small_n ? small[ns++] : jobs[njobs++]) = all[j];
    if (ns) launch_small<Ar, LOOP, MODE>(small, ns, st);
    if (njobs == 0) return (int)hipGetLastError();
  } else {
    for (int j = 0; j < nall; ++j) jobs[njobs++] = all[j];
  }
  T* part[HAF_MAX_JOBS];
  double* part_abs[HAF_MAX_JOBS];
  HafReduceJobs<T> rj{};
  int bsum = 0;
  for (int j = 0; j < njobs; ++j) {
    part[j] = reinterpret_cast<T*>(jobs[j].ws);
    part_abs[j] = reinterpret_cast<double*>(part[j] + (size_t)jobs[j].B * haf_nparts(jobs[j].N / 2));
    rj.part[j] = part[j];
    rj.part_abs[j] = part_abs[j];
    rj.out[j] = jobs[j].out;
    rj.aux[j] = jobs[j].aux;
    rj.n[j] = jobs[j].N / 2;
    rj.b_begin[j] = bsum;
    bsum += jobs[j].B;
  }
  rj.njobs = njobs;
  rj.b_begin[njobs] = bsum;
  launch_phases<Ar, LOOP, MODE>(jobs, njobs, part, part_abs, st, std::make_integer_sequence<int, HAF_MAX_PAIRS>{});
  hipLaunchKernelGGL((haf_reduce_kernel<Ar, MODE>), dim3((bsum + HAF_REDUCE_WAVES - 1) / HAF_REDUCE_WAVES),
                     dim3(64 * HAF_REDUCE_WAVES), 0, st, rj);
  return (int)hipGetLastError();
}

// explicit instantiations live in haf_<tier>_<kind>.hip, one translation unit
// each so that the (heavily unrolled) kernels compile in parallel
extern template int run_haf_multi<ArF64, false, HM_PLAIN>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArF64, true, HM_PLAIN>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArF64, false, HM_KAPPA>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArF64, true, HM_KAPPA>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArF64, false, HM_CERT>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArF64, true, HM_CERT>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArDD, false, HM_PLAIN>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArDD, true, HM_PLAIN>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArDD, false, HM_DDCERT>(const HafJob*, int, hipStream_t);
extern template int run_haf_multi<ArDD, true, HM_DDCERT>(const HafJob*, int, hipStream_t);

}  // namespace gbsk

// sieve.hip -- repeated-row loop hafnian by the finite-difference sieve, fp64 and
// certified fp64 (core/repeated.cu:45-150 and :197-337 compute the same sums
// one pattern per thread; cpu_ref/repeated.py:61-122 states the identity).
//
//   lhaf(A, gamma, n) = (N! 2^N)^-1 sum_{v <= n} (-1)^|v| prod C(n_k, v_k) S_N(x_v),
//   x_v = n - 2v,  S_N = sum_{k <= K} c_k sigma^k mu^(N - 2k),  K = N / 2,
//   sigma = x^T A x,  mu = x . gamma,  c_k = C(N, 2k) (2k - 1)!!.
//
// Work split (sieve.hpp states the index space). One wave per workgroup; a
// workgroup owns 64 J consecutive outer configurations of one pattern, lane l the
// J configurations starting at l J. Per outer configuration a lane forms the
// outer parts of sigma and mu once (O(m^2)); per inner configuration (a table in
// LDS built once per workgroup) it completes sigma with q fused multiply-adds
// and evaluates S_N by a two-variable Horner scheme in (sigma, mu^2):
//   H = c_K;  for k = K-1..0:  W = W mu^2,  H = H sigma + c_k W;  S = H (mu if N odd)
// (2 complex products + 1 real-scaled add per k, against the reference's 3).
// Lane sums, a fixed butterfly and an in-order reduction of the chunk partials
// give a value that depends on the pattern only.
//
// Certified tier (same value path, bit for bit): cert.hpp's model. sigma and mu
// are integer-weighted dot products, bounded a priori by gamma(L) times their
// magnitude mass; the Horner value of S at the computed (sigma, mu^2) is within
// gamma(6K + 3) P(|sigma|, |mu^2|) of its exact value on those arguments
// (P = sum c_k x^k y^(K-k), the magnitude polynomial, every monomial passing
// through at most 2K + 1 complex roundings of <= 3u each), and the exact S moves
// by at most P(|sigma| + e_sigma, |mu^2| + e_m) - P(|sigma|, |mu^2|) when its
// arguments move by (e_sigma, e_m). Both P are evaluated in round-to-nearest with
// the difference's own slack charged (gamma(12K + 32) covers Horner rounding and
// both evaluations). Coefficients and binomial weights carry their relative
// rounding. The kernel refuses nothing the fp64 tier evaluates: an overflowing
// pattern yields a non-finite bound, which the reduction turns into +inf.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "cert.hpp"
#include "common.hpp"
#include "errors.hpp"
#include "kernels.hpp"
#include "sieve.hpp"

namespace gbsk {

// C(n, k) for n <= 64, exact in uint64 and rounded once to binary64 (exact
// below 2^53)
struct BinomTab {
  double v[65 * 65];
};
constexpr BinomTab make_binom_tab() {
  BinomTab t{};
  uint64_t row[65] = {};
  for (int n = 0; n <= 64; ++n) {
    row[n] = 1;
    for (int k = n - 1; k >= 1; --k) row[k] += row[k - 1];
    for (int k = 0; k <= n; ++k) t.v[n * 65 + k] = (double)row[k];
  }
  return t;
}
__constant__ BinomTab kBinom = make_binom_tab();

// gamma_k = k u / (1 - k u) <= k u (1 + 1e-6) for k u <= 1e-7
__device__ inline double rep_gamma(double k) { return k * CERT_U * 1.000001; }
constexpr double REP_ABS_UP = 1.0 + 4.0 * CERT_U;  // |z|_1 computed, rounded up

struct SieveArgs {
  const cplx* A;
  const cplx* g;
  const int* reps;
  const RepPat* pats;
  cplx* part;
  double* part_e;
  int M;
  int p0, p1;   // pattern range of this launch
  int64_t wg0;  // batch-wide index of this launch's first workgroup
  int mm;       // largest m of the range (dynamic LDS layout)
};

// dynamic LDS: packed upper triangle of the compressed B (B_kk = A_kk,
// B_kl = A_kl + A_lk), its 1-norms (certified), the lanes' outer digits
static size_t sieve_lds_bytes(int mm, bool cert) {
  const size_t P2 = (size_t)mm * (mm + 1) / 2;
  return P2 * (sizeof(cplx) + (cert ? sizeof(double) : 0)) + 64 * (size_t)mm + 16;
}

template <bool CERT>
__device__ __forceinline__ void sieve_chunk_body(const SieveArgs& a, const int p, const int chunk) {
  extern __shared__ __align__(16) unsigned char dyn[];
  __shared__ cplx gs[REP_MAX_M];
  __shared__ double gN[REP_MAX_M];
  __shared__ int nk[REP_MAX_M], dmap[REP_MAX_M], offk[REP_MAX_M];
  __shared__ double coef[REP_MAX_N / 2 + 1];
  __shared__ double xI[REP_MAX_Q * REP_TI_MAX];
  __shared__ cplx qI[REP_TI_MAX], muI[REP_TI_MAX];
  __shared__ double cI[REP_TI_MAX], mqI[REP_TI_MAX], mmuI[REP_TI_MAX];
  const int P2 = a.mm * (a.mm + 1) / 2;
  cplx* Bs = reinterpret_cast<cplx*>(dyn);
  double* BsN = reinterpret_cast<double*>(dyn + (size_t)P2 * sizeof(cplx));
  unsigned char* dig = dyn + (size_t)P2 * (sizeof(cplx) + (CERT ? sizeof(double) : 0));

  const int lane = threadIdx.x;
  const RepPat P = a.pats[p];
  const int M = a.M, m = P.m, q = P.q, TI = P.TI, N = P.N, K = N / 2;

  // nonzero counts in ascending index order
  const int ni = lane < M ? a.reps[(size_t)p * M + lane] : 0;
  const uint64_t nz = __ballot(ni > 0);
  if (ni > 0) {
    const int k = __popcll(nz & ((1ull << lane) - 1ull));
    nk[k] = ni;
    dmap[k] = lane;
  }
  if (lane < m) offk[lane] = lane * m - lane * (lane - 1) / 2;
  __syncthreads();
  for (int k = 0; k < m; ++k) {
    const int l = k + lane;
    if (l < m) {
      const int dk = dmap[k], dl = dmap[l];
      const cplx v = (l == k) ? a.A[dk * M + dk] : c_add(a.A[dk * M + dl], a.A[dl * M + dk]);
      Bs[offk[k] + (l - k)] = v;
      if constexpr (CERT) BsN[offk[k] + (l - k)] = cert_norm1(v);
    }
  }
  if (lane < m) {
    const cplx v = a.g[dmap[lane]];
    gs[lane] = v;
    gN[lane] = cert_norm1(v);
  }
  if (lane <= K) {
    double df = 1.0;  // (2k - 1)!!
    for (int t = 3; t < 2 * lane; t += 2) df *= (double)t;
    coef[lane] = kBinom.v[N * 65 + 2 * lane] * df;
  }
  __syncthreads();

  // inner table: configuration i of the first q digits
  if (lane < TI) {
    int r = lane, par = 0;
    double c = 1.0;
    for (int j = 0; j < q; ++j) {
      const int rad = nk[j] + 1, v = r % rad;
      r /= rad;
      xI[j * REP_TI_MAX + lane] = (double)(nk[j] - 2 * v);
      c *= kBinom.v[nk[j] * 65 + v];
      par ^= v & 1;
    }
    cplx s{0.0, 0.0}, mu{0.0, 0.0};
    double ms = 0.0, mmu = 0.0;
    for (int j = 0; j < q; ++j) {
      const double xj = xI[j * REP_TI_MAX + lane];
      cplx t = c_scale(Bs[offk[j]], xj);
      double mt = 0.0;
      if constexpr (CERT) mt = fabs(xj) * BsN[offk[j]];
      for (int l = j + 1; l < q; ++l) {
        const double xl = xI[l * REP_TI_MAX + lane];
        t = c_axpy(xl, Bs[offk[j] + l - j], t);
        if constexpr (CERT) mt = fma(fabs(xl), BsN[offk[j] + l - j], mt);
      }
      s = c_axpy(xj, t, s);
      mu = c_axpy(xj, gs[j], mu);
      if constexpr (CERT) {
        ms = fma(fabs(xj), mt, ms);
        mmu = fma(fabs(xj), gN[j], mmu);
      }
    }
    qI[lane] = s;
    muI[lane] = mu;
    cI[lane] = par ? -c : c;
    mqI[lane] = ms;
    mmuI[lane] = mmu;
  }
  __syncthreads();

  const int J = P.J;
  const int64_t T = P.T, Th = (T + 1) / 2;
  const int64_t ob = (int64_t)chunk * 64 * J + (int64_t)lane * J;
  const int64_t oe = min(ob + (int64_t)J, P.TOh);
  // certified constants (see the header)
  const double g_sig = rep_gamma(2.0 * m + 16.0), g_mu = rep_gamma(m + 4.0);
  const double g_hor = rep_gamma(12.0 * K + 32.0), rho_c = rep_gamma(K + 2.0), rho_k = rep_gamma(2.0 * m + 2.0);
  cplx acc{0.0, 0.0};
  double eacc = 0.0;
  if (ob < oe) {
    {
      int64_t o = ob;
      for (int k = q; k < m; ++k) {
        const int rad = nk[k] + 1;
        dig[(k - q) * 64 + lane] = (unsigned char)(o % rad);
        o /= rad;
      }
    }
    for (int64_t o = ob; o < oe; ++o) {
      // outer parts: s_O = x_O^T B x_O, r_j = sum_k B_jk x_k (j inner), mu_O, weight
      cplx sO{0.0, 0.0}, muO{0.0, 0.0};
      double msO = 0.0, mmuO = 0.0, cO = 1.0;
      int par = 0;
      cplx r[REP_MAX_Q];
      double mr[REP_MAX_Q];
#pragma unroll
      for (int j = 0; j < REP_MAX_Q; ++j) {
        r[j] = cplx{0.0, 0.0};
        mr[j] = 0.0;
      }
      for (int k = q; k < m; ++k) {
        const int v = dig[(k - q) * 64 + lane];
        const double xk = (double)(nk[k] - 2 * v);
        cO *= kBinom.v[nk[k] * 65 + v];
        par ^= v & 1;
        cplx t = c_scale(Bs[offk[k]], xk);
        double mt = 0.0;
        if constexpr (CERT) mt = fabs(xk) * BsN[offk[k]];
        for (int l = k + 1; l < m; ++l) {
          const double xl = (double)(nk[l] - 2 * dig[(l - q) * 64 + lane]);
          t = c_axpy(xl, Bs[offk[k] + l - k], t);
          if constexpr (CERT) mt = fma(fabs(xl), BsN[offk[k] + l - k], mt);
        }
        sO = c_axpy(xk, t, sO);
        muO = c_axpy(xk, gs[k], muO);
        if constexpr (CERT) {
          msO = fma(fabs(xk), mt, msO);
          mmuO = fma(fabs(xk), gN[k], mmuO);
        }
#pragma unroll
        for (int j = 0; j < REP_MAX_Q; ++j) {
          if (j < q) {
            r[j] = c_axpy(xk, Bs[offk[j] + k - j], r[j]);
            if constexpr (CERT) mr[j] = fma(fabs(xk), BsN[offk[j] + k - j], mr[j]);
          }
        }
      }
      if (par) cO = -cO;
      const int64_t tb = o * TI;
      for (int i = 0; i < TI; ++i) {
        const int64_t t = tb + i;
        if (t >= Th) break;  // the half ends inside this outer configuration
        cplx sig = c_add(sO, qI[i]);
        double msig = 0.0;
        if constexpr (CERT) msig = msO + mqI[i];
#pragma unroll
        for (int j = 0; j < REP_MAX_Q; ++j) {
          if (j < q) {
            const double xj = xI[j * REP_TI_MAX + i];
            sig = c_axpy(xj, r[j], sig);
            if constexpr (CERT) msig = fma(fabs(xj), mr[j], msig);
          }
        }
        const cplx mu = c_add(muO, muI[i]);
        const cplx m2 = c_mul(mu, mu);
        cplx H{coef[K], 0.0}, W{1.0, 0.0};
        double x = 0.0, y = 0.0, xb = 0.0, yb = 0.0, h = 0.0, hb = 0.0, Y = 1.0, Yb = 1.0, emu = 0.0;
        if constexpr (CERT) {
          const double esig = g_sig * msig;
          emu = g_mu * (mmuO + mmuI[i]);
          const double em = cert_mul(mu, emu, mu, emu);
          x = cert_norm1(sig) * REP_ABS_UP;
          y = cert_norm1(m2) * REP_ABS_UP;
          xb = x + esig;
          yb = y + em;
          h = hb = coef[K];
        }
        for (int k = K - 1; k >= 0; --k) {
          W = c_mul(W, m2);
          H = c_axpy(coef[k], W, c_mul(H, sig));
          if constexpr (CERT) {
            Y *= y;
            Yb *= yb;
            h = fma(h, x, coef[k] * Y);
            hb = fma(hb, xb, coef[k] * Yb);
          }
        }
        double eS = 0.0;
        if constexpr (CERT) eS = (hb - h) + g_hor * (hb + h) + rho_c * hb;
        if (N & 1) {
          if constexpr (CERT) eS = cert_mul(H, eS, mu, emu);
          H = c_mul(H, mu);
        }
        double kap = cO * cI[i];
        if (2 * t + 1 != T) kap *= 2.0;  // t and its mirror T - 1 - t
        acc = c_axpy(kap, H, acc);
        if constexpr (CERT) eacc = cert_round(acc, eacc + fabs(kap) * fma(rho_k, cert_norm1(H), eS));
      }
      for (int k = q; k < m; ++k) {  // next outer configuration
        unsigned char* d = dig + (k - q) * 64 + lane;
        if (*d < nk[k]) {
          ++*d;
          break;
        }
        *d = 0;
      }
    }
  }
  if constexpr (CERT) cert_group_sum<64>(acc, eacc);
  else acc = group_sum<64>(acc);
  if (lane == 0) {
    a.part[P.part + chunk] = acc;
    if constexpr (CERT) a.part_e[P.part + chunk] = eacc;
  }
}

// the last pattern of [lo, hi] whose first workgroup is <= wg
__device__ inline int sieve_find_pattern(const RepPat* pats, int lo, int hi, int64_t wg) {
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (pats[mid].wg_begin <= wg) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// host-planned launch: workgroup blockIdx.x of patterns [p0, p1)
template <bool CERT>
__global__ __launch_bounds__(64) void sieve_chunk_kernel(const SieveArgs a) {
  const int64_t wg = a.wg0 + blockIdx.x;
  const int p = sieve_find_pattern(a.pats, a.p0, a.p1 - 1, wg);
  sieve_chunk_body<CERT>(a, p, (int)(wg - a.pats[p].wg_begin));
}

// device-planned launch: a fixed grid walks the *total work items the plan
// kernels counted (every workgroup exits once its index passes the total)
template <bool CERT>
__global__ __launch_bounds__(64) void sieve_chunk_loop_kernel(const SieveArgs a, const int64_t* __restrict__ total) {
  const int64_t tot = *total;
  for (int64_t wg = blockIdx.x; wg < tot; wg += gridDim.x) {
    const int p = sieve_find_pattern(a.pats, a.p0, a.p1 - 1, wg);
    sieve_chunk_body<CERT>(a, p, (int)(wg - a.pats[p].wg_begin));
    __syncthreads();
  }
}

// device plan: one thread per pattern; partial slots at a fixed stride chunk_cap
__global__ void sieve_plan_kernel(const int* __restrict__ reps, int M, int B, int chunk_cap, RepPat* __restrict__ pats) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  RepPat P;
  sieve_plan_one(reps + (size_t)b * M, M, chunk_cap, P);
  P.part = (int64_t)b * chunk_cap;
  P.wg_begin = 0;
  pats[b] = P;
}

// exclusive scan of the patterns' workgroup counts (one workgroup of 1024
// threads, contiguous segments per thread) and their total
__global__ __launch_bounds__(1024) void sieve_scan_kernel(RepPat* __restrict__ pats, int B, int64_t* __restrict__ total) {
  __shared__ int64_t seg[1024];
  const int t = threadIdx.x;
  const int per = (B + 1023) / 1024;
  const int b0 = t * per, b1 = min(B, b0 + per);
  int64_t s = 0;
  for (int b = b0; b < b1; ++b) s += pats[b].chunks;
  seg[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t run = 0;
    for (int i = 0; i < 1024; ++i) {
      const int64_t v = seg[i];
      seg[i] = run;
      run += v;
    }
    *total = run;
  }
  __syncthreads();
  int64_t run = seg[t];
  for (int b = b0; b < b1; ++b) {
    pats[b].wg_begin = run;
    run += pats[b].chunks;
  }
}

// one wave per pattern: chunk partials in order (lane-strided, then a fixed
// butterfly), then the division by N! 2^N
template <bool CERT>
__global__ __launch_bounds__(256) void sieve_reduce_kernel(const RepPat* __restrict__ pats, int B,
                                                           const cplx* __restrict__ part,
                                                           const double* __restrict__ part_e, cplx* __restrict__ out,
                                                           double* __restrict__ bound) {
  const int p = blockIdx.x * 4 + (threadIdx.x >> 6);  // wave-uniform
  const int lane = threadIdx.x & 63;
  if (p >= B) return;
  const RepPat P = pats[p];
  if (P.status != REP_EVAL) {
    if (lane == 0) {
      out[p] = P.status == REP_ONE ? cplx{1.0, 0.0} : cplx{NAN, NAN};
      if (CERT) bound[p] = P.status == REP_ONE ? 0.0 : INFINITY;
    }
    return;
  }
  cplx s{0.0, 0.0};
  double e = 0.0;
  for (int c = lane; c < P.chunks; c += 64) {
    s = c_add(s, part[P.part + c]);
    if constexpr (CERT) e = cert_round(s, e + part_e[P.part + c]);
  }
  if constexpr (CERT) cert_group_sum<64>(s, e);
  else s = group_sum<64>(s);
  if (lane != 0) return;
  double f = 1.0;  // N!, relative error <= gamma(N)
  for (int t = 2; t <= P.N; ++t) f *= (double)t;
  const double D = ldexp(f, P.N);
  const cplx v{s.re / D, s.im / D};
  out[p] = v;
  if constexpr (CERT) {
    const double gN = rep_gamma(P.N + 2.0);
    const double b = cert_finish(fma(gN, cert_norm1(s), e * (1.0 + gN)) / D + CERT_U * cert_norm1(v), P.depth);
    const bool ok = (b == b) && b < INFINITY && v.re == v.re && v.im == v.im && fabs(v.re) < INFINITY &&
                    fabs(v.im) < INFINITY;
    bound[p] = ok ? b : INFINITY;
  }
}

int sieve_plan(int M, const int* reps, int B, std::vector<RepPat>& plan, int64_t* parts) {
  plan.assign(B > 0 ? (size_t)B : 0, RepPat{});
  int64_t wg = 0;
  for (int b = 0; b < B; ++b) {
    RepPat& P = plan[b];
    sieve_plan_one(reps + (size_t)b * M, M, REP_CHUNKS_MAX, P);
    P.wg_begin = wg;
    P.part = wg;
    wg += P.chunks;
  }
  *parts = wg;
  return 0;
}

size_t sieve_partial_bytes(bool cert, int64_t parts) {
  return (size_t)parts * (sizeof(cplx) + (cert ? sizeof(double) : 0));
}

// d_part: sieve_partial_bytes(cert, parts) bytes; plan: host copy of d_plan
int sieve_launch(bool cert, const cplx* d_A, const cplx* d_g, int M, const int* d_reps, int B,
                 const std::vector<RepPat>& plan, const RepPat* d_plan, int64_t parts, void* d_part, cplx* d_out,
                 double* d_bound, hipStream_t st) {
  if (B <= 0) return 0;
  if (cert && !d_bound) return GBSK_ERR_ARG;
  SieveArgs a{};
  a.A = d_A;
  a.g = d_g;
  a.reps = d_reps;
  a.pats = d_plan;
  a.part = reinterpret_cast<cplx*>(d_part);
  a.part_e = reinterpret_cast<double*>(reinterpret_cast<char*>(d_part) + (size_t)parts * sizeof(cplx));
  a.M = M;
  // launches of at most 2^26 workgroups (64 threads each: 2^32 threads)
  const int64_t max_wg = (int64_t)1 << 26;
  int p = 0;
  while (p < B) {
    int64_t wg = 0;
    int mm = 1, p1 = p;
    while (p1 < B && (wg + plan[p1].chunks <= max_wg || p1 == p)) {
      wg += plan[p1].chunks;
      if (plan[p1].chunks) mm = std::max(mm, plan[p1].m);
      ++p1;
    }
    if (wg > 0) {
      a.p0 = p;
      a.p1 = p1;
      a.wg0 = plan[p].wg_begin;
      a.mm = mm;
      const size_t lds = sieve_lds_bytes(mm, cert);
      if (cert)
        hipLaunchKernelGGL(sieve_chunk_kernel<true>, dim3((unsigned)wg), dim3(64), lds, st, a);
      else
        hipLaunchKernelGGL(sieve_chunk_kernel<false>, dim3((unsigned)wg), dim3(64), lds, st, a);
      const int rc = (int)hipGetLastError();
      if (rc) return rc;
    }
    p = p1;
  }
  const unsigned grid = (unsigned)(((int64_t)B + 3) / 4);
  if (cert)
    hipLaunchKernelGGL(sieve_reduce_kernel<true>, dim3(grid), dim3(256), 0, st, d_plan, B, a.part, a.part_e, d_out,
                       d_bound);
  else
    hipLaunchKernelGGL(sieve_reduce_kernel<false>, dim3(grid), dim3(256), 0, st, d_plan, B, a.part, a.part_e, d_out,
                       d_bound);
  return (int)hipGetLastError();
}

// Device-planned sieve over a device reps table (B x M): plan, scan and a
// fixed grid of chunk workgroups, then the reduction; no host round trip.
// d_plan: B RepPat; d_total: one int64; d_part: B * REP_DEV_CHUNKS partials
// (plus as many doubles for the certified tier); mm >= every pattern's m.
size_t sieve_device_partial_bytes(bool cert, int B) {
  return (size_t)B * REP_DEV_CHUNKS * (sizeof(cplx) + (cert ? sizeof(double) : 0));
}

int sieve_launch_device(bool cert, const cplx* d_A, const cplx* d_g, int M, const int* d_reps, int B, int mm,
                        RepPat* d_plan, int64_t* d_total, void* d_part, cplx* d_out, double* d_bound,
                        hipStream_t st) {
  if (B <= 0) return 0;
  if (M > REP_MAX_M || mm < 1 || mm > REP_MAX_M) return GBSK_ERR_DIM;
  if (cert && !d_bound) return GBSK_ERR_ARG;
  hipLaunchKernelGGL(sieve_plan_kernel, dim3((B + 255) / 256), dim3(256), 0, st, d_reps, M, B, REP_DEV_CHUNKS, d_plan);
  hipLaunchKernelGGL(sieve_scan_kernel, dim3(1), dim3(1024), 0, st, d_plan, B, d_total);
  SieveArgs a{};
  a.A = d_A;
  a.g = d_g;
  a.reps = d_reps;
  a.pats = d_plan;
  a.part = reinterpret_cast<cplx*>(d_part);
  a.part_e = reinterpret_cast<double*>(reinterpret_cast<char*>(d_part) + (size_t)B * REP_DEV_CHUNKS * sizeof(cplx));
  a.M = M;
  a.p0 = 0;
  a.p1 = B;
  a.wg0 = 0;
  a.mm = mm;
  const size_t lds = sieve_lds_bytes(mm, cert);
  const unsigned grid = 4096;  // 16 workgroups per CU; each loops over the counted work
  if (cert)
    hipLaunchKernelGGL(sieve_chunk_loop_kernel<true>, dim3(grid), dim3(64), lds, st, a, (const int64_t*)d_total);
  else
    hipLaunchKernelGGL(sieve_chunk_loop_kernel<false>, dim3(grid), dim3(64), lds, st, a, (const int64_t*)d_total);
  const unsigned rgrid = (unsigned)(((int64_t)B + 3) / 4);
  if (cert)
    hipLaunchKernelGGL(sieve_reduce_kernel<true>, dim3(rgrid), dim3(256), 0, st, d_plan, B, a.part, a.part_e, d_out,
                       d_bound);
  else
    hipLaunchKernelGGL(sieve_reduce_kernel<false>, dim3(rgrid), dim3(256), 0, st, d_plan, B, a.part, a.part_e, d_out,
                       d_bound);
  return (int)hipGetLastError();
}

}  // namespace gbsk

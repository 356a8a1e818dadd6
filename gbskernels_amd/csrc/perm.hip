// perm.hip -- Glynn / BB-FG Gray-code permanent on gfx950 (fp64, kappa, dd).
//
// Algorithm (reference: core/permanent.cu:30-76, cpu_ref/permanent.py:108-139):
//   perm(A) = 2^-(n-1) sum_{delta, delta_0=+1} (prod_k delta_k) prod_r sum_c delta_c A[r,c]
// walked in reflected Gray-code order, so consecutive terms differ in one
// delta_col and the row sums change by +-2 A[:,col].
//
// MI355X mapping (this file's own design, not the reference's one-thread-per-
// matrix): the 2^(n-1) term space of one matrix is cut into contiguous,
// 2^M-aligned Gray ranges, one per lane; for large n the ranges of one matrix
// also span several workgroups ("chunks"; 2^M comes from a per-size table,
// perm_tune.hpp). Because every range starts at a
// multiple of 2^M, the flipped column ctz(i) of a lane's j-th step is ctz(j) for
// ALL lanes: the whole wave walks the same column sequence in lockstep. The row
// sums live in VGPRs (fully unrolled over a compile-time n); columns 1..HOT
// (flipped every 2nd, 4th, ... step) are cached in VGPRs, the rarer columns are
// read from an LDS copy of A^T with a wave-uniform address (a broadcast read).
// Lane partials are combined by a butterfly over the group's lanes, chunk
// partials by a fixed-order second pass. The chunking depends only on n and
// the tier family, so a matrix's value never depends on what else is in the
// batch.
#include "cert.hpp"
#include "common.hpp"
#include "perm.hpp"
#include "perm_tune.hpp"

namespace gbsk {

// ------------------------------------------------------------- split table
// Read with relaxed atomics by every launch; written only by the tuning entry
// point (documented: not while evaluations are in flight).
static int g_lane_log2[2][PERM_MAX_N + 1];
[[maybe_unused]] static const bool g_lane_log2_init = [] {
  for (int f = 0; f < 2; ++f)
    for (int n = 0; n <= PERM_MAX_N; ++n) g_lane_log2[f][n] = kPermLaneLog2[f][n];
  return true;
}();

template <int N, int DD>
static PermSplit split_of() {
  using C = PermCfg<N, DD>;
  const int t = __atomic_load_n(&g_lane_log2[DD][N], __ATOMIC_RELAXED);
  const int M = t < C::MLIM ? t : C::MLIM;
  return PermSplit{M, 1 << (C::MLIM - M)};
}

// rounding depth of the certificate arithmetic (cert.hpp): per lane 8 * 2^M + 16 N,
// plus the butterfly and the chunk sum
__device__ inline double perm_cert_depth(int N, int M, int W) {
  return 8.0 * (double)(1u << M) + 16.0 * N + 4.0 * W + 64.0;
}

// Per-lane Gray range [start, start + 2^M): initial row sums at gray(start),
// then 2^M - 1 single-column updates. Returns the signed, unscaled partial.
// Shared by the fp64 plain and kappa modes (identical value arithmetic).
// CERT: also carry a running bound (1-norm, round-to-nearest with a final
// inflation, see cert.hpp) of |value - exact|: er[r] for the row sums, ep for a
// product, E for the lane's partial.
template <int N, int HOT, bool KAPPA, bool CERT>
__device__ inline void glynn_range_fp64(const cplx* __restrict__ AT,  // LDS, AT[c*N + r]
                                        uint32_t start, int M, cplx& acc_out, double& abs_out) {
  cplx rs[N];
  double er[N];
  const uint32_t g = start ^ (start >> 1);
#pragma unroll
  for (int r = 0; r < N; ++r) {
    rs[r] = AT[r];  // column 0, delta_0 = +1 (exact)
    er[r] = 0.0;
  }
#pragma unroll 1
  for (int c = 1; c < N; ++c) {
    const double s = ((g >> (c - 1)) & 1u) ? -1.0 : 1.0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
      rs[r] = c_axpy(s, AT[c * N + r], rs[r]);
      if (CERT) er[r] = cert_round(rs[r], er[r]);
    }
  }
  // hot columns: 1..HOT
  cplx hot[HOT > 0 ? HOT : 1][N];
#pragma unroll
  for (int h = 0; h < HOT; ++h)
#pragma unroll
    for (int r = 0; r < N; ++r) hot[h][r] = AT[(h + 1) * N + r];

  double ep = 0.0;  // bound of the last product (CERT)
  auto product = [&]() {
    // four interleaved chains (rows r = q mod 4), then (t0 t1)(t2 t3): fixed order
    cplx t[4];
    double et[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      t[q] = q < N ? rs[q] : cplx{1.0, 0.0};
      et[q] = q < N ? er[q] : 0.0;
    }
#pragma unroll
    for (int r = 4; r < N; ++r) {
      if (CERT) et[r & 3] = cert_mul(t[r & 3], et[r & 3], rs[r], er[r]);
      t[r & 3] = c_mul(t[r & 3], rs[r]);
    }
    if constexpr (N == 1) {
      ep = et[0];
      return t[0];
    }
    if constexpr (N == 2) {
      if (CERT) ep = cert_mul(t[0], et[0], t[1], et[1]);
      return c_mul(t[0], t[1]);
    }
    if constexpr (N == 3) {
      const cplx t01 = c_mul(t[0], t[1]);
      if (CERT) ep = cert_mul(t01, cert_mul(t[0], et[0], t[1], et[1]), t[2], et[2]);
      return c_mul(t01, t[2]);
    }
    const cplx t01 = c_mul(t[0], t[1]), t23 = c_mul(t[2], t[3]);
    if (CERT) ep = cert_mul(t01, cert_mul(t[0], et[0], t[1], et[1]), t23, cert_mul(t[2], et[2], t[3], et[3]));
    return c_mul(t01, t23);
  };

  cplx acc = product();  // term j = 0: start is even, sign +
  double aacc = KAPPA ? c_abs(acc) : 0.0;
  double E = CERT ? ep : 0.0;
  const uint32_t nsteps = 1u << M;
  constexpr int U = HOT;  // static block length 2^U: steps t = 1..2^U-1 use hot columns
  if constexpr (U == 0) {
    for (uint32_t j = 1; j < nsteps; ++j) {
      const int k = __builtin_ctz(j);
      const uint32_t i = start + j;
      const double step = ((i >> (k + 1)) & 1u) ? 2.0 : -2.0;
      const cplx* col = AT + (k + 1) * N;
#pragma unroll
      for (int r = 0; r < N; ++r) {
        rs[r] = c_axpy(step, col[r], rs[r]);
        if (CERT) er[r] = cert_round(rs[r], er[r]);
      }
      cplx p = product();
      acc = (j & 1u) ? c_sub(acc, p) : c_add(acc, p);
      if (KAPPA) aacc += c_abs(p);
      if (CERT) E = cert_round(acc, E + ep);
    }
  } else {
    // blocks of 2^U steps: steps t < 2^U flip hot column ctz(t) + 1, step 2^U a
    // cold column. The last block stops before its cold step (j = 2^M lies
    // outside the range); it is peeled so that the loop body has no exit test
    // (a data-dependent break there cost ~60 registers and half the occupancy)
    const uint32_t nblocks = nsteps >> U;  // >= 1: M >= U
    auto body = [&](uint32_t jb, int t) __attribute__((always_inline)) {
        const uint32_t j = (jb << U) + (uint32_t)t;
        const uint32_t i = start + j;
        int k;
        if (t < (1 << U)) {
          k = __builtin_ctz((unsigned)t);
        } else {
          k = __builtin_ctz(j);
        }
        const double step = ((i >> (k + 1)) & 1u) ? 2.0 : -2.0;
        if (t < (1 << U)) {
#pragma unroll
          for (int r = 0; r < N; ++r) rs[r] = c_axpy(step, hot[__builtin_ctz((unsigned)t)][r], rs[r]);
        } else {
          const cplx* col = AT + (k + 1) * N;
#pragma unroll
          for (int r = 0; r < N; ++r) rs[r] = c_axpy(step, col[r], rs[r]);
        }
        if (CERT) {
#pragma unroll
          for (int r = 0; r < N; ++r) er[r] = cert_round(rs[r], er[r]);
        }
        cplx p = product();
        acc = (t & 1) ? c_sub(acc, p) : c_add(acc, p);
        if (KAPPA) aacc += c_abs(p);
        if (CERT) E = cert_round(acc, E + ep);
    };
    for (uint32_t jb = 0; jb + 1 < nblocks; ++jb) {
#pragma unroll
      for (int t = 1; t <= (1 << U); ++t) body(jb, t);
    }
#pragma unroll
    for (int t = 1; t < (1 << U); ++t) body(nblocks - 1, t);
  }
  acc_out = acc;
  abs_out = CERT ? E : aacc;
}

// Same walk in complex double-double (reference: core/permanent_dd.cu:34-73,
// cpu_ref/permanent.py:142-189). The +-2 A entries are exact in binary64, so a
// row-sum update is one dd + double add per component.
template <int N>
__device__ inline void glynn_range_dd(const cplx* __restrict__ AT, uint32_t start, int M, ddc& acc_out) {
  ddc rs[N];
  const uint32_t g = start ^ (start >> 1);
#pragma unroll
  for (int r = 0; r < N; ++r) rs[r] = ddc_from(AT[r]);
#pragma unroll 1
  for (int c = 1; c < N; ++c) {
    const double s = ((g >> (c - 1)) & 1u) ? -1.0 : 1.0;
#pragma unroll
    for (int r = 0; r < N; ++r) rs[r] = ddc_add_scaled(rs[r], s, AT[c * N + r]);
  }
  auto product = [&]() {
    // two interleaved chains (even / odd rows), then their product
    ddc t0 = rs[0];
#pragma unroll
    for (int r = 2; r < N; r += 2) t0 = ddc_mul(t0, rs[r]);
    if constexpr (N == 1) return t0;
    ddc t1 = rs[1];
#pragma unroll
    for (int r = 3; r < N; r += 2) t1 = ddc_mul(t1, rs[r]);
    return ddc_mul(t0, t1);
  };
  ddc acc = product();
  const uint32_t nsteps = 1u << M;
  for (uint32_t j = 1; j < nsteps; ++j) {
    const int k = __builtin_ctz(j);
    const uint32_t i = start + j;
    const double step = ((i >> (k + 1)) & 1u) ? 2.0 : -2.0;
    const cplx* col = AT + (k + 1) * N;
#pragma unroll
    for (int r = 0; r < N; ++r) rs[r] = ddc_add_scaled(rs[r], step, col[r]);
    ddc p = product();
    acc = (j & 1u) ? ddc_sub(acc, p) : ddc_add(acc, p);
  }
  acc_out = acc;
}

// Compiler barrier on a bound accumulator (emits no instruction). Without it
// LLVM sinks each operation's error computation to where the bound is next
// read, keeping three intermediates per row sum alive across the whole step:
// the certified-dd walk then spilled 900-1800 VGPRs to scratch (40-120x the dd
// walk's time); with it the bound is folded in as the values are formed.
__device__ __forceinline__ void pin(double& v) { asm volatile("" : "+v"(v)); }

// glynn_range_dd with a running bound (certified double-double). The values
// are glynn_range_dd's bit for bit. E bounds the error of every row sum (each
// row receives the same updates, so one bound serves all rows: each step adds
// the largest of the N update errors); products and the signed accumulation
// propagate it with the 1-norm product rule plus each operation's own error.
template <int N>
__device__ inline void glynn_range_ddcert(const cplx* __restrict__ AT, uint32_t start, int M, ddc& acc_out,
                                          double& e_out) {
  ddc rs[N];
  const uint32_t g = start ^ (start >> 1);
  double E = 0.0;
#pragma unroll
  for (int r = 0; r < N; ++r) rs[r] = ddc_from(AT[r]);
#pragma unroll 1
  for (int c = 1; c < N; ++c) {
    const double s = ((g >> (c - 1)) & 1u) ? -1.0 : 1.0;
    double m = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
      double e;
      rs[r] = cddc_add_scaled(rs[r], s, AT[c * N + r], &e);
      m = fmax(m, e);
      pin(m);
    }
    E += m;
  }
  auto product = [&](double& ep) {
    ddc t0 = rs[0];
    double e0 = E;
#pragma unroll
    for (int r = 2; r < N; r += 2) {
      double eo;
      const ddc t = cddc_mul(t0, rs[r], &eo);
      e0 = cddc_mul_bound(t0, e0, rs[r], E, eo);
      pin(e0);
      t0 = t;
    }
    if constexpr (N == 1) {
      ep = e0;
      return t0;
    }
    ddc t1 = rs[1];
    double e1 = E;
#pragma unroll
    for (int r = 3; r < N; r += 2) {
      double eo;
      const ddc t = cddc_mul(t1, rs[r], &eo);
      e1 = cddc_mul_bound(t1, e1, rs[r], E, eo);
      pin(e1);
      t1 = t;
    }
    double eo;
    const ddc t = cddc_mul(t0, t1, &eo);
    ep = cddc_mul_bound(t0, e0, t1, e1, eo);
    return t;
  };
  double ea;
  ddc acc = product(ea);
  const uint32_t nsteps = 1u << M;
  for (uint32_t j = 1; j < nsteps; ++j) {
    const int k = __builtin_ctz(j);
    const uint32_t i = start + j;
    const double step = ((i >> (k + 1)) & 1u) ? 2.0 : -2.0;
    const cplx* col = AT + (k + 1) * N;
    double m = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r) {
      double e;
      rs[r] = cddc_add_scaled(rs[r], step, col[r], &e);
      m = fmax(m, e);
      pin(m);
    }
    E += m;
    double ep, eo;
    const ddc p = product(ep);
    acc = (j & 1u) ? cddc_sub(acc, p, &eo) : cddc_add(acc, p, &eo);
    ea = ea + ep + eo;
  }
  acc_out = acc;
  e_out = ea;
}

// butterfly sum over G lanes of (ddc value, bound): group_sum<G>(ddc)'s order
template <int G>
__device__ inline void cddc_group_sum(ddc& v, double& e) {
#pragma unroll
  for (int m = 1; m < G; m <<= 1) {
    ddc o;
    o.re.hi = shfl_xor_d(v.re.hi, m);
    o.re.lo = shfl_xor_d(v.re.lo, m);
    o.im.hi = shfl_xor_d(v.im.hi, m);
    o.im.lo = shfl_xor_d(v.im.lo, m);
    const double eo = shfl_xor_d(e, m);
    const bool low = (threadIdx.x & m) == 0;
    double r;
    v = low ? cddc_add(v, o, &r) : cddc_add(o, v, &r);
    e = e + eo + r;
  }
}

// rounding depth of the certified-dd bound: the row bound's 2^M-step
// accumulation, the product chains, the butterfly and the chunk sum
__device__ inline double perm_ddcert_depth(int N, int M, int W) {
  return 8.0 * (double)(1u << M) + 16.0 * N + 8.0 * W + 128.0;
}

// One workgroup = one wave = MPW matrices x LPM lanes (or one chunk of one
// matrix when the matrix is split over W workgroups).
//   out      : final values (W == 1) or chunk partials [B][W] (W > 1)
//   aux      : kappa sum|term| (final or partial), may be null in plain mode
//   M, W     : the run-time split (log2 terms per lane, workgroups per matrix)
// fp64, kappa and dd ask for at least two waves per SIMD: without the hint the
// fp64 kernel takes 256 VGPRs plus a few AGPRs at n = 15..25 (one wave per
// SIMD); with it n = 14..18 fit in 202..250 VGPRs and n = 19..25 spill 1..6
// VGPRs. The certified modes carry a bound per row sum and would spill dozens
// of VGPRs to scratch: they keep one wave per SIMD and AGPR spill space.
template <int MODE>
constexpr int perm_min_waves() { return (MODE == PERM_CERT || MODE == PERM_DDCERT) ? 1 : 2; }

template <int N, int MODE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(perm_min_waves<MODE>()))) void perm_kernel(const cplx* __restrict__ mats, int B, cplx* __restrict__ out,
                                                  double* __restrict__ aux, int g8, int M, int W) {
  using C = PermCfg<N, perm_is_dd(MODE) ? 1 : 0>;
  constexpr int LPM = C::LPM, MPW = 64 / LPM;
  __shared__ cplx AT[MPW][N * N];
  const int lane = threadIdx.x;
  const int gidx = lane / LPM, sub = lane % LPM;
  // workgroup -> (matrix group, chunk). Workgroups are dealt to the 8 XCDs in
  // id order, so for a split matrix (W > 1) the id is decoded such that all W
  // chunks of one group share an XCD (its L2 then fetches the matrix once).
  const int64_t wg = blockIdx.x;
  int64_t grp;
  int chunk;
  if (W == 1) {
    grp = wg;
    chunk = 0;
  } else {
    // g8 = min(8, groups): with fewer than 8 groups no padding workgroups
    const int64_t k = wg / g8;
    grp = (k / W) * g8 + wg % g8;
    chunk = (int)(k % W);
  }
  const int64_t mat0 = grp * MPW;
  if (mat0 >= B) return;  // padding workgroup (uniform exit, before any barrier)
  // stage MPW matrices transposed into LDS
  for (int e = lane; e < MPW * N * N; e += 64) {
    const int q = e / (N * N), rc = e % (N * N);
    const int r = rc / N, c = rc % N;
    const int64_t b = mat0 + q;
    AT[q][c * N + r] = (b < B) ? mats[b * (N * N) + rc] : cplx{0.0, 0.0};
  }
  __syncthreads();
  const int64_t b = mat0 + gidx;
  const uint32_t start = ((uint32_t)chunk * LPM + (uint32_t)sub) << M;
  if constexpr (perm_is_dd(MODE)) {
    constexpr bool DDCERT = MODE == PERM_DDCERT;
    ddc acc;
    double ea = 0.0;
    if constexpr (DDCERT) {
      glynn_range_ddcert<N>(AT[gidx], start, M, acc, ea);
      cddc_group_sum<LPM>(acc, ea);
    } else {
      glynn_range_dd<N>(AT[gidx], start, M, acc);
      acc = group_sum<LPM>(acc);
    }
    if (sub == 0 && b < B) {
      if (W == 1) {
        const double scale = 1.0 / (double)(1ull << (N - 1));
        const ddc sc = ddc_mul_d(acc, scale);  // exact (power of two)
        out[b] = ddc_to_c(sc);
        if constexpr (DDCERT)
          aux[b] = cert_finish(fma(ea, scale, cddc_collapse_err(sc) + 4.0 * CERT_UFL), perm_ddcert_depth(N, M, W));
      } else {
        // chunk partials stay double-double: hi in out, lo in aux-as-cplx,
        // then (certified-dd) the bounds after the B * W lo entries
        out[b * W + chunk] = cplx{acc.re.hi, acc.im.hi};
        reinterpret_cast<cplx*>(aux)[b * W + chunk] = cplx{acc.re.lo, acc.im.lo};
        if constexpr (DDCERT) reinterpret_cast<double*>(reinterpret_cast<cplx*>(aux) + (size_t)B * W)[b * W + chunk] = ea;
      }
    }
  } else {
    constexpr bool KAPPA = (MODE == PERM_KAPPA);
    constexpr bool CERT = (MODE == PERM_CERT);
    cplx acc;
    double aacc;
    glynn_range_fp64<N, C::HOT, KAPPA, CERT>(AT[gidx], start, M, acc, aacc);
    if constexpr (CERT) {
      cert_group_sum<LPM>(acc, aacc);
    } else {
      acc = group_sum<LPM>(acc);
      if (KAPPA) aacc = group_sum<LPM>(aacc);
    }
    if (sub == 0 && b < B) {
      if (W == 1) {
        const double scale = 1.0 / (double)(1ull << (N - 1));
        out[b] = c_scale(acc, scale);
        if (KAPPA) aux[b] = aacc * scale;
        if (CERT) aux[b] = cert_finish(aacc * scale + CERT_UFL, perm_cert_depth(N, M, W));
      } else {
        out[b * W + chunk] = acc;
        if (KAPPA || CERT) aux[b * W + chunk] = aacc;
      }
    }
  }
}

// Fixed-order (chunk 0, 1, ..., W-1) reduction of chunk partials.
template <int N, int MODE>
__global__ void perm_reduce_kernel(const cplx* __restrict__ part, const double* __restrict__ part_aux, int B,
                                   cplx* __restrict__ out, double* __restrict__ aux, int M, int W) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  const double scale = 1.0 / (double)(1ull << (N - 1));
  if constexpr (perm_is_dd(MODE)) {
    const cplx* lo = reinterpret_cast<const cplx*>(part_aux);
    const double* pe = reinterpret_cast<const double*>(lo + (size_t)B * W);
    ddc s = ddc_zero();
    double e = 0.0;
    for (int w = 0; w < W; ++w) {
      cplx h = part[b * W + w], l = lo[b * W + w];
      if constexpr (MODE == PERM_DDCERT) {
        double r;
        s = cddc_add(s, ddc{dd{h.re, l.re}, dd{h.im, l.im}}, &r);
        e = e + pe[b * W + w] + r;
      } else {
        s = ddc_add(s, ddc{dd{h.re, l.re}, dd{h.im, l.im}});
      }
    }
    const ddc sc = ddc_mul_d(s, scale);
    out[b] = ddc_to_c(sc);
    if constexpr (MODE == PERM_DDCERT)
      aux[b] = cert_finish(fma(e, scale, cddc_collapse_err(sc) + 4.0 * CERT_UFL), perm_ddcert_depth(N, M, W));
  } else {
    cplx s{0.0, 0.0};
    double a = 0.0;
    for (int w = 0; w < W; ++w) {
      s = c_add(s, part[b * W + w]);
      if (MODE == PERM_KAPPA) a += part_aux[b * W + w];
      if (MODE == PERM_CERT) a = cert_round(s, a + part_aux[b * W + w]);
    }
    out[b] = c_scale(s, scale);
    if (MODE == PERM_KAPPA) aux[b] = a * scale;
    if (MODE == PERM_CERT) aux[b] = cert_finish(a * scale + CERT_UFL, perm_cert_depth(N, M, W));
  }
}

static size_t perm_chunk_bytes(int mode) { return mode == PERM_DDCERT ? 16 + 16 + 8 : 16 + 16; }

template <int N, int MODE>
static int launch_perm_n(const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, size_t ws_size,
                         hipStream_t st) {
  using C = PermCfg<N, perm_is_dd(MODE) ? 1 : 0>;
  constexpr int MPW = 64 / C::LPM;
  // the split is read once: the workspace check and the launches agree
  const PermSplit sp = split_of<N, perm_is_dd(MODE) ? 1 : 0>();
  const int M = sp.M, W = sp.W;
  if (W > 1 && ws_size < (size_t)B * W * perm_chunk_bytes(MODE)) return GBSK_ERR_WORKSPACE;
  // at most 2^30 workgroups per launch: large batches run as consecutive
  // sub-batches (same stream, the workspace reused), matrices of a sub-batch a
  // multiple of 8 groups so the XCD mapping is unchanged
  const int64_t max_groups = (((int64_t)1 << 30) / W) / 8 * 8;
  const int64_t max_b = max_groups * MPW;
  for (int64_t b0 = 0; b0 < B; b0 += max_b) {
    const int Bs = (int)((B - b0) < max_b ? (B - b0) : max_b);
    const cplx* in = d_in + b0 * N * N;
    cplx* out = d_out + b0;
    double* aux = d_aux ? d_aux + b0 : nullptr;
    const int64_t groups = ((int64_t)Bs + MPW - 1) / MPW;
    const int g8 = groups < 8 ? (int)groups : 8;
    const int64_t grid = (W == 1) ? groups : ((groups + g8 - 1) / g8) * g8 * W;
    if (W == 1) {
      hipLaunchKernelGGL((perm_kernel<N, MODE>), dim3((unsigned)grid), dim3(64), 0, st, in, Bs, out, aux, g8, M, W);
    } else {
      // partials: values [Bs][W] cplx, then [Bs][W] of 16 bytes (kappa / bound
      // doubles, or the dd low parts), then (certified dd) [Bs][W] bounds
      cplx* part = reinterpret_cast<cplx*>(ws);
      double* part_aux = reinterpret_cast<double*>(part + (size_t)Bs * W);
      hipLaunchKernelGGL((perm_kernel<N, MODE>), dim3((unsigned)grid), dim3(64), 0, st, in, Bs, part, part_aux, g8,
                         M, W);
      hipLaunchKernelGGL((perm_reduce_kernel<N, MODE>), dim3((unsigned)((Bs + 255) / 256)), dim3(256), 0, st, part,
                         part_aux, Bs, out, aux, M, W);
    }
  }
  return (int)hipGetLastError();
}

// n == 0: perm = 1 (empty product); kappa indicator 1, certificate 0 (exact)
__global__ void perm_n0_kernel(int B, cplx* out, double* aux, double auxval) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  out[b] = cplx{1.0, 0.0};
  if (aux) aux[b] = auxval;
}



template <int MODE, int... Ns>
static int dispatch_perm(int n, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, size_t ws_size,
                         hipStream_t st, std::integer_sequence<int, Ns...>) {
  int rc = GBSK_ERR_DIM;
  ((n == Ns ? (rc = launch_perm_n<Ns, MODE>(d_in, B, d_out, d_aux, ws, ws_size, st), 0) : 0), ...);
  return rc;
}

template <int... Ns>
static PermSplit split_dispatch(int n, int family, std::integer_sequence<int, Ns...>) {
  PermSplit r{0, 1};
  ((n == Ns ? (r = family ? split_of<Ns, 1>() : split_of<Ns, 0>(), 0) : 0), ...);
  return r;
}

template <int... Is>
constexpr std::integer_sequence<int, (Is + 1)...> shift1(std::integer_sequence<int, Is...>) {
  return {};
}
using PermSizes = decltype(shift1(std::make_integer_sequence<int, PERM_MAX_N>{}));

PermSplit perm_split(int n, int family) {
  if (n < 1 || n > PERM_MAX_N) return PermSplit{0, 1};
  return split_dispatch(n, family ? 1 : 0, PermSizes{});
}

int perm_lane_log2_get(int family, int n) {
  if ((family != 0 && family != 1) || n < 0 || n > PERM_MAX_N) return GBSK_ERR_ARG;
  return __atomic_load_n(&g_lane_log2[family][n], __ATOMIC_RELAXED);
}

// at least 2^LOGMIN terms per lane (16 fp64, 4 dd: the per-lane start costs n^2
// updates) and at most 2^27 (every n <= 28 then runs one workgroup per matrix)
int perm_lane_log2_set(int family, int n, int log2_terms) {
  if ((family != 0 && family != 1) || n < 0 || n > PERM_MAX_N) return GBSK_ERR_ARG;
  if (log2_terms < (family ? 2 : 4) || log2_terms > 27) return GBSK_ERR_ARG;
  __atomic_store_n(&g_lane_log2[family][n], log2_terms, __ATOMIC_RELAXED);
  return 0;
}

size_t perm_workspace_bytes(int n, int mode, int B) {
  if (n < 1 || n > PERM_MAX_N || B <= 0) return 0;
  const PermSplit sp = perm_split(n, perm_is_dd(mode) ? 1 : 0);
  return sp.W > 1 ? (size_t)B * sp.W * perm_chunk_bytes(mode) : 0;
}

int perm_launch(int n, int mode, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, size_t ws_size,
                hipStream_t st) {
  if (B <= 0) return 0;
  if (n < 0 || n > PERM_MAX_N) return GBSK_ERR_DIM;
  if (n == 0) {
    hipLaunchKernelGGL(perm_n0_kernel, dim3((B + 255) / 256), dim3(256), 0, st, B, d_out,
                       (mode == PERM_KAPPA || mode == PERM_CERT || mode == PERM_DDCERT) ? d_aux : nullptr,
                       mode == PERM_KAPPA ? 1.0 : 0.0);
    return (int)hipGetLastError();
  }
  switch (mode) {
    case PERM_FP64: return dispatch_perm<PERM_FP64>(n, d_in, B, d_out, d_aux, ws, ws_size, st, PermSizes{});
    case PERM_KAPPA: return dispatch_perm<PERM_KAPPA>(n, d_in, B, d_out, d_aux, ws, ws_size, st, PermSizes{});
    case PERM_DD: return dispatch_perm<PERM_DD>(n, d_in, B, d_out, d_aux, ws, ws_size, st, PermSizes{});
    case PERM_CERT: return dispatch_perm<PERM_CERT>(n, d_in, B, d_out, d_aux, ws, ws_size, st, PermSizes{});
    case PERM_DDCERT: return dispatch_perm<PERM_DDCERT>(n, d_in, B, d_out, d_aux, ws, ws_size, st, PermSizes{});
  }
  return GBSK_ERR_ARG;
}

}  // namespace gbsk

// sieve.hpp -- repeated-row loop hafnian (finite-difference sieve): the host
// plan shared by the launcher and the kernels.
//
// A pattern n (M counts) stands for the matrix with row/column i of the base A
// repeated n_i times and loop weight gamma_i on each copy. Its m nonzero counts,
// in ascending index order, are the digits of a mixed-radix index
//   t = sum_k v_k R_k,  R_0 = 1,  R_{k+1} = R_k (n_k + 1),  T = R_m terms,
// and t <-> T - 1 - t is the v <-> n - v pairing, whose two terms are equal: the
// kernels evaluate t < ceil(T / 2), weight 2 (weight 1 for the self-paired
// t = (T - 1) / 2 when T is odd). The first q digits form the inner block
// (TI = R_q <= 64 configurations, tabulated once per workgroup); the rest are
// the outer configurations o = t / TI, J consecutive ones per lane.
#pragma once

#include <stdint.h>

namespace gbsk {

constexpr int REP_MAX_M = 32;                       // base-matrix side
constexpr int REP_MAX_N = 64;                       // photons N = sum n_i
constexpr double REP_MAX_TERMS = 1099511627776.0;   // 2^40 sieve terms per pattern
constexpr int REP_TI_MAX = 64;                      // inner-block configurations
constexpr int REP_MAX_Q = 6;                        // inner digits (2^6 = 64)
constexpr int REP_CHUNKS_MAX = 1 << 20;             // workgroups per pattern (host plan)
constexpr int REP_DEV_CHUNKS = 32;                  // workgroups per pattern (device plan)

enum RepStatus { REP_EVAL = 0, REP_ONE = 1, REP_REFUSE = 2 };

struct RepPat {
  int64_t T;         // sieve terms prod (n_k + 1)
  int64_t TOh;       // outer configurations holding some t < ceil(T / 2)
  int64_t wg_begin;  // first workgroup of this pattern (over the whole batch)
  int64_t part;      // first chunk-partial slot
  double depth;      // rounding depth of the certified bound expression
  int status;        // RepStatus
  int m, N, q, TI;   // nonzero counts, photons, inner digits, inner configurations
  int J;             // outer configurations per lane per workgroup
  int chunks;        // workgroups of this pattern
};

// Plan of one pattern n (M counts): everything but wg_begin / part. The same
// rules on the host (chunk_cap = REP_CHUNKS_MAX) and on the device (a small cap,
// so that partial slots can have a fixed stride); the value depends on the plan.
__host__ __device__ inline void sieve_plan_one(const int* n, int M, int chunk_cap, RepPat& P) {
  P.status = REP_EVAL;
  P.T = P.TOh = 0;
  P.m = P.N = P.q = 0;
  P.TI = P.J = 1;
  P.chunks = 0;
  P.depth = 0.0;
  int N = 0, m = 0;
  double Td = 1.0;
  bool bad = M > REP_MAX_M;
  for (int i = 0; i < M && !bad; ++i) {
    if (n[i] < 0 || n[i] > REP_MAX_N) bad = true;
    N += n[i];
    m += n[i] > 0;
    Td *= (double)(n[i] + 1);
  }
  if (bad || N > REP_MAX_N || Td > REP_MAX_TERMS) {
    P.status = REP_REFUSE;
    return;
  }
  if (N == 0) {
    P.status = REP_ONE;
    return;
  }
  const int64_t T = (int64_t)Td, Th = (T + 1) / 2;
  int q = 0, TI = 1, k = 0;
  for (int i = 0; i < M && q == k; ++i) {  // inner digits: the first nonzero counts
    if (n[i] <= 0) continue;
    ++k;
    const int64_t nxt = (int64_t)TI * (n[i] + 1);
    if (nxt > REP_TI_MAX || (Th + nxt - 1) / nxt < 64) break;
    TI = (int)nxt;
    ++q;
  }
  const int64_t TOh = (Th + TI - 1) / TI;
  const int64_t jlane = 256 / TI > 1 ? 256 / TI : 1;
  int64_t J = (TOh + 63) / 64 < jlane ? (TOh + 63) / 64 : jlane;
  if (J < 1) J = 1;
  const int64_t jcap = (TOh + 64LL * chunk_cap - 1) / (64LL * chunk_cap);
  if (J < jcap) J = jcap;
  const int64_t chunks = (TOh + 64 * J - 1) / (64 * J);
  P.T = T;
  P.TOh = TOh;
  P.m = m;
  P.N = N;
  P.q = q;
  P.TI = TI;
  P.J = (int)J;
  P.chunks = (int)chunks;
  P.depth = 4.0 * m + 8.0 * (N / 2) + 256.0 + 4.0 * (double)J * TI + 4.0 * (double)chunks;
}

}  // namespace gbsk

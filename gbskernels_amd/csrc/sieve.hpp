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
constexpr int REP_CHUNKS_MAX = 1 << 20;             // workgroups per pattern

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

}  // namespace gbsk

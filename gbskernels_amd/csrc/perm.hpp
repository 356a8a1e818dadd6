// perm.hpp -- launch configuration and entry points of the Glynn permanent kernels.
#pragma once

#include <stddef.h>

#include <utility>

#include "common.hpp"
#include "errors.hpp"

namespace gbsk {

constexpr int PERM_MAX_N = 28;  // 2^27 Gray terms; 32-bit term indices
enum PermMode { PERM_FP64 = 0, PERM_DD = 1, PERM_KAPPA = 2, PERM_CERT = 3, PERM_DDCERT = 4 };
// double-double modes share one work split (so their values are identical)
constexpr bool perm_is_dd(int mode) { return mode == PERM_DD || mode == PERM_DDCERT; }

// Work split of one n x n permanent (depends on n and the tier family only,
// never on B). Compile-time part:
//   LPM  lanes per matrix (64 / LPM matrices share one wave),
//   MLIM log2 of the Gray terms per lane when one workgroup holds the matrix,
//   HOT  columns 1..HOT cached in VGPRs (fp64 only).
// Run-time part (PermSplit): M = log2 of the Gray terms per lane, from the
// tuning table (perm_tune.hpp, autotune.py) and capped at MLIM, and
// W = 2^(MLIM - M) workgroups ("chunks") per matrix.
template <int N, int DD>
struct PermCfg {
  static constexpr int LOGT = N - 1;
  static constexpr int LOGMIN = DD ? 2 : 4;  // at least 2^LOGMIN terms per lane
  static constexpr int LOGL_ = LOGT - LOGMIN < 0 ? 0 : (LOGT - LOGMIN > 6 ? 6 : LOGT - LOGMIN);
  static constexpr int LOGL = LOGL_;
  static constexpr int LPM = 1 << LOGL;
  static constexpr int MLIM = LOGT - LOGL;
  static constexpr int HOT_ = DD ? 0 : (N >= 20 ? 1 : 2);
  static constexpr int HOT = HOT_ < MLIM ? HOT_ : MLIM;
};

struct PermSplit {
  int M;  // log2 Gray terms per lane
  int W;  // workgroups per matrix
};
// family 0: fp64 / kappa / certified (one shared split: bit-identical values);
// family 1: dd / certified-dd
PermSplit perm_split(int n, int family);
// tuning table access; set returns 0 or GBSK_ERR_ARG (out of range)
int perm_lane_log2_get(int family, int n);
int perm_lane_log2_set(int family, int n, int log2_terms);

// Workspace bytes the launch needs for chunk partials (0 when W == 1).
size_t perm_workspace_bytes(int n, int mode, int B);
// Stream-ordered launch; d_aux: kappa sum|term| (PERM_KAPPA), else unused. The
// workspace is checked against the split the launch reads (GBSK_ERR_WORKSPACE).
int perm_launch(int n, int mode, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, size_t ws_size,
                hipStream_t st);

}  // namespace gbsk

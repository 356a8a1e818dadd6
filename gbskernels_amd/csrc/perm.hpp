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

// Work split of one n x n permanent (depends on n and the tier only, never on B):
//   LPM lanes per matrix (64 / LPM matrices share one wave),
//   W   workgroups (chunks) per matrix,
//   M   log2 of the Gray terms per lane,
//   HOT columns 1..HOT cached in VGPRs (fp64 only).
template <int N, int DD>
struct PermCfg {
  static constexpr int LOGT = N - 1;
  static constexpr int LOGMIN = DD ? 2 : 4;  // at least 2^LOGMIN terms per lane
  static constexpr int LOGL_ = LOGT - LOGMIN < 0 ? 0 : (LOGT - LOGMIN > 6 ? 6 : LOGT - LOGMIN);
  static constexpr int LOGL = LOGL_;
  static constexpr int LPM = 1 << LOGL;
  // at most 2^MAXM terms per lane, else split over workgroups: short enough
  // that one large matrix spreads over hundreds of waves (latency of a small
  // batch), long enough that the per-lane start (n^2 row-sum terms) stays a
  // few percent of the walk
  static constexpr int MAXM = DD ? 6 : 9;
  static constexpr int LOGW = LOGT - LOGL - MAXM > 0 ? LOGT - LOGL - MAXM : 0;
  static constexpr int W = 1 << LOGW;
  static constexpr int M = LOGT - LOGL - LOGW;
  static constexpr int HOT_ = DD ? 0 : (N >= 20 ? 1 : 2);
  static constexpr int HOT = HOT_ < M ? HOT_ : M;
};

// Workspace bytes the launch needs for chunk partials (0 when W == 1).
size_t perm_workspace_bytes(int n, int mode, int B);
// Stream-ordered launch; d_aux: kappa sum|term| (PERM_KAPPA), else unused.
int perm_launch(int n, int mode, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, hipStream_t st);

}  // namespace gbsk

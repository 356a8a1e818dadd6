// tor.hip -- torontonian of real, physical-domain O on gfx950 (fp64, kappa).
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
//    each eliminates its included modes from A (right-looking Cholesky,
//    cooperative over the wave's lanes) to get the Schur complement on the
//    remaining modes;
//  * T2 <= 6 middle modes: the wave walks their include/exclude tree together,
//    one Schur complement per tree level in LDS, every update spread over the
//    64 lanes; each of the 2^T2 leaves becomes one lane's task (its Schur
//    complement on the last R modes, the running determinant, the count);
//  * R <= 5 trailing modes: each lane walks its task's 2^R leaves alone with a
//    fully unrolled prefix-Cholesky (reference algorithm) whose factor lives in
//    VGPRs.
// Partials: lanes -> butterfly -> one partial per workgroup -> fixed-order sum.
#include "common.hpp"
#include "kernels.hpp"

namespace gbsk {

constexpr int TOR_MAX_MODES = 20;  // dim 40
constexpr int TOR_R = 5;           // per-lane trailing modes
constexpr int TOR_T2 = 6;          // cooperative middle modes (64 tasks)

struct TorSplit {
  int R, T2, T1;
};
// lanes first (T2 middle modes -> up to 64 tasks), then per-lane depth R, then
// leading patterns as separate workgroups
__host__ __device__ inline TorSplit tor_split(int nm) {
  TorSplit s;
  s.T2 = nm < TOR_T2 ? nm : TOR_T2;
  const int rest = nm - s.T2;
  s.R = rest < TOR_R ? rest : TOR_R;
  s.T1 = rest - s.R;
  return s;
}

// packed lower-triangular index
__device__ constexpr int tri(int r, int c) { return r * (r + 1) / 2 + c; }

// ----------------------------------------------------------- per-lane bottom
// Prefix-Cholesky walk over the R trailing modes of one task. M is the task's
// 2R x 2R Schur complement (packed lower, mode-major), L the factor prefix.
template <int R, int D, unsigned MASK, int C>
struct Bottom {
  static __device__ __forceinline__ void run(const double* __restrict__ M, double (&L)[2 * R][2 * R], double det,
                                             int count, int nm, bool& bad, double& acc, double& aacc) {
    // exclude mode D
    Bottom<R, D + 1, MASK, C>::run(M, L, det, count, nm, bad, acc, aacc);
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
        double s = gr >= gc ? M[tri(gr, gc)] : M[tri(gc, gr)];
#pragma unroll
        for (int u = 0; u < c; ++u) s = fma(-L[r][u], L[c][u], s);
        if (c < r) {
          L[r][c] = s / L[c][c];
        } else {
          bad |= !(s > 0.0);
          L[r][r] = sqrt(s);
        }
      }
    }
    const double q = L[2 * C][2 * C] * L[2 * C + 1][2 * C + 1];
    Bottom<R, D + 1, MASK | (1u << D), C + 1>::run(M, L, det * (q * q), count + 1, nm, bad, acc, aacc);
  }
};
template <int R, unsigned MASK, int C>
struct Bottom<R, R, MASK, C> {
  static __device__ __forceinline__ void run(const double* __restrict__, double (&)[2 * R][2 * R], double det,
                                             int count, int nm, bool&, double& acc, double& aacc) {
    const double t = 1.0 / sqrt(det);
    acc += ((nm - count) & 1) ? -t : t;
    aacc += t;
  }
};

template <int R>
__device__ inline void bottom_walk(const double* M, double det, int count, int nm, bool& bad, double& acc,
                                   double& aacc) {
  double L[2 * R > 0 ? 2 * R : 1][2 * R > 0 ? 2 * R : 1];
  if constexpr (R == 0) {
    const double t = 1.0 / sqrt(det);
    acc += ((nm - count) & 1) ? -t : t;
    aacc += t;
  } else {
    Bottom<R, 0, 0u, 0>::run(M, L, det, count, nm, bad, acc, aacc);
  }
}

// ----------------------------------------------------- cooperative helpers
// Eliminate pivot row p of the D x D symmetric matrix X (full storage, LDS) in
// place: X -= X[:,p] X[p,:] / X[p][p] on rows/cols != p. Returns the pivot.
__device__ inline double eliminate(double* X, int D, int p) {
  const double piv = X[p * D + p];
  __syncthreads();
  for (int e = threadIdx.x; e < D * D; e += 64) {
    const int i = e / D, j = e % D;
    if (i != p && j != p) X[e] = fma(-X[i * D + p] / piv, X[p * D + j], X[e]);
  }
  __syncthreads();
  return piv;
}

// Y (Dy x Dy) = X (Dx x Dx) with mode `md` (rows 2md, 2md+1) removed.
__device__ inline void drop_mode(const double* X, int Dx, int md, double* Y) {
  const int Dy = Dx - 2;
  for (int e = threadIdx.x; e < Dy * Dy; e += 64) {
    int i = e / Dy, j = e % Dy;
    const int si = i < 2 * md ? i : i + 2, sj = j < 2 * md ? j : j + 2;
    Y[e] = X[si * Dx + sj];
  }
  __syncthreads();
}

// ---------------------------------------------------------------- kernel
// workgroup = (instance b, leading pattern p1); out partials [B][2^T1]
template <int R, bool KAPPA>
__global__ __launch_bounds__(64) void tor_real_kernel(const cplx* __restrict__ mats, int nm, int B,
                                                      double* __restrict__ part, double* __restrict__ part_abs) {
  extern __shared__ double lds[];
  const TorSplit sp = tor_split(nm);
  const int T1 = sp.T1, T2 = sp.T2;
  const int dim = 2 * nm;
  const int lane = threadIdx.x;
  const int64_t wg = blockIdx.x;
  const int64_t b = wg >> T1;
  const unsigned p1 = (unsigned)(wg & ((1 << T1) - 1));
  // LDS layout: work matrix (dim^2), level stack (T2+1 levels), task slots
  double* W = lds;
  double* lev = W + dim * dim;
  int levoff[TOR_T2 + 2];
  {
    int o = 0;
    for (int d = 0; d <= T2; ++d) {
      levoff[d] = o;
      const int Dd = 2 * (nm - T1 - d);
      o += Dd * Dd;
    }
    levoff[T2 + 1] = o;
  }
  constexpr int SLOT = (2 * R) * (2 * R + 1) / 2;
  double* slots = lev + levoff[T2 + 1];
  double* slot_det = slots + 64 * SLOT;
  int* slot_cnt = reinterpret_cast<int*>(slot_det + 64);
  __shared__ int bad_flag;
  if (lane == 0) bad_flag = 0;
  __syncthreads();

  // A = I - O, mode-major: W[(2i+s)][(2j+t)] = delta - O[i + s*nm][j + t*nm]
  const cplx* O = mats + b * dim * dim;
  for (int e = lane; e < dim * dim; e += 64) {
    const int r = e / dim, c = e % dim;
    const int gr = (r >> 1) + (r & 1) * nm, gc = (c >> 1) + (c & 1) * nm;
    const cplx o = O[gr * dim + gc];
    W[e] = (r == c ? 1.0 : 0.0) - o.re;
    if (o.im != 0.0) bad_flag = 1;  // complex O: not this kernel's domain
  }
  __syncthreads();

  // leading modes: eliminate the included ones (mode order, x then p pivot)
  double det = 1.0;
  int count = 0;
  bool bad = false;
  for (int j = 0; j < T1; ++j) {
    if ((p1 >> j) & 1u) {
      const double a = eliminate(W, dim, 2 * j);
      const double c = eliminate(W, dim, 2 * j + 1);
      bad |= !(a > 0.0) || !(c > 0.0);
      det *= a * c;
      ++count;
    }
  }
  // level 0 = W restricted to modes T1..nm-1
  {
    const int D0 = 2 * (nm - T1);
    for (int e = lane; e < D0 * D0; e += 64) {
      const int i = e / D0, j = e % D0;
      lev[levoff[0] + e] = W[(2 * T1 + i) * dim + 2 * T1 + j];
    }
    __syncthreads();
  }

  // cooperative walk over the T2 middle modes
  double dets[TOR_T2 + 1];
  int cnts[TOR_T2 + 1];
  int state[TOR_T2 + 1];
  dets[0] = det;
  cnts[0] = count;
  int ntask = 0;
  int d = 0;
  state[0] = 0;
  for (;;) {
    if (d == T2) {
      // leaf: the task is this level's matrix (the trailing R modes)
      const int Dd = 2 * R;
      const double* X = lev + levoff[d];
      for (int e = lane; e < SLOT; e += 64) {
        int r = 0;
        while (tri(r + 1, 0) <= e) ++r;
        const int c = e - tri(r, 0);
        slots[ntask * SLOT + e] = X[r * Dd + c];
      }
      if (lane == 0) {
        slot_det[ntask] = dets[d];
        slot_cnt[ntask] = cnts[d];
      }
      ++ntask;
      if (d == 0) break;
      --d;
      continue;
    }
    const int Dd = 2 * (nm - T1 - d);
    double* X = lev + levoff[d];
    double* Y = lev + levoff[d + 1];
    if (state[d] == 0) {  // exclude middle mode d: drop its rows
      state[d] = 1;
      drop_mode(X, Dd, 0, Y);
      dets[d + 1] = dets[d];
      cnts[d + 1] = cnts[d];
      state[d + 1] = 0;
      ++d;
      continue;
    }
    if (state[d] == 1) {  // include: eliminate rows 0, 1 into a copy
      state[d] = 2;
      // two pivots (x then p) of mode d, Schur complement formed directly in Y
      const double a = X[0];
      const double b01 = X[1];
      const double c = X[Dd + 1] - b01 * b01 / a;
      const int Dy = Dd - 2;
      for (int e = lane; e < Dy * Dy; e += 64) {
        const int i = e / Dy + 2, j = e % Dy + 2;
        const double xi0 = X[i * Dd], x0j = X[j];
        double v = fma(-xi0 / a, x0j, X[i * Dd + j]);
        const double xi1 = fma(-xi0 / a, b01, X[i * Dd + 1]);
        const double x1j = fma(-b01 / a, x0j, X[Dd + j]);
        v = fma(-xi1 / c, x1j, v);
        Y[e] = v;
      }
      __syncthreads();
      bad |= !(a > 0.0) || !(c > 0.0);
      dets[d + 1] = dets[d] * (a * c);
      cnts[d + 1] = cnts[d] + 1;
      state[d + 1] = 0;
      ++d;
      continue;
    }
    if (d == 0) break;
    --d;
  }
  __syncthreads();

  // per-lane trailing walks
  double acc = 0.0, aacc = 0.0;
  if (lane < ntask) {
    bottom_walk<R>(slots + lane * SLOT, slot_det[lane], slot_cnt[lane], nm, bad, acc, aacc);
  }
  if (bad) bad_flag = 1;
  acc = group_sum<64>(acc);
  if (KAPPA) aacc = group_sum<64>(aacc);
  __syncthreads();
  if (lane == 0) {
    const size_t np = (size_t)1 << T1;
    part[b * np + p1] = bad_flag ? NAN : acc;
    if (KAPPA) part_abs[b * np + p1] = bad_flag ? NAN : aacc;
  }
}

template <bool KAPPA>
__global__ void tor_reduce_kernel(const double* __restrict__ part, const double* __restrict__ part_abs, int np, int B,
                                  cplx* __restrict__ out, double* __restrict__ aux) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double s = 0.0, a = 0.0;
  for (int i = 0; i < np; ++i) {
    s += part[b * np + i];
    if (KAPPA) a += part_abs[b * np + i];
  }
  out[b] = cplx{s, 0.0};
  if (KAPPA) aux[b] = a;
}

__global__ void tor_n0_kernel(int B, cplx* out, double* aux) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  out[b] = cplx{1.0, 0.0};
  if (aux) aux[b] = 1.0;
}

static size_t tor_lds_bytes(int nm) {
  const TorSplit sp = tor_split(nm);
  const int dim = 2 * nm;
  size_t lev = 0;
  for (int d = 0; d <= sp.T2; ++d) {
    const size_t Dd = 2 * (nm - sp.T1 - d);
    lev += Dd * Dd;
  }
  const size_t slot = (size_t)(2 * sp.R) * (2 * sp.R + 1) / 2;
  return (size_t)dim * dim * 8 + lev * 8 + 64 * slot * 8 + 64 * 8 + 64 * 4;
}

template <bool KAPPA>
static int run_tor_real(int nm, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, hipStream_t st) {
  const TorSplit sp = tor_split(nm);
  const int np = 1 << sp.T1;
  double* part = reinterpret_cast<double*>(ws);
  double* part_abs = part + (size_t)B * np;
  const size_t lds = tor_lds_bytes(nm);
  const dim3 grid((unsigned)((size_t)B * np));
  switch (sp.R) {
#define TOR_CASE(RR)                                                                                            \
  case RR:                                                                                                      \
    hipLaunchKernelGGL((tor_real_kernel<RR, KAPPA>), grid, dim3(64), lds, st, d_in, nm, B, part, part_abs); \
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
  hipLaunchKernelGGL((tor_reduce_kernel<KAPPA>), dim3((B + 255) / 256), dim3(256), 0, st, part, part_abs, np, B, d_out,
                     d_aux);
  return (int)hipGetLastError();
}

size_t tor_workspace_bytes(int tier, int dim, int B) {
  if (dim < 2 || (dim & 1) || dim > 2 * TOR_MAX_MODES || B <= 0) return 0;
  if (tier != GBSK_FP64 && tier != GBSK_KAPPA) return 0;
  const TorSplit sp = tor_split(dim / 2);
  return (size_t)B * ((size_t)1 << sp.T1) * 2 * sizeof(double);
}

int tor_max_dim(int tier) {
  switch (tier) {
    case GBSK_FP64:
    case GBSK_KAPPA: return 2 * TOR_MAX_MODES;
  }
  return 0;
}

int tor_launch(int tier, int dim, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, hipStream_t st) {
  if (B <= 0) return 0;
  if (dim < 0 || (dim & 1) || dim > 2 * TOR_MAX_MODES) return GBSK_ERR_DIM;
  if (tier != GBSK_FP64 && tier != GBSK_KAPPA) return GBSK_ERR_FUNC;
  const int nm = dim / 2;
  if (nm == 0) {
    hipLaunchKernelGGL(tor_n0_kernel, dim3((B + 255) / 256), dim3(256), 0, st, B, d_out,
                       tier == GBSK_KAPPA ? d_aux : nullptr);
    return (int)hipGetLastError();
  }
  if (!ws) return GBSK_ERR_WORKSPACE;
  return tier == GBSK_KAPPA ? run_tor_real<true>(nm, d_in, B, d_out, d_aux, ws, st)
                            : run_tor_real<false>(nm, d_in, B, d_out, nullptr, ws, st);
}

}  // namespace gbsk

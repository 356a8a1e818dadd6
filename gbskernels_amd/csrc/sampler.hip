// sampler.hip -- the resident chain-rule GBS sampler (reference
// core/sampler_session.cu:64-128, core/sampler_gather.cu, core/sampler_draw.cu,
// sampling/sampler.py:214-244).
//
// Mode k (1..M) of every draw is drawn from P(j | prefix) ∝ max(Re haf(A_k[idx]), 0) / j!,
// j = 0..cutoff, where A_k (2k x 2k) is the reduced A-matrix of the first k modes
// and idx repeats mode i (and its conjugate i + k) prefix_i times and mode k - 1
// (and 2k - 1) j times. That hafnian is the repeated-row loop hafnian with zero
// loop weights and the counts (prefix, j, prefix, j): this chain evaluates it
// with the sieve (sieve.hip), planned on the device from the resident state, so
// no submatrix is ever gathered and the expansion size is not capped at 20 (the
// sieve's own limits apply: 2k <= 32, N <= 64, 2^40 terms; a weight past them
// fails the call instead of being taken as zero).
//
// Per mode: the counts table from the state (one thread per (draw, j)), the
// device-planned sieve, then one thread per draw normalises its cutoff + 1
// weights and inverts the CDF with a Philox4x32-10 uniform whose counter is
// (draw, mode) -- the same clip / normalise / "count cumsum < u" rule as the
// reference. Nothing returns to the host until the final samples.
#include <hip/hip_runtime.h>

#include "../../include/gbsk.h"
#include "common.hpp"
#include "errors.hpp"
#include "kernels.hpp"
#include "philox.hpp"
#include "sieve.hpp"

namespace gbsk {

__global__ void sampler_fill_kernel(int* __restrict__ p, size_t n, int v) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void sampler_zero_c_kernel(cplx* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = cplx{0.0, 0.0};
}

// counts over A_k's 2k indices for pattern t = d (C + 1) + j:
// (prefix_0 .. prefix_{k-2}, j, prefix_0 .. prefix_{k-2}, j)
__global__ void sampler_reps_kernel(const int* __restrict__ state, int M, int k, int D, int C, int* __restrict__ reps) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)D * (C + 1)) return;
  const int64_t d = t / (C + 1);
  const int j = (int)(t % (C + 1));
  int* row = reps + t * 2 * k;
  const int* pre = state + d * M;
  for (int i = 0; i < k - 1; ++i) {
    row[i] = pre[i];
    row[k + i] = pre[i];
  }
  row[k - 1] = j;
  row[2 * k - 1] = j;
}

// one thread per draw: weights max(Re haf, 0) / j!, inverse CDF at a Philox
// uniform with counter (draw, mode); a NaN weight (a pattern past the sieve's
// domain) raises the flag
__global__ void sampler_draw_kernel(const cplx* __restrict__ haf, const double* __restrict__ invfac, int C, int D,
                                    int M, int col, uint32_t key0, uint32_t key1, int* __restrict__ state,
                                    int* __restrict__ flag) {
  const int64_t d = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= D) return;
  const cplx* w = haf + d * (C + 1);
  double sum = 0.0;
  bool bad = false;
  for (int j = 0; j <= C; ++j) {
    const double h = w[j].re;
    bad |= !(h == h);
    sum += (h > 0.0 ? h : 0.0) * invfac[j];
  }
  if (bad) *flag = 1;
  int photon = 0;
  if (sum > 0.0) {
    const uint32_t ctr[4] = {(uint32_t)d, (uint32_t)(d >> 32), (uint32_t)col, 0u};
    const uint32_t key[2] = {key0, key1};
    uint32_t r[4];
    philox4x32_10(ctr, key, r);
    const double u = philox_unit(r[0], r[1]);
    double c = 0.0;
    for (int j = 0; j <= C; ++j) {
      const double h = w[j].re;
      c += ((h > 0.0 ? h : 0.0) * invfac[j]) / sum;
      if (c < u) ++photon;
    }
  }
  state[d * M + col] = photon;
}

}  // namespace gbsk

using namespace gbsk;

extern "C" {

void gbs_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out) {
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]}, k[2] = {key[0], key[1]}, o[4];
  philox4x32_10(c, k, o);
  for (int i = 0; i < 4; ++i) out[i] = o[i];
}

int gbs_sampler_sample(const gbs_complex* Ak, int ak_total, const int* off, int M, int num_draws, int cutoff,
                       int maxn, const double* invfac, unsigned long long seed, int* out) {
  (void)maxn;  // the reference's gather stride; the sieve chain has none
  if (M < 0 || num_draws < 0 || cutoff < 0 || ak_total < 0) return GBSK_ERR_ARG;
  if (num_draws == 0 || M == 0) return 0;
  if (!Ak || !off || !invfac || !out) return GBSK_ERR_ARG;
  if (2 * M > REP_MAX_M) return GBSK_ERR_DIM;
  for (int k = 1; k <= M; ++k)
    if (off[k - 1] < 0 || (int64_t)off[k - 1] + 4LL * k * k > ak_total) return GBSK_ERR_ARG;
  const int D = num_draws, C = cutoff;
  const int64_t B64 = (int64_t)D * (C + 1);
  if (B64 > (1LL << 30)) return GBSK_ERR_DIM;
  const int B = (int)B64;
  hipStream_t st = nullptr;
  int rc = (int)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (rc) return rc;
  // one allocation, carved into 256-byte aligned pieces
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t b_ak = al((size_t)ak_total * sizeof(cplx)), b_inv = al((size_t)(C + 1) * sizeof(double));
  const size_t b_g = al((size_t)2 * M * sizeof(cplx)), b_state = al((size_t)D * M * sizeof(int));
  const size_t b_reps = al((size_t)B * 2 * M * sizeof(int)), b_plan = al((size_t)B * sizeof(RepPat));
  const size_t b_tot = al(sizeof(int64_t) + sizeof(int)), b_val = al((size_t)B * sizeof(cplx));
  const size_t b_part = al(sieve_device_partial_bytes(false, B));
  const size_t total = b_ak + b_inv + b_g + b_state + b_reps + b_plan + b_tot + b_val + b_part;
  char* base = nullptr;
  rc = (int)hipMalloc(&base, total);
  if (rc) {
    (void)hipStreamDestroy(st);
    return rc;
  }
  char* p = base;
  cplx* d_ak = (cplx*)p;
  p += b_ak;
  double* d_inv = (double*)p;
  p += b_inv;
  cplx* d_g = (cplx*)p;
  p += b_g;
  int* d_state = (int*)p;
  p += b_state;
  int* d_reps = (int*)p;
  p += b_reps;
  RepPat* d_plan = (RepPat*)p;
  p += b_plan;
  int64_t* d_total = (int64_t*)p;
  int* d_flag = (int*)(p + sizeof(int64_t));
  p += b_tot;
  cplx* d_val = (cplx*)p;
  p += b_val;
  void* d_part = p;
  if (ak_total) rc = (int)hipMemcpyAsync(d_ak, Ak, (size_t)ak_total * sizeof(cplx), hipMemcpyHostToDevice, st);
  if (!rc) rc = (int)hipMemcpyAsync(d_inv, invfac, (size_t)(C + 1) * sizeof(double), hipMemcpyHostToDevice, st);
  if (!rc) {
    hipLaunchKernelGGL(sampler_zero_c_kernel, dim3((2 * M + 255) / 256), dim3(256), 0, st, d_g, 2 * M);
    hipLaunchKernelGGL(sampler_fill_kernel, dim3((unsigned)(((size_t)D * M + 255) / 256)), dim3(256), 0, st, d_state,
                       (size_t)D * M, 0);
    hipLaunchKernelGGL(sampler_fill_kernel, dim3(1), dim3(64), 0, st, d_flag, (size_t)1, 0);
    rc = (int)hipGetLastError();
  }
  const uint32_t key0 = (uint32_t)seed, key1 = (uint32_t)(seed >> 32);
  for (int k = 1; !rc && k <= M; ++k) {
    hipLaunchKernelGGL(sampler_reps_kernel, dim3((unsigned)((B64 + 255) / 256)), dim3(256), 0, st, d_state, M, k, D, C,
                       d_reps);
    rc = sieve_launch_device(false, d_ak + off[k - 1], d_g, 2 * k, d_reps, B, 2 * k, d_plan, d_total, d_part, d_val,
                             nullptr, st);
    if (!rc) {
      hipLaunchKernelGGL(sampler_draw_kernel, dim3((unsigned)((D + 255) / 256)), dim3(256), 0, st, d_val, d_inv, C, D,
                         M, k - 1, key0, key1, d_state, d_flag);
      rc = (int)hipGetLastError();
    }
  }
  int flag = 0;
  if (!rc) rc = (int)hipMemcpyAsync(out, d_state, (size_t)D * M * sizeof(int), hipMemcpyDeviceToHost, st);
  if (!rc) rc = (int)hipMemcpyAsync(&flag, d_flag, sizeof(int), hipMemcpyDeviceToHost, st);
  const int rs = (int)hipStreamSynchronize(st);
  (void)hipFree(base);
  (void)hipStreamDestroy(st);
  if (rc || rs) return rc ? rc : rs;
  return flag ? GBSK_ERR_DIM : 0;
}

}  // extern "C"

// haf.hip -- dispatch of the power-trace hafnian / loop hafnian launchers
// (kernels in haf_impl.hpp, instantiated in haf_*.hip).
#include "haf_impl.hpp"

namespace gbsk {

// built-in schedule of the fused small-size kernel (measured on MI355X)
int haf_small_max_n = 12;
int haf_small_target_wgs = 0;  // 0: per-bucket packing from the cost model

int haf_schedule_get(int* small_max_n, int* small_target_wgs) {
  if (small_max_n) *small_max_n = haf_small_max_n;
  if (small_target_wgs) *small_target_wgs = haf_small_target_wgs;
  return 0;
}

int haf_schedule_set(int small_max_n, int small_target_wgs) {
  if (small_max_n < 0 || small_max_n > HAF_SMALL_MAX_N || small_target_wgs < 0) return GBSK_ERR_ARG;
  haf_small_max_n = small_max_n;
  haf_small_target_wgs = small_target_wgs;
  return 0;
}

// N == 0: value 1 (kappa indicator 1, certificate 0); odd N: haf 0 exactly, lhaf NaN
__global__ void haf_trivial_kernel(int N, bool loop, int B, cplx* out, double* aux, bool kappa) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  if (N == 0) {
    out[b] = cplx{1.0, 0.0};
    if (aux) aux[b] = kappa ? 1.0 : 0.0;
  } else if (!loop) {
    out[b] = cplx{0.0, 0.0};
    if (aux) aux[b] = 0.0;
  } else {
    out[b] = cplx{NAN, NAN};
    if (aux) aux[b] = NAN;
  }
}

size_t haf_workspace_bytes(bool, int tier, int N, int B) {
  if (N < 2 || (N & 1) || N > HAF_MAX_N || B <= 0) return 0;
  const size_t elem = (tier == GBSK_DD || tier == GBSK_DD_CERTIFIED) ? sizeof(ddc) : sizeof(cplx);
  return (size_t)B * haf_nparts(N / 2) * (elem + sizeof(double));
}

int haf_max_dim(bool, int tier) {
  switch (tier) {
    case GBSK_FP64:
    case GBSK_DD:
    case GBSK_KAPPA:
    case GBSK_CERTIFIED:
    case GBSK_DD_CERTIFIED: return HAF_MAX_N;
  }
  return 0;
}

static int run_multi(bool loop, int tier, const HafJob* jobs, int njobs, hipStream_t st) {
  switch (tier) {
    case GBSK_DD:
      return loop ? run_haf_multi<ArDD, true, HM_PLAIN>(jobs, njobs, st)
                  : run_haf_multi<ArDD, false, HM_PLAIN>(jobs, njobs, st);
    case GBSK_KAPPA:
      return loop ? run_haf_multi<ArF64, true, HM_KAPPA>(jobs, njobs, st)
                  : run_haf_multi<ArF64, false, HM_KAPPA>(jobs, njobs, st);
    case GBSK_CERTIFIED:
      return loop ? run_haf_multi<ArF64, true, HM_CERT>(jobs, njobs, st)
                  : run_haf_multi<ArF64, false, HM_CERT>(jobs, njobs, st);
    case GBSK_DD_CERTIFIED:
      return loop ? run_haf_multi<ArDD, true, HM_DDCERT>(jobs, njobs, st)
                  : run_haf_multi<ArDD, false, HM_DDCERT>(jobs, njobs, st);
  }
  return loop ? run_haf_multi<ArF64, true, HM_PLAIN>(jobs, njobs, st)
              : run_haf_multi<ArF64, false, HM_PLAIN>(jobs, njobs, st);
}

int haf_launch_multi(bool loop, int tier, const HafJob* jobs, int njobs, hipStream_t st) {
  if (njobs < 0 || njobs > HAF_MAX_JOBS || (njobs && !jobs)) return GBSK_ERR_ARG;
  if (tier < GBSK_FP64 || tier > GBSK_DD_CERTIFIED) return GBSK_ERR_FUNC;
  const bool aux = tier == GBSK_KAPPA || tier == GBSK_CERTIFIED || tier == GBSK_DD_CERTIFIED;
  HafJob work[HAF_MAX_JOBS];
  int nw = 0;
  auto flush = [&]() {
    const int rc = nw ? run_multi(loop, tier, work, nw, st) : 0;
    nw = 0;
    return rc;
  };
  for (int j = 0; j < njobs; ++j) {
    const HafJob& J = jobs[j];
    if (J.B <= 0) continue;
    if (J.N < 0 || J.N > HAF_MAX_N) return GBSK_ERR_DIM;
    if (!J.in || !J.out || (aux && !J.aux)) return GBSK_ERR_ARG;
    if (J.N == 0 || (J.N & 1)) {  // empty matrix or odd size: closed forms
      hipLaunchKernelGGL(haf_trivial_kernel, dim3((J.B + 255) / 256), dim3(256), 0, st, J.N, loop, J.B, J.out,
                         aux ? J.aux : nullptr, tier == GBSK_KAPPA);
      continue;
    }
    if (!J.ws) return GBSK_ERR_WORKSPACE;
    // a job's share of one launch stays below 2^26 workgroups (at most
    // HAF_MAX_JOBS jobs per launch: below 2^30); larger buckets are split into
    // consecutive instance ranges, each with its own slice of the workspace
    int nch_max = 1;
    for (int m = 1; m <= J.N / 2; ++m) {
      const int c = haf_nchunks(J.N / 2, 2 * m);
      nch_max = c > nch_max ? c : nch_max;
    }
    const int64_t max_b = ((int64_t)1 << 26) / nch_max;
    const size_t per_inst = haf_workspace_bytes(loop, tier, J.N, 1);
    for (int64_t b0 = 0; b0 < J.B; b0 += max_b) {
      const int Bs = (int)((J.B - b0) < max_b ? (J.B - b0) : max_b);
      if (nw == HAF_MAX_JOBS) {
        const int rc = flush();
        if (rc) return rc;
      }
      work[nw++] = HafJob{J.in + b0 * J.N * J.N, J.out + b0, aux ? J.aux + b0 : nullptr,
                          (char*)J.ws + (size_t)b0 * per_inst, J.N, Bs};
    }
  }
  const int rc = flush();
  if (rc) return rc;
  return (int)hipGetLastError();
}

// Arithmetic-model self-check: the certified hafnian charges one rounding of
// u|result| to the Newton step's division by the index j (ArF64::div_int), which
// holds only if that division is correctly rounded. out[i] = div_int(a[i], j_i)
// with j_i = 1 + i mod 32, the expression the value path evaluates.
__global__ void haf_div_probe_kernel(const double* __restrict__ a, int64_t n, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[i] = ArF64::div_int(cplx{a[i], 0.0}, (int)(1 + i % 32)).re;
}

int haf_div_probe(const double* d_a, int64_t n, double* d_out, hipStream_t st) {
  if (n <= 0) return 0;
  const int64_t blocks = (n + 255) / 256;
  hipLaunchKernelGGL(haf_div_probe_kernel, dim3((unsigned)(blocks < 65536 ? blocks : 65536)), dim3(256), 0, st, d_a,
                     n, d_out);
  return (int)hipGetLastError();
}

int haf_launch(bool loop, int tier, int N, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws,
               hipStream_t st) {
  if (B <= 0) return 0;
  if (N < 0 || N > HAF_MAX_N) return GBSK_ERR_DIM;
  const HafJob j{d_in, d_out, d_aux, ws, N, B};
  return haf_launch_multi(loop, tier, &j, 1, st);
}

}  // namespace gbsk

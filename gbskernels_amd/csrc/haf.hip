// haf.hip -- dispatch of the power-trace hafnian / loop hafnian launchers
// (kernels in haf_impl.hpp, instantiated in haf_*.hip).
#include "haf_impl.hpp"

namespace gbsk {

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

int haf_launch(bool loop, int tier, int N, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws,
               hipStream_t st) {
  if (B <= 0) return 0;
  if (N < 0 || N > HAF_MAX_N) return GBSK_ERR_DIM;
  if (tier < GBSK_FP64 || tier > GBSK_DD_CERTIFIED) return GBSK_ERR_FUNC;
  if (N == 0 || (N & 1)) {
    const bool aux = tier == GBSK_KAPPA || tier == GBSK_CERTIFIED || tier == GBSK_DD_CERTIFIED;
    hipLaunchKernelGGL(haf_trivial_kernel, dim3((B + 255) / 256), dim3(256), 0, st, N, loop, B, d_out,
                       aux ? d_aux : nullptr, tier == GBSK_KAPPA);
    return (int)hipGetLastError();
  }
  if (!ws) return GBSK_ERR_WORKSPACE;
  if (tier == GBSK_DD) return loop ? run_haf<ArDD, true, HM_PLAIN>(N, d_in, B, d_out, nullptr, ws, st)
                                   : run_haf<ArDD, false, HM_PLAIN>(N, d_in, B, d_out, nullptr, ws, st);
  if (tier == GBSK_KAPPA) return loop ? run_haf<ArF64, true, HM_KAPPA>(N, d_in, B, d_out, d_aux, ws, st)
                                      : run_haf<ArF64, false, HM_KAPPA>(N, d_in, B, d_out, d_aux, ws, st);
  if (tier == GBSK_DD_CERTIFIED) return loop ? run_haf<ArDD, true, HM_DDCERT>(N, d_in, B, d_out, d_aux, ws, st)
                                             : run_haf<ArDD, false, HM_DDCERT>(N, d_in, B, d_out, d_aux, ws, st);
  if (tier == GBSK_CERTIFIED) return loop ? run_haf<ArF64, true, HM_CERT>(N, d_in, B, d_out, d_aux, ws, st)
                                          : run_haf<ArF64, false, HM_CERT>(N, d_in, B, d_out, d_aux, ws, st);
  return loop ? run_haf<ArF64, true, HM_PLAIN>(N, d_in, B, d_out, nullptr, ws, st)
              : run_haf<ArF64, false, HM_PLAIN>(N, d_in, B, d_out, nullptr, ws, st);
}


}  // namespace gbsk

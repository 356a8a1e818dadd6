// kernels.hpp -- launcher entry points of every kernel family (device pointers,
// stream-ordered; return 0 or an error code from errors.hpp).
#pragma once

#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "common.hpp"
#include "errors.hpp"
#include "perm.hpp"
#include "sieve.hpp"

namespace gbsk {

// power-trace hafnian / loop hafnian; tier codes as in include/gbsk.h
size_t haf_workspace_bytes(bool loop, int tier, int N, int B);
int haf_launch(bool loop, int tier, int N, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws,
               hipStream_t st);
int haf_max_dim(bool loop, int tier);
// Several uniform hafnian buckets (same function and tier, different sizes) in
// one launch per subset size: each bucket keeps its own kernels' arithmetic and
// work split (values identical to haf_launch), the launches just carry more
// workgroups. ws: haf_workspace_bytes(loop, tier, N, B) bytes per job.
struct HafJob {
  const cplx* in;
  cplx* out;
  double* aux;
  void* ws;
  int N, B;
};
constexpr int HAF_MAX_JOBS = 16;
int haf_launch_multi(bool loop, int tier, const HafJob* jobs, int njobs, hipStream_t st);

// torontonian; dim = matrix side 2n
size_t tor_workspace_bytes(int tier, int dim, int B);
int tor_launch(int tier, int dim, const cplx* d_in, int B, cplx* d_out, double* d_aux, void* ws, hipStream_t st);
int tor_max_dim(int tier);
// one real instance of up to tor_single_max_modes() modes (input still complex
// storage, imaginary parts zero); real path only, NaN / +inf off the domain
int tor_single_max_modes();
size_t tor_single_workspace_bytes(int tier, int nm);
int tor_single_launch(int tier, int nm, const cplx* d_in, cplx* d_out, double* d_aux, void* ws, hipStream_t st);

// fused small-size hafnian kernel: sizes N <= small_max_n (0..12), instances
// per workgroup chosen to keep at least small_target_wgs workgroups (values
// never depend on either)
int haf_schedule_get(int* small_max_n, int* small_target_wgs);
int haf_schedule_set(int small_max_n, int small_target_wgs);

// torontonian schedule knobs (family 0: fp64 / kappa / certified, 1: dd /
// certified-dd): grouped small-count kernel up to grouped_max modes (changes
// last bits), breadth-first middle walk up to bfs_max modes (same values)
int tor_schedule_get(int family, int* grouped_max, int* bfs_max);
int tor_schedule_set(int family, int grouped_max, int bfs_max);

// repeated-row loop hafnian (sieve): the host plan from the (B x M) host reps
// table (one RepPat per pattern, copied to d_plan by the caller), then the chunk
// and reduction launches; d_part holds sieve_partial_bytes(cert, parts) bytes
int sieve_plan(int M, const int* reps, int B, std::vector<RepPat>& plan, int64_t* parts);
size_t sieve_partial_bytes(bool cert, int64_t parts);
int sieve_launch(bool cert, const cplx* d_A, const cplx* d_g, int M, const int* d_reps, int B,
                 const std::vector<RepPat>& plan, const RepPat* d_plan, int64_t parts, void* d_part, cplx* d_out,
                 double* d_bound, hipStream_t st);
// the same from a device reps table, planned on the device (no host round
// trip); mm >= every pattern's nonzero-count m; buffers: B RepPat, one int64,
// sieve_device_partial_bytes(cert, B)
size_t sieve_device_partial_bytes(bool cert, int B);
int sieve_launch_device(bool cert, const cplx* d_A, const cplx* d_g, int M, const int* d_reps, int B, int mm,
                        RepPat* d_plan, int64_t* d_total, void* d_part, cplx* d_out, double* d_bound,
                        hipStream_t st);

// arithmetic-model self-checks (device pointers): q = a / b and s = sqrt|a| as
// the certified torontonian evaluates them; h = a / (1 + i mod 32) as the
// certified hafnian's Newton step does
int tor_div_sqrt_probe(const double* d_a, const double* d_b, int64_t n, double* d_q, double* d_s, hipStream_t st);
int haf_div_probe(const double* d_a, int64_t n, double* d_out, hipStream_t st);
// r = 1 / sqrt(a) as the torontonian's trailing walk evaluates it (rsqrt_nr)
int tor_rsqrt_probe(const double* d_a, int64_t n, double* d_r, hipStream_t st);

}  // namespace gbsk

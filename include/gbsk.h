/* gbsk.h -- C ABI of libgbsk.so, the MI355X (gfx950) batched backend for the
 * permanent, hafnian, loop hafnian and torontonian.
 *
 * This ABI replaces the reference's nanobind module bindings/gbskernels_ext.cpp
 * and the host layer it binds, core/host_api.cu. The functions in sections 1-3
 * keep the names, argument meaning and error convention of the extern "C"
 * functions that bindings/gbskernels_ext.cpp:24-68 declares; each cites the
 * core/host_api.cu definition it replaces. Section 4 is new (no reference
 * counterpart): stream-ordered evaluation on device-resident buffers.
 *
 * Conventions (as in the reference):
 *  - matrices are C-contiguous row-major complex128 stacks (B, d, d), i.e. the
 *    numpy / std::complex<double> layout; gbs_complex is layout-compatible;
 *  - perm/haf/lhaf take the matrix side; tor takes the MODE count n (matrices
 *    are 2n x 2n, xxpp ordering);
 *  - return 0 on success, a nonzero hipError_t on a HIP failure, or one of the
 *    GBSK_ERR_* codes below for a rejected argument; B <= 0 is a no-op;
 *  - per-element failures are values, not errors: NaN value (outside a kernel's
 *    domain, e.g. an odd loop hafnian), +inf certificate bound (refusal);
 *  - every call is synchronous and reentrant: host entry points serialise on
 *    a process-wide context (one HIP stream, grow-only buffers) and never
 *    synchronise the whole device.
 */
#ifndef GBSK_H
#define GBSK_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  double re, im;
} gbs_complex;

#define GBSK_ERR_ARG 10001
#define GBSK_ERR_DIM 10002
#define GBSK_ERR_FUNC 10003
#define GBSK_ERR_WORKSPACE 10004
#define GBSK_ERR_ALLOC 10005

/* function codes (core/host_api.cu:474-482: 0 perm, 1 haf, 2 lhaf, 3 tor) */
#define GBSK_PERM 0
#define GBSK_HAF 1
#define GBSK_LHAF 2
#define GBSK_TOR 3
/* tier codes: 0 fp64 and 1 dd as in the reference session; 2-4 are the
 * reference's separate *_kappa / *_certified / *_dd_certified entry points */
#define GBSK_FP64 0
#define GBSK_DD 1
#define GBSK_KAPPA 2
#define GBSK_CERTIFIED 3
#define GBSK_DD_CERTIFIED 4

/* ---- 1. one-shot host entry points: H2D -> launch -> D2H ----------------- */
/* core/host_api.cu:154-158 */
int gbs_perm_host(const gbs_complex* in, int n, int B, gbs_complex* out);
/* core/host_api.cu:159-161 */
int gbs_perm_dd_host(const gbs_complex* in, int n, int B, gbs_complex* out);
/* core/host_api.cu:162-167 */
int gbs_haf_host(const gbs_complex* in, int N, int B, gbs_complex* out);
/* core/host_api.cu:168-170 */
int gbs_haf_dd_host(const gbs_complex* in, int N, int B, gbs_complex* out);
/* core/host_api.cu:171-173 */
int gbs_lhaf_host(const gbs_complex* in, int N, int B, gbs_complex* out);
/* core/host_api.cu:174-176 */
int gbs_lhaf_dd_host(const gbs_complex* in, int N, int B, gbs_complex* out);
/* core/host_api.cu:177-229 (real-domain recursive path with off-domain fallback) */
int gbs_tor_host(const gbs_complex* in, int n, int B, gbs_complex* out);
/* core/host_api.cu:230-232 (real-domain DD) */
int gbs_tor_dd_host(const gbs_complex* in, int n, int B, gbs_complex* out);

/* ---- 2. fp64 value + per-element auxiliary ------------------------------ */
/* kappa: aux = sum|term| (core/host_api.cu:237-248) */
int gbs_perm_kappa_host(const gbs_complex* in, int n, int B, gbs_complex* out, double* aux);
int gbs_haf_kappa_host(const gbs_complex* in, int N, int B, gbs_complex* out, double* aux);
int gbs_lhaf_kappa_host(const gbs_complex* in, int N, int B, gbs_complex* out, double* aux);
int gbs_tor_kappa_host(const gbs_complex* in, int n, int B, gbs_complex* out, double* aux);
/* certified: aux = rigorous bound on |out - exact| (core/host_api.cu:252-269) */
int gbs_perm_certified_host(const gbs_complex* in, int n, int B, gbs_complex* out, double* aux);
int gbs_haf_certified_host(const gbs_complex* in, int N, int B, gbs_complex* out, double* aux);
int gbs_lhaf_certified_host(const gbs_complex* in, int N, int B, gbs_complex* out, double* aux);
int gbs_tor_certified_host(const gbs_complex* in, int n, int B, gbs_complex* out, double* aux);
int gbs_perm_dd_certified_host(const gbs_complex* in, int n, int B, gbs_complex* out, double* aux);
int gbs_haf_dd_certified_host(const gbs_complex* in, int N, int B, gbs_complex* out, double* aux);
/* the reference probes these optionally (gbskernels/__init__.py:415-416) */
int gbs_lhaf_dd_certified_host(const gbs_complex* in, int N, int B, gbs_complex* out, double* aux);
int gbs_tor_dd_certified_host(const gbs_complex* in, int n, int B, gbs_complex* out, double* aux);

/* ---- 3. device-resident session (core/host_api.cu:474-658) --------------- */
typedef struct gbs_session gbs_session;
int gbs_session_open(gbs_session** out);
/* dim = matrix side (tor: 2n); func GBSK_PERM..GBSK_TOR; prec GBSK_FP64|GBSK_DD */
int gbs_session_evaluate(gbs_session* s, int func, int prec, const gbs_complex* in, int dim, int B,
                         gbs_complex* out);
int gbs_session_close(gbs_session* s);
size_t gbs_session_reallocs(const gbs_session* s);
/* result left on the device in a fresh buffer owned by the caller */
int gbs_session_evaluate_resident(gbs_session* s, int func, int prec, const gbs_complex* in, int dim, int B,
                                  gbs_complex** d_result);
int gbs_dev_to_host(const gbs_complex* d, gbs_complex* h, int n);
int gbs_dev_free(gbs_complex* d);

/* ---- 4. stream-ordered evaluation on device buffers (no reference counterpart)
 * d_in/d_out/d_aux/d_ws are device pointers; `stream` is a hipStream_t (NULL =
 * the legacy default stream). Nothing is allocated or synchronised; the call
 * returns after the launches are queued. dim = matrix side (tor: 2n). d_aux
 * receives the tier's per-element double (kappa / certificate), else unused.
 * Workspace: at least gbs_workspace_bytes(func, tier, dim, B) bytes. */
size_t gbs_workspace_bytes(int func, int tier, int dim, int B);
int gbs_eval_device(int func, int tier, const gbs_complex* d_in, int dim, int B, gbs_complex* d_out, double* d_aux,
                    void* d_ws, size_t ws_bytes, void* stream);
/* largest supported matrix side for (func, tier), 0 if the pair is unsupported */
int gbs_max_dim(int func, int tier);
const char* gbs_error_string(int rc);
/* 1 when the library was compiled for a GPU (always: there is no host build) */
int gbs_is_device_build(void);

#ifdef __cplusplus
}
#endif

#endif /* GBSK_H */

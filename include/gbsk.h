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
#define GBSK_ERR_SCRATCH 10006 /* gbs_eval_device_noalloc: batch needs more than its fixed scratch */

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
/* gbs_eval_device without a workspace argument and without any allocation.
 * No counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). The workspace is a fixed device array of
 * GBSK_NOALLOC_SCRATCH_BYTES that each GPU the library runs on holds until the
 * process ends; a batch needing more (gbs_workspace_bytes) returns
 * GBSK_ERR_SCRATCH. Calls that overlap on one GPU (any streams, any threads)
 * share that array and race: serialise them. Values are those of
 * gbs_eval_device, bit for bit. */
#define GBSK_NOALLOC_SCRATCH_BYTES (64u << 20)
int gbs_eval_device_noalloc(int func, int tier, const gbs_complex* d_in, int dim, int B, gbs_complex* d_out,
                            double* d_aux, void* stream);
/* largest supported matrix side for (func, tier), 0 if the pair is unsupported */
int gbs_max_dim(int func, int tier);
const char* gbs_error_string(int rc);
/* 1 when the library was compiled for a GPU (always: there is no host build) */
int gbs_is_device_build(void);
/* Self-check of the arithmetic model the certified fp64 tier relies on (no
 * reference counterpart; the reference states the same assumption for its
 * directed-rounding layer, core/certified_rounding.cuh:11-52). The certified
 * bounds charge one rounding of u|result| to each division and square root,
 * i.e. they hold if those are correctly rounded. For host arrays of n doubles:
 * q[i] = a[i] / b[i] and s[i] = sqrt(|a[i]|) evaluated by the certified
 * torontonian's own device expressions, h[i] = a[i] / (1 + i mod 32) by the
 * certified hafnian's Newton-step division. A caller compares them bit for bit
 * with IEEE-754 round-to-nearest results. Synchronous; allocates 5 n doubles. */
int gbs_check_div_sqrt(const double* a, const double* b, long long n, double* q, double* s, double* h);
/* Companion check (no reference counterpart): r[i] = 1 / sqrt(a[i]) as the
 * torontonian's trailing walk evaluates its pivots' reciprocal square roots
 * (the hardware estimate refined by one Newton step for a in [2^-1000,
 * 2^1000], the correctly rounded expression elsewhere). The certified tier
 * charges it 2u|r|; a caller compares with an extended-precision 1 / sqrt(a).
 * Synchronous; allocates 2 n doubles. */
int gbs_check_rsqrt(const double* a, long long n, double* r);

/* ---- 5. heterogeneous batches (no reference counterpart) ------------------
 * Replaces the per-bucket loop of Workspace._ragged / _dispatch_bucket
 * (gbskernels/__init__.py:813-852), which pays one synchronous round trip per
 * (function, size) bucket. Item i is a dim[i] x dim[i] matrix (tor: dim = 2n)
 * at in + off[i] (in gbs_complex units) evaluated by func[i] at tier[i]
 * (GBSK_FP64..GBSK_DD_CERTIFIED); out[i] receives its value and aux[i] (may be
 * NULL) its kappa / certificate. Items are grouped into (func, tier, dim)
 * buckets that run on `nstreams` streams after one packed host-to-device copy;
 * with use_graphs the device part is captured into a hipGraph, cached by the
 * bucket signature (the multiset of (func, tier, dim, count)). Every bucket runs
 * the kernels of a uniform call, so values are bit-identical to evaluating each
 * item alone. budget_bytes (0 = no limit) caps the device bytes the handle's
 * buffers hold: a batch that needs more returns GBSK_ERR_WORKSPACE. A handle is reentrant (calls are
 * serialised on it); distinct handles run concurrently. */
typedef struct gbs_batch gbs_batch;
typedef struct {
  unsigned long long batches, items, buckets, graphs_captured, graph_replays;
  unsigned long long device_reallocs; /* times a grow-only device buffer grew */
} gbs_batch_stats;
int gbs_batch_create(gbs_batch** out, size_t budget_bytes, int nstreams, int use_graphs);
int gbs_batch_run(gbs_batch* b, int n, const int* func, const int* tier, const int* dim, const gbs_complex* in,
                  const long long* off, gbs_complex* out, double* aux);
int gbs_batch_get_stats(gbs_batch* b, gbs_batch_stats* st);
int gbs_batch_destroy(gbs_batch* b);

/* ---- 6. continuous-batching server (no reference counterpart) -------------
 * Requests are submitted one at a time, from any thread, while earlier ones are
 * on the device. A dispatcher thread forms a batch from everything pending
 * whenever one of `slots` batch slots is free and the batch's device bytes fit
 * in what the other slots leave of budget_bytes (each slot keeps grow-only
 * buffers, and what they hold counts; idle slots give theirs back when a
 * request would not fit otherwise); batches in different slots overlap on the
 * device. Each bucket's results are copied back and
 * completed as soon as its own kernels finish. Tier GBSK_AUTO evaluates with
 * kappa and re-evaluates in double-double when sum|term| / |value| >=
 * kappa_max (the reference's auto tier, gbskernels/__init__.py:240-297). The
 * request's matrix is copied at submission. Timestamps are gbs_now_ns(). */
#define GBSK_AUTO 5
typedef struct gbs_server gbs_server;
typedef struct {
  size_t budget_bytes;  /* device bytes the slots' buffers may hold together; 0 = no limit */
  int slots;            /* batches in flight at once, 1..8 */
  int streams_per_slot; /* 1..16 */
  int max_batch;        /* requests per batch; 0 = no limit */
  double kappa_max;     /* GBSK_AUTO escalation threshold (the reference uses 1e8) */
} gbs_server_config;
typedef struct {
  long long arrival_ns, admitted_ns, done_ns;
  int tier_used; /* tier of the returned value (GBSK_AUTO -> GBSK_KAPPA or GBSK_DD) */
  int status;    /* 0, or the error that failed the request */
} gbs_request_info;
typedef struct {
  unsigned long long submitted, completed, batches, buckets, escalated, max_in_flight;
  /* largest device bytes the slots' buffers held together, admitted batches
   * included (counted only under a budget; never above budget_bytes) */
  unsigned long long device_bytes_max;
} gbs_server_stats;
long long gbs_now_ns(void);
int gbs_server_create(gbs_server** out, const gbs_server_config* cfg);
/* arrival_ns <= 0: now. *ticket identifies the request in gbs_server_wait. */
int gbs_server_submit(gbs_server* s, int func, int tier, int dim, const gbs_complex* in, long long arrival_ns,
                      long long* ticket);
/* blocks until the request is done; releases the ticket. aux and info may be NULL. */
int gbs_server_wait(gbs_server* s, long long ticket, gbs_complex* out, double* aux, gbs_request_info* info);
/* open-loop replay: request i is submitted from the calling thread at its due
 * time t0 + arrival_offset_ns[i] (t0 = the call's start; offsets ascending),
 * which is also its arrival_ns, however late the submission; then every result
 * is collected. Returns when all are done. aux and info may be NULL. */
int gbs_server_replay(gbs_server* s, int n, const int* func, const int* tier, const int* dim, const gbs_complex* in,
                      const long long* off, const long long* arrival_offset_ns, gbs_complex* out, double* aux,
                      gbs_request_info* info);
int gbs_server_get_stats(gbs_server* s, gbs_server_stats* st);
/* waits for every submitted request, then stops the dispatcher and frees everything */
int gbs_server_destroy(gbs_server* s);

/* ---- 7. repeated-row loop hafnian (finite-difference sieve) ---------------
 * One base matrix A (M x M) and loop vector gamma (M) shared by the call; reps
 * is a (B x M) int32 table of repetition counts. out[b] is the loop hafnian of A
 * with row/column i repeated reps[b][i] times and gamma_i on each copy's
 * diagonal (gamma = 0: the hafnian of the expansion). Per element: N = sum of the
 * counts; N = 0 gives 1 (bound 0); M > 32, N > 64, a negative count or more than
 * 2^40 sieve terms prod(n_i + 1) give NaN (bound +inf). The certified entry
 * point's values are bit-identical to the plain one's. */
/* core/host_api.cu:411-431 (bindings/gbskernels_ext.cpp:49) */
int gbs_lhaf_repeated_host(const gbs_complex* A, const gbs_complex* gamma, int M, const int* reps, int B,
                           gbs_complex* out);
/* core/host_api.cu:435-458 (bindings/gbskernels_ext.cpp:50) */
int gbs_lhaf_repeated_cert_host(const gbs_complex* A, const gbs_complex* gamma, int M, const int* reps, int B,
                                gbs_complex* out, double* bound);

/* ---- 8. single-large torontonian -------------------------------------------
 * One real (2n x 2n) O, row-major doubles, xxpp order, n <= 32 modes (dim 64),
 * split across the whole GPU. Real physical domain only: if I - O_S is not SPD
 * for some subset the value is NaN (bound +inf). g is the reference's subtree
 * exponent, accepted for compatibility and checked (0 <= g <= n); the work split
 * is this library's own and the value does not depend on g. */
/* core/host_api.cu:262-289 (bindings/gbskernels_ext.cpp:257-268) */
int gbs_tor_recursive_single_host(const double* O, int n, int g, double* out);
/* core/host_api.cu:294-344 (bindings/gbskernels_ext.cpp:271-282) */
int gbs_tor_single_certified_host(const double* O, int n, int g, double* out, double* bound);
/* core/host_api.cu:350-404 (bindings/gbskernels_ext.cpp:285-295) */
int gbs_tor_single_ddcertified_host(const double* O, int n, int g, double* out, double* bound);

/* ---- 9. resident chain-rule GBS sampler -------------------------------------
 * Ak: the M reduced A-matrices A_1..A_M concatenated (A_k is 2k x 2k row-major,
 * starting at Ak + off[k - 1]); invfac[j] = 1 / j!, j = 0..cutoff. Draws
 * num_draws photon patterns mode by mode, P(j | prefix) proportional to
 * max(Re haf(A_k[prefix, j]), 0) / j! over j <= cutoff, with every conditional
 * hafnian evaluated on the device by the repeated-row sieve and uniforms from
 * Philox4x32-10 at counter (draw, mode) under key = seed. out: num_draws x M
 * int32, row-major. M <= 16; maxn is accepted for compatibility (unused).
 * Returns GBSK_ERR_DIM if a conditional fell outside the sieve's domain. */
/* core/sampler_session.cu:96-128 (bindings/gbskernels_ext.cpp:153-167) */
int gbs_sampler_sample(const gbs_complex* Ak, int ak_total, const int* off, int M, int num_draws, int cutoff,
                       int maxn, const double* invfac, unsigned long long seed, int* out);
/* the generator the sampler draws from (host evaluation; no reference
 * counterpart): out = Philox4x32-10(ctr[4], key[2]) */
void gbs_philox4x32_10(const unsigned* ctr, const unsigned* key, unsigned* out);

/* ---- 10. tuning: the permanent's Gray-code chunk shape (no reference
 * counterpart; it replaces the reference's fixed crossovers PERM_COOP_CROSSOVER
 * / PERM_COOP_GROUPS, core/host_api.cu:65-72) ---------------------------------
 * One n x n permanent walks 2^(n-1) Gray terms: LPM lanes (fixed per n) times W
 * workgroups times 2^k terms per lane. The table entry t gives k = min(t, the
 * one-workgroup limit) and W = 2^(limit - k). family 0: fp64 / kappa /
 * certified (one split, so their values stay bit-identical); family 1: dd /
 * certified-dd. A value's last bits depend on the split (its summation order),
 * never on the batch. The built-in table comes from gbskernels_amd/autotune.py;
 * setting an entry is for tuning runs and must not overlap evaluations.
 * log2_terms: 4..27 (family 0), 2..27 (family 1). */
int gbs_perm_get_lane_log2(int family, int n, int* log2_terms);
int gbs_perm_set_lane_log2(int family, int n, int log2_terms);
/* the split a launch of n x n permanents uses now (n = 1..28) */
int gbs_perm_split(int family, int n, int* lane_log2, int* workgroups);

/* ---- 11. tuning: the hafnian's fused small-size kernel (no reference
 * counterpart; it replaces the reference's small-cap dispatch,
 * core/host_api.cu:162-167). Matrices of side N <= small_max_n (0..12) of every
 * tier run in one fused launch (all subset sizes and the final sum); the dd
 * loop hafnian stops at N = 10 and certified dd at N = 8 whatever small_max_n
 * says. Several instances share a workgroup: as many as keep at least
 * small_target_wgs workgroups, or, with small_target_wgs = 0 (built in), as
 * many as a cost model of the kernel's rounds and occupancy picks per bucket;
 * larger sides run one launch per subset size.
 * Both paths give identical values and bounds. The built-in values come from measurements
 * with gbskernels_amd/autotune.py; setting them is for tuning runs and must not
 * overlap evaluations. */
int gbs_haf_get_schedule(int* small_max_n, int* small_target_wgs);
int gbs_haf_set_schedule(int small_max_n, int small_target_wgs);

/* ---- 12. tuning: the batched torontonian's schedule (no reference
 * counterpart; it replaces the reference's fixed per-size dispatch,
 * core/host_api.cu:177-229). family 0: fp64 / kappa / certified, 1: dd /
 * certified-dd. grouped_max_modes (0..10): mode counts up to which several
 * instances share a wave (lane groups of 2^(n - R) lanes, R = the family's
 * per-lane trailing modes); this split differs from the one-instance-per-wave
 * kernel's, so it moves values' last bits. bfs_max_modes (0..32): up to this
 * mode count the one-instance-per-wave kernel walks its middle include/exclude
 * tree breadth-first (one pass per level), depth-first above; both give
 * identical values. The built-in schedule comes from measurements with
 * gbskernels_amd/autotune.py; setting it is for tuning runs and must not
 * overlap evaluations. */
int gbs_tor_get_schedule(int family, int* grouped_max_modes, int* bfs_max_modes);
int gbs_tor_set_schedule(int family, int grouped_max_modes, int bfs_max_modes);

#ifdef __cplusplus
}
#endif

#endif /* GBSK_H */

// api.hip -- the C ABI (include/gbsk.h): dispatch, one-shot host entry points,
// sessions and stream-ordered device evaluation.
//
// Replaces core/host_api.cu. Differences by design:
//  * no hipMalloc/hipFree per call: one process-wide context owns a HIP stream
//    and grow-only device buffers, serialised by a mutex (ctypes drops the GIL,
//    so concurrent Python threads are safe);
//  * no device-wide synchronisation: every call waits on its own stream only;
//  * the torontonian's real-domain test and the recursive/complex dispatch run
//    on the device, not as an O(B d^2) host scan.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "../../include/gbsk.h"
#include "common.hpp"
#include "errors.hpp"
#include "kernels.hpp"

namespace gbsk {

static_assert(sizeof(gbs_complex) == sizeof(cplx), "gbs_complex layout");

// kernel dimension the launchers take: matrix side for perm/haf/lhaf, modes for tor
static size_t ws_bytes(int func, int tier, int dim, int B) {
  switch (func) {
    case GBSK_PERM:
      return perm_workspace_bytes(dim, tier == GBSK_DD ? PERM_DD : (tier == GBSK_DD_CERTIFIED ? PERM_DDCERT : PERM_FP64), B);
    case GBSK_HAF:
    case GBSK_LHAF: return haf_workspace_bytes(func == GBSK_LHAF, tier, dim, B);
    case GBSK_TOR: return tor_workspace_bytes(tier, dim, B);
  }
  return 0;
}

static int launch(int func, int tier, const cplx* d_in, int dim, int B, cplx* d_out, double* d_aux, void* ws,
                  size_t ws_size, hipStream_t st) {
  if (B <= 0) return 0;
  if (!d_in || !d_out) return GBSK_ERR_ARG;
  if (ws_size < ws_bytes(func, tier, dim, B)) return GBSK_ERR_WORKSPACE;
  switch (func) {
    case GBSK_PERM:
      switch (tier) {
        case GBSK_FP64: return perm_launch(dim, PERM_FP64, d_in, B, d_out, nullptr, ws, ws_size, st);
        case GBSK_DD: return perm_launch(dim, PERM_DD, d_in, B, d_out, nullptr, ws, ws_size, st);
        case GBSK_KAPPA:
          if (!d_aux) return GBSK_ERR_ARG;
          return perm_launch(dim, PERM_KAPPA, d_in, B, d_out, d_aux, ws, ws_size, st);
        case GBSK_CERTIFIED:
          if (!d_aux) return GBSK_ERR_ARG;
          return perm_launch(dim, PERM_CERT, d_in, B, d_out, d_aux, ws, ws_size, st);
        case GBSK_DD_CERTIFIED:
          if (!d_aux) return GBSK_ERR_ARG;
          return perm_launch(dim, PERM_DDCERT, d_in, B, d_out, d_aux, ws, ws_size, st);
        default: return GBSK_ERR_FUNC;
      }
    case GBSK_HAF:
    case GBSK_LHAF:
      if ((tier == GBSK_KAPPA || tier >= GBSK_CERTIFIED) && !d_aux) return GBSK_ERR_ARG;
      return haf_launch(func == GBSK_LHAF, tier, dim, d_in, B, d_out, d_aux, ws, st);
    case GBSK_TOR:
      if ((tier == GBSK_KAPPA || tier >= GBSK_CERTIFIED) && !d_aux) return GBSK_ERR_ARG;
      return tor_launch(tier, dim, d_in, B, d_out, d_aux, ws, st);
  }
  return GBSK_ERR_FUNC;
}

// ------------------------------------------------------------ host context
struct DevBuf {
  void* p = nullptr;
  size_t cap = 0;
  int grow(size_t need, size_t* grew = nullptr) {
    if (need <= cap) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, need) != hipSuccess) {
      p = nullptr;
      return (int)hipErrorOutOfMemory;
    }
    cap = need;
    if (grew) ++*grew;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct Ctx {
  std::mutex mu;
  hipStream_t stream = nullptr;
  DevBuf in, out, aux, ws;
  size_t reallocs = 0;  // times `in` grew: the reuse witness of a session
  int init() {
    if (stream) return 0;
    return (int)hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
  }
};

static Ctx& global_ctx() {
  static Ctx* c = new Ctx();  // intentionally leaked: HIP may be torn down first at exit
  return *c;
}

// The one-shot path shared by every host entry point. `dim` is the matrix side.
static int run_host(Ctx& c, int func, int tier, const gbs_complex* h_in, int dim, int B, gbs_complex* h_out,
                    double* h_aux, bool count_reallocs) {
  if (B <= 0) return 0;
  if (!h_in || !h_out || dim < 0) return GBSK_ERR_ARG;
  const bool want_aux = (tier == GBSK_KAPPA || tier == GBSK_CERTIFIED || tier == GBSK_DD_CERTIFIED);
  if (want_aux && !h_aux) return GBSK_ERR_ARG;
  int rc = c.init();
  const size_t in_bytes = (size_t)B * dim * dim * sizeof(cplx);
  const size_t wsb = ws_bytes(func, tier, dim, B);
  if (!rc) rc = c.in.grow(in_bytes > 0 ? in_bytes : 16, count_reallocs ? &c.reallocs : nullptr);
  if (!rc) rc = c.out.grow((size_t)B * sizeof(cplx));
  if (!rc) rc = c.aux.grow((size_t)B * sizeof(double));
  if (!rc && wsb) rc = c.ws.grow(wsb);
  if (!rc && in_bytes) rc = (int)hipMemcpyAsync(c.in.p, h_in, in_bytes, hipMemcpyHostToDevice, c.stream);
  if (!rc)
    rc = launch(func, tier, (const cplx*)c.in.p, dim, B, (cplx*)c.out.p, (double*)c.aux.p, c.ws.p, c.ws.cap,
                c.stream);
  if (!rc) rc = (int)hipMemcpyAsync(h_out, c.out.p, (size_t)B * sizeof(cplx), hipMemcpyDeviceToHost, c.stream);
  if (!rc && want_aux)
    rc = (int)hipMemcpyAsync(h_aux, c.aux.p, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  int rs = (int)hipStreamSynchronize(c.stream);
  return rc ? rc : rs;
}

static int host_call(int func, int tier, const gbs_complex* in, int kdim, int B, gbs_complex* out, double* aux) {
  Ctx& c = global_ctx();
  std::lock_guard<std::mutex> g(c.mu);
  const int dim = (func == GBSK_TOR) ? 2 * kdim : kdim;
  if (kdim < 0) return GBSK_ERR_DIM;
  return run_host(c, func, tier, in, dim, B, out, aux, false);
}

// repeated-row sieve, one shot: A, gamma, reps and the plan go up in the input
// buffer (256-byte aligned pieces), chunk partials live in the workspace
static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
static int sieve_host(bool cert, const gbs_complex* hA, const gbs_complex* hg, int M, const int* hreps, int B,
                      gbs_complex* hout, double* hbound) {
  if (B <= 0) return 0;
  if (M < 0 || !hout || (cert && !hbound) || (M > 0 && (!hA || !hg || !hreps))) return GBSK_ERR_ARG;
  std::vector<RepPat> plan;
  int64_t parts = 0;
  int rc = sieve_plan(M, hreps, B, plan, &parts);
  if (rc) return rc;
  const size_t a_b = (size_t)M * M * sizeof(cplx), g_b = (size_t)M * sizeof(cplx);
  const size_t r_b = (size_t)B * M * sizeof(int), p_b = (size_t)B * sizeof(RepPat);
  const size_t o_a = 0, o_g = align256(a_b), o_r = o_g + align256(g_b), o_p = o_r + align256(r_b);
  const size_t in_b = o_p + align256(p_b), ws_b = sieve_partial_bytes(cert, parts);
  Ctx& c = global_ctx();
  std::lock_guard<std::mutex> g(c.mu);
  rc = c.init();
  if (!rc) rc = c.in.grow(in_b);
  if (!rc) rc = c.out.grow((size_t)B * sizeof(cplx));
  if (!rc) rc = c.aux.grow((size_t)B * sizeof(double));
  if (!rc) rc = c.ws.grow(ws_b > 0 ? ws_b : 16);
  char* din = (char*)c.in.p;
  if (!rc && a_b) rc = (int)hipMemcpyAsync(din + o_a, hA, a_b, hipMemcpyHostToDevice, c.stream);
  if (!rc && g_b) rc = (int)hipMemcpyAsync(din + o_g, hg, g_b, hipMemcpyHostToDevice, c.stream);
  if (!rc && r_b) rc = (int)hipMemcpyAsync(din + o_r, hreps, r_b, hipMemcpyHostToDevice, c.stream);
  if (!rc) rc = (int)hipMemcpyAsync(din + o_p, plan.data(), p_b, hipMemcpyHostToDevice, c.stream);
  if (!rc)
    rc = sieve_launch(cert, (const cplx*)(din + o_a), (const cplx*)(din + o_g), M, (const int*)(din + o_r), B, plan,
                      (const RepPat*)(din + o_p), parts, c.ws.p, (cplx*)c.out.p, (double*)c.aux.p, c.stream);
  if (!rc) rc = (int)hipMemcpyAsync(hout, c.out.p, (size_t)B * sizeof(cplx), hipMemcpyDeviceToHost, c.stream);
  if (!rc && cert)
    rc = (int)hipMemcpyAsync(hbound, c.aux.p, (size_t)B * sizeof(double), hipMemcpyDeviceToHost, c.stream);
  const int rs = (int)hipStreamSynchronize(c.stream);  // also keeps `plan` alive for its copy
  return rc ? rc : rs;
}

// single-large torontonian, one shot: the real (2n x 2n) O goes up as complex
// storage; the subtree count g is validated only (the split is the engine's)
static int tor_single_host(int tier, const double* hO, int n, int g, double* out, double* bound) {
  const bool cert = tier == GBSK_CERTIFIED || tier == GBSK_DD_CERTIFIED;
  if ((!hO && n > 0) || !out || (cert && !bound)) return GBSK_ERR_ARG;
  if (n < 0 || n > tor_single_max_modes()) return GBSK_ERR_DIM;
  if (g < 0 || g > n) return GBSK_ERR_ARG;
  const int dim = 2 * n;
  std::vector<cplx> h((size_t)dim * dim);
  for (size_t i = 0; i < h.size(); ++i) h[i] = cplx{hO[i], 0.0};
  Ctx& c = global_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int rc = c.init();
  const size_t wsb = tor_single_workspace_bytes(tier, n);
  if (!rc) rc = c.in.grow(h.empty() ? 16 : h.size() * sizeof(cplx));
  if (!rc) rc = c.out.grow(sizeof(cplx));
  if (!rc) rc = c.aux.grow(sizeof(double));
  if (!rc && wsb) rc = c.ws.grow(wsb);
  if (!rc && !h.empty()) rc = (int)hipMemcpyAsync(c.in.p, h.data(), h.size() * sizeof(cplx), hipMemcpyHostToDevice, c.stream);
  if (!rc) rc = tor_single_launch(tier, n, (const cplx*)c.in.p, (cplx*)c.out.p, (double*)c.aux.p, c.ws.p, c.stream);
  cplx v{0.0, 0.0};
  double e = 0.0;
  if (!rc) rc = (int)hipMemcpyAsync(&v, c.out.p, sizeof(cplx), hipMemcpyDeviceToHost, c.stream);
  if (!rc && cert) rc = (int)hipMemcpyAsync(&e, c.aux.p, sizeof(double), hipMemcpyDeviceToHost, c.stream);
  const int rs = (int)hipStreamSynchronize(c.stream);  // also keeps `h`, `v`, `e` alive for their copies
  if (rc || rs) return rc ? rc : rs;
  *out = v.re;
  if (cert) *bound = e;
  return 0;
}

// ------------------------------------------------- allocation-free evaluation
// gbs_eval_device_noalloc's workspace: a file-scope device array that every GPU
// the library runs on holds until the process ends. Its name is a host-side
// stand-in, not a device address: host code uses only the per-device address
// hipGetSymbolAddress returns, fetched once per device (with that device
// current) on the entry point's first call there. The kernels write every
// workspace byte they read, so nothing in it needs clearing; calls that overlap
// on one GPU share the array and race.
constexpr size_t SCRATCH_BYTES = (size_t)GBSK_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int SCRATCH_DEVICES = 64;
static std::mutex scratch_mu;
static void* scratch_ptr[SCRATCH_DEVICES];

static int scratch_address(void** out) {
  int dev = 0;
  int rc = (int)hipGetDevice(&dev);
  if (rc) return rc;
  if (dev < 0 || dev >= SCRATCH_DEVICES) return GBSK_ERR_ARG;
  std::lock_guard<std::mutex> g(scratch_mu);
  if (!scratch_ptr[dev]) {
    void* p = nullptr;
    rc = (int)hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc) return rc;
    if (!p) return GBSK_ERR_ARG;
    scratch_ptr[dev] = p;
  }
  *out = scratch_ptr[dev];
  return 0;
}

}  // namespace gbsk

using namespace gbsk;

struct gbs_session {
  Ctx ctx;
};

extern "C" {

int gbs_perm_host(const gbs_complex* m, int n, int b, gbs_complex* o) { return host_call(GBSK_PERM, GBSK_FP64, m, n, b, o, nullptr); }
int gbs_perm_dd_host(const gbs_complex* m, int n, int b, gbs_complex* o) { return host_call(GBSK_PERM, GBSK_DD, m, n, b, o, nullptr); }
int gbs_haf_host(const gbs_complex* m, int N, int b, gbs_complex* o) { return host_call(GBSK_HAF, GBSK_FP64, m, N, b, o, nullptr); }
int gbs_haf_dd_host(const gbs_complex* m, int N, int b, gbs_complex* o) { return host_call(GBSK_HAF, GBSK_DD, m, N, b, o, nullptr); }
int gbs_lhaf_host(const gbs_complex* m, int N, int b, gbs_complex* o) { return host_call(GBSK_LHAF, GBSK_FP64, m, N, b, o, nullptr); }
int gbs_lhaf_dd_host(const gbs_complex* m, int N, int b, gbs_complex* o) { return host_call(GBSK_LHAF, GBSK_DD, m, N, b, o, nullptr); }
int gbs_tor_host(const gbs_complex* m, int n, int b, gbs_complex* o) { return host_call(GBSK_TOR, GBSK_FP64, m, n, b, o, nullptr); }
int gbs_tor_dd_host(const gbs_complex* m, int n, int b, gbs_complex* o) { return host_call(GBSK_TOR, GBSK_DD, m, n, b, o, nullptr); }

int gbs_perm_kappa_host(const gbs_complex* m, int n, int b, gbs_complex* o, double* a) { return host_call(GBSK_PERM, GBSK_KAPPA, m, n, b, o, a); }
int gbs_haf_kappa_host(const gbs_complex* m, int N, int b, gbs_complex* o, double* a) { return host_call(GBSK_HAF, GBSK_KAPPA, m, N, b, o, a); }
int gbs_lhaf_kappa_host(const gbs_complex* m, int N, int b, gbs_complex* o, double* a) { return host_call(GBSK_LHAF, GBSK_KAPPA, m, N, b, o, a); }
int gbs_tor_kappa_host(const gbs_complex* m, int n, int b, gbs_complex* o, double* a) { return host_call(GBSK_TOR, GBSK_KAPPA, m, n, b, o, a); }

int gbs_perm_certified_host(const gbs_complex* m, int n, int b, gbs_complex* o, double* e) { return host_call(GBSK_PERM, GBSK_CERTIFIED, m, n, b, o, e); }
int gbs_haf_certified_host(const gbs_complex* m, int N, int b, gbs_complex* o, double* e) { return host_call(GBSK_HAF, GBSK_CERTIFIED, m, N, b, o, e); }
int gbs_lhaf_certified_host(const gbs_complex* m, int N, int b, gbs_complex* o, double* e) { return host_call(GBSK_LHAF, GBSK_CERTIFIED, m, N, b, o, e); }
int gbs_tor_certified_host(const gbs_complex* m, int n, int b, gbs_complex* o, double* e) { return host_call(GBSK_TOR, GBSK_CERTIFIED, m, n, b, o, e); }
int gbs_perm_dd_certified_host(const gbs_complex* m, int n, int b, gbs_complex* o, double* e) { return host_call(GBSK_PERM, GBSK_DD_CERTIFIED, m, n, b, o, e); }
int gbs_haf_dd_certified_host(const gbs_complex* m, int N, int b, gbs_complex* o, double* e) { return host_call(GBSK_HAF, GBSK_DD_CERTIFIED, m, N, b, o, e); }
int gbs_lhaf_dd_certified_host(const gbs_complex* m, int N, int b, gbs_complex* o, double* e) { return host_call(GBSK_LHAF, GBSK_DD_CERTIFIED, m, N, b, o, e); }
int gbs_tor_dd_certified_host(const gbs_complex* m, int n, int b, gbs_complex* o, double* e) { return host_call(GBSK_TOR, GBSK_DD_CERTIFIED, m, n, b, o, e); }

// --- session -----------------------------------------------------------------
int gbs_session_open(gbs_session** out) {
  if (!out) return GBSK_ERR_ARG;
  *out = nullptr;
  gbs_session* s = new (std::nothrow) gbs_session();
  if (!s) return GBSK_ERR_ALLOC;
  int rc = s->ctx.init();
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return 0;
}

int gbs_session_evaluate(gbs_session* s, int func, int prec, const gbs_complex* in, int dim, int B, gbs_complex* out) {
  if (!s) return GBSK_ERR_ARG;
  if (func < GBSK_PERM || func > GBSK_TOR || (prec != GBSK_FP64 && prec != GBSK_DD)) return GBSK_ERR_FUNC;
  if (func == GBSK_TOR && (dim & 1)) return GBSK_ERR_DIM;
  std::lock_guard<std::mutex> g(s->ctx.mu);
  return run_host(s->ctx, func, prec, in, dim, B, out, nullptr, true);
}

int gbs_session_evaluate_resident(gbs_session* s, int func, int prec, const gbs_complex* in, int dim, int B,
                                  gbs_complex** d_result) {
  if (!s || !d_result) return GBSK_ERR_ARG;
  *d_result = nullptr;
  if (B <= 0) return 0;
  if (!in || dim < 0) return GBSK_ERR_ARG;
  if (func < GBSK_PERM || func > GBSK_TOR || (prec != GBSK_FP64 && prec != GBSK_DD)) return GBSK_ERR_FUNC;
  if (func == GBSK_TOR && (dim & 1)) return GBSK_ERR_DIM;
  Ctx& c = s->ctx;
  std::lock_guard<std::mutex> g(c.mu);
  const size_t in_bytes = (size_t)B * dim * dim * sizeof(cplx);
  const size_t wsb = ws_bytes(func, prec, dim, B);
  int rc = c.in.grow(in_bytes > 0 ? in_bytes : 16, &c.reallocs);
  if (!rc && wsb) rc = c.ws.grow(wsb);
  if (!rc) rc = c.aux.grow((size_t)B * sizeof(double));
  void* res = nullptr;
  if (!rc) rc = (int)hipMalloc(&res, (size_t)B * sizeof(cplx));
  if (!rc && in_bytes) rc = (int)hipMemcpyAsync(c.in.p, in, in_bytes, hipMemcpyHostToDevice, c.stream);
  if (!rc) rc = launch(func, prec, (const cplx*)c.in.p, dim, B, (cplx*)res, (double*)c.aux.p, c.ws.p, c.ws.cap, c.stream);
  int rs = (int)hipStreamSynchronize(c.stream);
  if (!rc) rc = rs;
  if (rc) {
    if (res) (void)hipFree(res);
    return rc;
  }
  *d_result = (gbs_complex*)res;
  return 0;
}

int gbs_session_close(gbs_session* s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> g(s->ctx.mu);
    if (s->ctx.stream) (void)hipStreamSynchronize(s->ctx.stream);
    s->ctx.in.release();
    s->ctx.out.release();
    s->ctx.aux.release();
    s->ctx.ws.release();
    if (s->ctx.stream) (void)hipStreamDestroy(s->ctx.stream);
    s->ctx.stream = nullptr;
  }
  delete s;
  return 0;
}

size_t gbs_session_reallocs(const gbs_session* s) { return s ? s->ctx.reallocs : 0; }

int gbs_dev_to_host(const gbs_complex* d, gbs_complex* h, int n) {
  if (n <= 0) return 0;
  if (!d || !h) return GBSK_ERR_ARG;
  return (int)hipMemcpy(h, d, (size_t)n * sizeof(gbs_complex), hipMemcpyDeviceToHost);
}

int gbs_dev_free(gbs_complex* d) {
  if (d) (void)hipFree(d);
  return 0;
}

// --- stream-ordered device evaluation -----------------------------------------
size_t gbs_workspace_bytes(int func, int tier, int dim, int B) { return ws_bytes(func, tier, dim, B); }

int gbs_eval_device(int func, int tier, const gbs_complex* d_in, int dim, int B, gbs_complex* d_out, double* d_aux,
                    void* d_ws, size_t ws_size, void* stream) {
  if (dim < 0) return GBSK_ERR_DIM;
  if (func == GBSK_TOR && (dim & 1)) return GBSK_ERR_DIM;
  const int cap = gbs_max_dim(func, tier);
  if (cap == 0) return GBSK_ERR_FUNC;  // (func, tier) not built
  if (dim > cap) return GBSK_ERR_DIM;
  return launch(func, tier, (const cplx*)d_in, dim, B, (cplx*)d_out, d_aux, d_ws, ws_size, (hipStream_t)stream);
}

int gbs_eval_device_noalloc(int func, int tier, const gbs_complex* d_in, int dim, int B, gbs_complex* d_out,
                            double* d_aux, void* stream) {
  if (dim < 0) return GBSK_ERR_DIM;
  if (func == GBSK_TOR && (dim & 1)) return GBSK_ERR_DIM;
  const int cap = gbs_max_dim(func, tier);
  if (cap == 0) return GBSK_ERR_FUNC;
  if (dim > cap) return GBSK_ERR_DIM;
  if (B <= 0) return 0;
  if (ws_bytes(func, tier, dim, B) > SCRATCH_BYTES) return GBSK_ERR_SCRATCH;
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc) return rc;
  return launch(func, tier, (const cplx*)d_in, dim, B, (cplx*)d_out, d_aux, ws, SCRATCH_BYTES, (hipStream_t)stream);
}

int gbs_perm_get_lane_log2(int family, int n, int* log2_terms) {
  if (!log2_terms) return GBSK_ERR_ARG;
  const int v = perm_lane_log2_get(family, n);
  if (v == GBSK_ERR_ARG) return v;
  *log2_terms = v;
  return 0;
}

int gbs_perm_set_lane_log2(int family, int n, int log2_terms) { return perm_lane_log2_set(family, n, log2_terms); }

int gbs_haf_get_schedule(int* small_max_n, int* small_target_wgs) {
  return haf_schedule_get(small_max_n, small_target_wgs);
}

int gbs_haf_set_schedule(int small_max_n, int small_target_wgs) {
  return haf_schedule_set(small_max_n, small_target_wgs);
}

int gbs_tor_get_schedule(int family, int* grouped_max_modes, int* bfs_max_modes) {
  return tor_schedule_get(family, grouped_max_modes, bfs_max_modes);
}

int gbs_tor_set_schedule(int family, int grouped_max_modes, int bfs_max_modes) {
  return tor_schedule_set(family, grouped_max_modes, bfs_max_modes);
}

int gbs_perm_split(int family, int n, int* lane_log2, int* workgroups) {
  if ((family != 0 && family != 1) || n < 1 || n > PERM_MAX_N || !lane_log2 || !workgroups) return GBSK_ERR_ARG;
  const PermSplit sp = perm_split(n, family);
  *lane_log2 = sp.M;
  *workgroups = sp.W;
  return 0;
}

int gbs_max_dim(int func, int tier) {
  switch (func) {
    case GBSK_PERM: return (tier >= GBSK_FP64 && tier <= GBSK_DD_CERTIFIED) ? PERM_MAX_N : 0;
    case GBSK_HAF:
    case GBSK_LHAF: return haf_max_dim(func == GBSK_LHAF, tier);
    case GBSK_TOR: return tor_max_dim(tier);
  }
  return 0;
}

const char* gbs_error_string(int rc) {
  switch (rc) {
    case 0: return "success";
    case GBSK_ERR_ARG: return "invalid argument";
    case GBSK_ERR_DIM: return "matrix dimension outside the kernel's supported range";
    case GBSK_ERR_FUNC: return "unknown or unsupported (function, tier)";
    case GBSK_ERR_WORKSPACE: return "workspace smaller than gbs_workspace_bytes()";
    case GBSK_ERR_ALLOC: return "host allocation failed";
    case GBSK_ERR_SCRATCH:
      return "batch needs more workspace than gbs_eval_device_noalloc's fixed scratch (GBSK_NOALLOC_SCRATCH_BYTES); "
             "split the batch or use gbs_eval_device";
  }
  return hipGetErrorString((hipError_t)rc);
}

int gbs_is_device_build(void) { return 1; }

int gbs_check_div_sqrt(const double* a, const double* b, long long n, double* q, double* s, double* h) {
  if (n <= 0) return 0;
  if (!a || !b || !q || !s || !h) return GBSK_ERR_ARG;
  Ctx& c = global_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int rc = c.init();
  const size_t bytes = (size_t)n * sizeof(double);
  double* d = nullptr;  // a, b, q, s, h
  if (!rc) rc = (int)hipMalloc(&d, 5 * bytes);
  if (rc) return rc;
  double *da = d, *db = d + n, *dq = d + 2 * n, *ds = d + 3 * n, *dh = d + 4 * n;
  rc = (int)hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, c.stream);
  if (!rc) rc = (int)hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, c.stream);
  if (!rc) rc = tor_div_sqrt_probe(da, db, n, dq, ds, c.stream);
  if (!rc) rc = haf_div_probe(da, n, dh, c.stream);
  if (!rc) rc = (int)hipMemcpyAsync(q, dq, bytes, hipMemcpyDeviceToHost, c.stream);
  if (!rc) rc = (int)hipMemcpyAsync(s, ds, bytes, hipMemcpyDeviceToHost, c.stream);
  if (!rc) rc = (int)hipMemcpyAsync(h, dh, bytes, hipMemcpyDeviceToHost, c.stream);
  const int rs = (int)hipStreamSynchronize(c.stream);
  (void)hipFree(d);
  return rc ? rc : rs;
}

int gbs_check_rsqrt(const double* a, long long n, double* r) {
  if (n <= 0) return 0;
  if (!a || !r) return GBSK_ERR_ARG;
  Ctx& c = global_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  int rc = c.init();
  const size_t bytes = (size_t)n * sizeof(double);
  double* d = nullptr;  // a, r
  if (!rc) rc = (int)hipMalloc(&d, 2 * bytes);
  if (rc) return rc;
  rc = (int)hipMemcpyAsync(d, a, bytes, hipMemcpyHostToDevice, c.stream);
  if (!rc) rc = tor_rsqrt_probe(d, n, d + n, c.stream);
  if (!rc) rc = (int)hipMemcpyAsync(r, d + n, bytes, hipMemcpyDeviceToHost, c.stream);
  const int rs = (int)hipStreamSynchronize(c.stream);
  (void)hipFree(d);
  return rc ? rc : rs;
}

int gbs_tor_recursive_single_host(const double* O, int n, int g, double* out) {
  return tor_single_host(GBSK_FP64, O, n, g, out, nullptr);
}
int gbs_tor_single_certified_host(const double* O, int n, int g, double* out, double* bound) {
  return tor_single_host(GBSK_CERTIFIED, O, n, g, out, bound);
}
int gbs_tor_single_ddcertified_host(const double* O, int n, int g, double* out, double* bound) {
  return tor_single_host(GBSK_DD_CERTIFIED, O, n, g, out, bound);
}

int gbs_lhaf_repeated_host(const gbs_complex* A, const gbs_complex* gamma, int M, const int* reps, int B,
                           gbs_complex* out) {
  return sieve_host(false, A, gamma, M, reps, B, out, nullptr);
}
int gbs_lhaf_repeated_cert_host(const gbs_complex* A, const gbs_complex* gamma, int M, const int* reps, int B,
                                gbs_complex* out, double* bound) {
  return sieve_host(true, A, gamma, M, reps, B, out, bound);
}

}  // extern "C"

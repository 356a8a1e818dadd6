// batch.hip -- heterogeneous batches: many instances of mixed function, tier
// and size evaluated by one call, and the continuous-batching server on top.
//
// The reference evaluates a ragged batch one uniform bucket at a time, each a
// synchronous host round trip (gbskernels/__init__.py:813-852 over
// core/host_api.cu:537-590). Here a batch is planned once:
//  * items are grouped into (function, tier, dim) buckets, longest (canonical
//    flops) first;
//  * inputs are packed bucket-contiguous into pinned host memory and copied to
//    the device in one transfer;
//  * buckets are launched on a pool of streams (event fork / join), so small and
//    large sizes run concurrently and fill the device;
//  * results come back in one transfer and are scattered to item order;
//  * optionally the device part (copy in, launches, copy out) is captured into a
//    hipGraph, cached by bucket signature, and replayed.
// Every bucket runs the same kernels as a uniform call, so an item's value is
// bit-identical alone, in a bucket, or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "../../include/gbsk.h"
#include "batch.hpp"
#include "common.hpp"
#include "kernels.hpp"

namespace gbsk {

// canonical flops of one evaluation (reference algorithm as written; the cost
// model that orders buckets longest first)
double canonical_flops(int func, int tier, int dim) {
  double f = 0.0;
  if (func == GBSK_PERM) {
    f = dim > 0 ? std::ldexp(1.0, dim - 1) * (8.0 * dim - 4.0) : 1.0;
  } else if (func == GBSK_HAF || func == GBSK_LHAF) {
    const int n = dim / 2;
    double c = 1.0;  // C(n, m)
    for (int m = 1; m <= n; ++m) {
      c = c * (n - m + 1) / m;
      const double s = 2.0 * m;
      f += c * ((n - 1) * 8.0 * s * s * s + n * 2.0 * s + 4.0 * n * (n + 1) + 4.0 * n + 2.0);
      if (func == GBSK_LHAF) f += c * n * (8.0 * s * s + 8.0 * s + 6.0);
    }
  } else {
    const int nm = dim / 2;
    double c = 1.0;
    for (int k = 1; k <= nm; ++k) {
      c = c * (nm - k + 1) / k;
      f += c * (8.0 * k * k + 4.0);  // two appended Cholesky rows + the leaf
    }
    f += 3.0 * std::ldexp(1.0, nm);
  }
  if (tier == GBSK_DD || tier == GBSK_DD_CERTIFIED) f *= 10.0;
  if (tier == GBSK_CERTIFIED) f *= 2.0;
  return f;
}

struct Bucket {
  int func, tier, dim;
  std::vector<int> items;
  size_t in_off = 0, ws_off = 0, ws_bytes = 0;
  size_t res_off = 0;  // bytes: values (16 B each) then auxiliaries (8 B each)
  double cost = 0.0;
  int stream = 0;     // index into the owner's stream pool
  bool done = false;  // per-bucket completion (server slots)
  // hafnian buckets of one (function, tier, size class) share their launches:
  // the leader (first in launch order) launches all members, itself included
  int leader = -1;           // index of the group's leader, -1 if not grouped
  std::vector<int> members;  // (leader only) the group, leader first
};

struct GraphEntry {
  std::vector<long long> sig;
  hipGraphExec_t exec = nullptr;
  unsigned long long gen = 0;
};

}  // namespace gbsk

using namespace gbsk;

struct gbs_batch {
  std::mutex mu;
  int nstreams = 4;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> joins;
  hipEvent_t ev_in = nullptr;
  void *d_in = nullptr, *d_res = nullptr, *d_ws = nullptr;
  size_t cap_in = 0, cap_res = 0, cap_ws = 0;
  cplx* h_in = nullptr;
  char* h_res = nullptr;  // per bucket: values then auxiliaries (one copy back each)
  size_t hcap_in = 0, hcap_res = 0;
  size_t res_total = 0;
  size_t budget = 0;
  bool slack = true;  // no memory budget: buffers may grow past a batch's need
  bool use_graphs = false;
  std::vector<GraphEntry> graphs;  // small LRU keyed by bucket signature
  unsigned long long gen = 0;
  unsigned long long grows = 0;     // device buffer reallocations (invalidate graphs)
  gbs_batch_stats stats{};
  hipEvent_t ev_done = nullptr;  // completion of the submitted batch
  std::vector<Bucket> cur;       // plan of the submitted batch
  int cur_n = 0;
  bool per_bucket = false;              // server slot: per-bucket copies back and events
  std::vector<hipEvent_t> bucket_ev;    // grow-only pool, one per bucket of the plan
};

namespace gbsk {

// Without a memory budget, buffers grow to at least twice their old capacity:
// hipFree waits for the whole device, so a server slot that reallocated on
// every larger batch would stall the other slots' streams each time. Under a
// budget they grow to exactly what a batch needs.
static int grow_dev(void** p, size_t* cap, size_t need, bool slack, unsigned long long* grows) {
  if (need <= *cap) return 0;
  if (*p) (void)hipFree(*p);
  if (slack && need < 2 * *cap) need = 2 * *cap;
  *p = nullptr;
  *cap = 0;
  if (hipMalloc(p, need) != hipSuccess) {
    *p = nullptr;
    return (int)hipErrorOutOfMemory;
  }
  *cap = need;
  ++*grows;
  return 0;
}

template <class T>
static int grow_pinned(T** p, size_t* cap, size_t need_elems, bool slack, unsigned long long* grows) {
  if (need_elems <= *cap) return 0;
  if (*p) (void)hipHostFree(*p);
  if (slack && need_elems < 2 * *cap) need_elems = 2 * *cap;
  *p = nullptr;
  *cap = 0;
  if (hipHostMalloc((void**)p, need_elems * sizeof(T), hipHostMallocDefault) != hipSuccess) {
    *p = nullptr;
    return GBSK_ERR_ALLOC;
  }
  *cap = need_elems;
  ++*grows;
  return 0;
}

static void drop_graphs(gbs_batch* b) {
  for (auto& g : b->graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  b->graphs.clear();
}

// device bytes a batch object holds (its grow-only buffers)
static size_t held_bytes(const gbs_batch* b) { return b->cap_in + b->cap_res + b->cap_ws; }

// frees a batch object's device buffers (its streams must be idle); cached
// graphs refer to them and are dropped
static void release_device(gbs_batch* b) {
  for (void** p : {&b->d_in, &b->d_res, &b->d_ws}) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
  }
  b->cap_in = b->cap_res = b->cap_ws = 0;
  drop_graphs(b);
}

// the device bytes a batch object holds after growing its buffers to a batch
// of in_b input, res_b result and ws_b workspace bytes (batch_submit's grow_dev
// requests, no slack)
static size_t footprint(const gbs_batch* b, size_t in_b, size_t res_b, size_t ws_b) {
  return std::max(b->cap_in, in_b + 16) + std::max(b->cap_res, res_b + 16) + std::max(b->cap_ws, ws_b + 256);
}


// plan: buckets (longest first) with packed offsets; returns bytes needed
// Buckets are ordered longest first (lpt: the makespan of one batch) or shortest
// first (!lpt: mean completion time of a served batch) and each is given to the
// least-loaded of nstreams streams.
static int plan(int n, const int* func, const int* tier, const int* dim, std::vector<Bucket>& out, size_t* in_elems,
                size_t* ws_total, size_t* res_total, bool lpt, int nstreams) {
  std::map<long long, int> index;
  out.clear();
  for (int i = 0; i < n; ++i) {
    if (func[i] < GBSK_PERM || func[i] > GBSK_TOR || tier[i] < GBSK_FP64 || tier[i] > GBSK_DD_CERTIFIED)
      return GBSK_ERR_FUNC;
    const int cap = gbs_max_dim(func[i], tier[i]);
    if (cap == 0) return GBSK_ERR_FUNC;  // tier not built for this function
    if (dim[i] < 0 || dim[i] > cap || (func[i] == GBSK_TOR && (dim[i] & 1))) return GBSK_ERR_DIM;
    const long long key = ((long long)func[i] << 40) | ((long long)tier[i] << 32) | (long long)dim[i];
    auto it = index.find(key);
    if (it == index.end()) {
      index.emplace(key, (int)out.size());
      Bucket bk;
      bk.func = func[i];
      bk.tier = tier[i];
      bk.dim = dim[i];
      out.push_back(bk);
      out.back().items.push_back(i);
    } else {
      out[it->second].items.push_back(i);
    }
  }
  for (auto& bk : out) bk.cost = canonical_flops(bk.func, bk.tier, bk.dim) * (double)bk.items.size();
  if (lpt)
    std::stable_sort(out.begin(), out.end(), [](const Bucket& a, const Bucket& c) { return a.cost > c.cost; });
  else
    std::stable_sort(out.begin(), out.end(), [](const Bucket& a, const Bucket& c) { return a.cost < c.cost; });
  // group hafnian buckets: same function and tier, at most HAF_MAX_JOBS each
  // (one launch per subset size for the whole group)
  {
    std::map<long long, int> open;
    for (int k = 0; k < (int)out.size(); ++k) {
      Bucket& bk = out[k];
      if (bk.func != GBSK_HAF && bk.func != GBSK_LHAF) continue;
      const long long key = ((long long)bk.func << 8) | (long long)bk.tier;
      auto it = open.find(key);
      if (it != open.end() && (int)out[it->second].members.size() < HAF_MAX_JOBS) {
        bk.leader = it->second;
        out[it->second].members.push_back(k);
      } else {
        bk.leader = k;
        bk.members.assign(1, k);
        open[key] = k;
      }
    }
  }
  std::vector<double> load((size_t)(nstreams < 1 ? 1 : nstreams), 0.0);
  size_t ie = 0, re = 0, ws = 0;
  for (int kb = 0; kb < (int)out.size(); ++kb) {
    Bucket& bk = out[kb];
    if (bk.leader >= 0 && bk.leader != kb) {
      bk.stream = out[bk.leader].stream;  // members run in their leader's launches
    } else {
      double cost = 0.0;
      if (bk.leader == kb)
        for (int m : bk.members) cost += out[m].cost;
      else
        cost = bk.cost;
      const size_t k = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
      bk.stream = (int)k;
      load[k] += cost;
    }
    bk.in_off = ie;
    ie += bk.items.size() * (size_t)bk.dim * bk.dim;
    // results: a group's members back to back (one copy brings the group home)
    if (bk.leader < 0 || bk.leader == kb) {
      const std::vector<int> one(1, kb);
      for (int m : (bk.leader == kb ? bk.members : one)) {
        out[m].res_off = re;
        re += (out[m].items.size() * (sizeof(cplx) + sizeof(double)) + 15) & ~(size_t)15;
      }
    }
    bk.ws_bytes = gbs_workspace_bytes(bk.func, bk.tier, bk.dim, (int)bk.items.size());
    bk.ws_off = ws;
    ws += (bk.ws_bytes + 255) & ~(size_t)255;
  }
  *in_elems = ie;
  *ws_total = ws;
  *res_total = re;
  return 0;
}

// tiers whose kernels write a per-element double (kappa sum|term| or bound)
static bool tier_has_aux(int tier) { return tier == GBSK_KAPPA || tier == GBSK_CERTIFIED || tier == GBSK_DD_CERTIFIED; }

// device bytes a batch occupies (inputs, outputs, auxiliaries, workspace)
static size_t batch_bytes(size_t in_elems, size_t res_total, size_t ws_total) {
  return in_elems * sizeof(cplx) + 16 + res_total + 16 + ws_total + 256;  // the buffers' requests
}

static int enqueue_device(gbs_batch* b, const std::vector<Bucket>& bks, size_t in_elems) {
  hipStream_t s0 = b->streams[0];
  int rc = 0;
  if (in_elems) rc = (int)hipMemcpyAsync(b->d_in, b->h_in, in_elems * sizeof(cplx), hipMemcpyHostToDevice, s0);
  if (!rc) rc = (int)hipEventRecord(b->ev_in, s0);
  std::vector<char> waited((size_t)b->nstreams, 0);
  waited[0] = 1;
  for (size_t k = 0; !rc && k < bks.size(); ++k) {
    const Bucket& bk = bks[k];
    hipStream_t s = b->streams[bk.stream];
    if (!waited[bk.stream]) {
      rc = (int)hipStreamWaitEvent(s, b->ev_in, 0);
      waited[bk.stream] = 1;
    }
    if (bk.leader >= 0 && bk.leader != (int)k) continue;  // launched with its group
    auto res_of = [&](const Bucket& x) { return (char*)b->d_res + x.res_off; };
    if (!rc && bk.leader == (int)k) {
      HafJob jobs[HAF_MAX_JOBS];
      int nj = 0;
      for (int m : bk.members) {
        const Bucket& x = bks[m];
        const size_t c = x.items.size();
        jobs[nj++] = HafJob{(const cplx*)b->d_in + x.in_off, (cplx*)res_of(x), (double*)(res_of(x) + c * sizeof(cplx)),
                            (char*)b->d_ws + x.ws_off, x.dim, (int)c};
      }
      rc = haf_launch_multi(bk.func == GBSK_LHAF, bk.tier, jobs, nj, s);
    } else if (!rc) {
      const size_t cnt = bk.items.size();
      rc = gbs_eval_device(bk.func, bk.tier, (const gbs_complex*)((cplx*)b->d_in + bk.in_off), bk.dim, (int)cnt,
                           (gbs_complex*)res_of(bk), (double*)(res_of(bk) + cnt * sizeof(cplx)),
                           (char*)b->d_ws + bk.ws_off, bk.ws_bytes, (void*)s);
    }
    if (!rc && b->per_bucket) {
      // a bucket's (or a whole group's, laid out back to back) values and
      // auxiliaries come back in one copy on the launching stream, then the
      // events of its buckets fire
      const std::vector<int> one(1, (int)k);
      const std::vector<int>& mem = bk.leader == (int)k ? bk.members : one;
      const Bucket& last = bks[mem.back()];
      const size_t beg = bk.res_off, end = last.res_off + last.items.size() * (sizeof(cplx) + sizeof(double));
      rc = (int)hipMemcpyAsync(b->h_res + beg, (char*)b->d_res + beg, end - beg, hipMemcpyDeviceToHost, s);
      for (int m : mem)
        if (!rc) rc = (int)hipEventRecord(b->bucket_ev[m], s);
    }
  }
  if (b->per_bucket) return rc;
  // join only the streams this batch forked (under capture an unforked stream
  // is not part of the graph)
  for (int k = 1; !rc && k < b->nstreams; ++k) {
    if (!waited[k]) continue;
    rc = (int)hipEventRecord(b->joins[k], b->streams[k]);
    if (!rc) rc = (int)hipStreamWaitEvent(s0, b->joins[k], 0);
  }
  if (!rc && b->res_total)
    rc = (int)hipMemcpyAsync(b->h_res, b->d_res, b->res_total, hipMemcpyDeviceToHost, s0);
  return rc;
}

// Plan, pack and queue one batch on b's streams; completion is b->ev_done.
// ptrs[i] is item i's matrix (host memory, dim[i]^2 entries).
static int batch_submit(gbs_batch* b, int n, const int* func, const int* tier, const int* dim,
                        const cplx* const* ptrs) {
  b->cur.clear();
  b->cur_n = 0;
  if (n <= 0) return 0;
  size_t in_elems = 0, ws_total = 0, res_total = 0;
  int rc = plan(n, func, tier, dim, b->cur, &in_elems, &ws_total, &res_total, !b->per_bucket, b->nstreams);
  if (rc) return rc;
  b->res_total = res_total;
  const size_t need = batch_bytes(in_elems, res_total, ws_total);
  if (b->budget && need > b->budget) return GBSK_ERR_WORKSPACE;
  // under a budget the grow-only buffers must not together exceed it: buffers
  // that grew for different earlier batches are given back first
  if (b->budget && footprint(b, in_elems * sizeof(cplx), res_total, ws_total) > b->budget) {
    for (hipStream_t st : b->streams) (void)hipStreamSynchronize(st);
    release_device(b);
  }
  // grow-only buffers; any reallocation changes addresses a captured graph
  // refers to, so cached graphs are dropped
  const unsigned long long g0 = b->grows;
  const bool slack = b->slack;
  rc = grow_dev(&b->d_in, &b->cap_in, in_elems * sizeof(cplx) + 16, slack, &b->grows);
  if (!rc) rc = grow_dev(&b->d_res, &b->cap_res, res_total + 16, slack, &b->grows);
  if (!rc) rc = grow_dev(&b->d_ws, &b->cap_ws, ws_total + 256, slack, &b->grows);
  b->stats.device_reallocs += b->grows - g0;
  if (!rc) rc = grow_pinned(&b->h_in, &b->hcap_in, in_elems + 1, slack, &b->grows);
  if (!rc) rc = grow_pinned(&b->h_res, &b->hcap_res, res_total + 16, slack, &b->grows);
  if (rc) return rc;
  if (b->grows != g0) drop_graphs(b);
  const std::vector<Bucket>& bks = b->cur;
  if (b->per_bucket)
    while (b->bucket_ev.size() < bks.size()) {
      hipEvent_t e = nullptr;
      rc = (int)hipEventCreateWithFlags(&e, hipEventDisableTiming);
      if (rc) return rc;
      b->bucket_ev.push_back(e);
    }
  // pack bucket-contiguous
  for (const Bucket& bk : bks) {
    const size_t m2 = (size_t)bk.dim * bk.dim;
    cplx* dst = b->h_in + bk.in_off;
    for (size_t j = 0; j < bk.items.size(); ++j) std::memcpy(dst + j * m2, ptrs[bk.items[j]], m2 * sizeof(cplx));
  }
  // device part: replay a captured graph of the same signature, else launch
  bool done = false;
  if (b->use_graphs && !b->per_bucket) {
    std::vector<long long> sig;
    sig.reserve(bks.size() * 4 + 1);
    sig.push_back((long long)in_elems);
    sig.push_back((long long)res_total);
    for (const Bucket& bk : bks) {
      sig.push_back(bk.func);
      sig.push_back(bk.tier);
      sig.push_back(bk.dim);
      sig.push_back((long long)bk.items.size());
    }
    GraphEntry* hit = nullptr;
    for (auto& g : b->graphs)
      if (g.sig == sig) hit = &g;
    if (!hit) {
      hipGraph_t graph = nullptr;
      rc = (int)hipStreamBeginCapture(b->streams[0], hipStreamCaptureModeThreadLocal);
      if (!rc) {
        int rc2 = enqueue_device(b, bks, in_elems);
        rc = (int)hipStreamEndCapture(b->streams[0], &graph);
        if (rc2) rc = rc2;
      }
      GraphEntry ge;
      ge.sig = sig;
      if (!rc) rc = (int)hipGraphInstantiate(&ge.exec, graph, nullptr, nullptr, 0);
      if (graph) (void)hipGraphDestroy(graph);
      if (rc) return rc;
      if (b->graphs.size() >= 16) {
        auto oldest = std::min_element(b->graphs.begin(), b->graphs.end(),
                                       [](const GraphEntry& x, const GraphEntry& y) { return x.gen < y.gen; });
        (void)hipGraphExecDestroy(oldest->exec);
        b->graphs.erase(oldest);
      }
      b->graphs.push_back(ge);
      hit = &b->graphs.back();
      ++b->stats.graphs_captured;
    } else {
      ++b->stats.graph_replays;
    }
    hit->gen = ++b->gen;
    rc = (int)hipGraphLaunch(hit->exec, b->streams[0]);
    done = true;
  }
  if (!done && !rc) rc = enqueue_device(b, bks, in_elems);
  if (!rc && !b->per_bucket) rc = (int)hipEventRecord(b->ev_done, b->streams[0]);
  if (!rc) b->cur_n = n;
  return rc;
}

// Wait for the submitted batch and scatter its results to item order.
static int batch_complete(gbs_batch* b, gbs_complex* out, double* aux) {
  if (b->cur_n == 0) return 0;
  int rc = (int)hipEventSynchronize(b->ev_done);
  if (rc) return rc;
  for (const Bucket& bk : b->cur) {
    const bool has_aux = tier_has_aux(bk.tier);
    for (size_t j = 0; j < bk.items.size(); ++j) {
      const char* res = b->h_res + bk.res_off;
      out[bk.items[j]] = ((const gbs_complex*)res)[j];
      if (aux) aux[bk.items[j]] = has_aux ? ((const double*)(res + bk.items.size() * sizeof(cplx)))[j] : 0.0;
    }
  }
  ++b->stats.batches;
  b->stats.items += (unsigned long long)b->cur_n;
  b->stats.buckets += (unsigned long long)b->cur.size();
  b->cur_n = 0;
  return 0;
}

static int run_batch(gbs_batch* b, int n, const int* func, const int* tier, const int* dim, const gbs_complex* in,
                     const long long* off, gbs_complex* out, double* aux) {
  if (n <= 0) return 0;
  if (!func || !tier || !dim || !in || !off || !out) return GBSK_ERR_ARG;
  std::vector<const cplx*> ptrs((size_t)n);
  for (int i = 0; i < n; ++i) ptrs[i] = (const cplx*)in + off[i];
  int rc = batch_submit(b, n, func, tier, dim, ptrs.data());
  if (!rc) rc = batch_complete(b, out, aux);
  if (rc) (void)hipStreamSynchronize(b->streams[0]);
  return rc;
}

}  // namespace gbsk

// ---------------------------------------------------------------------------
// continuous-batching server

namespace gbsk {

static long long now_ns() {
  return (long long)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct Req {
  int func = 0, tier = 0, dim = 0;
  int run_tier = 0;  // tier of the pass being evaluated
  std::vector<cplx> a;        // owned copy of the matrix (gbs_server_submit)
  const cplx* src = nullptr;  // or the caller's matrix, alive until done (replay)
  const cplx* mat() const { return src ? src : a.data(); }
  long long arrival = 0, admitted = 0, done = 0;
  gbs_complex out{0.0, 0.0};
  double aux = 0.0;
  int tier_used = 0, status = 0;
  bool finished = false;
};

// A batch slot: FREE -> ASSIGNED (dispatcher) -> SUBMITTED (the slot's worker
// thread packed and queued it) -> FREE (dispatcher, all buckets complete).
// Each slot has its own worker so that packing and launching, the host cost of
// a batch, run for several batches at once.
enum SlotState { SLOT_FREE = 0, SLOT_ASSIGNED = 1, SLOT_SUBMITTED = 2 };
struct Slot {
  gbs_batch* b = nullptr;
  std::vector<Req*> reqs;  // local item index -> request
  size_t reserved = 0;     // admitted device bytes
  size_t remaining = 0;    // buckets not yet complete
  int state = SLOT_FREE;
  std::thread worker;
};


}  // namespace gbsk

struct gbs_server {
  gbs_server_config cfg{};
  int device = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_slot;
  std::deque<Req*> pending;
  std::map<long long, Req*> table;
  long long next_ticket = 1;
  std::vector<Slot> slots;
  bool stop = false;
  bool stop_workers = false;
  std::thread th;
  gbs_server_stats stats{};
  unsigned long long in_flight = 0;
};

namespace gbsk {

static void finish_req(Req* r, long long t) {
  r->done = t;
  r->finished = true;
}

// Completes bucket k of a slot: results to the requests, or an auto pass that
// needs double-double back to the queue (front). Called with s->mu held.
static void complete_bucket(gbs_server* s, Slot& sl, Bucket& bk, int err, std::vector<Req*>& requeue) {
  const long long t = now_ns();
  for (size_t j = 0; j < bk.items.size(); ++j) {
    Req* r = sl.reqs[bk.items[j]];
    if (err) {
      r->status = err;
      finish_req(r, t);
      continue;
    }
    const char* res = sl.b->h_res + bk.res_off;
    const gbs_complex v = ((const gbs_complex*)res)[j];
    const double ax = ((const double*)(res + bk.items.size() * sizeof(cplx)))[j];
    if (r->tier == GBSK_AUTO && r->run_tier == GBSK_KAPPA) {
      r->out = v;
      r->aux = ax;
      r->tier_used = GBSK_KAPPA;
      // kappa = sum|term| / max(|value|, 1e-300) >= kappa_max, as the auto tier
      const double kappa = ax / std::max(std::hypot(v.re, v.im), 1e-300);
      const bool risky = kappa >= s->cfg.kappa_max;
      bool dd_ok = r->dim <= gbs_max_dim(r->func, GBSK_DD);
      if (dd_ok && r->func == GBSK_TOR) {  // the dd torontonian is real-domain
        const cplx* m = r->mat();
        for (size_t e = 0; e < (size_t)r->dim * r->dim && dd_ok; ++e) dd_ok = m[e].im == 0.0;
      }
      if (risky && dd_ok) {
        r->run_tier = GBSK_DD;
        requeue.push_back(r);
        ++s->stats.escalated;
        continue;
      }
      finish_req(r, t);
      continue;
    }
    r->out = v;
    if (r->tier != GBSK_AUTO) {
      r->aux = tier_has_aux(r->tier) ? ax : 0.0;
      r->tier_used = r->tier;
    } else {
      r->tier_used = GBSK_DD;  // aux keeps the first pass's sum|term|
    }
    finish_req(r, t);
  }
  bk.done = true;
}

// release a slot whose batch is complete or failed (s->mu held)
static void free_slot(gbs_server* s, Slot& sl) {
  (void)s;
  sl.state = SLOT_FREE;
  sl.reserved = 0;
  sl.reqs.clear();
}

// Device bytes the slots other than `self` hold or will hold: a busy slot its
// admitted footprint (its buffers grow to at most that), a free slot what its
// buffers hold now. Called with s->mu held.
static size_t others_bytes(const gbs_server* s, const Slot* self) {
  size_t t = 0;
  for (const Slot& o : s->slots)
    if (&o != self) t += o.state == SLOT_FREE ? held_bytes(o.b) : o.reserved;
  return t;
}

// The slot's worker: packs and queues each batch the dispatcher assigns.
static void slot_worker(gbs_server* s, Slot* sl) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_slot.wait(lk, [s, sl] { return sl->state == SLOT_ASSIGNED || s->stop_workers; });
    if (sl->state != SLOT_ASSIGNED) return;
    const int n = (int)sl->reqs.size();
    std::vector<int> fn(n), tr(n), dm(n);
    std::vector<const cplx*> ptrs(n);
    for (int i = 0; i < n; ++i) {
      fn[i] = sl->reqs[i]->func;
      tr[i] = sl->reqs[i]->run_tier;
      dm[i] = sl->reqs[i]->dim;
      ptrs[i] = sl->reqs[i]->mat();
    }
    lk.unlock();  // pack and launch without holding the server lock
    const int rc = batch_submit(sl->b, n, fn.data(), tr.data(), dm.data(), ptrs.data());
    if (rc)
      for (hipStream_t st : sl->b->streams) (void)hipStreamSynchronize(st);
    lk.lock();
    if (rc) {
      // nothing of a failed submission is awaited: fail its requests
      const long long t = now_ns();
      for (Req* r : sl->reqs) {
        r->status = rc;
        finish_req(r, t);
      }
      for (auto& bk : sl->b->cur) bk.done = true;
      sl->remaining = 0;
      free_slot(s, *sl);
      s->cv_done.notify_all();
    } else {
      sl->remaining = sl->b->cur.size();
      s->stats.buckets += (unsigned long long)sl->remaining;
      sl->state = SLOT_SUBMITTED;
    }
    s->cv_work.notify_all();
  }
}

static void dispatcher(gbs_server* s) {
  (void)hipSetDevice(s->device);
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    bool progress = false;
    bool any_busy = false;
    // 1. completions of submitted batches, bucket by bucket
    for (Slot& sl : s->slots) {
      if (sl.state == SLOT_ASSIGNED) any_busy = true;
      if (sl.state != SLOT_SUBMITTED) continue;
      std::vector<Req*> requeue;
      bool fin = false;
      for (size_t k = 0; k < sl.b->cur.size(); ++k) {
        Bucket& bk = sl.b->cur[k];
        if (bk.done) continue;
        const hipError_t q = hipEventQuery(sl.b->bucket_ev[k]);
        if (q == hipErrorNotReady) continue;
        complete_bucket(s, sl, bk, q == hipSuccess ? 0 : (int)q, requeue);
        fin = true;
        --sl.remaining;
      }
      for (auto it = requeue.rbegin(); it != requeue.rend(); ++it) s->pending.push_front(*it);
      if (fin) {
        progress = true;
        s->cv_done.notify_all();
      }
      if (sl.remaining == 0)
        free_slot(s, sl);
      else
        any_busy = true;
    }
    // 2. admission: everything pending that fits, into each free slot. Under a
    // budget, what every slot's buffers hold counts (they are grow-only), so
    // the slots together never hold more than budget_bytes of device memory.
    for (Slot& sl : s->slots) {
      if (sl.state != SLOT_FREE || s->pending.empty()) continue;
      std::vector<Req*> take;
      std::map<long long, int> counts;
      const size_t budget = s->cfg.budget_bytes;
      size_t others = budget ? others_bytes(s, &sl) : 0;
      size_t in_b = 0, res_b = 0, ws_b = 0, fp = 0;
      bool released = false;
      while (!s->pending.empty() && (s->cfg.max_batch <= 0 || (int)take.size() < s->cfg.max_batch)) {
        Req* r = s->pending.front();
        const long long key = ((long long)r->func << 40) | ((long long)r->run_tier << 32) | (long long)r->dim;
        int& c = counts[key];
        // the plan's packing: inputs back to back, per bucket 16-byte aligned
        // results and a 256-byte aligned workspace slice
        const size_t w1 = gbs_workspace_bytes(r->func, r->run_tier, r->dim, c + 1);
        const size_t w0 = c ? gbs_workspace_bytes(r->func, r->run_tier, r->dim, c) : 0;
        const size_t a_in = in_b + (size_t)r->dim * r->dim * sizeof(cplx);
        const size_t a_res = res_b + sizeof(cplx) + sizeof(double) + (c ? 0 : 16);
        const size_t a_ws = ws_b + (w1 > w0 ? w1 - w0 : 0) + (c ? 0 : 256);
        const size_t a_fp = budget ? footprint(sl.b, a_in, a_res, a_ws) : 0;
        if (budget && others + a_fp > budget) {
          if (take.empty() && !released) {
            // idle slots (this one included) give their buffers back, then retry
            for (Slot& o : s->slots)
              if (o.state == SLOT_FREE && held_bytes(o.b)) release_device(o.b);
            released = true;
            others = others_bytes(s, &sl);
            continue;
          }
          bool busy = false;
          for (const Slot& o : s->slots) busy |= (&o != &sl && o.state != SLOT_FREE);
          if (take.empty() && !busy) {  // can never fit
            s->pending.pop_front();
            r->status = GBSK_ERR_WORKSPACE;
            finish_req(r, now_ns());
            s->cv_done.notify_all();
            progress = true;
            continue;
          }
          break;
        }
        in_b = a_in;
        res_b = a_res;
        ws_b = a_ws;
        fp = a_fp;
        ++c;
        take.push_back(r);
        s->pending.pop_front();
      }
      if (take.empty()) continue;
      const long long t = now_ns();
      for (Req* r : take)
        if (!r->admitted) r->admitted = t;
      sl.reqs = std::move(take);
      sl.reserved = fp;
      if (budget) {
        const size_t total = others + fp;
        if (total > s->stats.device_bytes_max) s->stats.device_bytes_max = total;
      }
      sl.state = SLOT_ASSIGNED;
      ++s->stats.batches;
      progress = true;
      any_busy = true;
      unsigned long long inf = 0;
      for (const Slot& o : s->slots)
        if (o.state != SLOT_FREE) inf += (unsigned long long)o.reqs.size();
      if (inf > s->stats.max_in_flight) s->stats.max_in_flight = inf;
      s->cv_slot.notify_all();
    }
    if (s->stop && s->pending.empty() && !any_busy) break;
    if (!progress) {
      if (!any_busy) {
        s->cv_work.wait(lk, [s] { return s->stop || !s->pending.empty(); });
      } else {
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
      }
    }
  }
}

}  // namespace gbsk

extern "C" {

int gbs_batch_create(gbs_batch** out, size_t budget_bytes, int nstreams, int use_graphs) {
  if (!out) return GBSK_ERR_ARG;
  *out = nullptr;
  gbs_batch* b = new (std::nothrow) gbs_batch();
  if (!b) return GBSK_ERR_ALLOC;
  b->nstreams = nstreams < 1 ? 1 : (nstreams > 16 ? 16 : nstreams);
  b->budget = budget_bytes;
  b->slack = budget_bytes == 0;
  b->use_graphs = use_graphs != 0;
  b->streams.resize(b->nstreams, nullptr);
  b->joins.resize(b->nstreams, nullptr);
  int rc = 0;
  for (int k = 0; !rc && k < b->nstreams; ++k) {
    rc = (int)hipStreamCreateWithFlags(&b->streams[k], hipStreamNonBlocking);
    if (!rc) rc = (int)hipEventCreateWithFlags(&b->joins[k], hipEventDisableTiming);
  }
  if (!rc) rc = (int)hipEventCreateWithFlags(&b->ev_in, hipEventDisableTiming);
  if (!rc) rc = (int)hipEventCreateWithFlags(&b->ev_done, hipEventDisableTiming);
  if (rc) {
    gbs_batch_destroy(b);
    return rc;
  }
  *out = b;
  return 0;
}

int gbs_batch_run(gbs_batch* b, int n, const int* func, const int* tier, const int* dim, const gbs_complex* in,
                  const long long* off, gbs_complex* out, double* aux) {
  if (!b) return GBSK_ERR_ARG;
  std::lock_guard<std::mutex> g(b->mu);
  return run_batch(b, n, func, tier, dim, in, off, out, aux);
}

int gbs_batch_get_stats(gbs_batch* b, gbs_batch_stats* s) {
  if (!b || !s) return GBSK_ERR_ARG;
  std::lock_guard<std::mutex> g(b->mu);
  *s = b->stats;
  return 0;
}

int gbs_batch_destroy(gbs_batch* b) {
  if (!b) return 0;
  {
    std::lock_guard<std::mutex> g(b->mu);
    for (auto s : b->streams)
      if (s) (void)hipStreamSynchronize(s);
    drop_graphs(b);
    if (b->d_in) (void)hipFree(b->d_in);
    if (b->d_res) (void)hipFree(b->d_res);
    if (b->d_ws) (void)hipFree(b->d_ws);
    if (b->h_in) (void)hipHostFree(b->h_in);
    if (b->h_res) (void)hipHostFree(b->h_res);
    for (auto e : b->joins)
      if (e) (void)hipEventDestroy(e);
    if (b->ev_in) (void)hipEventDestroy(b->ev_in);
    if (b->ev_done) (void)hipEventDestroy(b->ev_done);
    for (auto e : b->bucket_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto s : b->streams)
      if (s) (void)hipStreamDestroy(s);
  }
  delete b;
  return 0;
}

long long gbs_now_ns(void) { return now_ns(); }

int gbs_server_create(gbs_server** out, const gbs_server_config* cfg) {
  if (!out) return GBSK_ERR_ARG;
  *out = nullptr;
  gbs_server_config c{};
  if (cfg) c = *cfg;
  if (c.slots < 1) c.slots = 1;
  if (c.slots > 8) c.slots = 8;
  if (c.streams_per_slot < 1) c.streams_per_slot = 1;
  if (c.streams_per_slot > 16) c.streams_per_slot = 16;
  if (!(c.kappa_max > 0.0)) c.kappa_max = 1e8;
  gbs_server* s = new (std::nothrow) gbs_server();
  if (!s) return GBSK_ERR_ALLOC;
  s->cfg = c;
  int rc = (int)hipGetDevice(&s->device);
  s->slots.resize((size_t)c.slots);
  for (size_t k = 0; !rc && k < s->slots.size(); ++k) {
    rc = gbs_batch_create(&s->slots[k].b, 0, c.streams_per_slot, 0);
    if (!rc) {
      s->slots[k].b->per_bucket = true;
      s->slots[k].b->slack = c.budget_bytes == 0;  // a server budget bounds what slots hold
    }
  }
  if (!rc) {
    try {
      for (auto& sl : s->slots) sl.worker = std::thread(slot_worker, s, &sl);
      s->th = std::thread(dispatcher, s);
    } catch (...) {
      rc = GBSK_ERR_ALLOC;
    }
  }
  if (rc) {
    {
      std::lock_guard<std::mutex> g(s->mu);
      s->stop = true;
      s->stop_workers = true;
    }
    s->cv_slot.notify_all();
    s->cv_work.notify_all();
    if (s->th.joinable()) s->th.join();
    for (auto& sl : s->slots)
      if (sl.worker.joinable()) sl.worker.join();
    for (auto& sl : s->slots)
      if (sl.b) gbs_batch_destroy(sl.b);
    delete s;
    return rc;
  }
  *out = s;
  return 0;
}

}  // extern "C"

namespace gbsk {

// a validated request; copy: own the matrix, else borrow the caller's
static int make_req(int func, int tier, int dim, const gbs_complex* in, long long arrival_ns, bool copy, Req** out) {
  if (!in && dim > 0) return GBSK_ERR_ARG;
  if (func < GBSK_PERM || func > GBSK_TOR || tier < GBSK_FP64 || tier > GBSK_AUTO) return GBSK_ERR_FUNC;
  const int run = tier == GBSK_AUTO ? GBSK_KAPPA : tier;
  const int cap = gbs_max_dim(func, run);
  if (cap == 0) return GBSK_ERR_FUNC;
  if (dim < 0 || dim > cap || (func == GBSK_TOR && (dim & 1))) return GBSK_ERR_DIM;
  Req* r = new (std::nothrow) Req();
  if (!r) return GBSK_ERR_ALLOC;
  r->func = func;
  r->tier = tier;
  r->run_tier = run;
  r->dim = dim;
  r->arrival = arrival_ns > 0 ? arrival_ns : now_ns();
  const size_t m2 = (size_t)dim * dim;
  if (copy) {
    try {
      r->a.resize(m2);
    } catch (...) {
      delete r;
      return GBSK_ERR_ALLOC;
    }
    if (m2) std::memcpy(r->a.data(), in, m2 * sizeof(cplx));
  } else {
    r->src = (const cplx*)in;
  }
  *out = r;
  return 0;
}

// queue requests under one lock; tickets[i] receives request i's ticket
static int enqueue_reqs(gbs_server* s, Req* const* rs, int n, long long* tickets) {
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) return GBSK_ERR_ARG;
    for (int i = 0; i < n; ++i) {
      tickets[i] = s->next_ticket++;
      s->table.emplace(tickets[i], rs[i]);
      s->pending.push_back(rs[i]);
    }
    s->stats.submitted += (unsigned long long)n;
  }
  s->cv_work.notify_one();
  return 0;
}

}  // namespace gbsk

extern "C" {

int gbs_server_submit(gbs_server* s, int func, int tier, int dim, const gbs_complex* in, long long arrival_ns,
                      long long* ticket) {
  if (!s || !ticket) return GBSK_ERR_ARG;
  Req* r = nullptr;
  int rc = make_req(func, tier, dim, in, arrival_ns, true, &r);
  if (rc) return rc;
  rc = enqueue_reqs(s, &r, 1, ticket);
  if (rc) delete r;
  return rc;
}

int gbs_server_wait(gbs_server* s, long long ticket, gbs_complex* out, double* aux, gbs_request_info* info) {
  if (!s || !out) return GBSK_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->table.find(ticket);
  if (it == s->table.end()) return GBSK_ERR_ARG;
  Req* r = it->second;
  s->cv_done.wait(lk, [r] { return r->finished; });
  *out = r->out;
  if (aux) *aux = r->aux;
  if (info) {
    info->arrival_ns = r->arrival;
    info->admitted_ns = r->admitted;
    info->done_ns = r->done;
    info->tier_used = r->tier_used;
    info->status = r->status;
  }
  const int st = r->status;
  s->table.erase(it);
  ++s->stats.completed;
  delete r;
  return st;
}

int gbs_server_replay(gbs_server* s, int n, const int* func, const int* tier, const int* dim, const gbs_complex* in,
                      const long long* off, const long long* arrival_offset_ns, gbs_complex* out, double* aux,
                      gbs_request_info* info) {
  if (!s) return GBSK_ERR_ARG;
  if (n <= 0) return 0;
  if (!func || !tier || !dim || !in || !off || !arrival_offset_ns || !out) return GBSK_ERR_ARG;
  std::vector<long long> tickets((size_t)n, 0);
  std::vector<Req*> rs;
  rs.reserve(1024);
  int first = 0;
  const long long t0 = now_ns();
  int i = 0;
  // every request already due is queued under one lock; the matrices are
  // borrowed from `in`, which outlives the call
  while (i < n && !first) {
    long long t = now_ns();
    if (t < t0 + arrival_offset_ns[i]) {
      std::this_thread::yield();
      continue;
    }
    int j = i;
    rs.clear();
    while (j < n && t0 + arrival_offset_ns[j] <= t && j - i < 4096) {
      Req* r = nullptr;
      const int rc = make_req(func[j], tier[j], dim[j], in + off[j], t0 + arrival_offset_ns[j], false, &r);
      if (rc) {
        first = rc;
        break;
      }
      rs.push_back(r);
      ++j;
    }
    if (!rs.empty()) {
      const int rc = enqueue_reqs(s, rs.data(), (int)rs.size(), tickets.data() + i);
      if (rc) {
        for (Req* r : rs) delete r;
        first = rc;
        break;
      }
    }
    i += (int)rs.size();
  }
  for (int j = 0; j < i; ++j) {
    const int rc = gbs_server_wait(s, tickets[j], out + j, aux ? aux + j : nullptr, info ? info + j : nullptr);
    if (rc && !first) first = rc;
  }
  return first;
}

int gbs_server_get_stats(gbs_server* s, gbs_server_stats* st) {
  if (!s || !st) return GBSK_ERR_ARG;
  std::lock_guard<std::mutex> g(s->mu);
  *st = s->stats;
  return 0;
}

int gbs_server_destroy(gbs_server* s) {
  if (!s) return 0;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->th.joinable()) s->th.join();  // drains: every slot is free afterwards
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop_workers = true;
  }
  s->cv_slot.notify_all();
  for (auto& sl : s->slots)
    if (sl.worker.joinable()) sl.worker.join();
  for (auto& sl : s->slots)
    if (sl.b) gbs_batch_destroy(sl.b);
  for (auto& kv : s->table) delete kv.second;
  delete s;
  return 0;
}

}  // extern "C"


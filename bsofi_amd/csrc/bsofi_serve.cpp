// Serving loop: continuous batching of BSOF/I requests from DQMC chains.
//
// A chain is a fixed lattice (lx x ly), p, U, dtau, t and a current HS
// field s (p x n, +-1).  e^{-dtau K} of each lattice lives on the device
// once and is shared by every chain on that lattice.  A request flips a set
// of fields of one chain and asks for G = A(s)^{-1}; it ships only its field
// snapshot (p*n bytes) to the device, where the p-cyclic A is assembled in
// place by the generator kernel.
//
// Scheduling.  Requests wait in FIFO order.  Each call to step():
//   1. retires finished slots (HIP event query; no host blocking),
//   2. admits queued requests into every free slot as long as the device
//      memory they need (A, tau, plan workspace) fits the HBM budget,
//      forming one batch per slot (one plan, one launch sequence, matrices
//      of different n and p mixed),
//   3. enqueues each new batch on its slot's stream.
// Slots run concurrently on separate streams, so requests that arrive while
// others are on the GPU start as soon as a slot and memory free up and
// overlap with the work in flight.  Every matrix's arithmetic is independent
// of its batch (per-matrix tasks, no cross-matrix reductions), so a result
// is bit-identical alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "bsofi.h"
#include "bsofi_internal.h"

namespace {

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Lattice {
  int lx, ly;
  double t, dtau;
  double* d_expk = nullptr;
};

struct Chain {
  int lattice;
  int n, p;
  double ep, em;
  std::vector<signed char> fields;  // host copy of the current field
};

enum ReqState { QUEUED = 0, RUNNING = 1, DONE = 2 };

struct Request {
  long long id;
  int chain;
  int n, p;
  std::vector<signed char> fields;  // snapshot after this request's flips
  double t_submit = 0, t_admit = 0, t_done = 0;
  int state = QUEUED;
  int info = 0;
  double* dA = nullptr;
  double* dtau = nullptr;
  signed char* dfields = nullptr;
  size_t result_bytes = 0;  // G and tau: held until release
  size_t work_bytes = 0;    // field snapshot and plan workspace: held while running
  int slot = -1;
};

struct Slot {
  bsofi_ctx_t ctx = nullptr;
  bsofi_plan_t plan = nullptr;
  hipEvent_t done = nullptr;
  std::vector<long long> reqs;
  size_t ws_bytes = 0;     // work bytes of the batch, counted in `used` while it runs
  void* plan_ws = nullptr;  // the plan's workspace, from the server's pool
  size_t plan_ws_bytes = 0;
  bool busy = false;
};

size_t result_bytes(int n, int p) {
  const size_t N = (size_t)n * p;
  return N * N * 8 + ((N * 8 + 255) & ~(size_t)255);
}

size_t work_bytes(int n, int p) {
  const size_t N = (size_t)n * p;
  bsofi_matrix_t m{};
  m.n = n;
  m.p = p;
  m.lda = (int)N;
  m.A = (double*)1;
  m.tau = (double*)1;
  return ((size_t)p * n + 256) + bsofi_plan_workspace_bytes(1, &m, BSOFI_PHASE_ALL);
}

}  // namespace

struct bsofi_server_s {
  int device = 0;
  size_t budget = 0, used = 0;
  // Freed request buffers (G, tau, field snapshots), keyed by exact size:
  // chains have fixed shapes, so steady-state serving reuses them and never
  // calls hipMalloc / hipFree (hipFree synchronises the whole device and
  // would stall every slot in flight).  Counted against the budget.
  std::map<size_t, std::vector<void*>> pool;
  size_t pool_bytes = 0;
  int max_batch = 64;
  std::vector<Lattice> lattices;
  std::vector<Chain> chains;
  std::vector<Slot> slots;
  std::deque<long long> queue;
  std::map<long long, std::unique_ptr<Request>> reqs;
  long long next_id = 1;
  long long completed = 0;
  std::mutex mu;
};

static void* pool_get(bsofi_server_t s, size_t bytes) {
  auto it = s->pool.find(bytes);
  if (it != s->pool.end() && !it->second.empty()) {
    void* p = it->second.back();
    it->second.pop_back();
    s->pool_bytes -= bytes;
    return p;
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  return p;
}

static void pool_put(bsofi_server_t s, void* p, size_t bytes) {
  if (!p) return;
  s->pool[bytes].push_back(p);
  s->pool_bytes += bytes;
}

// Free pooled blocks (largest first) until `extra` more bytes fit.
static void pool_trim(bsofi_server_t s, size_t extra) {
  while (s->pool_bytes > 0 && s->used + s->pool_bytes + extra > s->budget) {
    auto it = std::prev(s->pool.end());
    while (it->second.empty() && it != s->pool.begin()) --it;
    if (it->second.empty()) break;
    (void)hipFree(it->second.back());
    it->second.pop_back();
    s->pool_bytes -= it->first;
  }
}

extern "C" {

int bsofi_server_create(int device, size_t hbm_budget_bytes, int nslots, int max_batch, bsofi_server_t* out) {
  if (!out) return -5;
  *out = nullptr;
  if (nslots < 1 || nslots > 8) return -3;
  if (max_batch < 1) return -4;
  if (bsofi_device_count() <= device || device < 0) return BSOFI_ERR_NO_DEVICE;
  auto* s = new bsofi_server_s();
  s->device = device;
  s->budget = hbm_budget_bytes;
  s->max_batch = max_batch;
  s->slots.resize(nslots);
  for (auto& sl : s->slots) {
    if (bsofi_ctx_create(&sl.ctx, device, nullptr) != 0 || hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) !=
                                                               hipSuccess) {
      bsofi_server_destroy(s);
      return BSOFI_ERR_NO_DEVICE;
    }
    // The slots already overlap one another on their own streams; a second
    // lane per slot would double the streams sharing the device's hardware
    // queues (measured: capacity 242 -> 200 requests/s with 4 slots).
    bsofi_ctx_set_lanes(sl.ctx, 1);
  }
  *out = s;
  return 0;
}

int bsofi_server_destroy(bsofi_server_t s) {
  if (!s) return 0;
  (void)hipSetDevice(s->device);
  for (auto& sl : s->slots) {
    if (sl.ctx) bsofi_ctx_synchronize(sl.ctx);
    if (sl.plan) bsofi_plan_destroy(sl.plan);
    if (sl.plan_ws) (void)hipFree(sl.plan_ws);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.ctx) bsofi_ctx_destroy(sl.ctx);
  }
  for (auto& kv : s->reqs) {
    Request& r = *kv.second;
    if (r.dA) (void)hipFree(r.dA);
    if (r.dtau) (void)hipFree(r.dtau);
    if (r.dfields) (void)hipFree(r.dfields);
  }
  for (auto& kv : s->pool)
    for (void* p : kv.second) (void)hipFree(p);
  for (auto& L : s->lattices)
    if (L.d_expk) (void)hipFree(L.d_expk);
  delete s;
  return 0;
}

int bsofi_server_add_chain(bsofi_server_t s, int lx, int ly, int p, double U, double dtau, double t,
                           unsigned long long seed, int instance, int* chain_id) {
  if (!s || !chain_id) return -1;
  const int n = lx * ly;
  if (bsofi_check_args(n, p, n * p) != 0) return bsofi_check_args(n, p, n * p);
  std::lock_guard<std::mutex> lk(s->mu);
  (void)hipSetDevice(s->device);
  int li = -1;
  for (size_t i = 0; i < s->lattices.size(); ++i) {
    const Lattice& L = s->lattices[i];
    if (L.lx == lx && L.ly == ly && L.t == t && L.dtau == dtau) li = (int)i;
  }
  if (li < 0) {
    Lattice L{lx, ly, t, dtau, nullptr};
    std::vector<double> ek((size_t)n * n);
    bsofi_dqmc_expk(lx, ly, t, dtau, ek.data());
    if (hipMalloc((void**)&L.d_expk, ek.size() * 8) != hipSuccess) return BSOFI_ERR_ALLOC;
    if (hipMemcpy(L.d_expk, ek.data(), ek.size() * 8, hipMemcpyHostToDevice) != hipSuccess) return BSOFI_ERR_NO_DEVICE;
    s->used += ek.size() * 8;
    s->lattices.push_back(L);
    li = (int)s->lattices.size() - 1;
  }
  Chain c;
  c.lattice = li;
  c.n = n;
  c.p = p;
  const double lam = std::acosh(std::exp(U * dtau / 2.0));
  c.ep = std::exp(lam);
  c.em = std::exp(-lam);
  c.fields.resize((size_t)p * n);
  bsofi_dqmc_fields(seed, instance, p, n, c.fields.data());
  s->chains.push_back(std::move(c));
  *chain_id = (int)s->chains.size() - 1;
  return 0;
}

int bsofi_server_submit(bsofi_server_t s, int chain, const int* flips, int nflips, long long* request_id) {
  if (!s || !request_id) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  if (chain < 0 || chain >= (int)s->chains.size()) return -2;
  Chain& c = s->chains[chain];
  const int total = c.n * c.p;
  if (nflips < 0 || (nflips > 0 && !flips)) return -4;
  for (int k = 0; k < nflips; ++k)
    if (flips[k] < 0 || flips[k] >= total) return -3;
  // a request larger than the whole budget could never be admitted and
  // would block the FIFO head for good
  if (result_bytes(c.n, c.p) + work_bytes(c.n, c.p) > s->budget) return BSOFI_ERR_ALLOC;
  for (int k = 0; k < nflips; ++k) c.fields[flips[k]] = (signed char)(-c.fields[flips[k]]);
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->chain = chain;
  r->n = c.n;
  r->p = c.p;
  r->fields = c.fields;
  r->result_bytes = result_bytes(c.n, c.p);
  r->work_bytes = work_bytes(c.n, c.p);
  r->t_submit = now_s();
  *request_id = r->id;
  s->queue.push_back(r->id);
  s->reqs[r->id] = std::move(r);
  return 0;
}

// Returns a slot's device buffers to the pool and forgets its batch (after
// it retired, or after a launch that failed part way).
static void slot_clear(bsofi_server_t s, Slot& sl) {
  if (sl.plan) bsofi_plan_destroy(sl.plan);
  sl.plan = nullptr;
  pool_put(s, sl.plan_ws, sl.plan_ws_bytes);
  sl.plan_ws = nullptr;
  sl.plan_ws_bytes = 0;
  s->used -= sl.ws_bytes;
  sl.ws_bytes = 0;
  sl.reqs.clear();
  sl.busy = false;
}

static int launch_slot(bsofi_server_t s, Slot& sl, std::vector<Request*>& batch) {
  hipStream_t st = (hipStream_t)bsofi_ctx_stream(sl.ctx);
  std::vector<bsofi_matrix_t> mats(batch.size());
  for (size_t i = 0; i < batch.size(); ++i) {
    Request& r = *batch[i];
    const size_t N = (size_t)r.n * r.p;
    r.dA = (double*)pool_get(s, N * N * 8);
    r.dtau = (double*)pool_get(s, N * 8);
    r.dfields = (signed char*)pool_get(s, r.fields.size());
    if (!r.dA || !r.dtau || !r.dfields) return BSOFI_ERR_ALLOC;
    mats[i].A = r.dA;
    mats[i].tau = r.dtau;
    mats[i].n = r.n;
    mats[i].p = r.p;
    mats[i].lda = (int)N;
  }
  // The plan's workspace comes from the pool too, so that it is counted in
  // the budget (the requests' work bytes cover it: a batch's workspace is at
  // most the sum of its matrices' single-matrix workspaces).
  sl.plan_ws_bytes = bsofi_plan_workspace_bytes((int)mats.size(), mats.data(), BSOFI_PHASE_ALL);
  sl.plan_ws = pool_get(s, sl.plan_ws_bytes);
  if (!sl.plan_ws) {
    sl.plan_ws_bytes = 0;
    return BSOFI_ERR_ALLOC;
  }
  // The plan first (its creation waits for the slot's stream, which is idle
  // now), then the inputs are built and the plan runs behind them.
  int rc = bsofi_plan_create(sl.ctx, (int)mats.size(), mats.data(), BSOFI_PHASE_ALL, sl.plan_ws, &sl.plan);
  if (rc) return rc;
  for (size_t i = 0; i < batch.size(); ++i) {
    Request& r = *batch[i];
    const Chain& c = s->chains[r.chain];
    if (hipMemcpyAsync(r.dfields, r.fields.data(), r.fields.size(), hipMemcpyHostToDevice, st) != hipSuccess)
      return BSOFI_ERR_NO_DEVICE;
    rc = bsofi_dqmc_build(sl.ctx, r.dA, r.n * r.p, r.n, r.p, s->lattices[c.lattice].d_expk, r.dfields, c.ep, c.em);
    if (rc) return rc;
  }
  rc = bsofi_plan_execute(sl.plan);
  if (rc) return rc;
  if (hipEventRecord(sl.done, st) != hipSuccess) return BSOFI_ERR_NO_DEVICE;
  sl.busy = true;
  return 0;
}

// A batch whose launch failed: its requests end with info = rc (the caller
// sees them in poll() and releases them), their buffers go back to the pool
// once the slot's stream has drained, and the budget they held is returned.
static void fail_batch(bsofi_server_t s, Slot& sl, std::vector<Request*>& batch, int rc) {
  (void)bsofi_ctx_synchronize(sl.ctx);
  const double t = now_s();
  for (Request* r : batch) {
    const size_t N = (size_t)r->n * r->p;
    pool_put(s, r->dA, N * N * 8);
    pool_put(s, r->dtau, N * 8);
    pool_put(s, r->dfields, r->fields.size());
    r->dA = nullptr;
    r->dtau = nullptr;
    r->dfields = nullptr;
    s->used -= r->result_bytes;
    r->result_bytes = 0;
    r->state = DONE;
    r->info = rc;
    r->t_done = t;
  }
  slot_clear(s, sl);
}

int bsofi_server_step(bsofi_server_t s, int* inflight) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  (void)hipSetDevice(s->device);
  // 1. retire finished slots
  for (auto& sl : s->slots) {
    if (!sl.busy) continue;
    hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) return BSOFI_ERR_NO_DEVICE;
    std::vector<int> info(sl.reqs.size(), 0);
    bsofi_plan_info(sl.plan, info.data());
    const double t = now_s();
    for (size_t i = 0; i < sl.reqs.size(); ++i) {
      Request& r = *s->reqs[sl.reqs[i]];
      r.state = DONE;
      r.info = info[i];
      r.t_done = t;
      pool_put(s, r.dfields, r.fields.size());
      r.dfields = nullptr;
      s->completed++;
    }
    slot_clear(s, sl);  // the batch's work bytes; results stay until release
  }
  // 2./3. admit into free slots within the budget, FIFO
  for (size_t si = 0; si < s->slots.size() && !s->queue.empty(); ++si) {
    Slot& sl = s->slots[si];
    if (sl.busy) continue;
    std::vector<Request*> batch;
    size_t add = 0;
    while (!s->queue.empty() && (int)batch.size() < s->max_batch) {
      Request& r = *s->reqs[s->queue.front()];
      const size_t need = r.result_bytes + r.work_bytes;
      if (s->used + add + need > s->budget) break;
      add += need;
      batch.push_back(&r);
      s->queue.pop_front();
    }
    if (batch.empty()) break;  // head does not fit: wait for memory
    pool_trim(s, add);
    const double t = now_s();
    for (Request* r : batch) {
      r->state = RUNNING;
      r->t_admit = t;
      r->slot = (int)si;
      sl.reqs.push_back(r->id);
    }
    s->used += add;
    sl.ws_bytes = 0;
    for (Request* r : batch) sl.ws_bytes += r->work_bytes;
    int rc = launch_slot(s, sl, batch);
    if (rc) fail_batch(s, sl, batch, rc);
  }
  if (inflight) {
    int k = 0;
    for (auto& sl : s->slots) k += (int)sl.reqs.size();
    *inflight = k;
  }
  return (int)s->queue.size();
}

// Completed requests not yet released: ids and latencies (seconds from
// submit to completion as observed by step()).
int bsofi_server_poll(bsofi_server_t s, long long* ids, double* latency, double* queue_wait, int* info, int max) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  int k = 0;
  for (auto& kv : s->reqs) {
    if (k >= max) break;
    Request& r = *kv.second;
    if (r.state != DONE) continue;
    if (ids) ids[k] = r.id;
    if (latency) latency[k] = r.t_done - r.t_submit;
    if (queue_wait) queue_wait[k] = r.t_admit - r.t_submit;
    if (info) info[k] = r.info;
    ++k;
  }
  return k;
}

// Device pointers of a finished request's G (N x N, ld N) and tau.
int bsofi_server_result(bsofi_server_t s, long long id, double** dG, double** dtau, int* n, int* p) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return -2;
  Request& r = *it->second;
  if (r.state != DONE) return 1;
  if (dG) *dG = r.dA;
  if (dtau) *dtau = r.dtau;
  if (n) *n = r.n;
  if (p) *p = r.p;
  return 0;
}

int bsofi_server_release(bsofi_server_t s, long long id) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return -2;
  Request& r = *it->second;
  if (r.state != DONE) return 1;
  const size_t N = (size_t)r.n * r.p;
  pool_put(s, r.dA, N * N * 8);
  pool_put(s, r.dtau, N * 8);
  r.dA = nullptr;
  r.dtau = nullptr;
  s->used -= r.result_bytes;
  s->reqs.erase(it);
  return 0;
}

int bsofi_server_stats(bsofi_server_t s, long long* completed, int* queued, size_t* used_bytes) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  if (completed) *completed = s->completed;
  if (queued) *queued = (int)s->queue.size();
  if (used_bytes) *used_bytes = s->used;
  return 0;
}

int bsofi_server_request_bytes(int n, int p, size_t* bytes) {
  if (!bytes) return -3;
  if (bsofi_check_args(n, p, n * p)) return bsofi_check_args(n, p, n * p);
  *bytes = result_bytes(n, p) + work_bytes(n, p);
  return 0;
}

}  // extern "C"

// Reference-compatible C and Fortran entry points of libbsofi_amd.
//
// Every function here keeps the name, argument list, array layouts and
// info convention of its SGo-Go/BSOFI counterpart (cited per function) and
// runs the computation on the GPU through the native API of
// bsofi_engine.hip.  Host arrays are staged through device buffers that
// the library keeps between calls, via pinned host buffers and a few copy
// threads; nothing is computed on the CPU.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <sys/time.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "bsofi.h"
#include "bsofi_internal.h"

namespace {

// ---------------------------------------------------------------------------
// Parallel copies between pinned staging buffers and the caller's arrays
// ---------------------------------------------------------------------------
// A small pool of host threads: run(count, fn) calls fn(i) for i < count,
// spread over the workers and the calling thread, and returns when all are
// done.  The caller's arrays are ordinary pageable memory (registering them
// is not available), so host-array transfers go through pinned buffers; the
// host-side copy into or out of them is the part threads speed up.
class CopyPool {
 public:
  explicit CopyPool(int workers) {
    for (int t = 0; t < workers; ++t) th_.emplace_back([this] { worker(); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
  }
  void run(int count, const std::function<void(int)>& fn) {
    {
      std::unique_lock<std::mutex> lk(mu_);
      done_.wait(lk, [&] { return active_ == 0; });  // no worker still inside the previous job
      fn_ = &fn;
      count_ = count;
      next_.store(0);
      ++gen_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [&] { return active_ == 0; });
  }

 private:
  void drain() {
    for (int i = next_.fetch_add(1); i < count_; i = next_.fetch_add(1)) (*fn_)(i);
  }
  void worker() {
    long seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(mu_);
      cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      ++active_;
      lk.unlock();
      drain();
      lk.lock();
      if (--active_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> th_;
  std::mutex mu_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* fn_ = nullptr;
  int count_ = 0;
  std::atomic<int> next_{0};
  int active_ = 0;
  long gen_ = 0;
  bool stop_ = false;
};

// A column-major rectangle of A: columns [c0, c1), rows [r0, r1).
struct Rect {
  int c0, c1, r0, r1;
};

// The non-zero blocks of a p-cyclic input: block column k rows k and k+1
// (the diagonal and subdiagonal blocks), and the corner block A_BLK(0, L-1).
// The rest of the device copy is zero-filled: the reference's calls read
// those blocks as the zeros the p-cyclic contract puts there.
void rects_input(int n, int L, std::vector<Rect>& out) {
  const int N = n * L;
  for (int k = 0; k < L; ++k) out.push_back({k * n, (k + 1) * n, k * n, std::min((k + 2) * n, N)});
  out.push_back({(L - 1) * n, L * n, 0, n});
}

// What BSOF and BSTRI leave in A (V + R, V + X): block column k holds data
// in rows [0, (k+2) n) only -- R or X above, the reflectors in its diagonal
// and subdiagonal blocks -- and the caller's zeros below.
void rects_factor(int n, int L, std::vector<Rect>& out) {
  const int N = n * L;
  for (int k = 0; k < L; ++k) out.push_back({k * n, (k + 1) * n, 0, std::min((k + 2) * n, N)});
}

void rects_full(int N, std::vector<Rect>& out) { out.push_back({0, N, 0, N}); }

// ---------------------------------------------------------------------------
// Device staging for host-array calls
// ---------------------------------------------------------------------------
constexpr size_t kSlotBytes = (size_t)64 << 20;  // two pinned slots of 64 MiB

struct Staging {
  bsofi_ctx_t ctx = nullptr;
  double* dA = nullptr;
  size_t capA = 0;
  double* dtau = nullptr;
  size_t capTau = 0;
  char* ws = nullptr;
  size_t capWs = 0;
  char* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  std::unique_ptr<CopyPool> pool;
  int failed = 0;

  int ensure_ctx() {
    if (ctx) return 0;
    if (failed) return BSOFI_ERR_NO_DEVICE;
    int rc = bsofi_ctx_create(&ctx, 0, nullptr);
    if (rc) {
      failed = 1;
      static int once = 0;
      if (!once++) fprintf(stderr, "bsofi_amd: no usable GPU (code %d); BSOF/I calls return info=%d\n", rc, rc);
    }
    return rc;
  }
  // pinned slots, their events and the copy threads, on first use
  int ensure_pinned() {
    if (pin[0]) return 0;
    for (int s = 0; s < 2; ++s) {
      if (hipHostMalloc((void**)&pin[s], kSlotBytes, hipHostMallocDefault) != hipSuccess ||
          hipEventCreateWithFlags(&ev[s], hipEventDisableTiming) != hipSuccess) {
        release_pinned();
        return BSOFI_ERR_ALLOC;
      }
    }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    pool.reset(new CopyPool((int)std::min(15u, hw - 1)));
    return 0;
  }
  void release_pinned() {
    for (int s = 0; s < 2; ++s) {
      if (pin[s]) (void)hipHostFree(pin[s]);
      if (ev[s]) (void)hipEventDestroy(ev[s]);
      pin[s] = nullptr;
      ev[s] = nullptr;
    }
    pool.reset();
  }
  static int grow(void** p, size_t* cap, size_t need) {
    if (*cap >= need && *p) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (hipMalloc(p, need) != hipSuccess) {
      *p = nullptr;
      return BSOFI_ERR_ALLOC;
    }
    *cap = need;
    return 0;
  }
  void release() {
    if (ctx) (void)bsofi_ctx_synchronize(ctx);
    if (dA) (void)hipFree(dA);
    if (dtau) (void)hipFree(dtau);
    if (ws) (void)hipFree(ws);
    dA = dtau = nullptr;
    ws = nullptr;
    capA = capTau = capWs = 0;
    release_pinned();
    if (ctx) bsofi_ctx_destroy(ctx);
    ctx = nullptr;
  }
};

std::mutex g_mu;
Staging g_stage;

inline int ld_dev(int N) { return ((N + 15) / 16) * 16; }

// Pieces of at most one slot: consecutive columns of one rectangle.
void chunks(const std::vector<Rect>& rects, std::vector<Rect>& out) {
  for (const Rect& r : rects) {
    const size_t col_bytes = (size_t)(r.r1 - r.r0) * 8;
    if (col_bytes == 0) continue;
    const int per = (int)std::max<size_t>(1, kSlotBytes / col_bytes);
    for (int c = r.c0; c < r.c1; c += per) out.push_back({c, std::min(r.c1, c + per), r.r0, r.r1});
  }
}

// Host -> device for the rectangles: the pool packs chunk i into a pinned
// slot while the DMA of chunk i - 1 runs from the other.
int upload(Staging& st, const std::vector<Rect>& rects, const double* A, int lda, double* dA, int ldda) {
  hipStream_t s = (hipStream_t)bsofi_ctx_stream(st.ctx);
  std::vector<Rect> ch;
  chunks(rects, ch);
  for (size_t i = 0; i < ch.size(); ++i) {
    const Rect& r = ch[i];
    const int slot = (int)(i & 1);
    const size_t rows = (size_t)(r.r1 - r.r0);
    if (rows * 8 > kSlotBytes) {  // a column larger than a slot: direct from pageable memory
      if (hipMemcpy2DAsync(dA + (size_t)r.c0 * ldda + r.r0, (size_t)ldda * 8, A + (size_t)r.c0 * lda + r.r0,
                           (size_t)lda * 8, rows * 8, r.c1 - r.c0, hipMemcpyHostToDevice, s) != hipSuccess)
        return BSOFI_ERR_NO_DEVICE;
      continue;
    }
    if (i >= 2 && hipEventSynchronize(st.ev[slot]) != hipSuccess) return BSOFI_ERR_NO_DEVICE;
    double* buf = (double*)st.pin[slot];
    st.pool->run(r.c1 - r.c0, [&](int j) {
      memcpy(buf + (size_t)j * rows, A + (size_t)(r.c0 + j) * lda + r.r0, rows * 8);
    });
    if (hipMemcpy2DAsync(dA + (size_t)r.c0 * ldda + r.r0, (size_t)ldda * 8, buf, rows * 8, rows * 8, r.c1 - r.c0,
                         hipMemcpyHostToDevice, s) != hipSuccess ||
        hipEventRecord(st.ev[slot], s) != hipSuccess)
      return BSOFI_ERR_NO_DEVICE;
  }
  return 0;
}

// Device -> host: the DMA of chunk i into one slot runs while the pool
// unpacks chunk i - 1 from the other into the caller's array.
int download(Staging& st, const std::vector<Rect>& rects, double* A, int lda, const double* dA, int ldda) {
  hipStream_t s = (hipStream_t)bsofi_ctx_stream(st.ctx);
  std::vector<Rect> ch;
  chunks(rects, ch);
  auto unpack = [&](size_t i) -> int {
    const Rect& r = ch[i];
    const size_t rows = (size_t)(r.r1 - r.r0);
    if (rows * 8 > kSlotBytes) return 0;
    const int slot = (int)(i & 1);
    if (hipEventSynchronize(st.ev[slot]) != hipSuccess) return BSOFI_ERR_NO_DEVICE;
    const double* buf = (const double*)st.pin[slot];
    st.pool->run(r.c1 - r.c0, [&](int j) {
      memcpy(A + (size_t)(r.c0 + j) * lda + r.r0, buf + (size_t)j * rows, rows * 8);
    });
    return 0;
  };
  for (size_t i = 0; i < ch.size(); ++i) {
    const Rect& r = ch[i];
    const size_t rows = (size_t)(r.r1 - r.r0);
    if (rows * 8 > kSlotBytes) {
      if (hipMemcpy2DAsync(A + (size_t)r.c0 * lda + r.r0, (size_t)lda * 8, dA + (size_t)r.c0 * ldda + r.r0,
                           (size_t)ldda * 8, rows * 8, r.c1 - r.c0, hipMemcpyDeviceToHost, s) != hipSuccess)
        return BSOFI_ERR_NO_DEVICE;
    } else {
      const int slot = (int)(i & 1);
      // the slot's previous chunk (i - 2) was unpacked at step i - 1
      if (hipMemcpy2DAsync(st.pin[slot], rows * 8, dA + (size_t)r.c0 * ldda + r.r0, (size_t)ldda * 8, rows * 8,
                           r.c1 - r.c0, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipEventRecord(st.ev[slot], s) != hipSuccess)
        return BSOFI_ERR_NO_DEVICE;
    }
    if (i >= 1) {
      int rc = unpack(i - 1);
      if (rc) return rc;
    }
  }
  if (!ch.empty()) {
    int rc = unpack(ch.size() - 1);
    if (rc) return rc;
  }
  return bsofi_ctx_synchronize(st.ctx);
}

// Upload A (and tau), run the phases, download A (and tau).  Only what the
// phases read and write moves: a call that starts with BSOF reads the
// p-cyclic blocks, one that starts later the V + R / V + X region; a call
// that ends with BSOI writes the dense G, one that ends earlier the
// V + R / V + X region.  dwork, when large enough, supplies all device
// memory (A copy, tau, workspace).
int run_host(Staging& st, int n, int L, double* A, int lda, double* tau, int phases, bool tau_in, bool tau_out,
             double* dwork, long long ldwork, int* info) {
  int rc = st.ensure_ctx();
  if (!rc) rc = st.ensure_pinned();
  if (rc) {
    *info = rc;
    return rc;
  }
  const int N = n * L, ldda = ld_dev(N);
  bsofi_matrix_t m{};
  m.n = n;
  m.p = L;
  m.lda = ldda;
  const size_t bytesA = (size_t)ldda * N * 8, bytesT = (((size_t)N * 8) + 255) & ~(size_t)255;
  // plan workspace is computed with the device leading dimension
  m.A = (double*)1;
  m.tau = (double*)1;
  const size_t wsb = bsofi_plan_workspace_bytes(1, &m, phases);
  double* dA;
  double* dtau;
  char* ws;
  const uintptr_t base = dwork ? (((uintptr_t)dwork + 255) & ~(uintptr_t)255) : 0;
  const size_t need = dwork ? (size_t)(base - (uintptr_t)dwork) + ((bytesA + 255) & ~(size_t)255) + bytesT + wsb : 0;
  if (dwork && ldwork > 0 && (size_t)ldwork * 8 >= need) {
    dA = (double*)base;
    dtau = (double*)(base + ((bytesA + 255) & ~(size_t)255));
    ws = (char*)dtau + bytesT;
  } else {
    if ((rc = Staging::grow((void**)&st.dA, &st.capA, bytesA)) ||
        (rc = Staging::grow((void**)&st.dtau, &st.capTau, bytesT)) ||
        (rc = Staging::grow((void**)&st.ws, &st.capWs, std::max<size_t>(wsb, 256)))) {
      *info = rc;
      return rc;
    }
    dA = st.dA;
    dtau = st.dtau;
    ws = st.ws;
  }
  hipStream_t s = (hipStream_t)bsofi_ctx_stream(st.ctx);
  std::vector<Rect> in, out;
  if (phases & BSOFI_PHASE_BSOF) {
    if (hipMemset2DAsync(dA, (size_t)ldda * 8, 0, (size_t)N * 8, N, s) != hipSuccess) {
      *info = BSOFI_ERR_NO_DEVICE;
      return *info;
    }
    rects_input(n, L, in);
  } else {
    rects_factor(n, L, in);
  }
  if (phases & BSOFI_PHASE_BSOI)
    rects_full(N, out);
  else
    rects_factor(n, L, out);
  if ((rc = upload(st, in, A, lda, dA, ldda))) {
    *info = rc;
    return rc;
  }
  if (tau_in && hipMemcpyAsync(dtau, tau, (size_t)N * 8, hipMemcpyHostToDevice, s) != hipSuccess) {
    *info = BSOFI_ERR_NO_DEVICE;
    return *info;
  }
  m.A = dA;
  m.tau = dtau;
  bsofi_plan_t plan = nullptr;
  rc = bsofi_plan_create(st.ctx, 1, &m, phases, ws, &plan);
  if (rc == 0) rc = bsofi_plan_execute(plan);
  int inf = 0;
  if (rc == 0) rc = bsofi_plan_info(plan, &inf);
  bsofi_plan_destroy(plan);
  if (rc) {
    *info = rc;
    return rc;
  }
  if ((rc = download(st, out, A, lda, dA, ldda))) {
    *info = rc;
    return rc;
  }
  if (tau_out) {
    if (hipMemcpyAsync(tau, dtau, (size_t)N * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        bsofi_ctx_synchronize(st.ctx) != 0) {
      *info = BSOFI_ERR_NO_DEVICE;
      return *info;
    }
  }
  *info = inf;
  return inf;
}

// LAPACK-style argument checks shared by the host-array entry points.
int check_host(int n, int L, const double* A, int lda, const double* tau, bool need_tau) {
  if (n < 1) return -1;
  if (L < 2) return -2;
  if (!A) return -3;
  if ((long long)lda < (long long)n * L) return -4;
  if (need_tau && !tau) return -5;
  if (n > BSOFI_MAX_N) return BSOFI_ERR_UNSUPPORTED;
  return 0;
}

int host_lwork(int n, int L) {
  // Host workspace the GPU path needs is one double; the value returned
  // keeps callers written for the reference (which size it from LAPACK's
  // dgeqrf/dormqr queries) allocating a comparable, sufficient array.
  if (n < 1 || L < 2) return 1;
  const long long N = (long long)n * L;
  const long long v = 4LL * n * n + 64LL * std::max<long long>(N, 2LL * n);
  return (int)std::min<long long>(v, INT_MAX);
}

int device_lwork(int n, int L) {
  if (n < 1 || L < 2 || n > BSOFI_MAX_N) return 0;
  const int N = n * L, ldda = ld_dev(N);
  bsofi_matrix_t m{};
  m.n = n;
  m.p = L;
  m.lda = ldda;
  m.A = (double*)1;
  m.tau = (double*)1;
  const size_t bytes = (size_t)ldda * N * 8 + (((size_t)N * 8 + 255) & ~(size_t)255) +
                       bsofi_plan_workspace_bytes(1, &m, BSOFI_PHASE_ALL) + 512;
  const size_t d = (bytes + 7) / 8;
  return d > (size_t)INT_MAX ? INT_MAX : (int)d;
}

// ---------------------------------------------------------------------------
// H-API handle: {magic, n, L, lda, context*} in the caller's int array.
// ---------------------------------------------------------------------------
const int kMagic = 0x42534f46;  // "BSOF"

struct HState {
  Staging st;
};

struct HHandle {
  int magic;
  int n, L, lda;
  HState* state;
};
static_assert(sizeof(HHandle) <= 32, "H-API handle must fit the reference's 8-integer array");

HState* unpack(const int* h) {
  if (!h) return nullptr;
  HHandle hh;
  memcpy(&hh, h, sizeof hh);
  if (hh.magic != kMagic) return nullptr;
  return hh.state;
}

}  // namespace

namespace bsofi {
static double fm_geqrf(double m, double n) {
  return m > n ? n * (n * (0.5 - (1. / 3.) * n + m) + m + 23. / 6.) : m * (m * (-0.5 - (1. / 3.) * m + n) + 2. * n + 23. / 6.);
}
static double fa_geqrf(double m, double n) {
  return m > n ? n * (n * (0.5 - (1. / 3.) * n + m) + 5. / 6.) : m * (m * (-0.5 - (1. / 3.) * m + n) + n + 5. / 6.);
}
static double f_geqrf(double m, double n) { return fm_geqrf(m, n) + fa_geqrf(m, n); }
static double f_ormqr(double m, double n, double k, char side) {
  if (side == 'L')
    return (2. * n * m * k - n * k * k + 2. * n * k) + (2. * n * m * k - n * k * k + n * k);
  return (2. * n * m * k - m * k * k + m * k + n * k - 0.5 * k * k + 0.5 * k) + (2. * n * m * k - m * k * k + m * k);
}
static double f_trtri(double n) { return n * (n * (1. / 6. * n + 0.5) + 1. / 3.) + n * (n * (1. / 6. * n - 0.5) + 1. / 3.); }
static double f_trmm(char side, double m, double n) {
  return side == 'L' ? 0.5 * n * m * (m + 1) + 0.5 * n * m * (m - 1) : 0.5 * m * n * (n + 1) + 0.5 * m * n * (n - 1);
}
static double f_gemm(double m, double n, double k) { return 2. * m * n * k; }
double flops_bsof(int n, int L) {
  return (L - 2) * (2 * f_ormqr(2. * n, n, n, 'L') + f_geqrf(2. * n, n)) + f_geqrf(2. * n, 2. * n);
}
double flops_bstri(int n, int L) {
  return f_trtri(3. * n) + f_trmm('R', (double)n * (L - 3), n) + (L - 3) * (f_trtri(n) + 2 * f_trmm('L', n, n)) +
         (double)(((L - 3) * (L + 2)) / 2) * f_gemm(n, n, n);
}
double flops_bsoi(int n, int L) {
  return (L - 2) * f_ormqr((double)L * n, 2. * n, n, 'R') + f_ormqr((double)L * n, 2. * n, 2. * n, 'R');
}
}  // namespace bsofi

extern "C" {

// ---------------------------------------------------------------------------
// CPU-named API (include/bsofi_cpu.h:24-41), computed on the GPU
// ---------------------------------------------------------------------------
int lapackXbsofiLWork(int n, int L, int lda) {
  (void)lda;
  return host_lwork(n, L);
}

int lapackXbsof(int n, int L, double A[], int lda, double tau[], double work[], int lwork, int* info) {
  if (!info) return -8;
  if (lwork == -1) {
    if (work) work[0] = host_lwork(n, L);
    *info = 0;
    return 0;
  }
  int a = check_host(n, L, A, lda, tau, true);
  if (!a && (!work || lwork < 1)) a = -7;
  if (a) return *info = a;
  std::lock_guard<std::mutex> lk(g_mu);
  return run_host(g_stage, n, L, A, lda, tau, BSOFI_PHASE_BSOF, false, true, nullptr, 0, info);
}

int lapackXbstri(int n, int L, double A[], int lda, int* info) {
  if (!info) return -5;
  int a = check_host(n, L, A, lda, nullptr, false);
  if (a) return *info = a;
  std::lock_guard<std::mutex> lk(g_mu);
  return run_host(g_stage, n, L, A, lda, nullptr, BSOFI_PHASE_BSTRI, false, false, nullptr, 0, info);
}

static int bsoi_host(int n, int L, double A[], int lda, double tau[], double work[], int lwork, int* info) {
  if (!info) return -8;
  if (lwork == -1) {
    if (work) work[0] = host_lwork(n, L);
    *info = 0;
    return 0;
  }
  int a = check_host(n, L, A, lda, tau, true);
  if (!a && (!work || lwork < 1)) a = -7;
  if (a) return *info = a;
  std::lock_guard<std::mutex> lk(g_mu);
  return run_host(g_stage, n, L, A, lda, tau, BSOFI_PHASE_BSOI, true, false, nullptr, 0, info);
}

int lapackXbsoi_ormqr(int n, int L, double A[], int lda, double tau[], double work[], int lwork, int* info) {
  return bsoi_host(n, L, A, lda, tau, work, lwork, info);
}

int lapackXbsoi_orgqr(int n, int L, double A[], int lda, double tau[], double work[], int lwork, int* info) {
  return bsoi_host(n, L, A, lda, tau, work, lwork, info);
}

// Fortran names (src/bsofi_callback.c:24-56)
int lapackxbsofilwork_(int* n, int* L, int* lda) { return lapackXbsofiLWork(*n, *L, *lda); }
int lapackxbsof_(int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork, int* info) {
  return lapackXbsof(*n, *L, A, *lda, tau, work, *lwork, info);
}
int lapackxbstri_(int* n, int* L, double A[], int* lda, int* info) { return lapackXbstri(*n, *L, A, *lda, info); }
int lapackxbsoi_ormqr_(int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork, int* info) {
  return lapackXbsoi_ormqr(*n, *L, A, *lda, tau, work, *lwork, info);
}
int lapackxbsoi_orgqr_(int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork, int* info) {
  return lapackXbsoi_orgqr(*n, *L, A, *lda, tau, work, *lwork, info);
}

// ---------------------------------------------------------------------------
// Hybrid API (include/bsofi_hybrid.h:22-52)
// ---------------------------------------------------------------------------
int hybridXbsofiLWork(int n, int L, int lda, int lworkHostDevice) {
  (void)lda;
  return lworkHostDevice == -2 ? device_lwork(n, L) : host_lwork(n, L);
}

static int hybrid_common(int phases, int n, int L, double A[], int lda, double tau[], double work[], int lwork,
                         double dwork[], int ldwork, int* info) {
  if (!info) return -11;
  if (lwork == -1 || lwork == -2) {
    if (work) work[0] = hybridXbsofiLWork(n, L, lda, lwork);
    *info = 0;
    return 0;
  }
  int a = check_host(n, L, A, lda, tau, true);
  if (!a && (!work || lwork < 1)) a = -8;
  if (a) return *info = a;
  std::lock_guard<std::mutex> lk(g_mu);
  const bool bsoi = phases == BSOFI_PHASE_BSOI;
  return run_host(g_stage, n, L, A, lda, tau, phases, bsoi, !bsoi, dwork, ldwork, info);
}

int hybridXbsoftri(void* handle, int n, int L, double A[], int lda, double tau[], double work[], int lwork,
                   double dwork[], int ldwork, int* info) {
  (void)handle;
  return hybrid_common(BSOFI_PHASE_BSOF | BSOFI_PHASE_BSTRI, n, L, A, lda, tau, work, lwork, dwork, ldwork, info);
}

int hybridXbsoi(void* handle, int n, int L, double A[], int lda, double tau[], double work[], int lwork,
                double dwork[], int ldwork, int* info) {
  (void)handle;
  return hybrid_common(BSOFI_PHASE_BSOI, n, L, A, lda, tau, work, lwork, dwork, ldwork, info);
}

int hybridxbsofilwork_(int* n, int* L, int* lda, int* lworkHostDevice) {
  return hybridXbsofiLWork(*n, *L, *lda, *lworkHostDevice);
}
int hybridxbsoftri_(int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork, double dwork[],
                    int* ldwork, int* info) {
  return hybridXbsoftri(nullptr, *n, *L, A, *lda, tau, work, *lwork, dwork, *ldwork, info);
}
int hybridxbsoi_(int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork, double dwork[],
                 int* ldwork, int* info) {
  return hybridXbsoi(nullptr, *n, *L, A, *lda, tau, work, *lwork, dwork, *ldwork, info);
}

// ---------------------------------------------------------------------------
// H-API (src/bsofi_callback.c:101-180)
// ---------------------------------------------------------------------------
int hybridxinith_(int* handle, int* n, int* L, int* lda) {
  if (!handle || !n || !L || !lda) return -1;
  if (check_host(*n, *L, (const double*)1, *lda, nullptr, false)) return -1;
  HState* s = new HState();
  if (s->st.ensure_ctx() || s->st.ensure_pinned()) {
    s->st.release();
    delete s;
    return -1;
  }
  // size the device copies for (n, L) now, as cb_hybridXinitH sizes dwork
  const int N = (*n) * (*L), ldda = ld_dev(N);
  bsofi_matrix_t m{};
  m.n = *n;
  m.p = *L;
  m.lda = ldda;
  m.A = (double*)1;
  m.tau = (double*)1;
  if (Staging::grow((void**)&s->st.dA, &s->st.capA, (size_t)ldda * N * 8) ||
      Staging::grow((void**)&s->st.dtau, &s->st.capTau, (((size_t)N * 8) + 255) & ~(size_t)255) ||
      Staging::grow((void**)&s->st.ws, &s->st.capWs,
                    std::max<size_t>(bsofi_plan_workspace_bytes(1, &m, BSOFI_PHASE_ALL), 256))) {
    s->st.release();
    delete s;
    return -1;
  }
  HHandle hh{kMagic, *n, *L, *lda, s};
  memcpy(handle, &hh, sizeof hh);
  return 0;
}

int hybridxbsofilworkh_(int* handle, int* n, int* L, int* lda) {
  (void)handle;
  return lapackXbsofiLWork(*n, *L, *lda);
}

int hybridxbsoftrih_(int* handle, int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork,
                     int* info) {
  if (!info) return -9;
  HState* s = unpack(handle);
  if (!s) return *info = BSOFI_ERR_HANDLE;
  int a = check_host(*n, *L, A, *lda, tau, true);
  if (!a && (!work || *lwork < 1)) a = -8;
  if (a) return *info = a;
  return run_host(s->st, *n, *L, A, *lda, tau, BSOFI_PHASE_BSOF | BSOFI_PHASE_BSTRI, false, true, nullptr, 0, info);
}

int hybridxbsoih_(int* handle, int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork,
                  int* info) {
  if (!info) return -9;
  HState* s = unpack(handle);
  if (!s) return *info = BSOFI_ERR_HANDLE;
  int a = check_host(*n, *L, A, *lda, tau, true);
  if (!a && (!work || *lwork < 1)) a = -8;
  if (a) return *info = a;
  // A and tau are read from the caller's arrays, as the reference's
  // cb_hybridXbsoiH reads them: a caller may change them between the calls
  return run_host(s->st, *n, *L, A, *lda, tau, BSOFI_PHASE_BSOI, true, false, nullptr, 0, info);
}

int hybridxfinalizeh_(int* handle) {
  HState* s = unpack(handle);
  if (!s) return BSOFI_ERR_HANDLE;
  s->st.release();
  delete s;
  memset(handle, 0, sizeof(HHandle));
  return 0;
}

// ---------------------------------------------------------------------------
// Timers (src/bsofi_callback.c:189-218)
// ---------------------------------------------------------------------------
void fgetwalltime_(double tv[2]) {
  struct timeval t;
  gettimeofday(&t, nullptr);
  tv[0] = (double)t.tv_sec;
  tv[1] = (double)t.tv_usec;
}

double felapsed_(double t1[2], double t0[2]) { return (t1[0] - t0[0]) + (t1[1] - t0[1]) * 1e-6; }

// ---------------------------------------------------------------------------
// Device-pointer API (include/bsofi_gpu.h:34-62)
// ---------------------------------------------------------------------------
static int run_device(int n, int L, double* dA, int ldda, double* htau, int phases, bool tau_in, bool tau_out,
                      double* dwork, long long ldwork, int* info) {
  int a = check_host(n, L, dA, ldda, (phases & (BSOFI_PHASE_BSOF | BSOFI_PHASE_BSOI)) ? htau : (double*)1, true);
  if (a) return *info = a;
  std::lock_guard<std::mutex> lk(g_mu);
  Staging& st = g_stage;
  int rc = st.ensure_ctx();
  if (rc) return *info = rc;
  const int N = n * L;
  bsofi_matrix_t m{};
  m.n = n;
  m.p = L;
  m.lda = ldda;
  m.A = dA;
  m.tau = (double*)1;
  const size_t wsb = std::max<size_t>(bsofi_plan_workspace_bytes(1, &m, phases), 256);
  const size_t bytesT = (((size_t)N * 8) + 255) & ~(size_t)255;
  if ((rc = Staging::grow((void**)&st.dtau, &st.capTau, bytesT))) return *info = rc;
  char* ws;
  if (dwork && ldwork > 0 && (size_t)ldwork * 8 >= wsb + 256) {
    ws = (char*)(((uintptr_t)dwork + 255) & ~(uintptr_t)255);
  } else {
    if ((rc = Staging::grow((void**)&st.ws, &st.capWs, wsb))) return *info = rc;
    ws = st.ws;
  }
  hipStream_t s = (hipStream_t)bsofi_ctx_stream(st.ctx);
  if (tau_in && hipMemcpyAsync(st.dtau, htau, (size_t)N * 8, hipMemcpyHostToDevice, s) != hipSuccess)
    return *info = BSOFI_ERR_NO_DEVICE;
  m.tau = st.dtau;
  bsofi_plan_t plan = nullptr;
  rc = bsofi_plan_create(st.ctx, 1, &m, phases, ws, &plan);
  if (rc == 0) rc = bsofi_plan_execute(plan);
  int inf = 0;
  if (rc == 0) rc = bsofi_plan_info(plan, &inf);
  bsofi_plan_destroy(plan);
  if (rc == 0 && tau_out) {
    if (hipMemcpyAsync(htau, st.dtau, (size_t)N * 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
        bsofi_ctx_synchronize(st.ctx))
      rc = BSOFI_ERR_NO_DEVICE;
  }
  *info = rc ? rc : inf;
  return *info;
}

static size_t device_ws_doubles(int n, int L, int ldda, int phases) {
  bsofi_matrix_t m{};
  m.n = n;
  m.p = L;
  m.lda = ldda;
  m.A = (double*)1;
  m.tau = (double*)1;
  return (bsofi_plan_workspace_bytes(1, &m, phases) + 512 + 7) / 8;
}

int gpuXbsof(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork, double* dT, int* info) {
  (void)dT;
  if (!info) return -9;
  if (lwork == -1 || lwork == -2) {
    // The T factors are recomputed from (V, tau) by the BSOI phase, so the
    // dT array of the reference carries nothing here; 1 double suffices.
    if (work) work[0] = 1;
    *info = 0;
    return 0;
  }
  return run_device(n, L, dA, ldda, tau, BSOFI_PHASE_BSOF, false, true, nullptr, 0, info);
}

int gpuXbstri_rseq(int n, int L, double dA[], int ldda, double dwork[], int ldwork, int* info) {
  if (!info) return -7;
  return run_device(n, L, dA, ldda, nullptr, BSOFI_PHASE_BSTRI, false, false, dwork, ldwork, info);
}

int gpuXbstri_rbatched(int n, int L, double dA[], int ldda, double dwork[], int ldwork, int* info) {
  return gpuXbstri_rseq(n, L, dA, ldda, dwork, ldwork, info);
}

static int gpu_bsoi(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork, double* dT,
                    double dwork[], int ldwork, int* info) {
  (void)dT;
  if (!info) return -11;
  if (lwork == -1 || lwork == -2) {
    if (work) {
      const size_t d = lwork == -2 ? device_ws_doubles(n, L, ldda, BSOFI_PHASE_BSOI) : 1;
      work[0] = (double)d;
    }
    *info = 0;
    return 0;
  }
  return run_device(n, L, dA, ldda, tau, BSOFI_PHASE_BSOI, true, false, dwork, ldwork, info);
}

int gpuXbsoi_ormqr(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork, double* dT,
                   double dwork[], int ldwork, int* info) {
  return gpu_bsoi(n, L, dA, ldda, tau, work, lwork, dT, dwork, ldwork, info);
}
int gpuXbsoi_orgqr(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork, double* dT,
                   double dwork[], int ldwork, int* info) {
  return gpu_bsoi(n, L, dA, ldda, tau, work, lwork, dT, dwork, ldwork, info);
}
int gpuXbsoi_batched(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork, double* dT,
                     double dwork[], int ldwork, int* info) {
  return gpu_bsoi(n, L, dA, ldda, tau, work, lwork, dT, dwork, ldwork, info);
}

// ---------------------------------------------------------------------------
// Flop counts and DQMC host helpers
// ---------------------------------------------------------------------------
void bsofi_flops(int n, int L, double out[3]) {
  out[0] = bsofi::flops_bsof(n, L);
  out[1] = bsofi::flops_bstri(n, L);
  out[2] = bsofi::flops_bsoi(n, L);
}

// e^{a C} for the adjacency C of a periodic chain of length len (entries
// count both directions, so len = 2 has 2 off the diagonal): a circulant
// with c_d = (1/len) sum_k exp(2a cos(2 pi k / len)) cos(2 pi k d / len).
static void chain_exp(int len, double a, double* c) {
  if (len == 1) {
    c[0] = 1.0;
    return;
  }
  const double pi = 3.14159265358979323846;
  for (int d = 0; d < len; ++d) {
    double s = 0.0;
    for (int k = 0; k < len; ++k) s += std::exp(2.0 * a * std::cos(2.0 * pi * k / len)) * std::cos(2.0 * pi * k * d / len);
    c[d] = s / len;
  }
}

int bsofi_dqmc_expk(int lx, int ly, double t, double dtau, double* expk) {
  if (lx < 1) return -1;
  if (ly < 1) return -2;
  if (!expk) return -5;
  // K = -t (I_y (x) C_x + C_y (x) I_x), site x + lx*y; the two terms
  // commute, so e^{-dtau K} = e^{dtau t C_y} (x) e^{dtau t C_x}.
  double* cx = new double[lx];
  double* cy = new double[ly];
  chain_exp(lx, dtau * t, cx);
  chain_exp(ly, dtau * t, cy);
  const int n = lx * ly;
  for (int yc = 0; yc < ly; ++yc)
    for (int xc = 0; xc < lx; ++xc)
      for (int yr = 0; yr < ly; ++yr)
        for (int xr = 0; xr < lx; ++xr) {
          const int r = xr + lx * yr, c = xc + lx * yc;
          expk[(size_t)c * n + r] = cx[((xr - xc) % lx + lx) % lx] * cy[((yr - yc) % ly + ly) % ly];
        }
  delete[] cx;
  delete[] cy;
  return 0;
}

static inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int bsofi_dqmc_fields(unsigned long long seed, int instance, int p, int n, signed char* fields) {
  if (p < 1) return -3;
  if (n < 1) return -4;
  if (!fields) return -5;
  const uint64_t h0 = splitmix64((uint64_t)seed + 0xD6E8FEB86659FD93ull * (uint64_t)(instance + 1));
  for (int l = 0; l < p; ++l)
    for (int i = 0; i < n; ++i) {
      const uint64_t h = splitmix64(h0 ^ ((uint64_t)l * 0xD1B54A32D192ED03ull + (uint64_t)i));
      fields[(size_t)l * n + i] = (h >> 63) ? 1 : -1;
    }
  return 0;
}

}  // extern "C"

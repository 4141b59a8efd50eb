// Planner, executor and native C API of libbsofi_amd.
//
// A plan is the whole BSOF/I of a batch of device-resident matrices laid
// out as an ordered list of kernel launches.  Each launch reads task and
// work-item tables from one device blob, so matrices of different n and p
// share launches: step s of the BSOF sweep of every matrix goes into the
// same pair of launches, row s of every BSTRI sweep into the same GEMM
// launch, and all BSOI row tiles into one scan launch.  The host/device load
// balancing of the reference (src/bsofi_hybrid_load_balancing.c, KAPPA_* and
// *_SWITCH_* of include/bsofi_params.h) has no counterpart: every matrix is
// device-resident and the plan is the schedule.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <type_traits>
#include <vector>

#include "bsofi.h"
#include "bsofi_internal.h"
#include "bsofi_kernels.hip"

namespace bsofi {

bool hip_ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  static int reported = 0;
  if (reported < 8) {
    fprintf(stderr, "bsofi_amd: %s failed: %s\n", what, hipGetErrorString(e));
    reported++;
  }
  return false;
}

#define HIPCHK(x)                                   \
  do {                                              \
    if (!hip_ok((x), #x)) return BSOFI_ERR_NO_DEVICE; \
  } while (0)

enum Family { F_PANEL = 0, F_REFL = 1, F_GEMM = 2, F_TRTRI = 3, F_EYE = 4, F_SCAN = 5, F_LARFT = 6, F_FILL = 7,
              F_DET = 8, F_FLOW = 9 };

struct Launch {
  int family = 0, variant = 0, grid = 0, ntasks = 0, ldcar = 0;
  int lane = 0;  // stream of a two-lane batch plan: 0 the context's, 1 its side stream
  int spw = 1;   // reflector launches: slabs per workgroup
  bool prefork = false;  // runs on the context's stream before a two-lane plan forks
  size_t task_off = 0, item_off = 0, shmem = 0;
  // panel and reflector launches of one task pass it in the kernel
  // arguments (the kernels then read no task table)
  bool one = false;
  PanelTask pt1{};
  ReflTask rt1{};
  // F_FLOW: blob offsets of the panel, reflector and matrix tables; the
  // zero-filled control region (queue head, stop word, counters, flags);
  // the plan's info words; the wait limit in wall-clock ticks
  size_t off2 = 0, off3 = 0, off4 = 0;
  int* ctl = nullptr;
  long long fill_count = 0;
  int* info = nullptr;
  int count = 0;
  long long wait_ticks = 0;
};

struct Blob {
  std::vector<char> h;
  template <class T>
  size_t push(const std::vector<T>& v) {
    size_t off = (h.size() + 255) & ~(size_t)255;
    h.resize(off + v.size() * sizeof(T));
    if (!v.empty()) memcpy(h.data() + off, v.data(), v.size() * sizeof(T));
    return off;
  }
};

// Per-matrix geometry of the blocked algorithm.
struct Geo {
  int n, p, lda, N, nb, S;
  int panel_variant;  // panel_factor_kernel instance
  int scan_mti, scan_ntw;
  int gemm_tile;  // BSTRI GEMM tile edge: 0 = the launch-size rule (row sweep by strips where it applies), else 32, 64 or 128
  int refl_spw;   // reflector slabs per workgroup: 0 = the launch-size rule, else 1, 2, 4 or 8
};

static int nb_for(int n) { return (2 * n <= 512) ? 32 : 16; }

// Panel variants: NW * 10000 + RPG * 100 + CPW for panel_factor_kernel,
// plus kPanelCyc for panel_cyc_kernel (the 8-wave 32-wide instances).
static constexpr int kPanelCyc = 100000;

// Choices a context can override per block size n (the autotuner's knobs):
// the sub-panel width nb of BSOF / BSOI and the rows per workgroup of the
// BSOI scan (16 * scan_mti).  Zero keeps the default.
struct Tune {
  int nb = 0, scan_mti = 0, gemm_tile = 0, refl_spw = 0;
};

// Sub-panel widths with a panel_factor_kernel instance for 2n rows.
static bool nb_supported(int n, int nb) {
  const int m = 2 * n;
  if (nb == 32) return m <= 512;
  if (nb == 16) return m <= 2048;
  if (nb == 8) return m <= 2048;
  return false;
}

// Scan forms: scan_mti = 1, 2, 4 is bsoi_scan_kernel (16x16x4 MFMA, 4 waves)
// on 16 * scan_mti rows; kScan8W is the same kernel on 16 rows with 8 waves
// (block sizes 512 < n <= 1024: 16 output tiles per wave instead of 32);
// kScan4 is bsoi_scan4_kernel (4x4x4 MFMA) on 32 rows, kScan4D the same
// kernel with its operands loaded two k-steps ahead (three buffers), kScan4S
// that form with two workgroups per CU (n <= 64: a quarter of the LDS
// staging, and registers for two waves per SIMD).
static constexpr int kScan8W = 8;
static constexpr int kScan4 = 9;
static constexpr int kScan4D = 10;
static constexpr int kScan4S = 12;
static bool scan_is4(int mti) { return mti == kScan4 || mti == kScan4D || mti == kScan4S; }
static int scan_rows(int mti) { return scan_is4(mti) ? 32 : (mti == kScan8W ? 16 : 16 * mti); }
static int scan_waves(int mti) { return mti == kScan8W ? 8 : 4; }
static int scan_ntw_for(int n, int mti) {
  const int w = scan_waves(mti);
  return ((2 * n + 15) / 16 + w - 1) / w;
}

static bool scan_variant_ok(int mti, int ntw) {
  // with 8 | n, or 4 | n for kScan4D (scan_ok); ntw <= 8 means n <= 256, the column count of the
  // kernel's LDS staging buffer for X_{r,k}
  if (mti == kScan4S) return ntw >= 1 && ntw <= 2;
  if (scan_is4(mti)) return ntw >= 1 && ntw <= 8;
  if (mti == kScan8W) return ntw >= 9 && ntw <= 16;
  if (mti == 1) return ntw >= 1 && ntw <= 16;
  if (mti == 2) return ntw >= 1 && ntw <= 8;
  if (mti == 4) return ntw >= 1 && ntw <= 4;
  return false;
}

// bsoi_scan4_kernel reads Q_k^T without column masks: 16 | 2n.
// The two-buffer loop runs n / 4 k-steps in pairs (8 | n), the three-buffer
// loop any n / 4 >= 2 (4 | n).
static bool scan_ok(int n, int mti, int ntw) {
  return scan_variant_ok(mti, ntw) && (mti != kScan4 || n % 8 == 0) &&
         (mti == kScan4 || !scan_is4(mti) || (n % 4 == 0 && n >= 8));
}

static Geo geometry(int n, int p, int lda, const Tune* tune = nullptr) {
  Geo g;
  g.n = n;
  g.p = p;
  g.lda = lda;
  g.N = n * p;
  g.nb = (tune && tune->nb) ? tune->nb : nb_for(n);
  g.S = (2 * n + g.nb - 1) / g.nb;
  const int m = 2 * n;
  // panel_factor_kernel<RPG, CPW, NW>: 8 waves of 4 columns for every
  // 32-wide sub-panel of 2n <= 512, so that a batch round of such matrices
  // takes one panel launch (over 512 rows here; the round's largest m picks
  // 128 or 256 where it can); 8 waves of 2 columns for 16-wide sub-panels
  // of 2n > 1024 (32 rows per lane: 4 waves of 4 columns would not fit the
  // registers); else 4 waves of nb / 4 columns over the next power-of-two
  // row count
  const bool p32 = g.nb == 32 && m <= 512;
  const bool p8 = p32 || (g.nb == 16 && m > 1024);
  const int rpg = p32 ? 8 : m <= 128 ? 2 : m <= 256 ? 4 : m <= 512 ? 8 : m <= 1024 ? 16 : 32;
  const int nw = p8 ? 8 : 4;
  g.panel_variant = nw * 10000 + rpg * 100 + g.nb / nw;
  // default scan form: the three-buffer 4x4x4 scan from n = 128 (one
  // workgroup per CU) and up to n = 64 (two per CU), for 4 | n; in between,
  // and for other n, the 16x16x4 form (batches of n = 100 ran faster on it,
  // a single matrix on the 4x4x4 form: the autotuner decides); 8 waves
  // beyond n = 512
  const bool by4 = n % 4 == 0 && n >= 8;
  g.scan_mti = (tune && tune->scan_mti)
                   ? tune->scan_mti
                   : (n <= 256 ? (by4 && n >= 128 ? kScan4D : (by4 && n <= 64 ? kScan4S : 2))
                               : (n <= 512 ? 1 : kScan8W));
  g.scan_ntw = scan_ntw_for(n, g.scan_mti);
  g.gemm_tile = tune ? tune->gemm_tile : 0;
  g.refl_spw = tune ? tune->refl_spw : 0;
  return g;
}

// Workspace layout of one matrix (doubles, each piece 256-byte aligned).
struct WsLayout {
  size_t T, Q, W, Z1, Z2, Y, total;  // byte offsets
};

static size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

static WsLayout ws_layout(const Geo& g, int phases) {
  WsLayout L{};
  size_t off = 0;
  const size_t n2 = (size_t)g.n * g.n;
  auto take = [&](size_t doubles) {
    size_t o = off;
    off += al256(doubles * sizeof(double));
    return o;
  };
  const bool need_T = phases & (BSOFI_PHASE_BSOF | BSOFI_PHASE_BSOI);
  // sized for any sub-panel width nb <= 32: S nb^2 <= 2n nb + nb^2
  L.T = need_T ? take((size_t)(g.p - 1) * (64 * (size_t)g.n + 1024)) : 0;
  L.Q = (phases & BSOFI_PHASE_BSOI) ? take((size_t)(g.p - 1) * 4 * n2) : 0;
  if (phases & BSOFI_PHASE_BSTRI) {
    L.W = take((size_t)(g.p - 1) * n2);
    L.Z1 = take((size_t)std::max(g.p - 2, 1) * n2);
    L.Z2 = take((size_t)std::max(g.p - 2, 1) * n2);
    L.Y = take((size_t)g.p * g.n * 32);
  }
  L.total = off;
  return L;
}

}  // namespace bsofi

using namespace bsofi;

struct bsofi_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipStream_t side = nullptr;  // second lane of batch plans, created on first use
  int lanes = 2;               // lanes of batch plans created on the context
  int schedule = 0;            // 1: BSOF as one device-scheduled launch where supported, 0: launch list
  double wait_s = 4.0;         // wait limit of the device-scheduled launch, seconds
  std::map<int, bsofi::Tune> tune;  // per block size n
  // Device buffers of destroyed plans (task blobs, workspaces), reused by
  // later plans on this context so that steady-state plan churn calls
  // neither hipMalloc nor hipFree (which synchronises the whole device).
  // A buffer is cached only after the context's stream has drained, and
  // reuse is ordered on that same stream.
  std::vector<std::pair<size_t, char*>> cache;
  size_t cache_bytes = 0;
};

namespace bsofi {
static constexpr size_t kCacheMaxEntries = 8;

// Sizes on a ladder of 8 steps per octave, so buffers of similar plans
// share a size and fit each other.
static size_t ladder(size_t bytes) {
  if (bytes <= 4096) return 4096;
  int e = 63 - __builtin_clzll((unsigned long long)bytes);  // 2^e <= bytes
  const size_t step = (size_t)1 << (e - 3);
  return (bytes + step - 1) / step * step;
}

// Smallest cached buffer of at least `bytes` (and at most 2x, so a huge
// buffer does not serve a tiny request), else a new allocation.  On return
// `bytes` is the size of the block, which ctx_free takes back.
static char* ctx_alloc(bsofi_ctx_s* c, size_t& bytes) {
  size_t best = SIZE_MAX;
  int bi = -1;
  for (size_t i = 0; i < c->cache.size(); ++i) {
    const size_t b = c->cache[i].first;
    if (b >= bytes && b <= 2 * bytes + 4096 && b < best) {
      best = b;
      bi = (int)i;
    }
  }
  if (bi >= 0) {
    char* ptr = c->cache[bi].second;
    c->cache_bytes -= best;
    c->cache.erase(c->cache.begin() + bi);
    bytes = best;
    return ptr;
  }
  bytes = ladder(bytes);
  char* ptr = nullptr;
  if (hipMalloc((void**)&ptr, bytes) != hipSuccess) return nullptr;
  return ptr;
}

// `bytes` is the block size ctx_alloc reported.
static void ctx_free(bsofi_ctx_s* c, char* ptr, size_t bytes) {
  if (!ptr) return;
  c->cache.push_back({bytes, ptr});
  c->cache_bytes += bytes;
  while (c->cache.size() > kCacheMaxEntries) {  // drop the oldest
    (void)hipFree(c->cache.front().second);
    c->cache_bytes -= c->cache.front().first;
    c->cache.erase(c->cache.begin());
  }
}

static void ctx_trim(bsofi_ctx_s* c) {
  for (auto& e : c->cache) (void)hipFree(e.second);
  c->cache.clear();
  c->cache_bytes = 0;
}
}  // namespace bsofi

struct bsofi_plan_s {
  bsofi_ctx_s* ctx = nullptr;
  std::vector<Launch> launches;
  Blob blob;
  char* d_blob = nullptr;
  size_t blob_bytes = 0;
  bool own_blob = true;  // false: the blob lives in caller-provided memory
  char* ws = nullptr;
  bool own_ws = false;
  size_t ws_bytes = 0;
  size_t ws_block = 0;  // size of the owned workspace block
  int* d_info = nullptr;
  double* d_det = nullptr;
  bool has_det = false;
  int count = 0;
  bool use_graph = false;
  int lanes = 1;
  hipEvent_t fork = nullptr, join = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int timed_family = -1;
  std::vector<hipEvent_t> ev;
  int ev_used = 0;
};

namespace bsofi {

// ---------------------------------------------------------------------------
// Planner
// ---------------------------------------------------------------------------
struct MatJob {
  Geo g;
  double* A;
  double* tau;
  int* info;
  double* det;  // sign, log|det|
  double *T, *Q, *W, *Z1, *Z2, *Y;
  int *fctr, *fdone;  // device-scheduled BSOF: column counters, sub-panel flags
};

// Sub-panels of one matrix's BSOF sweep (an upper bound over nb >= 8, so
// that workspace sizes do not depend on the tuning).
static size_t flow_subpanels_max(int n, int p) { return (size_t)p * ((2 * (size_t)n + 7) / 8); }

static double* blk(const MatJob& m, int i, int j) {
  return m.A + (size_t)m.g.n * ((size_t)j * m.g.lda + i);
}

static double* Tptr(const MatJob& m, int k, int j) {
  return m.T + ((size_t)k * m.g.S + j) * m.g.nb * m.g.nb;
}

// Workgroup i runs on XCD i mod 8 (each XCD with its own L2): reorders
// `items` of a launch of several tasks so that every item of task t runs on
// XCD t mod 8, in their order within each XCD (a task's workgroups then read
// its shared operands through one L2).
static void deal_tasks_to_xcds(std::vector<Item>& items, size_t ntasks) {
  constexpr int kXcd = 8;
  if (ntasks < 2) return;
  std::vector<std::vector<Item>> q(kXcd);
  for (const Item& it : items) q[it.task % kXcd].push_back(it);
  std::vector<Item> dealt;
  dealt.reserve(items.size());
  std::vector<size_t> head(kXcd, 0);
  while (dealt.size() < items.size())
    for (int x = 0; x < kXcd; ++x)
      if (head[x] < q[x].size()) dealt.push_back(q[x][head[x]++]);
  items.swap(dealt);
}

struct Planner {
  Blob& blob;
  std::vector<Launch>& out;
  int lane = 0;
  size_t cus = 256;  // compute units of the device
  Planner(Blob& b, std::vector<Launch>& o) : blob(b), out(o) {}

  // every launch of the planner goes through here, tagged with its lane
  void push(Launch L) {
    L.lane = lane;
    out.push_back(L);
  }

  template <class T>
  void emit(int family, int variant, const std::vector<T>& tasks, const std::vector<Item>& items, int ldcar = 0,
            size_t shmem = 0) {
    if (tasks.empty() || items.empty()) return;
    Launch L;
    L.family = family;
    L.variant = variant;
    L.task_off = blob.push(tasks);
    L.item_off = blob.push(items);
    L.ntasks = (int)tasks.size();
    L.grid = (int)items.size();
    L.ldcar = ldcar;
    L.shmem = shmem;
    // a panel launch of one task: its one workgroup takes the task from the
    // kernel arguments (reflector launches: emit_refl, which knows spw)
    if constexpr (std::is_same<T, PanelTask>::value) {
      if (tasks.size() == 1 && items.size() == 1 && items[0].task == 0) {
        L.one = true;
        L.pt1 = tasks[0];
      }
    }
    push(L);
  }

  // refl_apply_kernel instance for the largest m: 8 waves with the fewest
  // row tiles each that cover it, RT = 1 .. 4, or 8 beyond m = 512 (variant =
  // 100 + RT, m <= 128 RT); beyond m = 1024 refl_apply_stream_kernel
  // (variant 116), which holds no tile across its phases.  Every RT sums a task's tiles on the same chains
  // in the same order.  Below m = 256 the 8-wave instances replaced 4-wave
  // ones with twice the tiles per wave, half the MFMA chain per wave (the
  // uneven matrices one at a time +1.9 %, batches unchanged); RT = 3 for
  // m <= 384 gave C2 +0.6 %.
  static int refl_variant(const std::vector<ReflTask>& rt) {
    int mm = 0;
    for (auto& q : rt) mm = std::max(mm, q.m);
    return mm <= 512 ? 100 + std::max(1, (mm + 127) / 128) : (mm <= 1024 ? 108 : 116);
  }

  // Reflector tasks of one launch step.  The kernel instance follows from
  // the largest m of the step; every instance sums a task's partial
  // products in the same order, so a task's bits do not depend on the other
  // matrices of the launch.  In the 8-wave instances up to 4 tiles a workgroup
  // takes spw consecutive 16-column slabs of one task (variant 200 + RT): 1 while
  // the slabs fit one round of workgroups on the GPU (each workgroup is
  // latency-bound, so more slabs per workgroup would only lengthen the
  // launch), else as many as it takes to fit one round, at most 8, so that V
  // is loaded once per workgroup and each slab's loads overlap the previous
  // slab's work.
  struct ReflRound {
    std::vector<ReflTask> t;
    std::vector<int> slabs;
    size_t total = 0;
    int spw_set = 0;  // smallest tuned slabs-per-workgroup among the tasks (0: none tuned)
    void add(const ReflTask& q, int n, int tuned_spw = 0) {
      t.push_back(q);
      slabs.push_back(n);
      total += (size_t)n;
      if (tuned_spw > 0) spw_set = spw_set ? std::min(spw_set, tuned_spw) : tuned_spw;
    }
  };
  void emit_refl(const ReflRound& R) {
    int variant = refl_variant(R.t);
    const size_t round = cus * (variant == 101 ? 2 : 1);  // multi-slab workgroups per CU (VGPRs, LDS)
    int spw = variant <= 104 ? (int)std::min<size_t>(8, (R.total + round - 1) / round) : 1;
    if (R.spw_set && variant <= 104) spw = R.spw_set;  // the autotuner's choice for these block sizes
    if (spw > 1) variant += 100;
    std::vector<Item> items;
    for (size_t k = 0; k < R.t.size(); ++k)
      for (int s = 0; s < R.slabs[k]; s += spw) items.push_back({(int)k, s});
    // all workgroups of a task on one XCD, reading its V through one L2
    // (task by task, a task's slabs spread over every XCD and each fetched V
    // from HBM)
    deal_tasks_to_xcds(items, R.t.size());
    const size_t before = out.size();
    emit(F_REFL, variant, R.t, items);
    if (out.size() > before) {
      Launch& L = out.back();
      L.spw = std::max(spw, 1);
      // one task whose workgroup g takes slab g spw: the task goes in the
      // kernel arguments, and the kernel forms the item itself
      bool one = R.t.size() == 1 && variant != 116;
      for (size_t g = 0; one && g < items.size(); ++g)
        one = items[g].task == 0 && items[g].sub == (int)g * L.spw;
      if (one) {
        L.one = true;
        L.rt1 = R.t[0];
      }
    }
  }

  // -- BSOF: src/bsofi_cpu.c:86-118 ---------------------------------------
  void bsof(std::vector<MatJob>& mats) {
    struct Step { int k, j; };
    std::vector<std::vector<Step>> seq(mats.size());
    size_t rounds = 0;
    for (size_t i = 0; i < mats.size(); ++i) {
      const Geo& g = mats[i].g;
      for (int k = 0; k <= g.p - 2; ++k) {
        const int ncols = (k == g.p - 2) ? 2 * g.n : g.n;
        for (int j = 0; j * g.nb < ncols; ++j) seq[i].push_back({k, j});
      }
      rounds = std::max(rounds, seq[i].size());
    }
    for (size_t r = 0; r < rounds; ++r) {
      std::map<int, std::vector<PanelTask>> pt;
      std::map<int, std::vector<Item>> pi;
      ReflRound rr;
      for (size_t i = 0; i < mats.size(); ++i) {
        if (r >= seq[i].size()) continue;
        const MatJob& M = mats[i];
        const Geo& g = M.g;
        const int k = seq[i][r].k, j = seq[i][r].j;
        const int ncols = (k == g.p - 2) ? 2 * g.n : g.n;
        const int off = k * g.n + j * g.nb;  // row == col of the sub-panel corner
        const int m = 2 * g.n - j * g.nb;
        const int nbj = std::min(g.nb, ncols - j * g.nb);
        double* P = M.A + (size_t)off * g.lda + off;
        PanelTask t{P, M.tau + off, Tptr(M, k, j), M.info, g.lda, m, nbj, g.nb};
        const int v = g.panel_variant;
        pi[v].push_back({(int)pt[v].size(), 0});
        pt[v].push_back(t);
        // trailing columns of the panel + target 1 (block column k+1) are
        // contiguous; target 2 is the last block column.
        const int w1 = 2 * g.n - j * g.nb - nbj;
        const int w2 = (k < g.p - 2) ? g.n : 0;
        if (w1 + w2 > 0) {
          ReflTask q;
          q.V = P;
          q.T = Tptr(M, k, j);
          q.C1 = M.A + (size_t)(off + nbj) * g.lda + off;
          q.C2 = w2 ? M.A + (size_t)((g.p - 1) * g.n) * g.lda + off : nullptr;
          q.ldv = g.lda;
          q.ldt = g.nb;
          q.ldc = g.lda;
          q.m = m;
          q.nbj = nbj;
          q.w1 = w1;
          q.w2 = w2;
          q.trans = 1;
          rr.add(q, (w1 + w2 + 15) / 16, g.refl_spw);
        }
      }
      for (auto& kv : pt) emit(F_PANEL, panel_instance(kv.first, kv.second), kv.second, pi[kv.first]);
      emit_refl(rr);
    }
  }

  // -- BSOF as one device-scheduled launch (bsof_flow_kernel) --------------
  // Every matrix on the 32-wide 8-wave panel (2n <= 512), whose body the
  // flow kernel runs.
  static bool flow_ok(const std::vector<MatJob>& mats) {
    for (const MatJob& M : mats)
      if (M.g.nb != 32 || 2 * M.g.n > 512) return false;
    return !mats.empty();
  }

  // Queue order, per matrix, in phases r = 0, 1, ...: the QR of sub-panel r,
  // then the reflector slabs of sub-panel r-1 that sub-panel r+1 does not
  // need, then those of sub-panel r that it does (its lookahead); the
  // matrices' phases interleaved.  Every task's inputs come from tasks
  // earlier in that order.
  void flow_bsof(std::vector<MatJob>& mats, int* ctl, long long fill_count, int* info, int count,
                 long long wait_ticks) {
    struct Sub {
      int off, nbj, pref, rref, slabs, col1, col2, w1, w2;
    };
    std::vector<FlowMat> fm;
    std::vector<PanelTask> pts;
    std::vector<ReflTask> rts;
    std::vector<std::vector<Sub>> subs(mats.size());
    int maxm = 0;
    size_t maxG = 0;
    for (size_t i = 0; i < mats.size(); ++i) {
      MatJob& M = mats[i];
      const Geo& g = M.g;
      fm.push_back({M.fctr, M.fdone, g.n, g.p, g.nb, (g.n + g.nb - 1) / g.nb});
      for (int k = 0; k <= g.p - 2; ++k) {
        const int ncols = (k == g.p - 2) ? 2 * g.n : g.n;
        for (int j = 0; j * g.nb < ncols; ++j) {
          Sub sb{};
          sb.off = k * g.n + j * g.nb;
          const int m = 2 * g.n - j * g.nb;
          sb.nbj = std::min(g.nb, ncols - j * g.nb);
          double* P = M.A + (size_t)sb.off * g.lda + sb.off;
          sb.pref = (int)pts.size();
          pts.push_back(PanelTask{P, M.tau + sb.off, Tptr(M, k, j), M.info, g.lda, m, sb.nbj, g.nb});
          sb.w1 = 2 * g.n - j * g.nb - sb.nbj;
          sb.w2 = (k < g.p - 2) ? g.n : 0;
          sb.col1 = sb.off + sb.nbj;
          sb.col2 = (g.p - 1) * g.n;
          sb.rref = -1;
          if (sb.w1 + sb.w2 > 0) {
            ReflTask q;
            q.V = P;
            q.T = Tptr(M, k, j);
            q.C1 = M.A + (size_t)(sb.off + sb.nbj) * g.lda + sb.off;
            q.C2 = sb.w2 ? M.A + (size_t)((g.p - 1) * g.n) * g.lda + sb.off : nullptr;
            q.ldv = g.lda;
            q.ldt = g.nb;
            q.ldc = g.lda;
            q.m = m;
            q.nbj = sb.nbj;
            q.w1 = sb.w1;
            q.w2 = sb.w2;
            q.trans = 1;
            sb.rref = (int)rts.size();
            rts.push_back(q);
            sb.slabs = (sb.w1 + sb.w2 + 15) / 16;
          }
          maxm = std::max(maxm, m);
          subs[i].push_back(sb);
        }
      }
      maxG = std::max(maxG, subs[i].size());
    }
    // does slab sl of sub-panel s touch a column in [lo, hi)?
    auto touches = [](const Sub& sb, int sl, int lo, int hi) {
      for (int v = 16 * sl; v < 16 * sl + 16 && v < sb.w1 + sb.w2; ++v) {
        const int c = v < sb.w1 ? sb.col1 + v : sb.col2 + (v - sb.w1);
        if (c >= lo && c < hi) return true;
      }
      return false;
    };
    std::vector<FlowTask> ft;
    auto refl_tasks = [&](size_t i, size_t r, bool crit) {
      const std::vector<Sub>& S = subs[i];
      const Sub& sb = S[r];
      if (sb.rref < 0) return;
      const bool has_next = r + 1 < S.size();
      const int lo = has_next ? S[r + 1].off : 0, hi = has_next ? S[r + 1].off + S[r + 1].nbj : 0;
      for (int sl = 0; sl < sb.slabs; ++sl)
        if ((has_next && touches(sb, sl, lo, hi)) == crit)
          ft.push_back(FlowTask{1, (int)i, (int)r, sb.rref, sl, sb.col1, sb.col2, 0});
    };
    for (size_t r = 0; r <= maxG; ++r)
      for (size_t i = 0; i < mats.size(); ++i) {
        const size_t G = subs[i].size();
        if (r < G) ft.push_back(FlowTask{0, (int)i, (int)r, subs[i][r].pref, 0, subs[i][r].off, 0, 0});
        if (r >= 1 && r - 1 < G) refl_tasks(i, r - 1, false);
        if (r < G) refl_tasks(i, r, true);
      }
    Launch L;
    L.family = F_FLOW;
    L.variant = maxm <= 128 ? 21 : (maxm <= 256 ? 42 : 84);  // bsof_flow_kernel<RPG, RT>
    L.task_off = blob.push(ft);
    L.off2 = blob.push(pts);
    L.off3 = blob.push(rts);
    L.off4 = blob.push(fm);
    L.ntasks = (int)ft.size();
    L.grid = (int)std::min<size_t>(ft.size(), cus);
    L.prefork = true;
    L.ctl = ctl;
    L.fill_count = fill_count;
    L.info = info;
    L.count = count;
    L.wait_ticks = wait_ticks;
    push(L);
  }

  // The 8-wave panel instance holds 64 * RPG rows per column; a round takes
  // the smallest RPG (2 .. 8) that covers its largest m, a shorter chain per
  // column.  Rows beyond m enter only as zeros (fma(0, 0, s) = s) and are
  // never stored, so the factors do not depend on the instance.
  static int panel_instance(int variant, const std::vector<PanelTask>& t) {
    if (variant != 80804 && variant != kPanelCyc + 80804) return variant;
    int mm = 0;
    for (const PanelTask& q : t) mm = std::max(mm, q.m);
    return variant - 800 + 100 * std::max(2, (mm + 63) / 64);
  }

  // sign(det A) and log|det A| from R's diagonal and tau, before BSTRI
  // overwrites the diagonal blocks.
  void det(std::vector<MatJob>& mats) {
    std::vector<DetTask> dt;
    std::vector<Item> di;
    for (auto& M : mats) {
      di.push_back({(int)dt.size(), 0});
      dt.push_back({M.A, M.tau, M.det, M.g.lda, M.g.N});
    }
    emit(F_DET, 0, dt, di);
  }

  // -- BSTRI: src/bsofi_cpu.c:146-177, restructured by block rows ---------
  // GEMM tasks of one launch step: 128 x 128 tiles where a product still
  // yields a full round of them on the 256 CUs, 64 x 64 tiles otherwise (at
  // 192 the n = 256 row sweep's largest products took 128-wide tiles and the
  // plain inversion was 0.8 % slower).
  static bool gemm_big(int M, int N) { return (long long)(M / 128) * (N / 128) >= 256; }

  // A task's tile is the launch-size rule's (64, or 128 for big products;
  // 64-tile launches of fewer tiles than CUs become 32-tile launches) unless
  // its matrix's tuning forces 32, 64 or 128.  Every tile size keeps each
  // element's k order, so the choice never changes the bits.
  struct GemmBatch {
    std::vector<GemmTask> t64, t128, f32, f64;
    std::vector<Item> i64, i128, j32, j64;
    void add(GemmTask t, int forced = 0) {
      if (t.M <= 0 || t.N <= 0) return;
      const int tile = forced ? forced : (gemm_big(t.M, t.N) ? 128 : 64);
      auto& tasks = tile == 128 ? t128 : tile == 32 ? f32 : (forced ? f64 : t64);
      auto& items = tile == 128 ? i128 : tile == 32 ? j32 : (forced ? j64 : i64);
      t.tiles_n = (t.N + tile - 1) / tile;
      const int tm = (t.M + tile - 1) / tile;
      const int id = (int)tasks.size();
      tasks.push_back(t);
      for (int s = 0; s < tm * t.tiles_n; ++s) items.push_back({id, s});
    }
  };

  // A launch of fewer 64 x 64 tiles than CUs takes 32 x 32 tiles instead:
  // four times the workgroups, each a quarter of the MFMA chain.  The k order
  // of every element is the same for every tile size, so are the bits.
  void emit_gemm(const GemmBatch& b) {
    if (!b.i64.empty() && b.i64.size() < cus) {
      std::vector<GemmTask> t32 = b.t64;
      std::vector<Item> i32;
      for (size_t id = 0; id < t32.size(); ++id) {
        GemmTask& t = t32[id];
        t.tiles_n = (t.N + 31) / 32;
        const int tm = (t.M + 31) / 32;
        for (int s = 0; s < tm * t.tiles_n; ++s) i32.push_back({(int)id, s});
      }
      emit(F_GEMM, 2, t32, i32);
    } else {
      emit(F_GEMM, 0, b.t64, b.i64);
    }
    emit(F_GEMM, 1, b.t128, b.i128);
    emit(F_GEMM, 2, b.f32, b.j32);
    emit(F_GEMM, 0, b.f64, b.j64);
  }

  static GemmTask gt(const double* A, int lda, const double* B, int ldb, double* C, int ldc, int M, int N, int K,
                     double alpha, int amask, int bmask) {
    GemmTask t{};
    t.A = A;
    t.B = B;
    t.C = C;
    t.D = nullptr;
    t.alpha = alpha;
    t.beta = 0.0;
    t.lda = lda;
    t.ldb = ldb;
    t.ldc = ldc;
    t.ldd = 0;
    t.M = M;
    t.N = N;
    t.K = K;
    t.amask = amask;
    t.bmask = bmask;
    t.dcol0 = INT_MAX;
    t.tiles_n = 1;
    return t;
  }

  // The row sweep of matrices with n <= 256 as column-strip chains
  // (sweep_kernel; masked instances where 32 does not divide n, the last
  // strip partial): each strip of block column b runs its b products in
  // turn, the strips side by side, longest chain first, and the strips of
  // one block column on one XCD (they read the same W_i at the same time),
  // the block columns spread over the XCDs by work.
  // One launch per kernel instance: n <= 64 (4 waves, 4 workgroups per CU),
  // n <= 128 (8 waves, 2 per CU) and n <= 256 (8 waves, 2 row tiles each,
  // 1 per CU).  The n <= 128 instances take a group only while its strips
  // fit about two rounds of residency: a small-n step is short and latency-
  // bound, and many matrices' rounds fill the GPU better (the 512-matrix
  // n = 64 batch ran 0.7 % slower, the uneven batch 2.7 %, through strips;
  // one n = 64, p = 32 matrix 3 % faster).  The strips' elements are the
  // GEMM rounds' to the bit, so which matrices take them never changes a
  // result; a matrix whose tuning forces a GEMM tile keeps the rounds with
  // that tile.  Returns the matrices left for the rounds.
  static int sweep_variant(int n) { return (n % 32 ? 20 : 10) + (n <= 64 ? 1 : (n <= 128 ? 2 : 3)); }
  std::vector<MatJob*> emit_sweep(std::vector<MatJob>& mats) {
    std::vector<MatJob*> rest;
    std::map<int, std::vector<MatJob*>> groups;
    for (auto& M : mats) {
      const Geo& g = M.g;
      if (g.n > 256 || g.gemm_tile != 0)
        rest.push_back(&M);
      else
        groups[sweep_variant(g.n)].push_back(&M);
    }
    for (auto& kv : groups) {
      size_t strips = 0;
      for (MatJob* M : kv.second) strips += (size_t)(M->g.p - 1) * ((M->g.n + 31) / 32);
      const int size_class = kv.first % 10;  // 1: n <= 64, 2: n <= 128, 3: n <= 256
      const size_t per_cu = size_class == 1 ? 4 : 2;
      if (size_class != 3 && strips > 2 * per_cu * cus) {
        rest.insert(rest.end(), kv.second.begin(), kv.second.end());
        continue;
      }
      std::vector<SweepTask> st;
      for (MatJob* M : kv.second)
        for (int b = 1; b < M->g.p; ++b) st.push_back({M->A, M->W, M->Z2, M->g.lda, M->g.n, M->g.p, b});
      std::stable_sort(st.begin(), st.end(), [](const SweepTask& x, const SweepTask& y) { return x.b > y.b; });
      // Block columns to XCDs, longest first, each to the XCD with the least
      // work so far (b n^2 per strip): round-robin dealing left one XCD 70
      // steps of strips for C2's 63-step critical path, this gives 63.
      // Workgroup i runs on XCD i mod 8, so the XCDs' queues are padded to
      // one length with an empty task (b = 0, which returns at once) and
      // then interleaved.
      constexpr int kXcd = 8;
      std::vector<std::vector<Item>> q(kXcd);
      std::vector<double> load(kXcd, 0.0);
      for (size_t id = 0; id < st.size(); ++id) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        const int strips = (st[id].n + 31) / 32;
        for (int c = 0; c < strips; ++c) q[x].push_back({(int)id, c});
        load[x] += (double)st[id].b * strips * st[id].n * st[id].n;
      }
      size_t len = 0;
      for (auto& v : q) len = std::max(len, v.size());
      bool pad = false;
      for (auto& v : q) pad = pad || v.size() < len;
      if (pad) st.push_back({st[0].A, st[0].W, st[0].Z2, st[0].lda, st[0].n, st[0].p, 0});
      std::vector<Item> items;
      for (size_t r = 0; r < len; ++r)
        for (int x = 0; x < kXcd; ++x) items.push_back(r < q[x].size() ? q[x][r] : Item{(int)st.size() - 1, 0});
      emit(F_GEMM, kv.first, st, items);
    }
    return rest;
  }

  void bstri(std::vector<MatJob>& mats) {
    // (1) X_ii = R_ii^{-1} for every diagonal block: 32 x 32 diagonal
    // sub-blocks first, then block columns J = 1.. by two GEMMs each:
    //   Y = R_{0:J,J} X_JJ,   X_{0:J,J} = -X_{0:J,0:J} Y.
    std::vector<TrtriTask> tt;
    int maxJ = 0;
    for (auto& M : mats) {
      const Geo& g = M.g;
      const int nsb = (g.n + 31) / 32;
      maxJ = std::max(maxJ, nsb);
      for (int i = 0; i < g.p; ++i)
        for (int b = 0; b < nsb; ++b) {
          const int o = i * g.n + 32 * b;
          tt.push_back({M.A + (size_t)o * g.lda + o, M.info, g.lda, std::min(32, g.n - 32 * b), o + 1, 0});
        }
    }
    if (!tt.empty()) {
      Launch L;
      L.family = F_TRTRI;
      L.task_off = blob.push(tt);
      L.ntasks = (int)tt.size();
      L.grid = (int)tt.size();
      push(L);
    }
    for (int J = 1; J < maxJ; ++J) {
      GemmBatch b1, b2;
      for (auto& M : mats) {
        const Geo& g = M.g;
        if (32 * J >= g.n) continue;
        const int nbJ = std::min(32, g.n - 32 * J);
        for (int i = 0; i < g.p; ++i) {
          double* Bii = blk(M, i, i);
          double* Yi = M.Y + (size_t)i * g.n * 32;
          b1.add(gt(Bii + (size_t)32 * J * g.lda, g.lda, Bii + (size_t)32 * J * (g.lda + 1), g.lda, Yi, g.n, 32 * J,
                    nbJ, nbJ, 1.0, 0, 1),
                 g.gemm_tile);
          b2.add(gt(Bii, g.lda, Yi, g.n, Bii + (size_t)32 * J * g.lda, g.lda, 32 * J, nbJ, 32 * J, -1.0, 1, 0),
                 g.gemm_tile);
        }
      }
      emit_gemm(b1);
      emit_gemm(b2);
    }
    // (2) W_i = -X_ii R_{i,i+1} (i <= p-2) and Z1_i = R_{i,p-1} X_{p-1,p-1}
    // (i <= p-3), then Z2_i = -X_ii Z1_i: the two TRMMs of
    // src/bsofi_cpu.c:160,163,165 applied to every row at once.
    {
      GemmBatch b1, b2;
      for (auto& M : mats) {
        const Geo& g = M.g;
        const size_t n2 = (size_t)g.n * g.n;
        for (int i = 0; i <= g.p - 2; ++i)
          b1.add(gt(blk(M, i, i), g.lda, blk(M, i, i + 1), g.lda, M.W + i * n2, g.n, g.n, g.n, g.n, -1.0, 1, 0),
                 g.gemm_tile);
        for (int i = 0; i <= g.p - 3; ++i) {
          b1.add(gt(blk(M, i, g.p - 1), g.lda, blk(M, g.p - 1, g.p - 1), g.lda, M.Z1 + i * n2, g.n, g.n, g.n, g.n,
                    1.0, 0, 1),
                 g.gemm_tile);
          b2.add(gt(blk(M, i, i), g.lda, M.Z1 + i * n2, g.n, M.Z2 + i * n2, g.n, g.n, g.n, g.n, -1.0, 1, 0),
                 g.gemm_tile);
        }
      }
      emit_gemm(b1);
      emit_gemm(b2);
    }
    // (3) Row sweep, bottom-up: X_{i,i+1:p-1} = W_i X_{i+1,i+1:p-1}, plus
    // Z2_i on the last block column (src/bsofi_cpu.c:161-167).
    const std::vector<MatJob*> rest = emit_sweep(mats);
    int rounds = 0;
    for (MatJob* M : rest) rounds = std::max(rounds, M->g.p - 1);
    for (int s = 0; s < rounds; ++s) {
      GemmBatch b;
      for (MatJob* Mp : rest) {
        const MatJob& M = *Mp;
        const Geo& g = M.g;
        const int i = g.p - 2 - s;
        if (i < 0) continue;
        GemmTask q = gt(M.W + (size_t)i * g.n * g.n, g.n, blk(M, i + 1, i + 1), g.lda, blk(M, i, i + 1), g.lda, g.n,
                        (g.p - 1 - i) * g.n, g.n, 1.0, 0, 1);
        if (i <= g.p - 3) {
          q.D = M.Z2 + (size_t)i * g.n * g.n;
          q.ldd = g.n;
          q.dcol0 = (g.p - 2 - i) * g.n;
        }
        b.add(q, g.gemm_tile);
      }
      emit_gemm(b);
    }
  }

  // -- BSOI: src/bsofi_cpu.c:283-333 ----------------------------------------
  void bsoi(std::vector<MatJob>& mats, int max_shmem, bool diag_only) {
    // (1) T factors from (V, tau), (2) Q_k = I, (3) Q_k <- B_j Q_k for
    // j = last .. 0 (dorgqr's backward accumulation), (4) row scans.
    std::vector<PanelTask> lt;
    std::vector<Item> li;
    std::vector<EyeTask> et;
    int maxS = 0;
    for (auto& M : mats) {
      const Geo& g = M.g;
      for (int k = 0; k <= g.p - 2; ++k) {
        const int ncols = (k == g.p - 2) ? 2 * g.n : g.n;
        for (int j = 0; j * g.nb < ncols; ++j) {
          const int off = k * g.n + j * g.nb;
          PanelTask t{M.A + (size_t)off * g.lda + off, M.tau + off, Tptr(M, k, j), M.info, g.lda,
                      2 * g.n - j * g.nb, std::min(g.nb, ncols - j * g.nb), g.nb};
          li.push_back({(int)lt.size(), 0});
          lt.push_back(t);
        }
      }
      maxS = std::max(maxS, g.S);
      et.push_back({M.Q, (long long)4 * g.n * g.n, 2 * g.n, g.p - 1});
    }
    emit(F_LARFT, 0, lt, li);
    if (emit_qform(mats)) {
      bsoi_scans(mats, max_shmem, diag_only);
      return;
    }
    if (!et.empty()) {
      Launch L;
      L.family = F_EYE;
      L.task_off = blob.push(et);
      L.ntasks = (int)et.size();
      L.grid = 1024;
      push(L);
    }
    for (int r = 0; r < maxS; ++r) {
      ReflRound rr;
      for (auto& M : mats) {
        const Geo& g = M.g;
        for (int k = 0; k <= g.p - 2; ++k) {
          const int ncols = (k == g.p - 2) ? 2 * g.n : g.n;
          const int S = (ncols + g.nb - 1) / g.nb;
          const int j = S - 1 - r;
          if (j < 0) continue;
          const int off = k * g.n + j * g.nb;
          const int n2 = 2 * g.n;
          ReflTask q;
          q.V = M.A + (size_t)off * g.lda + off;
          q.T = Tptr(M, k, j);
          double* Qk = M.Q + (size_t)k * 4 * g.n * g.n;
          q.C1 = Qk + (size_t)(j * g.nb) * n2 + j * g.nb;
          q.C2 = nullptr;
          q.ldv = g.lda;
          q.ldt = g.nb;
          q.ldc = n2;
          q.m = n2 - j * g.nb;
          q.nbj = std::min(g.nb, ncols - j * g.nb);
          q.w1 = n2 - j * g.nb;
          q.w2 = 0;
          q.trans = 0;
          rr.add(q, (q.w1 + 15) / 16, g.refl_spw);
        }
      }
      emit_refl(rr);
    }
    bsoi_scans(mats, max_shmem, diag_only);
  }

  // Every Q_k of the plan in one launch (qform_kernel), where every matrix
  // has 2n <= 512 and 16 | nb: the identity fill and the S reflector rounds
  // of the launch list, which write and re-read Q_k once per round, with
  // the same bits.  Slabs heaviest first (a slab takes every H_j with
  // j nb <= its first column), and all slabs of one Q_k on one XCD, which
  // reads the V_j through one L2.
  bool emit_qform(const std::vector<MatJob>& mats) {
    int rt = 1;
    for (auto& M : mats) {
      const Geo& g = M.g;
      if (2 * g.n > 512 || (g.nb != 16 && g.nb != 32)) return false;
      rt = std::max(rt, (2 * g.n + 127) / 128);
    }
    std::vector<QformTask> qt;
    struct W { long long work; Item it; };
    std::vector<W> order;
    for (auto& M : mats) {
      const Geo& g = M.g;
      const int n2 = 2 * g.n;
      for (int k = 0; k <= g.p - 2; ++k) {
        const int ncols = (k == g.p - 2) ? n2 : g.n;
        const int S = (ncols + g.nb - 1) / g.nb;
        const int off = k * g.n;
        const int id = (int)qt.size();
        qt.push_back({M.A + (size_t)off * g.lda + off, Tptr(M, k, 0), M.Q + (size_t)k * 4 * g.n * g.n, g.lda, n2,
                      g.nb, S, ncols});
        for (int slab = 0; slab * 16 < n2; ++slab) {
          long long work = 0;
          for (int j = std::min(S - 1, (16 * slab) / g.nb); j >= 0; --j) work += n2 - j * g.nb;
          order.push_back({work, {id, slab}});
        }
      }
    }
    if (qt.empty()) return true;
    std::stable_sort(order.begin(), order.end(), [](const W& a, const W& b) { return a.work > b.work; });
    std::vector<Item> items;
    items.reserve(order.size());
    for (const W& w : order) items.push_back(w.it);
    deal_tasks_to_xcds(items, qt.size());
    emit(F_REFL, 400 + rt, qt, items);
    return true;
  }

  // (4) row scans, grouped by kernel instance; heaviest tiles first
  void bsoi_scans(std::vector<MatJob>& mats, int max_shmem, bool diag_only) {
    std::vector<std::pair<int, int>> keys;
    for (auto& M : mats) keys.push_back({M.g.scan_mti, M.g.scan_ntw});
    std::sort(keys.begin(), keys.end());
    keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
    for (auto key : keys) {
      std::vector<ScanTask> st;
      std::vector<Item> si;
      int maxn = 0;
      struct W { double cost; int task, sub; };
      std::vector<W> order;
      for (auto& M : mats) {
        const Geo& g = M.g;
        if (g.scan_mti != key.first || g.scan_ntw != key.second) continue;
        const int MT = scan_rows(g.scan_mti);
        const int id = (int)st.size();
        st.push_back({M.A, M.Q, (long long)4 * g.n * g.n, g.lda, g.n, g.p, 0, g.N, diag_only ? 1 : 0});
        maxn = std::max(maxn, g.n);
        const int tiles = (g.N + MT - 1) / MT;
        for (int s = 0; s < tiles; ++s) {
          const int rb = (s * MT) / g.n;
          const double cost = (double)g.n * g.n * (2.0 * std::max(0, g.p - 1 - rb) + std::min(rb, g.p - 1));
          order.push_back({cost, id, s});
        }
      }
      // Heaviest row panels first.  When every matrix of the group has the
      // same n and p, that order only interleaves identical matrices, so it
      // goes matrix by matrix instead (each one's block rows heaviest first):
      // the panels an XCD runs at once then belong to one matrix and share
      // its Q_k in L2 (interleaved, 512 matrices of n = 64 re-read each Q_k
      // from HBM about 30 times; C3 +5 %).  Uneven groups keep the cost order,
      // whose longest-first schedule matters more there (C4 -1.4 % otherwise).
      bool uniform = true;
      for (auto& t : st) uniform = uniform && t.n == st[0].n && t.p == st[0].p;
      if (uniform)
        std::stable_sort(order.begin(), order.end(), [](const W& a, const W& b) {
          return a.task != b.task ? a.task < b.task : a.cost > b.cost;
        });
      else
        std::stable_sort(order.begin(), order.end(), [](const W& a, const W& b) { return a.cost > b.cost; });
      // XCD-aware order.  Workgroup i runs on XCD i mod 8, each XCD with its
      // own L2.  Row panels of one block row (one cost) step through k in
      // lockstep and read the same Q_k, so whole block rows go to one XCD.
      // An XCD takes consecutive block rows, as many as it runs at once
      // (their X parts end within a few k of each other, so the panels drift
      // little apart and keep sharing Q_k in L2), and the rounds snake over
      // the XCDs (0..7, then 7..0), so the XCD with the heaviest rows of one
      // round gets the lightest of the next.
      constexpr int kXcd = 8;
      std::vector<std::vector<Item>> q(kXcd);
      const size_t per_round = std::max<size_t>(1, cus / kXcd);
      int xcd = 0, dir = 1;
      size_t filled = 0;
      for (size_t i = 0; i < order.size(); ++i) {
        const bool new_row = i == 0 || order[i].cost != order[i - 1].cost || order[i].task != order[i - 1].task;
        if (new_row && filled >= per_round) {
          filled = 0;
          xcd += dir;
          if (xcd == kXcd || xcd < 0) {
            dir = -dir;
            xcd += dir;
          }
        }
        q[xcd].push_back({order[i].task, order[i].sub});
        ++filled;
      }
      std::vector<size_t> head(kXcd, 0);
      for (size_t i = 0; i < order.size();) {
        for (int x = 0; x < kXcd && i < order.size(); ++x) {
          if (head[x] < q[x].size()) {
            si.push_back(q[x][head[x]++]);
            ++i;
          }
        }
      }
      const int MT = scan_rows(key.first);
      int ldcar = MT + 1;
      if ((size_t)maxn * ldcar * sizeof(double) > (size_t)max_shmem) ldcar = MT;
      emit(F_SCAN, key.first * 100 + key.second, st, si, ldcar, (size_t)maxn * ldcar * sizeof(double));
    }
  }
};

// ---------------------------------------------------------------------------
// Executor
// ---------------------------------------------------------------------------
// scan_mti kScan8W is bsoi_scan_kernel<1, NTW, 8>
template <int MTI, int NTW>
static void launch_scan(const Launch& L, const char* base, hipStream_t s) {
  constexpr int M = MTI == kScan8W ? 1 : MTI, NW = MTI == kScan8W ? 8 : 4;
  hipLaunchKernelGGL((bsoi_scan_kernel<M, NTW, NW>), dim3(L.grid), dim3(64 * NW), L.shmem, s,
                     (const ScanTask*)(base + L.task_off), (const Item*)(base + L.item_off), L.ldcar);
}

template <int MTI, int NTW>
static void set_scan_attr(size_t shmem) {
  constexpr int M = MTI == kScan8W ? 1 : MTI, NW = MTI == kScan8W ? 8 : 4;
  if (shmem > 65536)
    (void)hipFuncSetAttribute((const void*)bsoi_scan_kernel<M, NTW, NW>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
}

template <int W, int D, int XN = 256, int MINB = 1>
static void launch_scan4(const Launch& L, const char* base, hipStream_t s) {
  hipLaunchKernelGGL((bsoi_scan4_kernel<W, 4, D, XN, MINB>), dim3(L.grid), dim3(256), L.shmem, s,
                     (const ScanTask*)(base + L.task_off), (const Item*)(base + L.item_off), L.ldcar);
}

template <int W, int D, int XN = 256, int MINB = 1>
static void set_scan_attr4(size_t shmem) {
  if (shmem > 65536)
    (void)hipFuncSetAttribute((const void*)bsoi_scan4_kernel<W, 4, D, XN, MINB>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem);
}

// scan instances: variant = MTI * 100 + NTW (see scan_variant_ok); the
// 4x4x4 forms are kScan4 * 100 + NTW and kScan4D * 100 + NTW
#define SCAN_C(FN, M, W, ...) \
  case M * 100 + W: FN<M, W>(__VA_ARGS__); break;
#define SCAN4_C(FN, W, ...)                                  \
  case kScan4 * 100 + W: FN##4<W, 2>(__VA_ARGS__); break; \
  case kScan4D * 100 + W: FN##4<W, 3>(__VA_ARGS__); break;
#define SCAN4S_C(FN, W, ...) \
  case kScan4S * 100 + W: FN##4<W, 3, 64, 2>(__VA_ARGS__); break;
#define SCAN_SWITCH(FN, ...)                                                                                 \
  switch (variant) {                                                                                         \
    SCAN4_C(FN, 1, __VA_ARGS__) SCAN4_C(FN, 2, __VA_ARGS__) SCAN4_C(FN, 3, __VA_ARGS__)                      \
    SCAN4_C(FN, 4, __VA_ARGS__) SCAN4_C(FN, 5, __VA_ARGS__) SCAN4_C(FN, 6, __VA_ARGS__)                      \
    SCAN4_C(FN, 7, __VA_ARGS__) SCAN4_C(FN, 8, __VA_ARGS__)                                                  \
    SCAN4S_C(FN, 1, __VA_ARGS__) SCAN4S_C(FN, 2, __VA_ARGS__)                                                \
    SCAN_C(FN, 1, 1, __VA_ARGS__) SCAN_C(FN, 1, 2, __VA_ARGS__) SCAN_C(FN, 1, 3, __VA_ARGS__)                \
    SCAN_C(FN, 1, 4, __VA_ARGS__) SCAN_C(FN, 1, 5, __VA_ARGS__) SCAN_C(FN, 1, 6, __VA_ARGS__)                \
    SCAN_C(FN, 1, 7, __VA_ARGS__) SCAN_C(FN, 1, 8, __VA_ARGS__) SCAN_C(FN, 1, 9, __VA_ARGS__)                \
    SCAN_C(FN, 1, 10, __VA_ARGS__) SCAN_C(FN, 1, 11, __VA_ARGS__) SCAN_C(FN, 1, 12, __VA_ARGS__)             \
    SCAN_C(FN, 1, 13, __VA_ARGS__) SCAN_C(FN, 1, 14, __VA_ARGS__) SCAN_C(FN, 1, 15, __VA_ARGS__)             \
    SCAN_C(FN, 1, 16, __VA_ARGS__)                                                                           \
    SCAN_C(FN, 2, 1, __VA_ARGS__) SCAN_C(FN, 2, 2, __VA_ARGS__) SCAN_C(FN, 2, 3, __VA_ARGS__)                \
    SCAN_C(FN, 2, 4, __VA_ARGS__) SCAN_C(FN, 2, 5, __VA_ARGS__) SCAN_C(FN, 2, 6, __VA_ARGS__)                \
    SCAN_C(FN, 2, 7, __VA_ARGS__) SCAN_C(FN, 2, 8, __VA_ARGS__)                                              \
    SCAN_C(FN, 4, 1, __VA_ARGS__) SCAN_C(FN, 4, 2, __VA_ARGS__) SCAN_C(FN, 4, 3, __VA_ARGS__)                \
    SCAN_C(FN, 4, 4, __VA_ARGS__)                                                                            \
    SCAN_C(FN, 8, 9, __VA_ARGS__) SCAN_C(FN, 8, 10, __VA_ARGS__) SCAN_C(FN, 8, 11, __VA_ARGS__)              \
    SCAN_C(FN, 8, 12, __VA_ARGS__) SCAN_C(FN, 8, 13, __VA_ARGS__) SCAN_C(FN, 8, 14, __VA_ARGS__)             \
    SCAN_C(FN, 8, 15, __VA_ARGS__) SCAN_C(FN, 8, 16, __VA_ARGS__)                                            \
    default: break;                                                                                          \
  }

static void run_launch(const Launch& L, const char* base, hipStream_t s) {
  const int variant = L.variant;
  switch (L.family) {
    case F_PANEL: {
      auto t = (const PanelTask*)(base + L.task_off);
      auto it = (const Item*)(base + L.item_off);
      const PanelTask& o = L.pt1;
      // variant = RPG * 100 + CPW
#define PANEL_CASE(R, C, W)                                                                       \
  case W * 10000 + R * 100 + C:                                                                     \
    hipLaunchKernelGGL((panel_factor_kernel<R, C, W>), dim3(L.grid), dim3(64 * W), 0, s, t, it); \
    break;
#define PANEL_CYC_CASE(R)                                                                                 \
  case kPanelCyc + 80000 + R * 100 + 4:                                                                   \
    if (L.one)                                                                                            \
      hipLaunchKernelGGL((panel_cyc_kernel<R, 4, 8, true>), dim3(L.grid), dim3(512), 0, s, t, it, o);     \
    else                                                                                                  \
      hipLaunchKernelGGL((panel_cyc_kernel<R, 4, 8>), dim3(L.grid), dim3(512), 0, s, t, it, o);           \
    break;
      switch (L.variant) {
        PANEL_CYC_CASE(2) PANEL_CYC_CASE(3) PANEL_CYC_CASE(4) PANEL_CYC_CASE(5) PANEL_CYC_CASE(6)
        PANEL_CYC_CASE(7) PANEL_CYC_CASE(8)
        PANEL_CASE(2, 4, 4) PANEL_CASE(4, 4, 4) PANEL_CASE(8, 4, 4) PANEL_CASE(16, 4, 4)
        PANEL_CASE(2, 2, 4) PANEL_CASE(4, 2, 4) PANEL_CASE(8, 2, 4) PANEL_CASE(16, 2, 4) PANEL_CASE(32, 2, 4)
        PANEL_CASE(2, 4, 8) PANEL_CASE(3, 4, 8) PANEL_CASE(4, 4, 8) PANEL_CASE(5, 4, 8) PANEL_CASE(6, 4, 8)
        PANEL_CASE(7, 4, 8) PANEL_CASE(8, 4, 8) PANEL_CASE(32, 2, 8)
        default: break;
      }
#undef PANEL_CASE
#undef PANEL_CYC_CASE
      break;
    }
    case F_REFL: {
      auto it = (const Item*)(base + L.item_off);
      if (L.variant > 400 && L.variant <= 404) {
        auto qt = (const QformTask*)(base + L.task_off);
        switch (L.variant) {
          case 401: hipLaunchKernelGGL(qform_kernel<1>, dim3(L.grid), dim3(512), 0, s, qt, it); break;
          case 402: hipLaunchKernelGGL(qform_kernel<2>, dim3(L.grid), dim3(512), 0, s, qt, it); break;
          case 403: hipLaunchKernelGGL(qform_kernel<3>, dim3(L.grid), dim3(512), 0, s, qt, it); break;
          default: hipLaunchKernelGGL(qform_kernel<4>, dim3(L.grid), dim3(512), 0, s, qt, it); break;
        }
        break;
      }
      auto t = (const ReflTask*)(base + L.task_off);
      const ReflTask& o = L.rt1;
#define REFL_L(RT, MS)                                                                                       \
  do {                                                                                                       \
    if (L.one)                                                                                               \
      hipLaunchKernelGGL((refl_apply_kernel<RT, 8, MS, true>), dim3(L.grid), dim3(512), 0, s, t, it, L.spw, o); \
    else                                                                                                     \
      hipLaunchKernelGGL((refl_apply_kernel<RT, 8, MS>), dim3(L.grid), dim3(512), 0, s, t, it, L.spw, o);   \
  } while (0)
      switch (L.variant) {
        case 101: REFL_L(1, false); break;
        case 102: REFL_L(2, false); break;
        case 103: REFL_L(3, false); break;
        case 104: REFL_L(4, false); break;
        case 201: REFL_L(1, true); break;
        case 202: REFL_L(2, true); break;
        case 203: REFL_L(3, true); break;
        case 204: REFL_L(4, true); break;
        case 116: hipLaunchKernelGGL(refl_apply_stream_kernel, dim3(L.grid), dim3(512), 0, s, t, it); break;
        default: REFL_L(8, false); break;
      }
#undef REFL_L
      break;
    }
    case F_GEMM:
      if (L.variant >= 11) {
        auto st = (const SweepTask*)(base + L.task_off);
        auto si = (const Item*)(base + L.item_off);
        switch (L.variant) {
          case 11: hipLaunchKernelGGL((sweep_kernel<1, 4, 68, false>), dim3(L.grid), dim3(256), 0, s, st, si); break;
          case 12: hipLaunchKernelGGL((sweep_kernel<1, 8, 132, false>), dim3(L.grid), dim3(512), 0, s, st, si); break;
          case 13: hipLaunchKernelGGL((sweep_kernel<2, 8, 260, false>), dim3(L.grid), dim3(512), 0, s, st, si); break;
          case 21: hipLaunchKernelGGL((sweep_kernel<1, 4, 68, true>), dim3(L.grid), dim3(256), 0, s, st, si); break;
          case 22: hipLaunchKernelGGL((sweep_kernel<1, 8, 132, true>), dim3(L.grid), dim3(512), 0, s, st, si); break;
          default: hipLaunchKernelGGL((sweep_kernel<2, 8, 260, true>), dim3(L.grid), dim3(512), 0, s, st, si); break;
        }
      } else if (L.variant == 2)
        hipLaunchKernelGGL(gemm_kernel<1>, dim3(L.grid), dim3(256), 0, s, (const GemmTask*)(base + L.task_off),
                           (const Item*)(base + L.item_off));
      else if (L.variant == 1)
        hipLaunchKernelGGL(gemm_kernel<4>, dim3(L.grid), dim3(256), 0, s, (const GemmTask*)(base + L.task_off),
                           (const Item*)(base + L.item_off));
      else
        hipLaunchKernelGGL(gemm_kernel<2>, dim3(L.grid), dim3(256), 0, s, (const GemmTask*)(base + L.task_off),
                           (const Item*)(base + L.item_off));
      break;
    case F_TRTRI:
      hipLaunchKernelGGL(trtri_diag_kernel, dim3(L.grid), dim3(64), 0, s, (const TrtriTask*)(base + L.task_off));
      break;
    case F_EYE:
      hipLaunchKernelGGL(eye_kernel, dim3(L.grid), dim3(256), 0, s, (const EyeTask*)(base + L.task_off), L.ntasks);
      break;
    case F_DET:
      hipLaunchKernelGGL(det_kernel, dim3(L.grid), dim3(256), 0, s, (const DetTask*)(base + L.task_off));
      break;
    case F_LARFT:
      hipLaunchKernelGGL(larft_kernel, dim3(L.grid), dim3(256), 0, s, (const PanelTask*)(base + L.task_off),
                         (const Item*)(base + L.item_off));
      break;
    case F_SCAN: {
      // variants outside the instance table are rejected at plan time
      SCAN_SWITCH(launch_scan, L, base, s)
      break;
    }
    case F_FLOW: {
      const unsigned fill_blocks = (unsigned)std::min<long long>((L.fill_count + 255) / 256, 1 << 20);
      hipLaunchKernelGGL(fill_int_kernel, dim3(fill_blocks), dim3(256), 0, s, L.ctl, (int)L.fill_count, 0);
      auto ft = (const FlowTask*)(base + L.task_off);
      auto pt = (const PanelTask*)(base + L.off2);
      auto rt = (const ReflTask*)(base + L.off3);
      auto fm = (const FlowMat*)(base + L.off4);
#define FLOW_CASE(R, T)                                                                                     \
  case R * 10 + T:                                                                                          \
    hipLaunchKernelGGL((bsof_flow_kernel<R, T>), dim3(L.grid), dim3(512), 0, s, ft, L.ntasks, pt, rt, fm, \
                       L.ctl, L.wait_ticks);                                                                \
    break;
      switch (L.variant) {
        FLOW_CASE(2, 1) FLOW_CASE(4, 2) FLOW_CASE(8, 4)
        default: break;
      }
#undef FLOW_CASE
      hipLaunchKernelGGL(flow_check_kernel, dim3((L.count + 255) / 256), dim3(256), 0, s, L.ctl, L.info, L.count,
                         BSOFI_ERR_SCHEDULE);
      break;
    }
    default:
      break;
  }
}

// Whether run_launch has a kernel instance for a launch: a plan whose
// launches it could not issue is refused at creation instead of silently
// skipping work.
static bool known_variant(const Launch& L) {
  switch (L.family) {
    case F_PANEL: {
      static const int ok[] = {40204, 40404, 40804, 41604, 40202, 40402, 40802, 41602, 43202, 80204, 80304,
                               80404, 80504, 80604, 80704, 80804, 83202, kPanelCyc + 80204,
                               kPanelCyc + 80304, kPanelCyc + 80404, kPanelCyc + 80504, kPanelCyc + 80604,
                               kPanelCyc + 80704, kPanelCyc + 80804};
      return std::find(std::begin(ok), std::end(ok), L.variant) != std::end(ok);
    }
    case F_REFL:
      return (L.variant >= 101 && L.variant <= 104) || (L.variant >= 201 && L.variant <= 204) || L.variant == 108 ||
             L.variant == 116 || (L.variant >= 401 && L.variant <= 404);
    case F_GEMM:
      return (L.variant >= 0 && L.variant <= 2) || (L.variant >= 11 && L.variant <= 13) ||
             (L.variant >= 21 && L.variant <= 23);
    case F_FLOW:
      return L.variant == 21 || L.variant == 42 || L.variant == 84;
    case F_SCAN:
      return scan_variant_ok(L.variant / 100, L.variant % 100);
    default:
      return true;
  }
}

}  // namespace bsofi

static const bsofi::Tune* ctx_tune(const bsofi_ctx_s* ctx, int n) {
  auto it = ctx->tune.find(n);
  return it == ctx->tune.end() ? nullptr : &it->second;
}

// ===========================================================================
// C API
// ===========================================================================
extern "C" {

int bsofi_ctx_set_tuning(bsofi_ctx_t ctx, int n, int nb, int scan_mti) {
  if (!ctx || n < 1 || n > BSOFI_MAX_N || nb < 0 || scan_mti < 0) return -1;
  Tune t;
  t.nb = nb;
  t.scan_mti = scan_mti;
  const Geo g = geometry(n, 2, 2 * n, &t);
  if (!nb_supported(n, g.nb) || !scan_ok(n, g.scan_mti, g.scan_ntw)) return BSOFI_ERR_UNSUPPORTED;
  if (nb == 0 && scan_mti == 0)
    ctx->tune.erase(n);
  else
    ctx->tune[n] = t;
  return 0;
}

int bsofi_ctx_set_tuning_ex(bsofi_ctx_t ctx, int n, int nb, int scan_mti, int gemm_tile, int refl_spw) {
  if (!ctx || n < 1 || n > BSOFI_MAX_N || nb < 0 || scan_mti < 0) return -1;
  if (gemm_tile != 0 && gemm_tile != 32 && gemm_tile != 64 && gemm_tile != 128) return BSOFI_ERR_UNSUPPORTED;
  if (refl_spw != 0 && refl_spw != 1 && refl_spw != 2 && refl_spw != 4 && refl_spw != 8) return BSOFI_ERR_UNSUPPORTED;
  Tune t;
  t.nb = nb;
  t.scan_mti = scan_mti;
  t.gemm_tile = gemm_tile;
  t.refl_spw = refl_spw;
  const Geo g = geometry(n, 2, 2 * n, &t);
  if (!nb_supported(n, g.nb) || !scan_ok(n, g.scan_mti, g.scan_ntw)) return BSOFI_ERR_UNSUPPORTED;
  if (nb == 0 && scan_mti == 0 && gemm_tile == 0 && refl_spw == 0)
    ctx->tune.erase(n);
  else
    ctx->tune[n] = t;
  return 0;
}

int bsofi_ctx_get_tuning_ex(bsofi_ctx_t ctx, int n, int* nb, int* scan_mti, int* gemm_tile, int* refl_spw) {
  if (!ctx || !nb || !scan_mti || !gemm_tile || !refl_spw || n < 1) return -1;
  const Geo g = geometry(n, 2, 2 * n, ctx_tune(ctx, n));
  *nb = g.nb;
  *scan_mti = g.scan_mti;
  *gemm_tile = g.gemm_tile;
  *refl_spw = g.refl_spw;
  return 0;
}

int bsofi_ctx_get_tuning(bsofi_ctx_t ctx, int n, int* nb, int* scan_mti) {
  if (!ctx || !nb || !scan_mti || n < 1) return -1;
  const Geo g = geometry(n, 2, 2 * n, ctx_tune(ctx, n));
  *nb = g.nb;
  *scan_mti = g.scan_mti;
  return 0;
}

int bsofi_ctx_clear_tuning(bsofi_ctx_t ctx) {
  if (!ctx) return -1;
  ctx->tune.clear();
  return 0;
}

int bsofi_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

int bsofi_check_args(int n, int L, int lda) {
  if (n < 1) return -1;
  if (L < 2) return -2;
  if ((long long)lda < (long long)n * L) return -4;
  if (n > BSOFI_MAX_N) return BSOFI_ERR_UNSUPPORTED;
  return 0;
}

int bsofi_ctx_create(bsofi_ctx_t* ctx, int device, void* hip_stream) {
  if (!ctx) return -1;
  *ctx = nullptr;
  int count = bsofi_device_count();
  if (count <= 0 || device < 0 || device >= count) return BSOFI_ERR_NO_DEVICE;
  HIPCHK(hipSetDevice(device));
  auto* c = new bsofi_ctx_s();
  c->device = device;
  if (hip_stream) {
    c->stream = (hipStream_t)hip_stream;
  } else {
    // A blocking stream: ordered against the legacy default stream, so data
    // that a caller (torch, hipMemcpy) produced there is visible to our
    // kernels without an explicit synchronisation.  Default priority.
    if (!hip_ok(hipStreamCreate(&c->stream), "hipStreamCreate")) {
      delete c;
      return BSOFI_ERR_NO_DEVICE;
    }
    c->own_stream = true;
  }
  *ctx = c;
  return 0;
}

int bsofi_ctx_destroy(bsofi_ctx_t c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  ctx_trim(c);
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  if (c->side) (void)hipStreamDestroy(c->side);
  delete c;
  return 0;
}

int bsofi_ctx_set_stream(bsofi_ctx_t c, void* s) {
  if (!c) return -1;
  (void)hipSetDevice(c->device);
  // cached buffers were last used on the old stream: let it drain first
  (void)hipStreamSynchronize(c->stream);
  if (c->own_stream) {
    (void)hipStreamDestroy(c->stream);
    c->own_stream = false;
  }
  c->stream = (hipStream_t)s;
  return 0;
}

size_t bsofi_ctx_cached_bytes(bsofi_ctx_t c) { return c ? c->cache_bytes : 0; }

int bsofi_ctx_set_lanes(bsofi_ctx_t c, int lanes) {
  if (!c || (lanes != 1 && lanes != 2)) return -1;
  c->lanes = lanes;
  return 0;
}

int bsofi_ctx_set_schedule(bsofi_ctx_t c, int mode) {
  if (!c || (mode != 0 && mode != 1)) return -1;
  c->schedule = mode;
  return 0;
}

int bsofi_ctx_set_wait_limit(bsofi_ctx_t c, double seconds) {
  if (!c || !(seconds >= 0.0) || seconds > 3600.0) return -1;
  c->wait_s = seconds;
  return 0;
}

int bsofi_ctx_release_cache(bsofi_ctx_t c) {
  if (!c) return -1;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  ctx_trim(c);
  return 0;
}

void* bsofi_ctx_stream(bsofi_ctx_t c) { return c ? (void*)c->stream : nullptr; }

int bsofi_ctx_synchronize(bsofi_ctx_t c) {
  if (!c) return -1;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

size_t bsofi_workspace_bytes(int n, int p) {
  if (bsofi_check_args(n, p, n * p) != 0) return 0;
  return ws_layout(geometry(n, p, n * p), BSOFI_PHASE_ALL).total;
}

// Plan workspace: info words, then (sign, log|det|) per matrix, then (with
// BSOF) the control region of the device-scheduled sweep (stop word and
// queue head in the first 1 KB, then per matrix N column counters and its
// sub-panel flags), then the matrices' own pieces.
static size_t plan_header_bytes(int count) {
  return al256((size_t)count * sizeof(int)) + al256((size_t)count * 2 * sizeof(double));
}

static size_t flow_region_ints(int count, const bsofi_matrix_t* mats, int phases) {
  if (!(phases & BSOFI_PHASE_BSOF) || count == 0) return 0;
  size_t ints = 256;  // stop word at 1, queue head at 16
  for (int i = 0; i < count; ++i)
    ints += (size_t)mats[i].n * mats[i].p + flow_subpanels_max(mats[i].n, mats[i].p);
  return ints;
}

size_t bsofi_plan_workspace_bytes(int count, const bsofi_matrix_t* mats, int phases) {
  size_t total = plan_header_bytes(count) + al256(flow_region_ints(count, mats, phases) * sizeof(int));
  for (int i = 0; i < count; ++i) total += ws_layout(geometry(mats[i].n, mats[i].p, mats[i].lda), phases).total;
  return total;
}

}  // extern "C"

// Plan creation.  blob_dev, when given, is device memory of blob_cap bytes
// for the task tables (else they come from the context's buffer cache); a
// plan whose tables do not fit returns BSOFI_ERR_SCRATCH.  one_lane keeps
// The plan's fork/join events (two-lane plans).
static void plan_free_events(bsofi_plan_s* P) {
  hipEvent_t* ev[] = {&P->fork, &P->join};
  for (hipEvent_t* e : ev)
    if (*e) {
      (void)hipEventDestroy(*e);
      *e = nullptr;
    }
}

// every launch on the context's stream (no side stream or events).
static int plan_create(bsofi_ctx_t ctx, int count, const bsofi_matrix_t* mats, int phases, void* workspace,
                       char* blob_dev, size_t blob_cap, bool one_lane, bsofi_plan_t* out) {
  if (!ctx || !out || count < 0 || (count > 0 && !mats)) return -1;
  *out = nullptr;
  for (int i = 0; i < count; ++i) {
    int a = bsofi_check_args(mats[i].n, mats[i].p, mats[i].lda);
    if (a) return a;
    if (!mats[i].A || !mats[i].tau) return -3;
    const Geo g = geometry(mats[i].n, mats[i].p, mats[i].lda, ctx_tune(ctx, mats[i].n));
    if (!nb_supported(g.n, g.nb) || !scan_ok(g.n, g.scan_mti, g.scan_ntw)) return BSOFI_ERR_UNSUPPORTED;
  }
  HIPCHK(hipSetDevice(ctx->device));
  auto* P = new bsofi_plan_s();
  P->ctx = ctx;
  P->count = count;
  P->ws_bytes = bsofi_plan_workspace_bytes(count, mats, phases);
  if (workspace) {
    P->ws = (char*)workspace;
  } else if (P->ws_bytes) {
    P->ws_block = P->ws_bytes;
    P->ws = ctx_alloc(ctx, P->ws_block);
    if (!P->ws) {
      delete P;
      return BSOFI_ERR_ALLOC;
    }
    P->own_ws = true;
  }
  P->d_info = (int*)P->ws;
  P->d_det = (double*)(P->ws + al256((size_t)count * sizeof(int)));
  P->has_det = (phases & BSOFI_PHASE_BSOF) != 0;
  std::vector<MatJob> jobs(count);
  const size_t flow_ints = flow_region_ints(count, mats, phases);
  int* flow_ctl = (int*)(P->ws + plan_header_bytes(count));
  size_t off = plan_header_bytes(count) + al256(flow_ints * sizeof(int));
  size_t flow_off = 256;
  for (int i = 0; i < count; ++i) {
    MatJob& J = jobs[i];
    J.g = geometry(mats[i].n, mats[i].p, mats[i].lda, ctx_tune(ctx, mats[i].n));
    // the 32-wide 8-wave sub-panels take the cyclic panel (panel_cyc_kernel,
    // the bits of panel_factor_kernel without its per-column barrier)
    if (J.g.panel_variant == 80804) J.g.panel_variant += kPanelCyc;
    J.A = mats[i].A;
    J.tau = mats[i].tau;
    J.info = P->d_info + i;
    J.det = P->d_det + 2 * i;
    const WsLayout L = ws_layout(J.g, phases);
    char* b = P->ws + off;
    J.T = (double*)(b + L.T);
    J.Q = (double*)(b + L.Q);
    J.W = (double*)(b + L.W);
    J.Z1 = (double*)(b + L.Z1);
    J.Z2 = (double*)(b + L.Z2);
    J.Y = (double*)(b + L.Y);
    off += L.total;
    if (flow_ints) {
      J.fctr = flow_ctl + flow_off;
      flow_off += (size_t)J.g.N;
      J.fdone = flow_ctl + flow_off;
      flow_off += flow_subpanels_max(J.g.n, J.g.p);
    }
  }
  int max_shmem = 65536;
  (void)hipDeviceGetAttribute(&max_shmem, hipDeviceAttributeMaxSharedMemoryPerBlock, ctx->device);
  Planner pl(P->blob, P->launches);
  int cus = 256;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && cus > 0)
    pl.cus = (size_t)cus;
  if (count > 0) {
    // info words start at INT_MAX ("no zero pivot seen"); the fill launch
    // addresses the workspace directly (task_off holds no blob offset).
    Launch L;
    L.family = F_FILL;
    L.ntasks = count;
    L.grid = 1;
    P->launches.push_back(L);
  }
  // A batch runs in two lanes: the matrices are split into two groups of
  // about equal work (longest first, each to the lighter lane), and each
  // group's launches go to a stream of its own, so that one lane's
  // latency-bound launches (panel QRs, short reflector steps, the tails of
  // wide launches) run beside the other lane's work.  A matrix's launches
  // stay in one lane and in order, and its kernels and tasks are the ones it
  // gets in any other grouping, so its result is the same bits.
  P->lanes = (count >= 2 && !one_lane) ? ctx->lanes : 1;
  std::vector<int> lane_of(count, 0);
  if (P->lanes == 2) {
    auto work = [&](int i) {
      const double n = jobs[i].g.n, p = jobs[i].g.p;
      return n * n * n * p * (p + 4.0);  // scan ~ n^3 p^2, BSOF and BSTRI ~ n^3 p
    };
    std::vector<int> ord(count);
    for (int i = 0; i < count; ++i) ord[i] = i;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return work(a) > work(b); });
    double load[2] = {0.0, 0.0};
    for (int i : ord) {
      const int l = load[1] < load[0] ? 1 : 0;
      lane_of[i] = l;
      load[l] += work(i);
    }
  }
  // BSOF of every matrix as one device-scheduled launch on the context's
  // stream ahead of the lanes, where every matrix has a flow-kernel panel
  const bool flow = (phases & BSOFI_PHASE_BSOF) && ctx->schedule == 1 && Planner::flow_ok(jobs);
  if (flow) {
    int rate_khz = 100000;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, ctx->device) != hipSuccess ||
        rate_khz <= 0)
      rate_khz = 100000;
    const long long ticks = (long long)(ctx->wait_s * 1e3 * rate_khz);
    pl.flow_bsof(jobs, flow_ctl, (long long)flow_ints, P->d_info, count, ticks);
    pl.det(jobs);
    for (auto& L : P->launches) L.prefork = true;
  }
  for (int l = 0; l < P->lanes; ++l) {
    std::vector<MatJob> grp;
    for (int i = 0; i < count; ++i)
      if (lane_of[i] == l) grp.push_back(jobs[i]);
    if (grp.empty()) continue;
    pl.lane = l;
    if ((phases & BSOFI_PHASE_BSOF) && !flow) {
      pl.bsof(grp);
      pl.det(grp);
    }
    if (phases & BSOFI_PHASE_BSTRI) pl.bstri(grp);
    if (phases & BSOFI_PHASE_BSOI) pl.bsoi(grp, max_shmem, (phases & BSOFI_PHASE_DIAG_ONLY) != 0);
  }
  if (P->lanes == 2) {
    bool ok = ctx->side || hip_ok(hipStreamCreate(&ctx->side), "hipStreamCreate(side)");
    ok = ok && hip_ok(hipEventCreateWithFlags(&P->fork, hipEventDisableTiming), "hipEventCreate");
    ok = ok && hip_ok(hipEventCreateWithFlags(&P->join, hipEventDisableTiming), "hipEventCreate");
    if (!ok) {
      plan_free_events(P);
      if (P->own_ws) ctx_free(ctx, P->ws, P->ws_block);
      delete P;
      return BSOFI_ERR_NO_DEVICE;
    }
  }
  for (auto& L : P->launches) {
    if (!known_variant(L)) {
      fprintf(stderr, "bsofi_amd: no kernel instance for family %d variant %d\n", L.family, L.variant);
      plan_free_events(P);
      if (P->own_ws) ctx_free(ctx, P->ws, P->ws_block);
      delete P;
      return BSOFI_ERR_UNSUPPORTED;
    }
    if (L.family == F_SCAN) {
      const int variant = L.variant;
      SCAN_SWITCH(set_scan_attr, L.shmem)
    }
  }
  P->blob_bytes = std::max<size_t>(P->blob.h.size(), 256);
  if (blob_dev) {
    if (P->blob_bytes > blob_cap) {
      plan_free_events(P);
      if (P->own_ws) ctx_free(ctx, P->ws, P->ws_block);
      delete P;
      return BSOFI_ERR_SCRATCH;
    }
    P->d_blob = blob_dev;
    P->own_blob = false;
  } else {
    P->d_blob = ctx_alloc(ctx, P->blob_bytes);
  }
  if (!P->d_blob) {
    plan_free_events(P);
    if (P->own_ws) ctx_free(ctx, P->ws, P->ws_block);
    delete P;
    return BSOFI_ERR_ALLOC;
  }
  if (!P->blob.h.empty() &&
      !hip_ok(hipMemcpyAsync(P->d_blob, P->blob.h.data(), P->blob.h.size(), hipMemcpyHostToDevice, ctx->stream),
              "hipMemcpyAsync(plan)")) {
    plan_free_events(P);
    if (P->own_blob) ctx_free(ctx, P->d_blob, P->blob_bytes);
    if (P->own_ws) ctx_free(ctx, P->ws, P->ws_block);
    delete P;
    return BSOFI_ERR_NO_DEVICE;
  }
  (void)hipStreamSynchronize(ctx->stream);
  *out = P;
  return 0;
}

extern "C" {

int bsofi_plan_create(bsofi_ctx_t ctx, int count, const bsofi_matrix_t* mats, int phases, void* workspace,
                      bsofi_plan_t* out) {
  return plan_create(ctx, count, mats, phases, workspace, nullptr, 0, false, out);
}

// Issues the plan's launches.  A two-lane plan forks the side stream off the
// context's stream right after the info-word fill (an event the side stream
// waits for) and joins it back at the end, so that work the caller queues on
// the context's stream afterwards sees both lanes done; the two lanes'
// launches are issued alternately, so that neither waits on the host for the
// other's.  Under stream capture the same calls record a graph with two
// branches.
static void run_one(bsofi_plan_s* P, const Launch& L, hipStream_t s) {
  bool timed = (L.family == P->timed_family);
  if (timed) {
    // a launch whose events cannot be created runs untimed
    while ((size_t)P->ev_used + 2 > P->ev.size()) {
      hipEvent_t e = nullptr;
      if (hipEventCreate(&e) != hipSuccess) break;
      P->ev.push_back(e);
    }
    timed = (size_t)P->ev_used + 2 <= P->ev.size();
    if (timed) (void)hipEventRecord(P->ev[P->ev_used], s);
  }
  run_launch(L, P->d_blob, s);
  if (timed) {
    (void)hipEventRecord(P->ev[P->ev_used + 1], s);
    P->ev_used += 2;
  }
}

static void run_all(bsofi_plan_s* P) {
  hipStream_t st[2] = {P->ctx->stream, P->lanes == 2 ? P->ctx->side : P->ctx->stream};
  std::vector<const Launch*> lane[2];
  for (const Launch& L : P->launches) {
    if (L.family == F_FILL) {
      hipLaunchKernelGGL(fill_int_kernel, dim3((L.ntasks + 255) / 256), dim3(256), 0, st[0], P->d_info, L.ntasks,
                         INT_MAX);
      continue;
    }
    if (L.prefork) {
      run_one(P, L, st[0]);
      continue;
    }
    lane[L.lane & 1].push_back(&L);
  }
  if (P->lanes == 2) {
    (void)hipEventRecord(P->fork, st[0]);
    (void)hipStreamWaitEvent(st[1], P->fork, 0);
  }
  std::vector<const Launch*> order;
  order.reserve(lane[0].size() + lane[1].size());
  for (size_t i = 0; i < std::max(lane[0].size(), lane[1].size()); ++i)
    for (int l = 0; l < 2; ++l)
      if (i < lane[l].size()) order.push_back(lane[l][i]);
  for (const Launch* Lp : order) run_one(P, *Lp, st[Lp->lane & 1]);
  if (P->lanes == 2) {
    (void)hipEventRecord(P->join, st[1]);
    (void)hipStreamWaitEvent(st[0], P->join, 0);
  }
}

int bsofi_plan_execute(bsofi_plan_t P) {
  if (!P) return -1;
  HIPCHK(hipSetDevice(P->ctx->device));
  hipStream_t s = P->ctx->stream;
  if (P->use_graph && P->timed_family < 0) {
    if (!P->exec) {
      HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      run_all(P);
      HIPCHK(hipStreamEndCapture(s, &P->graph));
      HIPCHK(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    }
    HIPCHK(hipGraphLaunch(P->exec, s));
  } else {
    run_all(P);
  }
  HIPCHK(hipGetLastError());
  return 0;
}

int bsofi_plan_use_graph(bsofi_plan_t P, int on) {
  if (!P) return -1;
  P->use_graph = on != 0;
  return 0;
}

int bsofi_plan_num_launches(bsofi_plan_t P) { return P ? (int)P->launches.size() : -1; }

int bsofi_plan_info(bsofi_plan_t P, int* info) {
  if (!P || !info) return -1;
  if (P->count == 0) return 0;
  HIPCHK(hipSetDevice(P->ctx->device));
  HIPCHK(hipMemcpyAsync(info, P->d_info, sizeof(int) * P->count, hipMemcpyDeviceToHost, P->ctx->stream));
  HIPCHK(hipStreamSynchronize(P->ctx->stream));
  for (int i = 0; i < P->count; ++i)
    if (info[i] == INT_MAX) info[i] = 0;
  return 0;
}

int bsofi_plan_det(bsofi_plan_t P, double* sign, double* logabs) {
  if (!P || !sign || !logabs) return -1;
  if (!P->has_det) return -2;
  if (P->count == 0) return 0;
  HIPCHK(hipSetDevice(P->ctx->device));
  std::vector<double> h(2 * (size_t)P->count);
  HIPCHK(hipMemcpyAsync(h.data(), P->d_det, h.size() * sizeof(double), hipMemcpyDeviceToHost, P->ctx->stream));
  HIPCHK(hipStreamSynchronize(P->ctx->stream));
  for (int i = 0; i < P->count; ++i) {
    sign[i] = h[2 * i];
    logabs[i] = h[2 * i + 1];
  }
  return 0;
}

int bsofi_det(bsofi_ctx_t ctx, const double* dA, int lda, int n, int p, const double* dtau, double* sign,
              double* logabs) {
  if (!ctx || !dA || !dtau || !sign || !logabs) return -1;
  const int a = bsofi_check_args(n, p, lda);
  if (a) return a;
  HIPCHK(hipSetDevice(ctx->device));
  // the task record and the result share one cached device buffer
  size_t bytes = 256 + sizeof(DetTask);
  char* buf = ctx_alloc(ctx, bytes);
  if (!buf) return BSOFI_ERR_ALLOC;
  DetTask t{dA, dtau, (double*)buf, lda, n * p};
  double h[2] = {0, 0};
  hipError_t e = hipMemcpyAsync(buf + 256, &t, sizeof t, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(det_kernel, dim3(1), dim3(256), 0, ctx->stream, (const DetTask*)(buf + 256));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(h, buf, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx_free(ctx, buf, bytes);
  if (!hip_ok(e, "det")) return BSOFI_ERR_NO_DEVICE;
  *sign = h[0];
  *logabs = h[1];
  return 0;
}

int bsofi_plan_time_family(bsofi_plan_t P, int family) {
  if (!P) return -1;
  P->timed_family = family;
  P->ev_used = 0;
  return 0;
}

int bsofi_plan_family_time(bsofi_plan_t P, double* ms, int* launches) {
  if (!P) return -1;
  HIPCHK(hipSetDevice(P->ctx->device));
  HIPCHK(hipStreamSynchronize(P->ctx->stream));
  double total = 0.0;
  for (int k = 0; k + 1 < P->ev_used; k += 2) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, P->ev[k], P->ev[k + 1]) == hipSuccess) total += t;
  }
  if (ms) *ms = total;
  if (launches) *launches = P->ev_used / 2;
  P->ev_used = 0;
  return 0;
}

int bsofi_plan_destroy(bsofi_plan_t P) {
  if (!P) return 0;
  (void)hipSetDevice(P->ctx->device);
  (void)hipStreamSynchronize(P->ctx->stream);
  if (P->exec) (void)hipGraphExecDestroy(P->exec);
  if (P->graph) (void)hipGraphDestroy(P->graph);
  plan_free_events(P);
  for (auto e : P->ev) (void)hipEventDestroy(e);
  // the stream has drained: the buffers go back to the context for reuse
  if (P->own_blob) ctx_free(P->ctx, P->d_blob, P->blob_bytes);
  if (P->own_ws) ctx_free(P->ctx, P->ws, P->ws_block);
  delete P;
  return 0;
}

int bsofi_invert_batched(bsofi_ctx_t ctx, int count, const bsofi_matrix_t* mats, int phases, int* info) {
  bsofi_plan_t P = nullptr;
  int rc = bsofi_plan_create(ctx, count, mats, phases, nullptr, &P);
  if (rc) return rc;
  rc = bsofi_plan_execute(P);
  if (rc == 0 && info) rc = bsofi_plan_info(P, info);
  bsofi_plan_destroy(P);
  return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The no-allocation entry point's scratch: one file-scope device array, of
// which every GPU the library runs on holds its own copy (at its own
// address) until the process ends.  Host code never passes `scratch` itself
// to a kernel or a copy: that name is only a host-side handle.  The device
// address comes from hipGetSymbolAddress, once per device with that device
// current, and is kept per device.
// ---------------------------------------------------------------------------
static __device__ __align__(256) char scratch[BSOFI_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, char*> g_scratch_ptr;

// Device address of `scratch` on the context's device (which is current).
static char* scratch_address(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) return it->second;
  char* ptr = nullptr;
  if (!hip_ok(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") || !ptr)
    return nullptr;
  g_scratch_ptr[device] = ptr;
  return ptr;
}

extern "C" {

int bsofi_invert_batched_noalloc(bsofi_ctx_t ctx, int count, const bsofi_matrix_t* mats, int phases, int* info) {
  if (!ctx || count < 0 || (count > 0 && !mats)) return -1;
  for (int i = 0; i < count; ++i) {
    const int a = bsofi_check_args(mats[i].n, mats[i].p, mats[i].lda);
    if (a) return a;
  }
  const size_t wsb = al256(bsofi_plan_workspace_bytes(count, mats, phases));
  if (wsb >= (size_t)BSOFI_NOALLOC_SCRATCH_BYTES) {
    fprintf(stderr, "bsofi_amd: bsofi_invert_batched_noalloc needs %zu bytes of workspace, more than its %zu-byte "
                    "scratch; use bsofi_invert_batched\n", wsb, (size_t)BSOFI_NOALLOC_SCRATCH_BYTES);
    return BSOFI_ERR_SCRATCH;
  }
  HIPCHK(hipSetDevice(ctx->device));
  char* base = scratch_address(ctx->device);
  if (!base) return BSOFI_ERR_NO_DEVICE;
  bsofi_plan_t P = nullptr;
  // workspace at the start of the scratch, the task tables behind it
  int rc = plan_create(ctx, count, mats, phases, base, base + wsb, (size_t)BSOFI_NOALLOC_SCRATCH_BYTES - wsb, true,
                       &P);
  if (rc) return rc;
  rc = bsofi_plan_execute(P);
  if (rc == 0 && info) rc = bsofi_plan_info(P, info);
  bsofi_plan_destroy(P);
  return rc;
}

int bsofi_gather_blocks(bsofi_ctx_t ctx, const double* dA, int lda, int n, int p, double* dblocks) {
  if (!ctx) return -1;
  HIPCHK(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(gather_blocks_kernel, dim3(1024), dim3(256), 0, ctx->stream, dA, lda, n, p, dblocks);
  HIPCHK(hipGetLastError());
  return 0;
}

int bsofi_pcyclic_residual(bsofi_ctx_t ctx, const double* dblocks, const double* dG, int ldg, int n, int p,
                           double* resid) {
  if (!ctx || !resid) return -1;
  HIPCHK(hipSetDevice(ctx->device));
  size_t acc_bytes = 256;
  double* acc = (double*)ctx_alloc(ctx, acc_bytes);
  if (!acc) return BSOFI_ERR_ALLOC;
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, ctx->stream, (int*)acc, 6, 0);
  const int N = n * p;
  const int tiles = p * ((n + 63) / 64) * ((N + 63) / 64);
  hipLaunchKernelGGL(pcyclic_residual_kernel, dim3(tiles), dim3(256), 0, ctx->stream, dblocks, dG, ldg, n, p, acc);
  hipLaunchKernelGGL(sumsq_kernel, dim3(1024), dim3(256), 0, ctx->stream, dblocks, n, n, 2 * n * p, acc + 1);
  hipLaunchKernelGGL(sumsq_kernel, dim3(1024), dim3(256), 0, ctx->stream, dG, ldg, N, N, acc + 2);
  double h[3] = {0, 0, 0};
  hipError_t e = hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  ctx_free(ctx, (char*)acc, acc_bytes);
  if (!hip_ok(e, "residual")) return BSOFI_ERR_NO_DEVICE;
  *resid = std::sqrt(h[0]) / (std::sqrt(h[1]) * std::sqrt(h[2]));
  return 0;
}

int bsofi_debug_gemm(bsofi_ctx_t ctx, int M, int N, int K, double alpha, const double* dA, int lda,
                     const double* dB, int ldb, double beta, double* dC, int ldc, int amask, int bmask) {
  if (!ctx) return -1;
  if (M < 0 || N < 0 || K < 0) return -2;
  HIPCHK(hipSetDevice(ctx->device));
  GemmTask t{};
  t.A = dA;
  t.B = dB;
  t.C = dC;
  t.D = nullptr;
  t.alpha = alpha;
  t.beta = beta;
  t.lda = lda;
  t.ldb = ldb;
  t.ldc = ldc;
  t.M = M;
  t.N = N;
  t.K = K;
  t.amask = amask;
  t.bmask = bmask;
  t.dcol0 = INT_MAX;
  // the same kernel choice as the planner's GemmBatch
  const bool big = Planner::gemm_big(M, N);
  const int tile = big ? 128 : 64;
  t.tiles_n = (N + tile - 1) / tile;
  const int tiles = ((M + tile - 1) / tile) * t.tiles_n;
  if (tiles == 0) return 0;
  std::vector<Item> items(tiles);
  for (int s2 = 0; s2 < tiles; ++s2) items[s2] = {0, s2};
  const size_t ib = al256(sizeof(GemmTask));
  size_t bytes = ib + items.size() * sizeof(Item);
  char* d = ctx_alloc(ctx, bytes);
  if (!d) return BSOFI_ERR_ALLOC;
  hipError_t e = hipMemcpyAsync(d, &t, sizeof t, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d + ib, items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) {
    if (big)
      hipLaunchKernelGGL(gemm_kernel<4>, dim3(tiles), dim3(256), 0, ctx->stream, (const GemmTask*)d,
                         (const Item*)(d + ib));
    else
      hipLaunchKernelGGL(gemm_kernel<2>, dim3(tiles), dim3(256), 0, ctx->stream, (const GemmTask*)d,
                         (const Item*)(d + ib));
    e = hipStreamSynchronize(ctx->stream);
  }
  ctx_free(ctx, d, bytes);
  return hip_ok(e, "debug_gemm") ? 0 : BSOFI_ERR_NO_DEVICE;
}

int bsofi_fp64_rate(bsofi_ctx_t ctx, int kind, double* tflops) {
  if (!ctx || !tflops || kind < 0 || kind > 2) return -1;
  HIPCHK(hipSetDevice(ctx->device));
  int cus = 0;
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  const int grid = 2 * cus;
  const int iters = kind == 2 ? 40000 : (kind == 1 ? 40000 : 10000);
  size_t bytes = (size_t)grid * 256 * sizeof(double);
  double* out = (double*)ctx_alloc(ctx, bytes);
  if (!out) return BSOFI_ERR_ALLOC;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float best = 1e30f;
  for (int rep = 0; rep < 3 && e == hipSuccess; ++rep) {
    (void)hipEventRecord(e0, ctx->stream);
    if (kind == 0)
      hipLaunchKernelGGL(fp64_rate_kernel<0>, dim3(grid), dim3(256), 0, ctx->stream, out, iters, 1.0, 0.5);
    else if (kind == 1)
      hipLaunchKernelGGL(fp64_rate_kernel<1>, dim3(grid), dim3(256), 0, ctx->stream, out, iters, 1.0, 0.5);
    else
      hipLaunchKernelGGL(fp64_rate_kernel<2>, dim3(grid), dim3(256), 0, ctx->stream, out, iters, 0.999, 1e-3);
    (void)hipEventRecord(e1, ctx->stream);
    e = hipEventSynchronize(e1);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (rep > 0) best = std::min(best, ms);  // the first launch warms up
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  ctx_free(ctx, (char*)out, bytes);
  if (!hip_ok(e, "fp64_rate")) return BSOFI_ERR_NO_DEVICE;
  const double per_wave = kind == 0 ? 8.0 * 2048 : (kind == 1 ? 8.0 * 512 : 8.0 * 64 * 2);
  *tflops = (double)grid * 4 * iters * per_wave / (best * 1e-3) / 1e12;
  return 0;
}

int bsofi_dqmc_build(bsofi_ctx_t ctx, double* dA, int lda, int n, int p, const double* dexpk,
                     const signed char* dfields, double ep, double em) {
  if (!ctx) return -1;
  HIPCHK(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(dqmc_build_kernel, dim3(4096), dim3(256), 0, ctx->stream, dA, lda, n, p, dexpk, dfields, ep,
                     em);
  HIPCHK(hipGetLastError());
  return 0;
}

int bsofi_memcpy(void* dst, const void* src, size_t bytes) {
  HIPCHK(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
  return 0;
}

int bsofi_dqmc_flip(bsofi_ctx_t ctx, signed char* dfields, const int* dflips, int count) {
  if (!ctx) return -1;
  if (count <= 0) return 0;
  HIPCHK(hipSetDevice(ctx->device));
  hipLaunchKernelGGL(hs_flip_kernel, dim3((count + 255) / 256), dim3(256), 0, ctx->stream, dfields, dflips, count);
  HIPCHK(hipGetLastError());
  return 0;
}

}  // extern "C"

// Host-side checks of libbsofi_amd, linked against the stand-in runtime of
// fake_hip.cpp and built with AddressSanitizer / UBSan: the planner's task
// tables for many shapes and both schedules, the C ABI (argument errors,
// staging, the handle API including a double finalize), the no-allocation
// entry point, and the serving loop's admission under an HBM budget,
// release, oversized requests and a failing device allocation.  Exits
// non-zero on the first failed check; the sanitizers abort on any memory or
// undefined-behaviour error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "bsofi.h"

extern "C" void fake_hip_fail_after(long k);
extern "C" long fake_hip_launches();
extern "C" long fake_hip_live_allocs();
extern "C" int hip_malloc_for_tests(void** ptr, size_t bytes);
extern "C" int hip_free_for_tests(void* ptr);

#define CHECK(c)                                                            \
  do {                                                                      \
    if (!(c)) {                                                             \
      fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
      exit(1);                                                              \
    }                                                                       \
  } while (0)

static double* dev_doubles(size_t count) {
  double* p = nullptr;  // device memory of the stand-in runtime is host memory
  CHECK(hip_malloc_for_tests((void**)&p, count * sizeof(double)) == 0);
  memset(p, 0, count * sizeof(double));
  return p;
}

static void planner_shapes() {
  const std::vector<std::vector<std::pair<int, int>>> shapes = {
      {{1, 2}}, {{1, 9}}, {{4, 3}}, {{36, 10}}, {{64, 32}}, {{100, 7}, {9, 3}, {256, 4}}, {{144, 5}}, {{300, 3}},
      {{513, 2}}, {{1024, 2}}, {{600, 2}, {36, 4}}, {{7, 2}, {33, 3}, {130, 2}, {200, 3}},
      std::vector<std::pair<int, int>>(40, {16, 6})};
  const long live0 = fake_hip_live_allocs();
  for (int schedule = 0; schedule <= 1; ++schedule)
    for (int lanes = 1; lanes <= 2; ++lanes)
      for (int graph = 0; graph <= 1; ++graph) {
        bsofi_ctx_t ctx = nullptr;
        CHECK(bsofi_ctx_create(&ctx, 0, nullptr) == 0);
        CHECK(bsofi_ctx_set_schedule(ctx, schedule) == 0);
        CHECK(bsofi_ctx_set_lanes(ctx, lanes) == 0);
        CHECK(bsofi_ctx_set_tuning_ex(ctx, 64, 16, 4, 32, 2) == 0);
        CHECK(bsofi_ctx_set_tuning_ex(ctx, 64, 16, 4, 48, 0) == BSOFI_ERR_UNSUPPORTED);
        CHECK(bsofi_ctx_set_tuning_ex(ctx, 36, 0, 10, 0, 0) == 0);  // three-buffer 4x4x4 scan: 4 | n
        CHECK(bsofi_ctx_set_tuning_ex(ctx, 100, 0, 9, 0, 0) == BSOFI_ERR_UNSUPPORTED);  // two buffers: 8 | n
        CHECK(bsofi_ctx_set_tuning_ex(ctx, 300, 0, 10, 0, 0) == BSOFI_ERR_UNSUPPORTED);  // n <= 256
        for (const auto& sh : shapes) {
          std::vector<bsofi_matrix_t> mats;
          for (auto [n, p] : sh) {
            bsofi_matrix_t m{};
            m.n = n;
            m.p = p;
            m.lda = n * p + (n % 3);
            m.A = dev_doubles((size_t)m.lda * n * p);
            m.tau = dev_doubles((size_t)n * p);
            mats.push_back(m);
          }
          for (int phases : {BSOFI_PHASE_ALL, BSOFI_PHASE_BSOF, BSOFI_PHASE_BSTRI | BSOFI_PHASE_BSOI,
                             BSOFI_PHASE_ALL | BSOFI_PHASE_DIAG_ONLY}) {
            bsofi_plan_t plan = nullptr;
            CHECK(bsofi_plan_create(ctx, (int)mats.size(), mats.data(), phases, nullptr, &plan) == 0);
            CHECK(bsofi_plan_num_launches(plan) > 0);
            CHECK(bsofi_plan_use_graph(plan, graph) == 0);
            CHECK(bsofi_plan_time_family(plan, 5) == 0);
            CHECK(bsofi_plan_execute(plan) == 0);
            std::vector<int> info(mats.size(), -7);
            CHECK(bsofi_plan_info(plan, info.data()) == 0);
            double ms = -1;
            int launches = -1;
            CHECK(bsofi_plan_family_time(plan, &ms, &launches) == 0);
            if (phases & BSOFI_PHASE_BSOF) {
              std::vector<double> sg(mats.size()), lg(mats.size());
              CHECK(bsofi_plan_det(plan, sg.data(), lg.data()) == 0);
            }
            CHECK(bsofi_plan_destroy(plan) == 0);
          }
          if (lanes == 1 && graph == 0) {
            std::vector<int> info(mats.size());
            const int rc = bsofi_invert_batched_noalloc(ctx, (int)mats.size(), mats.data(), BSOFI_PHASE_ALL,
                                                        info.data());
            CHECK(rc == 0 || rc == BSOFI_ERR_SCRATCH);
          }
          for (auto& m : mats) {
            hip_free_for_tests(m.A);
            hip_free_for_tests(m.tau);
          }
        }
        CHECK(bsofi_ctx_release_cache(ctx) == 0);
        CHECK(bsofi_ctx_destroy(ctx) == 0);
      }
  CHECK(fake_hip_live_allocs() == live0);  // every plan buffer went back
  // bad arguments
  bsofi_ctx_t ctx = nullptr;
  CHECK(bsofi_ctx_create(&ctx, 0, nullptr) == 0);
  bsofi_matrix_t bad{};
  bad.n = 1025;
  bad.p = 2;
  bad.lda = 2050;
  bad.A = bad.tau = (double*)16;
  bsofi_plan_t plan = nullptr;
  CHECK(bsofi_plan_create(ctx, 1, &bad, BSOFI_PHASE_ALL, nullptr, &plan) == BSOFI_ERR_UNSUPPORTED);
  bad.n = 4;
  bad.lda = 7;
  CHECK(bsofi_plan_create(ctx, 1, &bad, BSOFI_PHASE_ALL, nullptr, &plan) == -4);
  CHECK(bsofi_plan_create(ctx, 1, nullptr, BSOFI_PHASE_ALL, nullptr, &plan) == -1);
  CHECK(bsofi_ctx_set_wait_limit(ctx, -1.0) == -1);
  CHECK(bsofi_ctx_set_schedule(ctx, 2) == -1);
  CHECK(bsofi_ctx_destroy(ctx) == 0);
  printf("planner shapes: ok (%ld launches)\n", fake_hip_launches());
}

static void abi_calls() {
  const int n = 5, L = 4, N = n * L, lda = N + 3;
  std::vector<double> A((size_t)lda * N, 0.0), tau(N, 0.0), work(64, 0.0);
  for (int i = 0; i < N; ++i) A[(size_t)i * lda + i] = 1.0;
  int info = 0;
  // workspace queries and argument errors, LAPACK style
  CHECK(lapackXbsof(n, L, A.data(), lda, tau.data(), work.data(), -1, &info) == 0 && work[0] >= 1);
  CHECK(lapackXbsof(0, L, A.data(), lda, tau.data(), work.data(), 64, &info) == -1 && info == -1);
  CHECK(lapackXbsof(n, 1, A.data(), lda, tau.data(), work.data(), 64, &info) == -2);
  CHECK(lapackXbsof(n, L, nullptr, lda, tau.data(), work.data(), 64, &info) == -3);
  CHECK(lapackXbsof(n, L, A.data(), N - 1, tau.data(), work.data(), 64, &info) == -4);
  CHECK(lapackXbsof(n, L, A.data(), lda, nullptr, work.data(), 64, &info) == -5);
  CHECK(lapackXbsof(n, L, A.data(), lda, tau.data(), nullptr, 64, &info) == -7);
  CHECK(lapackXbsof(1025, 2, A.data(), 4096, tau.data(), work.data(), 64, &info) == BSOFI_ERR_UNSUPPORTED);
  // the host-array paths stage through pinned buffers and copy threads
  CHECK(lapackXbsof(n, L, A.data(), lda, tau.data(), work.data(), 64, &info) == 0);
  CHECK(lapackXbstri(n, L, A.data(), lda, &info) == 0);
  CHECK(lapackXbsoi_ormqr(n, L, A.data(), lda, tau.data(), work.data(), 64, &info) == 0);
  CHECK(hybridXbsofiLWork(n, L, lda, -2) > 0 && hybridXbsofiLWork(n, L, lda, -1) > 0);
  // caller device workspace: too small (staged internally) and large enough
  const int ldw = hybridXbsofiLWork(n, L, lda, -2);
  double* dwork = nullptr;
  CHECK(hip_malloc_for_tests((void**)&dwork, (size_t)ldw * 8) == 0);
  CHECK(hybridXbsoftri(nullptr, n, L, A.data(), lda, tau.data(), work.data(), 64, dwork, 8, &info) == 0);
  CHECK(hybridXbsoftri(nullptr, n, L, A.data(), lda, tau.data(), work.data(), 64, dwork, ldw, &info) == 0);
  CHECK(hybridXbsoi(nullptr, n, L, A.data(), lda, tau.data(), work.data(), 64, dwork, ldw, &info) == 0);
  // device-pointer API
  CHECK(gpuXbsof(n, L, dwork, N, tau.data(), work.data(), 64, nullptr, &info) == 0);
  CHECK(gpuXbstri_rseq(n, L, dwork, N, nullptr, 0, &info) == 0);
  CHECK(gpuXbsoi_ormqr(n, L, dwork, N, tau.data(), work.data(), 64, nullptr, nullptr, 0, &info) == 0);
  hip_free_for_tests(dwork);
  // handle API: init, the two calls, finalize twice, calls on a dead handle
  int h[10] = {0};
  int nn = n, LL = L, llda = lda, lw = 64;
  CHECK(hybridxinith_(h, &nn, &LL, &llda) == 0);
  CHECK(hybridxbsofilworkh_(h, &nn, &LL, &llda) >= 1);
  CHECK(hybridxbsoftrih_(h, &nn, &LL, A.data(), &llda, tau.data(), work.data(), &lw, &info) == 0);
  CHECK(hybridxbsoih_(h, &nn, &LL, A.data(), &llda, tau.data(), work.data(), &lw, &info) == 0);
  CHECK(hybridxfinalizeh_(h) == 0);
  CHECK(hybridxfinalizeh_(h) == BSOFI_ERR_HANDLE);
  CHECK(hybridxbsoftrih_(h, &nn, &LL, A.data(), &llda, tau.data(), work.data(), &lw, &info) == BSOFI_ERR_HANDLE);
  int bad_h[10] = {0};
  CHECK(hybridxbsoih_(bad_h, &nn, &LL, A.data(), &llda, tau.data(), work.data(), &lw, &info) == BSOFI_ERR_HANDLE);
  int nbig = 2000;
  CHECK(hybridxinith_(h, &nbig, &LL, &llda) == -1);
  double t0[2], t1[2];
  fgetwalltime_(t0);
  fgetwalltime_(t1);
  CHECK(felapsed_(t1, t0) >= 0.0);
  printf("ABI: ok\n");
}

static void serving() {
  size_t rb64 = 0, rb36 = 0;
  CHECK(bsofi_server_request_bytes(64, 20, &rb64) == 0);
  CHECK(bsofi_server_request_bytes(36, 10, &rb36) == 0);
  const size_t expk = 64 * 64 * 8 + 36 * 36 * 8;
  const size_t budget = expk + 3 * rb64 + rb64 / 2;  // three n = 64 requests at a time
  bsofi_server_t srv = nullptr;
  CHECK(bsofi_server_create(0, budget, 4, 8, &srv) == 0);
  int c64 = -1, c36 = -1, cbig = -1;
  CHECK(bsofi_server_add_chain(srv, 8, 8, 20, 4.0, 0.125, 1.0, 7, 0, &c64) == 0);
  CHECK(bsofi_server_add_chain(srv, 6, 6, 10, 4.0, 0.125, 1.0, 7, 1, &c36) == 0);
  long long completed = 0;
  int queued = 0;
  size_t used0 = 0;
  CHECK(bsofi_server_stats(srv, &completed, &queued, &used0) == 0);
  // a request larger than the whole budget is refused at submit
  CHECK(bsofi_server_add_chain(srv, 16, 16, 80, 4.0, 0.125, 1.0, 7, 2, &cbig) == 0);
  long long rid = 0;
  const int flips[3] = {1, 5, 9};
  CHECK(bsofi_server_submit(srv, cbig, flips, 3, &rid) == BSOFI_ERR_ALLOC);
  CHECK(bsofi_server_submit(srv, c64, flips, -1, &rid) == -4);
  const int bad_flip = 64 * 20;
  CHECK(bsofi_server_submit(srv, c64, &bad_flip, 1, &rid) == -3);
  CHECK(bsofi_server_stats(srv, &completed, &queued, &used0) == 0);
  // ten requests: the budget admits three n = 64 ones at a time, the rest wait
  for (int k = 0; k < 10; ++k) CHECK(bsofi_server_submit(srv, k % 3 ? c64 : c36, flips, 3, &rid) == 0);
  int inflight = 0, max_queue = 0, done = 0;
  for (int it = 0; it < 1000 && done < 10; ++it) {
    const int q = bsofi_server_step(srv, &inflight);
    CHECK(q >= 0);
    max_queue = q > max_queue ? q : max_queue;
    size_t used = 0;
    CHECK(bsofi_server_stats(srv, &completed, &queued, &used) == 0);
    CHECK(used <= budget);
    long long ids[16];
    double lat[16], qw[16];
    int inf[16];
    const int got = bsofi_server_poll(srv, ids, lat, qw, inf, 16);
    for (int i = 0; i < got; ++i) {
      CHECK(inf[i] == 0);
      double *g = nullptr, *t = nullptr;
      int rn = 0, rp = 0;
      CHECK(bsofi_server_result(srv, ids[i], &g, &t, &rn, &rp) == 0 && g && t);
      CHECK(bsofi_server_release(srv, ids[i]) == 0);
      CHECK(bsofi_server_release(srv, ids[i]) == -2);  // released once only
      ++done;
    }
  }
  CHECK(done == 10);
  CHECK(max_queue > 0);  // requests waited for memory instead of failing
  size_t used = 0;
  CHECK(bsofi_server_stats(srv, &completed, &queued, &used) == 0);
  CHECK(used == used0 && queued == 0 && completed == 10);  // release gave the budget back
  // a device allocation that fails while admitting: the batch's requests end
  // with info = the error, their budget comes back, the server goes on
  int c100 = -1;
  CHECK(bsofi_server_add_chain(srv, 10, 10, 6, 4.0, 0.125, 1.0, 7, 3, &c100) == 0);
  CHECK(bsofi_server_stats(srv, &completed, &queued, &used0) == 0);
  long long r1 = 0, r2 = 0;
  CHECK(bsofi_server_submit(srv, c100, flips, 3, &r1) == 0);
  CHECK(bsofi_server_submit(srv, c100, flips, 3, &r2) == 0);
  fake_hip_fail_after(1);
  CHECK(bsofi_server_step(srv, &inflight) >= 0);
  fake_hip_fail_after(-1);
  long long ids[16];
  int inf[16];
  int failed = 0;
  for (int it = 0; it < 100; ++it) {
    bsofi_server_step(srv, &inflight);
    const int got = bsofi_server_poll(srv, ids, nullptr, nullptr, inf, 16);
    for (int i = 0; i < got; ++i) {
      CHECK(inf[i] == BSOFI_ERR_ALLOC || inf[i] == 0);
      failed += inf[i] != 0;
      CHECK(bsofi_server_release(srv, ids[i]) == 0);
    }
    CHECK(bsofi_server_stats(srv, &completed, &queued, &used) == 0);
    if (queued == 0 && inflight == 0 && got == 0) break;
  }
  CHECK(failed >= 1);
  CHECK(bsofi_server_stats(srv, &completed, &queued, &used) == 0);
  CHECK(used == used0);
  CHECK(bsofi_server_submit(srv, c100, flips, 3, &r1) == 0);  // and it serves again
  for (int it = 0; it < 100; ++it) {
    bsofi_server_step(srv, &inflight);
    if (bsofi_server_poll(srv, ids, nullptr, nullptr, inf, 16) == 1) {
      CHECK(inf[0] == 0);
      CHECK(bsofi_server_release(srv, ids[0]) == 0);
      break;
    }
  }
  CHECK(bsofi_server_destroy(srv) == 0);
  CHECK(bsofi_server_destroy(nullptr) == 0);
  printf("serving: ok (longest queue %d)\n", max_queue);
}

int main() {
  planner_shapes();
  abi_calls();
  serving();
  printf("host harness: all checks passed\n");
  return 0;
}

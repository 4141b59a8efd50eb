/*
 * bsofi.h -- C ABI of libbsofi_amd.so, an MI355X (gfx950) implementation of
 * BSOF/I: structured orthogonal factorization and explicit inversion of
 * block p-cyclic matrices (N = n*L, column-major, FP64, in place).
 *
 * Part 1 keeps every entry point of SGo-Go/BSOFI's C and Fortran API with
 * the same names, argument meaning, A/tau/workspace layout and info codes;
 * each line cites the reference declaration it replaces.  All of them run
 * on the GPU; there is no CPU fallback.  When no GPU is usable they return
 * info = BSOFI_ERR_NO_DEVICE and leave the outputs untouched.
 *
 * Part 2 is the device-pointer API (the reference's gpuX* family).
 * Part 3 is the native API: contexts with their own HIP stream, batched
 * inversion of many matrices of different n and p in one launch sequence,
 * reusable plans that can be captured in a HIP graph, the DQMC generator,
 * and the serving loop.  Part 3 has no counterpart in the reference.
 *
 * Argument errors follow LAPACK: info = -i for the i-th argument.
 * info > 0 reports the 1-based global index of the first exactly-zero
 * diagonal entry of R (R singular), as dtrtri does for one block.
 */
#ifndef BSOFI_AMD_H
#define BSOFI_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif
#if defined(__GNUC__)
#pragma GCC visibility push(default)
#endif

#define BSOFI_ERR_NO_DEVICE (-1000) /* no usable GPU / HIP runtime error */
#define BSOFI_ERR_ALLOC (-1001)     /* device allocation failed */
#define BSOFI_ERR_UNSUPPORTED (-1002) /* n beyond the supported range */
#define BSOFI_ERR_HANDLE (-1003)    /* uninitialised or finalized H-API handle */
#define BSOFI_ERR_SCHEDULE (-1004)  /* the device-scheduled BSOF sweep hit its wait limit */
#define BSOFI_ERR_SCRATCH (-1005)   /* the call needs more than the no-allocation scratch holds */

/* ======================================================================
 * Part 1: reference-compatible entry points (host arrays)
 * ====================================================================== */

/* include/bsofi_cpu.h:24 -- host workspace (doubles) for the calls below. */
int lapackXbsofiLWork(int n, int L, int lda);
/* include/bsofi_cpu.h:26-28 -- BSOF: A <- V + R, tau (N). */
int lapackXbsof(int n, int L, double A[], int lda, double tau[], double work[], int lwork,
                int* info);
/* include/bsofi_cpu.h:38-41 -- BSTRI: upper block triangle R <- R^{-1}. */
int lapackXbstri(int n, int L, double A[], int lda, int* info);
/* include/bsofi_cpu.h:34-36 -- BSOI: A <- G = R^{-1} Q^T (reflector form). */
int lapackXbsoi_ormqr(int n, int L, double A[], int lda, double tau[], double work[], int lwork,
                      int* info);
/* include/bsofi_cpu.h:30-32 -- BSOI with explicit Q_k (same result). */
int lapackXbsoi_orgqr(int n, int L, double A[], int lda, double tau[], double work[], int lwork,
                      int* info);
/* bsofi.h:36-37 of the reference: the default BSOI is the dormqr one. */
#define lapackXbsoi lapackXbsoi_ormqr

/* include/bsofi_cpu.h:48-72 / src/bsofi_callback.c:24-56 -- Fortran names. */
int lapackxbsofilwork_(int* n, int* L, int* lda);
int lapackxbsof_(int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork,
                 int* info);
int lapackxbstri_(int* n, int* L, double A[], int* lda, int* info);
int lapackxbsoi_ormqr_(int* n, int* L, double A[], int* lda, double tau[], double work[],
                       int* lwork, int* info);
int lapackxbsoi_orgqr_(int* n, int* L, double A[], int* lda, double tau[], double work[],
                       int* lwork, int* info);

/* include/bsofi_hybrid.h:22 -- lworkHostDevice = -1: host doubles,
 * -2: device doubles for the dwork argument below. */
int hybridXbsofiLWork(int n, int L, int lda, int lworkHostDevice);
/* include/bsofi_hybrid.h:32-37 -- BSOF + BSTRI.  The first argument took a
 * cuBLAS handle; it is kept as an opaque pointer-sized value and ignored.
 * dwork (device, ldwork doubles) is used as the whole device workspace when
 * ldwork >= hybridXbsofiLWork(n, L, lda, -2); otherwise the library
 * allocates its own.  On exit A = V + R^{-1} and tau as in the reference. */
int hybridXbsoftri(void* handle, int n, int L, double A[], int lda, double tau[], double work[],
                   int lwork, double dwork[], int ldwork, int* info);
/* include/bsofi_hybrid.h:39-43 -- BSOI; A = G on exit. */
int hybridXbsoi(void* handle, int n, int L, double A[], int lda, double tau[], double work[],
                int lwork, double dwork[], int ldwork, int* info);
/* include/bsofi_hybrid.h:39-40 -- aliases kept by the reference. */
#define hybridXbsof lapackXbsof
#define hybridXbstri lapackXbstri

/* include/bsofi_hybrid.h:56-74 / src/bsofi_callback.c:62-87 -- Fortran. */
int hybridxbsofilwork_(int* n, int* L, int* lda, int* lworkHostDevice);
int hybridxbsoftri_(int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork,
                    double dwork[], int* ldwork, int* info);
int hybridxbsoi_(int* n, int* L, double A[], int* lda, double tau[], double work[], int* lwork,
                 double dwork[], int* ldwork, int* info);

/* bsofi.h:56-70 / include/bsofi_hybrid.h:77-91 / src/bsofi_callback.c:101-180
 * -- the simplified handle API.  The handle lives in the caller's integer
 * array (at least 8 Fortran integers, as the reference's 32-byte struct):
 * {magic, n, L, lda, context pointer}, 24 bytes.  init allocates a device
 * context with its own stream, workspace and pinned staging buffers;
 * bsoftriH uploads the 2L non-zero blocks of the p-cyclic A (the rest of the
 * device copy is zero, as the contract has it), runs BSOF and BSTRI, and
 * writes V + R^{-1} (block column k: rows below (k+2) n untouched, they stay
 * the caller's zeros) and tau back; bsoiH reads A (V + R^{-1}) and tau from
 * the caller's arrays again, as the reference does, and writes A = G back.
 * finalize on a finalized (zeroed) handle returns BSOFI_ERR_HANDLE. */
int hybridxinith_(int* handle, int* n, int* L, int* lda);
int hybridxbsofilworkh_(int* handle, int* n, int* L, int* lda);
int hybridxbsoftrih_(int* handle, int* n, int* L, double A[], int* lda, double tau[],
                     double work[], int* lwork, int* info);
int hybridxbsoih_(int* handle, int* n, int* L, double A[], int* lda, double tau[], double work[],
                  int* lwork, int* info);
int hybridxfinalizeh_(int* handle);

/* src/bsofi_callback.c:189-218 -- wall-clock helpers for Fortran callers. */
void fgetwalltime_(double tv[2]);
double felapsed_(double t1[2], double t0[2]);

/* ======================================================================
 * Part 2: device-pointer API (include/bsofi_gpu.h:34-62)
 * dA is device memory with leading dimension ldda; tau and work are host
 * arrays as in the reference.  dT is accepted for signature compatibility
 * and not used: the BSOI phase rebuilds the T factors from the reflectors in
 * dA and tau (a -2 query on gpuXbsof returns 1).  dwork may be NULL, or
 * device memory of (-2 query) doubles used as the whole workspace.
 * ====================================================================== */
int gpuXbsof(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork,
             double* dT, int* info);
int gpuXbstri_rseq(int n, int L, double dA[], int ldda, double dwork[], int ldwork, int* info);
int gpuXbstri_rbatched(int n, int L, double dA[], int ldda, double dwork[], int ldwork,
                       int* info);
int gpuXbsoi_ormqr(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork,
                   double* dT, double dwork[], int ldwork, int* info);
int gpuXbsoi_orgqr(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork,
                   double* dT, double dwork[], int ldwork, int* info);
int gpuXbsoi_batched(int n, int L, double dA[], int ldda, double tau[], double work[], int lwork,
                     double* dT, double dwork[], int ldwork, int* info);
#define gpuXbstri gpuXbstri_rseq
#define gpuXbsoi gpuXbsoi_batched

/* ======================================================================
 * Part 3: native API (no counterpart in the reference)
 * ====================================================================== */
typedef struct bsofi_ctx_s* bsofi_ctx_t;
typedef struct bsofi_plan_s* bsofi_plan_t;

/* One p-cyclic matrix resident in device memory. */
typedef struct {
  double* A;   /* device, lda x N, column-major, in place */
  double* tau; /* device, N doubles */
  int n, p, lda;
  int pad;
} bsofi_matrix_t;

#define BSOFI_PHASE_BSOF 1
#define BSOFI_PHASE_BSTRI 2
#define BSOFI_PHASE_BSOI 4
#define BSOFI_PHASE_ALL 7
/* With BSOFI_PHASE_BSOI: produce only the diagonal blocks G_ll (the
 * equal-time Green's functions of DQMC), in place in the diagonal blocks of
 * A; the other blocks of A are left holding intermediate values.  The
 * diagonal blocks are bit-identical to those of the full inverse.  No
 * counterpart in the reference. */
#define BSOFI_PHASE_DIAG_ONLY 8

/* Contexts.  stream = NULL makes the context create its own stream. */
int bsofi_device_count(void);
int bsofi_ctx_create(bsofi_ctx_t* ctx, int device, void* hip_stream);
int bsofi_ctx_destroy(bsofi_ctx_t ctx);
int bsofi_ctx_set_stream(bsofi_ctx_t ctx, void* hip_stream);
void* bsofi_ctx_stream(bsofi_ctx_t ctx);
int bsofi_ctx_synchronize(bsofi_ctx_t ctx);
/* Tuning (the autotuner's knobs), per block size n, for plans created on
 * the context afterwards: nb = sub-panel width of BSOF/BSOI (8, 16 or 32;
 * 32 needs 2n <= 512), scan_mti = form of the BSOI scan: 1, 2 or 4 =
 * 16-row multiples per workgroup on the 16x16x4 FP64 MFMA with 4 waves
 * (1 needs n <= 512, 2 needs n <= 256, 4 needs n <= 128), 8 = 16 rows with
 * 8 waves (512 < n <= 1024, the default there), 9 = 32 rows on the 4x4x4
 * FP64 MFMA with two operand buffers (needs n <= 256 and n % 8 == 0), 10 =
 * the same with three buffers, loading two k-steps ahead (n <= 256,
 * n % 4 == 0; the default for n >= 128; the same bits as 9), 12 = form 10
 * with two workgroups per CU (8 <= n <= 64, n % 4 == 0; the default there;
 * the same bits as 10).  0
 * keeps the default; unsupported pairs return BSOFI_ERR_UNSUPPORTED.  No
 * counterpart in the reference, whose static CPU/GPU split (KAPPA_*,
 * *_SWITCH_*) these choices replace. */
int bsofi_ctx_set_tuning(bsofi_ctx_t ctx, int n, int nb, int scan_mti);
int bsofi_ctx_get_tuning(bsofi_ctx_t ctx, int n, int* nb, int* scan_mti);
/* The same with two more knobs: gemm_tile = edge of the BSTRI GEMM tiles
 * (0: 128 where a product yields a full round of such tiles, else 64, and
 * 32 for launches of fewer 64-tiles than CUs; or 32, 64, 128 fixed) and
 * refl_spw = 16-column slabs per workgroup of the reflector applications up
 * to m = 512 (0: as many as fit one round of workgroups, at most 8; or 1, 2,
 * 4, 8 fixed; a launch that mixes block sizes takes the smallest setting).
 * Neither changes the bits.  set_tuning keeps them at 0. */
int bsofi_ctx_set_tuning_ex(bsofi_ctx_t ctx, int n, int nb, int scan_mti, int gemm_tile, int refl_spw);
int bsofi_ctx_get_tuning_ex(bsofi_ctx_t ctx, int n, int* nb, int* scan_mti, int* gemm_tile, int* refl_spw);
int bsofi_ctx_clear_tuning(bsofi_ctx_t ctx);
/* Lanes of batch plans created on the context afterwards: 2 (the default)
 * splits a plan of two or more matrices into two groups of about equal work
 * that run on two streams (the context's and a second one it creates),
 * forked after the plan's first launch and joined before it returns to the
 * context's stream; 1 keeps every launch on the context's stream.  Results
 * are the same bits either way.  Callers that already run several plans at
 * once on several contexts (the serving loop) want 1.  Returns -1 for other
 * values.  No counterpart in the reference. */
int bsofi_ctx_set_lanes(bsofi_ctx_t ctx, int lanes);
/* How plans created on the context afterwards run BSOF.  0 (the default):
 * one launch per sweep step (panel, then reflectors), scheduled on the host.
 * 1: where every matrix of the plan has 2n <= 512 and 32-wide sub-panels, the
 * whole sweep of all matrices is ONE persistent launch that schedules itself
 * on the device: workgroups take sub-panel QRs and reflector-slab updates
 * from a device queue ordered so that every task's inputs come from earlier
 * tasks, and wait on per-column counters and per-sub-panel flags for them
 * (the next sub-panel's QR starts as soon as its own columns are updated,
 * beside the updates of the rest).  Both give the same bits; on MI355X the
 * device-scope release / acquire every task needs (the XCDs' L2 caches are
 * not coherent) costs more than the launches it saves, so 1 is measured
 * slower and is not the default.  The
 * reference's CPU/GPU load balancing (src/bsofi_hybrid_load_balancing.c,
 * KAPPA_* and *_SWITCH_*) is what this replaces.  Returns -1 for other
 * values. */
int bsofi_ctx_set_schedule(bsofi_ctx_t ctx, int mode);
/* Wait limit of the device-scheduled sweep: a workgroup that waits longer
 * than this for a task's inputs stops the sweep, and every matrix of the
 * plan reports info = BSOFI_ERR_SCHEDULE instead of the GPU hanging
 * (default 4 s; 0 makes any wait fail, for tests of that path).  Returns -1
 * outside [0, 3600]. */
int bsofi_ctx_set_wait_limit(bsofi_ctx_t ctx, double seconds);
/* Destroyed plans return their device buffers (task tables, workspace) to
 * the context, and later plans on it reuse them, so that creating and
 * destroying plans in a loop allocates nothing once warm.  These report and
 * free that cache (at most 8 buffers).  No counterpart in the reference. */
size_t bsofi_ctx_cached_bytes(bsofi_ctx_t ctx);
int bsofi_ctx_release_cache(bsofi_ctx_t ctx);
/* Achievable FP64 rate of this GPU (TFLOP/s) for one instruction kind run
 * back to back on every CU: 0 = v_mfma_f64_16x16x4f64 (what the kernels
 * use), 1 = v_mfma_f64_4x4x4f64, 2 = v_fma_f64.  A calibration for roofline
 * figures next to the datasheet peak.  No counterpart in the reference. */
int bsofi_fp64_rate(bsofi_ctx_t ctx, int kind, double* tflops);
/* Supported block sizes: 1 <= n <= BSOFI_MAX_N, L >= 2. */
#define BSOFI_MAX_N 1024
int bsofi_check_args(int n, int L, int lda);

/* Device workspace (bytes) one matrix of order (n, p) needs. */
size_t bsofi_workspace_bytes(int n, int p);

/* Plans: the launch sequence for a batch and a phase set, built once and
 * executed many times (optionally as a HIP graph).  workspace may be NULL
 * (the plan allocates it) or device memory of bsofi_plan_workspace_bytes. */
size_t bsofi_plan_workspace_bytes(int count, const bsofi_matrix_t* mats, int phases);
int bsofi_plan_create(bsofi_ctx_t ctx, int count, const bsofi_matrix_t* mats, int phases,
                      void* workspace, bsofi_plan_t* plan);
int bsofi_plan_execute(bsofi_plan_t plan);
int bsofi_plan_use_graph(bsofi_plan_t plan, int on);
int bsofi_plan_num_launches(bsofi_plan_t plan);
/* Per-matrix status after execution (host copy, count ints). */
int bsofi_plan_info(bsofi_plan_t plan, int* info);
/* sign(det A) (+1, -1, or 0 if R is singular) and log|det A| of every matrix
 * of a plan that ran BSOF, taken from R's diagonal and tau right after BSOF
 * (before BSTRI overwrites R): each reflector with tau != 0 contributes
 * det = -1.  Host arrays of the plan's count.  Returns -2 for a plan without
 * the BSOF phase.  (The reference's examples/ftest_cpu/ftest_cpu.f90:134-138
 * flips its sign on tau == 0 instead, i.e. it yields (-1)^N sign(det A).) */
int bsofi_plan_det(bsofi_plan_t plan, double* sign, double* logabs);
/* The same for one matrix already factored by BSOF on the device. */
int bsofi_det(bsofi_ctx_t ctx, const double* dA, int lda, int n, int p, const double* dtau, double* sign,
              double* logabs);
/* Time every launch of one kernel family with HIP events on the plan's
 * stream; read the accumulated milliseconds and launch count after a sync.
 * family: 0 panel QR, 1 reflector apply, 2 tile GEMM, 3 diag TRTRI,
 * 4 identity fill, 5 BSOI scan, 6 T factor, 8 determinant, 9 the
 * device-scheduled BSOF sweep. */
int bsofi_plan_time_family(bsofi_plan_t plan, int family);
int bsofi_plan_family_time(bsofi_plan_t plan, double* ms, int* launches);
int bsofi_plan_destroy(bsofi_plan_t plan);

/* One-shot batched inversion on device (builds, runs and frees a plan). */
int bsofi_invert_batched(bsofi_ctx_t ctx, int count, const bsofi_matrix_t* mats, int phases,
                         int* info);
/* bsofi_invert_batched without any device allocation, for callers that
 * cannot allocate (HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The plan's workspace
 * and task tables live in one fixed device array of
 * BSOFI_NOALLOC_SCRATCH_BYTES per GPU (held from when the library is loaded
 * until the process ends); a call whose workspace
 * (bsofi_plan_workspace_bytes, rounded up to 256) plus task tables exceed it
 * returns BSOFI_ERR_SCRATCH, and bsofi_invert_batched remains the way to run
 * it.  Calls no hipMalloc / hipFree and creates no stream or event (every
 * launch on the context's stream); the planner's host-side tables use the
 * host heap.  Synchronous, like bsofi_invert_batched, and the same bits.
 * Every call on one GPU shares that GPU's scratch: calls that overlap on one
 * GPU (from several threads or contexts) race and must be serialised by the
 * caller. */
#define BSOFI_NOALLOC_SCRATCH_BYTES (128u << 20)
int bsofi_invert_batched_noalloc(bsofi_ctx_t ctx, int count, const bsofi_matrix_t* mats, int phases,
                                 int* info);

/* Relative residual ||A G - I||_F / (||A||_F ||G||_F) computed on the
 * device from the 2p non-zero blocks of the original A (blocks in the
 * compact layout written by bsofi_gather_blocks) and G. */
int bsofi_gather_blocks(bsofi_ctx_t ctx, const double* dA, int lda, int n, int p,
                        double* dblocks);
int bsofi_pcyclic_residual(bsofi_ctx_t ctx, const double* dblocks, const double* dG, int ldg,
                           int n, int p, double* resid);

/* C = alpha * op(A) * B + beta * C with the library's FP64 MFMA tile GEMM;
 * amask / bmask read A / B as upper triangular.  Synchronous; for tests. */
int bsofi_debug_gemm(bsofi_ctx_t ctx, int M, int N, int K, double alpha, const double* dA, int lda,
                     const double* dB, int ldb, double beta, double* dC, int ldc, int amask, int bmask);

/* Synchronous copy between any host/device pointers (hipMemcpyDefault). */
int bsofi_memcpy(void* dst, const void* src, size_t bytes);

/* Flop counts of include/bsofi_flops.h:19-30 (out[0..2] = BSOF, BSTRI,
 * BSOI). */
void bsofi_flops(int n, int L, double out[3]);

/* DQMC generator.  e^{-dtau K} of the periodic lx x ly lattice (hopping t)
 * into expk (n x n, column-major, n = lx*ly), by the closed form of the
 * lattice Fourier modes. */
int bsofi_dqmc_expk(int lx, int ly, double t, double dtau, double* expk);
/* HS field s_{l,i} in {+1,-1} (l < p, i < n) from the counter-based hash
 * keyed (seed, instance, l, i); host and device produce the same values. */
int bsofi_dqmc_fields(unsigned long long seed, int instance, int p, int n, signed char* fields);
/* Device: A = p-cyclic DQMC matrix from device expk (n x n) and device
 * fields (p x n int8), with e^{+-lambda} given. */
int bsofi_dqmc_build(bsofi_ctx_t ctx, double* dA, int lda, int n, int p, const double* dexpk,
                     const signed char* dfields, double exp_plus, double exp_minus);
/* Device: flip count HS fields at flat indices l*n + i. */
int bsofi_dqmc_flip(bsofi_ctx_t ctx, signed char* dfields, const int* dflips, int count);

/* Serving loop (continuous batching of DQMC inversion requests).
 * A chain owns a lattice (lx x ly, hopping t, dtau), p, U and an HS field
 * seeded by (seed, instance); e^{-dtau K} of each lattice is cached on the
 * device.  submit() flips fields (flat indices l*n + i) and queues a
 * request that ships only its field snapshot; step() retires finished
 * batches, admits queued requests FIFO into free stream slots while their
 * device memory (G, tau, workspace) fits the HBM budget, and launches one
 * batched plan per slot.  Results stay on the device until release(). */
typedef struct bsofi_server_s* bsofi_server_t;
int bsofi_server_create(int device, size_t hbm_budget_bytes, int nslots, int max_batch, bsofi_server_t* srv);
int bsofi_server_destroy(bsofi_server_t srv);
int bsofi_server_add_chain(bsofi_server_t srv, int lx, int ly, int p, double U, double dtau, double t,
                           unsigned long long seed, int instance, int* chain_id);
int bsofi_server_submit(bsofi_server_t srv, int chain, const int* flips, int nflips, long long* request_id);
/* Returns the queue length (>= 0) or an error (< 0). */
int bsofi_server_step(bsofi_server_t srv, int* inflight);
int bsofi_server_poll(bsofi_server_t srv, long long* ids, double* latency_s, double* queue_wait_s, int* info,
                      int max);
int bsofi_server_result(bsofi_server_t srv, long long id, double** dG, double** dtau, int* n, int* p);
int bsofi_server_release(bsofi_server_t srv, long long id);
int bsofi_server_stats(bsofi_server_t srv, long long* completed, int* queued, size_t* used_bytes);
int bsofi_server_request_bytes(int n, int p, size_t* bytes);

#if defined(__GNUC__)
#pragma GCC visibility pop
#endif
#ifdef __cplusplus
}
#endif
#endif /* BSOFI_AMD_H */

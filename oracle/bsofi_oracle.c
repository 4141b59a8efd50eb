/*
 * bsofi_oracle.c -- CPU oracle for BSOF/I.  TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, and only as the checker or the timed CPU baseline.  The
 * product (bsofi_amd/libbsofi_amd.so) never links, loads or calls it.
 *
 * What it restates
 * ----------------
 * 1. The reference's BSOF/I call sequence (SGo-Go/BSOFI src/bsofi_cpu.c):
 *      oracle_bsofi_lwork   <- lapackXbsofiLWork   src/bsofi_cpu.c:31-42
 *      oracle_bsof          <- lapackXbsof         src/bsofi_cpu.c:86-118
 *      oracle_bstri         <- lapackXbstri        src/bsofi_cpu.c:146-177
 *      oracle_bsoi_ormqr    <- lapackXbsoi_ormqr   src/bsofi_cpu.c:211-248
 *      oracle_bsoi_orgqr    <- lapackXbsoi_orgqr   src/bsofi_cpu.c:283-333
 *      cpyzero              <- lapackXcpyzero      src/bsofi_lapack.c:260-281
 *    Each makes the same LAPACK/BLAS calls, with the same shapes, offsets and
 *    order, as the reference function it names.
 *
 * 2. The LAPACK routines those calls reach.  LAPACK is a third-party
 *    dependency that the reference does not vendor (its makefiles pin MKL
 *    composer_xe_2013 / Netlib LAPACK 3.4.2 in the config/platforms/ makefiles).
 *    Three back ends are selectable at run time (oracle_set_backend):
 *      BACKEND_MKL      : libmkl_rt.so (the reference's default LAPACK,
 *                         config/makefile-general-settings.mk:76), dlopen'd.
 *      BACKEND_OPENBLAS : scipy's bundled OpenBLAS (symbols prefixed scipy_).
 *      BACKEND_BUILTIN  : plain-C restatements of the published unblocked
 *                         LAPACK 3.x algorithms (dlarfg/dnrm2/dlapy2, dgeqr2,
 *                         dorm2r, dorg2r, dtrti2, dtrmm, dgemm, dlacpy), below.
 *    The blocked MKL routines and the unblocked restatements apply the same
 *    reflectors; they agree to rounding, which the tests check.
 *
 * Parity pins (tests/test_oracle.py): the builtin and MKL back ends against
 * each other; the MKL back end bit-for-bit against oracle/_ref (the
 * reference's own src/bsofi_cpu.c compiled against the same MKL, one
 * thread); the golden fixtures under tests/golden/; and the closed forms
 * (A*G = I, Q^T Q = I, Q*R = A, p = 2 reducing to one dgeqrf, R^{-1} by
 * numpy, the sign of det A by slogdet).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* LAPACK function table                                              */
/* ------------------------------------------------------------------ */
typedef void (*geqrf_t)(const int*, const int*, double*, const int*, double*,
                        double*, const int*, int*);
typedef void (*ormqr_t)(const char*, const char*, const int*, const int*,
                        const int*, double*, const int*, double*, double*,
                        const int*, double*, const int*, int*);
typedef void (*orgqr_t)(const int*, const int*, const int*, double*, const int*,
                        double*, double*, const int*, int*);
typedef void (*trtri_t)(const char*, const char*, const int*, double*,
                        const int*, int*);
typedef void (*trmm_t)(const char*, const char*, const char*, const char*,
                       const int*, const int*, const double*, double*,
                       const int*, double*, const int*);
typedef void (*gemm_t)(const char*, const char*, const int*, const int*,
                       const int*, const double*, double*, const int*,
                       double*, const int*, const double*, double*,
                       const int*);
typedef void (*lacpy_t)(const char*, const int*, const int*, double*,
                        const int*, double*, const int*);

typedef struct {
  geqrf_t geqrf;
  ormqr_t ormqr;
  orgqr_t orgqr;
  trtri_t trtri;
  trmm_t trmm;
  gemm_t gemm;
  lacpy_t lacpy;
} lapack_table;

enum { BACKEND_BUILTIN = 0, BACKEND_MKL = 1, BACKEND_OPENBLAS = 2 };

static lapack_table g_la;
static int g_backend = -1;
static void* g_handle = NULL;
static char g_backend_path[1024] = "";

/* ------------------------------------------------------------------ */
/* Builtin restatement of the LAPACK 3.x unblocked algorithms          */
/* ------------------------------------------------------------------ */
#define AI(a, lda, i, j) ((a)[(size_t)(j) * (size_t)(lda) + (size_t)(i)])

/* dlapy2: sqrt(x^2 + y^2) without destructive overflow (LAPACK dlapy2). */
static double b_dlapy2(double x, double y) {
  double xa = fabs(x), ya = fabs(y);
  double w = xa > ya ? xa : ya, z = xa < ya ? xa : ya;
  if (z == 0.0 || w > DBL_MAX) return w;
  return w * sqrt(1.0 + (z / w) * (z / w));
}

/* dnrm2: scaled sum of squares (reference BLAS dnrm2, LAPACK <= 3.9). */
static double b_dnrm2(int n, const double* x, int incx) {
  if (n < 1 || incx < 1) return 0.0;
  if (n == 1) return fabs(x[0]);
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < n; i++) {
    double v = x[(size_t)i * incx];
    if (v != 0.0) {
      double a = fabs(v);
      if (scale < a) {
        ssq = 1.0 + ssq * (scale / a) * (scale / a);
        scale = a;
      } else {
        ssq += (a / scale) * (a / scale);
      }
    }
  }
  return scale * sqrt(ssq);
}

/* dlarfg: generate H with H*(alpha;x) = (beta;0), H = I - tau*v*v', v(1)=1,
 * beta = -sign(alpha)*||(alpha;x)||, tau = (beta-alpha)/beta, x <- x/(alpha-beta);
 * tau = 0 when x = 0.  Safe rescaling loop as in LAPACK dlarfg. */
static void b_dlarfg(int n, double* alpha, double* x, int incx, double* tau) {
  if (n <= 1) { *tau = 0.0; return; }
  double xnorm = b_dnrm2(n - 1, x, incx);
  if (xnorm == 0.0) { *tau = 0.0; return; }
  double beta = -copysign(b_dlapy2(*alpha, xnorm), *alpha);
  double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  int knt = 0;
  if (fabs(beta) < safmin) {
    double rsafmn = 1.0 / safmin;
    do {
      knt++;
      for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= rsafmn;
      beta *= rsafmn;
      *alpha *= rsafmn;
    } while (fabs(beta) < safmin && knt < 20);
    xnorm = b_dnrm2(n - 1, x, incx);
    beta = -copysign(b_dlapy2(*alpha, xnorm), *alpha);
  }
  *tau = (beta - *alpha) / beta;
  double sc = 1.0 / (*alpha - beta);
  for (int i = 0; i < n - 1; i++) x[(size_t)i * incx] *= sc;
  for (int j = 0; j < knt; j++) beta *= safmin;
  *alpha = beta;
}

/* dlarf, side 'L': C <- (I - tau v v') C, v stored with unit first entry. */
static void b_dlarf_left(int m, int n, const double* v, double tau, double* c,
                         int ldc, double* work) {
  if (tau == 0.0) return;
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    for (int i = 0; i < m; i++) s += v[i] * AI(c, ldc, i, j);
    work[j] = s;
  }
  for (int j = 0; j < n; j++) {
    double w = tau * work[j];
    for (int i = 0; i < m; i++) AI(c, ldc, i, j) -= v[i] * w;
  }
}

/* dlarf, side 'R': C <- C (I - tau v v'). */
static void b_dlarf_right(int m, int n, const double* v, double tau, double* c,
                          int ldc, double* work) {
  if (tau == 0.0) return;
  for (int i = 0; i < m; i++) work[i] = 0.0;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) work[i] += AI(c, ldc, i, j) * v[j];
  for (int j = 0; j < n; j++) {
    double w = tau * v[j];
    for (int i = 0; i < m; i++) AI(c, ldc, i, j) -= work[i] * w;
  }
}

/* dgeqr2 (also serves as dgeqrf): Householder QR, column by column. */
static void b_dgeqrf(const int* pm, const int* pn, double* a, const int* plda,
                     double* tau, double* work, const int* plwork, int* info) {
  int m = *pm, n = *pn, lda = *plda, k = m < n ? m : n;
  *info = 0;
  if (*plwork == -1) { work[0] = n > 1 ? n : 1; return; }
  double* w = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double* v = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  for (int i = 0; i < k; i++) {
    int len = m - i;
    int row1 = (i + 1 < m) ? i + 1 : m - 1;
    b_dlarfg(len, &AI(a, lda, i, i), &AI(a, lda, row1, i), 1, &tau[i]);
    if (i < n - 1) {
      v[0] = 1.0;
      for (int r = 1; r < len; r++) v[r] = AI(a, lda, i + r, i);
      b_dlarf_left(len, n - i - 1, v, tau[i], &AI(a, lda, i, i + 1), lda, w);
    }
  }
  free(w);
  free(v);
}

/* dorm2r (also serves as dormqr). */
static void b_dormqr(const char* side, const char* trans, const int* pm,
                     const int* pn, const int* pk, double* a, const int* plda,
                     double* tau, double* c, const int* pldc, double* work,
                     const int* plwork, int* info) {
  int m = *pm, n = *pn, k = *pk, lda = *plda, ldc = *pldc;
  int left = (side[0] == 'L' || side[0] == 'l');
  int notran = (trans[0] == 'N' || trans[0] == 'n');
  *info = 0;
  if (*plwork == -1) { work[0] = left ? (n > 1 ? n : 1) : (m > 1 ? m : 1); return; }
  int nq = left ? m : n;
  double* v = (double*)malloc(sizeof(double) * (size_t)(nq > 0 ? nq : 1));
  double* w = (double*)malloc(sizeof(double) * (size_t)((m > n ? m : n) + 1));
  int forward = (left && !notran) || (!left && notran);
  for (int t = 0; t < k; t++) {
    int i = forward ? t : k - 1 - t;
    int len = nq - i;
    v[0] = 1.0;
    for (int r = 1; r < len; r++) v[r] = AI(a, lda, i + r, i);
    if (left)
      b_dlarf_left(len, n, v, tau[i], &AI(c, ldc, i, 0), ldc, w);
    else
      b_dlarf_right(m, len, v, tau[i], &AI(c, ldc, 0, i), ldc, w);
  }
  free(v);
  free(w);
}

/* dorg2r (also serves as dorgqr). */
static void b_dorgqr(const int* pm, const int* pn, const int* pk, double* a,
                     const int* plda, double* tau, double* work,
                     const int* plwork, int* info) {
  int m = *pm, n = *pn, k = *pk, lda = *plda;
  *info = 0;
  if (*plwork == -1) { work[0] = n > 1 ? n : 1; return; }
  double* v = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  double* w = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int j = k; j < n; j++) {
    for (int l = 0; l < m; l++) AI(a, lda, l, j) = 0.0;
    AI(a, lda, j, j) = 1.0;
  }
  for (int i = k - 1; i >= 0; i--) {
    if (i < n - 1) {
      int len = m - i;
      v[0] = 1.0;
      for (int r = 1; r < len; r++) v[r] = AI(a, lda, i + r, i);
      b_dlarf_left(len, n - i - 1, v, tau[i], &AI(a, lda, i, i + 1), lda, w);
    }
    for (int r = i + 1; r < m; r++) AI(a, lda, r, i) *= -tau[i];
    AI(a, lda, i, i) = 1.0 - tau[i];
    for (int l = 0; l < i; l++) AI(a, lda, l, i) = 0.0;
  }
  free(v);
  free(w);
}

/* dtrti2 (also serves as dtrtri), uplo 'U', diag 'N' only. */
static void b_dtrtri(const char* uplo, const char* diag, const int* pn,
                     double* a, const int* plda, int* info) {
  int n = *pn, lda = *plda;
  (void)uplo;
  (void)diag;
  *info = 0;
  for (int j = 0; j < n; j++)
    if (AI(a, lda, j, j) == 0.0) { *info = j + 1; return; }
  for (int j = 0; j < n; j++) {
    AI(a, lda, j, j) = 1.0 / AI(a, lda, j, j);
    double ajj = -AI(a, lda, j, j);
    /* x <- T x, T = A(0:j,0:j) upper (already inverted), x = A(0:j, j) */
    for (int l = 0; l < j; l++) {
      double xl = AI(a, lda, l, j);
      if (xl != 0.0) {
        for (int i = 0; i < l; i++) AI(a, lda, i, j) += xl * AI(a, lda, i, l);
        AI(a, lda, l, j) = xl * AI(a, lda, l, l);
      }
    }
    for (int i = 0; i < j; i++) AI(a, lda, i, j) *= ajj;
  }
}

/* dtrmm, uplo 'U', transa 'N', diag 'N'; sides 'L' and 'R'. */
static void b_dtrmm(const char* side, const char* uplo, const char* transa,
                    const char* diag, const int* pm, const int* pn,
                    const double* palpha, double* a, const int* plda,
                    double* b, const int* pldb) {
  int m = *pm, n = *pn, lda = *plda, ldb = *pldb;
  double alpha = *palpha;
  (void)uplo;
  (void)transa;
  (void)diag;
  if (side[0] == 'L' || side[0] == 'l') {
    /* B <- alpha * A * B, A m x m upper */
    for (int j = 0; j < n; j++)
      for (int i = 0; i < m; i++) {
        double s = 0.0;
        for (int l = i; l < m; l++) s += AI(a, lda, i, l) * AI(b, ldb, l, j);
        AI(b, ldb, i, j) = alpha * s;
      }
  } else {
    /* B <- alpha * B * A, A n x n upper; column j of result uses A(0:j, j) */
    for (int j = n - 1; j >= 0; j--)
      for (int i = 0; i < m; i++) {
        double s = 0.0;
        for (int l = 0; l <= j; l++) s += AI(b, ldb, i, l) * AI(a, lda, l, j);
        AI(b, ldb, i, j) = alpha * s;
      }
  }
}

/* dgemm, transa 'N', transb 'N' or 'T'. */
static void b_dgemm(const char* transa, const char* transb, const int* pm,
                    const int* pn, const int* pk, const double* palpha,
                    double* a, const int* plda, double* b, const int* pldb,
                    const double* pbeta, double* c, const int* pldc) {
  int m = *pm, n = *pn, k = *pk, lda = *plda, ldb = *pldb, ldc = *pldc;
  double alpha = *palpha, beta = *pbeta;
  int tb = (transb[0] == 'T' || transb[0] == 't');
  (void)transa;
  double* col = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  for (int j = 0; j < n; j++) {
    for (int i = 0; i < m; i++) col[i] = 0.0;
    for (int l = 0; l < k; l++) {
      double blj = tb ? AI(b, ldb, j, l) : AI(b, ldb, l, j);
      if (blj != 0.0)
        for (int i = 0; i < m; i++) col[i] += AI(a, lda, i, l) * blj;
    }
    for (int i = 0; i < m; i++)
      AI(c, ldc, i, j) = (beta == 0.0 ? 0.0 : beta * AI(c, ldc, i, j)) + alpha * col[i];
  }
  free(col);
}

static void b_dlacpy(const char* uplo, const int* pm, const int* pn, double* a,
                     const int* plda, double* b, const int* pldb) {
  int m = *pm, n = *pn, lda = *plda, ldb = *pldb;
  (void)uplo;
  for (int j = 0; j < n; j++)
    for (int i = 0; i < m; i++) AI(b, ldb, i, j) = AI(a, lda, i, j);
}

/* ------------------------------------------------------------------ */
/* Back-end selection                                                  */
/* ------------------------------------------------------------------ */
static void* resolve(const char* prefix, const char* name) {
  char buf[128];
  snprintf(buf, sizeof buf, "%s%s", prefix, name);
  return dlsym(g_handle, buf);
}

/* Returns 0 on success, -1 if the library or a symbol is missing. */
int oracle_set_backend(int backend, const char* path) {
  if (g_handle) { dlclose(g_handle); g_handle = NULL; }
  g_backend_path[0] = 0;
  if (backend == BACKEND_BUILTIN) {
    g_la.geqrf = b_dgeqrf; g_la.ormqr = b_dormqr; g_la.orgqr = b_dorgqr;
    g_la.trtri = b_dtrtri; g_la.trmm = b_dtrmm; g_la.gemm = b_dgemm;
    g_la.lacpy = b_dlacpy;
    g_backend = backend;
    snprintf(g_backend_path, sizeof g_backend_path, "builtin");
    return 0;
  }
  const char* prefix = (backend == BACKEND_OPENBLAS) ? "scipy_" : "";
  if (!path || !path[0]) return -1;
  /* MKL's default Intel OpenMP layer gives wrong results when libgomp (which
   * torch loads) is in the same process; the GNU layer shares libgomp. */
  if (backend == BACKEND_MKL) setenv("MKL_THREADING_LAYER", "GNU", 0);
  g_handle = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!g_handle) return -1;
  g_la.geqrf = (geqrf_t)resolve(prefix, "dgeqrf_");
  g_la.ormqr = (ormqr_t)resolve(prefix, "dormqr_");
  g_la.orgqr = (orgqr_t)resolve(prefix, "dorgqr_");
  g_la.trtri = (trtri_t)resolve(prefix, "dtrtri_");
  g_la.trmm = (trmm_t)resolve(prefix, "dtrmm_");
  g_la.gemm = (gemm_t)resolve(prefix, "dgemm_");
  g_la.lacpy = (lacpy_t)resolve(prefix, "dlacpy_");
  if (!g_la.geqrf || !g_la.ormqr || !g_la.orgqr || !g_la.trtri || !g_la.trmm ||
      !g_la.gemm || !g_la.lacpy) {
    dlclose(g_handle);
    g_handle = NULL;
    g_backend = -1;
    return -1;
  }
  g_backend = backend;
  snprintf(g_backend_path, sizeof g_backend_path, "%s", path);
  return 0;
}

int oracle_get_backend(void) { return g_backend; }
const char* oracle_get_backend_path(void) { return g_backend_path; }

/* Sets the thread count of the MKL back end (mkl_set_num_threads); returns
 * -1 when the active back end has no such control. */
int oracle_set_threads(int nthreads) {
  if (!g_handle) return -1;
  void (*f)(const int*) = (void (*)(const int*))dlsym(g_handle, "MKL_Set_Num_Threads");
  if (f) { f(&nthreads); return 0; }
  void (*g)(int) = (void (*)(int))dlsym(g_handle, "scipy_openblas_set_num_threads64_");
  if (!g) g = (void (*)(int))dlsym(g_handle, "scipy_openblas_set_num_threads");
  if (g) { g(nthreads); return 0; }
  return -1;
}

static int ensure_backend(void) {
  if (g_backend >= 0) return 0;
  return oracle_set_backend(BACKEND_BUILTIN, NULL);
}

/* ------------------------------------------------------------------ */
/* Char-argument wrappers, as src/bsofi_lapack.c:130-250               */
/* ------------------------------------------------------------------ */
static void xgeqrf(int m, int n, double* A, int lda, double* tau, double* work,
                   int lwork, int* info) {
  g_la.geqrf(&m, &n, A, &lda, tau, work, &lwork, info);
}
static void xormqr(char side, char trans, int m, int n, int k, double* a,
                   int lda, double* tau, double* c, int ldc, double* work,
                   int lwork, int* info) {
  char s[2] = {side, 0}, t[2] = {trans, 0};
  g_la.ormqr(s, t, &m, &n, &k, a, &lda, tau, c, &ldc, work, &lwork, info);
}
static void xorgqr(int m, int n, int k, double* a, int lda, double* tau,
                   double* work, int lwork, int* info) {
  g_la.orgqr(&m, &n, &k, a, &lda, tau, work, &lwork, info);
}
static void xtrtri(char uplo, char diag, int n, double* A, int lda, int* info) {
  char u[2] = {uplo, 0}, d[2] = {diag, 0};
  g_la.trtri(u, d, &n, A, &lda, info);
}
static void xtrmm(char side, char uplo, char transa, char diag, int m, int n,
                  double alpha, double* A, int lda, double* B, int ldb) {
  char s[2] = {side, 0}, u[2] = {uplo, 0}, t[2] = {transa, 0}, d[2] = {diag, 0};
  g_la.trmm(s, u, t, d, &m, &n, &alpha, A, &lda, B, &ldb);
}
static void xgemm(char transa, char transb, int m, int n, int k, double alpha,
                  double* a, int lda, double* b, int ldb, double beta,
                  double* c, int ldc) {
  char ta[2] = {transa, 0}, tb[2] = {transb, 0};
  g_la.gemm(ta, tb, &m, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
}
static void xlacpy(char uplo, int m, int n, double* A, int lda, double* B,
                   int ldb) {
  char u[2] = {uplo, 0};
  g_la.lacpy(u, &m, &n, A, &lda, B, &ldb);
}

/* lapackXcpyzero('L'): move the strict lower triangle (trapezoid) of A into
 * B and zero it in A (src/bsofi_lapack.c:260-281, case 'L'). */
static void cpyzero_lower(int m, int n, double* A, int lda, double* B, int ldb) {
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < m; i++) {
      AI(B, ldb, i, j) = AI(A, lda, i, j);
      AI(A, lda, i, j) = 0.0;
    }
}

/* ------------------------------------------------------------------ */
/* BSOF/I, following src/bsofi_cpu.c call for call                     */
/* ------------------------------------------------------------------ */
#define BLK(A, lda, n, i, j) ((A) + (size_t)(n) * ((size_t)(j) * (size_t)(lda) + (size_t)(i)))

/* lapackXbsof, src/bsofi_cpu.c:86-118 */
int oracle_bsof(int n, int L, double* A, int lda, double* tau0, double* work,
                int lwork, int* info) {
  if (ensure_backend()) return -1000;
  double* tau = tau0;
  *info = 0;
  if (lwork == -1) {
    double max_work0;
    xgeqrf(2 * n, 2 * n, A, lda, tau, work, lwork, info);
    max_work0 = work[0];
    xormqr('L', 'T', 2 * n, n, n, A, lda, tau, A, lda, work, lwork, info);
    if (max_work0 > work[0]) work[0] = max_work0;
    return *info;
  }
  for (int i = 0; i < L - 2; i++) {
    double* Rii = BLK(A, lda, n, i, i);
    double* Riip1 = BLK(A, lda, n, i, i + 1);
    double* Rin = BLK(A, lda, n, i, L - 1);
    xgeqrf(2 * n, n, Rii, lda, tau, work, lwork, info);
    xormqr('L', 'T', 2 * n, n, n, Rii, lda, tau, Riip1, lda, work, lwork, info);
    xormqr('L', 'T', 2 * n, n, n, Rii, lda, tau, Rin, lda, work, lwork, info);
    tau += n;
  }
  xgeqrf(2 * n, 2 * n, BLK(A, lda, n, L - 2, L - 2), lda, tau, work, lwork, info);
  return *info;
}

/* lapackXbstri, src/bsofi_cpu.c:146-177 (row version). */
int oracle_bstri(int n, int L, double* A, int lda, int* info) {
  if (ensure_backend()) return -1000;
  int N = n * L;
  *info = 0;
  if (L <= 2) {
    xtrtri('U', 'N', N, BLK(A, lda, n, 0, 0), lda, info);
  } else {
    xtrtri('U', 'N', 3 * n, BLK(A, lda, n, L - 3, L - 3), lda, info);
    if (L > 3) {
      xtrmm('R', 'U', 'N', 'N', n * (L - 3), n, 1.0, BLK(A, lda, n, L - 1, L - 1),
            lda, BLK(A, lda, n, 0, L - 1), lda);
      for (int i = L - 4; i >= 0; i--) {
        xtrtri('U', 'N', n, BLK(A, lda, n, i, i), lda, info);
        xtrmm('L', 'U', 'N', 'N', n, n, -1.0, BLK(A, lda, n, i, i), lda,
              BLK(A, lda, n, i, L - 1), lda);
        xtrmm('L', 'U', 'N', 'N', n, n, -1.0, BLK(A, lda, n, i, i), lda,
              BLK(A, lda, n, i, i + 1), lda);
        xgemm('N', 'N', n, n * (L - i - 2), n, 1.0, BLK(A, lda, n, i, i + 1), lda,
              BLK(A, lda, n, i + 1, i + 2), lda, 1.0, BLK(A, lda, n, i, i + 2), lda);
        xtrmm('R', 'U', 'N', 'N', n, n, 1.0, BLK(A, lda, n, i + 1, i + 1), lda,
              BLK(A, lda, n, i, i + 1), lda);
      }
    }
  }
  return *info;
}

/* lapackXbsoi_ormqr, src/bsofi_cpu.c:211-248 (the default lapackXbsoi,
 * bsofi.h:36-37). */
int oracle_bsoi_ormqr(int n, int L, double* A, int lda, double* tau0,
                      double* work, int lwork, int* info) {
  if (ensure_backend()) return -1000;
  int N = n * L, ldq = 2 * n;
  double* R1i = A;
  double* tau = tau0 + (size_t)(L - 2) * n;
  int lworkX = lwork - 4 * n * n;
  double* Q = work;
  double* workX = work + 4 * (size_t)n * n;
  *info = 0;
  if (lwork == -1) {
    xormqr('R', 'T', N, 2 * n, 2 * n, Q, ldq, tau, R1i, lda, work, -1, info);
    work[0] += 4 * n * n;
    return *info;
  }
  R1i = BLK(A, lda, n, 0, L - 2);
  cpyzero_lower(2 * n, 2 * n, BLK(A, lda, n, L - 2, L - 2), lda, Q, ldq);
  xormqr('R', 'T', N, 2 * n, 2 * n, Q, ldq, tau, R1i, lda, workX, lworkX, info);
  for (int i = L - 3; i >= 0; i--) {
    R1i = BLK(A, lda, n, 0, i);
    tau -= n;
    cpyzero_lower(2 * n, n, BLK(A, lda, n, i, i), lda, Q, ldq);
    xormqr('R', 'T', N, 2 * n, n, Q, ldq, tau, R1i, lda, workX, lworkX, info);
  }
  return *info;
}

/* lapackXbsoi_orgqr, src/bsofi_cpu.c:283-333 (explicit Q_k + GEMMs). */
int oracle_bsoi_orgqr(int n, int L, double* A, int lda, double* tau0,
                      double* work, int lwork, int* info) {
  if (ensure_backend()) return -1000;
  int N = n * L, ldq = 2 * n;
  double* tau = tau0 + (size_t)(L - 2) * n;
  int lworkX = lwork;
  double* workX = work;
  double* Q = workX;
  lworkX -= 2 * n * ldq;
  workX += (size_t)2 * n * ldq;
  double* hX = workX;
  lworkX -= 2 * n * lda;
  workX += (size_t)2 * n * lda;
  *info = 0;
  if (lwork == -1) {
    xorgqr(2 * n, 2 * n, 2 * n, Q, ldq, tau, work, -1, info);
    work[0] += 2 * n * (ldq + lda);
    return *info;
  }
  cpyzero_lower(2 * n, 2 * n, BLK(A, lda, n, L - 2, L - 2), lda, Q, ldq);
  xorgqr(2 * n, 2 * n, 2 * n, Q, ldq, tau, workX, lworkX, info);
  xlacpy('A', N, 2 * n, BLK(A, lda, n, 0, L - 2), lda, hX, lda);
  xgemm('N', 'T', N, 2 * n, 2 * n, 1.0, hX, lda, Q, ldq, 0.0,
        BLK(A, lda, n, 0, L - 2), lda);
  for (int i = L - 3; i >= 0; i--) {
    tau -= n;
    cpyzero_lower(2 * n, n, BLK(A, lda, n, i, i), lda, Q, ldq);
    xorgqr(2 * n, 2 * n, n, Q, ldq, tau, workX, lworkX, info);
    xlacpy('A', (i + 2) * n, n, BLK(A, lda, n, 0, i), lda, hX, lda);
    xlacpy('A', N, n, BLK(A, lda, n, 0, i + 1), lda, hX + (size_t)n * lda, lda);
    xgemm('N', 'T', (i + 1) * n, 2 * n, 2 * n, 1.0, hX, lda, Q, ldq, 0.0,
          BLK(A, lda, n, 0, i), lda);
    xgemm('N', 'T', N - (i + 1) * n, 2 * n, n, 1.0,
          hX + (size_t)n * lda + (size_t)(i + 1) * n, lda, Q + (size_t)n * ldq, ldq,
          0.0, BLK(A, lda, n, i + 1, i), lda);
  }
  return *info;
}

/* lapackXbsofiLWork, src/bsofi_cpu.c:31-42 (with lapackXbsoi = _ormqr as
 * bsofi.h:36-37 sets it for every caller that includes bsofi.h). */
int oracle_bsofi_lwork(int n, int L, int lda) {
  double work[1];
  int info, total;
  double* A = NULL;
  double* tau = NULL;
  oracle_bsof(n, L, A, lda, tau, work, -1, &info);
  total = (int)work[0];
  oracle_bsoi_ormqr(n, L, A, lda, tau, work, -1, &info);
  if (total < work[0]) total = (int)work[0];
  return total;
}

/* Full inversion in place: bsof -> bstri -> bsoi (ormqr or orgqr), the
 * sequence of examples/ftest_cpu/ftest_cpu.f90:133-140.  Returns the info of
 * the first failing phase.  times[3] (optional) receive per-phase wall time. */
int oracle_invert(int n, int L, double* A, int lda, double* tau, int use_orgqr,
                  double* times) {
  if (ensure_backend()) return -1000;
  int info = 0;
  int lw = oracle_bsofi_lwork(n, L, lda);
  double wq[1];
  oracle_bsoi_orgqr(n, L, NULL, lda, NULL, wq, -1, &info);
  if (use_orgqr && wq[0] > lw) lw = (int)wq[0];
  if (lw < 1) lw = 1;
  double* work = (double*)malloc(sizeof(double) * (size_t)lw);
  if (!work) return -1001;
  struct timespec t0, t1, t2, t3;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  oracle_bsof(n, L, A, lda, tau, work, lw, &info);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (info == 0) oracle_bstri(n, L, A, lda, &info);
  clock_gettime(CLOCK_MONOTONIC, &t2);
  if (info == 0) {
    if (use_orgqr)
      oracle_bsoi_orgqr(n, L, A, lda, tau, work, lw, &info);
    else
      oracle_bsoi_ormqr(n, L, A, lda, tau, work, lw, &info);
  }
  clock_gettime(CLOCK_MONOTONIC, &t3);
  if (times) {
    times[0] = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);
    times[1] = (t2.tv_sec - t1.tv_sec) + 1e-9 * (t2.tv_nsec - t1.tv_nsec);
    times[2] = (t3.tv_sec - t2.tv_sec) + 1e-9 * (t3.tv_nsec - t2.tv_nsec);
  }
  free(work);
  return info;
}

/* Relative residual ||A*G - I||_F / (||A||_F ||G||_F) with A given dense
 * (examples/ftest_cpu/ftest_cpu.f90:53-100), computed in long double to be
 * independent of the back end. */
double oracle_residual_dense(int N, const double* A, int lda, const double* G,
                             int ldg) {
  long double na = 0, ng = 0, nr = 0;
  long double* col = (long double*)malloc(sizeof(long double) * (size_t)N);
  for (int j = 0; j < N; j++) {
    for (int i = 0; i < N; i++) col[i] = 0;
    for (int l = 0; l < N; l++) {
      long double g = AI(G, ldg, l, j);
      if (g != 0)
        for (int i = 0; i < N; i++) col[i] += (long double)AI(A, lda, i, l) * g;
    }
    col[j] -= 1;
    for (int i = 0; i < N; i++) nr += col[i] * col[i];
  }
  for (int j = 0; j < N; j++)
    for (int i = 0; i < N; i++) {
      na += (long double)AI(A, lda, i, j) * AI(A, lda, i, j);
      ng += (long double)AI(G, ldg, i, j) * AI(G, ldg, i, j);
    }
  free(col);
  return (double)(sqrtl(nr) / (sqrtl(na) * sqrtl(ng)));
}

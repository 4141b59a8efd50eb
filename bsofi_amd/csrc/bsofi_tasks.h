// Task descriptors shared by the host planner and the gfx950 kernels.
//
// The planner (bsofi_engine.cpp) turns a batch of p-cyclic matrices into an
// ordered list of launches.  Every launch reads a table of task records and
// a table of per-workgroup records {task, sub} from one device blob, so a
// launch can mix matrices of different n and p, and workgroups can be
// ordered heaviest-first.  All matrices are FP64, column-major.
#pragma once
#include <stdint.h>

namespace bsofi {

// Householder QR of one m x nbj sub-panel (columns of one BSOF panel),
// LAPACK dgeqrf conventions (dlarfg sign rule, unit-lower reflectors stored
// below the diagonal, tau per column), plus the nbj x nbj upper triangular
// T of the compact WY form H_1...H_nbj = I - V T V^T.
struct PanelTask {
  double* P;      // top-left of the sub-panel in A (its diagonal corner)
  double* tau;    // tau for the sub-panel's first column
  double* T;      // T factor out, column-major, leading dimension ldt
  int* info;      // per-matrix status word (unused by the panel today)
  int lda, m, nbj, ldt;
};

// Block reflector application: C <- op(I - V T V^T) C over the m rows of the
// reflectors, for the column set {C1 .. C1+w1} U {C2 .. C2+w2}.
// trans = 1 applies (I - V T V^T)^T = I - V T^T V^T (used by BSOF, Q_k^T
// from the left); trans = 0 applies I - V T V^T (used to build Q_k).
struct ReflTask {
  const double* V;  // reflector block top-left (unit diagonal implicit)
  const double* T;
  double* C1;
  double* C2;
  int ldv, ldt, ldc;
  int m, nbj, w1, w2, trans;
};

// Explicit Q_k = H_0 H_1 ... H_{S-1} (dorgqr, backward accumulation) of one
// BSOF step from its sub-panels' block reflectors: sub-panel j has
// V_j = V0 + j nb (ldv + 1) (m = n2 - j nb rows, nbj = min(nb, ncols - j nb)
// columns) and T_j = T0 + j nb nb (leading dimension nb); Q is n2 x n2 with
// leading dimension n2.  A workgroup builds one 16-column slab of Q from
// the identity, every H_j that reaches it applied in registers.
struct QformTask {
  const double* V0;
  const double* T0;
  double* Q;
  int ldv, n2, nb, S, ncols;
};

// Batched tile GEMM: C = alpha*A*B + beta*C (+ D on columns >= dcol0).
// Triangular operands are read through masks: amask = 1 treats A as upper
// triangular (A[i][k] = 0 for i > k); bmask = 1 treats B as upper
// triangular (B[k][j] = 0 for k > j).  Both stay valid for the lower part
// of the array holding other data (the Householder vectors).
struct GemmTask {
  const double* A;
  const double* B;
  double* C;
  const double* D;
  double alpha, beta;
  int lda, ldb, ldc, ldd;
  int M, N, K;
  int amask, bmask, dcol0, tiles_n;
};

// BSTRI's row sweep for block column b >= 1 of one matrix as independent
// 32-column strips (Item::sub = strip; the last one partial where 32 does
// not divide n): X_{i,b} = W_i X_{i+1,b} (+ Z2_i for
// b = p - 1, i <= p - 3) for i = b - 1 .. 0, from the diagonal block X_{b,b}
// (upper triangular; A holds reflectors below its diagonal).  X_{i,b} is
// A + (b n) lda + i n; W_i is W + i n n and Z2_i is Z2 + i n n (leading
// dimension n).  A task with b = 0 is empty.
struct SweepTask {
  double* A;
  const double* W;
  const double* Z2;
  int lda, n, p, b;
};

// Inversion of one diagonal sub-block (<= 32 x 32) of an upper triangular
// block in place.  gidx0 is the 1-based global column index of the
// sub-block's first diagonal element, reported through *info (atomicMin) if
// a diagonal entry is exactly zero.
struct TrtriTask {
  double* X;
  int* info;
  int ld, nb, gidx0, pad;
};

// One row tile of the BSOI scan: rows [r0, r0 + rows) of matrix A (N = n*p)
// computed as G = X Q_{p-2}^T ... Q_0^T with X = R^{-1} read in place and
// the explicit Q_k (2n x 2n each, stride qstride) from the workspace.
struct ScanTask {
  double* A;
  const double* Q;
  long long qstride;
  int lda, n, p, r0, rows;
  int diag_only;  // 1: produce only the diagonal blocks G_ll
};

// Identity fill of count square matrices of order dim, stride apart.
struct EyeTask {
  double* Q;
  long long stride;
  int dim, count;
};

// sign(det A) and log|det A| of one factored matrix (after BSOF): R's
// diagonal A[i + i lda] and tau[i], i < N; out[0] = sign, out[1] = log|det|.
struct DetTask {
  const double* A;
  const double* tau;
  double* out;
  int lda, N;
};

// Device-scheduled BSOF sweep (bsof_flow_kernel).  One persistent launch
// runs every sub-panel QR (kind 0) and every application of a sub-panel's
// reflectors to one 16-column slab (kind 1) of all matrices of a plan;
// workgroups take tasks from a device queue in a dependency-respecting order
// and wait, with a time limit, for what each task reads:
//   * per matrix, one counter per column of A: how many sub-panels'
//     reflectors have been applied to that column so far;
//   * per matrix, one flag per sub-panel: its QR (V, T, tau) is stored.
// A task on column c of sub-panel index g (step-major order of the BSOF
// sweep) needs the counter of c to have reached the number of sub-panels
// before g whose reflectors touch c, a closed form of (n, p, nb, c, g).
struct FlowMat {
  int* ctr;   // N column counters
  int* done;  // one flag per sub-panel
  int n, p, nb;
  int s0;     // sub-panels per step in steps 0 .. p-3 (ceil(n / nb))
};

struct FlowTask {
  int kind;  // 0: sub-panel QR, 1: reflector slab
  int mat;   // FlowMat index
  int g;     // sub-panel index within its matrix
  int ref;   // PanelTask (kind 0) or ReflTask (kind 1) index
  int sub;   // kind 1: 16-column slab of the ReflTask
  int col0;  // kind 0: first column of the sub-panel; kind 1: column of C1's first column
  int col2;  // kind 1: column of C2's first column
  int pad;
};

// Work-item record: which task a workgroup serves and which piece of it.
struct Item {
  int task;
  int sub;
};

}  // namespace bsofi

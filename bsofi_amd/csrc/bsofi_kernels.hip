// gfx950 (CDNA4) kernels for BSOF/I.  FP64 throughout.
//
// Kernel map (reference call site each one replaces):
//   panel_factor_kernel  dgeqrf of the stacked 2n x n pair, one sub-panel
//                        at a time (src/bsofi_cpu.c:107, :115)
//   refl_apply_kernel    dormqr('L','T') of Q_k^T onto the neighbouring and
//                        last block columns (src/bsofi_cpu.c:108-111), and
//                        dorgqr's backward accumulation of the explicit Q_k
//                        (src/bsofi_cpu.c:308, :323)
//   gemm_kernel          the TRMM/GEMM calls of BSTRI (src/bsofi_cpu.c:160-167)
//                        and the blocked TRTRI updates
//   trtri_diag_kernel    dtrti2 of <= 32 x 32 diagonal sub-blocks
//                        (src/bsofi_cpu.c:155-162)
//   bsoi_scan_kernel     G = X Q_{p-2}^T ... Q_0^T (src/bsofi_cpu.c:283-333)
//                        as one row-panel scan per workgroup
//   dqmc_build_kernel    the p-cyclic DQMC matrix from e^{-dtau K} and HS
//                        fields (replaces include/bench/bsofi_random.h)
//
// MFMA: v_mfma_f64_16x16x4_f64.  Operand layouts (64-lane wave, lane l):
//   A (16 x 4):  lane l holds A[l % 16][l / 16]
//   B (4 x 16):  lane l holds B[l / 16][l % 16]
//   C (16 x 16): accumulator register r of lane l holds C[4 * r + l / 16][l % 16]
//                (crow() below; checked on the device by the GEMM tests)
#include <hip/hip_runtime.h>
#include <float.h>
#include <math.h>

#include "bsofi_tasks.h"

namespace bsofi {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Row, inside a 16 x 16 accumulator tile, of register r of a lane whose
// lane / 16 is lk.
__device__ __forceinline__ int crow(int lk, int r) { return 4 * r + lk; }

__device__ __forceinline__ d4 d4zero() {
  d4 z = {0.0, 0.0, 0.0, 0.0};
  return z;
}

// ---------------------------------------------------------------------------
// Sub-panel Householder QR (K1)
// ---------------------------------------------------------------------------
// 256 threads.  Lane layout: cg = lane & 3 is the column group (CPG columns
// each), rg = (lane >> 2) + 16 * wave is the row group; thread (rg, cg)
// holds rows rg + 64 q (q < RPG) of its CPG columns in registers.  Per
// column c one fused reduction produces, for every column j of the
// sub-panel, sum_{r > c} x_r a_rj with x = column c below the pivot:
//   j == c : ||x||^2          (dlarfg)
//   j >  c : the dot products of H_c's application
//   j <  c : v_j^T v_c        (T factor, dlarft)
// so each column costs one block-wide reduction and one barrier.
// Reduce v[0..CPG) over the 16 lanes sharing lane & 3 (xor 32, 16, 8, 4).
// Each stage halves the entries a lane keeps ("transpose reduce"), so CPG
// values cost CPG - 1 + (stages left) shuffles instead of 4 * CPG.  On
// return the lane holds the wave sum of entry idx.  Stage sizes are
// template arguments so every array index is a compile-time constant.
template <int CNT, int MASK, int CPG>
__device__ __forceinline__ void transpose_stage(double (&v)[CPG], int lane, int& idx) {
  if constexpr (MASK >= 4) {
    if constexpr (CNT >= 2) {
      constexpr int half = CNT / 2;
      const bool bit = (lane & MASK) != 0;
#pragma unroll
      for (int t = 0; t < half; ++t) {
        const double send = bit ? v[t] : v[t + half];
        const double keep = bit ? v[t + half] : v[t];
        v[t] = keep + __shfl_xor(send, MASK);
      }
      if (bit) idx += half;
      transpose_stage<half, MASK / 2, CPG>(v, lane, idx);
    } else {
      v[0] += __shfl_xor(v[0], MASK);
      transpose_stage<CNT, MASK / 2, CPG>(v, lane, idx);
    }
  }
}

template <int CPG>
__device__ __forceinline__ double wave_transpose_reduce(double (&v)[CPG], int lane, int& idx) {
  idx = 0;
  transpose_stage<CPG, 32, CPG>(v, lane, idx);
  return v[0];
}

__device__ double block_sum_256(double v, double* scratch) {
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  const double r = scratch[0] + scratch[1] + scratch[2] + scratch[3];
  __syncthreads();
  return r;
}

__device__ double block_max_256(double v, double* scratch) {
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) scratch[wave] = v;
  __syncthreads();
  const double r = fmax(fmax(scratch[0], scratch[1]), fmax(scratch[2], scratch[3]));
  __syncthreads();
  return r;
}

// T of the compact WY form (dlarft, forward, columnwise) from the Gram
// entries G[j][c] = v_j^T v_c (j < c) and tau; lane i owns row i of T:
//   T[c][c] = tau_c,  T[i][c] = -tau_c * sum_{j=i}^{c-1} T[i][j] G[j][c].
// Row i is kept in registers (the loops are unrolled), so the only memory
// traffic is the broadcast read of G.
template <int NB>
__device__ void build_T(const double (*G)[NB + 1], const double* taus, int nbj, double* __restrict__ Tout, int ldt,
                        int lane) {
  if (lane >= NB) return;
  const int i = lane;
  double Trow[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) Trow[j] = 0.0;
  for (int c = 0; c < nbj; ++c) {
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j < c) s = fma(Trow[j], G[j][c], s);  // Trow[j] = 0 for j < i
    const double v = (i < c) ? -taus[c] * s : (i == c ? taus[c] : 0.0);
#pragma unroll
    for (int j = 0; j < NB; ++j)
      if (j == c) Trow[j] = v;
  }
#pragma unroll
  for (int c = 0; c < NB; ++c)
    if (c < nbj) Tout[(size_t)c * ldt + i] = (i < nbj) ? Trow[c] : 0.0;
}

// Broadcast lane (lane & ~3) | cc of each quad to the whole quad (DPP).
template <int CC>
__device__ __forceinline__ double quad_bcast_c(double v) {
  constexpr int ctrl = CC | (CC << 2) | (CC << 4) | (CC << 6);
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), ctrl, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), ctrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ double quad_bcast(double v, int cc) {
  switch (cc) {
    case 0: return quad_bcast_c<0>(v);
    case 1: return quad_bcast_c<1>(v);
    case 2: return quad_bcast_c<2>(v);
    default: return quad_bcast_c<3>(v);
  }
}

// ||x|| over the block with scaling by max|x| (two block reductions); the
// rare path of the panel's dlarfg, kept out of line so the unrolled column
// loop stays small enough for the instruction cache.
template <int RPG>
__device__ __noinline__ double scaled_norm_256(const double* x, double* scratch) {
  double lm = 0.0;
  for (int q = 0; q < RPG; ++q) lm = fmax(lm, fabs(x[q]));
  const double amax = block_max_256(lm, scratch);
  double ls = 0.0;
  if (amax > 0.0) {
    for (int q = 0; q < RPG; ++q) {
      const double y = x[q] / amax;
      ls = fma(y, y, ls);
    }
  }
  const double s2 = block_sum_256(ls, scratch);
  return amax > 0.0 ? amax * sqrt(s2) : 0.0;
}

// dlarfg from alpha and ||x||: beta, tau and the scale that turns x into
// the stored reflector, with LAPACK's safe-minimum rescaling loop.
__device__ __noinline__ void dlarfg_scalar(double alpha, double xnorm, double* tau_o, double* beta_o,
                                           double* scale_o) {
  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  double tau, beta, scale;
  if (xnorm == 0.0) {
    tau = 0.0;
    beta = alpha;
    scale = 1.0;
  } else {
    beta = -copysign(hypot(alpha, xnorm), alpha);
    double al = alpha, sc_acc = 1.0;
    int knt = 0;
    if (fabs(beta) < safmin) {
      const double rsafmn = 1.0 / safmin;
      do {
        knt++;
        sc_acc *= rsafmn;
        beta *= rsafmn;
        al *= rsafmn;
      } while (fabs(beta) < safmin && knt < 20);
      beta = -copysign(hypot(al, xnorm * sc_acc), al);
    }
    tau = (beta - al) / beta;
    scale = sc_acc / (al - beta);
    for (int q2 = 0; q2 < knt; ++q2) beta *= safmin;
  }
  *tau_o = tau;
  *beta_o = beta;
  *scale_o = scale;
}

template <int RPG, int CPG>
__global__ __launch_bounds__(256) void panel_factor_kernel(const PanelTask* __restrict__ tasks,
                                                           const Item* __restrict__ items) {
  constexpr int NB = 4 * CPG;
  const Item it = items[blockIdx.x];
  const PanelTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cg = lane & 3;
  const int rg = (lane >> 2) + 16 * wave;  // rows rg + 64 q; only q = 0 rows can lie at or above a pivot (c < 32)
  const int m = t.m, nbj = t.nbj, lda = t.lda;

  __shared__ double red[2][NB][4];  // [buffer][column][wave] partial sums
  __shared__ double prow[2][NB];
  __shared__ double G[NB][NB + 1];
  __shared__ double taus[NB];
  __shared__ double scratch[4];

  double a[RPG][CPG];
#pragma unroll
  for (int q = 0; q < RPG; ++q) {
    const int row = rg + 64 * q;
#pragma unroll
    for (int jj = 0; jj < CPG; ++jj) {
      const int col = cg * CPG + jj;
      a[q][jj] = (row < m && col < nbj) ? t.P[(size_t)col * lda + row] : 0.0;
    }
  }

  for (int cc = 0; cc < 4; ++cc) {
#pragma unroll
    for (int jc = 0; jc < CPG; ++jc) {
      const int c = cc * CPG + jc;
      if (c < nbj) {
        const int buf = c & 1;
        // x = column c below the pivot at my rows (zero at and above it)
        double x[RPG];
#pragma unroll
        for (int q = 0; q < RPG; ++q) x[q] = __shfl(a[q][jc], (lane & ~3) | cc);
        x[0] = (rg > c) ? x[0] : 0.0;
        double P[CPG];
#pragma unroll
        for (int jj = 0; jj < CPG; ++jj) {
          double acc = x[0] * a[0][jj];
#pragma unroll
          for (int q = 1; q < RPG; ++q) acc = fma(x[q], a[q][jj], acc);
          P[jj] = acc;
        }
        int idx;
        const double ws = wave_transpose_reduce<CPG>(P, lane, idx);
        if ((lane & 4) == 0) red[buf][cg * CPG + idx][wave] = ws;
        if (rg == c) {
#pragma unroll
          for (int jj = 0; jj < CPG; ++jj) prow[buf][cg * CPG + jj] = a[0][jj];
        }
        __syncthreads();

        const double xn2 = (red[buf][c][0] + red[buf][c][1]) + (red[buf][c][2] + red[buf][c][3]);
        const double alpha = prow[buf][c];

        // dlarfg.  The fast path takes ||x|| = sqrt(sum x^2); a sum of
        // squares outside [2^-960, 2^960] is recomputed with scaling.
        const double xnorm = (!(xn2 > 0x1p-960) || !(xn2 < 0x1p960)) ? scaled_norm_256<RPG>(x, scratch) : sqrt(xn2);
        double tau, beta, scale;
        dlarfg_scalar(alpha, xnorm, &tau, &beta, &scale);

        // w_j = tau v_c^T a_j (j > c); G[j][c] = v_j^T v_c (j < c)
        double w[CPG];
#pragma unroll
        for (int jj = 0; jj < CPG; ++jj) {
          const int j = cg * CPG + jj;
          const double tot = (red[buf][j][0] + red[buf][j][1]) + (red[buf][j][2] + red[buf][j][3]);
          const double d = fma(scale, tot, prow[buf][j]);
          w[jj] = (j > c) ? tau * d : 0.0;
          if (j < c && rg == 0) G[j][c] = d;
        }
        // a <- (I - tau v v^T) a with v = 0 above the pivot, 1 at it, x*scale below
        const double v0 = (rg > c) ? x[0] * scale : (rg == c ? 1.0 : 0.0);
#pragma unroll
        for (int jj = 0; jj < CPG; ++jj) a[0][jj] = fma(-v0, w[jj], a[0][jj]);
#pragma unroll
        for (int q = 1; q < RPG; ++q) {
          const double vq = x[q] * scale;
#pragma unroll
          for (int jj = 0; jj < CPG; ++jj) a[q][jj] = fma(-vq, w[jj], a[q][jj]);
        }
        if (cg == cc) {
          a[0][jc] = (rg > c) ? v0 : (rg == c ? beta : a[0][jc]);
#pragma unroll
          for (int q = 1; q < RPG; ++q) a[q][jc] = x[q] * scale;
        }
        if (tid == 0) {
          t.tau[c] = tau;
          taus[c] = tau;
        }
      }
    }
  }

#pragma unroll
  for (int q = 0; q < RPG; ++q) {
    const int row = rg + 64 * q;
#pragma unroll
    for (int jj = 0; jj < CPG; ++jj) {
      const int col = cg * CPG + jj;
      if (row < m && col < nbj) t.P[(size_t)col * lda + row] = a[q][jj];
    }
  }
  __syncthreads();
  if (wave == 0) build_T<NB>(G, taus, nbj, t.T, t.ldt, lane);
}

// ---------------------------------------------------------------------------
// Block reflector application (K2)
// ---------------------------------------------------------------------------
// One workgroup = one 16-column slab of the target columns, all m rows.
//   phase 1: W = V^T C        (nbj x 16, K = m split over the 4 waves)
//   phase 2: W <- -op(T) W    (LDS, VALU)
//   phase 3: C <- C + V W     (m x 16, K = nbj)
__device__ __forceinline__ double vload(const double* __restrict__ V, int ldv, int m, int nbj, int row, int col) {
  if (row >= m || col >= nbj || row < col) return 0.0;
  if (row == col) return 1.0;
  return V[(size_t)col * ldv + row];
}

// The target slab lives in MFMA accumulator registers for the whole kernel:
// wave w owns row tiles w, w+4, ... (RT of them, m <= 64 RT).  The C layout
// of a 16 x 16 accumulator tile is exactly four consecutive B fragments of
// a K = 16 block (register r = rows 4r..4r+3), so phase 1 multiplies
// straight out of the accumulators and only V is read from memory.
template <int RT>
__global__ __launch_bounds__(256) void refl_apply_kernel(const ReflTask* __restrict__ tasks,
                                                         const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const ReflTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m = t.m, nbj = t.nbj;

  __shared__ double Wp[4][32][17];
  __shared__ double Wf[32][17];
  __shared__ double Tl[32][33];

  // the column this lane serves in B / C fragments
  const int vcol = it.sub * 16 + li;
  double* cptr = nullptr;
  if (vcol < t.w1)
    cptr = t.C1 + (size_t)vcol * t.ldc;
  else if (vcol - t.w1 < t.w2)
    cptr = t.C2 + (size_t)(vcol - t.w1) * t.ldc;

  for (int e = tid; e < 32 * 32; e += 256) {
    const int i = e & 31, j = e >> 5;
    Tl[i][j] = (i < nbj && j < nbj) ? t.T[(size_t)j * t.ldt + i] : 0.0;
  }
  d4 acc[RT];
#pragma unroll
  for (int s = 0; s < RT; ++s) {
    const int rt = wave + 4 * s;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = rt * 16 + crow(lk, r);
      acc[s][r] = (cptr && row < m) ? cptr[row] : 0.0;
    }
  }
  // phase 1: W = V^T C
  const bool two = nbj > 16;
  d4 w0 = d4zero(), w1 = d4zero();
#pragma unroll
  for (int s = 0; s < RT; ++s) {
    const int rt = wave + 4 * s;
    if (rt * 16 < m) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + 4 * r + lk;
        const double a0 = vload(t.V, t.ldv, m, nbj, row, li);
        const double a1 = two ? vload(t.V, t.ldv, m, nbj, row, 16 + li) : 0.0;
        w0 = mfma(a0, acc[s][r], w0);
        if (two) w1 = mfma(a1, acc[s][r], w1);
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Wp[wave][crow(lk, r)][li] = w0[r];
    Wp[wave][16 + crow(lk, r)][li] = two ? w1[r] : 0.0;
  }
  __syncthreads();
  // phase 2: Wf = -op(T) W
  __shared__ double W[32][17];
  for (int e = tid; e < 32 * 16; e += 256) {
    const int i = e >> 4, j = e & 15;
    W[i][j] = (Wp[0][i][j] + Wp[1][i][j]) + (Wp[2][i][j] + Wp[3][i][j]);
  }
  __syncthreads();
  for (int e = tid; e < 32 * 16; e += 256) {
    const int i = e >> 4, j = e & 15;
    double s2 = 0.0;
    if (i < nbj) {
      if (t.trans) {  // T^T: sum_{k <= i} T[k][i] W[k][j]
#pragma unroll 4
        for (int k = 0; k <= i; ++k) s2 = fma(Tl[k][i], W[k][j], s2);
      } else {        // T: sum_{k >= i} T[i][k] W[k][j]
#pragma unroll 4
        for (int k = i; k < nbj; ++k) s2 = fma(Tl[i][k], W[k][j], s2);
      }
    }
    Wf[i][j] = -s2;
  }
  __syncthreads();
  // phase 3: C += V Wf
#pragma unroll
  for (int s = 0; s < RT; ++s) {
    const int rt = wave + 4 * s;
    if (rt * 16 < m) {
#pragma unroll
      for (int k0 = 0; k0 < 32; k0 += 4) {
        if (k0 < nbj) {
          const double av = vload(t.V, t.ldv, m, nbj, rt * 16 + li, k0 + lk);
          acc[s] = mfma(av, Wf[k0 + lk][li], acc[s]);
        }
      }
    }
  }
  if (cptr) {
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      const int rt = wave + 4 * s;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + crow(lk, r);
        if (row < m) cptr[row] = acc[s][r];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// T factor from stored reflectors and tau (dlarft), one workgroup per
// sub-panel: G = V^T V on MFMA (K = m split over the waves), then the
// recurrence of build_T.  Used by BSOI, whose inputs are only A (holding V)
// and tau, as in the reference's lapackXbsoi.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void larft_kernel(const PanelTask* __restrict__ tasks,
                                                    const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const PanelTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m = t.m, nbj = t.nbj;
  __shared__ double Gp[4][32][33];
  __shared__ double G[32][33];
  __shared__ double taus[32];
  const int per = ((m + 15) / 16) * 4;
  const int kb = wave * per, ke = min(m, kb + per);
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4zero();
  for (int k0 = kb; k0 < ke; k0 += 4) {
    const int row = k0 + lk;
    double f[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) f[a] = (row < ke) ? vload(t.P, t.lda, m, nbj, row, a * 16 + li) : 0.0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = mfma(f[a], f[b], acc[a][b]);
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) Gp[wave][a * 16 + crow(lk, r)][b * 16 + li] = acc[a][b][r];
  if (tid < 32) taus[tid] = (tid < nbj) ? t.tau[tid] : 0.0;
  __syncthreads();
  for (int e = tid; e < 32 * 32; e += 256) {
    const int i = e >> 5, j = e & 31;
    G[i][j] = (Gp[0][i][j] + Gp[1][i][j]) + (Gp[2][i][j] + Gp[3][i][j]);
  }
  __syncthreads();
  if (wave == 0) build_T<32>(G, taus, nbj, t.T, t.ldt, lane);
}

__global__ void fill_int_kernel(int* __restrict__ p, int count, int value) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) p[e] = value;
}

// ---------------------------------------------------------------------------
// Tile GEMM (K4): 64 x 64 output tile per workgroup, K staged through LDS in
// chunks of 16, 2 x 2 waves of 32 x 32 (2 x 2 MFMA tiles each).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void gemm_kernel(const GemmTask* __restrict__ tasks,
                                                   const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const GemmTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int tm = it.sub / t.tiles_n, tn = it.sub % t.tiles_n;
  const int i0 = tm * 64, j0 = tn * 64;
  const int wi = wave & 1, wj = wave >> 1;

  __shared__ double As[64][17];
  __shared__ double Bs[64][17];

  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4zero();

  const int M = t.M, N = t.N, K = t.K;
  for (int k0 = 0; k0 < K; k0 += 16) {
    // A tile: 64 rows x 16 k
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int i = tid & 63, kk = (tid >> 6) + 4 * r;
      const int gi = i0 + i, gk = k0 + kk;
      double v = 0.0;
      if (gi < M && gk < K && !(t.amask && gi > gk)) v = t.A[(size_t)gk * t.lda + gi];
      As[i][kk] = v;
    }
    // B tile: 16 k x 64 cols
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int kk = tid & 15, j = (tid >> 4) + 16 * r;
      const int gk = k0 + kk, gj = j0 + j;
      double v = 0.0;
      if (gk < K && gj < N && !(t.bmask && gk > gj)) v = t.B[(size_t)gj * t.ldb + gk];
      Bs[j][kk] = v;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 16; ks += 4) {
      double af[2], bf[2];
#pragma unroll
      for (int a = 0; a < 2; ++a) af[a] = As[wi * 32 + a * 16 + li][ks + lk];
#pragma unroll
      for (int b = 0; b < 2; ++b) bf[b] = Bs[wj * 32 + b * 16 + li][ks + lk];
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = mfma(af[a], bf[b], acc[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = j0 + wj * 32 + b * 16 + li;
      if (col >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + wi * 32 + a * 16 + crow(lk, r);
        if (row >= M) continue;
        double v = t.alpha * acc[a][b][r];
        double* cp = t.C + (size_t)col * t.ldc + row;
        if (t.beta != 0.0) v = fma(t.beta, *cp, v);
        if (t.D && col >= t.dcol0) v += t.D[(size_t)(col - t.dcol0) * t.ldd + row];
        *cp = v;
      }
    }
}

// ---------------------------------------------------------------------------
// In-place inverse of a <= 32 x 32 upper triangular sub-block (dtrti2).
// 64 threads; the strictly lower part of the array is left untouched.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void trtri_diag_kernel(const TrtriTask* __restrict__ tasks) {
  const TrtriTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x, nb = t.nb;
  __shared__ double S[32][33];
  __shared__ double xn[32];
  __shared__ int bad;
  if (tid == 0) bad = 0;
  for (int e = tid; e < 32 * 32; e += 64) {
    const int i = e & 31, j = e >> 5;
    S[i][j] = (i <= j && i < nb && j < nb) ? t.X[(size_t)j * t.ld + i] : 0.0;
  }
  __syncthreads();
  if (tid < nb && S[tid][tid] == 0.0) {
    atomicMin(t.info, t.gidx0 + tid);
    atomicExch(&bad, 1);
  }
  __syncthreads();
  if (bad) return;
  for (int j = 0; j < nb; ++j) {
    if (tid == 0) S[j][j] = 1.0 / S[j][j];
    __syncthreads();
    const double ajj = -S[j][j];
    double v = 0.0;
    if (tid < j) {
      for (int l = tid; l < j; ++l) v = fma(S[tid][l], S[l][j], v);
      xn[tid] = v * ajj;
    }
    __syncthreads();
    if (tid < j) S[tid][j] = xn[tid];
    __syncthreads();
  }
  for (int e = tid; e < 32 * 32; e += 64) {
    const int i = e & 31, j = e >> 5;
    if (i <= j && j < nb) t.X[(size_t)j * t.ld + i] = S[i][j];
  }
}

// ---------------------------------------------------------------------------
// Identity fill of the explicit-Q buffers (a kernel rather than a memset so
// the launch sequence can be captured in a graph).
// ---------------------------------------------------------------------------
__global__ void eye_kernel(const EyeTask* __restrict__ tasks, int ntasks) {
  for (int ti = 0; ti < ntasks; ++ti) {
    const EyeTask t = tasks[ti];
    const long long per = (long long)t.dim * t.dim;
    const long long total = per * t.count;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (long long)gridDim.x * blockDim.x) {
      const long long mi = e / per, r = e - mi * per;
      const int j = (int)(r / t.dim), i = (int)(r - (long long)j * t.dim);
      t.Q[mi * t.stride + (long long)j * t.dim + i] = (i == j) ? 1.0 : 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// BSOI row-panel scan (K5).
// ---------------------------------------------------------------------------
// A workgroup owns MT = 16*MTI consecutive rows of G.  For k = p-2 .. 0 it
// computes [new_k | new_{k+1}] = [X_{r,k} | carry] Q_k^T on FP64 MFMA,
// stores new_{k+1} as the final G block column k+1 and keeps new_k (the
// carry, MT x n) in LDS.  Rows whose block row r exceeds k have X_{r,k} = 0,
// so only the carry half (K = n) is multiplied; the diagonal block X_{k,k}
// is read through an upper-triangular mask because A still holds the
// reflectors below its diagonal.  The 2n output columns are split in
// 16-column tiles dealt round-robin to the 4 waves (NTW tiles per wave).
template <int MTI, int NTW>
__global__ __launch_bounds__(256) void bsoi_scan_kernel(const ScanTask* __restrict__ tasks,
                                                        const Item* __restrict__ items, int ldcar) {
  constexpr int MT = 16 * MTI;
  extern __shared__ double carry[];  // carry[kk * ldcar + i], kk < n, i < MT
  const Item it = items[blockIdx.x];
  const ScanTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int n = t.n, p = t.p, lda = t.lda, N = n * p, n2 = 2 * n;
  const int R0 = t.r0 + it.sub * MT;
  double* __restrict__ A = t.A;

  // rows served by this lane in A fragments: R0 + a*16 + li
  int rrow[MTI], rblk[MTI], rin[MTI];
#pragma unroll
  for (int a = 0; a < MTI; ++a) {
    rrow[a] = R0 + a * 16 + li;
    rblk[a] = rrow[a] < N ? rrow[a] / n : p;
    rin[a] = rrow[a] - rblk[a] * n;
  }
  const int ntiles = (n2 + 15) / 16;

  for (int e = tid; e < n * MT; e += 256) {
    const int i = e % MT, kk = e / MT;
    const int row = R0 + i;
    double v = 0.0;
    if (row < N) {
      const int rb = row / n;
      if (!(rb == p - 1 && row - rb * n > kk)) v = A[(size_t)((p - 1) * n + kk) * lda + row];
    }
    carry[kk * ldcar + i] = v;
  }
  __syncthreads();

  const int rbmin = R0 / n;
  int bcol[NTW];
#pragma unroll
  for (int b = 0; b < NTW; ++b) bcol[b] = (wave + 4 * b) * 16 + li;
  for (int k = p - 2; k >= 0; --k) {
    const double* __restrict__ Qk = t.Q + (size_t)k * t.qstride;
    d4 acc[MTI][NTW];
#pragma unroll
    for (int a = 0; a < MTI; ++a)
#pragma unroll
      for (int b = 0; b < NTW; ++b) acc[a][b] = d4zero();

    // Fragments of the next k-step are loaded while the MFMAs of the
    // current one run (two register buffers, ping-pong).
    double a0[MTI], b0[NTW], a1[MTI], b1[NTW];
    if (rbmin <= k) {
      // [X_{r,k}] part: K rows 0..n-1 of Q_k^T
      const double* __restrict__ Xk = A + (size_t)k * n * lda;
      auto load = [&](double (&af)[MTI], double (&bf)[NTW], int k0) {
        const int kk = k0 + lk;
#pragma unroll
        for (int a = 0; a < MTI; ++a) {
          double v = 0.0;
          if (kk < n && rblk[a] <= k && !(rblk[a] == k && rin[a] > kk)) v = Xk[(size_t)kk * lda + rrow[a]];
          af[a] = v;
        }
#pragma unroll
        for (int b = 0; b < NTW; ++b) bf[b] = (kk < n && bcol[b] < n2) ? Qk[(size_t)kk * n2 + bcol[b]] : 0.0;
      };
      load(a0, b0, 0);
      for (int k0 = 0; k0 < n; k0 += 8) {
        if (k0 + 4 < n) load(a1, b1, k0 + 4);
#pragma unroll
        for (int a = 0; a < MTI; ++a)
#pragma unroll
          for (int b = 0; b < NTW; ++b) acc[a][b] = mfma(a0[a], b0[b], acc[a][b]);
        if (k0 + 8 < n) load(a0, b0, k0 + 8);
        if (k0 + 4 < n) {
#pragma unroll
          for (int a = 0; a < MTI; ++a)
#pragma unroll
            for (int b = 0; b < NTW; ++b) acc[a][b] = mfma(a1[a], b1[b], acc[a][b]);
        }
      }
    }
    // [carry] part: K rows n..2n-1 of Q_k^T
    {
      const double* __restrict__ Qc = Qk + (size_t)n * n2;
      auto load = [&](double (&af)[MTI], double (&bf)[NTW], int k0) {
        const int kk = k0 + lk;
#pragma unroll
        for (int a = 0; a < MTI; ++a) af[a] = (kk < n) ? carry[kk * ldcar + a * 16 + li] : 0.0;
#pragma unroll
        for (int b = 0; b < NTW; ++b) bf[b] = (kk < n && bcol[b] < n2) ? Qc[(size_t)kk * n2 + bcol[b]] : 0.0;
      };
      load(a0, b0, 0);
      for (int k0 = 0; k0 < n; k0 += 8) {
        if (k0 + 4 < n) load(a1, b1, k0 + 4);
#pragma unroll
        for (int a = 0; a < MTI; ++a)
#pragma unroll
          for (int b = 0; b < NTW; ++b) acc[a][b] = mfma(a0[a], b0[b], acc[a][b]);
        if (k0 + 8 < n) load(a0, b0, k0 + 8);
        if (k0 + 4 < n) {
#pragma unroll
          for (int a = 0; a < MTI; ++a)
#pragma unroll
            for (int b = 0; b < NTW; ++b) acc[a][b] = mfma(a1[a], b1[b], acc[a][b]);
        }
      }
    }
    __syncthreads();
    // epilogue
#pragma unroll
    for (int b = 0; b < NTW; ++b) {
      const int tile = wave + 4 * b;
      if (tile >= ntiles) continue;
      const int col = tile * 16 + li;
      if (col >= n2) continue;
#pragma unroll
      for (int a = 0; a < MTI; ++a) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int i = a * 16 + crow(lk, r);
          if (col < n) {
            carry[col * ldcar + i] = acc[a][b][r];
          } else {
            const int row = R0 + i;
            if (row < N) A[(size_t)((k + 1) * n + (col - n)) * lda + row] = acc[a][b][r];
          }
        }
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < n * MT; e += 256) {
    const int i = e % MT, kk = e / MT;
    const int row = R0 + i;
    if (row < N) A[(size_t)kk * lda + row] = carry[kk * ldcar + i];
  }
}

// ---------------------------------------------------------------------------
// DQMC matrix builder: A = I + corner B_1 - subdiagonal B_l, with
// B_l[:, i] = expk[:, i] * (s_{l,i} > 0 ? ep : em).
// ---------------------------------------------------------------------------
__global__ void dqmc_build_kernel(double* __restrict__ A, int lda, int n, int p,
                                  const double* __restrict__ expk, const signed char* __restrict__ fields,
                                  double ep, double em) {
  const long long N = (long long)n * p;
  const long long total = N * N;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long col = e / N, row = e - col * N;
    const int cb = (int)(col / n), rb = (int)(row / n);
    const int ci = (int)(col - (long long)cb * n), ri = (int)(row - (long long)rb * n);
    double v = (row == col) ? 1.0 : 0.0;
    if (rb == 0 && cb == p - 1) {
      v = expk[(size_t)ci * n + ri] * (fields[ci] > 0 ? ep : em);
    } else if (rb == cb + 1) {
      v = -(expk[(size_t)ci * n + ri] * (fields[(size_t)rb * n + ci] > 0 ? ep : em));
    }
    A[col * lda + row] = v;
  }
}

// Flip the HS field of (slice l, site i) for each listed flip.
__global__ void hs_flip_kernel(signed char* __restrict__ fields, const int* __restrict__ flips, int count) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) fields[flips[e]] = (signed char)(-fields[flips[e]]);
}

// ---------------------------------------------------------------------------
// Residual of a p-cyclic inverse: sum over the tile of (A G - I)^2, with A
// given by its 2p non-zero blocks (diag block i and off-diagonal block i,
// whose block column is i-1, or p-1 for i = 0).  One workgroup computes a
// 64 x 64 tile of row block ib of A*G and adds its squared residual to
// *out.  Blocks are stored compactly: blocks[(2*ib + s) * n * n], s = 0 for
// the diagonal block, 1 for the off-diagonal one.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pcyclic_residual_kernel(const double* __restrict__ blocks,
                                                               const double* __restrict__ G, int ldg,
                                                               int n, int p, double* out) {
  const int N = n * p;
  const int tiles_r = (n + 63) / 64, tiles_c = (N + 63) / 64;
  const int ib = blockIdx.x / (tiles_r * tiles_c);
  const int rem = blockIdx.x % (tiles_r * tiles_c);
  const int i0 = (rem / tiles_c) * 64, j0 = (rem % tiles_c) * 64;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int wi = wave & 1, wj = wave >> 1;
  __shared__ double As[64][17];
  __shared__ double Bs[64][17];
  __shared__ double part[4];
  d4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4zero();
  const int offb = (ib == 0) ? p - 1 : ib - 1;
  for (int s = 0; s < 2; ++s) {
    const double* Ab = blocks + (size_t)(2 * ib + s) * n * n;
    const int gb = (s == 0) ? ib : offb;  // block row of G multiplied
    for (int k0 = 0; k0 < n; k0 += 16) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = tid & 63, kk = (tid >> 6) + 4 * r;
        const int gi = i0 + i, gk = k0 + kk;
        As[i][kk] = (gi < n && gk < n) ? Ab[(size_t)gk * n + gi] : 0.0;
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int kk = tid & 15, j = (tid >> 4) + 16 * r;
        const int gk = k0 + kk, gj = j0 + j;
        Bs[j][kk] = (gk < n && gj < N) ? G[(size_t)gj * ldg + gb * n + gk] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < 16; ks += 4) {
        double af[2], bf[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) af[a] = As[wi * 32 + a * 16 + li][ks + lk];
#pragma unroll
        for (int b = 0; b < 2; ++b) bf[b] = Bs[wj * 32 + b * 16 + li][ks + lk];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
          for (int b = 0; b < 2; ++b) acc[a][b] = mfma(af[a], bf[b], acc[a][b]);
      }
      __syncthreads();
    }
  }
  double ss = 0.0;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int col = j0 + wj * 32 + b * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + wi * 32 + a * 16 + crow(lk, r);
        if (row < n && col < N) {
          double v = acc[a][b][r];
          if (ib * n + row == col) v -= 1.0;
          ss = fma(v, v, ss);
        }
      }
    }
  for (int m2 = 32; m2 >= 1; m2 >>= 1) ss += __shfl_xor(ss, m2);
  if (lane == 0) part[wave] = ss;
  __syncthreads();
  if (tid == 0) atomicAdd(out, (part[0] + part[1]) + (part[2] + part[3]));
}

// Sum of squares of an m x n column-major array (for Frobenius norms).
__global__ void sumsq_kernel(const double* __restrict__ X, int ld, int m, int n, double* out) {
  double s = 0.0;
  const long long total = (long long)m * n;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long j = e / m, i = e - j * m;
    const double v = X[j * ld + i];
    s = fma(v, v, s);
  }
  for (int m2 = 32; m2 >= 1; m2 >>= 1) s += __shfl_xor(s, m2);
  __shared__ double part[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) part[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += part[w];
    atomicAdd(out, t);
  }
}

// Gather the 2p non-zero blocks of a p-cyclic A into the compact layout
// used by pcyclic_residual_kernel.
__global__ void gather_blocks_kernel(const double* __restrict__ A, int lda, int n, int p,
                                     double* __restrict__ blocks) {
  const long long per = (long long)n * n, total = per * 2 * p;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long bi = e / per, r = e - bi * per;
    const int ib = (int)(bi >> 1), s = (int)(bi & 1);
    const int j = (int)(r / n), i = (int)(r - (long long)j * n);
    const int cb = (s == 0) ? ib : (ib == 0 ? p - 1 : ib - 1);
    blocks[e] = A[(size_t)(cb * n + j) * lda + ib * n + i];
  }
}

}  // namespace bsofi

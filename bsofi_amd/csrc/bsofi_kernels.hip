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
template <int CPG>
__device__ __forceinline__ double wave_transpose_reduce(double (&v)[CPG], int lane, int& idx) {
  // Reduce v[0..CPG) over the 16 lanes sharing lane & 3 (xor 4, 8, 16, 32).
  // On return the lane holds the wave sum of entry idx.
  int cnt = CPG;
  idx = 0;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const int mask = 32 >> s;
    if (cnt >= 2) {
      const int half = cnt >> 1;
      const bool bit = (lane & mask) != 0;
#pragma unroll
      for (int t = 0; t < CPG / 2; ++t) {
        if (t < half) {
          const double send = bit ? v[t] : v[t + half];
          const double keep = bit ? v[t + half] : v[t];
          v[t] = keep + __shfl_xor(send, mask);
        }
      }
      if (bit) idx += half;
      cnt = half;
    } else {
      v[0] += __shfl_xor(v[0], mask);
    }
  }
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
template <int NB>
__device__ void build_T(const double (*G)[NB + 1], double (*Ts)[NB + 1], const double* taus, int nbj,
                        double* __restrict__ Tout, int ldt, int lane) {
  if (lane >= NB) return;
  const int i = lane;
  for (int c = 0; c < nbj; ++c) {
    double v = 0.0;
    if (i < c) {
      double s = 0.0;
      for (int j = i; j < c; ++j) s = fma(Ts[i][j], G[j][c], s);
      v = -taus[c] * s;
    } else if (i == c) {
      v = taus[c];
    }
    Ts[i][c] = v;
  }
  for (int c = 0; c < nbj; ++c) Tout[(size_t)c * ldt + i] = (i < nbj) ? Ts[i][c] : 0.0;
}

template <int RPG, int CPG>
__global__ __launch_bounds__(256) void panel_factor_kernel(const PanelTask* __restrict__ tasks,
                                                           const Item* __restrict__ items) {
  constexpr int NB = 4 * CPG;
  const Item it = items[blockIdx.x];
  const PanelTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cg = lane & 3;
  const int rg = (lane >> 2) + 16 * wave;
  const int m = t.m, nbj = t.nbj, lda = t.lda;

  __shared__ double red[2][4][NB];
  __shared__ double prow[2][NB];
  __shared__ double G[NB][NB + 1];
  __shared__ double Ts[NB][NB + 1];
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

  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  const int src_base = lane & ~3;

  for (int cc = 0; cc < 4; ++cc) {
#pragma unroll
    for (int jc = 0; jc < CPG; ++jc) {
      const int c = cc * CPG + jc;
      if (c < nbj) {
        const int buf = c & 1;
        // x = column c at my rows, from the lane (rg, cc) that owns it.
        double x[RPG];
#pragma unroll
        for (int q = 0; q < RPG; ++q) x[q] = __shfl(a[q][jc], src_base | cc);
        double P[CPG];
#pragma unroll
        for (int jj = 0; jj < CPG; ++jj) P[jj] = 0.0;
#pragma unroll
        for (int q = 0; q < RPG; ++q) {
          if (rg + 64 * q > c) {
#pragma unroll
            for (int jj = 0; jj < CPG; ++jj) P[jj] = fma(x[q], a[q][jj], P[jj]);
          }
        }
        int idx;
        const double ws = wave_transpose_reduce<CPG>(P, lane, idx);
        if ((lane & 4) == 0) red[buf][wave][cg * CPG + idx] = ws;
        if (rg == c) {
#pragma unroll
          for (int jj = 0; jj < CPG; ++jj) prow[buf][cg * CPG + jj] = a[0][jj];
        }
        __syncthreads();

        double tot[CPG], pr[CPG];
#pragma unroll
        for (int jj = 0; jj < CPG; ++jj) {
          const int j = cg * CPG + jj;
          tot[jj] = (red[buf][0][j] + red[buf][1][j]) + (red[buf][2][j] + red[buf][3][j]);
          pr[jj] = prow[buf][j];
        }
        double xn2 = (red[buf][0][c] + red[buf][1][c]) + (red[buf][2][c] + red[buf][3][c]);
        const double alpha = prow[buf][c];

        // dlarfg.  The fast path takes ||x|| = sqrt(sum x^2); a sum of
        // squares outside [2^-960, 2^960] is recomputed with scaling.
        double xnorm;
        if (!(xn2 > 0x1p-960) || !(xn2 < 0x1p960)) {
          double lm = 0.0;
#pragma unroll
          for (int q = 0; q < RPG; ++q)
            if (rg + 64 * q > c) lm = fmax(lm, fabs(x[q]));
          const double amax = block_max_256(lm, scratch);
          double ls = 0.0;
          if (amax > 0.0) {
#pragma unroll
            for (int q = 0; q < RPG; ++q)
              if (rg + 64 * q > c) {
                const double y = x[q] / amax;
                ls = fma(y, y, ls);
              }
          }
          const double s2 = block_sum_256(ls, scratch);
          xnorm = amax > 0.0 ? amax * sqrt(s2) : 0.0;
        } else {
          xnorm = sqrt(xn2);
        }
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

        double w[CPG];
#pragma unroll
        for (int jj = 0; jj < CPG; ++jj) {
          const int j = cg * CPG + jj;
          const double d = fma(scale, tot[jj], pr[jj]);
          w[jj] = (j > c) ? tau * d : 0.0;
          if (j < c && rg == 0) G[j][c] = d;
        }
#pragma unroll
        for (int q = 0; q < RPG; ++q) {
          const int row = rg + 64 * q;
          if (row > c) {
            const double v = x[q] * scale;
#pragma unroll
            for (int jj = 0; jj < CPG; ++jj) a[q][jj] = fma(-v, w[jj], a[q][jj]);
            if (cg == cc) a[q][jc] = v;
          } else if (row == c) {
#pragma unroll
            for (int jj = 0; jj < CPG; ++jj) a[q][jj] -= w[jj];
            if (cg == cc) a[q][jc] = beta;
          }
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
  if (wave == 0) build_T<NB>(G, Ts, taus, nbj, t.T, t.ldt, lane);
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

__global__ __launch_bounds__(256) void refl_apply_kernel(const ReflTask* __restrict__ tasks,
                                                         const Item* __restrict__ items) {
  const Item it = items[blockIdx.x];
  const ReflTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m = t.m, nbj = t.nbj;

  __shared__ double Wp[4][32][17];
  __shared__ double W[32][17];
  __shared__ double Tl[32][33];

  // the column this lane serves in B / C fragments
  const int vcol = it.sub * 16 + li;
  double* cptr = nullptr;
  if (vcol < t.w1)
    cptr = t.C1 + (size_t)vcol * t.ldc;
  else if (vcol - t.w1 < t.w2)
    cptr = t.C2 + (size_t)(vcol - t.w1) * t.ldc;

  // phase 1
  const int per = ((m + 15) / 16) * 4;
  const int kb = wave * per, ke = min(m, kb + per);
  d4 acc0 = d4zero(), acc1 = d4zero();
  const bool two = nbj > 16;
  for (int k0 = kb; k0 < ke; k0 += 4) {
    const int row = k0 + lk;
    const double bv = (cptr && row < ke) ? cptr[row] : 0.0;
    const double a0 = (row < ke) ? vload(t.V, t.ldv, m, nbj, row, li) : 0.0;
    acc0 = mfma(a0, bv, acc0);
    if (two) {
      const double a1 = (row < ke) ? vload(t.V, t.ldv, m, nbj, row, 16 + li) : 0.0;
      acc1 = mfma(a1, bv, acc1);
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Wp[wave][crow(lk, r)][li] = acc0[r];
    Wp[wave][16 + crow(lk, r)][li] = two ? acc1[r] : 0.0;
  }
  for (int e = tid; e < 32 * 32; e += 256) {
    const int i = e & 31, j = e >> 5;
    Tl[i][j] = (i < nbj && j < nbj) ? t.T[(size_t)j * t.ldt + i] : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < 32 * 16; e += 256) {
    const int i = e >> 4, j = e & 15;
    W[i][j] = (Wp[0][i][j] + Wp[1][i][j]) + (Wp[2][i][j] + Wp[3][i][j]);
  }
  __syncthreads();
  // phase 2: Wp[0] <- -op(T) W
  for (int e = tid; e < 32 * 16; e += 256) {
    const int i = e >> 4, j = e & 15;
    double s = 0.0;
    if (i < nbj) {
      if (t.trans) {  // T^T: sum_{k <= i} T[k][i] W[k][j]
        for (int k = 0; k <= i; ++k) s = fma(Tl[k][i], W[k][j], s);
      } else {        // T: sum_{k >= i} T[i][k] W[k][j]
        for (int k = i; k < nbj; ++k) s = fma(Tl[i][k], W[k][j], s);
      }
    }
    Wp[0][i][j] = -s;
  }
  __syncthreads();
  // phase 3
  const int mt = (m + 15) / 16;
  for (int rt = wave; rt < mt; rt += 4) {
    d4 acc;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = rt * 16 + crow(lk, r);
      acc[r] = (cptr && row < m) ? cptr[row] : 0.0;
    }
    for (int k0 = 0; k0 < nbj; k0 += 4) {
      const double av = vload(t.V, t.ldv, m, nbj, rt * 16 + li, k0 + lk);
      const double bv = Wp[0][k0 + lk][li];
      acc = mfma(av, bv, acc);
    }
    if (cptr) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = rt * 16 + crow(lk, r);
        if (row < m) cptr[row] = acc[r];
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
  __shared__ double Ts[32][33];
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
  if (wave == 0) build_T<32>(G, Ts, taus, nbj, t.T, t.ldt, lane);
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
  for (int k = p - 2; k >= 0; --k) {
    const double* __restrict__ Qk = t.Q + (size_t)k * t.qstride;
    d4 acc[MTI][NTW];
#pragma unroll
    for (int a = 0; a < MTI; ++a)
#pragma unroll
      for (int b = 0; b < NTW; ++b) acc[a][b] = d4zero();

    int bcol[NTW];
#pragma unroll
    for (int b = 0; b < NTW; ++b) bcol[b] = (wave + 4 * b) * 16 + li;

    if (rbmin <= k) {
      // [X_{r,k}] part: K rows 0..n-1 of Q_k^T
      const size_t colbase = (size_t)k * n * lda;
      for (int k0 = 0; k0 < n; k0 += 4) {
        const int kk = k0 + lk;
        double af[MTI], bf[NTW];
#pragma unroll
        for (int a = 0; a < MTI; ++a) {
          double v = 0.0;
          if (kk < n && rblk[a] <= k && !(rblk[a] == k && rin[a] > kk))
            v = A[colbase + (size_t)kk * lda + rrow[a]];
          af[a] = v;
        }
#pragma unroll
        for (int b = 0; b < NTW; ++b)
          bf[b] = (kk < n && bcol[b] < n2) ? Qk[(size_t)kk * n2 + bcol[b]] : 0.0;
#pragma unroll
        for (int a = 0; a < MTI; ++a)
#pragma unroll
          for (int b = 0; b < NTW; ++b) acc[a][b] = mfma(af[a], bf[b], acc[a][b]);
      }
    }
    // [carry] part: K rows n..2n-1 of Q_k^T
    for (int k0 = 0; k0 < n; k0 += 4) {
      const int kk = k0 + lk;
      double af[MTI], bf[NTW];
#pragma unroll
      for (int a = 0; a < MTI; ++a) af[a] = (kk < n) ? carry[kk * ldcar + a * 16 + li] : 0.0;
#pragma unroll
      for (int b = 0; b < NTW; ++b)
        bf[b] = (kk < n && bcol[b] < n2) ? Qk[(size_t)(n + kk) * n2 + bcol[b]] : 0.0;
#pragma unroll
      for (int a = 0; a < MTI; ++a)
#pragma unroll
        for (int b = 0; b < NTW; ++b) acc[a][b] = mfma(af[a], bf[b], acc[a][b]);
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

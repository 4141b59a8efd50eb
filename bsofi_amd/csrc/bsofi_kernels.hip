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

#include <type_traits>
#include <utility>

#include "bsofi.h"
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

// Global-memory views of the task pointers.  The pointers reach the kernels
// inside task records, so the compiler cannot tell that they address global
// memory and would emit flat loads and stores.  Flat operations count on the
// LDS wait counter too: every wait for an LDS read would also wait for all
// HBM loads in flight, and stores get a full drain in front of them.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> glb(T* p) {
  return (gptr<T>)p;
}

// ---------------------------------------------------------------------------
// Sub-panel Householder QR (K1)
// ---------------------------------------------------------------------------
// NW waves (4 or 8); wave w owns columns w CPW .. w CPW + CPW - 1 of the
// sub-panel (NB = NW CPW columns), lane l holds rows l + 64 q (q < RPG) of
// them in registers.  Per column c:
//   * the wave owning c has formed its reflector (dlarfg) and put v_c
//     (zero above row c, one at row c) and tau_c in LDS; one barrier, then
//     every wave reads v_c at its rows;
//   * each wave forms d_j = v_c^T a_j for its own columns j and reduces them
//     over its 64 lanes (permlane swaps + DPP):  j > c : H_c's dot products,
//     j < c : v_j^T v_c (the Gram entries of the T factor, dlarft);
//   * the owner of c + 1 applies H_c to that column first, forms its
//     reflector (norm reduction + dlarfg) and publishes it, then every wave
//     applies H_c to the rest of its unfinished columns.
// So a column costs one barrier, no cross-wave reduction, and only the
// owner of the next column spends instructions on its norm and dlarfg.

// 64-bit DPP move (two 32-bit halves); CTRL is a GFX9 DPP control code.
template <int CTRL>
__device__ __forceinline__ double dpp64(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xffffffffLL), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Cross-half exchanges of gfx950: with (x', y') = swap(x, y),
//   permlane32: x'[l] = y[l - 32] for l >= 32, y'[l] = x[l + 32] for l < 32;
//   permlane16: the same between odd rows of x and even rows of y (rows of
//   16 lanes).  So x' + y' is, for lanes whose stage bit is clear,
//   x[l] + x[partner] and, for the others, y[partner] + y[l].
template <int S>
__device__ __forceinline__ double swap_add(double x, double y) {
  const unsigned long long bx = __builtin_bit_cast(unsigned long long, x);
  const unsigned long long by = __builtin_bit_cast(unsigned long long, y);
  unsigned lo0, lo1, hi0, hi1;
  if constexpr (S == 0) {
    const auto l = __builtin_amdgcn_permlane32_swap((unsigned)bx, (unsigned)by, false, false);
    const auto h = __builtin_amdgcn_permlane32_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    lo0 = l[0]; lo1 = l[1]; hi0 = h[0]; hi1 = h[1];
  } else {
    const auto l = __builtin_amdgcn_permlane16_swap((unsigned)bx, (unsigned)by, false, false);
    const auto h = __builtin_amdgcn_permlane16_swap((unsigned)(bx >> 32), (unsigned)(by >> 32), false, false);
    lo0 = l[0]; lo1 = l[1]; hi0 = h[0]; hi1 = h[1];
  }
  const double x2 = __builtin_bit_cast(double, ((unsigned long long)hi0 << 32) | lo0);
  const double y2 = __builtin_bit_cast(double, ((unsigned long long)hi1 << 32) | lo1);
  return x2 + y2;
}

// Value of the partner lane at stages 2..5 (lane ^ 8 by row_ror:8, the
// mirror inside each 8-lane half which flips bit 2, lane ^ 2, lane ^ 1).
template <int S>
__device__ __forceinline__ double dpp_partner(double v) {
  if constexpr (S == 2) return dpp64<0x128>(v);
  else if constexpr (S == 3) return dpp64<0x141>(v);
  else if constexpr (S == 4) return dpp64<0x4E>(v);
  else return dpp64<0xB1>(v);
}

// "Transpose reduce" of v[0..CNT) over the 64 lanes of a wave.  Stage S
// pairs lanes that differ in bit 32 >> S (and hold the same entries); while
// more than one entry is left a stage halves the entries a lane keeps (the
// lane with the bit set keeps the upper half), afterwards stages just add.
// Every lane then holds the full sum of one entry; which one is fixed by its
// bits of the halving stages (see wave_reduce_lane).
template <int CNT, int S, int CPW>
__device__ __forceinline__ void wave_reduce_stage(double (&v)[CPW], int lane) {
  if constexpr (S < 6) {
    if constexpr (CNT >= 2) {
      constexpr int half = CNT / 2;
      if constexpr (S < 2) {
#pragma unroll
        for (int t = 0; t < half; ++t) v[t] = swap_add<S>(v[t], v[t + half]);
      } else {
        const bool bit = (lane & (32 >> S)) != 0;
#pragma unroll
        for (int t = 0; t < half; ++t) {
          const double send = bit ? v[t] : v[t + half];
          const double keep = bit ? v[t + half] : v[t];
          v[t] = keep + dpp_partner<S>(send);
        }
      }
      wave_reduce_stage<half, S + 1, CPW>(v, lane);
    } else {
      if constexpr (S < 2) v[0] = swap_add<S>(v[0], v[0]);
      else v[0] += dpp_partner<S>(v[0]);
      wave_reduce_stage<CNT, S + 1, CPW>(v, lane);
    }
  }
}

// A lane holding the full sum of entry jj after wave_reduce_stage<CPW, 0>:
// the bits of jj, most significant first, are its bits 32, 16, 8, ...
template <int CPW>
__device__ __forceinline__ constexpr int wave_reduce_lane(int jj) {
  int lane = 0, bitv = 32;
  for (int h = CPW / 2; h >= 1; h /= 2, bitv /= 2)
    if (jj & h) lane |= bitv;
  return lane;
}

__device__ __forceinline__ double readlane64(double v, int l) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// Wave-wide sum and max, the same bits in every lane (lane 0's tree).
__device__ __forceinline__ double wave_sum(double v, int lane) {
  double t[1] = {v};
  wave_reduce_stage<1, 0, 1>(t, lane);
  return readlane64(t[0], 0);
}

__device__ __forceinline__ double wave_max(double v) {
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  return readlane64(v, 0);
}

// Workgroup barrier that orders LDS accesses only: global stores issued
// before it stay in flight (__syncthreads also waits for them).
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// T of the compact WY form (dlarft, forward, columnwise) from the Gram
// entries G[j][c] = v_j^T v_c (read for j < c only) and tau (zero for the
// columns at and beyond nbj), by all 256 threads of the workgroup.
// Recursive doubling: with T1, T2 the factors of two adjacent column
// groups, H1 H2 = I - V T V^T with
//   T = [[T1, -T1 (V1^T V2) T2], [0, T2]],
// so levels b = 1, 2, 4, ... NB/2 each form the off-diagonal blocks of all
// merges of width 2b in two barrier-separated products.  This equals the
// column recurrence T[0:c, c] = -tau_c T[0:c, 0:c] V^T v_c of dlarft up to
// rounding.  Tl and X are LDS scratch of NB x (NB + 1).  Must be called by
// every thread of the block.  Writes the nbj x nbj factor, Tout[c * ldt + i]
// for i, c < nbj (ldt >= nbj may be below NB).  Its barriers order LDS only,
// so global stores issued before the call stay in flight.
template <int NB, int NT = 256>
__device__ __forceinline__ void build_T(const double (*G)[NB + 1], const double* taus, int nbj,
                                        gptr<double> Tout, int ldt, double (*Tl)[NB + 1],
                                        double (*X)[NB + 1]) {
  static_assert(NB * NB / 4 <= NT, "one thread per entry of the widest level");
  const int tid = threadIdx.x;
  for (int e = tid; e < NB * NB; e += NT) {
    const int i = e / NB, j = e % NB;
    Tl[i][j] = (i == j) ? taus[i] : 0.0;
  }
  lds_sync();
#pragma unroll
  for (int b = 1; b < NB; b *= 2) {
    const int e = tid;
    const bool act = e < NB * b / 2;
    const int mi = e / (b * b), loc = e % (b * b);
    const int r = loc / b, cq = loc % b;
    const int s0 = mi * 2 * b;
    if (act) {  // X = (V1^T V2) T2
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < b; ++j) acc = fma(G[s0 + r][s0 + b + j], Tl[s0 + b + j][s0 + b + cq], acc);
      X[s0 + r][s0 + b + cq] = acc;
    }
    lds_sync();
    if (act) {  // T12 = -T1 X
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < b; ++j) acc = fma(Tl[s0 + r][s0 + j], X[s0 + j][s0 + b + cq], acc);
      Tl[s0 + r][s0 + b + cq] = -acc;
    }
    lds_sync();
  }
  for (int e = tid; e < nbj * nbj; e += NT) {
    const int c = e / nbj, i = e % nbj;
    Tout[(size_t)c * ldt + i] = Tl[i][c];
  }
}

// ||x|| with scaling by max|x|, over one wave (every wave holds all of x):
// the rare path of the panel's dlarfg, taken when the plain sum of squares
// leaves the safe range.
template <int RPG>
__device__ __forceinline__ double scaled_norm_wave(const double (&x)[RPG], int lane) {
  double lm = 0.0;
#pragma unroll
  for (int q = 0; q < RPG; ++q) lm = fmax(lm, fabs(x[q]));
  const double amax = wave_max(lm);
  if (!(amax > 0.0)) return 0.0;
  double ls = 0.0;
#pragma unroll
  for (int q = 0; q < RPG; ++q) {
    const double y = x[q] / amax;
    ls = fma(y, y, ls);
  }
  return amax * sqrt(wave_sum(ls, lane));
}

// sqrt and 1/x for arguments known to be normal, finite and positive (sqrt)
// or nonzero (1/x) and far from the exponent limits, as in the panel's fast
// path (|x| within [2^-960, 2^962]).  From the hardware approximations
// v_rsq_f64 / v_rcp_f64, one Goldschmidt iteration and a Newton correction
// for sqrt, two Newton iterations for 1/x: a few dependent FMAs, within an
// ulp or two of the correctly rounded value, without the range scaling the
// general library routines carry.
__device__ __forceinline__ double sqrt_normal(double s) {
  const double y = __builtin_amdgcn_rsq(s);
  double g = s * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  const double d = fma(-g, g, s);
  return fma(d, h, g);
}

__device__ __forceinline__ double rcp_normal(double x) {
  double z = __builtin_amdgcn_rcp(x);
  double e = fma(-x, z, 1.0);
  z = fma(z, e, z);
  e = fma(-x, z, 1.0);
  return fma(z, e, z);
}

struct Larfg {
  double tau, beta, scale;
};

// dlarfg from alpha and ||x||: beta, tau and the scale that turns x into
// the stored reflector, with LAPACK's safe-minimum rescaling loop.
__device__ __forceinline__ Larfg dlarfg_scalar(double alpha, double xnorm) {
  const double safmin = DBL_MIN / (DBL_EPSILON * 0.5);
  Larfg r;
  if (xnorm == 0.0) {
    r.tau = 0.0;
    r.beta = alpha;
    r.scale = 1.0;
  } else {
    double beta = -copysign(hypot(alpha, xnorm), alpha);
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
    r.tau = (beta - al) / beta;
    r.scale = sc_acc / (al - beta);
    for (int q2 = 0; q2 < knt; ++q2) beta *= safmin;
    r.beta = beta;
  }
  return r;
}

// Entry of the CPW transposed-reduced values that lane l holds after
// wave_reduce_stage<CPW, 0>: its bits 32, 16, ... of the halving stages.
template <int CPW>
__device__ __forceinline__ int wave_reduce_entry(int lane) {
  int jj = 0, bitv = 32;
  for (int h = CPW / 2; h >= 1; h /= 2, bitv /= 2)
    if (lane & bitv) jj |= h;
  return jj;
}

// The body of one sub-panel QR by one workgroup of 64 NW threads (all of
// them must call it); the kernels below and the device-scheduled BSOF sweep
// run it.
template <int RPG, int CPW, int NW>
__device__ __forceinline__ void panel_body(const PanelTask& t) {
  constexpr int NB = NW * CPW;
  constexpr int NT = 64 * NW;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m = t.m, nbj = t.nbj, lda = t.lda;
  const int col0 = wave * CPW;
  const gptr<double> Pg = glb(t.P);

  __shared__ double vs[2][RPG][64];  // v of the current column, double-buffered
  __shared__ double G[NB][NB + 1];
  __shared__ double taus[NB];
  __shared__ double Tl[NB][NB + 1];
  __shared__ double Xl[NB][NB + 1];

  double a[RPG][CPW];
  // The column loop runs first on dlarfg's fast path alone, with no branch
  // in it; `bad` records whether any column left the fast path's range.
  // Then, and only then, the panel -- still intact in memory, it is written
  // back at the end -- is factored again with LAPACK's scaled path.
  auto run = [&](auto safe_tag) -> bool {
    constexpr bool SAFE = decltype(safe_tag)::value;
    bool bad = false;
    for (int e = tid; e < NB * (NB + 1); e += NT) (&G[0][0])[e] = 0.0;
    if (tid < NB) taus[tid] = 0.0;
#pragma unroll
    for (int q = 0; q < RPG; ++q) {
      const int row = lane + 64 * q;
#pragma unroll
      for (int jj = 0; jj < CPW; ++jj) {
        const int col = col0 + jj;
        a[q][jj] = (row < m && col < nbj) ? Pg[(size_t)col * lda + row] : 0.0;
      }
    }
    // vmcnt(0): the panel is in registers before anything else.  Without it
    // the compiler cannot rule out loads still in flight at the final
    // stores (a path with no column step) and drains the memory counter
    // before each of them, which serialises the stores.
    __builtin_amdgcn_s_waitcnt(0x0F70);

    // Reflector of column c, local column jc of the calling wave, whose
    // entries at and below row c are final: dlarfg, v_c to vs[buf], tau_c
    // to taus[c]; the column becomes beta on the diagonal and v below it.
    auto reflector = [&](int jc, int c, int buf) {
      const double alpha = readlane64(a[0][jc], c);  // row c < 64: lane c, q = 0
      double x[RPG];
      x[0] = (lane > c) ? a[0][jc] : 0.0;
#pragma unroll
      for (int q = 1; q < RPG; ++q) x[q] = a[q][jc];
      // Both chains start with an fma, never a product the compiler could
      // contract into the final add: the panel instances (RPG) then round
      // identically, the extra rows of a larger RPG adding exact zeros.
      double n0 = fma(x[0], x[0], 0.0), n1 = 0.0;
#pragma unroll
      for (int q = 1; q < RPG; ++q) {
        if (q & 1) n1 = fma(x[q], x[q], n1);
        else n0 = fma(x[q], x[q], n0);
      }
      const double xn2 = wave_sum(n0 + n1, lane);
      // In range, beta = -sign(alpha) sqrt(alpha^2 + ||x||^2) and, with
      // r = 1 / (beta (alpha - beta)),
      //   tau = (beta - alpha) / beta = -(alpha - beta)^2 r,  scale = beta r.
      // A sum of squares outside [2^-960, 2^960] (including x = 0) or
      // |alpha| >= 2^480 takes LAPACK's scaled path instead.
      const bool inrange = xn2 > 0x1p-960 && xn2 < 0x1p960 && fabs(alpha) < 0x1p480;
      double beta = -copysign(sqrt_normal(fma(alpha, alpha, xn2)), alpha);
      const double amb = alpha - beta;
      const double rr = rcp_normal(beta * amb);
      double tau = -(amb * amb) * rr;
      double scale = beta * rr;
      if constexpr (SAFE) {
        if (!inrange) {
          const Larfg lf = dlarfg_scalar(alpha, scaled_norm_wave<RPG>(x, lane));
          tau = lf.tau;
          beta = lf.beta;
          scale = lf.scale;
        }
      } else {
        bad |= !inrange;
      }
      const double v0 = (lane > c) ? x[0] * scale : (lane == c ? 1.0 : 0.0);
      vs[buf][0][lane] = v0;
      a[0][jc] = (lane > c) ? v0 : (lane == c ? beta : a[0][jc]);
#pragma unroll
      for (int q = 1; q < RPG; ++q) {
        const double v = x[q] * scale;
        vs[buf][q][lane] = v;
        a[q][jc] = v;
      }
      if (lane == 0) taus[c] = tau;
    };

    if (wave == 0 && nbj > 0) reflector(0, 0, 0);
    for (int cw = 0; cw < NW; ++cw) {
#pragma unroll
      for (int jc = 0; jc < CPW; ++jc) {
        const int c = cw * CPW + jc;
        if (c < nbj) {
          const int buf = c & 1;
          __syncthreads();
          double v[RPG];
#pragma unroll
          for (int q = 0; q < RPG; ++q) v[q] = vs[buf][q][lane];
          const double tau = taus[c];
          // d_j = v_c^T a_j (two chains per column)
          double P[CPW];
#pragma unroll
          for (int jj = 0; jj < CPW; ++jj) {
            double s0 = fma(v[0], a[0][jj], 0.0), s1 = 0.0;  // fma start, as for the norm
#pragma unroll
            for (int q = 1; q < RPG; ++q) {
              if (q & 1) s1 = fma(v[q], a[q][jj], s1);
              else s0 = fma(v[q], a[q][jj], s0);
            }
            P[jj] = s0 + s1;
          }
          wave_reduce_stage<CPW, 0, CPW>(P, lane);
          // Gram entries v_j^T v_c of this wave's finished columns j < c
          {
            const int jj = wave_reduce_entry<CPW>(lane);
            if ((lane & (64 / CPW - 1)) == 0 && col0 + jj < c) G[col0 + jj][c] = P[0];
          }
          double w[CPW];
#pragma unroll
          for (int jj = 0; jj < CPW; ++jj) w[jj] = tau * readlane64(P[0], wave_reduce_lane<CPW>(jj));
          // column c + 1 first; its owner forms and publishes its reflector
          const int jn = (jc + 1) % CPW;
          const int wn = cw + (jc + 1) / CPW;
          if (col0 + jn > c && col0 + jn < nbj) {
#pragma unroll
            for (int q = 0; q < RPG; ++q) a[q][jn] = fma(-v[q], w[jn], a[q][jn]);
          }
          if (wave == wn && c + 1 < nbj) reflector(jn, c + 1, buf ^ 1);
#pragma unroll
          for (int jj = 0; jj < CPW; ++jj) {
            if (jj == jn) continue;
            if (col0 + jj > c && col0 + jj < nbj) {
#pragma unroll
              for (int q = 0; q < RPG; ++q) a[q][jj] = fma(-v[q], w[jj], a[q][jj]);
            }
          }
        }
      }
    }
    return __syncthreads_or(bad) != 0;
  };
  if (run(std::false_type{})) run(std::true_type{});

  // G and taus are complete (the final barrier of run).  The panel's stores
  // go out first and stay in flight under build_T, whose barriers order LDS
  // only.
#pragma unroll
  for (int q = 0; q < RPG; ++q) {
    const int row = lane + 64 * q;
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int col = col0 + jj;
      if (row < m && col < nbj) Pg[(size_t)col * lda + row] = a[q][jj];
    }
  }
  if (tid < nbj) glb(t.tau)[tid] = taus[tid];
  build_T<NB, NT>(G, taus, nbj, glb(t.T), t.ldt, Tl, Xl);
}

template <int RPG, int CPW, int NW>
__global__ __launch_bounds__(64 * NW) void panel_factor_kernel(const PanelTask* __restrict__ tasks,
                                                               const Item* __restrict__ items) {
  panel_body<RPG, CPW, NW>(tasks[items[blockIdx.x].task]);
}

// A launch's task: from the task table through the workgroup's item, or,
// in the ONE instances (launches of one task), from the kernel arguments.
// The single-matrix path launches one task at a time: its first data load
// then waits for the kernel argument load alone, not for the item and task
// reads behind it.
template <bool ONE, class T>
__device__ __forceinline__ const T& launch_task(const T* __restrict__ tasks, const Item* __restrict__ items,
                                                const T& one) {
  if constexpr (ONE) return one;
  else return tasks[items[blockIdx.x].task];
}

// f(std::integral_constant<int, i>{}) for i = 0 .. N-1 in order.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__device__ __forceinline__ void static_for(F& f) {
  static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Sub-panel Householder QR, cyclic form (K1c).  The same arithmetic as
// panel_body -- the same FMA chains and the same butterfly reduction tree
// (stages pair lanes 32, 16, ... apart, whichever lane keeps the sum; the
// adds are commutative), so the same bits -- without a workgroup barrier per
// column:
//   * column c belongs to wave c mod NW (local column c / NW), so the owner
//     of the next column is never the wave that has just formed one;
//   * every v_c and tau_c stays in LDS (Vl, taus) and a counter `ready`
//     (columns published) replaces the barrier: a wave waits for column c
//     only when it needs v_c, and may lag the chain by any number of
//     columns;
//   * the owner of c + 1 applies H_c to that column first, with a
//     reduction of that one dot product, and forms its reflector at raised
//     wave priority; only then does it apply H_c to its other columns (one
//     transposed reduction of all its dot products, which also yields the
//     Gram entries v_j^T v_c of its finished columns j < c for T).
// Scheduling barriers keep the compiler from moving the other columns'
// work into the hand-off chain (without them it did, and the chain ran four
// times longer).  Rows beyond m add exact zeros, so the bits do not depend
// on the instance (RPG) or on the timing of the waves either.  A column
// leaving the fast path's range makes the whole panel run again on LAPACK's
// scaled path, as in panel_body.
template <int RPG, int CPW, int NW>
__device__ __forceinline__ void panel_cyc_body(const PanelTask& t) {
  constexpr int NB = NW * CPW;
  constexpr int NT = 64 * NW;
  static_assert(NB < 64, "rows c and c + 1 of column c + 1 are lanes of the first row group");
  static_assert(2 * NB * (NB + 1) <= NB * RPG * 64, "build_T's scratch fits in Vl");
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int m = t.m, nbj = t.nbj, lda = t.lda;
  const gptr<double> Pg = glb(t.P);

  __shared__ double Vl[NB][RPG][64];  // v_c at rows lane + 64 q
  __shared__ double G[NB][NB + 1];
  __shared__ double taus[NB];
  __shared__ int ready;
  // polls of `ready` before a wave gives up on a column: far beyond any
  // real wait (a whole sub-panel takes well under a million clocks), so
  // reached only through a defect, which then ends in an error code, not a
  // hung GPU
  constexpr int kMaxPolls = 1 << 22;

  double a[RPG][CPW];  // column wave + NW jj, rows lane + 64 q
  auto run = [&](auto safe_tag) -> bool {
    constexpr bool SAFE = decltype(safe_tag)::value;
    bool bad = false, stuck = false;
    for (int e = tid; e < NB * (NB + 1); e += NT) (&G[0][0])[e] = 0.0;
    if (tid < NB) taus[tid] = 0.0;
    if (tid == 0) ready = 0;
#pragma unroll
    for (int q = 0; q < RPG; ++q) {
      const int row = lane + 64 * q;
#pragma unroll
      for (int jj = 0; jj < CPW; ++jj) {
        const int col = wave + NW * jj;
        a[q][jj] = (row < m && col < nbj) ? Pg[(size_t)col * lda + row] : 0.0;
      }
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0), as in panel_body
    __syncthreads();                     // G, taus and ready cleared before any publish

    double vo[RPG];  // v and tau of the last reflector this wave formed
    double tau_o = 0.0;
    // Reflector of column c (local column jc of the calling wave) from
    // alpha = a[c] and ||x||^2 over the rows below: v_c to Vl[c] and vo,
    // tau_c to taus[c]; the column becomes beta on the diagonal and v below
    // it; then ready = c + 1 (workgroup-scope release: the LDS writes are
    // complete before any wave can see the count).
    auto reflector = [&](int jc, int c, double alpha, double xn2) {
      const bool inrange = xn2 > 0x1p-960 && xn2 < 0x1p960 && fabs(alpha) < 0x1p480;
      double beta = -copysign(sqrt_normal(fma(alpha, alpha, xn2)), alpha);
      const double amb = alpha - beta;
      const double rr = rcp_normal(beta * amb);
      double tau = -(amb * amb) * rr;
      double scale = beta * rr;
      if constexpr (SAFE) {
        if (!inrange) {
          double x[RPG];
          x[0] = (lane > c) ? a[0][jc] : 0.0;
#pragma unroll
          for (int q = 1; q < RPG; ++q) x[q] = a[q][jc];
          const Larfg lf = dlarfg_scalar(alpha, scaled_norm_wave<RPG>(x, lane));
          tau = lf.tau;
          beta = lf.beta;
          scale = lf.scale;
        }
      } else {
        bad |= !inrange;
      }
      const double v0 = (lane > c) ? a[0][jc] * scale : (lane == c ? 1.0 : 0.0);
      Vl[c][0][lane] = v0;
      vo[0] = v0;
      a[0][jc] = (lane > c) ? v0 : (lane == c ? beta : a[0][jc]);
#pragma unroll
      for (int q = 1; q < RPG; ++q) {
        const double v = a[q][jc] * scale;
        Vl[c][q][lane] = v;
        vo[q] = v;
        a[q][jc] = v;
      }
      tau_o = tau;
      if (lane == 0) taus[c] = tau;
      __hip_atomic_store(&ready, c + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
      __builtin_amdgcn_sched_barrier(0);
    };
    // ||x||^2 of local column jc below row c, summed from the column
    auto plain_norm2 = [&](int jc, int c) {
      const double x0 = (lane > c) ? a[0][jc] : 0.0;
      double n0 = fma(x0, x0, 0.0), n1 = 0.0;
#pragma unroll
      for (int q = 1; q < RPG; ++q) {
        if (q & 1) n1 = fma(a[q][jc], a[q][jc], n1);
        else n0 = fma(a[q][jc], a[q][jc], n0);
      }
      return wave_sum(n0 + n1, lane);
    };

    // A wave with no column below nbj has nothing to do in the loop.
    if (wave < nbj) {
      if (wave == 0) reflector(0, 0, readlane64(a[0][0], 0), plain_norm2(0, 0));
      auto column_round = [&](auto jc_tag) {
        constexpr int jc = decltype(jc_tag)::value;
        for (int o = 0; o < NW; ++o) {
          const int c = jc * NW + o;  // the column whose reflector is applied
          if (c >= nbj) break;
          const bool crit = c + 1 < nbj && wave == (o + 1) % NW;
          double v[RPG];
          double tau;
          if (wave == o) {
#pragma unroll
            for (int q = 0; q < RPG; ++q) v[q] = vo[q];
            tau = tau_o;
          } else {
            int seen = __builtin_amdgcn_readfirstlane(
                __hip_atomic_load(&ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            for (int polls = 0; seen <= c; ++polls) {
              if (polls >= kMaxPolls) {
                stuck = true;
                break;
              }
              if (!crit) __builtin_amdgcn_s_sleep(1);
              seen = __builtin_amdgcn_readfirstlane(
                  __hip_atomic_load(&ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP));
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
            for (int q = 0; q < RPG; ++q) v[q] = Vl[c][q][lane];
            tau = taus[c];
          }
          // the next column, if this wave owns it: local index jc
          // (o < NW - 1) or jc + 1 (o = NW - 1)
          auto critical = [&](auto jn_tag) {
            constexpr int jn = decltype(jn_tag)::value;
            if constexpr (jn < CPW) {
              __builtin_amdgcn_sched_barrier(0);
              __builtin_amdgcn_s_setprio(3);
              const int c1 = c + 1;
              double s0 = fma(v[0], a[0][jn], 0.0), s1 = 0.0;
#pragma unroll
              for (int q = 1; q < RPG; ++q) {
                if (q & 1) s1 = fma(v[q], a[q][jn], s1);
                else s0 = fma(v[q], a[q][jn], s0);
              }
              const double w = tau * wave_sum(s0 + s1, lane);
#pragma unroll
              for (int q = 0; q < RPG; ++q) a[q][jn] = fma(-v[q], w, a[q][jn]);
              reflector(jn, c1, readlane64(a[0][jn], c1), plain_norm2(jn, c1));
              __builtin_amdgcn_s_setprio(0);
              __builtin_amdgcn_sched_barrier(0);
            }
          };
          if (crit) {
            if (o < NW - 1) critical(std::integral_constant<int, jc>{});
            else critical(std::integral_constant<int, jc + 1>{});
          }
          // the rest: every local column's dot product in one transposed
          // reduction; Gram entries of the finished columns, H_c applied to
          // the unfinished ones (the critical column is done already)
          double P[CPW];
#pragma unroll
          for (int jj = 0; jj < CPW; ++jj) {
            double s0 = fma(v[0], a[0][jj], 0.0), s1 = 0.0;
#pragma unroll
            for (int q = 1; q < RPG; ++q) {
              if (q & 1) s1 = fma(v[q], a[q][jj], s1);
              else s0 = fma(v[q], a[q][jj], s0);
            }
            P[jj] = s0 + s1;
          }
          wave_reduce_stage<CPW, 0, CPW>(P, lane);
          {
            const int jj = wave_reduce_entry<CPW>(lane);
            if ((lane & (64 / CPW - 1)) == 0 && wave + NW * jj < c) G[wave + NW * jj][c] = P[0];
          }
          const int skip = crit ? (o < NW - 1 ? jc : jc + 1) : -1;
#pragma unroll
          for (int jj = 0; jj < CPW; ++jj) {
            if (jj != skip && wave + NW * jj > c && wave + NW * jj < nbj) {
              const double w = tau * readlane64(P[0], wave_reduce_lane<CPW>(jj));
#pragma unroll
              for (int q = 0; q < RPG; ++q) a[q][jj] = fma(-v[q], w, a[q][jj]);
            }
          }
        }
      };
      static_for<CPW>(column_round);
    }
    const bool any_stuck = __syncthreads_or(stuck) != 0;
    if (any_stuck && tid == 0) *t.info = BSOFI_ERR_SCHEDULE;
    return __syncthreads_or(bad && !any_stuck) != 0;
  };
  if (run(std::false_type{})) run(std::true_type{});

  // G and taus are complete (the final barrier of run) and Vl is free for
  // build_T's scratch.  The panel's stores go out first and stay in flight
  // under build_T, whose barriers order LDS only.
#pragma unroll
  for (int q = 0; q < RPG; ++q) {
    const int row = lane + 64 * q;
#pragma unroll
    for (int jj = 0; jj < CPW; ++jj) {
      const int col = wave + NW * jj;
      if (row < m && col < nbj) Pg[(size_t)col * lda + row] = a[q][jj];
    }
  }
  if (tid < nbj) glb(t.tau)[tid] = taus[tid];
  double (*Tl)[NB + 1] = reinterpret_cast<double (*)[NB + 1]>(&Vl[0][0][0]);
  build_T<NB, NT>(G, taus, nbj, glb(t.T), t.ldt, Tl, Tl + NB);
}

template <int RPG, int CPW, int NW, bool ONE = false>
__global__ __launch_bounds__(64 * NW) void panel_cyc_kernel(const PanelTask* __restrict__ tasks,
                                                            const Item* __restrict__ items,
                                                            const PanelTask one) {
  panel_cyc_body<RPG, CPW, NW>(launch_task<ONE>(tasks, items, one));
}

// ---------------------------------------------------------------------------
// Block reflector application (K2)
// ---------------------------------------------------------------------------
// One workgroup = one 16-column slab of the target columns, all m rows (in
// the multi-slab form, spw consecutive slabs one after the other).
//   phase 1: W = V^T C        (nbj x 16, K = m split over the 4 or 8 waves)
//   phase 2: W <- -op(T) W    (partials summed in LDS, the product on MFMA)
//   phase 3: C <- C + V W     (m x 16, K = nbj)
__device__ __forceinline__ double vload(gptr<const double> V, int ldv, int m, int nbj, int row, int col) {
  if (row >= m || col >= nbj || row < col) return 0.0;
  if (row == col) return 1.0;
  return V[(size_t)col * ldv + row];
}

// The target slab lives in MFMA accumulator registers: NW waves (4 or 8;
// the engine launches the 8-wave form, whose sums the 4-wave form repeats
// in the same order); wave w owns row tiles w, w+NW, ... (RT of them,
// m <= 16 NW RT).  Eight
// waves halve the loads each lane has in flight at entry (a wave keeps at
// most 63 outstanding) and put two waves per SIMD.  The C layout of a
// 16 x 16 accumulator tile is exactly four consecutive B fragments of a
// K = 16 block (register r = rows 4r..4r+3), so phase 1 multiplies straight
// out of the accumulators.
// Global memory is only touched in "row-major lanes" layouts where the 16
// lanes of a quarter-wave read 16 consecutive rows of one column (128-byte
// segments): the slab as C[rt*16 + li][4r + lk] and V as
// V[rt*16 + li][4 kq + lk] (phase 3's A fragments).  A per-wave LDS tile
// turns them into the accumulator layout and phase 1's V^T fragments, and
// turns the result back for the store.
// MS: a workgroup takes slabs it.sub .. it.sub + spw - 1 of its task, one
// after the other: V (in registers) and T (in LDS) are loaded once, and the
// next slab's C is loaded while the current one goes through phases 2 and 3
// and is stored (a workgroup barrier waits for LDS operations only, so the
// loads stay in flight across it).  Without MS a workgroup takes the one slab
// it.sub.  Every slab goes through the same arithmetic either way.
// The body for slab it.sub of task t (or slabs it.sub .. it.sub + spw - 1
// with MS) by one workgroup of 64 NW threads.
template <int RT, int NW, bool MS = false>
__device__ __forceinline__ void refl_body(const ReflTask& t, const Item it, int spw) {
  static_assert(!MS || RT <= 4, "the multi-slab form holds V in registers");
  constexpr int NT = 64 * NW;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m = t.m, nbj = t.nbj;
  const int sub_end = MS ? min(it.sub + spw, (t.w1 + t.w2 + 15) / 16) : it.sub + 1;

  __shared__ double Wf[32][17];
  __shared__ double Tl[32][33];
  // per wave: the transpose tile X (16 x 33) in phases 1 and 3, the wave's
  // share of W (32 x 17) between them
  __shared__ double XW[NW][32 * 17];
  auto Wp = [&](int w) { return reinterpret_cast<double (*)[17]>(XW[w]); };

  // the slab column of register r in the row-major-lanes layout
  auto columns = [&](int sub, gptr<double>(&cp)[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int vcol = sub * 16 + 4 * r + lk;
      cp[r] = nullptr;
      if (vcol < t.w1)
        cp[r] = glb(t.C1) + (size_t)vcol * t.ldc;
      else if (vcol - t.w1 < t.w2)
        cp[r] = glb(t.C2) + (size_t)(vcol - t.w1) * t.ldc;
    }
  };
  gptr<double> cp[4];
  columns(it.sub, cp);
  const gptr<const double> Vg = glb(t.V);
  // T goes to registers first and into LDS once the C and V loads are out,
  // so that their latencies overlap (the vector memory counter is in order)
  constexpr int TE = 32 * 32 / NT;
  double tv[TE];
#pragma unroll
  for (int q = 0; q < TE; ++q) {
    const int e = tid + q * NT, i = e & 31, j = e >> 5;
    tv[q] = (i < nbj && j < nbj) ? glb(t.T)[(size_t)j * t.ldt + i] : 0.0;
  }
  // Registers hold V up front for up to 4 row tiles per wave; beyond that
  // (8 waves x 8 tiles, m <= 1024) V tiles are loaded when used (same
  // layout).
  constexpr bool PREV = RT <= 4;
  double cl[RT][4];              // C[rt*16 + li][4r + lk]
  double vb[PREV ? RT : 1][8];   // V[rt*16 + li][4kq + lk]
  auto load_c = [&](const gptr<double>(&c)[4]) {
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      const int row = (wave + NW * s) * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) cl[s][r] = (c[r] && row < m) ? c[r][row] : 0.0;
    }
  };
  load_c(cp);
  if constexpr (PREV) {
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      const int row = (wave + NW * s) * 16 + li;
#pragma unroll
      for (int kq = 0; kq < 8; ++kq) vb[s][kq] = vload(Vg, t.ldv, m, nbj, row, 4 * kq + lk);
    }
  }
#pragma unroll
  for (int q = 0; q < TE; ++q) {
    const int e = tid + q * NT;
    Tl[e & 31][e >> 5] = tv[q];
  }
  auto vtile = [&](int s, int kq) -> double {
    if constexpr (PREV) return vb[s][kq];
    else return vload(Vg, t.ldv, m, nbj, (wave + NW * s) * 16 + li, 4 * kq + lk);
  };
  double (*X)[33] = reinterpret_cast<double (*)[33]>(XW[wave]);
  const bool two = nbj > 16;
  __shared__ double W[32][17];

  for (int sub = it.sub; sub < sub_end; ++sub) {
    // phase 1: W = V^T C.  The partial sums follow one order for both wave
    // counts, so that a task gets the same bits from either kernel: row tile
    // t goes to chain (t mod 4, (t / 4) mod 2), each chain adds its tiles in
    // increasing t, and W = sum over q = 0..3 of (chain (q,0) + chain (q,1)).
    // With 4 waves, wave q holds both chains of q (accumulators a and b, by
    // the parity of s); with 8 waves, wave w holds the one chain
    // (w mod 4, w / 4), and phase 2 adds waves q and q + 4 first.
    d4 acc[RT];
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      // slab tile -> accumulator layout C[4r + lk][li]
#pragma unroll
      for (int r = 0; r < 4; ++r) X[li][4 * r + lk] = cl[s][r];
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[s][r] = X[4 * r + lk][li];
    }
    d4 w0a = d4zero(), w0b = d4zero(), w1a = d4zero(), w1b = d4zero();
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      // V tile -> V^T fragments V[4r + lk][li], V[4r + lk][16 + li]
#pragma unroll
      for (int kq = 0; kq < 8; ++kq) X[li][4 * kq + lk] = vtile(s, kq);
      double va[4][2];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        va[r][0] = X[4 * r + lk][li];
        va[r][1] = X[4 * r + lk][16 + li];
      }
      if ((wave + NW * s) * 16 < m) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (NW == 4 && (s & 1)) {
            w0b = mfma(va[r][0], acc[s][r], w0b);
            if (two) w1b = mfma(va[r][1], acc[s][r], w1b);
          } else {
            w0a = mfma(va[r][0], acc[s][r], w0a);
            if (two) w1a = mfma(va[r][1], acc[s][r], w1a);
          }
        }
      }
    }
    const d4 w0 = NW == 8 ? w0a : w0a + w0b, w1 = NW == 8 ? w1a : w1a + w1b;
    {
      double (*Wq)[17] = Wp(wave);  // the wave's X tile is done with
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Wq[crow(lk, r)][li] = w0[r];
        Wq[16 + crow(lk, r)][li] = two ? w1[r] : 0.0;
      }
    }
    // the next slab's C is in flight from here to its phase 1
    gptr<double> cn[4];
    const bool more = MS && sub + 1 < sub_end;
    if (more) {
      columns(sub + 1, cn);
      load_c(cn);
    }
    __syncthreads();
    // phase 2: Wf = -op(T) W
    for (int e = tid; e < 32 * 16; e += NT) {
      const int i = e >> 4, j = e & 15;
      double v = 0.0;
      if constexpr (NW == 8) {
#pragma unroll
        for (int q = 0; q < 4; ++q) v += Wp(q)[i][j] + Wp(q + 4)[i][j];
      } else {
#pragma unroll
        for (int q = 0; q < NW; ++q) v += Wp(q)[i][j];
      }
      W[i][j] = v;
    }
    __syncthreads();
    // op(T) W on MFMA: waves 0 and 1 form rows 0-15 and 16-31 (K = 32).
    // Tl is zero outside the nbj x nbj triangle, so no masking is needed.
    if (wave < 2) {
      const int i = wave * 16 + li;
      d4 tw = d4zero();
#pragma unroll
      for (int kc = 0; kc < 8; ++kc) {
        const int k = 4 * kc + lk;
        tw = mfma(t.trans ? Tl[k][i] : Tl[i][k], W[k][li], tw);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) Wf[wave * 16 + crow(lk, r)][li] = -tw[r];
    }
    __syncthreads();
    // phase 3: C += V Wf, then back to the row-major-lanes layout and out
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      const int rt = wave + NW * s;
      if (rt * 16 < m) {
#pragma unroll
        for (int kq = 0; kq < 8; ++kq)
          if (4 * kq < nbj) acc[s] = mfma(vtile(s, kq), Wf[4 * kq + lk][li], acc[s]);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) X[crow(lk, r)][li] = acc[s][r];
      const int row = rt * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double v = X[li][4 * r + lk];
        if (cp[r] && row < m) cp[r][row] = v;
      }
    }
    if (more) {
#pragma unroll
      for (int r = 0; r < 4; ++r) cp[r] = cn[r];
    }
  }
}

template <int RT, int NW, bool MS = false, bool ONE = false>
__global__ __launch_bounds__(64 * NW) void refl_apply_kernel(const ReflTask* __restrict__ tasks,
                                                             const Item* __restrict__ items, int spw,
                                                             const ReflTask one) {
  // one task: workgroup g takes slabs g spw .. g spw + spw - 1
  const Item it = ONE ? Item{0, (int)blockIdx.x * spw} : items[blockIdx.x];
  refl_body<RT, NW, MS>(launch_task<ONE>(tasks, items, one), it, spw);
}

// ---------------------------------------------------------------------------
// Explicit Q_k in one launch (BSOI's dorgqr)
// ---------------------------------------------------------------------------
// Q_k = H_0 ... H_{S-1} applied to the identity backward (j = S-1 .. 0),
// H_j = I - V_j T_j V_j^T acting on rows and columns >= j nb.  One
// workgroup (8 waves) per 16-column slab of Q_k keeps the slab's n2 rows
// in MFMA accumulators (RT row tiles per wave, wave w holding tiles w + 8 s)
// from the identity to the end, and applies every H_j that reaches the slab
// (j nb <= its first column) with refl_body's arithmetic: the reflector
// rounds of one launch each, from and to memory, give the same bits.  Round
// j counts row tiles from its first row (t0 = j nb / 16), so relative tile
// t = t_abs - t0 belongs to chain ((t mod 4), (t / 4) mod 2); every tile a
// wave holds has the same t mod 8, so each wave forms one chain, in
// increasing t, and phase 2 adds the chains in refl_body's order.
template <int RT>
__global__ __launch_bounds__(512) void qform_kernel(const QformTask* __restrict__ tasks,
                                                    const Item* __restrict__ items) {
  constexpr int NW = 8, NT = 512;
  const Item it = items[blockIdx.x];
  const QformTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int n2 = t.n2, nb = t.nb, slab = it.sub;

  __shared__ double Wf[32][17];
  __shared__ double Tl[32][33];
  __shared__ double XW[NW][32 * 17];
  __shared__ double W[32][17];
  auto Wp = [&](int w) { return reinterpret_cast<double (*)[17]>(XW[w]); };
  double (*X)[33] = reinterpret_cast<double (*)[33]>(XW[wave]);

  d4 acc[RT];  // C[4r + lk][li] of row tile wave + 8 s: the identity to start
#pragma unroll
  for (int s = 0; s < RT; ++s)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = (wave + NW * s) * 16 + 4 * r + lk, col = 16 * slab + li;
      acc[s][r] = (row == col && col < n2) ? 1.0 : 0.0;
    }

  const int jmax = min(t.S - 1, (16 * slab) / nb);
  for (int j = jmax; j >= 0; --j) {
    const int m = n2 - j * nb, nbj = min(nb, t.ncols - j * nb), t0 = (j * nb) / 16;
    const gptr<const double> Vg = glb(t.V0) + (size_t)j * nb * (t.ldv + 1);
    const gptr<const double> Tg = glb(t.T0) + (size_t)j * nb * nb;
    const bool two = nbj > 16;
    // T to registers, V tiles of the rows this reflector reaches, then T to LDS
    constexpr int TE = 32 * 32 / NT;
    double tv[TE];
#pragma unroll
    for (int q = 0; q < TE; ++q) {
      const int e = tid + q * NT, i = e & 31, c = e >> 5;
      tv[q] = (i < nbj && c < nbj) ? Tg[(size_t)c * nb + i] : 0.0;
    }
    double vb[RT][8];  // V[t * 16 + li][4 kq + lk], t = wave + 8 s - t0 (relative tile)
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      const int trel = wave + NW * s - t0;
#pragma unroll
      for (int kq = 0; kq < 8; ++kq)
        vb[s][kq] = trel >= 0 ? vload(Vg, t.ldv, m, nbj, trel * 16 + li, 4 * kq + lk) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < TE; ++q) {
      const int e = tid + q * NT;
      Tl[e & 31][e >> 5] = tv[q];
    }
    // phase 1: W = V^T C, this wave's one chain
    d4 w0a = d4zero(), w1a = d4zero();
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      const int trel = wave + NW * s - t0;
      if (trel < 0) continue;
#pragma unroll
      for (int kq = 0; kq < 8; ++kq) X[li][4 * kq + lk] = vb[s][kq];
      double va[4][2];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        va[r][0] = X[4 * r + lk][li];
        va[r][1] = X[4 * r + lk][16 + li];
      }
      if (trel * 16 < m) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          w0a = mfma(va[r][0], acc[s][r], w0a);
          if (two) w1a = mfma(va[r][1], acc[s][r], w1a);
        }
      }
    }
    {
      double (*Wq)[17] = Wp(wave);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        Wq[crow(lk, r)][li] = w0a[r];
        Wq[16 + crow(lk, r)][li] = two ? w1a[r] : 0.0;
      }
    }
    __syncthreads();
    // phase 2: W = sum over q of (chain (q, 0) + chain (q, 1)); chain c is
    // in wave (c + t0) mod 8
    for (int e = tid; e < 32 * 16; e += NT) {
      const int i = e >> 4, c = e & 15;
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) v += Wp((q + t0) & 7)[i][c] + Wp((q + 4 + t0) & 7)[i][c];
      W[i][c] = v;
    }
    __syncthreads();
    // Wf = -T W (waves 0 and 1: rows 0-15 and 16-31, K = 32)
    if (wave < 2) {
      const int i = wave * 16 + li;
      d4 tw = d4zero();
#pragma unroll
      for (int kc = 0; kc < 8; ++kc) {
        const int k = 4 * kc + lk;
        tw = mfma(Tl[i][k], W[k][li], tw);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) Wf[wave * 16 + crow(lk, r)][li] = -tw[r];
    }
    __syncthreads();
    // phase 3: C += V Wf on the rows this reflector reaches
#pragma unroll
    for (int s = 0; s < RT; ++s) {
      const int trel = wave + NW * s - t0;
      if (trel >= 0 && trel * 16 < m) {
#pragma unroll
        for (int kq = 0; kq < 8; ++kq)
          if (4 * kq < nbj) acc[s] = mfma(vb[s][kq], Wf[4 * kq + lk][li], acc[s]);
      }
    }
    __syncthreads();  // LDS (Tl, W, Wf, the wave tiles) free for the next reflector
  }
  // the slab, back to the row-major-lanes layout and out
#pragma unroll
  for (int s = 0; s < RT; ++s) {
    const int rt = wave + NW * s;
#pragma unroll
    for (int r = 0; r < 4; ++r) X[crow(lk, r)][li] = acc[s][r];
    const int row = rt * 16 + li;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double v = X[li][4 * r + lk];
      const int col = 16 * slab + 4 * r + lk;
      if (row < n2 && col < n2) glb(t.Q)[(size_t)col * n2 + row] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Device-scheduled BSOF sweep
// ---------------------------------------------------------------------------
// Number of sub-panels before g (step-major index within one matrix) whose
// reflectors act on column c.  Column c lies in block column b:
//   * b = p-1 (the last block column) is target 2 of every sub-panel of
//     steps 0 .. p-3: indexes [0, s0 (p-2));
//   * 1 <= b <= p-2 is target 1 of every sub-panel of step b-1:
//     [s0 (b-1), s0 b);
//   * then, inside the step kc whose panel holds c (kc = b, or p-2 for the
//     last two block columns), every earlier sub-panel of that step:
//     [s0 kc, s0 kc + jc), jc the sub-panel holding c.
// (Steps 0 .. p-3 have s0 sub-panels each, so step k starts at s0 k.)
__device__ __forceinline__ int flow_touch_before(const FlowMat& M, int c, int g) {
  const int n = M.n, p = M.p, s0 = M.s0;
  const int b = c / n;
  int a1 = 0, b1 = 0;
  if (b == p - 1) {
    b1 = (p - 2) * s0;
  } else if (b >= 1) {
    a1 = (b - 1) * s0;
    b1 = b * s0;
  }
  const int kc = b <= p - 3 ? b : p - 2;
  const int a2 = kc * s0, b2 = a2 + (c - kc * n) / M.nb;
  return max(0, min(b1, g) - a1) + max(0, min(b2, g) - a2);
}

__device__ __forceinline__ int ld_agent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(int* p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Column of a task that lane `lane` of wave 0 watches (-1: none): the
// sub-panel's own columns for a QR, the slab's 16 columns for a reflector
// task.
__device__ __forceinline__ int flow_column(const FlowTask& f, const PanelTask* pt, const ReflTask* rt, int lane) {
  if (f.kind == 0) return lane < pt[f.ref].nbj ? f.col0 + lane : -1;
  if (lane >= 16) return -1;
  const ReflTask& t = rt[f.ref];
  const int vcol = f.sub * 16 + lane;
  if (vcol < t.w1) return f.col0 + vcol;
  if (vcol - t.w1 < t.w2) return f.col2 + (vcol - t.w1);
  return -1;
}

// One persistent launch of 512-thread workgroups.  A workgroup takes the
// next task from the queue (ctl[16] is its head, ft[0 .. ntasks) the
// tasks); wave 0 waits until every column the task touches has its counter
// at the task's count (and, for a reflector task, until its sub-panel's QR
// is stored), then the workgroup runs the task with the panel / reflector
// bodies of the per-step kernels (the same arithmetic, so the same bits),
// waits for its stores, and wave 0 publishes the new counters or the
// sub-panel flag.  Publication is a device-scope release (the XCD's L2 is
// written back, so a workgroup on another XCD sees the data), the wait a
// device-scope acquire (stale L1 / L2 lines are dropped).  The queue lists
// tasks in an order where every task's inputs come from earlier tasks, so
// the earliest unfinished task can always run and the sweep cannot
// deadlock, whatever number of workgroups is resident.  Every wait is
// bounded by wait_ticks of the constant wall clock: past it the workgroup
// sets ctl[1], every workgroup stops taking tasks and leaves, and
// flow_check_kernel reports the error through the matrices' info words.
template <int RPG, int RT>
__global__ __launch_bounds__(512) void bsof_flow_kernel(const FlowTask* __restrict__ ft, int ntasks,
                                                        const PanelTask* __restrict__ pt,
                                                        const ReflTask* __restrict__ rt,
                                                        const FlowMat* __restrict__ fm, int* ctl,
                                                        long long wait_ticks) {
  __shared__ int s_task, s_ok;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  for (;;) {
    if (tid == 0) s_task = ld_agent(ctl + 1) != 0 ? ntasks : atomicAdd(ctl + 16, 1);
    __syncthreads();
    const int i = s_task;
    if (i >= ntasks) break;
    const FlowTask f = ft[i];
    const FlowMat M = fm[f.mat];
    if (wave == 0) {
      const int c = flow_column(f, pt, rt, lane);
      const int need = c >= 0 ? flow_touch_before(M, c, f.g) : 0;
      const long long t0 = wall_clock64();
      bool ok = true;
      for (;;) {
        bool ready = c < 0 || ld_agent(M.ctr + c) >= need;
        if (f.kind == 1) ready = ready && ld_agent(M.done + f.g) != 0;
        if (__all(ready)) break;
        if (ld_agent(ctl + 1) != 0 || wall_clock64() - t0 > wait_ticks) {
          ok = false;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if (lane == 0) {
        if (!ok) st_agent(ctl + 1, 1);
        s_ok = ok;
      }
      // what the counters published is visible to this workgroup
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    if (!s_ok) break;
    if (f.kind == 0)
      panel_body<RPG, 4, 8>(pt[f.ref]);
    else
      refl_body<RT, 8, false>(rt[f.ref], Item{f.ref, f.sub}, 1);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores have reached L2
    __syncthreads();
    if (wave == 0) {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      // The fence's L2 write-back (buffer_wbl2) is counted in vmcnt, but the
      // compiler drops the wait that should follow it when no other memory
      // operation is outstanding; without this wait a flag could reach memory
      // before the data it publishes and a workgroup on another XCD could read
      // stale values.
      __builtin_amdgcn_s_waitcnt(0x0F70);
      if (f.kind == 0) {
        if (lane == 0) st_agent(M.done + f.g, 1);
      } else {
        const int c = flow_column(f, pt, rt, lane);
        if (c >= 0) st_agent(M.ctr + c, flow_touch_before(M, c, f.g) + 1);
      }
    }
  }
}

// After a flow launch: a sweep that stopped on a wait limit marks every
// matrix of the plan with info = code.
__global__ void flow_check_kernel(const int* __restrict__ ctl, int* __restrict__ info, int count, int code) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count && ctl[1] != 0) info[e] = code;
}

// Reflector application for slabs taller than the register-resident forms
// hold (m > 1024, block sizes n > 512): the same three phases and the same
// summation order as refl_apply_kernel with 8 waves (wave w takes row tiles
// w, w + 8, ... in increasing order into one chain per 16 columns of W), but
// no tile is held between phases: phase 1 streams the slab's row tiles
// through the accumulator layout, phase 3 loads each tile again, adds V Wf
// and stores it.  A task therefore gets the bits it gets from the resident
// form at any m, for twice the slab's read traffic.
__global__ __launch_bounds__(512) void refl_apply_stream_kernel(const ReflTask* __restrict__ tasks,
                                                                const Item* __restrict__ items) {
  constexpr int NW = 8, NT = 512;
  const Item it = items[blockIdx.x];
  const ReflTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int m = t.m, nbj = t.nbj;

  __shared__ double Wf[32][17];
  __shared__ double Tl[32][33];
  __shared__ double XW[NW][32 * 17];
  __shared__ double W[32][17];
  auto Wp = [&](int w) { return reinterpret_cast<double (*)[17]>(XW[w]); };

  gptr<double> cp[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int vcol = it.sub * 16 + 4 * r + lk;
    cp[r] = nullptr;
    if (vcol < t.w1)
      cp[r] = glb(t.C1) + (size_t)vcol * t.ldc;
    else if (vcol - t.w1 < t.w2)
      cp[r] = glb(t.C2) + (size_t)(vcol - t.w1) * t.ldc;
  }
  const gptr<const double> Vg = glb(t.V);
  for (int e = tid; e < 32 * 32; e += NT) {
    const int i = e & 31, j = e >> 5;
    Tl[i][j] = (i < nbj && j < nbj) ? glb(t.T)[(size_t)j * t.ldt + i] : 0.0;
  }
  double (*X)[33] = reinterpret_cast<double (*)[33]>(XW[wave]);
  const bool two = nbj > 16;
  // one row tile of the slab into the accumulator layout C[4r + lk][li]
  auto load_tile = [&](int rt, d4& acc) {
    const int row = rt * 16 + li;
    double cl[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) cl[r] = (cp[r] && row < m) ? cp[r][row] : 0.0;
#pragma unroll
    for (int r = 0; r < 4; ++r) X[li][4 * r + lk] = cl[r];
#pragma unroll
    for (int r = 0; r < 4; ++r) acc[r] = X[4 * r + lk][li];
  };

  // phase 1: W = V^T C, wave w's chain over tiles w, w + 8, ...
  d4 w0 = d4zero(), w1 = d4zero();
  for (int rt = wave; rt * 16 < m; rt += NW) {
    d4 acc;
    load_tile(rt, acc);
    const int row = rt * 16 + li;
#pragma unroll
    for (int kq = 0; kq < 8; ++kq) X[li][4 * kq + lk] = vload(Vg, t.ldv, m, nbj, row, 4 * kq + lk);
    double va[4][2];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      va[r][0] = X[4 * r + lk][li];
      va[r][1] = X[4 * r + lk][16 + li];
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      w0 = mfma(va[r][0], acc[r], w0);
      if (two) w1 = mfma(va[r][1], acc[r], w1);
    }
  }
  {
    double (*Wq)[17] = Wp(wave);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      Wq[crow(lk, r)][li] = w0[r];
      Wq[16 + crow(lk, r)][li] = two ? w1[r] : 0.0;
    }
  }
  __syncthreads();
  // phase 2: Wf = -op(T) W, partial chains summed as refl_apply_kernel's
  for (int e = tid; e < 32 * 16; e += NT) {
    const int i = e >> 4, j = e & 15;
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) v += Wp(q)[i][j] + Wp(q + 4)[i][j];
    W[i][j] = v;
  }
  __syncthreads();
  if (wave < 2) {
    const int i = wave * 16 + li;
    d4 tw = d4zero();
#pragma unroll
    for (int kc = 0; kc < 8; ++kc) {
      const int k = 4 * kc + lk;
      tw = mfma(t.trans ? Tl[k][i] : Tl[i][k], W[k][li], tw);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) Wf[wave * 16 + crow(lk, r)][li] = -tw[r];
  }
  __syncthreads();
  // phase 3: C += V Wf tile by tile (each tile loaded again)
  for (int rt = wave; rt * 16 < m; rt += NW) {
    d4 acc;
    load_tile(rt, acc);
    const int row = rt * 16 + li;
#pragma unroll
    for (int kq = 0; kq < 8; ++kq)
      if (4 * kq < nbj) acc = mfma(vload(Vg, t.ldv, m, nbj, row, 4 * kq + lk), Wf[4 * kq + lk][li], acc);
#pragma unroll
    for (int r = 0; r < 4; ++r) X[crow(lk, r)][li] = acc[r];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double v = X[li][4 * r + lk];
      if (cp[r] && row < m) cp[r][row] = v;
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
    for (int a = 0; a < 2; ++a) f[a] = (row < ke) ? vload(glb((const double*)t.P), t.lda, m, nbj, row, a * 16 + li) : 0.0;
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
  if (tid < 32) taus[tid] = (tid < nbj) ? glb(t.tau)[tid] : 0.0;
  __syncthreads();
  for (int e = tid; e < 32 * 32; e += 256) {
    const int i = e >> 5, j = e & 31;
    G[i][j] = (Gp[0][i][j] + Gp[1][i][j]) + (Gp[2][i][j] + Gp[3][i][j]);
  }
  __syncthreads();
  build_T<32>(G, taus, nbj, glb(t.T), t.ldt, &Gp[0][0], &Gp[1][0]);
}

// ---------------------------------------------------------------------------
// Determinant from the factorization A = Q R of BSOF (one workgroup per
// matrix).  Every Householder reflector with tau != 0 has determinant -1
// (tau v^T v = 2 for dlarfg's reflectors), one with tau = 0 is the
// identity, so
//   sign(det A) = (-1)^#{tau_i != 0} prod sign(R_ii),
//   log|det A|  = sum log|R_ii|.
// A zero R_ii gives sign 0 and log -inf.  Fixed reduction order, so the
// result is the same bits in a batch and alone.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void det_kernel(const DetTask* __restrict__ tasks) {
  const DetTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x;
  __shared__ double ls[256];
  __shared__ int par[256];
  double lg = 0.0;
  int flips = 0, zero = 0;
  for (int i = tid; i < t.N; i += 256) {
    const double r = glb(t.A)[(size_t)i * t.lda + i];
    zero |= (r == 0.0);
    flips ^= (r < 0.0) ^ (glb(t.tau)[i] != 0.0);
    lg += log(fabs(r));
  }
  ls[tid] = lg;
  par[tid] = flips | (zero << 1);
  __syncthreads();
  for (int h = 128; h >= 1; h >>= 1) {
    if (tid < h) {
      ls[tid] += ls[tid + h];
      par[tid] = ((par[tid] ^ par[tid + h]) & 1) | ((par[tid] | par[tid + h]) & 2);
    }
    __syncthreads();
  }
  if (tid == 0) {
    glb(t.out)[0] = (par[0] & 2) ? 0.0 : ((par[0] & 1) ? -1.0 : 1.0);
    glb(t.out)[1] = ls[0];
  }
}

__global__ void fill_int_kernel(int* __restrict__ p, int count, int value) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < count) p[e] = value;
}

// ---------------------------------------------------------------------------
// Tile GEMM (K4): C = alpha op(A) op(B) + beta C (+ D), 32 WT x 32 WT output
// tile per workgroup (WT = 1: 32 x 32, 2: 64 x 64, 4: 128 x 128), each wave a
// 16 WT x 16 WT quarter (WT x WT MFMA tiles).  K runs in chunks of 16
// through two LDS buffers: the global loads of chunk c + 1 are in flight
// while chunk c is multiplied, one barrier per chunk.  Global reads are
// unit-stride across threads (rows of A, k of B).
// ---------------------------------------------------------------------------
template <int WT>
__global__ __launch_bounds__(256, 2) void gemm_kernel(const GemmTask* __restrict__ tasks,
                                                   const Item* __restrict__ items) {
  constexpr int TM = 32 * WT;          // tile edge
  constexpr int AK = 256 / TM;         // k rows of A per pass of the block
  constexpr int NA = 16 / AK;          // A values per thread per chunk
  constexpr int NBV = TM / 16;         // B values per thread per chunk
  const Item it = items[blockIdx.x];
  const GemmTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int tm = it.sub / t.tiles_n, tn = it.sub % t.tiles_n;
  const int i0 = tm * TM, j0 = tn * TM;
  const int wi = wave & 1, wj = wave >> 1;
  const int M = t.M, N = t.N, K = t.K;
  const gptr<const double> Ag = glb(t.A), Bg = glb(t.B), Dg = glb(t.D);
  const gptr<double> Cg = glb(t.C);

  __shared__ double As[2][TM][17];
  __shared__ double Bs[2][TM][17];

  const int ia = tid % TM, ka = tid / TM;
  const int kb = tid & 15, jb = tid >> 4;
  double ra[NA], rb[NBV];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const int gi = i0 + ia, gk = k0 + ka + AK * r;
      ra[r] = (gi < M && gk < K && !(t.amask && gi > gk)) ? Ag[(size_t)gk * t.lda + gi] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NBV; ++r) {
      const int gk = k0 + kb, gj = j0 + jb + 16 * r;
      rb[r] = (gk < K && gj < N && !(t.bmask && gk > gj)) ? Bg[(size_t)gj * t.ldb + gk] : 0.0;
    }
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int r = 0; r < NA; ++r) As[buf][ia][ka + AK * r] = ra[r];
#pragma unroll
    for (int r = 0; r < NBV; ++r) Bs[buf][jb + 16 * r][kb] = rb[r];
  };

  d4 acc[WT][WT];
#pragma unroll
  for (int a = 0; a < WT; ++a)
#pragma unroll
    for (int b = 0; b < WT; ++b) acc[a][b] = d4zero();

  const int nch = (K + 15) / 16;
  fetch(0);
  stash(0);
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    const int buf = c & 1;
    if (c + 1 < nch) fetch(16 * (c + 1));
#pragma unroll
    for (int ks = 0; ks < 16; ks += 4) {
      double af[WT], bf[WT];
#pragma unroll
      for (int a = 0; a < WT; ++a) af[a] = As[buf][wi * 16 * WT + a * 16 + li][ks + lk];
#pragma unroll
      for (int b = 0; b < WT; ++b) bf[b] = Bs[buf][wj * 16 * WT + b * 16 + li][ks + lk];
#pragma unroll
      for (int a = 0; a < WT; ++a)
#pragma unroll
        for (int b = 0; b < WT; ++b) acc[a][b] = mfma(af[a], bf[b], acc[a][b]);
    }
    if (c + 1 < nch) stash(buf ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < WT; ++a)
#pragma unroll
    for (int b = 0; b < WT; ++b) {
      const int col = j0 + wj * 16 * WT + b * 16 + li;
      if (col >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + wi * 16 * WT + a * 16 + crow(lk, r);
        if (row >= M) continue;
        double v = t.alpha * acc[a][b][r];
        const gptr<double> cp = Cg + (size_t)col * t.ldc + row;
        if (t.beta != 0.0) v = fma(t.beta, *cp, v);
        if (t.D && col >= t.dcol0) v += Dg[(size_t)(col - t.dcol0) * t.ldd + row];
        *cp = v;
      }
    }
}

// ---------------------------------------------------------------------------
// BSTRI row sweep by column strips (K4s)
// ---------------------------------------------------------------------------
// One workgroup (8 waves) per 32-column strip of a block column b >= 1 runs
// the strip's whole chain X_{i,b} = W_i X_{i+1,b} (+ Z2_i) for
// i = b - 1 .. 0: the strip of X_{i+1,b} (n x 32) is the product's B, in
// LDS, and W_i streams from L2 into the A fragments, so the sweep's p - 1
// dependent GEMM launches become one launch whose strips run side by side.
// Every element is gemm_kernel's: one v_mfma_f64_16x16x4f64 chain over k in
// steps of 4, increasing, times alpha = 1, plus Z2_i on the last block
// column, the diagonal block read through the same upper-triangular mask;
// so the bits are the same (for n not a multiple of 32, MASK, rows, columns
// and k beyond n enter as zeros, as in gemm_kernel).  The planner deals the strips of one block
// column to one XCD: they start together and step through the same W_i,
// which each XCD's L2 then fetches once.  NW waves; wave w holds row tiles
// w, w + NW (RTW of them: n <= 16 NW RTW, 32 | n) and both 16-column tiles
// of the strip.  The strip is double-buffered in LDS with a row stride of
// SLD >= n doubles (SLD = 4 mod 32, a stride of 8 banks: the fragment reads
// and the result writes are free of bank conflicts): step i reads one buffer and writes
// its result to the other, one barrier a step.  For n > 128 the 2 x 65 KiB
// keep one workgroup per CU, so that a long chain never shares its CU's
// matrix cores; the n <= 64 and n <= 128 instances fit four and two.
template <int RTW, int NW, int SLD, bool MASK>
__global__ __launch_bounds__(64 * NW) void sweep_kernel(const SweepTask* __restrict__ tasks,
                                                        const Item* __restrict__ items) {
  constexpr int PD = 8;  // k-steps of W_i in flight
  const Item it = items[blockIdx.x];
  const SweepTask t = tasks[it.task];
  if (t.b <= 0) return;  // an empty task (the planner's padding)
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int n = t.n, p = t.p, lda = t.lda;
  const int b = t.b, c0 = it.sub * 32;
  // k runs to n32 in chunks of 4 PD: past n both operands are zero, exact
  // no-ops on the accumulators (which are never -0), as in gemm_kernel's
  // zero-padded last k chunk
  const int n32 = MASK ? (n + 31) & ~31 : n;
  __shared__ double Bs[2][32][SLD];  // B[k][col] at Bs[buf][col][k]
  const gptr<double> A = glb(t.A);
  for (int e = tid; e < 32 * n32; e += NW * 64) {
    const int k = e % n32, col = e / n32, gc = c0 + col;
    Bs[0][col][k] = (gc < n && k <= gc) ? A[(size_t)(b * n + gc) * lda + b * n + k] : 0.0;
    if (MASK) Bs[1][col][k] = 0.0;  // rows the results never reach stay zero
  }
  // Row tiles of this wave (uniform), and each lane's row of W_i: a tile
  // beyond n reads tile 0's rows (in bounds) and is never multiplied; a row
  // beyond n (MASK) reads row n - 1 and takes zero.
  bool has[RTW], rok[RTW];
  int wrow[RTW];
#pragma unroll
  for (int s = 0; s < RTW; ++s) {
    const int rt = wave + NW * s;
    has[s] = rt * 16 < n;
    const int row = (has[s] ? rt * 16 : 0) + li;
    rok[s] = row < n;
    wrow[s] = MASK ? min(row, n - 1) : row;
  }
  // A fragments W_i[rt 16 + li][k + lk], PD k-steps ahead across steps
  const gptr<const double> W = glb(t.W);
  const size_t nn = (size_t)n * n;
  // (MASK: the zero for a row or k beyond n is selected when the fragment
  // is used, not when it arrives, so that the loads stay in flight)
  double a[PD][RTW];
  bool am[PD][RTW];
  auto load = [&](int u, const gptr<const double> Wb, int k) {
    const int kk = k + lk;
#pragma unroll
    for (int s = 0; s < RTW; ++s) {
      if (MASK) {
        a[u][s] = Wb[(size_t)min(kk, n - 1) * n + wrow[s]];
        am[u][s] = kk < n && rok[s];
      } else {
        a[u][s] = Wb[(size_t)kk * n + wrow[s]];
      }
    }
  };
#pragma unroll
  for (int u = 0; u < PD; ++u) load(u, W + (size_t)(b - 1) * nn, 4 * u);
  __syncthreads();
  int cur = 0;
  for (int i = b - 1; i >= 0; --i) {
    const gptr<const double> Wi = W + (size_t)i * nn;
    d4 acc[RTW][2];
#pragma unroll
    for (int s = 0; s < RTW; ++s) acc[s][0] = acc[s][1] = d4zero();
    for (int k0 = 0; k0 < n32; k0 += 4 * PD) {
      // the chunk the slots load next: the rest of W_i, then the start of
      // W_{i-1}; after W_0 they reload W_0's first rows (in bounds, never
      // used): the loads stay unconditional, so the compiler keeps PD of
      // them in flight instead of draining them at a branch
      const bool in_i = k0 + 4 * PD < n32;
      const gptr<const double> Wb = in_i ? Wi : (i > 0 ? Wi - nn : W);
      const int kb = in_i ? k0 + 4 * PD : 0;
#pragma unroll
      for (int u = 0; u < PD; ++u) {
        const int k = k0 + 4 * u;
        const double b0 = Bs[cur][li][k + lk], b1 = Bs[cur][16 + li][k + lk];
#pragma unroll
        for (int s = 0; s < RTW; ++s) {
          if (has[s]) {
            const double av = MASK ? (am[u][s] ? a[u][s] : 0.0) : a[u][s];
            acc[s][0] = mfma(av, b0, acc[s][0]);
            acc[s][1] = mfma(av, b1, acc[s][1]);
          }
        }
        load(u, Wb, kb + 4 * u);
      }
    }
    const bool dz = b == p - 1 && i <= p - 3;
    const gptr<const double> Zi = glb(t.Z2) + (size_t)i * nn;
#pragma unroll
    for (int s = 0; s < RTW; ++s) {
      if (!has[s]) continue;
      const int rt = wave + NW * s;
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        const int col = ct * 16 + li;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = rt * 16 + crow(lk, r);
          // a row or column beyond n holds an exact zero (zero operands)
          const bool ok = !MASK || (c0 + col < n && row < n);
          double v = acc[s][ct][r];
          if (dz && ok) v += Zi[(size_t)(c0 + col) * n + row];
          if (ok) A[(size_t)(b * n + c0 + col) * lda + i * n + row] = v;
          Bs[cur ^ 1][col][row] = v;
        }
      }
    }
    __syncthreads();
    cur ^= 1;
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
    S[i][j] = (i <= j && i < nb && j < nb) ? glb(t.X)[(size_t)j * t.ld + i] : 0.0;
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
    if (i <= j && j < nb) glb(t.X)[(size_t)j * t.ld + i] = S[i][j];
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
      glb(t.Q)[mi * t.stride + (long long)j * t.dim + i] = (i == j) ? 1.0 : 0.0;
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
// 16-column tiles dealt round-robin to the NW waves (NTW tiles per wave).
// With diag_only the workgroup stops once its rows' diagonal blocks are
// done (block row r is finished at k = r - 1) and writes only G_rr; the
// diagonal blocks are the same bits as in the full inverse.
template <int MTI, int NTW, int NW = 4>
__global__ __launch_bounds__(64 * NW) void bsoi_scan_kernel(const ScanTask* __restrict__ tasks,
                                                            const Item* __restrict__ items, int ldcar) {
  constexpr int MT = 16 * MTI;
  constexpr int NT = 64 * NW;
  extern __shared__ double carry[];  // carry[kk * ldcar + i], kk < n, i < MT
  const Item it = items[blockIdx.x];
  const ScanTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int n = t.n, p = t.p, lda = t.lda, N = n * p, n2 = 2 * n;
  const int R0 = t.r0 + it.sub * MT;
  const gptr<double> A = glb(t.A);

  // rows served by this lane in A fragments: R0 + a*16 + li
  int rrow[MTI], rblk[MTI], rin[MTI];
#pragma unroll
  for (int a = 0; a < MTI; ++a) {
    rrow[a] = R0 + a * 16 + li;
    rblk[a] = rrow[a] < N ? rrow[a] / n : p;
    rin[a] = rrow[a] - rblk[a] * n;
  }
  const int ntiles = (n2 + 15) / 16;

  for (int e = tid; e < n * MT; e += NT) {
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
  const bool diag = t.diag_only != 0;
  const int kstop = diag ? max(rbmin - 1, 0) : 0;
  int bcol[NTW];
#pragma unroll
  for (int b = 0; b < NTW; ++b) bcol[b] = (wave + NW * b) * 16 + li;
  for (int k = p - 2; k >= kstop; --k) {
    const gptr<const double> Qk = glb(t.Q) + (size_t)k * t.qstride;
    d4 acc[MTI][NTW];
#pragma unroll
    for (int a = 0; a < MTI; ++a)
#pragma unroll
      for (int b = 0; b < NTW; ++b) acc[a][b] = d4zero();

    // Fragments of the next k-step are loaded while the MFMAs of the
    // current one run (two register buffers, ping-pong).  The 8-wave form
    // (16 tiles per wave, two waves per SIMD) has no registers for the second
    // buffer and loads each k-step's fragments right before its MFMAs; the
    // k order, and so every sum, is the same.
    constexpr bool DB = NW == 4;
    double a0[MTI], b0[NTW], a1[DB ? MTI : 1], b1[DB ? NTW : 1];
    auto run_part = [&](auto load) {
      if constexpr (DB) {
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
      } else {
        for (int k0 = 0; k0 < n; k0 += 4) {
          load(a0, b0, k0);
#pragma unroll
          for (int a = 0; a < MTI; ++a)
#pragma unroll
            for (int b = 0; b < NTW; ++b) acc[a][b] = mfma(a0[a], b0[b], acc[a][b]);
        }
      }
    };
    if (rbmin <= k) {
      // [X_{r,k}] part: K rows 0..n-1 of Q_k^T
      const gptr<const double> Xk = A + (size_t)k * n * lda;
      auto load = [&](auto& af, auto& bf, int k0) {
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
      run_part(load);
    }
    // [carry] part: K rows n..2n-1 of Q_k^T
    {
      const gptr<const double> Qc = Qk + (size_t)n * n2;
      auto load = [&](auto& af, auto& bf, int k0) {
        const int kk = k0 + lk;
#pragma unroll
        for (int a = 0; a < MTI; ++a) af[a] = (kk < n) ? carry[kk * ldcar + a * 16 + li] : 0.0;
#pragma unroll
        for (int b = 0; b < NTW; ++b) bf[b] = (kk < n && bcol[b] < n2) ? Qc[(size_t)kk * n2 + bcol[b]] : 0.0;
      };
      run_part(load);
    }
    __syncthreads();
    // epilogue
#pragma unroll
    for (int b = 0; b < NTW; ++b) {
      const int tile = wave + NW * b;
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
            if (row < N && (!diag || row / n == k + 1))
              A[(size_t)((k + 1) * n + (col - n)) * lda + row] = acc[a][b][r];
          }
        }
      }
    }
    __syncthreads();
  }
  // block column 0 (the carry after k = 0), for diag_only only its diagonal
  if (!diag || rbmin == 0) {
    for (int e = tid; e < n * MT; e += NT) {
      const int i = e % MT, kk = e / MT;
      const int row = R0 + i;
      if (row < N && (!diag || row < n)) A[(size_t)kk * lda + row] = carry[kk * ldcar + i];
    }
  }
}

// ---------------------------------------------------------------------------
// BSOI row-panel scan on v_mfma_f64_4x4x4f64 (K5, 4x4x4 form).
// ---------------------------------------------------------------------------
// Software pipeline of a k-loop over DEPTH register buffers (the 4x4x4
// scan): step i uses buffer i mod DEPTH, whose operands were loaded
// DEPTH - 1 steps earlier.  ld(u, step) loads buffer u, st(u) runs it.
// pipe_prologue loads steps 0 .. DEPTH - 2 (those below S); pipe_steps runs
// steps s + I .. s + R - 1 from a loop head at s (a multiple of DEPTH), each
// first loading step s + i + DEPTH - 1 when it is below S (always in the
// steady loop, STEADY = true skips the test); pipe_tail runs the R = S - s
// last steps, 2 <= R <= 2 DEPTH - 2 (R < DEPTH - 1 only when the loop never
// ran: all S >= 2 steps were loaded by the prologue), one unrolled sequence
// each.
template <int I, int DEPTH, class Ld>
__device__ __forceinline__ void pipe_prologue(Ld& ld, int S) {
  if constexpr (I < DEPTH - 1) {
    if (I < 2 || I < S) ld(I, I);  // S >= 2 (n >= 8)
    pipe_prologue<I + 1, DEPTH>(ld, S);
  }
}

template <int I, int R, int DEPTH, bool STEADY, class Ld, class St>
__device__ __forceinline__ void pipe_steps(Ld& ld, St& st, int s, int S) {
  if constexpr (I < R) {
    if constexpr (STEADY) {
      ld((I + DEPTH - 1) % DEPTH, s + I + DEPTH - 1);
    } else if constexpr (I + DEPTH - 1 < R) {
      ld((I + DEPTH - 1) % DEPTH, s + I + DEPTH - 1);
    }
    st(I % DEPTH);
    pipe_steps<I + 1, R, DEPTH, STEADY>(ld, st, s, S);
  }
}

template <int R, int DEPTH, class Ld, class St>
__device__ __forceinline__ void pipe_tail(Ld& ld, St& st, int s, int rem) {
  if constexpr (R <= 2 * DEPTH - 2) {
    if (rem == R) {
      pipe_steps<0, R, DEPTH, false>(ld, st, s, s + R);
      return;
    }
    pipe_tail<R + 1, DEPTH>(ld, st, s, rem);
  }
}

// The same scan as bsoi_scan_kernel for a 32-row panel, on the 4-block
// 4x4x4 FP64 MFMA (measured here at 1.5x the FP64 rate of the 16x16x4
// form).  For lane l = 16 x + 4 blk + y the instruction takes
//   A: A_blk[y][x],  B: B_blk[x][y],  and accumulates D_blk[x][y],
// the four blocks blk independent (checked on the device by the layout
// probe behind this kernel and by the parity tests).  Here all four blocks
// share the same 4 rows and cover the four 4-column quarters of a 16-column
// tile: B is then exactly the 16x16x4 B fragment (Q_k^T read as in
// bsoi_scan_kernel), and A is the 4 x 4 piece of the rows (the same in the
// four blocks).  NW waves; wave w holds 8 row groups x NTW column tiles
// w, w + NW, ... of accumulators, one double each (64 at NW = 4, in AGPRs:
// one workgroup per CU, one wave per SIMD, which measured faster than two
// waves per SIMD with half the tiles each, as each A value then feeds 8
// MFMAs).  The carry sits in LDS with row 4 g + y at y * 8 + g, so that a
// lane's 8 A values of a k-step are contiguous.
//
// DEPTH is the number of register buffers of the k-loop: 2 loads each
// k-step's operands one step ahead, 3 two steps ahead (then every lane's
// Q_k^T loads address one uniform row base plus a fixed per-lane offset).
// Both run the same MFMAs in the same order, so the sums are the same.
// (Four buffers, three steps ahead, measured the same as three at n = 100
// and 144, where the registers allow them.)  XN is the largest n the X
// staging buffer holds (256), or 64 for the form that runs two workgroups
// per CU (MINB = 2: a quarter of the LDS, at most 256 registers per lane):
// batches of small matrices, whose k-loops are too short for one wave per
// SIMD to keep the matrix cores busy on its own.
template <int NTW, int NW, int DEPTH = 2, int XN = 256, int MINB = 1>
__global__ __launch_bounds__(64 * NW, MINB) void bsoi_scan4_kernel(const ScanTask* __restrict__ tasks,
                                                                   const Item* __restrict__ items, int ldcar) {
  static_assert(DEPTH == 2 || DEPTH == 3, "two or three operand buffers");
  constexpr int NT = 64 * NW;
  constexpr int MT = 32;
  constexpr int XLD = 36;  // column stride of Xs: the A-fragment reads hit 16 distinct bank pairs
  extern __shared__ double carry[];  // carry[kk * ldcar + (i & 3) * 8 + (i >> 2)] = row i, column kk
  // X_{r,k} (the panel's 32 rows of block column k of A), staged by direct
  // global -> LDS loads during block k + 1's epilogue: Xs[kk * XLD + i].
  // The planner takes this form for n <= XN only (NTW <= XN / 32).
  __shared__ double Xs[XN * XLD];
  static_assert(NTW * 32 <= XN, "n <= XN: X_{r,k} must fit Xs");
  const Item it = items[blockIdx.x];
  const ScanTask t = tasks[it.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int x = lane >> 4, y = lane & 3, l16 = lane & 15;
  const int n = t.n, p = t.p, lda = t.lda, N = n * p, n2 = 2 * n;
  const int R0 = t.r0 + it.sub * MT;
  const gptr<double> A = glb(t.A);
  auto cidx = [&](int kk, int i) { return kk * ldcar + (i & 3) * 8 + (i >> 2); };
  // One wave instruction moves column kk of X_{r,kx} (32 rows, 256 bytes):
  // lane l loads bytes 4l .. 4l + 3, and the instruction writes them to
  // Xs + kk * XLD + l / 2 doubles.  Rows past the matrix end read row N - 1
  // (their products are masked).
  const int srow = min(R0 + (lane >> 1), N - 1);
  auto stage_col = [&](int kx, int kk) {
    const char* src = reinterpret_cast<const char*>(t.A + ((size_t)kx * n + kk) * lda + srow) + 4 * (lane & 1);
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                     (__attribute__((address_space(3))) void*)(Xs + kk * XLD), 4, 0, 0);
  };

  // This lane's A values are rows R0 + 4 g + y.  Row i of X_{r,k} takes part
  // in k-step column kk (of block column k) iff i < N and i lies in block row
  // k or above, on or above the diagonal: i <= min(k n + kk, N - 1).

  for (int e = tid; e < n * MT; e += NT) {
    const int i = e % MT, kk = e / MT;
    const int row = R0 + i;
    double v = 0.0;
    if (row < N) {
      const int rb = row / n;
      if (!(rb == p - 1 && row - rb * n > kk)) v = A[(size_t)((p - 1) * n + kk) * lda + row];
    }
    carry[cidx(kk, i)] = v;
  }
  __syncthreads();

  const int rbmin = R0 / n;
  const bool diag = t.diag_only != 0;
  const int kstop = diag ? max(rbmin - 1, 0) : 0;
  // the X part of block k runs when rbmin <= k: stage the first one now
  if (rbmin <= p - 2) {
    for (int kk = wave; kk < n; kk += NW) stage_col(p - 2, kk);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's columns are in LDS
  }
  __syncthreads();
  // Q_k^T columns of this lane's tiles.  When the 16-column tiles do not
  // divide evenly over the waves, a wave's last tiles lie beyond 2n: they
  // read the last column (in bounds) and their results are never stored.
  int bcol[NTW];
  unsigned boff[NTW];  // byte offset of (row x, column bcol[b]) in a 4-row band of Q_k^T
#pragma unroll
  for (int b = 0; b < NTW; ++b) {
    bcol[b] = min((wave + NW * b) * 16 + l16, n2 - 1);
    boff[b] = (unsigned)(x * n2 + bcol[b]) * 8u;
  }
  for (int k = p - 2; k >= kstop; --k) {
    const gptr<const double> Qk = glb(t.Q) + (size_t)k * t.qstride;
    double acc[8][NTW];
#pragma unroll
    for (int g = 0; g < 8; ++g)
#pragma unroll
      for (int b = 0; b < NTW; ++b) acc[g][b] = 0.0;

    // Two register buffers: the loads of k-step s + 1 (Q_k^T from L2, the
    // A fragments from LDS) are in flight while the 8 NTW MFMAs of step s
    // run; with one wave per SIMD nothing else covers their latency, which
    // is why X_{r,k} comes from LDS and not from HBM.  With 4 | n (and the
    // clamped columns) every index of the carry part and of Q_k^T is in
    // range, so only the X part of a workgroup whose rows reach the diagonal
    // block (or the matrix end) needs masks, applied to the A values just
    // before the MFMAs (bit g: row group g).  (A select right after a load,
    // or a predicated load, makes the compiler wait for the loads early.)
    double af[DEPTH][8], bf[DEPTH][NTW];
    int msk[DEPTH];
    auto step = [&](int u, auto masked_tag) {
      constexpr bool MASKED = decltype(masked_tag)::value;
      double a[8];
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        if constexpr (MASKED) a[g] = (msk[u] >> g & 1) ? af[u][g] : 0.0;
        else a[g] = af[u][g];
      }
#pragma unroll
      for (int b = 0; b < NTW; ++b)
#pragma unroll
        for (int g = 0; g < 8; ++g) acc[g][b] = __builtin_amdgcn_mfma_f64_4x4x4f64(a[g], bf[u][b], acc[g][b], 0, 0, 0);
    };
    // n / 4 k-steps, in pairs with two buffers (8 | n), any number >= 2 with
    // three (4 | n): the loop body has no branch
    // around its loads, so the wait before each step can leave the next
    // step's loads in flight.
    auto run_part = [&](auto load, auto masked_tag) {
      if constexpr (DEPTH == 2) {
        load(0, 0);
        int k0 = 0;
        for (; k0 + 8 < n; k0 += 8) {
          load(1, k0 + 4);
          step(0, masked_tag);
          load(0, k0 + 8);
          step(1, masked_tag);
        }
        load(1, k0 + 4);
        step(0, masked_tag);
        step(1, masked_tag);
      } else {
        // S = n / 4 >= 2 k-steps; step s uses buffer s mod DEPTH, loaded
        // DEPTH - 1 steps ahead.  At the loop head buffers 0 .. DEPTH - 2
        // hold steps s .. s + DEPTH - 2 (s a multiple of DEPTH).
        const int S = n >> 2;
        auto st = [&](int u) { step(u, masked_tag); };
        auto ld = [&](int u, int sx) { load(u, 4 * sx); };
        pipe_prologue<0, DEPTH>(ld, S);
        int s = 0;
        for (; s + 2 * DEPTH - 1 <= S; s += DEPTH) pipe_steps<0, DEPTH, DEPTH, true>(ld, st, s, S);
        pipe_tail<2, DEPTH>(ld, st, s, S - s);  // S - s >= 2
      }
    };
    // Operand b of the 4-row band of Q_k^T that starts at row k0 of Q (Q_k
    // itself, or its rows n .. 2n - 1 for the carry part, qrow0 = n).  The
    // three-buffer form reads through a buffer resource over Q_k: a uniform
    // band offset in an SGPR and a fixed per-lane offset, no per-step address
    // arithmetic (out-of-range offsets read 0; none occur with 4 | n).
    const __amdgpu_buffer_rsrc_t qrs =
        __builtin_amdgcn_make_buffer_rsrc((void*)(t.Q + (size_t)k * t.qstride), (short)0, 8 * n2 * n2, 0x00020000);
    auto qload = [&](gptr<const double> Q, int qrow0, int k0, int b) -> double {
      if constexpr (DEPTH == 2) {
        return Q[(size_t)(k0 + x) * n2 + bcol[b]];
      } else {
        return __builtin_bit_cast(double,
                                  __builtin_amdgcn_raw_buffer_load_b64(qrs, boff[b], 8 * (qrow0 + k0) * n2, 0));
      }
    };
    if (rbmin <= k) {
      // [X_{r,k}] part: K rows 0..n-1 of Q_k^T, X_{r,k} from Xs
      if (R0 + MT <= N && (R0 + MT - 1) / n < k) {  // every row below block row k's diagonal: no masks
        run_part(
            [&](int u, int k0) {
              const int kk = k0 + x;
              const double* xr = Xs + kk * XLD + y;
#pragma unroll
              for (int g = 0; g < 8; ++g) af[u][g] = xr[4 * g];
#pragma unroll
              for (int b = 0; b < NTW; ++b) bf[u][b] = qload(Qk, 0, k0, b);
            },
            std::false_type{});
      } else {
        run_part(
            [&](int u, int k0) {
              const int kk = k0 + x;
              const double* xr = Xs + kk * XLD + y;
#pragma unroll
              for (int g = 0; g < 8; ++g) af[u][g] = xr[4 * g];
              // bit g: 4 g <= lim, i.e. row groups 0 .. lim / 4
              const int lim = min(k * n + kk, N - 1) - R0 - y;
              const int mk = lim < 0 ? 0 : (lim >= 28 ? 0xFF : (1 << (lim / 4 + 1)) - 1);
#pragma unroll
              for (int b = 0; b < NTW; ++b) bf[u][b] = qload(Qk, 0, k0, b);
              msk[u] = mk;
            },
            std::true_type{});
      }
    }
    // [carry] part: K rows n..2n-1 of Q_k^T
    {
      const gptr<const double> Qc = Qk + (size_t)n * n2;
      run_part(
          [&](int u, int k0) {
            const int kk = k0 + x;
            const double* cr = carry + kk * ldcar + y * 8;
#pragma unroll
            for (int g = 0; g < 8; ++g) af[u][g] = cr[g];
#pragma unroll
            for (int b = 0; b < NTW; ++b) bf[u][b] = qload(Qc, n, k0, b);
          },
          std::false_type{});
    }
    // epilogue: lane holds rows 4 g + x, column tile * 16 + l16.  First the
    // G block (columns n..2n-1 of the product), which touches no LDS, so a
    // wave stores it as soon as its own k-loop ends; then, once every wave is
    // done with Xs and the carry, X_{r,k-1} into Xs when block k - 1 has an
    // X part (wave w: columns w, w + 4, ...), then the new carry, so that
    // the stores' and the staging loads' latencies pass under the carry
    // writes; the last barrier waits for both.  (Mixed into the k-loop, a
    // direct-to-LDS load among the loads of Q_k^T makes the compiler drain
    // every load at each k-step.)
#pragma unroll
    for (int b = 0; b < NTW; ++b) {
      const int col = (wave + NW * b) * 16 + l16;
      if (col < n || col >= n2) continue;
#pragma unroll
      for (int g = 0; g < 8; ++g) {
        const int row = R0 + 4 * g + x;
        if (row < N && (!diag || row / n == k + 1)) A[(size_t)((k + 1) * n + (col - n)) * lda + row] = acc[g][b];
      }
    }
    __syncthreads();
    const bool stage = k - 1 >= kstop && rbmin <= k - 1;
    if (stage)
      for (int kk = wave; kk < n; kk += NW) stage_col(k - 1, kk);
#pragma unroll
    for (int b = 0; b < NTW; ++b) {
      const int col = (wave + NW * b) * 16 + l16;
      if (col >= n) continue;
#pragma unroll
      for (int g = 0; g < 8; ++g) carry[cidx(col, 4 * g + x)] = acc[g][b];
    }
    if (stage) __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's columns of X_{r,k-1} are in LDS
    __syncthreads();
  }
  // block column 0 (the carry after k = 0), for diag_only only its diagonal
  if (!diag || rbmin == 0) {
    for (int e = tid; e < n * MT; e += NT) {
      const int i = e % MT, kk = e / MT;
      const int row = R0 + i;
      if (row < N && (!diag || row < n)) A[(size_t)kk * lda + row] = carry[cidx(kk, i)];
    }
  }
}

// ---------------------------------------------------------------------------
// FP64 rate calibration: independent chains of one instruction kind on
// every CU (KIND 0: v_mfma_f64_16x16x4f64, 1: v_mfma_f64_4x4x4f64 with its
// 4 blocks, 2: v_fma_f64), 8 chains per wave; the result is stored so
// nothing is dropped.  Flops per iteration per wave: 8 x 2048, 8 x 512,
// 8 x 64 x 2.
// ---------------------------------------------------------------------------
template <int KIND>
__global__ __launch_bounds__(256) void fp64_rate_kernel(double* __restrict__ out, int iters, double a, double b) {
  const double x = a + threadIdx.x * 1e-9;
  double s = 0.0;
  if constexpr (KIND == 0) {
    d4 acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = d4zero();
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = mfma(x, b, acc[c]);
#pragma unroll
    for (int c = 0; c < 8; ++c) s += (acc[c][0] + acc[c][1]) + (acc[c][2] + acc[c][3]);
  } else {
    double acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = KIND == 1 ? 0.0 : x + c;
    for (int i = 0; i < iters; ++i)
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        if constexpr (KIND == 1)
          acc[c] = __builtin_amdgcn_mfma_f64_4x4x4f64(x, b, acc[c], 0, 0, 0);
        else
          acc[c] = fma(acc[c], a, b);
      }
#pragma unroll
    for (int c = 0; c < 8; ++c) s += acc[c];
  }
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
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

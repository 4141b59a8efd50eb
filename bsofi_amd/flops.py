"""Flop counts of BSOF/I, restating the reference's accounting.

FLOPS_BSOF/BSTRI/BSOI follow include/bsofi_flops.h:19-30, built on the
LAWN-41 formulas of include/bench/flops.h (FMULS_/FADDS_ GEQRF, UNMQR,
TRTRI, TRMM, GEMM at lines 42-146).  These are the numerators of every
GFLOP/s and roofline figure the project reports.
"""
from __future__ import annotations


def _fmuls_geqrf(m, n):
    if m > n:
        return n * (n * (0.5 - (1.0 / 3.0) * n + m) + m + 23.0 / 6.0)
    return m * (m * (-0.5 - (1.0 / 3.0) * m + n) + 2.0 * n + 23.0 / 6.0)


def _fadds_geqrf(m, n):
    if m > n:
        return n * (n * (0.5 - (1.0 / 3.0) * n + m) + 5.0 / 6.0)
    return m * (m * (-0.5 - (1.0 / 3.0) * m + n) + n + 5.0 / 6.0)


def dgeqrf(m, n):
    return _fmuls_geqrf(float(m), float(n)) + _fadds_geqrf(float(m), float(n))


def dormqr(m, n, k, side):
    m, n, k = float(m), float(n), float(k)
    if side == "L":
        mul = 2.0 * n * m * k - n * k * k + 2.0 * n * k
        add = 2.0 * n * m * k - n * k * k + n * k
    else:
        mul = 2.0 * n * m * k - m * k * k + m * k + n * k - 0.5 * k * k + 0.5 * k
        add = 2.0 * n * m * k - m * k * k + m * k
    return mul + add


def dtrtri(n):
    n = float(n)
    return n * (n * (1.0 / 6.0 * n + 0.5) + 1.0 / 3.0) + n * (n * (1.0 / 6.0 * n - 0.5) + 1.0 / 3.0)


def dtrmm(side, m, n):
    m, n = float(m), float(n)
    if side == "L":
        mul = 0.5 * n * m * (m + 1)
        add = 0.5 * n * m * (m - 1)
    else:
        mul = 0.5 * m * n * (n + 1)
        add = 0.5 * m * n * (n - 1)
    return mul + add


def dgemm(m, n, k):
    return 2.0 * float(m) * float(n) * float(k)


def bsof(n, L):
    return (L - 2) * (2 * dormqr(2 * n, n, n, "L") + dgeqrf(2 * n, n)) + dgeqrf(2 * n, 2 * n)


def bstri(n, L):
    # (((L)-3)*((L)+2)/2) is integer arithmetic in the C macro.
    return (dtrtri(3 * n) + dtrmm("R", n * (L - 3), n)
            + (L - 3) * (dtrtri(n) + 2 * dtrmm("L", n, n))
            + (((L - 3) * (L + 2)) // 2) * dgemm(n, n, n))


def bsoi(n, L):
    return (L - 2) * dormqr(L * n, 2 * n, n, "R") + dormqr(L * n, 2 * n, 2 * n, "R")


def bsofi(n, L):
    """Total work of one inversion (the roofline numerator)."""
    return bsof(n, L) + bstri(n, L) + bsoi(n, L)


def bytes_min(n, L):
    """Write G, read the 2p non-zero blocks of A, write tau (8-byte words)."""
    N = n * L
    return 8.0 * (N * N + 2 * L * n * n + N)


def families(n, L, nb=32):
    """Algorithmic flops per inversion of each kernel family of this build's
    plan (the numerators of the per-family roofline figures):

    * panel -- the sub-panel QRs: dgeqrf(m, nbj) of every sub-panel of the
      BSOF sweep (m = 2n - j nb rows, nbj <= nb columns);
    * refl  -- block reflector applications, 4 m nbj w each (W = V^T C and
      C -= V (T W), leading order): the BSOF updates of the w = w1 + w2
      trailing and last-block-column columns, and BSOI's formation of the
      explicit Q_k (applied to the identity, w = m);
    * gemm  -- BSTRI, counted as the reference counts it (bstri(n, L));
    * scan  -- BSOI, counted as the reference counts it (bsoi(n, L)).
    """
    panel = refl = 0.0
    for k in range(L - 1):
        ncols = 2 * n if k == L - 2 else n
        for j in range(0, (ncols + nb - 1) // nb):
            m = 2 * n - j * nb
            nbj = min(nb, ncols - j * nb)
            panel += dgeqrf(m, nbj)
            w = (2 * n - j * nb - nbj) + (n if k < L - 2 else 0)
            refl += 4.0 * m * nbj * w  # BSOF update
            refl += 4.0 * m * nbj * m  # Q_k formation in BSOI
    return {"panel": panel, "refl": refl, "gemm": bstri(n, L), "scan": bsoi(n, L)}

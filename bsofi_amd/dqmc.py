"""Seeded Hubbard-DQMC p-cyclic matrices (host side, numpy).

A block p-cyclic matrix A (N = n*p, column-major) has identity diagonal
blocks, B_1 in the top-right corner and -B_l on the block subdiagonal
(l = 2..p), with B_l = e^{-dtau K} diag(e^{lambda s_{l,i}}):

* K is the nearest-neighbour hopping matrix of a periodic Lx x Ly square
  lattice, K_ij = -t for neighbours, so its eigenvalues are
  -2t(cos kx + cos ky);
* lambda = acosh(e^{U dtau / 2});
* s_{l,i} = +-1 comes from a counter-based hash keyed by
  (seed, instance, l, i), so any slice of the field can be regenerated
  without sequential state.  The same hash is implemented in the native
  library (bsofi_amd/csrc/bsofi_dqmc.h) for the device-side generator.

The reference's own generators (include/bench/bsofi_random.h:24-44,
examples/ctest/ctest.c:38-53) fill the blocks with time-seeded U[0,1]
numbers; these matrices replace them with the physical B-matrices the
method is built for, at a fixed seed.
"""
from __future__ import annotations

import math

import numpy as np

DEFAULT_SEED = 20261203
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def hs_fields(seed: int, instance: int, p: int, n: int) -> np.ndarray:
    """s[l, i] in {+1, -1}, l = 0..p-1 (time slice), i = 0..n-1 (site)."""
    l = np.arange(p, dtype=np.uint64)[:, None]
    i = np.arange(n, dtype=np.uint64)[None, :]
    with np.errstate(over="ignore"):
        h0 = _splitmix64(np.uint64(seed) + np.uint64(0xD6E8FEB86659FD93) * np.uint64(instance + 1))
        h = _splitmix64(h0 ^ (l * np.uint64(0xD1B54A32D192ED03) + i))
    return np.where((h >> np.uint64(63)) == 1, 1.0, -1.0)


def hopping_matrix(lx: int, ly: int, t: float = 1.0) -> np.ndarray:
    """K for the periodic lx x ly lattice, site index x + lx*y."""
    n = lx * ly
    K = np.zeros((n, n))
    for y in range(ly):
        for x in range(lx):
            s = x + lx * y
            for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
                r = (x + dx) % lx + lx * ((y + dy) % ly)
                if r != s:
                    K[s, r] -= t
    return K


def expk(lx: int, ly: int, t: float = 1.0, dtau: float = 0.125) -> np.ndarray:
    """e^{-dtau K} by the symmetric eigendecomposition of K."""
    K = hopping_matrix(lx, ly, t)
    w, U = np.linalg.eigh(K)
    return np.asfortranarray((U * np.exp(-dtau * w)) @ U.T)


def hs_lambda(U: float, dtau: float) -> float:
    return math.acosh(math.exp(U * dtau / 2.0))


def lattice_side(n: int) -> int:
    s = int(round(math.sqrt(n)))
    if s * s != n:
        raise ValueError(f"n={n} is not a square lattice size")
    return s


def assemble(ek: np.ndarray, fields: np.ndarray, lam: float, p: int, lda: int | None = None,
             out: np.ndarray | None = None) -> np.ndarray:
    """Column-major A (lda x N) from e^{-dtau K} and the +-1 fields.

    B_l[:, i] = ek[:, i] * e^{lambda s_{l,i}}; the two scale factors are the
    only transcendental values, so any implementation handed the same ek and
    e^{+-lambda} reproduces A bit for bit."""
    n = ek.shape[0]
    N = n * p
    lda = lda or N
    A = out if out is not None else np.zeros((lda, N), order="F")
    A[:] = 0.0
    A[np.arange(N), np.arange(N)] = 1.0
    ep, em = math.exp(lam), math.exp(-lam)
    for l in range(1, p + 1):
        scale = np.where(fields[l - 1] > 0, ep, em)
        B = ek * scale[None, :]
        if l == 1:
            A[0:n, (p - 1) * n:p * n] = B
        else:
            A[(l - 1) * n:l * n, (l - 2) * n:(l - 1) * n] = -B
    return A


def dqmc_matrix(lx: int, p: int, U: float = 4.0, dtau: float = 0.125, t: float = 1.0,
                seed: int = DEFAULT_SEED, instance: int = 0, ly: int | None = None,
                lda: int | None = None):
    """(A, meta) for one seeded DQMC p-cyclic matrix; n = lx*ly."""
    ly = ly or lx
    n = lx * ly
    ek = expk(lx, ly, t, dtau)
    s = hs_fields(seed, instance, p, n)
    lam = hs_lambda(U, dtau)
    A = assemble(ek, s, lam, p, lda)
    return A, dict(n=n, p=p, lx=lx, ly=ly, U=U, dtau=dtau, t=t, seed=seed,
                   instance=instance, lam=lam, expk=ek, fields=s)


def random_pcyclic(n: int, p: int, seed: int = 0, identity_diag: bool = True, lda=None):
    """U[0,1] blocks in the pattern of examples/ctest/ctest.c:38-53 (identity
    diagonal) or include/bench/bsofi_random.h:24-44 (random diagonal), from a
    fixed numpy seed instead of time()."""
    rng = np.random.default_rng(seed)
    N = n * p
    lda = lda or N
    A = np.zeros((lda, N), order="F")
    for k in range(p):
        r = ((k + 1) % p) * n
        if identity_diag:
            A[k * n:(k + 1) * n, k * n:(k + 1) * n] += np.eye(n)
        else:
            A[k * n:(k + 1) * n, k * n:(k + 1) * n] = rng.random((n, n))
        A[r:r + n, k * n:(k + 1) * n] = rng.random((n, n))
    return A

"""CPU oracle for BSOF/I -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg import this package, and only as the checker or as the
timed CPU baseline.  The product (``bsofi_amd``) never imports it.

Two libraries are wrapped:

* ``liboracle.so`` -- ``bsofi_oracle.c``, a C restatement of the reference's
  CPU call sequence (SGo-Go/BSOFI ``src/bsofi_cpu.c:31-333``) with LAPACK
  resolved at run time: MKL (the reference's default), scipy's OpenBLAS, or a
  plain-C restatement of the unblocked LAPACK algorithms.
* ``_ref/libbsofi_ref.so`` -- the reference's own ``src/bsofi_cpu.c``,
  ``src/bsofi_lapack.c`` and ``src/bsofi_callback.c`` compiled against MKL by
  ``oracle/Makefile`` (only where the reference sources exist; the built file
  travels with the tree).
"""
from __future__ import annotations

import ctypes as C
import glob
import math
import os

import numpy as np

# MKL's default Intel OpenMP threading layer computes wrong results when
# libgomp is also in the process (torch loads it): use the GNU layer, which
# shares libgomp.  Must be set before MKL initialises.
os.environ.setdefault("MKL_THREADING_LAYER", "GNU")

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbsofi_ref.so")

BACKEND_BUILTIN, BACKEND_MKL, BACKEND_OPENBLAS = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def mkl_path():
    for p in ("/opt/conda/lib/libmkl_rt.so", "/opt/conda/lib/libmkl_rt.so.1"):
        if os.path.exists(p):
            return p
    return None


def openblas_path():
    try:
        import scipy

        d = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs")
        c = sorted(glob.glob(os.path.join(d, "libscipy_openblas*.so")))
        return c[0] if c else None
    except Exception:  # pragma: no cover - scipy is in the image
        return None


_ACTIVE = {"backend": None}


class Oracle:
    """ctypes view of liboracle.so.  All matrices are Fortran-ordered float64.

    The library holds one process-wide LAPACK back end; every instance
    re-selects its own before each call, so instances with different back
    ends can be used side by side."""

    def __init__(self, backend: str = "auto", threads: int | None = None):
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle` (or __graft_entry__.build())")
        self.lib = C.CDLL(LIB_PATH, mode=C.RTLD_LOCAL)
        L = self.lib
        L.oracle_set_backend.argtypes = [C.c_int, C.c_char_p]
        L.oracle_set_threads.argtypes = [C.c_int]
        L.oracle_get_backend_path.restype = C.c_char_p
        for f in ("oracle_bsof", "oracle_bsoi_ormqr", "oracle_bsoi_orgqr"):
            getattr(L, f).argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, _ip]
        L.oracle_bstri.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip]
        L.oracle_bsofi_lwork.argtypes = [C.c_int, C.c_int, C.c_int]
        L.oracle_invert.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, C.c_int, _dp]
        L.oracle_residual_dense.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_int]
        L.oracle_residual_dense.restype = C.c_double
        self._threads = None if threads is None else int(threads)
        self.set_backend(backend)

    # -- back end -----------------------------------------------------------
    def set_backend(self, backend: str):
        if backend == "auto":
            backend = "mkl" if mkl_path() else "builtin"
        if backend == "builtin":
            rc = self.lib.oracle_set_backend(BACKEND_BUILTIN, None)
        elif backend == "mkl":
            rc = self.lib.oracle_set_backend(BACKEND_MKL, (mkl_path() or "").encode())
        elif backend == "openblas":
            rc = self.lib.oracle_set_backend(BACKEND_OPENBLAS, (openblas_path() or "").encode())
        else:
            raise ValueError(backend)
        if rc != 0:
            raise RuntimeError(f"oracle back end {backend!r} unavailable")
        self.backend = backend
        _ACTIVE["backend"] = backend
        if getattr(self, "_threads", None) is not None:
            self.lib.oracle_set_threads(self._threads)
        return self

    def _activate(self):
        if _ACTIVE["backend"] != self.backend:
            self.set_backend(self.backend)

    def set_threads(self, n: int) -> bool:
        self._threads = int(n)
        self._activate()
        return self.lib.oracle_set_threads(int(n)) == 0

    @property
    def backend_path(self) -> str:
        return self.lib.oracle_get_backend_path().decode()

    # -- the reference call sequence ------------------------------------------
    def lwork(self, n, L, lda):
        self._activate()
        return self.lib.oracle_bsofi_lwork(n, L, lda)

    def bsof(self, n, L, A, tau):
        self._activate()
        lw = max(self.lwork(n, L, A.shape[0]), 1)
        work = np.zeros(lw)
        info = C.c_int(0)
        self.lib.oracle_bsof(n, L, _ptr(A), A.shape[0], _ptr(tau), _ptr(work), lw, C.byref(info))
        return info.value

    def bstri(self, n, L, A):
        self._activate()
        info = C.c_int(0)
        self.lib.oracle_bstri(n, L, _ptr(A), A.shape[0], C.byref(info))
        return info.value

    def bsoi(self, n, L, A, tau, variant="ormqr"):
        self._activate()
        lw = max(self.lwork(n, L, A.shape[0]), 4 * n * n + 2 * n * (2 * n + A.shape[0]) + 64 * n)
        work = np.zeros(lw)
        info = C.c_int(0)
        f = self.lib.oracle_bsoi_ormqr if variant == "ormqr" else self.lib.oracle_bsoi_orgqr
        f(n, L, _ptr(A), A.shape[0], _ptr(tau), _ptr(work), lw, C.byref(info))
        return info.value

    def invert(self, n, L, A, variant="ormqr"):
        """In-place G = A^{-1}; returns (tau, info, phase_times[3])."""
        self._activate()
        assert A.flags.f_contiguous and A.dtype == np.float64
        tau = np.zeros(n * L)
        times = np.zeros(3)
        info = self.lib.oracle_invert(n, L, _ptr(A), A.shape[0], _ptr(tau),
                                      1 if variant == "orgqr" else 0, _ptr(times))
        return tau, info, times

    def residual(self, A, G):
        A = np.asfortranarray(A)
        G = np.asfortranarray(G)
        return self.lib.oracle_residual_dense(A.shape[0], _ptr(A), A.shape[0], _ptr(G), G.shape[0])


class Reference:
    """ctypes view of oracle/_ref/libbsofi_ref.so: the reference's own CPU path."""

    def __init__(self, threads: int | None = None):
        if not os.path.exists(REF_PATH):
            raise RuntimeError(f"{REF_PATH} not built (reference sources absent when building)")
        self.lib = C.CDLL(REF_PATH, mode=C.RTLD_LOCAL)
        L = self.lib
        for f in ("lapackXbsof", "lapackXbsoi_ormqr", "lapackXbsoi_orgqr"):
            getattr(L, f).argtypes = [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, C.c_int, _ip]
            getattr(L, f).restype = C.c_int
        L.lapackXbstri.argtypes = [C.c_int, C.c_int, _dp, C.c_int, _ip]
        L.lapackXbsofiLWork.argtypes = [C.c_int, C.c_int, C.c_int]
        if threads is not None:
            set_mkl_threads(threads)

    def lwork(self, n, L, lda):
        return self.lib.lapackXbsofiLWork(n, L, lda)

    def invert(self, n, L, A, variant="ormqr", times=None):
        """The sequence of examples/ftest_cpu/ftest_cpu.f90:129-140, in place."""
        import time

        lda = A.shape[0]
        lw = max(self.lwork(n, L, lda), 1)
        if variant == "orgqr":
            lw = max(lw, 4 * n * n + 2 * n * (2 * n + lda) + 64 * n * 4)
        work = np.zeros(lw)
        tau = np.zeros(n * L)
        info = C.c_int(0)
        t0 = time.perf_counter()
        self.lib.lapackXbsof(n, L, _ptr(A), lda, _ptr(tau), _ptr(work), lw, C.byref(info))
        t1 = time.perf_counter()
        self.lib.lapackXbstri(n, L, _ptr(A), lda, C.byref(info))
        t2 = time.perf_counter()
        f = self.lib.lapackXbsoi_ormqr if variant == "ormqr" else self.lib.lapackXbsoi_orgqr
        f(n, L, _ptr(A), lda, _ptr(tau), _ptr(work), lw, C.byref(info))
        t3 = time.perf_counter()
        if times is not None:
            times[:] = [t1 - t0, t2 - t1, t3 - t2]
        return tau, info.value

    def bsof_bstri(self, n, L, A):
        lda = A.shape[0]
        lw = max(self.lwork(n, L, lda), 1)
        work = np.zeros(lw)
        tau = np.zeros(n * L)
        info = C.c_int(0)
        self.lib.lapackXbsof(n, L, _ptr(A), lda, _ptr(tau), _ptr(work), lw, C.byref(info))
        self.lib.lapackXbstri(n, L, _ptr(A), lda, C.byref(info))
        return tau, info.value


def set_mkl_threads(n: int) -> bool:
    p = mkl_path()
    if not p:
        return False
    lib = C.CDLL(p, mode=C.RTLD_GLOBAL)
    f = getattr(lib, "MKL_Set_Num_Threads", None)
    if f is None:
        return False
    x = C.c_int(int(n))
    f(C.byref(x))
    return True


def ref_available() -> bool:
    return os.path.exists(REF_PATH) and mkl_path() is not None


def dense_inverse(A):
    """numpy LU inverse (the reference's own comparison uses dgetrf/dgetri,
    examples/ftest_cpu/ftest_cpu.f90:174-186)."""
    return np.linalg.inv(A)


def residual(A, G):
    """||A G - I||_F / (||A||_F ||G||_F), examples/ftest_cpu/ftest_cpu.f90:97."""
    R = A @ G
    R[np.diag_indices_from(R)] -= 1.0
    return np.linalg.norm(R) / (np.linalg.norm(A) * np.linalg.norm(G))


def pcyclic_blocks(A, n, p):
    """The 2p non-zero blocks of a p-cyclic A: (diag[i], off[i]) with
    off[i] = A_BLK(i, i-1) for i >= 1 and A_BLK(0, p-1) for i = 0."""
    diag = [np.array(A[i * n:(i + 1) * n, i * n:(i + 1) * n]) for i in range(p)]
    off = [np.array(A[i * n:(i + 1) * n, ((i - 1) % p) * n:((i - 1) % p + 1) * n]) for i in range(p)]
    return diag, off


def residual_blocks(diag, off, G, n, p):
    """||A G - I||_F / (||A||_F ||G||_F) (examples/ftest_cpu/ftest_cpu.f90:97)
    with A given by its 2p non-zero blocks: block row i of A G is
    diag[i] G_i + off[i] G_{i-1} (G_{-1} = G_{p-1}), 4 n N^2 flops instead of
    the dense product's 2 N^3."""
    N = n * p
    ss = 0.0
    for i in range(p):
        j = (i - 1) % p
        R = diag[i] @ G[i * n:(i + 1) * n, :] + off[i] @ G[j * n:(j + 1) * n, :]
        R[np.arange(n), i * n + np.arange(n)] -= 1.0
        ss += float(np.sum(R * R))
    na = math.sqrt(sum(float(np.sum(d * d)) + float(np.sum(o * o)) for d, o in zip(diag, off)))
    return math.sqrt(ss) / (na * np.linalg.norm(G[:N, :N]))


def det_sign(tau, A_after_bsof):
    """Sign of det(A) read from tau and diag(R), as
    examples/ftest_cpu/ftest_cpu.f90:133-138 does it: flip for every tau==0
    and every negative diagonal entry.  The formula equals (-1)^N sign(det A)
    (each non-trivial reflector has determinant -1)."""
    s = 1.0
    d = np.diag(A_after_bsof)
    for i in range(len(tau)):
        if tau[i] == 0.0:
            s = -s
        if d[i] < 0.0:
            s = -s
    return s

"""bsofi_amd -- BSOF/I (structured orthogonal factorization and inversion of
block p-cyclic matrices) on AMD Instinct MI355X.

The compute lives in ``libbsofi_amd.so`` (HIP kernels for gfx950 behind the C
ABI of ``include/bsofi.h``).  This module is the host-side Python view of
that ABI:

* the reference's entry points with their names and argument meaning
  (``lapackXbsof``, ``lapackXbstri``, ``lapackXbsoi_ormqr``,
  ``lapackXbsoi_orgqr``, ``hybridXbsoftri``, ``hybridXbsoi`` and the handle
  API ``hybridXinitH`` / ``hybridXbsoftriH`` / ``hybridXbsoiH`` /
  ``hybridXfinalizeH``, SGo-Go/BSOFI include/bsofi_cpu.h:24-41,
  include/bsofi_hybrid.h:22-91, src/bsofi_callback.c:101-180) on numpy
  arrays;
* the native device API (contexts, plans, batched inversion) on device
  pointers, typically from torch tensors.

There is no CPU fallback: if the library is missing, importing the module
raises, and on a machine without a GPU every compute call returns
``BSOFI_ERR_NO_DEVICE``.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

from . import flops  # noqa: F401  (re-export)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsofi_amd.so")

BSOFI_ERR_NO_DEVICE = -1000
BSOFI_ERR_ALLOC = -1001
BSOFI_ERR_UNSUPPORTED = -1002
BSOFI_ERR_HANDLE = -1003
BSOFI_ERR_SCHEDULE = -1004
BSOFI_ERR_SCRATCH = -1005
NOALLOC_SCRATCH_BYTES = 128 << 20
BSOFI_MAX_N = 1024

PHASE_BSOF, PHASE_BSTRI, PHASE_BSOI, PHASE_ALL, PHASE_DIAG_ONLY = 1, 2, 4, 7, 8
FAMILY = {"panel": 0, "refl": 1, "gemm": 2, "trtri": 3, "eye": 4, "scan": 5, "larft": 6, "det": 8, "flow": 9}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class bsofi_matrix_t(C.Structure):
    _fields_ = [("A", _vp), ("tau", _vp), ("n", C.c_int), ("p", C.c_int), ("lda", C.c_int), ("pad", C.c_int)]


def _load():
    # One HIP runtime per process: torch bundles its own libamdhip64 (same
    # soname, different file).  Loading torch first makes this library bind
    # to the runtime already in the process instead of mapping a second one.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run __graft_entry__.build() (make -C bsofi_amd/csrc)")
    lib = C.CDLL(LIB_PATH, mode=C.RTLD_LOCAL)
    i, d = C.c_int, C.c_double
    sig = {
        "lapackXbsofiLWork": (i, [i, i, i]),
        "lapackXbsof": (i, [i, i, _dp, i, _dp, _dp, i, _ip]),
        "lapackXbstri": (i, [i, i, _dp, i, _ip]),
        "lapackXbsoi_ormqr": (i, [i, i, _dp, i, _dp, _dp, i, _ip]),
        "lapackXbsoi_orgqr": (i, [i, i, _dp, i, _dp, _dp, i, _ip]),
        "hybridXbsofiLWork": (i, [i, i, i, i]),
        "hybridXbsoftri": (i, [_vp, i, i, _dp, i, _dp, _dp, i, _vp, i, _ip]),
        "hybridXbsoi": (i, [_vp, i, i, _dp, i, _dp, _dp, i, _vp, i, _ip]),
        "hybridxinith_": (i, [_ip, _ip, _ip, _ip]),
        "hybridxbsofilworkh_": (i, [_ip, _ip, _ip, _ip]),
        "hybridxbsoftrih_": (i, [_ip, _ip, _ip, _dp, _ip, _dp, _dp, _ip, _ip]),
        "hybridxbsoih_": (i, [_ip, _ip, _ip, _dp, _ip, _dp, _dp, _ip, _ip]),
        "hybridxfinalizeh_": (i, [_ip]),
        "gpuXbsof": (i, [i, i, _vp, i, _dp, _dp, i, _vp, _ip]),
        "gpuXbstri_rseq": (i, [i, i, _vp, i, _vp, i, _ip]),
        "gpuXbsoi_ormqr": (i, [i, i, _vp, i, _dp, _dp, i, _vp, _vp, i, _ip]),
        "bsofi_device_count": (i, []),
        "bsofi_check_args": (i, [i, i, i]),
        "bsofi_ctx_create": (i, [C.POINTER(_vp), i, _vp]),
        "bsofi_ctx_destroy": (i, [_vp]),
        "bsofi_ctx_set_stream": (i, [_vp, _vp]),
        "bsofi_ctx_stream": (_vp, [_vp]),
        "bsofi_ctx_synchronize": (i, [_vp]),
        "bsofi_ctx_set_tuning": (i, [_vp, i, i, i]),
        "bsofi_plan_det": (i, [_vp, _dp, _dp]),
        "bsofi_det": (i, [_vp, _vp, i, i, i, _vp, _dp, _dp]),
        "bsofi_ctx_cached_bytes": (C.c_size_t, [_vp]),
        "bsofi_ctx_set_lanes": (i, [_vp, i]),
        "bsofi_ctx_set_schedule": (i, [_vp, i]),
        "bsofi_ctx_set_tuning_ex": (i, [_vp, i, i, i, i, i]),
        "bsofi_ctx_get_tuning_ex": (i, [_vp, i, _ip, _ip, _ip, _ip]),
        "bsofi_ctx_set_wait_limit": (i, [_vp, d]),
        "bsofi_fp64_rate": (i, [_vp, i, _dp]),
        "bsofi_ctx_release_cache": (i, [_vp]),
        "bsofi_ctx_get_tuning": (i, [_vp, i, C.POINTER(i), C.POINTER(i)]),
        "bsofi_ctx_clear_tuning": (i, [_vp]),
        "bsofi_workspace_bytes": (C.c_size_t, [i, i]),
        "bsofi_plan_workspace_bytes": (C.c_size_t, [i, C.POINTER(bsofi_matrix_t), i]),
        "bsofi_plan_create": (i, [_vp, i, C.POINTER(bsofi_matrix_t), i, _vp, C.POINTER(_vp)]),
        "bsofi_plan_execute": (i, [_vp]),
        "bsofi_plan_use_graph": (i, [_vp, i]),
        "bsofi_plan_num_launches": (i, [_vp]),
        "bsofi_plan_info": (i, [_vp, _ip]),
        "bsofi_plan_time_family": (i, [_vp, i]),
        "bsofi_plan_family_time": (i, [_vp, _dp, _ip]),
        "bsofi_plan_destroy": (i, [_vp]),
        "bsofi_invert_batched": (i, [_vp, i, C.POINTER(bsofi_matrix_t), i, _ip]),
        "bsofi_invert_batched_noalloc": (i, [_vp, i, C.POINTER(bsofi_matrix_t), i, _ip]),
        "bsofi_gather_blocks": (i, [_vp, _vp, i, i, i, _vp]),
        "bsofi_pcyclic_residual": (i, [_vp, _vp, _vp, i, i, i, _dp]),
        "bsofi_flops": (None, [i, i, _dp]),
        "bsofi_debug_gemm": (i, [_vp, i, i, i, d, _vp, i, _vp, i, d, _vp, i, i, i]),
        "bsofi_dqmc_expk": (i, [i, i, d, d, _dp]),
        "bsofi_dqmc_fields": (i, [C.c_ulonglong, i, i, i, C.POINTER(C.c_byte)]),
        "bsofi_dqmc_build": (i, [_vp, _vp, i, i, i, _vp, _vp, d, d]),
        "bsofi_dqmc_flip": (i, [_vp, _vp, _vp, i]),
        "bsofi_memcpy": (i, [_vp, _vp, C.c_size_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _p(a):
    return a.ctypes.data_as(_dp)


def _check_host(A, n, L):
    if not (isinstance(A, np.ndarray) and A.dtype == np.float64 and A.flags.f_contiguous):
        raise TypeError("A must be a Fortran-ordered float64 numpy array")
    if A.shape[1] < n * L or A.shape[0] < n * L:
        raise ValueError("A is smaller than N x N")


# -- reference entry points on host arrays -----------------------------------
def lapackXbsofiLWork(n, L, lda):
    return lib.lapackXbsofiLWork(n, L, lda)


def lapackXbsof(n, L, A, tau, work=None):
    """BSOF in place (src/bsofi_cpu.c:86-118 semantics); returns info."""
    _check_host(A, n, L)
    work = work if work is not None else np.zeros(max(lapackXbsofiLWork(n, L, A.shape[0]), 1))
    info = C.c_int(0)
    lib.lapackXbsof(n, L, _p(A), A.shape[0], _p(tau), _p(work), work.size, C.byref(info))
    return info.value


def lapackXbstri(n, L, A):
    _check_host(A, n, L)
    info = C.c_int(0)
    lib.lapackXbstri(n, L, _p(A), A.shape[0], C.byref(info))
    return info.value


def lapackXbsoi_ormqr(n, L, A, tau, work=None):
    _check_host(A, n, L)
    work = work if work is not None else np.zeros(max(lapackXbsofiLWork(n, L, A.shape[0]), 1))
    info = C.c_int(0)
    lib.lapackXbsoi_ormqr(n, L, _p(A), A.shape[0], _p(tau), _p(work), work.size, C.byref(info))
    return info.value


def lapackXbsoi_orgqr(n, L, A, tau, work=None):
    _check_host(A, n, L)
    work = work if work is not None else np.zeros(max(lapackXbsofiLWork(n, L, A.shape[0]), 1))
    info = C.c_int(0)
    lib.lapackXbsoi_orgqr(n, L, _p(A), A.shape[0], _p(tau), _p(work), work.size, C.byref(info))
    return info.value


lapackXbsoi = lapackXbsoi_ormqr


def hybridXbsofiLWork(n, L, lda, lworkHostDevice):
    return lib.hybridXbsofiLWork(n, L, lda, lworkHostDevice)


def hybridXbsoftri(n, L, A, tau, work=None, dwork_ptr=None, ldwork=0):
    _check_host(A, n, L)
    work = work if work is not None else np.zeros(max(hybridXbsofiLWork(n, L, A.shape[0], -1), 1))
    info = C.c_int(0)
    lib.hybridXbsoftri(None, n, L, _p(A), A.shape[0], _p(tau), _p(work), work.size, dwork_ptr, ldwork,
                       C.byref(info))
    return info.value


def hybridXbsoi(n, L, A, tau, work=None, dwork_ptr=None, ldwork=0):
    _check_host(A, n, L)
    work = work if work is not None else np.zeros(max(hybridXbsofiLWork(n, L, A.shape[0], -1), 1))
    info = C.c_int(0)
    lib.hybridXbsoi(None, n, L, _p(A), A.shape[0], _p(tau), _p(work), work.size, dwork_ptr, ldwork,
                    C.byref(info))
    return info.value


class HandleAPI:
    """The simplified handle API of src/bsofi_callback.c:101-180, called the
    way a Fortran caller does (every argument by reference; the handle is an
    integer array, README.md:37-52 of the reference)."""

    def __init__(self, n, L, lda, nints=10):
        self.h = (C.c_int * nints)()
        self.n, self.L, self.lda = C.c_int(n), C.c_int(L), C.c_int(lda)
        self.rc = lib.hybridxinith_(self.h, C.byref(self.n), C.byref(self.L), C.byref(self.lda))

    def lwork(self):
        return lib.hybridxbsofilworkh_(self.h, C.byref(self.n), C.byref(self.L), C.byref(self.lda))

    def bsoftri(self, A, tau, work):
        info = C.c_int(0)
        lw = C.c_int(work.size)
        lib.hybridxbsoftrih_(self.h, C.byref(self.n), C.byref(self.L), _p(A), C.byref(self.lda), _p(tau), _p(work),
                             C.byref(lw), C.byref(info))
        return info.value

    def bsoi(self, A, tau, work):
        info = C.c_int(0)
        lw = C.c_int(work.size)
        lib.hybridxbsoih_(self.h, C.byref(self.n), C.byref(self.L), _p(A), C.byref(self.lda), _p(tau), _p(work),
                          C.byref(lw), C.byref(info))
        return info.value

    def finalize(self):
        return lib.hybridxfinalizeh_(self.h)


# -- native device API --------------------------------------------------------
def device_count():
    return lib.bsofi_device_count()


def flop_counts(n, L):
    out = (C.c_double * 3)()
    lib.bsofi_flops(n, L, out)
    return tuple(out)


class Context:
    """A device context with its own stream (or a caller-supplied one)."""

    def __init__(self, device=0, stream=None):
        h = _vp()
        rc = lib.bsofi_ctx_create(C.byref(h), device, stream)
        if rc:
            raise RuntimeError(f"bsofi_ctx_create failed ({rc})")
        self.h = h

    @property
    def stream(self):
        return lib.bsofi_ctx_stream(self.h)

    def set_stream(self, stream):
        lib.bsofi_ctx_set_stream(self.h, stream)

    def set_lanes(self, lanes):
        """1 or 2 streams for batch plans created afterwards (default 2)."""
        if lanes not in (1, 2):
            raise ValueError("lanes must be 1 or 2")
        rc = lib.bsofi_ctx_set_lanes(self.h, lanes)
        if rc:
            raise RuntimeError(f"bsofi_ctx_set_lanes failed ({rc})")

    def set_schedule(self, mode):
        """1: BSOF as one device-scheduled launch where supported, 0: one
        launch per sweep step."""
        rc = lib.bsofi_ctx_set_schedule(self.h, mode)
        if rc:
            raise ValueError(f"bsofi_ctx_set_schedule({mode}) failed ({rc})")

    def set_wait_limit(self, seconds):
        rc = lib.bsofi_ctx_set_wait_limit(self.h, float(seconds))
        if rc:
            raise ValueError(f"bsofi_ctx_set_wait_limit({seconds}) failed ({rc})")

    def synchronize(self):
        rc = lib.bsofi_ctx_synchronize(self.h)
        if rc:
            raise RuntimeError(f"synchronize failed ({rc})")

    def set_tuning(self, n, nb=0, scan_mti=0, gemm_tile=0, refl_spw=0):
        """Sub-panel width, scan form, BSTRI GEMM tile and reflector slabs per
        workgroup for block size n (0 = default)."""
        rc = lib.bsofi_ctx_set_tuning_ex(self.h, n, nb, scan_mti, gemm_tile, refl_spw)
        if rc:
            raise ValueError(f"unsupported tuning n={n} nb={nb} scan_mti={scan_mti} gemm_tile={gemm_tile} "
                             f"refl_spw={refl_spw} ({rc})")

    def get_tuning_ex(self, n):
        v = [C.c_int(0) for _ in range(4)]
        rc = lib.bsofi_ctx_get_tuning_ex(self.h, n, *[C.byref(x) for x in v])
        if rc:
            raise RuntimeError(f"get_tuning_ex failed ({rc})")
        return tuple(x.value for x in v)

    def get_tuning(self, n):
        nb, mti = C.c_int(0), C.c_int(0)
        rc = lib.bsofi_ctx_get_tuning(self.h, n, C.byref(nb), C.byref(mti))
        if rc:
            raise RuntimeError(f"get_tuning failed ({rc})")
        return nb.value, mti.value

    def clear_tuning(self):
        lib.bsofi_ctx_clear_tuning(self.h)

    def close(self):
        if self.h:
            lib.bsofi_ctx_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter shutdown the HIP runtime may already be torn down:
        # leave the memory to the process exit instead of calling into it
        # (sys.is_finalizing is bound when the class is defined: the module's
        # globals, sys among them, may already be gone when this runs)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def matrices(descs):
    """descs: iterable of (A_ptr, tau_ptr, n, p, lda) -> ctypes array."""
    descs = list(descs)
    arr = (bsofi_matrix_t * max(len(descs), 1))()
    for k, (a, t, n, p, lda) in enumerate(descs):
        arr[k] = bsofi_matrix_t(a, t, n, p, lda, 0)
    return arr, len(descs)


class Plan:
    """The launch sequence of a batch (see bsofi_plan_create)."""

    def __init__(self, ctx: Context, descs, phases=PHASE_ALL, workspace=None):
        self.mats, self.count = matrices(descs)
        h = _vp()
        rc = lib.bsofi_plan_create(ctx.h, self.count, self.mats, phases, workspace, C.byref(h))
        if rc:
            raise RuntimeError(f"bsofi_plan_create failed ({rc})")
        self.h = h
        self.ctx = ctx

    def execute(self):
        rc = lib.bsofi_plan_execute(self.h)
        if rc:
            raise RuntimeError(f"bsofi_plan_execute failed ({rc})")

    def use_graph(self, on=True):
        lib.bsofi_plan_use_graph(self.h, 1 if on else 0)

    @property
    def num_launches(self):
        return lib.bsofi_plan_num_launches(self.h)

    def info(self):
        out = (C.c_int * max(self.count, 1))()
        rc = lib.bsofi_plan_info(self.h, out)
        if rc:
            raise RuntimeError(f"bsofi_plan_info failed ({rc})")
        return list(out)[: self.count]

    def det(self):
        """[(sign(det A), log|det A|)] per matrix, from the last execution."""
        n = max(self.count, 1)
        sg, lg = (C.c_double * n)(), (C.c_double * n)()
        rc = lib.bsofi_plan_det(self.h, sg, lg)
        if rc:
            raise RuntimeError(f"bsofi_plan_det failed ({rc})")
        return [(sg[k], lg[k]) for k in range(self.count)]

    def time_family(self, family):
        lib.bsofi_plan_time_family(self.h, FAMILY[family] if isinstance(family, str) else family)

    def family_time(self):
        ms = C.c_double(0)
        n = C.c_int(0)
        lib.bsofi_plan_family_time(self.h, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def close(self):
        if getattr(self, "h", None):
            lib.bsofi_plan_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter shutdown the HIP runtime may already be torn down:
        # leave the memory to the process exit instead of calling into it
        # (sys.is_finalizing is bound when the class is defined: the module's
        # globals, sys among them, may already be gone when this runs)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def invert_batched(ctx: Context, descs, phases=PHASE_ALL):
    mats, count = matrices(descs)
    info = (C.c_int * max(count, 1))()
    rc = lib.bsofi_invert_batched(ctx.h, count, mats, phases, info)
    if rc:
        raise RuntimeError(f"bsofi_invert_batched failed ({rc})")
    return list(info)[:count]


def det(ctx: Context, dA_ptr, lda, n, p, dtau_ptr):
    """(sign(det A), log|det A|) of a matrix factored by BSOF on the device."""
    sg, lg = C.c_double(0), C.c_double(0)
    rc = lib.bsofi_det(ctx.h, dA_ptr, lda, n, p, dtau_ptr, C.byref(sg), C.byref(lg))
    if rc:
        raise RuntimeError(f"bsofi_det failed ({rc})")
    return sg.value, lg.value


def fp64_rate(ctx: Context, kind):
    """Achievable FP64 TFLOP/s: kind 0 = MFMA 16x16x4, 1 = MFMA 4x4x4, 2 = FMA."""
    out = C.c_double(0)
    rc = lib.bsofi_fp64_rate(ctx.h, kind, C.byref(out))
    if rc:
        raise RuntimeError(f"bsofi_fp64_rate failed ({rc})")
    return out.value


def pcyclic_residual(ctx: Context, blocks_ptr, G_ptr, ldg, n, p):
    r = C.c_double(0)
    rc = lib.bsofi_pcyclic_residual(ctx.h, blocks_ptr, G_ptr, ldg, n, p, C.byref(r))
    if rc:
        raise RuntimeError(f"bsofi_pcyclic_residual failed ({rc})")
    return r.value


def gather_blocks(ctx: Context, A_ptr, lda, n, p, blocks_ptr):
    rc = lib.bsofi_gather_blocks(ctx.h, A_ptr, lda, n, p, blocks_ptr)
    if rc:
        raise RuntimeError(f"bsofi_gather_blocks failed ({rc})")


def copy_to_host(dptr, n_doubles):
    out = np.empty(n_doubles)
    rc = lib.bsofi_memcpy(out.ctypes.data, dptr, out.nbytes)
    if rc:
        raise RuntimeError(f"bsofi_memcpy failed ({rc})")
    return out


def dqmc_expk(lx, ly, t=1.0, dtau=0.125):
    n = lx * ly
    out = np.zeros((n, n), order="F")
    rc = lib.bsofi_dqmc_expk(lx, ly, t, dtau, _p(out))
    if rc:
        raise ValueError(rc)
    return out


def dqmc_fields(seed, instance, p, n):
    out = np.zeros((p, n), dtype=np.int8)
    rc = lib.bsofi_dqmc_fields(seed, instance, p, n, out.ctypes.data_as(C.POINTER(C.c_byte)))
    if rc:
        raise ValueError(rc)
    return out

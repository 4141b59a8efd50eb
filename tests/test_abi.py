"""The drop-in boundary on the CPU: the C ABI library loads, exports every
function include/bsofi.h declares, validates arguments the LAPACK way and,
without a GPU, fails loudly (info = BSOFI_ERR_NO_DEVICE) instead of
falling back to a CPU path.  No compute call reaches a GPU here."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import bsofi_amd as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bsofi.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|double|size_t|void\*)\s*\*?\s*([A-Za-z_]\w*)\s*\(", txt, flags=re.M)
    return sorted(set(names))


def exported_symbols():
    out = subprocess.run(["nm", "-D", "--defined-only", B.LIB_PATH], capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_declares_reference_entry_points():
    names = set(declared_functions())
    for ref_name in ["lapackXbsofiLWork", "lapackXbsof", "lapackXbstri", "lapackXbsoi_ormqr", "lapackXbsoi_orgqr",
                     "lapackxbsof_", "lapackxbstri_", "lapackxbsoi_ormqr_", "lapackxbsoi_orgqr_", "lapackxbsofilwork_",
                     "hybridXbsofiLWork", "hybridXbsoftri", "hybridXbsoi", "hybridxbsofilwork_", "hybridxbsoftri_",
                     "hybridxbsoi_", "hybridxinith_", "hybridxbsofilworkh_", "hybridxbsoftrih_", "hybridxbsoih_",
                     "hybridxfinalizeh_", "fgetwalltime_", "felapsed_", "gpuXbsof", "gpuXbstri_rseq",
                     "gpuXbstri_rbatched", "gpuXbsoi_ormqr", "gpuXbsoi_orgqr", "gpuXbsoi_batched"]:
        assert ref_name in names, ref_name


def test_library_exports_every_declared_function():
    exported = exported_symbols()
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_only_c_abi_is_exported():
    assert not [s for s in exported_symbols() if s.startswith("_Z")]


def test_product_library_does_not_link_the_oracle():
    out = subprocess.run(["ldd", B.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "mkl" not in out and "bsofi_ref" not in out


def test_argument_checks_follow_lapack():
    A = np.zeros((8, 8), order="F")
    tau = np.zeros(8)
    assert B.lapackXbsof(0, 4, A, tau) == -1
    assert B.lib.bsofi_check_args(2, 1, 8) == -2
    assert B.lib.bsofi_check_args(2, 4, 7) == -4
    assert B.lib.bsofi_check_args(B.BSOFI_MAX_N + 1, 2, 4096) == B.BSOFI_ERR_UNSUPPORTED
    info = C.c_int(99)
    w = np.zeros(4)
    rc = B.lib.lapackXbsof(2, 4, A.ctypes.data_as(C.POINTER(C.c_double)), 8, None,
                           w.ctypes.data_as(C.POINTER(C.c_double)), 4, C.byref(info))
    assert rc == -5 and info.value == -5
    rc = B.lib.lapackXbsof(2, 4, A.ctypes.data_as(C.POINTER(C.c_double)), 8,
                           tau.ctypes.data_as(C.POINTER(C.c_double)), w.ctypes.data_as(C.POINTER(C.c_double)), 0,
                           C.byref(info))
    assert info.value == -7


def test_fortran_names_take_pointers():
    n, L, lda, info = C.c_int(0), C.c_int(4), C.c_int(8), C.c_int(0)
    A = np.zeros((8, 8), order="F")
    tau = np.zeros(8)
    w = np.zeros(4)
    lw = C.c_int(4)
    dp = C.POINTER(C.c_double)
    B.lib.lapackxbsof_.argtypes = None
    rc = B.lib.lapackxbsof_(C.byref(n), C.byref(L), A.ctypes.data_as(dp), C.byref(lda), tau.ctypes.data_as(dp),
                            w.ctypes.data_as(dp), C.byref(lw), C.byref(info))
    assert rc == -1 and info.value == -1
    B.lib.lapackxbsofilwork_.argtypes = None
    assert B.lib.lapackxbsofilwork_(C.byref(C.c_int(4)), C.byref(C.c_int(5)), C.byref(C.c_int(20))) >= 1


def test_workspace_queries():
    n, L = 16, 8
    N = n * L
    w = np.zeros(1)
    info = C.c_int(5)
    dp = C.POINTER(C.c_double)
    B.lib.lapackXbsof(n, L, None, N, None, w.ctypes.data_as(dp), -1, C.byref(info))
    assert info.value == 0 and w[0] >= 1
    assert B.lapackXbsofiLWork(n, L, N) >= 1
    assert B.hybridXbsofiLWork(n, L, N, -1) >= 1
    # the device query covers a device copy of A plus the workspace
    assert B.hybridXbsofiLWork(n, L, N, -2) >= N * N + N
    assert B.lib.bsofi_workspace_bytes(n, L) > 0


@pytest.mark.skipif(B.device_count() > 0, reason="checks the no-GPU behaviour")
def test_no_gpu_means_loud_failure_not_fallback():
    A0 = np.asfortranarray(np.eye(8))
    A = A0.copy(order="F")
    tau = np.zeros(8)
    assert B.lapackXbsof(2, 4, A, tau) == B.BSOFI_ERR_NO_DEVICE
    assert np.array_equal(A, A0)  # untouched: nothing computed on the CPU
    assert B.lapackXbstri(2, 4, A) == B.BSOFI_ERR_NO_DEVICE
    assert B.lapackXbsoi_ormqr(2, 4, A, tau) == B.BSOFI_ERR_NO_DEVICE
    h = B.HandleAPI(2, 4, 8)
    assert h.rc == -1  # cb_hybridXinitH's device-failure code
    with pytest.raises(RuntimeError):
        B.Context(0)


def test_handle_misuse_is_reported():
    h = (C.c_int * 10)()
    A = np.zeros((8, 8), order="F")
    tau = np.zeros(8)
    w = np.zeros(8)
    info = C.c_int(0)
    dp = C.POINTER(C.c_double)
    n, L, lda, lw = C.c_int(2), C.c_int(4), C.c_int(8), C.c_int(8)
    B.lib.hybridxbsoftrih_(h, C.byref(n), C.byref(L), A.ctypes.data_as(dp), C.byref(lda), tau.ctypes.data_as(dp),
                           w.ctypes.data_as(dp), C.byref(lw), C.byref(info))
    assert info.value == B.BSOFI_ERR_HANDLE
    assert B.lib.hybridxfinalizeh_(h) == B.BSOFI_ERR_HANDLE


def test_walltime_helpers():
    t0 = (C.c_double * 2)()
    t1 = (C.c_double * 2)()
    B.lib.fgetwalltime_.argtypes = [C.c_void_p]
    B.lib.felapsed_.argtypes = [C.c_void_p, C.c_void_p]
    B.lib.felapsed_.restype = C.c_double
    B.lib.fgetwalltime_(t0)
    B.lib.fgetwalltime_(t1)
    assert 0.0 <= B.lib.felapsed_(t1, t0) < 5.0


def test_wrappers_alive_at_interpreter_exit_stay_quiet():
    """Context, Plan and Server objects still alive when the interpreter
    tears its modules down skip the HIP cleanup without an error: the
    finalizing check is bound when the class is defined, not looked up
    through the module's globals (which may already be gone)."""
    import sys

    code = (
        f"import sys; sys.path.insert(0, {ROOT!r})\n"
        "import builtins\n"
        "import bsofi_amd as B\n"
        "from bsofi_amd import serve\n"
        "objs = []\n"
        "for cls in (B.Context, B.Plan, serve.Server):\n"
        "    o = object.__new__(cls)\n"
        "    o.h = None\n"
        "    objs.append(o)\n"
        "builtins._bsofi_keep = objs\n"
    )
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Exception ignored" not in r.stderr and "Traceback" not in r.stderr, r.stderr

"""GPU parity: the HIP path against the oracle on the same seeded inputs.

Tolerances (FP64 throughout):
* every inverse: ||A G - I||_F / (||A||_F ||G||_F) <= 100 N eps;
* beta = p * dtau <= 4: ||G - G_ref||_F / ||G_ref||_F <= 1e-10;
* intermediates (A = V + R after BSOF, tau, A = V + R^{-1} after BSTRI) are
  compared with 1e-10 relative Frobenius error: the blocked GPU reflectors
  and LAPACK's apply the same Householder transformations in a different
  order, so they agree to rounding, not bit for bit.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from bsofi_amd import dqmc

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
GOLD_CASES = sorted(p for p in glob.glob(os.path.join(GOLD, "*.npz")) if "checksums" not in p)


@pytest.fixture(scope="module")
def B():
    import bsofi_amd

    if bsofi_amd.device_count() < 1:
        pytest.fail("GPU tests need a GPU (the HIP path has no CPU fallback)")
    return bsofi_amd


@pytest.fixture(scope="module")
def O():
    return oracle.Oracle("auto", threads=1)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def make(kind, a, p, seed=7):
    if kind == "dqmc":
        A, m = dqmc.dqmc_matrix(a, p)
        return A, m["n"]
    return dqmc.random_pcyclic(a, p, seed=seed, identity_diag=(kind == "rand_eye")), a


PHASE_CASES = [("dqmc", 2, 2), ("dqmc", 2, 3), ("dqmc", 2, 4), ("dqmc", 3, 5), ("dqmc", 4, 8), ("dqmc", 6, 4),
               ("dqmc", 8, 6), ("dqmc", 10, 3), ("rand_eye", 1, 5), ("rand_eye", 17, 4), ("rand_full", 33, 3),
               ("dqmc", 12, 3), ("rand_eye", 130, 2),
               # 2n above 512: the m <= 1024 variants of the panel and reflector kernels
               ("rand_eye", 300, 3), ("dqmc", 20, 2),
               # 8 | n in [128, 256]: the 4x4x4 scan by default
               ("rand_eye", 128, 3), ("rand_full", 160, 2),
               # BSTRI's row sweep by column strips with a partial last row-tile pair (n = 224: 14 tiles
               # over 8 waves) and chains of 1 .. 4
               ("rand_eye", 224, 5),
               # n = 144, p = 16: BSOI's Q_k formation steps hold more slabs than the GPU has CUs, so a
               # single-matrix plan takes the multi-slab reflector form
               ("dqmc", 12, 16)]


@pytest.mark.parametrize("kind,a,p", PHASE_CASES)
def test_phases_match_oracle(B, O, kind, a, p, eps):
    A0, n = make(kind, a, p)
    N = n * p
    print(f"case {kind} n={n} p={p}", flush=True)
    Ao = A0.copy(order="F")
    tau_o = np.zeros(N)
    assert O.bsof(n, p, Ao, tau_o) == 0
    Ag = A0.copy(order="F")
    tau_g = np.zeros(N)
    assert B.lapackXbsof(n, p, Ag, tau_g) == 0
    assert rel(tau_g, tau_o) <= 1e-10
    assert rel(Ag, Ao) <= 1e-10
    Xo = Ao.copy(order="F")
    O.bstri(n, p, Xo)
    Xg = Ao.copy(order="F")
    assert B.lapackXbstri(n, p, Xg) == 0
    assert rel(Xg, Xo) <= 1e-10
    Go = Xo.copy(order="F")
    O.bsoi(n, p, Go, tau_o)
    Gg = Xo.copy(order="F")
    assert B.lapackXbsoi_ormqr(n, p, Gg, tau_o.copy()) == 0
    assert rel(Gg, Go) <= 1e-10
    assert oracle.residual(A0, Gg) <= 100 * N * eps
    Gg2 = Xo.copy(order="F")
    assert B.lapackXbsoi_orgqr(n, p, Gg2, tau_o.copy()) == 0
    assert np.array_equal(Gg2, Gg)


@pytest.mark.parametrize("scale", [1e-300, 1e-160, 1e160])
def test_scaled_inputs_match_oracle(B, O, scale):
    """Entries whose squares leave the panel's fast dlarfg range (sums of
    squares outside [2^-960, 2^960]) take LAPACK's scaled path; 1e-300 also
    runs dlarfg's safe-minimum rescaling loop (|beta| < safmin)."""
    A0, n = make("rand_eye", 17, 4)
    p = 4
    N = n * p
    A0 = A0 * scale
    print(f"case scaled {scale:g} n={n} p={p}", flush=True)

    def srel(a, b):  # relative error, norms taken after rescaling by max|b|
        s = np.abs(b).max()
        return rel(a / s, b / s)

    Ao = A0.copy(order="F")
    tau_o = np.zeros(N)
    assert O.bsof(n, p, Ao, tau_o) == 0
    Ag = A0.copy(order="F")
    tau_g = np.zeros(N)
    assert B.lapackXbsof(n, p, Ag, tau_g) == 0
    assert np.all(np.isfinite(Ag)) and np.all(np.isfinite(tau_g))
    assert rel(tau_g, tau_o) <= 1e-10
    # V (below the diagonal) is scale-free, R (on and above it) carries the scale
    low = np.tril(np.ones((N, N), dtype=bool), -1)
    assert srel(np.where(low, Ag, 0.0), np.where(low, Ao, 0.0)) <= 1e-10
    assert srel(np.where(low, 0.0, Ag), np.where(low, 0.0, Ao)) <= 1e-10


@pytest.mark.parametrize("path", GOLD_CASES, ids=[os.path.basename(p)[:-4] for p in GOLD_CASES])
def test_golden_fixtures_through_handle_api(B, path, eps):
    g = np.load(path)
    n, p = int(g["n"]), int(g["p"])
    N = n * p
    A = np.asfortranarray(g["A"].copy())
    tau = np.zeros(N)
    h = B.HandleAPI(n, p, N)
    assert h.rc == 0
    work = np.zeros(max(h.lwork(), 1))
    assert h.bsoftri(A, tau, work) == 0
    assert rel(tau, g["tau"]) <= 1e-10
    assert rel(A, g["A_bstri"]) <= 1e-10  # observable state after bsoftriH: V + R^{-1}
    assert h.bsoi(A, tau, work) == 0
    assert h.finalize() == 0
    assert rel(A, g["G"]) <= 1e-10
    assert oracle.residual(g["A"], A) <= 100 * N * eps


def test_c1_checksums(B, eps):
    z = np.load(os.path.join(GOLD, "dqmc_l8_p16_checksums.npz"))
    n, p = int(z["n"]), int(z["p"])
    N = n * p
    A0, _ = dqmc.dqmc_matrix(8, p)
    A = A0.copy(order="F")
    tau = np.zeros(N)
    assert B.hybridXbsoftri(n, p, A, tau) == 0
    assert rel(tau, z["tau"]) <= 1e-10
    assert B.hybridXbsoi(n, p, A, tau) == 0
    assert abs(np.linalg.norm(A) - float(z["G_fro"])) <= 1e-10 * float(z["G_fro"])
    idx = z["sample_idx"]
    assert np.allclose(A[idx[:, 0], idx[:, 1]], z["sample_val"], rtol=1e-9, atol=1e-12 * np.abs(A).max())
    Gd = np.stack([A[k * n:(k + 1) * n, k * n:(k + 1) * n] for k in range(p)])
    assert rel(Gd, z["G_diag_blocks"]) <= 1e-10
    assert oracle.residual(A0, A) <= 100 * N * eps


@pytest.mark.parametrize("lx,p,pad", [(3, 4, 5), (12, 3, 5), (12, 3, 6)])
def test_lda_larger_than_N(B, O, eps, lx, p, pad):
    """Padding rows are left alone (n = 144: the 4x4x4 scan; odd and even
    lda)."""
    A0, m = dqmc.dqmc_matrix(lx, p)
    n = m["n"]
    N = n * p
    lda = N + pad
    print(f"case lda n={n} p={p} lda={lda}", flush=True)
    Ab = np.zeros((lda, N), order="F")
    Ab[:N] = A0
    Ab[N:] = 7.0  # padding rows must be left untouched
    tau = np.zeros(N)
    assert B.lapackXbsof(n, p, Ab, tau) == 0
    assert B.lapackXbstri(n, p, Ab) == 0
    assert B.lapackXbsoi_ormqr(n, p, Ab, tau) == 0
    assert np.all(Ab[N:] == 7.0)
    G = A0.copy(order="F")
    O.invert(n, p, G)
    assert rel(Ab[:N], G) <= 1e-10


def test_singular_R_reports_info(B):
    n, p = 4, 4
    A = dqmc.random_pcyclic(n, p, seed=5)
    A[:, 2] = 0.0  # a zero column makes R(2,2) exactly zero
    tau = np.zeros(n * p)
    assert B.lapackXbsof(n, p, A, tau) == 0
    info = B.lapackXbstri(n, p, A)
    assert info == 3


def test_handle_bsoi_reads_callers_arrays(B, O):
    """bsoiH reads A and tau from the caller's arrays, as the reference's
    cb_hybridXbsoiH does: a caller that replaces them between bsoftriH and
    bsoiH (here by another matrix's V + X and tau from the oracle) gets the
    inverse of that other matrix."""
    lx, p = 4, 5
    A0, m = dqmc.dqmc_matrix(lx, p)
    A1, _ = dqmc.dqmc_matrix(lx, p, instance=9)
    n = m["n"]
    N = n * p
    X1 = A1.copy(order="F")
    tau1 = np.zeros(N)
    assert O.bsof(n, p, X1, tau1) == 0
    assert O.bstri(n, p, X1) == 0
    G1 = A1.copy(order="F")
    O.invert(n, p, G1)
    A = A0.copy(order="F")
    tau = np.zeros(N)
    h = B.HandleAPI(n, p, N, nints=8)
    assert h.rc == 0
    w = np.zeros(16)
    assert h.bsoftri(A, tau, w) == 0
    A[:] = X1
    tau[:] = tau1
    assert h.bsoi(A, tau, w) == 0
    assert h.finalize() == 0
    assert h.finalize() == B.BSOFI_ERR_HANDLE  # a finalized handle is refused, not freed twice
    assert rel(A, G1) <= 1e-10


@pytest.mark.parametrize("lx,p", [(6, 7), (16, 6)])
def test_handle_api_bits_equal_device_path(B, lx, p):
    """The host-array handle API (pinned staging; only the p-cyclic blocks go
    up, V + X and then G come down) gives the bits of the same inversion on
    a device-resident copy."""
    import torch

    A0, m = dqmc.dqmc_matrix(lx, p, instance=3)
    n = m["n"]
    N = n * p
    A = A0.copy(order="F")
    tau = np.zeros(N)
    h = B.HandleAPI(n, p, N)
    work = np.zeros(max(h.lwork(), 1))
    assert h.bsoftri(A, tau, work) == 0
    assert h.bsoi(A, tau, work) == 0
    h.finalize()
    ctx = B.Context(0)
    d = torch.from_numpy(np.ascontiguousarray(A0.T)).cuda()
    t = torch.zeros(N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, N)]) == [0]
    G = d.cpu().numpy().T
    ctx.close()
    assert np.array_equal(A, G)


def test_hybrid_with_caller_device_workspace(B, O):
    import torch

    lx, p = 6, 4
    A0, m = dqmc.dqmc_matrix(lx, p)
    n = m["n"]
    N = n * p
    ldw = B.hybridXbsofiLWork(n, p, N, -2)
    dwork = torch.empty(ldw, dtype=torch.float64, device="cuda")
    A = A0.copy(order="F")
    tau = np.zeros(N)
    assert B.hybridXbsoftri(n, p, A, tau, dwork_ptr=dwork.data_ptr(), ldwork=ldw) == 0
    assert B.hybridXbsoi(n, p, A, tau, dwork_ptr=dwork.data_ptr(), ldwork=ldw) == 0
    G = A0.copy(order="F")
    O.invert(n, p, G)
    assert rel(A, G) <= 1e-10


def test_device_pointer_api(B, O, eps):
    import ctypes as C

    import torch

    lx, p = 4, 6
    A0, m = dqmc.dqmc_matrix(lx, p)
    n = m["n"]
    N = n * p
    ldda = N + 3
    dA = torch.zeros((N, ldda), dtype=torch.float64, device="cuda")  # row r of dA = column r of A
    dA[:, :N] = torch.from_numpy(A0.T.copy())
    tau = np.zeros(N)
    work = np.zeros(4)
    info = C.c_int(0)
    dp = C.POINTER(C.c_double)
    B.lib.gpuXbsof(n, p, dA.data_ptr(), ldda, tau.ctypes.data_as(dp), work.ctypes.data_as(dp), 4, None,
                   C.byref(info))
    assert info.value == 0
    B.lib.gpuXbstri_rseq(n, p, dA.data_ptr(), ldda, None, 0, C.byref(info))
    assert info.value == 0
    B.lib.gpuXbsoi_ormqr(n, p, dA.data_ptr(), ldda, tau.ctypes.data_as(dp), work.ctypes.data_as(dp), 4, None,
                         None, 0, C.byref(info))
    assert info.value == 0
    torch.cuda.synchronize()
    G = dA[:, :N].cpu().numpy().T
    Gr = A0.copy(order="F")
    O.invert(n, p, Gr)
    assert rel(G, Gr) <= 1e-10
    assert oracle.residual(A0, G) <= 100 * N * eps


def test_debug_gemm_matches_numpy(B):
    import torch

    ctx = B.Context(0)
    rng = np.random.default_rng(1)
    # the last one has >= 192 128 x 128 tiles (the big-tile kernel), the rest use 64 x 64 tiles
    for (M, N, K, am, bm) in [(16, 16, 4, 0, 0), (100, 70, 90, 0, 0), (70, 130, 70, 1, 0), (50, 50, 50, 0, 1),
                              (5, 3, 2, 0, 0), (128, 256, 16, 0, 0), (256, 300, 100, 0, 1), (200, 520, 37, 1, 0),
                              (300, 1000, 256, 1, 1), (2048, 1664, 64, 0, 0)]:
        A = rng.standard_normal((M, K))
        Bm = rng.standard_normal((K, N))
        Cm = rng.standard_normal((M, N))
        dA = torch.from_numpy(A.T.copy()).cuda()
        dB = torch.from_numpy(Bm.T.copy()).cuda()
        dC = torch.from_numpy(Cm.T.copy()).cuda()
        rc = B.lib.bsofi_debug_gemm(ctx.h, M, N, K, -1.5, dA.data_ptr(), M, dB.data_ptr(), K, 0.5, dC.data_ptr(), M,
                                    am, bm)
        assert rc == 0
        ref = -1.5 * (np.triu(A) if am else A) @ (np.triu(Bm) if bm else Bm) + 0.5 * Cm
        assert np.abs(dC.cpu().numpy().T - ref).max() <= 1e-13 * np.abs(ref).max()
    ctx.close()


def test_nan_input_runs_through(B, O):
    """A NaN in A, which neither the reference nor LAPACK checks for: the GPU
    path ends with info = 0, as the reference's LAPACK path does on the same
    input, and the NaN reaches G; nothing hangs on it (the panel's norm
    checks send a NaN column down LAPACK's scaled dlarfg path, whose loops
    are bounded)."""
    A0, m = dqmc.dqmc_matrix(3, 4)
    n, p = m["n"], 4
    N = n * p
    A0 = A0.copy(order="F")
    A0[5, 2] = np.nan
    G = A0.copy(order="F")
    _, info_ref, _ = O.invert(n, p, G)
    A = A0.copy(order="F")
    tau = np.zeros(N)
    h = B.HandleAPI(n, p, N)
    work = np.zeros(max(h.lwork(), 1))
    assert h.bsoftri(A, tau, work) == info_ref == 0
    assert h.bsoi(A, tau, work) == 0
    h.finalize()
    assert np.isnan(A).any() and np.isnan(G).any()

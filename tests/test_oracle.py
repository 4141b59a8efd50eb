"""Pins the CPU oracle (oracle/) before anything is compared against it.

* the MKL back end against the reference's own CPU path built from the
  reference sources (oracle/_ref), bit for bit, one thread;
* every back end against the golden fixtures written from oracle/_ref;
* the plain-C LAPACK restatement against MKL (agreement to rounding);
* closed forms that hold whatever the implementation: A G = I, Q^T Q = I
  and Q R = A for the factorization, R X = I for BSTRI, p = 2 reducing to a
  single dgeqrf of the 2n x 2n matrix (scipy's LAPACK), numpy's inverse,
  the sign of det A;
* planted mistakes in a copy of the oracle's outputs (a flipped tau, a
  transposed Q, a dropped block) must fail those checks.
"""
import glob
import os

import numpy as np
import pytest

import oracle
from bsofi_amd import dqmc

GOLD = os.path.join(os.path.dirname(__file__), "golden")
CASES = sorted(p for p in glob.glob(os.path.join(GOLD, "*.npz")) if "checksums" not in p)
BACKENDS = ["builtin"] + (["mkl"] if oracle.mkl_path() else []) + (["openblas"] if oracle.openblas_path() else [])


def load(path):
    z = np.load(path)
    return {k: z[k] for k in z.files}


def tol_for(N, eps):
    return 100 * N * eps


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("path", CASES, ids=[os.path.basename(p)[:-4] for p in CASES])
def test_oracle_matches_golden(path, backend, eps):
    g = load(path)
    n, p = int(g["n"]), int(g["p"])
    N = n * p
    O = oracle.Oracle(backend, threads=1)
    A = np.asfortranarray(g["A"].copy())
    tau = np.zeros(N)
    assert O.bsof(n, p, A, tau) == 0
    rt = 1e-15 * N if backend == "mkl" else 1e-12
    assert np.linalg.norm(tau - g["tau"]) <= rt * np.linalg.norm(g["tau"]) + 1e-300
    assert np.linalg.norm(A - g["A_bsof"]) <= rt * np.linalg.norm(g["A_bsof"])
    O.bstri(n, p, A)
    assert np.linalg.norm(A - g["A_bstri"]) <= 1e-10 * np.linalg.norm(g["A_bstri"])
    O.bsoi(n, p, A, tau)
    assert np.linalg.norm(A - g["G"]) <= 1e-10 * np.linalg.norm(g["G"])
    assert oracle.residual(g["A"], A) <= tol_for(N, eps)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("lx,p", [(2, 2), (2, 3), (3, 4), (4, 6), (6, 5)])
def test_mkl_oracle_bit_identical_to_reference(lx, p):
    A0, meta = dqmc.dqmc_matrix(lx, p)
    n = meta["n"]
    O = oracle.Oracle("mkl", threads=1)
    R = oracle.Reference(threads=1)
    A1 = A0.copy(order="F")
    t1, i1, _ = O.invert(n, p, A1)
    A2 = A0.copy(order="F")
    t2, i2 = R.invert(n, p, A2)
    assert i1 == i2 == 0
    assert np.array_equal(A1, A2) and np.array_equal(t1, t2)
    # the explicit-Q variant reaches the same inverse
    A3 = A0.copy(order="F")
    O.invert(n, p, A3, variant="orgqr")
    A4 = A0.copy(order="F")
    R.invert(n, p, A4, variant="orgqr")
    assert np.array_equal(A3, A4)


@pytest.mark.skipif(not oracle.mkl_path(), reason="MKL absent")
@pytest.mark.parametrize("lx,p", [(2, 5), (4, 4), (8, 6)])
def test_builtin_lapack_agrees_with_mkl(lx, p, eps):
    A0, meta = dqmc.dqmc_matrix(lx, p)
    n = meta["n"]
    N = n * p
    G1 = A0.copy(order="F")
    t1, _, _ = oracle.Oracle("builtin").invert(n, p, G1)
    G2 = A0.copy(order="F")
    t2, _, _ = oracle.Oracle("mkl", threads=1).invert(n, p, G2)
    assert np.linalg.norm(t1 - t2) <= 1e3 * N * eps * np.linalg.norm(t2)
    assert np.linalg.norm(G1 - G2) <= 1e3 * N * eps * np.linalg.norm(G2)


def householder_q(V, tau):
    """Q = H_1 ... H_k from LAPACK-style reflectors (numpy, independent)."""
    m, k = V.shape
    Q = np.eye(m)
    for i in range(k - 1, -1, -1):
        v = np.zeros(m)
        v[i] = 1.0
        v[i + 1:] = V[i + 1:, i]
        Q = Q - tau[i] * np.outer(v, v @ Q)
    return Q


@pytest.mark.parametrize("backend", BACKENDS)
def test_p2_is_one_dgeqrf(backend):
    scipy_lapack = pytest.importorskip("scipy.linalg.lapack")
    A0, meta = dqmc.dqmc_matrix(3, 2)
    n = meta["n"]
    A = A0.copy(order="F")
    tau = np.zeros(2 * n)
    oracle.Oracle(backend, threads=1).bsof(n, 2, A, tau)
    qr, tau_s, work, info = scipy_lapack.dgeqrf(A0.copy(order="F"))
    assert np.allclose(A, qr, rtol=0, atol=1e-13) and np.allclose(tau, tau_s, rtol=0, atol=1e-14)


@pytest.mark.parametrize("backend", BACKENDS)
def test_bsof_factorization_invariants(backend, eps):
    """Q_0 ... Q_{p-2} R = A rebuilt from V and tau, Q orthogonal, R with the
    p-cyclic zero pattern (src/bsofi_cpu.c:44-85 describes the factors)."""
    lx, p = 3, 5
    A0, meta = dqmc.dqmc_matrix(lx, p)
    n = meta["n"]
    N = n * p
    A = A0.copy(order="F")
    tau = np.zeros(N)
    oracle.Oracle(backend, threads=1).bsof(n, p, A, tau)
    Q = np.eye(N)
    for k in range(p - 1):
        w = n if k < p - 2 else 2 * n
        V = A[k * n:k * n + 2 * n, k * n:k * n + w]
        Qk = householder_q(V, tau[k * n:k * n + w])
        E = np.eye(N)
        E[k * n:k * n + 2 * n, k * n:k * n + 2 * n] = Qk
        Q = Q @ E
    R = np.triu(A)
    for i in range(p):
        for j in range(p):
            if j not in (i, i + 1, p - 1) and j > i:
                R[i * n:(i + 1) * n, j * n:(j + 1) * n] = 0.0  # untouched input zeros
    assert np.linalg.norm(Q.T @ Q - np.eye(N)) <= 100 * N * eps
    assert np.linalg.norm(Q @ R - A0) <= 100 * N * eps * np.linalg.norm(A0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_bstri_is_inverse_of_R(backend, eps):
    lx, p = 2, 6
    A0, meta = dqmc.dqmc_matrix(lx, p)
    n = meta["n"]
    N = n * p
    O = oracle.Oracle(backend, threads=1)
    A = A0.copy(order="F")
    tau = np.zeros(N)
    O.bsof(n, p, A, tau)
    R = np.triu(A)
    X = A.copy(order="F")
    O.bstri(n, p, X)
    X = np.triu(X)
    assert np.linalg.norm(R @ X - np.eye(N)) <= 100 * N * eps * np.linalg.norm(R) * np.linalg.norm(X)
    assert np.allclose(X, np.linalg.inv(R), rtol=0, atol=1e-10 * np.abs(X).max())


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("kind", ["dqmc", "rand"])
def test_inverse_against_numpy(backend, kind, eps):
    if kind == "dqmc":
        A0, meta = dqmc.dqmc_matrix(4, 5)
        n, p = meta["n"], 5
    else:
        n, p = 7, 4
        A0 = dqmc.random_pcyclic(n, p, seed=3)
    G = A0.copy(order="F")
    oracle.Oracle(backend, threads=1).invert(n, p, G)
    Gi = np.linalg.inv(A0)
    assert np.linalg.norm(G - Gi) <= 1e-10 * np.linalg.norm(Gi)
    assert oracle.residual(A0, G) <= 100 * n * p * eps


def test_det_sign_matches_slogdet():
    """The sign rule of examples/ftest_cpu/ftest_cpu.f90:133-138 equals
    (-1)^N sign(det A) (every reflector with tau != 0 has determinant -1)."""
    for lx, p in [(2, 3), (3, 4), (2, 7)]:
        A0, meta = dqmc.dqmc_matrix(lx, p)
        n = meta["n"]
        N = n * p
        A = A0.copy(order="F")
        tau = np.zeros(N)
        oracle.Oracle("auto", threads=1).bsof(n, p, A, tau)
        s = oracle.det_sign(tau, A)
        sign, _ = np.linalg.slogdet(A0)
        assert s == sign * (-1) ** N


def test_residual_helpers_agree():
    A0, meta = dqmc.dqmc_matrix(2, 4)
    G = np.linalg.inv(A0)
    G[3, 5] += 1e-6  # a residual well above rounding, so both helpers see the same value
    r1 = oracle.residual(A0, G)
    r2 = oracle.Oracle("builtin").residual(A0, G)
    assert r1 > 1e-9 and abs(r1 - r2) <= 1e-6 * r1


# ---- planted mistakes: each must be caught by the checks above ---------------
def _oracle_outputs():
    A0, meta = dqmc.dqmc_matrix(3, 4)
    n, p = meta["n"], 4
    O = oracle.Oracle("builtin")
    A = A0.copy(order="F")
    tau = np.zeros(n * p)
    O.bsof(n, p, A, tau)
    return A0, A, tau, n, p, O


def test_planted_tau_sign_error_is_caught(eps):
    A0, A, tau, n, p, O = _oracle_outputs()
    bad = tau.copy()
    bad[n + 2] = -bad[n + 2]
    X = A.copy(order="F")
    O.bstri(n, p, X)
    O.bsoi(n, p, X, bad)
    assert oracle.residual(A0, X) > 1e3 * 100 * n * p * eps


def test_planted_transposed_q_is_caught(eps):
    A0, A, tau, n, p, O = _oracle_outputs()
    X = A.copy(order="F")
    O.bstri(n, p, X)
    G = X.copy(order="F")
    O.bsoi(n, p, G, tau)
    # G computed with Q instead of Q^T on the last pair: apply Q twice more
    Vk = A[(p - 2) * n:p * n, (p - 2) * n:p * n]
    Qk = householder_q(Vk, tau[(p - 2) * n:p * n])
    G2 = G.copy()
    G2[:, (p - 2) * n:] = G2[:, (p - 2) * n:] @ Qk @ Qk
    assert oracle.residual(A0, G2) > 1e3 * 100 * n * p * eps


def test_planted_dropped_block_is_caught(eps):
    A0, A, tau, n, p, O = _oracle_outputs()
    X = A.copy(order="F")
    O.bstri(n, p, X)
    X[0:n, (p - 1) * n:] = 0.0  # drop the last-column block of row 0
    O.bsoi(n, p, X, tau)
    assert oracle.residual(A0, X) > 1e3 * 100 * n * p * eps


def test_c1_checksums_reproduced_by_oracle():
    z = np.load(os.path.join(GOLD, "dqmc_l8_p16_checksums.npz"))
    n, p = int(z["n"]), int(z["p"])
    A0, _ = dqmc.dqmc_matrix(8, p)
    G = A0.copy(order="F")
    tau, info, _ = oracle.Oracle("auto", threads=1).invert(n, p, G)
    assert info == 0
    assert np.linalg.norm(tau - z["tau"]) <= 1e-13 * np.linalg.norm(z["tau"])
    assert abs(np.linalg.norm(G) - float(z["G_fro"])) <= 1e-11 * float(z["G_fro"])
    idx = z["sample_idx"]
    assert np.allclose(G[idx[:, 0], idx[:, 1]], z["sample_val"], rtol=1e-10, atol=1e-12)

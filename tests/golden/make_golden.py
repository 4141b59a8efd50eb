"""Writes the golden fixtures of tests/golden/ from the reference's own CPU
path (oracle/_ref/libbsofi_ref.so: SGo-Go/BSOFI src/bsofi_cpu.c,
src/bsofi_lapack.c and src/bsofi_callback.c compiled against MKL by
oracle/Makefile), run with one MKL thread so the bits are reproducible.

The reference ships no golden vectors (its inputs are time()-seeded), so
these fixtures are the reference's outputs on fixed inputs: each case stores
the input A, A and tau after lapackXbsof, A after lapackXbstri, and
G = A^{-1} after lapackXbsoi_ormqr.  Inputs are seeded Hubbard-DQMC
matrices (bsofi_amd.dqmc) and U[0,1] p-cyclic matrices in the patterns of
examples/ctest/ctest.c:38-53 and include/bench/bsofi_random.h:24-44.

    python tests/golden/make_golden.py      (needs /root/reference built into oracle/_ref)
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from bsofi_amd import dqmc  # noqa: E402

CASES = [
    # name, kind, n (or lattice side), p
    ("dqmc_l2_p2", "dqmc", 2, 2),
    ("dqmc_l2_p3", "dqmc", 2, 3),
    ("dqmc_l2_p4", "dqmc", 2, 4),
    ("dqmc_l2_p8", "dqmc", 2, 8),
    ("dqmc_l3_p5", "dqmc", 3, 5),
    ("dqmc_l4_p4", "dqmc", 4, 4),
    ("dqmc_l6_p3", "dqmc", 6, 3),
    ("rand_n5_p3_eye", "rand_eye", 5, 3),
    ("rand_n6_p4_full", "rand_full", 6, 4),
    ("rand_n1_p6_eye", "rand_eye", 1, 6),
]

# C1 of the benchmark set (8x8 lattice, p = 16): too large to store, kept
# as checksums of the reference's G and tau.
C1 = ("dqmc_l8_p16", 8, 16)


def make_input(kind, a, p, seed=7):
    if kind == "dqmc":
        A, meta = dqmc.dqmc_matrix(a, p)
        return A, meta["n"]
    if kind == "rand_eye":
        return dqmc.random_pcyclic(a, p, seed=seed, identity_diag=True), a
    return dqmc.random_pcyclic(a, p, seed=seed, identity_diag=False), a


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libbsofi_ref.so is not built")
    R = oracle.Reference(threads=1)
    meta = {"generator": "oracle/_ref (reference src/bsofi_cpu.c + MKL, 1 thread)",
            "mkl": oracle.mkl_path(), "cases": []}
    for name, kind, a, p in CASES:
        A0, n = make_input(kind, a, p)
        N = n * p
        A = A0.copy(order="F")
        tau, info = R.bsof_bstri(n, p, A.copy(order="F"))  # warm path check
        A = A0.copy(order="F")
        lw = max(R.lwork(n, p, N), 1)
        work = np.zeros(lw)
        import ctypes as C
        tau = np.zeros(N)
        inf = C.c_int(0)
        dp = C.POINTER(C.c_double)
        R.lib.lapackXbsof(n, p, A.ctypes.data_as(dp), N, tau.ctypes.data_as(dp), work.ctypes.data_as(dp), lw,
                          C.byref(inf))
        A_bsof = A.copy(order="F")
        R.lib.lapackXbstri(n, p, A.ctypes.data_as(dp), N, C.byref(inf))
        A_bstri = A.copy(order="F")
        R.lib.lapackXbsoi_ormqr(n, p, A.ctypes.data_as(dp), N, tau.ctypes.data_as(dp), work.ctypes.data_as(dp), lw,
                                C.byref(inf))
        G = A.copy(order="F")
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), A=A0, A_bsof=A_bsof, tau=tau, A_bstri=A_bstri, G=G,
                            n=n, p=p)
        meta["cases"].append({"name": name, "kind": kind, "n": n, "p": p, "info": inf.value})
        print(name, n, p, "resid %.2e" % oracle.residual(A0, G))
    name, lx, p = C1
    A0, m = dqmc.dqmc_matrix(lx, p)
    n = m["n"]
    G = A0.copy(order="F")
    tau, info = R.invert(n, p, G)
    rng = np.random.default_rng(11)
    idx = rng.integers(0, n * p, size=(256, 2))
    np.savez_compressed(os.path.join(HERE, f"{name}_checksums.npz"), n=n, p=p, tau=tau, G_fro=np.linalg.norm(G),
                        G_colsum=G.sum(axis=0), G_rowsum=G.sum(axis=1), sample_idx=idx,
                        sample_val=G[idx[:, 0], idx[:, 1]], G_diag_blocks=np.stack(
                            [G[k * n:(k + 1) * n, k * n:(k + 1) * n] for k in range(p)]))
    meta["cases"].append({"name": name, "kind": "dqmc", "n": n, "p": p, "info": info, "checksums_only": True})
    with open(os.path.join(HERE, "golden_meta.json"), "w") as f:
        json.dump(meta, f, indent=1)


if __name__ == "__main__":
    main()

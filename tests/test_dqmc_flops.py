"""The seeded DQMC generator (host C and numpy twins) and the flop formulas."""
import math

import numpy as np
import pytest

import bsofi_amd as B
from bsofi_amd import dqmc, flops


@pytest.mark.parametrize("lx,ly", [(1, 1), (2, 2), (3, 3), (4, 2), (8, 8), (6, 6), (16, 16)])
def test_expk_closed_form(lx, ly):
    t, dtau = 1.0, 0.125
    E = B.dqmc_expk(lx, ly, t, dtau)
    assert np.allclose(E, E.T, atol=1e-15)
    # eigenvalues of e^{-dtau K}: exp(2 dtau t (cos kx + cos ky)) (K = -t adjacency)
    kx = 2 * np.pi * np.arange(lx) / lx
    ky = 2 * np.pi * np.arange(ly) / ly
    cx = np.where(lx > 1, 2 * np.cos(kx), 0.0) if lx > 2 else (np.array([0.0]) if lx == 1 else 2 * np.cos(kx))
    cy = np.where(ly > 1, 2 * np.cos(ky), 0.0) if ly > 2 else (np.array([0.0]) if ly == 1 else 2 * np.cos(ky))
    lam = np.sort(np.exp(dtau * t * (cx[:, None] + cy[None, :])).ravel())
    assert np.allclose(np.sort(np.linalg.eigvalsh(E)), lam, rtol=1e-13)
    assert np.allclose(E, dqmc.expk(lx, ly, t, dtau), atol=1e-14)


def test_hs_fields_host_twins_identical():
    for seed, inst, p, n in [(dqmc.DEFAULT_SEED, 0, 16, 64), (1, 7, 80, 256), (123, 255, 10, 36)]:
        a = B.dqmc_fields(seed, inst, p, n).astype(np.float64)
        b = dqmc.hs_fields(seed, inst, p, n)
        assert np.array_equal(a, b)
        assert set(np.unique(a)) <= {-1.0, 1.0}
        assert abs(a.mean()) < 0.2


def test_lambda_value():
    assert abs(dqmc.hs_lambda(4.0, 0.125) - 0.7369045906) < 1e-9


def test_assembled_pattern():
    A, m = dqmc.dqmc_matrix(2, 4)
    n, p = m["n"], 4
    ek, s = m["expk"], m["fields"]
    ep, em = math.exp(m["lam"]), math.exp(-m["lam"])
    for k in range(p):
        assert np.array_equal(A[k * n:(k + 1) * n, k * n:(k + 1) * n], np.eye(n))
    assert np.array_equal(A[0:n, (p - 1) * n:], ek * np.where(s[0] > 0, ep, em)[None, :])
    for l in range(2, p + 1):
        assert np.array_equal(A[(l - 1) * n:l * n, (l - 2) * n:(l - 1) * n], -(ek * np.where(s[l - 1] > 0, ep, em)))
    # nothing else is non-zero
    mask = np.zeros_like(A, dtype=bool)
    for k in range(p):
        mask[k * n:(k + 1) * n, k * n:(k + 1) * n] = True
        r = (k + 1) % p
        mask[r * n:(r + 1) * n, k * n:(k + 1) * n] = True
    assert not A[~mask].any()


def test_flops_match_published_figures():
    # FLOPS_BSOFI of include/bsofi_flops.h:19-30 evaluated by hand at these sizes
    assert abs(flops.bsofi(64, 16) - 5.228e8) / 5.228e8 < 1e-3
    assert abs(flops.bsofi(256, 64) - 4.958e11) / 4.958e11 < 1e-3
    assert abs(flops.bsofi(64, 32) - 1.996e9) / 1.996e9 < 1e-3
    for n, p in [(64, 16), (256, 64), (36, 10), (100, 77)]:
        c = B.flop_counts(n, p)
        assert np.allclose(c, (flops.bsof(n, p), flops.bstri(n, p), flops.bsoi(n, p)), rtol=1e-12)

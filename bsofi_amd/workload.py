"""Seeded Hubbard-DQMC p-cyclic matrices built directly in device memory.

A matrix is (lattice side lx, p, U, instance): n = lx^2, B_l =
e^{-dtau K} diag(e^{lambda s_{l,i}}), lambda = acosh(e^{U dtau / 2}),
s = +-1 from the counter-based generator keyed (seed, instance, l, i) --
the same inputs as bsofi_amd.dqmc, built by the library's device generator
(bsofi_dqmc_build).  The 2p non-zero blocks are kept beside A so that the
residual ||A G - I||_F / (||A||_F ||G||_F) can be computed on the device
after A has been overwritten by G.
"""
from __future__ import annotations

import math

import numpy as np

from . import dqmc


class DeviceMatrices:
    """specs: list of (lx, p, U) or (lx, p, U, instance)."""

    def __init__(self, B, ctx, specs, seed=dqmc.DEFAULT_SEED, dtau=0.125, t=1.0, first_instance=0,
                 keep_blocks=True):
        import torch

        self.B, self.ctx = B, ctx
        self.specs = []
        self.A, self.tau, self.blocks, self.fields, self.ek, self.scal = [], [], [], [], [], []
        ek_cache = {}
        for k, sp in enumerate(specs):
            lx, p, U = sp[0], sp[1], sp[2]
            inst = sp[3] if len(sp) > 3 else first_instance + k
            n = lx * lx
            N = n * p
            if lx not in ek_cache:
                ek = B.dqmc_expk(lx, lx, t, dtau)
                ek_cache[lx] = torch.from_numpy(ek.T.copy()).cuda()  # column-major n x n
            lam = dqmc.hs_lambda(U, dtau)
            f = B.dqmc_fields(seed, inst, p, n)
            self.specs.append((lx, n, p, U, inst))
            self.ek.append(ek_cache[lx])
            self.scal.append((math.exp(lam), math.exp(-lam)))
            self.fields.append(torch.from_numpy(f.copy()).cuda())
            self.A.append(torch.empty((N, N), dtype=torch.float64, device="cuda"))
            self.tau.append(torch.zeros(N, dtype=torch.float64, device="cuda"))
            if keep_blocks:
                self.blocks.append(torch.empty(2 * p * n * n, dtype=torch.float64, device="cuda"))
        self.keep_blocks = keep_blocks
        torch.cuda.synchronize()
        self.rebuild()

    def __len__(self):
        return len(self.A)

    def rebuild(self, idx=None):
        B = self.B
        for k in (range(len(self.A)) if idx is None else idx):
            _, n, p, _, _ = self.specs[k]
            ep, em = self.scal[k]
            rc = B.lib.bsofi_dqmc_build(self.ctx.h, self.A[k].data_ptr(), n * p, n, p, self.ek[k].data_ptr(),
                                        self.fields[k].data_ptr(), ep, em)
            if rc:
                raise RuntimeError(f"dqmc_build failed ({rc})")
            if self.keep_blocks:
                B.gather_blocks(self.ctx, self.A[k].data_ptr(), n * p, n, p, self.blocks[k].data_ptr())
        self.ctx.synchronize()

    def desc(self, k):
        _, n, p, _, _ = self.specs[k]
        return (self.A[k].data_ptr(), self.tau[k].data_ptr(), n, p, n * p)

    def descs(self, idx=None):
        return [self.desc(k) for k in (range(len(self.A)) if idx is None else idx)]

    def residual(self, k):
        _, n, p, _, _ = self.specs[k]
        return self.B.pcyclic_residual(self.ctx, self.blocks[k].data_ptr(), self.A[k].data_ptr(), n * p, n, p)

    def release(self):
        import torch

        self.A, self.tau, self.blocks, self.fields = [], [], [], []
        torch.cuda.empty_cache()


def uneven_specs(count=256, seed=dqmc.DEFAULT_SEED, sides=(6, 8, 10, 12, 16), p_lo=10, p_hi=80, U_lo=2.0,
                 U_hi=6.0):
    """The uneven mixed workload: lattice side uniform over `sides`, p
    uniform on {p_lo..p_hi}, U uniform on [U_lo, U_hi], from one seeded
    generator; instance k uses HS fields keyed (seed, 5000 + k)."""
    rng = np.random.default_rng(seed)
    lx = rng.choice(np.array(sides), size=count)
    p = rng.integers(p_lo, p_hi + 1, size=count)
    U = rng.uniform(U_lo, U_hi, size=count)
    return [(int(a), int(b), float(c), 5000 + k) for k, (a, b, c) in enumerate(zip(lx, p, U))]

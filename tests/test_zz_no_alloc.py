"""GPU: bsofi_invert_batched_noalloc, the entry point that allocates no
device memory (its workspace and task tables live in the library's fixed
per-device scratch array), against bsofi_invert_batched on the same inputs:
the same bits, on two calls in a row (the second smaller, so it runs on
what the first left in the scratch), and a clear error for a workspace
larger than the scratch."""
import ctypes as C

import numpy as np
import pytest

import oracle
from bsofi_amd import dqmc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def B():
    import bsofi_amd

    if bsofi_amd.device_count() < 1:
        pytest.fail("GPU tests need a GPU")
    return bsofi_amd


def to_dev(A):
    import torch

    return torch.from_numpy(np.ascontiguousarray(A.T)).cuda()


def from_dev(dA, N):
    return dA.cpu().numpy().T[:N, :N]


def run(B, ctx, specs, noalloc):
    import torch

    inputs = [dqmc.dqmc_matrix(lx, p, instance=200 + k)[0] for k, (lx, p) in enumerate(specs)]
    dims = [(lx * lx, p) for lx, p in specs]
    devs = [to_dev(A) for A in inputs]
    taus = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for n, p in dims]
    torch.cuda.synchronize()
    mats, count = B.matrices([(d.data_ptr(), t.data_ptr(), n, p, n * p) for d, t, (n, p) in zip(devs, taus, dims)])
    info = (C.c_int * count)()
    f = B.lib.bsofi_invert_batched_noalloc if noalloc else B.lib.bsofi_invert_batched
    rc = f(ctx.h, count, mats, B.PHASE_ALL, info)
    assert rc == 0, rc
    assert list(info) == [0] * count
    return inputs, [from_dev(d, n * p) for d, (n, p) in zip(devs, dims)], dims


def test_noalloc_matches_invert_batched(B, eps):
    ctx = B.Context(0)
    for specs in ([(12, 8), (8, 12), (16, 4), (3, 5)], [(6, 5)]):  # the second call smaller than the first
        print(f"noalloc case {specs}", flush=True)
        cached = B.lib.bsofi_ctx_cached_bytes(ctx.h)
        inputs, got, dims = run(B, ctx, specs, True)
        assert B.lib.bsofi_ctx_cached_bytes(ctx.h) == cached  # nothing came from the context's buffers
        _, ref, _ = run(B, ctx, specs, False)
        for k, (a, b) in enumerate(zip(got, ref)):
            assert np.array_equal(a, b), f"matrix {k} {dims[k]} differs from bsofi_invert_batched"
            n, p = dims[k]
            assert oracle.residual(inputs[k], a) <= 100 * n * p * eps
    ctx.close()


def test_noalloc_rejects_workspace_beyond_scratch(B):
    import torch

    ctx = B.Context(0)
    n, p = 256, 64  # about 0.3 GB of workspace
    mats, count = B.matrices([(1, 1, n, p, n * p)])  # never dereferenced: refused before any device work
    info = (C.c_int * 1)()
    assert B.lib.bsofi_invert_batched_noalloc(ctx.h, count, mats, B.PHASE_ALL, info) == B.BSOFI_ERR_SCRATCH
    torch.cuda.synchronize()
    ctx.close()

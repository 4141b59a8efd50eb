"""GPU: batched plans, HIP-graph replay, and the serving loop.

* a batch of matrices of different n and p in one plan gives, for every
  matrix, exactly the bits it gets alone (per-matrix tasks, no cross-matrix
  reductions), and matches the oracle to the stated tolerance;
* a plan captured in a HIP graph and replayed several times, with eager
  work and synchronisations in between, reproduces the eager result bit for
  bit on every replay;
* the device DQMC generator reproduces the host generator bit for bit;
* the serving loop returns, for a request served alone and the same request
  served in a mix of concurrent requests from other chains, the same G.
"""
import math

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


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def to_dev(A):
    import torch

    return torch.from_numpy(np.ascontiguousarray(A.T)).cuda()  # row-major (N, lda) == column-major A


def from_dev(dA, N):
    return dA.cpu().numpy().T[:N, :N]


# (lattice side, p): n from 1 to 256, so that reflector steps of one to four
# row tiles per wave share rounds, and the n = 256 matrix's 8-row-group panels
# and 4-tile reflector steps run beside small matrices
MIX = [(2, 3), (3, 7), (6, 5), (8, 4), (4, 12), (10, 3), (1, 9), (12, 3), (16, 3)]


def test_batched_mix_equals_alone_and_oracle(B, eps):
    import torch

    ctx = B.Context(0)
    inputs, devs, taus = [], [], []
    for k, (lx, p) in enumerate(MIX):
        A0, m = dqmc.dqmc_matrix(lx, p, instance=k)
        inputs.append((A0, m["n"], p))
        devs.append(to_dev(A0))
        taus.append(torch.zeros(m["n"] * p, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    descs = [(d.data_ptr(), t.data_ptr(), n, p, n * p) for d, t, (_, n, p) in zip(devs, taus, inputs)]
    info = B.invert_batched(ctx, descs)
    assert info == [0] * len(MIX)
    mixed = [from_dev(d, n * p).copy() for d, (_, n, p) in zip(devs, inputs)]
    O = oracle.Oracle("auto", threads=1)
    for k, (A0, n, p) in enumerate(inputs):
        print(f"batched case {k}: n={n} p={p}", flush=True)
        d = to_dev(A0)
        t = torch.zeros(n * p, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)]) == [0]
        alone = from_dev(d, n * p)
        assert np.array_equal(alone, mixed[k]), f"case {k} differs between alone and mix"
        G = A0.copy(order="F")
        O.invert(n, p, G)
        assert rel(alone, G) <= 1e-10
        assert oracle.residual(A0, alone) <= 100 * n * p * eps
    ctx.close()


def test_many_slab_reflector_launch_same_bits(B, eps):
    """24 matrices of n = 144 in one plan: each BSOF reflector launch holds
    600 slabs of m = 288 rows (more workgroups than CUs, matrices in both
    lanes of the plan), against 25 slabs alone.  The bits must agree."""
    import torch

    ctx = B.Context(0)
    cases = [dqmc.dqmc_matrix(12, 3, instance=k)[0] for k in range(24)]
    n, p = 144, 3
    devs = [to_dev(A0) for A0 in cases]
    taus = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for _ in cases]
    torch.cuda.synchronize()
    info = B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p) for d, t in zip(devs, taus)])
    assert info == [0] * len(cases)
    for k in (0, 23):
        print(f"many-slab case {k}: n={n} p={p}", flush=True)
        d = to_dev(cases[k])
        t = torch.zeros(n * p, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)]) == [0]
        alone = from_dev(d, n * p)
        assert np.array_equal(alone, from_dev(devs[k], n * p)), f"case {k} differs between alone and mix"
        assert oracle.residual(cases[k], alone) <= 100 * n * p * eps
    ctx.close()


def test_large_same_size_batch_same_bits(B, eps):
    """64 matrices of n = 144, p = 4 in one plan (two lanes, BSTRI steps of a
    few hundred GEMM tasks, multi-slab reflector launches): three of them
    get the bits they get alone, and check out against the oracle."""
    import torch

    ctx = B.Context(0)
    n, p = 144, 4
    cases = [dqmc.dqmc_matrix(12, p, instance=100 + k)[0] for k in range(64)]
    devs = [to_dev(A0) for A0 in cases]
    taus = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for _ in cases]
    torch.cuda.synchronize()
    info = B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p) for d, t in zip(devs, taus)])
    assert info == [0] * len(cases)
    for k in (0, 37, 63):
        print(f"large-batch case {k}: n={n} p={p}", flush=True)
        d = to_dev(cases[k])
        t = torch.zeros(n * p, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)]) == [0]
        alone = from_dev(d, n * p)
        assert np.array_equal(alone, from_dev(devs[k], n * p)), f"case {k} differs between alone and batch"
        assert oracle.residual(cases[k], alone) <= 100 * n * p * eps
    ctx.close()


def test_one_lane_same_bits(B):
    """A batch plan in one lane (every launch on the context's stream) and in
    two (the default) gives the same bits, eagerly and as a graph."""
    import torch

    inputs = [dqmc.dqmc_matrix(lx, p, instance=k)[0] for k, (lx, p) in enumerate(MIX)]
    sizes = [(lx * lx, p) for lx, p in MIX]
    out = {}
    for lanes in (1, 2):
        for graph in (False, True):
            ctx = B.Context(0)
            ctx.set_lanes(lanes)
            devs = [to_dev(A0) for A0 in inputs]
            taus = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for n, p in sizes]
            torch.cuda.synchronize()
            plan = B.Plan(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p) for d, t, (n, p) in zip(devs, taus, sizes)])
            plan.use_graph(graph)
            print(f"lanes={lanes} graph={graph}", flush=True)
            plan.execute()
            ctx.synchronize()
            assert plan.info() == [0] * len(MIX)
            plan.close()
            out[lanes, graph] = [from_dev(d, n * p).copy() for d, (n, p) in zip(devs, sizes)]
            ctx.close()
    for key in out:
        for k in range(len(MIX)):
            assert np.array_equal(out[key][k], out[1, False][k]), f"case {k} differs with lanes/graph {key}"


@pytest.mark.parametrize("lx,p,diag", [(12, 5, False), (8, 6, False), (12, 4, True)])
def test_single_matrix_plan_graph_and_lanes_same_bits(B, eps, lx, p, diag):
    """A single-matrix plan (which runs in one lane whatever the context's
    setting): the same bits with the context at one lane or two, eagerly and
    over graph replays with eager work between them, and the inverse (or its
    diagonal blocks) checks out against the oracle."""
    import torch

    A0, m = dqmc.dqmc_matrix(lx, p, instance=7)
    n = m["n"]
    phases = B.PHASE_ALL | (B.PHASE_DIAG_ONLY if diag else 0)
    out = {}
    for lanes in (1, 2):
        for graph in (False, True):
            ctx = B.Context(0)
            ctx.set_lanes(lanes)
            d = to_dev(A0)
            t = torch.zeros(n * p, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            plan = B.Plan(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)], phases=phases)
            plan.use_graph(graph)
            reps = 3 if graph else 1
            for r in range(reps):
                print(f"n={n} p={p} diag={diag} lanes={lanes} graph={graph} replay={r}", flush=True)
                d.copy_(to_dev(A0))
                t.zero_()
                torch.cuda.synchronize()
                plan.execute()
                ctx.synchronize()
                assert plan.info() == [0]
                got = from_dev(d, n * p).copy()
                if (lanes, graph) in out:
                    assert np.array_equal(got, out[lanes, graph])
                out[lanes, graph] = got
            plan.close()
            ctx.close()
    for key in out:
        assert np.array_equal(out[key], out[1, False]), f"lanes/graph {key} differ from one lane, eager"
    G = out[2, True]
    if diag:
        full = A0.copy(order="F")
        oracle.Oracle("auto", threads=1).invert(n, p, full)
        for l in range(p):
            blk = np.s_[l * n:(l + 1) * n, l * n:(l + 1) * n]
            assert rel(G[blk], full[blk]) <= 1e-10
    else:
        assert oracle.residual(A0, G) <= 100 * n * p * eps


def test_graph_replays_match_eager(B):
    import torch

    ctx = B.Context(0)
    cases = [(4, 6), (8, 5), (3, 9)]
    inputs = [dqmc.dqmc_matrix(lx, p, instance=20 + k)[0] for k, (lx, p) in enumerate(cases)]
    dims = [(lx * lx, p) for lx, p in cases]
    devs = [to_dev(A) for A in inputs]
    taus = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for n, p in dims]
    torch.cuda.synchronize()
    descs = [(d.data_ptr(), t.data_ptr(), n, p, n * p) for d, t, (n, p) in zip(devs, taus, dims)]
    eager = B.Plan(ctx, descs)
    eager.execute()
    ctx.synchronize()
    ref = [from_dev(d, n * p).copy() for d, (n, p) in zip(devs, dims)]
    graph = B.Plan(ctx, descs)
    graph.use_graph(True)
    for rep in range(4):
        for d, A in zip(devs, inputs):
            d.copy_(to_dev(A))
        torch.cuda.synchronize()
        graph.execute()
        ctx.synchronize()
        # eager work between replays
        x = torch.randn(256, 256, dtype=torch.float64, device="cuda")
        (x @ x).sum().item()
        for k, (d, (n, p)) in enumerate(zip(devs, dims)):
            assert np.array_equal(from_dev(d, n * p), ref[k]), f"replay {rep}, matrix {k}"
        assert graph.info() == [0, 0, 0]
    eager.close()
    graph.close()
    ctx.close()


def test_device_generator_matches_host(B):
    import torch

    ctx = B.Context(0)
    for lx, p, inst in [(2, 3, 0), (8, 16, 5), (6, 10, 9)]:
        A0, m = dqmc.dqmc_matrix(lx, p, instance=inst)
        n, N = m["n"], m["n"] * p
        ek = B.dqmc_expk(lx, lx)
        dek = torch.from_numpy(m["expk"].T.copy()).cuda()  # same expk as the host twin
        f = B.dqmc_fields(dqmc.DEFAULT_SEED, inst, p, n)
        df = torch.from_numpy(f.copy()).cuda()
        dA = torch.full((N, N), 3.0, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        lam = m["lam"]
        assert B.lib.bsofi_dqmc_build(ctx.h, dA.data_ptr(), N, n, p, dek.data_ptr(), df.data_ptr(), math.exp(lam),
                                      math.exp(-lam)) == 0
        ctx.synchronize()
        assert np.array_equal(from_dev(dA, N), A0)
        assert np.allclose(ek, m["expk"], atol=1e-14)
    ctx.close()


def test_device_residual_matches_host(B):
    import torch

    ctx = B.Context(0)
    lx, p = 4, 6
    A0, m = dqmc.dqmc_matrix(lx, p)
    n, N = m["n"], m["n"] * p
    dA = to_dev(A0)
    blocks = torch.empty(2 * p * n * n, dtype=torch.float64, device="cuda")
    tau = torch.zeros(N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    B.gather_blocks(ctx, dA.data_ptr(), N, n, p, blocks.data_ptr())
    assert B.invert_batched(ctx, [(dA.data_ptr(), tau.data_ptr(), n, p, N)]) == [0]
    G = from_dev(dA, N)
    r_dev = B.pcyclic_residual(ctx, blocks.data_ptr(), dA.data_ptr(), N, n, p)
    r_host = oracle.residual(A0, G)
    assert abs(r_dev - r_host) <= 0.5 * r_host + 1e-18
    G2 = G.copy()
    G2[5, 7] += 1e-7
    dA.copy_(to_dev(G2))
    torch.cuda.synchronize()
    r_dev2 = B.pcyclic_residual(ctx, blocks.data_ptr(), dA.data_ptr(), N, n, p)
    assert abs(r_dev2 - oracle.residual(A0, G2)) <= 1e-3 * r_dev2
    ctx.close()


def test_serving_alone_equals_mix(B, eps):
    import torch

    from bsofi_amd import serve

    def grab(srv, rid):
        g, t, n, p = srv.result(rid)
        N = n * p
        return B.copy_to_host(g, N * N).reshape(N, N).T.copy()  # column-major N x N

    flips = np.array([3, 17, 40, 41, 99], dtype=np.int32)
    # alone
    s1 = serve.Server(budget_bytes=8 << 30, slots=2, max_batch=8)
    c = s1.add_chain(4, 8, instance=3)
    rid = s1.submit(c, flips)
    got = None
    for _ in range(100000):
        s1.step()
        done = s1.poll()
        if done:
            assert done[0][3] == 0
            got = grab(s1, rid)
            s1.release(rid)
            break
    assert got is not None
    s1.close()
    # in a mix: other chains (different n, p) submitted before and after
    s2 = serve.Server(budget_bytes=8 << 30, slots=3, max_batch=4)
    others = [s2.add_chain(lx, p, instance=50 + k) for k, (lx, p) in enumerate([(6, 5), (3, 11), (8, 4)])]
    c2 = s2.add_chain(4, 8, instance=3)
    ids = [s2.submit(o, np.array([1, 2], dtype=np.int32)) for o in others]
    rid2 = s2.submit(c2, flips)
    ids += [s2.submit(o, np.array([5], dtype=np.int32)) for o in others]
    results = {}
    for _ in range(200000):
        s2.step()
        for r, lat, qw, info in s2.poll():
            assert info == 0
            if r == rid2:
                results[r] = grab(s2, r)
            s2.release(r)
            results.setdefault(r, None)
        if len(results) == len(ids) + 1:
            break
    assert rid2 in results and results[rid2] is not None
    assert np.array_equal(results[rid2], got)
    s2.close()
    # and it is the inverse of the chain's matrix after the flips
    A0, m = dqmc.dqmc_matrix(4, 8, instance=3)
    s = m["fields"].ravel().copy()
    s[flips] *= -1
    A = dqmc.assemble(m["expk"], s.reshape(8, 16), m["lam"], 8)
    assert oracle.residual(A, got) <= 100 * 128 * eps


def test_tuned_variants_match_oracle(B, eps):
    """Every (nb, scan_mti) pair the autotuner may pick gives the oracle's
    inverse: the nb = 16 panel instances, the 64-row scan tiles, the 16-row
    scan tiles below n = 256, the 32-row 4x4x4 scan (scan_mti 10) where
    4 divides n (n = 16, 64, 100, 144), and for n = 576 both sub-panel
    widths (8 and 16) under the 8-wave scan."""
    import torch

    from bsofi_amd import autotune

    O = oracle.Oracle("auto", threads=1)
    ctx = B.Context(0)
    for lx, p in [(4, 5), (8, 4), (10, 3), (12, 3), (24, 2)]:
        n = lx * lx
        A0, _ = dqmc.dqmc_matrix(lx, p, instance=70 + lx)
        G = A0.copy(order="F")
        O.invert(n, p, G)
        for nb, mti in autotune.candidates(n):
            print(f"tuned case n={n} p={p} nb={nb} scan_mti={mti}", flush=True)
            ctx.set_tuning(n, nb, mti)
            assert ctx.get_tuning(n) == (nb, mti)
            d = to_dev(A0)
            t = torch.zeros(n * p, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)]) == [0]
            got = from_dev(d, n * p)
            assert rel(got, G) <= 1e-10, (n, nb, mti)
            assert oracle.residual(A0, got) <= 100 * n * p * eps
    with pytest.raises(ValueError):
        ctx.set_tuning(300, 32, 0)  # 2n > 512 has no 32-wide panel
    with pytest.raises(ValueError):
        ctx.set_tuning(144, 0, 4)  # 64-row scan tiles need n <= 128
    ctx.clear_tuning()
    assert ctx.get_tuning(64) == (32, 12)  # the 4x4x4 scan, two workgroups per CU, for 4 | n <= 64
    assert ctx.get_tuning(144) == (32, 10)  # the three-buffer 4x4x4 scan for 4 | n in [128, 256]
    ctx.close()


def test_scan_forms_same_bits(B):
    """Every scan form the autotuner may pick for a sub-panel width gives the
    same bits: the 16-, 32- and 64-row tiles on the 16x16x4 MFMA and the
    32-row 4x4x4 forms (one or two workgroups per CU) run each element's k
    in the same order of 4-term steps, and the two FP64 MFMAs evaluate a
    step the same way."""
    import torch

    from bsofi_amd import autotune

    ctx = B.Context(0)
    for lx, p in [(8, 4), (10, 3), (12, 3), (16, 3)]:
        n = lx * lx
        A0, _ = dqmc.dqmc_matrix(lx, p, instance=70 + lx)
        res = {}
        for nb, mti in autotune.candidates(n):
            print(f"scan form case n={n} p={p} nb={nb} scan_mti={mti}", flush=True)
            ctx.set_tuning(n, nb, mti)
            d = to_dev(A0)
            t = torch.zeros(n * p, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)]) == [0]
            res[(nb, mti)] = from_dev(d, n * p)
        for (nb, mti), got in res.items():
            ref = res[min(k for k in res if k[0] == nb)]
            assert np.array_equal(got, ref), (n, nb, mti)
        ctx.clear_tuning()
    ctx.close()


def test_scan4_buffer_depths_same_bits(B):
    """The 4x4x4 scan with operands loaded one (scan_mti 9) or two (10)
    k-steps ahead, and the latter with two workgroups per CU (12, n <= 64),
    run the same MFMAs in the same order: the same bits, full and diagonal-only, for block sizes whose 32-row panels
    straddle block rows (n = 16, 24, 144) and whose k-step counts S = n / 4
    reach every tail of the deeper loops (S = 2 .. 64), in one batch beside
    a matrix on the 16x16x4 scan.  Forms 10 and 12 also take n = 4 mod 8
    (n = 12, 20, 28, 36, 100: a half column tile at the end of Q_k^T); every
    result of theirs matches the oracle at 1e-10 and has its diagonal-only
    blocks equal to the full inverse's."""
    import torch

    O = oracle.Oracle("auto", threads=1)
    cases = [dqmc.dqmc_matrix(4, 7, instance=120)[0], dqmc.dqmc_matrix(8, 4, instance=121)[0],
             dqmc.dqmc_matrix(12, 3, instance=122)[0], dqmc.dqmc_matrix(16, 2, instance=123)[0],
             dqmc.dqmc_matrix(10, 3, instance=124)[0], dqmc.random_pcyclic(8, 5, seed=125),
             dqmc.random_pcyclic(32, 4, seed=126), dqmc.random_pcyclic(24, 3, seed=127),
             dqmc.dqmc_matrix(6, 6, instance=128)[0], dqmc.random_pcyclic(20, 4, seed=129),
             dqmc.random_pcyclic(12, 5, seed=130), dqmc.random_pcyclic(40, 3, seed=131),
             dqmc.random_pcyclic(28, 3, seed=132)]
    dims = [(16, 7), (64, 4), (144, 3), (256, 2), (100, 3), (8, 5), (32, 4), (24, 3), (36, 6), (20, 4),
            (12, 5), (40, 3), (28, 3)]

    def takes(form, n):
        return n % 8 == 0 if form == 9 else (n % 4 == 0 and (form == 10 or n <= 64))

    diag_only = B.PHASE_ALL | B.PHASE_DIAG_ONLY
    out = {}
    for form in (9, 10, 12):
        for phases in (B.PHASE_ALL, diag_only):
            ctx = B.Context(0)
            devs, taus = [], []
            for A0, (n, p) in zip(cases, dims):
                assert A0.shape == (n * p, n * p)
                if takes(form, n):
                    ctx.set_tuning(n, 0, form)
                    assert ctx.get_tuning(n)[1] == form
                devs.append(to_dev(A0))
                taus.append(torch.zeros(n * p, dtype=torch.float64, device="cuda"))
            torch.cuda.synchronize()
            print(f"scan form {form} phases {phases}: " + ", ".join(f"n={n} p={p}" for n, p in dims), flush=True)
            assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)
                                          for d, t, (n, p) in zip(devs, taus, dims)],
                                    phases=phases) == [0] * len(dims)
            out[(form, phases)] = [from_dev(d, n * p) for d, (n, p) in zip(devs, dims)]
            ctx.close()
    for phases in (B.PHASE_ALL, diag_only):
        for i, (n, p) in enumerate(dims):
            for f1, f2 in ((9, 10), (10, 12)):
                if takes(f1, n) and takes(f2, n):
                    assert np.array_equal(out[(f1, phases)][i], out[(f2, phases)][i]), (n, p, f1, f2, phases)
    for form in (10, 12):
        for A0, (n, p), G, D in zip(cases, dims, out[(form, B.PHASE_ALL)], out[(form, diag_only)]):
            if not takes(form, n):
                continue
            R = A0.copy(order="F")
            O.invert(n, p, R)
            assert rel(G, R) <= 1e-10, (form, n, p)
            for l in range(p):
                s = slice(l * n, (l + 1) * n)
                assert np.array_equal(D[s, s], G[s, s]), (form, n, p, l)


def test_det_matches_slogdet(B):
    """sign(det A) and log|det A| from the device factorization: against
    numpy's LU-based slogdet, against the reference's tau/diag(R) loop
    (which yields (-1)^N sign(det A)), through bsofi_det on a BSOF-only
    plan, and the same bits in a batch as alone."""
    import torch

    cases = [dqmc.dqmc_matrix(4, 6, instance=31)[0], dqmc.dqmc_matrix(3, 5, instance=32)[0],
             dqmc.random_pcyclic(7, 4, seed=5, identity_diag=False), dqmc.random_pcyclic(12, 3, seed=6)]
    dims = [(16, 6), (9, 5), (7, 4), (12, 3)]
    O = oracle.Oracle("auto", threads=1)
    ctx = B.Context(0)
    devs = [to_dev(A) for A in cases]
    taus = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for n, p in dims]
    torch.cuda.synchronize()
    descs = [(d.data_ptr(), t.data_ptr(), n, p, n * p) for d, t, (n, p) in zip(devs, taus, dims)]
    plan = B.Plan(ctx, descs)
    plan.execute()
    batch = plan.det()
    plan.close()
    for k, (A0, (n, p)) in enumerate(zip(cases, dims)):
        N = n * p
        print(f"det case {k}: n={n} p={p}", flush=True)
        s_ref, l_ref = np.linalg.slogdet(A0)
        sg, lg = batch[k]
        assert sg == s_ref
        assert abs(lg - l_ref) <= 1e-10 * max(1.0, abs(l_ref))
        Ao = A0.copy(order="F")
        tau_o = np.zeros(N)
        assert O.bsof(n, p, Ao, tau_o) == 0
        assert sg == (-1) ** N * oracle.det_sign(tau_o, Ao)
        # alone, BSOF only, then the standalone device call on the factor
        d = to_dev(A0)
        t = torch.zeros(N, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        one = B.Plan(ctx, [(d.data_ptr(), t.data_ptr(), n, p, N)], phases=B.PHASE_BSOF)
        one.execute()
        assert one.det()[0] == batch[k]
        one.close()
        assert B.det(ctx, d.data_ptr(), N, n, p, t.data_ptr()) == batch[k]
    ctx.close()


def test_diag_only_blocks_equal_full_inverse(B):
    """BSOFI_PHASE_DIAG_ONLY: the diagonal blocks G_ll equal, bit for bit,
    those of the full inverse -- for block sizes whose scan tiles straddle
    block rows (n = 9, 36, 100), every scan form (n = 144 under the 32-row
    4x4x4 scan), and in a batch."""
    import torch

    cases = [(3, 5), (6, 4), (10, 3), (4, 7), (8, 3), (12, 3)]
    ctx = B.Context(0)
    for mti in (0, 1, 2, 4, 9, 10, 12):
        mats = []
        for k, (lx, p) in enumerate(cases):
            n = lx * lx
            if mti:
                try:
                    ctx.set_tuning(n, 0, mti)
                except ValueError:
                    ctx.set_tuning(n, 0, 0)
            A0, _ = dqmc.dqmc_matrix(lx, p, instance=90 + k)
            mats.append((A0, n, p))
        full = [to_dev(A0) for A0, _, _ in mats]
        diag = [to_dev(A0) for A0, _, _ in mats]
        tf = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for _, n, p in mats]
        td = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for _, n, p in mats]
        torch.cuda.synchronize()
        assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)
                                      for d, t, (_, n, p) in zip(full, tf, mats)]) == [0] * len(mats)
        assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)
                                      for d, t, (_, n, p) in zip(diag, td, mats)],
                                phases=B.PHASE_ALL | B.PHASE_DIAG_ONLY) == [0] * len(mats)
        for k, ((A0, n, p), f, d) in enumerate(zip(mats, full, diag)):
            print(f"diag-only case n={n} p={p} scan_mti={mti or 'default'}", flush=True)
            G = from_dev(f, n * p)
            D = from_dev(d, n * p)
            for l in range(p):
                s = slice(l * n, (l + 1) * n)
                assert np.array_equal(D[s, s], G[s, s]), (k, l)
        ctx.clear_tuning()
    ctx.close()


def test_fp64_rate_calibration(B):
    """bsofi_fp64_rate returns a plausible positive rate for every kind
    (MI355X datasheet FP64 peak 78.6 TFLOP/s; a reading above twice that
    would mean a miscounted kernel)."""
    ctx = B.Context(0)
    for kind in range(3):
        r = B.fp64_rate(ctx, kind)
        print(f"fp64 rate kind {kind}: {r:.1f} TFLOP/s", flush=True)
        assert 1.0 < r < 160.0
    with pytest.raises(RuntimeError):
        B.fp64_rate(ctx, 3)
    ctx.close()


def _run_plan(B, inputs, sizes, schedule, graph=False, reps=1, wait=None, phases=None):
    """Invert `inputs` (host arrays) in one plan on a fresh context with the
    given BSOF schedule; returns (results, info of the last execution)."""
    import torch

    ctx = B.Context(0)
    ctx.set_schedule(schedule)
    if wait is not None:
        ctx.set_wait_limit(wait)
    devs = [to_dev(A0) for A0 in inputs]
    taus = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for n, p in sizes]
    torch.cuda.synchronize()
    kw = {} if phases is None else {"phases": phases}
    plan = B.Plan(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p) for d, t, (n, p) in zip(devs, taus, sizes)], **kw)
    plan.use_graph(graph)
    out, info = None, None
    for r in range(reps):
        for d, A0 in zip(devs, inputs):
            d.copy_(to_dev(A0))
        torch.cuda.synchronize()
        plan.execute()
        ctx.synchronize()
        info = plan.info()
        got = [from_dev(d, n * p).copy() for d, (n, p) in zip(devs, sizes)]
        if out is not None:
            for a, b in zip(got, out):
                assert np.array_equal(a, b), f"replay {r} differs"
        out = got
    plan.close()
    ctx.close()
    return out, info


@pytest.mark.parametrize("case", ["single_n256", "single_small", "mix", "graph"])
def test_device_schedule_same_bits(B, eps, case):
    """BSOF as one device-scheduled launch (workgroups take sub-panel QRs and
    reflector slabs from a device queue, waiting on column counters) gives
    the bits of the launch-per-step schedule: one n = 256 matrix, small
    single matrices (n = 1 .. 144), the mixed batch, and graph replays."""
    if case == "single_n256":
        specs = [[(16, 8)]]
    elif case == "single_small":
        specs = [[(1, 4)], [(3, 5)], [(6, 3)], [(8, 6)], [(12, 4)]]
    elif case == "mix":
        specs = [[(lx, p) for lx, p in MIX]]
    else:
        specs = [[(4, 6), (8, 5), (16, 2)]]
    for sp in specs:
        inputs = [dqmc.dqmc_matrix(lx, p, instance=60 + k)[0] for k, (lx, p) in enumerate(sp)]
        sizes = [(lx * lx, p) for lx, p in sp]
        print(f"schedule case {case}: {sizes}", flush=True)
        graph = case == "graph"
        ref, info0 = _run_plan(B, inputs, sizes, 0, graph=graph, reps=3 if graph else 1)
        got, info1 = _run_plan(B, inputs, sizes, 1, graph=graph, reps=3 if graph else 1)
        assert info0 == [0] * len(sp) and info1 == [0] * len(sp)
        for k, (a, b) in enumerate(zip(got, ref)):
            assert np.array_equal(a, b), f"matrix {k} {sizes[k]}: device schedule differs"
            assert oracle.residual(inputs[k], a) <= 100 * sizes[k][0] * sizes[k][1] * eps


@pytest.mark.parametrize("scale", [1.0, 1e-300, 1e160])
def test_cyclic_panel_same_bits_as_barrier_form(B, scale):
    """The launch-per-step BSOF runs the cyclic sub-panel QR (no barrier per
    column), the device-scheduled sweep the barrier form: the same FMA chains
    and reduction tree, so the same A and tau bit for bit -- on the fast
    dlarfg path (scale 1) and on LAPACK's scaled path (1e-300, 1e160, where
    the whole panel runs again with the scaled norm).  Block sizes 16, 100
    (its last sub-panels have 4 and 8 columns: waves without a column) and
    256 (512-row panels), alone and in one plan."""
    import torch

    sp = [(4, 6), (10, 3), (16, 3)]
    inputs = [dqmc.dqmc_matrix(lx, p, instance=90 + k)[0] * scale for k, (lx, p) in enumerate(sp)]
    sizes = [(lx * lx, p) for lx, p in sp]

    def bsof(schedule, idx):
        ctx = B.Context(0)
        ctx.set_schedule(schedule)
        devs = [to_dev(inputs[k]) for k in idx]
        taus = [torch.zeros(sizes[k][0] * sizes[k][1], dtype=torch.float64, device="cuda") for k in idx]
        torch.cuda.synchronize()
        plan = B.Plan(ctx, [(d.data_ptr(), t.data_ptr(), sizes[k][0], sizes[k][1], sizes[k][0] * sizes[k][1])
                            for d, t, k in zip(devs, taus, idx)], phases=B.PHASE_BSOF)
        plan.execute()
        ctx.synchronize()
        assert plan.info() == [0] * len(idx)
        out = [(from_dev(d, sizes[k][0] * sizes[k][1]).copy(), t.cpu().numpy()) for d, t, k in zip(devs, taus, idx)]
        plan.close()
        ctx.close()
        return out

    print(f"cyclic vs barrier panel, scale {scale:g}: {sizes}", flush=True)
    ref = bsof(1, [0, 1, 2])
    for idx in ([0, 1, 2], [0], [1], [2]):
        got = bsof(0, idx)
        for (a, t), k in zip(got, idx):
            assert np.all(np.isfinite(a)) and np.all(np.isfinite(t))
            assert np.array_equal(a, ref[k][0]), f"{sizes[k]} scale {scale:g}: A differs"
            assert np.array_equal(t, ref[k][1]), f"{sizes[k]} scale {scale:g}: tau differs"


def test_device_schedule_wait_limit_reports_error(B):
    """A wait limit of zero makes the first unmet dependency stop the
    device-scheduled sweep: every workgroup leaves, every matrix of the plan
    reports BSOFI_ERR_SCHEDULE, and the context runs the next plan normally
    with the limit restored."""
    sp = [(16, 4), (8, 6)]
    inputs = [dqmc.dqmc_matrix(lx, p, instance=80 + k)[0] for k, (lx, p) in enumerate(sp)]
    sizes = [(lx * lx, p) for lx, p in sp]
    _, info = _run_plan(B, inputs, sizes, 1, wait=0.0, phases=B.PHASE_BSOF)
    assert info == [B.BSOFI_ERR_SCHEDULE] * len(sp)
    ref, info0 = _run_plan(B, inputs, sizes, 0)
    got, info1 = _run_plan(B, inputs, sizes, 1, wait=4.0)
    assert info0 == [0, 0] and info1 == [0, 0]
    for a, b in zip(got, ref):
        assert np.array_equal(a, b)


def test_tuning_knobs_same_bits(B, eps):
    """The BSTRI GEMM tile (32, 64, 128 or the launch-size rule) and the
    reflector slabs per workgroup (1, 2, 4, 8 or the rule) change only speed:
    every value gives the default's bits, for a batch of mixed block sizes
    whose steps span several tile and slab choices.  For n = 256 the rule
    runs the row sweep as column-strip chains (sweep_kernel) and a forced
    tile runs it as GEMM rounds: the same bits, over chains of 1 .. 8."""
    sp = [(8, 8), (8, 6), (12, 4), (16, 3), (6, 9), (4, 12), (16, 9)]
    inputs = [dqmc.dqmc_matrix(lx, p, instance=300 + k)[0] for k, (lx, p) in enumerate(sp)]
    sizes = [(lx * lx, p) for lx, p in sp]
    ref, info = _run_plan(B, inputs, sizes, 0)
    assert info == [0] * len(sp)
    for k, (A0, (n, p)) in enumerate(zip(inputs, sizes)):
        assert oracle.residual(A0, ref[k]) <= 100 * n * p * eps
    import torch

    for knob in [("gemm_tile", t) for t in (32, 64, 128)] + [("refl_spw", s) for s in (1, 2, 4, 8)]:
        print(f"knob {knob}", flush=True)
        ctx = B.Context(0)
        for n, _ in sizes:
            ctx.set_tuning(n, **{knob[0]: knob[1]})
            assert ctx.get_tuning_ex(n)[2 if knob[0] == "gemm_tile" else 3] == knob[1]
        devs = [to_dev(A0) for A0 in inputs]
        taus = [torch.zeros(n * p, dtype=torch.float64, device="cuda") for n, p in sizes]
        torch.cuda.synchronize()
        assert B.invert_batched(ctx, [(d.data_ptr(), t.data_ptr(), n, p, n * p)
                                      for d, t, (n, p) in zip(devs, taus, sizes)]) == [0] * len(sp)
        for k, (d, (n, p)) in enumerate(zip(devs, sizes)):
            assert np.array_equal(from_dev(d, n * p), ref[k]), f"{knob}: matrix {k} differs"
        with pytest.raises(ValueError):
            ctx.set_tuning(64, gemm_tile=48)
        ctx.close()

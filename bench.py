"""BSOF/I benchmark on one MI355X.

Default workload (the configuration the headline metric is quoted on):
plain single-matrix BSOF/I of a 16x16-lattice Hubbard-DQMC p-cyclic matrix,
n = 256, p = 64 (N = 16384, beta = 8), FP64, with A already resident in HBM
when the timed region starts.  One step = one full inversion
(BSOF -> BSTRI -> BSOI) of one matrix; each timed step inverts a different
seeded matrix (K device buffers built by the device generator beforehand).

Printed as one JSON line (rank 0):
* value       whole-job inversions/s over the K timed steps;
* p50/p99     per-inversion latency from HIP events on the library stream;
* roofline    the dominant kernel (BSOI row scan, ~82% of the flops):
              FLOPS_BSOI(n,p) (include/bsofi_flops.h:27-28) per launch over
              its average launch duration (HIP events around every scan
              launch in the timed region), against the FP64 MFMA peak;
* cpu_baseline the reference's own CPU path (oracle/_ref: src/bsofi_cpu.c
              built against MKL) on the host cores, one full inversion of the
              same matrix (instance 0) as the bounded sample;
* batched     (--batched) the 512 x (n=64, p=32) uniform batch in one plan,
              captured in a HIP graph, against the plain path on the same
              matrix size.
Every GPU result is checked: ||A G - I||_F / (||A||_F ||G||_F) <= 100 N eps,
computed on the device from the 2p non-zero blocks of A.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "p-cyclic inversions/s at resid≤100Nε (batched vs plain, 1 GPU); p50/p99 ms"
FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (matrix and vector)
EPS = float(np.finfo(np.float64).eps)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# BSOF schedule of every context the benchmark creates (bsofi_ctx_set_schedule):
# 0 = one launch per sweep step (the library default), 1 = one device-scheduled launch per plan
SCHEDULE = 0


def new_ctx(B):
    ctx = B.Context(0)
    ctx.set_schedule(SCHEDULE)
    return ctx


def fp64_rates(B):
    """This GPU's FP64 rates for back-to-back instructions of one kind on
    every CU (bsofi_fp64_rate): the ceilings the kernels actually face."""
    ctx = B.Context(0)
    try:
        return {k: round(B.fp64_rate(ctx, i), 1) for i, k in enumerate(("mfma_16x16x4", "mfma_4x4x4", "vector_fma"))}
    finally:
        ctx.close()


def measured_traffic(n, p):
    """HBM bytes per scan launch from the committed rocprofv3 --pmc summary
    for this (n, p) (FETCH_SIZE + WRITE_SIZE, scaled by a known-size copy),
    or None when no such summary exists."""
    path = os.path.join(ROOT, "profiles", f"bsoi_scan_n{n}_p{p}_hbm_traffic.txt")
    try:
        with open(path) as f:
            for ln in f:
                if ln.startswith("traffic_bytes_per_launch:"):
                    return float(ln.split(":")[1])
    except OSError:
        pass
    return None


class DeviceSet:
    """K seeded DQMC matrices (and their 2p non-zero blocks) on the device."""

    def __init__(self, B, ctx, lx, p, count, seed, U=4.0, dtau=0.125, t=1.0, first_instance=0):
        import torch

        from bsofi_amd import dqmc

        self.n, self.p = lx * lx, p
        self.N = self.n * p
        self.lda = self.N
        self.count = count
        ek = B.dqmc_expk(lx, lx, t, dtau)
        lam = dqmc.hs_lambda(U, dtau)
        self.ep, self.em = math.exp(lam), math.exp(-lam)
        self.d_ek = torch.from_numpy(ek.T.copy()).cuda()  # column-major n x n
        self.A, self.tau, self.blocks, self.fields = [], [], [], []
        for k in range(count):
            f = B.dqmc_fields(seed, first_instance + k, p, self.n)
            df = torch.from_numpy(f.copy()).cuda()
            A = torch.empty((self.N, self.lda), dtype=torch.float64, device="cuda")
            self.A.append(A)
            self.fields.append(df)
            self.tau.append(torch.zeros(self.N, dtype=torch.float64, device="cuda"))
            self.blocks.append(torch.empty(2 * p * self.n * self.n, dtype=torch.float64, device="cuda"))
        self.B, self.ctx = B, ctx
        torch.cuda.synchronize()
        self.rebuild()

    def rebuild(self, idx=None):
        B = self.B
        for k in (range(self.count) if idx is None else idx):
            rc = B.lib.bsofi_dqmc_build(self.ctx.h, self.A[k].data_ptr(), self.lda, self.n, self.p,
                                        self.d_ek.data_ptr(), self.fields[k].data_ptr(), self.ep, self.em)
            assert rc == 0, rc
            B.gather_blocks(self.ctx, self.A[k].data_ptr(), self.lda, self.n, self.p, self.blocks[k].data_ptr())
        self.ctx.synchronize()

    def desc(self, k):
        return (self.A[k].data_ptr(), self.tau[k].data_ptr(), self.n, self.p, self.lda)

    def residual(self, k):
        return self.B.pcyclic_residual(self.ctx, self.blocks[k].data_ptr(), self.A[k].data_ptr(), self.lda, self.n,
                                       self.p)


def scan_kernel_name(n, scan_mti=0):
    """The BSOI scan kernel a plan on block size n runs (the library's
    default scan form unless scan_mti overrides it)."""
    by4 = n % 4 == 0 and n >= 8
    form = scan_mti or (10 if 128 <= n <= 256 and by4 else (12 if n <= 64 and by4 else (2 if n <= 256 else 1)))
    return ("bsoi_scan4_kernel (v_mfma_f64_4x4x4f64)" if form in (9, 10, 12)
            else "bsoi_scan_kernel (v_mfma_f64_16x16x4f64)")


def bench_plain(B, lx, p, steps, warmup, seed, nb=0, scan_mti=0, phases=None, families=False, graph=False):
    import torch

    phases = B.PHASE_ALL if phases is None else phases
    ctx = new_ctx(B)
    if nb or scan_mti:
        ctx.set_tuning(lx * lx, nb, scan_mti)
    stream = torch.cuda.ExternalStream(ctx.stream)
    nbuf = steps + 1
    log(f"[bench] building {nbuf} device matrices n={lx * lx} p={p}")
    ds = DeviceSet(B, ctx, lx, p, nbuf, seed)
    plans = [B.Plan(ctx, [ds.desc(k)], phases=phases) for k in range(nbuf)]
    log(f"[bench] plan: {plans[0].num_launches} launches per inversion")
    if graph:
        # each plan is captured into a HIP graph on its first execution: do
        # that before the timing, then rebuild the inputs it overwrote
        for pl in plans:
            pl.use_graph(True)
            pl.execute()
        ctx.synchronize()
        ds.rebuild()
    # warm-up on buffer `steps` (rebuilt between warm-up steps, outside timing)
    for w in range(warmup):
        plans[steps].execute()
        ctx.synchronize()
        ds.rebuild([steps])
        log(f"[bench] warmup {w + 1}/{warmup} done")
    for pl in plans[:steps]:
        pl.time_family(-1 if graph else "scan")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(steps + 1)]
    torch.cuda.synchronize()
    ctx.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        ev[0].record(stream)
        for k in range(steps):
            plans[k].execute()
            ev[k + 1].record(stream)
    ctx.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([ev[k].elapsed_time(ev[k + 1]) for k in range(steps)])
    scan_ms, scan_n = 0.0, 0
    for pl in plans[:steps]:
        ms, cnt = pl.family_time()
        scan_ms += ms
        scan_n += cnt
    infos = [pl.info()[0] for pl in plans[:steps]]
    # the residual needs the whole G; diagonal-only runs are checked by the
    # tests (bit-equal to the full inverse's diagonal blocks)
    res = [ds.residual(k) for k in range(steps)] if phases == B.PHASE_ALL else [0.0]
    # every kernel family timed on its own (HIP events around each of its
    # launches on the library's stream), one extra inversion each, outside
    # the timed region
    fam = {}
    if families:
        spare = plans[steps]
        for name in ("panel", "refl", "gemm", "trtri", "larft", "eye", "det", "flow", "scan"):
            ds.rebuild([steps])
            spare.time_family(name)
            spare.execute()
            ctx.synchronize()
            ms, cnt = spare.family_time()
            if cnt:
                fam[name] = (ms, cnt)
        spare.time_family(-1)
    for pl in plans:
        pl.close()
    out = dict(n=ds.n, p=p, N=ds.N, wall=wall, lat=lat, scan_ms=scan_ms / max(scan_n, 1), scan_launches=scan_n,
               infos=infos, resid=res, families=fam, nb=ctx.get_tuning(ds.n)[0])
    del ds
    ctx.close()
    torch.cuda.empty_cache()
    return out


def bench_batched(B, lx, p, count, reps, seed):
    """count matrices in one plan (HIP graph), reps executions (inputs
    rebuilt between executions, outside the timed interval)."""
    import torch

    ctx = new_ctx(B)
    stream = torch.cuda.ExternalStream(ctx.stream)
    log(f"[bench] batched: building {count} matrices n={lx * lx} p={p}")
    ds = DeviceSet(B, ctx, lx, p, count, seed, first_instance=1000)
    plan = B.Plan(ctx, [ds.desc(k) for k in range(count)])
    plan.use_graph(True)
    plan.execute()  # warm-up + graph instantiation
    ctx.synchronize()
    ds.rebuild()
    times = []
    for r in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            e0.record(stream)
            plan.execute()
            e1.record(stream)
        ctx.synchronize()
        torch.cuda.synchronize()
        times.append(e0.elapsed_time(e1))
        if r + 1 < reps:
            ds.rebuild()
    res = max(ds.residual(k) for k in range(0, count, max(1, count // 16)))
    infos = plan.info()
    plan.close()
    del ds
    ctx.close()
    torch.cuda.empty_cache()
    return dict(ms=float(np.median(times)), resid=res, info_bad=sum(1 for i in infos if i != 0))


def bench_uneven(B, count, seed, reps, tune_p, log_fn, save_tune=None):
    """The uneven mixed workload: `count` matrices with lattice side from
    {6, 8, 10, 12, 16}, p from U{10..80}, U from U[2, 6] (seeded draw).
    (1) the autotuner picks nb, scan form, GEMM tile and reflector slabs per
        workgroup per n, on trial batches of the workload's own matrices,
        then keeps each n's choice only where it beats the defaults in the
        whole plan;
    (2) all matrices in ONE plan on the tuned context, captured in a HIP
        graph, `reps` executions on rebuilt inputs;
    (3) the plain path on the same matrices: one single-matrix plan each on
        an untuned context, executed back to back.
    Every result of (2) is residual-checked on the device, and a sample of
    (3)."""
    import torch

    from bsofi_amd import autotune, flops, workload

    specs = workload.uneven_specs(count, seed)
    ctx = new_ctx(B)
    t0 = time.perf_counter()
    # trial batch per block size: every matrix of that size in the workload
    # (a smaller batch makes different launches and picks different forms)
    table = autotune.tune(B, ctx, sorted({sp[0] for sp in specs}), p=tune_p, batch=len(specs), reps=1,
                          log=log_fn, specs=specs)
    tune_s = time.perf_counter() - t0
    log_fn(f"[bench] uneven: building {count} matrices")
    mats = workload.DeviceMatrices(B, ctx, specs, seed=seed)
    # each tuned block size against the defaults, in the whole plan
    t0 = time.perf_counter()
    autotune.refine(B, ctx, table, mats.descs(), mats.rebuild, reps=3, log=log_fn)
    tune_s += time.perf_counter() - t0
    mats.rebuild()
    if save_tune:
        autotune.save(table, save_tune)
    tot_flops = sum(flops.bsofi(n, p) for (_, n, p, _, _) in mats.specs)
    stream = torch.cuda.ExternalStream(ctx.stream)
    plan = B.Plan(ctx, mats.descs())
    plan.use_graph(True)
    plan.execute()  # warm-up + graph instantiation
    ctx.synchronize()
    times = []
    for r in range(reps):
        mats.rebuild()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            e0.record(stream)
            plan.execute()
            e1.record(stream)
        ctx.synchronize()
        times.append(e0.elapsed_time(e1))
    infos = plan.info()
    res = [mats.residual(k) for k in range(len(mats))]
    bounds = [100 * n * p * EPS for (_, n, p, _, _) in mats.specs]
    plan.close()
    batched_ms = float(np.median(times))
    # plain path: one plan per matrix on an untuned context
    mats.rebuild()
    pctx = new_ctx(B)
    pstream = torch.cuda.ExternalStream(pctx.stream)
    plans = [B.Plan(pctx, [mats.desc(k)]) for k in range(len(mats))]
    plans[0].execute()  # warm-up of the kernels, then fresh input
    pctx.synchronize()
    mats.rebuild([0])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with torch.cuda.stream(pstream):
        e0.record(pstream)
        for pl in plans:
            pl.execute()
        e1.record(pstream)
    pctx.synchronize()
    plain_ms = e0.elapsed_time(e1)
    pinfo = [pl.info()[0] for pl in plans]
    pres = [mats.residual(k) for k in range(0, len(mats), 16)]
    pbounds = [bounds[k] for k in range(0, len(mats), 16)]
    for pl in plans:
        pl.close()
    pctx.close()
    mats.release()
    ctx.close()
    torch.cuda.empty_cache()
    ns = [n for (_, n, _, _, _) in mats.specs]
    ps = [p for (_, _, p, _, _) in mats.specs]
    return dict(batched_ms=batched_ms, plain_ms=plain_ms, tot_flops=tot_flops, tune=table, tune_s=tune_s,
                resid_ok=all(r <= b for r, b in zip(res, bounds)) and all(i == 0 for i in infos),
                plain_resid_ok=all(r <= b for r, b in zip(pres, pbounds)) and all(i == 0 for i in pinfo),
                resid_max_rel=max(r / b for r, b in zip(res, bounds)), n_range=(min(ns), max(ns)),
                p_range=(min(ps), max(ps)))


def bench_serving(B, seed, nchains, budget_gb, slots, max_batch, cap_requests, requests, log_fn):
    """Serving: `nchains` DQMC chains (n in {64, 144, 256}, p in {20, 40, 80}),
    each request flips 1-5 % of its chain's HS fields.  (1) capacity: a
    closed burst of cap_requests requests, completions / wall; (2) Poisson
    arrivals at 80 % of that capacity, p50 / p99 latency (arrival to
    completion) and throughput.  Admission by the HBM budget."""
    from bsofi_amd import serve

    srv = serve.Server(budget_bytes=int(budget_gb * (1 << 30)), slots=slots, max_batch=max_batch)
    dims = serve.c5_chains(nchains, seed)
    chains = [srv.add_chain(lx, p, instance=7000 + k) for k, (lx, p) in enumerate(dims)]
    # warm-up: one request per distinct shape
    seen = {}
    for cid, (lx, p) in zip(chains, dims):
        seen.setdefault((lx, p), cid)
    serve.run(srv, list(seen.values()), len(seen), 1e9, seed=seed + 1, max_seconds=120)
    lat_c, _, info_c, wall_c = serve.run(srv, chains, cap_requests, 1e9, seed=seed + 2, max_seconds=120)
    capacity = len(lat_c) / wall_c
    rate = 0.8 * capacity
    log_fn(f"[bench] serving: capacity {capacity:.1f} req/s, Poisson at {rate:.1f} req/s")
    lat, waits, infos, wall = serve.run(srv, chains, requests, rate, seed=seed + 3, max_seconds=120)
    srv.close()
    return dict(capacity=capacity, rate=rate, done=len(lat), wall=wall, p50=float(np.percentile(lat, 50)) * 1e3,
                p99=float(np.percentile(lat, 99)) * 1e3, qwait_p50=float(np.percentile(waits, 50)) * 1e3,
                info_bad=sum(1 for v in list(infos.values()) + list(info_c.values()) if v != 0),
                shapes=sorted(set(dims)))


def bench_e2e(B, lx, p, seed):
    """C2 through the host-array handle API (hybridxbsoftrih_ + hybridxbsoih_):
    the PCIe-inclusive time of one inversion (A copied in, G copied out)."""
    from bsofi_amd import dqmc

    A0, meta = dqmc.dqmc_matrix(lx, p, seed=seed, instance=0)
    n, N = meta["n"], meta["n"] * p
    A = np.asfortranarray(A0)
    tau = np.zeros(N)
    H = B.HandleAPI(n, p, N)
    work = np.zeros(max(1, H.lwork()))
    t0 = time.perf_counter()
    i1 = H.bsoftri(A, tau, work)
    i2 = H.bsoi(A, tau, work)
    dt = time.perf_counter() - t0
    H.finalize()
    return dict(seconds=dt, info=(i1, i2))


def host_info():
    """The host the CPU baseline runs on: the cores this process may use
    (affinity), the cgroup CPU quota, physical cores among them, and the CPU
    model string as /proc/cpuinfo reports it."""
    aff = sorted(os.sched_getaffinity(0))
    quota = None
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()
        quota = None if q == "max" else float(q) / float(per)
    except (OSError, ValueError):
        pass
    model, phys = "unknown", set()
    try:
        for blk in open("/proc/cpuinfo").read().split("\n\n"):
            d = {}
            for ln in blk.splitlines():
                if ":" in ln:
                    k, v = ln.split(":", 1)
                    d[k.strip()] = v.strip()
            if "processor" in d and int(d["processor"]) in aff:
                model = d.get("model name", model)
                phys.add((d.get("physical id"), d.get("core id")))
    except OSError:
        pass
    return dict(cores=len(aff), cgroup_quota_cpus=quota, physical_cores=len(phys) or len(aff), model=model,
                nproc=os.cpu_count())


def cpu_impl(threads):
    import oracle

    if oracle.ref_available():
        return oracle.Reference(threads=threads), "reference"
    return oracle.Oracle("auto", threads=threads), "port"


def cpu_invert(impl, kind, n, p, A):
    if kind == "reference":
        return impl.invert(n, p, A)[1]
    return impl.invert(n, p, A)[1]


def cpu_baseline(lx, p, seed, threads, reps=5):
    """The reference's CPU path (oracle/_ref) on the same matrix (instance 0),
    BLAS threads = `threads`: one warm-up inversion, then the median of
    `reps` (each on a fresh copy)."""
    from bsofi_amd import dqmc

    impl, kind = cpu_impl(threads)
    A0, meta = dqmc.dqmc_matrix(lx, p, seed=seed, instance=0)
    n = meta["n"]
    times, info = [], 0
    for r in range(reps + 1):
        A = A0.copy(order="F")
        t0 = time.perf_counter()
        info = cpu_invert(impl, kind, n, p, A)
        if r > 0:
            times.append(time.perf_counter() - t0)
    return dict(seconds=float(np.median(times)), kind=kind, info=info, reps=reps)


def cpu_batch_baseline(specs, threads, seed):
    """A batched workload on the host, both ways the reference's CPU path can
    run it: threaded (BLAS threads = all cores, one matrix after another) and
    throughput (one matrix per core at a time, one BLAS thread each, `threads`
    host threads calling the reference concurrently).  specs: (lx, p, U,
    instance).  Inputs are built before the timing; returns inversions/s of
    each mode over the sample."""
    from concurrent.futures import ThreadPoolExecutor

    from bsofi_amd import dqmc

    mats = [(dqmc.dqmc_matrix(lx, p, U=U, seed=seed, instance=inst)[0], lx * lx, p) for lx, p, U, inst in specs]
    out = {}
    for mode in ("threaded", "throughput"):
        impl, kind = cpu_impl(threads if mode == "threaded" else 1)
        work = [A.copy(order="F") for A, _, _ in mats]
        t0 = time.perf_counter()
        if mode == "threaded":
            infos = [cpu_invert(impl, kind, n, p, A) for A, (_, n, p) in zip(work, mats)]
        else:
            with ThreadPoolExecutor(max_workers=threads) as ex:
                infos = list(ex.map(lambda a: cpu_invert(impl, kind, a[1][1], a[1][2], a[0]), zip(work, mats)))
        dt = time.perf_counter() - t0
        out[mode] = dict(inversions_per_s=len(mats) / dt, seconds=dt, info_nonzero=sum(1 for i in infos if i))
        del work
    best = max(out, key=lambda m: out[m]["inversions_per_s"])
    return dict(out, best=best, kind=kind, sample=len(mats))


def family_roofline(r, flops):
    """Per-inversion time, launches and achieved FP64 rate of each kernel
    family of the plain run (flops.families: algorithmic flops of the panel
    QRs, the reflector applications, BSTRI and the BSOI scan)."""
    fl = flops.families(r["n"], r["p"], r.get("nb", 32))
    out = {}
    for name, (ms, cnt) in r.get("families", {}).items():
        e = {"ms_per_inversion": round(ms, 4), "launches": cnt}
        if name in fl and ms > 0:
            e["flops"] = fl[name]
            e["tflops"] = round(fl[name] / (ms * 1e-3) / 1e12, 3)
            e["frac"] = round(fl[name] / (ms * 1e-3) / (FP64_PEAK_TFLOPS * 1e12), 4)
        out[name] = e
    return out


def cpu_batch_line(cb, sample, threads):
    best = cb[cb["best"]]
    return {
        "value": round(best["inversions_per_s"], 3),
        "unit": "inversions/s",
        "mode": cb["best"],
        "threaded_inversions_per_s": round(cb["threaded"]["inversions_per_s"], 3),
        "throughput_inversions_per_s": round(cb["throughput"]["inversions_per_s"], 3),
        "cores": threads,
        "kind": cb["kind"],
        "sample": f"{sample}: threaded = MKL {threads} threads, one matrix after another; throughput = "
                  f"{threads} host threads, one matrix each at a time, MKL 1 thread",
        "info_nonzero": cb["threaded"]["info_nonzero"] + cb["throughput"]["info_nonzero"],
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--lx", type=int, default=16, help="lattice side (n = lx^2)")
    ap.add_argument("--p", type=int, default=64)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--batched", type=int, default=1, help="also run the 512 x (n=64,p=32) batch")
    ap.add_argument("--uneven", type=int, default=256, help="matrices in the uneven batch (0 = skip)")
    ap.add_argument("--serve", type=int, default=1, help="run the serving loop")
    ap.add_argument("--serve-requests", type=int, default=400)
    ap.add_argument("--e2e", type=int, default=1, help="also time one inversion through the host-array H-API")
    ap.add_argument("--diag", type=int, default=1, help="also time the diagonal-blocks-only mode on the same size")
    ap.add_argument("--cpu-baseline", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--schedule", type=int, default=0,
                    help="BSOF schedule: 1 = one device-scheduled launch per plan, 0 = one launch per sweep step")
    ap.add_argument("--save-tune", default=None, help="write the uneven run's autotuning table (JSON) here")
    ap.add_argument("--graph", type=int, default=0, help="replay each plain plan as a HIP graph (captured beforehand)")
    ap.add_argument("--nb", type=int, default=0, help="sub-panel width for the plain run (0 = default)")
    ap.add_argument("--scan-mti", type=int, default=0, help="scan tile height / 16 for the plain run (0 = default)")
    ap.add_argument("--calibrate-copy", type=float, default=0.0,
                    help="only run y = -x over this many GiB of FP64 (3 launches): the known-size kernel that "
                         "scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: one GPU only; multi-GPU runs are out of scope here (the path would shard by "
              "row panels of G, each GPU forming every Q_k and scanning its own rows).", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy > 0:
        import torch

        count = int(args.calibrate_copy * (1 << 30)) // 8
        x = torch.ones(count, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            torch.neg(x, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes_read": count * 8, "calibrate_copy_bytes_written": count * 8,
                          "launches": 3}), flush=True)
        return

    import bsofi_amd as B
    from bsofi_amd import flops

    global SCHEDULE
    SCHEDULE = args.schedule
    if B.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    r = bench_plain(B, args.lx, args.p, args.steps, args.warmup, args.seed, args.nb, args.scan_mti, families=True,
                    graph=bool(args.graph))
    n, p, N = r["n"], r["p"], r["N"]
    value = args.steps / r["wall"]
    tot_flops = flops.bsofi(n, p)
    bsoi_flops = flops.bsoi(n, p)
    scan_s = r["scan_ms"] * 1e-3
    achieved = bsoi_flops / scan_s / 1e12 if scan_s > 0 else 0.0
    resid_max = max(r["resid"])
    bound = 100 * N * EPS
    line = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "inversions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * r["wall"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded Hubbard-DQMC p-cyclic matrices (U=4, dtau=0.125, t=1), one per step",
        "config": {"workload": f"plain single-matrix BSOF/I, {args.lx}x{args.lx} lattice n={n}, p={p}, N={N}, "
                               "device-resident A, full inverse G formed on the GPU",
                   "n": n, "p": p, "N": N, "batch": 1},
        "p50_ms": round(float(np.percentile(r["lat"], 50)), 3),
        "p99_ms": round(float(np.percentile(r["lat"], 99)), 3),
        "gflops": round(tot_flops / (r["wall"] / args.steps) / 1e9, 1),
        "frac_fp64_peak": round(tot_flops / (r["wall"] / args.steps) / (FP64_PEAK_TFLOPS * 1e12), 4),
        "resid_max": resid_max,
        "resid_bound": bound,
        "resid_ok": bool(resid_max <= bound and all(i == 0 for i in r["infos"])),
        "roofline": {
            "kernel": scan_kernel_name(n, args.scan_mti),
            "bound": "mfma",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": measured_traffic(n, p),
            "traffic_source": f"profiles/bsoi_scan_n{n}_p{p}_hbm_traffic.txt (rocprofv3 --pmc FETCH_SIZE, WRITE_SIZE)",
            "achievable_fp64_tflops": fp64_rates(B),
            "flops_per_launch": bsoi_flops,
            "avg_launch_ms": round(r["scan_ms"], 4),
            "launches": r["scan_launches"],
            "families": family_roofline(r, flops),
        },
    }
    if args.batched:
        nb_lx, nb_p, count = 8, 32, 512
        b = bench_batched(B, nb_lx, nb_p, count, 3, args.seed)
        pl = bench_plain(B, nb_lx, nb_p, 5, 1, args.seed)
        plain_ips = 5 / pl["wall"]
        b_ips = count / (b["ms"] * 1e-3)
        line["batched"] = {
            "workload": f"{count} x (n={nb_lx * nb_lx}, p={nb_p}) in one plan, HIP graph",
            "inversions_per_s": round(b_ips, 1),
            "ms_per_batch": round(b["ms"], 3),
            "gflops": round(count * flops.bsofi(nb_lx * nb_lx, nb_p) / (b["ms"] * 1e-3) / 1e9, 1),
            "plain_same_size_inversions_per_s": round(plain_ips, 1),
            "speedup_vs_plain": round(b_ips / plain_ips, 2),
            "resid_max_sampled": b["resid"],
            "info_nonzero": b["info_bad"],
        }
    if args.diag:
        dg = bench_plain(B, args.lx, args.p, 3, 1, args.seed, phases=B.PHASE_ALL | B.PHASE_DIAG_ONLY)
        line["diag_only"] = {
            "workload": f"same n={n}, p={p} matrices, only the diagonal blocks G_ll written (DQMC equal-time "
                        "Green's functions), BSOFI_PHASE_DIAG_ONLY",
            "inversions_per_s": round(3 / dg["wall"], 3),
            "ms_per_step": round(1e3 * dg["wall"] / 3, 3),
            "scan_ms": round(dg["scan_ms"], 3),
        }
    if args.uneven:
        u = bench_uneven(B, args.uneven, args.seed, 2, 24, log, args.save_tune)
        b_ips = args.uneven / (u["batched_ms"] * 1e-3)
        p_ips = args.uneven / (u["plain_ms"] * 1e-3)
        line["uneven"] = {
            "workload": f"{args.uneven} matrices, lattice side in {{6,8,10,12,16}} (n={u['n_range'][0]}.."
                        f"{u['n_range'][1]}), p in U{{10..80}} (drawn {u['p_range'][0]}..{u['p_range'][1]}), "
                        "U in U[2,6]; one plan, HIP graph, autotuned nb / scan tiles",
            "inversions_per_s": round(b_ips, 2),
            "ms_per_batch": round(u["batched_ms"], 2),
            "gflops": round(u["tot_flops"] / (u["batched_ms"] * 1e-3) / 1e9, 1),
            "frac_fp64_peak": round(u["tot_flops"] / (u["batched_ms"] * 1e-3) / (FP64_PEAK_TFLOPS * 1e12), 4),
            "plain_inversions_per_s": round(p_ips, 2),
            "plain_ms_all": round(u["plain_ms"], 1),
            "speedup_vs_plain": round(b_ips / p_ips, 2),
            "resid_ok": u["resid_ok"],
            "plain_resid_ok_sampled": u["plain_resid_ok"],
            "resid_max_over_bound": u["resid_max_rel"],
            "autotune": {str(k): f"nb={v['nb']} scan_mti={v['scan_mti']} gemm_tile={v['gemm_tile']} "
                                 f"refl_spw={v['refl_spw']} ({v.get('kept', 'tuned')})" for k, v in u["tune"].items()},
            "autotune_s": round(u["tune_s"], 1),
        }
    if args.serve:
        sv = bench_serving(B, args.seed, 64, 200, 4, 16, max(64, args.serve_requests // 2), args.serve_requests, log)
        line["serving"] = {
            "workload": "64 DQMC chains, n in {64,144,256}, p in {20,40,80}; each request flips 1-5% of its "
                        "chain's HS fields; Poisson arrivals at 80% of the measured capacity; 200 GB HBM budget",
            "capacity_req_per_s": round(sv["capacity"], 2),
            "rate_req_per_s": round(sv["rate"], 2),
            "throughput_req_per_s": round(sv["done"] / sv["wall"], 2),
            "p50_ms": round(sv["p50"], 2),
            "p99_ms": round(sv["p99"], 2),
            "queue_wait_p50_ms": round(sv["qwait_p50"], 2),
            "requests": sv["done"],
            "info_nonzero": sv["info_bad"],
        }
    if args.e2e:
        e = bench_e2e(B, args.lx, args.p, args.seed)
        line["e2e_host_api"] = {
            "path": "hybridxinith_ (outside the timing), then hybridxbsoftrih_ + hybridxbsoih_ on host arrays: the "
                    "2p blocks of A in, V + X out, V + X in again, G out, through pinned double-buffered staging",
            "seconds": round(e["seconds"], 3),
            "inversions_per_s": round(1.0 / e["seconds"], 3),
            "info": list(e["info"]),
        }
    if args.cpu_baseline:
        host = host_info()
        threads = args.cpu_threads or host["cores"]
        log(f"[bench] cpu baseline: {threads} threads on {host}")
        c = cpu_baseline(args.lx, args.p, args.seed, threads)
        line["cpu_baseline"] = {
            "value": round(1.0 / c["seconds"], 5),
            "unit": "inversions/s",
            "cores": threads,
            "kind": c["kind"],
            "sample": f"median of {c['reps']} full inversions (after one warm-up) of the same n={n}, p={p} "
                      f"matrix (instance 0), MKL {threads} threads (every core this host exposes)",
            "host": host,
            "blas": "MKL (libmkl_rt)" if c["kind"] == "reference" else "the oracle's LAPACK back end",
            "gflops": round(tot_flops / c["seconds"] / 1e9, 2),
        }
        if args.batched:
            cb = cpu_batch_baseline([(8, 32, 4.0, 1000 + k) for k in range(64)], threads, args.seed)
            line["cpu_baseline"]["batched_uniform"] = cpu_batch_line(cb, "64 of the 512 (n=64, p=32) matrices", threads)
        if args.uneven:
            from bsofi_amd import workload

            cu = cpu_batch_baseline(list(workload.uneven_specs(args.uneven, args.seed)[:16]), threads, args.seed)
            line["cpu_baseline"]["uneven"] = cpu_batch_line(cu, f"the first 16 of the {args.uneven} uneven matrices",
                                                            threads)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Autotuner for the per-size launch choices of a context.

For every block size n it times, on the device, a small batch of seeded
DQMC matrices of that n under each supported pair

    nb       -- sub-panel width of BSOF / BSOI (16 or 32 where 2n <= 512,
                else 8 or 16): narrower panels shorten each column step of
                the panel QR but add reflector launches;
    scan_mti -- rows per workgroup of the BSOI scan, 16 * scan_mti (1, 2, 4):
                taller tiles reuse each Q_k fragment over more rows but
                leave fewer workgroups for small batches,

keeps the fastest (median of `reps` executions of one plan), and sets it on
the context with bsofi_ctx_set_tuning, so plans created afterwards use it.
The table can be saved as JSON and applied again without re-timing.
"""
from __future__ import annotations

import json
import time

import numpy as np

from .workload import DeviceMatrices


def candidates(n):
    nbs = [nb for nb, ok in ((32, 2 * n <= 512), (16, True), (8, True)) if ok][:2]
    # scan forms: 16 * m rows on the 16x16x4 MFMA with 4 waves (m = 1, 2, 4),
    # 8: 16 rows with 8 waves (n > 512), or 9: 32 rows on the 4x4x4 MFMA
    # (from n = 128: below, its one workgroup per CU starves real batches,
    # which the tuner's small trial batch does not show)
    mtis = [m for m, ok in ((1, n <= 512), (2, n <= 256), (4, n <= 128), (8, n > 512),
                            (9, 128 <= n <= 256 and n % 8 == 0)) if ok]
    return [(nb, m) for nb in nbs for m in mtis]


def time_plan(B, ctx, descs, reps, rebuild):
    """Median device time (ms) of one plan over `reps` executions, each on
    freshly rebuilt inputs (rebuild runs outside the timed interval)."""
    import torch

    stream = torch.cuda.ExternalStream(ctx.stream)
    plan = B.Plan(ctx, descs)
    try:
        plan.execute()  # warm-up
        ctx.synchronize()
        ts = []
        for _ in range(reps):
            rebuild()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            with torch.cuda.stream(stream):
                e0.record(stream)
                plan.execute()
                e1.record(stream)
            ctx.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts))
    finally:
        plan.close()


def tune(B, ctx, sides, p=24, batch=8, reps=3, log=None):
    """Tune every lattice side in `sides` (n = side^2).  Returns
    {n: {"nb": .., "scan_mti": .., "ms": .., "all": {"nb,mti": ms}}}."""
    table = {}
    for lx in sides:
        n = lx * lx
        mats = DeviceMatrices(B, ctx, [(lx, p, 4.0, 9000 + k) for k in range(batch)], keep_blocks=False)
        res = {}
        for nb, mti in candidates(n):
            ctx.set_tuning(n, nb, mti)
            t0 = time.perf_counter()
            res[(nb, mti)] = time_plan(B, ctx, mats.descs(), reps, mats.rebuild)
            mats.rebuild()
            if log:
                log(f"[tune] n={n} nb={nb} scan_mti={mti}: {res[(nb, mti)]:.3f} ms "
                    f"({time.perf_counter() - t0:.1f} s)")
        best = min(res, key=res.get)
        ctx.set_tuning(n, *best)
        table[n] = {"nb": best[0], "scan_mti": best[1], "ms": res[best],
                    "all": {f"{a},{b}": round(v, 4) for (a, b), v in res.items()}}
        mats.release()
    return table


def apply(ctx, table):
    for n, row in table.items():
        ctx.set_tuning(int(n), int(row["nb"]), int(row["scan_mti"]))


def save(table, path):
    with open(path, "w") as f:
        json.dump({str(k): v for k, v in table.items()}, f, indent=1)


def load(path):
    with open(path) as f:
        return {int(k): v for k, v in json.load(f).items()}

"""Autotuner for the per-size launch choices of a context.

For every block size n it times, on the device, a small batch of seeded
DQMC matrices of that n under each supported pair

    nb       -- sub-panel width of BSOF / BSOI (16 or 32 where 2n <= 512,
                else 8 or 16): narrower panels shorten each column step of
                the panel QR but add reflector launches;
    scan_mti -- rows per workgroup of the BSOI scan, 16 * scan_mti (1, 2, 4):
                taller tiles reuse each Q_k fragment over more rows but
                leave fewer workgroups for small batches,

keeps the fastest (median of `reps` executions of one plan), then does the
same for two knobs that change only speed -- the BSTRI GEMM tile edge
(32, 64, 128, or the launch-size rule) and the reflector applications'
16-column slabs per workgroup (1, 2, 4, 8, or the launch-size rule) -- and
sets the result on the context with bsofi_ctx_set_tuning_ex, so plans created
afterwards use it.  Trial batches can be drawn from the workload's own
(n, p, U) draw.  The table can be saved as JSON and applied again without
re-timing (replacing the reference's tune/hybrid/tune_hybrid.c, which timed
its CPU/GPU split constants).
"""
from __future__ import annotations

import json
import time

import numpy as np

from .workload import DeviceMatrices


def candidates(n):
    nbs = [nb for nb, ok in ((32, 2 * n <= 512), (16, True), (8, True)) if ok][:2]
    # scan forms: 16 * m rows on the 16x16x4 MFMA with 4 waves (m = 1, 2, 4),
    # 8: 16 rows with 8 waves (n > 512), 10: 32 rows on the 4x4x4 MFMA with
    # operands loaded two k-steps ahead (4 | n), 12: the same with two
    # workgroups per CU (n <= 64).  Form 9, the 4x4x4 scan with two buffers
    # (8 | n), gives the same bits as 10 and ran slower at every size
    # measured, so the tuner leaves it out.
    by4 = n % 4 == 0 and n >= 8
    mtis = [m for m, ok in ((1, n <= 512), (2, n <= 256), (4, n <= 128), (8, n > 512),
                            (10, n <= 256 and by4), (12, n <= 64 and by4)) if ok]
    return [(nb, m) for nb in nbs for m in mtis]


def time_plan(B, ctx, descs, reps, rebuild, stat="median"):
    """Median (stat="median") or shortest (stat="min") device time (ms) of
    one plan over `reps` executions, each on freshly rebuilt inputs (rebuild
    runs outside the timed interval)."""
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
        return float(np.min(ts) if stat == "min" else np.median(ts))
    finally:
        plan.close()


GEMM_TILES = (0, 32, 64, 128)
REFL_SPWS = (0, 1, 2, 4, 8)


def tune(B, ctx, sides, p=24, batch=8, reps=3, log=None, specs=None, margin=0.03):
    """Tune every lattice side in `sides` (n = side^2) on the device.

    Trial batch per n: with `specs` (the workload's own (lx, p, U, instance)
    draw), up to `batch` of its matrices with that lattice side, so the p
    values and U of the workload are the ones timed; otherwise `batch`
    matrices of order p.  Coordinate search, keeping the fastest at each
    stage: (nb, scan form) over candidates(n), then the BSTRI GEMM tile, then
    the reflector slabs per workgroup (these two keep their launch-size rule
    unless a fixed value is faster by more than `margin`).  None of the knobs
    changes the bits.
    Returns {n: {"nb", "scan_mti", "gemm_tile", "refl_spw", "ms", "all"}}."""
    table = {}
    for lx in sides:
        n = lx * lx
        if specs:
            trial = [tuple(sp) for sp in specs if sp[0] == lx][:batch]
        else:
            trial = [(lx, p, 4.0, 9000 + k) for k in range(batch)]
        mats = DeviceMatrices(B, ctx, trial, keep_blocks=False)
        res = {}

        def trial_ms(nb, mti, tile, spw):
            key = (nb, mti, tile, spw)
            if key not in res:
                ctx.set_tuning(n, nb, mti, tile, spw)
                t0 = time.perf_counter()
                res[key] = time_plan(B, ctx, mats.descs(), reps, mats.rebuild)
                mats.rebuild()
                if log:
                    log(f"[tune] n={n} nb={nb} scan_mti={mti} gemm_tile={tile} refl_spw={spw}: "
                        f"{res[key]:.3f} ms ({time.perf_counter() - t0:.1f} s)")
            return res[key]

        def pick(keys):
            # the launch-size rule (knob 0) stays unless a fixed value beats
            # it by more than `margin`: the rule adapts to the width of a
            # real batch's launches, which a small trial batch does not show
            base = keys[0]
            best_k = min(keys, key=lambda k: trial_ms(*k))
            return best_k if trial_ms(*best_k) < (1.0 - margin) * trial_ms(*base) else base

        best = min(((nb, mti, 0, 0) for nb, mti in candidates(n)), key=lambda k: trial_ms(*k))
        best = pick([(best[0], best[1], t, 0) for t in GEMM_TILES])
        best = pick([(best[0], best[1], best[2], s) for s in REFL_SPWS])
        ctx.set_tuning(n, *best)
        table[n] = {"nb": best[0], "scan_mti": best[1], "gemm_tile": best[2], "refl_spw": best[3], "ms": res[best],
                    "trial": [list(t[:2]) for t in trial],
                    "all": {",".join(str(v) for v in k): round(v, 4) for k, v in res.items()}}
        mats.release()
    return table


def refine(B, ctx, table, descs, rebuild, reps=3, log=None, margin=0.005):
    """Second pass on the whole workload (`descs`, already on the device,
    `rebuild` restoring their inputs).  Block sizes tuned in isolation share
    the GPU in the real plan, so for each tuned n (largest first) the whole
    plan is timed with up to three rows for that n: the tuned one, the
    library's defaults, and the fastest trial row with another scan form;
    the defaults stay unless another row is faster by more than `margin`.
    Each row's time is the shortest of `reps` executions: run-to-run noise
    on a whole plan only ever adds time (±1.5 % between executions, against
    a margin of 0.5 %).
    Then, on the row kept, 2 and 4 reflector slabs per workgroup (block
    sizes up to n = 256, whose reflector steps take the multi-slab form):
    a trial batch of one block size sees a different launch mix from the
    whole plan's, where a count the trial rejected can pay.  A count is kept
    only if faster by more than `margin` again.  Updates `table` in place
    (row["kept"]: "tuned", "default" or "alternative", with "+spw2" / "+spw4"
    where a slab count was kept) and the context's tuning."""
    for n in sorted(table, key=int, reverse=True):
        row = table[n]
        tuned = (int(row["nb"]), int(row["scan_mti"]), int(row.get("gemm_tile", 0)), int(row.get("refl_spw", 0)))
        ctx.set_tuning(int(n), 0, 0, 0, 0)
        default = ctx.get_tuning_ex(int(n))
        cands = {"default": default}
        if tuned != default:
            cands["tuned"] = tuned
        forms = {default[1], tuned[1]}
        for key, _ in sorted(row.get("all", {}).items(), key=lambda kv: kv[1]):
            k = tuple(int(v) for v in key.split(","))
            if k[1] not in forms:
                cands["alternative"] = k
                break
        times = {}
        for name, k in cands.items():
            ctx.set_tuning(int(n), *k)
            times[name] = time_plan(B, ctx, descs, reps, rebuild, stat="min")
            rebuild()
        best = min(times, key=times.get)
        if times[best] >= (1.0 - margin) * times["default"]:
            best = "default"
        k = cands[best]
        kept, best_t = best, times[best]
        if int(n) <= 256:
            for spw in (2, 4):
                if k[3] == spw:
                    continue
                ks = (k[0], k[1], k[2], spw)
                ctx.set_tuning(int(n), *ks)
                name = f"{best}+spw{spw}"
                cands[name] = ks
                times[name] = time_plan(B, ctx, descs, reps, rebuild, stat="min")
                rebuild()
                if times[name] < (1.0 - margin) * best_t:
                    k, kept, best_t = ks, name, times[name]
        ctx.set_tuning(int(n), *k)
        row.update(nb=k[0], scan_mti=k[1], gemm_tile=k[2], refl_spw=k[3], kept=kept,
                   whole_ms={name: round(t, 3) for name, t in times.items()})
        if log:
            log(f"[tune] whole workload, n={n}: " + ", ".join(f"{name} {cands[name]} {t:.2f} ms"
                                                             for name, t in times.items()) + f" -> {kept}")
    return table


def apply(ctx, table):
    for n, row in table.items():
        ctx.set_tuning(int(n), int(row["nb"]), int(row["scan_mti"]), int(row.get("gemm_tile", 0)),
                       int(row.get("refl_spw", 0)))


def save(table, path):
    with open(path, "w") as f:
        json.dump({str(k): v for k, v in table.items()}, f, indent=1)


def load(path):
    with open(path) as f:
        return {int(k): v for k, v in json.load(f).items()}

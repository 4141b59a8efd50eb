"""Serving driver over the native continuous-batching loop
(bsofi_server_* in libbsofi_amd.so).

Simulates DQMC chains that send inversion requests: each chain has a fixed
lattice and p; every request flips a fraction of that chain's HS fields (the
flips are the only data shipped), and asks for the full G.  Arrivals are a
Poisson process at a fixed rate; the loop calls step() continuously,
collects completions, releases results and reports throughput and latency
percentiles.
"""
from __future__ import annotations

import ctypes as C
import sys
import time

import numpy as np

from . import lib

_vp = C.c_void_p
_ll = C.c_longlong

_SIG = {
    "bsofi_server_create": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.c_int, C.POINTER(_vp)]),
    "bsofi_server_destroy": (C.c_int, [_vp]),
    "bsofi_server_add_chain": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         C.c_ulonglong, C.c_int, C.POINTER(C.c_int)]),
    "bsofi_server_submit": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(_ll)]),
    "bsofi_server_step": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "bsofi_server_poll": (C.c_int, [_vp, C.POINTER(_ll), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int), C.c_int]),
    "bsofi_server_result": (C.c_int, [_vp, _ll, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]),
    "bsofi_server_release": (C.c_int, [_vp, _ll]),
    "bsofi_server_stats": (C.c_int, [_vp, C.POINTER(_ll), C.POINTER(C.c_int), C.POINTER(C.c_size_t)]),
    "bsofi_server_request_bytes": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_size_t)]),
}
for _name, (_res, _args) in _SIG.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args


class Server:
    def __init__(self, budget_bytes=200 << 30, slots=4, max_batch=64, device=0):
        h = _vp()
        rc = lib.bsofi_server_create(device, budget_bytes, slots, max_batch, C.byref(h))
        if rc:
            raise RuntimeError(f"bsofi_server_create failed ({rc})")
        self.h = h
        self.chain_dims = []

    def add_chain(self, lx, p, U=4.0, dtau=0.125, t=1.0, seed=20261203, instance=0, ly=None):
        cid = C.c_int(-1)
        rc = lib.bsofi_server_add_chain(self.h, lx, ly or lx, p, U, dtau, t, seed, instance, C.byref(cid))
        if rc:
            raise RuntimeError(f"add_chain failed ({rc})")
        self.chain_dims.append((lx * (ly or lx), p))
        return cid.value

    def submit(self, chain, flips):
        flips = np.ascontiguousarray(flips, dtype=np.int32)
        rid = _ll(0)
        rc = lib.bsofi_server_submit(self.h, chain, flips.ctypes.data_as(C.POINTER(C.c_int)), int(flips.size),
                                     C.byref(rid))
        if rc:
            raise RuntimeError(f"submit failed ({rc})")
        return rid.value

    def step(self):
        inflight = C.c_int(0)
        q = lib.bsofi_server_step(self.h, C.byref(inflight))
        if q < 0:
            raise RuntimeError(f"step failed ({q})")
        return q, inflight.value

    def poll(self, max_n=4096):
        ids = (_ll * max_n)()
        lat = (C.c_double * max_n)()
        qw = (C.c_double * max_n)()
        info = (C.c_int * max_n)()
        k = lib.bsofi_server_poll(self.h, ids, lat, qw, info, max_n)
        return [(ids[i], lat[i], qw[i], info[i]) for i in range(k)]

    def result(self, rid):
        g, t = _vp(), _vp()
        n, p = C.c_int(0), C.c_int(0)
        rc = lib.bsofi_server_result(self.h, rid, C.byref(g), C.byref(t), C.byref(n), C.byref(p))
        if rc:
            raise RuntimeError(f"result failed ({rc})")
        return g.value, t.value, n.value, p.value

    def release(self, rid):
        return lib.bsofi_server_release(self.h, rid)

    def stats(self):
        c, q, u = _ll(0), C.c_int(0), C.c_size_t(0)
        lib.bsofi_server_stats(self.h, C.byref(c), C.byref(q), C.byref(u))
        return c.value, q.value, u.value

    def close(self):
        if getattr(self, "h", None):
            lib.bsofi_server_destroy(self.h)
            self.h = None

    def __del__(self, _finalizing=sys.is_finalizing):
        # at interpreter shutdown the HIP runtime may already be torn down:
        # leave the memory to the process exit instead of calling into it
        # (sys.is_finalizing is bound when the class is defined: the module's
        # globals, sys among them, may already be gone when this runs)
        if _finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


def request_bytes(n, p):
    """Device bytes one request of block size n and p holds while it runs
    (G, tau, field snapshot, plan workspace): what admission counts against
    the HBM budget."""
    out = C.c_size_t(0)
    rc = lib.bsofi_server_request_bytes(n, p, C.byref(out))
    if rc:
        raise ValueError(f"request_bytes({n}, {p}) failed ({rc})")
    return out.value


def c5_chains(nchains=64, seed=20261203):
    """Chain sizes of the serving workload: n in {64, 144, 256} (lattice
    sides 8, 12, 16), p in {20, 40, 80}, drawn from a seeded generator."""
    rng = np.random.default_rng(seed)
    sides = rng.choice([8, 12, 16], size=nchains)
    ps = rng.choice([20, 40, 80], size=nchains)
    return [(int(s), int(p)) for s, p in zip(sides, ps)]


def run(server: Server, chains, nrequests, rate_per_s, seed=1, flip_lo=0.01, flip_hi=0.05, on_done=None,
        max_seconds=120.0):
    """Poisson arrivals at rate_per_s from the given chain ids; returns
    per-request latencies (s) and the wall time."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=nrequests)
    arrivals = np.cumsum(gaps)
    who = rng.integers(0, len(chains), size=nrequests)
    lat, waits, infos = {}, {}, {}
    t0 = time.perf_counter()
    sent = 0
    done = 0
    while done < nrequests:
        now = time.perf_counter() - t0
        while sent < nrequests and arrivals[sent] <= now:
            cid = chains[who[sent]]
            n, p = server.chain_dims[cid]
            k = max(1, int(round(rng.uniform(flip_lo, flip_hi) * n * p)))
            flips = rng.choice(n * p, size=k, replace=False)
            server.submit(cid, flips)
            sent += 1
        server.step()
        for rid, l, qw, info in server.poll():
            lat[rid], waits[rid], infos[rid] = l, qw, info
            if on_done is not None:
                on_done(server, rid)
            server.release(rid)
            done += 1
        if time.perf_counter() - t0 > max_seconds:
            break
    wall = time.perf_counter() - t0
    return np.array(list(lat.values())), np.array(list(waits.values())), infos, wall

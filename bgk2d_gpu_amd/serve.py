"""Serving loop: continuous batching of many BGK2D solves on one GPU.

Requests (driven cavity, Poiseuille channel, Taylor-Green; any size; FP32
or FP16 storage) arrive over time.  Arrived requests are admitted in FIFO
order while their device footprint fits the HBM budget and a slot is free;
every admitted request stays resident and all of them advance together,
`chunk` steps per replayed graph of the persistent batch kernel
(bgk_batch_run).  Between chunks finished requests are checked, reported and
retired, which frees memory for the next arrivals.  Every `check_every`
steps of its own, each member is checked for non-finite values or |u| > 0.5
and retired as diverged at that step.  The same checks run at the same steps
in the sequential baseline, so a request's outcome is the same alone and in
the mix.
"""
import ctypes as C
import json
import math
import os
import time
from dataclasses import dataclass, field

import numpy as np

from .solver import BC, BgkError, BgkParams, IMPL, PREC, Solver, load

KINDS = ("cavity", "poiseuille", "tgv")


@dataclass
class Request:
    rid: int
    kind: str
    side: int
    re: float
    steps: int
    prec: str
    arrival: float = 0.0
    tblock: int = 0  # bgk_params_t.tblock: 0 automatic, 1 one step per launch, 2 two
    forcing_y: int = 0  # bgk_params_t.forcing_y: a channel's body force along y

    def geometry(self):
        """(lx, ly, bc, svisc, u0, amplitude): velocity scale U = 0.1 for the
        cavity lid and the channel centre-line, 0.01 for Taylor-Green; the
        channel is 4:1 long; nu = U * (characteristic length) / Re"""
        if self.kind == "poiseuille":
            lx, ly = self.side, max(8, self.side // 4)
            return lx, ly, "channel", 0.1 * ly / self.re, 0.1, 0.1
        if self.kind == "tgv":
            return self.side, self.side, "periodic", 0.01 * self.side / self.re, 0.0, 0.01
        return self.side, self.side, "driven", 0.1 * self.side / self.re, 0.1, 0.1

    @property
    def lups(self):
        lx, ly = self.geometry()[:2]
        return lx * ly * self.steps

    def params(self, device=0, form=1, rounding=1, tblock=None):
        """bgk_params_t of the solve (tblock: the request's unless given).
        The batch and bgk_run choose one or two steps per launch by the same
        rule; every path gives the same bits."""
        tblock = self.tblock if tblock is None else tblock
        lx, ly, bc, svisc, u0, _ = self.geometry()
        narrow = self.prec in ("fp16", "bf16")
        return BgkParams(lx=lx, ly=ly, bc=BC[bc], impl=IMPL["fused"], prec=PREC[self.prec],
                         radius=16, device=device, form=form if narrow else 0,
                         rounding=rounding if narrow else 0, tblock=tblock,
                         forcing_y=self.forcing_y, svisc=float(svisc),
                         u0=float(u0), scale=1024.0 if narrow else 1.0, cte_s=0.1)


def make_stream(n, seed=20261203, rate=200.0, side=(128, 4096), re=(100.0, 1e4),
                steps=(1000, 50000), fp16_frac=0.5):
    """Seeded request stream: exponential inter-arrival times (rate per s),
    kind uniform, side and Re log-uniform, steps uniform, FP16 w.p. fp16_frac"""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for r in range(n):
        t += rng.exponential(1.0 / rate)
        s = int(round(math.exp(rng.uniform(math.log(side[0]), math.log(side[1])))))
        out.append(Request(r, KINDS[int(rng.integers(0, 3))], s,
                           float(math.exp(rng.uniform(math.log(re[0]), math.log(re[1])))),
                           int(rng.integers(steps[0], steps[1] + 1)),
                           "fp16" if rng.random() < fp16_frac else "single", t))
    return out


@dataclass
class Outcome:
    rid: int
    arrival: float
    start: float = -1.0
    done: float = -1.0
    status: str = "pending"   # "ok" | "diverged@<step>" | "rejected"
    diag: list = field(default_factory=list)
    pops: object = None

    @property
    def latency(self):
        return self.done - self.arrival


class Batch:
    """ctypes view of bgk_batch_*"""

    def __init__(self, budget_bytes, max_members=256, device=0):
        self.L = load()
        L = self.L
        P, V, I32, I64, D = C.POINTER, C.c_void_p, C.c_int32, C.c_int64, C.c_double
        for name, args in (("bgk_batch_create", [I32, D, I32, P(V)]), ("bgk_batch_destroy", [V]),
                           ("bgk_batch_add", [V, P(BgkParams), I32, D, I64, I32, P(I32)]),
                           ("bgk_batch_ctx", [V, I32, P(V)]), ("bgk_batch_remaining", [V, I32, P(I64)]),
                           ("bgk_batch_remove", [V, I32]), ("bgk_batch_hold", [V, I32, I32]),
                           ("bgk_batch_run", [V, I32, P(I32)]),
                           ("bgk_batch_used", [V, P(D)]), ("bgk_batch_last_ms", [V, P(D)]),
                           ("bgk_footprint", [P(BgkParams), P(D)])):
            getattr(L, name).argtypes = args
            getattr(L, name).restype = C.c_int
        L.bgk_batch_last_error.argtypes = [V]
        L.bgk_batch_last_error.restype = C.c_char_p
        h = C.c_void_p()
        rc = L.bgk_batch_create(device, float(budget_bytes), max_members, C.byref(h))
        if rc != 0:
            raise BgkError(rc, "bgk_batch_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.L.bgk_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, p, init_v, amplitude, nsteps, tile=0):
        """slot index, or None when the budget or the slots are full"""
        slot = C.c_int32(-1)
        rc = self.L.bgk_batch_add(self.h, C.byref(p), init_v, float(amplitude), int(nsteps), tile,
                                  C.byref(slot))
        if rc == -2:
            return None
        if rc != 0:
            raise BgkError(rc, self.L.bgk_batch_last_error(self.h).decode())
        return slot.value

    def ctx(self, slot):
        h = C.c_void_p()
        rc = self.L.bgk_batch_ctx(self.h, slot, C.byref(h))
        if rc != 0:
            raise BgkError(rc, "bgk_batch_ctx")
        return h

    def run(self, chunk):
        fin = C.c_int32(0)
        rc = self.L.bgk_batch_run(self.h, int(chunk), C.byref(fin))
        if rc != 0:
            raise BgkError(rc, self.L.bgk_batch_last_error(self.h).decode())
        return fin.value

    def remaining(self, slot):
        r = C.c_int64(0)
        self.L.bgk_batch_remaining(self.h, slot, C.byref(r))
        return r.value

    def remove(self, slot):
        self.L.bgk_batch_remove(self.h, slot)

    def hold(self, slot, held):
        rc = self.L.bgk_batch_hold(self.h, slot, 1 if held else 0)
        if rc != 0:
            raise BgkError(rc, self.L.bgk_batch_last_error(self.h).decode())

    def last_ms(self):
        v = C.c_double(0)
        self.L.bgk_batch_last_ms(self.h, C.byref(v))
        return v.value


def _ctx_diag(L, ctx):
    out = (C.c_double * 4)()
    L.bgk_diagno(ctx, out)
    return list(out)


def _ctx_ok(L, ctx, umax=0.5):
    return L.bgk_check(ctx, float(umax)) == 0


def _ctx_raw(L, ctx, prec, shape):
    from .solver import STORE_DTYPE
    out = np.zeros((9,) + shape, dtype=STORE_DTYPE[prec])
    for k in range(9):
        L.bgk_download(ctx, k, out[k].ctypes.data)
    return out


TILE_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "bgk2d_gpu_amd", "tiles.json")


def autotune(prec, side, steps=64, chunk=32, device=0, cache=True):
    """Pick the batch tile shape (64x4, 128x2, 256x1, 32x8) that runs a
    cavity of this storage kind and size class fastest; cached per
    (storage kind, power-of-two size class)."""
    key = f"{prec}:{1 << max(7, int(round(math.log2(max(side, 2)))))}"
    table = {}
    if cache and os.path.exists(TILE_CACHE):
        try:
            table = json.load(open(TILE_CACHE))
        except (OSError, ValueError):
            table = {}
    if key in table:
        return table[key]
    n = 1 << max(7, int(round(math.log2(max(side, 2)))))
    req = Request(0, "cavity", n, 1000.0, steps, prec)
    best, best_ms = 1, float("inf")
    for tile in (1, 2, 3, 4):
        b = Batch(budget_bytes=64e9, max_members=1, device=device)
        slot = b.add(req.params(device), 0, 0.1, steps, tile)
        b.run(chunk)  # warm-up and graph build
        t0 = time.perf_counter()
        while b.run(chunk) == 0 and b.remaining(slot) > 0:
            pass
        ms = time.perf_counter() - t0
        b.close()
        if ms < best_ms:
            best, best_ms = tile, ms
    table[key] = best
    if cache:
        os.makedirs(os.path.dirname(TILE_CACHE), exist_ok=True)
        json.dump(table, open(TILE_CACHE, "w"))
    return best


def serve_batched(reqs, budget_bytes=200e9, chunk=100, check_every=1000, max_members=512,
                  tiles=None, keep=(), device=0, clock=time.perf_counter, fill_nodes=None):
    """Continuous batching of `reqs` (arrival times relative to the start).
    fill_nodes=None advances every admitted request each chunk (the most
    aggregate throughput).  With fill_nodes, each chunk advances only the
    admitted requests with the least remaining work (nodes x steps left)
    whose lattices together hold at least fill_nodes nodes and holds the
    others (shortest remaining work first: short requests are not slowed
    by long ones sharing every step; a lower p50 for less throughput).
    Returns (outcomes by rid, stats)."""
    assert check_every % chunk == 0
    L = load()
    b = Batch(budget_bytes, max_members, device)
    pending = sorted(reqs, key=lambda r: r.arrival)
    out = {r.rid: Outcome(r.rid, r.arrival) for r in reqs}
    active = {}  # slot -> (request, steps done)
    busy = 0.0
    t0 = clock()
    qi = 0
    while qi < len(pending) or active:
        now = clock() - t0
        # admit arrived requests in FIFO order while they fit
        while qi < len(pending) and pending[qi].arrival <= now:
            r = pending[qi]
            tile = (tiles or {}).get(r.prec, 0)
            slot = b.add(r.params(device), 0, r.geometry()[5], r.steps, tile)
            if slot is None:
                if not active:  # larger than the whole budget
                    out[r.rid].status = "rejected"
                    qi += 1
                    continue
                break
            out[r.rid].start = clock() - t0
            active[slot] = [r, 0, False]  # request, steps done, held
            qi += 1
        if not active:
            if qi < len(pending):
                time.sleep(max(0.0, min(0.01, pending[qi].arrival - (clock() - t0))))
            continue
        if fill_nodes:
            def work(slot):
                lx, ly = active[slot][0].geometry()[:2]
                return lx * ly * b.remaining(slot), lx * ly
            order = sorted(active, key=lambda sl: work(sl)[0])
            nodes, run = 0, set()
            for sl in order:
                run.add(sl)
                nodes += work(sl)[1]
                if nodes >= fill_nodes:
                    break
            for sl in active:
                held = sl not in run
                if held != active[sl][2]:
                    b.hold(sl, held)
                    active[sl][2] = held
        b.run(chunk)
        busy += b.last_ms() / 1e3
        done_now = clock() - t0
        for slot in list(active):
            if active[slot][2]:  # held this chunk
                continue
            r, done = active[slot][:2]
            rem = b.remaining(slot)
            done = r.steps - rem
            active[slot][1] = done
            ctx = b.ctx(slot)
            finished = rem == 0
            if finished or done % check_every == 0:
                if not _ctx_ok(L, ctx):
                    out[r.rid].status = f"diverged@{done}"
                    out[r.rid].done = done_now
                    b.remove(slot)
                    del active[slot]
                    continue
            if finished:
                o = out[r.rid]
                o.status = "ok"
                o.done = done_now
                o.diag = _ctx_diag(L, ctx)
                if r.rid in keep:
                    lx, ly = r.geometry()[:2]
                    o.pops = _ctx_raw(L, ctx, r.prec, (ly + 2, lx + 2))
                b.remove(slot)
                del active[slot]
    makespan = clock() - t0
    b.close()
    return out, dict(makespan=makespan, busy=busy)


def serve_sequential(reqs, check_every=1000, keep=(), device=0, clock=time.perf_counter):
    """FIFO, one request at a time, each with bgk_run in check_every-step
    pieces (the same kernels and the same checks as the batch)."""
    out = {r.rid: Outcome(r.rid, r.arrival) for r in reqs}
    t0 = clock()
    for r in sorted(reqs, key=lambda x: x.arrival):
        now = clock() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        o = out[r.rid]
        o.start = clock() - t0
        lx, ly, bc, svisc, u0, amp = r.geometry()
        narrow = r.prec in ("fp16", "bf16")
        s = Solver(lx, ly, bc=bc, impl="fused", prec=r.prec, svisc=svisc, u0=u0, amplitude=amp,
                   scale=1024.0 if narrow else 1.0, form=1 if narrow else 0,
                   rounding=1 if narrow else 0, tblock=r.tblock, forcing_y=r.forcing_y,
                   device=device)
        done, status = 0, "ok"
        while done < r.steps:
            n = min(check_every, r.steps - done)
            s.run(n)
            done += n
            if not s.check(0.5):
                status = f"diverged@{done}"
                break
        s.sync()
        o.done = clock() - t0
        o.status = status
        if status == "ok":
            o.diag = list(s.diagno())
            if r.rid in keep:
                o.pops = s.raw()
        s.close()
    return out, dict(makespan=clock() - t0)


def summarize(reqs, outcomes, stats):
    lat = np.array([outcomes[r.rid].latency for r in reqs if outcomes[r.rid].done >= 0])
    ok = [r for r in reqs if outcomes[r.rid].status == "ok"]
    div = [r for r in reqs if outcomes[r.rid].status.startswith("diverged")]
    lups = sum(r.lups for r in ok)
    # diverged requests count the steps they ran
    for r in div:
        st = int(outcomes[r.rid].status.split("@")[1])
        lx, ly = r.geometry()[:2]
        lups += lx * ly * st
    return dict(n=len(reqs), ok=len(ok), diverged=len(div),
                p50_s=float(np.percentile(lat, 50)) if len(lat) else None,
                p99_s=float(np.percentile(lat, 99)) if len(lat) else None,
                makespan_s=stats["makespan"], mlups=lups / stats["makespan"] / 1e6)

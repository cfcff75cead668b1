"""Batch members against the oracle, bit for bit.

Members of every storage kind the batch serves (FP32; FP16 and BF16 with
the serving settings: scale 1024, deviation form, stochastic rounding) and
every boundary variant it serves (driven, periodic, channel with the body
force along x and along y, inflow, couette, open), of uneven sizes and step
counts, advance together through bgk_batch_run: one-step groups, and
two-step groups large enough to fill the chip.  Each member's populations
equal the oracle's for the same parameters."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as po
from bgk2d_gpu_amd import serve
from bgk2d_gpu_amd.solver import BgkParams, BC, IMPL, PREC, STORE_DTYPE, load

pytestmark = pytest.mark.gpu

ORC = {"single": "single", "fp16": "mixed1", "bf16": "bf16"}


def params(lx, ly, bc, prec, svisc, u0, forcing_y=0, tblock=0):
    narrow = prec != "single"
    return BgkParams(lx=lx, ly=ly, bc=BC[bc], impl=IMPL["fused"], prec=PREC[prec], radius=16,
                     form=1 if narrow else 0, rounding=1 if narrow else 0, tblock=tblock,
                     forcing_y=forcing_y, svisc=svisc, u0=u0, scale=1024.0 if narrow else 1.0,
                     cte_s=0.1)


def oracle(lx, ly, bc, prec, svisc, u0, steps, amp, forcing_y=0):
    narrow = prec != "single"
    o = po.Oracle(lx, ly, bc=bc, impl="fused", prec=ORC[prec], svisc=svisc, u0=u0,
                  amplitude=amp, scale=1024.0 if narrow else 1.0, form=1 if narrow else 0,
                  rounding=1 if narrow else 0, forcing_y=forcing_y)
    o.step(steps)
    return o.raw()


def run_batch(members, chunk):
    """members: (lx, ly, bc, prec, svisc, u0, steps, amp, forcing_y, tblock)"""
    L = load()
    b = serve.Batch(budget_bytes=60e9, max_members=len(members))
    slots = []
    for lx, ly, bc, prec, sv, u0, steps, amp, fy, tb in members:
        slot = b.add(params(lx, ly, bc, prec, sv, u0, fy, tb), 0, amp, steps)
        assert slot is not None
        slots.append(slot)
    while any(b.remaining(s) > 0 for s in slots):
        b.run(chunk)
    out = []
    for (lx, ly, bc, prec, *_), s in zip(members, slots):
        raw = np.zeros((9, ly + 2, lx + 2), dtype=STORE_DTYPE[prec])
        ctx = b.ctx(s)
        for k in range(9):
            assert L.bgk_download(ctx, k, raw[k].ctypes.data) == 0
        out.append(raw)
    b.close()
    return out


def check(members, got):
    for m, g in zip(members, got):
        lx, ly, bc, prec, sv, u0, steps, amp, fy, tb = m
        ref = oracle(lx, ly, bc, prec, sv, u0, steps, amp, fy)
        u = np.uint32 if prec == "single" else np.uint16
        inner = (slice(None), slice(1, -1), slice(1, -1))
        diff = g[inner].view(u) != ref[inner].view(u)
        assert not diff.any(), f"{m}: {int(diff.sum())} differing elements"


def test_one_step_groups_equal_oracle():
    """18 members: three storage kinds x six variants, sizes and step
    counts all different (chunks of 10 steps, members ending mid-chunk)"""
    members = []
    variants = [("driven", 0.05, 0.1, 0.1), ("periodic", 0.05, 0.0, 0.01), ("channel", 0.05, 0.1, 0.1),
                ("inflow", 0.05, 0.1, 0.1), ("couette", 0.05, 0.1, 0.1), ("open", 0.05, 0.05, 0.1)]
    for a, prec in enumerate(("single", "fp16", "bf16")):
        for v, (bc, sv, u0, amp) in enumerate(variants):
            lx, ly = 60 + 13 * v + 7 * a, 40 + 5 * v + 3 * a
            members.append((lx, ly, bc, prec, sv, u0, 23 + 4 * v + a, amp, 0, 0))
    members.append((90, 50, "channel", "single", 0.05, 0.1, 31, 0.1, 1, 0))
    members.append((90, 50, "channel", "fp16", 0.05, 0.1, 29, 0.1, 1, 0))
    got = run_batch(members, chunk=10)
    check(members, got)


def test_two_step_groups_equal_oracle():
    """groups that fill the chip with strips of >= 8 rows advance two
    steps per launch (every storage kind by default); the odd
    remainder one step per launch"""
    members = [(1024, 800, "driven", "single", 0.05, 0.1, 40 + 2 * i, 0.1, 0, 0) for i in range(4)]
    members += [(1024, 800, "periodic", "fp16", 0.05, 0.0, 30 + i, 0.01, 0, 2) for i in range(4)]
    got = run_batch(members, chunk=20)
    check(members, got)

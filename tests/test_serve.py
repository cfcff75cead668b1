"""Serving loop: the request stream, continuous batching and its parity
with solves run alone."""
import numpy as np
import pytest

from bgk2d_gpu_amd import serve
from bgk2d_gpu_amd.serve import Request


def test_stream_is_seeded_and_follows_its_distributions():
    a = serve.make_stream(400, seed=5)
    b = serve.make_stream(400, seed=5)
    assert [(r.kind, r.side, r.steps, r.prec) for r in a] == [(r.kind, r.side, r.steps, r.prec) for r in b]
    gaps = np.diff([0.0] + [r.arrival for r in a])
    assert abs(gaps.mean() - 1 / 200) < 0.2 / 200 * 5  # exponential, mean 5 ms
    sides = np.log([r.side for r in a])
    assert np.log(128) - 0.01 <= sides.min() and sides.max() <= np.log(4096) + 0.01
    assert set(r.kind for r in a) == set(serve.KINDS)
    assert 0.35 < np.mean([r.prec == "fp16" for r in a]) < 0.65


def test_request_geometry():
    r = Request(0, "poiseuille", 512, 1000.0, 100, "single")
    lx, ly, bc, svisc, u0, amp = r.geometry()
    assert (lx, ly, bc) == (512, 128, "channel") and svisc == pytest.approx(0.1 * 128 / 1000)
    r = Request(1, "tgv", 256, 100.0, 100, "fp16")
    assert r.geometry()[2] == "periodic" and r.geometry()[5] == 0.01
    assert r.lups == 256 * 256 * 100


@pytest.mark.gpu
def test_batched_members_equal_solves_alone():
    """mixed sizes, kinds and storage kinds, admitted at different times
    and finishing at different steps: each result is bit-identical to the
    same request run alone"""
    reqs = [Request(0, "cavity", 200, 400.0, 730, "single", 0.0),
            Request(1, "tgv", 150, 300.0, 410, "fp16", 0.0),
            Request(2, "poiseuille", 256, 200.0, 960, "single", 0.0),
            Request(3, "cavity", 97, 100.0, 333, "fp16", 0.0),
            Request(4, "tgv", 300, 1000.0, 500, "bf16", 0.0),
            Request(5, "cavity", 513, 1000.0, 250, "single", 0.0),
            # the FORCING_Y channel (body force along y) beside the x-forced one
            Request(9, "poiseuille", 256, 200.0, 470, "single", 0.0, forcing_y=1),
            Request(10, "poiseuille", 300, 200.0, 390, "fp16", 0.0, forcing_y=1),
            # members of >= 2^22 nodes advance with bgk_run's kernels (the
            # two-step kernel for FP32) before the batch kernels
            Request(6, "cavity", 2050, 1000.0, 37, "single", 0.0),
            Request(7, "poiseuille", 4100, 500.0, 23, "single", 0.0),
            Request(8, "tgv", 2048, 500.0, 30, "fp16", 0.0)]
    keep = {r.rid for r in reqs}
    bo, _ = serve.serve_batched(reqs, budget_bytes=60e9, chunk=10, check_every=100, keep=keep,
                                tiles={"fp16": 3})
    so, _ = serve.serve_sequential(reqs, check_every=100, keep=keep)
    for r in reqs:
        assert bo[r.rid].status == so[r.rid].status == "ok"
        a, b = bo[r.rid].pops, so[r.rid].pops
        assert np.array_equal(a[:, 1:-1, 1:-1], b[:, 1:-1, 1:-1]), r
        np.testing.assert_array_equal(bo[r.rid].diag, so[r.rid].diag)


@pytest.mark.gpu
def test_two_step_batch_members_equal_solves_alone():
    """groups large enough to fill the chip advance two steps per launch
    (every storage kind by default), their last partial
    chunk one step per launch: each result is bit-identical to the same
    request run alone"""
    reqs = [Request(i, "cavity", 1024, 1000.0, 64 + 3 * i, "single", 0.0) for i in range(4)]
    reqs += [Request(4 + i, "tgv", 1024, 500.0, 41 + 2 * i, "fp16", 0.0, tblock=2) for i in range(4)]
    reqs += [Request(17, "poiseuille", 4096, 500.0, 30, "bf16", 0.0, tblock=2),
             Request(18, "poiseuille", 4000, 500.0, 27, "bf16", 0.0, tblock=2),
             Request(19, "cavity", 300, 1000.0, 50, "single", 0.0)]
    reqs += [Request(20 + i, "poiseuille", 2048, 500.0, 44 + i, "single", 0.0, forcing_y=i & 1)
             for i in range(4)]
    keep = {r.rid for r in reqs}
    bo, _ = serve.serve_batched(reqs, budget_bytes=60e9, chunk=10, check_every=100, keep=keep)
    so, _ = serve.serve_sequential(reqs, check_every=100, keep=keep)
    for r in reqs:
        assert bo[r.rid].status == so[r.rid].status == "ok"
        a, b = bo[r.rid].pops, so[r.rid].pops
        assert np.array_equal(a[:, 1:-1, 1:-1], b[:, 1:-1, 1:-1]), r
        np.testing.assert_array_equal(bo[r.rid].diag, so[r.rid].diag)


@pytest.mark.gpu
def test_shortest_first_holds_members_and_keeps_results():
    """with a fill threshold the loop holds the requests with the most work
    left: the short ones finish first, and every result is still
    bit-identical to the request run alone"""
    reqs = [Request(0, "cavity", 512, 1000.0, 300, "single", 0.0),
            Request(1, "tgv", 96, 300.0, 120, "fp16", 0.0),
            Request(2, "poiseuille", 256, 200.0, 90, "single", 0.0),
            Request(3, "cavity", 128, 100.0, 60, "bf16", 0.0)]
    keep = {r.rid for r in reqs}
    bo, _ = serve.serve_batched(reqs, budget_bytes=60e9, chunk=10, check_every=100, keep=keep,
                                fill_nodes=20000)
    so, _ = serve.serve_sequential(reqs, check_every=100, keep=keep)
    for r in reqs:
        assert bo[r.rid].status == so[r.rid].status == "ok"
        assert np.array_equal(bo[r.rid].pops[:, 1:-1, 1:-1], so[r.rid].pops[:, 1:-1, 1:-1]), r
    # the 512^2 cavity (most work) is held until the others are done
    assert bo[0].done >= max(bo[r.rid].done for r in reqs[1:])


@pytest.mark.gpu
def test_admission_respects_budget_and_divergence_matches():
    """a budget that holds about two requests at a time: everything still
    completes, and an unstable request (tau near 0.5) diverges at the same
    step in the batch and alone"""
    from bgk2d_gpu_amd.solver import load
    import ctypes as C
    reqs = [Request(i, "cavity", 256, 300.0, 400, "single", 0.0) for i in range(5)]
    reqs.append(Request(5, "cavity", 64, 1e6, 5000, "single", 0.0))  # nu tiny: unstable
    p = reqs[0].params()
    fp = C.c_double(0)
    L = load()
    L.bgk_footprint.argtypes = [C.POINTER(serve.BgkParams), C.POINTER(C.c_double)]
    L.bgk_footprint(C.byref(p), C.byref(fp))
    bo, _ = serve.serve_batched(reqs, budget_bytes=2.5 * fp.value, chunk=20, check_every=100)
    so, _ = serve.serve_sequential(reqs, check_every=100)
    for r in reqs:
        assert bo[r.rid].status == so[r.rid].status, (r, bo[r.rid].status, so[r.rid].status)
    assert bo[5].status.startswith("diverged@")
    assert all(bo[i].status == "ok" for i in range(5))

"""HIP path vs the reference and the oracle, through the C ABI.

Bar: bit-exact populations (the arithmetic is the reference's, operation
for operation, with no FMA contraction), FP64 diagnostics within 1e-12
relative of the oracle's long-double sums.
"""
import os

import numpy as np
import pytest

import pyoracle as po
from bgk2d_gpu_amd import Solver

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "ref_fixtures.npz"))
VARIANTS = sorted({k.split("/")[0] for k in FIX.files})
GPU_PREC = {"SINGLE": "single", "DOUBLE": "double", "MIXED1": "fp16", "MIXED2": "mixed2"}
ORC_PREC = {"single": "single", "double": "double", "fp16": "mixed1", "mixed2": "mixed2",
            "bf16": "bf16"}


def as_float(raw, prec):
    if prec == "fp16":
        return raw.view(np.float16).astype(np.float64)
    if prec == "bf16":
        return (raw.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    return raw.astype(np.float64)


def same_bits(a, b, prec):
    """bit equality, with every NaN equal to every NaN (x86 and gfx950
    produce different default NaN patterns)"""
    fa, fb = as_float(a, prec), as_float(b, prec)
    nan = np.isnan(fa) & np.isnan(fb)
    u = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}[a.dtype.itemsize]
    return (a.view(u) == b.view(u)) | nan


def fixture_solver(name, impl_override=None, **kw):
    p, impl, case = name.split("@")[0].split("_")
    lx, ly, ns, rad = (int(x) for x in FIX[name + "/params"])
    sv, u0 = (float(x) for x in FIX[name + "/fparams"])
    bc, obst = po.CASE[case]
    impl = impl_override or ("fused" if impl == "FUSED" else "original")
    s = Solver(lx, ly, bc=bc, impl=impl, prec=GPU_PREC[p], svisc=sv, u0=u0, obstacle=obst,
               radius=rad, mixed1_quirk=1 if p == "MIXED1" else 0,
               les=int(case in po.LES_CASES), cte_s=po.CTE_S_REF,
               trick1=int(case in po.TRICK1_CASES),
               forcing_y=int(case in po.FORCING_Y_CASES),
               noshift=int(case in po.NOSHIFT_CASES), **kw)
    return s, ns


@pytest.mark.parametrize("name", VARIANTS)
def test_reference_fixture_phases(name):
    """boundaries -> propagation -> collision per step: whole planes,
    halos included, equal the reference's"""
    print(f"case {name} phases", flush=True)
    s, ns = fixture_solver(name)
    s.step_phases(ns)
    got, ref = s.raw(), FIX[name + "/pops"]
    ok = same_bits(got, ref, s.prec)
    assert ok.all(), f"{int((~ok).sum())} differing elements"


@pytest.mark.parametrize("name", [v for v in VARIANTS if FIX.__contains__(v + "/obs")])
def test_reference_obstacle_mask(name):
    """bgk_build_obs makes the mask the reference's build_obs made"""
    s, _ = fixture_solver(name)
    mask, _ = s.obstacle()
    np.testing.assert_array_equal(mask, FIX[name + "/obs"])
    s.close()


@pytest.mark.parametrize("name", [v for v in VARIANTS if FIX.__contains__(v + "/drag")])
def test_reference_draglift(name):
    """bgk_draglift after the reference's steps against its draglift output
    (SRC/draglift.F90:55-101, fort.66: force / (u_inflow^2 radius)): the
    reference sums the per-link terms serially in its compute kind, the
    library in FP64; SINGLE relative 2e-6, DOUBLE 1e-12 of the drag."""
    s, ns = fixture_solver(name)
    s.run(ns)
    lx, ly, _, rad = (int(x) for x in FIX[name + "/params"])
    ic, jc = 2 * lx // 5, ly // 2
    fx, fy = s.draglift((ic - rad - 3, ic + rad + 3, jc - rad - 3, jc + rad + 3))
    norm = 1.0 / (float(np.float32(0.1)) ** 2 * rad)
    _, cd, cl = FIX[name + "/drag"][-1]
    tol = 2e-6 if name.startswith("SINGLE") else 1e-12
    assert abs(fx * norm - cd) <= tol * abs(cd)
    assert abs(fy * norm - cl) <= tol * abs(cd)
    s.close()


@pytest.mark.parametrize("name", VARIANTS)
def test_reference_fixture_run(name):
    """bgk_run (boundary folded into the fused pull where possible):
    interior equals the reference's"""
    print(f"case {name} run", flush=True)
    s, ns = fixture_solver(name)
    s.run(ns)
    got, ref = s.raw(), FIX[name + "/pops"]
    ok = same_bits(got[:, 1:-1, 1:-1], ref[:, 1:-1, 1:-1], s.prec)
    assert ok.all(), f"{int((~ok).sum())} differing elements"


NS_FOLDED = [v for v in VARIANTS
             if v.split("_")[1] == "FUSED" and v.split("_")[2].split("@")[0] in po.NOSHIFT_CASES
             and po.CASE[v.split("_")[2].split("@")[0]][0] != "step"]


@pytest.mark.parametrize("tb", [1, 2])
@pytest.mark.parametrize("name", NS_FOLDED)
def test_noshift_folded_kernels(name, tb):
    """NOSHIFT (whole populations, SRC/col_MC.F90:70-74, the cte1 = 1
    inflow/outflow equilibria of SRC/bcond.inflow.F90:48-52) through the
    folded kernels, one step (tblock 1) and two steps (tblock 2) per pass:
    interior equals the reference's NOSHIFT build"""
    print(f"case {name} noshift tblock={tb}", flush=True)
    s, ns = fixture_solver(name, tblock=tb)
    s.run(ns)
    got, ref = s.raw(), FIX[name + "/pops"]
    ok = same_bits(got[:, 1:-1, 1:-1], ref[:, 1:-1, 1:-1], s.prec)
    assert ok.all(), f"{int((~ok).sum())} differing elements"
    s.close()


CASES = [
    # (bc, obstacle, lx, ly, svisc, u0, steps)
    ("driven", 0, 131, 77, 0.05, 0.1, 40),
    ("periodic", 0, 96, 70, 0.02, 0.1, 40),
    ("channel", 0, 100, 33, 0.1, 0.08, 40),
    ("inflow", 1, 150, 64, 0.04, 0.1, 40),
    ("couette", 0, 70, 45, 0.08, 0.05, 40),
    ("step", 0, 120, 40, 0.1, 0.1, 30),
    ("driven", 1, 90, 66, 0.05, 0.1, 30),
    ("open", 1, 160, 64, 0.04, 0.05, 40),
    ("open", 0, 97, 45, 0.06, 0.05, 40),
]


@pytest.mark.parametrize("prec", ["single", "double", "fp16", "bf16", "mixed2"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}{'+obs' if c[1] else ''}")
def test_against_oracle(case, prec):
    bc, obst, lx, ly, sv, u0, ns = case
    print(f"case {bc} obs={obst} {lx}x{ly} {prec}", flush=True)
    scale = 8.0 if prec in ("fp16", "bf16") else 1.0
    o = po.Oracle(lx, ly, bc=bc, impl="fused", prec=ORC_PREC[prec], svisc=sv, u0=u0,
                  obstacle=obst, radius=9, scale=scale)
    o.step(ns)
    ref = o.raw()
    for impl, tb in (("fused", 1), ("fused", 2), ("original", 0)):
        s = Solver(lx, ly, bc=bc, impl=impl, prec=prec, svisc=sv, u0=u0, obstacle=obst,
                   radius=9, scale=scale, tblock=tb)
        s.run(ns)
        got = s.raw()
        inner = (slice(None), slice(1, -1), slice(1, -1))
        ok = same_bits(got[inner], ref[inner], prec)
        if bc == "step" and impl == "original":
            ok = ok[:, ly // 2:, :]  # FUSED reads never-written halos inside the step
        assert ok.all(), f"{impl}/tblock={tb}: {int((~ok).sum())} differing elements"
        if impl == "fused":
            d, dref = s.diagno(), o.diagno(mode=0)
            np.testing.assert_allclose(d, dref, rtol=1e-12, atol=1e-17)
            for got_p, ref_p in ((s.varm(), o.varm()), (s.prof_j(lx // 2), o.prof_j(lx // 2)),
                                 (s.prof_i(ly // 2), o.prof_i(ly // 2))):
                for g, r in zip(got_p, ref_p):
                    np.testing.assert_allclose(g, r, rtol=1e-12, atol=1e-17)
            np.testing.assert_array_equal(s.probe(lx // 2, ly // 2), o.probe(lx // 2, ly // 2))
        s.close()


OPTION_CASES = [
    # (bc, obstacle, prec, form, rounding, init_v)
    ("periodic", 0, "fp16", 1, 1, 1),
    ("periodic", 0, "bf16", 1, 1, 0),
    ("driven", 0, "fp16", 0, 1, 0),
    ("inflow", 1, "fp16", 1, 1, 0),
    ("channel", 0, "bf16", 0, 1, 0),
    ("couette", 0, "single", 1, 0, 0),
    ("periodic", 0, "double", 1, 0, 1),
    ("step", 0, "fp16", 1, 1, 0),
    ("open", 1, "fp16", 1, 1, 0),
    ("open", 0, "bf16", 1, 1, 0),
    ("inflow", 0, "bf16", 1, 1, 0),
    ("couette", 0, "fp16", 1, 1, 0),
]


@pytest.mark.parametrize("case", OPTION_CASES, ids=lambda c: "-".join(map(str, c)))
def test_options_against_oracle(case):
    """deviation form, stochastic FP16/BF16 rounding and the pressure-
    consistent Taylor-Green start, bit-exact against the oracle"""
    bc, obst, prec, form, rounding, init_v = case
    lx, ly, ns = 130, 70, 37
    o = po.Oracle(lx, ly, bc=bc, impl="fused", prec=ORC_PREC[prec], svisc=0.05, u0=0.05,
                  obstacle=obst, radius=9, scale=64.0, form=form, rounding=rounding,
                  init_v=init_v, amplitude=0.05)
    o.step(ns)
    ref = o.raw()
    for impl, tb in (("fused", 1), ("fused", 2), ("original", 0)):
        s = Solver(lx, ly, bc=bc, impl=impl, prec=prec, svisc=0.05, u0=0.05, obstacle=obst,
                   radius=9, scale=64.0, form=form, rounding=rounding, init_v=init_v,
                   amplitude=0.05, tblock=tb)
        s.run(ns)
        got = s.raw()
        ok = same_bits(got[:, 1:-1, 1:-1], ref[:, 1:-1, 1:-1], prec)
        if bc == "step" and impl == "original":
            ok = ok[:, ly // 2:, :]
        assert ok.all(), f"{impl}/tblock={tb}: {int((~ok).sum())} differing elements"
        s.close()


TB_CASES = [
    # (bc, obstacle, lx, ly): several strips of 256 columns, a partial last
    # strip, and heights that do not divide into the strip rows
    ("driven", 0, 600, 301),
    ("periodic", 0, 1030, 77),
    ("channel", 0, 513, 130),
    ("inflow", 1, 700, 200),
    ("couette", 0, 300, 97),
    ("driven", 1, 257, 513),
    ("open", 1, 700, 200),
]


@pytest.mark.parametrize("prec", ["single", "double", "fp16", "bf16", "mixed2"])
@pytest.mark.parametrize("case", TB_CASES, ids=lambda c: f"{c[0]}{'+obs' if c[1] else ''}-{c[2]}x{c[3]}")
def test_two_steps_per_pass_equal_one(case, prec):
    """k_fused_tb (two steps per HBM pass) against k_fused (one step per
    pass), bit for bit, through graph replays, an odd remainder step and
    diagnostics in between"""
    bc, obst, lx, ly = case
    print(f"case {bc} obs={obst} {lx}x{ly} {prec}", flush=True)
    kw = dict(bc=bc, prec=prec, svisc=0.03, u0=0.08, obstacle=obst, radius=11,
              scale=256.0 if prec in ("fp16", "bf16") else 1.0,
              form=1 if prec in ("fp16", "bf16") else 0,
              rounding=1 if prec in ("fp16", "bf16") else 0)
    a = Solver(lx, ly, tblock=1, **kw)
    b = Solver(lx, ly, tblock=2, **kw)
    for n in (67, 3, 2):
        a.run(n)
        b.run(n)
        np.testing.assert_allclose(a.diagno(), b.diagno(), rtol=0, atol=0)
        ok = same_bits(a.raw()[:, 1:-1, 1:-1], b.raw()[:, 1:-1, 1:-1], prec)
        assert ok.all(), f"after {n}: {int((~ok).sum())} differing elements"
    a.close()
    b.close()


def test_phases_equal_run_fused_driven_bitexact():
    """the folded kernel and the three-phase path agree bit for bit"""
    a = Solver(257, 129, bc="driven", impl="fused", svisc=0.02)
    b = Solver(257, 129, bc="driven", impl="fused", svisc=0.02)
    a.run(101)
    b.step_phases(101)
    assert np.array_equal(a.raw()[:, 1:-1, 1:-1], b.raw()[:, 1:-1, 1:-1])


def test_graph_replays_match_oracle_with_eager_work_between():
    """bgk_run replays HIP graphs of 32 steps; replays interleaved with
    diagnostics and eager steps stay bit-exact with the oracle"""
    for bc, impl in (("channel", "fused"), ("step", "fused"), ("driven", "original")):
        s = Solver(96, 40, bc=bc, impl=impl, svisc=0.05, u0=0.05)
        o = po.Oracle(96, 40, bc=bc, impl=impl, svisc=0.05, u0=0.05)
        for n in (64, 5, 33, 96, 1):
            s.run(n)
            o.step(n)
            s.diagno()
            got, ref = s.raw(), o.raw()
            ok = same_bits(got[:, 1:-1, 1:-1], ref[:, 1:-1, 1:-1], "single")
            assert ok.all(), f"{bc}/{impl} after {n}: {int((~ok).sum())} differ"


def test_upload_download_roundtrip_and_restart():
    """save_raw/restore_raw semantics: a state uploaded into a fresh
    context continues bit-identically"""
    a = Solver(64, 48, bc="periodic", svisc=0.05)
    a.run(20)
    snap = a.raw()
    a.run(15)
    b = Solver(64, 48, bc="periodic", svisc=0.05, init=False)
    b.set_raw(snap)
    assert np.array_equal(b.raw(), snap)
    b.run(15)
    assert np.array_equal(a.raw()[:, 1:-1, 1:-1], b.raw()[:, 1:-1, 1:-1])


def test_divergence_check():
    s = Solver(32, 32, bc="periodic", svisc=0.05)
    s.run(5)
    assert s.check(0.5)
    raw = s.raw()
    raw[2, 10, 10] = np.nan
    s.set_raw(raw)
    assert not s.check(0.5)


def test_invalid_arguments_fail_loudly():
    from bgk2d_gpu_amd import BgkError
    with pytest.raises(BgkError):
        Solver(1, 10)
    with pytest.raises(BgkError):
        Solver(10, 10, bc="step")  # stepx = l/20 = 0
    with pytest.raises(BgkError):
        Solver(32, 32, prec="fp16", scale=3.0)  # scale must be a power of two
    s = Solver(40, 40, init=False)
    with pytest.raises(BgkError):
        s.run(1)  # before init
    mask = np.zeros((40, 40), dtype=np.int32)
    mask[0, 5] = 1  # touches the first interior row
    with pytest.raises(BgkError):
        s.set_obstacle(mask)


def test_cavity_re100_256_against_reference_profile_and_ghia():
    """256^2, Re=100 (nu=0.256), 50,000 steps: the configuration of
    UTIL/bgk.256.R100.input; compare the centre-line profile with the
    reference-held TEST/DrivenCavity/prof_j.0000.dat last block and Ghia"""
    s = Solver(256, 256, bc="driven", impl="fused", prec="single", svisc=0.256, u0=0.1)
    s.run(50000)
    u, v, r = s.prof_j(128)
    ref = np.loadtxt(os.path.join(HERE, "golden", "prof_j_R100_256_t50000.dat"), comments="#")
    np.testing.assert_allclose(u, ref[:, 1], atol=2e-3 * 0.1 + 1e-4)
    gy = np.loadtxt(os.path.join(HERE, "golden", "ghia_y.dat"))
    um = 0.5 * (s.prof_j(128)[0] + s.prof_j(129)[0]) / 0.1
    y = (np.arange(1, 257) - 0.5) / 256
    sel = (gy[:, 0] > 0.02) & (gy[:, 0] < 0.98)
    assert np.max(np.abs(np.interp(gy[sel, 0], y, um) - gy[sel, 1])) < 0.015
    d = s.diagno()
    assert abs(d[2] - 1.0) < 1e-4


@pytest.mark.gpu
def test_c1_cavity_256_original_matches_reference_builds():
    """C1: lid-driven cavity 256^2, ORIGINAL stream + collide, Re = 100
    (svisc = 0.256, u0 = 0.1), 10,000 steps on the per-phase kernels: the
    interior populations carry the SHA-256 of the reference's own SINGLE
    and DOUBLE builds (tests/golden/make_cavity_reference.py), and FP32 differs from FP64
    by Linf 7.4e-7 in velocity (stated tolerance 1e-5), relative L2 1.4e-5"""
    import hashlib
    import json
    with open(os.path.join(HERE, "golden", "c1_reference.json")) as f:
        ref = json.load(f)
    cx = np.array([1, 1, 1, 0, -1, -1, -1, 0, 0])
    cy = np.array([-1, 1, 0, 1, -1, 1, 0, -1, 0])
    pops = {}
    for prec in ("single", "double"):
        s = Solver(256, 256, bc="driven", impl="original", prec=prec, svisc=0.256, u0=0.1)
        for _ in range(10000):
            s.boundaries()
            s.propagation()
            s.collision()
        raw = s.raw()
        s.close()
        assert hashlib.sha256(np.ascontiguousarray(raw[:, 1:-1, 1:-1]).tobytes()).hexdigest() == \
            ref["sha256_" + prec], prec
        pops[prec] = raw

    def fields(p):
        f = p[:, 1:-1, 1:-1].astype(np.float64)
        rho = f.sum(0) + 1.0
        return rho, (cx[:, None, None] * f).sum(0) / rho, (cy[:, None, None] * f).sum(0) / rho
    ra, ua, va = fields(pops["single"])
    rb, ub, vb = fields(pops["double"])
    du = np.concatenate([(ua - ub).ravel(), (va - vb).ravel()])
    linf = np.abs(du).max()
    rel = np.linalg.norm(du) / np.linalg.norm(np.concatenate([ub.ravel(), vb.ravel()]))
    assert linf == pytest.approx(ref["single_vs_double"]["uv_linf"], rel=1e-9) and linf < 1e-5
    assert rel == pytest.approx(ref["single_vs_double"]["uv_rel_l2"], rel=1e-9)


def _reference_digest_case(name, bc, n, m, svisc, u0, steps, precs):
    """run the HIP path (bgk_run, default kernels) for each storage kind
    and compare the interior's SHA-256 with the reference build's; returns
    the populations by kind"""
    import hashlib
    import json
    with open(os.path.join(HERE, "golden", f"{name}_reference.json")) as f:
        ref = json.load(f)
    pops = {}
    for prec in precs:
        print(f"case {name} {prec}", flush=True)
        s = Solver(n, m, bc=bc, impl="fused", prec=prec, svisc=svisc, u0=u0)
        s.run(steps)
        raw = s.raw()
        s.close()
        got = hashlib.sha256(np.ascontiguousarray(raw[:, 1:-1, 1:-1]).tobytes()).hexdigest()
        assert got == ref["sha256_" + prec], (name, prec)
        pops[prec] = raw
    return ref, pops


def _uv_diff(a, b):
    cx = np.array([1, 1, 1, 0, -1, -1, -1, 0, 0])
    cy = np.array([-1, 1, 0, 1, -1, 1, 0, -1, 0])
    out = []
    for p in (a, b):
        f = p[:, 1:-1, 1:-1].astype(np.float64)
        rho = f.sum(0) + 1.0
        out.append(((cx[:, None, None] * f).sum(0) / rho, (cy[:, None, None] * f).sum(0) / rho))
    du = np.concatenate([(out[0][0] - out[1][0]).ravel(), (out[0][1] - out[1][1]).ravel()])
    ref = np.concatenate([out[1][0].ravel(), out[1][1].ravel()])
    return np.abs(du).max(), np.linalg.norm(du) / np.linalg.norm(ref)


@pytest.mark.gpu
def test_c2_cavity_4096_fused_matches_reference_builds():
    """C2 size: lid-driven cavity 4096^2, FUSED, Re = 1e4 (svisc = 0.04096,
    u0 = 0.1), 2,000 steps through bgk_run (the two-step kernel): the
    interior populations carry the SHA-256 of the reference's own SINGLE
    and DOUBLE builds (tests/golden/make_cavity_reference.py c2short); FP32
    against FP64 as recorded there"""
    ref, pops = _reference_digest_case("c2short", "driven", 4096, 4096, 0.04096, 0.1, 2000,
                                       ("single", "double"))
    linf, rel = _uv_diff(pops["single"], pops["double"])
    assert linf == pytest.approx(ref["single_vs_double"]["uv_linf"], rel=1e-9)
    assert rel == pytest.approx(ref["single_vs_double"]["uv_rel_l2"], rel=1e-9)


@pytest.mark.gpu
def test_c2_cavity_4096_full_length_matches_reference_builds():
    """C2 at its full length: lid-driven cavity 4096^2, FUSED, Re = 1e4,
    20,000 steps through bgk_run: the interior populations carry the
    SHA-256 of the reference's own SINGLE and DOUBLE builds after 20,000
    steps (tests/golden/make_cavity_reference.py c2, about four hours of
    CPU for the two builds); FP32 against FP64 as recorded there"""
    ref, pops = _reference_digest_case("c2", "driven", 4096, 4096, 0.04096, 0.1, 20000,
                                       ("single", "double"))
    linf, rel = _uv_diff(pops["single"], pops["double"])
    assert linf == pytest.approx(ref["single_vs_double"]["uv_linf"], rel=1e-9)
    assert rel == pytest.approx(ref["single_vs_double"]["uv_rel_l2"], rel=1e-9)


@pytest.mark.gpu
def test_c3_tgv_8192_fused_matches_reference_build():
    """C3 size: Taylor-Green 8192^2, FUSED, tau = 0.8 (svisc = 0.1), the
    reference's initial field (amplitude 0.1), 5,000 steps through bgk_run:
    the interior populations carry the SHA-256 of the reference's own SINGLE
    build (tests/golden/make_cavity_reference.py c3)"""
    _reference_digest_case("c3", "periodic", 8192, 8192, 0.1, 0.1, 5000, ("single",))


@pytest.mark.gpu
def test_half_storage_tgv_1024_matches_reference_mixed1_build():
    """FP16 storage at scale 1 in the reference's operation order (the
    MIXED1 build: real*2 populations, real*4 arithmetic): Taylor-Green
    1024^2, tau = 0.8, amplitude 0.1, 500 steps through bgk_run gives the
    SHA-256 of the reference's MIXED1 build (make_cavity_reference.py
    c3half)"""
    import hashlib
    import json
    with open(os.path.join(HERE, "golden", "c3half_reference.json")) as f:
        ref = json.load(f)
    s = Solver(1024, 1024, bc="periodic", impl="fused", prec="fp16", svisc=0.1, u0=0.1, mixed1_quirk=1)
    s.run(500)
    raw = s.raw()
    s.close()
    assert hashlib.sha256(np.ascontiguousarray(raw[:, 1:-1, 1:-1]).tobytes()).hexdigest() == ref["sha256_mixed1"]


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["single", "fp16"])
@pytest.mark.parametrize("bc", ["periodic", "channel", "inflow"])
@pytest.mark.parametrize("edge", ["bottom", "top", "left", "right"])
def test_two_step_obstacle_at_periodic_edges(bc, edge, prec):
    """a small obstacle on the second or second-to-last row or column: along
    a periodic direction the strip blocks at the far edge update wrapped
    nodes next to it, so they too apply the obstacle rule; the result
    equals the one-step kernel bit for bit (fp32: 256-column strips,
    16-bit kinds: 512)"""
    l, m = 600, 260
    mask = np.zeros((m, l), dtype=np.int32)
    if edge == "bottom":
        mask[1:4, 300:306] = 1      # rows 2..4
    elif edge == "top":
        mask[m - 4:m - 1, 100:106] = 1  # rows m-3..m-1
    elif edge == "left":
        mask[120:126, 1:4] = 1      # columns 2..4
    else:
        mask[40:46, l - 4:l - 1] = 1  # columns l-3..l-1
    out = []
    narrow = prec == "fp16"
    for tb in (1, 2):
        print(f"case {bc} {edge} {prec} tblock={tb}", flush=True)
        s = Solver(l, m, bc=bc, impl="fused", prec=prec, svisc=0.05, u0=0.05, tblock=tb,
                   scale=256.0 if narrow else 1.0, form=int(narrow), rounding=int(narrow))
        s.set_obstacle(mask)
        s.run(41)
        out.append(s.raw())
        s.close()
    a, b = out
    u = np.uint16 if narrow else np.uint32
    assert np.array_equal(a[:, 1:-1, 1:-1].view(u), b[:, 1:-1, 1:-1].view(u))

"""Pin the oracle against the reference's own code.

tests/golden/ref_fixtures.npz holds, for 50 build variants (precision x
ORIGINAL/FUSED x LDC/TGV/POF/VKS/COU/BFS, with the LES, TRICK1, SQUARE,
HALF_CIRCLE, FORCING_Y, NOSHIFT and DRAG flags), the final populations and
text diagnostics written by the reference's hot-path Fortran sources
compiled unchanged (oracle/Makefile, tests/golden/make_golden.py), the
build_obs masks of the newer obstacle variants and draglift's output.  The oracle must
reproduce every stored population bit for bit, halos included.
"""
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "ref_fixtures.npz"))
VARIANTS = sorted({k.split("/")[0] for k in FIX.files})
PREC = {"SINGLE": "single", "DOUBLE": "double", "MIXED1": "mixed1", "MIXED2": "mixed2"}
UINT = {np.dtype(np.float64): np.uint64, np.dtype(np.float32): np.uint32,
        np.dtype(np.uint16): np.uint16}


def run_oracle(name):
    p, impl, case = name.split("@")[0].split("_")
    lx, ly, ns, rad = (int(x) for x in FIX[name + "/params"])
    sv, u0 = (float(x) for x in FIX[name + "/fparams"])
    bc, obst = po.CASE[case]
    o = po.Oracle(lx, ly, bc=bc, impl="fused" if impl == "FUSED" else "original",
                  prec=PREC[p], svisc=sv, u0=u0, obstacle=obst, radius=rad,
                  mixed1_quirk=1 if p == "MIXED1" else 0, les=int(case in po.LES_CASES),
                  cte_s=po.CTE_S_REF, trick1=int(case in po.TRICK1_CASES),
                  forcing_y=int(case in po.FORCING_Y_CASES), noshift=int(case in po.NOSHIFT_CASES))
    o.step(ns)
    return o


def test_fixture_inventory():
    assert len({v.split("@")[0] for v in VARIANTS}) == 50
    cases = {v.split("@")[0].split("_")[2] for v in VARIANTS}
    assert cases == {"LDC", "TGV", "POF", "VKS", "COU", "BFS", "LDCLES", "VKSLES", "LDCT1",
                     "VKSSQ", "VKSHC", "POFY", "LDCNS", "TGVNS", "POFNS", "VKSNS", "BFSNS",
                     "COUNS", "VKSD"}


@pytest.mark.parametrize("name", [v for v in VARIANTS if FIX.__contains__(v + "/obs")])
def test_obstacle_mask_of_build_obs(name):
    """the mask the oracle builds equals the one the reference's build_obs made"""
    o = run_oracle(name)
    mask, _ = o.obstacle()
    np.testing.assert_array_equal(mask, FIX[name + "/obs"])


@pytest.mark.parametrize("name", [v for v in VARIANTS if FIX.__contains__(v + "/drag")])
def test_draglift_against_reference(name):
    """draglift (SRC/draglift.F90:55-101) of the last step: the reference
    sums 2 c_k a_k over the fluid-solid links of the box (cylinder centre
    (2l/5, m/2) +- (radius + 3)) serially in its compute kind and prints
    force / (u_inflow^2 radius) with u_inflow = 0.1 (a default-real
    literal); the oracle sums the same per-link terms in long double.
    SINGLE: relative 2e-6 (the reference's fp32 sum); DOUBLE: 1e-12.  The
    lift is at the rounding level of the drag (symmetric flow)."""
    o = run_oracle(name)
    lx, ly, ns, rad = (int(x) for x in FIX[name + "/params"])
    ic, jc = 2 * lx // 5, ly // 2
    fx, fy = o.draglift((ic - rad - 3, ic + rad + 3, jc - rad - 3, jc + rad + 3))
    u_inflow = float(np.float32(0.1))
    norm = 1.0 / (u_inflow * u_inflow * rad)
    step, cd, cl = FIX[name + "/drag"][-1]
    assert step == ns
    tol = 2e-6 if name.startswith("SINGLE") else 1e-12
    assert abs(fx * norm - cd) <= tol * abs(cd)
    assert abs(fy * norm - cl) <= tol * abs(cd)


@pytest.mark.parametrize("name", VARIANTS)
def test_populations_bitexact(name):
    o = run_oracle(name)
    got, ref = o.raw(), FIX[name + "/pops"]
    assert got.dtype == ref.dtype and got.shape == ref.shape
    u = UINT[got.dtype]
    diff = got.view(u) != ref.view(u)
    assert not diff.any(), f"{int(diff.sum())} differing elements"


@pytest.mark.parametrize("name", VARIANTS)
def test_constants(name):
    p, impl, case = name.split("@")[0].split("_")
    lx, ly, ns, rad = (int(x) for x in FIX[name + "/params"])
    sv, u0 = (float(x) for x in FIX[name + "/fparams"])
    bc, obst = po.CASE[case]
    o = po.Oracle(lx, ly, bc=bc, prec=PREC[p], svisc=sv, u0=u0, obstacle=obst, radius=rad,
                  mixed1_quirk=1 if p == "MIXED1" else 0)
    c = o.constants()
    omega, fgrad, u00 = FIX[name + "/consts"]
    assert c["fgrad"] == fgrad and c["u00"] == u00
    # an LES build overwrites the module variable omega node by node
    # (SRC/col_MC.F90:229), so its final value is the last node's rate
    if case not in po.LES_CASES:
        assert c["omega"] == omega


@pytest.mark.parametrize("name", VARIANTS)
def test_diagnostics_text(name):
    """diagno (reference-order fp accumulation) against the reference's
    printed e14.6 values; profiles and probe against its text output."""
    o = run_oracle(name)
    ref = FIX[name + "/diagno"][-1]
    got = o.diagno(mode=1)
    np.testing.assert_allclose(got, ref[1:], rtol=6e-6, atol=1e-30)
    if name.startswith("MIXED"):
        # varm/prof_*/probe in the reference add stored values in the storage
        # kind (half for MIXED1, sp for MIXED2) before promoting; the oracle
        # and the HIP path do every per-node sum in the compute kind.
        return
    lx, ly = (int(x) for x in FIX[name + "/params"][:2])
    u, v, r = o.varm()
    vm = FIX[name + "/varm"]
    # row sums: fp32 serial accumulation in the reference vs long double here
    atol = 3e-8 if name.startswith("SINGLE") else 1e-15
    np.testing.assert_allclose(np.c_[u, v, r], vm[-ly:, 1:], rtol=6e-6, atol=atol)
    u, v, r = o.prof_j(lx // 2)
    pj = FIX[name + "/prof_j"]
    np.testing.assert_allclose(np.c_[u, v, r], pj[-ly:, 1:], rtol=6e-6, atol=1e-30)
    u, v, r = o.prof_i(ly // 2)
    pi = FIX[name + "/prof_i"]
    np.testing.assert_allclose(np.c_[u, v, r], pi[-lx:, 1:], rtol=6e-6, atol=1e-30)
    pr = FIX[name + "/probe"][-1]
    np.testing.assert_allclose(o.probe(lx // 2, ly // 2), pr[1:], rtol=6e-6, atol=1e-30)


def _c1():
    import json
    with open(os.path.join(HERE, "golden", "c1_reference.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("prec", ["single", "double"])
def test_c1_cavity_256_original_matches_reference_builds(prec):
    """C1: lid-driven cavity 256^2, ORIGINAL, Re = 100, 10,000 steps: the
    oracle's interior populations carry the same SHA-256 as the reference's
    own SINGLE / DOUBLE builds (tests/golden/make_cavity_reference.py)"""
    import hashlib
    ref = _c1()
    o = po.Oracle(256, 256, bc="driven", impl="original", prec=prec, svisc=0.256, u0=0.1)
    o.step(10000)
    got = hashlib.sha256(np.ascontiguousarray(o.raw()[:, 1:-1, 1:-1]).tobytes()).hexdigest()
    assert got == ref["sha256_" + prec]

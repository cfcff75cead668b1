"""Pin the oracle with what the mathematics fixes, independently of the
reference's operation order: storage rounding against numpy, BGK collision
against an independent numpy D2Q9 implementation, streaming as a
permutation, conservation laws, equilibrium fixed point, and closed-form
flows (Taylor-Green decay, Poiseuille profile, Ghia cavity, cylinder
symmetry)."""
import os

import numpy as np
import pytest

import pyoracle as po

HERE = os.path.dirname(os.path.abspath(__file__))
# D2Q9 in the oracle's order 01 03 05 08 10 12 14 17 19
CX = np.array([1, 1, 1, 0, -1, -1, -1, 0, 0])
CY = np.array([-1, 1, 0, 1, -1, 1, 0, -1, 0])
W = np.array([1 / 36, 1 / 36, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 9, 1 / 9, 4 / 9])
OPP = np.array([5, 4, 6, 7, 1, 0, 2, 3, 8])


def feq_shifted(rho, ux, uy):
    """textbook D2Q9 equilibrium minus the weights, shape (9, ...)"""
    cu = CX[:, None, None] * ux + CY[:, None, None] * uy
    usq = ux * ux + uy * uy
    return W[:, None, None] * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * usq) - W[:, None, None]


def moments(f):
    rho = f.sum(0) + 1.0
    ux = (CX[:, None, None] * f).sum(0) / rho
    uy = (CY[:, None, None] * f).sum(0) / rho
    return rho, ux, uy


# ----------------------------------------------------------- rounding

def test_f16_decode_exhaustive():
    L = po.lib()
    bits = np.arange(65536, dtype=np.uint16)
    ref = bits.view(np.float16).astype(np.float32)
    got = np.array([L.orc_f16_to_f32(int(b)) for b in bits], dtype=np.float32)
    nan = np.isnan(ref)
    assert np.array_equal(np.isnan(got), nan)
    assert np.array_equal(got[~nan].view(np.uint32), ref[~nan].view(np.uint32))


def test_f16_encode_against_numpy():
    L = po.lib()
    rng = np.random.default_rng(7)
    vals = np.concatenate([
        rng.standard_normal(4000).astype(np.float32) * np.float32(1e-3),
        (rng.random(2000) * 7e4).astype(np.float32),
        (rng.standard_normal(2000) * 1e-6).astype(np.float32),   # subnormal halves
        np.array([0.0, -0.0, 65504, 65519.99, 65520, 65536, 6.1035156e-05, 5.96e-08,
                  2.98e-08, 2.99e-08, np.inf, -np.inf, 2147483647.0], dtype=np.float32),
        # exact ties between adjacent halves
        (np.arange(1, 200, dtype=np.float32) * np.float32(2 ** -11) + np.float32(1.0)),
    ])
    with np.errstate(over="ignore"):
        ref = vals.astype(np.float16).view(np.uint16)
    got = np.array([L.orc_f32_to_f16(float(v)) for v in vals], dtype=np.uint16)
    assert np.array_equal(got, ref)


def test_bf16_round_nearest_even():
    L = po.lib()
    rng = np.random.default_rng(3)
    vals = np.concatenate([rng.standard_normal(5000).astype(np.float32) * np.float32(1e-2),
                           # ties: low 16 bits exactly 0x8000
                           (rng.integers(0x3c000000, 0x40000000, 500, dtype=np.uint32)
                            & np.uint32(0xFFFF0000) | np.uint32(0x8000)).view(np.float32)])
    got = np.array([L.orc_f32_to_bf16(float(v)) for v in vals], dtype=np.uint16)
    # independent rule: pick the nearer of the two neighbouring bf16 values
    b = vals.view(np.uint32)
    lo = (b >> 16).astype(np.uint16)
    hi = lo + np.uint16(1)
    flo = (lo.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    fhi = (hi.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    x = vals.astype(np.float64)
    dlo, dhi = np.abs(x - flo), np.abs(fhi - x)
    pick_hi = (dhi < dlo) | ((dhi == dlo) & (lo % 2 == 1))
    ref = np.where(pick_hi, hi, lo)
    assert np.array_equal(got, ref)
    back = np.array([L.orc_bf16_to_f32(int(g)) for g in got[:10]], dtype=np.float32)
    assert np.array_equal(back.view(np.uint32), got[:10].astype(np.uint32) << 16)


def test_sr_conversions_match_the_hardware_table():
    """the oracle's stochastic-rounding rule against gfx950's own
    v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32 on 53 k recorded conversions
    (tests/golden/make_sr_table.py): half subnormals, the normal range,
    overflow, binary32 subnormals, inf and NaN, edge and random words"""
    L = po.lib()
    t = np.load(os.path.join(HERE, "golden", "sr_convert.npz"))
    f16 = np.array([L.orc_sr_f16(int(x), int(r)) for x, r in zip(t["x"], t["r"])], dtype=np.uint32)
    bf16 = np.array([L.orc_sr_bf16(int(x), int(r)) for x, r in zip(t["x"], t["r"])], dtype=np.uint32)
    assert len(t["x"]) > 50000
    assert np.array_equal(f16, t["f16"])
    assert np.array_equal(bf16, t["bf16"])


def test_sr_conversions_are_unbiased():
    """averaged over every value of the offset bits the rule uses (the top
    13 bits of the word for FP16, 16 for BF16), the rounded value equals
    the input exactly in the normal range: a wrong shift, mask or
    truncation direction fails this"""
    L = po.lib()
    rng = np.random.default_rng(11)
    xs = np.concatenate([(rng.standard_normal(6) * 100).astype(np.float32),
                         np.array([1.0 + 2 ** -20, -3.14159, 65000.5, 2 ** -14 * 1.37], dtype=np.float32)])
    k16 = np.arange(1 << 13, dtype=np.uint64) << np.uint64(19)
    kb = np.arange(1 << 16, dtype=np.uint64) << np.uint64(16)
    for x in xs:
        b = int(np.float32(x).view(np.uint32))
        h = np.array([L.orc_sr_f16(b, int(r)) for r in k16], dtype=np.uint16)
        assert np.float64(h.view(np.float16).astype(np.float64).mean()) == np.float64(x)
        assert len(np.unique(h)) <= 2
        g = np.array([L.orc_sr_bf16(b, int(r)) for r in kb[::64]], dtype=np.uint32)
        m = (g << 16).view(np.float32).astype(np.float64).mean()
        # every 64th offset: the mean is exact up to 64 units of the dropped field
        assert abs(m - float(x)) <= 64 * 2.0 ** -16 * abs(float(x)) * 2 ** -7 + 1e-300


# ----------------------------------------------------------- collision

def _random_state(o, amp=0.02, seed=0):
    rng = np.random.default_rng(seed)
    m2, l2 = o.plane_shape
    rho = 1 + amp * rng.standard_normal((m2, l2))
    ux = amp * rng.standard_normal((m2, l2))
    uy = amp * rng.standard_normal((m2, l2))
    f = feq_shifted(rho, ux, uy) + 1e-3 * amp * rng.standard_normal((9, m2, l2))
    o.set_raw(f)
    return f


def _pull_periodic(f):
    """streaming on a periodic interior: f_k(x) <- f_k(x - c_k)"""
    g = np.empty_like(f)
    for k in range(9):
        g[k] = np.roll(np.roll(f[k], CX[k], axis=1), CY[k], axis=0)
    return g


def test_collision_matches_textbook_bgk():
    """one FUSED periodic step = pull + BGK relaxation to the textbook
    equilibrium; checked on the interior with an independent numpy code"""
    l, m = 12, 10
    o = po.Oracle(l, m, bc="periodic", impl="fused", prec="double", svisc=0.1, tgv=0)
    f0 = _random_state(o)
    omega = o.constants()["omega"]
    assert omega == pytest.approx(2 / (6 * 0.1 + 1), rel=1e-15)
    inner = f0[:, 1:-1, 1:-1]
    pulled = _pull_periodic(inner)
    rho, ux, uy = moments(pulled)
    expect = pulled - omega * (pulled - feq_shifted(rho, ux, uy))
    o.step(1)
    got = o.pops()[:, 1:-1, 1:-1]
    np.testing.assert_allclose(got, expect, rtol=0, atol=2e-15)


def test_omega_one_gives_equilibrium():
    l, m = 10, 8
    o = po.Oracle(l, m, bc="periodic", impl="original", prec="double", svisc=1 / 6, tgv=0)
    assert o.constants()["omega"] == 1.0
    f0 = _random_state(o, seed=5)
    pulled = _pull_periodic(f0[:, 1:-1, 1:-1])
    o.step(1)
    np.testing.assert_allclose(o.pops()[:, 1:-1, 1:-1], feq_shifted(*moments(pulled)),
                               rtol=0, atol=2e-15)


@pytest.mark.parametrize("forcing_y", [0, 1])
def test_channel_force_adds_momentum_not_mass(forcing_y):
    """CHANNEL forcing adds 6*fgrad*rho of momentum per node along x (along
    y in the FORCING_Y build), none across it and no mass
    (SRC/col.F90:210-237); checked on a uniform rest state."""
    l, m = 8, 6
    o = po.Oracle(l, m, bc="channel", impl="fused", prec="double", svisc=0.1, u0=0.05,
                  forcing_y=forcing_y)
    fgrad = o.constants()["fgrad"]
    assert fgrad == pytest.approx(4 * 0.05 * 0.1 / (3 * m * m), rel=1e-15)
    o.step(1)
    f = o.pops()[:, 2:-2, 2:-2]  # away from the walls
    rho, ux, uy = moments(f)
    along, across = (uy, ux) if forcing_y else (ux, uy)
    np.testing.assert_allclose(rho, 1.0, atol=1e-15)
    np.testing.assert_allclose(along * rho, 6 * fgrad, rtol=1e-12)
    np.testing.assert_allclose(across, 0.0, atol=1e-18)


def test_equilibrium_is_fixed_point():
    l, m = 10, 8
    o = po.Oracle(l, m, bc="periodic", impl="fused", prec="double", svisc=0.07, tgv=0)
    shape = o.plane_shape
    f = feq_shifted(np.full(shape, 1.01), np.full(shape, 0.03), np.full(shape, -0.02))
    o.set_raw(f)
    o.step(5)
    np.testing.assert_allclose(o.pops()[:, 1:-1, 1:-1], f[:, 1:-1, 1:-1], atol=2e-16)


def test_movef_is_permutation():
    """ORIGINAL streaming is an exact copy: after boundaries+propagation the
    b set equals the pulled a set bit for bit (tiny grid brute force)."""
    l, m = 6, 5
    o = po.Oracle(l, m, bc="periodic", impl="original", prec="single", tgv=0)
    rng = np.random.default_rng(1)
    a = rng.standard_normal((9,) + o.plane_shape).astype(np.float32)
    o.set_raw(a)
    o.boundaries()
    o.propagation()
    a_bc = o.raw("a")
    b = o.raw("b")
    for k in range(8):
        for j in range(1, m + 1):
            for i in range(1, l + 1):
                assert b[k, j, i] == a_bc[k, j - CY[k], i - CX[k]]
        # periodic halos are exact copies of the opposite interior edge
        rolled = np.roll(np.roll(a[k, 1:-1, 1:-1], CX[k], axis=1), CY[k], axis=0)
        assert np.array_equal(rolled, b[k, 1:-1, 1:-1])


@pytest.mark.parametrize("bc", ["driven", "periodic", "channel", "inflow", "couette", "step"])
@pytest.mark.parametrize("prec", ["single", "double", "mixed1", "bf16"])
def test_original_equals_fused(bc, prec):
    """both implementations perform the same arithmetic (SRC/col.F90 vs
    SRC/col_MC.F90); FUSED+step differs only inside the solid step, where
    the reference reads never-written halos of the second buffer."""
    kw = dict(bc=bc, prec=prec, svisc=0.08, u0=0.05)
    a = po.Oracle(40, 20, impl="original", **kw)
    b = po.Oracle(40, 20, impl="fused", **kw)
    a.step(30)
    b.step(30)
    ra, rb = a.raw()[:, 1:-1, 1:-1], b.raw()[:, 1:-1, 1:-1]
    if bc == "step":
        stepx, stepy = 40 // 20, 20 // 2
        ra, rb = ra[:, stepy:, :], rb[:, stepy:, :]  # rows above the step
    assert np.array_equal(ra, rb)


# ---------------------------------------------------------- conservation

@pytest.mark.parametrize("bc", ["periodic", "driven", "couette"])
def test_mass_conserved(bc):
    o = po.Oracle(24, 20, bc=bc, impl="fused", prec="double", svisc=0.05, u0=0.1)
    m0 = o.pops()[:, 1:-1, 1:-1].sum()
    o.step(200)
    m1 = o.pops()[:, 1:-1, 1:-1].sum()
    assert abs(m1 - m0) < 1e-11


def test_obstacle_mass_and_symmetry():
    """uniform inflow past a vertically centred cylinder with periodic y:
    the flow is mirror-symmetric about j = m/2 (u even, v odd)."""
    l, m = 48, 24
    o = po.Oracle(l, m, bc="inflow", obstacle=1, radius=4, impl="fused",
                  prec="double", svisc=0.05)
    o.step(120)
    f = o.pops()
    rho, ux, uy = moments(f[:, 1:-1, 1:-1])
    j = np.arange(1, m + 1)
    mirror = (m - j) % m  # j -> m - j on the periodic ring, 0-based row index
    mirror_idx = np.where(mirror == 0, m, mirror) - 1
    obs, bbox = o.obs()
    fluid = obs == 0
    # solid nodes hold values that depend on bcond_obs's serial j-outer
    # order (SRC/bcond.obs.F90:50-59), which is not mirror symmetric
    np.testing.assert_allclose(ux[fluid], ux[mirror_idx, :][fluid], atol=1e-12)
    np.testing.assert_allclose(uy[fluid], -uy[mirror_idx, :][fluid], atol=1e-12)
    assert obs.sum() > 0 and bbox[0] == 2 * l // 5 - 4


# ---------------------------------------------------------- closed forms

def test_taylor_green_decay():
    """<u^2+v^2>(t) = (A^2/2) exp(-4 nu k^2 t), k = 2 pi / L
    (TEST/TaylorGreenVortex/plot.gnu:171 uses 0.005 = 0.1^2/2)."""
    L, nu, A, T = 48, 0.05, 0.01, 300
    o = po.Oracle(L, L, bc="periodic", impl="fused", prec="double", svisc=nu, amplitude=A)
    e0 = o.diagno()[3]
    assert e0 == pytest.approx(A * A / 2, rel=1e-10)
    k = 2 * np.pi / L
    o.step(T)
    e1 = o.diagno()[3]
    # includes the start-up transient of the equilibrium initial state
    assert e1 == pytest.approx(A * A / 2 * np.exp(-4 * nu * k * k * T), rel=1e-2)
    o.step(T)
    # decay rate over a later window: discretisation error O(k^2) only
    assert o.diagno()[3] / e1 == pytest.approx(np.exp(-4 * nu * k * k * T), rel=5e-4)


def test_poiseuille_profile():
    """steady channel flow under the body force reaches u(y) = 4 u0 y'(1-y'),
    y' = (j - 1/2)/m (TEST/PoiseuilleFlow/plot.gnu:172)"""
    l, m, u0 = 4, 16, 0.01
    o = po.Oracle(l, m, bc="channel", impl="fused", prec="double", svisc=1 / 6, u0=u0)
    o.step(12000)
    u, v, r = o.prof_j(l // 2)
    y = (np.arange(1, m + 1) - 0.5) / m
    np.testing.assert_allclose(u, 4 * u0 * y * (1 - y), rtol=0, atol=2e-3 * u0)
    assert o.diagno()[3] == pytest.approx(8 * u0 * u0 / 15, rel=2e-2)


@pytest.mark.slow
def test_cavity_ghia_re100():
    """Lid-driven cavity at Re = U N / nu = 100 against Ghia et al. 1982
    (TEST/DrivenCavity/ghia.{x,y}.dat) on the centre lines."""
    n, u0 = 64, 0.1
    o = po.Oracle(n, n, bc="driven", impl="fused", prec="single", svisc=u0 * n / 100, u0=u0)
    o.step(16000)
    gy = np.loadtxt(os.path.join(HERE, "golden", "ghia_y.dat"))
    gx = np.loadtxt(os.path.join(HERE, "golden", "ghia_x.dat"))
    # vertical centre line x = 1/2 lies between columns n/2 and n/2+1
    u = 0.5 * (o.prof_j(n // 2)[0] + o.prof_j(n // 2 + 1)[0]) / u0
    y = (np.arange(1, n + 1) - 0.5) / n
    v = 0.5 * (o.prof_i(n // 2)[1] + o.prof_i(n // 2 + 1)[1]) / u0
    x = (np.arange(1, n + 1) - 0.5) / n
    sel = (gy[:, 0] > 0.02) & (gy[:, 0] < 0.98)
    assert np.max(np.abs(np.interp(gy[sel, 0], y, u) - gy[sel, 1])) < 0.03
    sel = (gx[:, 0] > 0.02) & (gx[:, 0] < 0.98)
    assert np.max(np.abs(np.interp(gx[sel, 0], x, v) - gx[sel, 1])) < 0.03


# ---------------------------------------------- options beyond the reference

def test_deviation_form_is_textbook_bgk():
    """form 1 keeps the density deviation apart from 1; the step is still the
    textbook BGK update (checked in double against the numpy code)"""
    l, m = 12, 10
    o = po.Oracle(l, m, bc="periodic", impl="fused", prec="double", svisc=0.1, tgv=0, form=1)
    f0 = _random_state(o, seed=3)
    omega = o.constants()["omega"]
    pulled = _pull_periodic(f0[:, 1:-1, 1:-1])
    rho, ux, uy = moments(pulled)
    expect = pulled - omega * (pulled - feq_shifted(rho, ux, uy))
    o.step(1)
    np.testing.assert_allclose(o.pops()[:, 1:-1, 1:-1], expect, rtol=0, atol=2e-16)


def test_stochastic_rounding_brackets_and_is_unbiased():
    """the same fp32 collision results stored by nearest-even and by
    stochastic rounding differ by at most one half-precision step, and the
    stochastic errors average out over positions (unbiased)"""
    l, m = 64, 64
    shape = (m + 2, l + 2)
    rng = np.random.default_rng(11)
    feq = feq_shifted(1 + 1e-3 * rng.standard_normal(shape), 0.01 + 1e-3 * rng.standard_normal(shape),
                      1e-3 * rng.standard_normal(shape))
    raw_in = feq.astype(np.float32).astype(np.float16).view(np.uint16)
    out = {}
    for r in (0, 1):
        # omega = 1: the stored result is the equilibrium of the pulled state
        o = po.Oracle(l, m, bc="periodic", impl="fused", prec="mixed1", svisc=1 / 6,
                      rounding=r, tgv=0)
        o.set_raw(raw_in)
        o.step(1)
        out[r] = o.raw()[:, 1:-1, 1:-1].view(np.float16).astype(np.float64)
    rn, sr = out[0], out[1]
    r16 = rn.astype(np.float16)
    up = np.nextafter(r16, np.float16(np.inf)).astype(np.float64) - rn
    down = rn - np.nextafter(r16, np.float16(-np.inf)).astype(np.float64)
    ulp = np.maximum(up, down)  # the larger step at binade boundaries
    assert np.all(np.abs(sr - rn) <= ulp * 1.0001)
    assert (sr != rn).mean() > 0.2  # the rule does change values
    dec_in = raw_in.view(np.float16).astype(np.float64)
    rho, ux, uy = moments(_pull_periodic(dec_in[:, 1:-1, 1:-1]))
    exact = feq_shifted(rho, ux, uy)
    bias_sr = np.abs((sr - exact).mean(axis=(1, 2)))
    assert np.all(bias_sr < 0.05 * ulp.mean(axis=(1, 2)))


def test_pressure_consistent_taylor_green_matches_analytic():
    """init_v = 1: density of the vortex's pressure field; the solution then
    follows the decaying vortex without an acoustic transient (positions use
    the default-real pi of SRC/init.F90, hence the 1e-10 on the density)"""
    L, nu, A, T = 48, 0.05, 0.01, 200
    o = po.Oracle(L, L, bc="periodic", impl="fused", prec="double", svisc=nu, amplitude=A,
                  init_v=1)
    f = o.pops()[:, 1:-1, 1:-1]
    rho, ux, uy = moments(f)
    x = (np.arange(1, L + 1) - 0.5) / L
    k = 2 * np.pi
    np.testing.assert_allclose(rho - 1, 0.75 * A * A * (np.cos(2 * k * x)[None, :] +
                                                          np.cos(2 * k * x)[:, None]), atol=1e-10)
    o.step(T)
    rho, ux, uy = moments(o.pops()[:, 1:-1, 1:-1])
    dec = np.exp(-nu * 2 * (k / L) ** 2 * T)
    ua = A * np.sin(k * x)[None, :] * np.cos(k * x)[:, None] * dec
    err = np.sqrt(((ux - ua) ** 2).sum() / (ua ** 2).sum())
    # O(k^2) lattice error of the decay rate at L = 48 (exponent 0.34)
    assert err < 4e-3


# ------------------------------------------------ open channel, drag/lift

@pytest.mark.parametrize("prec", ["single", "double"])
@pytest.mark.parametrize("impl", ["fused", "original"])
def test_open_free_slip_equals_mirrored_periodic(prec, impl):
    """A free-slip (specular) wall at y = 1/2 is the mirror plane of a
    periodic domain twice as high holding the field and its mirror image
    (v -> -v: 01<->03, 08<->17, 10<->12).  One step of OPEN on (l, m) and of
    the reference-pinned PERIODIC variant on (l, 2m) agree bit for bit on
    every row, away from the columns the x boundaries touch."""
    l, m = 30, 12
    a = po.Oracle(l, m, bc="open", impl=impl, prec=prec, svisc=0.05, u0=0.04)
    f = _random_state(a, amp=0.03, seed=5)
    mir = np.array([1, 0, 2, 7, 5, 4, 6, 3, 8])
    g = np.zeros((9, 2 * m + 2, l + 2))
    g[:, 1:m + 1] = f[:, 1:m + 1]
    g[:, m + 1:2 * m + 1] = f[mir][:, m:0:-1]  # row m+r <- mirror of row m+1-r
    b = po.Oracle(l, 2 * m, bc="periodic", impl=impl, prec=prec, svisc=0.05)
    b.set_raw(g)
    a.set_raw(f)  # rounded to the storage kind the same way as g
    a.step(1)
    b.step(1)
    ra, rb = a.raw(), b.raw()
    assert np.array_equal(ra[:, 1:m + 1, 2:l], rb[:, 1:m + 1, 2:l])


def test_open_channel_steady_flux_and_profile():
    """Empty OPEN channel at steady state: the inlet column carries the
    parabola 4 umax y'(1-y') (to 1.5 %, away from the wall rows, where the
    free-slip flow does not vanish); the same mass flux through every
    column; zero x-gradient at the outflow; a profile symmetric about the
    centre line; and, the walls being free-slip (no shear), a nearly flat
    profile downstream with the wall rows moving like the centre (a no-slip
    wall would hold them back)."""
    l, m, u0 = 120, 24, 0.05
    o = po.Oracle(l, m, bc="open", impl="fused", prec="double", svisc=0.1, u0=u0)
    o.step(10000)
    f = o.pops()[:, 1:-1, 1:-1]
    rho, ux, uy = moments(f)
    y = (np.arange(1, m + 1) - 0.5) / m
    np.testing.assert_allclose(ux[2:-2, 0], (4 * u0 * y * (1 - y))[2:-2], rtol=1.5e-2)
    flux = (rho * ux).sum(0)
    np.testing.assert_allclose(flux, flux.mean(), rtol=1e-9)
    np.testing.assert_allclose(ux[:, -1], ux[:, -2], atol=1e-3 * u0)
    np.testing.assert_allclose(ux, ux[::-1], rtol=0, atol=1e-12)
    np.testing.assert_allclose(uy, -uy[::-1], rtol=0, atol=1e-12)
    assert np.argmax(ux[:, 0]) in (m // 2 - 1, m // 2) and np.all(np.diff(ux[:m // 2, 0]) > 0)
    mid = ux[:, l // 2]
    assert mid.std() < 0.02 * mid.mean()
    assert ux[0, l // 2] > 0.95 * ux[m // 2, l // 2]
    o.step(500)
    f2 = o.pops()[:, 1:-1, 1:-1]
    assert np.abs(f2 - f).max() < 1e-9  # steady


def test_draglift_square_block_closed_form():
    """Fluid nodes next to an s x s solid block, with a_k the same at every
    node: each axis direction has s links into the block, each diagonal
    2s - 1, so Fx = 2[(a01+a03-a10-a12)(2s-1) + (a05-a14) s] and
    Fy = 2[(a03+a12-a01-a10)(2s-1) + (a08-a17) s]."""
    l, m, s, x0, y0 = 20, 16, 4, 7, 6
    o = po.Oracle(l, m, bc="periodic", impl="fused", prec="double", svisc=0.1)
    mask = np.zeros((m, l), dtype=np.int32)
    mask[y0 - 1:y0 - 1 + s, x0 - 1:x0 - 1 + s] = 1
    o.set_obs(mask)
    a = np.array([0.3, -0.7, 1.1, 0.45, -0.2, 0.9, -1.3, 0.05, 0.6])
    o.set_raw(np.broadcast_to(a[:, None, None], (9, m + 2, l + 2)))
    fx, fy = o.draglift((1, l, 1, m))
    nd, na = 2 * s - 1, s
    a01, a03, a05, a08, a10, a12, a14, a17 = a[:8]
    assert fx == pytest.approx(2 * ((a01 + a03 - a10 - a12) * nd + (a05 - a14) * na), rel=1e-14)
    assert fy == pytest.approx(2 * ((a03 + a12 - a01 - a10) * nd + (a08 - a17) * na), rel=1e-14)
    # a box that misses the block's neighbours sees no force
    assert np.array_equal(o.draglift((15, 20, 1, 16)), [0.0, 0.0])


def test_open_symmetric_cylinder_has_no_lift():
    """a cylinder centred on the channel's mirror line in the symmetric
    OPEN inflow: the lift vanishes to round-off, the drag does not"""
    l, m, R = 80, 40, 5
    o = po.Oracle(l, m, bc="open", impl="fused", prec="double", svisc=0.05, u0=0.05)
    jj, ii = np.mgrid[1:m + 1, 1:l + 1]
    mask = ((ii - 25) ** 2 + (jj - (m + 1) / 2) ** 2 <= R * R).astype(np.int32)
    o.set_obs(mask)
    o.step(800)
    fx, fy = o.draglift((25 - R - 3, 25 + R + 3, 1, m))
    assert fx > 0
    assert abs(fy) < 1e-9 * fx


def test_les_without_smagorinsky_term_is_bgk():
    """cteS = 0 leaves Ts = 1/(2 omega) + (1/omega)/2 = 1/omega: the LES
    collision reduces to plain BGK (to round-off of that identity)"""
    kw = dict(bc="driven", impl="fused", prec="double", svisc=0.02, u0=0.1)
    a = po.Oracle(40, 30, **kw)
    b = po.Oracle(40, 30, les=1, cte_s=0.0, **kw)
    a.step(200)
    b.step(200)
    np.testing.assert_allclose(b.pops(), a.pops(), rtol=0, atol=1e-13)


def test_les_only_adds_dissipation():
    """the Smagorinsky viscosity is non-negative: a Taylor-Green vortex
    loses kinetic energy at least as fast with LES, and faster the larger
    cteS"""
    kw = dict(bc="periodic", impl="fused", prec="double", svisc=0.002, amplitude=0.05)
    e = []
    for les, cs in ((0, 0.0), (1, 0.1), (1, 0.3)):
        o = po.Oracle(48, 48, les=les, cte_s=cs, **kw)
        o.step(400)
        e.append(o.diagno()[3])
    assert e[0] > e[1] > e[2]

"""The Fortran driver (bgk2d_hip): bgk.input semantics, output formats and
save/restore, on top of the C ABI through ISO_C_BINDING."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as po
from bgk2d_gpu_amd import build

INPUT = """&parameters
lx = {lx}
ly = {ly}
svisc={svisc}
u0={u0}
itfin={itfin}
ivtim=50000
isignal={isignal}
itsave={itsave}
icheck={icheck}
irestart={irestart}
init_v=0
ipad=0
radius={radius}        /
&gpu
case_name='{case}'
precision='{prec}'
implementation='{impl}'
phases={phases}
{extra}
/
"""


def write_input(d, **kw):
    args = dict(lx=64, ly=48, svisc=0.05, u0=0.1, itfin=300, isignal=100, itsave=300,
                icheck=100, irestart=0, case="LDC", prec="SINGLE", impl="FUSED", phases=0,
                radius=16, extra="")
    args.update(kw)
    with open(os.path.join(d, "bgk.input"), "w") as f:
        f.write(INPUT.format(**args))
    return args


def run_driver(d, timeout=120):
    return subprocess.run([build.DRIVER], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True, timeout=timeout)


def read_save(path, dtype, l, m):
    """Fortran unformatted sequential file: 4-byte length markers per record"""
    raw = open(path, "rb").read()
    pos, recs = 0, []
    while pos < len(raw):
        n = int(np.frombuffer(raw[pos:pos + 4], dtype=np.int32)[0])
        recs.append(raw[pos + 4:pos + 4 + n])
        pos += n + 8
    itime = int(np.frombuffer(recs[0], dtype=np.int32)[0])
    pops = np.stack([np.frombuffer(r, dtype=dtype).reshape(m + 2, l + 2) for r in recs[1:10]])
    return itime, pops


def test_driver_built_and_linked():
    assert os.path.exists(build.DRIVER)
    out = subprocess.run(["ldd", build.DRIVER], stdout=subprocess.PIPE, text=True).stdout
    assert "libbgk2d_hip.so" in out and "not found" not in out.split("libbgk2d_hip.so")[1].split("\n")[0]


def test_driver_rejects_unknown_case_before_gpu(tmp_path):
    write_input(str(tmp_path), case="XYZ")
    r = run_driver(str(tmp_path))
    assert r.returncode != 0 and "unknown case_name" in r.stdout


@pytest.mark.parametrize("kw", [dict(itsave=300), dict(itsave=1000, irestart=1)])
def test_driver_refuses_scaled_16bit_save_or_restore_before_gpu(tmp_path, kw):
    """the save/restore layout has no scale field: FP16/BF16 with scale != 1
    cannot write or read it"""
    write_input(str(tmp_path), prec="FP16", extra="scale=1024.d0", **kw)
    r = run_driver(str(tmp_path))
    assert r.returncode != 0 and "scale 1 only" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("phases", [0, 1])
@pytest.mark.parametrize("case,prec,impl,extra", [("LDC", "SINGLE", "FUSED", ""),
                                                  ("TGV", "DOUBLE", "ORIGINAL", ""),
                                                  ("VKS", "SINGLE", "FUSED", ""),
                                                  ("POF", "MIXED1", "FUSED", ""),
                                                  ("POF", "SINGLE", "FUSED", "forcing_y=1"),
                                                  ("LDC", "SINGLE", "FUSED", "noshift=1"),
                                                  ("VKS", "DOUBLE", "ORIGINAL", "noshift=1")])
def test_driver_matches_oracle(tmp_path, case, prec, impl, extra, phases):
    d = str(tmp_path)
    a = write_input(d, case=case, prec=prec, impl=impl, phases=phases, lx=80, ly=40,
                    svisc=0.05, itfin=300, itsave=300, extra=extra)
    r = run_driver(d)
    assert r.returncode == 0, r.stdout
    bc, obst = po.CASE[case]
    oprec = {"SINGLE": "single", "DOUBLE": "double", "MIXED1": "mixed1"}[prec]
    o = po.Oracle(80, 40, bc=bc, impl=impl.lower(), prec=oprec, svisc=0.05, u0=0.1,
                  obstacle=obst, radius=16 if not obst else 16, mixed1_quirk=int(prec == "MIXED1"),
                  forcing_y=int(extra == "forcing_y=1"), noshift=int(extra == "noshift=1"))
    o.step(300)
    it, pops = read_save(os.path.join(d, "save.00000300.bin"), po.STORE_DTYPE[oprec], 80, 40)
    assert it == 300
    ref = o.raw()
    inner = (slice(None), slice(1, -1), slice(1, -1))
    assert np.array_equal(pops[inner].view(np.uint8), ref[inner].view(np.uint8))
    if phases == 1:
        # the per-phase path reproduces the reference's halos too
        assert np.array_equal(pops.view(np.uint8), ref.view(np.uint8))
    diag = np.loadtxt(os.path.join(d, "diagno.dat"))
    np.testing.assert_allclose(diag[-1, 1:], o.diagno(), rtol=1e-5, atol=1e-12)
    assert diag[-1, 0] == 300
    perf = open(os.path.join(d, "bgk.perf")).read()
    assert "Mlups" in perf


@pytest.mark.gpu
def test_driver_restart_continues_identically(tmp_path):
    d1, d2 = str(tmp_path / "a"), str(tmp_path / "b")
    os.makedirs(d1)
    os.makedirs(d2)
    write_input(d1, itfin=400, itsave=200, icheck=200)
    assert run_driver(d1).returncode == 0
    os.symlink(os.path.join(d1, "save.00000200.bin"), os.path.join(d2, "restore.bin"))
    write_input(d2, itfin=400, itsave=200, icheck=200, irestart=1)
    r = run_driver(d2)
    assert r.returncode == 0, r.stdout
    _, a = read_save(os.path.join(d1, "save.00000400.bin"), np.float32, 64, 48)
    _, b = read_save(os.path.join(d2, "save.00000400.bin"), np.float32, 64, 48)
    assert np.array_equal(a[:, 1:-1, 1:-1], b[:, 1:-1, 1:-1])


@pytest.mark.gpu
@pytest.mark.parametrize("phases", [0, 1])
def test_driver_cylinder_channel_and_draglift(tmp_path, phases):
    """case CYL: open channel past a cylinder at a chosen centre; the saved
    populations equal the oracle's and drag.lift.dat holds F / (u0^2 R)"""
    d = str(tmp_path)
    l, m, r, xc, yc, u0 = 120, 48, 6, 30, 25, 0.05
    write_input(d, case="CYL", lx=l, ly=m, svisc=0.05, u0=u0, itfin=200, itsave=200,
                icheck=100, radius=r, phases=phases, extra=f"obs_x={xc}\nobs_y={yc}\ndrag=1")
    res = run_driver(d)
    assert res.returncode == 0, res.stdout
    o = po.Oracle(l, m, bc="open", impl="fused", prec="single", svisc=0.05, u0=u0)
    jj, ii = np.mgrid[1:m + 1, 1:l + 1]
    o.set_obs(((ii - xc) ** 2 + (jj - yc) ** 2 <= r * r).astype(np.int32))
    o.step(200)
    _, pops = read_save(os.path.join(d, "save.00000200.bin"), np.float32, l, m)
    ref = o.raw()
    assert np.array_equal(pops[:, 1:-1, 1:-1].view(np.uint32), ref[:, 1:-1, 1:-1].view(np.uint32))
    dl = np.loadtxt(os.path.join(d, "drag.lift.dat"))
    assert dl.shape == (2, 3) and dl[-1, 0] == 200
    f = o.draglift((xc - r - 3, xc + r + 3, yc - r - 3, yc + r + 3)) / (u0 * u0 * r)
    np.testing.assert_allclose(dl[-1, 1:], f, rtol=1e-7)


def read_vtk(path):
    """legacy VTK as the reference writes it: text lines, then big-endian
    FP32 blocks after each 'VECTORS'/'LOOKUP_TABLE' line"""
    raw = open(path, "rb").read()
    pos, header, blocks = 0, [], {}
    npts = None
    while pos < len(raw):
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode().strip()  # (Aw) right-justifies: leading blanks
        pos = end + 1
        header.append(line)
        if line.startswith("POINT_DATA"):
            npts = int(line.split()[1])
        if line.startswith("VECTORS"):
            n = 3 * npts
            blocks[line.split()[1]] = np.frombuffer(raw[pos:pos + 4 * n], dtype=">f4").reshape(npts, 3)
            pos += 4 * n
        elif line.startswith("LOOKUP_TABLE"):
            name = header[-2].split()[1]
            blocks[name] = np.frombuffer(raw[pos:pos + 4 * npts], dtype=">f4")
            pos += 4 * npts
    return header, blocks


@pytest.mark.gpu
def test_driver_vtk_files(tmp_path):
    """tec_xy / tec_om at every ivtim step, in the layout of
    SRC/vtk_xy.binary.F90 and SRC/vtk_om.binary.F90"""
    d = str(tmp_path)
    l, m = 40, 30
    with open(os.path.join(d, "bgk.input"), "w") as f:
        f.write(INPUT.format(lx=l, ly=m, svisc=0.05, u0=0.1, itfin=100, isignal=100, itsave=100,
                             icheck=100, irestart=0, case="LDC", prec="SINGLE", impl="FUSED",
                             phases=0, radius=16, extra="").replace("ivtim=50000", "ivtim=50"))
    r = run_driver(d)
    assert r.returncode == 0, r.stdout
    assert os.path.exists(os.path.join(d, "tec_xy.00000050.vtk"))
    hdr, b = read_vtk(os.path.join(d, "tec_xy.00000100.vtk"))
    assert hdr[:4] == ["# vtk DataFile Version 2.0", "Campo", "BINARY", "DATASET STRUCTURED_POINTS"]
    raw = open(os.path.join(d, "tec_xy.00000100.vtk"), "rb").read()
    # the reference's edit descriptors: (A23) and (A21) right-justify
    assert b"\n VECTORS velocity float\n" in raw and b"      SCALARS w float\nLOOKUP_TABLE" in raw
    assert hdr[4].split() == ["DIMENSIONS", str(l), str(m), "1"]
    assert hdr[7].split() == ["POINT_DATA", str(l * m)]
    o = po.Oracle(l, m, bc="driven", impl="fused", prec="single", svisc=0.05, u0=0.1)
    o.step(100)
    a = o.raw()[:, 1:-1, 1:-1].reshape(9, -1)
    us = a[0] + a[1] + a[2] - a[4] - a[5] - a[6]
    ws = a[1] + a[3] + a[5] - a[0] - a[4] - a[7]
    den = (a[0] + a[1] + a[2] + a[3] + a[4] + a[5] + a[6] + a[7] + a[8]) + np.float32(1)
    u, w = us / den, ws / den
    assert np.array_equal(b["velocity"][:, 0], u) and np.array_equal(b["velocity"][:, 1], w)
    assert not b["velocity"][:, 2].any()
    assert np.array_equal(b["w"], w) and np.array_equal(b["u"], u) and np.array_equal(b["rho"], den)
    hdr, b = read_vtk(os.path.join(d, "tec_om.00000100.vtk"))
    assert hdr[4].split() == ["DIMENSIONS", str(l - 2), str(m - 2), "1"]
    U, Wf = u.reshape(m, l), w.reshape(m, l)
    vort = (np.float32(0.5) * (U[2:, 1:-1] - U[:-2, 1:-1])
            - np.float32(0.5) * (Wf[1:-1, 2:] - Wf[1:-1, :-2]))
    assert np.array_equal(b["vorticity"], vort.ravel())
    assert b["stream"].shape == ((l - 2) * (m - 2),) and np.isfinite(b["stream"]).all()

"""Generate the golden fixtures from the reference's own code.

Runs each oracle/_ref/ref_<PREC>_<IMPL>_<CASE> binary (the reference's
hot-path Fortran sources compiled unchanged by oracle/Makefile, driven by
oracle/ref_harness.F90) on a small grid and stores, per variant, the inputs,
the final populations with halos (raw storage bits) and the text diagnostics
the reference itself writes (diagno.dat, u_med.dat, prof_i/j.dat, probe.dat
formats).  Only data is stored; no reference source.

    make -C oracle ref && python tests/golden/make_golden.py
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

# (variant, lx, ly, svisc, u0, nsteps, radius)
CASES = [
    ("SINGLE_ORIG_LDC", 24, 20, 0.1, 0.1, 60, 16),
    ("SINGLE_FUSED_LDC", 24, 20, 0.1, 0.1, 60, 16),
    ("DOUBLE_ORIG_LDC", 24, 20, 0.1, 0.1, 60, 16),
    ("DOUBLE_FUSED_LDC", 24, 20, 0.1, 0.1, 60, 16),
    ("MIXED1_FUSED_LDC", 24, 20, 0.1, 0.1, 60, 16),
    ("MIXED2_FUSED_LDC", 24, 20, 0.1, 0.1, 60, 16),
    ("SINGLE_FUSED_TGV", 24, 24, 0.05, 0.1, 40, 16),
    ("SINGLE_ORIG_TGV", 24, 24, 0.05, 0.1, 40, 16),
    ("DOUBLE_FUSED_TGV", 24, 24, 0.05, 0.1, 40, 16),
    ("MIXED1_FUSED_TGV", 24, 24, 0.05, 0.1, 40, 16),
    ("SINGLE_FUSED_POF", 32, 12, 0.1, 0.1, 80, 16),
    ("SINGLE_ORIG_POF", 32, 12, 0.1, 0.1, 80, 16),
    ("DOUBLE_FUSED_POF", 32, 12, 0.1, 0.1, 80, 16),
    ("SINGLE_FUSED_VKS", 40, 24, 0.05, 0.1, 60, 4),
    ("SINGLE_ORIG_VKS", 40, 24, 0.05, 0.1, 60, 4),
    ("DOUBLE_FUSED_VKS", 40, 24, 0.05, 0.1, 60, 4),
    ("SINGLE_FUSED_COU", 24, 16, 0.1, 0.1, 60, 16),
    ("SINGLE_ORIG_COU", 24, 16, 0.1, 0.1, 60, 16),
    ("SINGLE_FUSED_BFS", 60, 20, 0.1, 0.1, 50, 16),
    ("SINGLE_ORIG_BFS", 60, 20, 0.1, 0.1, 50, 16),
    # wider grids: non-square Taylor-Green tables, two wavefronts per row
    ("DOUBLE_FUSED_TGV@96x70", 96, 70, 0.02, 0.1, 3, 16),
    ("SINGLE_FUSED_TGV@96x70", 96, 70, 0.02, 0.1, 3, 16),
    ("MIXED1_FUSED_TGV@96x70", 96, 70, 0.02, 0.1, 3, 16),
    ("SINGLE_FUSED_LDC@130x66", 130, 66, 0.02, 0.1, 20, 16),
    ("SINGLE_FUSED_VKS@100x60", 100, 60, 0.04, 0.1, 30, 8),
    # Smagorinsky LES (-DLES, cteS = 0.1) at low viscosity
    ("SINGLE_FUSED_LDCLES", 24, 20, 0.005, 0.1, 60, 16),
    ("SINGLE_ORIG_LDCLES", 24, 20, 0.005, 0.1, 60, 16),
    ("DOUBLE_FUSED_LDCLES", 24, 20, 0.005, 0.1, 60, 16),
    ("MIXED2_FUSED_LDCLES", 24, 20, 0.005, 0.1, 60, 16),
    ("SINGLE_FUSED_VKSLES", 40, 24, 0.005, 0.1, 60, 4),
    # TRICK1 corner order of bcond_driven (square cavity), SQUARE and
    # HALF_CIRCLE obstacles of build_obs
    ("SINGLE_FUSED_LDCT1", 24, 24, 0.1, 0.1, 60, 16),
    ("SINGLE_ORIG_LDCT1", 24, 24, 0.1, 0.1, 60, 16),
    ("DOUBLE_FUSED_LDCT1", 24, 24, 0.1, 0.1, 60, 16),
    ("SINGLE_FUSED_VKSSQ", 40, 24, 0.05, 0.1, 60, 4),
    ("SINGLE_ORIG_VKSSQ", 40, 24, 0.05, 0.1, 60, 4),
    ("SINGLE_FUSED_VKSHC", 40, 24, 0.05, 0.1, 60, 5),
    ("DOUBLE_FUSED_VKSHC", 40, 24, 0.05, 0.1, 60, 5),
    # FORCING_Y: the channel's body force along y, against the walls
    ("SINGLE_FUSED_POFY", 32, 12, 0.1, 0.1, 80, 16),
    ("SINGLE_ORIG_POFY", 32, 12, 0.1, 0.1, 80, 16),
    ("DOUBLE_FUSED_POFY", 32, 12, 0.1, 0.1, 80, 16),
    # NOSHIFT: whole populations f stored instead of the deviations f - w
    ("SINGLE_FUSED_LDCNS", 24, 20, 0.1, 0.1, 60, 16),
    ("SINGLE_ORIG_LDCNS", 24, 20, 0.1, 0.1, 60, 16),
    ("DOUBLE_FUSED_LDCNS", 24, 20, 0.1, 0.1, 60, 16),
    ("SINGLE_FUSED_TGVNS", 24, 24, 0.05, 0.1, 40, 16),
    ("DOUBLE_FUSED_TGVNS", 24, 24, 0.05, 0.1, 40, 16),
    ("MIXED1_FUSED_TGVNS", 24, 24, 0.05, 0.1, 40, 16),
    ("SINGLE_FUSED_POFNS", 32, 12, 0.1, 0.1, 80, 16),
    ("SINGLE_ORIG_POFNS", 32, 12, 0.1, 0.1, 80, 16),
    ("SINGLE_FUSED_VKSNS", 40, 24, 0.05, 0.1, 60, 4),
    ("SINGLE_ORIG_VKSNS", 40, 24, 0.05, 0.1, 60, 4),
    ("SINGLE_FUSED_BFSNS", 60, 20, 0.1, 0.1, 50, 16),
    ("SINGLE_ORIG_BFSNS", 60, 20, 0.1, 0.1, 50, 16),
    ("SINGLE_FUSED_COUNS", 24, 16, 0.1, 0.1, 60, 16),
    # DRAG: draglift (SRC/draglift.F90) of the last step, fort.66
    ("SINGLE_FUSED_VKSD", 100, 60, 0.02, 0.1, 400, 8),
    ("DOUBLE_FUSED_VKSD", 100, 60, 0.02, 0.1, 400, 8),
]


def _read_text(path):
    rows = []
    if not os.path.exists(path):
        return np.zeros((0, 0))
    with open(path) as f:
        for line in f:
            t = line.split()
            if not t or t[0].startswith("#"):
                continue
            try:
                rows.append([float(x) for x in t])
            except ValueError:
                continue
    return np.array(rows)


def run_variant(name, lx, ly, svisc, u0, nsteps, radius):
    exe = os.path.join(ROOT, "oracle", "_ref", "ref_" + name.split("@")[0])
    with tempfile.TemporaryDirectory() as td:
        dump = os.path.join(td, "dump.bin")
        subprocess.run([exe, str(lx), str(ly), repr(svisc), repr(u0), str(nsteps),
                        str(radius), dump], cwd=td, check=True,
                       stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        d = pyoracle.read_ref_dump(dump)
        d["diagno"] = _read_text(os.path.join(td, "fort.63"))
        d["varm"] = _read_text(os.path.join(td, "fort.62"))
        d["prof_i"] = _read_text(os.path.join(td, "fort.61"))
        d["prof_j"] = _read_text(os.path.join(td, "fort.64"))
        d["probe"] = _read_text(os.path.join(td, "fort.68"))
        d["drag"] = _read_text(os.path.join(td, "fort.66"))
    return d


def main():
    out = {}
    for name, lx, ly, svisc, u0, nsteps, radius in CASES:
        d = run_variant(name, lx, ly, svisc, u0, nsteps, radius)
        key = name
        out[key + "/params"] = np.array([lx, ly, nsteps, radius], dtype=np.int64)
        out[key + "/fparams"] = np.array([svisc, u0])
        out[key + "/pops"] = d["pops"]
        out[key + "/consts"] = np.array([d["omega"], d["fgrad"], d["u00"]])
        for t in ("diagno", "varm", "prof_i", "prof_j", "probe"):
            out[key + "/" + t] = d[t]
        if d["drag"].size:
            out[key + "/drag"] = d["drag"]
        if d["obs"] is not None:
            out[key + "/obs"] = d["obs"]
        print(f"{name}: {d['pops'].dtype} {d['pops'].shape}")
    path = os.path.join(HERE, "ref_fixtures.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()

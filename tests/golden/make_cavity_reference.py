"""Cavity fixtures from the reference's own code, built SINGLE and DOUBLE
(oracle/_ref/ref_{SINGLE,DOUBLE}_{ORIG,FUSED}_LDC):

  c1: 256^2, ORIGINAL stream + collide, Re = 100 (svisc = 0.256, u0 = 0.1),
      10,000 steps (seconds);
  c2: 4096^2, FUSED, Re = 1e4 (svisc = 0.04096, u0 = 0.1), 20,000 steps
      (about 1.5 hours, the two builds on one core each).

Stores only digests and derived numbers (<case>_reference.json): the
SHA-256 of each build's interior populations (raw storage bits,
population-major, rows j = 1..m, columns i = 1..l), the velocity/density
fields' Linf and relative L2 differences between the two builds, and the
total mass of each.  Nothing of the reference's source.

    make -C oracle ref && python tests/golden/make_cavity_reference.py [c1|c2]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle  # noqa: E402

# name: (build case, implementation, l, m, svisc, u0, steps, builds)
CASES = {
    "c1": ("LDC", "ORIG", 256, 256, 0.256, 0.1, 10000, ("SINGLE", "DOUBLE")),
    "c2": ("LDC", "FUSED", 4096, 4096, 0.04096, 0.1, 20000, ("SINGLE", "DOUBLE")),
    "c2short": ("LDC", "FUSED", 4096, 4096, 0.04096, 0.1, 2000, ("SINGLE", "DOUBLE")),
    # Taylor-Green at the C3 size with the reference's own initial field
    # (amplitude 0.1, SRC/init.F90); the SINGLE build only (DOUBLE: ~2 h)
    "c3": ("TGV", "FUSED", 8192, 8192, 0.1, 0.1, 5000, ("SINGLE",)),
    # half-precision storage (the MIXED1 build: real*2 populations, real*4
    # arithmetic) at 1024^2; the reference's real*2 runs at ~8 MLUPS here
    "c3half": ("TGV", "FUSED", 1024, 1024, 0.1, 0.1, 500, ("MIXED1",)),
}
# D2Q9 in the order 01 03 05 08 10 12 14 17 19
CX = np.array([1, 1, 1, 0, -1, -1, -1, 0, 0])
CY = np.array([-1, 1, 0, 1, -1, 1, 0, -1, 0])


def digest(pops):
    """SHA-256 of the interior populations, raw bits"""
    return hashlib.sha256(np.ascontiguousarray(pops[:, 1:-1, 1:-1]).tobytes()).hexdigest()


def fields(pops):
    """rho, u, v in FP64 from the shifted populations (f - w)"""
    f = pops[:, 1:-1, 1:-1]
    f = (f.view(np.float16) if f.dtype == np.uint16 else f).astype(np.float64)  # real*2 bits
    rho = f.sum(0) + 1.0
    return rho, (CX[:, None, None] * f).sum(0) / rho, (CY[:, None, None] * f).sum(0) / rho


def compare(a, b):
    """Linf and relative L2 of (u, v) and Linf of rho between two runs"""
    ra, ua, va = fields(a)
    rb, ub, vb = fields(b)
    du = np.concatenate([(ua - ub).ravel(), (va - vb).ravel()])
    ref = np.concatenate([ub.ravel(), vb.ravel()])
    return dict(uv_linf=float(np.abs(du).max()), uv_rel_l2=float(np.linalg.norm(du) / np.linalg.norm(ref)),
                rho_linf=float(np.abs(ra - rb).max()))


def run(precs, case, impl, l, m, svisc, u0, steps):
    """the builds of `precs`, in parallel processes; their populations"""
    with tempfile.TemporaryDirectory() as td:
        procs = []
        for prec in precs:
            exe = os.path.join(ROOT, "oracle", "_ref", f"ref_{prec}_{impl}_{case}")
            wd = os.path.join(td, prec)
            os.makedirs(wd)
            procs.append(subprocess.Popen([exe, str(l), str(m), repr(svisc), repr(u0), str(steps), "16",
                                           os.path.join(wd, "dump.bin")], cwd=wd,
                                          stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL))
        for p in procs:
            if p.wait() != 0:
                raise RuntimeError("reference run failed")
        return [pyoracle.read_ref_dump(os.path.join(td, prec, "dump.bin"))["pops"] for prec in precs]


def main():
    name = sys.argv[1] if len(sys.argv) > 1 else "c1"
    case, impl, l, m, svisc, u0, steps, precs = CASES[name]
    pops = run(precs, case, impl, l, m, svisc, u0, steps)
    title = {"LDC": "lid-driven cavity", "TGV": "Taylor-Green vortex (amplitude 0.1)"}[case]
    out = dict(case=f"{title} {l}x{m}, {'ORIGINAL' if impl == 'ORIG' else impl}, "
                    f"svisc={svisc}, u0={u0}, {steps} steps")
    for prec, p in zip(precs, pops):
        out["sha256_" + prec.lower()] = digest(p)
        out["mass_" + prec.lower()] = float(fields(p)[0].sum())
    if len(pops) == 2:
        out["single_vs_double"] = compare(pops[0], pops[1])
    with open(os.path.join(HERE, f"{name}_reference.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()

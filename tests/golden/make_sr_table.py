"""Record gfx950's stochastic-rounding conversions for the oracle's tests.

Builds sr_probe.hip, runs it on the GPU over a seeded table of binary32
inputs (every exponent that matters for FP16 and BF16: half subnormals,
the normal half range, overflow, binary32 subnormals, inf and NaN) times
random words (fixed edge words and random ones), and writes
sr_convert.npz: x (binary32 bits), r, f16, bf16 (uint32 arrays).

    python tests/golden/make_sr_table.py [out.npz]   # on a machine with the GPU
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def table(seed=20261204):
    rng = np.random.default_rng(seed)
    xs = set()
    for _ in range(2500):  # around the half range
        e = int(rng.integers(99, 145))
        xs.add((int(rng.integers(0, 2)) << 31) | (e << 23) | int(rng.integers(0, 1 << 23)))
    for _ in range(600):  # anywhere, plus the ends of the binary32 range
        e = int(rng.choice([0, 1, 2, 253, 254, int(rng.integers(0, 255))]))
        xs.add((int(rng.integers(0, 2)) << 31) | (e << 23) | int(rng.integers(0, 1 << 23)))
    special = [0x00000000, 0x80000000, 0x3f800000, 0xbf800000, 0x477fe000, 0x477fefff, 0x477ff000,
               0x477fffff, 0x47800000, 0x38800000, 0x387fffff, 0x33800000, 0x33000000, 0x32ffffff,
               0x7f7fffff, 0x7f7f8000, 0x7f7f7fff, 0xff7fffff, 0x00000001, 0x80000001, 0x007fffff,
               0x7f800000, 0xff800000, 0x7f800001, 0x7fbfffff, 0x7fc00000, 0x7fc00001, 0xffc00000,
               0x7fa00000]
    xs.update(special)
    words = [0, 1, 0x1fff, 0x2000, 0xffff, 0x10000, 0x7ffff, 0x7fffffff, 0x80000000, 0xffffffff,
             0xffffe000, 0xfff80000, 0xffff0000]
    rows = []
    for x in sorted(xs):
        for r in words:
            rows.append((x, r))
        for r in rng.integers(0, 1 << 32, size=4, dtype=np.uint64):
            rows.append((x, int(r)))
    return rows


def main():
    rows = table()
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "sr_probe")
        subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops", "-o", exe, os.path.join(HERE, "sr_probe.hip")], check=True)
        src, dst = os.path.join(d, "in.txt"), os.path.join(d, "out.txt")
        with open(src, "w") as f:
            f.writelines(f"{x:08x} {r:08x}\n" for x, r in rows)
        subprocess.run([exe, src, dst], check=True)
        out = np.array([[int(v, 16) for v in line.split()] for line in open(dst)], dtype=np.uint32)
    path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(HERE, "sr_convert.npz")
    np.savez_compressed(path, x=out[:, 0], r=out[:, 1], f16=out[:, 2], bf16=out[:, 3])
    print(f"{path}: {len(out)} conversions")


if __name__ == "__main__":
    sys.exit(main())

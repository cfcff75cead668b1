"""Build recipe of the native parts: libbgk2d_hip.so (HIP, gfx950) and the
Fortran driver bgk2d_hip (flang, ISO_C_BINDING).  Builds in-tree so the
built files travel with the source tree."""
import os
import shutil
import subprocess

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
LIB_DIR = os.path.join(PKG, "lib")
LIB = os.path.join(LIB_DIR, "libbgk2d_hip.so")
DRIVER = os.path.join(LIB_DIR, "bgk2d_hip")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")

# -ffp-contract=off: every product and sum separately rounded, matching the
# reference's x86-64 build bit for bit.  Packed FP32 ops are disabled.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wno-unused-value", "-Wno-unused-command-line-argument"]

CSRC = os.path.join(PKG, "csrc")
# the C ABI, then the step loop of each storage kind and the batch kernels
# (explicit instantiations, compiled in parallel)
SOURCES = [os.path.join(CSRC, f) for f in ("bgk_api.hip", "bgk_kind_f32.hip", "bgk_kind_f64.hip",
                                           "bgk_kind_f16.hip", "bgk_kind_bf16.hip", "bgk_kind_mixed2.hip",
                                           "bgk_kind_batch.hip")]
DEPS = SOURCES + [os.path.join(CSRC, f) for f in ("bgk_device.hpp", "bgk_kernels.hpp", "bgk_serve.hpp",
                                                  "bgk_host.hpp")] + \
    [os.path.join(ROOT, "include", "bgk2d.h")]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        msg = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{msg[-6000:]}")
    return r.stdout


def build_lib(verbose=False, jobs=None):
    """compile every translation unit (in parallel), then link"""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(LIB_DIR, "obj")
    os.makedirs(objdir, exist_ok=True)
    objs = [os.path.join(objdir, os.path.basename(src)[:-4] + ".o") for src in SOURCES]
    jobs = jobs or min(len(SOURCES), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        for f in [ex.submit(_run, [HIPCC] + HIP_FLAGS + ["-c", "-o", o, src]) for src, o in zip(SOURCES, objs)]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_driver(verbose=False):
    """Fortran driver: bgk2d_gpu_amd/driver/*.F90 linked against the library."""
    drv = os.path.join(PKG, "driver")
    srcs = [os.path.join(drv, f) for f in ("bgk_hip.F90", "bgk2d_hip.F90")]
    if not all(os.path.exists(s) for s in srcs) or not os.path.exists(FLANG):
        return None
    moddir = os.path.join(LIB_DIR, "mod")
    os.makedirs(moddir, exist_ok=True)
    _run([FLANG, "-O2", "-module-dir", moddir, "-o", DRIVER] + srcs +
         ["-L" + LIB_DIR, "-lbgk2d_hip", "-Wl,-rpath,$ORIGIN"])
    if verbose:
        print("built", DRIVER)
    return DRIVER


def build_all(verbose=False):
    build_lib(verbose)
    build_driver(verbose)


def have_lib():
    return os.path.exists(LIB)


if __name__ == "__main__":
    build_all(verbose=True)

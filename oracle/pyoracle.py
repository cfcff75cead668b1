"""ctypes binding of the oracle (liboracle.so) and reader for reference dumps.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

BC = {"driven": 0, "periodic": 1, "channel": 2, "inflow": 3, "couette": 4, "step": 5,
      "open": 6}
PREC = {"single": 0, "double": 1, "mixed1": 2, "mixed2": 3, "bf16": 4}
IMPL = {"original": 0, "fused": 1}
# reference build case -> (bc, obstacle)
CASE = {"LDC": ("driven", 0), "POF": ("channel", 0), "TGV": ("periodic", 0),
        "VKS": ("inflow", 1), "COU": ("couette", 0), "BFS": ("step", 0),
        "LDCLES": ("driven", 0), "VKSLES": ("inflow", 1), "LDCT1": ("driven", 0),
        "VKSSQ": ("inflow", 2), "VKSHC": ("inflow", 3), "POFY": ("channel", 0),
        "LDCNS": ("driven", 0), "TGVNS": ("periodic", 0), "POFNS": ("channel", 0),
        "VKSNS": ("inflow", 1), "BFSNS": ("step", 0), "COUNS": ("couette", 0),
        "VKSD": ("inflow", 1)}
# reference build cases compiled with -DNOSHIFT
NOSHIFT_CASES = {"LDCNS", "TGVNS", "POFNS", "VKSNS", "BFSNS", "COUNS"}
# reference build cases compiled with -DTRICK1
TRICK1_CASES = {"LDCT1"}
# reference build cases compiled with -DFORCING_Y
FORCING_Y_CASES = {"POFY"}
# reference build cases compiled with -DLES (cteS = 0.1, a default-real literal)
LES_CASES = {"LDCLES", "VKSLES"}
CTE_S_REF = 0.10000000149011612
STORE_DTYPE = {"single": np.float32, "double": np.float64, "mixed1": np.uint16,
               "mixed2": np.float32, "bf16": np.uint16}


class OrcParams(C.Structure):
    _fields_ = [("lx", C.c_int32), ("ly", C.c_int32), ("bc", C.c_int32),
                ("obstacle", C.c_int32), ("impl", C.c_int32), ("prec", C.c_int32),
                ("radius", C.c_int32), ("init_v", C.c_int32),
                ("mixed1_quirk", C.c_int32), ("tgv", C.c_int32),
                ("form", C.c_int32), ("rounding", C.c_int32),
                ("les", C.c_int32), ("trick1", C.c_int32), ("forcing_y", C.c_int32),
                ("noshift", C.c_int32),
                ("svisc", C.c_double), ("u0", C.c_double),
                ("amplitude", C.c_double), ("scale", C.c_double), ("cte_s", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [P(OrcParams)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_setup.argtypes = [C.c_void_p]
        L.orc_setup.restype = C.c_int
        for fn in ("orc_boundaries", "orc_propagation", "orc_collision"):
            getattr(L, fn).argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int32]
        L.orc_storage_bytes.argtypes = [C.c_void_p]
        L.orc_constants.argtypes = [C.c_void_p, P(C.c_double)]
        L.orc_get_pop.argtypes = [C.c_void_p, C.c_int32, P(C.c_double)]
        L.orc_get_raw.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.orc_get_raw_b.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.orc_set_raw.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.orc_get_obs.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_int32)]
        L.orc_set_obs.argtypes = [C.c_void_p, P(C.c_int32)]
        L.orc_draglift.argtypes = [C.c_void_p] + [C.c_int32] * 4 + [P(C.c_double)]
        L.orc_diagno.argtypes = [C.c_void_p, C.c_int32, P(C.c_double)]
        L.orc_varm.argtypes = [C.c_void_p] + [P(C.c_double)] * 3
        L.orc_prof_i.argtypes = [C.c_void_p, C.c_int32] + [P(C.c_double)] * 3
        L.orc_prof_j.argtypes = [C.c_void_p, C.c_int32] + [P(C.c_double)] * 3
        L.orc_probe.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_double)]
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_f32_to_bf16.argtypes = [C.c_float]
        L.orc_f32_to_bf16.restype = C.c_uint16
        L.orc_bf16_to_f32.argtypes = [C.c_uint16]
        L.orc_bf16_to_f32.restype = C.c_float
        for f in (L.orc_sr_f16, L.orc_sr_bf16):
            f.argtypes = [C.c_uint32, C.c_uint32]
            f.restype = C.c_uint16
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One CPU solve; mirrors the reference's module state `storage`."""

    def __init__(self, lx, ly, bc="driven", impl="fused", prec="single", svisc=0.1,
                 u0=0.1, obstacle=0, radius=16, amplitude=0.1, scale=1.0,
                 mixed1_quirk=0, init_v=0, tgv=None, form=0, rounding=0, les=0, cte_s=0.1,
                 trick1=0, forcing_y=0, noshift=0):
        L = lib()
        if tgv is None:
            tgv = 1 if bc == "periodic" else 0
        self.p = OrcParams(lx, ly, BC[bc], obstacle, IMPL[impl], PREC[prec], radius,
                           init_v, mixed1_quirk, tgv, form, rounding, les, trick1, forcing_y, noshift, svisc, u0, amplitude,
                           scale, cte_s)
        self.prec = prec
        self.l, self.m = lx, ly
        self.h = L.orc_create(C.byref(self.p))
        if not self.h:
            raise MemoryError("orc_create failed")
        if L.orc_setup(self.h) != 0:
            raise ValueError("orc_setup rejected the parameters")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    @property
    def plane_shape(self):
        return (self.m + 2, self.l + 2)

    def step(self, n=1):
        lib().orc_step(self.h, n)

    def boundaries(self):
        lib().orc_boundaries(self.h)

    def propagation(self):
        lib().orc_propagation(self.h)

    def collision(self):
        lib().orc_collision(self.h)

    def constants(self):
        out = np.zeros(4)
        lib().orc_constants(self.h, _dp(out))
        return dict(omega=out[0], fgrad=out[1], u00=out[2], force=out[3])

    def pops(self):
        """decoded current populations, shape (9, m+2, l+2), float64"""
        out = np.zeros((9,) + self.plane_shape)
        for k in range(9):
            lib().orc_get_pop(self.h, k, _dp(out[k]))
        return out

    def raw(self, which="a"):
        dt = STORE_DTYPE[self.prec]
        out = np.zeros((9,) + self.plane_shape, dtype=dt)
        fn = lib().orc_get_raw if which == "a" else lib().orc_get_raw_b
        for k in range(9):
            fn(self.h, k, out[k].ctypes.data)
        return out

    def set_raw(self, arr):
        dt = STORE_DTYPE[self.prec]
        arr = np.ascontiguousarray(arr, dtype=dt)
        for k in range(9):
            lib().orc_set_raw(self.h, k, arr[k].ctypes.data)

    def obs(self):
        out = np.zeros((self.m, self.l), dtype=np.int32)
        bbox = np.zeros(4, dtype=np.int32)
        lib().orc_get_obs(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)),
                          bbox.ctypes.data_as(C.POINTER(C.c_int32)))
        return out, bbox

    def obstacle(self):
        """(mask (m, l) int32, bbox [imin, imax, jmin, jmax])"""
        mask = np.zeros((self.m, self.l), dtype=np.int32)
        bbox = np.zeros(4, dtype=np.int32)
        lib().orc_get_obs(self.h, mask.ctypes.data_as(C.POINTER(C.c_int32)),
                          bbox.ctypes.data_as(C.POINTER(C.c_int32)))
        return mask, bbox

    def set_obs(self, mask):
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        assert mask.shape == (self.m, self.l)
        lib().orc_set_obs(self.h, mask.ctypes.data_as(C.POINTER(C.c_int32)))

    def draglift(self, box):
        out = np.zeros(2)
        lib().orc_draglift(self.h, *(int(b) for b in box), _dp(out))
        return out

    def diagno(self, mode=0):
        out = np.zeros(4)
        lib().orc_diagno(self.h, mode, _dp(out))
        return out

    def varm(self):
        u, v, r = np.zeros(self.m), np.zeros(self.m), np.zeros(self.m)
        lib().orc_varm(self.h, _dp(u), _dp(v), _dp(r))
        return u, v, r

    def prof_i(self, jc):
        u, v, r = np.zeros(self.l), np.zeros(self.l), np.zeros(self.l)
        lib().orc_prof_i(self.h, jc, _dp(u), _dp(v), _dp(r))
        return u, v, r

    def prof_j(self, ic):
        u, v, r = np.zeros(self.m), np.zeros(self.m), np.zeros(self.m)
        lib().orc_prof_j(self.h, ic, _dp(u), _dp(v), _dp(r))
        return u, v, r

    def probe(self, i0, j0):
        out = np.zeros(3)
        lib().orc_probe(self.h, i0, j0, _dp(out))
        return out


def read_ref_dump(path):
    """Read a dump written by oracle/ref_harness.F90."""
    with open(path, "rb") as f:
        hdr = np.frombuffer(f.read(16), dtype=np.int32)
        l, m, es, nsteps = (int(x) for x in hdr)
        dt = {2: np.uint16, 4: np.float32, 8: np.float64}[es]
        n = (l + 2) * (m + 2)
        pops = np.frombuffer(f.read(9 * n * es), dtype=dt).reshape(9, m + 2, l + 2).copy()
        consts = np.frombuffer(f.read(24), dtype=np.float64).copy()
        rest = f.read(4 * l * m)  # obstacle builds: the build_obs mask
        obs = np.frombuffer(rest, dtype=np.int32).reshape(m, l).copy() if len(rest) == 4 * l * m else None
    return dict(l=l, m=m, esize=es, nsteps=nsteps, pops=pops,
                omega=consts[0], fgrad=consts[1], u00=consts[2], obs=obs)

"""ctypes view of libbgk2d_hip.so for tests, benchmarks and the serving loop.

Method names follow the reference's subroutines (boundaries, propagation,
collision, diagno, varm, prof_i, prof_j, probe).  There is no fallback: if
the HIP library is missing or no GPU is visible, construction raises.
"""
import ctypes as C
import os

import numpy as np

from . import build as _build

BC = {"driven": 0, "periodic": 1, "channel": 2, "inflow": 3, "couette": 4, "step": 5,
      "open": 6}
PREC = {"single": 0, "double": 1, "fp16": 2, "mixed2": 3, "bf16": 4}
IMPL = {"original": 0, "fused": 1}
STORE_DTYPE = {"single": np.float32, "double": np.float64, "fp16": np.uint16,
               "mixed2": np.float32, "bf16": np.uint16}
ERRORS = {-1: "invalid argument", -2: "out of memory", -3: "HIP error", -4: "diverged",
          -5: "bad call order"}

# every entry point declared in include/bgk2d.h
EXPORTS = ("bgk_create", "bgk_destroy", "bgk_last_error", "bgk_version", "bgk_constants",
           "bgk_build_obs", "bgk_set_obstacle", "bgk_get_obstacle", "bgk_init",
           "bgk_init_noalloc", "bgk_boundaries", "bgk_propagation", "bgk_collision", "bgk_run", "bgk_sync",
           "bgk_diagno", "bgk_varm", "bgk_prof_i", "bgk_prof_j", "bgk_probe", "bgk_draglift",
           "bgk_fields", "bgk_check",
           "bgk_download", "bgk_upload", "bgk_storage_bytes", "bgk_stream", "bgk_last_run_ms",
           "bgk_batch_create", "bgk_batch_destroy", "bgk_batch_last_error", "bgk_footprint",
           "bgk_batch_add", "bgk_batch_ctx", "bgk_batch_remaining", "bgk_batch_remove", "bgk_batch_hold",
           "bgk_batch_used", "bgk_batch_run", "bgk_batch_last_ms")


class BgkParams(C.Structure):
    _fields_ = [("lx", C.c_int32), ("ly", C.c_int32), ("ipad", C.c_int32), ("jpad", C.c_int32),
                ("bc", C.c_int32), ("obstacle", C.c_int32), ("impl", C.c_int32),
                ("prec", C.c_int32), ("radius", C.c_int32), ("device", C.c_int32),
                ("mixed1_quirk", C.c_int32), ("form", C.c_int32),
                ("rounding", C.c_int32), ("tblock", C.c_int32),
                ("les", C.c_int32), ("trick1", C.c_int32), ("forcing_y", C.c_int32),
                ("noshift", C.c_int32),
                ("svisc", C.c_double), ("u0", C.c_double), ("scale", C.c_double),
                ("cte_s", C.c_double)]


class BgkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def load(path=None):
    """Load the HIP library (in-tree build).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or _build.LIB
    if not os.path.exists(path):
        raise OSError(f"{path} is not built; run bgk2d_gpu_amd.build.build_all()")
    L = C.CDLL(path)
    P, V, I32, D = C.POINTER, C.c_void_p, C.c_int32, C.c_double
    sig = {
        "bgk_create": ([P(BgkParams), P(V)], C.c_int),
        "bgk_destroy": ([V], C.c_int),
        "bgk_last_error": ([V], C.c_char_p),
        "bgk_version": ([], C.c_int),
        "bgk_constants": ([V, P(D)], C.c_int),
        "bgk_build_obs": ([V], C.c_int),
        "bgk_set_obstacle": ([V, P(I32)], C.c_int),
        "bgk_get_obstacle": ([V, P(I32), P(I32)], C.c_int),
        "bgk_init": ([V, I32, D], C.c_int),
        "bgk_init_noalloc": ([V, I32, D], C.c_int),
        "bgk_boundaries": ([V], C.c_int),
        "bgk_propagation": ([V], C.c_int),
        "bgk_collision": ([V, I32], C.c_int),
        "bgk_run": ([V, I32], C.c_int),
        "bgk_sync": ([V], C.c_int),
        "bgk_diagno": ([V, P(D)], C.c_int),
        "bgk_varm": ([V, P(D), P(D), P(D)], C.c_int),
        "bgk_prof_i": ([V, I32, P(D), P(D), P(D)], C.c_int),
        "bgk_prof_j": ([V, I32, P(D), P(D), P(D)], C.c_int),
        "bgk_probe": ([V, I32, I32, P(D)], C.c_int),
        "bgk_draglift": ([V, I32, I32, I32, I32, P(D)], C.c_int),
        "bgk_fields": ([V, P(C.c_float), P(C.c_float), P(C.c_float)], C.c_int),
        "bgk_check": ([V, D], C.c_int),
        "bgk_download": ([V, I32, V], C.c_int),
        "bgk_upload": ([V, I32, V], C.c_int),
        "bgk_storage_bytes": ([V], C.c_int),
        "bgk_stream": ([V], V),
        "bgk_last_run_ms": ([V, P(D)], C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One resident BGK2D solve on the GPU (a bgk_ctx)."""

    def __init__(self, lx, ly, bc="driven", impl="fused", prec="single", svisc=0.1, u0=0.1,
                 obstacle=0, radius=16, scale=1.0, device=0, mixed1_quirk=0, init=True,
                 init_v=0, amplitude=0.1, form=0, rounding=0, tblock=0, les=0, cte_s=0.1,
                 trick1=0, forcing_y=0, noshift=0):
        self.L = load()
        self.lx, self.ly = int(lx), int(ly)
        self.prec = prec
        self.bc = bc
        self.impl = impl
        self.p = BgkParams(lx=self.lx, ly=self.ly, bc=BC[bc], obstacle=int(obstacle),
                           impl=IMPL[impl], prec=PREC[prec], radius=int(radius),
                           device=int(device), mixed1_quirk=int(mixed1_quirk), form=int(form),
                           rounding=int(rounding), tblock=int(tblock), les=int(les),
                           trick1=int(trick1), forcing_y=int(forcing_y), noshift=int(noshift),
                           svisc=float(svisc), u0=float(u0), scale=float(scale),
                           cte_s=float(cte_s))
        h = C.c_void_p()
        rc = self.L.bgk_create(C.byref(self.p), C.byref(h))
        if rc != 0:
            raise BgkError(rc, f"bgk_create({lx}x{ly}, {bc}, {impl}, {prec})")
        self.h = h
        self.itime = 0
        if obstacle:
            self._chk(self.L.bgk_build_obs(self.h))
        if init:
            self.init(init_v, amplitude)

    def _chk(self, rc):
        if rc != 0:
            raise BgkError(rc, self.L.bgk_last_error(self.h).decode())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.bgk_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def plane_shape(self):
        return (self.ly + 2, self.lx + 2)

    # ---- the reference's phases
    def init(self, init_v=0, amplitude=0.1):
        self._chk(self.L.bgk_init(self.h, int(init_v), float(amplitude)))
        self.itime = 0

    def boundaries(self):
        self._chk(self.L.bgk_boundaries(self.h))

    def propagation(self):
        self._chk(self.L.bgk_propagation(self.h))

    def collision(self, itime=None):
        self.itime += 1
        self._chk(self.L.bgk_collision(self.h, self.itime if itime is None else int(itime)))

    def step_phases(self, n=1):
        for _ in range(n):
            self.boundaries()
            self.propagation()
            self.collision()

    def run(self, n):
        self._chk(self.L.bgk_run(self.h, int(n)))
        self.itime += int(n)

    def sync(self):
        self._chk(self.L.bgk_sync(self.h))

    def last_run_ms(self):
        out = C.c_double()
        self._chk(self.L.bgk_last_run_ms(self.h, C.byref(out)))
        return out.value

    # ---- diagnostics
    def constants(self):
        out = np.zeros(4)
        self._chk(self.L.bgk_constants(self.h, _dp(out)))
        return dict(omega=out[0], fgrad=out[1], u00=out[2], force=out[3])

    def diagno(self):
        out = np.zeros(4)
        self._chk(self.L.bgk_diagno(self.h, _dp(out)))
        return out

    def check(self, umax=0.5):
        rc = self.L.bgk_check(self.h, float(umax))
        return rc == 0

    def varm(self):
        u, v, r = np.zeros(self.ly), np.zeros(self.ly), np.zeros(self.ly)
        self._chk(self.L.bgk_varm(self.h, _dp(u), _dp(v), _dp(r)))
        return u, v, r

    def prof_i(self, jc):
        u, v, r = np.zeros(self.lx), np.zeros(self.lx), np.zeros(self.lx)
        self._chk(self.L.bgk_prof_i(self.h, int(jc), _dp(u), _dp(v), _dp(r)))
        return u, v, r

    def prof_j(self, ic):
        u, v, r = np.zeros(self.ly), np.zeros(self.ly), np.zeros(self.ly)
        self._chk(self.L.bgk_prof_j(self.h, int(ic), _dp(u), _dp(v), _dp(r)))
        return u, v, r

    def probe(self, i0, j0):
        out = np.zeros(3)
        self._chk(self.L.bgk_probe(self.h, int(i0), int(j0), _dp(out)))
        return out

    # ---- state transfer
    def raw(self):
        dt = STORE_DTYPE[self.prec]
        out = np.zeros((9,) + self.plane_shape, dtype=dt)
        for k in range(9):
            self._chk(self.L.bgk_download(self.h, k, out[k].ctypes.data))
        return out

    def set_raw(self, arr):
        dt = STORE_DTYPE[self.prec]
        arr = np.ascontiguousarray(arr, dtype=dt)
        assert arr.shape == (9,) + self.plane_shape
        for k in range(9):
            self._chk(self.L.bgk_upload(self.h, k, arr[k].ctypes.data))

    def draglift(self, box):
        """(Fx, Fy) on the obstacle from the fluid nodes of box = (i0, i1, j0, j1)"""
        out = np.zeros(2)
        self._chk(self.L.bgk_draglift(self.h, *(int(b) for b in box), _dp(out)))
        return out

    def fields(self):
        """(u, v, rho) of the VTK writers, float32 arrays of shape (ly, lx)"""
        out = [np.zeros((self.ly, self.lx), dtype=np.float32) for _ in range(3)]
        fp = [o.ctypes.data_as(C.POINTER(C.c_float)) for o in out]
        self._chk(self.L.bgk_fields(self.h, *fp))
        return tuple(out)

    def obstacle(self):
        out = np.zeros((self.ly, self.lx), dtype=np.int32)
        bbox = np.zeros(4, dtype=np.int32)
        self._chk(self.L.bgk_get_obstacle(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)),
                                          bbox.ctypes.data_as(C.POINTER(C.c_int32))))
        return out, bbox

    def set_obstacle(self, mask):
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        assert mask.shape == (self.ly, self.lx)
        self._chk(self.L.bgk_set_obstacle(self.h, mask.ctypes.data_as(C.POINTER(C.c_int32))))

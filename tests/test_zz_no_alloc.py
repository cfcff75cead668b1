"""bgk_init_noalloc: bgk_init with its trig tables in a static device array
instead of a hipMalloc'd buffer.  Its populations equal bgk_init's bit for
bit, on two calls in a row (the second on a smaller lattice, so the array
still holds the first call's longer tables), and lattices whose tables
exceed the array are refused."""
import ctypes as C

import numpy as np
import pytest

from bgk2d_gpu_amd import Solver
from bgk2d_gpu_amd.solver import BgkError

pytestmark = pytest.mark.gpu

NARROW = ("fp16", "bf16")


def make(lx, ly, prec, bc):
    narrow = prec in NARROW
    return Solver(lx, ly, bc=bc, prec=prec, svisc=0.05, u0=0.0 if bc == "periodic" else 0.1,
                  scale=1024.0 if narrow else 1.0, form=1 if narrow else 0,
                  rounding=1 if narrow else 0, init=False)


def noalloc(s, init_v, amp):
    rc = s.L.bgk_init_noalloc(s.h, C.c_int32(init_v), C.c_double(amp))
    if rc != 0:
        raise BgkError(rc, s.L.bgk_last_error(s.h).decode())


@pytest.mark.parametrize("prec", ["single", "double", "fp16", "bf16", "mixed2"])
@pytest.mark.parametrize("bc,init_v,amp", [("periodic", 0, 0.03), ("periodic", 1, 0.05),
                                           ("driven", 0, 0.1)])
def test_noalloc_init_equals_init(prec, bc, init_v, amp):
    for lx, ly in ((700, 500), (300, 200)):  # second call smaller
        print(f"case {prec} {bc} init_v={init_v} {lx}x{ly}", flush=True)
        a, b = make(lx, ly, prec, bc), make(lx, ly, prec, bc)
        try:
            a.init(init_v, amp)
            b.init(1 - init_v if bc == "periodic" else 0, 2 * amp)  # other contents first
            noalloc(b, init_v, amp)
            ra, rb = a.raw(), b.raw()
            assert np.array_equal(ra.view(np.uint8), rb.view(np.uint8)), (prec, bc, init_v, lx, ly)
            # and it steps like a context bgk_init set up
            a.run(20)
            b.run(20)
            assert np.array_equal(a.raw().view(np.uint8), b.raw().view(np.uint8))
        finally:
            a.close()
            b.close()


def test_noalloc_init_refuses_tables_beyond_the_array():
    # fp64 compute: 2 (lx + ly + 4) doubles > 1 MiB once lx + ly + 4 > 65536
    s = make(65600, 4, "double", "periodic")
    try:
        with pytest.raises(BgkError) as e:
            noalloc(s, 0, 0.01)
        assert e.value.code == -1 and "scratch" in str(e.value)
        s.init(0, 0.01)  # the context stays usable
    finally:
        s.close()
    a, b = make(65500, 4, "double", "periodic"), make(65500, 4, "double", "periodic")  # just fits
    try:
        a.init(0, 0.02)
        b.init(1, 0.01)
        noalloc(b, 0, 0.02)
        assert np.array_equal(a.raw().view(np.uint8), b.raw().view(np.uint8))
    finally:
        a.close()
        b.close()

"""CPU checks of the drop-in boundary: the library loads and exports every
entry point include/bgk2d.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

import bgk2d_gpu_amd as pkg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "bgk2d.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bgk_\w+)\s*\(", txt)))


def test_header_declares_expected_api():
    names = declared()
    assert set(pkg.EXPORTS) == set(names), set(pkg.EXPORTS) ^ set(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(pkg.build.LIB)
    for n in declared():
        assert hasattr(lib, n), n
    assert lib.bgk_version() == 4


def test_struct_layout_matches_header():
    # bgk_params_t: 18 int32 then 4 doubles, no padding
    assert ctypes.sizeof(pkg.solver.BgkParams) == 18 * 4 + 4 * 8
    assert pkg.solver.BgkParams.svisc.offset == 18 * 4


@pytest.mark.parametrize("change", [
    dict(forcing_y=1, bc=0),      # FORCING_Y outside the channel
    dict(forcing_y=2),            # not a flag
    dict(trick1=1, bc=2),         # TRICK1 outside the driven cavity
    dict(trick1=1, bc=0, ly=48),  # TRICK1 on a non-square cavity
    dict(les=2), dict(form=2), dict(rounding=-1), dict(tblock=3),
    dict(obstacle=4), dict(svisc=0.0), dict(bc=7), dict(lx=1),
])
def test_invalid_parameters_rejected_without_gpu(change):
    """bgk_create checks the parameters before it touches a device"""
    L = pkg.load()
    kw = dict(lx=64, ly=64, bc=2, impl=1, prec=0, svisc=0.1, u0=0.1, scale=1.0, cte_s=0.1)
    kw.update(change)
    p = pkg.solver.BgkParams(**kw)
    h = ctypes.c_void_p()
    assert L.bgk_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert not h.value


def test_null_arguments_rejected_without_gpu():
    L = pkg.load()
    assert L.bgk_create(None, None) == -1
    assert L.bgk_destroy(None) == -1
    assert L.bgk_constants(None, None) == -1
    for f, args in (("bgk_batch_hold", (None, 0, 1)), ("bgk_batch_remove", (None, 0)),
                    ("bgk_batch_run", (None, 10, None)), ("bgk_batch_destroy", (None,)),
                    ("bgk_draglift", (None, 1, 2, 1, 2, None)), ("bgk_fields", (None, None, None, None))):
        assert getattr(L, f)(*args) == -1, f


@pytest.mark.parametrize("lx,ly,prec", [(32768, 32768, 0), (24000, 24000, 1), (46400, 46400, 2)])
def test_lattice_beyond_32bit_plane_offsets_rejected_without_gpu(lx, ly, prec):
    """The hot kernels address a population plane with 32-bit byte offsets:
    bgk_create, bgk_footprint and bgk_batch_add refuse a plane of 4 GiB or
    more (32768^2 FP32, 24000^2 FP64, 46400^2 FP16) before any device call,
    with the reason in bgk_last_error(NULL)."""
    L = pkg.load()
    p = pkg.solver.BgkParams(lx=lx, ly=ly, bc=0, impl=1, prec=prec, svisc=0.1, u0=0.1, scale=1.0,
                             cte_s=0.1)
    h = ctypes.c_void_p()
    assert L.bgk_create(ctypes.byref(p), ctypes.byref(h)) == -1
    assert not h.value
    L.bgk_last_error.restype = ctypes.c_char_p
    assert b"4 GiB" in L.bgk_last_error(None)
    d = ctypes.c_double(0)
    assert L.bgk_footprint(ctypes.byref(p), ctypes.byref(d)) == -1


def test_largest_admissible_plane_has_a_footprint():
    """just below the limit the footprint is given (no device needed)"""
    L = pkg.load()
    p = pkg.solver.BgkParams(lx=32000, ly=32000, bc=0, impl=1, prec=0, svisc=0.1, u0=0.1, scale=1.0,
                             cte_s=0.1)
    d = ctypes.c_double(0)
    assert L.bgk_footprint(ctypes.byref(p), ctypes.byref(d)) == 0
    assert 18 * 32000 * 32000 * 4 < d.value < 18 * 32002 * 32064 * 4 + 32000 * 32000 + 2 ** 24

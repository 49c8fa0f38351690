"""The C-ABI library: it loads without a GPU, exports every entry point
include/lbfast_hip.h declares, and its pure host functions agree with the
reference formulas (LBFAST.f90:212-219, 258-328).  No compute call is made
here (there is no GPU in this container)."""
import ctypes
import os
import re

import numpy as np
import pytest

from lbfast_amd import hip
from lbfast_amd.params import LbfParams, make_params, derive, PREC_FP32

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lbfast_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = hip.load()
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(hip.EXPORTS)


def test_params_struct_layout_matches_header():
    # offsets of include/lbfast_hip.h's lbf_params as laid out by the C compiler
    assert (LbfParams.fx.offset, LbfParams.tau1.offset, LbfParams.store_scale.offset) == (72, 160, 208)
    assert (LbfParams.bounce_back.offset, LbfParams.smagorinsky.offset, LbfParams.repulsive_flux.offset) == (56, 60, 64)
    assert (LbfParams.a_rep.offset, LbfParams.q_th.offset, LbfParams.cos_opp_t.offset) == (288, 296, 304)
    assert (LbfParams.openbc.offset, LbfParams.openbc_type.offset, LbfParams.reserved_i2.offset) == (320, 324, 348)
    assert (LbfParams.openbc_press.offset, LbfParams.openbc_u.offset, LbfParams.openbc_v.offset,
            LbfParams.openbc_w.offset) == (352, 400, 448, 496)
    assert ctypes.sizeof(LbfParams) == 544
    L = hip.load()
    p = LbfParams()
    L.lbf_params_default(ctypes.byref(p))
    assert (p.nx, p.lattice, p.highorder, p.precision) == (64, 27, 1, PREC_FP32)
    assert abs(p.omega - 1.0 / (3 * 0.1 + 0.5)) < 1e-15


def test_params_derive_matches_reference_formulas():
    L = hip.load()
    p = make_params((8, 8, 8), twocomponent=True, densratio=True, visc1=0.07, visc2=0.2,
                    sigma=0.03, width=4.0, tau_diff=0.12)
    q = p.copy()
    q.tau1 = q.omega = q.beta = q.kapp = q.sharp_c = q.tau2 = 0
    assert L.lbf_params_derive(ctypes.byref(q)) == 0
    assert abs(q.tau1 - (3 * 0.07 + 0.5)) < 1e-15 and abs(q.omega - 1 / q.tau1) < 1e-15
    assert abs(q.beta - 12 * 0.03 / 4.0) < 1e-15 and abs(q.kapp - 1.5 * 0.03 * 4.0) < 1e-15
    assert abs(q.sharp_c - 4 * 0.12 / 4.0) < 1e-15 and abs(q.tau2 - (3 * 0.2 + 0.5)) < 1e-15


def test_create_without_gpu_fails_cleanly():
    L = hip.load()
    p = make_params((16, 16, 16))
    h = ctypes.c_void_p()
    rc = L.lbf_create(ctypes.byref(p), ctypes.byref(h))
    if rc == 0:  # a GPU is visible: release it and stop here
        L.lbf_destroy(h)
        pytest.skip("GPU present")
    assert rc in (3, 4) and not h.value
    assert L.lbf_global_last_error()


def test_create_rejects_bad_switches():
    L = hip.load()
    for kw in (dict(lattice=15), dict(lattice=19, highorder=True)):
        p = make_params((16, 16, 16), **{k: v for k, v in kw.items() if k != "highorder"})
        if "highorder" in kw:
            p.highorder = 1
        h = ctypes.c_void_p()
        assert L.lbf_create(ctypes.byref(p), ctypes.byref(h)) == 1


def test_create_rejects_bad_open_boundaries():
    """Open faces exist only in the obstacle build (boundary_cds_module.f90:
    49-57), take 32/64-bit storage and the types 0, 1 and 2; each violation
    is refused before the device is touched."""
    L = hip.load()
    cases = [
        (dict(internal_obstacles=False), {}, b"INTERNAL_OBSTACLES"),
        (dict(internal_obstacles=True, precision=3), {}, b"FP64, FP32 or MIXED"),
        (dict(internal_obstacles=True), {"type": 3}, b"openbc_type 3"),
        (dict(internal_obstacles=True), {"type": -1}, b"openbc_type -1"),
    ]
    for kw, extra, msg in cases:
        p = make_params((16, 16, 16), pbc=(0, 1, 1), **kw)
        p.openbc = 1
        if "type" in extra:
            p.openbc_type[1] = extra["type"]
        h = ctypes.c_void_p()
        assert L.lbf_create(ctypes.byref(p), ctypes.byref(h)) == 1 and not h.value
        assert msg in L.lbf_global_last_error(), (kw, extra, L.lbf_global_last_error())


@pytest.mark.parametrize("n,prec,ok", [
    ((8192, 8192, 8), 0, False),    # FP64: 10 * 8192 * 8192 * 8 B per plane = 5.4 GB
    ((16384, 8192, 2), 1, False),   # FP32: 10 * 16384 * 8192 * 4 B = 5.4 GB
    ((8192, 8192, 2), 3, True),     # FP16S: 10 * 8192 * 8192 * 2 B = 1.34 GB; aux 7 * 4 B: 1.9 GB
    ((4096, 4096, 2), 0, True),     # FP64: 1.34 GB per plane
])
def test_create_rejects_planes_beyond_32bit_offsets(n, prec, ok):
    # The kernels address one x-y plane of a field set with 32-bit byte
    # offsets; lbf_create refuses a plane of 4 GiB or more before it touches
    # the device, so this runs without a GPU.
    L = hip.load()
    p = make_params(n, precision=prec)
    h = ctypes.c_void_p()
    if ok:
        # accepted sizes would allocate gigabytes on a GPU: check them only
        # where lbf_create stops at the missing device
        rc = L.lbf_create(ctypes.byref(p), ctypes.byref(h))
        if rc == 0:
            L.lbf_destroy(h)
            pytest.skip("GPU present")
        assert rc in (3, 4), L.lbf_global_last_error()
        return
    rc = L.lbf_create(ctypes.byref(p), ctypes.byref(h))
    if rc == 0:
        L.lbf_destroy(h)
    else:
        assert rc == 2 and not h.value
        assert b"32-bit plane offsets" in L.lbf_global_last_error()


def test_namelist_reader():
    from lbfast_amd.params import read_namelists, params_from_namelist
    deck = """&simulation
   nsteps=4000
   lx=256
   ly=128, lz=64   ! comment
   pbc_x=1
   pbc_y=0
&end
&fluid
   fz=0.1e-6
   visc1= 0.05e0 !0.025e0
   visc2= 0.4e0
   rho_r=0.5e0
   rho_b=1.5e0
   tau_diff=0.1e0
   width=4.0e0
   sigma=0.00e0
   center= 128.0, 64.0, 32.0
&end
"""
    nml = read_namelists(deck)
    assert nml["simulation"]["lx"] == 256 and nml["simulation"]["lz"] == 64
    assert nml["fluid"]["center"] == [128.0, 64.0, 32.0]
    p = params_from_namelist(deck, internal_obstacles=True)
    assert (p.nx, p.ny, p.nz) == (256, 128, 64) and tuple(p.pbc) == (1, 0, 1)
    assert abs(p.fz - 1e-7) < 1e-20 and abs(p.rho_b - 1.5) < 1e-15

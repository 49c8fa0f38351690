"""The Fortran host on the CPU: its namelist reading, compile-time switches,
derived parameters and the iso_c_binding layout of lbf_params.

The host prints its parameter block before it touches the device, so these
run without a GPU (the run then stops at lbf_create with 'no HIP device').
The derived parameters are those of LBFAST.f90:212-219; the switch
implications those of defines.h:38-74.
"""
import ctypes
import os
import subprocess

import pytest

from lbfast_amd.params import LbfParams, make_params, PREC_FP64, PREC_FP32, PREC_FP16S
from tests import fhost

pytestmark = pytest.mark.skipif(not os.path.exists(fhost.exe("tg27h_fp32")),
                                reason="Fortran host not built (amdflang absent)")


def _ran_or_no_device(r):
    assert r.returncode == 0 or (r.returncode == 2 and "no HIP device" in r.stdout), r.stdout + r.stderr


def test_deck_and_derived_parameters(tmp_path):
    sim = {"nsteps": 0, "lx": 24, "ly": 20, "lz": 16, "lprint": False, "lraw": False, "iprobe": 1, "jprobe": 10,
           "kprobe": 8, "pbc_x": 1, "pbc_y": 1, "pbc_z": 1}
    fl = {"visc1": 0.07, "visc2": 0.03, "sigma": 0.02, "width": 4.5, "tau_diff": 0.11, "rho_r": 1.0, "rho_b": 0.1,
          "radius": 5.0, "center": [12.0, 10.0, 8.0], "fx": 1e-06}
    r = fhost.run("laplace27h_fp64", tmp_path, fhost.deck(sim, fl))
    _ran_or_no_device(r)
    h = fhost.header(r.stdout)
    assert h["n"] == (24, 20, 16)
    assert (h["lattice"], h["highorder"], h["twocomponent"], h["precision"]) == (27, 1, 1, PREC_FP64)
    assert (h["densratio"], h["incomp"], h["obstacles"], h["bounce_back"]) == (1, 1, 0, 0)
    p = make_params((24, 20, 16), lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.07,
                    visc2=0.03, sigma=0.02, width=4.5, tau_diff=0.11, rho_r=1.0, rho_b=0.1, precision=PREC_FP64)
    for k in ("tau1", "omega", "tau2", "beta", "kapp"):
        assert abs(h[k] - getattr(p, k)) <= 1e-7 * abs(getattr(p, k)), k


@pytest.mark.parametrize("variant,want", [
    ("tg27h_fp32", dict(lattice=27, highorder=1, twocomponent=0, precision=PREC_FP32, obstacles=0, densratio=0)),
    ("tg19_fp32", dict(lattice=19, highorder=0, twocomponent=0, precision=PREC_FP32)),
    ("laplace27h_fp16", dict(lattice=27, twocomponent=1, densratio=1, incomp=1, precision=PREC_FP16S)),
    ("pois19_bb_fp64", dict(lattice=19, twocomponent=0, obstacles=1, bounce_back=1, pbc=11, precision=PREC_FP64)),
])
def test_switch_implications(tmp_path, variant, want):
    sim = {"nsteps": 0, "lx": 16, "ly": 8, "lz": 8, "lprint": False, "pbc_x": 0 if "pois" in variant else 1}
    fl = {"visc1": 0.1, "width": 4.0, "radius": 3.0}
    if want["twocomponent"]:
        fl.update(sigma=0.01, visc2=0.1)  # names of the &fluid group only in two-component builds
    r = fhost.run(variant, tmp_path, fhost.deck(sim, fl))
    _ran_or_no_device(r)
    h = fhost.header(r.stdout)
    for k, v in want.items():
        assert h[k] == v, (k, h[k], v)


def test_open_face_keys_reach_the_parameters(tmp_path):
    """The reference's openbc keys (LBFAST.f90:33-37): one (low, high) pair
    per axis, passed on in the face order x(1) x(2) y(1) y(2) z(1) z(2)."""
    sim = {"nsteps": 0, "lx": 12, "ly": 16, "lz": 8, "lprint": False, "pbc_x": 0, "pbc_y": 0, "pbc_z": 1,
           "openbc": 1, "openbc_type_x": [0, 2], "openbc_type_y": [1, 1], "openbc_press_y": [1e-4, 0.0]}
    r = fhost.run("pois19_fp64", tmp_path, fhost.deck(sim, {"visc1": 0.1}))
    _ran_or_no_device(r)
    h = fhost.header(r.stdout)
    assert h["openbc"] == 1 and h["openbc_type"] == (0, 2, 1, 1, 0, 0)


def test_namelist_rejects_names_of_other_builds(tmp_path):
    # sigma belongs to &fluid only when TWOCOMPONENT is compiled (LBFAST.f90:39-59)
    text = fhost.deck({"nsteps": 0, "lx": 8, "ly": 8, "lz": 8, "lprint": False}, {"visc1": 0.1, "sigma": 0.01})
    r = fhost.run("tg27h_fp32", tmp_path, text)
    assert r.returncode != 0 and "sigma" in r.stdout + r.stderr


def test_command_line_forms(tmp_path):
    text = fhost.deck({"nsteps": 0, "lx": 8, "ly": 8, "lz": 8, "lprint": False}, {"visc1": 0.1, "uwall": 0.01})
    r = fhost.run("tg27h_fp32", tmp_path, text, args=["1", "1", "1", "input.inp"])
    _ran_or_no_device(r)
    assert fhost.header(r.stdout)["n"] == (8, 8, 8)
    r = fhost.run("tg27h_fp32", tmp_path, text, args=["2", "1", "1", "input.inp"])
    assert r.returncode == 2 and "decomposition must be 1 1 1" in r.stdout
    r = fhost.run("tg27h_fp32", tmp_path, text, args=["missing.inp"])
    assert r.returncode == 2 and "does not exist" in r.stdout


LAYOUT_PROGRAM = """
program layout
  use iso_c_binding
  use lbfast_hip
  implicit none
  type(lbf_params), target :: p
  integer(c_intptr_t) :: b
  b = transfer(c_loc(p), b)
  write (6, '(a,i0)') 'size ', c_sizeof(p)
  write (6, '(a,i0)') 'nx ', transfer(c_loc(p%nx), b) - b
  write (6, '(a,i0)') 'pbc ', transfer(c_loc(p%pbc), b) - b
  write (6, '(a,i0)') 'precision ', transfer(c_loc(p%precision), b) - b
  write (6, '(a,i0)') 'bounce_back ', transfer(c_loc(p%bounce_back), b) - b
  write (6, '(a,i0)') 'repulsive_flux ', transfer(c_loc(p%repulsive_flux), b) - b
  write (6, '(a,i0)') 'fx ', transfer(c_loc(p%fx), b) - b
  write (6, '(a,i0)') 'wettab_r ', transfer(c_loc(p%wettab_r), b) - b
  write (6, '(a,i0)') 'tau1 ', transfer(c_loc(p%tau1), b) - b
  write (6, '(a,i0)') 'sharp_c ', transfer(c_loc(p%sharp_c), b) - b
  write (6, '(a,i0)') 'store_scale ', transfer(c_loc(p%store_scale), b) - b
  write (6, '(a,i0)') 'a_rep ', transfer(c_loc(p%a_rep), b) - b
  write (6, '(a,i0)') 'reserved_d ', transfer(c_loc(p%reserved_d), b) - b
  write (6, '(a,i0)') 'openbc ', transfer(c_loc(p%openbc), b) - b
  write (6, '(a,i0)') 'openbc_type ', transfer(c_loc(p%openbc_type), b) - b
  write (6, '(a,i0)') 'openbc_press ', transfer(c_loc(p%openbc_press), b) - b
  write (6, '(a,i0)') 'openbc_w ', transfer(c_loc(p%openbc_w), b) - b
end program
"""


@pytest.mark.skipif(not os.path.exists("/opt/rocm/bin/amdflang"), reason="amdflang absent")
def test_fortran_params_layout_matches_c(tmp_path):
    src = tmp_path / "layout.f90"
    src.write_text(LAYOUT_PROGRAM)
    mod = os.path.join(fhost.ROOT, "lbfast_amd", "fortran", "lbfast_hip.f90")
    fl = "/opt/rocm/bin/amdflang"
    subprocess.run([fl, "-c", mod, "-J", str(tmp_path), "-o", str(tmp_path / "m.o")], check=True)
    libdir = os.path.join(fhost.ROOT, "lbfast_amd")
    subprocess.run([fl, str(src), str(tmp_path / "m.o"), "-I", str(tmp_path), "-o", str(tmp_path / "layout"),
                    f"-L{libdir}", "-llbfast_hip", f"-Wl,-rpath,{libdir}"], check=True)
    out = subprocess.run([str(tmp_path / "layout")], capture_output=True, text=True, check=True).stdout
    got = {k: int(v) for k, v in (line.split() for line in out.strip().splitlines())}
    assert got.pop("size") == ctypes.sizeof(LbfParams) == 544
    for k, off in got.items():
        assert getattr(LbfParams, k).offset == off, k

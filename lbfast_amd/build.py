"""Build liblbfast_hip.so in-tree with hipcc for gfx950, and the Fortran host.

Each translation unit is compiled in parallel, then linked.  Flags:
-O3, gfx950 only, and packed-FP32 instruction selection disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``).

The Fortran host (fortran/lbfast_host.f90 over the iso_c_binding module
fortran/lbfast_hip.f90) is compiled with amdflang once per switch set in
FORTRAN_VARIANTS, as the reference is compiled once per defines.h
selection; the executables go to fortran/bin/ and find the library through
an $ORIGIN-relative rpath.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblbfast_hip.so")
SOURCES = ["lbf_fused.hip", "lbf_stages.hip", "lbf_api.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FBIN = os.path.join(FDIR, "bin")
# name -> preprocessor switches of lbfast_defines.h
FORTRAN_VARIANTS = {
    "tg27h_fp32": ["-DTAYLORGREEN", "-DLATTICE=27", "-DHIGHORDER", "-DPRC=4"],
    "tg27h_fp64": ["-DTAYLORGREEN", "-DLATTICE=27", "-DHIGHORDER", "-DPRC=8"],
    "tg19_fp32": ["-DTAYLORGREEN", "-DLATTICE=19", "-DPRC=4"],
    "laplace27h_fp32": ["-DLAPLACE", "-DLATTICE=27", "-DHIGHORDER", "-DPRC=4"],
    "laplace27h_fp64": ["-DLAPLACE", "-DLATTICE=27", "-DHIGHORDER", "-DPRC=8"],
    "laplace27h_fp16": ["-DLAPLACE", "-DLATTICE=27", "-DHIGHORDER", "-DSTORE_FP16"],
    "pois19_bb_fp64": ["-DPOISEUILLE", "-DLATTICE=19", "-DBOUNCE_BACK", "-DPRC=8"],
    "pois19_fp64": ["-DPOISEUILLE", "-DLATTICE=19", "-DPRC=8", "-DDOXDMF"],
    "twopois27h_fp64": ["-DTWOPOISEUILLE", "-DLATTICE=27", "-DHIGHORDER", "-DPRC=8", "-DDOXDMF", "-DWRITEPRESS"],
    "lamb27h_fp32": ["-DLAMBTEST", "-DLATTICE=27", "-DHIGHORDER", "-DPRC=4", "-DPRINTPHI"],
    "twopois27h_bcphi_fp64": ["-DTWOPOISEUILLE", "-DLATTICE=27", "-DHIGHORDER", "-DPRC=8", "-DBCPHIFLUX"],
}
ARCH = "gfx950"
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-Wno-unused-result"]


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, "-c", "-x", "hip", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    if os.path.exists(FLANG):
        build_fortran(verbose)
    elif verbose:
        print("amdflang not found: Fortran host not built")
    return LIB


def _flang(args: list, what: str) -> None:
    r = subprocess.run([FLANG, *args], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for {what}:\n{r.stderr[-6000:]}")


def _fortran_variant(name: str) -> str:
    mod = os.path.join(BUILD, "fortran", name)
    os.makedirs(mod, exist_ok=True)
    defs = FORTRAN_VARIANTS[name]
    common = ["-cpp", "-O2", *defs, f"-I{FDIR}", "-J", mod]
    m_o = os.path.join(mod, "lbfast_hip.o")
    h_o = os.path.join(mod, "lbfast_host.o")
    _flang([*common, "-c", os.path.join(FDIR, "lbfast_hip.f90"), "-o", m_o], f"{name}: module")
    _flang([*common, "-c", os.path.join(FDIR, "lbfast_host.f90"), "-o", h_o], f"{name}: host")
    exe = os.path.join(FBIN, "lbfast_host_" + name)
    _flang([m_o, h_o, "-o", exe + ".tmp", f"-L{HERE}", "-llbfast_hip", "-Wl,-rpath,$ORIGIN/../.."],
           f"{name}: link")
    os.replace(exe + ".tmp", exe)
    return exe


def build_fortran(verbose: bool = False) -> list:
    """Compile the Fortran host for every switch set in FORTRAN_VARIANTS."""
    os.makedirs(FBIN, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(FORTRAN_VARIANTS), os.cpu_count() or 4)) as ex:
        exes = list(ex.map(_fortran_variant, FORTRAN_VARIANTS))
    if verbose:
        print("built", len(exes), "Fortran host variants in", FBIN)
    return exes


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

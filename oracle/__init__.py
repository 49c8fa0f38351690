"""The oracle: TEST INFRASTRUCTURE ONLY.

A CPU restatement of the reference time step (lbf_oracle.cpp) wrapped with
ctypes.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
may import this package, and only as the checker / CPU baseline -- never as
the thing measured or shipped.  The product library (lbfast_amd) does not
import it and fails loudly when its HIP extension is missing.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from lbfast_amd.params import LbfParams, HFIELDS, FIELD_PHI

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liblbf_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++, OpenMP)."""
    if force and os.path.exists(_LIB):
        os.remove(_LIB)
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        P = ctypes.POINTER(LbfParams)
        dp = ctypes.POINTER(ctypes.c_double)
        bp = ctypes.POINTER(ctypes.c_int8)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [P]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for name in ("orc_set_field", "orc_get_field"):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
            getattr(L, name).restype = ctypes.c_int
        L.orc_set_isfluid.argtypes = [ctypes.c_void_p, bp]
        for name in ("orc_init_state", "orc_begin_step", "orc_end_step"):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_stage.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_populations.argtypes = [ctypes.c_void_p, dp]
        L.orc_classify_isfluid.argtypes = [P, bp, bp]
        L.orc_derive.argtypes = [P]
        L.orc_minv.restype = ctypes.c_double
        L.orc_minv.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_m.restype = ctypes.c_double
        L.orc_m.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_time_steps.restype = ctypes.c_double
        L.orc_time_steps.argtypes = [ctypes.c_void_p, ctypes.c_int]
        _lib = L
    return _lib


def minv(k: int, l: int) -> float:
    return lib().orc_minv(k, l)


def mmat(k: int, l: int) -> float:
    return lib().orc_m(k, l)


def classify_isfluid(p: LbfParams, raw=None) -> np.ndarray:
    shape = (p.nz, p.ny, p.nx)
    out = np.empty(shape, dtype=np.int8)
    if raw is None:
        inp = None
    else:
        inp = np.ascontiguousarray(raw, dtype=np.int8)
    lib().orc_classify_isfluid(ctypes.byref(p),
                               inp.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)) if inp is not None else None,
                               out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
    return out


class Oracle:
    """CPU reference simulator (precision from ``p.precision``: FP64, FP32 or MIXED)."""

    def __init__(self, p: LbfParams):
        self.p = p.copy()
        self.h = lib().orc_create(ctypes.byref(self.p))
        if not self.h:
            raise ValueError("oracle rejected parameters: " + self.p.describe())
        self.shape = (p.nz, p.ny, p.nx)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_destroy(h)
            self.h = None

    def set_field(self, fid: int, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        rc = lib().orc_set_field(self.h, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc:
            raise RuntimeError(f"orc_set_field({fid}) -> {rc}")

    def get_field(self, fid: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        rc = lib().orc_get_field(self.h, fid, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc:
            raise RuntimeError(f"orc_get_field({fid}) -> {rc}")
        return out

    def set_isfluid(self, isf):
        a = np.ascontiguousarray(isf, dtype=np.int8).reshape(self.shape)
        lib().orc_set_isfluid(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))

    def load(self, fields: dict, isfluid=None):
        """Upload initial fields (dict from lbfast_amd.setups) and run the
        reference's initialisation (O := S, bcs_phi, halos)."""
        names = ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")
        if isfluid is not None:
            self.set_isfluid(isfluid)
        for fid, nm in zip(HFIELDS, names):
            self.set_field(fid, fields[nm])
        if "phi" in fields:
            self.set_field(FIELD_PHI, fields["phi"])
        lib().orc_init_state(self.h)

    def step(self, n: int = 1):
        lib().orc_step(self.h, int(n))

    def time_steps(self, n: int) -> float:
        return lib().orc_time_steps(self.h, int(n))

    def begin_step(self):
        lib().orc_begin_step(self.h)

    def stage(self, sid: int):
        rc = lib().orc_stage(self.h, sid)
        if rc:
            raise RuntimeError(f"orc_stage({sid}) -> {rc}")

    def end_step(self):
        lib().orc_end_step(self.h)

    def populations(self) -> np.ndarray:
        nl = 27
        out = np.empty((nl,) + self.shape, dtype=np.float64)
        rc = lib().orc_populations(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc:
            raise RuntimeError("orc_populations -> %d" % rc)
        return out

    def hfields(self) -> np.ndarray:
        return np.stack([self.get_field(f) for f in HFIELDS])

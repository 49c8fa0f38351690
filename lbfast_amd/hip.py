"""ctypes binding of liblbfast_hip.so -- the host-side mirror of the
reference driver interface (lb_cuda_driver_module.f90) over the C-ABI of
include/lbfast_hip.h.

Loading fails loudly when the in-tree library is missing: there is no CPU
fallback on the product path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .params import (LbfParams, HFIELDS, HFIELD_NAMES, FIELD_PHI, FIELD_UOLD, STEP_ORDER)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblbfast_hip.so")
_lib = None

# exported symbols (include/lbfast_hip.h); tests check the library exports all
EXPORTS = (
    "lbf_abi_version", "lbf_params_default", "lbf_params_derive", "lbf_classify_isfluid",
    "lbf_create", "lbf_destroy", "lbf_last_error", "lbf_global_last_error",
    "lbf_upload_field", "lbf_download_field", "lbf_upload_isfluid", "lbf_init_state",
    "lbf_step", "lbf_begin_step", "lbf_stage", "lbf_end_step", "lbf_reduce_fp64",
    "lbf_synchronize", "lbf_get_step", "lbf_field_device_ptr", "lbf_last_step_timing",
    "lbf_set_tile", "lbf_set_timing", "lbf_reduce_fp64_noalloc",
    "lbf_batch_create", "lbf_batch_destroy", "lbf_batch_last_error", "lbf_batch_step", "lbf_batch_capture",
    "lbf_batch_set_tile", "lbf_batch_autotune", "lbf_batch_synchronize", "lbf_batch_timing",
    "lbf_batch_set_order", "lbf_batch_tiles",
)


class LbfError(RuntimeError):
    pass


ABI_VERSION = 2  # LBF_ABI_VERSION of include/lbfast_hip.h (lbf_params with the open-boundary block)


def load(path: str = LIB_PATH):
    """Load the HIP library (once).  Raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LbfError(f"{path} is missing: build it with `python -m lbfast_amd.build` "
                       "(there is no CPU fallback)")
    L = ctypes.CDLL(path)
    L.lbf_abi_version.restype = ctypes.c_int32
    if L.lbf_abi_version() != ABI_VERSION:
        raise LbfError(f"{path} has ABI version {L.lbf_abi_version()}, these bindings expect {ABI_VERSION}: "
                       "rebuild it with `python -m lbfast_amd.build`")
    P = ctypes.POINTER(LbfParams)
    vp, ip8 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int8)
    L.lbf_abi_version.restype = ctypes.c_int32
    L.lbf_params_default.argtypes = [P]
    L.lbf_params_derive.argtypes = [P]
    L.lbf_classify_isfluid.argtypes = [P, ip8, ip8]
    L.lbf_create.argtypes = [P, ctypes.POINTER(vp)]
    L.lbf_destroy.argtypes = [vp]
    L.lbf_last_error.argtypes = [vp]
    L.lbf_last_error.restype = ctypes.c_char_p
    L.lbf_global_last_error.restype = ctypes.c_char_p
    L.lbf_upload_field.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
    L.lbf_download_field.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
    L.lbf_upload_isfluid.argtypes = [vp, ip8]
    for n in ("lbf_init_state", "lbf_begin_step", "lbf_end_step", "lbf_synchronize"):
        getattr(L, n).argtypes = [vp]
    L.lbf_step.argtypes = [vp, ctypes.c_int32]
    L.lbf_stage.argtypes = [vp, ctypes.c_int]
    L.lbf_reduce_fp64.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.lbf_reduce_fp64_noalloc.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    L.lbf_get_step.argtypes = [vp]
    L.lbf_get_step.restype = ctypes.c_int32
    L.lbf_field_device_ptr.argtypes = [vp, ctypes.c_int]
    L.lbf_field_device_ptr.restype = vp
    L.lbf_last_step_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.lbf_set_tile.argtypes = [vp, ctypes.c_int32, ctypes.c_int32]
    L.lbf_set_timing.argtypes = [vp, ctypes.c_int32]
    i32, pi32 = ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)
    L.lbf_batch_create.argtypes = [ctypes.POINTER(vp), i32, ctypes.POINTER(vp)]
    L.lbf_batch_destroy.argtypes = [vp]
    L.lbf_batch_last_error.argtypes = [vp]
    L.lbf_batch_last_error.restype = ctypes.c_char_p
    L.lbf_batch_step.argtypes = [vp, i32]
    L.lbf_batch_capture.argtypes = [vp, i32]
    L.lbf_batch_set_tile.argtypes = [vp, i32, i32, i32]
    L.lbf_batch_autotune.argtypes = [vp, i32, pi32, pi32]
    L.lbf_batch_set_order.argtypes = [vp, pi32]
    L.lbf_batch_tiles.argtypes = [vp, pi32, pi32, pi32, pi32]
    L.lbf_batch_synchronize.argtypes = [vp]
    L.lbf_batch_timing.argtypes = [vp, ctypes.POINTER(ctypes.c_double)]
    _lib = L
    return L


def classify_isfluid(p: LbfParams, raw=None) -> np.ndarray:
    """LBFAST.f90:258-328 classification (host function of the library)."""
    L = load()
    out = np.empty((p.nz, p.ny, p.nx), dtype=np.int8)
    inp = None if raw is None else np.ascontiguousarray(raw, dtype=np.int8)
    rc = L.lbf_classify_isfluid(ctypes.byref(p),
                                None if inp is None else inp.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)),
                                out.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)))
    if rc:
        raise LbfError(f"lbf_classify_isfluid -> {rc}")
    return out


class Domain:
    """One LB domain on the current HIP device (one context, one stream)."""

    def __init__(self, p: LbfParams):
        self.L = load()
        self.p = p.copy()
        h = ctypes.c_void_p()
        rc = self.L.lbf_create(ctypes.byref(self.p), ctypes.byref(h))
        if rc:
            raise LbfError(f"lbf_create -> {rc}: {self.L.lbf_global_last_error().decode()}")
        self.h = h
        self.shape = (p.nz, p.ny, p.nx)

    def close(self):
        if getattr(self, "h", None):
            self.L.lbf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc:
            raise LbfError(f"{what} -> {rc}: {self.L.lbf_last_error(self.h).decode()}")

    # ---- data movement ----------------------------------------------------
    def upload(self, fid: int, a, dtype=np.float64):
        a = np.ascontiguousarray(a, dtype=dtype).reshape(self.shape)
        self._chk(self.L.lbf_upload_field(self.h, fid, a.ctypes.data, 0 if dtype == np.float64 else 1),
                  f"upload({fid})")

    def download(self, fid: int, dtype=np.float64) -> np.ndarray:
        out = np.empty(self.shape, dtype=dtype)
        self._chk(self.L.lbf_download_field(self.h, fid, out.ctypes.data, 0 if dtype == np.float64 else 1),
                  f"download({fid})")
        return out

    def upload_isfluid(self, isf):
        a = np.ascontiguousarray(isf, dtype=np.int8).reshape(self.shape)
        self._chk(self.L.lbf_upload_isfluid(self.h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int8))),
                  "upload_isfluid")

    def load(self, fields: dict, isfluid=None):
        """Upload initial fields (from lbfast_amd.setups) and initialise."""
        if isfluid is not None:
            self.upload_isfluid(isfluid)
        for fid, nm in zip(HFIELDS, HFIELD_NAMES):
            self.upload(fid, fields[nm])
        if "phi" in fields and self.p.twocomponent:
            self.upload(FIELD_PHI, fields["phi"])
        self.init_state()

    def init_state(self):
        self._chk(self.L.lbf_init_state(self.h), "init_state")

    # ---- time stepping ----------------------------------------------------
    def step(self, n: int = 1):
        self._chk(self.L.lbf_step(self.h, int(n)), "step")

    def run(self, nsteps: int):
        """Run like the reference loop, which advances in FLIP/FLOP pairs
        (integrator_module.f90:266): 2*ceil(nsteps/2) steps."""
        self.step(2 * ((nsteps + 1) // 2))

    def begin_step(self):
        self._chk(self.L.lbf_begin_step(self.h), "begin_step")

    def stage(self, sid: int):
        self._chk(self.L.lbf_stage(self.h, sid), f"stage({sid})")

    def end_step(self):
        self._chk(self.L.lbf_end_step(self.h), "end_step")

    def step_staged(self, n: int = 1):
        """The reference stage sequence, one kernel per stage."""
        for _ in range(n):
            self.begin_step()
            for sid in STEP_ORDER:
                self.stage(sid)
            self.end_step()

    def synchronize(self):
        self._chk(self.L.lbf_synchronize(self.h), "synchronize")

    def set_tile(self, ty: int = 0, tz: int = 0):
        self._chk(self.L.lbf_set_tile(self.h, ty, tz), "set_tile")

    def set_timing(self, on: bool = True):
        """Bracket every fused launch with a HIP event pair (lbf_set_timing)."""
        self._chk(self.L.lbf_set_timing(self.h, 1 if on else 0), "set_timing")

    def timing(self):
        """(total ms, fused-kernel ms) of the last step call; the second is -1
        unless set_timing(True) was called."""
        t, f = ctypes.c_double(), ctypes.c_double()
        self._chk(self.L.lbf_last_step_timing(self.h, ctypes.byref(t), ctypes.byref(f)), "timing")
        return t.value, f.value

    def reduce(self) -> np.ndarray:
        out = (ctypes.c_double * 8)()
        self._chk(self.L.lbf_reduce_fp64(self.h, out), "reduce")
        return np.array(out[:])

    def reduce_noalloc(self) -> np.ndarray:
        """lbf_reduce_fp64_noalloc: the same sums through the library's own
        device array (no allocation)."""
        out = (ctypes.c_double * 8)()
        self._chk(self.L.lbf_reduce_fp64_noalloc(self.h, out), "reduce_noalloc")
        return np.array(out[:])

    @property
    def step_count(self) -> int:
        return self.L.lbf_get_step(self.h)

    def hfields(self, dtype=np.float64) -> np.ndarray:
        return np.stack([self.download(f, dtype) for f in HFIELDS])

    def uold(self) -> np.ndarray:
        return np.stack([self.download(f) for f in FIELD_UOLD])

    def device_ptr(self, fid: int) -> int:
        return self.L.lbf_field_device_ptr(self.h, fid) or 0


class Batch:
    """Independent domains advanced together by one launch per kernel and step
    (lbf_batch_*).  The domains must outlive the batch."""

    def __init__(self, domains):
        self.L = load()
        self.domains = list(domains)
        arr = (ctypes.c_void_p * len(self.domains))(*[d.h.value for d in self.domains])
        h = ctypes.c_void_p()
        rc = self.L.lbf_batch_create(arr, len(self.domains), ctypes.byref(h))
        if rc:
            raise LbfError(f"lbf_batch_create -> {rc}: {self.L.lbf_global_last_error().decode()}")
        self.h = h

    def _chk(self, rc, what):
        if rc:
            raise LbfError(f"{what} -> {rc}: {self.L.lbf_batch_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            self.L.lbf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        self._chk(self.L.lbf_batch_step(self.h, n), "lbf_batch_step")

    def capture(self, n: int):
        self._chk(self.L.lbf_batch_capture(self.h, n), "lbf_batch_capture")

    def set_tile(self, ty: int, tz: int, tz_prep: int = None):
        self._chk(self.L.lbf_batch_set_tile(self.h, ty, tz, tz if tz_prep is None else tz_prep), "lbf_batch_set_tile")

    def autotune(self, trial_steps: int = 4):
        ty, tz = ctypes.c_int32(), ctypes.c_int32()
        self._chk(self.L.lbf_batch_autotune(self.h, trial_steps, ctypes.byref(ty), ctypes.byref(tz)),
                  "lbf_batch_autotune")
        return ty.value, tz.value

    def set_order(self, order):
        arr = (ctypes.c_int32 * len(self.domains))(*[int(x) for x in order])
        self._chk(self.L.lbf_batch_set_order(self.h, arr), "lbf_batch_set_order")

    def tiles(self):
        """(row tile, per-domain fused z chunks, per-domain prep z chunks, table order)."""
        n = len(self.domains)
        ty = ctypes.c_int32()
        tz, tzp, order = (ctypes.c_int32 * n)(), (ctypes.c_int32 * n)(), (ctypes.c_int32 * n)()
        self._chk(self.L.lbf_batch_tiles(self.h, ctypes.byref(ty), tz, tzp, order), "lbf_batch_tiles")
        return ty.value, list(tz), list(tzp), list(order)

    def synchronize(self):
        self._chk(self.L.lbf_batch_synchronize(self.h), "lbf_batch_synchronize")

    def timing(self) -> float:
        ms = ctypes.c_double()
        self._chk(self.L.lbf_batch_timing(self.h, ctypes.byref(ms)), "lbf_batch_timing")
        return ms.value

"""Host-side parameter block: a ctypes mirror of ``lbf_params`` (include/lbfast_hip.h).

It carries the ``&simulation`` / ``&fluid`` namelist values the time step reads
(LBFAST.f90:32-60), the ``defines.h`` switches (LATTICE, HIGHORDER,
TWOCOMPONENT, DENSRATIO, INTERFACE_INCOMP, INTERNAL_OBSTACLES, PRC/STRPRC) and
the derived quantities of LBFAST.f90:212-219.
"""
from __future__ import annotations

import ctypes
import re

PREC_FP64, PREC_FP32, PREC_MIXED, PREC_FP16S, PREC_BF16S = 0, 1, 2, 3, 4
PREC_NAMES = {"fp64": PREC_FP64, "fp32": PREC_FP32, "mixed": PREC_MIXED,
              "fp16s": PREC_FP16S, "bf16s": PREC_BF16S}

# field ids (include/lbfast_hip.h)
FIELD_P, FIELD_U, FIELD_V, FIELD_W = 1, 2, 3, 4
FIELD_PXX, FIELD_PYY, FIELD_PZZ, FIELD_PXY, FIELD_PXZ, FIELD_PYZ = 5, 6, 7, 8, 9, 10
FIELD_PHI = 11
FIELD_UOLD = (12, 13, 14)
FIELD_FORCE = (15, 16, 17)
FIELD_AUX = tuple(range(18, 25))
FIELD_LAPPHI, FIELD_DIVTHN = 25, 26
FIELD_O = tuple(range(27, 37))
FIELD_PHI_O = 37
FIELD_RHOPHI = 40
FIELD_REP_PARTNER = (41, 42, 43)  # locaux 3..5
FIELD_REP_FLUX = (44, 45, 46)     # locaux 6..8
FIELD_REP_MASK = 47
HFIELDS = tuple(range(1, 11))
HFIELD_NAMES = ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")

STAGE_NORM, STAGE_PHI_INT_BOUNDARY, STAGE_DIV_THETA_N = 1, 2, 3
STAGE_MOMENTS, STAGE_FUSED, STAGE_UPDATE_PHI, STAGE_LB_INT_BOUNDARY = 4, 5, 6, 7
STAGE_THINFILM_SCAN, STAGE_REPULSIVE_FLUX, STAGE_PHI_LAGRANGE = 8, 9, 10
STAGE_OPEN_BC_PHI, STAGE_OPEN_BC_H = 11, 12
# integrator_module.f90:278-524 (one FLIP half step)
# integrator_module.f90:266-545; the repulsive stages, the phase-mass
# correction, the open faces and LB_int_boundary are no-ops unless
# REPULSIVE_FLUX / BCPHIFLUX / openbc / INTERNAL_OBSTACLES without BOUNCE_BACK
# (bcs_mesoscopic_phifields runs PHI_int_boundary, the open faces' phase field,
# then the BCPHIFLUX correction: boundary_cds_module.f90:1347-1887;
# bcs_mesoscopic_hfields runs LB_int_boundary, then the open faces: 49-1311)
STEP_ORDER = (STAGE_NORM, STAGE_PHI_INT_BOUNDARY, STAGE_OPEN_BC_PHI, STAGE_PHI_LAGRANGE, STAGE_DIV_THETA_N,
              STAGE_THINFILM_SCAN, STAGE_REPULSIVE_FLUX, STAGE_MOMENTS, STAGE_FUSED, STAGE_UPDATE_PHI,
              STAGE_LB_INT_BOUNDARY, STAGE_OPEN_BC_H)


class LbfParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("lattice", ctypes.c_int32), ("highorder", ctypes.c_int32),
        ("twocomponent", ctypes.c_int32), ("densratio", ctypes.c_int32),
        ("interface_incomp", ctypes.c_int32), ("internal_obstacles", ctypes.c_int32),
        ("buoyancy_forcing", ctypes.c_int32), ("pbc", ctypes.c_int32 * 3),
        ("precision", ctypes.c_int32), ("bounce_back", ctypes.c_int32),
        ("smagorinsky", ctypes.c_int32), ("repulsive_flux", ctypes.c_int32), ("bcphiflux", ctypes.c_int32),
        ("fx", ctypes.c_double), ("fy", ctypes.c_double), ("fz", ctypes.c_double),
        ("visc1", ctypes.c_double), ("visc2", ctypes.c_double),
        ("rho_r", ctypes.c_double), ("rho_b", ctypes.c_double),
        ("tau_diff", ctypes.c_double), ("sigma", ctypes.c_double), ("width", ctypes.c_double),
        ("wettab_r", ctypes.c_double),
        ("tau1", ctypes.c_double), ("omega", ctypes.c_double), ("tau2", ctypes.c_double),
        ("beta", ctypes.c_double), ("kapp", ctypes.c_double), ("sharp_c", ctypes.c_double),
        ("store_scale", ctypes.c_double * 10),
        ("a_rep", ctypes.c_double), ("q_th", ctypes.c_double), ("cos_opp_t", ctypes.c_double),
        ("reserved_d", ctypes.c_double),
        # open boundaries; faces x(1) x(2) y(1) y(2) z(1) z(2)
        ("openbc", ctypes.c_int32), ("openbc_type", ctypes.c_int32 * 6), ("reserved_i2", ctypes.c_int32),
        ("openbc_press", ctypes.c_double * 6), ("openbc_u", ctypes.c_double * 6),
        ("openbc_v", ctypes.c_double * 6), ("openbc_w", ctypes.c_double * 6),
    ]

    @property
    def shape(self):
        return (self.nx, self.ny, self.nz)

    @property
    def nodes(self) -> int:
        return self.nx * self.ny * self.nz

    def copy(self) -> "LbfParams":
        c = LbfParams()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(LbfParams))
        return c

    def describe(self) -> str:
        return (f"D3Q{self.lattice}{'h' if self.highorder else ''} "
                f"{'2c' if self.twocomponent else '1c'}{' densratio' if self.densratio else ''}"
                f"{' obstacles' if self.internal_obstacles else ''} {self.nx}x{self.ny}x{self.nz} "
                f"prec={self.precision}")


def derive(p: LbfParams) -> LbfParams:
    """Derived parameters exactly as LBFAST.f90:212-219."""
    p.tau1 = p.visc1 * 3.0 + 0.5
    p.omega = 1.0 / p.tau1
    if p.twocomponent:
        p.beta = 12.0 * p.sigma / p.width
        p.kapp = 1.5 * p.sigma * p.width
        p.sharp_c = 4.0 * p.tau_diff / p.width
        p.tau2 = p.visc2 * 3.0 + 0.5
    return p


def make_params(n=(64, 64, 64), lattice=27, highorder=True, twocomponent=False,
                densratio=False, internal_obstacles=False, pbc=(1, 1, 1), precision=PREC_FP32,
                fx=0.0, fy=0.0, fz=0.0, visc1=0.1, visc2=0.1, rho_r=1.0, rho_b=1.0,
                tau_diff=0.1, sigma=0.0, width=4.0, wettab_r=90.0,
                interface_incomp=None, buoyancy_forcing=False, bounce_back=False, smagorinsky=False,
                repulsive_flux=False, a_rep=0.0, q_th=0.125, cos_opp_t=-0.8, bcphiflux=False,
                openbc=False, openbc_type=(0,) * 6, openbc_press=(0.0,) * 6, openbc_u=(0.0,) * 6,
                openbc_v=(0.0,) * 6, openbc_w=(0.0,) * 6) -> LbfParams:
    """Build a parameter block. ``interface_incomp`` defaults to ``densratio``
    because defines.h:72-74 defines INTERFACE_INCOMP whenever DENSRATIO is set.
    The open-boundary sequences are per face: x(1), x(2), y(1), y(2), z(1), z(2)."""
    p = LbfParams()
    p.nx, p.ny, p.nz = n
    p.lattice = lattice
    p.highorder = int(bool(highorder) and lattice == 27)
    p.twocomponent = int(bool(twocomponent))
    p.densratio = int(bool(densratio) and bool(twocomponent))
    if interface_incomp is None:
        interface_incomp = densratio
    p.interface_incomp = int(bool(interface_incomp) and p.densratio)
    p.internal_obstacles = int(bool(internal_obstacles))
    p.buoyancy_forcing = int(bool(buoyancy_forcing))
    for i in range(3):
        p.pbc[i] = int(pbc[i])
    p.precision = precision
    p.fx, p.fy, p.fz = fx, fy, fz
    p.visc1, p.visc2 = visc1, visc2
    p.rho_r, p.rho_b = rho_r, rho_b
    p.tau_diff, p.sigma, p.width = tau_diff, sigma, width
    p.wettab_r = wettab_r
    p.bounce_back = int(bool(bounce_back))
    p.smagorinsky = int(bool(smagorinsky))
    p.repulsive_flux = int(bool(repulsive_flux) and p.twocomponent)
    p.a_rep, p.q_th, p.cos_opp_t = a_rep, q_th, cos_opp_t
    p.bcphiflux = int(bool(bcphiflux))
    p.openbc = int(bool(openbc))
    for f in range(6):
        p.openbc_type[f] = int(openbc_type[f])
        p.openbc_press[f], p.openbc_u[f] = openbc_press[f], openbc_u[f]
        p.openbc_v[f], p.openbc_w[f] = openbc_v[f], openbc_w[f]
    return derive(p)


_NML_RE = re.compile(r"&(\w+)(.*?)(?:&end|/)", re.S | re.I)


def read_namelists(text: str) -> dict:
    """Parse the ``&simulation`` / ``&fluid`` namelist groups of a reference
    input deck (e.g. laplace.inp) into {group: {name: value}}.  Fortran
    comments (``!``) are stripped; logicals map to bool, numbers to int/float,
    comma lists to lists."""
    out = {}
    clean = "\n".join(line.split("!")[0] for line in text.splitlines())
    for grp, body in _NML_RE.findall(clean):
        vals = {}
        for m in re.finditer(r"(\w+)\s*=\s*([^=]+?)(?=,?\s*\w+\s*=|\s*$)", body.strip(), re.S):
            key, raw = m.group(1).lower(), m.group(2).strip().rstrip(",")
            items = [s.strip() for s in raw.split(",") if s.strip()]
            conv = [_conv(s) for s in items]
            vals[key] = conv[0] if len(conv) == 1 else conv
        out[grp.lower()] = vals
    return out


def _conv(s: str):
    ls = s.lower()
    if ls in (".true.", "t", ".t."):
        return True
    if ls in (".false.", "f", ".f."):
        return False
    try:
        return int(s)
    except ValueError:
        pass
    try:
        return float(ls.replace("d", "e"))
    except ValueError:
        return s.strip("'\"")


def params_from_namelist(text: str, lattice=27, highorder=True, twocomponent=True,
                         densratio=True, internal_obstacles=False,
                         precision=PREC_FP32, bounce_back=False, smagorinsky=False,
                         repulsive_flux=False) -> LbfParams:
    """Parameter block from a reference input deck plus defines.h switches."""
    nml = read_namelists(text)
    sim, fl = nml.get("simulation", {}), nml.get("fluid", {})
    return make_params(
        n=(sim.get("lx", 64), sim.get("ly", 64), sim.get("lz", 64)),
        lattice=lattice, highorder=highorder, twocomponent=twocomponent, densratio=densratio,
        internal_obstacles=internal_obstacles,
        pbc=(sim.get("pbc_x", 1), sim.get("pbc_y", 1), sim.get("pbc_z", 1)), precision=precision,
        fx=fl.get("fx", 0.0), fy=fl.get("fy", 0.0), fz=fl.get("fz", 0.0),
        visc1=fl.get("visc1", 0.1), visc2=fl.get("visc2", fl.get("visc1", 0.1)),
        rho_r=fl.get("rho_r", 1.0), rho_b=fl.get("rho_b", 1.0),
        tau_diff=fl.get("tau_diff", 0.1), sigma=fl.get("sigma", 0.0), width=fl.get("width", 4.0),
        wettab_r=fl.get("wettab_r", 90.0), bounce_back=bounce_back, smagorinsky=smagorinsky,
        repulsive_flux=repulsive_flux, a_rep=fl.get("a_rep", 0.0), q_th=fl.get("q_th", 0.125),
        cos_opp_t=fl.get("cosoppt", -0.8), openbc=bool(sim.get("openbc", 0)),
        **{f"openbc_{k}": _faces(sim, k) for k in ("type", "press", "u", "v", "w")})


def _faces(sim: dict, k: str) -> tuple:
    """openbc_<k>_x(1:2), _y(1:2), _z(1:2) of a deck as six values per face,
    x(1) x(2) y(1) y(2) z(1) z(2); an absent value is 0 (the declared defaults)."""
    out = []
    for ax in "xyz":
        v = sim.get(f"openbc_{k}_{ax}", [0, 0])
        v = list(v) if isinstance(v, (list, tuple)) else [v]  # a scalar sets element (1) only, as in Fortran
        out += [(int if k == "type" else float)(x) for x in (v + [0, 0])[:2]]
    return tuple(out)

"""Shared workload builders and comparison helpers for the tests."""
from __future__ import annotations

import numpy as np

from lbfast_amd import setups
from lbfast_amd.params import (make_params, PREC_FP64, PREC_FP32, PREC_MIXED, HFIELDS, FIELD_PHI,
                               FIELD_PHI_O, FIELD_UOLD)


def case(name: str, n=(32, 32, 32), precision=PREC_FP64, seed=1):
    """(params, fields, raw isfluid or None) for a named test workload.
    Suffixes: ``_bb`` BOUNCE_BACK, ``_smag`` SMAGORINSKI, ``_bcphi`` BCPHIFLUX."""
    for suffix, key in (("_bb", "bounce_back"), ("_smag", "smagorinsky"), ("_bcphi", "bcphiflux")):
        if name.endswith(suffix):
            p, f, raw = case(name[: -len(suffix)], n, precision, seed)
            setattr(p, key, 1)
            return p, f, raw
    if name == "film27h_rep":  # two facing interfaces 4 nodes apart, REPULSIVE_FLUX
        p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True,
                        visc1=0.1, visc2=0.1, rho_r=1.0, rho_b=0.1, tau_diff=0.1, sigma=0.01, width=2.0,
                        precision=precision, repulsive_flux=True, a_rep=0.05, q_th=0.125, cos_opp_t=-0.8)
        return p, setups.thin_film(n, 4.0, 2.0, axis=0, noise=1e-5, seed=seed), None
    if name == "tg19":  # Taylor-Green, D3Q19, one component
        p = make_params(n, lattice=19, highorder=False, visc1=(0.8 - 0.5) / 3, precision=precision)
        return p, setups.taylor_green(n, 0.02), None
    if name == "tg27":
        p = make_params(n, lattice=27, highorder=False, visc1=0.05 / 3, precision=precision)
        return p, setups.taylor_green(n, 0.04), None
    if name == "tg27h":
        p = make_params(n, lattice=27, highorder=True, visc1=0.05 / 3, precision=precision)
        return p, setups.taylor_green(n, 0.04), None
    if name == "tg27h_force":  # body force exercises the forcing terms
        p = make_params(n, lattice=27, highorder=True, visc1=0.05, precision=precision,
                        fx=1e-5, fy=-2e-5, fz=3e-5)
        return p, setups.taylor_green(n, 0.04), None
    if name == "tg19_force":
        p = make_params(n, lattice=19, highorder=False, visc1=0.1, precision=precision, fx=2e-5, fz=-1e-5)
        return p, setups.taylor_green(n, 0.03), None
    if name in ("drop27h", "drop19", "drop27"):
        lat = 19 if name == "drop19" else 27
        R = n[0] / 4.0
        p = make_params(n, lattice=lat, highorder=(name == "drop27h"), twocomponent=True, densratio=True,
                        visc1=0.1, visc2=0.1, rho_r=1.0, rho_b=0.1, tau_diff=0.1, sigma=0.01, width=3.0,
                        precision=precision)
        f = setups.droplet(n, R, 3.0, [n[0] / 2 + 0.5, n[1] / 2 + 0.5, n[2] / 2 + 0.5], 0.01, 1.0, 0.1,
                           noise=1e-4, seed=seed)
        return p, f, None
    if name in ("drop27h_bulk", "bubble27h_bulk"):
        # a droplet (or, inverted, a bubble) off centre whose phi is exactly 0
        # or 1 away from the interface
        R = 8.0
        p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True,
                        visc1=0.1, visc2=0.1, rho_r=1.0, rho_b=0.1, tau_diff=0.1, sigma=0.01, width=3.0,
                        precision=precision)
        f = setups.droplet(n, R, 3.0, [28.5, n[1] / 2 + 0.5, n[2] / 2 + 0.5], 0.01, 1.0, 0.1,
                           noise=1e-4, seed=seed)
        phi = np.where(f["phi"] < 1e-4, 0.0, np.where(f["phi"] > 1.0 - 1e-4, 1.0, f["phi"]))
        if name == "bubble27h_bulk":
            phi = 1.0 - phi
        rhophi = 1.0 * phi + (1.0 - phi) * 0.1
        f["phi"] = phi
        f["P"] = phi * 0.02 / R / (rhophi / 3.0)
        return p, f, None
    if name == "drop27h_nodens":
        R = n[0] / 4.0
        p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=False,
                        visc1=0.1, visc2=0.05, rho_r=1.0, rho_b=1.0, tau_diff=0.1, sigma=0.01, width=3.0,
                        precision=precision)
        f = setups.droplet(n, R, 3.0, [n[0] / 2, n[1] / 2, n[2] / 2], 0.01, 1.0, 1.0, densratio=False,
                           noise=1e-4, seed=seed)
        return p, f, None
    if name in ("chan27h", "chan19", "duct27h"):
        # walls in y (duct27h: in x and y, so that walls meet along edges),
        # spheres, two layers, body force
        lat = 19 if name == "chan19" else 27
        pbc = (0, 0, 1) if name == "duct27h" else (1, 0, 1)
        p = make_params(n, lattice=lat, highorder=(lat == 27), twocomponent=True, densratio=True,
                        internal_obstacles=True, pbc=pbc, visc1=0.05, visc2=0.1, rho_r=2.0, rho_b=1.0,
                        tau_diff=0.1, sigma=0.005, width=4.0, fx=0.0 if name == "duct27h" else 1e-5,
                        fz=1e-5 if name == "duct27h" else 0.0, precision=precision)
        raw = setups.random_spheres(n, fill=0.15, rmin=2.0, rmax=4.0, seed=seed, margin=2)
        f = setups.two_layer_channel(n, raw, frac=0.4, width=4.0, rho_r=2.0, rho_b=1.0, axis=1)
        return p, f, raw
    if name == "twolayer27h":  # walls in x, force in z, the two-layer start of TWOPOISEUILLE, no spheres
        p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, internal_obstacles=True,
                        pbc=(0, 1, 1), visc1=0.05, visc2=0.4, rho_r=0.5, rho_b=1.5, tau_diff=0.1, sigma=0.0,
                        width=4.0, fz=1e-6, precision=precision)
        f = setups.two_layer_channel(n, None, frac=0.5, width=4.0, rho_r=0.5, rho_b=1.5, axis=0)
        # velocity noise off the wall columns, so that every moment carries
        # signal rather than rounding residue
        rng = np.random.default_rng(seed)
        uvw = rng.uniform(-1e-4, 1e-4, (3, n[2], n[1], n[0]))
        uvw[:, :, :, 0] = 0.0
        uvw[:, :, :, -1] = 0.0
        f.update(setups._moments(f["P"], *uvw))
        return p, f, np.ones((n[2], n[1], n[0]), dtype=np.int8)
    if name in ("duct27h_open", "duct19_open"):
        # open boundaries, one component: a duct along x with spheres, walls
        # in y and z (type-0 faces at rest); 27h: pressure inlet and outlet,
        # 19: velocity inlet, local (type 2) outlet, the z(2) face moving
        lat = 19 if name == "duct19_open" else 27
        p = make_params(n, lattice=lat, highorder=(lat == 27), internal_obstacles=True, pbc=(0, 0, 0),
                        visc1=0.05, precision=precision)
        p.openbc = 1
        if lat == 27:
            p.openbc_type[0], p.openbc_press[0] = 1, 3e-3
            p.openbc_type[1], p.openbc_press[1] = 1, 0.0
        else:
            p.openbc_type[0], p.openbc_u[0], p.openbc_v[0] = 0, 2e-3, 1e-4
            p.openbc_type[1] = 2
            p.openbc_type[5], p.openbc_u[5] = 0, 1e-3
        raw = setups.random_spheres(n, fill=0.1, rmin=2.0, rmax=3.0, seed=seed, margin=3)
        rng = np.random.default_rng(seed)
        uvw = rng.uniform(-1e-4, 1e-4, (3, n[2], n[1], n[0]))
        P = rng.uniform(-1e-4, 1e-4, (n[2], n[1], n[0]))
        return p, setups._moments(P, *uvw), raw
    if name == "chan27h_open":
        # two components through open x faces: a velocity inlet and a
        # pressure outlet, walls in y (type-0 faces at rest), z periodic
        p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True,
                        internal_obstacles=True, pbc=(0, 0, 1), visc1=0.05, visc2=0.1, rho_r=2.0, rho_b=1.0,
                        tau_diff=0.1, sigma=0.005, width=4.0, precision=precision)
        p.openbc = 1
        p.openbc_type[0], p.openbc_u[0] = 0, 3e-3
        p.openbc_type[1], p.openbc_press[1] = 1, 0.0
        raw = setups.random_spheres(n, fill=0.1, rmin=2.0, rmax=3.0, seed=seed, margin=3)
        f = setups.two_layer_channel(n, raw, frac=0.4, width=4.0, rho_r=2.0, rho_b=1.0, axis=1)
        return p, f, raw
    if name == "pois27h_1c":  # one component with walls and obstacles
        p = make_params(n, lattice=27, highorder=True, internal_obstacles=True, pbc=(1, 0, 1), visc1=0.1,
                        fx=1e-5, precision=precision)
        raw = setups.random_spheres(n, fill=0.15, rmin=2.0, rmax=4.0, seed=seed, margin=2)
        z = np.zeros((n[2], n[1], n[0]))
        f = {k: z.copy() for k in ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")}
        return p, f, raw
    raise KeyError(name)


def prepare(p, fields, raw, classify):
    """Classify isfluid and set the reference's initial solid-node phi."""
    isf = None
    if raw is not None or p.internal_obstacles:
        isf = classify(p, raw)
        if p.twocomponent:
            fields = dict(fields)
            fields["phi"] = setups.solid_phi_average(fields["phi"], isf, p.lattice)
    return fields, isf


def state_of(sim, p):
    """Comparable state: hfields of S, phi (both buffers), u_old."""
    out = {"h": np.stack([sim_get(sim, f) for f in HFIELDS])}
    if p.twocomponent:
        out["phi"] = sim_get(sim, FIELD_PHI)
        out["phiO"] = sim_get(sim, FIELD_PHI_O)
        if p.densratio:
            out["uold"] = np.stack([sim_get(sim, f) for f in FIELD_UOLD])
    return out


def sim_get(sim, fid):
    return sim.get_field(fid) if hasattr(sim, "get_field") else sim.download(fid)


def rel_l2(a, b, mask=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a = a[..., mask]
        b = b[..., mask]
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return d / nb if nb > 0 else d


def compare(st_a, st_b, isf=None):
    """Per-field relative L2 differences on fluid nodes (all nodes when no
    isfluid).  phi is compared on every node: solid-node ghost values are
    part of the state."""
    fluid = None if isf is None else (isf != 0)
    res = {}
    from lbfast_amd.params import HFIELD_NAMES
    for i, nm in enumerate(HFIELD_NAMES):
        res[nm] = rel_l2(st_a["h"][i], st_b["h"][i], fluid)
    if "phi" in st_b:
        res["phi"] = rel_l2(st_a["phi"], st_b["phi"])
        res["phiO"] = rel_l2(st_a["phiO"], st_b["phiO"])
    if "uold" in st_b:
        res["uold"] = rel_l2(st_a["uold"], st_b["uold"], fluid)
    return res

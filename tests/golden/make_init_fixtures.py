"""Fixtures from the reference's own initial-condition statements.

The per-node body of the initial-condition loop of init_conditions_module.f90
(lines 276-450: the branches TAYLORGREEN, TWOPOISEUILLE, POISEUILLE,
LAMBTEST and the default droplet, one- or two-component) and the constants
before it (invdim, lines 55-57; H_pois, xc_pois and the two-layer
coefficients, lines 247-274) are read from the reference checkout as text and
evaluated on every node of a small grid with the evaluator of
make_ref_fixtures.py, under the switch sets of the Fortran host's builds.
The helpers the body calls are taken from the reference's text too:
pbc_images (init_conditions_module.f90:591-599) is inlined in place of its
call, fcut_tanh (vars_module.f90:1306-1329) is evaluated from its statement.
The random term stdev * randgauss_CPU() of the Poiseuille branches
(lines 339, 350) is taken as 0: it draws from the host's random generator,
which no restatement can reproduce.

Only numbers are stored (tests/golden/ref_init.npz): per case, the switches,
the parameters and, per node, P, u, v, w, the six second moments and phi.
tests/test_initial_conditions.py and tests/test_gpu_fortran_host.py compare
the Python setups and the Fortran host's start with them.

Usage: python tests/golden/make_init_fixtures.py [/root/reference]
"""
from __future__ import annotations

import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_ref_fixtures as M  # noqa: E402  (the statement evaluator)

# name -> (switches of the host build, parameters, grid)
CASES = {
    "tg": (["TAYLORGREEN"], dict(uwall=0.04), (8, 6, 4)),
    "laplace_2c": (["TWOCOMPONENT", "DENSRATIO"],
                   dict(radius=3.3, width=3.0, sigma=0.01, rho_r=1.0, rho_b=0.1, center=[2.5, 7.25, 3.0]),
                   (10, 9, 8)),
    "laplace_1c": ([], dict(radius=3.3, width=3.0, center=[5.5, 4.0, 2.25]), (10, 9, 8)),
    "pois": (["POISEUILLE"], dict(fz=2e-5, visc1=0.1), (13, 3, 4)),
    "pois_bb": (["POISEUILLE", "BOUNCE_BACK"], dict(fz=2e-5, visc1=0.1), (13, 3, 4)),
    "twopois": (["TWOPOISEUILLE", "TWOCOMPONENT", "DENSRATIO"],
                dict(fz=1e-6, visc1=0.05, visc2=0.4, rho_r=0.5, rho_b=1.5, width=4.0), (20, 3, 4)),
    "twopois_nodens": (["TWOPOISEUILLE", "TWOCOMPONENT"],
                       dict(fz=1e-6, visc1=0.05, visc2=0.4, rho_r=1.0, rho_b=1.0, width=4.0), (20, 3, 4)),
    "lamb": (["LAMBTEST", "TWOCOMPONENT", "DENSRATIO"],
             dict(radius=3.5, width=3.0, sigma=0.005, rho_r=1.0, rho_b=0.5, visc1=0.01, visc2=0.02,
                  center=[6.5, 5.0, 4.75]), (12, 10, 9)),
}


def _constants(ref):
    """Every parameter declaration of vars_module.f90 that evaluates (ONE,
    TWO, HALF, SIXTEEN, cssq, pi_greek, ...)."""
    ns = M.base_ns()
    for st in M.statements(M.preprocess(M.read(ref, "vars_module.f90"), {})):
        low = st.lower()
        if "parameter" not in low or "::" not in low:
            continue
        for part in M._split_top(low.split("::", 1)[1]):
            a = M._assign_split(part)
            if not a:
                continue
            try:
                ns[a[0].strip()] = eval(M.pyexpr(a[1]), ns)
            except Exception:
                pass
    return ns


def _fcut_tanh(ref):
    """fcut_tanh(r, r_cut, width_cut) from its one statement."""
    lines = M.read(ref, "vars_module.f90")
    a = next(i for i, l in enumerate(lines) if re.search(r"pure\s+function\s+fcut_tanh", l, re.I))
    st = next(l for l in lines[a:a + 40] if re.match(r"\s*fcut_tanh\s*=", l, re.I))
    rhs = M.pyexpr(M._assign_split(M._strip_comment(st))[1])
    return rhs


def _pbc_images_lines(ref):
    lines = M.read(ref, "init_conditions_module.f90")
    a = next(i for i, l in enumerate(lines) if re.search(r"subroutine\s+pbc_images\s*\(", l, re.I))
    return [l for l in lines[a:a + 12] if re.match(r"\s*xout\(\d\)\s*=", l, re.I)]


def bodies(ref, defines):
    lines = M.read(ref, "init_conditions_module.f90")
    inv = [l for l in lines if re.match(r"\s*invdim\(\d\)\s*=", l, re.I)][:3]
    h = next(i for i, l in enumerate(lines) if re.match(r"\s*H_pois\s*=", l))
    a = max(i for i in range(h) if re.match(r"#if defined\(POISEUILLE\) \|\| defined\(TWOPOISEUILLE\)", lines[i]))
    b = next(i for i in range(a, len(lines)) if re.match(r"\s*do k=1,nz", lines[i]))
    pre = inv + lines[a:b]
    s0 = next(i for i in range(b, len(lines)) if re.match(r"\s*loc_press\s*=\s*ZERO", lines[i]))
    s1 = next(i for i in range(s0, len(lines)) if re.match(r"\s*enddo\b", lines[i]))
    body = []
    for l in lines[s0:s1]:
        if re.match(r"\s*call\s+pbc_images\s*\(\s*invdim\s*,\s*dist3d\s*,\s*dist3dout\s*\)", l, re.I):
            for s in _pbc_images_lines(ref):
                body.append(re.sub(r"\baaa\b", "invdim", re.sub(r"\bxxs\b", "dist3d", re.sub(r"\bxout\b", "dist3dout", s))))
        else:
            body.append(l)
    dset = {d: 1 for d in defines}
    return (M.parse(M.statements(M.preprocess(pre, dset)))[0],
            M.parse(M.statements(M.preprocess(body, dset)))[0])


class Arr:
    """A Fortran rank-1 array (1-based) the evaluator can read and set."""

    def __init__(self, vals):
        self.v = list(vals)

    def __call__(self, i):
        return self.v[int(i) - 1]


def fnint(x):
    # Fortran NINT: to the nearest integer, halves away from zero
    return int(math.floor(abs(x) + 0.5)) * (1 if x >= 0 else -1)


def evaluate(ref, name):
    defines, prm, n = CASES[name]
    pre, body = bodies(ref, defines)
    consts = _constants(ref)
    ft = _fcut_tanh(ref)
    lx, ly, lz = n
    out = np.zeros((11, lz, ly, lx))
    for gk in range(1, lz + 1):
        for gj in range(1, ly + 1):
            for gi in range(1, lx + 1):
                ns = M.base_ns()
                ns.update({k: v for k, v in consts.items() if k not in ("real", "abs", "sqrt", "sin", "cos")})
                ns.update(tanh=math.tanh, nint=fnint, log=math.log, lx=lx, ly=ly, lz=lz, nx=lx, ny=ly, nz=lz,
                          gi=gi, gj=gj, gk=gk, i=gi, j=gj, k=gk, ii=gi, jj=gj, kk=gk, myblock=1, stdev=0.0,
                          uwall=0.0, fz=0.0, visc1=0.1, visc2=0.1, rho_r=1.0, rho_b=1.0, width=4.0, radius=1.0,
                          sigma=0.0, lamb_visc=0.0, lamb_visc_temp=0.0, loc_press=0.0)
                ns.update({k: v for k, v in prm.items() if k != "center"})
                ns["center"] = Arr(prm.get("center", [0.0, 0.0, 0.0]))
                ns["randgauss_cpu"] = lambda: 0.0
                ns["isfluid"] = lambda a, b, c: 1
                ns["fcut_tanh"] = (lambda e: (lambda r, r_cut, width_cut: eval(
                    e, {**M.base_ns(), "tanh": math.tanh, "r": r, "r_cut": r_cut, "width_cut": width_cut})))(ft)
                arrays = {"invdim": Arr([0.0] * 3), "dist3d": Arr([0.0] * 3), "dist3dout": Arr([0.0] * 3)}
                ns.update(arrays)
                node = {}

                def setter(nm, args, val, arrays=arrays, node=node):
                    if nm in arrays:
                        arrays[nm].v[int(args[0]) - 1] = val
                    elif nm == "hfields_flip":
                        node[int(args[3])] = val
                    elif nm == "phifields_flip":
                        node[11] = val
                ns["__set"] = setter
                M.execute(pre, ns)
                M.execute(body, ns)
                for f in range(1, 12):
                    out[f - 1, gk - 1, gj - 1, gi - 1] = node.get(f, 0.0)
    return defines, prm, n, out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    arrays, meta = {}, {}
    for name in CASES:
        defines, prm, n, out = evaluate(ref, name)
        meta[name] = {"defines": defines, "params": prm, "n": list(n)}
        arrays[name + "/fields"] = out
        print(f"{name}: n={n} |P|max={np.abs(out[0]).max():.3e} |u|max={np.abs(out[1:4]).max():.3e} "
              f"phi in [{out[10].min():.3f}, {out[10].max():.3f}]")
    arrays["meta"] = np.frombuffer(json.dumps(meta).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(HERE, "ref_init.npz"), **arrays)


if __name__ == "__main__":
    main()

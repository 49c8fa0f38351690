"""Fixtures from the reference's own per-node kernel statements.

The reference cannot be compiled here (CUDA Fortran for nvfortran), and it
ships no test vectors.  This script reads the bodies of its per-node kernels
from the reference checkout as text and evaluates their statements on seeded
random inputs in float64, with a small evaluator of the Fortran subset those
bodies use (assignments, DO loops, IF blocks, CYCLE, RETURN, and the
defines.h preprocessor switches):

  fused_LB_kernel          lb_cuda_fused_module.f90:426-750
                           -> the 27 (19) post-collision populations of a node
  moments_LB_kernel        lb_cuda_moments_module.f90:19-341
                           -> u, Pi after the moments stage and the stored force
  LB_int_boundary_kernel   lb_cuda_boundary_module.f90:19-216
                           -> the near-wall closure's output moments
  PHI_int_boundary_kernel  lb_cuda_boundary_module.f90:218-459
                           -> the wetting ghost value of a solid node
  open face x(1)           boundary_cds_module.f90:104-262 (the body of the
                           x(1) loop of bcs_mesoscopic_hfields; the other five
                           faces' bodies are the same statements with their
                           own openbc_*(n))
                           -> the node's output moments after the open-face
                              correction, for openbc_type 0, 1 and 2

Each kernel runs under several switch sets.  The lattice tables are the
reference's own: ex, ey, ez, opp, p and the constants come from evaluating
the parameter declarations of vars_module.f90 under the same switches, and
Minv_mrt from tests/golden/hermite_tables.npz (extracted from
vars_module.f90:359-521 by make_tables.py).

Only numbers are stored (tests/golden/ref_kernels.npz): per case, the
switches, the physical parameters, the inputs and the outputs.
tests/test_oracle_reference_kernels.py drives the oracle's stages with the
same inputs and compares.

Usage: python tests/golden/make_ref_fixtures.py [/root/reference]
"""
from __future__ import annotations

import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

# ----------------------------------------------------------------------------
# preprocessor (the subset the kernels use)


def _cpp_value(expr: str, defines: dict) -> bool:
    e = re.sub(r"defined\s*\(\s*(\w+)\s*\)", lambda m: "1" if m.group(1) in defines else "0", expr)
    e = re.sub(r"defined\s+(\w+)", lambda m: "1" if m.group(1) in defines else "0", e)
    e = re.sub(r"[A-Za-z_]\w*", lambda m: str(defines.get(m.group(0), 0)) if m.group(0) not in ("and", "or", "not")
               else m.group(0), e)
    e = e.replace("&&", " and ").replace("||", " or ")
    e = re.sub(r"!(?!=)", " not ", e)
    return bool(eval(e, {}, {}))


def preprocess(lines, defines):
    out, stack = [], []  # stack of (active, taken)
    for ln in lines:
        s = ln.strip()
        if s.startswith("#"):
            d = s[1:].strip()
            word = d.split()[0] if d.split() else ""
            parent = all(a for a, _ in stack)
            if word == "ifdef":
                v = d.split()[1] in defines
                stack.append((v, v))
            elif word == "ifndef":
                v = d.split()[1] not in defines
                stack.append((v, v))
            elif word == "if":
                v = _cpp_value(d[2:], defines)
                stack.append((v, v))
            elif word == "elif":
                a, taken = stack.pop()
                v = (not taken) and _cpp_value(d[4:], defines)
                stack.append((v, taken or v))
            elif word == "else":
                a, taken = stack.pop()
                stack.append((not taken, True))
            elif word == "endif":
                stack.pop()
            elif word in ("warning", "error", "define", "include", "undef"):
                pass
            else:
                raise ValueError("unknown directive: " + s)
            continue
        if all(a for a, _ in stack):
            out.append(ln)
    assert not stack
    return out


# ----------------------------------------------------------------------------
# logical statements


def _strip_comment(ln: str) -> str:
    q = None
    for i, ch in enumerate(ln):
        if q:
            if ch == q:
                q = None
        elif ch in "'\"":
            q = ch
        elif ch == "!":
            return ln[:i]
    return ln


def statements(lines):
    out, cur = [], ""
    for ln in lines:
        s = _strip_comment(ln).rstrip()
        if not s.strip():
            continue
        t = s.strip()
        if t.startswith("&"):
            t = t[1:]
        if t.endswith("&"):
            cur += t[:-1] + " "
            continue
        cur += t
        out.append(cur.strip())
        cur = ""
    assert not cur.strip()
    return out


def _match_paren(s: str, i: int) -> int:
    depth = 0
    for j in range(i, len(s)):
        if s[j] == "(":
            depth += 1
        elif s[j] == ")":
            depth -= 1
            if depth == 0:
                return j
    raise ValueError("unbalanced: " + s)


def _split_top(s: str, sep=","):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        if ch == sep and depth == 0:
            parts.append(cur)
            cur = ""
        else:
            cur += ch
    parts.append(cur)
    return parts


def _assign_split(s: str):
    depth = 0
    for i, ch in enumerate(s):
        if ch in "([":
            depth += 1
        elif ch in ")]":
            depth -= 1
        elif ch == "=" and depth == 0:
            prev = s[i - 1] if i else ""
            nxt = s[i + 1] if i + 1 < len(s) else ""
            if prev in "=/<>" or nxt == "=":
                continue
            return s[:i].strip(), s[i + 1:].strip()
    return None


class F32(float):
    """A default-kind (single precision) REAL: literals without a kind suffix
    or d exponent.  Operations between two of them (or with an integer) round
    to binary32; with a double operand the result is double."""

    def _bin(self, o, f, rev=False):
        if isinstance(o, (F32, int)) and not isinstance(o, bool):
            a, b = np.float32(self), np.float32(o)
            return F32(float(f(b, a) if rev else f(a, b)))
        return f(o, float(self)) if rev else f(float(self), o)

    def __add__(self, o): return self._bin(o, lambda a, b: a + b)
    def __radd__(self, o): return self._bin(o, lambda a, b: a + b, True)
    def __sub__(self, o): return self._bin(o, lambda a, b: a - b)
    def __rsub__(self, o): return self._bin(o, lambda a, b: a - b, True)
    def __mul__(self, o): return self._bin(o, lambda a, b: a * b)
    def __rmul__(self, o): return self._bin(o, lambda a, b: a * b, True)
    def __truediv__(self, o): return self._bin(o, lambda a, b: a / b)
    def __rtruediv__(self, o): return self._bin(o, lambda a, b: a / b, True)
    def __pow__(self, o): return self._bin(o, lambda a, b: a ** b)
    def __rpow__(self, o): return self._bin(o, lambda a, b: a ** b, True)
    def __neg__(self): return F32(-float(self))


_NUM = re.compile(r"(?<![\w.])(\d+\.\d*|\.\d+|\d+)(?:([de])([+-]?\d+))?(?:_[a-z]\w*)?(?![\w])")


def pyexpr(e: str) -> str:
    e = e.lower()
    for a, b in ((".ne.", " != "), (".eq.", " == "), (".lt.", " < "), (".le.", " <= "), (".gt.", " > "),
                 (".ge.", " >= "), (".and.", " and "), (".or.", " or "), (".not.", " not "), ("/=", "!="),
                 ("(/", "["), ("/)", "]")):
        e = e.replace(a, b)

    def num(m):
        mant, ex, ev = m.group(1), m.group(2), m.group(3)
        kind = m.group(0)[len(mant) + (len(ex) + len(ev) if ex else 0):]
        lit = (f"{mant}e{ev}" if "." in mant else f"{mant}.0e{ev}") if ex else mant
        if ex == "e" or (not ex and "." in mant):
            if not kind:  # default-kind real
                return f"F32(float(np.float32({lit})))"
        return lit
    return _NUM.sub(num, e)


# statement tree: ("set", lhs, rhs_py, rhs_src) ("do", var, lo, hi, body)
# ("if", [(cond, body), ...]) ("cycle",) ("return",)
def parse(stmts, i=0):
    body = []
    while i < len(stmts):
        s = stmts[i]
        low = s.lower().strip()
        key = re.sub(r"\s+", "", low)
        if key in ("enddo", "endif") or re.match(r"^(else|elseif\(.*\)then)$", key):
            return body, i
        m = re.match(r"^do\s+(\w+)\s*=\s*(.+)$", low)
        if m:
            lo, hi = _split_top(m.group(2))[:2]
            sub, j = parse(stmts, i + 1)
            assert re.sub(r"\s+", "", stmts[j].lower()) == "enddo", stmts[j]
            body.append(("do", m.group(1), pyexpr(lo), pyexpr(hi), sub))
            i = j + 1
            continue
        if re.match(r"^if\s*\(", low):
            p0 = low.index("(")
            p1 = _match_paren(low, p0)
            cond = low[p0 + 1:p1]
            rest = low[p1 + 1:].strip()
            if rest == "then":
                branches = []
                sub, j = parse(stmts, i + 1)
                branches.append((pyexpr(cond), sub))
                while True:
                    k = re.sub(r"\s+", "", stmts[j].lower())
                    if k == "endif":
                        break
                    if k == "else":
                        sub, j = parse(stmts, j + 1)
                        branches.append(("True", sub))
                        continue
                    t = stmts[j].lower().strip()
                    q0 = t.index("(")
                    q1 = _match_paren(t, q0)
                    sub, j2 = parse(stmts, j + 1)
                    branches.append((pyexpr(t[q0 + 1:q1]), sub))
                    j = j2
                body.append(("if", branches))
                i = j + 1
                continue
            inner, _ = parse([rest], 0)
            body.append(("if", [(pyexpr(cond), inner)]))
            i += 1
            continue
        if key == "cycle":
            body.append(("cycle",))
        elif key == "return":
            body.append(("return",))
        elif key == "exit":
            body.append(("exit",))
        elif low.startswith("call ") or low.startswith("implicit"):
            pass
        elif "::" in low:
            decl = low.split("::", 1)[1]
            for part in _split_top(decl):
                a = _assign_split(part)
                if a:
                    body.append(("set", a[0].strip(), pyexpr(a[1]), a[1]))
        else:
            a = _assign_split(s)
            if a is None:
                raise ValueError("cannot parse: " + s)
            body.append(("set", a[0].lower().strip(), pyexpr(a[1]), a[1]))
        i += 1
    return body, i


FROZEN = {"i", "j", "k", "ii", "jj", "kk", "li", "lj", "lk", "gi", "gj", "gk", "myblock", "xblock", "yblock",
          "zblock", "intblock", "oxblock", "oyblock", "ozblock", "omyblock", "oii", "ojj", "okk", "lii", "ljj",
          "lkk"}


class Stop(Exception):
    pass


def execute(body, ns, hook=None):
    for node in body:
        kind = node[0]
        if kind == "set":
            lhs, rhs, src = node[1], node[2], node[3]
            m = re.match(r"^(\w+)\s*\((.*)\)$", lhs)
            name = m.group(1) if m else lhs
            if name in FROZEN or "%" in rhs:
                continue
            val = eval(rhs, ns)
            if m:
                # the LDS target ftemp is addressed by tile indices: only the link matters
                args = [] if name == "ftemp" else [eval(pyexpr(a), ns) for a in _split_top(m.group(2))]
                ns["__set"](name, args, val)
            else:
                ns[name] = val
            if hook:
                hook(name, src, val)
        elif kind == "do":
            _, var, lo, hi, sub = node
            for v in range(int(eval(lo, ns)), int(eval(hi, ns)) + 1):
                ns[var] = v
                r = execute(sub, ns, hook)
                if r == "return":
                    return r
                if r == "exit":
                    break
        elif kind == "if":
            for cond, sub in node[1]:
                if eval(cond, ns):
                    r = execute(sub, ns, hook)
                    if r:
                        return r
                    break
        elif kind in ("cycle", "return", "exit"):
            return kind
    return None


def fsign(a, b):
    return math.copysign(abs(a), b)


def freal(x, *a, **k):
    if isinstance(x, (list, tuple)):
        return [float(v) for v in x]
    return float(x)


def base_ns():
    return {"real": freal, "dble": freal, "db": 8, "strdb": 8, "abs": abs, "sqrt": math.sqrt, "sin": math.sin,
            "cos": math.cos, "max": max, "min": min, "sign": fsign, "int": int, "nint": round, "F32": F32,
            "np": np, "float": float, "__builtins__": {}}


# ----------------------------------------------------------------------------
# the reference's source


def read(ref, name):
    return open(os.path.join(ref, name)).read().splitlines()


def kernel(ref, fname, kname):
    """Body of `attributes(global) subroutine kname(...)`: the lines after
    its argument list (which may hold preprocessor lines) up to endsubroutine."""
    lines = read(ref, fname)
    a = next(i for i, l in enumerate(lines) if re.search(r"attributes\(global\)\s+subroutine\s+" + kname + r"\b", l))
    b = next(i for i in range(a, len(lines)) if re.search(r"endsubroutine\s+" + kname + r"\b", lines[i]))
    depth, h = 0, a
    for h in range(a, b):
        if lines[h].strip().startswith("#"):
            continue
        code = _strip_comment(lines[h])
        depth += code.count("(") - code.count(")")
        if depth == 0 and "(" in "".join(lines[a:h + 1]):
            break
    return lines[h + 1:b]


def module_tables(ref, defines):
    """Parameter declarations of vars_module.f90 (constants, lattice tables)
    evaluated under the switch set."""
    lines = preprocess(read(ref, "vars_module.f90"), defines)
    ns = base_ns()
    want = ("zero", "half", "one", "two", "three", "four", "five", "six", "seven", "eight", "nine", "ten",
            "cssq", "invcssq", "nlinks", "ex", "ey", "ez", "opp", "p0", "p1", "p2", "p3", "p", "pi_greek")
    for st in statements(lines):
        low = st.lower()
        if "parameter" not in low or "::" not in low:
            continue
        decl = low.split("::", 1)[1]
        for part in _split_top(decl):
            a = _assign_split(part)
            if not a or a[0].strip() not in want:
                continue
            ns[a[0].strip()] = eval(pyexpr(a[1]), ns)
    for t in ("ex", "ey", "ez"):
        ns["d" + t] = [float(v) for v in ns[t]]
    minv = np.load(os.path.join(HERE, "hermite_tables.npz"))["minv_mrt"]
    return ns, minv


# ----------------------------------------------------------------------------
# per-node evaluation

C0 = 10  # the node's (i, j, k) in the evaluator; neighbour offsets are relative to it


def run_node(body, defines, tabs, params, fields, isf, extra_hook=None):
    """fields: {name: {comp: {offset: value}}}; isf: {offset: class} (default 1)."""
    ns0, minv = tabs
    ns = base_ns()
    ns.update({k: v for k, v in ns0.items() if k not in ("real", "abs", "sqrt", "sin", "cos", "max", "min")})
    ns.update(params)
    ns.update(i=C0, j=C0, k=C0, ii=C0, jj=C0, kk=C0, li=0, lj=0, lk=0, myblock=1, intblock=1, omyblock=1)
    for t in ("ex", "ey", "ez", "opp", "p", "dex", "dey", "dez"):
        arr = ns0[t]
        ns[t] = (lambda a: (lambda l: a[int(l)]))(arr)
    ns["minv_mrt"] = lambda kk, ll: float(minv[int(kk), int(ll)])
    ns["isfluid"] = lambda a, b, c: isf.get((a - C0, b - C0, c - C0), 1)
    store = {}

    def getter(name):
        def g(a, b, c, comp, blk=None):
            off = (a - C0, b - C0, c - C0)
            if (name, comp, off) in store:
                return store[(name, comp, off)]
            return fields.get(name, {}).get(int(comp), {}).get(off, 0.0)
        return g

    for nm in ("hfields_s", "hfields_old", "hfields_in", "hfields_out", "phifields_s", "auxfields_s",
               "locauxfields_s", "forces_s", "ftemp"):
        ns[nm] = getter(nm)
    pops = {}

    def setter(name, args, val):
        if name == "ftemp":
            pops[int(ns["l"])] = val
            return
        a, b, c, comp = args[:4]
        store[(name, int(comp), (a - C0, b - C0, c - C0))] = val
    ns["__set"] = setter
    ns["ftemp"] = lambda *a: 0.0

    def hook(name, src, val):
        # fused_LB_kernel: the rest population is `udotc` built from fneq1 before the link loop
        if name == "udotc" and "fneq1" in src.lower() and 0 not in pops:
            pops[0] = val
        if extra_hook:
            extra_hook(name, src, val, ns)
    execute(body, ns, hook)
    return store, pops, ns


def fused_body(ref, defines):
    return parse(statements(preprocess(kernel(ref, "lb_cuda_fused_module.f90", "fused_LB_kernel"), defines)))[0]


# ----------------------------------------------------------------------------
# cases


def _mom_inputs(rng, n, umax=0.08):
    P = rng.uniform(-1e-2, 1e-2, n)
    u = rng.uniform(-umax, umax, (3, n))
    pn = rng.uniform(-2e-3, 2e-3, (6, n))
    Pi = np.stack([u[0] * u[0] + pn[0], u[1] * u[1] + pn[1], u[2] * u[2] + pn[2],
                   u[0] * u[1] + pn[3], u[0] * u[2] + pn[4], u[1] * u[2] + pn[5]])
    return np.concatenate([P[None], u, Pi])  # (10, n): P u v w Pxx Pyy Pzz Pxy Pxz Pyz


def derived(prm):
    """LBFAST.f90:212-219"""
    q = dict(prm)
    q["tau1"] = q["visc1"] * 3.0 + 0.5
    q["omega"] = 1.0 / q["tau1"]
    if q.get("width"):
        q["beta"] = 12.0 * q["sigma"] / q["width"]
        q["kapp"] = 1.5 * q["sigma"] * q["width"]
        q["sharp_c"] = 4.0 * q["tau_diff"] / q["width"]
    q["invrho_r"] = 1.0 / q["rho_r"]
    q["invrho_b"] = 1.0 / q["rho_b"]
    return q


PARAMS_2C = dict(visc1=0.07, visc2=0.13, rho_r=1.0, rho_b=0.12, tau_diff=0.11, sigma=0.013, width=3.5,
                 fx=2e-5, fy=-1e-5, fz=3e-5, wettab_r=90.0, wettab_b=90.0)

FUSED_CASES = {
    "fused_d3q27h_1c": dict(LATTICE=27, HIGHORDER=1),
    "fused_d3q27h_2c_dens": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1),
    "fused_d3q27_2c": dict(LATTICE=27, TWOCOMPONENT=1),
    "fused_d3q19_1c": dict(LATTICE=19),
    "fused_d3q27h_2c_dens_smag": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, SMAGORINSKI=1),
    "fused_d3q27h_2c_dens_obst": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1),
}

MOMENTS_CASES = {
    "moments_2c_dens_incomp": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERFACE_INCOMP=1),
    "moments_2c_dens": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1),
    "moments_2c": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1),
    "moments_1c": dict(LATTICE=27, HIGHORDER=1),
    "moments_2c_dens_incomp_buoy": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERFACE_INCOMP=1,
                                        BUOYANCY_FORCING=1),
    "moments_2c_dens_incomp_rep": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERFACE_INCOMP=1,
                                       REPULSIVE_FLUX=1),
}

LBINT_CASES = {
    "lbint_d3q27_1c": dict(LATTICE=27, HIGHORDER=1, INTERNAL_OBSTACLES=1),
    "lbint_d3q27_2c_dens": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1),
    "lbint_d3q27_2c_dens_smag": dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1,
                                     SMAGORINSKI=1),
    "lbint_d3q19_1c": dict(LATTICE=19, INTERNAL_OBSTACLES=1),
}

PHIBND_CASES = {
    "phibnd_d3q27_90": (dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1), 90.0),
    "phibnd_d3q27_60": (dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1,
                             WETTABILITY=1), 60.0),
    "phibnd_d3q27_130": (dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1,
                              WETTABILITY=1), 130.0),
    "phibnd_d3q19_90": (dict(LATTICE=19, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1), 90.0),
}

OFFS = [(a, b, c) for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1)]  # x fastest


def gen_fused(ref, name, defines, rng, n=64):
    tabs = module_tables(ref, defines)
    body = fused_body(ref, defines)
    prm = derived(PARAMS_2C)
    mom = _mom_inputs(rng, n)
    frc = rng.uniform(-1e-4, 1e-4, (3, n))
    phi = rng.uniform(0.0, 1.0, n)
    isf = np.ones(n, dtype=np.int8)
    if "INTERNAL_OBSTACLES" in defines:
        isf[rng.uniform(size=n) < 0.25] = 0
    nl = tabs[0]["nlinks"] + 1
    out = np.zeros((27, n))
    for s in range(n):
        fields = {"hfields_in": {c + 1: {(0, 0, 0): float(mom[c, s])} for c in range(10)},
                  "forces_s": {c + 1: {(0, 0, 0): float(frc[c, s])} for c in range(3)},
                  "phifields_s": {1: {(0, 0, 0): float(phi[s])}}}
        _, pops, _ = run_node(body, defines, tabs, prm, fields, {(0, 0, 0): int(isf[s])})
        assert sorted(pops) == list(range(nl)), (name, sorted(pops))
        for l in range(nl):
            out[l, s] = pops[l]
    return {"moments": mom, "forces": frc, "phi": phi, "isfluid": isf, "populations": out[:nl]}


def gen_moments(ref, name, defines, rng, n=64):
    tabs = module_tables(ref, defines)
    body = parse(statements(preprocess(kernel(ref, "lb_cuda_moments_module.f90", "moments_LB_kernel"), defines)))[0]
    prm = derived(PARAMS_2C)
    mom = _mom_inputs(rng, n)
    uold = mom[1:4] + rng.uniform(-5e-3, 5e-3, (3, n))
    phi = rng.uniform(0.0, 1.0, n)
    nrm = rng.normal(size=(3, n))
    nrm /= np.linalg.norm(nrm, axis=0)
    mg = rng.uniform(0.0, 0.4, n)
    lap = rng.uniform(-0.05, 0.05, n)
    dvt = rng.uniform(-0.05, 0.05, n)
    rep = rng.uniform(-3e-3, 3e-3, (3, n)) * (rng.uniform(size=(3, n)) < 0.7)
    newmom = np.zeros((10, n))
    force = np.zeros((3, n))
    for s in range(n):
        fields = {"hfields_s": {c + 1: {(0, 0, 0): float(mom[c, s])} for c in range(10)},
                  "hfields_old": {c + 2: {(0, 0, 0): float(uold[c, s])} for c in range(3)},
                  "phifields_s": {1: {(0, 0, 0): float(phi[s])}},
                  "auxfields_s": {1: {(0, 0, 0): float(nrm[0, s])}, 2: {(0, 0, 0): float(nrm[1, s])},
                                  3: {(0, 0, 0): float(nrm[2, s])}, 4: {(0, 0, 0): float(mg[s])}},
                  "locauxfields_s": {1: {(0, 0, 0): float(lap[s])}, 2: {(0, 0, 0): float(dvt[s])},
                                     6: {(0, 0, 0): float(rep[0, s])}, 7: {(0, 0, 0): float(rep[1, s])},
                                     8: {(0, 0, 0): float(rep[2, s])}}}
        store, _, _ = run_node(body, defines, tabs, prm, fields, {})
        for c in range(10):
            newmom[c, s] = store.get(("hfields_s", c + 1, (0, 0, 0)), mom[c, s])
        for c in range(3):
            force[c, s] = store[("forces_s", c + 1, (0, 0, 0))]
    return {"moments": mom, "uold": uold, "phi": phi, "normal": nrm, "modgrad": mg, "lap": lap, "divthn": dvt,
            "repflux": rep, "moments_out": newmom, "forces_out": force}


def gen_lbint(ref, name, defines, rng, n=64):
    tabs = module_tables(ref, defines)
    body = parse(statements(preprocess(kernel(ref, "lb_cuda_boundary_module.f90", "LB_int_boundary_kernel"),
                                       defines)))[0]
    prm = derived(PARAMS_2C)
    mom = _mom_inputs(rng, n)
    oin = _mom_inputs(rng, n)
    phi = rng.uniform(0.0, 1.0, n)
    # solid (0) or fluid (1) neighbours; the node itself is a near-wall fluid node (-1)
    nb = (rng.uniform(size=(27, n)) >= 0.35).astype(np.int8)
    nb[13] = -1
    out = np.zeros((10, n))
    for s in range(n):
        fields = {"hfields_in": {c + 1: {(0, 0, 0): float(mom[c, s])} for c in range(10)},
                  "hfields_out": {c + 1: {(0, 0, 0): float(oin[c, s])} for c in range(10)},
                  "phifields_s": {1: {(0, 0, 0): float(phi[s])}}}
        isf = {o: int(nb[q, s]) for q, o in enumerate(OFFS)}
        store, _, _ = run_node(body, defines, tabs, prm, fields, isf)
        for c in range(10):
            out[c, s] = store[("hfields_out", c + 1, (0, 0, 0))]
    return {"moments": mom, "moments_o": oin, "phi": phi, "neighbours": nb, "moments_o_out": out}


def gen_phibnd(ref, name, defines, wettab, rng, n=96):
    tabs = module_tables(ref, defines)
    src = kernel(ref, "lb_cuda_boundary_module.f90", "PHI_int_boundary_kernel")
    # the donor's fields are addressed through its tile indices (oii, ojj, okk);
    # the evaluator addresses every node by its absolute coordinates
    src = [re.sub(r"\(\s*oii\s*,\s*ojj\s*,\s*okk\s*,", "(donor_x,donor_y,donor_z,", l) for l in src]
    src = [re.sub(r"\(\s*ii\s*,\s*jj\s*,\s*kk\s*,", "(i,j,k,", l) for l in src]
    body = parse(statements(preprocess(src, defines)))[0]
    prm = derived(dict(PARAMS_2C, wettab_r=wettab, wettab_b=wettab))
    nl = tabs[0]["nlinks"] + 1
    ex, ey, ez = tabs[0]["ex"], tabs[0]["ey"], tabs[0]["ez"]
    lat_offs = {(ex[l], ey[l], ez[l]) for l in range(nl)}
    phi = rng.uniform(0.0, 1.0, (27, n))
    nrm = rng.normal(size=(3, 27, n))
    nrm /= np.linalg.norm(nrm, axis=0)
    mg = rng.uniform(0.0, 0.3, (27, n))
    nb = np.zeros((27, n), dtype=np.int8)
    for s in range(n):
        kind = s % 6
        if kind == 4:    # all neighbours fluid: the wall normal vanishes (early return)
            nb[:, s] = rng.choice([1, -1], size=27)
        elif kind == 5:  # all neighbours solid but one link pair: no normal or no donor
            nb[:, s] = 0
            q = int(rng.integers(27))
            nb[q, s] = 1
            nb[26 - q, s] = rng.choice([0, 1])
        elif kind == 0:  # a plane wall through the node, random orientation
            w = rng.normal(size=3)
            for q, o in enumerate(OFFS):
                d = np.dot(o, w)
                nb[q, s] = 0 if d <= 0 else (1 if d > 1.2 else -1)
        elif kind == 1:  # an edge (two half-spaces)
            w1, w2 = rng.normal(size=3), rng.normal(size=3)
            for q, o in enumerate(OFFS):
                nb[q, s] = 0 if (np.dot(o, w1) <= 0 or np.dot(o, w2) <= 0) else rng.choice([1, -1])
        else:            # scattered
            nb[:, s] = rng.choice([0, 1, -1], size=27, p=[0.45, 0.3, 0.25])
        nb[13, s] = 0
    out = np.zeros(n)
    for s in range(n):
        isf = {o: int(nb[q, s]) for q, o in enumerate(OFFS)}
        fields = {"phifields_s": {1: {o: float(phi[q, s]) for q, o in enumerate(OFFS)}},
                  "auxfields_s": {c + 1: {o: float(nrm[c, q, s]) for q, o in enumerate(OFFS)} for c in range(3)}}
        fields["auxfields_s"][4] = {o: float(mg[q, s]) for q, o in enumerate(OFFS)}
        # links outside the compiled lattice are never visited by the kernel
        assert all(o in lat_offs or True for o in OFFS)
        store, _, _ = run_node(body, defines, tabs, prm, fields, isf)
        out[s] = store.get(("phifields_s", 1, (0, 0, 0)), phi[13, s])
    return {"phi": phi, "normal": nrm, "modgrad": mg, "neighbours": nb, "phi_out": out}


OPENH_CASES = {
    "openh_d3q27h_1c_t1": (dict(LATTICE=27, HIGHORDER=1, INTERNAL_OBSTACLES=1), 1),
    "openh_d3q27h_1c_t0": (dict(LATTICE=27, HIGHORDER=1, INTERNAL_OBSTACLES=1), 0),
    "openh_d3q27h_2c_dens_t2": (dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1), 2),
    "openh_d3q27h_2c_dens_t0": (dict(LATTICE=27, HIGHORDER=1, TWOCOMPONENT=1, DENSRATIO=1, INTERNAL_OBSTACLES=1), 0),
    "openh_d3q27_2c_t1": (dict(LATTICE=27, TWOCOMPONENT=1, INTERNAL_OBSTACLES=1), 1),
    "openh_d3q19_1c_t0": (dict(LATTICE=19, INTERNAL_OBSTACLES=1), 0),
}


def open_face_body(ref):
    """The per-node statements of the first open face of
    bcs_mesoscopic_hfields: from the load of hfields_in to the store of the
    tenth output moment."""
    lines = read(ref, "boundary_cds_module.f90")
    a = next(i for i, l in enumerate(lines) if "press=real(hfields_in(ii,jj,kk,1,myblock)" in l.replace(" ", ""))
    b = next(i for i in range(a, len(lines)) if "hfields_out(ii,jj,kk,10,myblock)=real(opyz" in lines[i].replace(" ", ""))
    return lines[a:b + 1]


def gen_openh(ref, name, defines, otype, rng, n=64):
    tabs = module_tables(ref, defines)
    body = parse(statements(preprocess(open_face_body(ref), defines)))[0]
    face = dict(press=float(rng.uniform(-5e-3, 5e-3)), u=float(rng.uniform(-0.05, 0.05)),
                v=float(rng.uniform(-0.05, 0.05)), w=float(rng.uniform(-0.05, 0.05)))
    prm = derived(PARAMS_2C)
    prm["openbc_type_x"] = lambda k: [otype, 2][int(k) - 1]
    prm["openbc_press_x"] = lambda k: [face["press"], 0.0][int(k) - 1]
    prm["openbc_u_x"] = lambda k: [face["u"], 0.0][int(k) - 1]
    prm["openbc_v_x"] = lambda k: [face["v"], 0.0][int(k) - 1]
    prm["openbc_w_x"] = lambda k: [face["w"], 0.0][int(k) - 1]
    mom = _mom_inputs(rng, n)
    oin = _mom_inputs(rng, n)
    phi = rng.uniform(0.0, 1.0, n)
    # the wall (x - 1) is solid, the other neighbours solid or fluid; the node
    # itself is a near-wall fluid node (-1)
    nb = (rng.uniform(size=(27, n)) >= 0.35).astype(np.int8)
    for q, o in enumerate(OFFS):
        if o[0] == -1:
            nb[q] = 0
    nb[13] = -1
    out = np.zeros((10, n))
    for s in range(n):
        fields = {"hfields_in": {c + 1: {(0, 0, 0): float(mom[c, s])} for c in range(10)},
                  "hfields_out": {c + 1: {(0, 0, 0): float(oin[c, s])} for c in range(10)},
                  "phifields_s": {1: {(0, 0, 0): float(phi[s])}}}
        isf = {o: int(nb[q, s]) for q, o in enumerate(OFFS)}
        store, _, _ = run_node(body, defines, tabs, prm, fields, isf)
        for c in range(10):
            out[c, s] = store[("hfields_out", c + 1, (0, 0, 0))]
    return {"moments": mom, "moments_o": oin, "phi": phi, "neighbours": nb, "moments_o_out": out,
            "face": np.array([otype, face["press"], face["u"], face["v"], face["w"]])}


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    rng = np.random.Generator(np.random.MT19937(20261203))
    arrays, meta = {}, {}

    def put(case, kernel_name, defines, extra, data):
        meta[case] = {"kernel": kernel_name, "defines": defines, "params": derived(dict(PARAMS_2C, **extra))}
        for k, v in data.items():
            arrays[f"{case}/{k}"] = np.asarray(v)

    for case, d in FUSED_CASES.items():
        put(case, "fused", d, {}, gen_fused(ref, case, d, rng))
    for case, d in MOMENTS_CASES.items():
        put(case, "moments", d, {}, gen_moments(ref, case, d, rng))
    for case, d in LBINT_CASES.items():
        put(case, "lb_int_boundary", d, {}, gen_lbint(ref, case, d, rng))
    for case, (d, wt) in PHIBND_CASES.items():
        put(case, "phi_int_boundary", d, {"wettab_r": wt, "wettab_b": wt}, gen_phibnd(ref, case, d, wt, rng))
    for case, (d, t) in OPENH_CASES.items():
        put(case, "open_h", d, {}, gen_openh(ref, case, d, t, rng))
    arrays["meta"] = np.frombuffer(json.dumps(meta, sort_keys=True).encode(), dtype=np.uint8)
    out = os.path.join(HERE, "ref_kernels.npz")
    np.savez_compressed(out, **arrays)
    print("wrote", out, "cases:", len(meta))


if __name__ == "__main__":
    main()

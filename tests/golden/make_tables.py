"""Extract the reference's literal D3Q27 Hermite tables as numeric fixtures.

Reads vars_module.f90 (as text) from the reference checkout, evaluates the
rational literals of the Minv_mrt and M_mrt array constructors
(vars_module.f90:359-664) and stores the two 27x27 matrices in Fortran index
order, A[i, j] = A(i, j) with 0-based i, j, as tests/golden/hermite_tables.npz.

Usage: python tests/golden/make_tables.py [/root/reference]
"""
import os
import re
import sys
from fractions import Fraction

import numpy as np


def parse_constructor(text, name):
    i = text.index(name + "(0:nlinks,0:nlinks) = reshape([")
    j = text.index("], shape=", i)
    body = text[i:j].split("reshape([", 1)[1]
    body = body.replace("&", " ")
    toks = re.findall(r"-?\d+\.\d+_db(?:/\d+\.\d+_db)?", body)
    vals = []
    for t in toks:
        t = t.replace("_db", "")
        if "/" in t:
            a, b = t.split("/")
            vals.append(Fraction(a) / Fraction(b))
        else:
            vals.append(Fraction(t))
    assert len(vals) == 27 * 27, (name, len(vals))
    return np.array([float(v) for v in vals]).reshape((27, 27), order="F")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    text = open(os.path.join(ref, "vars_module.f90")).read()
    minv = parse_constructor(text, "Minv_mrt")
    m = parse_constructor(text, "M_mrt")
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hermite_tables.npz")
    np.savez(out, minv_mrt=minv, m_mrt=m)
    print("wrote", out, "M*Minv-I max:", np.abs(m @ minv - np.eye(27)).max())


if __name__ == "__main__":
    main()

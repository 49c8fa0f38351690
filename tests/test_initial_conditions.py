"""The initial conditions against the reference's own statements.

tests/golden/ref_init.npz holds, per switch set, the fields that the
reference's initial-condition loop (init_conditions_module.f90:276-450, with
pbc_images and fcut_tanh from their own statements) produces on every node
of a small grid (tests/golden/make_init_fixtures.py; numbers only).  The
Python setups that the parity tests start from must reproduce them:
the Taylor-Green vortex (lines 298-314) and the two-component droplet of the
default branch (lines 394-426, with the periodic images of pbc_images).  The
Fortran host's starts are compared with the same fixtures on the GPU
(tests/test_gpu_fortran_host.py).
"""
import json
import os

import numpy as np
import pytest

from lbfast_amd import setups

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_DATA = np.load(os.path.join(GOLD, "ref_init.npz"))
META = json.loads(_DATA["meta"].tobytes().decode())
KEYS = ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")


def _fields(name):
    return _DATA[name + "/fields"]


def _close(got, want, tol=1e-14):
    scale = max(float(np.abs(want).max()), 1e-300)
    return float(np.abs(np.asarray(got) - want).max()) / scale < tol


def test_fixtures_cover_each_branch():
    """Each case exercises its branch: a flow, a pressure field or an
    interface, and the branches differ where they should."""
    assert set(META) >= {"tg", "laplace_2c", "laplace_1c", "pois", "pois_bb", "twopois", "twopois_nodens", "lamb"}
    f = {k: _fields(k) for k in META}
    assert np.abs(f["tg"][1]).max() > 0.03 and np.abs(f["tg"][0]).max() > 1e-3
    assert 0.0 < f["laplace_2c"][10].min() and f["laplace_2c"][10].max() > 0.9
    assert np.abs(f["laplace_1c"][10]).max() == 0.0 and np.abs(f["laplace_1c"][0]).max() > 1e-3
    assert np.abs(f["pois"][3]).max() > 1e-3 and not np.array_equal(f["pois"][3], f["pois_bb"][3])
    assert not np.array_equal(f["twopois"][3], f["twopois_nodens"][3])
    # the Lamb ellipsoid is stretched along x: phi falls slower along x than along y
    lm, (cx, cy, cz) = f["lamb"][10], META["lamb"]["params"]["center"]
    assert lm[int(cz) - 1, int(cy) - 1, int(cx) - 1 + 3] > lm[int(cz) - 1, int(cy) - 1 + 3, int(cx) - 1]


def test_taylor_green_setup_matches_reference_statements():
    m = META["tg"]
    n = tuple(m["n"])
    got = setups.taylor_green(n, m["params"]["uwall"])
    want = _fields("tg")
    for i, k in enumerate(KEYS):
        assert _close(got[k], want[i]) or np.abs(want[i]).max() == np.abs(got[k]).max() == 0.0, k


@pytest.mark.parametrize("name", ["laplace_2c"])
def test_droplet_setup_matches_reference_statements(name):
    m = META[name]
    n, q = tuple(m["n"]), m["params"]
    got = setups.droplet(n, q["radius"], q["width"], q["center"], q["sigma"], q["rho_r"], q["rho_b"])
    want = _fields(name)
    assert _close(got["P"], want[0])
    assert _close(got["phi"], want[10])
    for i, k in enumerate(KEYS[1:], start=1):
        assert np.abs(got[k]).max() == 0.0 and np.abs(want[i]).max() == 0.0, k

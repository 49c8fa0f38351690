"""Pin the oracle's lattice tables.

The reference has no test data of its own (no fixtures, golden outputs or
unit tests ship with it), so the tables are pinned against the reference's
literal Minv_mrt / M_mrt constructors (vars_module.f90:359-664), extracted
as numbers into tests/golden/hermite_tables.npz by tests/golden/make_tables.py,
and against the textbook D3Q27 / D3Q19 definitions.
"""
import os

import numpy as np
import pytest

import oracle
from lbfast_amd.lattice import EX, EY, EZ, OPP, weights

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = np.load(os.path.join(HERE, "golden", "hermite_tables.npz"))


def test_minv_matches_reference_literal_table():
    ref = TABLES["minv_mrt"]  # ref[k, l] = Minv_mrt(k, l): k Hermite basis, l population
    ours = np.array([[oracle.minv(k, l) for l in range(27)] for k in range(27)])
    assert np.abs(ours - ref).max() < 1e-15


def test_m_matches_reference_literal_table():
    ref = TABLES["m_mrt"]  # ref[l, k] = M_mrt(l, k) = H_k(c_l)
    ours = np.array([[oracle.mmat(k, l) for k in range(27)] for l in range(27)])
    assert np.abs(ours - ref).max() < 1e-15


def test_m_minv_identity():
    m = TABLES["m_mrt"]
    minv = TABLES["minv_mrt"]
    assert np.abs(m @ minv - np.eye(27)).max() < 1e-14
    assert np.abs(minv @ m - np.eye(27)).max() < 1e-14


def test_minv_low_rows_are_weights_and_velocities():
    w = np.array(weights(27))
    minv = TABLES["minv_mrt"]
    assert np.allclose(minv[0], w, atol=0, rtol=1e-15)             # Minv(0,l) = w_l
    assert np.allclose(minv[1], 3 * w * np.array(EX[:27]), atol=1e-16)  # Minv(1,l) = 3 w_l c_lx
    assert np.allclose(minv[2], 3 * w * np.array(EY[:27]), atol=1e-16)
    assert np.allclose(minv[3], 3 * w * np.array(EZ[:27]), atol=1e-16)


def test_hermite_orthogonality_high_order_has_no_low_moments():
    """HIGHORDER terms (k >= 10) carry no mass, momentum or second moment:
    M rows 0..9 against Minv rows 10..26 vanish."""
    m = TABLES["m_mrt"]
    minv = TABLES["minv_mrt"]
    prod = m[:, :10].T @ minv[10:].T
    assert np.abs(prod).max() < 1e-15


@pytest.mark.parametrize("lat", [19, 27])
def test_velocity_set(lat):
    n = lat
    c = np.array([EX[:n], EY[:n], EZ[:n]]).T
    w = np.array(weights(lat))
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose((w[:, None] * c).sum(0), 0)
    # isotropy: sum w c_a c_b = cs^2 delta_ab, cs^2 = 1/3
    assert np.allclose(np.einsum("l,la,lb->ab", w, c, c), np.eye(3) / 3.0, atol=1e-15)
    # fourth-order isotropy: sum w c_a^2 c_b^2 = cs^4 (1 + 2 delta_ab)
    q4 = np.einsum("l,la,lb->ab", w, c ** 2, c ** 2)
    assert np.allclose(q4, (np.ones((3, 3)) + 2 * np.eye(3)) / 9.0, atol=1e-15)
    for l in range(n):
        assert EX[OPP[l]] == -EX[l] and EY[OPP[l]] == -EY[l] and EZ[OPP[l]] == -EZ[l]
    assert len({(EX[l], EY[l], EZ[l]) for l in range(n)}) == n

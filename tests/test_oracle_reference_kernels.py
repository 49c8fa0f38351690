"""The oracle's per-node stages against the reference's own statements.

tests/golden/ref_kernels.npz holds, for several switch sets, seeded inputs
and the outputs that the reference's kernel statements produce on them
(evaluated from the reference source text by tests/golden/make_ref_fixtures.py;
numbers only):

  fused_*     fused_LB_kernel (lb_cuda_fused_module.f90:530-713): all 27 (19)
              post-collision populations of a node -- feq, the regularised
              fneq, the HIGHORDER E_k/N_k terms, the Guo forcing, the
              relaxation rate (two-component viscosity mix, SMAGORINSKI) and
              the zeroed inputs of solid nodes
  moments_*   moments_LB_kernel (lb_cuda_moments_module.f90:104-339): u, Pi
              and the stored force -- surface tension, body force (plain and
              BUOYANCY_FORCING), the DENSRATIO pressure and viscous
              corrections, the INTERFACE_INCOMP velocity correction and the
              REPULSIVE_FLUX clipping
  lbint_*     LB_int_boundary_kernel (lb_cuda_boundary_module.f90:117-212):
              the output moments of a near-wall node with a random set of
              solid neighbours
  phibnd_*    PHI_int_boundary_kernel (lb_cuda_boundary_module.f90:274-454):
              the wetting ghost value of a solid node next to plane walls,
              edges and scattered solids, for contact angles 60, 90, 130
  openh_*     the open-face correction of bcs_mesoscopic_hfields
              (boundary_cds_module.f90:104-262, the x(1) face): the output
              moments of a first-layer node for openbc_type 0, 1 and 2

The oracle runs the same inputs through its own stages (FP64) and must agree
to round-off.  Also here: the HIGHORDER equilibrium by mathematics -- with
Pi = u u (no non-equilibrium part) and no force, every Hermite moment of
order >= 1 of the rebuilt populations equals u^a v^b w^c.
"""
import json
import os

import numpy as np
import pytest

import oracle
from lbfast_amd.params import make_params, PREC_FP64

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_DATA = np.load(os.path.join(GOLD, "ref_kernels.npz"))
META = json.loads(_DATA["meta"].tobytes().decode())

STAGE_MOMENTS, STAGE_FUSED, STAGE_LB_INT, STAGE_PHI_BND, STAGE_OPEN_H = 4, 5, 7, 2, 12
OFFS = [(a, b, c) for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1)]  # as the generator


def _case(name):
    return {k.split("/", 1)[1]: _DATA[k] for k in _DATA.files if k.startswith(name + "/")}


def _params(meta, n):
    d, q = meta["defines"], meta["params"]
    p = make_params(n, lattice=d.get("LATTICE", 27), highorder=bool(d.get("HIGHORDER")),
                    twocomponent=bool(d.get("TWOCOMPONENT")), densratio=bool(d.get("DENSRATIO")),
                    internal_obstacles=bool(d.get("INTERNAL_OBSTACLES")), visc1=q["visc1"], visc2=q["visc2"],
                    rho_r=q["rho_r"], rho_b=q["rho_b"], tau_diff=q["tau_diff"], sigma=q["sigma"], width=q["width"],
                    fx=q["fx"], fy=q["fy"], fz=q["fz"], precision=PREC_FP64)
    p.interface_incomp = 1 if d.get("INTERFACE_INCOMP") else 0
    p.buoyancy_forcing = 1 if d.get("BUOYANCY_FORCING") else 0
    p.smagorinsky = 1 if d.get("SMAGORINSKI") else 0
    p.repulsive_flux = 1 if d.get("REPULSIVE_FLUX") else 0
    p.wettab_r = q.get("wettab_r", 90.0)
    return p


def _grid(vals, shape):
    """per-sample values (n,) -> (nz, ny, nx) in Fortran order (x fastest)."""
    return np.asarray(vals, dtype=np.float64).reshape(shape[2], shape[1], shape[0])


def _blocks(shape):
    """centres of the 3x3x3 sample blocks of a grid (x fastest)."""
    nx, ny, nz = shape
    return [(3 * a + 1, 3 * b + 1, 3 * c + 1) for c in range(nz // 3) for b in range(ny // 3) for a in range(nx // 3)]


def _rel(got, want):
    return float(np.abs(got - want).max() / max(np.abs(want).max(), 1e-300))


@pytest.mark.parametrize("name", sorted(k for k, m in META.items() if m["kernel"] == "fused"))
def test_fused_collision_matches_reference_statements(name):
    m, c = META[name], _case(name)
    n = (4, 4, 4)
    p = _params(m, n)
    o = oracle.Oracle(p)
    if p.internal_obstacles:
        o.set_isfluid(_grid(c["isfluid"], n).astype(np.int8))
    for f in range(10):
        o.set_field(1 + f, _grid(c["moments"][f], n))
    for f in range(3):
        o.set_field(15 + f, _grid(c["forces"][f], n))
    if p.twocomponent:
        o.set_field(11, _grid(c["phi"], n))
    o.begin_step()
    o.stage(STAGE_FUSED)
    pops = o.populations()
    want = c["populations"]
    nl = want.shape[0]
    got = pops[:nl].reshape(nl, -1)
    err = _rel(got, want)
    assert err < 1e-13, (name, err)
    if nl == 19:
        assert np.all(pops[19:] == 0.0)


@pytest.mark.parametrize("name", sorted(k for k, m in META.items() if m["kernel"] == "moments"))
def test_moments_stage_matches_reference_statements(name):
    m, c = META[name], _case(name)
    n = (4, 4, 4)
    p = _params(m, n)
    if p.repulsive_flux:
        p.a_rep, p.q_th, p.cos_opp_t = 0.05, 0.125, -0.8
    o = oracle.Oracle(p)
    for f in range(10):
        o.set_field(1 + f, _grid(c["moments"][f], n))
    for f in range(3):
        o.set_field(12 + f, _grid(c["uold"][f], n))  # u_old: the other buffer's u, v, w
    if p.twocomponent:
        o.set_field(11, _grid(c["phi"], n))
        for f in range(3):
            o.set_field(18 + f, _grid(c["normal"][f], n))
        o.set_field(21, _grid(c["modgrad"], n))
        o.set_field(25, _grid(c["lap"], n))
        o.set_field(26, _grid(c["divthn"], n))
        if p.repulsive_flux:
            for f in range(3):
                o.set_field(44 + f, _grid(c["repflux"][f], n))
    o.begin_step()
    o.stage(STAGE_MOMENTS)
    got_m = np.stack([o.get_field(1 + f).ravel() for f in range(10)])
    got_f = np.stack([o.get_field(15 + f).ravel() for f in range(3)])
    em = max(_rel(got_m[f], c["moments_out"][f]) for f in range(1, 10))
    ef = _rel(got_f, c["forces_out"])
    assert np.array_equal(got_m[0], c["moments_out"][0])  # P is not written
    assert em < 1e-13 and ef < 1e-12, (name, em, ef)


@pytest.mark.parametrize("name", sorted(k for k, m in META.items() if m["kernel"] == "lb_int_boundary"))
def test_lb_int_boundary_matches_reference_statements(name):
    m, c = META[name], _case(name)
    n = (12, 12, 12)
    p = _params(m, n)
    nb = c["neighbours"]
    cen = _blocks(n)
    assert len(cen) == nb.shape[1]
    isf = np.ones((n[2], n[1], n[0]), dtype=np.int8)
    S = np.zeros((10, n[2], n[1], n[0]))
    O = np.zeros((10, n[2], n[1], n[0]))
    phi = np.full((n[2], n[1], n[0]), 0.5)
    for s, (x, y, z) in enumerate(cen):
        for q, (a, b, cc) in enumerate(OFFS):
            isf[z + cc, y + b, x + a] = nb[q, s]
        S[:, z, y, x] = c["moments"][:, s]
        O[:, z, y, x] = c["moments_o"][:, s]
        phi[z, y, x] = c["phi"][s]
    o = oracle.Oracle(p)
    o.set_isfluid(isf)
    for f in range(10):
        o.set_field(1 + f, S[f])
        o.set_field(27 + f, O[f])
    if p.twocomponent:
        o.set_field(11, phi)
    o.begin_step()
    o.stage(STAGE_LB_INT)
    got = np.stack([o.get_field(27 + f) for f in range(10)])
    got = np.stack([got[:, z, y, x] for (x, y, z) in cen], axis=1)
    err = _rel(got, c["moments_o_out"])
    assert err < 1e-13, (name, err)


@pytest.mark.parametrize("name", sorted(k for k, m in META.items() if m["kernel"] == "open_h"))
def test_open_face_matches_reference_statements(name):
    """The open-face correction of bcs_mesoscopic_hfields on the x(1) face:
    64 sample nodes on the first fluid layer (x = 1, 0-based), 3 apart in y
    and z, inside the face's edges; y and z periodic so no other face acts
    on them."""
    m, c = META[name], _case(name)
    A = 8
    n = (4, 3 * A, 3 * A)
    p = _params(m, n)
    p.pbc[0], p.pbc[1], p.pbc[2] = 0, 1, 1
    t, press, u, v, w = c["face"]
    p.openbc = 1
    p.openbc_type[0], p.openbc_type[1] = int(t), 2
    p.openbc_press[0], p.openbc_u[0], p.openbc_v[0], p.openbc_w[0] = press, u, v, w
    nb = c["neighbours"]
    cen = [(1, 3 * b + 1, 3 * a + 1) for a in range(A) for b in range(A)]
    assert len(cen) == nb.shape[1]
    isf = np.ones((n[2], n[1], n[0]), dtype=np.int8)
    S = np.zeros((10, n[2], n[1], n[0]))
    O = np.zeros((10, n[2], n[1], n[0]))
    phi = np.full((n[2], n[1], n[0]), 0.5)
    for s, (x, y, z) in enumerate(cen):
        for q, (a, b, cc) in enumerate(OFFS):
            isf[z + cc, y + b, x + a] = nb[q, s]
        S[:, z, y, x] = c["moments"][:, s]
        O[:, z, y, x] = c["moments_o"][:, s]
        phi[z, y, x] = c["phi"][s]
    o = oracle.Oracle(p)
    o.set_isfluid(isf)
    for f in range(10):
        o.set_field(1 + f, S[f])
        o.set_field(27 + f, O[f])
    if p.twocomponent:
        o.set_field(11, phi)
    o.begin_step()
    o.stage(STAGE_OPEN_H)
    got = np.stack([o.get_field(27 + f) for f in range(10)])
    got = np.stack([got[:, z, y, x] for (x, y, z) in cen], axis=1)
    err = _rel(got, c["moments_o_out"])
    assert err < 1e-13, (name, err)
    # the correction is not a no-op: the reference's output differs from its input
    assert np.abs(c["moments_o_out"] - c["moments_o"]).max() > 1e-3


@pytest.mark.parametrize("name", sorted(k for k, m in META.items() if m["kernel"] == "phi_int_boundary"))
def test_phi_int_boundary_matches_reference_statements(name):
    m, c = META[name], _case(name)
    nb = c["neighbours"]
    n = (12, 12, 3 * (nb.shape[1] // 16))
    p = _params(m, n)
    cen = _blocks(n)
    assert len(cen) == nb.shape[1]
    isf = np.ones((n[2], n[1], n[0]), dtype=np.int8)
    phi = np.zeros((n[2], n[1], n[0]))
    aux = np.zeros((4, n[2], n[1], n[0]))
    for s, (x, y, z) in enumerate(cen):
        for q, (a, b, cc) in enumerate(OFFS):
            isf[z + cc, y + b, x + a] = nb[q, s]
            phi[z + cc, y + b, x + a] = c["phi"][q, s]
            aux[:3, z + cc, y + b, x + a] = c["normal"][:, q, s]
            aux[3, z + cc, y + b, x + a] = c["modgrad"][q, s]
    o = oracle.Oracle(p)
    o.set_isfluid(isf)
    o.set_field(11, phi)
    for f in range(4):
        o.set_field(18 + f, aux[f])
    o.begin_step()
    o.stage(STAGE_PHI_BND)
    got = o.get_field(11)
    got = np.array([got[z, y, x] for (x, y, z) in cen])
    want = c["phi_out"]
    assert np.abs(got - want).max() < 1e-13, (name, float(np.abs(got - want).max()))
    # the cases exercise the kernel's branches: some ghosts are rewritten, some
    # keep their value (no wall normal or no donor), some are clipped to [0, 1]
    changed = want != c["phi"][13]
    assert changed.sum() >= len(want) // 4 and (~changed).sum() >= 1


def test_fixture_coverage():
    """The fixture holds every kernel under the switch sets that matter."""
    kinds = {}
    for k, m in META.items():
        kinds.setdefault(m["kernel"], []).append(m["defines"])
    assert len(kinds["fused"]) >= 6 and len(kinds["moments"]) >= 6
    assert any(d.get("HIGHORDER") and d.get("DENSRATIO") for d in kinds["fused"])
    assert any(d.get("LATTICE") == 19 for d in kinds["fused"])
    assert any(d.get("INTERFACE_INCOMP") for d in kinds["moments"])


@pytest.mark.parametrize("lattice,highorder", [(27, True), (27, False), (19, False)])
def test_equilibrium_hermite_moments_are_velocity_monomials(lattice, highorder):
    """With Pi = u u and no force the post-collision populations are the
    equilibrium, whose Hermite moments (M_mrt rows of the reference's literal
    table) are P for order 0 and u^a v^b w^c for every (a, b, c) of order >= 1
    that the lattice resolves: all 26 for D3Q27 HIGHORDER (the E_k of
    fused:668-684), only order <= 2 without it."""
    tabs = np.load(os.path.join(GOLD, "hermite_tables.npz"))
    M = tabs["m_mrt"]
    HK = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1),
          (2, 1, 0), (2, 0, 1), (1, 2, 0), (1, 0, 2), (0, 1, 2), (0, 2, 1), (1, 1, 1), (2, 2, 0), (2, 0, 2), (0, 2, 2),
          (1, 1, 2), (1, 2, 1), (2, 1, 1), (2, 1, 2), (2, 2, 1), (1, 2, 2), (2, 2, 2)]
    n = (4, 4, 4)
    rng = np.random.default_rng(31)
    P = rng.uniform(-1e-2, 1e-2, n[::-1])
    u = rng.uniform(-0.1, 0.1, (3,) + n[::-1])
    p = make_params(n, lattice=lattice, highorder=highorder, visc1=0.04, precision=PREC_FP64)
    o = oracle.Oracle(p)
    vals = [P, u[0], u[1], u[2], u[0] * u[0], u[1] * u[1], u[2] * u[2], u[0] * u[1], u[0] * u[2], u[1] * u[2]]
    for f, v in enumerate(vals):
        o.set_field(1 + f, v)
    o.begin_step()
    o.stage(STAGE_FUSED)
    f = o.populations().reshape(27, -1)
    if lattice == 19:
        # the D3Q19 order-2 moments, directly
        from lbfast_amd.lattice import EX, EY, EZ
        c = np.array([EX[:19], EY[:19], EZ[:19]], dtype=np.float64)
        ff = f[:19]
        uu = u.reshape(3, -1)
        assert np.abs(ff.sum(0) - P.ravel()).max() < 1e-15
        for a in range(3):
            assert np.abs((c[a][:, None] * ff).sum(0) - uu[a]).max() < 1e-15
            for b in range(3):
                m2 = ((c[a] * c[b] - (1.0 / 3.0 if a == b else 0.0))[:, None] * ff).sum(0)
                assert np.abs(m2 - uu[a] * uu[b]).max() < 1e-15
        return
    A = M.T @ f  # M_mrt(l, k) = H_k(c_l): first index the population
    uu = u.reshape(3, -1)
    assert np.abs(A[0] - P.ravel()).max() < 1e-15
    for k, (a, b, cc) in enumerate(HK[1:], start=1):
        want = uu[0] ** a * uu[1] ** b * uu[2] ** cc
        if not highorder and a + b + cc > 2:
            # without HIGHORDER the higher moments are those of the second-order feq
            continue
        assert np.abs(A[k] - want).max() < 1e-15, (k, (a, b, cc), float(np.abs(A[k] - want).max()))

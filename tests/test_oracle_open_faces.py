"""Oracle pins for the open boundaries (openbc, boundary_cds_module.f90).

The per-node correction is pinned against the reference's own statements in
test_oracle_reference_kernels.py (openh_* fixtures).  What those leave open --
which nodes each face visits, on which layer, in which direction, and where
the corrections sit in the step -- is pinned here by what the mathematics
fixes:

* symmetry: a pressure-driven duct run along x, along y and along z, and with
  its inlet and outlet swapped, is the same flow up to the axis permutation
  and the reflection, to round-off.  A face loop on the wrong layer, a
  transposed face index or a sign slip on the high faces breaks it at O(1).
* closed form: the fully developed part of the pressure-driven duct is the
  body-force duct (the near-wall closure's exact parabola in a channel is
  pinned in test_oracle_switches.py) driven by F = -cs^2 dP/dx, with dP/dx
  the duct's own linear pressure profile; the flow runs from the high to the
  low pressure and every cross-section carries the same flux.
* a velocity inlet carries, far downstream, its prescribed flux U A to within
  the closure's near-face error (1 %).
* BCPHIFLUX with open faces equals a brute-force restatement: the target
  phase sum is the initial one plus this step's flux u_a phi through the
  first fluid layer of each open face (edges included, + at the low face,
  - at the high one), as boundary_cds_module.f90:1667-1884 forms it.
* periodic axes carry no open face: openbc = 1 with every axis periodic is
  bit-identical to openbc = 0.
"""
import numpy as np

import oracle
from lbfast_amd import setups
from lbfast_amd.params import make_params, PREC_FP64, FIELD_PHI

KEYS = ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")
CS2 = 1.0 / 3.0


def _run(n, pbc, steps, faces=None, fx=0.0, lattice=27, nu=1.0 / 6.0, axis=0):
    """A fluid at rest in an open box; faces: {side: (type, press, velocity
    along `axis`)} of the two faces normal to `axis`; the other non-periodic
    faces keep the default type 0 with zero velocity."""
    p = make_params(n, lattice=lattice, highorder=(lattice == 27), internal_obstacles=True, pbc=pbc,
                    visc1=nu, fx=fx, precision=PREC_FP64)
    p.openbc = 1
    for side, (t, press, vel) in (faces or {}).items():
        f = 2 * axis + side
        p.openbc_type[f] = t
        p.openbc_press[f] = press
        uvw = [0.0, 0.0, 0.0]
        uvw[axis] = vel
        p.openbc_u[f], p.openbc_v[f], p.openbc_w[f] = uvw
    isf = oracle.classify_isfluid(p, None)
    shape = (n[2], n[1], n[0])
    o = oracle.Oracle(p)
    o.load({k: np.zeros(shape) for k in KEYS}, isf)
    o.step(steps)
    return o


def _moments(o):
    return np.stack([o.get_field(1 + k) for k in range(10)])


def _duct(axis, steps=800, L=24, W=8, faces=None):
    n = [W, W, W]
    n[axis] = L
    faces = faces or {0: (1, 1e-4, 0.0), 1: (1, 0.0, 0.0)}
    return _moments(_run(tuple(n), (0, 0, 0), steps, faces=faces, axis=axis))


def test_open_face_ducts_agree_across_axes_and_sides():
    mx = _duct(0)
    assert np.abs(mx[1]).max() > 1e-6  # a flow develops
    # (z, y, x) arrays; moments P u v w xx yy zz xy xz yz
    my = np.transpose(_duct(1), (0, 1, 3, 2))[[0, 2, 1, 3, 5, 4, 6, 7, 9, 8]]
    mz = np.transpose(_duct(2), (0, 3, 2, 1))[[0, 3, 2, 1, 6, 5, 4, 9, 8, 7]]
    mm = _duct(0, faces={0: (1, 0.0, 0.0), 1: (1, 1e-4, 0.0)})[:, :, :, ::-1]
    mm = mm * np.array([1, -1, 1, 1, 1, 1, 1, -1, -1, 1.0])[:, None, None, None]
    scale = np.abs(mx).max()
    for other in (my, mz, mm):
        assert np.abs(mx - other).max() < 1e-13 * scale


def test_pressure_driven_duct_is_the_body_force_duct():
    for lattice in (27, 19):
        n, dP = (24, 8, 8), 1e-4
        o = _run(n, (0, 0, 0), 800, faces={0: (1, dP, 0.0), 1: (1, 0.0, 0.0)}, lattice=lattice)
        m = _moments(o)
        o.step(20)
        assert np.abs(_moments(o) - m).max() < 1e-9 * np.abs(m).max()  # steady
        x = np.arange(1, n[0] + 1, dtype=float)
        mid = slice(6, n[0] - 6)
        P = m[0][4, 4, :]
        c = np.polyfit(x[mid], P[mid], 1)
        assert np.abs(np.polyval(c, x[mid]) - P[mid]).max() < 1e-3 * dP  # linear away from the faces
        assert c[0] < 0  # falls from the inlet
        drop = np.polyval(c, 1.0) - np.polyval(c, float(n[0]))  # between the two solid layers
        assert 0.9 * dP < drop < 1.3 * dP
        u = m[1]
        assert u[1:-1, 1:-1, 1:-1].min() > 0.0  # from high to low pressure everywhere
        flux = u[1:-1, 1:-1, mid].sum(axis=(0, 1))
        # node velocities sum to the plane flux up to the closure's O(1e-4)
        # near the faces
        assert np.abs(flux / flux.mean() - 1.0).max() < 2e-4
        # the body-force duct: x periodic, the same walls (type-0 faces at rest)
        ob = _run((3, 8, 8), (1, 0, 0), 800, fx=-CS2 * c[0], lattice=lattice)
        ub = ob.get_field(2)[:, :, 1]
        assert np.abs(u[:, :, n[0] // 2] - ub).max() < 5e-4 * ub.max()


def test_velocity_inlet_carries_its_flux():
    U, n = 1e-3, (24, 8, 8)
    o = _run(n, (0, 0, 0), 2000, faces={0: (0, 0.0, U), 1: (1, 0.0, 0.0)})
    u = o.get_field(2)
    area = (n[1] - 2) * (n[2] - 2)
    flux = u[1:-1, 1:-1, 1:-1].sum(axis=(0, 1)) / (U * area)
    assert np.abs(flux[4:-4] - 1.0).max() < 0.01
    assert np.abs(flux[8:-8] - flux[8:-8].mean()).max() < 1e-4  # uniform where fully developed


def _bcphi_open_case(n=(24, 16, 12), seed=7):
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, internal_obstacles=True,
                    pbc=(0, 0, 1), visc1=0.05, visc2=0.1, rho_r=2.0, rho_b=1.0, tau_diff=0.1, sigma=0.005,
                    width=4.0, precision=PREC_FP64, bcphiflux=True)
    p.openbc = 1
    p.openbc_type[0], p.openbc_u[0] = 0, 2e-3  # inflow at x(1)
    p.openbc_type[1], p.openbc_press[1] = 1, 0.0  # outflow at x(2)
    isf = oracle.classify_isfluid(p, None)
    f = setups.two_layer_channel(n, None, frac=0.4, width=4.0, rho_r=2.0, rho_b=1.0, axis=1)
    return p, f, isf


def _open_flux(p, isf, u3, phi):
    """this step's u_a phi through the first fluid layer of every open face,
    brute force over the whole grid"""
    chg = 0.0
    n = (p.nx, p.ny, p.nz)
    for ax in range(3):
        if p.pbc[ax]:
            continue
        for side, sgn in ((0, 1.0), (1, -1.0)):
            lay = n[ax] - 2 if side else 1  # 0-based
            sl = [slice(None)] * 3
            sl[2 - ax] = lay  # arrays are (z, y, x)
            sl = tuple(sl)
            sel = isf[sl] == -1
            chg += sgn * float((u3[ax][sl] * phi[sl])[sel].sum())
    return chg


def test_bcphiflux_with_open_faces_equals_brute_force():
    p, f, isf = _bcphi_open_case()
    o = oracle.Oracle(p)
    o.load(f, isf)
    fl = np.abs(isf) == 1
    ini = o.get_field(FIELD_PHI)[fl].sum() - _open_flux(p, isf, [o.get_field(2 + a) for a in range(3)],
                                                        o.get_field(FIELD_PHI))
    o.step(5)
    o.begin_step()
    for sid in (1, 2, 11):  # compute_norm, PHI_int_boundary, the open faces' phi
        o.stage(sid)
    before = o.get_field(FIELD_PHI)
    u3 = [o.get_field(2 + a) for a in range(3)]  # the step's input moments
    chg = _open_flux(p, isf, u3, before)
    assert abs(chg) > 1e-4  # phase is carried in and out
    o.stage(10)  # the correction
    after = o.get_field(FIELD_PHI)
    band = fl & (before > 0.5) & (before < 0.9)
    assert band.sum() > 10
    corr = (ini + chg - before[fl].sum()) / float(np.float32(band.sum()))
    want = np.where(band, before + corr, before)
    assert np.abs(after - want).max() < 1e-12
    assert abs(after[fl].sum() - (ini + chg)) < 1e-11 * ini


def test_periodic_axes_carry_no_open_face():
    n = (8, 8, 8)
    p = make_params(n, lattice=27, highorder=True, internal_obstacles=True, pbc=(1, 1, 1), visc1=0.1,
                    fx=1e-5, precision=PREC_FP64)
    raw = setups.random_spheres(n, fill=0.1, rmin=1.5, rmax=2.0, seed=2, margin=1)
    isf = oracle.classify_isfluid(p, raw)
    shape = (n[2], n[1], n[0])
    out = []
    for obc in (0, 1):
        p.openbc = obc
        p.openbc_type[0], p.openbc_press[0] = 1, 0.3
        o = oracle.Oracle(p)
        o.load({k: np.zeros(shape) for k in KEYS}, isf)
        o.step(10)
        out.append(_moments(o))
    assert np.abs(out[0][1]).max() > 0
    assert np.array_equal(out[0], out[1])

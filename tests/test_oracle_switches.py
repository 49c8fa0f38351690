"""Oracle pins for the optional switches: BOUNCE_BACK, SMAGORINSKI and
REPULSIVE_FLUX.

* The default near-wall closure (LB_int_boundary): a body-force channel
  relaxes to an exact parabola with the bulk curvature -F/nu and its zeros
  within 0.05 of the solid layers' centres.
* BOUNCE_BACK (lb_cuda_fused_module.f90:20-424): halfway bounce-back conserves
  the fluid mass exactly, and a body-force driven channel relaxes to the
  Poiseuille parabola with the walls half a node inside the solid layer
  (the H_pois = (lx - 2)/2 of init_conditions_module.f90:248-252).
* SMAGORINSKI (fused:603-616): the relaxation rate read back from the
  post-collision populations tends to 1/(1/2 + nu) as the non-equilibrium
  stress vanishes and falls as the stress grows; one value is checked
  against the closed form with the single-precision constant 0.053*18.
* REPULSIVE_FLUX (lb_cuda_repulsive_module.f90): on a thin symmetric film the
  scan pairs each interfacial node with its mirror node, the flux is the
  closed form A q w(q) along the normal, antisymmetric across the film; a
  thick film has no facing partner within the 7^3 window.
* BCPHIFLUX (lb_cuda_boundary_module.f90:461-566, boundary_cds_module.f90:
  1864-1884): the correction stage equals a brute-force restatement (sum of
  phi over |isfluid| = 1 nodes, count of 0.5 < phi < 0.9, shift of the
  counted nodes by (initial sum - sum) / count); over a run the phase mass
  stays at its initial value to round-off, where without it it drifts.
"""
import numpy as np
import pytest

import oracle
from lbfast_amd import setups
from lbfast_amd.params import (make_params, PREC_FP64, FIELD_PHI, FIELD_AUX, FIELD_REP_MASK, FIELD_REP_FLUX,
                               FIELD_REP_PARTNER)


def _channel(n, bb, fx=0.0, noise=0.0, spheres=True, seed=3, lattice=19, visc=0.1):
    p = make_params(n, lattice=lattice, highorder=(lattice == 27), internal_obstacles=True, pbc=(1, 0, 1),
                    visc1=visc, fx=fx, precision=PREC_FP64, bounce_back=bb)
    raw = setups.random_spheres(n, fill=0.1, rmin=2.0, rmax=3.0, seed=seed, margin=2) if spheres else None
    isf = oracle.classify_isfluid(p, raw)
    rng = np.random.default_rng(seed)
    shape = (n[2], n[1], n[0])
    f = {k: np.zeros(shape) for k in ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")}
    if noise:
        f["P"] = rng.uniform(-noise, noise, shape)
        for k in ("u", "v", "w"):
            f[k] = rng.uniform(-noise, noise, shape)
        f["pxx"], f["pyy"], f["pzz"] = f["u"] ** 2, f["v"] ** 2, f["w"] ** 2
        f["pxy"], f["pxz"], f["pyz"] = f["u"] * f["v"], f["u"] * f["w"], f["v"] * f["w"]
    fl = isf != 0
    for k in f:
        f[k] = np.where(fl, f[k], 0.0)
    return p, f, isf


@pytest.mark.parametrize("lattice", [19, 27])
def test_bounce_back_conserves_fluid_mass(lattice):
    n = (16, 16, 16)
    p, f, isf = _channel(n, True, noise=1e-3, lattice=lattice)
    o = oracle.Oracle(p)
    o.load(f, isf)
    fl = isf != 0
    m0 = o.get_field(1)[fl].sum()
    o.step(20)
    m1 = o.get_field(1)[fl].sum()
    assert abs(m1 - m0) < 1e-13 * fl.sum()
    # the second-order near-wall closure is not mass conserving: the check has teeth
    p2, f2, isf2 = _channel(n, False, noise=1e-3, lattice=lattice)
    o2 = oracle.Oracle(p2)
    o2.load(f2, isf2)
    o2.step(20)
    assert abs(o2.get_field(1)[fl].sum() - m0) > 1e-9


def test_bounce_back_poiseuille():
    ly = 18
    n = (4, ly, 4)
    nu, fx = 1.0 / 6.0, 1e-6
    p, f, isf = _channel(n, True, fx=fx, spheres=False, visc=nu, lattice=19)
    o = oracle.Oracle(p)
    o.load(f, isf)
    o.step(4000)
    u = o.get_field(2)[0, :, 0]  # u_x(y)
    gj = np.arange(1, ly + 1, dtype=float)
    H, yc = 0.5 * (ly - 2), 0.5 * (ly + 1)
    want = fx / (2 * nu) * (H * H - (gj - yc) ** 2)
    inner = slice(1, ly - 1)
    err = np.abs(u[inner] - want[inner]).max() / want.max()
    assert err < 0.02, err
    # the walls sit half a node inside the solid layer: the nodes next to the
    # walls carry the parabola's value there, not zero
    assert u[1] > 0.5 * want[1]


def test_near_wall_closure_channel_is_parabolic():
    """The default near-wall closure (LB_int_boundary, lb_cuda_boundary_module.
    f90:117-212) in a body-force channel, walls in y, run to steady state: the
    fluid profile, the closure's nodes next to the walls included, is a
    parabola to round-off with the bulk curvature -F/nu, and its zeros (the
    effective walls) lie within 0.05 of the solid layers' centres.  A closure
    that mis-sets the wall nodes' populations bends the profile near the
    walls or moves its zeros by O(1)."""
    for ly, nu in ((18, 1.0 / 6.0), (18, 0.05)):
        n, fx = (3, ly, 3), 1e-6
        p = make_params(n, lattice=19, internal_obstacles=True, pbc=(1, 0, 1), visc1=nu, fx=fx,
                        precision=PREC_FP64)
        isf = oracle.classify_isfluid(p, None)
        shape = (n[2], n[1], n[0])
        f = {k: np.zeros(shape) for k in ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")}
        o = oracle.Oracle(p)
        o.load(f, isf)
        o.step(int(3 * ly * ly / nu))
        u = o.get_field(2)[0, :, 0]
        o.step(50)
        assert np.abs(o.get_field(2)[0, :, 0] - u).max() < 1e-12 * u.max()  # steady
        gj = np.arange(1, ly + 1, dtype=float)
        inner = slice(1, ly - 1)
        c = np.polyfit(gj[inner], u[inner], 2)
        assert np.abs(np.polyval(c, gj[inner]) - u[inner]).max() < 1e-9 * u.max()
        assert abs(2.0 * c[0] / (-fx / nu) - 1.0) < 1e-9
        lo, hi = sorted(np.roots(c).real)
        assert abs(lo - 1.0) < 0.05 and abs(hi - ly) < 0.05, (lo, hi)


def _smag_omega(pneq_scale, nu=0.02):
    """Relaxation rate of the collision read back from the populations of one
    rest node with a prescribed non-equilibrium stress (D3Q27 HIGHORDER)."""
    n = (4, 4, 4)
    p = make_params(n, lattice=27, highorder=True, visc1=nu, precision=PREC_FP64, smagorinsky=True)
    shape = (n[2], n[1], n[0])
    f = {k: np.zeros(shape) for k in ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")}
    pn = {"pxx": 1.0, "pyy": -0.4, "pzz": -0.6, "pxy": 0.3, "pxz": -0.2, "pyz": 0.5}
    for k, v in pn.items():
        f[k] = np.full(shape, v * pneq_scale)
    o = oracle.Oracle(p)
    o.load(f, None)
    o.begin_step()
    o.stage(4)
    o.stage(5)
    pop = o.populations()[:, 0, 0, 0]
    from lbfast_amd.lattice import EX, EY
    ex, ey = np.asarray(EX, float), np.asarray(EY, float)
    pxy_post = (pop * ex * ey).sum()
    Q = sum(v * v for v in (pn["pxx"], pn["pyy"], pn["pzz"])) + 2 * sum(
        pn[k] ** 2 for k in ("pxy", "pxz", "pyz"))
    return 1.0 - pxy_post / (pn["pxy"] * pneq_scale), Q * pneq_scale ** 2


def test_smagorinsky_relaxation():
    nu = 0.02
    w0, _ = _smag_omega(1e-14, nu)
    assert abs(w0 - 1.0 / (0.5 + nu)) < 1e-6  # vanishing stress: tau = 1/2 + nu as written
    ws = [_smag_omega(s, nu)[0] for s in (1e-6, 1e-4, 1e-2)]
    assert ws[0] > ws[1] > ws[2]  # more stress, more eddy viscosity
    w, Q = _smag_omega(1e-4, nu)
    C = float(np.float32(0.053) * np.float32(18.0))
    tau = 0.5 + (3 * nu + np.sqrt((3 * nu) ** 2 + C * np.sqrt(2 * Q))) / 6.0
    assert abs(w - 1.0 / tau) < 1e-10


def _film(thickness, width=2.0, n=(24, 8, 8), a_rep=0.05):
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.1, visc2=0.1,
                    rho_r=1.0, rho_b=0.1, tau_diff=0.1, sigma=0.01, width=width, precision=PREC_FP64,
                    repulsive_flux=True, a_rep=a_rep, q_th=0.125, cos_opp_t=-0.8)
    f = setups.thin_film(n, thickness, width, axis=0)
    o = oracle.Oracle(p)
    o.load(f, None)
    o.begin_step()
    for s in (1, 2, 3, 8, 9):
        o.stage(s)
    return p, o


def test_repulsive_thin_film():
    p, o = _film(4.0)
    mask = o.get_field(FIELD_REP_MASK)
    xs = np.unique(np.nonzero(mask)[2])
    assert list(xs) == [10, 13]  # the two interfacial planes (phi ~ 0.7)
    assert mask[:, :, 10].all() and mask[:, :, 13].all()
    part = np.stack([o.get_field(i) for i in FIELD_REP_PARTNER])
    assert np.all(part[0][:, :, 10] == 14) and np.all(part[0][:, :, 13] == 11)  # mirror node, 1-based
    flux = np.stack([o.get_field(i) for i in FIELD_REP_FLUX])
    assert np.abs(flux[1:]).max() < 1e-14
    fx_l, fx_r = flux[0][:, :, 10], flux[0][:, :, 13]
    assert np.allclose(fx_l, -fx_r, rtol=0, atol=1e-15)
    # closed form for a symmetric pair: q1 = q2 = q; the facing factor is
    # -n1.n2 = |n|^2 with the stored normals n = g / (|g| + 1e-9)
    phi = o.get_field(FIELD_PHI)[0, 0, 10]
    q = phi * (1 - phi)
    n1, n2 = o.get_field(FIELD_AUX[0])[0, 0, 10], o.get_field(FIELD_AUX[0])[0, 0, 13]
    face = -n1 * n2
    ach = np.arccosh(1.0 / (2.0 * np.sqrt(q)))
    wdth = 1.0 / (1.0 + (p.width * ach) ** 4)
    f = p.a_rep * wdth * q * face
    want = f * min(1.0, phi / (1.5 * f))
    assert abs(fx_l[0, 0] - want) < 1e-12 * abs(want) + 1e-16
    assert fx_l[0, 0] > 0  # along the normal, oriented towards the partner


def test_repulsive_thick_film_has_no_partner():
    p, o = _film(12.0)
    assert o.get_field(FIELD_REP_MASK).sum() == 0
    assert np.abs(np.stack([o.get_field(i) for i in FIELD_REP_FLUX])).max() == 0


def test_repulsive_force_reaches_the_moments():
    runs = []
    for a in (0.0, 0.05):
        p, o = _film(4.0, a_rep=a)
        o.stage(4)  # moments: the repulsive flux enters the force
        runs.append(o.get_field(2))
    d = np.abs(runs[1] - runs[0])
    assert d.max() > 1e-6
    assert d[:, :, :8].max() < 1e-14 and d[:, :, 16:].max() < 1e-14  # only near the film


def _bcphi_case(n=(24, 20, 16), bcphi=True, seed=5):
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, internal_obstacles=True,
                    pbc=(1, 0, 1), visc1=0.05, visc2=0.1, rho_r=2.0, rho_b=1.0, tau_diff=0.1, sigma=0.005,
                    width=4.0, fx=1e-5, precision=PREC_FP64, bcphiflux=bcphi)
    raw = setups.random_spheres(n, fill=0.15, rmin=2.0, rmax=3.0, seed=seed, margin=2)
    isf = oracle.classify_isfluid(p, raw)
    f = setups.two_layer_channel(n, raw, frac=0.4, width=4.0, rho_r=2.0, rho_b=1.0, axis=1)
    f["phi"] = setups.solid_phi_average(f["phi"], isf, 27)
    return p, f, isf


def test_bcphiflux_stage_equals_brute_force():
    p, f, isf = _bcphi_case()
    o = oracle.Oracle(p)
    o.load(f, isf)
    fl = np.abs(isf) == 1
    ini = o.get_field(FIELD_PHI)[fl].sum()  # after the initial correction the sum is the initial one
    o.step(3)
    o.begin_step()
    o.stage(1)  # compute_norm
    o.stage(2)  # PHI_int_boundary
    before = o.get_field(FIELD_PHI)
    o.stage(10)  # the correction
    after = o.get_field(FIELD_PHI)
    band = fl & (before > 0.5) & (before < 0.9)
    assert band.sum() > 10
    corr = (ini - before[fl].sum()) / float(np.float32(band.sum()))
    want = np.where(band, before + corr, before)
    assert abs(corr) > 1e-12  # the phase update does drift between corrections
    # the oracle sums ~8k values in sequence, numpy pairwise: the sums differ
    # by ~1e-11, the corrections by ~1e-14
    assert np.abs(after - want).max() < 1e-12
    assert abs(after[fl].sum() - ini) < 1e-12 * ini


def test_bcphiflux_keeps_phase_mass():
    drift = {}
    for bcphi in (True, False):
        p, f, isf = _bcphi_case(bcphi=bcphi)
        o = oracle.Oracle(p)
        o.load(f, isf)
        fl = np.abs(isf) == 1
        m0 = o.get_field(FIELD_PHI)[fl].sum()
        worst = 0.0
        for _ in range(30):
            # the step's stage order (STEP_ORDER), sampling the phase mass
            # right after the correction stage
            o.begin_step()
            for sid in (1, 2, 10):
                o.stage(sid)
            worst = max(worst, abs(o.get_field(FIELD_PHI)[fl].sum() - m0) / m0)
            for sid in (3, 8, 9, 4, 5, 6, 7):
                o.stage(sid)
            o.end_step()
        drift[bcphi] = worst
    assert drift[True] < 1e-13
    assert drift[False] > 1e3 * max(drift[True], 1e-16)

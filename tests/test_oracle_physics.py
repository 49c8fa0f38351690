"""Pin the oracle's stages by closed forms and invariants.

Every check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign or index, a transposed operand, a wrong streaming
direction) fails it:
  * rest state and uniform flow are fixed points of the full step;
  * the rebuilt post-collision populations have the moments the regularised
    collision prescribes (P, u + a/2, u u + (1-omega) Pn + sym(u a));
  * streaming is the exact permutation f_l(x) <- f*_l(x - c_l) (checked by an
    independent numpy roll of the population dump);
  * total P is conserved and total momentum grows by exactly N F per step;
  * the isotropic D3Q27 stencils are exact on linear (gradient, divergence)
    and quadratic (laplacian) fields;
  * the moments stage of a one-component run is u + F/2, Pi + sym(F u);
  * the phase update leaves a constant phi at rest unchanged;
  * isfluid classification matches a brute-force restatement;
  * the wetting ghost update reduces to phi_ghost = phi_donor at 90 degrees
    and to phi_donor - |grad_t| cot(theta) 4 phi(1-phi) otherwise, on a flat wall,
    and to phi_donor - sqrt 2 |t| cot(theta) 4 phi(1-phi) on a 45-degree
    staircase wall (normal, donor and tangential part derived by hand);
  * the surface-tension force of the moments stage, (4 beta phi (phi-1)
    (phi-1/2) - kappa lap phi) grad phi with beta = 12 sigma/w and
    kappa = 1.5 sigma w, vanishes on the equilibrium tanh profile of width w
    (the profile the chemical potential's zero fixes: xi^2 = 2 kappa/beta =
    (w/2)^2), with a residual of second order in 1/w.
"""
import numpy as np
import pytest

import oracle
from lbfast_amd import setups
from lbfast_amd.lattice import EX, EY, EZ, weights
from lbfast_amd.params import (make_params, PREC_FP64, HFIELDS, HFIELD_NAMES, FIELD_PHI, FIELD_AUX,
                               FIELD_LAPPHI, FIELD_DIVTHN, FIELD_FORCE, STAGE_NORM, STAGE_DIV_THETA_N,
                               STAGE_MOMENTS, STAGE_FUSED, STAGE_UPDATE_PHI, STAGE_PHI_INT_BOUNDARY,
                               FIELD_PHI_O)

SHAPE = (12, 10, 8)


def _zeros(shape=SHAPE):
    nz = (shape[2], shape[1], shape[0])
    return {k: np.zeros(nz) for k in HFIELD_NAMES}


def _random_state(rng, shape=SHAPE, u_amp=0.03, two=False):
    nz = (shape[2], shape[1], shape[0])
    u, v, w = (rng.uniform(-u_amp, u_amp, nz) for _ in range(3))
    f = {"P": rng.uniform(-1e-2, 1e-2, nz), "u": u, "v": v, "w": w}
    for k, (a, b) in zip(("pxx", "pyy", "pzz", "pxy", "pxz", "pyz"),
                         ((u, u), (v, v), (w, w), (u, v), (u, w), (v, w))):
        f[k] = a * b + rng.uniform(-1e-4, 1e-4, nz)
    if two:
        f["phi"] = rng.uniform(0.0, 1.0, nz)
    return f


@pytest.mark.parametrize("lat,ho", [(19, False), (27, False), (27, True)])
def test_rest_state_fixed_point(lat, ho):
    p = make_params(SHAPE, lattice=lat, highorder=ho, visc1=0.05, precision=PREC_FP64)
    o = oracle.Oracle(p)
    o.load(_zeros())
    o.step(3)
    assert np.abs(o.hfields()).max() == 0.0


@pytest.mark.parametrize("lat,ho", [(19, False), (27, False), (27, True)])
def test_uniform_flow_fixed_point(lat, ho):
    p = make_params(SHAPE, lattice=lat, highorder=ho, visc1=0.05, precision=PREC_FP64)
    f = _zeros()
    u0 = (0.031, -0.017, 0.022)
    f["P"] += 0.004
    f["u"] += u0[0]; f["v"] += u0[1]; f["w"] += u0[2]
    f["pxx"] += u0[0] ** 2; f["pyy"] += u0[1] ** 2; f["pzz"] += u0[2] ** 2
    f["pxy"] += u0[0] * u0[1]; f["pxz"] += u0[0] * u0[2]; f["pyz"] += u0[1] * u0[2]
    o = oracle.Oracle(p)
    o.load(f)
    o.step(4)
    h = o.hfields()
    for i, nm in enumerate(HFIELD_NAMES):
        assert np.abs(h[i] - f[nm]).max() < 1e-15, nm


@pytest.mark.parametrize("lat,ho", [(19, False), (27, False), (27, True)])
def test_collision_moments_closure(lat, ho):
    """Moments of the rebuilt populations before streaming (fused:637-713)."""
    rng = np.random.Generator(np.random.PCG64(5))
    F = (2e-4, -3e-4, 1e-4)
    visc = 0.07
    p = make_params(SHAPE, lattice=lat, highorder=ho, visc1=visc, precision=PREC_FP64,
                    fx=F[0], fy=F[1], fz=F[2])
    f = _random_state(rng)
    o = oracle.Oracle(p)
    o.load(f)
    o.begin_step()
    o.stage(STAGE_MOMENTS)
    # state after the moments stage (u + F/2, Pi + sym(F u)) is what fused reads
    post = o.hfields()
    o.stage(STAGE_FUSED)
    pop = o.populations()[: lat]
    c = np.array([EX[:lat], EY[:lat], EZ[:lat]], dtype=float)
    om = 1.0 / (3 * visc + 0.5)
    P, u = post[0], post[1:4]
    Pi = np.array([[post[4], post[7], post[8]], [post[7], post[5], post[9]], [post[8], post[9], post[6]]])
    a = np.array(F)[:, None, None, None] * np.ones_like(P)
    m0 = pop.sum(0)
    m1 = np.einsum("al,l...->a...", c, pop)
    m2 = np.einsum("al,bl,l...->ab...", c, c, pop) - np.eye(3)[:, :, None, None, None] / 3.0 * m0
    uu = np.einsum("a...,b...->ab...", u, u)
    expect2 = uu + (1 - om) * (Pi - uu) + 0.5 * (np.einsum("a...,b...->ab...", u, a) +
                                                  np.einsum("a...,b...->ab...", a, u))
    assert np.abs(m0 - P).max() < 1e-15
    assert np.abs(m1 - (u + 0.5 * a)).max() < 1e-15
    assert np.abs(m2 - expect2).max() < 1e-15


@pytest.mark.parametrize("lat,ho", [(19, False), (27, True)])
def test_streaming_is_exact_permutation(lat, ho):
    rng = np.random.Generator(np.random.PCG64(9))
    p = make_params(SHAPE, lattice=lat, highorder=ho, visc1=0.05, precision=PREC_FP64, fx=1e-4)
    o = oracle.Oracle(p)
    o.load(_random_state(rng))
    o.begin_step()
    o.stage(STAGE_MOMENTS)
    o.stage(STAGE_FUSED)
    pop = o.populations()[:lat]
    o.end_step()
    out = o.hfields()  # the new S buffer = output of fused
    # pull: f_l(x) = f*_l(x - c_l); arrays are [z, y, x]
    streamed = np.stack([np.roll(pop[l], shift=(EZ[l], EY[l], EX[l]), axis=(0, 1, 2)) for l in range(lat)])
    c = np.array([EX[:lat], EY[:lat], EZ[:lat]], dtype=float)
    m0 = streamed.sum(0)
    m1 = np.einsum("al,l...->a...", c, streamed)
    m2 = np.einsum("al,bl,l...->ab...", c, c, streamed) - np.eye(3)[:, :, None, None, None] / 3.0 * m0
    assert np.abs(out[0] - m0).max() < 1e-15
    assert np.abs(out[1:4] - m1).max() < 1e-15
    for i, (a, b) in enumerate(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))):
        assert np.abs(out[4 + i] - m2[a, b]).max() < 1e-15


@pytest.mark.parametrize("lat,ho", [(19, False), (27, True)])
def test_mass_conserved_momentum_injected(lat, ho):
    rng = np.random.Generator(np.random.PCG64(3))
    F = np.array([3e-5, -1e-5, 2e-5])
    p = make_params(SHAPE, lattice=lat, highorder=ho, visc1=0.1, precision=PREC_FP64,
                    fx=F[0], fy=F[1], fz=F[2])
    o = oracle.Oracle(p)
    f = _random_state(rng)
    o.load(f)
    N = np.prod(SHAPE)
    m0 = o.hfields()
    for s in range(1, 4):
        o.step(1)
        h = o.hfields()
        assert abs(h[0].sum() - m0[0].sum()) < 1e-12
        for a in range(3):
            assert abs(h[1 + a].sum() - (m0[1 + a].sum() + s * N * F[a])) < 1e-12


def _two_params(shape=SHAPE, **kw):
    base = dict(lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.1, visc2=0.05,
                rho_r=1.0, rho_b=0.2, tau_diff=0.1, sigma=0.01, width=3.0, precision=PREC_FP64)
    base.update(kw)
    return make_params(shape, **base)


def test_gradient_exact_on_linear_and_laplacian_on_quadratic():
    shape = (16, 16, 16)
    p = _two_params(shape)
    o = oracle.Oracle(p)
    k, j, i = np.meshgrid(*(np.arange(n, dtype=float) for n in (16, 16, 16)), indexing="ij")
    g = np.array([0.011, -0.007, 0.005])
    phi = 0.3 + g[0] * i + g[1] * j + g[2] * k
    f = _zeros(shape)
    f["phi"] = phi
    o.load(f)
    o.begin_step()
    o.stage(STAGE_NORM)
    n_ = np.stack([o.get_field(FIELD_AUX[a]) for a in range(3)])
    mg = o.get_field(FIELD_AUX[3])
    inner = (slice(2, -2),) * 3
    grad = n_ * mg  # the moments stage reads n * |g| with n = g / (|g| + 1e-9)
    gn = np.linalg.norm(g)
    assert np.abs(mg[inner] - gn).max() < 1e-14
    for a in range(3):
        assert np.abs(grad[a][inner] - g[a] * gn / (gn + 1e-9)).max() < 1e-15
    assert np.abs(o.get_field(FIELD_LAPPHI)[inner]).max() < 1e-12
    # quadratic: lap(x^2 + 2 y^2 - z^2) = 2 + 4 - 2 = 4
    o2 = oracle.Oracle(p)
    f2 = _zeros(shape)
    f2["phi"] = 1e-3 * (i ** 2 + 2 * j ** 2 - k ** 2)
    o2.load(f2)
    o2.begin_step()
    o2.stage(STAGE_NORM)
    assert np.abs(o2.get_field(FIELD_LAPPHI)[inner] - 4e-3).max() < 1e-12


def test_divergence_exact_on_linear_theta():
    shape = (14, 12, 10)
    p = _two_params(shape)
    o = oracle.Oracle(p)
    o.load(_zeros(shape) | {"phi": np.full((10, 12, 14), 0.5)})
    k, j, i = np.meshgrid(*(np.arange(n, dtype=float) for n in (10, 12, 14)), indexing="ij")
    o.set_field(FIELD_AUX[4], 0.01 * i + 0.2 * j)   # theta_x
    o.set_field(FIELD_AUX[5], -0.03 * j + 0.1 * k)  # theta_y
    o.set_field(FIELD_AUX[6], 0.05 * k - 0.3 * i)   # theta_z
    o.begin_step()
    o.stage(STAGE_DIV_THETA_N)
    d = o.get_field(FIELD_DIVTHN)
    inner = (slice(2, -2),) * 3
    assert np.abs(d[inner] - (0.01 - 0.03 + 0.05)).max() < 1e-12


def test_moments_stage_one_component_closed_form():
    rng = np.random.Generator(np.random.PCG64(1))
    F = np.array([1e-4, 2e-4, -5e-5])
    p = make_params(SHAPE, lattice=27, highorder=True, visc1=0.1, precision=PREC_FP64,
                    fx=F[0], fy=F[1], fz=F[2])
    f = _random_state(rng)
    o = oracle.Oracle(p)
    o.load(f)
    o.begin_step()
    o.stage(STAGE_MOMENTS)
    h = o.hfields()
    u = np.stack([f["u"], f["v"], f["w"]]) + 0.5 * F[:, None, None, None]
    assert np.abs(h[1:4] - u).max() < 1e-16
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    for n_, (a, b) in zip(("pxx", "pyy", "pzz", "pxy", "pxz", "pyz"), pairs):
        expect = f[n_] + 0.5 * (F[a] * u[b] + F[b] * u[a])
        assert np.abs(h[HFIELD_NAMES.index(n_)] - expect).max() < 1e-16
    for a in range(3):
        assert np.abs(o.get_field(FIELD_FORCE[a]) - F[a]).max() == 0.0


def test_phase_update_constant_phi_at_rest():
    p = _two_params()
    f = _zeros()
    f["phi"] = np.full((SHAPE[2], SHAPE[1], SHAPE[0]), 0.37)
    o = oracle.Oracle(p)
    o.load(f)
    o.step(3)
    assert np.abs(o.get_field(FIELD_PHI) - 0.37).max() < 1e-15
    assert np.abs(o.hfields()).max() < 1e-15


def _classify_brute(p, raw):
    """LBFAST.f90:258-328 restated independently (pure Python)."""
    nz, ny, nx = raw.shape
    a = raw.astype(int).copy()
    if p.internal_obstacles:
        if not p.pbc[0]:
            a[:, :, 0] = 0; a[:, :, -1] = 0
        if not p.pbc[1]:
            a[:, 0, :] = 0; a[:, -1, :] = 0
        if not p.pbc[2]:
            a[0, :, :] = 0; a[-1, :, :] = 0

    def look(i, j, k):
        c, n = [i, j, k], [nx, ny, nz]
        for d in range(3):
            if c[d] < 0 or c[d] >= n[d]:
                if not p.pbc[d]:
                    return 1
                c[d] %= n[d]
        return a[c[2], c[1], c[0]]

    out = a.copy()
    nl = 27 if p.lattice == 27 else 19
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                if a[k, j, i] == 1 and any(abs(look(i + EX[l], j + EY[l], k + EZ[l])) != 1 for l in range(1, nl)):
                    out[k, j, i] = -1
    return out.astype(np.int8)


@pytest.mark.parametrize("lat", [19, 27])
def test_classification_matches_brute_force(lat):
    shape = (12, 9, 10)
    p = make_params(shape, lattice=lat, highorder=False, internal_obstacles=True, pbc=(1, 0, 1),
                    precision=PREC_FP64)
    rng = np.random.Generator(np.random.PCG64(11))
    raw = (rng.uniform(size=(10, 9, 12)) > 0.1).astype(np.int8)
    assert (oracle.classify_isfluid(p, raw) == _classify_brute(p, raw)).all()
    from lbfast_amd import hip
    assert (hip.classify_isfluid(p, raw) == _classify_brute(p, raw)).all()


@pytest.mark.parametrize("angle", [90.0, 60.0, 120.0])
def test_wetting_ghost_flat_wall(angle):
    """Solid slab x <= 2 (walls in x); phi = phi0 + gy*y + gz*z in the fluid,
    so the donor gradient is purely tangential.  A wall node's ghost value
    becomes phi_donor - |grad_t| cot(theta') 4 phi_g (1 - phi_g), theta' =
    180 - angle (lb_cuda_boundary_module.f90:409-447)."""
    shape = (12, 10, 10)
    p = _two_params(shape, internal_obstacles=True, pbc=(0, 1, 1), wettab_r=angle)
    raw = np.ones((10, 10, 12), dtype=np.int8)
    raw[:, :, :3] = 0  # solid x = 0, 1, 2 (plus the wall at x = nx-1)
    isf = oracle.classify_isfluid(p, raw)
    k, j, i = np.meshgrid(*(np.arange(n, dtype=float) for n in (10, 10, 12)), indexing="ij")
    gy, gz = 0.004, -0.003
    phi = np.full(i.shape, 0.4)
    f = _zeros(shape)
    f["phi"] = phi
    o = oracle.Oracle(p)
    o.load(f, isf)
    # impose a tangential donor gradient directly in aux (n |g| = (0, gy, gz))
    g = np.hypot(gy, gz)
    o.set_field(FIELD_AUX[0], np.zeros_like(phi))
    o.set_field(FIELD_AUX[1], np.full_like(phi, gy / g))
    o.set_field(FIELD_AUX[2], np.full_like(phi, gz / g))
    o.set_field(FIELD_AUX[3], np.full_like(phi, g))
    ghost0 = 0.25
    phis = phi.copy()
    phis[:, :, :3] = ghost0
    o.set_field(FIELD_PHI, phis)
    o.begin_step()
    o.stage(STAGE_PHI_INT_BOUNDARY)
    got = o.get_field(FIELD_PHI)[:, :, 2]  # the wall layer next to the fluid
    th = (180.0 - angle) * np.pi / 180.0
    cot = np.cos(th) / np.sin(th)
    expect = np.clip(0.4 - g * cot * 4 * ghost0 * (1 - ghost0), 0, 1)
    assert np.abs(got - expect).max() < 1e-13
    # deeper solid nodes have no fluid neighbour along the normal: unchanged
    assert np.abs(o.get_field(FIELD_PHI)[:, :, 0] - ghost0).max() == 0.0


@pytest.mark.parametrize("angle", [90.0, 60.0, 135.0])
def test_wetting_ghost_45_degree_wall(angle):
    """A 45-degree staircase wall, solid where i + j <= 14 (0-based, z
    periodic).  Hand derivation for a staircase node (i + j = 14) and the
    layer below it (i + j = 13): the weighted sum over the non-solid lattice
    neighbours points along -(1, 1, 0) (19/6 and 7/6 times it), so the unit
    wall normal is wn = -(1, 1, 0)/sqrt 2; of the fluid neighbours the
    diagonal (1, 1, 0) scores highest ((e . -wn) |e|^-1 = 1 against 0.816 for
    (1, 1, +-1) and 0.707 for the axes), so the donor sits sqrt 2 away along
    the normal with back-direction bs = (-1, -1, 0).  With donor gradient g,
    tangential part t = g - (g . wn) wn and b = t - |t| cot(theta') wn,
    b . bs = -sqrt 2 |t| cot(theta'), so the ghost value is
    phi_donor - sqrt 2 |t| cot(theta') 4 phi_g (1 - phi_g) -- the flat-wall
    rule with the donor distance sqrt 2.  Nodes with i + j <= 12 have no
    non-solid neighbour and keep their value."""
    shape = (16, 16, 4)
    p = _two_params(shape, internal_obstacles=True, pbc=(0, 0, 1), wettab_r=angle)
    k, j, i = np.meshgrid(*(np.arange(n) for n in (4, 16, 16)), indexing="ij")
    raw = np.where(i + j <= 14, 0, 1).astype(np.int8)
    isf = oracle.classify_isfluid(p, raw)
    f = _zeros(shape)
    f["phi"] = np.full(i.shape, 0.4)
    o = oracle.Oracle(p)
    o.load(f, isf)
    g = np.array([0.003, 0.001, -0.002])  # normal, in-plane tangential and z parts
    mg = np.linalg.norm(g)
    for c in range(3):
        o.set_field(FIELD_AUX[c], np.full(i.shape, g[c] / mg))
    o.set_field(FIELD_AUX[3], np.full(i.shape, mg))
    ghost0 = 0.25
    phis = np.where(isf == 0, ghost0, 0.4)
    o.set_field(FIELD_PHI, phis)
    o.begin_step()
    o.stage(STAGE_PHI_INT_BOUNDARY)
    got = o.get_field(FIELD_PHI)
    wn = -np.array([1.0, 1.0, 0.0]) / np.sqrt(2.0)
    t = g - g.dot(wn) * wn
    th = (180.0 - angle) * np.pi / 180.0
    cot = np.cos(th) / np.sin(th)
    expect = np.clip(0.4 - np.sqrt(2.0) * np.linalg.norm(t) * cot * 4 * ghost0 * (1 - ghost0), 0, 1)
    inside = (i >= 3) & (i <= 12) & (j >= 3) & (j <= 12)
    for layer in (14, 13):
        sel = inside & (i + j == layer)
        assert sel.sum() > 0 and (isf[sel] == 0).all()
        assert np.abs(got[sel] - expect).max() < 1e-14, (layer, got[sel][:3], expect)
    sel = inside & (i + j <= 12)
    assert np.abs(got[sel] - ghost0).max() == 0.0


def test_surface_tension_vanishes_on_the_equilibrium_profile():
    """A planar slab phi = fcut_tanh(|x - xc|, 96, w) at rest with equal
    densities: the force the moments stage stores is the surface-tension term
    alone.  Relative to the size of its bulk part, 4 beta phi(phi-1)(phi-1/2)
    grad phi, it falls as w^-2 (the stencils' truncation error): below 1 % at
    w = 16, by a factor 4 +- 1 per doubling of w.  A wrong beta or kappa
    coefficient, a sign slip between the two terms or a wrong Laplacian leaves
    an O(1) residual that does not fall with w."""
    ratio = {}
    for w in (8.0, 16.0, 32.0):
        n, sigma = (384, 3, 3), 0.01
        p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.1, visc2=0.1,
                        rho_r=1.0, rho_b=1.0, tau_diff=0.1, sigma=sigma, width=w, precision=PREC_FP64)
        shape = (n[2], n[1], n[0])
        x = np.arange(1, n[0] + 1, dtype=float)
        phi1 = setups.fcut_tanh(np.abs(x - 192.5), 96.0, w)
        f = _zeros(n)
        f["phi"] = np.broadcast_to(phi1, shape).copy()
        o = oracle.Oracle(p)
        o.load(f)
        o.begin_step()
        for st in (STAGE_NORM, STAGE_DIV_THETA_N, STAGE_MOMENTS):
            o.stage(st)
        fx = o.get_field(FIELD_FORCE[0])[0, 0]
        assert np.abs(o.get_field(FIELD_FORCE[1])).max() == 0.0
        beta = 12.0 * sigma / w
        gx = 0.5 * (np.roll(phi1, -1) - np.roll(phi1, 1))
        bulk = 4.0 * beta * phi1 * (phi1 - 1.0) * (phi1 - 0.5) * gx
        ratio[w] = np.abs(fx).max() / np.abs(bulk).max()
    assert ratio[16.0] < 0.01, ratio
    assert 3.0 < ratio[8.0] / ratio[16.0] < 5.0, ratio
    assert 3.0 < ratio[16.0] / ratio[32.0] < 5.0, ratio

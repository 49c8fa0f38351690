"""Initial fields for the benchmark workloads, computed on the host in FP64.

These restate the reference's host-side initial conditions
(init_conditions_module.f90:177-589) so the same arrays can be uploaded to
the GPU library and to any checker.  Arrays are numpy, shape (nz, ny, nx)
C-order, which is the reference's Fortran (i,j,k) order when flattened.

Every builder returns a dict with keys ``P u v w pxx pyy pzz pxy pxz pyz``
(the reference hfields order) and, for two-component runs, ``phi``; plus
``isfluid`` (int8, 1/0 before classification) when obstacles are present.
"""
from __future__ import annotations

import numpy as np

CSSQ = 1.0 / 3.0


def _grid(shape):
    nx, ny, nz = shape
    gk, gj, gi = np.meshgrid(np.arange(1, nz + 1, dtype=np.float64),
                             np.arange(1, ny + 1, dtype=np.float64),
                             np.arange(1, nx + 1, dtype=np.float64), indexing="ij")
    return gi, gj, gk


def _moments(P, u, v, w):
    # Pi = u u at start (init_conditions_module.f90:437-442)
    return {"P": P, "u": u, "v": v, "w": w, "pxx": u * u, "pyy": v * v, "pzz": w * w,
            "pxy": u * v, "pxz": u * w, "pyz": v * w}


def fcut_tanh(r, r0, width):
    """vars_module.f90:1306-1329: 0.5*(1 - tanh((r - r0)/(width/2)))."""
    return 0.5 * (1.0 - np.tanh((r - r0) / (0.5 * width)))


def taylor_green(shape, u0=0.02):
    """Taylor-Green vortex, one component (init_conditions_module.f90:298-314).

    x = 2 pi (gi-1)/lx; u = U0 sin x cos y cos z; v = -U0 cos x sin y cos z; w = 0;
    P = (U0^2/16)(cos 2x + cos 2y)(2 + cos 2z)/cs^2."""
    nx, ny, nz = shape
    gi, gj, gk = _grid(shape)
    x = 2.0 * np.pi * (gi - 1.0) / nx
    y = 2.0 * np.pi * (gj - 1.0) / ny
    z = 2.0 * np.pi * (gk - 1.0) / nz
    u = u0 * np.sin(x) * np.cos(y) * np.cos(z)
    v = -u0 * np.cos(x) * np.sin(y) * np.cos(z)
    w = np.zeros_like(u)
    P = ((u0 * u0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (2.0 + np.cos(2 * z))) / (1.0 * CSSQ)
    return _moments(P, u, v, w)


def _pbc_images(d, L):
    # init_conditions_module.f90:591-599: d - L*nint(d/L)
    return d - L * np.rint(d / L)


def droplet(shape, radius, width, center, sigma, rho_r, rho_b, densratio=True,
            noise=0.0, seed=20261202, ellipse_a=0.0):
    """Two-component droplet (init_conditions_module.f90:394-426).

    phi = fcut_tanh(r, R, w) = 0.5(1 + tanh(2(R - r)/w)); P = phi 2 sigma/R/(rhophi cs^2);
    velocity zero plus optional uniform noise U(-noise, noise) from a seeded
    generator.  ``ellipse_a`` > 0 perturbs the radius as R(1 + a P2(cos theta))
    (an l = 2 mode about the z axis)."""
    nx, ny, nz = shape
    gi, gj, gk = _grid(shape)
    dx = _pbc_images(gi - center[0], nx)
    dy = _pbc_images(gj - center[1], ny)
    dz = _pbc_images(gk - center[2], nz)
    r = np.sqrt(dx * dx + dy * dy + dz * dz)
    if ellipse_a != 0.0:
        ct = np.where(r > 0, dz / np.maximum(r, 1e-300), 1.0)
        p2 = 0.5 * (3.0 * ct * ct - 1.0)
        rad = radius * (1.0 + ellipse_a * p2)
    else:
        rad = radius
    phi = fcut_tanh(r, rad, width)
    rhophi = rho_r * phi + (1.0 - phi) * rho_b if densratio else np.ones_like(phi)
    P = phi * (sigma * 2.0) / radius / (rhophi * CSSQ)
    if noise > 0.0:
        rng = np.random.Generator(np.random.MT19937(seed))
        u = rng.uniform(-noise, noise, size=phi.shape)
        v = rng.uniform(-noise, noise, size=phi.shape)
        w = rng.uniform(-noise, noise, size=phi.shape)
    else:
        u = np.zeros_like(phi)
        v = np.zeros_like(phi)
        w = np.zeros_like(phi)
    out = _moments(P, u, v, w)
    out["phi"] = phi
    return out


def random_spheres(shape, fill=0.25, rmin=6.0, rmax=14.0, seed=42, wall_axis=1, margin=2):
    """Non-overlapping spheres by random sequential addition until the solid
    fraction of the slab reaches ``fill``.  Spheres stay ``margin`` nodes away
    from every face of the box so that no donor or wall-normal search reaches
    the halo.  Returns the raw isfluid (1 fluid, 0 solid), int8 (nz, ny, nx)."""
    nx, ny, nz = shape
    rng = np.random.Generator(np.random.MT19937(seed))
    isf = np.ones((nz, ny, nx), dtype=np.int8)
    gi, gj, gk = _grid(shape)
    total = nx * ny * nz
    solid = 0
    spheres = []
    tries = 0
    while solid < fill * total and tries < 200000:
        tries += 1
        r = rng.uniform(rmin, rmax)
        lo = r + margin + 1.0
        c = np.array([rng.uniform(lo, nx + 1 - lo), rng.uniform(lo, ny + 1 - lo),
                      rng.uniform(lo, nz + 1 - lo)])
        if any(np.linalg.norm(c - c2) < r + r2 + 1.0 for c2, r2 in spheres):
            continue
        spheres.append((c, r))
        x0, x1 = int(max(1, np.floor(c[0] - r))), int(min(nx, np.ceil(c[0] + r)))
        y0, y1 = int(max(1, np.floor(c[1] - r))), int(min(ny, np.ceil(c[1] + r)))
        z0, z1 = int(max(1, np.floor(c[2] - r))), int(min(nz, np.ceil(c[2] + r)))
        sub = ((gi[z0 - 1:z1, y0 - 1:y1, x0 - 1:x1] - c[0]) ** 2 +
               (gj[z0 - 1:z1, y0 - 1:y1, x0 - 1:x1] - c[1]) ** 2 +
               (gk[z0 - 1:z1, y0 - 1:y1, x0 - 1:x1] - c[2]) ** 2) <= r * r
        blk = isf[z0 - 1:z1, y0 - 1:y1, x0 - 1:x1]
        solid += int(np.count_nonzero(sub & (blk == 1)))
        blk[sub] = 0
    return isf


def thin_film(shape, thickness, width, axis=0, noise=0.0, seed=5):
    """A slab of phase phi = 1 between two flat tanh interfaces a distance
    ``thickness`` apart, centred in the box along ``axis`` (facing interfaces
    for REPULSIVE_FLUX); P = 0, velocity uniform noise U(-noise, noise)."""
    gi, gj, gk = _grid(shape)
    coord = (gi, gj, gk)[axis]
    L = shape[axis]
    xc = 0.5 * (L + 1)
    x1, x2 = xc - 0.5 * thickness, xc + 0.5 * thickness
    phi = 0.5 * (np.tanh(2.0 * (coord - x1) / width) - np.tanh(2.0 * (coord - x2) / width))
    z = np.zeros_like(phi)
    if noise > 0:
        rng = np.random.default_rng(seed)
        u, v, w = (rng.uniform(-noise, noise, phi.shape) for _ in range(3))
    else:
        u, v, w = z.copy(), z.copy(), z.copy()
    out = _moments(z.copy(), u, v, w)
    out["phi"] = phi
    return out


def two_layer_channel(shape, isfluid, frac=0.4, width=4.0, rho_r=10.0, rho_b=1.0, axis=1):
    """Dense phase (phi = 1) in the lower ``frac`` of the channel height along
    ``axis`` with a tanh interface; fluid at rest; P = 0."""
    nx, ny, nz = shape
    gi, gj, gk = _grid(shape)
    coord = (gi, gj, gk)[axis]
    L = shape[axis]
    phi = fcut_tanh(coord, frac * L, width)
    z = np.zeros_like(phi)
    out = _moments(z.copy(), z.copy(), z.copy(), z.copy())
    out["phi"] = phi
    return out


def solid_phi_average(phi, isf_classified, lattice=27):
    """Initial phi at solid nodes: weighted mean of in-range non-solid
    neighbours with weights 1, 1/sqrt2, 1/sqrt3 (init_conditions_module.f90:468-542)."""
    from .lattice import EX, EY, EZ
    nl = 27 if lattice == 27 else 19
    nz, ny, nx = phi.shape
    out = phi.copy()
    ks, js, is_ = np.nonzero(isf_classified == 0)
    for k, j, i in zip(ks, js, is_):
        num = 0.0
        den = 0.0
        for l in range(1, nl):
            ii, jj, kk = i + EX[l], j + EY[l], k + EZ[l]
            if not (0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz):
                continue
            if isf_classified[kk, jj, ii] == 0:
                continue
            d2 = EX[l] ** 2 + EY[l] ** 2 + EZ[l] ** 2
            s = 1.0 if d2 == 1 else (1.0 / np.sqrt(2.0) if d2 == 2 else 1.0 / np.sqrt(3.0))
            num += s * phi[kk, jj, ii]
            den += s
        # den > 0 always holds here for solids adjacent to fluid; fully
        # interior solid nodes get 0/0 in the reference -> keep 0 instead
        out[k, j, i] = num / den if den > 0 else 0.0
    return out

"""The separable form of the phase-field stencils.

The fast path evaluates the gradient, Laplacian and divergence stencils of
compute_norm_interface / compute_div_theta_n (lb_cuda_auxfields_module.f90:
72-130, 1110-1160) from per-plane partial sums (part3 / grad_lap in
lbfast_amd/csrc/lbf_common.h; the divergence in prep_body_z of
lbf_stages.hip).  This restates that arithmetic in numpy, in the kernels'
order, and checks it against the direct weighted sums over the 26 D3Q27
neighbours, with the weights taken from the lattice tables that
test_oracle_tables.py pins to the reference's literal tables:

  grad_i phi = 3 sum_l w_l c_li phi(x + c_l)
  lap phi    = 6 sum_l w_l (phi(x + c_l) - phi(x))
  div theta  = 3 sum_l w_l c_l . theta(x + c_l)

In float64 the two forms agree to rounding; a wrong smoothing weight or a
swapped difference direction fails.
"""
import numpy as np
import pytest

from lbfast_amd.lattice import EX, EY, EZ, weights

W27 = weights(27)


def shift(f, a, b, c):
    """f(x + a, y + b, z + c) on a periodic (z, y, x) array."""
    return np.roll(f, (-c, -b, -a), axis=(0, 1, 2))


def direct(phi, th):
    gx, gy, gz, lap, div = (np.zeros_like(phi) for _ in range(5))
    for l in range(27):
        a, b, c, w = int(EX[l]), int(EY[l]), int(EZ[l]), W27[l]
        s = shift(phi, a, b, c)
        gx += 3 * w * a * s
        gy += 3 * w * b * s
        gz += 3 * w * c * s
        lap += 6 * w * (s - phi)
        div += 3 * w * (a * shift(th[0], a, b, c) + b * shift(th[1], a, b, c) + c * shift(th[2], a, b, c))
    return gx, gy, gz, lap, div


def part3(f, z_off, u=(1.0, 4.0, 1.0)):
    """The partial sums of plane z + z_off at every node: v[b][a] is
    f(x + a - 1, y + b - 1, z + z_off); the 2-D smoothing s is taken of the
    differences to the plane's centre value c, which is returned beside it."""
    v = [[shift(f, aa - 1, bb - 1, z_off) for aa in range(3)] for bb in range(3)]
    u0, u1, u2 = u
    dx = u0 * (v[0][2] - v[0][0]) + u1 * (v[1][2] - v[1][0]) + u2 * (v[2][2] - v[2][0])
    dy = u0 * (v[2][0] - v[0][0]) + u1 * (v[2][1] - v[0][1]) + u2 * (v[2][2] - v[0][2])
    c = v[1][1]
    r0 = u0 * (v[0][0] - c) + u1 * (v[0][1] - c) + u2 * (v[0][2] - c)
    r1 = u0 * (v[1][0] - c) + u2 * (v[1][2] - c)
    r2 = u0 * (v[2][0] - c) + u1 * (v[2][1] - c) + u2 * (v[2][2] - c)
    s = u0 * r0 + u1 * r1 + u2 * r2
    return dx, dy, s, c


def separable(phi, th, u=(1.0, 4.0, 1.0)):
    m, o, p = (part3(phi, dz, u) for dz in (-1, 0, 1))
    dt = phi.dtype.type
    four, half, g3, l6 = dt(u[1]), dt(0.5), dt(3.0 / 216.0), dt(6.0 / 216.0)
    gx = g3 * (m[0] + four * o[0] + p[0])
    gy = g3 * (m[1] + four * o[1] + p[1])
    gz = g3 * (p[2] - m[2]) + half * (p[3] - m[3])
    lap = l6 * (m[2] + four * o[2] + p[2]) + ((m[3] - o[3]) + (p[3] - o[3]))
    # divergence: x difference of theta_x, y difference of theta_y, 2-D
    # smoothing of theta_z, per plane; planes k-1, k, k+1 combined as above
    ex = [part3(th[0], dz, u)[0] for dz in (-1, 0, 1)]
    ey = [part3(th[1], dz, u)[1] for dz in (-1, 0, 1)]
    sz = [part3(th[2], dz, u) for dz in (-1, 0, 1)]
    div = g3 * ((ex[0] + four * ex[1] + ex[2]) + (ey[0] + four * ey[1] + ey[2]) + (sz[2][2] - sz[0][2])) \
        + half * (sz[2][3] - sz[0][3])
    return gx, gy, gz, lap, div


def reference_order(phi, th):
    """The reference's own expression order (lb_cuda_auxfields_module.f90:
    72-130, 1128-1160): paired differences for the gradient and divergence,
    and 6 (c (p0 - 1) + p1 S1 + p2 S2 + p3 S3) for the Laplacian, evaluated
    in the arrays' precision."""
    dt = phi.dtype.type
    p0, p1, p2, p3 = dt(8 / 27), dt(2 / 27), dt(1 / 54), dt(1 / 216)
    F = lambda a, b, c: shift(phi, a, b, c)  # noqa: E731

    def paired(G, a):
        # sum_l w_l c_la G(x + c_l) / (p1, p2, p3 grouped), paired l / opposite
        axis = [(1, 0, 0), (0, 1, 0), (0, 0, 1)][a]
        acc1 = shift(G, *axis) - shift(G, *(-e for e in axis))
        acc2 = np.zeros_like(G)
        acc3 = np.zeros_like(G)
        for l in range(1, 27):
            e = (int(EX[l]), int(EY[l]), int(EZ[l]))
            if e[a] != 1:
                continue
            nz = sum(1 for q in e if q)
            d = shift(G, *e) - shift(G, *(-q for q in e))
            if nz == 2:
                acc2 = acc2 + d
            elif nz == 3:
                acc3 = acc3 + d
        return p1 * acc1 + p2 * acc2 + p3 * acc3
    gx, gy, gz = (dt(3) * paired(phi, a) for a in range(3))
    s1 = F(1, 0, 0) + F(-1, 0, 0) + F(0, 1, 0) + F(0, -1, 0) + F(0, 0, 1) + F(0, 0, -1)
    s2 = np.zeros_like(phi)
    s3 = np.zeros_like(phi)
    for l in range(1, 27):
        e = (int(EX[l]), int(EY[l]), int(EZ[l]))
        nz = sum(1 for q in e if q)
        if nz == 2:
            s2 = s2 + F(*e)
        elif nz == 3:
            s3 = s3 + F(*e)
    lap = dt(6) * (phi * (p0 - dt(1)) + (p1 * s1 + p2 * s2 + p3 * s3))
    div = (paired(th[0], 0) + paired(th[1], 1) + paired(th[2], 2)) * dt(3)
    return gx, gy, gz, lap, div


@pytest.fixture(scope="module")
def fields():
    rng = np.random.default_rng(7)
    phi = rng.random((7, 6, 5))
    th = rng.standard_normal((3, 7, 6, 5))
    return phi, th


def test_separable_equals_direct(fields):
    phi, th = fields
    for name, d, s in zip(("gx", "gy", "gz", "lap", "div"), direct(phi, th), separable(phi, th)):
        assert np.abs(d - s).max() <= 1e-13 * max(1.0, np.abs(d).max()), name


def test_weights_factor_as_outer_product():
    # w(a,b,c) = u_a u_b u_c / 216, u = (1, 4, 1): what the separable form rests on
    for l in range(27):
        u = [(1.0, 4.0, 1.0)[int(e) + 1] for e in (EX[l], EY[l], EZ[l])]
        assert abs(W27[l] - u[0] * u[1] * u[2] / 216.0) < 1e-16


def test_wrong_smoothing_weight_is_detected(fields):
    phi, th = fields
    bad = separable(phi, th, u=(1.0, 2.0, 1.0))
    errs = [np.abs(d - s).max() for d, s in zip(direct(phi, th), bad)]
    assert min(errs) > 1e-3


def test_swapped_difference_is_detected(fields):
    phi, th = fields
    good = direct(phi, th)
    # theta_x and theta_y exchanged: the divergence must change
    _, _, _, _, div_sw = separable(phi, np.stack([th[1], th[0], th[2]]))
    assert np.abs(div_sw - good[4]).max() > 1e-2


def test_float32_rounding_matches_direct_order(fields):
    # the kernels run in float32 with integer weights and one final scale:
    # the separable result stays within a few float32 ulps of the direct one
    phi, th = fields
    d = direct(phi, th)
    s32 = separable(phi.astype(np.float32), th.astype(np.float32))
    for name, a, b in zip(("gx", "gy", "gz", "lap", "div"), d, s32):
        assert np.abs(a - b.astype(np.float64)).max() < 2e-6 * max(1.0, np.abs(a).max()), name


def test_float32_at_least_as_accurate_as_reference_order():
    # On a smooth tanh interface (a droplet, R = 10, w = 3, phi up to 1), the
    # float32 separable form is at least as accurate as the reference's own
    # expression order in float32, against float64: the smoothing sums are
    # taken of differences, so nothing of size 36 or 216 cancels.  Summing the
    # raw values instead (the earlier form) is 4.6x less accurate in the z
    # gradient and 1.4x in the Laplacian, and fails here.
    n = 32
    g = np.arange(n) + 0.5 - n / 2
    z, y, x = np.meshgrid(g, g, g, indexing="ij")
    r = np.sqrt(x * x + y * y + z * z)
    phi = 0.5 * (1 + np.tanh(2 * (10.0 - r) / 3.0))
    gg = np.stack(direct(phi, np.zeros((3,) + phi.shape))[:3])
    th = phi * (1 - phi) * gg / (np.sqrt((gg * gg).sum(0)) + 1e-9)
    truth = direct(phi, th)
    p32, t32 = phi.astype(np.float32), th.astype(np.float32)
    sep = separable(p32, t32)
    ref = reference_order(p32, t32)
    band = (phi > 1e-3) & (phi < 1 - 1e-3)
    for name, t, s_, r_ in zip(("gx", "gy", "gz", "lap", "div"), truth, sep, ref):
        es = np.sqrt(np.mean((s_.astype(np.float64) - t)[band] ** 2))
        er = np.sqrt(np.mean((r_.astype(np.float64) - t)[band] ** 2))
        assert es <= 1.25 * er, (name, es, er)

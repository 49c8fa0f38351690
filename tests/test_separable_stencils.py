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
    """The three partial sums of plane z + z_off at every node: v[b][a] is
    f(x + a - 1, y + b - 1, z + z_off)."""
    v = [[shift(f, aa - 1, bb - 1, z_off) for aa in range(3)] for bb in range(3)]
    u0, u1, u2 = u
    dx = u0 * (v[0][2] - v[0][0]) + u1 * (v[1][2] - v[1][0]) + u2 * (v[2][2] - v[2][0])
    dy = u0 * (v[2][0] - v[0][0]) + u1 * (v[2][1] - v[0][1]) + u2 * (v[2][2] - v[0][2])
    r = [u0 * v[bb][0] + u1 * v[bb][1] + u2 * v[bb][2] for bb in range(3)]
    s = u0 * r[0] + u1 * r[1] + u2 * r[2]
    return dx, dy, s


def separable(phi, th, u=(1.0, 4.0, 1.0)):
    m, o, p = (part3(phi, dz, u) for dz in (-1, 0, 1))
    four, g3, l6 = u[1], 3.0 / 216.0, 6.0 / 216.0
    gx = g3 * (m[0] + four * o[0] + p[0])
    gy = g3 * (m[1] + four * o[1] + p[1])
    gz = g3 * (p[2] - m[2])
    lap = l6 * (m[2] + four * o[2] + p[2]) - 6.0 * phi
    # divergence: x difference of theta_x, y difference of theta_y, 2-D
    # smoothing of theta_z, per plane; planes k-1, k, k+1 combined as above
    ex = [part3(th[0], dz, u)[0] for dz in (-1, 0, 1)]
    ey = [part3(th[1], dz, u)[1] for dz in (-1, 0, 1)]
    sz = [part3(th[2], dz, u)[2] for dz in (-1, 0, 1)]
    div = g3 * ((ex[0] + four * ex[1] + ex[2]) + (ey[0] + four * ey[1] + ey[2]) + (sz[2] - sz[0]))
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

"""Long-horizon GPU checks.

* The stated FP32 bar: after 1000 steps rho, u and phi within 1e-5 relative
  L2 of the oracle run at the same precision (Taylor-Green 64^3 D3Q19 --
  the reference's CPU-runnable case); FP64 within 1e-11.  For a
  two-component D3Q27 HIGHORDER droplet at density ratio 100: rho_phi and
  phi within 1e-5, and u and P as accurate as the FP32 oracle against FP64.
* The reference's physics acceptance checks (statistics.f90), restated on
  the GPU results:
  - Taylor-Green: E(t)/E0 = exp(-6 nu k^2 t), k = 2 pi / L (statistics.f90:927-934);
    the fitted nu agrees with (tau - 1/2)/3.
  - Laplace: sigma_eff = (p_in - p_out) R / 2 with p = rho_phi cs^2 P and R
    from the phi(1-phi)-weighted interface positions along x through the
    centre (statistics.f90:160-268); sigma_eff agrees with sigma for the
    reference's density-matched deck and is steady at density ratio 100.
* Open boundaries to steady state: a pressure-driven duct (pressure faces
  in x, walls as velocity faces at rest) run 800 steps on the GPU equals the
  FP64 oracle's run (FP64) and, in FP32, is the body-force duct of the
  oracle's closed-form pin (developed profile within 5e-4 of the peak).
"""
import numpy as np
import pytest

import oracle
from lbfast_amd import hip, setups
from lbfast_amd.params import make_params, PREC_FP64, PREC_FP32, HFIELDS, FIELD_PHI

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32])
def test_taylor_green_1000_steps(prec):
    n = (64, 64, 64)
    tau = 0.8
    p = make_params(n, lattice=19, highorder=False, visc1=(tau - 0.5) / 3, precision=prec)
    f = setups.taylor_green(n, 0.02)
    print(f"[long] TG 64^3 D3Q19 prec={prec} 1000 steps", flush=True)
    o = oracle.Oracle(p)
    o.load(f)
    o.step(1000)
    d = hip.Domain(p)
    d.load(f)
    e0 = None
    ts, es = [], []
    for s in range(0, 1001, 100):
        if s:
            d.step(100)
        u = np.stack([d.download(i) for i in (2, 3, 4)])
        e = 0.5 * np.mean((u ** 2).sum(0))
        e0 = e if e0 is None else e0
        ts.append(s)
        es.append(e / e0)
    got_rho = 1.0 + d.download(1)
    got_u = np.stack([d.download(i) for i in (2, 3, 4)])
    d.close()
    want_rho = 1.0 + o.get_field(1)
    want_u = np.stack([o.get_field(i) for i in (2, 3, 4)])
    tol = 1e-11 if prec == PREC_FP64 else 1e-5
    er, eu = _rel(got_rho, want_rho), _rel(got_u, want_u)
    print(f"[long]   rel-L2 rho {er:.2e} u {eu:.2e}", flush=True)
    assert er < tol and eu < tol
    # kinetic-energy decay rate fit over log E >= -0.25 .. all points here
    k2 = 3 * (2 * np.pi / 64) ** 2  # |k|^2 of the TG mode: E ~ exp(-2 nu |k|^2 t)
    slope = np.polyfit(ts, np.log(es), 1)[0]
    nu_num = -slope / (2 * k2)
    nu = (tau - 0.5) / 3
    print(f"[long]   nu fit {nu_num:.5f} vs {nu:.5f}", flush=True)
    assert abs(nu_num / nu - 1) < 0.01


def test_fp32_mass_drift_no_larger_than_reference_weights():
    """One component, periodic, no force: sum P is conserved exactly in exact
    arithmetic.  In FP32 the GPU's drift over 1000 Taylor-Green steps (D3Q27
    HIGHORDER, 64^3) is at most twice the FP32 oracle's, whose drift comes
    from the rounded weights of the reference's table.  A reconstruction
    whose one-axis weights fl(1/6) + fl(2/3) + fl(1/6) exceed 1 by 3e-8
    drifts about ten times more and fails."""
    n = (64, 64, 64)
    p = make_params(n, lattice=27, highorder=True, visc1=0.05 / 3, precision=PREC_FP32)
    f = setups.taylor_green(n, 0.04)
    # a uniform offset makes the density deviation's sum large against its drift
    f = dict(f)
    f["P"] = f["P"] + 1e-2
    steps = 1000
    print(f"[long] mass drift TG 64^3 D3Q27h FP32 {steps} steps", flush=True)
    o = oracle.Oracle(p)
    o.load(f)
    s0 = float(np.sum(o.get_field(1)))
    o.step(steps)
    drift_o = abs(float(np.sum(o.get_field(1))) - s0) / np.prod(n)
    d = hip.Domain(p)
    d.load(f)
    g0 = d.reduce()[0]
    d.step(steps)
    drift_g = abs(d.reduce()[0] - g0) / np.prod(n)
    d.close()
    print(f"[long]   mean P drift: GPU {drift_g:.3e}, FP32 oracle {drift_o:.3e}", flush=True)
    assert drift_g <= 2.0 * drift_o + 1e-12, (drift_g, drift_o)


def test_droplet_1000_steps_fp32():
    """The stated FP32 bar on a droplet like the headline workload's (64^3,
    R = 16, w = 3, density ratio 100, sigma = 0.01, D3Q27 HIGHORDER
    DENSRATIO) after 1000 steps: rho_phi and phi within 1e-5 relative L2 of
    the FP32 oracle.  u and P are spurious currents and a pressure deviation
    near rest (|u| ~ 1e-5, P a difference of O(1) populations), where two
    correct FP32 evaluation orders differ by more than 1e-5; for them the
    criterion is accuracy: the GPU's error against the FP64 oracle is at most
    1.5 times the FP32 oracle's own error against the FP64 oracle."""
    n = (64, 64, 64)
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.1, visc2=0.1,
                    rho_r=1.0, rho_b=0.01, tau_diff=0.1, sigma=0.01, width=3.0, precision=PREC_FP32)
    f = setups.droplet(n, 16.0, 3.0, [32.5, 32.5, 32.5], 0.01, 1.0, 0.01, noise=1e-6, seed=20261202)
    steps = 1000
    print(f"[long] droplet 64^3 R=16 D3Q27h densratio 100 FP32 {steps} steps", flush=True)
    p64 = p.copy()
    p64.precision = PREC_FP64
    runs = {}
    for name, q in (("o32", p), ("o64", p64)):
        o = oracle.Oracle(q)
        o.load(f)
        o.step(steps)
        runs[name] = (o.get_field(FIELD_PHI), np.stack([o.get_field(i) for i in (2, 3, 4)]), o.get_field(1))
        del o
        print(f"[long]   oracle {name} done", flush=True)
    d = hip.Domain(p)
    d.load(f)
    d.step(steps)
    runs["gpu"] = (d.download(FIELD_PHI), np.stack([d.download(i) for i in (2, 3, 4)]), d.download(1))
    d.close()

    def rho(phi):
        return p.rho_r * phi + (1 - phi) * p.rho_b
    g, o32, o64 = runs["gpu"], runs["o32"], runs["o64"]
    res = {"rho": _rel(rho(g[0]), rho(o32[0])), "phi": _rel(g[0], o32[0]), "u": _rel(g[1], o32[1]),
           "P": _rel(g[2], o32[2])}
    acc = {"u": (_rel(g[1], o64[1]), _rel(o32[1], o64[1])), "P": (_rel(g[2], o64[2]), _rel(o32[2], o64[2])),
           "phi": (_rel(g[0], o64[0]), _rel(o32[0], o64[0]))}
    print(f"[long]   vs FP32 oracle {res}", flush=True)
    print(f"[long]   (GPU, FP32 oracle) vs FP64 oracle {acc}", flush=True)
    assert res["rho"] < 1e-5 and res["phi"] < 1e-5
    for k, (eg, eo) in acc.items():
        assert eg <= 1.5 * eo + 1e-9, (k, eg, eo)


def test_laplace_density_ratio_100_against_fp64_oracle():
    """The Laplace diagnostic at density ratio 100 comes from the model, not
    from the kernels: a 64^3 droplet (R = 16, w = 3, sigma = 0.01, rho_r/rho_b
    = 100) run 2000 steps in FP64 on the GPU and in the FP64 oracle gives the
    same fields to 1e-9 and the same sigma_eff/sigma to 1e-7."""
    n, R, steps = 64, 16.0, 2000
    shape = (n, n, n)
    center = [n / 2 + 0.5] * 3
    p = make_params(shape, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.1, visc2=0.1,
                    rho_r=1.0, rho_b=0.01, tau_diff=0.1, sigma=0.01, width=3.0, precision=PREC_FP64)
    f = setups.droplet(shape, R, 3.0, center, 0.01, 1.0, 0.01, noise=1e-6, seed=7)
    print(f"[long] laplace 64^3 R=16 density ratio 100 FP64, GPU vs oracle, {steps} steps", flush=True)
    o = oracle.Oracle(p)
    o.load(f)
    o.step(steps)
    d = hip.Domain(p)
    d.load(f)
    d.step(steps)
    s_gpu, r_gpu = _laplace_sigma(d, p, shape, center)
    s_orc, r_orc = _laplace_sigma(o, p, shape, center)
    e_phi = _rel(d.download(FIELD_PHI), o.get_field(FIELD_PHI))
    e_P = _rel(d.download(1), o.get_field(1))
    e_u = _rel(np.stack([d.download(i) for i in (2, 3, 4)]), np.stack([o.get_field(i) for i in (2, 3, 4)]))
    d.close()
    print(f"[long]   sigma_eff/sigma GPU {s_gpu / 0.01:.6f} oracle {s_orc / 0.01:.6f} R {r_gpu:.4f}; "
          f"rel-L2 phi {e_phi:.2e} P {e_P:.2e} u {e_u:.2e}", flush=True)
    assert e_phi < 1e-9 and e_P < 1e-9 and e_u < 1e-9
    assert abs(s_gpu / s_orc - 1) < 1e-7 and abs(r_gpu - r_orc) < 1e-7


def _laplace_sigma(d, p, n, center):
    get = d.download if hasattr(d, "download") else d.get_field
    phi = get(FIELD_PHI)
    P = get(1)
    nx = n[0]
    j, k = n[1] // 2 - 1, n[2] // 2 - 1  # gj = ly/2, gk = lz/2 (1-based)
    line = phi[k, j, :]
    gi = np.arange(1, nx + 1, dtype=float)
    wgt = line * (1 - line)
    left = (gi[: nx // 2] * wgt[: nx // 2]).sum() / wgt[: nx // 2].sum()
    right = (gi[nx // 2:] * wgt[nx // 2:]).sum() / wgt[nx // 2:].sum()
    rad = right - 0.5 * (right + left)
    ci = [int(round(c)) - 1 for c in center]
    rho_in = p.rho_r * phi[ci[2], ci[1], ci[0]] + (1 - phi[ci[2], ci[1], ci[0]]) * p.rho_b
    rho_out = p.rho_r * phi[0, 0, 0] + (1 - phi[0, 0, 0]) * p.rho_b
    p_in = rho_in * P[ci[2], ci[1], ci[0]] / 3.0
    p_out = rho_out * P[0, 0, 0] / 3.0
    return (p_in - p_out) * rad * 0.5, rad


def _laplace_run(n, R, w, rho_b, steps, sigma=0.01, chunks=4):
    center = [n / 2 + 0.5] * 3
    shape = (n, n, n)
    p = make_params(shape, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.1, visc2=0.1,
                    rho_r=1.0, rho_b=rho_b, tau_diff=0.1, sigma=sigma, width=w, precision=PREC_FP32)
    f = setups.droplet(shape, R, w, center, sigma, 1.0, rho_b, noise=1e-6, seed=7)
    d = hip.Domain(p)
    d.load(f)
    vals = []
    for s in range(chunks):
        d.step(steps // chunks)
        se, rad = _laplace_sigma(d, p, shape, center)
        vals.append(se / sigma)
        print(f"[long] laplace n={n} R={R} w={w} rho_b={rho_b} step {(s + 1) * steps // chunks} "
              f"sigma_eff/sigma {se / sigma:.4f} R_meas {rad:.3f}", flush=True)
    d.close()
    return np.array(vals)


def test_laplace_density_matched():
    """The reference's Laplace deck (laplace.inp: rho_r = rho_b = 1, width 4):
    sigma_eff within 10% of sigma (measured 0.94 at R = 16)."""
    v = _laplace_run(64, 16.0, 4.0, 1.0, 3000)
    assert np.all(np.isfinite(v))
    assert abs(v[-2:].mean() - 1) < 0.1


def test_laplace_density_ratio_100():
    """Config 3 (256^3, R = 40, w = 3, density ratio 100, 4000 steps).  With the
    reference's pressure diagnostic p = rho_phi cs^2 P the pressure jump settles
    about 20% above 2 sigma / R (measured 1.23; the FP64 oracle gives the
    GPU's value, test_laplace_density_ratio_100_against_fp64_oracle, so it is
    the model's); checked steady and inside [0.8, 1.35]."""
    v = _laplace_run(256, 40.0, 3.0, 0.01, 4000)
    assert np.all(np.isfinite(v))
    assert abs(v[-1] - v[-2]) < 0.03
    assert 0.8 < v[-1] < 1.35


_MOM = ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")


def _open_duct(prec, n=(24, 8, 8), dP=1e-4):
    p = make_params(n, lattice=27, highorder=True, internal_obstacles=True, pbc=(0, 0, 0), visc1=1.0 / 6.0,
                    precision=prec)
    p.openbc = 1
    p.openbc_type[0], p.openbc_press[0] = 1, dP
    p.openbc_type[1], p.openbc_press[1] = 1, 0.0
    isf = hip.classify_isfluid(p)
    shape = (n[2], n[1], n[0])
    return p, {k: np.zeros(shape) for k in _MOM}, isf


@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32])
def test_open_duct_steady_state(prec):
    n, steps = (24, 8, 8), 800
    print(f"[long] open duct {n} prec={prec} {steps} steps", flush=True)
    p, f, isf = _open_duct(prec, n)
    d = hip.Domain(p)
    d.load(f, isf)
    d.step(steps)
    got = np.stack([d.download(1 + k) for k in range(10)])
    d.step(20)
    later = np.stack([d.download(1 + k) for k in range(10)])
    d.close()
    p64, f, isf = _open_duct(PREC_FP64, n)
    fl = isf != 0
    assert np.abs(later - got)[:, fl].max() < (1e-9 if prec == PREC_FP64 else 1e-5) * np.abs(got[:, fl]).max()  # steady
    o = oracle.Oracle(p64)
    o.load(f, isf)
    o.step(steps)
    ref = np.stack([o.get_field(1 + k) for k in range(10)])
    fluid = isf != 0  # solid nodes' outputs are never read (as in the parity tests)
    if prec == PREC_FP64:
        assert _rel(got[:, fluid], ref[:, fluid]) < 1e-11
    # the developed profile is the body-force duct at the duct's own gradient
    x = np.arange(1, n[0] + 1, dtype=float)
    mid = slice(6, n[0] - 6)
    c = np.polyfit(x[mid], got[0][4, 4, mid], 1)
    pb = make_params((3, n[1], n[2]), lattice=27, highorder=True, internal_obstacles=True, pbc=(1, 0, 0),
                     visc1=1.0 / 6.0, fx=-c[0] / 3.0, precision=PREC_FP64)
    pb.openbc = 1
    ob = oracle.Oracle(pb)
    shape = (n[2], n[1], 3)
    ob.load({k: np.zeros(shape) for k in _MOM}, oracle.classify_isfluid(pb, None))
    ob.step(steps)
    ub = ob.get_field(2)[1:-1, 1:-1, 1]  # the fluid cross-section (walls at y, z = 0 and n - 1)
    u = got[1][1:-1, 1:-1, n[0] // 2]
    assert np.abs(u - ub).max() < 5e-4 * ub.max(), np.abs(u - ub).max() / ub.max()

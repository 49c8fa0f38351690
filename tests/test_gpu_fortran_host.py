"""The Fortran host on the GPU.

* The host drives the same C-ABI as the Python bindings: a host run and a
  hip.Domain run on the same inputs (lreadinit RAW files, or the analytic
  Poiseuille start) give the same RAW frames bit for bit.
* Its own initial conditions and diagnostics follow the physics: the
  Taylor-Green kinetic energy decays as exp(-6 nu k^2 t) (statistics.f90:63-86),
  a droplet obeys the Laplace law, pressure jump times radius independent of
  the radius (statistics.f90:184-275), the bounce-back
  channel keeps the Poiseuille parabola (init_conditions_module.f90:240-252).
* A restart (output/restart.raw) continues a one-component FP32 run exactly.
"""
import os

import numpy as np
import pytest

from lbfast_amd import hip, setups
from lbfast_amd.params import make_params, PREC_FP64, PREC_FP32, PREC_FP16S, FIELD_PHI
from tests import fhost

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(fhost.exe("tg27h_fp32")), reason="Fortran host not built")]

LAT = {"tg27h_fp32": (27, True, PREC_FP32), "tg27h_fp64": (27, True, PREC_FP64), "tg19_fp32": (19, False, PREC_FP32),
       "laplace27h_fp32": (27, True, PREC_FP32), "laplace27h_fp64": (27, True, PREC_FP64),
       "laplace27h_fp16": (27, True, PREC_FP16S), "pois19_bb_fp64": (19, False, PREC_FP64)}


def _run(variant, wd, sim, fl, **kw):
    print(f"[fortran host] {variant} n=({sim['lx']},{sim['ly']},{sim['lz']}) nsteps={sim['nsteps']} in {wd}",
          flush=True)
    r = fhost.run(variant, wd, fhost.deck(sim, fl), **kw)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "lbfast_host: done" in r.stdout
    return r.stdout


def _sim(n, nsteps, **kw):
    s = {"nsteps": nsteps, "stamp": nsteps, "stamp_term": nsteps, "lx": n[0], "ly": n[1], "lz": n[2],
         "lprint": True, "lraw": True, "pbc_x": 1, "pbc_y": 1, "pbc_z": 1}
    s.update(kw)
    return s


def _f32(a):
    return np.asarray(a, dtype=np.float32)


@pytest.mark.parametrize("variant", ["tg27h_fp32", "tg27h_fp64", "tg19_fp32"])
def test_readinit_one_component_equals_domain(tmp_path, variant):
    lat, ho, prec = LAT[variant]
    n, steps = (24, 20, 16), 10
    rng = np.random.default_rng(17)
    shape = (n[2], n[1], n[0])
    tg = setups.taylor_green(n, 0.03)
    init = {"P": _f32(tg["P"] + rng.uniform(-1e-4, 1e-4, shape))}
    for c in ("u", "v", "w"):
        init[c] = _f32(tg[c] + rng.uniform(-2e-3, 2e-3, shape))
    for nm, f in (("rho.raw", "P"), ("u.raw", "u"), ("v.raw", "v"), ("w.raw", "w")):
        fhost.write_raw(tmp_path, nm, init[f])
    # lreadinit replaces P, u, v, w; the second moments stay those of the
    # generated start, which is at rest for uwall = 0
    out = _run(variant, tmp_path, _sim(n, steps, lreadinit=True), {"visc1": 0.04, "fx": 1e-6, "uwall": 0.0})
    assert "TAYLORGREEN step" in out
    p = make_params(n, lattice=lat, highorder=ho, visc1=0.04, fx=1e-6, precision=prec)
    f = {k: np.zeros(shape) for k in ("pxx", "pyy", "pzz", "pxy", "pxz", "pyz")}
    f.update({k: v.astype(np.float64) for k, v in init.items()})
    d = hip.Domain(p)
    d.load(f)
    d.step(steps)
    want_rho = _f32(d.download(1))
    want_vel = np.stack([_f32(d.download(i)) for i in (2, 3, 4)], axis=-1)
    d.close()
    assert np.array_equal(fhost.read_frame(tmp_path, "rho", 1, n), want_rho)
    assert np.array_equal(fhost.read_frame(tmp_path, "vel", 1, n, 3), want_vel)
    assert np.array_equal(fhost.read_frame(tmp_path, "rho", 0, n), init["P"])


def test_readinit_two_component_equals_domain(tmp_path):
    variant = "laplace27h_fp32"
    n, steps = (32, 24, 20), 12
    fl = {"visc1": 0.05, "visc2": 0.08, "sigma": 0.01, "width": 4.0, "tau_diff": 0.1, "rho_r": 1.0, "rho_b": 0.1,
          "radius": 6.0, "center": [16.0, 12.0, 10.0]}
    drop = setups.droplet(n, 6.0, 4.0, [16.0, 12.0, 10.0], 0.01, 1.0, 0.1, noise=1e-4, seed=4)
    init = {k: _f32(drop[k]) for k in ("P", "u", "v", "w", "phi")}
    for nm, f in (("rho.raw", "P"), ("u.raw", "u"), ("v.raw", "v"), ("w.raw", "w"), ("phi.raw", "phi")):
        fhost.write_raw(tmp_path, nm, init[f])
    _run(variant, tmp_path, _sim(n, steps, lreadinit=True, iprobe=1, jprobe=1, kprobe=1), fl)
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.05, visc2=0.08,
                    sigma=0.01, width=4.0, tau_diff=0.1, rho_r=1.0, rho_b=0.1, precision=PREC_FP32)
    shape = (n[2], n[1], n[0])
    f = {k: np.zeros(shape) for k in ("pxx", "pyy", "pzz", "pxy", "pxz", "pyz")}  # the generated start is at rest
    f.update({k: v.astype(np.float64) for k, v in init.items()})
    d = hip.Domain(p)
    d.load(f)
    d.step(steps)
    phi = d.download(FIELD_PHI)
    want_rho = _f32(1.0 * phi + (1.0 - phi) * 0.1)  # rho_r phi + (1 - phi) rho_b, as the frame writer
    want_vel = np.stack([_f32(d.download(i)) for i in (2, 3, 4)], axis=-1)
    d.close()
    assert np.array_equal(fhost.read_frame(tmp_path, "rho", 1, n), want_rho)
    assert np.array_equal(fhost.read_frame(tmp_path, "vel", 1, n, 3), want_vel)


@pytest.mark.parametrize("variant", ["tg27h_fp64", "tg27h_fp32", "tg19_fp32"])
def test_taylor_green_decay(tmp_path, variant):
    lat, ho, prec = LAT[variant]
    n, nu, u0, steps = (32, 32, 32), 0.05, 0.02, 200
    out = _run(variant, tmp_path, _sim(n, steps, stamp_term=50, lprint=False), {"visc1": nu, "uwall": u0})
    rows = fhost.stats(out, "TAYLORGREEN")
    assert [r["step"] for r in rows] == [0, 50, 100, 150, 200]
    last = rows[-1]
    # the lattice solution decays at the viscosity of the collision, up to
    # O(k^2) lattice corrections: 2% of the decay exponent
    assert abs(np.log(last["E/E0"]) / np.log(last["theory"]) - 1.0) < 0.02, last
    # the host's own start (init_conditions_module.f90:298-314) equals the
    # Python one up to the last bits of sin/cos
    p = make_params(n, lattice=lat, highorder=ho, visc1=nu, precision=prec)
    d = hip.Domain(p)
    d.load(setups.taylor_green(n, u0))
    e0 = sum(float((d.download(i) ** 2).sum()) for i in (2, 3, 4))
    d.step(steps)
    e1 = sum(float((d.download(i) ** 2).sum()) for i in (2, 3, 4))
    d.close()
    assert abs(last["E/E0"] - e1 / e0) < 1e-6 * (e1 / e0)


def _laplace(tmp_path, variant, n=(40, 40, 40), steps=1000, R=10.0, rho_b=0.1):
    sigma = 0.01
    fl = {"visc1": 0.1, "visc2": 0.1, "sigma": sigma, "width": 4.0, "tau_diff": 0.1, "rho_r": 1.0, "rho_b": rho_b,
          "radius": R, "center": [n[0] / 2.0, n[1] / 2.0, n[2] / 2.0]}
    sim = _sim(n, steps, stamp_term=max(steps // 4, 2), iprobe=1, jprobe=1, kprobe=1)
    out = _run(variant, tmp_path, sim, fl)
    return fhost.stats(out, "LAPLACE"), fl


def test_laplace_law(tmp_path):
    # equal densities, as the reference's Laplace deck: the pressure jump
    # times the measured radius over two is the same for two radii (Laplace
    # law) and close to the surface tension of the deck
    eff = []
    for R in (8.0, 12.0):
        wd = tmp_path / f"r{int(R)}"
        wd.mkdir()
        rows, fl = _laplace(wd, "laplace27h_fp64", steps=600, R=R, rho_b=1.0)
        last = rows[-1]
        print(f"[fortran host]   R={R}: {last}", flush=True)
        assert abs(last["radius"] - R) < 0.5
        eff.append(last["sigma_eff"])
    assert abs(eff[0] / eff[1] - 1.0) < 0.05
    assert all(abs(e / fl["sigma"] - 1.0) < 0.1 for e in eff), eff


def test_laplace_precisions(tmp_path):
    # density ratio 10: the FP32 and 16-bit storage builds against FP64
    res = {}
    for variant in ("laplace27h_fp64", "laplace27h_fp32", "laplace27h_fp16"):
        wd = tmp_path / variant
        wd.mkdir()
        rows, fl = _laplace(wd, variant)
        res[variant] = rows[-1]
        print(f"[fortran host]   {variant}: {rows[-1]}", flush=True)
    ref = res["laplace27h_fp64"]
    assert abs(ref["radius"] - fl["radius"]) < 0.5
    assert abs(res["laplace27h_fp32"]["sigma_eff"] / ref["sigma_eff"] - 1.0) < 1e-4
    assert abs(res["laplace27h_fp16"]["sigma_eff"] / ref["sigma_eff"] - 1.0) < 5e-3


def test_laplace_start_equals_python_droplet(tmp_path):
    # The host's droplet (init_conditions_module.f90:394-426) equals the
    # Python one up to the last bits of tanh.  The comparison stops after two
    # steps: where |grad phi| vanishes (the droplet axis midway between
    # periodic images) the normal g / (|g| + 1e-9) amplifies last-bit
    # differences of the start by up to 1e9 per step, so longer runs are
    # compared from bitwise-identical starts (the lreadinit tests).
    n = (32, 32, 32)
    rows, fl = _laplace(tmp_path, "laplace27h_fp64", n=n, steps=2)
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.1, visc2=0.1,
                    sigma=0.01, width=4.0, tau_diff=0.1, rho_r=1.0, rho_b=0.1, precision=PREC_FP64)
    d = hip.Domain(p)
    d.load(setups.droplet(n, fl["radius"], 4.0, fl["center"], 0.01, 1.0, 0.1))
    want0 = d.download(FIELD_PHI)
    d.step(2)
    phi = d.download(FIELD_PHI)
    vel = np.stack([d.download(i) for i in (2, 3, 4)], axis=-1)
    d.close()
    got0 = fhost.read_frame(tmp_path, "rho", 0, n).astype(np.float64)
    got1 = fhost.read_frame(tmp_path, "rho", 1, n).astype(np.float64)
    # frames hold rho_r phi + (1 - phi) rho_b in real(4)
    assert np.abs(got0 - (want0 + (1 - want0) * 0.1)).max() < 1e-6
    assert np.abs(got1 - (phi + (1 - phi) * 0.1)).max() < 1e-6
    gv = fhost.read_frame(tmp_path, "vel", 1, n, 3).astype(np.float64)
    assert np.abs(gv - vel).max() < 1e-7 * np.abs(vel).max()
    assert [r["step"] for r in rows] == [0, 2]


def test_restart_continues_exactly(tmp_path):
    variant = "tg27h_fp32"
    n = (24, 24, 16)
    fl = {"visc1": 0.03, "uwall": 0.03}
    a, b = tmp_path / "split", tmp_path / "whole"
    a.mkdir()
    b.mkdir()
    _run(variant, a, _sim(n, 20, stamp=10, lwriterestart=True), fl)
    assert os.path.exists(a / "output" / "restart.raw")
    _run(variant, a, _sim(n, 20, stamp=10, lrestart=True), fl)
    _run(variant, b, _sim(n, 40, stamp=10), fl)
    for fr in (3, 4):  # frame numbering continues from the restart file
        assert np.array_equal(fhost.read_frame(a, "vel", fr, n, 3), fhost.read_frame(b, "vel", fr, n, 3))
        assert np.array_equal(fhost.read_frame(a, "rho", fr, n), fhost.read_frame(b, "rho", fr, n))


def test_poiseuille_bounce_back_channel(tmp_path):
    variant = "pois19_bb_fp64"
    n, nu, fz, steps = (20, 4, 8), 1.0 / 6.0, 1e-6, 400
    _run(variant, tmp_path, _sim(n, steps, pbc_x=0), {"visc1": nu, "fz": fz})
    vel = fhost.read_frame(tmp_path, "vel", 1, n, 3).astype(np.float64)
    gi = np.arange(1, n[0] + 1, dtype=np.float64)
    hp, xc = 0.5 * (n[0] - 2), 0.5 * (n[0] + 1)
    dist = gi - xc
    want = np.where(np.abs(dist) <= hp, fz / (2.0 * nu) * (hp * hp - dist * dist), 0.0)
    w = vel[0, 0, :, 2]
    inner = slice(1, n[0] - 1)
    assert np.abs(w[inner] - want[inner]).max() < 0.01 * want.max()
    # the same start through the Python bindings: identical frames
    p = make_params(n, lattice=19, highorder=False, internal_obstacles=True, pbc=(0, 1, 1), visc1=nu, fz=fz,
                    precision=PREC_FP64, bounce_back=True)
    isf = hip.classify_isfluid(p)
    shape = (n[2], n[1], n[0])
    wz = np.where(np.abs(isf) == 1, np.broadcast_to(want, shape), 0.0)
    f = {k: np.zeros(shape) for k in ("P", "u", "v", "pxx", "pyy", "pxy", "pxz", "pyz")}
    f["w"], f["pzz"] = wz, wz * wz
    d = hip.Domain(p)
    d.load(f, isf)
    d.step(steps)
    got = np.stack([_f32(d.download(i)) for i in (2, 3, 4)], axis=-1)
    d.close()
    assert np.array_equal(fhost.read_frame(tmp_path, "vel", 1, n, 3), got)

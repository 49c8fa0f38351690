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
* The end-of-run benchmark comparisons of statistics.f90:688-1114: the one-
  and two-fluid Poiseuille channels approach the analytic profiles the host
  writes beside the numerical ones (plot_poiseuille.dat,
  plot_twopoiseuille.dat), and the Lamb oscillation period the host derives
  from lamb.dat equals the same peak rules applied independently to that file.
* Output formats: XDMF sidecars describe the RAW frames, VTK image data holds
  the same values as the RAW frame, 2-D plane frames are slices of the 3-D
  frame of the same step (print_module.f90:2007-2660).
* BCPHIFLUX: the host's switch reaches the library, and the phase mass of a
  two-layer channel stays nearer its start than without it.
* Open boundaries: a deck with pressure faces drives the flow from the high
  to the low pressure, and its frames equal a hip.Domain run bit for bit.
* The starts of the Taylor-Green, Laplace, Poiseuille (with and without
  bounce-back), two-fluid Poiseuille and Lamb builds equal the fields of the
  reference's own initial-condition statements (tests/golden/ref_init.npz).
"""
import json
import os

import numpy as np
import pytest

from lbfast_amd import hip, setups
from lbfast_amd.params import make_params, PREC_FP64, PREC_FP32, PREC_FP16S, FIELD_PHI
from tests import fhost

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(fhost.exe("tg27h_fp32")), reason="Fortran host not built")]

LAT = {"tg27h_fp32": (27, True, PREC_FP32), "tg27h_fp64": (27, True, PREC_FP64), "tg19_fp32": (19, False, PREC_FP32),
       "pois19_fp64": (19, False, PREC_FP64), "twopois27h_fp64": (27, True, PREC_FP64),
       "lamb27h_fp32": (27, True, PREC_FP32),
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
    want_rho = _f32(phi)  # LAPLACE implies PRINTPHI (defines.h:38-43): the rho frame holds phi
    want_vel = np.stack([_f32(d.download(i)) for i in (2, 3, 4)], axis=-1)
    want_p = _f32(d.download(1))  # and WRITEPRESS: P in out_press
    d.close()
    assert np.array_equal(fhost.read_frame(tmp_path, "rho", 1, n), want_rho)
    assert np.array_equal(fhost.read_frame(tmp_path, "press", 1, n), want_p)
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
    # frames hold phi (PRINTPHI) in real(4)
    assert np.abs(got0 - want0).max() < 1e-6
    assert np.abs(got1 - phi).max() < 1e-6
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


_INIT = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_init.npz"))
_INIT_META = json.loads(_INIT["meta"].tobytes().decode())


@pytest.mark.parametrize("case,variant", [("tg", "tg27h_fp64"), ("laplace_2c", "laplace27h_fp64"),
                                          ("pois", "pois19_fp64"), ("pois_bb", "pois19_bb_fp64"),
                                          ("twopois", "twopois27h_fp64"), ("lamb", "lamb27h_fp32")])
def test_host_start_matches_reference_statements(tmp_path, case, variant):
    """Frame 0 of a host run (written after lbf_init_state) against the
    fields the reference's initial-condition statements give on the same
    grid (tests/golden/ref_init.npz; the Poiseuille noise term taken as 0,
    which the host leaves out): the print observable (phi under PRINTPHI,
    which LAPLACE, LAMBTEST and TWOPOISEUILLE imply, P otherwise), the
    velocity and, where written, P.  Frames are real(4): agreement to one
    float32 ulp.  Walls (solid nodes) are left out."""
    m = _INIT_META[case]
    n, q = tuple(m["n"]), m["params"]
    want = _INIT[case + "/fields"]
    two = "TWOCOMPONENT" in m["defines"]
    keys = ("visc1", "fz", "uwall") + (("visc2", "rho_r", "rho_b", "sigma", "width", "radius", "center")
                                        if two else ("width", "radius", "center"))
    fl = {k: q[k] for k in keys if k in q}
    fl.setdefault("visc1", 0.1)  # the viscosities do not enter the start
    if two:
        fl.setdefault("visc2", 0.1)
        fl["tau_diff"] = 0.1
    walls = case.startswith("pois") or case.startswith("twopois")
    _run(variant, tmp_path, _sim(n, 2, stamp=2, stamp_term=2, pbc_x=0 if walls else 1), fl)
    mask = np.ones((n[2], n[1], n[0]), dtype=bool)
    if walls:
        mask[:, :, 0] = mask[:, :, -1] = False

    def agree(got, ref):
        r = ref.astype(np.float32)
        tol = np.spacing(np.abs(r)) + 1e-30
        return bool(np.all(np.abs(got - r)[mask] <= tol[mask]))
    obs = want[10] if two else want[0]
    assert agree(fhost.read_frame(tmp_path, "rho", 0, n), obs), case
    vel = fhost.read_frame(tmp_path, "vel", 0, n, 3)
    for c in range(3):
        assert agree(vel[..., c], want[1 + c]), (case, c)
    if os.path.exists(tmp_path / "output" / "out_press_000000.raw"):
        assert agree(fhost.read_frame(tmp_path, "press", 0, n), want[0]), case


def test_open_faces_through_the_host(tmp_path):
    """Open boundaries from a deck: a pressure inlet and outlet on the y
    faces, the x walls as velocity faces at rest, z periodic, starting at
    rest.  The flow runs from y(1) to y(2), and the frames equal those of a
    hip.Domain run with the same parameters bit for bit."""
    variant = "pois19_fp64"
    n, nu, dP, steps = (12, 30, 4), 1.0 / 6.0, 1e-4, 300
    sim = _sim(n, steps, pbc_x=0, pbc_y=0, openbc=1, openbc_type_x=[0, 0], openbc_type_y=[1, 1],
               openbc_press_y=[dP, 0.0])
    out = _run(variant, tmp_path, sim, {"visc1": nu, "fz": 0.0})
    h = fhost.header(out)
    assert h["openbc"] == 1 and h["openbc_type"] == (0, 0, 1, 1, 0, 0)
    vel = fhost.read_frame(tmp_path, "vel", 1, n, 3).astype(np.float64)
    v = vel[..., 1]
    assert v[:, 2:-2, 2:-2].min() > 0.0  # from the high-pressure face y(1) to y(2)
    # mid-channel: mirror-symmetric across x (both x faces at rest; the
    # frames are float32, so to one float32 ulp), peaked between the walls
    mid = v[:, n[1] // 2, :]
    assert np.allclose(mid, mid[:, ::-1], rtol=2e-7, atol=1e-15)
    assert np.argmax(mid[0]) in (n[0] // 2 - 1, n[0] // 2)
    p = make_params(n, lattice=19, highorder=False, internal_obstacles=True, pbc=(0, 0, 1), visc1=nu,
                    precision=PREC_FP64)
    p.openbc = 1
    p.openbc_type[2], p.openbc_type[3] = 1, 1
    p.openbc_press[2], p.openbc_press[3] = dP, 0.0
    isf = hip.classify_isfluid(p)
    shape = (n[2], n[1], n[0])
    d = hip.Domain(p)
    d.load({k: np.zeros(shape) for k in ("P", "u", "v", "w", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz")}, isf)
    d.step(steps)
    got = np.stack([_f32(d.download(i)) for i in (2, 3, 4)], axis=-1)
    d.close()
    assert np.array_equal(fhost.read_frame(tmp_path, "vel", 1, n, 3), got)


def _poiseuille_line(out):
    line = [l for l in out.splitlines() if l.startswith("POISEUILLE ")][-1]
    tok = line.split()
    return {tok[i]: float(tok[i + 1]) for i in range(2, len(tok) - 1, 2)}


def test_poiseuille_profile_and_output_formats(tmp_path):
    """One-fluid channel with the near-wall closure (walls in x, force in z):
    the x profile written to plot_poiseuille.dat approaches the parabola with
    H = (lx-1)/2; the frames of the run come as RAW with XDMF sidecars, as
    VTK image data and as 2-D planes."""
    variant = "pois19_fp64"
    n, nu, fz, steps = (34, 4, 8), 1.0 / 6.0, 5e-6, 6000
    sim = _sim(n, steps, pbc_x=0, stamp=steps, stamp2D=steps, lvtk=True, nplanes=2)
    out = _run(variant, tmp_path, sim, {"visc1": nu, "fz": fz, "ndir": [1, 3], "npoint": [9, 5]}, )
    res = _poiseuille_line(out)
    print(f"[fortran host]   poiseuille {res}", flush=True)
    prof = np.loadtxt(tmp_path / "plot_poiseuille.dat")
    assert prof.shape == (n[0], 3)
    assert res["rel_l2"] < 0.05
    # the profile the host wrote is the column mean of its own last frame,
    # with the solid wall columns counted as zero
    vel = fhost.read_frame(tmp_path, "vel", 1, n, 3).astype(np.float64)
    assert np.allclose(prof[1:-1, 1], vel[..., 2].mean(axis=(0, 1))[1:-1], rtol=1e-6, atol=1e-12)
    assert prof[0, 1] == 0 and prof[-1, 1] == 0
    # XDMF sidecars name their RAW files and the grid, slowest axis first
    x = (tmp_path / "output" / "out_vel_000001.xdmf").read_text()
    assert "out_vel_000001.raw" in x and f'Dimensions="{n[2]} {n[1]} {n[0]} 3"' in x
    assert 'AttributeType="Vector"' in x and 'Endian="Little"' in x
    # VTK image data: the appended block holds the RAW frame's bytes
    raw = (tmp_path / "output" / "out_vel_000001.raw").read_bytes()
    vti = (tmp_path / "output" / "out_vel_000001.vti").read_bytes()
    head, rest = vti.split(b'<AppendedData encoding="raw">\n', 1)
    assert b'WholeExtent="1 34 1 4 1 8"' in head and b'NumberOfComponents="3"' in head
    assert rest[:1] == b"_"
    nbytes = int(np.frombuffer(rest[1:5], dtype=np.int32)[0])
    assert nbytes == len(raw) and rest[5:5 + nbytes] == raw
    assert rest[5 + nbytes:].startswith(b"\n  </AppendedData>")
    # 2-D planes: x = 9 and z = 5 slices of the 3-D frame of the same step
    rho = fhost.read_frame(tmp_path, "rho", 1, n)
    px = np.fromfile(tmp_path / "output" / "out_1_9_rho_000001.raw", dtype=np.float32).reshape(n[2], n[1])
    pz = np.fromfile(tmp_path / "output" / "out_3_5_vel_000001.raw", dtype=np.float32).reshape(n[1], n[0], 3)
    assert np.array_equal(px, rho[:, :, 8])
    assert np.array_equal(pz, fhost.read_frame(tmp_path, "vel", 1, n, 3)[4])
    assert (tmp_path / "output" / "out_1_9_rho_000001.xdmf").exists()


def test_two_fluid_poiseuille_profile(tmp_path):
    """twopoiseuille.inp's physics at 34 x 4 x 8: two layers split at the
    channel centre (nu 0.05 / 0.4, rho 0.5 / 1.5, contact angle 90), walls in
    x, force in z.  The host's profile equals the FP64 oracle's run of the
    same start; both sit at rel-L2 0.16 from the reference's sharp-interface
    analytic profile, the diffuse interface (w = 4) spanning a quarter of
    the channel."""
    import oracle
    variant = "twopois27h_fp64"
    n, steps = (34, 4, 8), 8000
    fl = {"fz": 1e-6, "tau_diff": 0.1, "width": 4.0, "sigma": 0.0, "visc1": 0.05, "visc2": 0.4,
          "rho_r": 0.5, "rho_b": 1.5, "wettab_r": 90.0}
    out = _run(variant, tmp_path, _sim(n, steps, pbc_x=0, stamp=steps, stamp_term=steps), fl, timeout=300)
    res = _poiseuille_line(out)
    prof = np.loadtxt(tmp_path / "plot_twopoiseuille.dat")
    assert prof.shape == (n[0], 3)
    # the same start in the oracle: walls by classification, the two-layer
    # phi with the solid-node averages of init_conditions_module.f90:468-542,
    # and the analytic w the host starts from
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, internal_obstacles=True,
                    pbc=(0, 1, 1), visc1=0.05, visc2=0.4, rho_r=0.5, rho_b=1.5, tau_diff=0.1, sigma=0.0, width=4.0,
                    fz=1e-6, precision=PREC_FP64)
    isf = oracle.classify_isfluid(p, None)
    shape = (n[2], n[1], n[0])
    phi = np.broadcast_to(setups.fcut_tanh(np.arange(1, n[0] + 1) - 0.5 * (n[0] + 1), 0.0, 4.0), shape).copy()
    phi = setups.solid_phi_average(phi, isf, 27)
    w0 = np.where(np.abs(isf) == 1, np.broadcast_to(prof[:, 2], shape), 0.0)
    f = {k: np.zeros(shape) for k in ("P", "u", "v", "pxx", "pyy", "pxy", "pxz", "pyz")}
    f.update({"w": w0, "pzz": w0 * w0, "phi": phi})
    o = oracle.Oracle(p)
    o.load(f, isf)
    o.step(steps)
    w_orc = np.where(isf != 0, o.get_field(4), 0.0).mean(axis=(0, 1))
    e_orc = float(np.abs(prof[:, 1] - w_orc).max() / np.abs(w_orc).max())
    print(f"[fortran host]   two-fluid poiseuille {res}; host vs oracle profile {e_orc:.2e}", flush=True)
    assert e_orc < 5e-8
    assert res["rel_l2"] < 0.2
    assert (tmp_path / "output" / "out_press_000001.xdmf").exists()


def test_lamb_period_analysis(tmp_path):
    """The Lamb test with lamb.inp's droplet (an ellipsoid with semi-axes 2R,
    R, R; R = 20, width 5, sigma = 0.03) in half of lamb.inp's box, 128^3,
    with nu = 0.01: lamb.dat every stamp_term steps, and at the end the peaks
    of the z half-axis at least 2000 steps apart, their mean spacing and
    omega_num, which the same rules applied here to lamb.dat reproduce.  The
    oscillation of this droplet is slower than the 2000-step peak spacing,
    so every peak is found."""
    variant = "lamb27h_fp32"
    n, steps, every = (128, 128, 128), 12000, 20
    fl = {"tau_diff": 0.1, "radius": 20.0, "width": 5.0, "center": [64.0, 64.0, 64.0], "sigma": 0.03,
          "visc1": 0.01, "visc2": 0.01, "rho_r": 1.0, "rho_b": 1.0, "wettab_r": 90.0}
    out = _run(variant, tmp_path, _sim(n, steps, stamp=steps, stamp_term=every, lraw=False), fl, timeout=300)
    rows = np.loadtxt(tmp_path / "lamb.dat")
    assert rows.shape[0] == steps // every + 1 and np.all(np.isfinite(rows))
    line = [l for l in out.splitlines() if l.startswith("LAMB omega_num omega_th")]
    assert line, out[-2000:]
    om_num, om_th = (float(x) for x in line[0].split()[-2:])
    step, dosc = rows[:, 0].astype(int), rows[:, 6]
    peaks = []
    for i in range(1, len(dosc) - 1):
        if dosc[i] > dosc[i - 1] and dosc[i] >= dosc[i + 1] and dosc[i] > 20.0:
            if not peaks or step[i] - peaks[-1] >= 2000:
                peaks.append(step[i])
    assert len(peaks) >= 3
    # every local maximum of the record is a peak: none falls inside the
    # 2000-step exclusion of the one before it
    maxima = [int(step[i]) for i in range(1, len(dosc) - 1)
              if dosc[i] > dosc[i - 1] and dosc[i] >= dosc[i + 1] and dosc[i] > 20.0 and step[i] < 9500]
    assert maxima == [q for q in peaks if q < 9500]
    tmean = (peaks[-1] - peaks[0]) / (len(peaks) - 1)
    print(f"[fortran host]   lamb peaks {peaks} omega_num {om_num:.6f} theory {om_th:.6f}", flush=True)
    assert abs(om_num - 2 * np.pi / tmean) < 1e-6 * om_num
    assert (tmp_path / "lamb_theory.dat").exists()


def test_bcphiflux_switch_reaches_the_library(tmp_path):
    drift = {}
    for variant in ("twopois27h_fp64", "twopois27h_bcphi_fp64"):
        wd = tmp_path / variant
        wd.mkdir()
        n, steps = (34, 4, 8), 400
        fl = {"fz": 1e-6, "tau_diff": 0.1, "width": 4.0, "sigma": 0.01, "visc1": 0.05, "visc2": 0.4,
              "rho_r": 0.5, "rho_b": 1.5, "wettab_r": 90.0}
        out = _run(variant, wd, _sim(n, steps, pbc_x=0, stamp=steps, stamp_term=20, lraw=False), fl)
        assert fhost.header(out)["bcphiflux"] == (1 if "bcphi" in variant else 0)
        mass = [r["phase_mass"] for r in fhost.stats(out, "LAPLACE")]
        drift[variant] = max(abs(m - mass[0]) for m in mass) / mass[0]
    print(f"[fortran host]   phase-mass drift {drift}", flush=True)
    assert drift["twopois27h_bcphi_fp64"] < drift["twopois27h_fp64"]

"""16-bit moment storage (FP16S / BF16S): scaled deviations, FP32 arithmetic.

* Codec, bit-exact: an FP16S upload stores half(scale * P), half(scale * u)
  and half(scale * (Pi - u_r u_r)) with u_r the stored-and-decoded velocity;
  a download decodes stored / scale.  The GPU round trip equals the same
  formula evaluated with numpy's IEEE binary16 (BF16S: bfloat16, round to
  nearest even, computed by bit manipulation).
* Time steps stay within stated error bounds of the FP64 run of the same
  problem (velocity relative L2 after 50 steps of a two-component D3Q27h
  droplet: FP16S < 3e-3, BF16S < 3e-2; phase field < 1e-4), including sizes
  whose planes hold an odd number of 16-bit values.
* A reduced two-phase Poiseuille channel with internal obstacles (the
  reference's porous-channel workload at 64x48x32: walls in y, 25 % spheres,
  dense lower 40 % at density ratio 10, body force 1e-6) after 2000 steps:
  velocity relative L2 against FP64 over the fluid nodes FP32 < 2e-4,
  FP16S < 1e-3, BF16S < 1e-2.  The time step rounds the 16-bit stores
  stochastically; with round-to-nearest the FP16S error of this run grows to
  5e-3 by 5000 steps as sub-ulp velocity increments are dropped.
"""
import numpy as np
import pytest

from lbfast_amd import hip, setups
from lbfast_amd.params import (make_params, PREC_FP64, PREC_FP32, PREC_FP16S, PREC_BF16S, HFIELDS, HFIELD_NAMES,
                               FIELD_PHI)
from tests import cases

pytestmark = pytest.mark.gpu

SCALES = np.array([16384.0, 1024.0, 1024.0, 1024.0] + [1048576.0] * 6)  # the library's default scales


def _bf16(x):
    """float32 -> bfloat16 (round to nearest even) -> float32."""
    b = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def _q(x, prec):
    x32 = np.asarray(x, dtype=np.float32)
    return x32.astype(np.float16).astype(np.float32) if prec == PREC_FP16S else _bf16(x32)


@pytest.mark.parametrize("prec", [PREC_FP16S, PREC_BF16S])
def test_codec_round_trip_bit_exact(prec):
    n = (24, 10, 7)
    p = make_params(n, lattice=27, highorder=True, precision=prec)
    rng = np.random.default_rng(4)
    shape = (n[2], n[1], n[0])
    u = [rng.uniform(-0.05, 0.05, shape) for _ in range(3)]
    f = {"P": rng.uniform(-1e-3, 1e-3, shape), "u": u[0], "v": u[1], "w": u[2]}
    pairs = {"pxx": (0, 0), "pyy": (1, 1), "pzz": (2, 2), "pxy": (0, 1), "pxz": (0, 2), "pyz": (1, 2)}
    for k, (i, j) in pairs.items():
        f[k] = u[i] * u[j] + rng.uniform(-1e-5, 1e-5, shape)
    d = hip.Domain(p)
    d.load(f)
    got = {nm: d.download(fid) for fid, nm in zip(HFIELDS, HFIELD_NAMES)}
    d.close()
    # expected, in FP32 arithmetic as the encoder does it
    s32 = SCALES.astype(np.float32)
    inv = (1.0 / SCALES).astype(np.float32)
    f32 = {k: v.astype(np.float32) for k, v in f.items()}
    want = {"P": _q(f32["P"] * s32[0], prec) * inv[0]}
    ur = []
    for c, nm in enumerate(("u", "v", "w")):
        want[nm] = _q(f32[nm] * s32[1 + c], prec) * inv[1 + c]
        ur.append(want[nm])
    for m, (k, (i, j)) in enumerate(pairs.items()):
        dev = _q((f32[k] - ur[i] * ur[j]) * s32[4 + m], prec) * inv[4 + m]
        want[k] = dev + ur[i] * ur[j]
    for nm in HFIELD_NAMES:
        w = want[nm].astype(np.float32)
        g = got[nm].astype(np.float32)
        assert np.array_equal(g, w), (nm, float(np.abs(g - w).max()))


def _droplet(n, prec):
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.05, visc2=0.05,
                    rho_r=1.0, rho_b=0.1, tau_diff=0.1, sigma=0.01, width=3.0, precision=prec)
    f = setups.droplet(n, min(n) / 3.0, 3.0, [n[0] / 2 + 0.5, n[1] / 2 + 0.5, n[2] / 2 + 0.5], 0.01, 1.0, 0.1,
                       noise=1e-4, seed=9)
    return p, f


# (96, 20, 12) and (130, 9, 7): full 64-node row segments (two 16-bit fields
# per LDS-DMA transfer) beside short ones (one field per transfer) in one domain
@pytest.mark.parametrize("n", [(32, 32, 32), (33, 17, 21), (96, 20, 12), (130, 9, 7)])
@pytest.mark.parametrize("prec,tol_u", [(PREC_FP16S, 3e-3), (PREC_BF16S, 3e-2)])
def test_sixteen_bit_steps_against_fp64(n, prec, tol_u):
    steps = 50
    print(f"[precision] prec={prec} n={n} steps={steps}", flush=True)
    res = {}
    for q in (PREC_FP64, prec):
        p, f = _droplet(n, q)
        d = hip.Domain(p)
        d.load(f)
        d.step(steps)
        res[q] = (np.stack([d.download(i) for i in (2, 3, 4)]), d.download(FIELD_PHI), d.download(1))
        d.close()
    (u64, phi64, P64), (u16, phi16, P16) = res[PREC_FP64], res[prec]
    eu = np.linalg.norm(u16 - u64) / np.linalg.norm(u64)
    ephi = np.linalg.norm(phi16 - phi64) / np.linalg.norm(phi64)
    print(f"[precision]   u {eu:.2e} phi {ephi:.2e}", flush=True)
    assert np.all(np.isfinite(u16)) and np.all(np.isfinite(P16))
    assert eu < tol_u
    assert ephi < 1e-4


@pytest.mark.parametrize("prec,tol_u", [(PREC_FP16S, 3e-3), (PREC_BF16S, 3e-2)])
def test_given_scales_against_default_scales(prec, tol_u):
    """The default scales are compiled into their own kernels; any other
    store_scale takes the kernels that read the scales as arguments.  With
    every scale doubled the run stays as close to FP64 as the default one
    (the stochastic rounding draws differ, so not bit for bit), and a
    download decodes with the given scales."""
    n, steps = (40, 24, 20), 50
    print(f"[precision] given scales prec={prec} n={n} steps={steps}", flush=True)
    res = {}
    for tag, q in (("fp64", PREC_FP64), ("default", prec), ("doubled", prec)):
        p, f = _droplet(n, q)
        if tag == "doubled":
            for i in range(10):
                p.store_scale[i] = 2.0 * SCALES[i]
        d = hip.Domain(p)
        d.load(f)
        if tag == "doubled":
            u0 = d.download(2)
            half_ulp = 2.0 ** -11 if prec == PREC_FP16S else 2.0 ** -8  # of the significand
            assert np.abs(u0 - f["u"]).max() <= half_ulp * np.abs(f["u"]).max() + 1e-12
        d.step(steps)
        res[tag] = np.stack([d.download(i) for i in (2, 3, 4)])
        d.close()
    e = {k: np.linalg.norm(res[k] - res["fp64"]) / np.linalg.norm(res["fp64"]) for k in ("default", "doubled")}
    print(f"[precision]   u error default {e['default']:.2e} doubled {e['doubled']:.2e}", flush=True)
    assert e["default"] < tol_u and e["doubled"] < tol_u
    assert e["doubled"] < 2.0 * e["default"] + 1e-5


@pytest.mark.parametrize("name", ["drop27h_bulk", "bubble27h_bulk"])
@pytest.mark.parametrize("prec,tol_u", [(PREC_FP16S, 3e-3), (PREC_BF16S, 3e-2)])
def test_bulk_planes_against_fp64(name, prec, tol_u):
    """The droplet (or bubble) with exact bulk values away from the
    interface: the 16-bit run stays within the droplet bounds of the FP64 run."""
    n, steps = (96, 40, 40), 50
    print(f"[precision] bulk planes {name} prec={prec} n={n} steps={steps}", flush=True)
    res = {}
    for q in (PREC_FP64, prec):
        p, f, _ = cases.case(name, n, q)
        d = hip.Domain(p)
        d.load(f)
        d.step(steps)
        res[q] = (np.stack([d.download(i) for i in (2, 3, 4)]), d.download(FIELD_PHI))
        d.close()
    (u64, phi64), (u16, phi16) = res[PREC_FP64], res[prec]
    eu = np.linalg.norm(u16 - u64) / np.linalg.norm(u64)
    ephi = np.linalg.norm(phi16 - phi64) / np.linalg.norm(phi64)
    print(f"[precision]   u {eu:.2e} phi {ephi:.2e}", flush=True)
    assert eu < tol_u and ephi < 1e-4


def test_porous_channel_against_fp64():
    n, steps = (64, 48, 32), 2000
    raw = setups.random_spheres(n, fill=0.25, rmin=3.0, rmax=6.0, seed=11, margin=2)
    res = {}
    for prec in (PREC_FP64, PREC_FP32, PREC_FP16S, PREC_BF16S):
        print(f"[precision] porous channel prec={prec} n={n} steps={steps}", flush=True)
        p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, internal_obstacles=True,
                        pbc=(1, 0, 1), visc1=0.05, visc2=0.05, rho_r=10.0, rho_b=1.0, tau_diff=0.1, sigma=0.005,
                        width=4.0, fx=1e-6, precision=prec)
        f = setups.two_layer_channel(n, raw, frac=0.4, width=4.0, rho_r=10.0, rho_b=1.0, axis=1)
        f, isf = cases.prepare(p, f, raw, hip.classify_isfluid)
        d = hip.Domain(p)
        d.load(f, isf)
        d.step(steps)
        res[prec] = (np.stack([d.download(i) for i in (2, 3, 4)]), d.download(FIELD_PHI), isf)
        d.close()
    u64, phi64, isf = res[PREC_FP64]
    fl = isf != 0
    for prec, tol in ((PREC_FP32, 2e-4), (PREC_FP16S, 1e-3), (PREC_BF16S, 1e-2)):
        u, phi, _ = res[prec]
        eu = np.linalg.norm((u - u64)[:, fl]) / np.linalg.norm(u64[:, fl])
        ephi = np.linalg.norm((phi - phi64)[fl]) / np.linalg.norm(phi64[fl])
        print(f"[precision]   prec={prec}: u {eu:.2e} phi {ephi:.2e}", flush=True)
        assert np.all(np.isfinite(u))
        assert eu < tol, (prec, eu)
        assert ephi < 1e-3, (prec, ephi)


def test_planar_channel_fp16s_rounding_varies_by_step():
    """A planar Poiseuille channel (walls in y, body force in x): w, Pxz and
    Pyz are exactly 0 at every node, so the siblings whose low bits feed the
    stochastic rounding of v, Pxy and Pxz are constant.  The rounding seed
    also mixes in the step number, so each store still draws new bits every
    step: the FP16S velocity stays within the 1e-3 bar of the FP64 run after
    the flow has developed (4000 steps, about 2.6 diffusion times)."""
    n, steps = (16, 18, 4), 4000
    res = {}
    for prec in (PREC_FP64, PREC_FP16S):
        print(f"[precision] planar channel prec={prec} n={n} steps={steps}", flush=True)
        p = make_params(n, lattice=27, highorder=True, internal_obstacles=True, pbc=(1, 0, 1), visc1=1.0 / 6.0,
                        fx=5e-6, precision=prec)
        z = np.zeros((n[2], n[1], n[0]))
        f = {k: z.copy() for k in HFIELD_NAMES}
        isf = hip.classify_isfluid(p, None)
        d = hip.Domain(p)
        d.load(f, isf)
        d.step(steps)
        res[prec] = (np.stack([d.download(i) for i in (2, 3, 4)]), isf)
        d.close()
    (u64, isf), (u16, _) = res[PREC_FP64], res[PREC_FP16S]
    fl = isf != 0
    eu = np.linalg.norm((u16 - u64)[:, fl]) / np.linalg.norm(u64[:, fl])
    umax = float(np.abs(u64[0]).max())
    print(f"[precision]   planar channel u_max {umax:.3e}, FP16S velocity rel-L2 {eu:.2e}", flush=True)
    assert umax > 5e-4
    assert eu < 1e-3

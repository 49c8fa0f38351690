#!/usr/bin/env python3
"""Benchmark of the LB time step on one MI355X.

Headline workload (the metric's configuration): two-component D3Q27
HIGHORDER DENSRATIO, 512^3 periodic, a Laplace droplet (R = 128, w = 3,
density ratio 100, sigma = 0.01), FP32 moments in row-interleaved field
sets.  A "step" is one full time step of lbf_step (phase-field stencils +
the fused stream/collide/moments/phase-update kernel) over the whole domain;
state stays resident in HBM.

Prints ONE JSON line:
  value           MLUPS = nodes * steps / t / 1e6 (solid nodes counted, LBFAST.f90:500)
  roofline        the fused kernel against HBM: algorithmic bytes per launch
                  (112 B/node FP32: 10 moments + 3 u_old + phi, read and written)
                  over its average launch time from HIP events on the context stream
  fp16s / bf16s   the same workload with moments stored as scaled 16-bit
                  deviations, FP32 arithmetic; speed-up over FP32
  taylor_green_1c the one-component Taylor-Green vortex, D3Q27 HIGHORDER, 512^3,
                  U0 = 0.04, tau = 0.55, FP32 (80 B/node), with its own roofline
  precision_error per-field relative L2 and Linf of FP64 / FP32 / FP16S / BF16S
                  storage against the FP64 oracle (droplet and porous channel)
  ensemble        96 droplets of uneven size, batched and graph-captured, against
                  the same domains stepped one by one
  fortran_host    the headline workload driven by the Fortran host (namelist
                  deck, its own initial condition and time loop) through the C-ABI
  cpu_baseline    the oracle (a CPU port of the reference algorithm) timed on
                  this host over a few steps of the same 512^3 workload
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from lbfast_amd import setups  # noqa: E402
from lbfast_amd.params import make_params, PREC_FP32, PREC_FP64, PREC_FP16S, PREC_BF16S  # noqa: E402

METRIC = "MLUPS & % HBM peak, D3Q27 2-comp 512^3; FP16-store vs FP32 at stated error"
HBM_PEAK_GBS = 8000.0
A100_2C_D3Q27H_SP_MLUPS = 800.0   # README.md:405, 1x A100, 2c 512^3 D3Q27h SP
A100_1C_D3Q27H_SP_MLUPS = 1000.0  # README.md:340, 1x A100, 1c 512^3 D3Q27h SP
BYTES_PER_LUP = {PREC_FP32: 112.0, PREC_FP64: 224.0, PREC_FP16S: 60.0, PREC_BF16S: 60.0}
BYTES_PER_LUP_1C = {PREC_FP32: 80.0, PREC_FP64: 160.0, PREC_FP16S: 40.0, PREC_BF16S: 40.0}
DTYPE = {PREC_FP32: "f32", PREC_FP64: "f64", PREC_FP16S: "f32 (f16 storage)", PREC_BF16S: "f32 (bf16 storage)"}
VEL_ERR_BAR = 1e-3  # velocity relative L2 of 16-bit storage against FP64
ENSEMBLE_STEPS = 2000


def droplet_params(n, precision):
    return make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True,
                       visc1=0.1, visc2=0.1, rho_r=1.0, rho_b=0.01, tau_diff=0.1, sigma=0.01, width=3.0,
                       precision=precision)


def droplet_fields(n, dtype=np.float32):
    R = n[0] / 4.0
    f = setups.droplet(n, R, 3.0, [n[0] / 2 + 0.5, n[1] / 2 + 0.5, n[2] / 2 + 0.5], 0.01, 1.0, 0.01,
                       noise=1e-6, seed=20261202)
    return {k: v.astype(dtype) for k, v in f.items()}


def tg_params(n, precision):
    """Taylor-Green vortex, one component, D3Q27 HIGHORDER: tau = 0.55 (nu = 0.05/3)."""
    return make_params(n, lattice=27, highorder=True, visc1=0.05 / 3.0, precision=precision)


def tg_fields(n, dtype=np.float32):
    return {k: v.astype(dtype) for k, v in setups.taylor_green(n, u0=0.04).items()}


def load_domain(p, fields):
    from lbfast_amd import hip
    d = hip.Domain(p)
    dt = np.float32 if next(iter(fields.values())).dtype == np.float32 else np.float64
    from lbfast_amd.params import HFIELDS, HFIELD_NAMES, FIELD_PHI
    for fid, nm in zip(HFIELDS, HFIELD_NAMES):
        d.upload(fid, fields[nm], dtype=dt)
    if p.twocomponent:
        d.upload(FIELD_PHI, fields["phi"], dtype=dt)
    d.init_state()
    return d


def time_domain(d, steps, warmup):
    d.set_timing(True)
    d.step(warmup)
    d.synchronize()
    t0 = time.perf_counter()
    d.step(steps)
    d.synchronize()
    wall = time.perf_counter() - t0
    total_ms, fused_ms = d.timing()
    return wall, total_ms, fused_ms


def _field_errors(got, ref, mask):
    """Relative L2 and Linf of every field over the nodes in mask."""
    out = {}
    for k in ref:
        a, b = got[k][mask], ref[k][mask]
        nb = float(np.linalg.norm(b))
        out[k] = [float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a - b)),
                  float(np.abs(a - b).max())]
    u = np.stack([got[k][mask] for k in ("u", "v", "w")])
    ur = np.stack([ref[k][mask] for k in ("u", "v", "w")])
    out["vel"] = [float(np.linalg.norm(u - ur) / np.linalg.norm(ur)), float(np.abs(u - ur).max())]
    return out


def precision_errors():
    """Per-field errors of FP64 / FP32 / FP16S / BF16S storage against the
    FP64 oracle (a CPU port of the reference stages) on two workloads:
    the Laplace droplet of the headline run at 64^3 (R = 16, density ratio
    100) after 400 steps, and the porous two-layer channel of the
    reference's Poiseuille deck (walls in y, 25 % spheres, dense lower 40 %
    at density ratio 10, body force 1e-6) at 64x48x32 after 2000 steps.
    Errors are taken over the fluid nodes; "vel" is the velocity vector."""
    import oracle
    from lbfast_amd import hip
    from lbfast_amd.params import HFIELDS, HFIELD_NAMES, FIELD_PHI, FIELD_RHOPHI
    names = list(HFIELD_NAMES)
    out = {"vs": "FP64 oracle (CPU port of the reference stages)", "bar_vel_rel_l2": VEL_ERR_BAR}
    n = (64, 64, 64)
    drop = (n, 400, droplet_params, droplet_fields(n, np.float64), None, "laplace_droplet_64^3_R16_densratio100")
    nc = (64, 48, 32)
    raw = setups.random_spheres(nc, fill=0.25, rmin=3.0, rmax=6.0, seed=11, margin=2)

    def chan_params(m, prec):
        return make_params(m, lattice=27, highorder=True, twocomponent=True, densratio=True, internal_obstacles=True,
                           pbc=(1, 0, 1), visc1=0.05, visc2=0.05, rho_r=10.0, rho_b=1.0, tau_diff=0.1, sigma=0.005,
                           width=4.0, fx=1e-6, precision=prec)
    fc = setups.two_layer_channel(nc, raw, frac=0.4, width=4.0, rho_r=10.0, rho_b=1.0, axis=1)
    isf = hip.classify_isfluid(chan_params(nc, PREC_FP64), raw)
    fc["phi"] = setups.solid_phi_average(fc["phi"], isf, 27)  # the reference's start on solid nodes
    chan = (nc, 2000, chan_params, fc, isf, "porous_channel_64x48x32_densratio10")
    for shape, steps, mkp, f, isf_, label in (drop, chan):
        print(f"[bench] precision: {label}, FP64 oracle {steps} steps", file=sys.stderr, flush=True)
        o = oracle.Oracle(mkp(shape, PREC_FP64))
        o.load(f, isf_)
        o.step(steps)
        ref = {nm: o.get_field(fid) for fid, nm in zip(HFIELDS, names)}
        ref["phi"] = o.get_field(FIELD_PHI)
        q = mkp(shape, PREC_FP64)
        ref["rhophi"] = q.rho_r * ref["phi"] + (1.0 - ref["phi"]) * q.rho_b
        del o
        mask = np.ones(ref["P"].shape, bool) if isf_ is None else isf_ != 0
        res = {"grid": list(shape), "steps": steps}
        for name, prec in (("fp64", PREC_FP64), ("fp32", PREC_FP32), ("fp16s", PREC_FP16S), ("bf16s", PREC_BF16S)):
            d = hip.Domain(mkp(shape, prec))
            d.load(f, isf_)
            d.step(steps)
            got = {nm: d.download(fid) for fid, nm in zip(HFIELDS, names)}
            got["phi"] = d.download(FIELD_PHI)
            got["rhophi"] = d.download(FIELD_RHOPHI)
            d.close()
            e = _field_errors(got, ref, mask)
            res[name] = {k: [float(f"{v[0]:.3e}"), float(f"{v[1]:.3e}")] for k, v in e.items()}
            if name in ("fp16s", "bf16s"):
                res[name]["meets_vel_bar"] = bool(e["vel"][0] <= VEL_ERR_BAR)
        out[label] = res
    out["columns"] = ["rel_l2", "linf"]
    return out


def fortran_host(n, steps=100):
    """The 512^3 droplet driven by the Fortran host (lbfast_amd/fortran): a
    namelist deck with the bench's parameters, the host's own droplet start
    (init_conditions_module.f90:394-426) and time loop; its printed MLUPS
    covers the time loop only, as LBFAST.f90:495-505."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "lbfast_amd", "fortran", "bin", "lbfast_host_laplace27h_fp32")
    if not os.path.exists(exe):
        return {"skipped": "Fortran host not built (amdflang absent at build time)"}
    c = n[0] / 2.0 + 0.5
    deck = (f"&simulation\n nsteps={steps}\n lx={n[0]}\n ly={n[1]}\n lz={n[2]}\n lprint=.false.\n"
            f" lraw=.false.\n iprobe=1\n jprobe=1\n kprobe=1\n&end\n"
            f"&fluid\n visc1=0.1\n visc2=0.1\n rho_r=1.0\n rho_b=0.01\n tau_diff=0.1\n sigma=0.01\n"
            f" width=3.0\n radius={n[0] / 4.0}\n center={c}, {c}, {c}\n&end\n")
    with tempfile.TemporaryDirectory() as wd:
        with open(os.path.join(wd, "input.inp"), "w") as f:
            f.write(deck)
        r = subprocess.run([exe, "input.inp"], cwd=wd, capture_output=True, text=True, timeout=600)
    line = [l for l in r.stdout.splitlines() if "MLUPS" in l]
    if r.returncode != 0 or not line:
        return {"error": (r.stdout + r.stderr)[-500:]}
    tok = line[-1].split()
    return {"value": float(tok[tok.index("MLUPS") + 1]), "unit": "MLUPS", "steps": int(tok[tok.index("steps") + 1]),
            "binary": os.path.basename(exe)}


def cpu_baseline(fields, n, steps=2):
    """The oracle on this host, on the headline workload itself (2c D3Q27h
    DENSRATIO droplet, FP32, the same 512^3 initial fields): one untimed
    step (first touch of the arrays), then `steps` timed steps."""
    import oracle
    p = droplet_params(n, PREC_FP32)
    o = oracle.Oracle(p)
    o.load(fields)
    o.time_steps(1)
    t = o.time_steps(steps)
    del o
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    mlups = n[0] * n[1] * n[2] * steps / t / 1e6
    return {"value": round(mlups, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"oracle (C++/OpenMP CPU port of the reference stages), FP32, 2c D3Q27h densratio "
                      f"droplet {n[0]}x{n[1]}x{n[2]}, {steps} steps after 1 untimed step, {t:.1f} s"}


def ensemble(prec, ndom=96, steps=ENSEMBLE_STEPS, seed=7):
    """Batched oscillating-droplet ensemble: ndom two-component D3Q27
    HIGHORDER periodic domains, edge L drawn from {32,48,64,96,128}, an
    ellipsoidal droplet R0 = 0.3 L with an l = 2 perturbation of amplitude
    U[0.02, 0.1], density ratio 50, sigma = 0.005, tau_diff = 0.1,
    nu = 0.001, w = 5.  Aggregate MLUPS of the batched step (autotuned tile,
    two-step captured graph) against the same domains stepped one by one."""
    from lbfast_amd import hip
    rng = np.random.default_rng(seed)
    Ls = rng.choice([32, 48, 64, 96, 128], size=ndom)
    amps = rng.uniform(0.02, 0.1, size=ndom)
    doms = []
    for L, a in zip(Ls, amps):
        n = (int(L),) * 3
        p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.001, visc2=0.001,
                        rho_r=1.0, rho_b=0.02, tau_diff=0.1, sigma=0.005, width=5.0, precision=prec)
        f = setups.droplet(n, 0.3 * L, 5.0, [L / 2 + 0.5] * 3, 0.005, 1.0, 0.02, noise=1e-6,
                           seed=int(rng.integers(1 << 30)), ellipse_a=float(a))
        d = hip.Domain(p)
        d.load(f)
        doms.append(d)
    nodes = int(sum(int(L) ** 3 for L in Ls))
    # the plain kernel looped over the domains
    for d in doms:
        d.step(2)
    for d in doms:
        d.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for d in doms:
            d.step(1)
    for d in doms:
        d.synchronize()
    looped = nodes * steps / (time.perf_counter() - t0) / 1e6
    b = hip.Batch(doms)
    ty, tz = b.autotune(4)
    _, tzs, _, order = b.tiles()
    classes = {int(L): tzs[i] for i, L in enumerate(Ls)}
    b.capture(2)
    b.step(2)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    batched = nodes * steps / wall / 1e6
    b.close()
    for d in doms:
        d.close()
    kind = "as given" if order == sorted(order) else (
        "largest first" if [int(Ls[d]) for d in order] == sorted([int(x) for x in Ls], reverse=True) else "alternating")
    return {"mlups_batched": round(batched, 1), "mlups_looped": round(looped, 1),
            "speedup": round(batched / looped, 2), "tile": [ty, tz],
            "z_chunk_by_edge": {str(k): v for k, v in sorted(classes.items())}, "order": kind}, nodes


def measured_traffic(workload):
    """HBM bytes per launch of the fused kernel measured with rocprofv3
    counters (profiles/hbm_traffic.txt), or None."""
    fn = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
    if not os.path.exists(fn):
        return None
    for line in open(fn):
        parts = line.split()
        if len(parts) == 4 and not line.startswith("#") and parts[0] == workload:
            return int(parts[2]) + int(parts[3])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    # the GPU clocks settle over the first ~8 launches (fused kernel 4.7 ->
    # 4.0 ms at 512^3), so the untimed warm-up covers them
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-precision", action="store_true", help="skip the 16-bit storage legs")
    ap.add_argument("--only", default="", help="run one leg: fp32 | fp16s | bf16s | 1c | ensemble "
                    "(the batched ensemble alone, --steps steps, for profiling)")
    ap.add_argument("--tile", default="", help="ty,tz override of the fused kernel tile")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB device copy (HBM counter calibration under rocprofv3)")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = 1 << 28  # 2^28 float32 = 1 GiB read + 1 GiB written per copy
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 3}))
        return

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on one GPU only (no multi-GPU sharding is implemented); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.only == "ensemble":
        ens = {}
        for name, prec in (("fp32", PREC_FP32), ("fp16s", PREC_FP16S)):
            print(f"[bench] batched ensemble {name}, {args.steps} steps", file=sys.stderr, flush=True)
            ens[name], nodes_e = ensemble(prec, steps=args.steps)
        print(json.dumps({"ensemble": {"domains": 96, "nodes": nodes_e, "steps": args.steps, **ens}}), flush=True)
        return

    n = (args.n, args.n, args.n)
    nodes = n[0] * n[1] * n[2]
    tile = tuple(int(x) for x in args.tile.split(",")) if args.tile else None
    precs = {"fp32": PREC_FP32, "fp16s": PREC_FP16S, "bf16s": PREC_BF16S}
    legs = [args.only] if args.only else ["fp32", "fp16s", "bf16s", "1c"]
    if args.no_precision and not args.only:
        legs = ["fp32", "1c"]
    bad = [x for x in legs if x not in precs and x != "1c"]
    if bad:
        ap.error(f"--only: unknown leg {bad[0]} (fp32 | fp16s | bf16s | 1c | ensemble)")
    two_legs = [x for x in legs if x in precs]
    fields = None
    if two_legs or not args.no_cpu:
        print(f"[bench] building {n} droplet fields", file=sys.stderr, flush=True)
        fields = droplet_fields(n, np.float32)

    def leg(p, f, bytes_per_node, label):
        d = load_domain(p, f)
        if tile:
            d.set_tile(*tile)
        wall, total_ms, fused_ms = time_domain(d, args.steps, args.warmup)
        d.close()
        mlups = nodes * args.steps / wall / 1e6
        fused_avg_ms = fused_ms / args.steps
        achieved = bytes_per_node * nodes / (fused_avg_ms * 1e-3) / 1e9
        print(f"[bench] {label}: {mlups:.1f} MLUPS wall {wall*1e3/args.steps:.3f} ms/step "
              f"events {total_ms/args.steps:.3f} ms/step fused {fused_avg_ms:.3f} ms", file=sys.stderr, flush=True)
        return {"mlups": mlups, "ms_per_step": wall * 1e3 / args.steps, "fused_ms": fused_avg_ms,
                "achieved_gbs": achieved, "step_gbs": bytes_per_node * mlups * 1e6 / 1e9,
                "bytes_per_node": bytes_per_node, "prec": p.precision}

    res = {}
    for name in two_legs:
        res[name] = leg(droplet_params(n, precs[name]), fields, BYTES_PER_LUP[precs[name]], f"2c {name}")
    if "1c" in legs:
        print(f"[bench] building {n} Taylor-Green fields", file=sys.stderr, flush=True)
        res["1c"] = leg(tg_params(n, PREC_FP32), tg_fields(n, np.float32), BYTES_PER_LUP_1C[PREC_FP32], "1c fp32")

    def roofline(r, workload):
        return {"bound": "hbm", "achieved": round(r["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(workload),
                "kernel": "fused_kernel (inline moments)",
                "bytes_per_node": r["bytes_per_node"], "kernel_ms": round(r["fused_ms"], 4)}

    w2c = f"laplace_droplet_2c_d3q27h_densratio_{args.n}^3"
    w1c = f"taylor_green_1c_d3q27h_{args.n}^3"
    head = two_legs[0] if two_legs else "1c"
    main_leg = res[head]
    one_c = head == "1c"
    workload = w1c if one_c else w2c
    out = {
        "metric": METRIC,
        "value": round(main_leg["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_leg["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": (round(main_leg["mlups"] / (A100_1C_D3Q27H_SP_MLUPS if one_c else A100_2C_D3Q27H_SP_MLUPS), 2)
                        if main_leg["prec"] == PREC_FP32 and args.n == 512 else None),
        "dtype": DTYPE[main_leg["prec"]],
        "data": "synthetic",
        "config": ({"workload": w1c, "lattice": "D3Q27 HIGHORDER", "components": 1, "nodes": nodes,
                    "u0": 0.04, "tau": 0.55, "storage": "fp32, row-interleaved field sets",
                    "baseline": "1x A100 1c 512^3 D3Q27h SP 1000 MLUPS (README.md:340)"} if one_c else
                   {"workload": w2c, "lattice": "D3Q27 HIGHORDER", "components": 2, "densratio": 100,
                    "nodes": nodes, "storage": {PREC_FP32: "fp32", PREC_FP16S: "fp16 scaled deviations",
                                                PREC_BF16S: "bf16 scaled deviations"}[main_leg["prec"]]
                    + ", row-interleaved field sets",
                    "baseline": "1x A100 2c 512^3 D3Q27h SP 800 MLUPS (README.md:405)"}),
        "roofline": roofline(main_leg, workload if main_leg["prec"] == PREC_FP32 else workload + "_" + head),
        "hbm_frac_step": round(main_leg["step_gbs"] / HBM_PEAK_GBS, 4),
    }
    for name in ("fp16s", "bf16s"):
        if name in res and "fp32" in res:
            out[name] = {"value": round(res[name]["mlups"], 1),
                         "speedup_vs_fp32": round(res[name]["mlups"] / res["fp32"]["mlups"], 3),
                         "kernel_ms": round(res[name]["fused_ms"], 4),
                         "kernel_gbs_alg": round(res[name]["achieved_gbs"], 1)}
    if "1c" in res and not one_c:
        r = res["1c"]
        out["taylor_green_1c"] = {"workload": w1c, "value": round(r["mlups"], 1), "unit": "MLUPS",
                                  "ms_per_step": round(r["ms_per_step"], 4),
                                  "vs_a100": round(r["mlups"] / A100_1C_D3Q27H_SP_MLUPS, 2) if args.n == 512 else None,
                                  "roofline": roofline(r, w1c)}
    if not args.no_precision and not args.only:
        print("[bench] precision errors vs the FP64 oracle", file=sys.stderr, flush=True)
        pe = precision_errors()
        out["precision_error"] = pe
        for name in ("fp16s", "bf16s"):
            if name in out:
                ch = pe["porous_channel_64x48x32_densratio10"][name]
                out[name]["vel_rel_l2_vs_fp64_oracle"] = ch["vel"][0]
                out[name]["meets_vel_bar"] = ch["meets_vel_bar"]
    if not args.no_ensemble and not args.only:
        print("[bench] batched ensemble", file=sys.stderr, flush=True)
        ens = {}
        for name, prec in (("fp32", PREC_FP32), ("fp16s", PREC_FP16S)):
            ens[name], nodes_e = ensemble(prec, steps=ENSEMBLE_STEPS)
        out["ensemble"] = {"domains": 96, "nodes": nodes_e, "steps": ENSEMBLE_STEPS,
                           "workload": "96 two-component D3Q27h droplets, L in {32,48,64,96,128}", **ens}
    if not args.only:
        print("[bench] Fortran host", file=sys.stderr, flush=True)
        out["fortran_host"] = fortran_host(n)
    if not args.no_cpu:
        print("[bench] CPU baseline (oracle, 512^3 sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(fields, n)
    del fields
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

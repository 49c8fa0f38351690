"""Batches of independent domains (lbf_batch_*): one launch per kernel and
step covers every domain's tiles.

Every node's arithmetic is independent of the tiling, so a domain advanced in
a batch must equal, bit for bit, the same domain advanced on its own (which
the parity tests hold to the oracle).  Also checked: replays of a captured
graph, interleaved with eager work, equal eager steps; odd step counts keep
the buffer bookkeeping; the autotuner leaves every domain's state as it was.
"""
import numpy as np
import pytest

from lbfast_amd import hip, setups
from lbfast_amd.params import make_params, PREC_FP32, PREC_FP16S, HFIELDS, FIELD_PHI, FIELD_UOLD

pytestmark = pytest.mark.gpu


def _two_component(n, seed, prec, sigma=0.005, rho_b=0.02, a=0.05):
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.02, visc2=0.05,
                    rho_r=1.0, rho_b=rho_b, tau_diff=0.1, sigma=sigma, width=5.0, precision=prec)
    L = min(n)
    f = setups.droplet(n, 0.3 * L, 5.0, [n[0] / 2 + 0.5, n[1] / 2 + 0.5, n[2] / 2 + 0.5], sigma, 1.0, rho_b,
                       noise=1e-5, seed=seed, ellipse_a=a)
    return p, f


def _domains(specs, prec):
    out = []
    for i, n in enumerate(specs):
        p, f = _two_component(n, 11 + i, prec, sigma=0.005 * (1 + 0.2 * i), rho_b=0.02 * (1 + i), a=0.02 + 0.02 * i)
        out.append((p, f))
    return out


def _state(d, p):
    fields = [d.download(f) for f in HFIELDS]
    if p.twocomponent:
        fields.append(d.download(FIELD_PHI))
        fields += [d.download(f) for f in FIELD_UOLD]
    return np.stack(fields)


SPECS = [(32, 32, 32), (48, 40, 36), (64, 24, 20), (20, 28, 44)]


@pytest.mark.parametrize("prec", [PREC_FP32, PREC_FP16S])
def test_batch_equals_single_domains(prec):
    print(f"[batch] equality prec={prec} {SPECS}", flush=True)
    doms = _domains(SPECS, prec)
    steps = 7
    want = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        d.step(steps)
        want.append(_state(d, p))
        d.close()
    ds = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        ds.append(d)
    b = hip.Batch(ds)
    b.step(3)
    b.step(4)  # odd then even counts
    b.synchronize()
    for (p, _), d, w in zip(doms, ds, want):
        got = _state(d, p)
        assert np.array_equal(got, w), float(np.abs(got - w).max())
    b.close()
    for d in ds:
        d.close()


def test_batch_one_component_and_tiles():
    print("[batch] one component, tile 8/16", flush=True)
    specs = [(32, 24, 20), (40, 40, 40), (24, 64, 16)]
    doms = []
    for i, n in enumerate(specs):
        p = make_params(n, lattice=27, highorder=True, visc1=0.02 + 0.01 * i, precision=PREC_FP32,
                        fx=1e-6 * i)
        doms.append((p, setups.taylor_green(n, 0.03)))
    want = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        d.step(6)
        want.append(_state(d, p))
        d.close()
    ds = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        ds.append(d)
    b = hip.Batch(ds)
    b.set_tile(8, 16)
    b.step(6)
    b.synchronize()
    for (p, _), d, w in zip(doms, ds, want):
        assert np.array_equal(_state(d, p), w)
    b.close()
    for d in ds:
        d.close()


def test_batch_graph_replays_equal_eager_steps():
    print("[batch] graph replays vs eager", flush=True)
    doms = _domains(SPECS[:3], PREC_FP32)
    runs = []
    for use_graph in (False, True):
        ds = []
        for p, f in doms:
            d = hip.Domain(p)
            d.load(f)
            ds.append(d)
        b = hip.Batch(ds)
        if use_graph:
            b.capture(2)
        snaps = []
        for rep in range(3):  # several replays with eager host work between them
            b.step(2)
            b.synchronize()
            snaps.append([_state(d, p) for (p, _), d in zip(doms, ds)])
        runs.append(snaps)
        b.close()
        for d in ds:
            d.close()
    for eager, graph in zip(runs[0], runs[1]):
        for a, g in zip(eager, graph):
            assert np.array_equal(a, g)


def test_batch_autotune_restores_state():
    print("[batch] autotune", flush=True)
    doms = _domains(SPECS[:2], PREC_FP32)
    results = []
    for tune in (False, True):
        ds = []
        for p, f in doms:
            d = hip.Domain(p)
            d.load(f)
            ds.append(d)
        b = hip.Batch(ds)
        if tune:
            ty, tz = b.autotune(2)
            assert ty in (8, 13) and tz in (16, 32, 64, 128)
            ty2, tzs, tzps, order = b.tiles()
            assert ty2 == ty and sorted(order) == list(range(len(ds)))
            for (p, _), z in zip(doms, tzs):
                assert 1 <= z <= p.nz
        b.step(4)
        b.synchronize()
        results.append([_state(d, p) for (p, _), d in zip(doms, ds)])
        assert all(d.L.lbf_get_step(d.h) == 4 for d in ds)
        b.close()
        for d in ds:
            d.close()
    for a, g in zip(results[0], results[1]):
        assert np.array_equal(a, g)


def test_batch_rejects_mixed_configurations():
    p1, f1 = _two_component((16, 16, 16), 1, PREC_FP32)
    p2 = make_params((16, 16, 16), lattice=19, highorder=False, precision=PREC_FP32)
    d1, d2 = hip.Domain(p1), hip.Domain(p2)
    d1.load(f1)
    d2.load(setups.taylor_green((16, 16, 16), 0.02))
    with pytest.raises(hip.LbfError):
        hip.Batch([d1, d2])
    d1.close()
    d2.close()


@pytest.mark.parametrize("prec", [PREC_FP32, PREC_FP16S])
def test_batch_interleaved_with_single_domain_steps(prec):
    # A member stepped on its own between batch calls swaps its buffers and
    # advances its step count; the next batch call rebuilds its tables (and
    # drops a captured graph) from the contexts' current state.
    print(f"[batch] interleaved single-domain steps prec={prec}", flush=True)
    doms = _domains(SPECS[:3], prec)
    want = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        d.step(6)
        want.append(_state(d, p))
        d.close()
    ds = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        ds.append(d)
    b = hip.Batch(ds)
    b.capture(2)
    b.step(3)
    ds[1].step(1)            # one member alone: odd, so its buffers swap
    b.step(2)                # the others have 5 steps, member 1 has 6
    ds[0].step(1)
    ds[2].step(1)            # now all have 6
    b.synchronize()
    for (p, _), d, w in zip(doms, ds, want):
        assert d.step_count == 6
        got = _state(d, p)
        assert np.array_equal(got, w), float(np.abs(got - w).max())
    b.close()
    for d in ds:
        d.close()


def test_mixed_size_batch_against_oracle():
    # Domains of three shapes advanced together, each compared with the FP32
    # oracle (a CPU port of the reference stages) of its own problem: the
    # batch is held to the reference's arithmetic, not only to single-domain
    # runs.  Bar: the FP32 parity bar of the single-domain tests -- rho_phi,
    # u and phi within 1e-5 relative L2 of the same-precision oracle.
    import oracle
    from tests.cases import rel_l2
    specs = [(32, 32, 32), (48, 48, 48), (40, 24, 32)]
    steps = 12
    print(f"[batch] mixed sizes vs oracle {specs} steps={steps}", flush=True)
    doms = _domains(specs, PREC_FP32)
    ds = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        ds.append(d)
    b = hip.Batch(ds)
    b.step(steps)
    b.synchronize()
    for (p, f), d in zip(doms, ds):
        o = oracle.Oracle(p)
        o.load(f)
        o.step(steps)
        phi_g, phi_o = d.download(FIELD_PHI), o.get_field(FIELD_PHI)
        err = {"u": rel_l2(np.stack([d.download(i) for i in (2, 3, 4)]), np.stack([o.get_field(i) for i in (2, 3, 4)])),
               "phi": rel_l2(phi_g, phi_o),
               "rho": rel_l2(p.rho_r * phi_g + (1 - phi_g) * p.rho_b, p.rho_r * phi_o + (1 - phi_o) * p.rho_b),
               "uold": rel_l2(np.stack([d.download(i) for i in FIELD_UOLD]),
                              np.stack([o.get_field(i) for i in FIELD_UOLD]))}
        print(f"[batch]   {p.nx}x{p.ny}x{p.nz}: {err}", flush=True)
        assert max(err.values()) < 1e-5, err
    b.close()
    for d in ds:
        d.close()


def test_batch_order_and_class_tiles_leave_results_unchanged():
    # The table order and per-domain z chunks change only which block
    # computes which node: every domain still equals its single-domain run.
    print("[batch] order and per-domain chunks", flush=True)
    doms = _domains(SPECS, PREC_FP32)
    want = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        d.step(5)
        want.append(_state(d, p))
        d.close()
    ds = []
    for p, f in doms:
        d = hip.Domain(p)
        d.load(f)
        ds.append(d)
    b = hip.Batch(ds)
    b.set_order([2, 0, 3, 1])
    assert b.tiles()[3] == [2, 0, 3, 1]
    b.step(2)
    b.set_order([3, 2, 1, 0])
    b.step(3)
    b.synchronize()
    for (p, _), d, w in zip(doms, ds, want):
        assert np.array_equal(_state(d, p), w)
    with pytest.raises(hip.LbfError):
        b.set_order([0, 0, 1, 2])
    b.close()
    for d in ds:
        d.close()

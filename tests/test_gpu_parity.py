"""GPU parity: the HIP time step against the oracle on the same inputs.

Bars: FP64 runs agree with the FP64 oracle to 1e-11 relative L2 per field
(fluid nodes; phi on all nodes, ghost values included); FP32 runs agree with
the FP32 oracle to 1e-5 (the FP32 tolerance stated for this path).  Both the
fused fast path (lbf_step) and the one-kernel-per-stage path (lbf_stage)
are checked.
"""
import sys

import numpy as np
import pytest

import oracle
from lbfast_amd import hip
from lbfast_amd.params import PREC_FP64, PREC_FP32, PREC_MIXED
from tests.cases import case, prepare, state_of, compare

pytestmark = pytest.mark.gpu

TOL = {PREC_FP64: 1e-11, PREC_FP32: 1e-5, PREC_MIXED: 1e-5}


def run_pair(name, prec, n, steps, path="fast", tile=None):
    p, f, raw = case(name, n, prec)
    f, isf = prepare(p, f, raw, hip.classify_isfluid)
    print(f"[parity] {name} prec={prec} n={n} steps={steps} path={path} tile={tile}", flush=True)
    o = oracle.Oracle(p)
    o.load(f, isf)
    o.step(steps)
    d = hip.Domain(p)
    if tile:
        d.set_tile(*tile)
    d.load(f, isf)
    if path == "fast":
        d.step(steps)
    else:
        d.step_staged(steps)
    d.synchronize()
    res = compare(state_of(d, p), state_of(o, p), isf)
    d.close()
    worst = max(res.values())
    print(f"[parity]   worst rel-L2 {worst:.3e}  {res}", flush=True)
    return res


@pytest.mark.parametrize("name", ["tg19", "tg27", "tg27h", "tg19_force", "tg27h_force"])
@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32])
def test_one_component_fast(name, prec):
    res = run_pair(name, prec, (32, 24, 20), 12)
    assert max(res.values()) < TOL[prec], res


@pytest.mark.parametrize("name", ["drop27h", "drop19", "drop27h_nodens"])
@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32])
def test_two_component_fast(name, prec):
    res = run_pair(name, prec, (32, 32, 32), 10)
    assert max(res.values()) < TOL[prec], res


@pytest.mark.parametrize("name", ["chan27h", "chan19", "pois27h_1c"])
@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32])
def test_obstacles_fast(name, prec):
    res = run_pair(name, prec, (40, 24, 32), 10)
    assert max(res.values()) < TOL[prec], res


@pytest.mark.parametrize("name", ["tg27h_force", "drop27h", "chan27h"])
@pytest.mark.parametrize("prec", [PREC_FP64, PREC_FP32, PREC_MIXED])
def test_stage_path(name, prec):
    res = run_pair(name, prec, (32, 24, 32), 6, path="staged")
    assert max(res.values()) < TOL[prec], res


@pytest.mark.parametrize("n", [(64, 16, 16), (96, 20, 10), (16, 40, 12), (130, 9, 7)])
def test_tiling_edges(n):
    """Partial x tiles (nx not a multiple of 64), ny not a multiple of the
    row tile, rows shorter than a wave."""
    res = run_pair("tg27h_force", PREC_FP64, n, 5)
    assert max(res.values()) < 1e-11, res


@pytest.mark.parametrize("tile", [(8, 3), (13, 5), (13, 64)])
def test_tile_shapes(tile):
    res = run_pair("drop27h", PREC_FP64, (32, 32, 32), 4, tile=tile)
    assert max(res.values()) < 1e-11, res

"""lbf_reduce_fp64_noalloc: the FP64 reductions through the library's own
device array instead of the context's buffers.  Called through the C-ABI
twice in a row, the second time on a smaller domain, each result must equal
lbf_reduce_fp64 on the same state bit for bit (the same kernels over the same
node-to-thread assignment; the bar for this integer-exact comparison of
floating-point sums is identity)."""
import numpy as np
import pytest

from lbfast_amd import hip, setups
from lbfast_amd.params import make_params, PREC_FP32, PREC_FP64

pytestmark = pytest.mark.gpu


def _droplet(n, prec):
    p = make_params(n, lattice=27, highorder=True, twocomponent=True, densratio=True, visc1=0.1, visc2=0.1,
                    rho_r=1.0, rho_b=0.1, tau_diff=0.1, sigma=0.01, width=3.0, precision=prec)
    f = setups.droplet(n, n[0] / 4, 3.0, [n[0] / 2 + 0.5, n[1] / 2 + 0.5, n[2] / 2 + 0.5], 0.01, 1.0, 0.1,
                       noise=1e-4, seed=5)
    return p, f


def test_noalloc_reduction_equals_reduce_fp64_twice():
    cases = [((72, 60, 80), PREC_FP32, 3),   # 345,600 nodes: every one of the 1024 records in use
             ((20, 12, 9), PREC_FP64, 2)]    # 2,160 nodes: 9 records (smaller input second)
    for n, prec, steps in cases:
        print(f"[noalloc] n={n} prec={prec} steps={steps}", flush=True)
        p, f = _droplet(n, prec)
        d = hip.Domain(p)
        d.load(f)
        d.step(steps)
        want = d.reduce()
        got = d.reduce_noalloc()
        again = d.reduce()
        d.close()
        print(f"[noalloc]   sums {got}", flush=True)
        assert got[7] == float(n[0] * n[1] * n[2])   # fluid node count
        assert np.array_equal(got, want), (got, want)
        assert np.array_equal(again, want)            # the context's own path is untouched

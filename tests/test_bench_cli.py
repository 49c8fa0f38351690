"""bench.py's command line and the inputs of its roofline object, on the CPU:
it refuses more than one GPU with a message before touching a device,
rejects unknown legs, and finds the counter-measured traffic of every
workload it reports in profiles/hbm_traffic.txt."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=300)


def test_more_than_one_gpu_exits_with_a_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and not r.stdout.strip()


def test_unknown_leg_is_rejected():
    r = _run("--only", "fp8")
    assert r.returncode == 2 and "unknown leg" in r.stderr


def test_traffic_of_the_reported_workloads_is_on_file():
    for w, per_node in (("laplace_droplet_2c_d3q27h_densratio_512^3", bench.BYTES_PER_LUP),
                        ("taylor_green_1c_d3q27h_512^3", bench.BYTES_PER_LUP_1C)):
        t = bench.measured_traffic(w)
        assert t is not None, w
        # counted HBM bytes cover at least the compulsory state traffic and
        # stay within twice it (halo re-reads, the phase prep's output)
        alg = per_node[bench.PREC_FP32] * 512 ** 3
        assert alg <= t < 2 * alg, (w, t, alg)
    assert bench.measured_traffic("no_such_workload") is None

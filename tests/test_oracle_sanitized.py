"""The FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

oracle/Makefile builds oracle/_san/liblbm_oracle.so with -fsanitize=address,undefined
(host code, serial); a child Python process with the sanitizer runtimes preloaded loads
it through LBM_ORACLE_SO and runs every small case, every thermal case, both wall
distance methods and the faithful loop structure.  Any out-of-bounds access, use after
free or undefined behaviour aborts the child with the sanitizer's report, which the test
shows.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN_SO = os.path.join(ROOT, "oracle", "_san", "liblbm_oracle.so")

CHILD = r"""
import sys
root = sys.argv[1]
sys.path[:0] = [root, root + '/tests', root + '/lbm-2d_amd']
import numpy as np
import oracle as O
assert O.ORACLE_SO.endswith('_san/liblbm_oracle.so'), O.ORACLE_SO
from cases import CASES, NARROW_CASES, SPEC_CASES, THERMAL_CASES, build_oracle, build_thermal_oracle, thermal_step
for name, c in sorted({**CASES, **SPEC_CASES, **NARROW_CASES}.items()):
    o = build_oracle(c)
    o.step(c.tau, c.u, min(c.steps, 30))
    assert np.isfinite(o.f).all(), name
    o.total_force(); o.mass()
    print('case', name, flush=True)
for name, tc in sorted(THERMAL_CASES.items()):
    o = build_thermal_oracle(tc)
    for _ in range(10):
        thermal_step(o, tc)
    assert np.isfinite(o.g).all(), name
    print('thermal', name, flush=True)
m = (np.random.default_rng(1).random((37, 53)) < 0.2).astype(np.uint8)
a, b = O.wall_distance(m, 'bfs'), O.wall_distance(m, 'sweep')
assert (a == b).all()
f = O.Oracle(40, 20, O.CYLINDER, les=True, cs=0.1, faithful=True)
f.place_cylinder(10, 10, 4)
f.init_equilibrium(1.0, 0.1, 0.0)
f.step(0.52, 0.1, 5)
print('faithful LES', flush=True)
print('SANITIZED_OK')
"""


def _runtime(name):
    out = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(not _runtime("libasan.so") or not _runtime("libubsan.so"), reason="no sanitizer runtimes")
def test_oracle_runs_clean_under_asan_and_ubsan():
    out = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "_san/liblbm_oracle.so"],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr[-3000:]
    env = dict(os.environ)
    env["LD_PRELOAD"] = _runtime("libasan.so") + ":" + _runtime("libubsan.so")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["LBM_ORACLE_SO"] = SAN_SO
    run = subprocess.run([sys.executable, "-c", CHILD, ROOT], capture_output=True,
                         text=True, env=env, timeout=900)
    assert run.returncode == 0 and "SANITIZED_OK" in run.stdout, (run.stdout[-2000:], run.stderr[-6000:])

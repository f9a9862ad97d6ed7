"""The reference's own case drivers, unchanged, on the GPU.

oracle/Makefile builds src/main.cpp, src/cavity.cpp, src/flat_plate.cpp and
src/step.cpp of the reference twice into oracle/_ref/drivers/: against the reference's
headers (<name>_cpu, the CPU solver) and against lbm-2d_amd/include (<name>_gpu: the
same source, its execute_time_step stepping on the GPU through liblbm2d.so).  Each pair
runs with the same arguments in its own directory and every file the drivers write is
compared: forces.jsonl (Cd, Cl per step at the writer's six decimals), meta.json (text),
the JSON frames and VTK frames (every value), the cavity's centerline.csv (six
significant digits).  With FP64 storage the GPU step equals the reference's to ~1e-12,
so printed values agree to one unit of their last digit; with LBM2D_STORAGE=f32 the
drivers run the FP32 form and Cd / Cl agree to 1e-4.

The binaries are built from the reference sources where those are present; a machine
without them runs the binaries built beforehand.
"""
import json
import math
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "drivers")

# driver -> (arguments, output directory it writes); >= 50 steps: the drivers save a
# frame every steps / 50 steps (src/main.cpp:40)
RUNS = {
    "main": (["100", "100", "--nx", "96", "--ny", "48"], "output/cylinder/re100"),
    "cavity": (["100", "64", "100"], "output/cavity/re100"),
    "flat_plate": (["1000", "10", "100"], "output/flatplate/re1000_aoa10"),
    "step": (["100", "100"], "output/step/re100"),
}

NUM = re.compile(r"-?(?:nan|inf|\d+\.?\d*(?:e[-+]?\d+)?)", re.I)


def _run(exe, args, cwd, env=None):
    e = dict(os.environ)
    e.setdefault("OMP_NUM_THREADS", "8")
    if env:
        e.update(env)
    out = subprocess.run([exe] + args, cwd=cwd, capture_output=True, text=True, timeout=300, env=e)
    assert out.returncode == 0, (exe, out.stdout[-2000:], out.stderr[-2000:])
    return out.stdout


def _numbers(text):
    return [float(t) for t in NUM.findall(text)]


def _close(a, b, atol, rtol=0.0):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    return abs(a - b) <= atol + rtol * abs(b)


def _forces(path):
    return [json.loads(line) for line in open(path)]


def _compare_tree(cpu_dir, gpu_dir, tol_fixed, tol_sig):
    """Every file under the CPU run's output exists in the GPU run's; numbers agree."""
    n_files = 0
    for dp, _, files in os.walk(cpu_dir):
        for fn in files:
            a = os.path.join(dp, fn)
            b = os.path.join(gpu_dir, os.path.relpath(a, cpu_dir))
            assert os.path.exists(b), b
            ta, tb = open(a).read(), open(b).read()
            if fn == "meta.json":
                assert ta == tb
            elif fn == "forces.jsonl":
                fa, fb = _forces(a), _forces(b)
                assert [r["step"] for r in fa] == [r["step"] for r in fb]
                for ra, rb in zip(fa, fb):
                    assert _close(ra["cd"], rb["cd"], tol_fixed) and _close(ra["cl"], rb["cl"], tol_fixed), (ra, rb)
            else:
                na, nb = _numbers(ta), _numbers(tb)
                assert len(na) == len(nb), a
                atol = tol_sig if fn.endswith(".csv") or fn.endswith(".vtk") else tol_fixed
                rtol = tol_sig if fn.endswith(".csv") or fn.endswith(".vtk") else 0.0
                bad = [(x, y) for x, y in zip(na, nb) if not _close(x, y, atol * 1e-3 if rtol else atol, rtol)]
                assert not bad, (a, bad[:5])
            n_files += 1
    return n_files


@pytest.mark.skipif(not os.path.isdir(BIN), reason="oracle/_ref/drivers not built (no reference sources)")
@pytest.mark.parametrize("driver", sorted(RUNS))
def test_reference_driver_runs_unchanged_on_the_gpu(driver, tmp_path):
    args, sub = RUNS[driver]
    for kind in ("cpu", "gpu"):
        (tmp_path / kind).mkdir()
        out = _run(os.path.join(BIN, f"{driver}_{kind}"), args, tmp_path / kind)
        (tmp_path / f"{kind}.out").write_text(out)
    n = _compare_tree(tmp_path / "cpu" / sub, tmp_path / "gpu" / sub, tol_fixed=1.01e-6, tol_sig=1.01e-5)
    print(f"\n{driver}: {n} output files agree")
    assert n >= 3
    # the summaries the drivers print agree as well (values at 4-6 digits)
    for key in ("Mean Cd", "Cl amplitude", "Mean Cl"):
        a = re.search(key + r"\s*=\s*(\S+)", (tmp_path / "cpu.out").read_text())
        b = re.search(key + r"\s*=\s*(\S+)", (tmp_path / "gpu.out").read_text())
        assert (a is None) == (b is None)
        if a:
            assert _close(float(a.group(1)), float(b.group(1)), 1.01e-4), (key, a.group(1), b.group(1))


@pytest.mark.skipif(not os.path.isdir(BIN), reason="oracle/_ref/drivers not built (no reference sources)")
@pytest.mark.parametrize("driver", ["main", "flat_plate"])
def test_reference_driver_in_fp32_storage(driver, tmp_path):
    args, sub = RUNS[driver]
    for kind, env in (("cpu", None), ("gpu", {"LBM2D_STORAGE": "f32"})):
        (tmp_path / kind).mkdir()
        _run(os.path.join(BIN, f"{driver}_{kind}"), args, tmp_path / kind, env)
    fa = _forces(tmp_path / "cpu" / sub / "forces.jsonl")
    fb = _forces(tmp_path / "gpu" / sub / "forces.jsonl")
    assert len(fa) == len(fb)
    worst = max(max(abs(ra["cd"] - rb["cd"]), abs(ra["cl"] - rb["cl"])) for ra, rb in zip(fa, fb))
    print(f"\n{driver} FP32: worst |dCd|, |dCl| = {worst:.2e}")
    assert worst <= 1e-4

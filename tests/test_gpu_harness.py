"""GPU runs of a program that drives the reference's harness entry points the
way integralImage/main.cpp does (argv[1] picks TestBlockScan, TestSerielScan
or TestIncreamentScan; the rest is GetArgs' <scan> <dtype in> <dtype out>
<n>), linked against libsat_amd.so only, plus the header templates (Test<>,
SerielScan::TestX) on non-square sizes."""
from __future__ import annotations

import re
import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
PROG = ROOT / "tests" / "cpp" / "_build" / "harness_main"


def say(*a):
    print(*a, file=sys.stderr, flush=True)


# (argv[1..5], expected Test<> runs): the reference's dtype codes (integralImage.h:22-32)
CASES = [
    (["0", "0", "0", "4", "2"], 2),   # ScanRow-BRLT, KS, u8 -> u32, 1024^2 and 2048^2
    (["0", "1", "6", "6", "1"], 1),   # ScanRow-BRLT, LF, f32 -> f32
    (["1", "0", "1", "5", "2"], 2),   # BRLT-ScanRow, KS, i8 -> i32
    (["1", "1", "7", "7", "1"], 1),   # BRLT-ScanRow, LF, f64 -> f64
    (["2", "1", "0", "7", "1"], 1),   # ScanRowColumn, LF, u8 -> f64
    (["2", "0", "5", "5", "2"], 2),   # ScanRowColumn, KS, i32 -> i32
    (["2", "0", "6", "4", "1"], 0),   # f32 -> u32: not in the reference's dispatch, runs nothing
]


@pytest.mark.parametrize("args,runs", CASES)
def test_reference_dispatch(args, runs):
    say("[case] harness_main " + " ".join(args))
    r = subprocess.run([str(PROG), *args], capture_output=True, text=True, timeout=180)
    out = r.stdout
    assert r.returncode == 0, out[-3000:] + r.stderr[-2000:]
    lines = re.findall(r"^(\d+), (\d+), total time = ([0-9.]+), ([0-9.]+)$", out, flags=re.M)
    assert len(lines) == runs, out[-3000:]
    for i, (w, h, _, ev) in enumerate(lines):
        assert int(w) == int(h) == 1024 * (i + 1)
        assert float(ev) > 0
    assert out.count("compare = successed") == runs and "compare = failed" not in out, out[-3000:]


def test_harness_templates_direct():
    say("[case] harness_main all")
    r = subprocess.run([str(PROG), "all"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0 and "harness templates: 0 failures" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]
    assert r.stdout.count("compare = successed") == 5

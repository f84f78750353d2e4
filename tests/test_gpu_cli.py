"""GPU runs of the native `sat` CLI (the reference's `sat <alg> <scan> <dt0>
<dt1> <n>` protocol) and of the C++ drop-in program (the reference's harness
pattern over include/sat/sat.hpp and gpu_data.hpp)."""
from __future__ import annotations

import subprocess
import sys
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "sat_amd" / "bin" / "sat"
DROPIN = ROOT / "tests" / "cpp" / "_build" / "dropin_test"


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.mark.parametrize("args", [["3", "0", "0", "4", "2"], ["0", "1", "6", "6", "1"], ["1", "0", "5", "5", "1"],
                                  ["2", "0", "0", "7", "1"]])
def test_cli_runs_and_checks(args, tmp_path):
    say("[case] sat " + " ".join(args))
    csv = tmp_path / "sweep.csv"
    r = subprocess.run([str(CLI), *args, "--check", "--csv", str(csv)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = int(args[4])
    assert r.stdout.count("compare = successed") == n
    assert r.stdout.count("total time = ") == n
    rows = csv.read_text().strip().splitlines()
    assert len(rows) == n and rows[0].startswith(",".join(args[:4]) + ",1024,1024,")


def test_cpp_dropin_program():
    say("[case] tests/cpp dropin_test")
    r = subprocess.run([str(DROPIN)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "dropin ok" in r.stdout, r.stdout[-3000:] + r.stderr[-2000:]


def test_sweep_harness(tmp_path):
    say("[case] sweep harness, 2 algorithms x 2 pairs x 2 sizes")
    from sat_amd import sweep

    rows = sweep.sweep(algs=(0, 3), pairs=[(0, 4), (6, 6)], n=2, csv=tmp_path / "s.csv", check=True,
                       log=lambda *m: say(*m))
    assert rows[0] == sweep.HEADER and len(rows) == 1 + 2 * 2 * 2
    for r in rows[1:]:
        f = r.split(",")
        assert int(f[4]) in (1024, 2048) and float(f[6]) > 0 and float(f[7]) > 0

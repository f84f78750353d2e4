"""CPU checks of the native `sat` CLI and the C++ drop-in test program: they
are built, print usage, and reject bad arguments before touching a device."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
CLI = ROOT / "sat_amd" / "bin" / "sat"
DROPIN = ROOT / "tests" / "cpp" / "_build" / "dropin_test"


def test_cli_built_and_usage():
    assert CLI.exists(), "build with python -m sat_amd.build"
    r = subprocess.run([str(CLI), "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage: sat <alg" in r.stderr


def test_cli_rejects_bad_arguments():
    # u16 -> f32 is not one of the dispatched pairs (ScanRow-BRLT.cu:274-376)
    r = subprocess.run([str(CLI), "0", "0", "2", "6", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "not supported" in r.stderr
    r = subprocess.run([str(CLI), "7", "0", "6", "6", "1"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2


def test_dropin_program_built():
    assert DROPIN.exists(), "build with python -m sat_amd.build (builds tests/cpp too)"

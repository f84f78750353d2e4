"""CPU checks of the tile-configuration table (no device is touched: the
table lives in host memory of the library)."""
from __future__ import annotations

import ctypes as C
import json
import math
from pathlib import Path

import numpy as np
import pytest


@pytest.fixture(scope="module")
def L():
    import sat_amd

    L = sat_amd.lib()
    yield L
    from sat_amd.tune import reset_table

    reset_table(L)


def py_class(w, h, p):
    tx, ty = math.ceil(w / 256), math.ceil(h / 64)
    cp = 0 if p <= 1 else 1 if p < 4 else 2 if p < 16 else 3
    return min(int(math.floor(math.log2(tx))), 8), min(int(math.floor(math.log2(ty))), 11), cp


def test_class_formula(L):
    rng = np.random.default_rng(5)
    for _ in range(500):
        w, h, p = int(rng.integers(1, 1 << 17)), int(rng.integers(1, 1 << 17)), int(rng.integers(1, 64))
        cx, cy, cp = C.c_int(), C.c_int(), C.c_int()
        L.sat_tune_class(w, h, p, C.byref(cx), C.byref(cy), C.byref(cp))
        assert (cx.value, cy.value, cp.value) == py_class(w, h, p), (w, h, p)


def test_tuned_table_loaded():
    """The package applies sat_amd/tuned_gfx950.json when it loads the library
    (checked in a fresh process: other tests reset the table)."""
    import subprocess
    import sys

    code = (
        "import json, sat_amd\n"
        "L = sat_amd.lib()\n"
        "doc = json.loads(sat_amd.TUNED_TABLE.read_text())\n"
        "assert doc['entries']\n"
        "for e in doc['entries']:\n"
        "    assert L.sat_tune_get(e['in_bytes'], e['out_bytes'], e['class_x'], e['class_y'], e['class_p']) == e['cfg'], e\n"
        "for e in doc['batch']:\n"
        "    assert L.sat_batch_tune_get(e['out_bytes']) == e['cfg'], e\n"
        "print('table ok')\n")
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "table ok" in r.stdout, r.stderr[-2000:]


def test_library_starts_from_shipped_table():
    """The library itself (loaded without the Python package) starts from the
    shipped table compiled into it, so C and C++ callers get the tuned tiles."""
    import subprocess
    import sys

    code = (
        "import ctypes as C, json, pathlib\n"
        "root = pathlib.Path('.').resolve()\n"
        "L = C.CDLL(str(root / 'sat_amd' / 'libsat_amd.so'))\n"
        "doc = json.loads((root / 'sat_amd' / 'tuned_gfx950.json').read_text())\n"
        "for e in doc['entries']:\n"
        "    assert L.sat_tune_get(e['in_bytes'], e['out_bytes'], e['class_x'], e['class_y'], e['class_p']) == e['cfg'], e\n"
        "for e in doc['batch']:\n"
        "    assert L.sat_batch_tune_get(e['out_bytes']) == e['cfg'], e\n"
        "assert L.sat_tune_set(4, 4, 5, 7, 0, -1) == 0 and L.sat_tune_get(4, 4, 5, 7, 0) == 0\n"
        "print('compiled table ok')\n")
    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "compiled table ok" in r.stdout, r.stderr[-2000:]


def test_tuned_table_include_is_current():
    """sat_amd/csrc/tuned_table.inc is what the build generates from the JSON."""
    from sat_amd import build as b

    before = b.TUNED_INC.read_text()
    b.write_tuned_table()
    assert b.TUNED_INC.read_text() == before


def test_builtin_choice(L):
    from sat_amd.tune import reset_table

    reset_table(L)
    assert L.sat_tile_config_count() == 8
    # narrow and tall: 8-wave double-height tiles; square and wide: 4-wave tiles
    assert L.sat_tune_pick(4, 4, 256, 65536, 1) == 1
    assert L.sat_tune_pick(4, 4, 8192, 8192, 1) == 0
    assert L.sat_tune_pick(1, 4, 65536, 256, 1) == 0
    assert L.sat_tune_pick(8, 8, 4096, 4096, 1) == 0
    assert L.sat_batch_tune_get(4) == 0


def test_set_get_roundtrip(L):
    from sat_amd.tune import reset_table

    reset_table(L)
    assert L.sat_tune_set(4, 4, 5, 7, 0, 2) == 0
    assert L.sat_tune_get(4, 4, 5, 7, 0) == 2
    assert L.sat_tune_pick(4, 4, 8192, 8192, 1) == 2  # class (5, 7, 0)
    assert L.sat_tune_pick(4, 4, 8192, 8192, 2) == 0  # other plane class untouched
    assert L.sat_tune_pick(8, 8, 8192, 8192, 1) == 0  # other output width untouched
    # an unset 1-byte-input class falls back to the 4-byte-input entry ...
    assert L.sat_tune_pick(1, 4, 8192, 8192, 1) == 2
    # ... until it has an entry of its own
    assert L.sat_tune_set(1, 4, 5, 7, 0, 4) == 0
    assert L.sat_tune_pick(1, 4, 8192, 8192, 1) == 4 and L.sat_tune_pick(4, 4, 8192, 8192, 1) == 2
    assert L.sat_tune_set(1, 4, 5, 7, 0, -1) == 0
    assert L.sat_tune_set(4, 4, 5, 7, 0, -1) == 0
    assert L.sat_tune_get(4, 4, 5, 7, 0) == 0
    assert L.sat_batch_tune_set(8, 1) == 0 and L.sat_batch_tune_get(8) == 1
    assert L.sat_batch_tune_set(8, -1) == 0 and L.sat_batch_tune_get(8) == 0


def test_invalid_arguments(L):
    assert L.sat_tune_set(4, 2, 0, 0, 0, 0) != 0
    assert L.sat_tune_set(2, 4, 0, 0, 0, 0) != 0
    assert L.sat_tune_set(8, 4, 0, 0, 0, 0) != 0  # input wider than output
    assert L.sat_tune_set(4, 4, 9, 0, 0, 0) != 0
    assert L.sat_tune_set(4, 4, 0, 12, 0, 0) != 0
    assert L.sat_tune_set(4, 4, 0, 0, 4, 0) != 0
    assert L.sat_tune_set(4, 4, 0, 0, 0, 8) != 0
    assert L.sat_tune_set(4, 4, -1, 0, 0, 0) != 0
    assert L.sat_tune_get(4, 16, 0, 0, 0) < 0
    assert L.sat_tune_pick(8, 4, 64, 64, 1) < 0
    assert L.sat_batch_tune_set(4, 7) != 0
    assert L.sat_batch_tune_set(4, 5) != 0  # the packed configuration is one-pass only


def test_workspace_fits_every_config(L):
    # workspace queries do not depend on the table (configuration 5's packed
    # tiles of 1-byte inputs included)
    b8 = L.sat_integral_workspace_bytes(3, 0, 4, 1000, 70000)
    for cfg in range(8):
        L.sat_tune_set(1, 4, *py_class(1000, 70000, 1), cfg)
        assert L.sat_integral_workspace_bytes(3, 0, 4, 1000, 70000) == b8
    L.sat_tune_set(1, 4, *py_class(1000, 70000, 1), -1)
    before = L.sat_integral_workspace_bytes(3, 6, 6, 1000, 70000)
    for cfg in range(8):
        L.sat_tune_set(4, 4, *py_class(1000, 70000, 1), cfg)
        assert L.sat_integral_workspace_bytes(3, 6, 6, 1000, 70000) == before
    L.sat_tune_set(4, 4, *py_class(1000, 70000, 1), -1)


def test_apply_table(L, tmp_path):
    from sat_amd.tune import apply_table

    doc = {"entries": [{"in_bytes": 1, "out_bytes": 4, "class_x": 2, "class_y": 10, "class_p": 0, "cfg": 3},
                       {"in_bytes": 8, "out_bytes": 8, "class_x": 0, "class_y": 0, "class_p": 3, "cfg": 1}],
           "batch": [{"out_bytes": 4, "cfg": 2}]}
    p = tmp_path / "t.json"
    p.write_text(json.dumps(doc))
    assert apply_table(p) == 3
    assert L.sat_tune_get(1, 4, 2, 10, 0) == 3 and L.sat_tune_get(8, 8, 0, 0, 3) == 1
    assert L.sat_batch_tune_get(4) == 2
    L.sat_tune_set(1, 4, 2, 10, 0, -1)
    L.sat_tune_set(8, 8, 0, 0, 3, -1)
    L.sat_batch_tune_set(4, -1)

"""The oracle reproduces the golden fixtures of tests/golden (brute-force and
closed-form values computed independently of it, see make_golden.py)."""
from __future__ import annotations

from pathlib import Path

import numpy as np
import pytest

G = np.load(Path(__file__).resolve().parent / "golden" / "golden.npz")  # allow_pickle defaults to False

SAT_CASES = ["u8_u32", "i8_i32", "i32_i32", "u32_u32", "f64_f64", "f32_f32", "u8_f64"]


@pytest.mark.parametrize("case", SAT_CASES)
def test_oracle_sat_fixtures(O, case):
    x, exp = G[f"sat_{case}_in"], G[f"sat_{case}_out"]
    for f in (O.serial, O.parallel):
        got = f(x, exp.dtype)
        assert got.dtype == exp.dtype and np.array_equal(got, exp), case


def test_oracle_pass_fixtures(O):
    x = G["rows_t_in"]
    assert np.array_equal(O.transpose(O.allrow(x, np.uint32)), G["rows_t_out"])
    assert np.array_equal(O.colscan(G["cols_in"], np.uint32), G["cols_out"])


def test_oracle_reverse_box_ones_fold(O):
    assert np.array_equal(O.reverse(G["reverse_in"]), G["reverse_out"])
    got = O.box_sums_f64(O.sat_f64(G["boxes_img"]), G["boxes_boxes"][None])[0]
    assert np.array_equal(got, G["boxes_out"])
    assert np.array_equal(O.serial(G["ones_in"], np.uint32), G["ones_out"])
    assert np.array_equal(O.serial(G["f32_fold_in"], np.float32), G["f32_fold_out"])
    b = G["bordered_out"]
    assert not b[0].any() and not b[:, 0].any()
    assert np.array_equal(O.serial(G["bordered_in"], np.uint32), b[1:, 1:])

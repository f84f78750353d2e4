"""The library reproduces the golden fixtures of tests/golden exactly, with
every algorithm (the fixtures' sums are exact in the output type)."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest
import torch

from test_gpu_sat import to_dev, to_np

pytestmark = pytest.mark.gpu

G = np.load(Path(__file__).resolve().parent / "golden" / "golden.npz")
TD = {np.dtype(np.uint32): torch.uint32, np.dtype(np.int32): torch.int32, np.dtype(np.float32): torch.float32,
      np.dtype(np.float64): torch.float64}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def sa():
    import sat_amd

    sat_amd.lib()
    return sat_amd


@pytest.mark.parametrize("case", ["u8_u32", "i8_i32", "i32_i32", "u32_u32", "f64_f64", "f32_f32", "u8_f64"])
@pytest.mark.parametrize("alg", [0, 1, 2, 3])
def test_sat_fixtures(dev, sa, case, alg):
    say(f"[case] golden sat {case} alg={alg}")
    x, exp = G[f"sat_{case}_in"], G[f"sat_{case}_out"]
    got = to_np(sa.integral(to_dev(x, dev), out_dtype=TD[exp.dtype], alg=alg))
    assert np.array_equal(got, exp)


def test_pass_fixtures(dev, sa):
    say("[case] golden passes")
    xd = to_dev(G["rows_t_in"], dev)
    assert np.array_equal(to_np(sa.block_scan(xd, out_dtype=torch.uint32)), G["rows_t_out"])
    assert np.array_equal(to_np(sa.seriel_scan(xd, out_dtype=torch.uint32)), G["rows_t_out"])
    assert np.array_equal(to_np(sa.scan_y(to_dev(G["cols_in"], dev), out_dtype=torch.uint32)), G["cols_out"])


def test_other_fixtures(dev, sa):
    say("[case] golden reverse / boxes / bordered / ones / f32 fold")
    assert np.array_equal(sa.integral_reverse(to_dev(G["reverse_in"], dev)).cpu().numpy(), G["reverse_out"])
    S = sa.integral(to_dev(G["boxes_img"], dev))
    b = sa.box_sums(S[None], torch.from_numpy(G["boxes_boxes"]).to(dev)[None]).cpu().numpy()[0]
    assert np.array_equal(b, G["boxes_out"])
    assert np.array_equal(to_np(sa.integral_bordered(to_dev(G["bordered_in"], dev))), G["bordered_out"])
    for alg in range(4):
        assert np.array_equal(to_np(sa.integral(to_dev(G["ones_in"], dev), alg=alg)), G["ones_out"])
        assert np.array_equal(to_np(sa.integral(to_dev(G["f32_fold_in"], dev), alg=alg)), G["f32_fold_out"])

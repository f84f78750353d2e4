"""A seeded sweep of random cases against the oracle: shapes (tall, wide and
square, up to 20000 rows or columns), every dtype pair, forced tile
configurations, plane stacks, the four algorithms and the reverse SAT.
Integer outputs bit-exact, floats within the bounds of tests/helpers.py."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest
import torch

from helpers import KIND, assert_sat_close
from test_gpu_sat import PAIRS, TD, to_dev, to_np

pytestmark = pytest.mark.gpu

IN_BYTES = {"u8": 1, "i8": 1, "u32": 4, "i32": 4, "f32": 4, "f64": 8}


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def sa():
    import sat_amd

    L = sat_amd.lib()
    yield sat_amd
    from sat_amd.tune import reset_table

    reset_table(L)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_random_sweep(O, dev, sa, seed):
    rng = np.random.default_rng(20261204 + seed)
    L = sa.lib()
    for case in range(100):
        tin, tout = PAIRS[int(rng.integers(len(PAIRS)))]
        r = rng.random()
        if r < 0.15:
            h, w = int(rng.integers(1, 20000)), int(rng.integers(1, 300))
        elif r < 0.3:
            h, w = int(rng.integers(1, 300)), int(rng.integers(1, 20000))
        else:
            h, w = int(rng.integers(1, 2000)), int(rng.integers(1, 2000))
        P = 1 if rng.random() < 0.8 else int(rng.integers(2, 4))
        alg = 3 if rng.random() < 0.7 else int(rng.integers(0, 3))
        cfg = int(rng.integers(0, 8))
        ob = 8 if tout == "f64" else 4
        say(f"[case] sweep {seed}/{case}: {tin}->{tout} {P}x{h}x{w} alg {alg} cfg {cfg}")
        cx, cy, cp = C.c_int(), C.c_int(), C.c_int()
        L.sat_tune_class(w, h, P, C.byref(cx), C.byref(cy), C.byref(cp))
        before = L.sat_tune_get(IN_BYTES[tin], ob, cx.value, cy.value, cp.value)
        assert L.sat_tune_set(IN_BYTES[tin], ob, cx.value, cy.value, cp.value, cfg) == 0
        x = O.gen(KIND[tin], (P, h, w) if P > 1 else (h, w), 910000 + 1000 * seed + case)
        y = to_np(sa.integral(to_dev(x, dev), out_dtype=TD[tout], alg=alg))
        if P > 1:
            for p in range(P):
                assert_sat_close(O, x[p], y[p], tout)
        else:
            assert_sat_close(O, x, y, tout)
        if tin == tout and tin in ("f32", "u32", "f64", "i32") and P == 1 and rng.random() < 0.3:
            rv = to_np(sa.integral_reverse(to_dev(x, dev)))
            assert_sat_close(O, np.ascontiguousarray(x[::-1, ::-1]), np.ascontiguousarray(rv[::-1, ::-1]), tout)
        assert L.sat_tune_set(IN_BYTES[tin], ob, cx.value, cy.value, cp.value, before) == 0
    sa.check_faults()

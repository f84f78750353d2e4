"""Pins the CPU oracle (oracle/sat_oracle.c) before it is trusted as the checker.

The reference ships no golden vectors (its Compare() returns true before
looking, integralImage.h:268), so the oracle is pinned by facts fixed
independently of it: closed forms (the authors' all-ones known answer,
integralImage.h:273), brute-force sums on tiny inputs, numpy's cumsum for
integer data, the serial == OpenMP bit identity the reference's TestCPU
asserts (integralImage.h:351), a hand-derived f32 case that only the
reference's fold order reproduces, adjointness of the reverse SAT, and the
published first output of splitmix64.
"""
from __future__ import annotations

import itertools

import numpy as np
import pytest

from oracle.oracle import NP, PAIRS


def brute_sat(x: np.ndarray) -> list[list[int]]:
    h, w = x.shape
    xi = [[int(v) for v in row] for row in x]
    return [[sum(xi[k][l] for k in range(i + 1) for l in range(j + 1)) for j in range(w)] for i in range(h)]


def test_splitmix64_published_value(O):
    # splitmix64 seeded with 0: first output 0xE220A8397B1DCDAF (Vigna's reference implementation)
    assert O.sm64_at(0, 0) == 0xE220A8397B1DCDAF
    a = O.gen("u8", (4, 5), 7)
    assert np.array_equal(a, O.gen("u8", (4, 5), 7))
    assert not np.array_equal(a, O.gen("u8", (4, 5), 8))


@pytest.mark.parametrize("tin,tout", PAIRS)
def test_all_ones_closed_form(O, tin, tout):
    h, w = 37, 53
    x = np.ones((h, w), NP[tin])
    exp = (np.arange(1, h + 1)[:, None] * np.arange(1, w + 1)[None, :]).astype(NP[tout])
    for f in (O.serial, O.parallel):
        assert np.array_equal(f(x, NP[tout]), exp)


def test_constant_wraps_mod_2_32(O):
    # 255 * (i+1)(j+1) exceeds 2^32 beyond (i+1)(j+1) > 16843009
    h = w = 4200
    x = np.full((h, w), 255, np.uint8)
    got = O.serial(x, np.uint32)
    i = np.array([0, 100, 4100, 4199], dtype=np.uint64)
    j = np.array([0, 4199, 4150, 4199], dtype=np.uint64)
    exp = ((255 * (i + 1) * (j + 1)) % (1 << 32)).astype(np.uint32)
    assert np.array_equal(got[i.astype(int), j.astype(int)], exp)
    assert int(255 * 4200 * 4200) >= 1 << 32


@pytest.mark.parametrize("tin,tout", PAIRS)
def test_brute_force_tiny(O, tin, tout):
    kind = {"u8": "u8", "i8": "i8", "u32": "u32_20", "i32": "i32_16", "f32": "u8", "f64": "i8"}[tin]
    for h, w in ((1, 1), (1, 9), (7, 1), (5, 6), (12, 11)):
        raw = O.gen(kind, (h, w), 31 * h + w)
        x = raw.astype(NP[tin])  # float inputs get integer values: every order is exact
        exp = brute_sat(x)
        got = O.serial(x, NP[tout])
        if tout in ("u32", "i32"):
            e = np.array(exp, dtype=np.int64) % (1 << 32)
            assert np.array_equal(got.astype(np.int64) % (1 << 32), e)
        else:
            assert np.array_equal(got.astype(np.float64), np.array(exp, dtype=np.float64))


@pytest.mark.parametrize("kind,tin", [("u8", "u8"), ("u32_20", "u32"), ("u32", "u32"), ("i8", "i8")])
def test_integer_matches_numpy_cumsum(O, kind, tin):
    x = O.gen(kind, (301, 257), 99)
    tout = np.int32 if tin == "i8" else np.uint32
    exp = np.cumsum(np.cumsum(x.astype(np.int64), axis=1), axis=0)
    exp = (exp % (1 << 32)).astype(np.uint32)
    got = O.serial(x, tout).view(np.uint32)
    assert np.array_equal(got, exp)
    assert np.array_equal(O.allrow(x, tout).view(np.uint32),
                          (np.cumsum(x.astype(np.int64), axis=1) % (1 << 32)).astype(np.uint32))
    assert np.array_equal(O.colscan(x, tout).view(np.uint32),
                          (np.cumsum(x.astype(np.int64), axis=0) % (1 << 32)).astype(np.uint32))


@pytest.mark.parametrize("tin,tout", PAIRS)
def test_serial_equals_parallel(O, tin, tout):
    kind = {"u8": "u8", "i8": "i8", "u32": "u32_20", "i32": "i32_16", "f32": "f32n", "f64": "f64n"}[tin]
    x = O.gen(kind, (257, 301), 5)
    a, b = O.serial(x, NP[tout]), O.parallel(x, NP[tout])
    assert a.tobytes() == b.tobytes()


def test_reference_fold_order_f32(O):
    """S(1,1) of [[1e8, 1], [-1e8, 1]] is 2 exactly; the reference's order
    (row left folds, then add the row above: integralImage.h:113-122) gives 0
    in f32, a column-first order gives 2."""
    x = np.array([[1e8, 1], [-1e8, 1]], dtype=np.float32)
    s = O.serial(x, np.float32)
    assert s[0, 1] == np.float32(1e8)
    assert s[1, 0] == 0.0
    assert s[1, 1] == 0.0
    assert O.sat_f64(x)[1, 1] == 2.0
    # a row prefix is a left fold in the accumulator type: (1e8 + 1) + 1 in f32
    r = O.allrow(np.array([[1e8, 1, 1, 8]], dtype=np.float32), np.float32)
    assert r[0].tolist() == [1e8, 1e8, 1e8, 1e8 + 8]


def test_conversion_before_add(O):
    # TSrc -> TDst before the add (integralImage.h:79): no u8 / i8 overflow
    x = np.array([[255, 255, 255]], dtype=np.uint8)
    assert O.serial(x, np.uint32).tolist() == [[255, 510, 765]]
    y = np.array([[-128, -128], [-128, -128]], dtype=np.int8)
    assert O.serial(y, np.int32).tolist() == [[-128, -256], [-256, -512]]


def test_transpose_and_passes(O):
    x = O.gen("u8", (13, 29), 3)
    t = O.transpose(O.allrow(x, np.uint32))
    assert t.shape == (29, 13)
    assert np.array_equal(t, np.cumsum(x.astype(np.uint32), axis=1).T)


def test_reverse_is_adjoint(O):
    x = O.gen("u32_20", (40, 70), 11)
    g = O.gen("u32_20", (40, 70), 12)
    # <SAT x, g> == <x, SAT^T g> modulo 2^32 (exact in integers)
    lhs = int(np.sum(O.serial(x, np.uint32).astype(np.uint64) * g.astype(np.uint64)) % (1 << 32))
    rhs = int(np.sum(x.astype(np.uint64) * O.reverse(g).astype(np.uint64)) % (1 << 32))
    assert lhs == rhs
    # mirror identity
    assert np.array_equal(O.reverse(x), O.serial(x[::-1, ::-1].copy(), np.uint32)[::-1, ::-1])
    # all-ones closed form (H-i)(W-j)
    ones = np.ones((9, 14), np.float64)
    exp = (9 - np.arange(9))[:, None] * (14 - np.arange(14))[None, :]
    assert np.array_equal(O.reverse(ones), exp.astype(np.float64))


def test_box_sums_brute_force(O):
    x = O.gen("f64n", (3, 20, 31), 4)
    sat = np.stack([O.sat_f64(x[p]) for p in range(3)])
    rng = np.random.default_rng(0)
    boxes = []
    for p in range(3):
        for _ in range(50):
            y0, y1 = sorted(rng.integers(0, 20, 2))
            x0, x1 = sorted(rng.integers(0, 31, 2))
            boxes.append((y0, x0, y1, x1))
    boxes = np.array(boxes, np.int32).reshape(3, 50, 4)
    got = O.box_sums_f64(sat, boxes)
    for p, b in itertools.product(range(3), range(50)):
        y0, x0, y1, x1 = boxes[p, b]
        assert got[p, b] == pytest.approx(float(np.sum(x[p, y0:y1 + 1, x0:x1 + 1])), rel=1e-12, abs=1e-12)
    bad = np.array([[[0, 0, 20, 3]]], np.int32)
    assert np.isnan(O.box_sums_f64(sat[:1], bad)[0, 0])


@pytest.mark.parametrize("kind,tin,tout", [("u8", np.uint8, np.uint32), ("i32_16", np.int32, np.int32),
                                           ("u32", np.uint32, np.uint32)])
def test_difference_checker(O, kind, tin, tout):
    """The full-size GPU tests check integer SATs by their 2-D difference
    (helpers.assert_inverts). It must accept the oracle's SAT, wraps included,
    and reject a SAT with one planted error or with a row dropped from the fold."""
    import torch

    from helpers import assert_inverts

    def t(a):
        return torch.from_numpy(a.view(np.int32) if a.dtype == np.uint32 else a)

    x = O.gen(kind, (300, 170), 99)
    S = O.serial(x, tout)
    assert_inverts(t(S), t(x), chunk=64)
    bad = S.copy()
    bad[123, 45] += 1
    with pytest.raises(AssertionError):
        assert_inverts(t(bad), t(x), chunk=64)
    shifted = np.zeros_like(S)
    shifted[1:] = S[:-1]  # the SAT of the image without its last row, shifted down
    with pytest.raises(AssertionError):
        assert_inverts(t(shifted), t(x), chunk=64)

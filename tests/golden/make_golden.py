"""Writes the golden fixtures of this directory (golden.npz).

The reference ships no golden vectors (its Compare() returns true before
comparing, integralImage.h:268), and its code cannot be built here. These
fixtures are therefore computed by brute force in pure Python integer / exact
rational arithmetic (every expected value is a direct double sum over the
definition, independent of the oracle and of the library) and from closed
forms, on small seeded inputs:

  sat_<pair>        inclusive 2-D prefix sums (IntegralImageSerial's
                    definition, integralImage.h:105-124), integer pairs
                    modulo 2^32 as the reference's uint accumulators wrap
  rows_t            the transposed row scan of BlockScan::blockScan /
                    SerielScan::serielScan (ScanRow-BRLT.cu:24-25,
                    IntegralImageAllRow + Transpose, integralImage.h:148-160,
                    280-288)
  cols              the column scan of IncrementScan::ScanY
                    (IntegralStride, integralImage.h:90-100)
  reverse           suffix sums (the SAT's adjoint)
  boxes             four-corner box sums
  bordered          the (H+1) x (W+1) zero-bordered layout
  ones              the authors' all-ones known answer (x+1)(y+1)
                    (integralImage.h:273)
  f32_fold          a case whose f32 result is fixed by the reference's
                    fold order: [[1e8, 1], [-1e8, 1]] gives S(1,1) = 0 in the
                    row-then-column order of IntegralImageSerial (an order
                    that adds 1 + 1 first would give 2)

Run: python tests/golden/make_golden.py  (writes tests/golden/golden.npz)
"""
from __future__ import annotations

from fractions import Fraction
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent / "golden.npz"
M32 = 1 << 32


def brute_sat(x, mod=None):
    h, w = x.shape
    xs = [[int(v) if mod is not None or x.dtype.kind in "iu" else Fraction(float(v)) for v in row] for row in x]
    out = []
    for i in range(h):
        row = []
        for j in range(w):
            s = sum(xs[k][l] for k in range(i + 1) for l in range(j + 1))
            row.append(s % mod if mod else s)
        out.append(row)
    return out


def main():
    rng = np.random.default_rng(20260801)
    f = {}
    # integer pairs: u8 -> u32, i8 -> i32, i32 -> i32, u32 -> u32 (with wrap)
    x = rng.integers(0, 256, (7, 9), dtype=np.uint8)
    f["sat_u8_u32_in"], f["sat_u8_u32_out"] = x, np.array(brute_sat(x, M32), dtype=np.uint32)
    x = rng.integers(-128, 128, (5, 13), dtype=np.int8)
    f["sat_i8_i32_in"] = x
    f["sat_i8_i32_out"] = np.array(brute_sat(x, M32), dtype=np.uint64).astype(np.uint32).view(np.int32)
    x = rng.integers(-(1 << 30), 1 << 30, (6, 5), dtype=np.int32)
    f["sat_i32_i32_in"] = x
    f["sat_i32_i32_out"] = np.array(brute_sat(x, M32), dtype=np.uint64).astype(np.uint32).view(np.int32)
    x = rng.integers(M32 - 1000, M32, (3, 4), dtype=np.uint64).astype(np.uint32)  # every partial sum wraps
    f["sat_u32_u32_in"], f["sat_u32_u32_out"] = x, np.array(brute_sat(x, M32), dtype=np.uint32)
    # f64: small integers stored in f64 (exact sums), and dyadic fractions
    x = (rng.integers(-64, 64, (6, 7)) / 8.0).astype(np.float64)
    f["sat_f64_f64_in"] = x
    f["sat_f64_f64_out"] = np.array([[float(v) for v in r] for r in brute_sat(x)], dtype=np.float64)
    # f32 with exactly representable partial sums
    x = (rng.integers(0, 16, (9, 11)) / 4.0).astype(np.float32)
    f["sat_f32_f32_in"] = x
    f["sat_f32_f32_out"] = np.array([[float(v) for v in r] for r in brute_sat(x)], dtype=np.float32)
    # u8 -> f32 / f64
    x = rng.integers(0, 256, (8, 6), dtype=np.uint8)
    f["sat_u8_f64_in"] = x
    f["sat_u8_f64_out"] = np.array(brute_sat(x), dtype=np.float64)
    # transposed row scan (pass semantics of blockScan / serielScan): out[x][y] = sum_{k<=x} in[y][k]
    x = rng.integers(0, 256, (6, 10), dtype=np.uint8)
    rows = np.array([[sum(int(v) for v in r[:j + 1]) for j in range(x.shape[1])] for r in x], dtype=np.uint32)
    f["rows_t_in"], f["rows_t_out"] = x, np.ascontiguousarray(rows.T)
    # column scan: out[y][x] = sum_{k<=y} in[k][x]
    x = rng.integers(0, 256, (8, 5), dtype=np.uint8)
    cols = np.array([[sum(int(x[k, j]) for k in range(i + 1)) for j in range(x.shape[1])] for i in range(x.shape[0])],
                    dtype=np.uint32)
    f["cols_in"], f["cols_out"] = x, cols
    # reverse SAT: out[y][x] = sum_{k>=y, l>=x} in[k][l]
    x = (rng.integers(-50, 50, (6, 7)) / 2.0).astype(np.float64)
    h, w = x.shape
    rev = [[sum(Fraction(float(x[k, l])) for k in range(i, h) for l in range(j, w)) for j in range(w)] for i in range(h)]
    f["reverse_in"], f["reverse_out"] = x, np.array([[float(v) for v in r] for r in rev], dtype=np.float64)
    # box sums over an f64 image: boxes {y0, x0, y1, x1} inclusive
    x = (rng.integers(-20, 20, (12, 15)) / 4.0).astype(np.float64)
    boxes = np.array([[0, 0, 11, 14], [0, 0, 0, 0], [3, 4, 7, 9], [11, 14, 11, 14], [5, 0, 8, 14], [0, 7, 11, 7]],
                     dtype=np.int32)
    sums = [float(sum(Fraction(float(x[k, l])) for k in range(b[0], b[2] + 1) for l in range(b[1], b[3] + 1)))
            for b in boxes]
    f["boxes_img"], f["boxes_boxes"], f["boxes_out"] = x, boxes, np.array(sums, dtype=np.float64)
    # bordered (H+1) x (W+1) layout, u8 -> u32
    x = rng.integers(0, 256, (4, 5), dtype=np.uint8)
    bord = np.zeros((5, 6), dtype=np.uint32)
    bord[1:, 1:] = np.array(brute_sat(x, M32), dtype=np.uint32)
    f["bordered_in"], f["bordered_out"] = x, bord
    # the authors' all-ones known answer
    h, w = 33, 70
    f["ones_in"] = np.ones((h, w), dtype=np.uint8)
    f["ones_out"] = (np.arange(1, h + 1)[:, None] * np.arange(1, w + 1)[None, :]).astype(np.uint32)
    # f32 fold order
    f["f32_fold_in"] = np.array([[1e8, 1.0], [-1e8, 1.0]], dtype=np.float32)
    f["f32_fold_out"] = np.array([[1e8, 1e8], [0.0, 0.0]], dtype=np.float32)
    np.savez(OUT, **f)
    print(f"wrote {OUT} ({len(f)} arrays)")


if __name__ == "__main__":
    main()

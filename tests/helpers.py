"""Shared helpers of the parity tests (CPU side only: comparisons and inputs)."""
from __future__ import annotations

import numpy as np

# generator kind per input dtype for random parity cases
KIND = {"u8": "u8", "i8": "i8", "u32": "u32_20", "i32": "i32_16", "f32": "f32", "f64": "f64n"}
INT_OUT = {"u32", "i32"}

# f32 tolerance: |gpu - exact| <= F32_K * 2^-24 * SAT(|x|)(y, x) + tiny, i.e. a
# small multiple of one rounding of the largest partial sum on the way. The
# reference's own f32 CPU left fold has an error bound that grows with the
# image size instead; see test_float_error_vs_cpu_reference.
# Over 12,900 random shapes (tall, wide, square up to 20000 x 600 and
# 3000^2), tile configurations and algorithms with U[0,1) and N(0,1) inputs
# the largest f32 error measured was 8.2 u * SAT(|x|) (99.9th percentile 6.8).
F32_K = 32.0
# f64: the same sweep over 14,211 cases measured at most 4.6 u * SAT(|x|)
# (99th percentile 2.6, 99.9th 3.4; 1.70 on the 4096^2 N(0,1) workload). The
# look-back's order of additions varies from run to run, so the bound keeps
# the same margin over the measured maximum as the f32 one.
F64_K = 16.0


def sat_abs_f64(O, x: np.ndarray) -> np.ndarray:
    return O.serial(np.abs(x.astype(np.float64)), np.float64)


def assert_sat_close(O, x: np.ndarray, got: np.ndarray, tout: str, exact: np.ndarray | None = None,
                     k: float | None = None) -> float:
    """Integer outputs: bit-exact against the oracle. Float outputs: within the
    stated bound of the FP64 oracle. Returns the max error in units of
    u * SAT(|x|) (0 for integers)."""
    if tout in INT_OUT:
        if exact is None:
            exact = O.serial(x, np.dtype({"u32": np.uint32, "i32": np.int32}[tout]))
        bad = np.argwhere(got != exact)
        assert bad.size == 0, f"{len(bad)} mismatches, first at {tuple(bad[0])}: {got[tuple(bad[0])]} != {exact[tuple(bad[0])]}"
        return 0.0
    ref = O.sat_f64(x) if exact is None else exact
    scale = sat_abs_f64(O, x)
    u = 2.0 ** -24 if tout == "f32" else 2.0 ** -53
    k = k if k is not None else (F32_K if tout == "f32" else F64_K)
    err = np.abs(got.astype(np.float64) - ref)
    bound = k * u * scale + 1e-300
    ratio = float(np.max(err / (u * scale + 1e-300))) if err.size else 0.0
    assert np.all(err <= bound), f"max error {ratio:.2f} u*SAT(|x|) exceeds {k}"
    return ratio


def assert_inverts(S, x, chunk: int = 2048) -> None:
    """Checks that the 2-D difference S(i,j) - S(i-1,j) - S(i,j-1) + S(i-1,j-1)
    (S = 0 outside) gives back the input x modulo 2^32 at every pixel, which
    holds exactly when S is the inclusive 2-D prefix sum of x with 32-bit
    wraparound (the difference map is a bijection on 32-bit images). S and x
    are (H, W) torch tensors on any one device: S of a 32-bit integer dtype, x
    of any integer dtype; the work is done `chunk` rows at a time in int64."""
    import torch

    M32 = 0xFFFFFFFF
    H, W = S.shape
    Si = S.view(torch.int32)
    if x.dtype == torch.uint32:
        x = x.view(torch.int32)
    prev = torch.zeros((1, W), dtype=torch.int64, device=S.device)
    for r0 in range(0, H, chunk):
        blk = Si[r0:r0 + chunk].to(torch.int64) & M32
        d = blk - torch.cat([prev, blk[:-1]], 0)
        d2 = d.clone()
        d2[:, 1:] -= d[:, :-1]
        d2 &= M32
        xe = x[r0:r0 + chunk].to(torch.int64) & M32
        if not torch.equal(d2, xe):
            bad = (d2 != xe).nonzero()[0].tolist()
            i, j = bad[0] + r0, bad[1]
            raise AssertionError(f"2-D difference differs from the input at ({i}, {j}): "
                                 f"{int(d2[bad[0], j])} != {int(xe[bad[0], j])}")
        prev = blk[-1:]

"""GPU parity at the full workload sizes (f32 8192^2, u32 8192^2, f64 4096^2,
u8->u32 16384^2, f32 65536x256 and 256x65536, 96 planes of 1024^2).

Integer outputs are checked completely through a size-independent property:
the 2-D difference S(i,j) - S(i-1,j) - S(i,j-1) + S(i-1,j-1) (with S = 0
outside the image) must give back the input modulo 2^32 at every pixel. The
difference map is a bijection on 32-bit images, so this holds exactly when S
is the inclusive 2-D prefix sum of the input with the reference's uint
wraparound (integralImage.h:105-124). The all-255 16384^2 image, which wraps
many times, is also compared with its closed form 255(i+1)(j+1) mod 2^32.
Float outputs are compared with the FP64 oracle within the bound of
tests/helpers.py.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from helpers import assert_inverts, assert_sat_close

pytestmark = pytest.mark.gpu

M32 = 0xFFFFFFFF  # 32-bit wraparound mask


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def sa():
    import sat_amd

    sat_amd.lib()
    return sat_amd


def gen_int(dev, kind: str, shape, seed: int) -> torch.Tensor:
    g = torch.Generator(device=dev).manual_seed(seed)
    if kind == "u8":
        return torch.randint(0, 256, shape, dtype=torch.uint8, device=dev, generator=g)
    if kind == "u32_20":
        return torch.randint(0, 1 << 20, shape, dtype=torch.int32, device=dev, generator=g).view(torch.uint32)
    if kind == "i32_16":
        return torch.randint(-(1 << 16), 1 << 16, shape, dtype=torch.int32, device=dev, generator=g)
    raise ValueError(kind)


@pytest.mark.parametrize("alg", [3, 0, 1, 2])
def test_u8_u32_16384_random(dev, sa, alg):
    say(f"[full] u8->u32 16384x16384 random, alg={alg}")
    x = gen_int(dev, "u8", (16384, 16384), 20260105)
    y = sa.integral(x, out_dtype=torch.uint32, alg=alg)
    torch.cuda.synchronize()
    assert_inverts(y, x)


def test_u8_u32_16384_all_255_closed_form(dev, sa):
    say("[full] u8->u32 16384x16384 all 255 (wraps), one-pass")
    H = W = 16384
    x = torch.full((H, W), 255, dtype=torch.uint8, device=dev)
    y = sa.integral(x, out_dtype=torch.uint32).view(torch.int32)
    torch.cuda.synchronize()
    j1 = torch.arange(1, W + 1, dtype=torch.int64, device=dev)[None, :]
    for r0 in range(0, H, 2048):
        i1 = torch.arange(r0 + 1, r0 + 2049, dtype=torch.int64, device=dev)[:, None]
        exp = (255 * i1 * j1) & M32
        got = y[r0:r0 + 2048].to(torch.int64) & M32
        assert torch.equal(got, exp), f"closed form differs in rows {r0}..{r0 + 2047}"


@pytest.mark.parametrize("alg", [3, 0, 1, 2])
def test_u32_8192_random(dev, sa, alg):
    say(f"[full] u32->u32 8192x8192 values < 2^20 (wraps), alg={alg}")
    x = gen_int(dev, "u32_20", (8192, 8192), 20260102)
    y = sa.integral(x, out_dtype=torch.uint32, alg=alg)
    torch.cuda.synchronize()
    assert_inverts(y, x)


def test_i32_8192_signed_wrap(dev, sa):
    say("[full] i32->i32 8192x8192 in [-2^16, 2^16), one-pass")
    x = gen_int(dev, "i32_16", (8192, 8192), 20260107)
    y = sa.integral(x, out_dtype=torch.int32)
    torch.cuda.synchronize()
    assert_inverts(y, x)


@pytest.mark.parametrize("alg", [3, 0, 1, 2])
def test_f32_8192(O, dev, sa, alg):
    say(f"[full] f32 8192x8192, alg={alg}")
    x = O.gen("f32", (8192, 8192), 20260101)
    y = sa.integral(torch.from_numpy(x).to(dev), alg=alg)
    # the two-pass compositions chain one carry per 256 (or 1024) elements along
    # each 8192-element pass, about 2 x (32 + 16) roundings on the way: 128 u
    # bounds them; the one-pass kernel keeps the default bound
    assert_sat_close(O, x, y.cpu().numpy(), "f32", k=None if alg == 3 else 128.0)


def test_f64_4096(O, dev, sa):
    say("[full] f64 4096x4096 N(0,1), one-pass")
    x = O.gen("f64n", (4096, 4096), 20260103)
    y = sa.integral(torch.from_numpy(x).to(dev))
    assert_sat_close(O, x, y.cpu().numpy(), "f64")


@pytest.mark.parametrize("shape", [(65536, 256), (256, 65536)])
def test_f32_tall_and_wide(O, dev, sa, shape):
    say(f"[full] f32 {shape[0]}x{shape[1]}, one-pass")
    x = O.gen("f32", shape, 20260104)
    y = sa.integral(torch.from_numpy(x).to(dev))
    assert_sat_close(O, x, y.cpu().numpy(), "f32")


def test_f32_96_planes_1024(O, dev, sa):
    say("[full] f32 32x3x1024x1024 planes, one-pass")
    x = O.gen("f32", (96, 1024, 1024), 20260108).reshape(32, 3, 1024, 1024)
    y = sa.integral(torch.from_numpy(x).to(dev)).cpu().numpy()
    for n, c in ((0, 0), (7, 1), (19, 2), (31, 2)):
        assert_sat_close(O, np.ascontiguousarray(x[n, c]), y[n, c], "f32")


@pytest.mark.parametrize("cfg", [0, 5])
def test_u8_u32_beyond_2_32_pixels(dev, sa, cfg):
    """65539 x 65792 u8 -> u32 (4.31e9 pixels, a 16 GiB table): element
    offsets past 2^32 on the vector path, in the 64-row tiles and in the
    packed 256-row tiles, checked exactly through the 2-D difference."""
    H, W = 65536 + 3, 65536 + 256
    say(f"[full] u8->u32 {H}x{W} ({H * W / 2**32:.2f} x 2^32 pixels), cfg {cfg}")
    L = sa.lib()
    import ctypes as C

    cx, cy, cp = C.c_int(), C.c_int(), C.c_int()
    L.sat_tune_class(W, H, 1, C.byref(cx), C.byref(cy), C.byref(cp))
    before = L.sat_tune_get(1, 4, cx.value, cy.value, cp.value)
    assert L.sat_tune_set(1, 4, cx.value, cy.value, cp.value, cfg) == 0
    try:
        x = gen_int(dev, "u8", (H, W), 20260107)
        y = sa.integral(x, out_dtype=torch.uint32)
        torch.cuda.synchronize()
        assert_inverts(y, x, chunk=1024)
        sa.check_faults()
    finally:
        L.sat_tune_set(1, 4, cx.value, cy.value, cp.value, before)
        x = y = None
        torch.cuda.empty_cache()

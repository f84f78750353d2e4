"""GPU parity of the benchmark workloads at their stated sizes, input bytes
and seeds (bench.py generates the same ones):

* u8 -> u32 2048 x 2048, U[0,255] (seed 20260106), every algorithm, bit-exact
  against the oracle's IntegralImageSerial (integralImage.h:105-124);
* f64 4096 x 4096, N(0,1) (seed 20260103), every algorithm, against the
  oracle's f64 IntegralImageSerial (the reference's summation order) with the
  error reported in units of u * SAT(|x|), u = 2^-53;
* the batched workload: 512 images with H and W uniform in [64, 2048], 50/50
  u8 -> u32 and f32 (seed 20260501), one launch, every image against the oracle;
* the differentiable workload: 32 x 3 x 1024^2 f32 N(0,1), 4096 boxes per
  plane: box sums and the fused gradient against FP64 on four planes.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from helpers import F32_K, F64_K, assert_sat_close

pytestmark = pytest.mark.gpu

# f64 error bounds in u * SAT(|x|) units: the one-pass kernel, and the
# two-pass compositions, which chain one carry per tile along each pass
F64_K_TWO_PASS = 2 * F64_K
# normwise relative error bar of the differentiable workload against FP64
DIFF_BAR = 1e-6


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def sa():
    import sat_amd

    sat_amd.lib()
    return sat_amd


@pytest.fixture(scope="module")
def bench():
    import bench as B

    return B


@pytest.mark.parametrize("alg", [0, 1, 2, 3])
def test_u8_u32_2048_bit_exact(O, dev, sa, alg):
    say(f"[config] u8->u32 2048x2048 seed 20260106, alg={alg}")
    x = O.gen("u8", (2048, 2048), 20260106)
    y = sa.integral(torch.from_numpy(x).to(dev), out_dtype=torch.uint32, alg=alg)
    got = y.view(torch.int32).cpu().numpy().view(np.uint32)
    sa.check_faults()
    exp = O.serial(x, np.uint32)
    bad = np.argwhere(got != exp)
    assert bad.size == 0, f"{len(bad)} mismatches, first at {tuple(bad[0])}"


@pytest.mark.parametrize("alg", [3, 0, 1, 2])
def test_f64_4096_all_algorithms(O, dev, sa, alg):
    say(f"[config] f64 4096x4096 N(0,1) seed 20260103, alg={alg}")
    x = O.gen("f64n", (4096, 4096), 20260103)
    y = sa.integral(torch.from_numpy(x).to(dev), alg=alg).cpu().numpy()
    sa.check_faults()
    k = F64_K if alg == 3 else F64_K_TWO_PASS
    ratio = assert_sat_close(O, x, y, "f64", k=k)
    ref = O.serial(x, np.float64)
    norm = float(np.linalg.norm(y - ref) / np.linalg.norm(ref))
    say(f"[config] f64 4096^2 alg={alg}: max error {ratio:.2f} u*SAT(|x|) (bound {k}), normwise {norm:.3g}, "
        f"bit-identical to the serial order {np.mean(y == ref):.3f}")
    assert norm < 64 * 2.0 ** -53


def test_batched_512_images_every_image(O, dev, sa, bench):
    say("[config] batched 512 images U[64,2048]^2, 50/50 u8->u32 / f32, seed 20260501")
    host, pairs, _ = bench.batch_inputs(O, dev)
    plan = sa.BatchPlan(pairs)
    plan.run()
    torch.cuda.synchronize()
    plan.check_faults()
    worst = 0.0
    for i, (x, (_, yd)) in enumerate(zip(host, pairs)):
        if x.dtype == np.uint8:
            got = yd.view(torch.int32).cpu().numpy().view(np.uint32)
            exp = O.serial(x, np.uint32)
            assert np.array_equal(got, exp), f"image {i} {x.shape} u8->u32 differs from the oracle"
        else:
            worst = max(worst, assert_sat_close(O, x, yd.cpu().numpy(), "f32", k=F32_K))
        if i % 128 == 127:
            say(f"[config] batched: {i + 1} images checked")
    say(f"[config] batched: f32 images within {worst:.2f} u*SAT(|x|)")


def test_diff_full_size_box_sums_and_gradient(O, dev, sa, bench):
    say("[config] diff 32x3x1024^2 f32 N(0,1), 4096 boxes per plane: box sums and fused gradient")
    x_np, boxes_np, g_np, x, boxes, g = bench.diff_inputs(O, dev)
    P, nbox = boxes_np.shape[:2]
    H, W = x.shape[-2:]
    xg = x.clone().requires_grad_(True)
    b = sa.box_filter(xg, boxes)
    b.backward(g.view(b.shape))
    sa.check_faults()
    bb = b.detach().view(P, nbox).cpu().numpy().astype(np.float64)
    gx = xg.grad.view(P, H, W).cpu().numpy().astype(np.float64)
    for p in (0, 31, 64, 95):
        S64 = O.sat_f64(x_np[p].astype(np.float64))
        ref = O.box_sums_f64(S64, boxes_np[p][None])[0]
        eb = float(np.linalg.norm(bb[p] - ref) / np.linalg.norm(ref))
        gref = bench.box_grad_f64(boxes_np[p], g_np[p], H, W)
        eg = float(np.linalg.norm(gx[p] - gref) / np.linalg.norm(gref))
        say(f"[config] diff plane {p}: box sums {eb:.3g}, gradient {eg:.3g} normwise vs FP64")
        assert eb <= DIFF_BAR, f"plane {p}: box sums {eb:.3g} > {DIFF_BAR}"
        assert eg <= DIFF_BAR, f"plane {p}: gradient {eg:.3g} > {DIFF_BAR}"

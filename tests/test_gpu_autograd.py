"""GPU checks of the differentiable SAT and the box-filter readout.

The SAT is the linear operator M = kron(L_H, L_W) (L lower-triangular ones) on
the row-major vector of the image; its adjoint is the reverse SAT. Gradients
are checked against that dense matrix, against torch.autograd.gradcheck
(finite differences, f64, 1x1x64x64) and against brute-force box indicators.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def sa():
    import sat_amd

    sat_amd.lib()
    return sat_amd


def make_boxes(planes, nbox, H, W, seed):
    """Random boxes {y0, x0, y1, x1} (inclusive): side uniform in [3, 63]
    (clipped to the image), centre uniform, box shifted to lie inside."""
    rng = np.random.default_rng(seed)
    sh = np.minimum(rng.integers(3, 64, (planes, nbox)), H)
    sw = np.minimum(rng.integers(3, 64, (planes, nbox)), W)
    cy = rng.integers(0, H, (planes, nbox))
    cx = rng.integers(0, W, (planes, nbox))
    y0 = np.clip(cy - sh // 2, 0, H - sh)
    x0 = np.clip(cx - sw // 2, 0, W - sw)
    return np.stack([y0, x0, y0 + sh - 1, x0 + sw - 1], -1).astype(np.int32)


def dense_sat_matrix(H, W):
    return np.kron(np.tril(np.ones((H, H))), np.tril(np.ones((W, W))))


def test_sat_backward_is_dense_adjoint(O, dev, sa):
    say("[case] SAT backward vs dense kron(L,L)^T")
    H, W = 40, 64
    M = dense_sat_matrix(H, W)
    x = torch.from_numpy(O.gen("f64n", (1, 1, H, W), 1)).to(dev).requires_grad_(True)
    g = O.gen("f64n", (1, 1, H, W), 2)
    y = sa.sat(x)
    assert np.allclose(y.detach().cpu().numpy().ravel(), M @ x.detach().cpu().numpy().ravel(), rtol=1e-12, atol=1e-12)
    y.backward(torch.from_numpy(g).to(dev))
    ref = M.T @ g.ravel()
    assert np.allclose(x.grad.cpu().numpy().ravel(), ref, rtol=1e-12, atol=1e-12)


def test_reverse_closed_form(O, dev, sa):
    say("[case] reverse SAT of ones = (H-i)(W-j)")
    H, W = 300, 517
    ones = torch.ones((H, W), dtype=torch.float32, device=dev)
    got = sa.integral_reverse(ones).cpu().numpy()
    exp = (H - np.arange(H))[:, None] * (W - np.arange(W))[None, :]
    assert np.array_equal(got, exp.astype(np.float32))


def test_gradcheck_sat_f64(dev, sa):
    say("[case] gradcheck sat f64 1x1x64x64")
    x = torch.randn(1, 1, 64, 64, dtype=torch.float64, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(sa.sat, (x,), eps=1e-6, atol=1e-8, rtol=1e-6, fast_mode=True)


def test_gradcheck_box_filter_f64(dev, sa):
    say("[case] gradcheck box_filter and box_sum f64 1x1x64x64")
    # A box sum is a difference of SAT values of magnitude ~|S| ~ 64, so the
    # central differences carry ~u*64/eps ~ 1e-8 rounding noise; the analytical
    # Jacobian is exactly 0 / +-1.
    boxes = torch.from_numpy(make_boxes(1, 50, 64, 64, 3)).to(dev)
    x = torch.randn(1, 1, 64, 64, dtype=torch.float64, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: sa.box_filter(t, boxes), (x,), eps=1e-6, atol=1e-6, rtol=1e-6)
    S = torch.randn(1, 1, 64, 64, dtype=torch.float64, device=dev, requires_grad=True)
    assert torch.autograd.gradcheck(lambda t: sa.box_sum(t, boxes), (S,), eps=1e-6, atol=1e-6, rtol=1e-6)


@pytest.mark.parametrize("dt", [torch.float32, torch.float64])
def test_box_filter_forward_backward(O, dev, sa, dt):
    say(f"[case] box filter fwd/bwd {dt}")
    P, H, W, nbox = 3, 200, 333, 500
    x_np = O.gen("f64n", (P, H, W), 5)
    boxes_np = make_boxes(P, nbox, H, W, 6)
    x = torch.from_numpy(x_np).to(dev).to(dt).requires_grad_(True)
    boxes = torch.from_numpy(boxes_np).to(dev)
    b = sa.box_filter(x, boxes)
    xs = x.detach().double().cpu().numpy()
    exp = np.array([[xs[p, y0:y1 + 1, x0:x1 + 1].sum() for (y0, x0, y1, x1) in boxes_np[p]] for p in range(P)])
    absx = np.abs(xs)
    scale = np.array([[absx[p, :y1 + 1, :x1 + 1].sum() for (y0, x0, y1, x1) in boxes_np[p]] for p in range(P)])
    u = 2.0 ** -24 if dt == torch.float32 else 2.0 ** -53
    assert np.all(np.abs(b.detach().double().cpu().numpy() - exp) <= 64 * u * scale + 1e-30)
    g_np = O.gen("f64n", (P, nbox), 7)
    b.backward(torch.from_numpy(g_np).to(dev).to(dt))
    ref = np.zeros((P, H, W))
    for p in range(P):
        for k, (y0, x0, y1, x1) in enumerate(boxes_np[p]):
            ref[p, y0:y1 + 1, x0:x1 + 1] += g_np[p, k]
    err = np.abs(x.grad.double().cpu().numpy() - ref)
    tol = (64 * u * np.abs(g_np).sum(axis=1)[:, None, None]) + 1e-30
    assert np.all(err <= tol)
    # the fused backward equals reverse SAT of the scattered corner impulses
    S = sa.integral(x.detach())
    Sg = S.clone().requires_grad_(True)
    sa.box_sum(Sg, boxes).backward(torch.from_numpy(g_np).to(dev).to(dt))
    via_impulses = sa.integral_reverse(Sg.grad)
    assert torch.allclose(via_impulses, x.grad, rtol=0, atol=float(tol.max()))


def test_box_invalid_boxes(dev, sa):
    say("[case] invalid boxes read nothing")
    S = torch.zeros(1, 10, 10, device=dev)
    boxes = torch.tensor([[[0, 0, 10, 3], [5, 5, 4, 6], [-1, 0, 2, 2], [0, 0, 9, 9]]], dtype=torch.int32, device=dev)
    out = sa.box_sums(S, boxes).cpu().numpy()
    assert np.isnan(out[0, :3]).all() and out[0, 3] == 0.0
    g = sa.box_grad(boxes, torch.ones(1, 4, device=dev), 10, 10).cpu().numpy()
    assert np.array_equal(g, np.ones((1, 10, 10), np.float32))


def test_managed_path_graph_replay(O, dev, sa):
    """A captured one-pass SAT (device-side epoch bump) replayed with new
    inputs, interleaved with eager calls on the same stream."""
    say("[case] graph capture of the one-pass SAT")
    H, W = 300, 700
    s = torch.cuda.Stream()
    x = torch.empty((H, W), dtype=torch.uint8, device=dev)
    y = torch.empty((H, W), dtype=torch.uint32, device=dev)
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        sa.integral(x.fill_(1), out=y)  # allocates the managed workspace of stream s
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        sa.integral(x, out=y)
    for rep in range(4):
        xn = O.gen("u8", (H, W), 300 + rep)
        xh = torch.from_numpy(xn).to(dev)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            x.copy_(xh)
            z = sa.integral(x)  # eager call on the same stream and workspace
            g.replay()
        torch.cuda.synchronize()
        exp = O.serial(xn, np.uint32)
        assert np.array_equal(y.view(torch.int32).cpu().numpy().view(np.uint32), exp)
        assert np.array_equal(z.view(torch.int32).cpu().numpy().view(np.uint32), exp)
    assert sa.lib().sat_workspace_fault(None, s.cuda_stream) == 0


def test_graph_replay_after_workspace_grows(O, dev, sa):
    """A graph captured on a small image keeps working after a larger eager
    call on the same stream grew the managed workspace (the captured launch
    still points at the old buffer, which the library retires instead of
    freeing)."""
    say("[case] graph replay after the managed workspace grew")
    H, W = 200, 300
    s = torch.cuda.Stream()
    x = torch.empty((H, W), dtype=torch.uint8, device=dev)
    y = torch.empty((H, W), dtype=torch.uint32, device=dev)
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        sa.integral(x.fill_(1), out=y)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        sa.integral(x, out=y)
    big_np = O.gen("u8", (4096, 6000), 310)  # a workspace many times the small one's
    big = torch.from_numpy(big_np).to(dev)
    with torch.cuda.stream(s):
        zb = sa.integral(big)
    torch.cuda.synchronize()
    zb_np = zb.view(torch.int32).cpu().numpy().view(np.uint32)
    assert np.array_equal(zb_np[-1], O.serial(big_np, np.uint32)[-1])
    del zb, zb_np
    for rep in range(4):
        xn = O.gen("u8", (H, W), 320 + rep)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            x.copy_(torch.from_numpy(xn).to(dev))
            g.replay()
            z = sa.integral(x)  # eager call on the grown workspace between replays
        torch.cuda.synchronize()
        exp = O.serial(xn, np.uint32)
        assert np.array_equal(y.view(torch.int32).cpu().numpy().view(np.uint32), exp), f"replay {rep}"
        assert np.array_equal(z.view(torch.int32).cpu().numpy().view(np.uint32), exp)
    assert sa.lib().sat_workspace_fault(None, s.cuda_stream) == 0


def test_box_grad_caller_workspace_many_planes(O, dev, sa):
    """The binned fused gradient (caller-provided workspace, many planes,
    boxes touching the borders) equals reverse SAT of the scattered corner
    impulses, both in f64."""
    import ctypes as C

    say("[case] box_grad binned, caller workspace, 24 planes")
    P, H, W, nbox = 24, 300, 520, 700
    boxes_np = make_boxes(P, nbox, H, W, 11)
    boxes_np[:, 0] = [0, 0, H - 1, W - 1]  # whole image
    boxes_np[:, 1] = [H - 1, W - 1, H - 1, W - 1]  # last pixel
    boxes = torch.from_numpy(boxes_np).to(dev)
    g = torch.from_numpy(O.gen("f64n", (P, nbox), 12)).to(dev)
    L = sa.lib()
    nb = L.sat_box_grad_workspace_bytes(7, W, H, P, nbox)
    assert nb > L.sat_box_grad_workspace_bytes(7, W, H, P, 0)  # room for the bins
    # garbage everywhere but the 256-byte header, which a caller zero-fills
    # once (it holds the fault word, kept across calls): the library clears
    # the carry area it needs and overwrites the bins
    ws = torch.full((nb,), 0xAB, dtype=torch.uint8, device=dev)
    ws[:256].zero_()
    gx = torch.empty((P, H, W), dtype=torch.float64, device=dev)
    for rep in range(2):
        st = L.sat_box_grad(7, boxes.data_ptr(), nbox, g.data_ptr(), gx.data_ptr(), W, W * H, W, H, P,
                            ws.data_ptr(), nb, torch.cuda.current_stream().cuda_stream)
        assert st == 0
        gs = sa.box_scatter(boxes, g, H, W)
        ref = sa.integral_reverse(gs)
        torch.cuda.synchronize()
        assert torch.allclose(gx, ref, rtol=0, atol=1e-9), float((gx - ref).abs().max())
    assert L.sat_workspace_fault(C.c_void_p(ws.data_ptr()), None) == 0

"""GPU parity of the full SAT and the pass-level drop-ins against the oracle.

Integer outputs must be bit-exact (including modulo-2^32 wraparound); float
outputs must lie within the bound of tests/helpers.py of the FP64 oracle.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from helpers import INT_OUT, KIND, assert_sat_close

pytestmark = pytest.mark.gpu

NP = {"u8": np.uint8, "i8": np.int8, "u32": np.uint32, "i32": np.int32, "f32": np.float32, "f64": np.float64}
TD = {"u8": torch.uint8, "i8": torch.int8, "u32": torch.uint32, "i32": torch.int32, "f32": torch.float32,
      "f64": torch.float64}
PAIRS = [("u8", "u32"), ("u8", "f32"), ("u8", "f64"), ("i8", "i32"), ("i8", "f32"), ("i8", "f64"), ("i32", "i32"),
         ("i32", "f64"), ("f32", "f32"), ("f32", "f64"), ("f64", "f64"), ("u32", "u32")]
SMALL = [(1, 1), (3, 5), (64, 256), (65, 257), (31, 300), (200, 17), (130, 520)]


def to_dev(x: np.ndarray, dev) -> torch.Tensor:
    if x.dtype == np.uint32:
        return torch.from_numpy(x.view(np.int32)).to(dev).view(torch.uint32)
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev)


def to_np(t: torch.Tensor) -> np.ndarray:
    if t.dtype == torch.uint32:
        return t.view(torch.int32).cpu().numpy().view(np.uint32)
    return t.cpu().numpy()


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def sa():
    import sat_amd

    sat_amd.lib()
    return sat_amd


@pytest.mark.parametrize("alg", [0, 1, 2, 3])
@pytest.mark.parametrize("tin,tout", PAIRS)
def test_integral_small(O, dev, sa, alg, tin, tout):
    say(f"[case] integral alg={alg} {tin}->{tout}")
    for k, (h, w) in enumerate(SMALL):
        x = O.gen(KIND[tin], (h, w), 1000 + 17 * k + alg)
        y = sa.integral(to_dev(x, dev), out_dtype=TD[tout], alg=alg)
        assert_sat_close(O, x, to_np(y), tout)


@pytest.mark.parametrize("alg", [0, 1, 2, 3])
@pytest.mark.parametrize("scan", [0, 1])
def test_scan_orders_float(O, dev, sa, alg, scan):
    say(f"[case] scan order alg={alg} scan={scan}")
    x = O.gen("f32", (300, 700), 77)
    y = sa.integral(to_dev(x, dev), alg=alg, scan=scan)
    assert_sat_close(O, x, to_np(y), "f32")


def test_all_ones_closed_form(O, dev, sa):
    """The authors' known answer (integralImage.h:273): all-ones -> (x+1)(y+1)."""
    h, w = 513, 1031
    x = np.ones((h, w), np.uint8)
    exp = ((np.arange(h, dtype=np.uint64)[:, None] + 1) * (np.arange(w, dtype=np.uint64)[None, :] + 1)).astype(np.uint32)
    for alg in range(4):
        say(f"[case] all-ones alg={alg}")
        y = to_np(sa.integral(to_dev(x, dev), alg=alg))
        assert np.array_equal(y, exp)


def test_pitched_buffers(O, dev, sa):
    """Row pitches larger than the width, odd element offsets (no vector path)."""
    h, w = 150, 333
    big = O.gen("f32", (h, w + 29), 5)
    x = to_dev(big, dev)[:, 3:3 + w]
    assert x.stride(0) == w + 29
    for alg in range(4):
        say(f"[case] pitched alg={alg}")
        y = sa.integral(x, alg=alg)
        assert_sat_close(O, big[:, 3:3 + w], to_np(y), "f32")


@pytest.mark.parametrize("tin,tout", PAIRS)
def test_pass_kernels(O, dev, sa, tin, tout):
    say(f"[case] passes {tin}->{tout}")
    ntout = NP[tout]
    for k, (h, w) in enumerate([(1, 1), (70, 300), (257, 65), (64, 512)]):
        x = O.gen(KIND[tin], (h, w), 300 + k)
        xd = to_dev(x, dev)
        rows = O.allrow(x, ntout)
        cols = O.colscan(x, ntout)
        rows_t = O.transpose(rows)
        for name, got, exp in (
            ("scan_x", sa.scan_x(xd, out_dtype=TD[tout]), rows),
            ("scan_y", sa.scan_y(xd, out_dtype=TD[tout]), cols),
            ("block_scan", sa.block_scan(xd, out_dtype=TD[tout]), rows_t),
            ("seriel_scan", sa.seriel_scan(xd, out_dtype=TD[tout]), rows_t),
        ):
            g = to_np(got)
            assert g.shape == exp.shape, name
            if tout in INT_OUT:
                assert np.array_equal(g, exp), f"{name} {h}x{w}"
            else:
                ref = (O.allrow(x, np.float64) if name != "scan_y" else O.colscan(x, np.float64))
                if name in ("block_scan", "seriel_scan"):
                    ref = O.transpose(ref)
                absx = np.abs(x.astype(np.float64))
                sc = O.allrow(absx, np.float64) if name != "scan_y" else O.colscan(absx, np.float64)
                if name in ("block_scan", "seriel_scan"):
                    sc = O.transpose(sc)
                u = 2.0 ** -24 if tout == "f32" else 2.0 ** -53
                err = np.abs(g.astype(np.float64) - ref)
                assert np.all(err <= 32 * u * sc + 1e-300), f"{name} {h}x{w}"


@pytest.mark.parametrize("dt", ["f32", "f64", "u32"])
def test_reverse(O, dev, sa, dt):
    say(f"[case] reverse {dt}")
    for k, (h, w) in enumerate([(1, 1), (77, 300), (64, 256), (300, 1000)]):
        x = O.gen({"f32": "f32", "f64": "f64n", "u32": "u32_20"}[dt], (h, w), 900 + k)
        y = to_np(sa.integral_reverse(to_dev(x, dev)))
        if dt == "u32":
            assert np.array_equal(y, O.reverse(x))
        else:
            ref = O.reverse(x.astype(np.float64))
            sc = O.reverse(np.abs(x.astype(np.float64)))
            u = 2.0 ** -24 if dt == "f32" else 2.0 ** -53
            assert np.all(np.abs(y - ref) <= 32 * u * sc + 1e-300)


def test_planes(O, dev, sa):
    say("[case] planes")
    x = O.gen("f32n", (2, 3, 100, 300), 4242)
    y = to_np(sa.integral(to_dev(x, dev)))
    for n in range(2):
        for c in range(3):
            assert_sat_close(O, x[n, c], y[n, c], "f32")


@pytest.mark.parametrize("tin,tout", [("u8", "u32"), ("i8", "i32"), ("f32", "f32"), ("f64", "f64")])
def test_integral_bordered(O, dev, sa, tin, tout):
    """OpenCV cv::integral layout: (H+1) x (W+1), zero first row and column."""
    say(f"[case] bordered {tin}->{tout}")
    for k, (h, w) in enumerate([(1, 1), (3, 5), (130, 520), (257, 2049)]):
        for alg in ((0, 1, 2, 3) if k == 2 else (3,)):
            x = O.gen(KIND[tin], (h, w), 3000 + 7 * k + alg)
            y = to_np(sa.integral_bordered(to_dev(x, dev), out_dtype=TD[tout], alg=alg))
            assert y.shape == (h + 1, w + 1)
            assert not y[0].any() and not y[:, 0].any()
            assert_sat_close(O, x, np.ascontiguousarray(y[1:, 1:]), tout)


def test_integral_bordered_pitched_planes(O, dev, sa):
    """Row pitch beyond W+1 and several planes; the border is written even
    over garbage, elements past W+1 in each row are left alone."""
    import ctypes as C

    say("[case] bordered pitched planes")
    P, h, w, pitch = 3, 70, 300, 320
    x = O.gen("u8", (P, h, w), 55)
    xd = to_dev(x, dev)
    out = torch.full((P, h + 1, pitch), -1, dtype=torch.int32, device=dev)
    st = sa.lib().sat_integral_bordered_planes(3, 0, 0, 4, xd.data_ptr(), w, h * w, out.data_ptr(), pitch,
                                               (h + 1) * pitch, w, h, P, None, 0,
                                               torch.cuda.current_stream().cuda_stream)
    assert st == 0
    o = out.cpu().numpy().view(np.uint32)
    for p in range(P):
        assert not o[p, 0, :w + 1].any() and not o[p, :, 0].any()
        assert np.array_equal(o[p, 1:, 1:w + 1], O.serial(x[p], np.uint32))
        assert (o[p, :, w + 1:] == 0xFFFFFFFF).all()


@pytest.mark.parametrize("nbytes", [1, 15, 16, 17, 4096, 1_000_003, (64 << 20) + 7])
def test_stream_copy(dev, sa, nbytes):
    """The bandwidth-reference copy moves every byte, tails included."""
    say(f"[stream_copy] {nbytes} bytes")
    g = torch.Generator(device="cpu").manual_seed(nbytes)
    src = torch.randint(0, 256, (nbytes,), dtype=torch.uint8, generator=g).to(dev)
    dst = torch.zeros(nbytes + 64, dtype=torch.uint8, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    sa.check(sa.lib().sat_stream_copy(src.data_ptr(), dst.data_ptr(), nbytes, st), "sat_stream_copy")
    torch.cuda.synchronize()
    assert torch.equal(dst[:nbytes], src)
    assert int(dst[nbytes:].count_nonzero()) == 0


def test_empty_inputs(dev, sa):
    """Empty images are no-ops with outputs of the right shape; the bordered
    layout of an empty image is its zero border alone."""
    say("[case] empty inputs")
    for shape in ((0, 5), (5, 0), (0, 0), (2, 0, 3), (0, 4, 4)):
        x = torch.zeros(shape, dtype=torch.float32, device=dev)
        y = sa.integral(x)
        assert y.shape == x.shape and y.dtype == torch.float32
        assert sa.integral_reverse(x).shape == x.shape
        b = sa.integral_bordered(x.to(torch.uint8), out_dtype=torch.uint32)
        assert b.shape == (*shape[:-2], shape[-2] + 1, shape[-1] + 1)
        if b.numel():
            assert int(b.view(torch.int32).count_nonzero()) == 0


def test_concurrent_streams(O, dev, sa):
    """Calls on different streams run concurrently, each on its own
    library-managed carry workspace (re-entrant across streams)."""
    say("[case] concurrent streams")
    xs = [O.gen("f32", (2048, 3000), 500 + k) for k in range(4)]
    streams = [torch.cuda.Stream() for _ in xs]
    outs = []
    for x, s in zip(xs, streams):
        xd = torch.from_numpy(x).to(dev)
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            for _ in range(3):  # several launches per stream, interleaved with the others
                y = sa.integral(xd)
        outs.append((x, y, xd))
    torch.cuda.synchronize()
    for x, y, _ in outs:
        assert_sat_close(O, x, to_np(y), "f32")
    for s in streams:
        assert sa.lib().sat_workspace_fault(None, s.cuda_stream) == 0


def test_caller_workspace_alignment(O, dev, sa):
    """A caller workspace at a 16-byte offset into a garbage-filled buffer
    (header zeroed) gives the oracle's results for 4- and 8-byte carries
    (8-byte payloads are 16-byte carry pairs), twice in a row; one at an
    8-byte offset is refused with SAT_ERR_INVALID before any launch."""
    import ctypes as C

    say("[case] caller workspace alignment")
    L = sa.lib()
    st = torch.cuda.current_stream().cuda_stream
    for tin, tout, (h, w) in (("f64", "f64", (700, 1500)), ("u32", "u32", (700, 1500)), ("f32", "f64", (300, 900))):
        nb = L.sat_integral_workspace_bytes(3, sa.sat_dtype(TD[tin]), sa.sat_dtype(TD[tout]), w, h)
        raw = torch.full((nb + 64,), 0x5A, dtype=torch.uint8, device=dev)
        base = raw.data_ptr()
        x = O.gen(KIND[tin], (h, w), 8100)
        xd = to_dev(x, dev)
        y = torch.empty((h, w), dtype=TD[tout], device=dev)
        args = lambda ws: (3, 0, sa.sat_dtype(TD[tin]), sa.sat_dtype(TD[tout]), C.c_void_p(xd.data_ptr()), w,
                           C.c_void_p(y.data_ptr()), w, w, h, C.c_void_p(ws), nb, C.c_void_p(st))
        assert base % 16 == 0
        assert L.sat_integral(*args(base + 8)) == 2  # SAT_ERR_INVALID
        ws = base + 16
        raw[16:16 + 256].zero_()
        for _ in range(2):
            y.view(torch.uint8).fill_(0)
            assert L.sat_integral(*args(ws)) == 0
            torch.cuda.synchronize()
            assert_sat_close(O, x, to_np(y), tout)
        assert L.sat_workspace_fault(C.c_void_p(ws), C.c_void_p(st)) == 0
    # the batch and the box gradient refuse an 8-byte-aligned workspace the same way
    x = torch.zeros((40, 300), dtype=torch.float32, device=dev)
    y = torch.empty_like(x)
    img = sa.sat_image(x.data_ptr(), y.data_ptr(), 300, 300, 300, 40, sa.sat_dtype(x.dtype), sa.sat_dtype(y.dtype))
    pimg = C.cast(C.pointer(img), C.c_void_p)
    nb = L.sat_batch_workspace_bytes(pimg, 1)
    raw = torch.zeros(nb + 64, dtype=torch.uint8, device=dev)
    assert L.sat_batch_prepare(pimg, 1, C.c_void_p(raw.data_ptr() + 8), nb, C.c_void_p(st)) == 2
    assert L.sat_batch_prepare(pimg, 1, C.c_void_p(raw.data_ptr() + 16), nb, C.c_void_p(st)) == 0
    boxes = torch.tensor([[0, 0, 9, 9]], dtype=torch.int32, device=dev)
    g = torch.ones(1, dtype=torch.float32, device=dev)
    gx = torch.empty((40, 300), dtype=torch.float32, device=dev)
    nb = L.sat_box_grad_workspace_bytes(sa.sat_dtype(torch.float32), 300, 40, 1, 1)
    raw = torch.zeros(nb + 64, dtype=torch.uint8, device=dev)
    args = lambda off: (sa.sat_dtype(torch.float32), C.c_void_p(boxes.data_ptr()), 1, C.c_void_p(g.data_ptr()),
                        C.c_void_p(gx.data_ptr()), 300, 300 * 40, 300, 40, 1, C.c_void_p(raw.data_ptr() + off), nb,
                        C.c_void_p(st))
    assert L.sat_box_grad(*args(8)) == 2
    assert L.sat_box_grad(*args(16)) == 0
    torch.cuda.synchronize()
    exp = torch.zeros_like(gx)
    exp[:10, :10] = 1
    assert torch.equal(gx, exp)


def test_random_shapes_and_pitches(O, dev, sa):
    """A seeded sweep of random shapes, input row pitches and offsets, dtype
    pairs and algorithms (through the C ABI with explicit pitches), each
    against the oracle: integers bit-exact, floats within the bound."""
    import ctypes as C

    rng = np.random.default_rng(20261203)
    L = sa.lib()
    st = torch.cuda.current_stream().cuda_stream
    for case in range(48):
        tin, tout = PAIRS[int(rng.integers(len(PAIRS)))]
        alg = int(rng.integers(4))
        h, w = int(rng.integers(1, 700)), int(rng.integers(1, 1100))
        pad, off = int(rng.integers(0, 40)), int(rng.integers(0, 8))
        say(f"[case] random #{case}: {tin}->{tout} alg={alg} {h}x{w} pitch {w + pad + off} offset {off}")
        big = O.gen(KIND[tin], (h, w + pad + off), 7000 + case)
        xd = to_dev(big, dev)
        y = torch.empty((h, w), dtype=TD[tout], device=dev)
        esz = big.dtype.itemsize
        st_code = L.sat_integral(alg, 0, sa.sat_dtype(TD[tin]), sa.sat_dtype(TD[tout]),
                                 C.c_void_p(xd.data_ptr() + off * esz), w + pad + off, C.c_void_p(y.data_ptr()), w,
                                 w, h, None, 0, C.c_void_p(st))
        assert st_code == 0, sa.STATUS.get(st_code, st_code)
        torch.cuda.synchronize()
        assert_sat_close(O, np.ascontiguousarray(big[:, off:off + w]), to_np(y), tout)


def test_random_pass_shapes(O, dev, sa):
    """Seeded random shapes for the four pass-level drop-ins (blockScan and
    serielScan write the transposed row scan, ScanX the row scan, ScanY the
    column scan), integer pairs bit-exact against the oracle's passes and
    float pairs within the bound."""
    rng = np.random.default_rng(20261204)
    for case in range(24):
        tin, tout = PAIRS[int(rng.integers(len(PAIRS)))]
        h, w = int(rng.integers(1, 600)), int(rng.integers(1, 900))
        say(f"[case] random passes #{case}: {tin}->{tout} {h}x{w}")
        x = O.gen(KIND[tin], (h, w), 9000 + case)
        xd = to_dev(x, dev)
        ntout = NP[tout]
        rows = O.allrow(x, ntout)
        for name, got, exp, col in (
            ("scan_x", sa.scan_x(xd, out_dtype=TD[tout]), rows, False),
            ("scan_y", sa.scan_y(xd, out_dtype=TD[tout]), O.colscan(x, ntout), True),
            ("block_scan", sa.block_scan(xd, out_dtype=TD[tout]), O.transpose(rows), False),
            ("seriel_scan", sa.seriel_scan(xd, out_dtype=TD[tout]), O.transpose(rows), False),
        ):
            g = to_np(got)
            assert g.shape == exp.shape, name
            if tout in INT_OUT:
                assert np.array_equal(g, exp), f"{name} {h}x{w}"
                continue
            absx = np.abs(x.astype(np.float64))
            ref = O.colscan(x, np.float64) if col else O.allrow(x, np.float64)
            sc = O.colscan(absx, np.float64) if col else O.allrow(absx, np.float64)
            if name in ("block_scan", "seriel_scan"):
                ref, sc = O.transpose(ref), O.transpose(sc)
            u = 2.0 ** -24 if tout == "f32" else 2.0 ** -53
            assert np.all(np.abs(g.astype(np.float64) - ref) <= 32 * u * sc + 1e-300), f"{name} {h}x{w}"


@pytest.mark.parametrize("tin", ["f32", "f64", "u8"])
def test_f64_outputs_odd_pitch_and_offset(O, dev, sa, tin):
    """8-byte outputs at an odd row pitch and a base one element past a
    16-byte boundary (rows only 8-byte aligned), on shapes with interior
    tiles, every algorithm: the results match the oracle and the padding
    around the image is never written."""
    import ctypes as C

    L = sa.lib()
    st = torch.cuda.current_stream().cuda_stream
    h, w, pad = 300, 777, 6
    pitch = w + pad + 1  # odd
    x = O.gen(KIND[tin], (h, w), 7300 + len(tin))
    xd = to_dev(x, dev)
    for alg in range(4):
        say(f"[case] {tin}->f64 alg={alg} {h}x{w} out pitch {pitch} offset 1")
        buf = torch.full((h, pitch), float("nan"), dtype=torch.float64, device=dev)
        code = L.sat_integral(alg, 0, sa.sat_dtype(TD[tin]), sa.sat_dtype(torch.float64), C.c_void_p(xd.data_ptr()),
                              w, C.c_void_p(buf.data_ptr() + 8), pitch, w, h, None, 0, C.c_void_p(st))
        assert code == 0, sa.STATUS.get(code, code)
        torch.cuda.synchronize()
        got = buf.cpu().numpy()
        assert np.isnan(got[:, 0]).all() and np.isnan(got[:, w + 1:]).all(), "padding written"
        assert_sat_close(O, x, np.ascontiguousarray(got[:, 1:w + 1]), "f64")

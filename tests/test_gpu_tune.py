"""GPU checks of the tile configurations: every configuration of the one-pass
SAT, the reverse SAT and the batch kernel matches the oracle (integer pairs
bit-exact, floats within the FP64 bound) — the result invariance the tuner
relies on — and the tuner's own measurement loop runs."""
from __future__ import annotations

import ctypes as C
import sys

import numpy as np
import pytest
import torch

from helpers import KIND, assert_sat_close
from test_gpu_sat import PAIRS, TD, to_dev, to_np

pytestmark = pytest.mark.gpu

SHAPES = [(1, 1), (3, 5), (64, 256), (130, 520), (257, 2049), (4100, 300), (96, 4100)]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def sa():
    import sat_amd

    L = sat_amd.lib()
    yield sat_amd
    from sat_amd.tune import reset_table

    reset_table(L)


def force(L, in_bytes, out_bytes, w, h, p, cfg):
    cx, cy, cp = C.c_int(), C.c_int(), C.c_int()
    L.sat_tune_class(w, h, p, C.byref(cx), C.byref(cy), C.byref(cp))
    assert L.sat_tune_set(in_bytes, out_bytes, cx.value, cy.value, cp.value, cfg) == 0
    assert L.sat_tune_pick(in_bytes, out_bytes, w, h, p) == cfg


CFGS = [0, 1, 2, 3, 4, 5, 6, 7]
BATCH_CFGS = [0, 1, 2, 3, 4]
IN_BYTES = {"u8": 1, "i8": 1, "u32": 4, "i32": 4, "f32": 4, "f64": 8}


@pytest.mark.parametrize("cfg", CFGS)
@pytest.mark.parametrize("tin,tout", PAIRS)
def test_integral_every_config(O, dev, sa, cfg, tin, tout):
    say(f"[case] cfg={cfg} {tin}->{tout}")
    L = sa.lib()
    ob = 8 if tout in ("f64",) else 4
    for k, (h, w) in enumerate(SHAPES):
        force(L, IN_BYTES[tin], ob, w, h, 1, cfg)
        x = O.gen(KIND[tin], (h, w), 4000 + 31 * k + cfg)
        y = sa.integral(to_dev(x, dev), out_dtype=TD[tout])
        assert_sat_close(O, x, to_np(y), tout)


@pytest.mark.parametrize("cfg", CFGS)
def test_planes_and_reverse_every_config(O, dev, sa, cfg):
    say(f"[case] planes + reverse cfg={cfg}")
    L = sa.lib()
    P, h, w = 5, 300, 700
    x = O.gen("u32_20", (P, h, w), 91 + cfg)
    force(L, 4, 4, w, h, P, cfg)
    y = to_np(sa.integral(to_dev(x, dev)))
    for p in range(P):
        assert np.array_equal(y[p], O.serial(x[p], np.uint32))
    xf = O.gen("f64n", (P, h, w), 92 + cfg)
    force(L, 8, 8, w, h, P, cfg)
    r = sa.integral_reverse(torch.from_numpy(xf).to(dev)).cpu().numpy()
    for p in range(P):
        flip = xf[p, ::-1, ::-1]
        assert_sat_close(O, np.ascontiguousarray(flip), np.ascontiguousarray(r[p, ::-1, ::-1]), "f64")


@pytest.mark.parametrize("cfg", BATCH_CFGS)
def test_batch_every_config(O, dev, sa, cfg):
    say(f"[case] batch cfg={cfg}")
    L = sa.lib()
    assert L.sat_batch_tune_set(4, cfg) == 0
    rng = np.random.default_rng(40 + cfg)
    pairs, exp = [], []
    for i in range(16):
        h, w = (int(v) for v in rng.integers(1, 900, 2))
        x = O.gen("u8", (h, w), 700 + 13 * i + cfg)
        xd, yd = torch.from_numpy(x).to(dev), torch.empty((h, w), dtype=torch.uint32, device=dev)
        pairs.append((xd, yd))
        exp.append(O.serial(x, np.uint32))
    plan = sa.BatchPlan(pairs)
    plan.run()
    torch.cuda.synchronize()
    for (xd, yd), e in zip(pairs, exp):
        assert np.array_equal(to_np(yd), e)
    assert L.sat_batch_tune_set(4, -1) == 0


def test_tuner_quick(sa, tmp_path):
    say("[case] tuner loop on two small shapes")
    from sat_amd import tune

    doc = tune.tune([(512, 2048, 1), (4096, 256, 1)], widths=((1, 4), (4, 4)), batch_out=(4,), batch_images=8,
                    reps=2, log=lambda *m: say(*m))
    ncfg = sa.lib().sat_tile_config_count()
    assert len(doc["entries"]) == 4 and all(0 <= e["cfg"] < ncfg for e in doc["entries"])
    assert {e["in_bytes"] for e in doc["entries"]} == {1, 4}
    assert doc["batch"][0]["cfg"] in BATCH_CFGS
    p = tmp_path / "tuned.json"
    import json

    p.write_text(json.dumps(doc))
    assert tune.apply_table(p) == 5


# The packed 1-byte tiles (configurations 5 and 7): partial chunks at the right edge
# (widths not a multiple of 4, below 4), partial 256- and 128-row tiles, rows that
# start at any byte (pitch and offset), negative i8 values, wraparound.
PACKED_CASES = [((1, 1), 1, 0), ((2, 3), 3, 1), ((5, 7), 9, 2), ((300, 257), 260, 3), ((257, 1030), 1033, 1),
                ((1000, 77), 80, 0), ((513, 4099), 4100, 0)]


@pytest.mark.parametrize("cfg", [5, 7])
@pytest.mark.parametrize("tin,tout", [("u8", "u32"), ("i8", "i32"), ("u8", "f32"), ("i8", "f32")])
def test_packed_edges_pitches_offsets(O, dev, sa, tin, tout, cfg):
    say(f"[case] packed tiles cfg {cfg} {tin}->{tout}")
    L = sa.lib()
    for k, ((h, w), pitch, off) in enumerate(PACKED_CASES):
        force(L, 1, 4, w, h, 1, cfg)
        x = O.gen(KIND[tin], (h, w), 5100 + k)
        buf = np.zeros(off + h * pitch + 8, dtype=x.dtype)
        buf[off:off + h * pitch].reshape(h, pitch)[:, :w] = x
        xd = torch.from_numpy(buf).to(dev)
        y = torch.empty((h, w), dtype=TD[tout], device=dev)
        tdt = {"u8": sa.U8, "i8": sa.I8}[tin]
        odt = {"u32": sa.U32, "i32": sa.I32, "f32": sa.F32}[tout]
        st = L.sat_integral(3, 0, tdt, odt, xd.data_ptr() + off, pitch, y.data_ptr(), w, w, h, None, 0,
                            torch.cuda.current_stream().cuda_stream)
        assert st == 0
        assert_sat_close(O, x, to_np(y), tout)
    sa.check_faults()


@pytest.mark.parametrize("cfg", [5, 7])
def test_packed_all_255_wraps(dev, sa, cfg):
    say(f"[case] packed tiles cfg {cfg} u8 all 255, 4500 x 4500 (wraps modulo 2^32)")
    L = sa.lib()
    H = W = 4500
    force(L, 1, 4, W, H, 1, cfg)
    x = torch.full((H, W), 255, dtype=torch.uint8, device=dev)
    y = sa.integral(x, out_dtype=torch.uint32).view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    i1 = torch.arange(1, H + 1, dtype=torch.int64, device=dev)[:, None]
    j1 = torch.arange(1, W + 1, dtype=torch.int64, device=dev)[None, :]
    assert torch.equal(y, (255 * i1 * j1) & 0xFFFFFFFF)
    xs = torch.full((H, W), -128, dtype=torch.int8, device=dev)
    ys = sa.integral(xs, out_dtype=torch.int32).to(torch.int64) & 0xFFFFFFFF
    assert torch.equal(ys, (-128 * i1 * j1) & 0xFFFFFFFF)


# Images of one tile column and at least 64 tile rows take a second vertical
# carry level (runs of 16 tile rows) for 4-byte payloads in the
# configurations whose vertical lanes follow one column each (1, 2, 3): tall
# shapes with partial last runs and tiles, one-pixel-wide images, planes, the
# reverse SAT and the column pass of algorithm 2; 8-byte outputs and two tile
# columns (one carry level) alongside.
TALL = [((5197, 256), 2), ((5197, 1), 2), ((9000, 255), 1), ((4100, 130), 3), ((16384 + 77, 200), 1),
        ((5197, 300), 2)]


@pytest.mark.parametrize("tin,tout", [("u32", "u32"), ("f32", "f32"), ("f64", "f64"), ("u8", "u32")])
def test_tall_two_level_chains(O, dev, sa, tin, tout):
    L = sa.lib()
    ob = 8 if tout == "f64" else 4
    for k, ((h, w), cfg) in enumerate(TALL):
        say(f"[case] two-level chains {tin}->{tout} {h}x{w} cfg {cfg}")
        force(L, IN_BYTES[tin], ob, w, h, 1, cfg)
        x = O.gen(KIND[tin], (h, w), 6100 + k)
        y = sa.integral(to_dev(x, dev), out_dtype=TD[tout])
        assert_sat_close(O, x, to_np(y), tout)
    sa.check_faults()


def test_tall_two_level_planes_reverse_and_column_pass(O, dev, sa):
    L = sa.lib()
    P, h, w = 3, 5197, 250
    say("[case] two-level chains: planes, reverse SAT, algorithm 2")
    x = O.gen("u32_20", (P, h, w), 6200)
    force(L, 4, 4, w, h, P, 2)
    y = to_np(sa.integral(to_dev(x, dev)))
    for p in range(P):
        assert np.array_equal(y[p], O.serial(x[p], np.uint32))
    for kind, dt, cfg in (("f32", "f32", 2), ("f32", "f32", 1), ("f64n", "f64", 1)):
        xf = O.gen(kind, (P, h, w), 6201)
        force(L, IN_BYTES[dt], IN_BYTES[dt], w, h, P, cfg)
        r = to_np(sa.integral_reverse(torch.from_numpy(xf).to(dev)))
        for p in range(P):
            flip = np.ascontiguousarray(xf[p, ::-1, ::-1])
            assert_sat_close(O, flip, np.ascontiguousarray(r[p, ::-1, ::-1]), dt)
    xr = O.gen("u32_20", (h, w), 6203)
    force(L, 4, 4, w, h, 1, 2)
    r = to_np(sa.integral_reverse(to_dev(xr, dev)))
    assert np.array_equal(r, O.reverse(xr))
    xu = O.gen("u32_20", (9000, 256), 6202)
    assert np.array_equal(to_np(sa.integral(to_dev(xu, dev), alg=2)), O.serial(xu, np.uint32))
    sa.check_faults()


@pytest.mark.parametrize("h", [20000 + 3, 131072 + 5])
@pytest.mark.parametrize("cfg", [0, 1, 2, 3])
def test_tall_caller_workspace_exact_size(O, dev, sa, cfg, h):
    """A caller workspace of exactly sat_integral_workspace_bytes, garbage
    after its zeroed header, serves a one-column image in every
    configuration, twice: with the second carry level (20003 rows in
    configurations 1-3) and without it (configuration 0, and 131077 rows,
    more tiles than resident workgroups, whose workspace still has room for
    the run words)."""
    import ctypes as C

    say(f"[case] tall image {h} x 256, caller workspace of the queried size, cfg {cfg}")
    L = sa.lib()
    w = 256
    force(L, 4, 4, w, h, 1, cfg)
    x = O.gen("u32_20", (h, w), 6300 + cfg)
    xd = to_dev(x, dev)
    y = torch.empty((h, w), dtype=torch.uint32, device=dev)
    nb = L.sat_integral_workspace_bytes(3, sa.U32, sa.U32, w, h)
    ws = torch.full((nb,), 0xA5, dtype=torch.uint8, device=dev)
    ws[:256].zero_()
    st = torch.cuda.current_stream().cuda_stream
    for _ in range(2):
        y.view(torch.uint8).zero_()
        assert L.sat_integral(3, 0, sa.U32, sa.U32, xd.data_ptr(), w, y.data_ptr(), w, w, h,
                              C.c_void_p(ws.data_ptr()), nb, C.c_void_p(st)) == 0
        torch.cuda.synchronize()
        assert np.array_equal(to_np(y), O.serial(x, np.uint32))
    assert L.sat_workspace_fault(C.c_void_p(ws.data_ptr()), C.c_void_p(st)) == 0


@pytest.mark.parametrize("cfg", [0, 1, 2, 5])
@pytest.mark.parametrize("tin,tout", [("u8", "u32"), ("f32", "f32"), ("f64", "f64")])
def test_one_row_and_one_column_extremes(O, dev, sa, cfg, tin, tout):
    """A single row of a million pixels (one tile row: a row chain of
    thousands of tiles, partial tile rows) and a single column of a million
    (one-pixel-wide tiles, a column chain of thousands of tiles in
    generations), in several configurations."""
    L = sa.lib()
    ob = 8 if tout == "f64" else 4
    for k, (h, w) in enumerate([(1, 1000003), (1000003, 1), (3, 333333)]):
        say(f"[case] extreme shape {h}x{w} {tin}->{tout} cfg {cfg}")
        force(L, IN_BYTES[tin], ob, w, h, 1, cfg)
        x = O.gen(KIND[tin], (h, w), 6400 + k)
        y = sa.integral(to_dev(x, dev), out_dtype=TD[tout])
        assert_sat_close(O, x, to_np(y), tout)
    sa.check_faults()

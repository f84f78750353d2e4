"""Autotuner for the tile configuration of the one-pass SAT and the batch kernel.

The library picks one of ``sat_tile_config_count()`` wave64 tile
configurations (tile shape and waves per SIMD) per launch from a table indexed
by input width, output width and shape class (``sat_tune_class``). This module measures every
configuration on the current device for representative shapes of each class,
checks that the configurations agree (integer pairs bit-for-bit, floating
point within the summation-order tolerance), and writes the fastest choice per
class as JSON; ``apply_table`` loads such a file into the library (the package
does so at load time for ``tuned_gfx950.json``).

    python -m sat_amd.tune                 # full table -> sat_amd/tuned_gfx950.json
    python -m sat_amd.tune --quick --out /tmp/t.json

The reference has no tuner: its block shapes are template parameters fixed by
the caller (e.g. ScanRow-BRLT.cu:24, ScanRowColumn.cu:16,144).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import sys
from pathlib import Path

CFG_NAMES = {0: "4 waves x 16 rows, 64x256 (32x256 for 8-byte outputs), 4 waves per SIMD",
             1: "8 waves x 16 rows, 128x256 (64x256), 4 waves per SIMD",
             2: "8 waves x 8 rows, 64x256 (32x256), up to 6 waves per SIMD",
             3: "8 waves x 8 rows, 64x256 (32x256), 8 waves per SIMD",
             4: "4 waves x 24 rows, 96x256 (48x256), 3 waves per SIMD",
             5: "1-byte inputs into 4-byte outputs: 4 waves x 64 packed rows, 256x256; other pairs as 1",
             6: "4 waves x 32 rows, 128x256 (64x256), 2 waves per SIMD",
             7: "1-byte inputs into 4-byte outputs: 4 waves x 32 packed rows, 128x256; other pairs as 0"}

# the packed configurations (1-byte inputs) and the configuration whose tile
# they run for wider inputs
PACKED_ALIAS = {5: 1, 7: 0}

# the batch kernel's configurations (0-4; the packed configurations are
# one-pass only)
BATCH_CFGS = 5

# (input bytes, output bytes) pairs the table is measured for, with the dtype
# pair timed for each: u8 -> u32, f32 -> f32, f64 -> f64. Other pairs take
# their output width's same-width entry.
WIDTHS = ((1, 4), (4, 4), (8, 8))

# (height, width, planes) shapes that stand for their class; the benchmark
# workloads (f32 8192^2, u8->u32 16384^2, f64 4096^2, 65536x256, 256x65536,
# 96 planes of 1024^2, ...) plus square, tall, wide and plane-stacked sweeps
FULL_SHAPES = [
    (2048, 2048, 1), (4096, 4096, 1), (8192, 8192, 1), (16384, 16384, 1),
    (1024, 1024, 1), (512, 512, 1), (256, 256, 1),
    (65536, 256, 1), (16384, 256, 1), (65536, 1024, 1), (16384, 1024, 1), (8192, 512, 1),
    (256, 65536, 1), (1024, 16384, 1), (2048, 32768, 1),
    (1024, 1024, 96), (512, 512, 64), (64, 64, 1024),
]
QUICK_SHAPES = [(8192, 8192, 1), (65536, 256, 1), (1024, 1024, 96)]


def classes_of(L, width: int, height: int, planes: int):
    cx, cy, cp = C.c_int(), C.c_int(), C.c_int()
    L.sat_tune_class(width, height, planes, C.byref(cx), C.byref(cy), C.byref(cp))
    return cx.value, cy.value, cp.value


def reset_table(L) -> None:
    """Every class back to unset (the library's built-in choice)."""
    for ob in (4, 8):
        L.sat_batch_tune_set(ob, -1)
        for ib in (1, 4, 8):
            if ib > ob:
                continue
            for cx in range(9):
                for cy in range(12):
                    for cp in range(4):
                        L.sat_tune_set(ib, ob, cx, cy, cp, -1)


def apply_table(path) -> int:
    """Load a tuned table (JSON written by :func:`tune`) into the library;
    returns the number of entries applied."""
    from . import lib

    L = lib()
    doc = json.loads(Path(path).read_text())
    n = 0
    for e in doc.get("entries", []):
        if L.sat_tune_set(int(e["in_bytes"]), int(e["out_bytes"]), int(e["class_x"]), int(e["class_y"]),
                          int(e["class_p"]), int(e["cfg"])) == 0:
            n += 1
    for e in doc.get("batch", []):
        if L.sat_batch_tune_set(int(e["out_bytes"]), int(e["cfg"])) == 0:
            n += 1
    return n


def _time(fn, reps: int, nbuf: int = 1) -> float:
    """Median microseconds of `reps` back-to-back groups of 5 calls fn(k),
    k cycling over nbuf buffer sets."""
    import torch

    fn(0)
    torch.cuda.synchronize()
    ts, k = [], 0
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        for _ in range(5):
            k += 1
            fn(k % nbuf)
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e) * 1000 / 5)
    ts.sort()
    return ts[len(ts) // 2]


_GRAPH_STREAM = None


def _time_graph(fn, reps: int, nbuf: int = 1) -> float:
    """Median microseconds per call of `reps` replays of a hipGraph holding
    max(8, nbuf) calls fn(k), k cycling over nbuf buffer sets. A replay
    launches every call without host work in between, so images whose
    kernels run shorter than a call's host path (a few tens of
    microseconds) are timed on the GPU rather than on the host. Captured
    calls advance the workspace's device epoch with a one-workgroup kernel
    before each launch, the same for every tile configuration."""
    import torch

    global _GRAPH_STREAM
    n = max(8, nbuf)
    if _GRAPH_STREAM is None:
        _GRAPH_STREAM = torch.cuda.Stream()  # one capture stream: one managed workspace
    st = _GRAPH_STREAM
    st.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(st):
        for k in range(nbuf):
            fn(k)  # the capture stream's managed workspace exists before capture
    torch.cuda.current_stream().wait_stream(st)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=st):
        for k in range(n):
            fn(k % nbuf)
    g.replay()
    torch.cuda.synchronize()
    ts = []
    for _ in range(reps):
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        g.replay()
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e) * 1000 / n)
    del g
    ts.sort()
    return ts[len(ts) // 2]


# timed calls rotate over input/output sets spanning at least this many bytes,
# so that no call finds its input in the 256 MiB MALL
ROTATE_BYTES = 1 << 30


def _tune_shape(L, sa, H, W, P, in_bytes, out_bytes, reps, log):
    import torch

    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(H * 131 + W * 7 + P + in_bytes)
    cx, cy, cp = classes_of(L, W, H, P)
    shape = (P, H, W)
    if (in_bytes, out_bytes) == (1, 4):
        # u8 -> u32: timed and compared bit for bit
        x = torch.randint(0, 256, shape, generator=g, dtype=torch.uint8).to(dev)
        out_dt, exact = torch.uint32, True
    elif out_bytes == 4:
        x = torch.rand(shape, generator=g).to(dev)
        out_dt, exact = torch.float32, False
    else:
        # integer values: every partial sum is exact in f64
        x = torch.randint(-(1 << 20), 1 << 20, shape, generator=g, dtype=torch.int32).to(dev).double()
        out_dt, exact = torch.float64, True
    nbuf = max(1, min(16, -(-ROTATE_BYTES // (x.numel() * (in_bytes + out_bytes)))))
    xs = [x] + [x.clone() for _ in range(nbuf - 1)]
    ys = [torch.empty(shape, dtype=out_dt, device=dev) for _ in range(nbuf)]
    times, ref = [], None
    for cfg in range(L.sat_tile_config_count()):
        if in_bytes != 1 and cfg in PACKED_ALIAS:
            # the packed configurations run their unpacked tile for wider
            # inputs: the same kernel, so the same time (ties pick the lower)
            times.append(times[PACKED_ALIAS[cfg]])
            continue
        sa.check(L.sat_tune_set(in_bytes, out_bytes, cx, cy, cp, cfg), "sat_tune_set")
        assert L.sat_tune_pick(in_bytes, out_bytes, W, H, P) == cfg
        sa.integral(x, out_dtype=out_dt, out=ys[0])
        torch.cuda.synchronize()
        got = ys[0].view(torch.int32) if out_dt == torch.uint32 else ys[0]
        if ref is None:
            ref = got.clone()
        elif exact:
            if not torch.equal(got, ref):
                raise AssertionError(f"tile configuration {cfg} changes exact results at {P}x{H}x{W}")
        else:
            scale = torch.cumsum(torch.cumsum(x.abs().double(), -1), -2)
            err = float(((got.double() - ref.double()).abs() / scale.clamp_min(1e-300)).max())
            if err > 64 * 2.0 ** -24:
                raise AssertionError(f"tile configuration {cfg} differs by {err:.2e} (relative) at {P}x{H}x{W}")
        times.append(_time_graph(lambda k: sa.integral(xs[k], out_dtype=out_dt, out=ys[k]), reps, nbuf))
    sa.check_faults()
    best = min(range(len(times)), key=lambda c: times[c])
    log(f"[tune] {P}x{H}x{W} {in_bytes}->{out_bytes}B class ({cx},{cy},{cp}): "
        + "  ".join(f"cfg{c} {t:8.1f} us" for c, t in enumerate(times)) + f"  -> cfg{best}")
    return {"in_bytes": in_bytes, "out_bytes": out_bytes, "class_x": cx, "class_y": cy, "class_p": cp, "cfg": best,
            "shape": [H, W], "planes": P, "us": [round(t, 2) for t in times]}


def _tune_batch(L, sa, out_bytes, n, reps, log):
    import numpy as np
    import torch

    dev = torch.device("cuda:0")
    rng = np.random.default_rng(20260400 + out_bytes)
    pairs = []
    for i in range(n):
        h, w = (int(v) for v in rng.integers(64, 2049, 2))
        if out_bytes == 4:
            xd = torch.rand((h, w), device=dev) if i % 2 else torch.randint(0, 256, (h, w), dtype=torch.uint8,
                                                                           device=dev)
            yd = torch.empty((h, w), dtype=torch.float32 if i % 2 else torch.uint32, device=dev)
        else:
            xd = torch.randn((h, w), dtype=torch.float64, device=dev)
            yd = torch.empty((h, w), dtype=torch.float64, device=dev)
        pairs.append((xd, yd))
    times = []
    outs = None
    for cfg in range(BATCH_CFGS):
        sa.check(L.sat_batch_tune_set(out_bytes, cfg), "sat_batch_tune_set")
        plan = sa.BatchPlan(pairs)
        plan.run()
        torch.cuda.synchronize()
        ints = [yd.clone() for (xd, yd) in pairs if yd.dtype == torch.uint32]
        if outs is None:
            outs = ints
        else:
            for a, b in zip(outs, ints):
                if not torch.equal(a.view(torch.int32), b.view(torch.int32)):
                    raise AssertionError(f"batch tile configuration {cfg} changes integer results")
        times.append(_time(lambda k: plan.run(), reps))
    best = min(range(len(times)), key=lambda c: times[c])
    log(f"[tune] batch of {n} out {out_bytes}B: " + "  ".join(f"cfg{c} {t:8.1f} us" for c, t in enumerate(times))
        + f"  -> cfg{best}")
    return {"out_bytes": out_bytes, "cfg": best, "images": n, "us": [round(t, 2) for t in times]}


def tune(shapes=None, widths=WIDTHS, batch_out=(4, 8), batch_images=512, reps=7, log=print) -> dict:
    """Measure every tile configuration per (input width, output width, shape
    class) on cuda:0 and return the table (classes not covered keep the
    fallbacks)."""
    import torch

    import sat_amd as sa

    L = sa.lib()
    shapes = FULL_SHAPES if shapes is None else shapes
    measured = {}
    try:
        for ib, ob in widths:
            for (H, W, P) in shapes:
                if H * W * P * (ib + ob) > (5 << 30):
                    continue  # keep each sweep image pair under 5 GiB
                e = _tune_shape(L, sa, H, W, P, ib, ob, reps, log)
                measured.setdefault((ib, ob, e["class_x"], e["class_y"], e["class_p"]), []).append(e)
        batch = [_tune_batch(L, sa, ob, batch_images, reps, log) for ob in batch_out] if batch_images else []
    finally:
        reset_table(L)
    torch.cuda.synchronize()
    entries = {}
    for key, es in measured.items():
        # several shapes in one class: the configuration with the least summed
        # slowdown relative to each shape's best
        ncfg = len(es[0]["us"])
        cost = [sum(e["us"][c] / min(e["us"]) for e in es) for c in range(ncfg)]
        best = min(range(ncfg), key=lambda c: cost[c])
        entries[key] = {"in_bytes": key[0], "out_bytes": key[1], "class_x": key[2], "class_y": key[3],
                        "class_p": key[4], "cfg": best,
                        "shapes": [[e["planes"], e["shape"][0], e["shape"][1]] for e in es],
                        "us": [e["us"] for e in es]}
    return {"arch": torch.cuda.get_device_properties(0).gcnArchName.split(":")[0],
            "generated_by": "python -m sat_amd.tune", "configs": CFG_NAMES,
            "entries": [entries[k] for k in sorted(entries)],
            "batch": batch}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=str(Path(__file__).resolve().parent / "tuned_gfx950.json"))
    ap.add_argument("--quick", action="store_true", help="three shapes, 4-byte outputs, small batch")
    ap.add_argument("--reps", type=int, default=7)
    a = ap.parse_args(argv)
    log = lambda *m: print(*m, file=sys.stderr, flush=True)  # noqa: E731
    if a.quick:
        doc = tune(QUICK_SHAPES, widths=((1, 4), (4, 4)), batch_out=(4,), batch_images=64, reps=a.reps, log=log)
    else:
        doc = tune(reps=a.reps, log=log)
    Path(a.out).write_text(json.dumps(doc, indent=1) + "\n")
    log(f"[tune] wrote {a.out}: {len(doc['entries'])} classes, {len(doc['batch'])} batch entries")


if __name__ == "__main__":
    main()

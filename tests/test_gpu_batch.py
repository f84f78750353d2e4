"""GPU parity of the batched queue (images of uneven sizes, mixed dtype pairs,
one launch) against the oracle, including hipGraph capture and replay."""
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


def _u32(t):
    return t.view(torch.int32).cpu().numpy().view(np.uint32)


def make_batch(O, dev, n, lo, hi, seed, kinds=("u8", "f32")):
    rng = np.random.default_rng(seed)
    imgs = []
    for i in range(n):
        h, w = (int(v) for v in rng.integers(lo, hi + 1, 2))
        kind = kinds[i % len(kinds)]
        if kind == "u8":
            x = O.gen("u8", (h, w), seed * 1000 + i)
            xd = torch.from_numpy(x).to(dev)
            yd = torch.empty((h, w), dtype=torch.uint32, device=dev)
        else:
            x = O.gen("f32", (h, w), seed * 1000 + i)
            xd = torch.from_numpy(x).to(dev)
            yd = torch.empty((h, w), dtype=torch.float32, device=dev)
        imgs.append((kind, x, xd, yd))
    return imgs


def check_batch(O, imgs):
    for kind, x, _, yd in imgs:
        if kind == "u8":
            assert np.array_equal(_u32(yd), O.serial(x, np.uint32)), x.shape
        else:
            got = yd.cpu().numpy().astype(np.float64)
            ref = O.sat_f64(x)
            assert np.all(np.abs(got - ref) <= 32 * 2.0 ** -24 * ref + 1e-30), x.shape


def test_batch_mixed_uneven(O, dev, sa):
    say("[case] batch 40 images 1..700 mixed u8->u32 / f32")
    imgs = make_batch(O, dev, 40, 1, 700, 7)
    plan = sa.BatchPlan([(xd, yd) for _, _, xd, yd in imgs])
    plan.run()
    check_batch(O, imgs)
    for _, _, _, yd in imgs:
        yd.zero_()
    plan.run()  # a second run of the same plan (new epoch)
    check_batch(O, imgs)
    assert sa.lib().sat_workspace_fault(plan.workspace.data_ptr(), None) == 0


def test_batch_other_pairs(O, dev, sa):
    say("[case] batch i8->i32, i32->i32, u32->u32")
    rng = np.random.default_rng(3)
    pairs = []
    exp = []
    for i in range(12):
        h, w = (int(v) for v in rng.integers(1, 400, 2))
        k = i % 3
        if k == 0:
            x = O.gen("i8", (h, w), 50 + i)
            xd, yd = torch.from_numpy(x).to(dev), torch.empty((h, w), dtype=torch.int32, device=dev)
            exp.append((yd, O.serial(x, np.int32)))
        elif k == 1:
            x = O.gen("i32_16", (h, w), 50 + i)
            xd, yd = torch.from_numpy(x).to(dev), torch.empty((h, w), dtype=torch.int32, device=dev)
            exp.append((yd, O.serial(x, np.int32)))
        else:
            x = O.gen("u32", (h, w), 50 + i)
            xd = torch.from_numpy(x.view(np.int32)).to(dev).view(torch.uint32)
            yd = torch.empty((h, w), dtype=torch.uint32, device=dev)
            exp.append((yd, O.serial(x, np.uint32)))
        pairs.append((xd, yd))
    plan = sa.BatchPlan(pairs)
    plan.run()
    for yd, e in exp:
        g = _u32(yd) if yd.dtype == torch.uint32 else yd.cpu().numpy()
        assert np.array_equal(g, e)


def test_batch_f64_outputs(O, dev, sa):
    say("[case] batch f32->f64 and f64->f64")
    pairs, refs = [], []
    for i, (h, w) in enumerate([(33, 700), (300, 41), (129, 257)]):
        x = O.gen("f64n", (h, w), 70 + i)
        xd = torch.from_numpy(x).to(dev)
        yd = torch.empty((h, w), dtype=torch.float64, device=dev)
        pairs.append((xd, yd))
        refs.append((yd, x))
    plan = sa.BatchPlan(pairs)
    plan.run()
    for yd, x in refs:
        got = yd.cpu().numpy()
        sc = O.serial(np.abs(x), np.float64)
        assert np.all(np.abs(got - O.sat_f64(x)) <= 32 * 2.0 ** -53 * sc + 1e-300)


def test_batch_graph_replay(O, dev, sa):
    """Capture the run in a hipGraph and replay it several times with new
    inputs in the same buffers and eager work in between; every replay must
    match."""
    say("[case] batch graph capture + replay")
    imgs = make_batch(O, dev, 12, 20, 600, 11)
    plan = sa.BatchPlan([(xd, yd) for _, _, xd, yd in imgs])
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        plan.run()  # warm-up on the capture stream
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        plan.run()
    for rep in range(4):
        fresh = []
        for i, (kind, x, xd, yd) in enumerate(imgs):
            x2 = O.gen("u8" if kind == "u8" else "f32", x.shape, 900 + 37 * rep + i)
            xd.copy_(torch.from_numpy(x2).to(dev))
            fresh.append((kind, x2, xd, yd))
        imgs = fresh
        _ = torch.ones(1000, device=dev).sum()  # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        check_batch(O, imgs)


@pytest.mark.parametrize("persistent", [1, 0])
def test_batch_both_launch_forms(O, dev, sa, persistent):
    """The persistent form (workgroups drawing queue items from per-XCD ticket
    counters) and the one-workgroup-per-item form give the oracle's results,
    on a batch with many more items than resident workgroups, run twice."""
    say(f"[case] batch launch form persistent={persistent}")
    L = sa.lib()
    before = L.sat_batch_persistent_get()
    assert L.sat_batch_persistent_set(persistent) == 0
    try:
        imgs = make_batch(O, dev, 96, 64, 1500, 23)
        plan = sa.BatchPlan([(xd, yd) for _, _, xd, yd in imgs])
        for rep in range(2):
            for _, _, _, yd in imgs:
                yd.zero_()
            plan.run()
            torch.cuda.synchronize()
            check_batch(O, imgs)
        plan.check_faults()
    finally:
        L.sat_batch_persistent_set(before)


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_device_epoch_wrap(O, dev, sa, persistent):
    """The device epoch range [2^29, 2^30 - 1] wraps. Carry words of a first
    run (epoch 2^29) are kept with their tags but garbage payloads; a run at
    epoch 2^30 - 1 must ignore them (another epoch), and the run after it,
    which wraps back to 2^29, must clear them before its tiles can read them.
    Both runs give the oracle's results and no look-back gives up."""
    say(f"[case] batch device epoch wrap persistent={persistent}")
    L = sa.lib()
    before = L.sat_batch_persistent_get()
    assert L.sat_batch_persistent_set(persistent) == 0
    try:
        base, emax = 1 << 29, (1 << 30) - 1
        imgs = make_batch(O, dev, 24, 1, 600, 31)
        plan = sa.BatchPlan([(xd, yd) for _, _, xd, yd in imgs])
        plan.run()
        torch.cuda.synchronize()
        check_batch(O, imgs)
        ws = plan.workspace
        assert ws.numel() % 8 == 0
        assert int(ws.view(torch.int32)[0]) == base  # header epoch after the first device-epoch run

        # The carry words lie between the 256-byte header and the item table,
        # which starts with image 0's input and output pointers. Only words
        # there are rewritten: the queue after the table holds words (empty
        # entries, all ones) whose upper half reads like an epoch tag.
        x0, y0 = imgs[0][2].data_ptr(), imgs[0][3].data_ptr()

        def plant(from_epoch, header):
            # every carry word tagged `from_epoch` -> tag of epoch 2^29 (status kept), garbage payload
            w = ws.cpu().numpy().view(np.uint64).copy()
            at = np.nonzero((w[:-1] == np.uint64(x0)) & (w[1:] == np.uint64(y0)))[0]
            assert len(at) == 1 and at[0] > 32
            idx = np.arange(w.size)
            tag = w >> np.uint64(32)
            st = tag & np.uint64(3)
            hit = ((idx >= 32) & (idx < int(at[0])) & ((tag >> np.uint64(2)) == np.uint64(from_epoch))
                   & ((st == np.uint64(1)) | (st == np.uint64(2))))
            assert int(hit.sum()) > 100
            planted = (((tag & np.uint64(3)) | np.uint64(base << 2)) << np.uint64(32)) | np.uint64(0x0BADF00D)
            w = np.where(hit, planted, w)
            ws.copy_(torch.from_numpy(w.view(np.uint8)).to(dev))
            ws.view(torch.int32)[0] = header
            torch.cuda.synchronize()

        plant(base, emax - 1)  # next run: epoch 2^30 - 1, no clear
        for _, _, _, yd in imgs:
            yd.zero_()
        plan.run()
        torch.cuda.synchronize()
        check_batch(O, imgs)
        assert int(ws.view(torch.int32)[0]) == emax
        plant(emax, emax)  # next run wraps to 2^29: must clear the planted words first
        for _, _, _, yd in imgs:
            yd.zero_()
        plan.run()
        torch.cuda.synchronize()
        check_batch(O, imgs)
        assert int(ws.view(torch.int32)[0]) == base
        plan.check_faults()
    finally:
        L.sat_batch_persistent_set(before)


def test_batch_many_tiny_images(O, dev, sa):
    """3000 images of 1 to 40 pixels a side (edge tiles only, many more
    images than workgroups per XCD), mixed u8 -> u32 / f32, one launch."""
    say("[case] batch of 3000 tiny images")
    imgs = make_batch(O, dev, 3000, 1, 40, 41)
    plan = sa.BatchPlan([(xd, yd) for _, _, xd, yd in imgs])
    plan.run()
    torch.cuda.synchronize()
    check_batch(O, imgs)
    plan.check_faults()


def _spread_batch(O, dev):
    """Three images that each hold more than an eighth of the batch's bytes
    (laid over all XCDs) beside 24 small ones (dealt whole), mixed u8 -> u32 /
    f32."""
    imgs = []
    for j, (h, w, kind) in enumerate(((1500, 1300, "f32"), (1100, 1700, "f32"), (2000, 1900, "u8"))):
        x = O.gen(kind, (h, w), 7000 + j)
        yd = torch.empty((h, w), dtype=torch.uint32 if kind == "u8" else torch.float32, device=dev)
        imgs.append((kind, x, torch.from_numpy(x).to(dev), yd))
    imgs += make_batch(O, dev, 24, 1, 260, 71)
    return imgs


@pytest.mark.parametrize("persistent", [0, 1])
def test_batch_large_images_over_all_xcds(O, dev, sa, persistent):
    """Images larger than one XCD's share of the bytes take consecutive queue
    positions over all XCDs (their look-backs read tiles of other XCDs); the
    persistent setting falls back to one workgroup per item for such a plan.
    Every image against the oracle, twice, and a lone image and two images
    (each above an XCD's share)."""
    say(f"[case] batch with images spread over all XCDs, persistent={persistent}")
    L = sa.lib()
    before = L.sat_batch_persistent_get()
    assert L.sat_batch_persistent_set(persistent) == 0
    try:
        imgs = _spread_batch(O, dev)
        plan = sa.BatchPlan([(xd, yd) for _, _, xd, yd in imgs])
        for rep in range(2):
            for _, _, _, yd in imgs:
                yd.zero_()
            plan.run()
            torch.cuda.synchronize()
            check_batch(O, imgs)
        plan.check_faults()
        for sub in (imgs[:1], imgs[1:3]):
            p = sa.BatchPlan([(xd, yd) for _, _, xd, yd in sub])
            for _, _, _, yd in sub:
                yd.zero_()
            p.run()
            torch.cuda.synchronize()
            check_batch(O, sub)
            p.check_faults()
    finally:
        L.sat_batch_persistent_set(before)


def test_batch_equal_images_one_more_than_xcds(O, dev, sa):
    """Nine images of one size: dealt whole, one XCD would hold two; the
    largest are laid over all XCDs until the rest deal evenly (mixed spread
    and dealt images in one queue). Every image against the oracle."""
    say("[case] batch of 9 equal images")
    imgs = []
    for j in range(9):
        kind = ("u8", "f32")[j % 2]
        x = O.gen(kind, (300, 530), 7100 + j)
        yd = torch.empty((300, 530), dtype=torch.uint32 if kind == "u8" else torch.float32, device=dev)
        imgs.append((kind, x, torch.from_numpy(x).to(dev), yd))
    plan = sa.BatchPlan([(xd, yd) for _, _, xd, yd in imgs])
    plan.run()
    torch.cuda.synchronize()
    check_batch(O, imgs)
    plan.check_faults()


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_batch_random_compositions(O, dev, sa, seed):
    """Seeded batches of k large images (sides 500..1400) and n small ones
    (sides 1..220), k in {0, 1, 2, 3, 9}, n in {0, 5, 30}, mixed u8 -> u32 /
    f32, each launch form: whatever mix of images laid over all XCDs and
    images dealt whole the planner picks, every image against the oracle."""
    rng = np.random.default_rng(9000 + seed)
    L = sa.lib()
    before = L.sat_batch_persistent_get()
    try:
        for case in range(4):
            k = int(rng.choice([0, 1, 2, 3, 9]))
            n = int(rng.choice([0, 5, 30])) if k else int(rng.choice([5, 30]))
            persistent = int(rng.integers(0, 2))
            say(f"[case] seed {seed} case {case}: {k} large + {n} small images, persistent={persistent}")
            assert L.sat_batch_persistent_set(persistent) == 0
            imgs = []
            for j in range(k):
                h, w = (int(v) for v in rng.integers(500, 1401, 2))
                kind = ("u8", "f32")[int(rng.integers(0, 2))]
                x = O.gen(kind, (h, w), 90000 + 1000 * seed + 100 * case + j)
                yd = torch.empty((h, w), dtype=torch.uint32 if kind == "u8" else torch.float32, device=dev)
                imgs.append((kind, x, torch.from_numpy(x).to(dev), yd))
            if n:
                imgs += make_batch(O, dev, n, 1, 220, 500 + 10 * seed + case)
            order = rng.permutation(len(imgs))
            imgs = [imgs[i] for i in order]
            plan = sa.BatchPlan([(xd, yd) for _, _, xd, yd in imgs])
            plan.run()
            torch.cuda.synchronize()
            check_batch(O, imgs)
            plan.check_faults()
    finally:
        L.sat_batch_persistent_set(before)

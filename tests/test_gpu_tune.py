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


CFGS = [0, 1, 2, 3, 4]
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


@pytest.mark.parametrize("cfg", CFGS)
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
    assert len(doc["entries"]) == 4 and all(0 <= e["cfg"] < 5 for e in doc["entries"])
    assert {e["in_bytes"] for e in doc["entries"]} == {1, 4}
    assert doc["batch"][0]["cfg"] in CFGS
    p = tmp_path / "tuned.json"
    import json

    p.write_text(json.dumps(doc))
    assert tune.apply_table(p) == 5

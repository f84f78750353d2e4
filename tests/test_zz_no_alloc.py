"""GPU check of sat_integral_noalloc, the one-pass SAT whose carry workspace
is the library's own device array: called through the C ABI, it must give
what sat_integral (the same one-pass algorithm with a library-managed
workspace) gives, on consecutive calls of decreasing size that reuse the
array, and refuse shapes whose workspace exceeds the array."""
from __future__ import annotations

import sys

import numpy as np
import pytest
import torch

from helpers import assert_sat_close

pytestmark = pytest.mark.gpu

U8, U32, F32 = 0, 4, 6


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def L():
    import sat_amd

    return sat_amd.lib()


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _noalloc(L, x, out_dtype, tin, tout):
    h, w = x.shape
    y = torch.empty((h, w), dtype=out_dtype, device=x.device)
    st = L.sat_integral_noalloc(0, tin, tout, x.data_ptr(), x.stride(0), y.data_ptr(), y.stride(0), w, h, _stream())
    assert st == 0, st
    return y


def _managed(L, x, out_dtype, tin, tout):
    h, w = x.shape
    y = torch.empty((h, w), dtype=out_dtype, device=x.device)
    st = L.sat_integral(3, 0, tin, tout, x.data_ptr(), x.stride(0), y.data_ptr(), y.stride(0), w, h, None, 0,
                        _stream())
    assert st == 0, st
    return y


def test_noalloc_matches_sat_integral_on_consecutive_calls(O, dev, L):
    import sat_amd

    # first call: u8 -> u32, bit-exact against sat_integral and the oracle
    say("[case] noalloc u8->u32 3000x2500, then f32 700x900, then u8->u32 37x5")
    x = O.gen("u8", (3000, 2500), 20261201)
    xd = torch.from_numpy(x).to(dev)
    a = _noalloc(L, xd, torch.uint32, U8, U32)
    b = _managed(L, xd, torch.uint32, U8, U32)
    torch.cuda.synchronize()
    assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    assert np.array_equal(a.view(torch.int32).cpu().numpy().view(np.uint32), O.serial(x, np.uint32))
    # second call, smaller: f32 within the FP64 bound, like sat_integral's
    xf = O.gen("f32", (700, 900), 20261202)
    xfd = torch.from_numpy(xf).to(dev)
    af = _noalloc(L, xfd, torch.float32, F32, F32)
    bf = _managed(L, xfd, torch.float32, F32, F32)
    torch.cuda.synchronize()
    assert_sat_close(O, xf, af.cpu().numpy(), "f32")
    assert_sat_close(O, xf, bf.cpu().numpy(), "f32")
    # third call, smaller still: the array keeps the earlier calls' words
    xs = O.gen("u8", (37, 5), 20261203)
    xsd = torch.from_numpy(xs).to(dev)
    s = _noalloc(L, xsd, torch.uint32, U8, U32)
    torch.cuda.synchronize()
    assert np.array_equal(s.view(torch.int32).cpu().numpy().view(np.uint32), O.serial(xs, np.uint32))
    sat_amd.check_faults()


def test_noalloc_refuses_oversized_and_bad_arguments(dev, L):
    say("[case] noalloc refuses a workspace beyond its array")
    cap = L.sat_integral_noalloc_max_bytes()
    assert cap == 32 << 20
    x = torch.zeros((4, 4), dtype=torch.float32, device=dev)
    y = torch.zeros((4, 4), dtype=torch.float32, device=dev)
    # 20000 x 20000 f32 needs more carry words than the array holds: refused
    # before anything is launched (the small buffers are never touched)
    assert L.sat_integral_workspace_bytes(3, F32, F32, 20000, 20000) > cap
    assert L.sat_integral_noalloc(0, F32, F32, x.data_ptr(), 20000, y.data_ptr(), 20000, 20000, 20000, _stream()) == 4
    assert L.sat_integral_noalloc(0, 2, U32, x.data_ptr(), 4, y.data_ptr(), 4, 4, 4, _stream()) == 3  # u16 pair
    assert L.sat_integral_noalloc(0, F32, F32, None, 4, y.data_ptr(), 4, 4, 4, _stream()) == 1
    assert L.sat_integral_noalloc(0, F32, F32, x.data_ptr(), 2, y.data_ptr(), 4, 4, 4, _stream()) == 2  # pitch < width

"""CPU checks of how look-back give-ups reach callers: the Python wrappers
read the library's host-visible fault flag and raise LookbackFault, and the
pass wrappers reject an `out` tensor the kernels would write out of bounds.
No GPU is used: the library is replaced by a stand-in object where a device
call would be made."""
from __future__ import annotations

import contextlib

import pytest
import torch


class FakeLib:
    def __init__(self, flag):
        self.flag = flag
        self.calls = []

    def sat_device_fault(self, clear):
        self.calls.append(clear)
        v = self.flag
        if clear:
            self.flag = 0
        return v


@pytest.fixture
def ops(monkeypatch):
    from sat_amd import torch_ops

    monkeypatch.setattr(torch.cuda, "device", lambda d: contextlib.nullcontext())
    monkeypatch.setattr(torch.cuda, "synchronize", lambda d=None: None)
    return torch_ops


def test_guard_raises_once_and_clears(ops, monkeypatch):
    fake = FakeLib(1)
    monkeypatch.setattr(ops, "lib", lambda: fake)
    t = torch.zeros(1)
    with pytest.raises(ops.LookbackFault, match="look-back gave up"):
        ops._guard(t)
    assert fake.calls == [1] and fake.flag == 0
    ops._guard(t)  # cleared: the next call goes through


def test_guard_silent_without_fault(ops, monkeypatch):
    fake = FakeLib(0)
    monkeypatch.setattr(ops, "lib", lambda: fake)
    ops._guard(torch.zeros(1))
    assert fake.calls == [1]


def test_check_faults_raises(ops, monkeypatch):
    fake = FakeLib(1)
    monkeypatch.setattr(ops, "lib", lambda: fake)
    with pytest.raises(ops.LookbackFault):
        ops.check_faults("cuda:0")
    ops.check_faults("cuda:0")


def test_guard_reports_library_errors(ops, monkeypatch):
    fake = FakeLib(-5)  # -SAT_ERR_HIP
    monkeypatch.setattr(ops, "lib", lambda: fake)
    with pytest.raises(ops.SatError, match="HIP runtime error"):
        ops._guard(torch.zeros(1))


def test_lookback_fault_is_a_sat_error(ops):
    import sat_amd

    assert issubclass(ops.LookbackFault, sat_amd.SatError)
    assert sat_amd.STATUS[sat_amd.SAT_ERR_LOOKBACK].startswith("a look-back gave up")


@pytest.mark.parametrize("bad", [
    dict(dtype=torch.float64),                   # 8-byte elements into a 4-byte pass output
    dict(shape=(5, 8)),                          # not transposed
    dict(shape=(4, 9)),                          # too small
    dict(strided=True),                          # column stride 2
    dict(device="meta"),                         # another device
])
def test_pass_out_rejects_bad_outputs(ops, bad):
    x = torch.zeros((5, 8), dtype=torch.float32)
    shape = bad.get("shape", (8, 5))
    if bad.get("strided"):
        out = torch.zeros((8, 10), dtype=torch.float32)[:, ::2]
    else:
        out = torch.zeros(shape, dtype=bad.get("dtype", torch.float32), device=bad.get("device", "cpu"))
    with pytest.raises(ops.SatError):
        ops._pass_out(x, out, torch.float32, (8, 5))


def test_pass_out_accepts_pitched_rows(ops):
    x = torch.zeros((5, 8), dtype=torch.float32)
    out = torch.zeros((8, 7), dtype=torch.float32)[:, :5]  # row pitch 7 >= 5
    assert ops._pass_out(x, out, torch.float32, (8, 5)) is out


def test_device_fault_abi_without_device():
    import sat_amd

    # no HIP device here: the call fails cleanly instead of touching memory
    assert sat_amd.lib().sat_device_fault(0) <= 0

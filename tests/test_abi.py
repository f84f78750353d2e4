"""CPU checks of the drop-in boundary: the C ABI library loads, exports every
symbol include/sat/sat.h declares, validates arguments before touching a
device, and the C++ header compiles. No compute call is made here."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "sat" / "sat.h"


def declared_symbols() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sat_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def L():
    import sat_amd

    return sat_amd.lib()


def test_exports_every_declared_symbol(L):
    syms = declared_symbols()
    assert len(syms) >= 20
    nm = subprocess.run(["nm", "-D", "--defined-only", str(ROOT / "sat_amd" / "libsat_amd.so")],
                        capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\s[TW]\s+(sat_\w+)$", nm, flags=re.M))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing
    import sat_amd

    assert set(sat_amd.SIGNATURES) == set(syms)


def test_enums_match_reference(L):
    import sat_amd

    # DataType (integralImage.h:22-32), SCAN (integralImage.h:34-37)
    assert (sat_amd.U8, sat_amd.I8, sat_amd.U32, sat_amd.I32, sat_amd.F32, sat_amd.F64) == (0, 1, 4, 5, 6, 7)
    assert (sat_amd.KOGGE_STONE, sat_amd.LF) == (0, 1)
    assert L.sat_abi_version() == 2
    assert [L.sat_dtype_size(d) for d in range(8)] == [1, 1, 2, 2, 4, 4, 4, 8]
    assert L.sat_status_string(4) == b"workspace too small"


def test_supported_pairs(L):
    ok = {(0, 4), (0, 6), (0, 7), (1, 5), (1, 6), (1, 7), (5, 5), (5, 7), (6, 6), (6, 7), (7, 7), (4, 4)}
    for a in range(8):
        for b in range(8):
            assert L.sat_dtype_pair_supported(a, b) == ((a, b) in ok), (a, b)


def test_validation_without_device(L):
    p = C.c_void_p(0x1000)
    # unsupported pair, bad scan, bad algorithm, null pointer, pitch < width
    assert L.sat_integral(3, 0, 2, 4, p, 64, p, 64, 64, 64, None, 0, None) == 3
    assert L.sat_integral(3, 7, 6, 6, p, 64, p, 64, 64, 64, None, 0, None) == 2
    assert L.sat_integral(9, 0, 6, 6, p, 64, p, 64, 64, 64, None, 0, None) == 2
    assert L.sat_integral(3, 0, 6, 6, None, 64, p, 64, 64, 64, None, 0, None) == 1
    assert L.sat_integral(3, 0, 6, 6, p, 32, p, 64, 64, 64, None, 0, None) == 2
    # empty images are a no-op
    assert L.sat_integral(3, 0, 6, 6, None, 0, None, 0, 0, 5, None, 0, None) == 0
    assert L.sat_block_scan(0, 0, 4, p, p, 64, 64, 64, 32, None) == 2  # heightStride < height
    assert L.sat_seriel_scan(6, 4, p, p, 64, 64, 64, 64, None) == 3
    assert L.sat_scan_x(0, 6, 6, None, p, 64, 64, 64, None) == 1
    assert L.sat_scan_y(0, 6, 5, p, p, 64, 64, 64, None) == 3
    assert L.sat_integral_reverse(0, 0, p, 64, p, 64, 64, 64, None, 0, None) == 3
    assert L.sat_box_sums(5, p, 64, 0, 64, 64, 1, p, 10, p, None) == 3
    assert L.sat_box_sums(6, None, 64, 0, 64, 64, 1, p, 10, p, None) == 1
    assert L.sat_box_grad(6, p, 10, p, p, 32, 0, 64, 64, 1, None, 0, None) == 2
    assert L.sat_batch_run(None, None) == 1
    assert L.sat_batch_run(p, None) == 2  # never prepared
    # planes with overlapping plane strides
    assert L.sat_integral_planes(3, 0, 6, 6, p, 64, 10, p, 64, 64 * 64, 64, 64, 2, None, 0, None) == 2
    # streaming copy: empty is a no-op; null, misaligned and overlapping buffers are refused
    assert L.sat_stream_copy(None, None, 0, None) == 0
    assert L.sat_stream_copy(None, p, 16, None) == 1
    assert L.sat_stream_copy(C.c_void_p(0x1008), C.c_void_p(0x9000), 16, None) == 2
    assert L.sat_stream_copy(p, C.c_void_p(0x1010), 64, None) == 2


def test_batch_validation(L):
    import sat_amd

    arr = (sat_amd.sat_image * 2)()
    for i in range(2):
        arr[i].in_, arr[i].out = 0x1000, 0x2000
        arr[i].in_pitch = arr[i].out_pitch = 100
        arr[i].width, arr[i].height = 100, 50
        arr[i].tin, arr[i].tout = 0, 4
    assert L.sat_batch_workspace_bytes(C.cast(arr, C.c_void_p), 2) > 0
    arr[1].tin, arr[1].tout = 6, 7  # 8-byte output mixed with 4-byte output
    assert L.sat_batch_workspace_bytes(C.cast(arr, C.c_void_p), 2) == 0
    arr[1].tin, arr[1].tout = 6, 6
    arr[1].in_pitch = 10
    assert L.sat_batch_workspace_bytes(C.cast(arr, C.c_void_p), 2) == 0


def test_workspace_sizes(L):
    # one-pass f32: the largest over the tile configurations is the 64x256 tile:
    # 256 column words + the corner word padded to a 64-byte line (8 words)
    # + 64 row words each
    tiles = (8192 // 64) * (8192 // 256)
    assert L.sat_integral_workspace_bytes(3, 6, 6, 8192, 8192) == 256 + tiles * (264 + 64) * 8
    # f64: 32x256 tiles, two words per carry
    tiles = (4096 // 32) * (4096 // 256)
    assert L.sat_integral_workspace_bytes(3, 7, 7, 4096, 4096) == 256 + tiles * (264 + 32) * 16
    # one tile column of 64-row tiles (configurations 2 and 3, one column per
    # vertical lane) adds the second vertical carry level: one word per column
    # for every run of 16 tile rows
    tiles = 65536 // 64
    assert L.sat_integral_workspace_bytes(3, 6, 6, 256, 65536) == 256 + (tiles * (264 + 64) + (tiles // 16) * 256) * 8
    # two tile columns: one level
    tiles = (65536 // 64) * 2
    assert L.sat_integral_workspace_bytes(3, 6, 6, 512, 65536) == 256 + tiles * (264 + 64) * 8
    # two-pass algorithms add a temporary of the output type
    assert L.sat_integral_workspace_bytes(2, 6, 6, 1000, 1000) >= 1024 * 1000 * 4
    assert L.sat_integral_workspace_bytes(0, 0, 4, 1000, 1000) >= 1024 * 1000 * 4
    assert L.sat_integral_workspace_bytes(3, 2, 4, 100, 100) == 0


def test_cpp_header_compiles(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('''#include "sat/sat.hpp"
int main() {
  const unsigned char* a = nullptr; unsigned* b = nullptr; float* f = nullptr;
  (void)BlockScan::blockScan<unsigned char, unsigned, 1, 32, 8>(a, b, 0, 0, 0, 0);
  (void)SerialScan::serielScan<unsigned char, unsigned, 32, 8, 1024>(a, b, 0, 0, 0, 0);
  (void)IncreamentScan::ScanX<unsigned char, unsigned, 0, 32, 32, 32>(a, b, 0, 0, 0);
  (void)IncrementScan::ScanY<float, float, 1, 32, 32>(f, f, 0, 0, 0);
  (void)sat::integral<unsigned char, unsigned>(a, 0, b, 0, 0, 0);
  return 0;
}
''')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-I", str(ROOT / "include"), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_product_does_not_use_the_oracle():
    for p in (ROOT / "sat_amd").rglob("*"):
        if p.suffix in (".py", ".hip", ".hpp", ".cpp", ".h"):
            t = p.read_text()
            assert "oracle" not in t.lower().replace("oracle-free", ""), p

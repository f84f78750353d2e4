"""TEST INFRASTRUCTURE ONLY: Python access to the CPU oracle.

Loads ``oracle/_build/libsat_oracle.so`` (built from ``oracle/sat_oracle.c``
by ``oracle/Makefile``), the C restatement of the reference's CPU integral
routines (pengdada/SAT ``integralImage/integralImage.h:74-160,280-288``), and
the splitmix64 input generators shared by tests and bench. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it, and
only as the checker; the ``sat_amd`` product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libsat_oracle.so"

_NAME = {
    np.dtype(np.uint8): "u8",
    np.dtype(np.int8): "i8",
    np.dtype(np.uint32): "u32",
    np.dtype(np.int32): "i32",
    np.dtype(np.float32): "f32",
    np.dtype(np.float64): "f64",
}
PAIRS = [
    ("u8", "u32"), ("u8", "f32"), ("u8", "f64"),
    ("i8", "i32"), ("i8", "f32"), ("i8", "f64"),
    ("i32", "i32"), ("i32", "f64"),
    ("f32", "f32"), ("f32", "f64"), ("f64", "f64"),
    ("u32", "u32"),
]
NP = {"u8": np.uint8, "i8": np.int8, "u32": np.uint32, "i32": np.int32, "f32": np.float32, "f64": np.float64}

_lib = None


def build() -> Path:
    """Compile the oracle library (gcc; a second or two)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _lib.sm64_at.restype = C.c_uint64
        _lib.sm64_at.argtypes = [C.c_uint64, C.c_uint64]
        _lib.oracle_threads.restype = C.c_int
    return _lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _name(dt) -> str:
    return _NAME[np.dtype(dt)]


def _pair_fn(prefix: str, tin, tout):
    fn = getattr(lib(), f"{prefix}_{_name(tin)}_{_name(tout)}")
    return fn


def _check2d(x: np.ndarray):
    if x.ndim != 2:
        raise ValueError("2-D array expected")
    return np.ascontiguousarray(x)


def serial(x: np.ndarray, tout) -> np.ndarray:
    """IntegralImageSerial (integralImage.h:105-124)."""
    x = _check2d(x)
    h, w = x.shape
    out = np.empty((h, w), dtype=tout)
    _pair_fn("oracle_serial", x.dtype, tout)(_p(x), _p(out), C.c_int64(w), C.c_int64(h))
    return out


def parallel(x: np.ndarray, tout) -> np.ndarray:
    """IntegralImageParallel (integralImage.h:129-146), OpenMP rows then columns."""
    x = _check2d(x)
    h, w = x.shape
    out = np.empty((h, w), dtype=tout)
    _pair_fn("oracle_parallel", x.dtype, tout)(_p(x), _p(out), C.c_int64(w), C.c_int64(h))
    return out


def allrow(x: np.ndarray, tout) -> np.ndarray:
    """IntegralImageAllRow (integralImage.h:148-160): row prefix sums only."""
    x = _check2d(x)
    h, w = x.shape
    out = np.empty((h, w), dtype=tout)
    _pair_fn("oracle_allrow", x.dtype, tout)(_p(x), _p(out), C.c_int64(w), C.c_int64(h))
    return out


def colscan(x: np.ndarray, tout) -> np.ndarray:
    """IntegralStride (integralImage.h:90-100) down every column."""
    x = _check2d(x)
    h, w = x.shape
    out = np.empty((h, w), dtype=tout)
    _pair_fn("oracle_colscan", x.dtype, tout)(_p(x), _p(out), C.c_int64(w), C.c_int64(h))
    return out


def transpose(x: np.ndarray) -> np.ndarray:
    """Transpose (integralImage.h:280-288)."""
    x = _check2d(x)
    h, w = x.shape
    out = np.empty((w, h), dtype=x.dtype)
    getattr(lib(), f"oracle_transpose_{_name(x.dtype)}")(_p(x), C.c_int64(w), C.c_int64(h), _p(out))
    return out


def reverse(x: np.ndarray) -> np.ndarray:
    """Suffix SAT: out[y][x] = sum over k >= y, l >= x (u32, f32, f64)."""
    x = _check2d(x)
    h, w = x.shape
    out = np.empty_like(x)
    getattr(lib(), f"oracle_reverse_{_name(x.dtype)}")(_p(x), _p(out), C.c_int64(w), C.c_int64(h))
    return out


def box_sums_f64(sat: np.ndarray, boxes: np.ndarray) -> np.ndarray:
    """Four-corner readout from f64 SAT planes (P, H, W); boxes (P, nbox, 4)."""
    sat = np.ascontiguousarray(sat, dtype=np.float64)
    if sat.ndim == 2:
        sat = sat[None]
    p, h, w = sat.shape
    boxes = np.ascontiguousarray(boxes, dtype=np.int32).reshape(p, -1, 4)
    nbox = boxes.shape[1]
    out = np.empty((p, nbox), dtype=np.float64)
    lib().oracle_box_sums_f64(_p(sat), C.c_int64(w), C.c_int64(h), C.c_int64(p), _p(boxes), C.c_int64(nbox), _p(out))
    return out


# ---------------------------------------------------------------------------
# splitmix64 counter-based generators (identical bytes for oracle and GPU runs)
# ---------------------------------------------------------------------------
def gen(kind: str, shape, seed: int) -> np.ndarray:
    """kind: 'u8' U[0,255], 'i8' U[-128,127], 'u32_20' U[0,2^20), 'u32' full,
    'i32_16' U[-2^15, 2^15), 'f32' U[0,1), 'f32n' / 'f64n' normal(0,1),
    'f64' U[0,1) as f64."""
    n = int(np.prod(shape))
    L = lib()
    s = C.c_uint64(seed & 0xFFFFFFFFFFFFFFFF)
    if kind == "u8":
        out = np.empty(n, np.uint8)
        L.sm64_u8(s, C.c_int64(n), _p(out))
    elif kind == "i8":
        out = np.empty(n, np.int8)
        L.sm64_i8(s, C.c_int64(n), _p(out))
    elif kind.startswith("u32"):
        bits = int(kind.split("_")[1]) if "_" in kind else 32
        out = np.empty(n, np.uint32)
        L.sm64_u32_bits(s, C.c_int64(n), C.c_int(bits), _p(out))
    elif kind == "i32_16":
        u = np.empty(n, np.uint32)
        L.sm64_u32_bits(s, C.c_int64(n), C.c_int(16), _p(u))
        out = (u.astype(np.int64) - 32768).astype(np.int32)
    elif kind == "f32":
        out = np.empty(n, np.float32)
        L.sm64_f32_unit(s, C.c_int64(n), _p(out))
    elif kind == "f64":
        out = np.empty(n, np.float32)
        L.sm64_f32_unit(s, C.c_int64(n), _p(out))
        out = out.astype(np.float64)
    elif kind == "f64n":
        out = np.empty(n, np.float64)
        L.sm64_f64_normal(s, C.c_int64(n), _p(out))
    elif kind == "f32n":
        out = np.empty(n, np.float32)
        L.sm64_f32_normal(s, C.c_int64(n), _p(out))
    else:
        raise ValueError(kind)
    return out.reshape(shape)


def sm64_at(seed: int, i: int) -> int:
    return int(lib().sm64_at(C.c_uint64(seed), C.c_uint64(i)))


def threads() -> int:
    return int(lib().oracle_threads())


def set_threads(n: int) -> None:
    lib().oracle_set_threads(C.c_int(n))


def sat_f64(x: np.ndarray) -> np.ndarray:
    """FP64 oracle of any real input: IntegralImageSerial with TDst = double."""
    x = _check2d(x)
    if x.dtype in (np.float64,):
        return serial(x, np.float64)
    if x.dtype == np.float32:
        return serial(x, np.float64)
    if x.dtype == np.uint8:
        return serial(x, np.float64)
    if x.dtype == np.int8:
        return serial(x, np.float64)
    if x.dtype == np.int32:
        return serial(x, np.float64)
    return serial(x.astype(np.float64), np.float64)

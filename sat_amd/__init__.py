"""sat_amd: summed-area tables on AMD Instinct MI355X (gfx950).

Python face of the C ABI in ``include/sat/sat.h`` (``libsat_amd.so``, built
from ``sat_amd/csrc`` by ``sat_amd/build.py``). PyTorch supplies device memory
and the current HIP stream; every computation runs in the library's HIP
kernels. If the library is missing, importing the ops raises: there is no
host or PyTorch fallback.

Reference names (pengdada/SAT) and what mirrors them here:

* ``BlockScan::blockScan``   (integralImage/ScanRow-BRLT.cu:24)  -> :func:`block_scan`
* ``SerielScan::serielScan`` (integralImage/BRLT-Scan.cu:15)     -> :func:`seriel_scan`
* ``IncrementScan::ScanX``   (integralImage/ScanRowColumn.cu:145) -> :func:`scan_x`
* ``IncrementScan::ScanY``   (integralImage/ScanRowColumn.cu:16)  -> :func:`scan_y`
* ``<NS>::Test`` two-pass SAT compositions                        -> :func:`integral` with ``alg`` 0..2
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("SAT_AMD_LIB", PKG / "libsat_amd.so"))

# DataType (integralImage.h:22-32), SCAN (integralImage.h:34-37), algorithm ids (main.cpp:20-34)
U8, I8, U16, I16, U32, I32, F32, F64 = range(8)
KOGGE_STONE, LF = 0, 1
SCANROW_BRLT, BRLT_SCANROW, SCANROW_COLUMN, SINGLE_PASS = 0, 1, 2, 3
ALG_NAMES = {SCANROW_BRLT: "ScanRow-BRLT", BRLT_SCANROW: "BRLT-ScanRow", SCANROW_COLUMN: "ScanRowColumn",
             SINGLE_PASS: "SinglePass"}

STATUS = {0: "ok", 1: "null pointer", 2: "invalid argument", 3: "unsupported dtype pair",
          4: "workspace too small", 5: "HIP runtime error", 6: "not supported by this build",
          7: "a look-back gave up waiting for a carry: results invalid"}
SAT_ERR_LOOKBACK = 7

# every function the C ABI exports: (restype, argtypes)
_i, _u32, _i64, _vp, _sz = C.c_int, C.c_uint32, C.c_int64, C.c_void_p, C.c_size_t
SIGNATURES = {
    "sat_abi_version": (_i, []),
    "sat_dtype_size": (_sz, [_i]),
    "sat_dtype_pair_supported": (_i, [_i, _i]),
    "sat_status_string": (C.c_char_p, [_i]),
    "sat_workspace_fault": (_i, [_vp, _vp]),
    "sat_device_fault": (_i, [_i]),
    "sat_release_workspaces": (_i, []),
    "sat_stream_copy": (_i, [_vp, _vp, _sz, _vp]),
    "sat_block_scan": (_i, [_i, _i, _i, _vp, _vp, _u32, _u32, _u32, _u32, _vp]),
    "sat_seriel_scan": (_i, [_i, _i, _vp, _vp, _u32, _u32, _u32, _u32, _vp]),
    "sat_scan_x": (_i, [_i, _i, _i, _vp, _vp, _u32, _u32, _u32, _vp]),
    "sat_scan_y": (_i, [_i, _i, _i, _vp, _vp, _u32, _u32, _u32, _vp]),
    "sat_scan_rows": (_i, [_i, _i, _i, _vp, _i64, _vp, _i64, _u32, _u32, _vp]),
    "sat_scan_cols": (_i, [_i, _i, _i, _vp, _i64, _vp, _i64, _u32, _u32, _vp]),
    "sat_integral_workspace_bytes": (_sz, [_i, _i, _i, _u32, _u32]),
    "sat_integral": (_i, [_i, _i, _i, _i, _vp, _i64, _vp, _i64, _u32, _u32, _vp, _sz, _vp]),
    "sat_integral_planes_workspace_bytes": (_sz, [_i, _i, _i, _u32, _u32, _u32]),
    "sat_integral_planes": (_i, [_i, _i, _i, _i, _vp, _i64, _i64, _vp, _i64, _i64, _u32, _u32, _u32, _vp, _sz, _vp]),
    "sat_integral_reverse_workspace_bytes": (_sz, [_i, _u32, _u32]),
    "sat_integral_reverse": (_i, [_i, _i, _vp, _i64, _vp, _i64, _u32, _u32, _vp, _sz, _vp]),
    "sat_integral_reverse_planes_workspace_bytes": (_sz, [_i, _u32, _u32, _u32]),
    "sat_integral_reverse_planes": (_i, [_i, _i, _vp, _i64, _i64, _vp, _i64, _i64, _u32, _u32, _u32, _vp, _sz, _vp]),
    "sat_batch_workspace_bytes": (_sz, [_vp, _i]),
    "sat_batch_prepare": (_i, [_vp, _i, _vp, _sz, _vp]),
    "sat_batch_run": (_i, [_vp, _vp]),
    "sat_batch_persistent_set": (_i, [_i]),
    "sat_batch_persistent_get": (_i, []),
    "sat_box_sums": (_i, [_i, _vp, _i64, _i64, _u32, _u32, _u32, _vp, _u32, _vp, _vp]),
    "sat_box_scatter": (_i, [_i, _vp, _u32, _vp, _vp, _i64, _i64, _u32, _u32, _u32, _vp]),
    "sat_box_grad_workspace_bytes": (_sz, [_i, _u32, _u32, _u32, _u32]),
    "sat_box_grad": (_i, [_i, _vp, _u32, _vp, _vp, _i64, _i64, _u32, _u32, _u32, _vp, _sz, _vp]),
    "sat_integral_noalloc_max_bytes": (_sz, []),
    "sat_integral_noalloc": (_i, [_i, _i, _i, _vp, _i64, _vp, _i64, _u32, _u32, _vp]),
    "sat_integral_bordered": (_i, [_i, _i, _i, _i, _vp, _i64, _vp, _i64, _u32, _u32, _vp, _sz, _vp]),
    "sat_integral_bordered_planes": (_i, [_i, _i, _i, _i, _vp, _i64, _i64, _vp, _i64, _i64, _u32, _u32, _u32, _vp,
                                          _sz, _vp]),
    "sat_tile_config_count": (_i, []),
    "sat_tune_class": (None, [_u32, _u32, _u32, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "sat_tune_get": (_i, [_i, _i, _i, _i, _i]),
    "sat_tune_set": (_i, [_i, _i, _i, _i, _i, _i]),
    "sat_tune_pick": (_i, [_i, _i, _u32, _u32, _u32]),
    "sat_batch_tune_get": (_i, [_i]),
    "sat_batch_tune_set": (_i, [_i, _i]),
}

# Tile-configuration table written by sat_amd.tune for this build's target
TUNED_TABLE = Path(__file__).resolve().parent / "tuned_gfx950.json"


class SatError(RuntimeError):
    pass


_lib = None


def lib() -> C.CDLL:
    """The loaded C ABI library. Raises when it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise SatError(f"{LIB_PATH} is missing: build it with `python -m sat_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        if TUNED_TABLE.exists():
            from .tune import apply_table

            apply_table(TUNED_TABLE)
    return _lib


def check(status: int, what: str = "sat call") -> None:
    if status != 0:
        raise SatError(f"{what} failed: {STATUS.get(status, status)} ({status})")


class sat_image(C.Structure):
    _fields_ = [("in_", C.c_void_p), ("out", C.c_void_p), ("in_pitch", C.c_int64), ("out_pitch", C.c_int64),
                ("width", C.c_uint32), ("height", C.c_uint32), ("tin", C.c_int32), ("tout", C.c_int32)]


from .torch_ops import *  # noqa: E402,F401,F403

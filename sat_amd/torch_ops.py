"""Tensor-level entry points and autograd Functions over the C ABI.

Tensors must live on a HIP device; pointers and the current stream are handed
to ``libsat_amd.so``. Leading dimensions of (..., H, W) tensors are treated as
independent planes (e.g. the N*C planes of an NCHW batch).
"""
from __future__ import annotations

import ctypes as C
from typing import Sequence

import torch

from . import (BRLT_SCANROW, F32, F64, I8, I32, KOGGE_STONE, SAT_ERR_LOOKBACK, SCANROW_BRLT, SCANROW_COLUMN,
               SINGLE_PASS, STATUS, U8, U32, SatError, check, lib, sat_image)

__all__ = [
    "sat_dtype", "default_out_dtype", "integral", "integral_bordered", "integral_reverse", "block_scan", "seriel_scan", "scan_x", "scan_y",
    "box_sums", "box_scatter", "box_grad", "BatchPlan", "sat", "box_sum", "box_filter", "SATFunction",
    "BoxSumFunction", "BoxFilterFunction", "SCANROW_BRLT", "BRLT_SCANROW", "SCANROW_COLUMN", "SINGLE_PASS",
    "check_faults", "LookbackFault",
]


class LookbackFault(SatError):
    """A look-back of an earlier launch on this device gave up waiting for a
    carry, so that launch's results are undefined (status SAT_ERR_LOOKBACK)."""


def _raise_fault(where: str) -> None:
    raise LookbackFault(f"{where}: {STATUS[SAT_ERR_LOOKBACK]} (status {SAT_ERR_LOOKBACK}); the results of the "
                        "faulting launch and of every call that read them are undefined")


def _guard(t: torch.Tensor) -> None:
    """Raises if a finished launch of the library on t's device gave up in a
    look-back since the last check: reads the library's sticky flag in pinned
    host memory, without synchronising, and clears it."""
    with torch.cuda.device(t.device):
        f = lib().sat_device_fault(1)
    if f < 0:
        check(-f, "sat_device_fault")
    if f:
        _raise_fault("an earlier sat_amd call")


def check_faults(device=None) -> None:
    """Synchronises `device` (default: the current one) and raises
    :class:`LookbackFault` if any look-back of this library on it gave up
    since the last check."""
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    torch.cuda.synchronize(dev)
    with torch.cuda.device(dev):
        f = lib().sat_device_fault(1)
    if f < 0:
        check(-f, "sat_device_fault")
    if f:
        _raise_fault("check_faults")

_TORCH2SAT = {torch.uint8: U8, torch.int8: I8, torch.int32: I32, torch.float32: F32, torch.float64: F64}
if hasattr(torch, "uint32"):
    _TORCH2SAT[torch.uint32] = U32
_DEFAULT_OUT = {torch.uint8: torch.uint32 if hasattr(torch, "uint32") else torch.int32, torch.int8: torch.int32,
                torch.int32: torch.int32, torch.float32: torch.float32, torch.float64: torch.float64}
if hasattr(torch, "uint32"):
    _DEFAULT_OUT[torch.uint32] = torch.uint32


def sat_dtype(dt: torch.dtype) -> int:
    try:
        return _TORCH2SAT[dt]
    except KeyError:
        raise SatError(f"dtype {dt} is not supported") from None


def default_out_dtype(dt: torch.dtype) -> torch.dtype:
    return _DEFAULT_OUT[dt]


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _dev_check(*ts: torch.Tensor) -> None:
    for t in ts:
        if not t.is_cuda:
            raise SatError("sat_amd operates on HIP device tensors only")
    for t in ts[1:]:
        if t.device != ts[0].device:
            raise SatError("all tensors of a call must be on one device")
    _guard(ts[0])


def _planes(x: torch.Tensor):
    """View (..., H, W) as (P, H, W) with unit column stride."""
    if x.dim() < 2:
        raise SatError("expected a tensor of at least 2 dimensions")
    h, w = x.shape[-2], x.shape[-1]
    p = 1
    for d in x.shape[:-2]:
        p *= d
    x3 = x.reshape(p, h, w)  # explicit plane count: also valid for empty images
    if x3.stride(-1) != 1 or (x3.shape[0] > 1 and x3.stride(0) < x3.stride(1) * h):
        x3 = x3.contiguous()
    return x3


def integral(x: torch.Tensor, out_dtype: torch.dtype | None = None, alg: int = SINGLE_PASS, scan: int = KOGGE_STONE,
             out: torch.Tensor | None = None) -> torch.Tensor:
    """Inclusive 2-D prefix sum over the last two dimensions (IntegralImageSerial,
    integralImage.h:105-124). ``alg`` 0..2 run the reference's two-pass
    compositions, 3 (default) the one-pass kernel."""
    _dev_check(x)
    out_dtype = out_dtype or default_out_dtype(x.dtype)
    x3 = _planes(x)
    p, h, w = x3.shape
    if out is None:
        out = torch.empty(x.shape, dtype=out_dtype, device=x.device)
    if (not isinstance(out, torch.Tensor) or not out.is_cuda or out.device != x.device or out.dtype != out_dtype
            or out.shape != x.shape or not out.is_contiguous()):
        raise SatError("`out` must be a contiguous tensor of the input's shape and the output dtype, on the "
                       "input's device")
    o3 = out.view(p, h, w)
    if x3.numel() == 0:
        return out
    st = lib().sat_integral_planes(alg, scan, sat_dtype(x.dtype), sat_dtype(out_dtype), x3.data_ptr(), x3.stride(1),
                                   x3.stride(0), o3.data_ptr(), o3.stride(1), o3.stride(0), w, h, p, None, 0,
                                   _stream(x))
    check(st, "sat_integral_planes")
    return out


def integral_bordered(x: torch.Tensor, out_dtype: torch.dtype | None = None, alg: int = SINGLE_PASS,
                      scan: int = KOGGE_STONE) -> torch.Tensor:
    """OpenCV ``cv::integral`` layout: (..., H+1, W+1) with a zero first row and
    column; out[..., y+1, x+1] is the inclusive 2-D prefix sum at (y, x)."""
    _dev_check(x)
    out_dtype = out_dtype or default_out_dtype(x.dtype)
    x3 = _planes(x)
    p, h, w = x3.shape
    out = torch.empty((*x.shape[:-2], h + 1, w + 1), dtype=out_dtype, device=x.device)
    o3 = out.view(p, h + 1, w + 1)
    if p == 0:
        return out
    st = lib().sat_integral_bordered_planes(alg, scan, sat_dtype(x.dtype), sat_dtype(out_dtype), x3.data_ptr(),
                                            x3.stride(1), x3.stride(0), o3.data_ptr(), o3.stride(1), o3.stride(0), w,
                                            h, p, None, 0, _stream(x))
    check(st, "sat_integral_bordered_planes")
    return out


def integral_reverse(x: torch.Tensor, scan: int = KOGGE_STONE) -> torch.Tensor:
    """Suffix sums over the last two dimensions: the adjoint of :func:`integral`."""
    _dev_check(x)
    x3 = _planes(x)
    p, h, w = x3.shape
    out = torch.empty(x.shape, dtype=x.dtype, device=x.device)
    if x3.numel() == 0:
        return out
    o3 = out.view(p, h, w)
    st = lib().sat_integral_reverse_planes(scan, sat_dtype(x.dtype), x3.data_ptr(), x3.stride(1), x3.stride(0),
                                           o3.data_ptr(), o3.stride(1), o3.stride(0), w, h, p, None, 0, _stream(x))
    check(st, "sat_integral_reverse_planes")
    return out


def _rows2d(x: torch.Tensor) -> torch.Tensor:
    _dev_check(x)
    if x.dim() != 2 or x.stride(1) != 1:
        raise SatError("expected a 2-D tensor with unit column stride")
    return x


def _pass_out(x: torch.Tensor, out: torch.Tensor | None, out_dtype: torch.dtype, shape: tuple) -> torch.Tensor:
    """The output of a pass: a new tensor, or `out` once it is checked to be a
    2-D tensor of `shape` and `out_dtype` on x's device with unit column stride
    (the C ABI sees only a pointer and a row pitch, so a wrong `out` would be
    written out of bounds)."""
    if out is None:
        return torch.empty(shape, dtype=out_dtype, device=x.device)
    if not isinstance(out, torch.Tensor) or out.device != x.device:
        raise SatError("`out` must be a tensor on the input's device")
    if out.dtype != out_dtype:
        raise SatError(f"`out` has dtype {out.dtype}, the pass writes {out_dtype}")
    if tuple(out.shape) != tuple(shape):
        raise SatError(f"`out` has shape {tuple(out.shape)}, the pass writes {tuple(shape)}")
    if out.dim() != 2 or (out.numel() > 0 and (out.stride(1) != 1 or out.stride(0) < shape[1])):
        raise SatError("`out` must be a 2-D tensor with unit column stride and rows that do not overlap")
    return out


def block_scan(x: torch.Tensor, out_dtype: torch.dtype | None = None, scan: int = KOGGE_STONE,
               out: torch.Tensor | None = None) -> torch.Tensor:
    """BlockScan::blockScan (ScanRow-BRLT.cu:24-25): out[x][y] = sum_{k<=x} in[y][k]."""
    x = _rows2d(x)
    h, w = x.shape
    out_dtype = out_dtype or default_out_dtype(x.dtype)
    out = _pass_out(x, out, out_dtype, (w, h))
    check(lib().sat_block_scan(scan, sat_dtype(x.dtype), sat_dtype(out_dtype), x.data_ptr(), out.data_ptr(), w,
                               x.stride(0), h, out.stride(0), _stream(x)), "sat_block_scan")
    return out


def seriel_scan(x: torch.Tensor, out_dtype: torch.dtype | None = None, out: torch.Tensor | None = None) -> torch.Tensor:
    """SerielScan::serielScan (BRLT-Scan.cu:14-15): same result as block_scan."""
    x = _rows2d(x)
    h, w = x.shape
    out_dtype = out_dtype or default_out_dtype(x.dtype)
    out = _pass_out(x, out, out_dtype, (w, h))
    check(lib().sat_seriel_scan(sat_dtype(x.dtype), sat_dtype(out_dtype), x.data_ptr(), out.data_ptr(), w,
                                x.stride(0), h, out.stride(0), _stream(x)), "sat_seriel_scan")
    return out


def scan_x(x: torch.Tensor, out_dtype: torch.dtype | None = None, scan: int = KOGGE_STONE,
           out: torch.Tensor | None = None) -> torch.Tensor:
    """IncrementScan::ScanX (ScanRowColumn.cu:144-145): row prefix sums, same layout."""
    x = _rows2d(x)
    h, w = x.shape
    out_dtype = out_dtype or default_out_dtype(x.dtype)
    out = _pass_out(x, out, out_dtype, (h, w))
    check(lib().sat_scan_rows(scan, sat_dtype(x.dtype), sat_dtype(out_dtype), x.data_ptr(), x.stride(0),
                              out.data_ptr(), out.stride(0), w, h, _stream(x)), "sat_scan_rows")
    return out


def scan_y(x: torch.Tensor, out_dtype: torch.dtype | None = None, scan: int = KOGGE_STONE,
           out: torch.Tensor | None = None) -> torch.Tensor:
    """IncrementScan::ScanY (ScanRowColumn.cu:15-16): column prefix sums."""
    x = _rows2d(x)
    h, w = x.shape
    out_dtype = out_dtype or default_out_dtype(x.dtype)
    out = _pass_out(x, out, out_dtype, (h, w))
    check(lib().sat_scan_cols(scan, sat_dtype(x.dtype), sat_dtype(out_dtype), x.data_ptr(), x.stride(0),
                              out.data_ptr(), out.stride(0), w, h, _stream(x)), "sat_scan_cols")
    return out


def _boxes(boxes: torch.Tensor, planes: int) -> tuple[torch.Tensor, int]:
    if boxes.dtype != torch.int32 or boxes.shape[-1] != 4:
        raise SatError("boxes must be int32 (..., 4) records {y0, x0, y1, x1}")
    b = boxes.reshape(planes, -1, 4).contiguous()
    return b, b.shape[1]


def box_sums(S: torch.Tensor, boxes: torch.Tensor) -> torch.Tensor:
    """Box sums from SAT planes S (..., H, W) (f32/f64); boxes (P, nbox, 4) int32,
    inclusive corners. Returns (P, nbox)."""
    _dev_check(S, boxes)
    s3 = _planes(S)
    p, h, w = s3.shape
    b, nbox = _boxes(boxes, p)
    out = torch.empty((p, nbox), dtype=S.dtype, device=S.device)
    check(lib().sat_box_sums(sat_dtype(S.dtype), s3.data_ptr(), s3.stride(1), s3.stride(0), w, h, p, b.data_ptr(),
                             nbox, out.data_ptr(), _stream(S)), "sat_box_sums")
    return out


def box_scatter(boxes: torch.Tensor, g: torch.Tensor, height: int, width: int) -> torch.Tensor:
    """Adjoint of :func:`box_sums` w.r.t. S: the corner impulses of every box."""
    _dev_check(boxes, g)
    p = g.shape[0]
    b, nbox = _boxes(boxes, p)
    g = g.reshape(p, nbox).contiguous()
    out = torch.zeros((p, height, width), dtype=g.dtype, device=g.device)
    check(lib().sat_box_scatter(sat_dtype(g.dtype), b.data_ptr(), nbox, g.data_ptr(), out.data_ptr(), width,
                                width * height, width, height, p, _stream(g)), "sat_box_scatter")
    return out


def box_grad(boxes: torch.Tensor, g: torch.Tensor, height: int, width: int) -> torch.Tensor:
    """grad of sum(g * box_sums(integral(x), boxes)) w.r.t. x, in one fused pass."""
    _dev_check(boxes, g)
    p = g.shape[0]
    b, nbox = _boxes(boxes, p)
    g = g.reshape(p, nbox).contiguous()
    out = torch.empty((p, height, width), dtype=g.dtype, device=g.device)
    check(lib().sat_box_grad(sat_dtype(g.dtype), b.data_ptr(), nbox, g.data_ptr(), out.data_ptr(), width,
                             width * height, width, height, p, None, 0, _stream(g)), "sat_box_grad")
    return out


class BatchPlan:
    """A prepared batch of (input, output) image pairs of uneven sizes, run by
    one persistent launch. Keep the tensors alive while the plan is in use."""

    def __init__(self, pairs: Sequence[tuple[torch.Tensor, torch.Tensor]]):
        if not pairs:
            raise SatError("empty batch")
        self.pairs = list(pairs)
        arr = (sat_image * len(self.pairs))()
        for i, (x, y) in enumerate(self.pairs):
            _dev_check(x, y)
            if x.dim() != 2 or y.shape != x.shape or x.stride(1) != 1 or y.stride(1) != 1:
                raise SatError("batch images must be 2-D, same shape in and out, unit column stride")
            arr[i].in_ = x.data_ptr()
            arr[i].out = y.data_ptr()
            arr[i].in_pitch = x.stride(0)
            arr[i].out_pitch = y.stride(0)
            arr[i].width = x.shape[1]
            arr[i].height = x.shape[0]
            arr[i].tin = sat_dtype(x.dtype)
            arr[i].tout = sat_dtype(y.dtype)
        self._arr = arr
        dev = self.pairs[0][0].device
        nbytes = lib().sat_batch_workspace_bytes(C.cast(arr, C.c_void_p), len(self.pairs))
        if nbytes == 0:
            raise SatError("invalid batch (dtype pairs, sizes or pitches)")
        self.workspace = torch.empty(nbytes, dtype=torch.uint8, device=dev)
        self.device = dev
        check(lib().sat_batch_prepare(C.cast(arr, C.c_void_p), len(self.pairs), self.workspace.data_ptr(), nbytes,
                                      torch.cuda.current_stream(dev).cuda_stream), "sat_batch_prepare")

    def run(self) -> None:
        _guard(self.workspace)
        check(lib().sat_batch_run(self.workspace.data_ptr(), torch.cuda.current_stream(self.device).cuda_stream),
              "sat_batch_run")

    def check_faults(self) -> None:
        """Synchronises the current stream and raises :class:`LookbackFault` if
        a look-back of a run of this plan gave up since the last check."""
        with torch.cuda.device(self.device):
            f = lib().sat_workspace_fault(self.workspace.data_ptr(), torch.cuda.current_stream(self.device).cuda_stream)
            lib().sat_device_fault(1)
        if f < 0:
            check(-f, "sat_workspace_fault")
        if f:
            _raise_fault("BatchPlan")


# ---------------------------------------------------------------------------
# autograd
# ---------------------------------------------------------------------------
class SATFunction(torch.autograd.Function):
    """y = SAT(x); the adjoint of a 2-D inclusive prefix sum is the 2-D suffix
    sum, so grad_x = reverse SAT(grad_y)."""

    @staticmethod
    def forward(ctx, x, scan=KOGGE_STONE):
        ctx.scan = scan
        return integral(x, scan=scan)

    @staticmethod
    def backward(ctx, gy):
        return integral_reverse(gy.contiguous(), scan=ctx.scan), None


class BoxSumFunction(torch.autograd.Function):
    """b = box_sums(S, boxes); grad_S = corner impulses of grad_b."""

    @staticmethod
    def forward(ctx, S, boxes):
        ctx.save_for_backward(boxes)
        ctx.hw = (S.shape[-2], S.shape[-1])
        ctx.sshape = S.shape
        return box_sums(S, boxes)

    @staticmethod
    def backward(ctx, gb):
        (boxes,) = ctx.saved_tensors
        gS = box_scatter(boxes, gb.contiguous(), *ctx.hw)
        return gS.reshape(ctx.sshape), None


class BoxFilterFunction(torch.autograd.Function):
    """b = box_sums(SAT(x), boxes) with a fused backward: grad_x is the reverse
    SAT of the corner impulses, built tile by tile from the box list.

    For f32 x the SAT the box sums read is computed in f64 (f32 in, f64 out)
    and the sums are rounded to f32 once: a box sum is a difference of four
    SAT values far larger than itself, so an f32 SAT would leave errors near
    1e-6 of the sums' norm. The f64 table is internal; the gradient
    accumulates in f64 too and is returned as f32."""

    @staticmethod
    def forward(ctx, x, boxes):
        S = integral(x, out_dtype=torch.float64 if x.dtype == torch.float32 else None)
        ctx.save_for_backward(boxes)
        ctx.xshape = x.shape
        return box_sums(S, boxes).to(x.dtype)

    @staticmethod
    def backward(ctx, gb):
        (boxes,) = ctx.saved_tensors
        h, w = ctx.xshape[-2], ctx.xshape[-1]
        gx = box_grad(boxes, gb.contiguous(), h, w)
        return gx.reshape(ctx.xshape), None


def sat(x: torch.Tensor, scan: int = KOGGE_STONE) -> torch.Tensor:
    return SATFunction.apply(x, scan)


def box_sum(S: torch.Tensor, boxes: torch.Tensor) -> torch.Tensor:
    return BoxSumFunction.apply(S, boxes)


def box_filter(x: torch.Tensor, boxes: torch.Tensor) -> torch.Tensor:
    return BoxFilterFunction.apply(x, boxes)

"""Benchmark: summed-area-table throughput on one MI355X.

Default workload: one-pass SAT of an f32 8192 x 8192 image (values U[0,1)),
inputs resident in HBM. A step is one SAT of one image. Throughput is
algorithmic bytes (H*W*(sizeof(in) + sizeof(out))) per second; the roofline is
MI355X's 8 TB/s HBM peak. Consecutive steps rotate over enough distinct
input/output pairs (at least 2 GiB in all) that no step finds its input in the
256 MiB MALL left there by an earlier step; the back-to-back rate on one pair
is reported beside it. The reference's CPU integral (the oracle's restatement
of IntegralImageParallel / IntegralImageSerial) is timed on the host cores of
the same machine as the CPU baseline.

Prints ONE JSON line. ``--gpus`` above 1 is not supported (one GPU only).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
ROTATE_BYTES = 2 << 30  # bytes the rotating buffer pairs span at least
PROFILE_WORKLOADS = ("f32_8192", "u32_8192", "f64_4096", "f32_65536x256", "f32_256x65536", "u8u32_16384",
                     "u8u32_2048", "batched", "diff", "copy")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_launches(fn, steps: int, warmup: int):
    """Average ms per call over `steps` back-to-back calls fn(k), k = 0.., on
    the current stream (where the library launches its kernels)."""
    import torch

    for k in range(warmup):
        fn(k)
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for k in range(steps):
        fn(warmup + k)
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def time_each(fn, steps: int, before=None):
    """Median of per-call event timings (ms); `before` runs untimed ahead of
    each call."""
    import torch

    ts = []
    for k in range(steps):
        if before is not None:
            before()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn(k)
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e))
    ts.sort()
    return ts[len(ts) // 2]


def n_rotate(pair_bytes: int) -> int:
    """Distinct buffer pairs a rotating timing loop uses: at least 4 and at
    least ROTATE_BYTES in all."""
    return max(4, math.ceil(ROTATE_BYTES / max(1, pair_bytes)))


def rotating(x, out_dtype, nbuf: int):
    """nbuf copies of the device tensor x and nbuf outputs of its shape."""
    import torch

    xs = [x] + [x.clone() for _ in range(nbuf - 1)]
    ys = [torch.empty(x.shape, dtype=out_dtype, device=x.device) for _ in range(nbuf)]
    return xs, ys


def abi_integral(sat_amd, xs, ys, alg: int):
    """fn(k): one call of sat_integral_planes on pair k % len(xs) through the C
    ABI, its arguments built once. The Python wrapper's checks take tens of
    microseconds per call, about as long as the kernels of the 30 us
    workloads, so timing through it would time the host; this path keeps the
    GPU busy (the headline, at 105 us a call, is timed through the wrapper)."""
    import torch

    L = sat_amd.lib()
    st = torch.cuda.current_stream().cuda_stream
    pre = []
    for x, y in zip(xs, ys):
        h, w = x.shape[-2], x.shape[-1]
        pre.append((alg, 0, sat_amd.sat_dtype(x.dtype), sat_amd.sat_dtype(y.dtype), x.data_ptr(), x.stride(-2),
                    x.stride(-2) * h, y.data_ptr(), y.stride(-2), y.stride(-2) * h, w, h, 1, None, 0, st))
    f = L.sat_integral_planes

    def call(k):
        s = f(*pre[k % len(pre)])
        if s:
            sat_amd.check(s, "sat_integral_planes")

    return call


def check_faults(sat_amd, what: str) -> None:
    """Fails the run if a look-back of the library-managed workspace of the
    current stream gave up (its results would be invalid)."""
    import torch

    f = sat_amd.lib().sat_workspace_fault(None, torch.cuda.current_stream().cuda_stream)
    if f != 0:
        raise SystemExit(f"[bench] {what}: look-back fault word raised ({f}); results invalid")


def cpu_baseline(x, tout, budget_s: float = 12.0):
    """Mean of the reference's 6-run protocol (GetAvgTime, integralImage.h:199-211)
    for the OpenMP path on all host cores and the serial path on one core."""
    from oracle import oracle as O

    import statistics

    res = {}
    ncores = os.cpu_count() or 1
    O.set_threads(ncores)
    res["omp_threads"] = O.threads()
    for name, fn, threads in (("parallel", O.parallel, ncores), ("serial", O.serial, 1)):
        fn(x, tout)  # warm-up
        ts = []
        for _ in range(6):
            t0 = time.perf_counter()
            fn(x, tout)
            ts.append(time.perf_counter() - t0)
        res[name] = (sum(ts) / len(ts), threads)
        res[name + "_median"] = statistics.median(ts)
        if sum(ts) > budget_s:
            break
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extra", action="store_true", help="only the headline workload")
    ap.add_argument("--profile", action="store_true",
                    help="profiling run of one workload (--profile-workload): --steps + 1 rotating calls, then "
                         "3 plain copies of 1 GiB (the counter calibration), no checks and no CPU baseline")
    ap.add_argument("--profile-workload", default="f32_8192", choices=PROFILE_WORKLOADS)
    ap.add_argument("--profile-alg", type=int, default=3, choices=[0, 1, 2, 3],
                    help="algorithm of a --profile run of a single-image workload (0-2: the reference's "
                         "two-pass compositions)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the SAT path here has no multi-GPU mode "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import sat_amd

    dev = torch.device("cuda:0")
    if args.profile:
        profile_run(sat_amd, dev, args)
        return

    from oracle import oracle as O

    H = W = 8192
    log(f"[bench] f32 {H}x{W} one-pass SAT: generating input")
    x_np = O.gen("f32", (H, W), 20260101)
    x = torch.from_numpy(x_np).to(dev)
    algo_bytes = H * W * (4 + 4)
    nbuf = n_rotate(algo_bytes)
    xs, ys = rotating(x, torch.float32, nbuf)

    def step(k):
        sat_amd.integral(xs[k % nbuf], alg=sat_amd.SINGLE_PASS, out=ys[k % nbuf])

    step(0)
    torch.cuda.synchronize()
    # parity of the measured output (not timed)
    log("[bench] checking the output against the FP64 oracle")
    ref = O.sat_f64(x_np)
    got = ys[0].cpu().numpy().astype(np.float64)
    rel = float(np.max(np.abs(got - ref) / np.maximum(ref, 1e-30)))
    del got, ref

    log(f"[bench] timing {args.steps} steps after {args.warmup} warm-up steps, rotating over {nbuf} "
        f"input/output pairs ({nbuf * algo_bytes / 2**30:.1f} GiB)")
    ms = time_launches(step, args.steps, args.warmup)
    check_faults(sat_amd, "headline")
    gbs = algo_bytes / (ms * 1e-3) / 1e9
    ms_b2b = time_launches(lambda k: sat_amd.integral(xs[0], out=ys[0]), args.steps, args.warmup)
    check_faults(sat_amd, "headline back-to-back")

    # measured ceiling: the library's streaming copy of the same bytes (the
    # SAT's 16-byte loads and nontemporal stores without any carries), timed
    # the same way
    log("[bench] timing the streaming-copy ceiling on the same bytes")
    lib = sat_amd.lib()
    cstream = torch.cuda.current_stream().cuda_stream

    def copy_step(k):
        sat_amd.check(lib.sat_stream_copy(xs[k % nbuf].data_ptr(), ys[k % nbuf].data_ptr(), H * W * 4, cstream),
                      "sat_stream_copy")

    cms = time_launches(copy_step, args.steps, args.warmup)
    cms_b2b = time_launches(lambda k: copy_step(0), args.steps, args.warmup)
    copy_gbs = algo_bytes / (cms * 1e-3) / 1e9
    del xs, ys
    torch.cuda.empty_cache()

    extra = {"f32_8192_SinglePass_back_to_back": {
        "ms": round(ms_b2b, 4), "GBs": round(algo_bytes / (ms_b2b * 1e-3) / 1e9, 1),
        "note": "one input/output pair, back-to-back (the MALL may hold part of it)"},
        "copy_ceiling_back_to_back": {"ms": round(cms_b2b, 5),
                                      "GBs": round(algo_bytes / (cms_b2b * 1e-3) / 1e9, 1)}}
    if not args.no_extra:
        extra.update(extra_configs(O, sat_amd, dev))

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle restatement of the reference CPU integral)")
        r = cpu_baseline(x_np, np.float32)
        dt, cores = r["parallel"]
        cpu = {"value": algo_bytes / dt / 1e9, "unit": "GB/s", "cores": cores, "kind": "port",
               "sample": f"IntegralImageParallel f32 {H}x{W}, OpenMP on all host cores, mean of 6 runs",
               "median_GBs": algo_bytes / r["parallel_median"] / 1e9, "nproc": os.cpu_count(),
               "omp_threads": r["omp_threads"]}
        if "serial" in r:
            dts, _ = r["serial"]
            cpu["serial_1core_GBs"] = algo_bytes / dts / 1e9
            cpu["serial_1core_median_GBs"] = algo_bytes / r["serial_median"] / 1e9
            # the faster of the two CPU paths (the reference's OpenMP column
            # phase walks columns with a stride of a whole row)
            cpu["fastest_cpu_path"] = ("IntegralImageSerial, 1 core" if dts < dt else
                                       "IntegralImageParallel, all cores")
            cpu["fastest_cpu_GBs"] = algo_bytes / min(dt, dts) / 1e9

    # HBM bytes per launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs of
    # this same workload, scaled by a plain copy of known size (the committed
    # profile; see its commands)
    traffic = None
    tfile = ROOT / "profiles" / "sat_f32_8192_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("traffic_bytes")

    out = {
        "metric": "SAT GB/s (in+out bytes) & % HBM peak",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (splitmix64 U[0,1) pixels)",
        "config": {"workload": "one-pass SAT, f32 8192x8192, row-major, inputs resident in HBM, "
                               f"steps rotating over {nbuf} input/output pairs "
                               f"({nbuf * algo_bytes / 2**30:.1f} GiB)",
                   "height": H, "width": W, "algorithm": "single-pass 2-D decoupled look-back",
                   "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
                   "f32_max_rel_err_vs_fp64": rel},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": algo_bytes,
                     "copy_ceiling": {"kernel": "sat_stream_copy, 256 MiB in + 256 MiB out, rotating like the SAT",
                                      "GBs": round(copy_gbs, 1), "ms": round(cms, 5),
                                      "sat_frac_of_copy": round(gbs / copy_gbs, 4)}},
        "cpu_baseline": cpu,
    }
    if extra:
        out["extra"] = extra
    print(json.dumps(out))


def make_boxes(rng, planes, nbox, H, W):
    """Boxes {y0, x0, y1, x1} (inclusive): side uniform in [3, 63], centre
    uniform over the image, box shifted to lie inside."""
    import numpy as np

    sh = np.minimum(rng.integers(3, 64, (planes, nbox)), H)
    sw = np.minimum(rng.integers(3, 64, (planes, nbox)), W)
    cy = rng.integers(0, H, (planes, nbox))
    cx = rng.integers(0, W, (planes, nbox))
    y0 = np.clip(cy - sh // 2, 0, H - sh)
    x0 = np.clip(cx - sw // 2, 0, W - sw)
    return np.stack([y0, x0, y0 + sh - 1, x0 + sw - 1], -1).astype(np.int32)


BATCH_SEED, BATCH_N = 20260501, 512


def batch_shapes(seed: int = BATCH_SEED, n: int = BATCH_N):
    """(h, w, kind) of the batched workload: H and W uniform in [64, 2048],
    kind 'u8' (u8 -> u32) or 'f32' with probability 1/2 each."""
    import numpy as np

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        h, w = (int(v) for v in rng.integers(64, 2049, 2))
        out.append((h, w, "u8" if rng.integers(0, 2) == 0 else "f32"))
    return out


def batch_inputs(O, dev, seed: int = BATCH_SEED, n: int = BATCH_N):
    """Host inputs, device (input, output) pairs and algorithmic bytes of the
    batched workload."""
    import torch

    pairs, host, algo = [], [], 0
    for i, (h, w, kind) in enumerate(batch_shapes(seed, n)):
        x = O.gen(kind, (h, w), 20260600 + i)
        y = torch.empty((h, w), dtype=torch.uint32 if kind == "u8" else torch.float32, device=dev)
        algo += h * w * (5 if kind == "u8" else 8)
        pairs.append((torch.from_numpy(x).to(dev), y))
        host.append(x)
    return host, pairs, algo


def batched_config(O, sat_amd, dev, reps=10):
    """512 images, H and W uniform in [64, 2048], 50/50 u8->u32 and f32: one
    batched launch against one plain launch per image on the same buffers,
    issued from Python and replayed as one captured graph of 512 launches."""
    import numpy as np
    import torch

    host, pairs, algo = batch_inputs(O, dev)
    n = len(pairs)
    plan = sat_amd.BatchPlan(pairs)
    plan.run()
    torch.cuda.synchronize()
    # parity of a sample (not timed): every 16th image against the oracle
    worst = 0.0
    for i in range(0, n, 16):
        x, (xd, yd) = host[i], pairs[i]
        if x.dtype == np.uint8:
            got = yd.view(torch.int32).cpu().numpy().view(np.uint32)
            if not np.array_equal(got, O.serial(x, np.uint32)):
                raise AssertionError(f"batched image {i} differs from the oracle")
        else:
            ref = O.sat_f64(x)
            worst = max(worst, float(np.max(np.abs(yd.cpu().numpy() - ref) / np.maximum(ref, 1e-30))))

    def per_image(k=0):
        for xd, yd in pairs:
            sat_amd.integral(xd, out=yd)

    ms_b = time_launches(lambda k: plan.run(), reps, 2)
    plan.check_faults()
    # the same queue run by the library's other launch form (persistent
    # workgroups drawing items from per-XCD ticket counters, or one workgroup
    # per item)
    lib = sat_amd.lib()
    mode = lib.sat_batch_persistent_get()
    lib.sat_batch_persistent_set(1 - mode)
    plan_alt = sat_amd.BatchPlan(pairs)
    lib.sat_batch_persistent_set(mode)
    ms_alt = time_launches(lambda k: plan_alt.run(), reps, 2)
    plan_alt.check_faults()
    del plan_alt
    ms_p = time_launches(per_image, max(2, reps // 2), 1)
    check_faults(sat_amd, "per-image launches")
    # the per-image path without Python: the same 512 launches captured once
    # and replayed as one graph (each call is an epoch bump and a SAT launch)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        per_image()  # grows this stream's managed workspace to the largest image first
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        per_image()
    ms_g = time_launches(lambda k: g.replay(), reps, 2)
    torch.cuda.synchronize()
    # the graph's results are the batch's results: bit-identical for u8 -> u32
    # images; f32 ones may differ in the last bits (a decoupled look-back adds
    # a varying number of tile aggregates before it meets an inclusive prefix)
    plan.run()
    torch.cuda.synchronize()
    ref = [pairs[i][1].clone() for i in range(0, n, 32)]
    g.replay()
    torch.cuda.synchronize()
    for j, i in enumerate(range(0, n, 32)):
        yd = pairs[i][1]
        same = (torch.equal(yd.view(torch.int32), ref[j].view(torch.int32)) if yd.dtype == torch.uint32 else
                torch.allclose(yd, ref[j], rtol=1e-5, atol=0))
        if not same:
            raise AssertionError(f"graph replay of image {i} differs from the batched result")
    del ref
    res = {"images": n, "algorithmic_bytes": algo, "batched_ms": round(ms_b, 4),
           "batched_launch_form": "persistent" if mode else "one workgroup per item",
           ("batched_one_workgroup_per_item_ms" if mode else "batched_persistent_ms"): round(ms_alt, 4),
           "per_image_launch_ms_python_loop": round(ms_p, 4),
           "per_image_launch_ms_captured_graph": round(ms_g, 4),
           "speedup_vs_per_image_python_loop": round(ms_p / ms_b, 2),
           "speedup_vs_per_image_captured_graph": round(ms_g / ms_b, 2),
           "batched_GBs": round(algo / (ms_b * 1e-3) / 1e9, 1),
           "pct_peak": round(100 * algo / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 1), "f32_max_rel_err_sample": worst,
           "batch_tile_config": int(sat_amd.lib().sat_batch_tune_get(4))}
    log(f"[bench] batched: {ms_b:.3f} ms vs per-image {ms_p:.3f} ms from Python ({ms_p / ms_b:.2f}x), "
        f"{ms_g:.3f} ms as one graph ({ms_g / ms_b:.2f}x)")
    del plan, pairs, g
    torch.cuda.empty_cache()
    return res


def diff_inputs(O, dev):
    import numpy as np
    import torch

    N, Cc, H, W, nbox = 32, 3, 1024, 1024, 4096
    P = N * Cc
    x_np = O.gen("f32n", (P, H, W), 20260701)
    rng = np.random.default_rng(20260702)
    boxes_np = make_boxes(rng, P, nbox, H, W)
    g_np = O.gen("f32n", (P, nbox), 20260703)
    x = torch.from_numpy(x_np).to(dev).view(N, Cc, H, W)
    boxes = torch.from_numpy(boxes_np).to(dev)
    g = torch.from_numpy(g_np).to(dev)
    return x_np, boxes_np, g_np, x, boxes, g


def diff_config(O, sat_amd, dev, reps=10):
    """N=32, C=3, 1024x1024 f32 N(0,1); box sums over 4096 boxes per plane;
    forward = SAT + box readout, backward = the fused reverse SAT of the box
    corners (box_filter), and the same through SAT autograd + box_sum."""
    import numpy as np
    import torch

    x_np, boxes_np, g_np, x, boxes, g = diff_inputs(O, dev)
    N, Cc, H, W = x.shape
    P, nbox = boxes_np.shape[:2]
    xg = x.clone().requires_grad_(True)

    def fwd(k=0):
        with torch.no_grad():
            return sat_amd.box_filter(x, boxes)

    def fwd_bwd(k=0):
        xg.grad = None
        b = sat_amd.box_filter(xg, boxes)
        b.backward(g.view(b.shape))

    def fwd_bwd_sat(k=0):
        xg.grad = None
        b = sat_amd.box_sum(sat_amd.sat(xg), boxes)
        b.backward(g.view(b.shape))

    b = fwd()
    fwd_bwd()
    torch.cuda.synchronize()
    # accuracy (not timed) on 4 planes: box sums and gradient against FP64
    errs_b, errs_g = [], []
    bb = b.view(P, nbox).cpu().numpy().astype(np.float64)
    gx = xg.grad.view(P, H, W).cpu().numpy().astype(np.float64)
    errs_s = []
    S32 = sat_amd.integral(x).view(P, H, W)
    for p in (0, 31, 64, 95):
        S64 = O.sat_f64(x_np[p].astype(np.float64))
        errs_s.append(float(np.linalg.norm(S32[p].cpu().numpy().astype(np.float64) - S64) / np.linalg.norm(S64)))
        ref = O.box_sums_f64(S64, boxes_np[p][None])[0]
        errs_b.append(float(np.linalg.norm(bb[p] - ref) / np.linalg.norm(ref)))
        gref = box_grad_f64(boxes_np[p], g_np[p], H, W)
        errs_g.append(float(np.linalg.norm(gx[p] - gref) / np.linalg.norm(gref)))
    ms_f = time_launches(fwd, reps, 2)
    ms_fb = time_launches(fwd_bwd, reps, 2)
    ms_fb_sat = time_launches(fwd_bwd_sat, reps, 2)
    check_faults(sat_amd, "diff")
    # comparator: torch cumsum on the same GPU
    xc = x.clone().requires_grad_(True)

    def t_fwd(k=0):
        with torch.no_grad():
            return torch.cumsum(torch.cumsum(x, -1), -2)

    def t_fwd_bwd(k=0):
        xc.grad = None
        y = torch.cumsum(torch.cumsum(xc, -1), -2)
        y.backward(torch.ones_like(y))

    ms_tf = time_launches(t_fwd, reps, 2)
    ms_tfb = time_launches(t_fwd_bwd, max(2, reps // 2), 1)
    # the forward's SAT reads f32 and writes the f64 table the box sums read
    algo = P * H * W * (4 + 8)
    res = {"planes": P, "shape": [N, Cc, H, W], "boxes_per_plane": nbox, "fwd_ms": round(ms_f, 4),
           "fwd_bwd_ms": round(ms_fb, 4), "fwd_bwd_over_fwd": round(ms_fb / ms_f, 3),
           "fwd_bwd_via_sat_autograd_ms": round(ms_fb_sat, 4),
           "fwd_sat_GBs_f32_in_f64_out": round(algo / (ms_f * 1e-3) / 1e9, 1),
           "sat_normwise_rel_err_f32_vs_fp64": max(errs_s), "grad_normwise_rel_err_f32_vs_fp64": max(errs_g),
           "box_sum_normwise_rel_err_f32_vs_fp64": max(errs_b),
           "torch_cumsum_fwd_ms": round(ms_tf, 4), "torch_cumsum_fwd_bwd_ms": round(ms_tfb, 4)}
    log(f"[bench] diff: fwd {ms_f:.3f} ms, fwd+bwd {ms_fb:.3f} ms ({ms_fb / ms_f:.2f}x), "
        f"err sat {max(errs_s):.2e} box {max(errs_b):.2e} grad {max(errs_g):.2e}")
    del x, xg, xc, b, S32
    torch.cuda.empty_cache()
    return res


def box_grad_f64(boxes, g, H, W):
    """FP64 gradient of sum_b g_b * boxsum_b(SAT(x)) w.r.t. x for one plane:
    the suffix sums of the boxes' corner impulses."""
    import numpy as np

    gref = np.zeros((H + 1, W + 1))
    y0, x0, y1, x1 = boxes.T
    gp = g.astype(np.float64)
    np.add.at(gref, (y0, x0), gp)
    np.add.at(gref, (y0, x1 + 1), -gp)
    np.add.at(gref, (y1 + 1, x0), -gp)
    np.add.at(gref, (y1 + 1, x1 + 1), gp)
    return np.cumsum(np.cumsum(gref, 0), 1)[:H, :W]


def f64_error(O, x_np, got):
    """Error of an f64 SAT against the oracle's f64 IntegralImageSerial (the
    reference's summation order): max |err| in units of u * SAT(|x|) with
    u = 2^-53, normwise relative error, and the share of bit-identical values."""
    import numpy as np

    ref = O.serial(x_np, np.float64)
    scale = O.serial(np.abs(x_np), np.float64)
    err = np.abs(got - ref)
    return {"max_err_in_u_sat_abs": float(np.max(err / (2.0 ** -53 * scale + 1e-300))),
            "normwise_rel_err": float(np.linalg.norm(err) / np.linalg.norm(ref)),
            "bit_identical_share": float(np.mean(got == ref))}


def extra_configs(O, sat_amd, dev):
    """Secondary workloads: the other algorithms, dtypes and shapes. Each is the
    average of calls rotating over at least 2 GiB of input/output pairs (as the
    headline), except the 2048^2 lines: those are median-timed one call at a
    time behind a 512 MiB write."""
    import numpy as np
    import torch

    res = {}

    flush_buf = torch.empty(512 << 20, dtype=torch.uint8, device=dev)

    def run(name, x, out_dtype, alg, in_b, out_b, steps=10, flush=False):
        pair = x.numel() * (in_b + out_b)
        nbuf = 1 if flush else n_rotate(pair)
        xs, ys = rotating(x, out_dtype, nbuf)
        f = abi_integral(sat_amd, xs, ys, alg)
        f(0)
        torch.cuda.synchronize()
        if flush:
            # small images fit in the 256 MiB MALL: a 512 MiB write before every
            # timed call keeps them from being served from it (and keeps the GPU
            # busy while the call is submitted, so no launch gap is timed)
            ms = time_each(f, steps, lambda: flush_buf.fill_(1))
        else:
            ms = time_launches(f, steps, 2)
        check_faults(sat_amd, name)
        gbs = pair / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "GBs": round(gbs, 1), "pct_peak": round(100 * gbs / HBM_PEAK_GBS, 1)}
        if nbuf > 1:
            res[name]["rotating_pairs"] = nbuf
        log(f"[bench] {name}: {ms:.4f} ms {gbs:.1f} GB/s")
        y = ys[0]
        del xs, ys
        return y

    xf = torch.from_numpy(O.gen("f32", (8192, 8192), 20260101)).to(dev)
    for alg in (0, 1, 2):
        run(f"f32_8192_{sat_amd.ALG_NAMES[alg]}", xf, torch.float32, alg, 4, 4)
    del xf
    xu = torch.from_numpy(O.gen("u32_20", (8192, 8192), 20260102).view(np.int32)).to(dev).view(torch.uint32)
    run("u32_8192_SinglePass", xu, torch.uint32, 3, 4, 4)
    del xu
    xd_np = O.gen("f64n", (4096, 4096), 20260103)
    xd = torch.from_numpy(xd_np).to(dev)
    for alg in (3, 0, 1, 2):
        y = run(f"f64_4096_{sat_amd.ALG_NAMES[alg]}", xd, torch.float64, alg, 8, 8)
        res[f"f64_4096_{sat_amd.ALG_NAMES[alg]}"]["error_vs_oracle"] = f64_error(O, xd_np, y.cpu().numpy())
        del y
    del xd
    for h, w in ((65536, 256), (256, 65536)):
        xt = torch.from_numpy(O.gen("f32", (h, w), 20260104)).to(dev)
        run(f"f32_{h}x{w}_SinglePass", xt, torch.float32, 3, 4, 4)
        del xt
    x8 = torch.from_numpy(O.gen("u8", (16384, 16384), 20260105)).to(dev)
    run("u8u32_16384_SinglePass", x8, torch.uint32, 3, 1, 4)
    del x8
    x2 = torch.from_numpy(O.gen("u8", (2048, 2048), 20260106)).to(dev)
    for alg in (0, 1, 2, 3):
        run(f"u8u32_2048_{sat_amd.ALG_NAMES[alg]}", x2, torch.uint32, alg, 1, 4, flush=True)
    del x2, flush_buf
    torch.cuda.empty_cache()
    res["batched"] = batched_config(O, sat_amd, dev)
    res["diff"] = diff_config(O, sat_amd, dev)
    return res


# ---------------------------------------------------------------------------
# Profiling runs: one workload, no checks, inputs from torch's generator so
# that no host-side oracle is needed
# ---------------------------------------------------------------------------
SINGLE = {  # name: (shape, input kind, output dtype name, in bytes, out bytes)
    "f32_8192": ((8192, 8192), "f32", "float32", 4, 4),
    "u32_8192": ((8192, 8192), "u32", "uint32", 4, 4),
    "f64_4096": ((4096, 4096), "f64", "float64", 8, 8),
    "f32_65536x256": ((65536, 256), "f32", "float32", 4, 4),
    "f32_256x65536": ((256, 65536), "f32", "float32", 4, 4),
    "u8u32_16384": ((16384, 16384), "u8", "uint32", 1, 4),
    "u8u32_2048": ((2048, 2048), "u8", "uint32", 1, 4),
    "copy": ((8192, 8192), "f32", "float32", 4, 4),
}


def _rand(kind, shape, dev):
    import torch

    g = torch.Generator(device=dev).manual_seed(20260999)
    if kind == "f32":
        return torch.rand(shape, generator=g, device=dev)
    if kind == "f64":
        return torch.randn(shape, generator=g, device=dev, dtype=torch.float64)
    if kind == "u8":
        return torch.randint(0, 256, shape, generator=g, device=dev, dtype=torch.uint8)
    if kind == "u32":
        return torch.randint(0, 1 << 20, shape, generator=g, device=dev, dtype=torch.int32).view(torch.uint32)
    raise ValueError(kind)


def profile_run(sat_amd, dev, args):
    import torch

    wl = args.profile_workload
    lib = sat_amd.lib()
    stream = torch.cuda.current_stream().cuda_stream
    if wl in SINGLE:
        shape, kind, odt, ib, ob = SINGLE[wl]
        x = _rand(kind, shape, dev)
        nbuf = n_rotate(x.numel() * (ib + ob))
        xs, ys = rotating(x, getattr(torch, odt), nbuf)
        if wl == "copy":
            fn = lambda k: sat_amd.check(lib.sat_stream_copy(  # noqa: E731
                xs[k % nbuf].data_ptr(), ys[k % nbuf].data_ptr(), x.numel() * ib, stream), "sat_stream_copy")
        else:
            fn = lambda k: sat_amd.integral(xs[k % nbuf], out_dtype=getattr(torch, odt),  # noqa: E731
                                            alg=args.profile_alg, out=ys[k % nbuf])
        for k in range(args.steps + 1):
            fn(k)
        what = f"{args.steps + 1} rotating calls over {nbuf} pairs"
    elif wl == "batched":
        from oracle import oracle as O

        _, pairs, _ = batch_inputs(O, dev)
        plan = sat_amd.BatchPlan(pairs)
        for _ in range(args.steps + 1):
            plan.run()
        what = f"{args.steps + 1} batched launches"
    else:  # diff
        from oracle import oracle as O

        _, _, _, x, boxes, g = diff_inputs(O, dev)
        xg = x.clone().requires_grad_(True)
        for _ in range(args.steps + 1):
            xg.grad = None
            b = sat_amd.box_filter(xg, boxes)
            b.backward(g.view(b.shape))
        what = f"{args.steps + 1} box_filter forward + backward"
    torch.cuda.synchronize()
    # counter calibration: plain copies of a known size (1 GiB each way)
    c0 = torch.empty(64 << 22, dtype=torch.float32, device=dev)
    c1 = torch.empty_like(c0)
    for _ in range(3):
        c1.copy_(c0)
    torch.cuda.synchronize()
    log(f"[bench] profile run of {wl} done: {what}, then 3 torch copies of {c0.numel() * 4} bytes")


if __name__ == "__main__":
    main()

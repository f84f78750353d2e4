"""Benchmark: summed-area-table throughput on one MI355X.

Default workload: one-pass SAT of an f32 8192 x 8192 image (values U[0,1)),
inputs resident in HBM. A step is one SAT of the image. Throughput is
algorithmic bytes (H*W*(sizeof(in) + sizeof(out))) per second; the roofline is
MI355X's 8 TB/s HBM peak. The reference's CPU integral (the oracle's
restatement of IntegralImageParallel / IntegralImageSerial) is timed on the
host cores of the same machine as the CPU baseline.

Prints ONE JSON line. ``--gpus`` above 1 is not supported (one GPU only).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_launches(fn, steps: int, warmup: int):
    """Average ms per call over `steps` back-to-back calls on the current stream."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def time_each(fn, steps: int, before=None):
    """Median of per-call event timings (ms); `before` runs untimed ahead of
    each call."""
    import torch

    ts = []
    for _ in range(steps):
        if before is not None:
            before()
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        fn()
        e.record()
        torch.cuda.synchronize()
        ts.append(s.elapsed_time(e))
    ts.sort()
    return ts[len(ts) // 2]


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
                    help="profiling run: the headline SAT --steps times, then 3 plain copies of the same "
                         "bytes (the counter calibration), no checks and no CPU baseline")
    ap.add_argument("--profile-alg", type=int, default=3, choices=[0, 1, 2, 3],
                    help="algorithm of the --profile run (0-2: the reference's two-pass compositions)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the SAT path here has no multi-GPU mode "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import sat_amd
    from oracle import oracle as O

    dev = torch.device("cuda:0")
    H = W = 8192
    log(f"[bench] f32 {H}x{W} one-pass SAT: generating input")
    x_np = O.gen("f32", (H, W), 20260101)
    x = torch.from_numpy(x_np).to(dev)
    y = torch.empty_like(x)

    alg = args.profile_alg if args.profile else sat_amd.SINGLE_PASS

    def step():
        sat_amd.integral(x, alg=alg, out=y)

    step()
    torch.cuda.synchronize()
    if args.profile:
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        y2 = torch.empty_like(x)
        for _ in range(3):
            y2.copy_(x)
        lib = sat_amd.lib()
        for _ in range(args.steps):
            sat_amd.check(lib.sat_stream_copy(x.data_ptr(), y2.data_ptr(), H * W * 4,
                                              torch.cuda.current_stream().cuda_stream), "sat_stream_copy")
        torch.cuda.synchronize()
        log(f"[bench] profile run done: {args.steps} SAT launches, 3 torch copies and {args.steps} "
            f"sat_stream_copy calls of {H * W * 4} bytes")
        return
    # parity of the measured output (not timed)
    log("[bench] checking the output against the FP64 oracle")
    ref = O.sat_f64(x_np)
    got = y.cpu().numpy().astype(np.float64)
    rel = float(np.max(np.abs(got - ref) / np.maximum(ref, 1e-30)))
    del got, ref

    log(f"[bench] timing {args.steps} steps after {args.warmup} warm-up steps")
    ms = time_launches(step, args.steps, args.warmup)
    algo_bytes = H * W * (4 + 4)
    gbs = algo_bytes / (ms * 1e-3) / 1e9

    # measured ceiling: the library's streaming copy of the same bytes (the
    # SAT's 16-byte loads and nontemporal stores without any carries)
    log("[bench] timing the streaming-copy ceiling on the same bytes")
    y2 = torch.empty_like(x)
    lib = sat_amd.lib()
    cstream = torch.cuda.current_stream().cuda_stream

    def copy_step():
        sat_amd.check(lib.sat_stream_copy(x.data_ptr(), y2.data_ptr(), H * W * 4, cstream), "sat_stream_copy")

    cms = time_launches(copy_step, args.steps, args.warmup)
    copy_gbs = algo_bytes / (cms * 1e-3) / 1e9
    del y2

    extra = {}
    if not args.no_extra:
        extra = extra_configs(O, sat_amd, dev)

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
        "config": {"workload": "one-pass SAT, f32 8192x8192, row-major, inputs resident in HBM",
                   "height": H, "width": W, "algorithm": "single-pass 2-D decoupled look-back",
                   "pct_hbm_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
                   "f32_max_rel_err_vs_fp64": rel},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": algo_bytes,
                     "copy_ceiling": {"kernel": "sat_stream_copy, 256 MiB in + 256 MiB out",
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


def batched_config(O, sat_amd, dev, reps=10):
    """512 images, H and W uniform in [64, 2048], 50/50 u8->u32 and f32: one
    batched launch against one plain launch per image, same buffers."""
    import numpy as np
    import torch

    rng = np.random.default_rng(20260501)
    n = 512
    pairs, host, algo = [], [], 0
    for i in range(n):
        h, w = (int(v) for v in rng.integers(64, 2049, 2))
        if rng.integers(0, 2) == 0:
            x = O.gen("u8", (h, w), 20260600 + i)
            y = torch.empty((h, w), dtype=torch.uint32, device=dev)
            algo += h * w * 5
        else:
            x = O.gen("f32", (h, w), 20260600 + i)
            y = torch.empty((h, w), dtype=torch.float32, device=dev)
            algo += h * w * 8
        pairs.append((torch.from_numpy(x).to(dev), y))
        host.append(x)
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

    def per_image():
        for xd, yd in pairs:
            sat_amd.integral(xd, out=yd)

    ms_b = time_launches(plan.run, reps, 2)
    ms_p = time_launches(per_image, max(2, reps // 2), 1)
    res = {"images": n, "algorithmic_bytes": algo, "batched_ms": round(ms_b, 4), "per_image_launch_ms": round(ms_p, 4),
           "speedup_vs_per_image": round(ms_p / ms_b, 2), "batched_GBs": round(algo / (ms_b * 1e-3) / 1e9, 1),
           "pct_peak": round(100 * algo / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS, 1), "f32_max_rel_err_sample": worst,
           "batch_tile_config": int(sat_amd.lib().sat_batch_tune_get(4))}
    log(f"[bench] batched: {ms_b:.3f} ms vs per-image {ms_p:.3f} ms ({ms_p / ms_b:.2f}x)")
    del plan, pairs
    torch.cuda.empty_cache()
    return res


def diff_config(O, sat_amd, dev, reps=10):
    """N=32, C=3, 1024x1024 f32 N(0,1); box sums over 4096 boxes per plane;
    forward = SAT + box readout, backward = the fused reverse SAT of the box
    corners (box_filter), and the same through SAT autograd + box_sum."""
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
    xg = x.clone().requires_grad_(True)

    def fwd():
        with torch.no_grad():
            return sat_amd.box_filter(x, boxes)

    def fwd_bwd():
        xg.grad = None
        b = sat_amd.box_filter(xg, boxes)
        b.backward(g.view(b.shape))

    def fwd_bwd_sat():
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
        gref = np.zeros((H + 1, W + 1))
        y0, x0, y1, x1 = boxes_np[p].T
        gp = g_np[p].astype(np.float64)
        np.add.at(gref, (y0, x0), gp)
        np.add.at(gref, (y0, x1 + 1), -gp)
        np.add.at(gref, (y1 + 1, x0), -gp)
        np.add.at(gref, (y1 + 1, x1 + 1), gp)
        gref = np.cumsum(np.cumsum(gref, 0), 1)[:H, :W]
        errs_g.append(float(np.linalg.norm(gx[p] - gref) / np.linalg.norm(gref)))
    ms_f = time_launches(fwd, reps, 2)
    ms_fb = time_launches(fwd_bwd, reps, 2)
    ms_fb_sat = time_launches(fwd_bwd_sat, reps, 2)
    # comparator: torch cumsum on the same GPU
    xc = x.clone().requires_grad_(True)

    def t_fwd():
        with torch.no_grad():
            return torch.cumsum(torch.cumsum(x, -1), -2)

    def t_fwd_bwd():
        xc.grad = None
        y = torch.cumsum(torch.cumsum(xc, -1), -2)
        y.backward(torch.ones_like(y))

    ms_tf = time_launches(t_fwd, reps, 2)
    ms_tfb = time_launches(t_fwd_bwd, max(2, reps // 2), 1)
    algo = P * H * W * 8
    res = {"planes": P, "shape": [N, Cc, H, W], "boxes_per_plane": nbox, "fwd_ms": round(ms_f, 4),
           "fwd_bwd_ms": round(ms_fb, 4), "fwd_bwd_over_fwd": round(ms_fb / ms_f, 3),
           "fwd_bwd_via_sat_autograd_ms": round(ms_fb_sat, 4),
           "fwd_sat_GBs": round(algo / (ms_f * 1e-3) / 1e9, 1),
           "sat_normwise_rel_err_f32_vs_fp64": max(errs_s), "grad_normwise_rel_err_f32_vs_fp64": max(errs_g),
           # a box sum is a difference of four SAT values ~30x its size: the
           # f32 SAT's rounding, amplified by that cancellation
           "box_sum_normwise_rel_err_f32_vs_fp64": max(errs_b),
           "torch_cumsum_fwd_ms": round(ms_tf, 4), "torch_cumsum_fwd_bwd_ms": round(ms_tfb, 4)}
    log(f"[bench] diff: fwd {ms_f:.3f} ms, fwd+bwd {ms_fb:.3f} ms ({ms_fb / ms_f:.2f}x), "
        f"err sat {max(errs_s):.2e} box {max(errs_b):.2e} grad {max(errs_g):.2e}")
    del x, xg, xc, b, S32
    torch.cuda.empty_cache()
    return res


def extra_configs(O, sat_amd, dev):
    """Secondary workloads: the other algorithms, dtypes and shapes. Each is the
    average of back-to-back calls (as the headline), except the 2048^2 lines:
    those are median-timed one call at a time behind a 512 MiB write."""
    import numpy as np
    import torch

    res = {}

    flush_buf = torch.empty(512 << 20, dtype=torch.uint8, device=dev)

    def run(name, x, out_dtype, alg, in_b, out_b, steps=10, flush=False):
        y = torch.empty(x.shape, dtype=out_dtype, device=dev)
        f = lambda: sat_amd.integral(x, out_dtype=out_dtype, alg=alg, out=y)  # noqa: E731
        f()
        torch.cuda.synchronize()
        if flush:
            # small images fit in the 256 MiB MALL: a 512 MiB write before every
            # timed call keeps them from being served from it (and keeps the GPU
            # busy while the call is submitted, so no launch gap is timed)
            ms = time_each(f, steps, lambda: flush_buf.fill_(1))
        else:
            ms = time_launches(f, steps, 2)
        gbs = x.numel() * (in_b + out_b) / (ms * 1e-3) / 1e9
        res[name] = {"ms": round(ms, 4), "GBs": round(gbs, 1), "pct_peak": round(100 * gbs / HBM_PEAK_GBS, 1)}
        log(f"[bench] {name}: {ms:.4f} ms {gbs:.1f} GB/s")
        return y

    xf = torch.from_numpy(O.gen("f32", (8192, 8192), 20260101)).to(dev)
    for alg in (0, 1, 2):
        run(f"f32_8192_{sat_amd.ALG_NAMES[alg]}", xf, torch.float32, alg, 4, 4)
    del xf
    xu = torch.from_numpy(O.gen("u32_20", (8192, 8192), 20260102).view(np.int32)).to(dev).view(torch.uint32)
    run("u32_8192_SinglePass", xu, torch.uint32, 3, 4, 4)
    del xu
    xd = torch.from_numpy(O.gen("f64n", (4096, 4096), 20260103)).to(dev)
    run("f64_4096_SinglePass", xd, torch.float64, 3, 8, 8)
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
    del x2
    torch.cuda.empty_cache()
    res["batched"] = batched_config(O, sat_amd, dev)
    res["diff"] = diff_config(O, sat_amd, dev)
    return res


if __name__ == "__main__":
    main()

"""Sweep harness over the `sat` CLI: the reference's benchmark protocol.

The reference benchmarks with one script per (algorithm, input, output type),
e.g. ``nvprof --print-gpu-trace --csv ./release/sat 0 0 6 6 20``
(integralImage/0066.sh), run together by test.sh: algorithms 0-2 x pairs
u8->u32 (0 4), u8->f32 (0 6), f32->f32 (6 6), f64->f64 (7 7), i32->i32
(5 5), square sizes 1024*i for i = 1..n. This module runs the same grid (plus
the one-pass algorithm 3) through ``sat_amd/bin/sat`` and gathers one CSV
(alg, scan, dt0, dt1, width, height, ms, GBs); with ``--rocprof DIR`` every
run goes under ``rocprofv3 --kernel-trace --stats -f csv``, which writes the
per-kernel trace CSV the reference's nvprof runs produced.

    python -m sat_amd.sweep --n 8 --csv sweep.csv
    python -m sat_amd.sweep --algs 3 --pairs 6:6 --n 16 --rocprof sweep_prof
"""
from __future__ import annotations

import argparse
import subprocess
import sys
from pathlib import Path

CLI = Path(__file__).resolve().parent / "bin" / "sat"
# (dt0, dt1) of the reference's scripts (integralImage/*.sh) plus i32 -> i32
REFERENCE_PAIRS = [(0, 4), (0, 6), (6, 6), (7, 7), (5, 5)]
HEADER = "alg,scan,dt0,dt1,width,height,ms,GBs"


def sweep(algs=(0, 1, 2, 3), pairs=REFERENCE_PAIRS, n=20, scan=0, csv=None, rocprof=None, check=False,
          log=print) -> list[str]:
    """Run the CLI over the grid; returns the CSV rows (header first)."""
    if not CLI.exists():
        raise FileNotFoundError(f"{CLI} is missing: build with python -m sat_amd.build")
    out = Path(csv) if csv else Path("sat_sweep.csv")
    out.write_text(HEADER + "\n")
    for alg in algs:
        for dt0, dt1 in pairs:
            cmd = [str(CLI), str(alg), str(scan), str(dt0), str(dt1), str(n), "--csv", str(out)]
            if check:
                cmd.append("--check")
            if rocprof:
                d = Path(rocprof) / f"{alg}{scan}{dt0}{dt1}"
                cmd = ["rocprofv3", "--kernel-trace", "--stats", "-f", "csv", "-d", str(d), "-o", "sat", "--", *cmd]
            log(f"[sweep] {' '.join(cmd)}")
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"sat {alg} {scan} {dt0} {dt1} {n} failed ({r.returncode}):\n"
                                   f"{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    rows = out.read_text().strip().splitlines()
    log(f"[sweep] {len(rows) - 1} rows in {out}")
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--algs", default="0,1,2,3")
    ap.add_argument("--pairs", default=",".join(f"{a}:{b}" for a, b in REFERENCE_PAIRS), help="dt0:dt1,...")
    ap.add_argument("--n", type=int, default=20, help="sizes 1024*i, i = 1..n")
    ap.add_argument("--scan", type=int, default=0)
    ap.add_argument("--csv", default="sat_sweep.csv")
    ap.add_argument("--rocprof", default=None, help="directory for rocprofv3 kernel traces")
    ap.add_argument("--check", action="store_true", help="cross-check each result against another algorithm")
    a = ap.parse_args(argv)
    algs = [int(v) for v in a.algs.split(",")]
    pairs = [tuple(int(v) for v in p.split(":")) for p in a.pairs.split(",")]
    sweep(algs, pairs, a.n, a.scan, a.csv, a.rocprof, a.check, log=lambda *m: print(*m, file=sys.stderr, flush=True))


if __name__ == "__main__":
    main()

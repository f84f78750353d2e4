"""Build the HIP library of this package in place: sat_amd/libsat_amd.so.

Every translation unit under sat_amd/csrc is compiled for gfx950 with hipcc
(in parallel), then linked into one shared library. Run as
``python -m sat_amd.build`` or through ``__graft_entry__.build()``.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_build"
LIB = PKG / "libsat_amd.so"
ARCH = os.environ.get("SAT_AMD_ARCH", "gfx950")

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 VALU forms are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build sat_amd")


def _compile(src: Path) -> Path:
    obj = OUT_DIR / (src.stem + ".o")
    cmd = [_hipcc(), *HIPCC_FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-4000:]}")
    return obj


TUNED_JSON = PKG / "tuned_gfx950.json"
TUNED_INC = CSRC / "tuned_table.inc"


def write_tuned_table() -> Path:
    """The autotuner's table (TUNED_JSON, written by ``python -m sat_amd.tune``)
    as C++ initialisers compiled into the library, so C and C++ callers start
    from the tuned tile configurations too. A table for another architecture
    than the build's is left out."""
    import json

    entries, batch, arch = [], [], None
    if TUNED_JSON.exists():
        doc = json.loads(TUNED_JSON.read_text())
        arch = doc.get("arch")
        if arch == ARCH:
            entries = [(int(e["in_bytes"]), int(e["out_bytes"]), int(e["class_x"]), int(e["class_y"]),
                        int(e["class_p"]), int(e["cfg"])) for e in doc.get("entries", [])]
            batch = [(int(e["out_bytes"]), int(e["cfg"])) for e in doc.get("batch", [])]
    lines = [f"// Generated by sat_amd/build.py from sat_amd/{TUNED_JSON.name} (arch {arch}, build {ARCH}).",
             "// {in_bytes, out_bytes, class_x, class_y, class_p, cfg}",
             "static const ShippedTune k_shipped_tune[] = {"]
    lines += [f"    {{{ib}, {ob}, {cx}, {cy}, {cp}, {cfg}}}," for ib, ob, cx, cy, cp, cfg in entries]
    lines += ["    {0, 0, 0, 0, 0, -1},  // end", "};", "// {out_bytes, cfg}", "static const ShippedBatch k_shipped_batch[] = {"]
    lines += [f"    {{{ob}, {cfg}}}," for ob, cfg in batch]
    lines += ["    {0, -1},  // end", "};", ""]
    text = "\n".join(lines)
    if not TUNED_INC.exists() or TUNED_INC.read_text() != text:
        TUNED_INC.write_text(text)
    return TUNED_INC


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    write_tuned_table()
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = jobs or min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(objs)} translation units", file=sys.stderr)
    build_cli(verbose)
    return LIB


CLI_SRC = PKG / "cli" / "sat_cli.cpp"
CLI = PKG / "bin" / "sat"


def build_cli(verbose: bool = True) -> Path:
    """The `sat` command-line driver (host code over the C ABI), linked against
    libsat_amd.so with an rpath to it."""
    CLI.parent.mkdir(exist_ok=True)
    tmp = CLI.with_suffix(".tmp")
    cmd = [_hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-Wno-unused-command-line-argument", str(CLI_SRC),
           f"-L{PKG}",
           "-lsat_amd", "-Wl,-rpath,$ORIGIN/..", "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"CLI build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, CLI)
    if verbose:
        print(f"built {CLI}", file=sys.stderr)
    return CLI


TESTS_CPP = PKG.parent / "tests" / "cpp"


def build_cpp_tests(verbose: bool = True) -> list:
    """C++ programs of the test suite (tests/cpp/*.cpp, host code over the C
    ABI and the C++ headers) -> tests/cpp/_build/<name>."""
    out = []
    if not TESTS_CPP.is_dir():
        return out
    (TESTS_CPP / "_build").mkdir(exist_ok=True)
    for src in sorted(TESTS_CPP.glob("*.cpp")):
        exe = TESTS_CPP / "_build" / src.stem
        cmd = [_hipcc(), f"--offload-arch={ARCH}", "-O2", "-std=c++17", "-Wno-unused-command-line-argument", str(src),
               f"-L{PKG}", "-lsat_amd", "-Wl,-rpath,$ORIGIN/../../../sat_amd", "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"test program {src.name} failed to build:\n{r.stderr[-4000:]}")
        out.append(exe)
    if verbose and out:
        print(f"built {len(out)} C++ test program(s) in {TESTS_CPP / '_build'}", file=sys.stderr)
    return out


if __name__ == "__main__":
    build()
    build_cpp_tests()

"""Builds libdamavand-gpu.so for gfx950 in-tree (damavand_amd/libdamavand-gpu.so).

    python -m damavand_amd.build [--jobs N] [--verbose] [--incremental] [--sanitize]

Every source under damavand_amd/csrc is compiled with hipcc (objects under
damavand_amd/_build/) and linked into one shared library exporting the C ABI
declared in include/damavand_gpu.h.  No GPU is needed to build.  By default
every unit is compiled; --incremental recompiles only units whose object is
older than the unit or than any header under csrc/ and include/.

--sanitize builds a second copy whose host code (planner, argument checks,
runtime) is instrumented with AddressSanitizer and UndefinedBehaviorSanitizer
(device code is not instrumented), into damavand_amd/_build_san/; load it with
DMV_LIBRARY=<path> and clang's ASan runtime preloaded (tests/run_sanitized.py).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libdamavand-gpu.so")
SAN_DIR = os.path.join(HERE, "_build_san")
SAN_LIB = os.path.join(SAN_DIR, "libdamavand-gpu.so")
# host-only instrumentation: each -fsanitize applies to the host pass alone
SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-omit-frame-pointer", "-Xarch_host", "-fno-sanitize-recover=undefined"]
ARCH = os.environ.get("DMV_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    f"--offload-arch={ARCH}",
    # no packed FP32 instructions in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep the per-mask switch cases of the tile interpreter apart: sinking
    # their common code merges register-array accesses into pointer phis,
    # which sends the tile's registers to scratch memory
    "-mllvm", "-simplifycfg-sink-common=false",
    # and keep them from being folded into select chains: with 8 registers
    # per thread the pair ops' small cases were otherwise speculated and
    # merged, indexing the register array by the runtime mask (a 16-way
    # compare-and-select per amplitude, twice the instructions of the op)
    "-mllvm", "-two-entry-phi-node-folding-threshold=0",
    "-I", os.path.join(ROOT, "include"),
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libdamavand-gpu)")


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src: str, verbose: bool, sanitize: bool = False) -> str:
    obj = os.path.join(SAN_DIR if sanitize else OBJ, os.path.basename(src) + ".o")
    cmd = [hipcc()] + COMMON + (SANITIZE if sanitize else []) + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compiling {os.path.basename(src)} failed:\n{r.stdout}\n{r.stderr}")
    # the host pass does not know the packed-fp32 feature and says so; anything else is shown
    noise = "'-packed-fp32-ops' is not a recognized feature"
    msg = "\n".join(line for line in r.stderr.splitlines() if noise not in line and "warning generated" not in line
                    and "warnings generated" not in line)
    if msg.strip() and verbose:
        print(msg, file=sys.stderr)
    return obj


def _deps(path: str, seen=None) -> set:
    """The file and every quoted #include it reaches (csrc/ and include/)."""
    seen = set() if seen is None else seen
    if path in seen or not os.path.exists(path):
        return seen
    seen.add(path)
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("#include") and '"' in line:
                name = line.split('"')[1]
                _deps(os.path.normpath(os.path.join(os.path.dirname(path), name)), seen)
    return seen


def _up_to_date(src: str) -> bool:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    if not os.path.exists(obj):
        return False
    t = os.path.getmtime(obj)
    return all(t >= os.path.getmtime(d) for d in _deps(src) | {__file__})


def build(jobs: int = 0, verbose: bool = False, incremental: bool = False, sanitize: bool = False) -> str:
    out_dir, lib = (SAN_DIR, SAN_LIB) if sanitize else (OBJ, LIB)
    os.makedirs(out_dir, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 1)
    todo = [s for s in srcs if not (incremental and not sanitize and _up_to_date(s))]
    # longest compiles first (the kernel instantiation units), so they overlap
    order = sorted(todo, key=lambda s: -os.path.getsize(s) if s.endswith(".hip") else 0)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(lambda s: _compile(s, verbose, sanitize), order))
    objs = [os.path.join(out_dir, os.path.basename(s) + ".o") for s in srcs]
    tmp = lib + ".tmp"
    link_san = ["-fsanitize=address,undefined", "-shared-libsan"] if sanitize else []
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}"] + link_san + ["-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"linking libdamavand-gpu.so failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, lib)
    return lib


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=0)
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--sanitize", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose, a.incremental, a.sanitize))


if __name__ == "__main__":
    main()

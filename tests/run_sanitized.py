"""Runs the CPU tests of the library's host code and of the oracle under
AddressSanitizer and UndefinedBehaviorSanitizer.

    python tests/run_sanitized.py [--no-build] [extra pytest arguments]

Builds the instrumented copies (``python -m damavand_amd.build --sanitize``:
host code of libdamavand-gpu with -fsanitize=address,undefined, device code
unchanged; ``make -C oracle san``), then runs tests/test_host.py,
tests/test_oracle.py, tests/test_qml_host.py and tests/test_tile_emulation.py
(the planner's plans, checked index by index) in a child process with
clang's ASan runtime preloaded and DMV_LIBRARY / DMV_ORACLE_LIBRARY pointing
at the instrumented builds.  Any sanitizer report fails the run
(halt_on_error); leak checking is off because the Python interpreter itself
is not instrumented.  No GPU is needed.
"""
import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def asan_runtime() -> str:
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not found:
        raise SystemExit("clang's ASan runtime (libclang_rt.asan-x86_64.so) was not found under /opt/rocm/llvm")
    return found[-1]


def main() -> int:
    args = sys.argv[1:]
    if "--no-build" in args:
        args.remove("--no-build")
    else:
        subprocess.run([sys.executable, "-m", "damavand_amd.build", "--sanitize"], cwd=ROOT, check=True)
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "san"], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = asan_runtime()
    env["LD_LIBRARY_PATH"] = "/opt/rocm/llvm/lib:" + env.get("LD_LIBRARY_PATH", "")
    env["DMV_LIBRARY"] = os.path.join(ROOT, "damavand_amd", "_build_san", "libdamavand-gpu.so")
    env["DMV_ORACLE_LIBRARY"] = os.path.join(ROOT, "oracle", "_san", "liboracle.so")
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    tests = [os.path.join(ROOT, "tests", t)
             for t in ("test_host.py", "test_oracle.py", "test_qml_host.py", "test_tile_emulation.py")]
    cmd = [sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-m", "not gpu"] + tests + args
    print("sanitized run:", " ".join(cmd), flush=True)
    return subprocess.run(cmd, cwd=ROOT, env=env).returncode


if __name__ == "__main__":
    sys.exit(main())

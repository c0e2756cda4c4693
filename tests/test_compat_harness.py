"""The compat layer driven from C exactly as the reference's Rust crate drives
it (tests/compat_harness.c): built against the installed header and library on
the CPU, run on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "damavand_amd")


def build_harness(out_dir):
    exe = os.path.join(out_dir, "compat_harness")
    cmd = ["gcc", "-O2", "-std=c11", "-Wall", "-D_DEFAULT_SOURCE", os.path.join(ROOT, "tests", "compat_harness.c"),
           "-I", os.path.join(ROOT, "include"), "-L", LIBDIR, "-ldamavand-gpu", f"-Wl,-rpath,{LIBDIR}", "-lm",
           "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_harness_builds_and_links(tmp_path):
    """Every compat symbol the harness (and the Rust crate) names resolves in the library."""
    exe = build_harness(str(tmp_path))
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_harness_runs(tmp_path):
    exe = build_harness(str(tmp_path))
    print(f"case compat harness {exe}", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "all checks passed" in r.stdout

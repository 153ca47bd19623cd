"""The headers are plain C: a C99 program that includes both, links
libatr_recon.so and calls the entry points that need no GPU builds with
-Wall -Werror -pedantic and runs clean (tests/c_consumer.c)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "aretomo3_amd")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_c_program_builds_and_runs(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libatr_recon.so")):
        pytest.skip("library not built")
    exe = str(tmp_path / "c_consumer")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
           "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c_consumer.c"),
           "-o", exe, "-L", LIBDIR, "-latr_recon", "-Wl,-rpath," + LIBDIR, "-lm"]
    build = subprocess.run(cmd, capture_output=True, text=True)
    assert build.returncode == 0, build.stderr
    run = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0 and run.stdout.strip() == "ok", run.stdout + run.stderr

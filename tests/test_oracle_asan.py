"""The CPU oracle under AddressSanitizer and UBSan: a host build of
oracle/vz_oracle.c (gcc, plain C) runs every entry point on small scenes
(the book scene, box scenes of 1 to 5001 spheres and of 1024 and 131072
(the SAH build and the SAH treelets), duplicate centres, FP32
and FP64 accumulation, row subsets, the ray recorder, brute force and BVH
traversal) from exact-size heap buffers; any out-of-bounds access or
undefined behaviour aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_asan")
    build = ["gcc", "-std=c11", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-I" + os.path.join(ROOT, "oracle"),
             os.path.join(ROOT, "tests", "oracle_asan_driver.c"), os.path.join(ROOT, "oracle", "vz_oracle.c"),
             "-o", exe, "-lm", "-lpthread"]
    b = subprocess.run(build, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "oracle sanitizer run: ok" in r.stdout

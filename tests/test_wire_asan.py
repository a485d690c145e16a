"""The wire-format decoders under AddressSanitizer and UBSan: a host build
of venusaur_amd/csrc/vz_wire.cpp (g++, no GPU code) decodes 200,000
truncated, extended and bit-flipped scene and request records from
exact-size heap buffers; any out-of-bounds read or undefined behaviour
aborts the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the ROCm headers")
def test_wire_decoders_under_sanitizers(tmp_path):
    exe = str(tmp_path / "wire_fuzz")
    build = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
             "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
             "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "wire_fuzz_asan.cpp"),
             os.path.join(ROOT, "venusaur_amd", "csrc", "vz_wire.cpp"), "-o", exe]
    b = subprocess.run(build, capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.stdout + r.stderr)[-3000:]
    assert "encode request: 0" in r.stdout
    done = [x for x in r.stdout.splitlines() if x.startswith("fuzz done")]
    assert done, r.stdout
    accepted = int(done[0].split()[2])
    assert accepted > 0  # the unmutated records still decode

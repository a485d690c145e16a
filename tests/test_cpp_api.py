"""The C++ drop-in API (venusaur_amd/core: Renderer, Scene, Camera,
HIPOutputBuffer, Exception) driven through the vz_render CLI, which follows
the reference's Init -> Draw -> Cleanup sequence (Core/Core.cpp:239-438)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from util import book_lookat, box_lookat, oracle_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "venusaur_amd", "bin", "vz_render")


def test_core_headers_compile(tmp_path):
    """Every public header of the C++ API compiles on its own (host only)."""
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    core = os.path.join(ROOT, "venusaur_amd", "core")
    for h in sorted(os.listdir(core)):
        src = tmp_path / f"inc_{h}.cpp"
        src.write_text(f'#include "{os.path.join(core, h)}"\nint main() {{ return 0; }}\n')
        r = subprocess.run([hipcc, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                            "-Wno-unused-command-line-argument", "--offload-arch=gfx950", str(src)],
                           capture_output=True, text=True)
        assert r.returncode == 0, (h, r.stderr[-2000:])


def test_cli_without_gpu_fails_loudly():
    from venusaur_amd.hipmem import hip
    import ctypes as C
    n = C.c_int(0)
    try:
        rc = hip().hipGetDeviceCount(C.byref(n))
    except Exception:
        rc = 1
    if rc == 0 and n.value > 0:
        pytest.skip("a GPU is visible")
    if not os.path.exists(CLI):
        pytest.skip("vz_render not built")
    r = subprocess.run([CLI, "--width", "8", "--height", "8"], capture_output=True, text=True)
    assert r.returncode == 1 and "VZ_ERROR_NO_DEVICE" in r.stderr


def _run_cli(tmp_path, *args):
    out = tmp_path / "accum.raw"
    r = subprocess.run([CLI, *args, "--accum", str(out), "--ppm", str(tmp_path / "img.ppm"),
                        "--pfm", str(tmp_path / "img.pfm")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return out, r.stdout


@pytest.mark.gpu
def test_renderer_progressive_frames_match_oracle(tmp_path, book):
    """Three Draw calls with an unchanged camera: progressive average of
    frames 1..3, bit-exact against the oracle run frame by frame."""
    w, h, spp, depth = 96, 64, 4, 8
    out, _ = _run_cli(tmp_path, "--width", str(w), "--height", str(h), "--spp", str(spp), "--depth",
                      str(depth), "--frames", "3")
    got = np.fromfile(out, np.float32).reshape(h, w, 4)
    s, m = book
    cam = dict(book_lookat(w, h))
    cam["aspect"] = float(np.float32(w) / np.float32(h))
    acc = None
    b = O.BVH(s)
    for f in (1, 2, 3):
        acc, img, _ = O.render(s, m, oracle_params(w, h, spp, depth, cam, subframe=f), bvh=b, accum=acc)
    assert np.array_equal(got, acc)
    # PPM is top-down: its first row is the oracle's last row
    raw = (tmp_path / "img.ppm").read_bytes()
    hdr = f"P6\n{w} {h}\n255\n".encode()
    pix = np.frombuffer(raw[len(hdr):], np.uint8).reshape(h, w, 3)
    assert np.abs(pix[0].astype(int) - img[h - 1, :, :3].astype(int)).max() <= 1
    pfm = (tmp_path / "img.pfm").read_bytes()
    hdr = f"PF\n{w} {h}\n-1.0\n".encode()
    lin = np.frombuffer(pfm[len(hdr):], "<f4").reshape(h, w, 3)
    assert np.array_equal(lin, acc[..., :3])


@pytest.mark.gpu
def test_renderer_box_scene_fp64(tmp_path):
    w, h, spp, depth = 64, 40, 6, 16
    out, line = _run_cli(tmp_path, "--scene", "box", "--spheres", "5000", "--seed", "17", "--width", str(w),
                         "--height", str(h), "--spp", str(spp), "--depth", str(depth), "--fp64")
    got = np.fromfile(out, np.float64).reshape(h, w, 4)
    s, m = O.box_scene(5000, 17)
    cam = dict(box_lookat(w, h))
    cam["aspect"] = float(np.float32(w) / np.float32(h))
    acc, _, st = O.render(s, m, oracle_params(w, h, spp, depth, cam, accum_fp64=True), bvh=O.BVH(s))
    assert np.array_equal(got, acc)
    assert f'"segments": {st.segments}' in line


@pytest.mark.gpu
@pytest.mark.parametrize("buffer", ["zero_copy", "p2p"])
def test_output_buffer_types_give_the_same_frame(tmp_path, buffer):
    """HIPOutputBuffer's ZERO_COPY (mapped pinned host memory written by the
    GPU) and HIP_P2P types hold the same 8-bit frame as HIP_DEVICE."""
    args = ["--width", "64", "--height", "40", "--spp", "4", "--depth", "8", "--frames", "2"]
    d = tmp_path / "device"
    z = tmp_path / buffer
    d.mkdir()
    z.mkdir()
    _run_cli(d, *args)
    _run_cli(z, *args, "--buffer", buffer)
    assert (z / "img.ppm").read_bytes() == (d / "img.ppm").read_bytes()
    assert (z / "accum.raw").read_bytes() == (d / "accum.raw").read_bytes()


@pytest.mark.gpu
def test_camera_move_restarts_the_progressive_average(tmp_path, book):
    """Renderer::Draw restarts the average when the camera changed
    (Core/Renderer.h:37-54): frame 1 at the book camera, then the camera moves
    and frames 2-3 must equal frames 1-2 rendered afresh from the new position."""
    w, h, spp, depth = 64, 40, 4, 8
    out, _ = _run_cli(tmp_path, "--width", str(w), "--height", str(h), "--spp", str(spp), "--depth", str(depth),
                      "--frames", "3", "--move-after", "1", "--move-to", "10,3,6")
    got = np.fromfile(out, np.float32).reshape(h, w, 4)
    s, m = book
    cam = dict(book_lookat(w, h))
    cam["lookfrom"] = [10, 3, 6]
    acc = None
    b = O.BVH(s)
    for f in (1, 2):
        acc, _, _ = O.render(s, m, oracle_params(w, h, spp, depth, cam, subframe=f), bvh=b, accum=acc)
    assert np.array_equal(got, acc)

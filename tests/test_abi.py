"""The C-ABI boundary without a GPU: library load, exported symbols, struct
layout against a C compiler, host helpers against the oracle, error paths
that need no device."""
import ctypes as C
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle as O
from venusaur_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(api.LIB_PATH)
    names = api.declared_symbols()
    assert len(names) >= 26
    missing = [n for n in names if not hasattr(L, n)]
    assert missing == []
    api.lib()  # the typed binding declares every one of them too


def test_struct_layout_matches_c_compiler(tmp_path):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = tmp_path / "abi_layout"
    subprocess.run([cc, "-std=c11", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "abi_layout.c"),
                    "-o", str(exe)], check=True)
    lay = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    assert lay["vzSphere"] == api.SPHERE_DT.itemsize == 16
    assert lay["vzMaterial"] == api.MATERIAL_DT.itemsize == 24
    assert lay["vzRay"] == api.RAY_DT.itemsize == 32
    assert lay["vzHit"] == api.HIT_DT.itemsize == 8
    for name, cls in [("vzLaunchParams", api.LaunchParams), ("vzStats", api.Stats), ("vzAccelInfo", api.AccelInfo),
                      ("vzBatchItem", api.BatchItem), ("vzTuning", api.Tuning), ("vzServeConfig", api.ServeConfig),
                      ("vzRequestResult", api.RequestResult), ("vzRequestDesc", api.RequestDesc)]:
        assert lay[name] == C.sizeof(cls), name
        for key, off in lay.items():
            if key.startswith(name + "."):
                field = key.split(".", 1)[1]
                assert getattr(cls, field).offset == off, key
    for key, off in lay.items():
        if key.startswith("vzMaterial."):
            assert api.MATERIAL_DT.fields[key.split(".")[1]][1] == off
        if key.startswith("vzRay."):
            f = key.split(".")[1]
            f = {"dir": "dx"}.get(f, f)
            assert api.RAY_DT.fields[f][1] == off


def test_book_scene_generator_matches_oracle_and_fixture(book):
    s, m = api.book_scene()
    assert s.tobytes() == book[0].tobytes() and m.tobytes() == book[1].tobytes()


@pytest.mark.parametrize("n,seed", [(1, 1), (999, 2), (100000, 3)])
def test_box_scene_generator_matches_oracle(n, seed):
    s, m = api.box_scene(n, seed)
    so, mo = O.box_scene(n, seed)
    assert s.tobytes() == so.tobytes() and m.tobytes() == mo.tobytes()


@pytest.mark.parametrize("args", [([13, 2, 3], [0, 0, 0], 20.0, 0.1, 10.0, 1920, 1080),
                                  ([0, 15, -120], [0, 0, 0], 40.0, 0.0, 10.0, 1024, 768),
                                  ([-7.5, 40.1, 3.3], [1, -2, 0.5], 55.0, 0.3, 4.0, 300, 500)])
def test_camera_matches_oracle(args):
    frm, at, vfov, ap, focus, w, h = args
    p = api.launch_params(w, h, 1, 4, frm, at, vfov, ap, focus)
    cam = O.camera_lookat(frm, at, vfov, float(np.float32(w) / np.float32(h)), ap, focus)
    for k in ("origin", "u", "v", "w"):
        assert list(getattr(p, k)) == list(getattr(cam, k)), k
    assert p.lens_radius == cam.lens_radius


def test_status_strings_and_null_arguments():
    L = api.lib()
    assert L.vzStatusString(0) == b"VZ_SUCCESS"
    assert L.vzStatusString(5) == b"VZ_ERROR_CAPACITY"
    assert L.vzCreate(0, None) == 1
    assert b"NULL" in L.vzGetLastError()
    assert L.vzSceneCreate(None, None, None, 0, None, None) == 1
    assert L.vzLaunch(None, None, None, None) == 1
    assert L.vzDestroy(None) == 1
    assert L.vzGenerateBookScene(1, None, None, 0, None) == 1
    n = C.c_uint32()
    assert L.vzGenerateBookScene(5489, None, None, 0, C.byref(n)) == 5 and n.value == 486


def test_request_footprint_grows_with_size_and_fp64():
    a = api.launch_params(256, 256, 16, 16, [0, 0, 1], [0, 0, 0], 40.0, 0.0, 1.0)
    b = api.launch_params(512, 512, 16, 16, [0, 0, 1], [0, 0, 0], 40.0, 0.0, 1.0)
    c = api.launch_params(512, 512, 2048, 16, [0, 0, 1], [0, 0, 0], 40.0, 0.0, 1.0, flags=api.ACCUM_FP64)
    fa, fb, fc = api.footprint(a), api.footprint(b), api.footprint(c)
    assert 0 < fa < fb < fc
    # the pass ring, min(spp, max_depth + 2) slots of RGB floats, dominates
    assert fb >= 16 * 3 * 4 * 512 * 512
    assert fc >= 18 * 3 * 4 * 512 * 512 + 3 * 8 * 512 * 512


def test_no_device_is_reported_not_faked():
    """Without a visible gfx950 the library refuses to create a context
    (there is no CPU fallback behind the ABI)."""
    from venusaur_amd.hipmem import hip
    count = C.c_int(0)
    try:
        rc = hip().hipGetDeviceCount(C.byref(count))
    except Exception:
        rc, count.value = 1, 0
    if rc == 0 and count.value > 0:
        pytest.skip("a GPU is visible")
    with pytest.raises(api.VzError) as e:
        api.Context(0)
    assert e.value.status == 8

"""The serving front end (venusaur_amd/bin/vz_serve) over a stream of
wire-format request records: every request's accumulator is bit-identical to
a plain launch of the decoded request, scenes stay resident under their key
(an inline copy of a resident scene is ignored), and a request that names a
scene neither resident nor inline fails alone."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from util import GpuTarget
from venusaur_amd import api

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVE = os.path.join(ROOT, "venusaur_amd", "bin", "vz_serve")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("realtime", [False, True])
def test_vz_serve_stream_matches_plain_launches(ctx, tmp_path, realtime):
    assert os.path.exists(SERVE), "venusaur_amd/bin/vz_serve is missing: build it with __graft_entry__.build()"
    book = api.book_scene()
    box = api.box_scene(3000, 77)
    arrays = {1: book, 2: box}
    # (key, inline scene, w, h, spp, depth, flags, lookfrom, lookat, vfov, aperture, subframe)
    spec = [
        (1, book, 64, 48, 8, 50, 0, [13, 2, 3], [0, 0, 0], 20.0, 0.1, 1),
        (2, box, 96, 64, 4, 16, 0, [0, 15, -120], [0, 0, 0], 40.0, 0.0, 1),
        (1, None, 40, 30, 16, 16, 0, [5, 3, 8], [0, 1, 0], 35.0, 0.0, 1),           # camera reuse of key 1
        (2, box, 32, 32, 1100, 8, api.ACCUM_FP64, [30, 20, -60], [0, 0, 0], 50.0, 0.0, 1),  # inline copy ignored
        (99, None, 16, 16, 2, 4, 0, [0, 1, -5], [0, 0, 0], 40.0, 0.0, 1),          # neither resident nor inline
        (2, None, 48, 40, 4, 16, 0, [0, 15, -120], [0, 0, 0], 40.0, 0.0, 2),       # subframe 2 on a cleared accum
    ]
    descs, stream = [], b""
    for tag, (key, inl, w, h, spp, depth, flags, frm, at, vfov, ap, sub) in enumerate(spec):
        d = api.request_desc(tag, key, w, h, spp, depth, frm, at, vfov, ap, 10.0, flags=flags, subframe=sub,
                             arrival_ms=float(tag))
        descs.append(d)
        stream += api.encode_request(d, *(inl if inl is not None else (None, None)))
    path = tmp_path / "stream.vzr"
    path.write_bytes(stream)
    out_dir = tmp_path / "accum"
    out_dir.mkdir()
    say(f"case: vz_serve over 6 wire-format requests (book + 3000-sphere box scene), realtime={realtime}")
    cmd = [SERVE, "--in", str(path), "--accum-dir", str(out_dir), "--budget-mb", "4096"]
    run = subprocess.run(cmd + (["--realtime"] if realtime else []), capture_output=True, text=True, timeout=120)
    say(run.stderr[-2000:])
    assert run.returncode == 1, run.stdout + run.stderr  # one request fails, by design
    lines = [json.loads(x) for x in run.stdout.strip().splitlines()]
    per, summary = lines[:-1], lines[-1]
    assert [x["tag"] for x in per] == list(range(6))
    assert per[4]["status"] == "VZ_ERROR_INVALID_VALUE" and "not resident" in per[4]["error"]
    assert (summary["requests"], summary["ok"], summary["scenes_resident"]) == (6, 5, 2)
    assert (summary["inline_scenes_built"], summary["inline_scenes_ignored"]) == (2, 1)
    assert summary["p50_ms"] > 0 and summary["p99_ms"] >= summary["p50_ms"]
    assert summary["realtime"] is realtime
    if realtime:  # each request is submitted no earlier than its arrival offset
        assert all(x["submit_ms"] >= x["arrival_ms"] for x in per if x["status"] == "VZ_SUCCESS")
    scenes = {k: api.Scene(ctx, s, m) for k, (s, m) in arrays.items()}
    try:
        for tag, d in enumerate(descs):
            if tag == 4:
                continue
            say(f"case: request {tag} against a plain launch")
            assert per[tag]["status"] == "VZ_SUCCESS", per[tag]
            fp64 = bool(d.flags & api.ACCUM_FP64)
            t = GpuTarget(d.width, d.height, fp64=fp64)
            p = api.request_launch_params(d, t.image.value, t.accum.value)
            st = scenes[d.scene_key].launch(p)
            got = np.fromfile(str(out_dir / f"{tag}.accum"), np.float64 if fp64 else np.float32)
            assert got.tobytes() == t.accum_host().tobytes(), tag
            assert per[tag]["segments"] == st.segments
    finally:
        for s in scenes.values():
            s.close()

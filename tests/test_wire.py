"""Wire format of the serving form (scene and request records of
include/vz_api.h), without a GPU: the library's encoder and decoder against
an independent decoder written from the documented layout, the published
FNV-1a 64 test vectors, a committed fixture, round trips, and every way a
record can be malformed."""
import os
import struct

import numpy as np
import pytest

from venusaur_amd import api, workloads

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def fnv1a64(b):
    """FNV-1a 64 (Fowler, Noll, Vo): offset basis 0xcbf29ce484222325, prime 0x100000001b3."""
    h = 0xCBF29CE484222325
    for x in b:
        h = ((h ^ x) * 0x100000001B3) & 0xFFFFFFFFFFFFFFFF
    return h


def test_fnv1a64_published_vectors():
    # test vectors of the FNV reference (isthe.com/chongo/tech/comp/fnv)
    assert fnv1a64(b"") == 0xCBF29CE484222325
    assert fnv1a64(b"a") == 0xAF63DC4C8601EC8C
    assert fnv1a64(b"foobar") == 0x85944171F73967E8


def parse_scene(b):
    """Independent decoder of a scene record (layout of include/vz_api.h)."""
    magic, version, n, reserved, checksum = struct.unpack_from("<4sIIIQ", b, 0)
    assert (magic, version, reserved) == (b"VZSC", 1, 0)
    payload = b[24:24 + 40 * n]
    assert len(payload) == 40 * n and checksum == fnv1a64(payload)
    sph = np.frombuffer(payload[:16 * n], api.SPHERE_DT)
    mat = np.frombuffer(payload[16 * n:], api.MATERIAL_DT)
    return sph, mat, 24 + 40 * n


def parse_request(b):
    f = struct.unpack_from("<4sIQQIIIIII3f3fffffII", b, 0)
    assert f[0] == b"VZRQ" and f[1] == 1 and f[-1] == 0
    keys = ["tag", "scene_key", "width", "height", "spp", "depth", "flags", "subframe"]
    out = dict(zip(keys, f[2:10]))
    out["lookfrom"], out["lookat"] = list(f[10:13]), list(f[13:16])
    out["vfov"], out["aperture"], out["focus"], out["arrival_ms"], out["scene_bytes"] = f[16:21]
    return out


def test_scene_record_layout(book):
    s, m = book
    rec = api.encode_scene(s, m)
    assert len(rec) == api.lib().vzWireSceneSize(len(s)) == 24 + 40 * len(s)
    sph, mat, used = parse_scene(rec)
    assert used == len(rec)
    assert sph.tobytes() == s.tobytes() and mat.tobytes() == m.tobytes()
    s2, m2, used2 = api.decode_scene(rec + b"trailing bytes are not part of the record")
    assert used2 == len(rec) and s2.tobytes() == s.tobytes() and m2.tobytes() == m.tobytes()


def test_request_fixture_matches_documented_layout():
    import importlib.util
    spec = importlib.util.spec_from_file_location("make_wire", os.path.join(GOLDEN, "make_wire.py"))
    mw = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mw)
    fixture = open(os.path.join(GOLDEN, "wire_request_small.vzrq"), "rb").read()
    assert len(fixture) == 96 + 24 + 3 * 40
    f = parse_request(fixture)
    assert (f["tag"], f["scene_key"], f["width"], f["height"], f["spp"], f["depth"]) == (
        0x0123456789ABCDEF, 42, 320, 180, 16, 12)
    assert (f["flags"], f["subframe"], f["scene_bytes"]) == (api.ACCUM_FP64, 3, 144)
    assert f["lookfrom"] == [13.0, 2.0, 3.0] and f["lookat"] == [0.0, 0.5, 0.0]
    assert (f["vfov"], f["aperture"], f["focus"], f["arrival_ms"]) == (20.0, np.float32(0.1), 10.0, 1234.5)
    sph, mat, _ = parse_scene(fixture[96:])
    assert sph.tobytes() == mw.SPHERES.tobytes() and mat.tobytes() == mw.MATERIALS.tobytes()
    # the library decodes the fixture to the same request, and encodes it back to the same bytes
    d, scene, used = api.decode_request(fixture)
    assert used == len(fixture) and scene == fixture[96:]
    assert (d.tag, d.scene_key, d.samples_per_pixel, d.subframe_index) == (0x0123456789ABCDEF, 42, 16, 3)
    assert api.encode_request(mw.desc(), mw.SPHERES, mw.MATERIALS) == fixture


def test_request_without_scene_and_launch_params():
    d = api.request_desc(5, 9, 64, 48, 4, 8, [13, 2, 3], [0, 0, 0], 20.0, 0.1, 10.0)
    rec = api.encode_request(d)
    assert len(rec) == 96 and parse_request(rec)["scene_bytes"] == 0
    d2, scene, used = api.decode_request(rec)
    assert scene == b"" and used == 96 and bytes(d2) == bytes(d)
    # launch params: the camera of vzCameraLookAt with aspect = width / height in float
    p = api.request_launch_params(d2, None, 4096)
    q = api.launch_params(64, 48, 4, 8, [13, 2, 3], [0, 0, 0], 20.0, 0.1, 10.0)
    q.accum = 4096
    assert bytes(p) == bytes(q)


def test_stream_round_trip(tmp_path):
    reqs = workloads.request_stream(16, 50.0, seed=3)
    for r in reqs:  # keep the scenes small: this test encodes every one of them
        r.n_small = min(r.n_small, 2000)
    path = tmp_path / "stream.vzr"
    size = workloads.write_stream(str(path), reqs)
    data = path.read_bytes()
    assert len(data) == size
    off, seen = 0, set()
    for i, r in enumerate(reqs):
        d, scene, used = api.decode_request(data[off:])
        assert d.tag == i and d.scene_key == r.scene_key
        assert (d.width, d.height, d.samples_per_pixel, d.max_depth) == (r.width, r.height, r.spp, r.depth)
        assert d.flags == (api.ACCUM_FP64 if r.fp64 else 0)
        assert d.arrival_ms == np.float32(r.arrival_s * 1e3)
        assert list(d.lookfrom) == [float(np.float32(x)) for x in r.lookfrom]
        if r.scene_key in seen:
            assert scene == b""
        else:
            s, m = r.scene_arrays()
            s2, m2, _ = api.decode_scene(scene)
            assert s2.tobytes() == s.tobytes() and m2.tobytes() == m.tobytes()
            seen.add(r.scene_key)
        off += used
    assert off == len(data)


def _raises(fn, *a):
    with pytest.raises(api.VzError) as e:
        fn(*a)
    assert e.value.status == 1  # VZ_ERROR_INVALID_VALUE
    return str(e.value)


def _small_scene():
    s = np.array([(0, -1000, 0, 1000), (0, 1, 0, 1), (4, 1, 0, 0.5)], api.SPHERE_DT)
    m = np.array([(0, (0.5, 0.5, 0.5), 0, 0), (2, (0, 0, 0), 0, 1.5), (1, (0.7, 0.6, 0.5), 0, 0)], api.MATERIAL_DT)
    return s, m


def test_every_truncation_is_rejected():
    s, m = _small_scene()
    d = api.request_desc(1, 2, 32, 32, 2, 4, [0, 1, -5], [0, 0, 0], 40.0)
    rec = api.encode_request(d, s, m)
    for k in range(len(rec)):
        _raises(api.decode_request, rec[:k])
    scene = api.encode_scene(s, m)
    for k in range(len(scene)):
        _raises(api.decode_scene, scene[:k])


def _patch(b, off, fmt, v):
    b = bytearray(b)
    struct.pack_into(fmt, b, off, v)
    return bytes(b)


def test_malformed_scene_records():
    s, m = _small_scene()
    rec = api.encode_scene(s, m)
    assert "not a scene" in _raises(api.decode_scene, b"VZSX" + rec[4:])
    assert "version" in _raises(api.decode_scene, _patch(rec, 4, "<I", 2))
    assert "count" in _raises(api.decode_scene, _patch(rec, 8, "<I", 0))
    assert "truncated" in _raises(api.decode_scene, _patch(rec, 8, "<I", 4))
    assert "reserved" in _raises(api.decode_scene, _patch(rec, 12, "<I", 1))
    flipped = bytearray(rec)
    flipped[24 + 5] ^= 0x01
    assert "checksum" in _raises(api.decode_scene, bytes(flipped))

    def with_payload(s2, m2):  # a record with a valid checksum over invalid values
        payload = s2.tobytes() + m2.tobytes()
        return struct.pack("<4sIIIQ", b"VZSC", 1, len(s2), 0, fnv1a64(payload)) + payload

    assert with_payload(s, m) == rec
    bad = s.copy()
    bad["cy"][1] = np.nan
    assert "sphere 1" in _raises(api.decode_scene, with_payload(bad, m))
    bad = s.copy()
    bad["r"][2] = 0.0
    _raises(api.decode_scene, with_payload(bad, m))
    badm = m.copy()
    badm["type"][0] = 3
    assert "unknown type" in _raises(api.decode_scene, with_payload(s, badm))
    badm = m.copy()
    badm["albedo"][1][2] = np.inf
    assert "non-finite" in _raises(api.decode_scene, with_payload(s, badm))
    # the encoder refuses to write such records
    _raises(api.encode_scene, bad, m)
    _raises(api.encode_scene, s, badm)


@pytest.mark.parametrize("field,value,msg", [
    ("width", 1, "width"), ("height", 1, "width"), ("samples_per_pixel", 0, "samples"),
    ("max_depth", 0, "max_depth"), ("max_depth", 256, "max_depth"), ("flags", api.MEGAKERNEL, "flags"),
    ("subframe_index", 0, "subframe"), ("vfov_deg", 0.0, "vfov"), ("vfov_deg", 180.0, "vfov"),
    ("focus_dist", 0.0, "focus"), ("aperture", -1.0, "aperture"), ("arrival_ms", -1.0, "arrival"),
    ("arrival_ms", float("inf"), "arrival"), ("samples_per_pixel", 2048, "FP64"),
])
def test_malformed_requests(field, value, msg):
    d = api.request_desc(1, 2, 32, 32, 2, 4, [0, 1, -5], [0, 0, 0], 40.0)
    good = api.encode_request(d)
    setattr(d, field, value)
    assert msg in _raises(api.encode_request, d)
    offsets = {"width": (24, "<I"), "height": (28, "<I"), "samples_per_pixel": (32, "<I"), "max_depth": (36, "<I"),
               "flags": (40, "<I"), "subframe_index": (44, "<I"), "vfov_deg": (72, "<f"), "aperture": (76, "<f"),
               "focus_dist": (80, "<f"), "arrival_ms": (84, "<f")}
    off, fmt = offsets[field]
    assert msg in _raises(api.decode_request, _patch(good, off, fmt, value))


def test_request_camera_and_scene_size_checks():
    d = api.request_desc(1, 2, 32, 32, 2, 4, [1, 2, 3], [1, 2, 3], 40.0)
    assert "lookfrom equals lookat" in _raises(api.encode_request, d)
    d = api.request_desc(1, 2, 32, 32, 2, 4, [0, 1, -5], [0, 0, 0], 40.0)
    s, m = _small_scene()
    rec = api.encode_request(d, s, m)
    assert "not a request" in _raises(api.decode_request, b"VZRX" + rec[4:])
    assert "reserved" in _raises(api.decode_request, _patch(rec, 92, "<I", 7))
    # inline size past the end of the record; smaller than its scene record;
    # larger than its scene record (bytes after it counted in)
    assert "truncated" in _raises(api.decode_request, _patch(rec, 88, "<I", len(rec) - 96 + 40))
    assert "truncated" in _raises(api.decode_request, _patch(rec, 88, "<I", len(rec) - 96 - 40))
    assert "does not match" in _raises(api.decode_request, _patch(rec, 88, "<I", len(rec) - 96 + 40) + b"\0" * 40)
    # a request with 2048 spp is valid with FP64 accumulation
    d.samples_per_pixel, d.flags = 2048, api.ACCUM_FP64
    assert api.decode_request(api.encode_request(d))[0].samples_per_pixel == 2048


SERVE = os.path.join(os.path.dirname(HERE), "venusaur_amd", "bin", "vz_serve")


def _serve(*args):
    import subprocess
    assert os.path.exists(SERVE), "venusaur_amd/bin/vz_serve is missing: build it with __graft_entry__.build()"
    return subprocess.run([SERVE, *args], capture_output=True, text=True, timeout=60)


def test_vz_serve_usage_errors(tmp_path):
    assert _serve().returncode == 2
    assert _serve("--in", str(tmp_path / "missing.vzr")).returncode == 2
    assert _serve("--in", "x", "--bogus").returncode == 2


def test_vz_serve_rejects_a_malformed_stream_before_any_gpu_work(tmp_path):
    """Each record is decoded before the device is touched: these fail with the
    record's own message, not a device error, with or without a GPU."""
    s, m = _small_scene()
    good = api.encode_request(api.request_desc(0, 1, 32, 32, 2, 4, [0, 1, -5], [0, 0, 0], 40.0), s, m)
    second = api.encode_request(api.request_desc(1, 1, 32, 32, 2, 4, [0, 2, -5], [0, 0, 0], 40.0))
    flipped = bytearray(good)
    flipped[96 + 30] ^= 0x01  # inside the inline scene's payload
    cases = {
        "empty": (b"", "no request record"),
        "truncated header": (good + second[:50], "record 1: truncated header"),
        "inline past the end": (good[:-8], "record 0: inline scene runs past the end"),
        "checksum": (bytes(flipped) + second, "record 0: scene record checksum mismatch"),
        "bad value": (good + _patch(second, 32, "<I", 0), "record 1: samples_per_pixel"),
    }
    for name, (data, msg) in cases.items():
        path = tmp_path / (name.replace(" ", "_") + ".vzr")
        path.write_bytes(data)
        r = _serve("--in", str(path))
        assert r.returncode == 1 and msg in r.stderr, (name, r.stderr)
        assert "vzCreate" not in r.stderr, name


def _fuzz_base():
    s, m = _small_scene()
    d = api.request_desc(3, 4, 32, 24, 2, 4, [0, 1, -5], [0, 0, 0], 40.0, arrival_ms=2.5)
    return api.encode_request(d, s, m)


def test_fuzzed_records_decode_or_raise():
    """Random byte edits, truncations and extensions of a valid request record
    (and pure garbage): the decoders either return a record that passes the
    same checks again on re-encoding, or raise VZ_ERROR_INVALID_VALUE; they
    never crash or read past the buffer they are given (the buffer ctypes
    hands over is exactly len(data) bytes)."""
    from hypothesis import given, settings, HealthCheck
    from hypothesis import strategies as st
    base = _fuzz_base()

    @settings(max_examples=1500, deadline=None, suppress_health_check=list(HealthCheck), derandomize=True)
    @given(st.lists(st.tuples(st.integers(0, len(base) - 1), st.integers(0, 255)), max_size=6),
           st.integers(0, len(base) + 16), st.binary(max_size=16))
    def check(edits, cut, tail):
        b = bytearray(base)
        for pos, val in edits:
            b[pos] = val
        data = bytes(b[:cut]) + (tail if cut >= len(base) else b"")
        try:
            d, scene, used = api.decode_request(data)
        except api.VzError as e:
            assert e.status == 1
            return
        assert used <= len(data) and used == 96 + len(scene)
        if scene:
            s, m, n = api.decode_scene(scene)
            assert n == len(scene)
            assert api.encode_request(d, s, m) == data[:used]
        else:
            assert api.encode_request(d) == data[:used]

    check()

    @settings(max_examples=500, deadline=None, suppress_health_check=list(HealthCheck), derandomize=True)
    @given(st.binary(max_size=400))
    def garbage(data):
        for fn in (api.decode_request, api.decode_scene):
            try:
                fn(data)
            except api.VzError as e:
                assert e.status == 1

    garbage()

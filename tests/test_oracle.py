"""Pins the CPU oracle (oracle/) before it is trusted as the parity target.

Each check is independent of the formula it checks: known-answer values
computed from the reference's own random.cuh (tests/golden/rng_kat.json and,
when built, oracle/_ref), the C++ standard's mt19937 check value, closed
forms and invariants of the geometry and the scattering laws, brute force
against the BVH, and the FP64 integrator against the FP32 one.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KAT = json.load(open(os.path.join(GOLDEN, "rng_kat.json")))
FIX = json.load(open(os.path.join(GOLDEN, "book_fixture.json")))


# ---- RNG (random.cuh:31-67) -------------------------------------------------
def test_rng_known_answers():
    for row in KAT["rows"]:
        t = O.tea4(row["v0"], row["v1"])
        assert t == row["tea4"]
        r, st = O.rnd_seq(t, 3)
        assert r == row["rnd"] and st == row["state"]
    s, st = O.rnd_seq(KAT["stream_seed"], 64)
    assert s == KAT["stream"] and st == KAT["stream_state"]


def test_rng_matches_reference_header():
    ref = O.ref_rng()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree not mounted)")
    import ctypes as C
    rng = np.random.default_rng(1)
    for v0, v1 in rng.integers(0, 2**32, size=(2000, 2), dtype=np.uint64):
        v0, v1 = int(v0), int(v1)
        assert ref.ref_tea4(v0, v1) == O.tea4(v0, v1)
        a, b = C.c_uint32(v0), C.c_uint32(v0)
        for _ in range(8):
            assert ref.ref_rnd(C.byref(a)) == O.lib().vzo_rnd(C.byref(b))
        assert a.value == b.value


def test_rnd_range_and_lcg_closed_form():
    # rnd = (state' & 0xFFFFFF) / 2^24 with state' = 1664525*state + 1013904223
    for s0 in (0, 1, 12345, 2**31 - 1, 2**32 - 1):
        r, st = O.rnd_seq(s0, 1)
        nxt = (1664525 * s0 + 1013904223) % 2**32
        assert st == nxt and r[0] == (nxt & 0xFFFFFF) / 2**24 and 0.0 <= r[0] < 1.0


def test_mt19937_standard_check_value():
    # [rand.predef]: the 10000th output of a default-constructed mt19937
    assert O.lib().vzo_mt19937_nth(5489, 10000) == 4123659995
    assert KAT["mt19937_5489_10000th"] == 4123659995


# ---- scenes (Scene.h:13-112) ---------------------------------------------------
def test_book_scene_fixture_and_structure(book):
    s, m = O.book_scene()
    fs, fm = book
    assert s.tobytes() == fs.tobytes() and m.tobytes() == fm.tobytes()
    assert len(s) == 486
    assert tuple(s[0]) == (0.0, -1000.0, 0.0, 1000.0) and m[0]["type"] == 0
    small = s[1:-3]
    assert np.all(small["cy"] == np.float32(0.2)) and np.all(small["r"] == np.float32(0.2))
    d = np.sqrt((small["cx"] - 4.0) ** 2 + (small["cz"]) ** 2)
    assert np.all(d > 0.9)
    assert [tuple(x) for x in s[-3:]] == [(0, 1, 0, 1), (-4, 1, 0, 1), (4, 1, 0, 1)]
    assert list(m[-3:]["type"]) == [2, 0, 1]
    metal = m[m["type"] == 1][:-1]
    assert np.all(metal["albedo"] >= 0.5) and np.all(metal["albedo"] < 1.0)
    assert np.all(metal["fuzz"] >= 0.0) and np.all(metal["fuzz"] < 0.5)
    lam = m[1:-3][m[1:-3]["type"] == 0]
    assert np.all(lam["albedo"] >= 0.0) and np.all(lam["albedo"] < 1.0)
    assert np.all(m[m["type"] == 2]["ir"] == np.float32(1.5))
    # material mix 0.8 / 0.15 / 0.05 over the 482 random spheres
    frac = np.bincount(m[1:-3]["type"], minlength=3) / len(m[1:-3])
    assert abs(frac[0] - 0.8) < 0.06 and abs(frac[1] - 0.15) < 0.05 and abs(frac[2] - 0.05) < 0.03


def test_box_scene_structure():
    s, m = O.box_scene(20000, 3)
    assert tuple(s[0]) == (0.0, -1000.0, 0.0, 1000.0)
    b = s[1:]
    assert b["cx"].min() >= -100 and b["cx"].max() < 100
    assert b["cy"].min() >= 0 and b["cy"].max() < 20
    assert b["r"].min() >= 0.0099 and b["r"].max() <= 0.5001
    lr = np.log(b["r"].astype(np.float64))
    assert abs(lr.mean() - (math.log(0.01) + math.log(0.5)) / 2) < 0.05  # log-uniform
    s2, m2 = O.box_scene(20000, 3)
    assert s.tobytes() == s2.tobytes() and m.tobytes() == m2.tobytes()


# ---- camera (Camera.cpp:24-37) ----------------------------------------------------
def test_camera_frame_closed_form():
    cam = O.camera_lookat([13, 2, 3], [0, 0, 0], 20.0, 1.5, 0.1, 10.0)
    U, V, W = (np.array(x, np.float64) for x in (cam.u, cam.v, cam.w))
    f = -np.array([13, 2, 3], np.float64)
    assert np.allclose(W / np.linalg.norm(W), f / np.linalg.norm(f), atol=1e-6)
    assert abs(np.linalg.norm(W) - 10.0) < 1e-5
    assert abs(U @ W) < 1e-4 and abs(V @ W) < 1e-4 and abs(U @ V) < 1e-4
    assert abs(U[1]) < 1e-7  # U is horizontal (W x world-up)
    h = math.tan(math.radians(20.0) / 2)
    assert abs(np.linalg.norm(V) - 10 * 2 * h) < 1e-4
    assert abs(np.linalg.norm(U) - 10 * 1.5 * 2 * h) < 1e-4
    assert cam.lens_radius == np.float32(0.05)


def test_camera_rays_land_in_their_pixel():
    # aperture 0: origin is the eye; the direction crosses the image plane at
    # W + (u/2)U + (v/2)V with u in [2x/(W-1)-1, 2(x+1)/(W-1)-1)
    w, h = 16, 9
    cam = O.camera_lookat([0, 1, 5], [0, 0, 0], 40.0, w / h, 0.0, 3.0)
    s = np.zeros(1, O.SPHERE_DT)
    s[0] = (0.0, -1e4, 0.0, 1.0)  # far below: every camera ray misses
    m = np.zeros(1, O.MATERIAL_DT)
    p = O.make_params(w, h, 1, 1, cam)
    rays = O.make_rayset(s, m, p, w * h, 0, 0)
    U, V, W = (np.array(x, np.float64) for x in (cam.u, cam.v, cam.w))
    M = np.stack([U / 2, V / 2, -np.array([r for r in [0, 0, 0]], np.float64) + 0], 1)
    for k, r in enumerate(rays):
        x, y = k % w, k // w
        assert (r["ox"], r["oy"], r["oz"]) == (0.0, 1.0, 5.0)
        d = np.array([r["dx"], r["dy"], r["dz"]], np.float64)
        # solve d = W + a*U/2 + b*V/2
        M = np.stack([U / 2, V / 2], 1)
        ab, *_ = np.linalg.lstsq(M, d - W, rcond=None)
        assert np.allclose(M @ ab + W, d, atol=1e-5)
        assert 2 * x / (w - 1) - 1 - 1e-5 <= ab[0] < 2 * (x + 1) / (w - 1) - 1 + 1e-5
        assert 2 * y / (h - 1) - 1 - 1e-5 <= ab[1] < 2 * (y + 1) / (h - 1) - 1 + 1e-5


# ---- sphere test (RayTracer.cu:229-270) ---------------------------------------------
def _hit32(o, d, s, tmin=0.001, tmax=1e16):
    import ctypes as C
    t = C.c_float()
    sp = np.zeros(1, O.SPHERE_DT)
    sp[0] = s
    ok = O.lib().vzo_hit_sphere_f32(np.asarray(o, np.float32).ctypes.data_as(C.c_void_p),
                                    np.asarray(d, np.float32).ctypes.data_as(C.c_void_p), tmin, tmax,
                                    sp.ctypes.data_as(C.c_void_p), C.byref(t))
    return (t.value if ok else None)


def test_sphere_axis_closed_forms():
    # ray along +x towards a sphere at the origin: t = |ox| - r (near root)
    for ox, r in ((-10.0, 1.0), (-3.0, 0.5), (-1000.0, 2.0)):
        assert _hit32([ox, 0, 0], [1, 0, 0], (0, 0, 0, r)) == pytest.approx(-ox - r, rel=1e-6)
    # from inside: the far root
    assert _hit32([0, 0, 0], [0, 1, 0], (0, 0, 0, 2.0)) == pytest.approx(2.0, rel=1e-7)
    # unnormalised direction scales t
    assert _hit32([-10, 0, 0], [2, 0, 0], (0, 0, 0, 1.0)) == pytest.approx(4.5, rel=1e-6)
    # miss, and a hit behind the origin rejected by tmin
    assert _hit32([-10, 5, 0], [1, 0, 0], (0, 0, 0, 1.0)) is None
    assert _hit32([10, 0, 0], [1, 0, 0], (0, 0, 0, 1.0)) is None
    # tmax bound
    assert _hit32([-10, 0, 0], [1, 0, 0], (0, 0, 0, 1.0), tmax=8.0) is None


def test_ground_sphere_against_fp64():
    # the r=1000 ground sphere: rays from just above the ground keep their
    # hits (compensated discriminant) and agree with FP64 to float precision
    rng = np.random.default_rng(7)
    s = (0.0, -1000.0, 0.0, 1000.0)
    import ctypes as C
    bad = 0
    for _ in range(3000):
        o = np.array([rng.uniform(-12, 12), rng.uniform(0.0, 2.0), rng.uniform(-12, 12)], np.float32)
        d = rng.normal(size=3).astype(np.float32)
        d[1] = -abs(d[1]) * rng.uniform(0.001, 1.0)
        t32 = _hit32(o, d, s)
        t64 = C.c_double()
        sp = np.zeros(1, O.SPHERE_DT)
        sp[0] = s
        ok64 = O.lib().vzo_hit_sphere_f64(o.astype(np.float64).ctypes.data_as(C.c_void_p),
                                          d.astype(np.float64).ctypes.data_as(C.c_void_p), 0.001, 1e16,
                                          sp.ctypes.data_as(C.c_void_p), C.byref(t64))
        if ok64:
            if t32 is None or abs(t32 - t64.value) > 1e-3 * max(1.0, t64.value):
                bad += 1
        elif t32 is not None:
            bad += 1
    assert bad == 0


# ---- scattering (RayTracer.cu:272-440, vec_math.h:558-570) -----------------------------
def _mat(t, albedo=(0.5, 0.5, 0.5), fuzz=0.0, ir=1.5):
    m = np.zeros(1, O.MATERIAL_DT)[0]
    m["type"], m["albedo"], m["fuzz"], m["ir"] = t, albedo, fuzz, ir
    return m


def test_metal_mirror_reflection_and_absorption():
    rng = np.random.default_rng(2)
    n = np.array([0, 1, 0], np.float32)
    for _ in range(200):
        d = rng.normal(size=3).astype(np.float32)
        d[1] = -abs(d[1]) - 0.05
        ok, nd, _ = O.scatter(_mat(1, fuzz=0.0), d, n, True, 9)
        u = d / np.linalg.norm(d)
        assert ok and np.allclose(nd, [u[0], -u[1], u[2]], atol=1e-6)  # |reflect| preserved, angle mirrored
    # grazing fuzzy metal gets absorbed sometimes, never continues below the surface
    absorbed = 0
    for seed in range(300):
        ok, nd, _ = O.scatter(_mat(1, fuzz=0.5), np.array([1, -0.01, 0], np.float32), n, True, seed)
        absorbed += not ok
        if ok:
            assert float(nd @ n) > 0
    assert absorbed > 0


def test_lambertian_hemisphere():
    n = np.array([0, 0, 1], np.float32)
    for seed in range(500):
        ok, nd, _ = O.scatter(_mat(0), np.array([0, 0, -1], np.float32), n, True, seed)
        assert ok and float(nd @ n) >= 0.0 and np.linalg.norm(nd) <= 2.0 + 1e-6


def test_dielectric_snell_tir_and_schlick():
    n = np.array([0, 1, 0], np.float32)
    rng = np.random.default_rng(4)
    seen_refract = seen_reflect = 0
    for k in range(400):
        th = rng.uniform(0.0, 1.5)
        d = np.array([math.sin(th), -math.cos(th), 0.0], np.float32)
        front = bool(k % 2)
        eta = 1 / 1.5 if front else 1.5
        ok, nd, _ = O.scatter(_mat(2), d, n, front, k)
        assert ok
        u = nd / np.linalg.norm(nd)
        sin_in = math.sin(th)
        if eta * sin_in > 1.0:  # total internal reflection
            assert u[1] > 0 and abs(u[0] - sin_in) < 1e-5
            seen_reflect += 1
        elif u[1] > 0:  # Schlick chose reflection
            assert abs(u[0] - sin_in) < 1e-5
            seen_reflect += 1
        else:  # refraction: sin(out) = eta sin(in)
            assert abs(u[0] - eta * sin_in) < 1e-5
            seen_refract += 1
    assert seen_refract > 100 and seen_reflect > 20
    # normal incidence: reflect with probability r0 = ((1-1.5)/(1+1.5))^2 = 0.04
    refl = sum(O.scatter(_mat(2), np.array([0, -1, 0], np.float32), n, True, s)[1][1] > 0 for s in range(4000))
    assert abs(refl / 4000 - 0.04) < 0.012


def test_sky_known_values():
    assert np.allclose(O.sky([0, 1, 0]), [0.5, 0.7, 1.0])
    assert np.allclose(O.sky([0, -3, 0]), [1.0, 1.0, 1.0])
    assert np.allclose(O.sky([5, 0, 0]), [0.75, 0.85, 1.0])


def test_encode8_known_values():
    assert O.encode8(0.0) == 0 and O.encode8(1.0) == 255 and O.encode8(2.0) == 255 and O.encode8(-1.0) == 0
    assert O.encode8(0.5) == int((1.055 * 0.5 ** (1 / 2.4) - 0.055) * 256)
    vals = [O.encode8(x) for x in np.linspace(0, 1, 200)]
    assert all(a <= b for a, b in zip(vals, vals[1:]))


# ---- LBVH ----------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 5, 64, 500, 4097, 30000])
def test_bvh_invariants_and_brute_force(n):
    s, m = O.box_scene(n - 1, 100 + n) if n > 1 else O.box_scene(0, 1)
    b = O.BVH(s)
    assert b.validate() == 0
    cam = O.camera_lookat([0, 15, -120], [0, 0, 0], 40.0, 2.0, 0.0, 10.0)
    p = O.make_params(64, 32, 1, 8, cam)
    rays = O.make_rayset(s, m, p, 1500, 1500, n)
    hits, nv, ns = b.trace(rays)
    assert hits.tobytes() == O.trace_brute(s, rays).tobytes()
    if n > 1:
        assert nv >= len(rays)


def test_bvh_duplicates_and_degenerate_extent():
    s, m = O.box_scene(700, 9)
    s["cx"][100:400] = 3.0
    s["cy"][100:400] = 4.0
    s["cz"][100:400] = 5.0
    b = O.BVH(s)
    assert b.validate() == 0
    rays = O.make_rayset(s, m, O.make_params(32, 16, 1, 4, O.camera_lookat([0, 15, -120], [0, 0, 0], 40.0,
                                                                             2.0, 0.0, 10.0)), 500, 2000, 3)
    assert b.trace(rays)[0].tobytes() == O.trace_brute(s, rays).tobytes()
    flat, fm = O.box_scene(50, 1)
    flat["cy"] = 1.0  # zero centroid extent along y
    assert O.BVH(flat).validate() == 0


def _geometric_scene(n):
    s, m = O.box_scene(n - 1, 3)
    x = np.float32(2.0) ** -np.minimum(np.arange(n), 125).astype(np.float32)
    s["cx"], s["cy"], s["cz"] = x, np.float32(0.5), np.float32(0.5)
    s["r"] = np.maximum(x * np.float32(0.01), np.float32(1e-35))
    return s, m


@pytest.mark.parametrize("n,kind", [(1, O.BUILD_LBVH), (2, O.BUILD_SAH), (500, O.BUILD_SAH),
                                    (O.SAH_MAX_PRIMS, O.BUILD_SAH), (O.SAH_MAX_PRIMS + 1, O.BUILD_LBVH)])
def test_builder_by_size(n, kind):
    s, m = O.box_scene(n - 1, 40 + n) if n > 1 else O.box_scene(0, 1)
    b = O.BVH(s)
    assert b.builder() == kind and b.validate() == 0
    if kind == O.BUILD_LBVH and n > 1:
        assert np.all(np.diff(b.morton().astype(np.int64)) >= 0)


@pytest.mark.parametrize("n", [200, 1024])
def test_sah_depth_cap(n):
    """Centres at x = 2^-min(i, 125): each binned split peels off a few, so the
    build reaches its depth cap (24) and ends with median splits; the tree
    stays within 24 + 10 internal levels and 35 wide levels, and its hits
    equal brute force."""
    s, m = _geometric_scene(n)
    b = O.BVH(s)
    assert b.builder() == O.BUILD_SAH and b.validate() == 0
    assert 24 + 1 < b.depth() <= 24 + 10 + 1 and b.wide_depth() <= 35
    cam = O.camera_lookat([0.5, 0.5, -3.0], [0.5, 0.5, 0.0], 40.0, 2.0, 0.0, 10.0)
    rays = O.make_rayset(s, m, O.make_params(64, 32, 1, 4, cam), 2000, 2000, n)
    assert b.trace(rays)[0].tobytes() == O.trace_brute(s, rays).tobytes()


@pytest.mark.parametrize("n", [O.TREELET_MIN_PRIMS - 1, O.TREELET_MIN_PRIMS])
def test_lbvh_treelets(n):
    """From TREELET_MIN_PRIMS spheres up the LBVH's treelets are SAH trees:
    the validator sorts each back to Morton order and checks the LBVH above
    and the SAH layout inside; hits equal brute force either way."""
    s, m = O.box_scene(n - 1, 11)
    b = O.BVH(s)
    assert b.builder() == O.BUILD_LBVH and b.validate() == 0
    keys = b.morton().astype(np.int64)
    assert (np.all(np.diff(keys) >= 0)) == (n < O.TREELET_MIN_PRIMS)
    assert np.array_equal(np.sort(b.leaf_prims()), np.arange(n))
    cam = O.camera_lookat([0, 15, -120], [0, 0, 0], 40.0, 2.0, 0.0, 10.0)
    rays = O.make_rayset(s, m, O.make_params(64, 32, 1, 8, cam), 300, 300, n)
    assert b.trace(rays)[0].tobytes() == O.trace_brute(s, rays).tobytes()


def test_sah_tree_on_the_book_scene(book):
    """The reason small scenes get the SAH build: the book scene's paths
    visit fewer than 6.5 wide nodes per segment (its Morton-order LBVH took
    7.3 on the 1080p frame), and the frame equals the brute-force render bit
    for bit."""
    s, m = book
    W, H = 320, 180
    p = O.make_params(W, H, 2, 50, O.camera_lookat([13, 2, 3], [0, 0, 0], 20.0, float(np.float32(W) / np.float32(H)),
                                                   0.1, 10.0))
    acc_s, _, st_s = O.render(s, m, p, bvh=O.BVH(s), want_image=False)
    assert st_s.node_visits / st_s.segments < 6.5
    assert acc_s.tobytes() == O.render(s, m, p, bvh=None, want_image=False)[0].tobytes()


def test_bvh_fixture_hashes(book):
    import hashlib
    b = O.BVH(book[0])
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    assert h(b.nodes()) == FIX["nodes_sha256"] and h(b.leaf_prims()) == FIX["leaf_sha256"]
    assert b.depth() == FIX["depth"]


def test_rayset_subset_fixture(book):
    rays = np.load(os.path.join(GOLDEN, "rayset_4096.npy"))
    want = np.load(os.path.join(GOLDEN, "rayset_4096_hits.npy"))
    b = O.BVH(book[0])
    assert b.trace(rays)[0].tobytes() == want.tobytes()
    assert O.trace_brute(book[0], rays).tobytes() == want.tobytes()
    t64, p64 = O.trace_brute_f64(book[0], rays)
    agree = (p64 == want["prim"]).mean()
    assert agree > 0.995  # FP32 vs FP64 closest primitive


# ---- integrator (RayTracer.cu:163-217) -----------------------------------------------
def _book_params(w, h, spp, depth, **kw):
    cam = O.camera_lookat([13, 2, 3], [0, 0, 0], 20.0, float(np.float32(w) / np.float32(h)), 0.1, 10.0)
    return O.make_params(w, h, spp, depth, cam, **kw)


def test_render_bvh_equals_brute_and_threads(book):
    s, m = book
    p = _book_params(48, 27, 4, 50)
    a1, _, st1 = O.render(s, m, p, bvh=O.BVH(s))
    a2, _, st2 = O.render(s, m, p, bvh=None)
    p1 = _book_params(48, 27, 4, 50, threads=1)
    a3, _, _ = O.render(s, m, p1, bvh=O.BVH(s))
    assert np.array_equal(a1, a2) and np.array_equal(a1, a3) and st1.segments == st2.segments


def test_c0_fixture(book):
    import hashlib
    s, m = book
    acc, _, st = O.render(s, m, _book_params(400, 225, 16, 50), bvh=O.BVH(s))
    assert hashlib.sha256(acc.tobytes()).hexdigest() == FIX["c0"]["accum_sha256"]
    assert st.segments == FIX["c0"]["segments"] and st.paths == 400 * 225 * 16


def test_depth_one_is_sky_or_black(book):
    s, m = book
    p = _book_params(40, 24, 2, 1)
    acc, _, st = O.render(s, m, p, bvh=O.BVH(s))
    assert st.segments == st.paths  # one segment per path
    rgb = acc[..., :3].reshape(-1, 3).astype(np.float64)
    # every sample is either black (its camera ray hit) or the sky colour of
    # its direction, whose blue is exactly 1: so blue * spp counts the sky
    # samples exactly, and r, g sit on the sky's lines through the same
    # samples: B - R = 0.5 * sum(t) / spp and B - G = 0.3 * sum(t) / spp
    b = rgb[:, 2] * 2
    assert np.all(b == np.round(b)) and b.min() == 0 and b.max() == 2
    assert np.allclose((rgb[:, 2] - rgb[:, 0]) / 0.5, (rgb[:, 2] - rgb[:, 1]) / 0.3, atol=1e-6)
    assert np.all(rgb <= 1.0 + 1e-6)


def test_single_sphere_no_hits_gives_exact_sky_average():
    # all camera rays miss a sphere far behind the camera: each pixel is the
    # average of sky(d) over its jittered samples, bounded by the sky values
    # at the pixel's extreme directions
    s = np.zeros(1, O.SPHERE_DT)
    s[0] = (0.0, 0.0, 1e5, 1.0)
    m = np.zeros(1, O.MATERIAL_DT)
    cam = O.camera_lookat([0, 0, 0], [0, 0, -1], 60.0, 2.0, 0.0, 1.0)
    p = O.make_params(20, 10, 8, 50, cam)
    acc, _, st = O.render(s, m, p, bvh=O.BVH(s))
    assert st.hits == 0 and st.segments == st.paths == 20 * 10 * 8
    rgb = acc[..., :3]
    assert np.allclose(rgb[..., 2], 1.0)
    # lerp((1,1,1),(0.5,0.7,1),t): r = 1 - 0.5 t, g = 1 - 0.3 t, so (1-r)/0.5 == (1-g)/0.3
    assert np.allclose((1 - rgb[..., 0]) / 0.5, (1 - rgb[..., 1]) / 0.3, atol=1e-5)
    # brighter (whiter) towards the bottom rows: t grows with the view direction's y
    t = (1 - rgb[..., 0]) / 0.5
    assert np.all(np.diff(t.mean(axis=1)) > 0)


def test_progressive_blend(book):
    s, m = book
    b = O.BVH(s)
    a1, _, _ = O.render(s, m, _book_params(24, 16, 2, 8, subframe=1), bvh=b)
    f1 = a1.copy()
    a2, _, _ = O.render(s, m, _book_params(24, 16, 2, 8, subframe=2), bvh=b, accum=a1)
    solo2, _, _ = O.render(s, m, _book_params(24, 16, 2, 8, subframe=2), bvh=b)
    # into a zero accumulator frame 2 writes 0 + (1/2)(avg - 0), so avg = 2*solo2
    # exactly; on top of frame 1 it blends 1/2 of that average in
    avg2 = np.float32(2.0) * solo2[..., :3]
    expect = f1[..., :3] + np.float32(0.5) * (avg2 - f1[..., :3])
    assert np.array_equal(a2[..., :3], expect.astype(np.float32))


def test_fp64_integrator_agrees_statistically(book):
    s, m = book
    b = O.BVH(s)
    p = _book_params(64, 36, 16, 50)
    a32, _, _ = O.render(s, m, p, bvh=b)
    a64, _ = O.render_f64(s, m, p, bvh=b)
    d = a32[..., :3].astype(np.float64) - a64[..., :3]
    rmse = float(np.sqrt(np.mean(d * d)))
    assert rmse < 0.02
    assert abs(a32[..., :3].mean() - a64[..., :3].mean()) < 0.01
    # FP64 accumulation of FP32 paths: sums agree with the FP32 sums closely
    pa = _book_params(64, 36, 16, 50, accum_fp64=True)
    acc_d, _, _ = O.render(s, m, pa, bvh=b)
    assert np.abs(acc_d[..., :3] - a32[..., :3]).max() < 1e-5


def test_plant_a_mistake_is_caught():
    # the brute-force check is sensitive: shrinking one sphere by one ulp
    # changes the closest hits of rays that graze it
    s, m = O.box_scene(2000, 5)
    b = O.BVH(s)
    rays = O.make_rayset(s, m, O.make_params(64, 32, 1, 4, O.camera_lookat([0, 15, -120], [0, 0, 0], 40.0, 2.0,
                                                                            0.0, 10.0)), 3000, 0, 0)
    good = b.trace(rays)[0]
    s2 = s.copy()
    hit_prims = np.unique(good["prim"][good["prim"] != O.MISS])
    s2["r"][hit_prims] *= np.float32(0.9)
    assert O.trace_brute(s2, rays).tobytes() != good.tobytes()


# ---- box-gated sphere test ----------------------------------------------------------
def test_grazing_roots_are_gated_by_the_sphere_box():
    """Five rays on which the ungated float32 sphere test reports a root on a
    tiny distant sphere (r ~ 0.01-0.03 at 70-250 units) that the ray passes
    outside: |oc|^2 - r^2 loses r^2 to rounding.  The root lies outside the
    sphere's box, so the gated test drops it, and LBVH traversal equals brute
    force (which it did not while the test was ungated)."""
    s, _ = O.box_scene(20000, 2024)
    rays = np.load(os.path.join(GOLDEN, "box_grazing_rays.npy"))
    fix = json.load(open(os.path.join(GOLDEN, "box_grazing.json")))
    for ray, p in zip(rays, fix["spurious_prim"]):
        o = [ray["ox"], ray["oy"], ray["oz"]]
        d = [ray["dx"], ray["dy"], ray["dz"]]
        t = _hit32(o, d, s[p])
        assert t is not None  # the ungated test reports a root ...
        pt = np.asarray(o, np.float64) + t * np.asarray(d, np.float64)
        c = np.array([s[p]["cx"], s[p]["cy"], s[p]["cz"]], np.float64)
        r = abs(float(s[p]["r"]))
        assert np.any(np.abs(pt - c) > r * 1.01)  # ... at a point outside the sphere's box
    hb, _, _ = O.BVH(s).trace(rays)
    hf = O.trace_brute(s, rays)
    assert hb.tobytes() == hf.tobytes()
    assert [float(x) for x in hb["t"]] == fix["t"] and [int(x) for x in hb["prim"]] == fix["prim"]
    assert not set(int(x) for x in hb["prim"]) & set(fix["spurious_prim"])


def test_box_scene_bvh_equals_brute_force():
    """The closest hit does not depend on the tree: LBVH traversal equals
    brute force bit for bit on the integrator's rays of a box scene of tiny
    spheres, where ungated grazing roots used to make them differ."""
    s, m = O.box_scene(20000, 2024)
    W, H = 160, 90
    cam = O.camera_lookat([0, 15, -120], [0, 0, 0], 40.0, float(np.float32(W) / np.float32(H)), 0.0, 10.0)
    p = O.make_params(W, H, 16, 16, cam)
    b = O.BVH(s)
    rays = O.make_rayset(s, m, p, 24000, 6000, 5, bvh=b)
    hb, _, _ = b.trace(rays)
    hf = O.trace_brute(s, rays)
    assert hb.tobytes() == hf.tobytes()


def _slot_boxes(node):
    """Dequantised slot boxes of a wide node: (lo[4,3], hi[4,3], valid[4])."""
    meta = int(node["meta"])
    lo, hi = np.zeros((4, 3)), np.zeros((4, 3))
    for k in range(3):
        s = math.ldexp(1.0, ((meta >> (8 * k)) & 0xFF) - 127)
        for c in range(4):
            lo[c, k] = np.float32(s * ((int(node["qlo"][k]) >> (8 * c)) & 0xFF) + float(node["o"][k]))
            hi[c, k] = np.float32(s * ((int(node["qhi"][k]) >> (8 * c)) & 0xFF) + float(node["o"][k]))
    return lo, hi, [(meta >> (28 + c)) & 1 for c in range(4)]


@pytest.mark.parametrize("offset", [(0.0, 0.0, 0.0), (5000.0, -3000.0, 7000.0)])
def test_wide_box_test_contains_exact_form(offset):
    """The traversal's slot test (one fma per plane, planes moved out by a
    rounding bound) accepts every slot that the exact-form robust slab test
    on the dequantised box accepts, with an entry distance no larger, and
    hardly any more.  Rays start inside, outside and exactly on slot planes,
    run along, near and off the axes, and end (tmax) at or just short of the
    exact entry distance; one scene sits far from the origin so that the
    plane offsets are large against the slot extents."""
    rng = np.random.default_rng(7)
    s, _ = O.box_scene(4000, 11)
    for k, off in enumerate(("cx", "cy", "cz")):
        s[off] = (s[off].astype(np.float64) + offset[k]).astype(np.float32)
    nodes = O.BVH(s).wide_nodes()
    exact_total = extra = 0
    for ni in rng.choice(len(nodes), 250, replace=False):
        node = nodes[ni]
        lo, hi, valid = _slot_boxes(node)
        for _ in range(12):
            c = int(rng.choice([i for i in range(4) if valid[i]]))
            ext = np.maximum(hi[c] - lo[c], 1e-3)
            kind = rng.integers(4)
            o = lo[c] + rng.uniform(-1.0, 2.0, 3) * ext
            if kind == 1:  # on a face of the slot box
                k = rng.integers(3)
                o[k] = (lo[c] if rng.integers(2) else hi[c])[k]
            elif kind == 2:  # far away
                o = (lo[c] + hi[c]) / 2 + rng.normal(size=3) * 200.0 * ext.max()
            tgt = lo[c] + rng.uniform(0.0, 1.0, 3) * (hi[c] - lo[c])
            corner = rng.integers(3) == 0
            if corner:
                tgt = np.where(rng.integers(2, size=3) == 1, hi[c], lo[c])
            d = tgt - o
            mode = rng.integers(4)
            if mode == 1:
                d[rng.integers(3)] = 0.0
            elif mode == 2:
                d[rng.integers(3)] = rng.choice([1e-9, -1e-9, 1e-25])
            if not np.any(d):
                continue
            o32, d32 = o.astype(np.float32), d.astype(np.float32)
            for tmin in (0.001, 0.0):
                (mf, me), tf, te = O.wide_box_check(node, o32, d32, tmin, 1e16)
                tmaxes = [1e16] + [float(x) for x in te[[j for j in range(4) if (me >> j) & 1]]]
                for tmax in tmaxes + [float(np.float32(t * (1 - 2.0 ** -22))) for t in tmaxes[1:]]:
                    (mf, me), tf, te = O.wide_box_check(node, o32, d32, tmin, tmax)
                    assert mf & me == me, (ni, o32, d32, tmin, tmax, mf, me)
                    for j in range(4):
                        if (me >> j) & 1:
                            assert tf[j] <= te[j], (ni, j, tf[j], te[j])
                    if tmax == 1e16 and kind != 1 and mode == 0 and not corner:  # tightness, generic rays
                        exact_total += bin(me).count("1")
                        extra += bin(mf & ~me).count("1")
    assert exact_total > 800
    assert extra <= 0.01 * exact_total, (extra, exact_total)

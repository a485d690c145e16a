"""GPU parity: libvz (HIP, gfx950) against the CPU oracle, through the C ABI.

Bit-exact: LBVH arrays (nodes, leaf order, Morton codes), per-ray hit
(prim, t) on the fixed ray sets, segment and visit counts, FP32/FP64
accumulators.  Images (8-bit sRGB) within 1 LSB (display powf).
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle as O
from util import (GpuTarget, book_lookat, box_lookat, geometric_scene, gpu_params, oracle_params, rmse, sha)
from venusaur_amd import api
from venusaur_amd.hipmem import DeviceBuffer

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = json.load(open(os.path.join(GOLDEN, "book_fixture.json")))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def compare_bvh(scene, spheres):
    nodes, leaf, morton = scene.download()
    ob = O.BVH(spheres)
    assert ob.validate() == 0
    assert np.array_equal(morton, ob.morton())
    assert np.array_equal(leaf, ob.leaf_prims())
    assert nodes.tobytes() == ob.nodes().tobytes()
    assert scene.info().depth == ob.depth()
    # the 4-wide tree every traversal walks, collapsed on the GPU
    wide = scene.download_wide()
    assert wide.tobytes() == ob.wide_nodes().tobytes()
    assert scene.info().num_wide_nodes == len(wide) and scene.info().wide_depth == ob.wide_depth()
    assert scene.info().builder == ob.builder()
    return ob


def test_lbvh_book_bit_exact(ctx, book):
    say("case: lbvh book scene")
    s, m = book
    sc = api.Scene(ctx, s, m)
    nodes, leaf, morton = sc.download()
    assert sha(nodes) == FIX["nodes_sha256"]
    assert sha(leaf) == FIX["leaf_sha256"]
    assert sha(morton) == FIX["morton_sha256"]
    assert sha(sc.download_wide()) == FIX["wide_nodes_sha256"]
    compare_bvh(sc, s)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 1024, 1025, 2047, 2048, 2049, 4095, 4097, 20000, 131071,
                               131072, 300000, 1048583])
def test_lbvh_sizes_bit_exact(ctx, n):
    """2 to 1024 spheres take the binned-SAH build, the others the LBVH, with
    SAH treelets from 131072 spheres up."""
    say(f"case: bvh box scene n={n}")
    s, m = api.box_scene(n - 1, 1234 + n) if n > 1 else api.box_scene(0, 1)
    sc = api.Scene(ctx, s, m)
    ob = compare_bvh(sc, s)
    assert ob.builder() == (O.BUILD_SAH if 2 <= n <= O.SAH_MAX_PRIMS else O.BUILD_LBVH)


@pytest.mark.parametrize("n", [200, 1024])
def test_sah_depth_cap_bit_exact(ctx, n):
    """Centres at x = 2^-min(i, 125): every SAH split peels off a few spheres,
    so the build reaches its depth cap and finishes with median splits."""
    say(f"case: sah depth cap n={n}")
    s, m = geometric_scene(n)
    sc = api.Scene(ctx, s, m)
    ob = compare_bvh(sc, s)
    assert ob.depth() > 25
    rays = O.make_rayset(s, m, O.make_params(64, 32, 1, 4, O.camera_lookat([0.5, 0.5, -3.0], [0.5, 0.5, 0.0],
                                                                           40.0, 2.0, 0.0, 10.0)), 2000, 2000, n)
    got, _ = _trace(sc, rays)
    assert got.tobytes() == ob.trace(rays)[0].tobytes()


def test_lbvh_duplicate_centres(ctx):
    say("case: lbvh duplicates")
    s, m = api.box_scene(999, 5)
    s["cx"][500:] = 1.0
    s["cy"][500:] = 2.0
    s["cz"][500:] = 3.0
    sc = api.Scene(ctx, s, m)
    compare_bvh(sc, s)


def _trace(scene, rays, stats=False):
    rb = DeviceBuffer.from_array(rays)
    hb = DeviceBuffer(len(rays) * 8)
    st = scene.trace(rb.value, len(rays), hb.value, stats=stats)
    return hb.download(api.HIT_DT), st


def test_trace_fixture_subset(ctx, book):
    say("case: trace 4096-ray fixture")
    s, m = book
    sc = api.Scene(ctx, s, m)
    rays = np.load(os.path.join(GOLDEN, "rayset_4096.npy"))
    want = np.load(os.path.join(GOLDEN, "rayset_4096_hits.npy"))
    got, _ = _trace(sc, rays)
    assert got.tobytes() == want.tobytes()


def test_trace_rayset_1m(ctx, book):
    say("case: trace 1M-ray set")
    s, m = book
    rs = FIX["rayset"]
    c = rs["config"]
    p = oracle_params(c["width"], c["height"], c["spp"], c["max_depth"], book_lookat(c["width"], c["height"]))
    ob = O.BVH(s)
    rays = O.make_rayset(s, m, p, rs["n_int"], rs["n_iso"], rs["iso_seed"], bvh=ob)
    assert sha(rays) == rs["rays_sha256"]
    sc = api.Scene(ctx, s, m)
    got, st = _trace(sc, rays, stats=True)
    assert sha(got) == rs["hits_sha256"]
    assert st.segments == len(rays)
    assert st.node_visits == rs["node_visits"]
    assert st.sphere_tests == rs["sphere_tests"]


def test_trace_grazing_rays_gated(ctx):
    say("case: trace grazing rays on the 20000-sphere box scene")
    s, m = api.box_scene(20000, 2024)
    sc = api.Scene(ctx, s, m)
    rays = np.load(os.path.join(GOLDEN, "box_grazing_rays.npy"))
    fix = json.load(open(os.path.join(GOLDEN, "box_grazing.json")))
    got, _ = _trace(sc, rays)
    assert [float(x) for x in got["t"]] == fix["t"] and [int(x) for x in got["prim"]] == fix["prim"]
    assert got.tobytes() == O.BVH(s).trace(rays)[0].tobytes()


def test_trace_box_scene_vs_brute(ctx):
    say("case: trace box scene vs brute force")
    s, m = api.box_scene(50000, 77)
    sc = api.Scene(ctx, s, m)
    cam = box_lookat(320, 180)
    p = oracle_params(320, 180, 2, 16, cam)
    ob = O.BVH(s)
    rays = O.make_rayset(s, m, p, 6000, 2000, 99, bvh=ob)
    got, st = _trace(sc, rays, stats=True)
    want, nv, ns = ob.trace(rays)
    assert got.tobytes() == want.tobytes()
    assert (st.node_visits, st.sphere_tests) == (nv, ns)
    brute = O.trace_brute(s, rays[:1500])
    assert got[:1500].tobytes() == brute.tobytes()


@pytest.mark.parametrize("mode", ["megakernel", "wavefront"])
def test_render_c0_bit_exact(ctx, book, mode):
    say(f"case: render C0 400x225x16 depth 50 {mode}")
    s, m = book
    cam = book_lookat(400, 225)
    sc = api.Scene(ctx, s, m)
    flags = api.MEGAKERNEL if mode == "megakernel" else 0
    tgt = GpuTarget(400, 225)
    st = sc.launch(tgt.bind(gpu_params(400, 225, 16, 50, cam, flags=flags)))
    acc = tgt.accum_host()
    assert sha(acc) == FIX["c0"]["accum_sha256"]
    assert st.segments == FIX["c0"]["segments"]
    assert st.paths == FIX["c0"]["paths"]
    want_acc, want_img, _ = O.render(s, m, oracle_params(400, 225, 16, 50, cam), bvh=O.BVH(s))
    assert np.array_equal(acc, want_acc)
    img = tgt.image_host()
    diff = np.abs(img.astype(int) - want_img.astype(int))
    assert diff.max() <= 1 and (diff > 0).mean() < 1e-3


@pytest.mark.parametrize("mode", ["megakernel", "wavefront"])
def test_render_fp64_accumulation(ctx, book, mode):
    say(f"case: fp64 accumulation {mode}")
    s, m = book
    cam = book_lookat(96, 64)
    sc = api.Scene(ctx, s, m)
    flags = api.ACCUM_FP64 | (api.MEGAKERNEL if mode == "megakernel" else 0)
    tgt = GpuTarget(96, 64, fp64=True)
    sc.launch(tgt.bind(gpu_params(96, 64, 24, 50, cam, flags=flags)))
    want, _, _ = O.render(s, m, oracle_params(96, 64, 24, 50, cam, accum_fp64=True), bvh=O.BVH(s))
    assert np.array_equal(tgt.accum_host(), want)


@pytest.mark.parametrize("mode", ["megakernel", "wavefront"])
def test_progressive_frames(ctx, book, mode):
    say(f"case: progressive frames {mode}")
    s, m = book
    cam = book_lookat(64, 48)
    sc = api.Scene(ctx, s, m)
    flags = api.MEGAKERNEL if mode == "megakernel" else 0
    tgt = GpuTarget(64, 48)
    acc_o = None
    ob = O.BVH(s)
    for frame in (1, 2, 3):
        sc.launch(tgt.bind(gpu_params(64, 48, 4, 8, cam, subframe=frame, flags=flags)))
        acc_o, _, _ = O.render(s, m, oracle_params(64, 48, 4, 8, cam, subframe=frame), bvh=ob, accum=acc_o)
        assert np.array_equal(tgt.accum_host(), acc_o), frame


def test_render_box_scene(ctx):
    say("case: render box scene 20k spheres")
    s, m = api.box_scene(20000, 11)
    cam = box_lookat(160, 90)
    sc = api.Scene(ctx, s, m)
    tgt = GpuTarget(160, 90)
    st = sc.launch(tgt.bind(gpu_params(160, 90, 4, 16, cam)))
    want, _, ost = O.render(s, m, oracle_params(160, 90, 4, 16, cam), bvh=O.BVH(s))
    assert np.array_equal(tgt.accum_host(), want)
    assert st.segments == ost.segments


def test_visit_counts_match_oracle(ctx, book):
    """The book scene rendered alone walks its 4-wide tree staged in LDS;
    rendered in a batch next to another scene it walks the same tree from
    memory.  Both visit counts match the oracle traversing that tree, and
    the binary tree would visit more nodes."""
    say("case: visit counters (4-wide tree from LDS alone, from memory in a batch)")
    s, m = book
    cam = book_lookat(80, 45)
    sc = api.Scene(ctx, s, m)
    tgt = GpuTarget(80, 45)
    st = sc.launch(tgt.bind(gpu_params(80, 45, 4, 50, cam, flags=api.COUNT_VISITS)))
    ob = O.BVH(s)
    _, _, ost = O.render(s, m, oracle_params(80, 45, 4, 50, cam), bvh=ob)
    assert (st.segments, st.node_visits, st.sphere_tests) == (ost.segments, ost.node_visits, ost.sphere_tests)
    _, _, obin = O.render(s, m, oracle_params(80, 45, 4, 50, cam, binary_tree=True), bvh=ob)
    assert obin.segments == ost.segments and ost.node_visits < obin.node_visits
    bs, bm = api.box_scene(3000, 12)
    sc2 = api.Scene(ctx, bs, bm)
    tgt2 = GpuTarget(40, 30)
    p1 = tgt.bind(gpu_params(80, 45, 4, 50, cam, flags=api.COUNT_VISITS))
    p2 = tgt2.bind(gpu_params(40, 30, 2, 16, box_lookat(40, 30), flags=api.COUNT_VISITS))
    bst = api.batch_launch(ctx, [(sc, p1), (sc2, p2)])
    w1 = ost
    _, _, w2 = O.render(bs, bm, oracle_params(40, 30, 2, 16, box_lookat(40, 30)), bvh=O.BVH(bs))
    assert (bst.segments, bst.node_visits, bst.sphere_tests) == (
        w1.segments + w2.segments, w1.node_visits + w2.node_visits, w1.sphere_tests + w2.sphere_tests)
    sc2.close()


def test_refit_after_motion(ctx):
    say("case: refit")
    s, m = api.box_scene(3000, 21)
    sc = api.Scene(ctx, s, m)
    s2 = s.copy()
    rng = np.random.default_rng(3)
    s2["cx"][1:] += rng.uniform(-0.5, 0.5, len(s) - 1).astype(np.float32)
    s2["cy"][1:] += rng.uniform(-0.5, 0.5, len(s) - 1).astype(np.float32)
    sc.update_spheres(s2)
    p = oracle_params(200, 100, 1, 4, box_lookat(200, 100))
    rays = O.make_rayset(s2, m, p, 4000, 4000, 5)
    got, _ = _trace(sc, rays)
    want = O.trace_brute(s2, rays)
    assert got.tobytes() == want.tobytes()
    sc.rebuild()
    compare_bvh(sc, s2)


def test_scene_from_device_pointers(ctx):
    say("case: scene created and updated from device pointers")
    s, m = api.box_scene(5000, 31)
    ds, dm = DeviceBuffer.from_array(s), DeviceBuffer.from_array(m)
    a = api.Scene(ctx, s, m)
    b = api.Scene.from_device(ctx, ds.value, dm.value, len(s))
    for x, y in zip(a.download(), b.download()):
        assert x.tobytes() == y.tobytes()
    # (the device buffers stay referenced until the calls that read them return)
    s2 = s.copy()
    s2["cx"][1:] += np.float32(0.125)
    ds2 = DeviceBuffer.from_array(s2)
    a.update_spheres(s2)
    b.update_spheres_device(ds2.value)
    for x, y in zip(a.download(), b.download()):
        assert x.tobytes() == y.tobytes()
    # invalid device input is rejected with the offending index; a rejected
    # update leaves the scene as it was
    bad = s2.copy()
    bad["r"][17] = 0.0
    dbad = DeviceBuffer.from_array(bad)
    with pytest.raises(api.VzError, match="sphere 17"):
        b.update_spheres_device(dbad.value)
    for x, y in zip(a.download(), b.download()):
        assert x.tobytes() == y.tobytes()
    with pytest.raises(api.VzError, match="sphere 17"):
        api.Scene.from_device(ctx, dbad.value, dm.value, len(s))
    badm = m.copy()
    badm["type"][5] = 9
    dbadm = DeviceBuffer.from_array(badm)
    with pytest.raises(api.VzError, match="material 5"):
        api.Scene.from_device(ctx, ds.value, dbadm.value, len(s))


def test_invalid_params_raise(ctx, book):
    say("case: invalid params")
    s, m = book
    sc = api.Scene(ctx, s, m)
    tgt = GpuTarget(8, 8)
    bad = gpu_params(8, 8, 1, 4, book_lookat(8, 8))
    bad.width = 1
    with pytest.raises(api.VzError) as e:
        sc.launch(tgt.bind(bad))
    assert e.value.status == 1
    bad = tgt.bind(gpu_params(8, 8, 2048, 4, book_lookat(8, 8)))
    with pytest.raises(api.VzError):
        sc.launch(bad)  # >1024 spp needs FP64 accumulation
    bad = tgt.bind(gpu_params(8, 8, 1, 0, book_lookat(8, 8)))
    with pytest.raises(api.VzError):
        sc.launch(bad)


def test_c0_error_vs_fp64_oracle(ctx, book):
    say("case: RMSE vs FP32 and FP64 oracles")
    s, m = book
    cam = book_lookat(200, 112)
    sc = api.Scene(ctx, s, m)
    tgt = GpuTarget(200, 112)
    sc.launch(tgt.bind(gpu_params(200, 112, 16, 50, cam)))
    ob = O.BVH(s)
    want32, _, _ = O.render(s, m, oracle_params(200, 112, 16, 50, cam), bvh=ob)
    want64, _ = O.render_f64(s, m, oracle_params(200, 112, 16, 50, cam), bvh=ob)
    got = tgt.accum_host()
    assert rmse(got, want32) == 0.0
    e64 = rmse(got, want64)
    say(f"RMSE vs FP64 oracle at 16 spp: {e64:.3e}")
    assert e64 < 0.05  # statistical: paths diverge chaotically, stated not bounded at 1e-3


def test_context_outlives_its_scenes_and_servers():
    say("case: context destroy with live scenes / servers")
    c = api.Context(0)
    s, m = api.box_scene(200, 3)
    sc = api.Scene(c, s, m)
    srv = api.Server(c, hbm_budget_bytes=1 << 30, max_active=4)
    st = api.lib().vzDestroy(c.h)
    assert api.STATUS_NAMES[st] == "VZ_ERROR_INVALID_VALUE"
    assert "1 scene(s) and 1 server(s)" in api.lib().vzGetLastError().decode()
    srv.close()
    sc.close()
    c.close()  # now succeeds
    # the Python context closes what it still owns before destroying itself
    c2 = api.Context(0)
    api.Scene(c2, s, m)
    keep = api.Scene(c2, s, m)
    c2.close()
    assert not keep.h


def test_two_contexts_on_two_threads(book):
    """Separate contexts may be used from different threads: each thread
    creates its own context and scene and renders; both results are
    bit-exact against the oracle."""
    import threading
    say("case: two contexts on two threads")
    s, m = book
    jobs = [(64, 48, 4, 8), (48, 40, 6, 16)]
    want = [O.render(s, m, oracle_params(w, h, spp, d, book_lookat(w, h)), bvh=O.BVH(s))[0] for w, h, spp, d in jobs]
    got, errors = [None, None], []

    def worker(k):
        try:
            w, h, spp, d = jobs[k]
            c = api.Context(0)
            sc = api.Scene(c, s, m)
            for _ in range(3):  # several launches, interleaved with the other thread's
                t = GpuTarget(w, h)
                sc.launch(t.bind(gpu_params(w, h, spp, d, book_lookat(w, h))))
                got[k] = t.accum_host()
            sc.close()
            c.close()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    th = [threading.Thread(target=worker, args=(k,)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert not errors, errors
    for k in range(2):
        assert np.array_equal(got[k], want[k]), k

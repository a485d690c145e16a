"""GPU parity at the sizes of the traversal-bound, batched and served
configurations, against the CPU oracle:

- 1M-sphere box scene (seed 2024): LBVH arrays bit-exact; a 64k-ray set
  (integrator rays from pixels of rows spread over the whole 1080p frame,
  plus isotropic rays) with (t, prim) and the node-visit and sphere-test
  counts bit-exact; every 64th row of the 1080p, 64 spp, depth 16 frame
  bit-exact.
- The 64-scene batch of the batched-form benchmark (1e3-2e5 spheres,
  256-1024 px, 4/16/64 spp, depth 16): every item bit-identical to its plain
  launch under the default and the benchmark's tuning, and two items
  bit-exact against the oracle on sampled rows.
- A 300-request stream of the serving distribution (Poisson arrivals, 20%
  camera reuse of resident scenes, 5% 512x512 at 2048 spp with FP64
  accumulation) through one server: every request bit-identical to its
  plain launch, two (one of them FP64) bit-exact against the oracle on
  sampled rows.
"""
import sys
import threading
import time

import numpy as np
import pytest

import oracle as O
from util import box_lookat, gpu_params, oracle_params, sha
from venusaur_amd import api, workloads
from venusaur_amd.hipmem import DeviceBuffer, synchronize

pytestmark = pytest.mark.gpu
N_C2, SEED_C2 = 1_000_000, 2024


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def c2(ctx):
    say("setup: 1M-sphere box scene (seed 2024) on the GPU and in the oracle")
    s, m = api.box_scene(N_C2, SEED_C2)
    sc = api.Scene(ctx, s, m)
    ob = O.BVH(s)
    yield s, m, sc, ob
    sc.close()


def test_c2_lbvh_bit_exact(c2):
    say("case: C2 LBVH arrays vs oracle")
    s, m, sc, ob = c2
    nodes, leaf, morton = sc.download()
    assert np.array_equal(morton, ob.morton())
    assert np.array_equal(leaf, ob.leaf_prims())
    assert nodes.tobytes() == ob.nodes().tobytes()
    assert sc.info().depth == ob.depth()
    assert sc.download_wide().tobytes() == ob.wide_nodes().tobytes()
    assert sc.info().wide_depth == ob.wide_depth()


def test_c2_rayset_bit_exact(c2):
    say("case: C2 64k-ray set (rows spread over the 1080p frame + isotropic) vs oracle")
    s, m, sc, ob = c2
    # one sample per pixel on every 120th row: 9 rows from bottom to top
    p = oracle_params(1920, 1080, 1, 16, box_lookat(1920, 1080), row_step=120)
    rays = O.make_rayset(s, m, p, 49152, 16384, SEED_C2, bvh=ob)
    assert np.all(rays["tmax"] > 0)  # the recorder filled all 49152 integrator slots
    assert len(np.unique(rays["oy"][:49152])) > 1000  # camera and bounce rays, not one pixel's
    rb = DeviceBuffer.from_array(rays)
    hb = DeviceBuffer(len(rays) * 8)
    st = sc.trace(rb.value, len(rays), hb.value, stats=True)
    got = hb.download(api.HIT_DT)
    want, nv, ns = ob.trace(rays)
    assert got.tobytes() == want.tobytes()
    assert (st.segments, st.node_visits, st.sphere_tests) == (len(rays), nv, ns)
    assert (got["prim"] != api.MISS).mean() > 0.3


def test_c2_frame_rows_bit_exact(c2):
    say("case: C2 1080p 64 spp depth 16 frame, every 64th row vs oracle")
    s, m, sc, ob = c2
    W, H, spp, depth = 1920, 1080, 64, 16
    cam = box_lookat(W, H)
    acc = DeviceBuffer(W * H * 16)
    p = gpu_params(W, H, spp, depth, cam)
    p.accum, p.image = acc.value, None
    st = sc.launch(p)
    got = acc.download(np.float32, (H, W, 4))
    op = oracle_params(W, H, spp, depth, cam, row_step=64)
    t0 = time.perf_counter()
    want, _, ost = O.render(s, m, op, bvh=ob, want_image=False)
    say(f"  oracle: {ost.segments} segments on {len(range(0, H, 64))} rows in {time.perf_counter() - t0:.1f} s")
    rows = np.arange(0, H, 64)
    assert np.array_equal(got[rows], want[rows])
    assert st.paths == W * H * spp and st.segments > st.paths


# ---- batched form at its benchmark size --------------------------------------------
def _scenes_for(ctx, reqs):
    scenes = {}
    for r in reqs:
        if r.scene_key not in scenes:
            s, m = r.scene_arrays()
            scenes[r.scene_key] = api.Scene(ctx, s, m)
    return scenes


def _oracle_rows(r, row_step):
    s, m = r.scene_arrays()
    cam = dict(lookfrom=list(r.lookfrom), lookat=list(r.lookat), vfov=r.vfov, aperture=0.0, focus=10.0,
               aspect=float(np.float32(r.width) / np.float32(r.height)))
    op = oracle_params(r.width, r.height, r.spp, r.depth, cam, accum_fp64=r.fp64, row_step=row_step)
    acc, _, _ = O.render(s, m, op, bvh=O.BVH(s), want_image=False)
    return acc, np.arange(0, r.height, row_step)


def test_c3_batch_64_scenes(ctx):
    say("case: C3 64-scene batch (benchmark mix): items vs plain launches and the oracle")
    reqs = workloads.batch_requests(64)
    scenes = _scenes_for(ctx, reqs)
    for r in reqs:
        r.bind()
    items = [(scenes[r.scene_key], r.params) for r in reqs]
    plain_seg = 0
    want = []
    for k, (sc, p) in enumerate(items):
        plain_seg += sc.launch(p).segments
        want.append(sha(reqs[k].accum_host()))
    say(f"  plain launches done: {plain_seg} segments")
    tunings = [None]
    t = api.Tuning()
    t.queue_capacity, t.tile_w, t.tile_h, t.slice_split = 6 << 22, 8, 8, 4  # a 24M queue, square tiles, split 4
    tunings.append(t)
    for tu in tunings:
        for r in reqs:
            r.accum.zero()
        st = api.batch_launch(ctx, items, tuning=tu)
        assert st.segments == plain_seg
        for k, r in enumerate(reqs):
            assert sha(r.accum_host()) == want[k], (k, None if tu is None else tu.as_dict())
        say(f"  batch ({'default' if tu is None else 'q=24M 8x8'}) identical to plain: {st.render_ms:.0f} ms")
    # two items against the oracle: the cheapest, and the largest scene among the small images
    cheap = min(range(64), key=lambda i: reqs[i].paths)
    small = [i for i in range(64) if reqs[i].width <= 512 and i != cheap]
    big = max(small, key=lambda i: reqs[i].n_small)
    for k in (cheap, big):
        r = reqs[k]
        acc, rows = _oracle_rows(r, 16)
        say(f"  item {k}: {r.n_small} spheres {r.width}px {r.spp} spp vs oracle rows y%16")
        assert np.array_equal(r.accum_host()[rows], acc[rows]), k
    for sc in scenes.values():
        sc.close()


# ---- served form at 300 requests ----------------------------------------------------
@pytest.mark.timeout(900)
def test_c4_stream_300_requests(ctx):
    say("case: C4 300-request stream through one server: each vs its plain launch and the oracle")
    reqs = workloads.request_stream(300, 150.0, seed=11)
    assert sum(r.fp64 for r in reqs) >= 5 and sum(r.reuse for r in reqs) >= 30
    scenes = _scenes_for(ctx, reqs)
    for r in reqs:
        r.bind()
    want = []
    for k, r in enumerate(reqs):
        if k % 50 == 0:
            say(f"  plain launch {k}/300")
        scenes[r.scene_key].launch(r.params)
        want.append(sha(r.accum_host()))
        r.accum.zero()
    synchronize()
    srv = api.Server(ctx, hbm_budget_bytes=32 << 30, max_active=256)
    ids = [None] * len(reqs)
    t0 = time.perf_counter()

    def client():
        for k, r in enumerate(reqs):
            now = time.perf_counter() - t0
            if now < r.arrival_s:
                time.sleep(r.arrival_s - now)
            ids[k] = srv.submit(scenes[r.scene_key], r.params)
    th = threading.Thread(target=client)
    th.start()
    th.join()
    say("  all submitted; draining")
    srv.drain()
    res = [srv.poll(i) for i in ids]
    srv.close()
    assert all(x is not None and x.status == 0 for x in res)
    for k, r in enumerate(reqs):
        assert sha(r.accum_host()) == want[k], k
    say(f"  served == plain for all 300 ({time.perf_counter() - t0:.1f} s)")
    fp64 = next(i for i, r in enumerate(reqs) if r.fp64)
    reuse = min((i for i, r in enumerate(reqs) if r.reuse), key=lambda i: reqs[i].paths)
    for k, step in ((fp64, 64), (reuse, 16)):
        r = reqs[k]
        acc, rows = _oracle_rows(r, step)
        say(f"  request {k}: {r.n_small} spheres {r.width}px {r.spp} spp fp64={r.fp64} vs oracle rows y%{step}")
        assert np.array_equal(r.accum_host()[rows], acc[rows]), k
    for sc in scenes.values():
        sc.close()


# ---- regression tests of the scene-table hazards ----------------------------------
def test_batch_of_2_then_80_scenes_on_one_context(ctx):
    """The captured graph reads the engine's scene table; a batch naming more
    scenes than the previous one must not move that table (it is allocated
    once at engine creation)."""
    say("case: 2-scene batch, then an 80-scene batch, on one context")
    small = [api.box_scene(200 + 10 * k, 500 + k) for k in range(80)]
    sc = [api.Scene(ctx, s, m) for s, m in small]
    cam = box_lookat(24, 16)

    def run(idx):
        tg, items = [], []
        for k in idx:
            acc = DeviceBuffer(24 * 16 * 16)
            p = gpu_params(24, 16, 2, 6, cam)
            p.accum, p.image = acc.value, None
            tg.append(acc)
            items.append((sc[k], p))
        want = []
        for (s_, p), a in zip(items, tg):
            s_.launch(p)
            want.append(a.download(np.float32, (16, 24, 4)))
            a.zero()
        api.batch_launch(ctx, items)
        for k, ((s_, p), a) in enumerate(zip(items, tg)):
            assert np.array_equal(a.download(np.float32, (16, 24, 4)), want[k]), (len(idx), k)
    run([0, 1])
    run(list(range(80)))
    run([2, 3])
    for s_ in sc:
        s_.close()


def test_server_scene_destroy_refused_then_recreated(ctx):
    """A scene named by queued or running requests cannot be destroyed; once
    they finish it can, and a new scene (possibly at the same address) is
    rendered from its own BVH, not a cached view of the old one."""
    say("case: server, scene destroy refused while in use, then a new scene")
    srv = api.Server(ctx, hbm_budget_bytes=4 << 30, max_active=8)
    sa, ma = api.box_scene(3000, 71)
    A = api.Scene(ctx, sa, ma)
    cam = box_lookat(256, 256)
    tg = []
    ids = []
    for _ in range(6):
        acc = DeviceBuffer(256 * 256 * 16)
        p = gpu_params(256, 256, 16, 8, cam)
        p.accum, p.image = acc.value, None
        tg.append(acc)
        ids.append(srv.submit(A, p))
    st = api.lib().vzSceneDestroy(A.h)
    assert st == 1 and "server request" in api.lib().vzGetLastError().decode()
    srv.drain()
    assert all(srv.poll(i).status == 0 for i in ids)
    A.close()  # now accepted
    for rnd in range(3):
        sb, mb = api.box_scene(2000 + rnd, 90 + rnd)
        B = api.Scene(ctx, sb, mb)
        acc = DeviceBuffer(64 * 48 * 16)
        p = gpu_params(64, 48, 4, 8, box_lookat(64, 48))
        p.accum, p.image = acc.value, None
        rid = srv.submit(B, p)
        srv.drain()
        assert srv.poll(rid).status == 0
        got = acc.download(np.float32, (48, 64, 4))
        want, _, _ = O.render(sb, mb, oracle_params(64, 48, 4, 8, box_lookat(64, 48)), bvh=O.BVH(sb))
        assert np.array_equal(got, want), rnd
        B.close()
    srv.close()


@pytest.mark.timeout(600)
def test_server_sees_more_than_4096_scenes(ctx):
    """4200 distinct scenes through one server, in waves whose scenes are
    destroyed after use: unreferenced scene entries are reused, so there is
    no lifetime limit on the scenes a server sees."""
    say("case: 4200 scenes through one server")
    srv = api.Server(ctx, hbm_budget_bytes=4 << 30, max_active=128)
    cam = box_lookat(8, 6)
    checked = 0
    for wave in range(42):
        if wave % 10 == 0:
            say(f"  wave {wave}/42")
        scenes, ids, accs, arrays = [], [], [], []
        for k in range(100):
            s, m = api.box_scene(6 + k % 5, 10_000 + wave * 100 + k)
            sc = api.Scene(ctx, s, m)
            acc = DeviceBuffer(8 * 6 * 16)
            p = gpu_params(8, 6, 1, 3, cam)
            p.accum, p.image = acc.value, None
            ids.append(srv.submit(sc, p))
            scenes.append(sc)
            accs.append(acc)
            arrays.append((s, m))
        srv.drain()
        assert all(srv.poll(i).status == 0 for i in ids)
        k = wave % 100
        s, m = arrays[k]
        want, _, _ = O.render(s, m, oracle_params(8, 6, 1, 3, cam), bvh=O.BVH(s))
        assert np.array_equal(accs[k].download(np.float32, (6, 8, 4)), want), (wave, k)
        checked += 1
        for sc in scenes:
            sc.close()
    srv.close()
    assert checked == 42

"""GPU tests of the batched and serving forms: every request's result is
bit-identical to the CPU oracle rendering that request alone, whatever else
shares the wavefront with it."""
import sys
import time

import numpy as np
import pytest

import oracle as O
from util import GpuTarget, book_lookat, box_lookat, gpu_params, oracle_params
from venusaur_amd import api

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scenes(ctx):
    bs, bm = api.book_scene()
    out = {"book": (api.Scene(ctx, bs, bm), bs, bm)}
    for n, seed in ((2000, 41), (12000, 42)):
        s, m = api.box_scene(n, seed)
        out[f"box{n}"] = (api.Scene(ctx, s, m), s, m)
    yield out
    for v in out.values():
        v[0].close()


# (scene, w, h, spp, depth, fp64)
MIX = [("book", 64, 48, 8, 50, False), ("box2000", 96, 64, 4, 16, False), ("box12000", 80, 80, 16, 16, False),
       ("book", 48, 32, 24, 16, True), ("box2000", 40, 30, 2, 4, False), ("box12000", 33, 17, 3, 16, False)]


def _cam(name, w, h):
    return book_lookat(w, h) if name == "book" else box_lookat(w, h)


def _oracle(sc, w, h, spp, depth, fp64, name, obvh):
    s, m = sc[1], sc[2]
    acc, _, st = O.render(s, m, oracle_params(w, h, spp, depth, _cam(name, w, h), accum_fp64=fp64), bvh=obvh)
    return acc, st


@pytest.fixture(scope="module")
def oracle_results(scenes):
    bvhs = {k: O.BVH(v[1]) for k, v in scenes.items()}
    return [_oracle(scenes[n], w, h, spp, d, f, n, bvhs[n]) for n, w, h, spp, d, f in MIX]


def _items(scenes):
    items, tgts = [], []
    for n, w, h, spp, d, f in MIX:
        t = GpuTarget(w, h, fp64=f)
        p = t.bind(gpu_params(w, h, spp, d, _cam(n, w, h), flags=api.ACCUM_FP64 if f else 0))
        items.append((scenes[n][0], p))
        tgts.append(t)
    return items, tgts


def test_batched_alone_equals_mix(ctx, scenes, oracle_results):
    say("case: batched mix of 6 requests")
    items, tgts = _items(scenes)
    st = api.batch_launch(ctx, items)
    for k, (t, (want, ost)) in enumerate(zip(tgts, oracle_results)):
        assert np.array_equal(t.accum_host(), want), MIX[k]
    assert st.segments == sum(o[1].segments for o in oracle_results)
    # replay: the captured graph again, same results
    for t in tgts:
        t.accum.zero()
    api.batch_launch(ctx, items)
    for k, (t, (want, _)) in enumerate(zip(tgts, oracle_results)):
        assert np.array_equal(t.accum_host(), want), MIX[k]


@pytest.mark.parametrize("q,tw,th,split", [(1 << 16, 8, 8, 4), (1 << 16, 8, 8, 1), (1 << 20, 16, 4, 2),
                                           (1 << 22, 64, 1, 4), (1 << 23, 32, 2, 1), (1 << 22, 8, 8, 8)])
def test_batched_any_tuning_same_result(ctx, scenes, oracle_results, q, tw, th, split):
    say(f"case: batched tuning q={q} tile={tw}x{th} split={split}")
    items, tgts = _items(scenes)
    t = api.Tuning()
    t.queue_capacity, t.tile_w, t.tile_h, t.iters_per_graph, t.slice_split = q, tw, th, 8, split
    api.batch_launch(ctx, items, tuning=t)
    for k, (tg, (want, _)) in enumerate(zip(tgts, oracle_results)):
        assert np.array_equal(tg.accum_host(), want), MIX[k]


def test_plain_batch_plain_on_one_context(ctx, scenes, oracle_results):
    """Plain launches use one queue slice per resident step workgroup and
    batches four; the engine re-partitions its queue between runs."""
    say("case: plain, batched (split 4), plain again on one context")
    items, tgts = _items(scenes)
    for rnd in range(2):
        for k, ((sc, p), tg) in enumerate(zip(items, tgts)):
            tg.accum.zero()
            sc.launch(p)
            assert np.array_equal(tg.accum_host(), oracle_results[k][0]), (rnd, MIX[k])
        if rnd == 0:
            for tg in tgts:
                tg.accum.zero()
            api.batch_launch(ctx, items)
            for k, (tg, (want, _)) in enumerate(zip(tgts, oracle_results)):
                assert np.array_equal(tg.accum_host(), want), MIX[k]


def test_batch_at_the_request_table_limit(ctx, scenes):
    """1024 items, the request table's size, in one batch with the queue split
    4 and 1 ways: every accumulator equals a plain launch of the same item.
    A 1025th item is refused with VZ_ERROR_CAPACITY.  The items have 48-96
    pixels, so one wave of the finish kernel's apply and regen stages covers
    several requests: this exercises their per-request waterfall loops (a
    wave that took the first lane's request for all lanes left pixels
    unwritten here)."""
    say("case: 1024-item batch (request-table limit), splits 4 and 1; 1025 refused")
    names = ("book", "box2000")
    items, tgts = [], []
    for k in range(1024):
        name = names[k % 2]
        w, h = 8 + k % 5, 6 + k % 3
        t = GpuTarget(w, h)
        items.append((scenes[name][0], t.bind(gpu_params(w, h, 1 + k % 3, 4 + k % 4, _cam(name, w, h)))))
        tgts.append(t)
    want = []
    for (sc, p), t in zip(items, tgts):
        sc.launch(p)
        want.append(t.accum_host())
    for split in (4, 1):
        for t in tgts:
            t.accum.zero()
        tu = api.Tuning()
        tu.slice_split = split
        api.batch_launch(ctx, items, tuning=tu)
        for k, t in enumerate(tgts):
            assert np.array_equal(t.accum_host(), want[k]), (split, k)
    extra = GpuTarget(8, 6)
    over = items + [(scenes["book"][0], extra.bind(gpu_params(8, 6, 1, 4, _cam("book", 8, 6))))]
    with pytest.raises(api.VzError) as err:
        api.batch_launch(ctx, over)
    assert err.value.status == 5  # VZ_ERROR_CAPACITY


@pytest.mark.parametrize("split", [0, 1])
def test_serving_continuous_batching(ctx, scenes, oracle_results, split):
    say(f"case: serving, staggered submissions, slice split {split or 'default'}")
    tuning = None
    if split:
        tuning = api.Tuning()
        tuning.slice_split = split
    srv = api.Server(ctx, hbm_budget_bytes=8 << 30, max_active=64, tuning=tuning)
    items, tgts = _items(scenes)
    ids = []
    for k, (sc, p) in enumerate(items):
        ids.append(srv.submit(sc, p))
        if k == 2:
            time.sleep(0.01)  # later requests arrive while the first ones run
    # a second copy of the whole mix, submitted while the first is in flight
    items2, tgts2 = _items(scenes)
    ids2 = [srv.submit(sc, p) for sc, p in items2]
    srv.drain()
    res = [srv.poll(i) for i in ids + ids2]
    assert all(r is not None and r.status == 0 for r in res)
    for k, (t, (want, ost)) in enumerate(zip(tgts + tgts2, oracle_results + oracle_results)):
        assert np.array_equal(t.accum_host(), want), k
        assert res[k].segments == ost.segments
        assert res[k].submit_ms <= res[k].admit_ms <= res[k].finish_ms
    srv.close()


def test_serving_reuses_slots(ctx, scenes):
    """300 small requests through a 64-entry table: slots are freed and
    reused about five times over, and every result equals a plain launch."""
    say("case: serving, 300 requests through 64 slots")
    names = ("book", "box2000", "box12000")
    items, tgts = [], []
    for k in range(300):
        name = names[k % 3]
        w, h = 12 + k % 7, 9 + k % 4
        t = GpuTarget(w, h)
        items.append((scenes[name][0], t.bind(gpu_params(w, h, 1 + k % 4, 3 + k % 6, _cam(name, w, h)))))
        tgts.append(t)
    want = []
    for (sc, p), t in zip(items, tgts):
        sc.launch(p)
        want.append(t.accum_host())
        t.accum.zero()
    srv = api.Server(ctx, hbm_budget_bytes=8 << 30, max_active=64)
    ids = [srv.submit(sc, p) for sc, p in items]
    srv.drain()
    res = [srv.poll(i) for i in ids]
    srv.close()
    assert all(r is not None and r.status == 0 for r in res)
    for k, t in enumerate(tgts):
        assert np.array_equal(t.accum_host(), want[k]), k


def test_serving_admission_by_footprint(ctx, scenes):
    say("case: serving admission budget")
    w, h, spp = 96, 64, 4
    p0 = gpu_params(w, h, spp, 16, box_lookat(w, h))
    fp = api.footprint(p0)
    srv = api.Server(ctx, hbm_budget_bytes=int(fp * 1.5), max_active=16)
    tg = [GpuTarget(w, h) for _ in range(3)]
    ids = [srv.submit(scenes["box2000"][0], t.bind(gpu_params(w, h, spp, 16, box_lookat(w, h)))) for t in tg]
    big = GpuTarget(512, 512)
    big_id = srv.submit(scenes["box2000"][0], big.bind(gpu_params(512, 512, 4, 16, box_lookat(512, 512))))
    srv.drain()
    r = [srv.poll(i) for i in ids]
    assert all(x.status == 0 for x in r)
    # only one request fits the budget at a time: admissions are serialised
    assert r[1].admit_ms >= r[0].finish_ms and r[2].admit_ms >= r[1].finish_ms
    assert srv.poll(big_id).status == 5  # larger than the whole budget
    first = tg[0].accum_host()
    assert all(np.array_equal(t.accum_host(), first) for t in tg[1:])
    srv.close()


def test_autotune_picks_a_candidate_and_keeps_results(ctx, scenes, oracle_results):
    say("case: autotune")
    items, tgts = _items(scenes)
    t = api.autotune(ctx, items)
    assert t.queue_capacity in (1 << 22, 1 << 23, 3 << 22, 6 << 22, 8 << 22, 12 << 22)
    assert (t.tile_w, t.tile_h) in ((32, 2), (16, 4), (8, 8))
    assert t.refill_lanes in (0, 25, 33, 49, 57)
    assert t.slice_split in (1, 2, 4, 8)
    # cached per kind of device: a fresh context on this GPU starts from it
    other = api.Context(0)
    try:
        d = other.default_tuning()
        assert (d.queue_capacity, d.tile_w, d.tile_h, d.refill_lanes, d.slice_split) == (
            t.queue_capacity, t.tile_w, t.tile_h, t.refill_lanes, t.slice_split)
    finally:
        other.close()
    for tg in tgts:
        tg.accum.zero()
    api.batch_launch(ctx, items, tuning=t)
    for k, (tg, (want, _)) in enumerate(zip(tgts, oracle_results)):
        assert np.array_equal(tg.accum_host(), want), MIX[k]
    assert ctx.default_tuning().queue_capacity == t.queue_capacity

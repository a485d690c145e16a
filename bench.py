"""Benchmark: Mray-segments/s of the plain method on the 1080p book scene.

Workload (one step): the book final scene (486 spheres, mt19937 seed 5489),
1920x1080, 256 samples per pixel, max depth 50, FP32, rendered by one
vzLaunch (LBVH resident, wavefront integrator).  Inputs are resident in HBM
before the timed region; the LBVH build is timed separately.

Prints ONE JSON line.  Besides value/ms_per_step it carries:
  roofline      the trace kernel (dominant): algorithmic bytes per launch /
                its mean launch time from HIP events recorded around every
                trace launch inside the timed steps, against 8 TB/s HBM
  cpu_baseline  the CPU oracle (same integrator, FP32) on a row sample of the
                same workload, on this host's cores
  checks        segment count vs the oracle on the sample rows, RMSE vs the
                oracle (FP32) and vs the FP64 oracle on a smaller row sample
  forms         the 1M-sphere traversal-bound scene; the batched form against
                64 sequential plain launches; the serving form against
                one-at-a-time plain launches on a Poisson stream at 3x the
                plain capacity (served) and at 0.8x (served_sustainable)
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import numpy as np  # noqa: E402

METRIC = "Mray-segments/s @1080p; batched/served vs plain, p50/p99 latency, RMSE≤1e-3"
HBM_PEAK_GBS = 8000.0
# VALU lane throughput of MI355X for unpacked FP32/INT32 operations: 256 CUs x
# 4 SIMDs x 16 lanes x 2.4 GHz (a wave64 instruction issues over 4 cycles);
# the published 157.3 TFLOP/s FP32 is this x 2 (FMA) x 2 (packed pairs)
VALU_PEAK_TLANE = 256 * 4 * 16 * 2.4e9 / 1e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--mode", choices=["wavefront", "megakernel"], default="wavefront")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=256)
    ap.add_argument("--depth", type=int, default=50)
    ap.add_argument("--cpu-row-step", type=int, default=8, help="CPU baseline renders every k-th row")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline (profiling runs)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (scales rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--batch-scenes", type=int, default=64, help="scenes in the batched-form comparison (0 = skip)")
    ap.add_argument("--serve-requests", type=int, default=2000,
                    help="requests in the serving comparison at 3x the plain capacity (0 = skip)")
    ap.add_argument("--serve-sustainable", type=int, default=0,
                    help="requests in a second serving comparison at 0.8x the plain capacity (0 = skip)")
    ap.add_argument("--batch-eager", action="store_true",
                    help="batched form without autotuner or hipGraph, each step launch timed (for profiling)")
    ap.add_argument("--batch-tuning", default="",
                    help="QUEUE,TW,TH,SPLIT[,REFILL] for --batch-eager (default: the context's default tuning), "
                         "e.g. the autotuner's pick from an unprofiled run")
    ap.add_argument("--c2-frames", type=int, default=1,
                    help="timed frames of the 1M-sphere traversal-bound scene (0 = skip)")
    ap.add_argument("--summarize", metavar="KERNEL",
                    help="no GPU: summarise rocprofv3 CSVs of profiled bench runs for KERNEL "
                         "(e.g. 'k_wf_step<false, true, true>') into --out-md / --out-json")
    ap.add_argument("--trace-csv", help="kernel_stats.csv of the --kernel-trace --stats run")
    ap.add_argument("--fetch-csv", help="counter_collection.csv of the --pmc FETCH_SIZE run")
    ap.add_argument("--write-csv", help="counter_collection.csv of the --pmc WRITE_SIZE run")
    ap.add_argument("--l2-csv", help="counter_collection.csv of the --pmc TCC_HIT_sum TCC_MISS_sum run")
    ap.add_argument("--summarize-sq", metavar="CSV",
                    help="no GPU: summarise a rocprofv3 --pmc CSV of SQ counters (SQ_INSTS_VALU "
                         "SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_BUSY_CYCLES ...) into --out-md and "
                         "--out-json, for the kernels named by --sq-map")
    ap.add_argument("--sq-map", default="", help="'KERNEL=key;KERNEL=key': workload key of each kernel")
    ap.add_argument("--out-md")
    ap.add_argument("--out-json")
    ap.add_argument("--commands", default="", help="the profiling commands, ' ;; '-separated, for the summary")
    return ap.parse_args()


def _kernel_name(n):
    import re
    m = re.search(r"(k_\w+|__amd_\w+)(<[^>]*>)?", n)
    return (m.group(1) + (m.group(2) or "")) if m else n[:40]


def summarize(a):
    """Summary of rocprofv3 CSVs for one kernel: the kernel-trace table, HBM
    traffic per dispatch from FETCH_SIZE / WRITE_SIZE scaled by the counters'
    reading for the run's own device-to-device copy of --calibrate-copy MiB,
    and the L2 hit rate (TCC_HIT / TCC_MISS).  Writes the markdown summary and
    the JSON that bench.py reads (traffic_bytes_per_launch, l2_hit_rate)."""
    import csv
    import statistics
    kern = a.summarize
    rows = list(csv.DictReader(open(a.trace_csv)))
    table = ["| kernel | calls | total ms | avg us | min us | max us | % |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for r in rows:
        table.append(f"| {_kernel_name(r['Name'])} | {r['Calls']} | {int(r['TotalDurationNs']) / 1e6:.3f} | "
                     f"{float(r['AverageNs']) / 1e3:.2f} | {int(r['MinNs']) / 1e3:.2f} | {int(r['MaxNs']) / 1e3:.2f} | "
                     f"{float(r['Percentage']):.1f} |")
    avg_ns = float({_kernel_name(r["Name"]): r for r in rows}[kern]["AverageNs"])

    def counter(path, name=None):
        vals = {"copy": [], "kern": []}
        for r in csv.DictReader(open(path)):
            if name and r["Counter_Name"] != name:
                continue
            k = _kernel_name(r["Kernel_Name"])
            if k == "__amd_rocclr_copyBuffer":
                vals["copy"].append(float(r["Counter_Value"]))
            elif k == kern:
                vals["kern"].append(float(r["Counter_Value"]))
        return vals

    f, w = counter(a.fetch_csv), counter(a.write_csv)
    true_kib = a.calibrate_copy * 1024
    ffac, wfac = f["copy"][0] / true_kib, w["copy"][0] / true_kib
    kf, kw = statistics.mean(f["kern"]), statistics.mean(w["kern"])
    fetch_b, write_b = kf * 1024 / ffac, kw * 1024 / wfac
    md = [f"# rocprofv3 summary: {kern}", "", "Commands (run from the repository root on the MI355X box):", "",
          "```", *[c.strip() for c in a.commands.split(" ;; ") if c.strip()], "```", "",
          "## Kernel time (kernel trace, --stats)", "", *table, "",
          "## HBM traffic of the dominant kernel (PMC)", "",
          f"- calibration: a {a.calibrate_copy} MiB device-to-device copy (`__amd_rocclr_copyBuffer`) reads "
          f"FETCH_SIZE = {f['copy'][0]:.1f} KiB and WRITE_SIZE = {w['copy'][0]:.1f} KiB, i.e. factors {ffac:.4f} "
          f"(fetch) and {wfac:.4f} (write) of the true bytes; kernel counters below are divided by these factors.",
          f"- `{kern}`: {len(f['kern'])} dispatches; mean FETCH_SIZE {kf:.1f} KiB, WRITE_SIZE {kw:.1f} KiB per dispatch",
          f"- scaled HBM traffic per dispatch: fetch {fetch_b / 1e6:.3f} MB + write {write_b / 1e6:.3f} MB = "
          f"{(fetch_b + write_b) / 1e6:.3f} MB",
          f"- mean dispatch duration (kernel trace): {avg_ns / 1e3:.2f} us -> measured HBM rate "
          f"{(fetch_b + write_b) / avg_ns:.1f} GB/s"]
    out = {}
    if a.l2_csv:
        h, m = counter(a.l2_csv, "TCC_HIT_sum")["kern"], counter(a.l2_csv, "TCC_MISS_sum")["kern"]
        hit, miss = statistics.mean(h), statistics.mean(m)
        out = {"l2_hit_rate": round(hit / (hit + miss), 4), "l2_requests_per_launch": round(hit + miss)}
        md += ["", "## L2 of the dominant kernel (PMC: TCC_HIT_sum, TCC_MISS_sum)", "",
               f"- `{kern}`: {len(h)} dispatches; mean {hit:.4g} hits and {miss:.4g} misses per dispatch: hit rate "
               f"{hit / (hit + miss):.3f}, {(hit + miss) / (avg_ns * 1e-9) / 1e9:.1f} G requests/s"]
    open(a.out_md, "w").write("\n".join(md) + "\n")
    out.update({"kernel": kern, "traffic_bytes_per_launch": round(fetch_b + write_b),
                "fetch_bytes_per_launch": round(fetch_b), "write_bytes_per_launch": round(write_b),
                "dispatches": len(f["kern"]), "avg_launch_ns_trace": avg_ns, "fetch_factor": ffac,
                "write_factor": wfac, "source": os.path.relpath(a.out_md, ROOT)})
    json.dump(out, open(a.out_json, "w"), indent=1)


def summarize_sq(a):
    """Per-dispatch means of SQ counters for the kernels of --sq-map, with
    the VALU issue fraction (instructions over the issue capacity of the
    busy cycles: 1024 SIMDs, one wave64 VALU instruction per 4 cycles, the
    busy-cycle counter summed over 32 shader engines) and the fraction of
    lanes active per VALU instruction."""
    import csv
    import statistics
    want = dict(x.split("=", 1) for x in a.sq_map.split(";") if x.strip())
    vals = {}
    for r in csv.DictReader(open(a.summarize_sq)):
        k = _kernel_name(r["Kernel_Name"])
        if k in want:
            vals.setdefault(k, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    out = json.load(open(a.out_json)) if os.path.exists(a.out_json) else {}
    rows = ["| kernel | workload | dispatches | SQ_INSTS_VALU | SQ_ACTIVE_INST_VALU | SQ_THREAD_CYCLES_VALU | "
            "SQ_BUSY_CYCLES | VALU issue / capacity | lanes active |", "|---|---|---:|---:|---:|---:|---:|---:|---:|"]
    for k, key in want.items():
        v = {c: statistics.mean(x) for c, x in vals[k].items()}
        n = len(vals[k]["SQ_INSTS_VALU"])
        issue = v["SQ_INSTS_VALU"] / (v["SQ_BUSY_CYCLES"] / 32 * 1024 / 4)
        lanes = v["SQ_THREAD_CYCLES_VALU"] / (64 * v["SQ_ACTIVE_INST_VALU"])
        out[key] = {"kernel": k, "dispatches": n, "insts_valu": v["SQ_INSTS_VALU"],
                    "active_inst_valu": v["SQ_ACTIVE_INST_VALU"], "thread_cycles_valu": v["SQ_THREAD_CYCLES_VALU"],
                    "busy_cycles": v["SQ_BUSY_CYCLES"], "valu_issue_frac": round(issue, 3),
                    "lanes_active": round(lanes, 3), "source": os.path.relpath(a.out_md, ROOT)}
        rows.append(f"| `{k}` | {key} | {n} | {v['SQ_INSTS_VALU']:.4g} | {v['SQ_ACTIVE_INST_VALU']:.4g} | "
                    f"{v['SQ_THREAD_CYCLES_VALU']:.4g} | {v['SQ_BUSY_CYCLES']:.4g} | {issue:.2f} | {lanes:.2f} |")
    md = [f"# rocprofv3 SQ counters of the step kernels", "", "Command (run from the repository root on the MI355X box):",
          "", "```", *[c.strip() for c in a.commands.split(" ;; ") if c.strip()], "```", "",
          "Means per dispatch.  VALU issue capacity of a dispatch = SQ_BUSY_CYCLES / 32 shader engines x 1024 "
          "SIMDs / 4 cycles per wave64 VALU instruction; lanes active = SQ_THREAD_CYCLES_VALU / "
          "(64 x SQ_ACTIVE_INST_VALU).", "", *rows]
    open(a.out_md, "w").write("\n".join(md) + "\n")
    json.dump(out, open(a.out_json, "w"), indent=1)


def _sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _progress(msg):
    """Progress of long phases on stderr (stdout carries the one JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _traffic(name, avg_ms):
    """Measured HBM traffic and L2 figures of a step kernel from its committed
    rocprofv3 summary (profiles/*.json), or Nones."""
    tf = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(tf):
        return {"traffic": None, "measured_hbm_gbs": None}
    j = json.load(open(tf))
    out = {"traffic": j.get("traffic_bytes_per_launch"),
           "measured_hbm_gbs": round(j["traffic_bytes_per_launch"] / (avg_ms * 1e-3) / 1e9, 1)}
    if j.get("l2_hit_rate") is not None:
        out["l2_hit_rate"] = j["l2_hit_rate"]
        out["l2_requests_per_launch"] = j["l2_requests_per_launch"]
    return out


def _sq(key):
    """Per-dispatch SQ counters of a step kernel on a workload (committed
    rocprofv3 --pmc summary, profiles/step_kernel_sq.json), or None."""
    tf = os.path.join(ROOT, "profiles", "step_kernel_sq.json")
    if not os.path.exists(tf):
        return None
    return json.load(open(tf)).get(key)


NODE_BYTES = 64  # one 4-wide node: four quantised child boxes and the child words
LDS_WIDE_NODES = 256  # a scene rendered alone with at most this many 4-wide nodes is staged in LDS


def record_model_bytes(st, node_bytes=NODE_BYTES):
    """Algorithmic bytes summed over a counting launch under a per-segment
    model with separate ray, hit and path records: the node bytes fetched (a
    4-wide node is 64 B; a binary node pair would be 32 B), 16 B per sphere
    tested, 16 B material per hit, and 144 B of queue traffic per segment
    (ray record written and read, hit record written and read, path state
    read and written)."""
    return node_bytes * st.node_visits + 16 * st.sphere_tests + 16 * st.hits + 144 * st.segments


def valu_roofline(sq, avg_ms):
    """Useful VALU lane operations per second of a step kernel: the lanes
    active over its VALU instructions per dispatch (SQ_THREAD_CYCLES_VALU, a
    function of the workload) over the live mean dispatch time, against the
    chip's lane throughput."""
    achieved = sq["thread_cycles_valu"] / (avg_ms * 1e-3) / 1e12
    return {"bound": "valu", "achieved": round(achieved, 2), "peak": round(VALU_PEAK_TLANE, 2), "unit": "Tlane-op/s",
            "frac": round(achieved / VALU_PEAK_TLANE, 4), "valu_issue_frac": sq["valu_issue_frac"],
            "lanes_active": sq["lanes_active"], "lane_ops_per_launch": sq["thread_cycles_valu"],
            "sq_source": sq["source"]}


def _build_scenes(ctx, reqs):
    from venusaur_amd import api
    scenes, build_ms = {}, []
    for k, r in enumerate(reqs):
        if k % 200 == 0:
            _progress(f"building scenes: request {k}/{len(reqs)}")
        if r.scene_key not in scenes:
            s, m = r.scene_arrays()
            scenes[r.scene_key] = api.Scene(ctx, s, m)
            build_ms.append(scenes[r.scene_key].info().build_ms)
    return scenes, build_ms


def _rmse_fp64(r):
    """RMSE of a bound request's GPU accumulator against the FP64 oracle."""
    import oracle as O
    from venusaur_amd import workloads
    s, m = O.box_scene(r.n_small, workloads.SCENE_SEED_BASE + r.scene_key)
    cam = O.camera_lookat(list(r.lookfrom), list(r.lookat), r.vfov, float(np.float32(r.width) / np.float32(r.height)),
                          0.0, 10.0)
    p = O.make_params(r.width, r.height, r.spp, r.depth, cam)
    acc64, _ = O.render_f64(s, m, p, bvh=O.BVH(s))
    d = r.accum_host()[..., :3].astype(np.float64) - acc64[..., :3]
    return float(np.sqrt(np.mean(d * d)))


def _cheapest(reqs, k):
    return sorted(range(len(reqs)), key=lambda i: (reqs[i].paths * reqs[i].n_small))[:k]


def run_batched(ctx, n_scenes, rmse_samples=2, eager=False, eager_tuning=""):
    """Batched form vs sequential plain launches of the same 64-scene mix.
    eager: skip the autotuner and the hipGraph (default tuning, every step
    launch timed with HIP events), the form rocprofv3 can profile; adds a
    counting launch and the step kernel's algorithmic bytes."""
    from venusaur_amd import api, workloads
    from venusaur_amd.hipmem import synchronize
    reqs = workloads.batch_requests(n_scenes)
    scenes, build_ms = _build_scenes(ctx, reqs)
    for r in reqs:
        r.bind()
    items = [(scenes[r.scene_key], r.params) for r in reqs]
    synchronize()
    t0 = time.perf_counter()
    segs_plain = sum(scenes[r.scene_key].launch(r.params).segments for r in reqs)
    t_plain = time.perf_counter() - t0
    want = [_sha(r.accum_host()) for r in reqs]
    if eager:
        tuning = ctx.default_tuning()
        if eager_tuning:
            v = [int(x) for x in eager_tuning.split(",")]
            tuning.queue_capacity, tuning.tile_w, tuning.tile_h, tuning.slice_split = v[:4]
            if len(v) > 4:
                tuning.refill_lanes = v[4]
        for r in reqs:
            r.params.flags |= api.TIME_STAGES
        items = [(scenes[r.scene_key], r.params) for r in reqs]
    else:
        tuning = api.autotune(ctx, items)  # untimed: picks queue capacity and tile shape
    api.batch_launch(ctx, items, tuning=tuning)  # warm the engine and its graph
    synchronize()
    t0 = time.perf_counter()
    st = api.batch_launch(ctx, items, tuning=tuning)
    t_batch = time.perf_counter() - t0
    step = None
    if eager:
        avg_ms = st.trace_ms / st.trace_launches
        counted = []
        for r in reqs:
            c = api.LaunchParams.from_buffer_copy(r.params)
            c.flags = (c.flags & ~api.TIME_STAGES) | api.COUNT_VISITS
            counted.append((scenes[r.scene_key], c))
        cst = api.batch_launch(ctx, counted, tuning=tuning)
        per_launch = step_bytes(cst) / st.trace_launches
        step = {"kernel": "k_wf_step (multi-scene)", "launches": st.trace_launches, "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": round(per_launch),
                "algorithmic_gbs": round(per_launch / (avg_ms * 1e-3) / 1e9, 1),
                "node_visits_per_segment": round(cst.node_visits / cst.segments, 2),
                "sphere_tests_per_segment": round(cst.sphere_tests / cst.segments, 2)}
        step.update(_traffic("step_kernel_traffic_batch64.json", avg_ms))
    same = all(_sha(r.accum_host()) == w for r, w in zip(reqs, want))
    rmse64 = {f"item{k}_{reqs[k].width}px_{reqs[k].spp}spp_{reqs[k].n_small}sph": _rmse_fp64(reqs[k])
              for k in _cheapest(reqs, rmse_samples)}
    out = {"workload": f"{n_scenes} box scenes (1e3-2e5 spheres), 256-1024 px square, 4/16/64 spp, depth 16",
           "segments": int(st.segments), "plain_sequential_mseg_s": round(segs_plain / t_plain / 1e6, 1),
           "batched_mseg_s": round(st.segments / t_batch / 1e6, 1),
           "speedup_vs_plain": round(t_plain / t_batch, 2), "plain_s": round(t_plain, 3),
           "batched_s": round(t_batch, 3), "bit_identical_to_plain": same, "segments_match": segs_plain == st.segments,
           "tuning": tuning.as_dict(), "lbvh_build_ms_mean": round(float(np.mean(build_ms)), 3),
           "rmse_vs_fp64_oracle": rmse64}
    if eager:
        out["mode"] = "eager (no hipGraph, every step launch timed)"
        out["step_kernel"] = step
    for sc in scenes.values():
        sc.close()
    return out


def step_bytes(st, node_bytes=NODE_BYTES):
    """Algorithmic bytes of the step kernel (trace + miss/shade stages) for a
    counting launch: per segment the 48 B path record read, 64 B per 4-wide
    BVH node fetched, 16 B per sphere tested; per hit 36 B of sphere, material
    and type re-read by the shade stage; per surviving path the 48 B record
    written to the next queue; per ended path its 12 B radiance written to the
    pass buffer."""
    return (48 * st.segments + node_bytes * st.node_visits + 16 * st.sphere_tests + 36 * st.hits
            + 48 * (st.segments - st.paths) + 12 * st.paths)


def run_c2(ctx, frames, seed=2024):
    """Traversal-bound plain method: 1M box spheres, 1080p, 64 spp, depth 16."""
    from venusaur_amd import api
    from venusaur_amd.hipmem import DeviceBuffer, synchronize
    W, H, SPP, DEPTH, N = 1920, 1080, 64, 16, 1_000_000
    t0 = time.perf_counter()
    s, m = api.box_scene(N, seed)
    t_gen = time.perf_counter() - t0
    sc = api.Scene(ctx, s, m)
    info = sc.info()
    p = api.launch_params(W, H, SPP, DEPTH, [0, 15, -120], [0, 0, 0], 40.0, 0.0, 10.0,
                          aspect=float(np.float32(W) / np.float32(H)))
    acc, img = DeviceBuffer(W * H * 16), DeviceBuffer(W * H * 4)
    p.accum, p.image = acc.value, img.value
    sc.launch(p)  # warmup
    pc = api.LaunchParams.from_buffer_copy(p)
    pc.flags = api.COUNT_VISITS
    cst = sc.launch(pc)
    pt = api.LaunchParams.from_buffer_copy(p)
    pt.flags = api.TIME_STAGES
    synchronize()
    t0 = time.perf_counter()
    stats = [sc.launch(pt) for _ in range(frames)]
    synchronize()
    wall = time.perf_counter() - t0
    segs = sum(x.segments for x in stats)
    launches = sum(x.trace_launches for x in stats)
    avg_ms = sum(x.trace_ms for x in stats) / launches
    per_launch = step_bytes(cst) * frames / launches
    out = {"workload": f"{N} box spheres + ground (seed {seed}), {W}x{H} {SPP}spp depth {DEPTH} plain (wavefront)",
           "frames": frames, "value": round(segs / wall / 1e6, 2), "unit": "Mray-segments/s",
           "ms_per_frame": round(wall / frames * 1e3, 2), "segments_per_frame": int(stats[0].segments),
           "lbvh_build_ms": round(info.build_ms, 3), "bvh_depth": int(info.depth),
           "wide_nodes": int(info.num_wide_nodes), "wide_depth": int(info.wide_depth),
           "host_generate_s": round(t_gen, 3), "node_visits_per_segment": round(cst.node_visits / cst.segments, 2),
           "sphere_tests_per_segment": round(cst.sphere_tests / cst.segments, 2),
           "step_kernel": {"launches": launches, "avg_launch_ms": round(avg_ms, 4),
                           "algorithmic_bytes_per_launch": round(per_launch),
                           "algorithmic_gbs": round(per_launch / (avg_ms * 1e-3) / 1e9, 1),
                           "algorithmic_frac_of_hbm": round(per_launch / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                           "record_model_frac_of_hbm": round(record_model_bytes(cst) * frames / launches
                                                               / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}}
    out["step_kernel"].update(_traffic("step_kernel_traffic_c2.json", avg_ms))
    sq = _sq("box1M_1080p_64spp")
    if sq:
        out["step_kernel"]["valu"] = valu_roofline(sq, avg_ms)
    acc.free()
    img.free()
    sc.close()
    return out


def plain_capacity(ctx):
    """Requests per second that one-at-a-time plain launches sustain on the
    stream's mix (24 requests of the serving distribution)."""
    from venusaur_amd import workloads
    from venusaur_amd.hipmem import synchronize
    probe = workloads.request_stream(24, 1.0, seed=12)
    pscenes, _ = _build_scenes(ctx, probe)
    for r in probe:
        r.bind()
    synchronize()
    t0 = time.perf_counter()
    for r in probe:
        pscenes[r.scene_key].launch(r.params)
    cap = len(probe) / (time.perf_counter() - t0)
    for sc in pscenes.values():
        sc.close()
    return cap


def run_served(ctx, n_requests, cap, load=3.0, rmse_samples=3, server_tuning=None):
    """Serving form vs one-at-a-time plain launches on the same Poisson stream,
    with arrivals at `load` times the plain capacity `cap` (requests/s)."""
    import threading
    from venusaur_amd import api, workloads
    from venusaur_amd.hipmem import synchronize
    rate = load * cap
    reqs = workloads.request_stream(n_requests, rate, seed=11)
    scenes, build_ms = _build_scenes(ctx, reqs)
    for r in reqs:
        r.bind()
    synchronize()
    # plain: one request at a time in arrival order
    lat_plain, segs = [], 0
    t_start = time.perf_counter()
    for k, r in enumerate(reqs):
        if k % 100 == 0:
            _progress(f"served comparison, plain launches: {k}/{len(reqs)}")
        now = time.perf_counter() - t_start
        if now < r.arrival_s:
            time.sleep(r.arrival_s - now)
        segs += scenes[r.scene_key].launch(r.params).segments
        lat_plain.append(time.perf_counter() - t_start - r.arrival_s)
    t_plain_end = time.perf_counter() - t_start
    want = [_sha(r.accum_host()) for r in reqs]
    # served: a client thread submits at the arrival times
    # the server runs with the serving defaults (48M-ray queue, 8x8 tiles, queue
    # split 4): its latency depends on the queue size, which the batch autotuner
    # does not weigh
    srv = api.Server(ctx, hbm_budget_bytes=64 << 30, max_active=512,
                     tuning=server_tuning if server_tuning is not None else api.Tuning())
    ids = [None] * len(reqs)
    t_srv0 = time.perf_counter()  # the server's clock starts at creation (same monotonic clock)
    t_start = time.perf_counter()

    def client():
        for k, r in enumerate(reqs):
            now = time.perf_counter() - t_start
            if now < r.arrival_s:
                time.sleep(r.arrival_s - now)
            ids[k] = srv.submit(scenes[r.scene_key], r.params)
    th = threading.Thread(target=client)
    th.start()
    th.join()
    while True:  # poll until every request has finished, reporting progress
        res = [srv.poll(i) for i in ids]
        left = sum(x is None for x in res)
        if left == 0:
            break
        _progress(f"served comparison, server: {len(ids) - left}/{len(ids)} requests done")
        time.sleep(10.0)
    srv.drain()
    off = t_start - t_srv0
    lat_srv = [(x.finish_ms / 1e3 - off) - r.arrival_s for x, r in zip(res, reqs)]
    t_srv_end = max(x.finish_ms for x in res) / 1e3 - off
    _progress("served comparison: checking results against the plain launches")
    same = all(_sha(r.accum_host()) == w for r, w in zip(reqs, want))
    ok = all(x.status == 0 for x in res)
    srv.close()
    # RMSE against the FP64 oracle on the cheapest requests
    rmse64 = {f"req{k}_{reqs[k].width}px_{reqs[k].spp}spp_{reqs[k].n_small}sph": _rmse_fp64(reqs[k])
              for k in _cheapest(reqs, rmse_samples)}
    for sc in scenes.values():
        sc.close()
    pct = lambda v, q: round(float(np.percentile(v, q)) * 1e3, 1)  # noqa: E731
    return {"workload": f"{n_requests} Poisson requests at {rate:.1f}/s ({load:g}x the measured plain capacity of "
                        f"{cap:.1f}/s), 20% camera reuse of resident scenes, 5% 512x512 at 2048 spp FP64",
            "segments": int(segs), "plain_mseg_s": round(segs / t_plain_end / 1e6, 1),
            "served_mseg_s": round(segs / t_srv_end / 1e6, 1),
            "speedup_vs_plain": round(t_plain_end / t_srv_end, 2),
            "plain_p50_ms": pct(lat_plain, 50), "plain_p99_ms": pct(lat_plain, 99),
            "served_p50_ms": pct(lat_srv, 50), "served_p99_ms": pct(lat_srv, 99),
            "bit_identical_to_plain": same, "all_ok": ok, "rmse_vs_fp64_oracle": rmse64,
            "scenes_resident": len(scenes), "lbvh_build_ms_mean": round(float(np.mean(build_ms)), 3)}


def main():
    a = parse()
    if a.summarize:
        summarize(a)
        return
    if a.summarize_sq:
        summarize_sq(a)
        return
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(sharding across GPUs is designed but not built)"}))
        sys.exit(2)
    if not (os.path.exists(os.path.join(ROOT, "venusaur_amd", "libvz.so"))
            and os.path.exists(os.path.join(ROOT, "oracle", "liboracle.so"))):
        import __graft_entry__
        __graft_entry__.build()
    from venusaur_amd import api
    from venusaur_amd.hipmem import DeviceBuffer, copy_d2d, synchronize

    if a.calibrate_copy:
        nb = a.calibrate_copy << 20
        src, dst = DeviceBuffer(nb), DeviceBuffer(nb)
        copy_d2d(dst, src, nb)
        src.free()
        dst.free()

    W, H, SPP, DEPTH = a.width, a.height, a.spp, a.depth
    aspect = float(np.float32(W) / np.float32(H))
    ctx = api.Context(0)
    spheres, mats = api.book_scene()
    scene = api.Scene(ctx, spheres, mats)
    build_ms = scene.info().build_ms
    flags = api.MEGAKERNEL if a.mode == "megakernel" else 0
    p = api.launch_params(W, H, SPP, DEPTH, [13, 2, 3], [0, 0, 0], 20.0, 0.1, 10.0, aspect=aspect)
    acc = DeviceBuffer(W * H * 16)
    img = DeviceBuffer(W * H * 4)
    p.accum, p.image = acc.value, img.value

    # untimed: warmup, then one counting pass for the algorithmic bytes
    _progress(f"book scene {W}x{H} {SPP} spp: warm-up, counting frame, {a.steps} timed frames")
    for _ in range(a.warmup):
        scene.launch(p)
    pc = api.LaunchParams.from_buffer_copy(p)
    pc.flags = flags | api.COUNT_VISITS
    cst = scene.launch(pc)

    # timed region: K full renders; trace launches bracketed by HIP events
    pt = api.LaunchParams.from_buffer_copy(p)
    pt.flags = flags | api.TIME_STAGES
    synchronize()
    t0 = time.perf_counter()
    stats = [scene.launch(pt) for _ in range(a.steps)]
    synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    segs = sum(s.segments for s in stats)
    value = segs / wall / 1e6
    gpu_acc = acc.download(np.float32, (H, W, 4))

    # roofline of the step kernel (trace + miss/shade stages, the dominant
    # kernel), from the algorithmic bytes of the counting launch; a scene of
    # at most LDS_WIDE_NODES 4-wide nodes rendered alone is staged in LDS
    seg_c, nv_c, ns_c = cst.segments, cst.node_visits, cst.sphere_tests
    lds_staged = scene.info().num_wide_nodes <= LDS_WIDE_NODES
    nb = NODE_BYTES
    bytes_step = step_bytes(cst, nb)
    trace_ms = sum(s.trace_ms for s in stats)
    launches = sum(s.trace_launches for s in stats)
    roof = None
    if a.mode == "wavefront" and trace_ms > 0:
        per_launch_bytes = bytes_step * a.steps / launches
        avg_ms = trace_ms / launches
        achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
        full = (W, H, SPP, DEPTH) == (1920, 1080, 256, 50)
        # measured HBM bytes per launch of the same kernel on the same
        # workload: rocprofv3 FETCH_SIZE + WRITE_SIZE, scaled by a copy
        # kernel of known size (profiles/k_wf_step_book_1080p_256spp.md)
        meas = _traffic("step_kernel_traffic.json", avg_ms) if full else {"traffic": None, "measured_hbm_gbs": None}
        # the book scene's nodes are staged in LDS, so its node reads (64 B
        # per visit) never reach HBM: the HBM-side algorithmic bytes leave
        # them out
        lds_nodes = cst.node_visits if lds_staged else 0
        hbm_alg = (bytes_step - NODE_BYTES * lds_nodes) * a.steps / launches
        svy = record_model_bytes(cst, nb) * a.steps / launches
        # with the nodes in LDS the algorithmic bytes are not HBM bytes: no
        # fraction of the HBM peak is given for them
        hbm = {"peak_gbs": HBM_PEAK_GBS, "node_source": "lds" if lds_staged else "hbm",
               "step_bytes_per_launch": round(per_launch_bytes), "step_bytes_gbs": round(achieved, 1),
               "step_bytes_frac": None if lds_staged else round(achieved / HBM_PEAK_GBS, 4),
               "record_model_bytes_per_segment": round(record_model_bytes(cst, nb) / seg_c, 1),
               "record_model_frac": None if lds_staged else round(svy / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
               "node_bytes_from_lds_per_launch": round(NODE_BYTES * lds_nodes * a.steps / launches),
               "hbm_algorithmic_bytes_per_launch": round(hbm_alg),
               "hbm_algorithmic_frac": round(hbm_alg / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        if meas.get("measured_hbm_gbs"):
            hbm["measured_hbm_gbs"] = meas["measured_hbm_gbs"]
            hbm["measured_hbm_frac"] = round(meas["measured_hbm_gbs"] / HBM_PEAK_GBS, 4)
            for k in ("l2_hit_rate", "l2_requests_per_launch"):
                if k in meas:
                    hbm[k] = meas[k]
        sq = _sq("book_1080p_256spp") if full else None
        if sq:  # the step kernel issues VALU at capacity: that is its roof
            roof = valu_roofline(sq, avg_ms)
        else:
            roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4)}
        roof.update({"traffic": meas.get("traffic"), "kernel": "k_wf_step", "launches": launches,
                     "avg_launch_ms": round(avg_ms, 4),
                     "trace_share_of_step": round(trace_ms / (sum(s.render_ms for s in stats)), 3),
                     "bytes_per_segment": round(bytes_step / seg_c, 1),
                     "node_visits_per_segment": round(nv_c / seg_c, 2),
                     "sphere_tests_per_segment": round(ns_c / seg_c, 2), "hbm": hbm})

    # CPU baseline: the oracle's FP32 integrator on every k-th row
    cpu = None
    checks = {"gpu_segments_per_step": stats[0].segments, "paths_per_step": stats[0].paths,
              "lbvh_build_ms": round(build_ms, 3)}
    # the same frame with its host traffic: scene from host arrays (upload and
    # LBVH build), render, 8-bit image back to host.  Reported beside value,
    # which keeps inputs resident in HBM.
    synchronize()
    t0 = time.perf_counter()
    scene2 = api.Scene(ctx, spheres, mats)
    st2 = scene2.launch(p)
    host_img = img.download(np.uint8, (H, W, 4))
    t_pcie = time.perf_counter() - t0
    scene2.close()
    checks["pcie_inclusive"] = {"mray_segments_per_s": round(st2.segments / t_pcie / 1e6, 1),
                                "frame_ms": round(t_pcie * 1e3, 2),
                                "host_bytes": int(spheres.nbytes + mats.nbytes + host_img.nbytes),
                                "what": "scene from host arrays (upload + LBVH build), one frame, uchar4 image to host"}
    if not a.no_cpu:
        _progress("CPU baseline (oracle) on sampled rows")
        import oracle as O
        threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or (os.cpu_count() or 1)
        threads = min(threads, os.cpu_count() or threads)
        cam = O.camera_lookat([13, 2, 3], [0, 0, 0], 20.0, aspect, 0.1, 10.0)
        op = O.make_params(W, H, SPP, DEPTH, cam, row_step=a.cpu_row_step, threads=threads)
        ob = O.BVH(spheres)
        cacc, _, cst_cpu = O.render(spheres, mats, op, bvh=ob, want_image=False)
        rows = np.arange(0, H, a.cpu_row_step)
        cpu_rate = cst_cpu.segments / cst_cpu.seconds / 1e6
        cpu = {"value": round(cpu_rate, 3), "unit": "Mray-segments/s", "cores": cst_cpu.threads,
               "kind": "port",
               "sample": f"rows y%{a.cpu_row_step}==0 ({len(rows)} of {H} rows, full {SPP} spp, depth {DEPTH}): "
                         f"{cst_cpu.segments} segments in {cst_cpu.seconds:.1f} s",
               "gpu_over_cpu": round(value / cpu_rate, 1), "cpu_model": _cpu_model(),
               "host_cpus": os.cpu_count()}
        checks["rmse_vs_cpu_fp32_sample_rows"] = float(np.sqrt(np.mean(
            (gpu_acc[rows, :, :3].astype(np.float64) - cacc[rows, :, :3]) ** 2)))
        checks["bit_exact_sample_rows"] = bool(np.array_equal(gpu_acc[rows], cacc[rows]))
        # FP64 oracle on a thinner sample
        r64 = 64
        op64 = O.make_params(W, H, SPP, DEPTH, cam, row_step=r64, threads=threads)
        acc64, _ = O.render_f64(spheres, mats, op64, bvh=ob)
        rows64 = np.arange(0, H, r64)
        checks["rmse_vs_fp64_oracle"] = float(np.sqrt(np.mean(
            (gpu_acc[rows64, :, :3].astype(np.float64) - acc64[rows64, :, :3]) ** 2)))
        checks["rmse_cpu_fp32_vs_fp64"] = float(np.sqrt(np.mean(
            (cacc[rows64, :, :3].astype(np.float64) - acc64[rows64, :, :3]) ** 2)))
        checks["fp64_sample"] = f"rows y%{r64}==0"

    forms = {}
    if a.c2_frames:
        _progress("traversal-bound scene (1M spheres)")
        forms["traversal_bound"] = run_c2(ctx, a.c2_frames)
    if a.batch_scenes:
        _progress(f"batched form ({a.batch_scenes} scenes)")
        forms["batched"] = run_batched(ctx, a.batch_scenes, eager=a.batch_eager, eager_tuning=a.batch_tuning)
    cap = None
    if a.serve_requests:
        _progress(f"served form ({a.serve_requests} requests)")
        cap = plain_capacity(ctx)
        forms["served"] = run_served(ctx, a.serve_requests, cap, load=3.0)
    if a.serve_sustainable:
        # the same comparison below the plain capacity, where latency is not queue growth
        _progress(f"served form at 0.8x the plain capacity ({a.serve_sustainable} requests)")
        cap = cap or plain_capacity(ctx)
        forms["served_sustainable"] = run_served(ctx, a.serve_sustainable, cap, load=0.8, rmse_samples=0)

    info = ctx.info()
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mray-segments/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(wall / a.steps * 1e3, 2),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (book scene from mt19937 seed 5489)",
        "config": {"workload": f"book scene {W}x{H} {SPP}spp depth {DEPTH} plain ({a.mode})",
                   "spheres": int(len(spheres)), "width": W, "height": H, "spp": SPP,
                   "max_depth": DEPTH, "device": info["name"], "cus": info["num_cus"]},
        "roofline": roof, "cpu_baseline": cpu, "checks": checks, "forms": forms,
    }
    print(json.dumps(out), flush=True)
    scene.close()
    ctx.close()


if __name__ == "__main__":
    main()

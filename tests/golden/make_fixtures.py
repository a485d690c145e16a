"""Writes the committed golden fixtures under tests/golden/.

Run from the repository root:  python tests/golden/make_fixtures.py

Sources:
  * RNG known answers: the table computed from the reference's own
    Core/random.cuh (compiled into oracle/_ref by oracle/Makefile); when that
    library is present the script re-derives every value from it and refuses
    to write a fixture that disagrees with the oracle.
  * Book-scene sphere list, LBVH hashes, the fixed ray set (4096-ray subset
    plus SHA-256 of the full 1,048,576-ray set and of its hit array) and the
    400x225 image hash: the CPU oracle (oracle/), which the tests pin
    independently (brute force, closed forms, invariants).
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as O  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

# Values computed from the reference's random.cuh (tea<4> then three rnd()).
RNG_KAT = [
    (0, 1, 1569133783, [0.592472672, 0.00853163004, 0.544892073], 1921744396),
    (1, 1, 1546757732, [0.0165778995, 0.761493087, 0.713731587], 968275741),
    (2, 1, 1720073555, [0.78959024, 0.126846135, 0.995700359], 4127123000),
    (1919, 1, 609908746, [0.832151473, 0.363278151, 0.996959388], 2482976955),
    (2073599, 1, 848512866, [0.924306452, 0.630872965, 0.25523299], 1531008755),
    (0, 2, 1523043173, [0.482505798, 0.397382677, 0.83317101], 2161461938),
    (12345, 7, 1964180806, [0.336363614, 0.0773377419, 0.0381569266], 252298407),
]

C1 = dict(width=1920, height=1080, spp=256, max_depth=50)
N_INT, N_ISO, ISO_SEED = 786432, 262144, 20240607


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def book_camera(w, h):
    return O.camera_lookat([13, 2, 3], [0, 0, 0], 20.0, float(np.float32(w) / np.float32(h)), 0.1, 10.0)


def rng_kat():
    ref = O.ref_rng()
    rows = []
    import ctypes as C
    for v0, v1, tea, rnds, state in RNG_KAT:
        t = O.tea4(v0, v1)
        r, st = O.rnd_seq(t, 3)
        assert t == tea and st == state, (v0, v1)
        assert all(abs(a - b) < 1e-8 for a, b in zip(r, rnds))
        if ref is not None:
            assert ref.ref_tea4(v0, v1) == t
            s = C.c_uint32(t)
            rr = [ref.ref_rnd(C.byref(s)) for _ in range(3)]
            assert rr == r and s.value == st
        rows.append({"v0": v0, "v1": v1, "tea4": t, "rnd": r, "state": st})
    # a longer stream: 64 lcg outputs from tea4(7, 3)
    s = O.tea4(7, 3)
    stream, st = O.rnd_seq(s, 64)
    if ref is not None:
        cs = C.c_uint32(s)
        assert [ref.ref_rnd(C.byref(cs)) for _ in range(64)] == stream
    return {"source": "Core/random.cuh tea<4>, lcg, rnd", "rows": rows,
            "stream_seed": s, "stream": stream, "stream_state": st,
            "mt19937_5489_10000th": O.lib().vzo_mt19937_nth(5489, 10000)}


def main():
    json.dump(rng_kat(), open(os.path.join(OUT, "rng_kat.json"), "w"), indent=1)

    s, m = O.book_scene()
    np.save(os.path.join(OUT, "book_spheres.npy"), s)
    np.save(os.path.join(OUT, "book_materials.npy"), m)
    bvh = O.BVH(s)
    assert bvh.validate() == 0
    meta = {"n": int(len(s)), "nodes_sha256": sha(bvh.nodes()), "leaf_sha256": sha(bvh.leaf_prims()),
            "morton_sha256": sha(bvh.morton()), "depth": int(bvh.depth()),
            "wide_nodes_sha256": sha(bvh.wide_nodes()), "wide_nodes": int(len(bvh.wide_nodes())),
            "wide_depth": int(bvh.wide_depth())}

    cam = book_camera(C1["width"], C1["height"])
    p = O.make_params(C1["width"], C1["height"], C1["spp"], C1["max_depth"], cam)
    rays = O.make_rayset(s, m, p, N_INT, N_ISO, ISO_SEED, bvh=bvh)
    hits, nv, ns = bvh.trace(rays)
    hits_b, nv_b, ns_b = bvh.trace_binary(rays)
    assert hits_b.tobytes() == hits.tobytes()
    sub = np.arange(0, len(rays), 256)
    brute = O.trace_brute(s, rays[sub])
    assert brute.tobytes() == hits[sub].tobytes()
    np.save(os.path.join(OUT, "rayset_4096.npy"), rays[sub])
    np.save(os.path.join(OUT, "rayset_4096_hits.npy"), hits[sub])
    meta["rayset"] = {"n_int": N_INT, "n_iso": N_ISO, "iso_seed": ISO_SEED, "config": C1,
                      "rays_sha256": sha(rays), "hits_sha256": sha(hits), "node_visits": int(nv),
                      "sphere_tests": int(ns), "hit_count": int((hits["prim"] != O.MISS).sum()),
                      "binary_node_visits": int(nv_b), "binary_sphere_tests": int(ns_b)}

    cam0 = book_camera(400, 225)
    p0 = O.make_params(400, 225, 16, 50, cam0)
    acc, img, st = O.render(s, m, p0, bvh=bvh)
    meta["c0"] = {"accum_sha256": sha(acc), "segments": int(st.segments), "paths": int(st.paths),
                  "mean_rgb": [float(x) for x in acc[..., :3].reshape(-1, 3).mean(0)]}
    json.dump(meta, open(os.path.join(OUT, "book_fixture.json"), "w"), indent=1)
    print(json.dumps(meta, indent=1))


if __name__ == "__main__":
    main()

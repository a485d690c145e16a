"""Writes box_grazing_rays.npy / box_grazing.json: five rays on which the
float32 sphere test, ungated, reports a grazing "hit" on a tiny distant
sphere that the ray does not enter (the rounding of |oc|^2 - r^2 exceeds
r^2), on the random-box scene with 20000 spheres + ground, seed 2024.

The rays were emitted by the integrator on that scene (160x90, 64 spp,
depth 16, camera (0,15,-120) -> origin, vfov 40; make_rayset with
n_int=300000, n_iso=50000, iso_seed=77) and picked out as the rays on which
LBVH traversal and brute force disagreed while the sphere test was not
gated by the sphere's own box.  With the gate the integrator emits other
rays, so they are kept here as literal float32 bit patterns
(ox, oy, oz, tmin, dx, dy, dz, tmax).  The expected (t, prim) are the
oracle's gated closest hits, asserted equal for its BVH and brute force.
Run from the repository root after `make -C oracle`.
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

RAY_BITS = [
    [0x41c33fa2, 0xc0714a30, 0xc2a67f9e, 0x3a83126f, 0xbe8d5086, 0x3d34cc20, 0xbcd4f240, 0x5a0e1bca],
    [0x41084b3b, 0x414b8688, 0xc2a186e4, 0x3a83126f, 0xbeb881c4, 0xbd097b70, 0x3f8280e4, 0x5a0e1bca],
    [0xc1c26e7e, 0x40c49142, 0xc2bc04c6, 0x3a83126f, 0x3e8e7d0d, 0x3d9820b0, 0x3ff2d1ac, 0x5a0e1bca],
    [0x42a9ca8a, 0x419ea165, 0xc28eee42, 0x3a83126f, 0xbf7defad, 0xbdbb5d45, 0xbdb393c0, 0x5a0e1bca],
    [0x403f49e0, 0x41012ee1, 0x425af828, 0x3a83126f, 0x3f5c98fd, 0xbd18cab6, 0x3f018b21, 0x5a0e1bca],
]
SPURIOUS_PRIM = [13969, 19967, 16941, 2250, 704]  # the ungated brute-force "hits"


def main():
    s, _ = O.box_scene(20000, 2024)
    rays = np.frombuffer(np.array(RAY_BITS, np.uint32).tobytes(), O.RAY_DT).copy()
    hb, _, _ = O.BVH(s).trace(rays)
    hf = O.trace_brute(s, rays)
    assert hb.tobytes() == hf.tobytes()
    np.save(os.path.join(HERE, "box_grazing_rays.npy"), rays)
    meta = {"scene": {"n_small": 20000, "seed": 2024}, "spurious_prim": SPURIOUS_PRIM,
            "t": [float(x) for x in hb["t"]], "prim": [int(x) for x in hb["prim"]]}
    json.dump(meta, open(os.path.join(HERE, "box_grazing.json"), "w"), indent=1)
    print(json.dumps(meta))


if __name__ == "__main__":
    main()

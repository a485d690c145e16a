"""Writes tests/golden/wire_request_small.vzrq: one wire-format request record
with a three-sphere inline scene, encoded by libvz.so.  tests/test_wire.py
checks the fixture field by field against the layout documented in
include/vz_api.h, with its own decoder and FNV-1a, and that encoding the same
request again gives the same bytes.

    python tests/golden/make_wire.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from venusaur_amd import api  # noqa: E402

SPHERES = np.array([(0.0, -1000.0, 0.0, 1000.0), (0.0, 1.0, 0.0, 1.0), (-4.0, 1.0, 0.5, -0.95)], api.SPHERE_DT)
MATERIALS = np.array([(0, (0.5, 0.5, 0.5), 0.0, 0.0), (2, (0.0, 0.0, 0.0), 0.0, 1.5),
                      (1, (0.7, 0.6, 0.5), 0.25, 0.0)], api.MATERIAL_DT)


def desc():
    return api.request_desc(0x0123456789ABCDEF, 42, 320, 180, 16, 12, [13.0, 2.0, 3.0], [0.0, 0.5, 0.0], 20.0,
                            0.1, 10.0, flags=api.ACCUM_FP64, subframe=3, arrival_ms=1234.5)


if __name__ == "__main__":
    rec = api.encode_request(desc(), SPHERES, MATERIALS)
    with open(os.path.join(HERE, "wire_request_small.vzrq"), "wb") as f:
        f.write(rec)
    print(len(rec), "bytes")

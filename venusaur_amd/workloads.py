"""Synthetic workloads of the batched and serving forms.

Batch: N scenes from the random-box generator with sphere counts
log-uniform in [1e3, 2e5], square resolutions from {256, 512, 768, 1024},
samples per pixel from {4, 16, 64}, max depth 16, camera (0,15,-120) looking
at the origin, vfov 40, aperture 0.

Stream: requests from the batch distribution with Poisson arrivals; a
fraction reuses an already resident scene with a random camera (lookfrom on
a sphere of radius U[15,150], vfov U[20,60]); a fraction are 512x512 at
2048 samples per pixel with FP64 accumulation.

Scene of a request: the random-box generator with seed SCENE_SEED_BASE +
scene_key.  write_stream() turns requests into wire-format request records
(include/vz_api.h), each scene inline in the first record that names it.
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import api
from .hipmem import DeviceBuffer

BOX_FROM = (0.0, 15.0, -120.0)
BOX_AT = (0.0, 0.0, 0.0)
SCENE_SEED_BASE = 1000


@dataclass
class Request:
    scene_key: int
    n_small: int
    width: int
    height: int
    spp: int
    depth: int
    lookfrom: tuple
    lookat: tuple
    vfov: float
    fp64: bool = False
    arrival_s: float = 0.0
    reuse: bool = False
    params: object = field(default=None, repr=False)
    accum: object = field(default=None, repr=False)
    image: object = field(default=None, repr=False)

    @property
    def paths(self):
        return self.width * self.height * self.spp

    def bind(self):
        """Allocates the request's accum/image buffers and launch params."""
        npx = self.width * self.height
        self.accum = DeviceBuffer(npx * (32 if self.fp64 else 16))
        self.image = DeviceBuffer(npx * 4)
        p = api.launch_params(self.width, self.height, self.spp, self.depth, list(self.lookfrom), list(self.lookat),
                              self.vfov, 0.0, 10.0, flags=api.ACCUM_FP64 if self.fp64 else 0,
                              aspect=float(np.float32(self.width) / np.float32(self.height)))
        p.accum, p.image = self.accum.value, self.image.value
        self.params = p
        return p

    def accum_host(self):
        return self.accum.download(np.float64 if self.fp64 else np.float32, (self.height, self.width, 4))

    def scene_arrays(self):
        return api.box_scene(self.n_small, SCENE_SEED_BASE + self.scene_key)

    def desc(self, tag):
        """The request as a wire-format vzRequestDesc (aperture 0, focus 10, subframe 1)."""
        return api.request_desc(tag, self.scene_key, self.width, self.height, self.spp, self.depth,
                                self.lookfrom, self.lookat, self.vfov, 0.0, 10.0,
                                flags=api.ACCUM_FP64 if self.fp64 else 0, arrival_ms=self.arrival_s * 1e3)


def _scene_size(rng):
    return int(round(math.exp(rng.uniform(math.log(1e3), math.log(2e5)))))


def batch_requests(n, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        res = int(rng.choice([256, 512, 768, 1024]))
        out.append(Request(scene_key=i, n_small=_scene_size(rng), width=res, height=res,
                           spp=int(rng.choice([4, 16, 64])), depth=16, lookfrom=BOX_FROM, lookat=BOX_AT, vfov=40.0))
    return out


def _random_camera(rng):
    r = rng.uniform(15.0, 150.0)
    v = rng.normal(size=3)
    v[1] = abs(v[1]) + 0.05  # stay above the ground sphere
    v = v / np.linalg.norm(v) * r
    return (float(v[0]), float(v[1]), float(v[2])), float(rng.uniform(20.0, 60.0))


def request_stream(n, rate_hz, seed=11, reuse_frac=0.2, fp64_frac=0.05):
    """n requests with Poisson arrivals at rate_hz (arrival_s ascending)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out, resident = [], []
    for i in range(n):
        t += rng.exponential(1.0 / rate_hz)
        u = rng.uniform()
        if u < fp64_frac:
            req = Request(scene_key=i, n_small=_scene_size(rng), width=512, height=512, spp=2048, depth=16,
                          lookfrom=BOX_FROM, lookat=BOX_AT, vfov=40.0, fp64=True)
        elif u < fp64_frac + reuse_frac and resident:
            base = resident[int(rng.integers(len(resident)))]
            frm, vfov = _random_camera(rng)
            res = int(rng.choice([256, 512, 768, 1024]))
            req = Request(scene_key=base.scene_key, n_small=base.n_small, width=res, height=res,
                          spp=int(rng.choice([4, 16, 64])), depth=16, lookfrom=frm, lookat=BOX_AT, vfov=vfov,
                          reuse=True)
        else:
            res = int(rng.choice([256, 512, 768, 1024]))
            req = Request(scene_key=i, n_small=_scene_size(rng), width=res, height=res,
                          spp=int(rng.choice([4, 16, 64])), depth=16, lookfrom=BOX_FROM, lookat=BOX_AT, vfov=40.0)
        req.arrival_s = t
        out.append(req)
        if not req.reuse and not req.fp64:
            resident.append(req)
    return out


def write_stream(path, requests):
    """Writes requests as wire-format request records (tag = index in the
    list); a scene goes inline in the first record that names its key.
    Returns the number of bytes written."""
    seen = set()
    total = 0
    with open(path, "wb") as f:
        for i, r in enumerate(requests):
            if r.scene_key in seen:
                rec = api.encode_request(r.desc(i))
            else:
                s, m = r.scene_arrays()
                rec = api.encode_request(r.desc(i), s, m)
                seen.add(r.scene_key)
            f.write(rec)
            total += len(rec)
    return total

"""Shared helpers for the parity tests: oracle/GPU renders of one config."""
import hashlib

import numpy as np

import oracle as O
from venusaur_amd import api
from venusaur_amd.hipmem import DeviceBuffer


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def book_lookat(w, h, aperture=0.1, focus=10.0):
    return dict(lookfrom=[13, 2, 3], lookat=[0, 0, 0], vfov=20.0, aperture=aperture, focus=focus,
                aspect=float(np.float32(w) / np.float32(h)))


def box_lookat(w, h):
    return dict(lookfrom=[0, 15, -120], lookat=[0, 0, 0], vfov=40.0, aperture=0.0, focus=10.0,
                aspect=float(np.float32(w) / np.float32(h)))


def oracle_params(w, h, spp, depth, cam, subframe=1, accum_fp64=False, row_step=1, binary_tree=False):
    c = O.camera_lookat(cam["lookfrom"], cam["lookat"], cam["vfov"], cam["aspect"],
                        cam["aperture"], cam["focus"])
    return O.make_params(w, h, spp, depth, c, subframe=subframe, accum_fp64=accum_fp64,
                         row_step=row_step, binary_tree=binary_tree)


def gpu_params(w, h, spp, depth, cam, subframe=1, flags=0):
    return api.launch_params(w, h, spp, depth, cam["lookfrom"], cam["lookat"], cam["vfov"],
                             cam["aperture"], cam["focus"], subframe=subframe, flags=flags,
                             aspect=cam["aspect"])


class GpuTarget:
    """accum + image device buffers for one request."""

    def __init__(self, w, h, fp64=False):
        self.w, self.h, self.fp64 = w, h, fp64
        self.accum = DeviceBuffer(w * h * (32 if fp64 else 16))
        self.image = DeviceBuffer(w * h * 4)
        self.accum.zero()
        self.image.zero()

    def bind(self, params):
        params.accum = self.accum.value
        params.image = self.image.value
        return params

    def accum_host(self):
        return self.accum.download(np.float64 if self.fp64 else np.float32, (self.h, self.w, 4))

    def image_host(self):
        return self.image.download(np.uint8, (self.h, self.w, 4))


def geometric_scene(n):
    """n spheres with centres at x = 2^-min(i, 125) (y = z = 0.5) and radii
    max(x / 100, 1e-35), all normal floats: a binned-SAH split can only peel
    a few of them off at a time, and the last ones share one centre."""
    s, m = O.box_scene(n - 1, 3)
    x = np.float32(2.0) ** -np.minimum(np.arange(n), 125).astype(np.float32)
    s["cx"] = x
    s["cy"] = np.float32(0.5)
    s["cz"] = np.float32(0.5)
    s["r"] = np.maximum(x * np.float32(0.01), np.float32(1e-35))
    return s, m


def rmse(a, b):
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    return float(np.sqrt(np.mean(d * d)))

"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  Nothing in the product
package (venusaur_amd/) imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_RNG_PATH = os.path.join(HERE, "_ref", "libref_rng.so")

SPHERE_DT = np.dtype([("cx", "<f4"), ("cy", "<f4"), ("cz", "<f4"), ("r", "<f4")])
MATERIAL_DT = np.dtype([("type", "<u4"), ("albedo", "<f4", 3), ("fuzz", "<f4"), ("ir", "<f4")])
RAY_DT = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                   ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT_DT = np.dtype([("t", "<f4"), ("prim", "<u4")])
NODE_DT = np.dtype([("o", "<f4", 3), ("meta", "<u4"), ("q", "<u4", 3), ("child", "<u4")])
WNODE_DT = np.dtype([("o", "<f4", 3), ("meta", "<u4"), ("qlo", "<u4", 3), ("qhi", "<u4", 3),
                     ("child", "<u4", 4), ("pad", "<u4", 2)])
MISS = 0xFFFFFFFF
BUILD_LBVH, BUILD_SAH = 0, 1
SAH_MAX_PRIMS = 1024
TREELET_MIN_PRIMS = 131072


class Camera(C.Structure):
    _fields_ = [("origin", C.c_float * 3), ("u", C.c_float * 3), ("v", C.c_float * 3),
                ("w", C.c_float * 3), ("lens_radius", C.c_float)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32),
                ("subframe", C.c_uint32), ("max_depth", C.c_uint32), ("cam", Camera),
                ("row_start", C.c_uint32), ("row_step", C.c_uint32), ("threads", C.c_uint32),
                ("accum_fp64", C.c_uint32), ("binary_tree", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("paths", C.c_uint64), ("hits", C.c_uint64),
                ("node_visits", C.c_uint64), ("sphere_tests", C.c_uint64),
                ("seconds", C.c_double), ("threads", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        vp, u32, f32 = C.c_void_p, C.c_uint32, C.c_float
        sig = {
            "vzo_tea4": (u32, [u32, u32]),
            "vzo_lcg": (u32, [C.POINTER(u32)]),
            "vzo_rnd": (f32, [C.POINTER(u32)]),
            "vzo_mt19937_nth": (u32, [u32, u32]),
            "vzo_book_scene": (u32, [u32, vp, vp, u32]),
            "vzo_box_scene": (None, [u32, u32, vp, vp]),
            "vzo_camera_lookat": (None, [vp, vp, f32, f32, f32, f32, C.POINTER(Camera)]),
            "vzo_hit_sphere_f32": (C.c_int, [vp, vp, f32, f32, vp, C.POINTER(f32)]),
            "vzo_hit_sphere_f64": (C.c_int, [vp, vp, C.c_double, C.c_double, vp,
                                             C.POINTER(C.c_double)]),
            "vzo_bvh_build": (vp, [vp, u32]),
            "vzo_bvh_free": (None, [vp]),
            "vzo_bvh_num_prims": (u32, [vp]),
            "vzo_bvh_num_nodes": (u32, [vp]),
            "vzo_bvh_nodes": (vp, [vp]),
            "vzo_bvh_leaf_prims": (vp, [vp]),
            "vzo_bvh_morton": (vp, [vp]),
            "vzo_bvh_node_boxes": (vp, [vp]),
            "vzo_bvh_depth": (u32, [vp]),
            "vzo_bvh_num_wide_nodes": (u32, [vp]),
            "vzo_bvh_wide_nodes": (vp, [vp]),
            "vzo_bvh_wide_depth": (u32, [vp]),
            "vzo_bvh_builder": (u32, [vp]),
            "vzo_wide_box_check": (None, [vp, vp, vp, f32, f32, vp, vp, vp]),
            "vzo_trace_binary": (None, [vp, vp, vp, u32, vp, C.POINTER(C.c_uint64),
                                        C.POINTER(C.c_uint64)]),
            "vzo_bvh_validate": (C.c_int, [vp, vp]),
            "vzo_trace": (None, [vp, vp, vp, u32, vp, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint64)]),
            "vzo_trace_brute": (None, [vp, u32, vp, u32, vp]),
            "vzo_trace_brute_f64": (None, [vp, u32, vp, u32, vp, vp]),
            "vzo_scatter_f32": (C.c_int, [vp, vp, vp, C.c_int, C.POINTER(u32), vp]),
            "vzo_sky_f32": (None, [vp, vp]),
            "vzo_render": (None, [vp, vp, vp, u32, C.POINTER(RenderParams), vp, vp,
                                  C.POINTER(Stats)]),
            "vzo_render_f64": (None, [vp, vp, vp, u32, C.POINTER(RenderParams), vp,
                                      C.POINTER(Stats)]),
            "vzo_make_rayset": (None, [vp, vp, vp, u32, C.POINTER(RenderParams), u32, u32, u32,
                                       vp]),
            "vzo_encode8": (C.c_uint8, [f32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ref_rng():
    """The reference's own random.cuh compiled into oracle/_ref, or None."""
    if not os.path.exists(REF_RNG_PATH):
        return None
    L = C.CDLL(REF_RNG_PATH)
    L.ref_tea4.restype = C.c_uint32
    L.ref_tea4.argtypes = [C.c_uint32, C.c_uint32]
    L.ref_lcg.restype = C.c_uint32
    L.ref_lcg.argtypes = [C.POINTER(C.c_uint32)]
    L.ref_rnd.restype = C.c_float
    L.ref_rnd.argtypes = [C.POINTER(C.c_uint32)]
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def tea4(a, b):
    return lib().vzo_tea4(a, b)


def rnd_seq(seed, n):
    s = C.c_uint32(seed)
    out = [lib().vzo_rnd(C.byref(s)) for _ in range(n)]
    return out, s.value


def book_scene(seed=5489):
    cap = 1024
    s = np.zeros(cap, SPHERE_DT)
    m = np.zeros(cap, MATERIAL_DT)
    n = lib().vzo_book_scene(seed, _p(s), _p(m), cap)
    return s[:n].copy(), m[:n].copy()


def box_scene(n_small, seed):
    s = np.zeros(n_small + 1, SPHERE_DT)
    m = np.zeros(n_small + 1, MATERIAL_DT)
    lib().vzo_box_scene(n_small, seed, _p(s), _p(m))
    return s, m


def camera_lookat(lookfrom, lookat, vfov, aspect, aperture, focus):
    cam = Camera()
    f = np.asarray(lookfrom, np.float32)
    a = np.asarray(lookat, np.float32)
    lib().vzo_camera_lookat(_p(f), _p(a), vfov, aspect, aperture, focus, C.byref(cam))
    return cam


def camera_dict(cam):
    return {"origin": list(cam.origin), "u": list(cam.u), "v": list(cam.v), "w": list(cam.w),
            "lens_radius": cam.lens_radius}


class BVH:
    def __init__(self, spheres):
        self.spheres = np.ascontiguousarray(spheres)
        self.h = lib().vzo_bvh_build(_p(self.spheres), len(self.spheres))
        if not self.h:
            raise RuntimeError("bvh build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().vzo_bvh_free(self.h)
            self.h = None

    @property
    def n(self):
        return lib().vzo_bvh_num_prims(self.h)

    def _arr(self, fn, dtype, count):
        ptr = fn(self.h)
        if count == 0:
            return np.zeros(0, dtype)
        buf = (C.c_uint8 * (count * np.dtype(dtype).itemsize)).from_address(ptr)
        return np.frombuffer(buf, dtype).copy()

    def nodes(self):
        return self._arr(lib().vzo_bvh_nodes, NODE_DT, lib().vzo_bvh_num_nodes(self.h))

    def leaf_prims(self):
        return self._arr(lib().vzo_bvh_leaf_prims, np.uint32, self.n)

    def morton(self):
        return self._arr(lib().vzo_bvh_morton, np.uint32, self.n)

    def node_boxes(self):
        return self._arr(lib().vzo_bvh_node_boxes, np.float32,
                         lib().vzo_bvh_num_nodes(self.h) * 6).reshape(-1, 6)

    def depth(self):
        return lib().vzo_bvh_depth(self.h)

    def wide_nodes(self):
        return self._arr(lib().vzo_bvh_wide_nodes, WNODE_DT, lib().vzo_bvh_num_wide_nodes(self.h))

    def wide_depth(self):
        return lib().vzo_bvh_wide_depth(self.h)

    def builder(self):
        """BUILD_LBVH, or BUILD_SAH for 2 to 1024 spheres."""
        return lib().vzo_bvh_builder(self.h)

    def validate(self):
        return lib().vzo_bvh_validate(self.h, _p(self.spheres))

    def trace(self, rays):
        """Closest hits over the 4-wide tree: (hits, wide node visits, sphere tests)."""
        rays = np.ascontiguousarray(rays, RAY_DT)
        hits = np.zeros(len(rays), HIT_DT)
        nv, ns = C.c_uint64(0), C.c_uint64(0)
        lib().vzo_trace(self.h, _p(self.spheres), _p(rays), len(rays), _p(hits), C.byref(nv),
                        C.byref(ns))
        return hits, nv.value, ns.value

    def trace_binary(self, rays):
        """Closest hits over the binary tree: (hits, node visits, sphere tests)."""
        rays = np.ascontiguousarray(rays, RAY_DT)
        hits = np.zeros(len(rays), HIT_DT)
        nv, ns = C.c_uint64(0), C.c_uint64(0)
        lib().vzo_trace_binary(self.h, _p(self.spheres), _p(rays), len(rays), _p(hits), C.byref(nv),
                               C.byref(ns))
        return hits, nv.value, ns.value


def trace_brute(spheres, rays):
    rays = np.ascontiguousarray(rays, RAY_DT)
    spheres = np.ascontiguousarray(spheres)
    hits = np.zeros(len(rays), HIT_DT)
    lib().vzo_trace_brute(_p(spheres), len(spheres), _p(rays), len(rays), _p(hits))
    return hits


def trace_brute_f64(spheres, rays):
    rays = np.ascontiguousarray(rays, RAY_DT)
    spheres = np.ascontiguousarray(spheres)
    t = np.zeros(len(rays), np.float64)
    p = np.zeros(len(rays), np.uint32)
    lib().vzo_trace_brute_f64(_p(spheres), len(spheres), _p(rays), len(rays), _p(t), _p(p))
    return t, p


def make_params(width, height, spp, max_depth, cam, subframe=1, row_start=0, row_step=1,
                threads=0, accum_fp64=False, binary_tree=False):
    """binary_tree: count visits over the binary tree (what the GPU walks for
    a scene of at most 513 spheres rendered alone); hits are the same."""
    p = RenderParams()
    p.width, p.height, p.spp, p.subframe, p.max_depth = width, height, spp, subframe, max_depth
    p.cam = cam
    p.row_start, p.row_step, p.threads = row_start, row_step, threads
    p.accum_fp64 = 1 if accum_fp64 else 0
    p.binary_tree = 1 if binary_tree else 0
    return p


def render(spheres, materials, params, bvh=None, accum=None, want_image=True):
    """FP32 integrator.  Returns (accum[H,W,4], image[H,W,4] or None, stats)."""
    npx = params.width * params.height
    if accum is None:
        accum = np.zeros((params.height, params.width, 4),
                         np.float64 if params.accum_fp64 else np.float32)
    img = np.zeros((params.height, params.width, 4), np.uint8) if want_image else None
    st = Stats()
    spheres = np.ascontiguousarray(spheres)
    materials = np.ascontiguousarray(materials)
    assert accum.size == npx * 4
    lib().vzo_render(bvh.h if bvh is not None else None, _p(spheres), _p(materials),
                     len(spheres), C.byref(params), _p(accum),
                     _p(img) if img is not None else None, C.byref(st))
    return accum, img, st


def render_f64(spheres, materials, params, bvh=None, accum=None):
    if accum is None:
        accum = np.zeros((params.height, params.width, 4), np.float64)
    st = Stats()
    spheres = np.ascontiguousarray(spheres)
    materials = np.ascontiguousarray(materials)
    lib().vzo_render_f64(bvh.h if bvh is not None else None, _p(spheres), _p(materials),
                         len(spheres), C.byref(params), _p(accum), C.byref(st))
    return accum, st


def make_rayset(spheres, materials, params, n_int, n_iso, iso_seed, bvh=None):
    out = np.zeros(n_int + n_iso, RAY_DT)
    spheres = np.ascontiguousarray(spheres)
    materials = np.ascontiguousarray(materials)
    lib().vzo_make_rayset(bvh.h if bvh is not None else None, _p(spheres), _p(materials),
                          len(spheres), C.byref(params), n_int, n_iso, iso_seed, _p(out))
    return out


def encode8(c):
    return lib().vzo_encode8(c)


def scatter(material, d, n, front, seed):
    m = np.zeros(1, MATERIAL_DT)
    m[0] = material
    d = np.asarray(d, np.float32)
    n = np.asarray(n, np.float32)
    out = np.zeros(3, np.float32)
    s = C.c_uint32(seed)
    ok = lib().vzo_scatter_f32(_p(m), _p(d), _p(n), int(front), C.byref(s), _p(out))
    return bool(ok), out, s.value


def sky(d):
    d = np.asarray(d, np.float32)
    out = np.zeros(3, np.float32)
    lib().vzo_sky_f32(_p(d), _p(out))
    return out


def wide_box_check(node, o, d, tmin, tmax):
    """Both slot-box tests of one wide node (a WNODE_DT record) for one ray:
    ((fast mask, exact mask), fast entry distances, exact entry distances).
    The traversal's test (fast) must accept every slot the exact-form robust
    test accepts, with an entry distance no larger."""
    node = np.ascontiguousarray(np.asarray(node, WNODE_DT).reshape(1))
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    masks = np.zeros(2, np.uint32)
    tf, te = np.zeros(4, np.float32), np.zeros(4, np.float32)
    lib().vzo_wide_box_check(_p(node), _p(o), _p(d), float(tmin), float(tmax), _p(masks), _p(tf), _p(te))
    return (int(masks[0]), int(masks[1])), tf, te

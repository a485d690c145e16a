"""ctypes binding of libvz.so (include/vz_api.h).

Host-side mirror of the reference's Renderer / Scene / Camera surface for
Python callers (tests, bench).  There is no CPU fallback: if libvz.so is
missing or no gfx950 device is visible, calls raise VzError.
"""
import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvz.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "vz_api.h")

VZ_SUCCESS = 0
STATUS_NAMES = {0: "VZ_SUCCESS", 1: "VZ_ERROR_INVALID_VALUE", 2: "VZ_ERROR_OUT_OF_MEMORY",
                3: "VZ_ERROR_HIP", 4: "VZ_ERROR_NOT_READY", 5: "VZ_ERROR_CAPACITY",
                6: "VZ_ERROR_UNSUPPORTED", 7: "VZ_ERROR_INTERNAL", 8: "VZ_ERROR_NO_DEVICE"}
LAMBERTIAN, METAL, DIELECTRIC = 0, 1, 2
MISS = 0xFFFFFFFF
ACCUM_FP64 = 0x1
MEGAKERNEL = 0x2
COUNT_VISITS = 0x4
TIME_STAGES = 0x8

SPHERE_DT = np.dtype([("cx", "<f4"), ("cy", "<f4"), ("cz", "<f4"), ("r", "<f4")])
MATERIAL_DT = np.dtype([("type", "<u4"), ("albedo", "<f4", 3), ("fuzz", "<f4"), ("ir", "<f4")])
RAY_DT = np.dtype([("ox", "<f4"), ("oy", "<f4"), ("oz", "<f4"), ("tmin", "<f4"),
                   ("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("tmax", "<f4")])
HIT_DT = np.dtype([("t", "<f4"), ("prim", "<u4")])
NODE_DT = np.dtype([("o", "<f4", 3), ("meta", "<u4"), ("q", "<u4", 3), ("child", "<u4")])
WNODE_DT = np.dtype([("o", "<f4", 3), ("meta", "<u4"), ("qlo", "<u4", 3), ("qhi", "<u4", 3),
                     ("child", "<u4", 4), ("pad", "<u4", 2)])


class VzError(RuntimeError):
    def __init__(self, status, msg):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {msg}")


class LaunchParams(C.Structure):
    _fields_ = [("image", C.c_void_p), ("accum", C.c_void_p), ("width", C.c_uint32),
                ("height", C.c_uint32), ("samples_per_pixel", C.c_uint32),
                ("subframe_index", C.c_uint32), ("origin", C.c_float * 3), ("u", C.c_float * 3),
                ("v", C.c_float * 3), ("w", C.c_float * 3), ("lens_radius", C.c_float),
                ("max_depth", C.c_uint32), ("flags", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("paths", C.c_uint64), ("node_visits", C.c_uint64),
                ("sphere_tests", C.c_uint64), ("render_ms", C.c_double), ("trace_ms", C.c_double),
                ("trace_launches", C.c_uint32), ("iterations", C.c_uint32), ("hits", C.c_uint64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class AccelInfo(C.Structure):
    _fields_ = [("num_prims", C.c_uint32), ("num_nodes", C.c_uint32), ("depth", C.c_uint32),
                ("node_bytes", C.c_uint32), ("device_bytes", C.c_uint64), ("build_ms", C.c_double),
                ("num_wide_nodes", C.c_uint32), ("wide_depth", C.c_uint32), ("builder", C.c_uint32)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("params", LaunchParams)]


class Tuning(C.Structure):
    _fields_ = [("queue_capacity", C.c_uint32), ("tile_w", C.c_uint32), ("tile_h", C.c_uint32),
                ("iters_per_graph", C.c_uint32), ("refill_lanes", C.c_uint32), ("slice_split", C.c_uint32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_active", C.c_uint32), ("tuning", Tuning)]


class RequestDesc(C.Structure):
    """vzRequestDesc: one render request of the wire format."""
    _fields_ = [("tag", C.c_uint64), ("scene_key", C.c_uint64), ("width", C.c_uint32), ("height", C.c_uint32),
                ("samples_per_pixel", C.c_uint32), ("max_depth", C.c_uint32), ("flags", C.c_uint32),
                ("subframe_index", C.c_uint32), ("lookfrom", C.c_float * 3), ("lookat", C.c_float * 3),
                ("vfov_deg", C.c_float), ("aperture", C.c_float), ("focus_dist", C.c_float),
                ("arrival_ms", C.c_float)]


class RequestResult(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("submit_ms", C.c_double), ("admit_ms", C.c_double),
                ("finish_ms", C.c_double), ("segments", C.c_uint64), ("status", C.c_uint32)]


_lib = None


def _declare(L):
    vp, u32, u64, f32, st = C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_int
    P = C.POINTER
    sig = {
        "vzCreate": (st, [C.c_int, P(vp)]),
        "vzDestroy": (st, [vp]),
        "vzDeviceInfo": (st, [vp, P(u32), P(u64), C.c_char_p, C.c_size_t]),
        "vzSceneCreate": (st, [vp, vp, vp, u32, vp, P(vp)]),
        "vzSceneCreateDevice": (st, [vp, vp, vp, u32, vp, P(vp)]),
        "vzSceneUpdateSpheresDevice": (st, [vp, vp, vp]),
        "vzSceneDestroy": (st, [vp]),
        "vzAccelBuild": (st, [vp, vp]),
        "vzSceneUpdateSpheres": (st, [vp, vp, vp]),
        "vzAccelRefit": (st, [vp, vp]),
        "vzAccelInfoGet": (st, [vp, P(AccelInfo)]),
        "vzAccelDownload": (st, [vp, vp, vp, vp, vp]),
        "vzAccelDownloadWide": (st, [vp, vp, vp]),
        "vzTraceClosest": (st, [vp, vp, u32, vp, vp, P(Stats)]),
        "vzTraceClosestNoAlloc": (st, [vp, vp, u32, vp, vp, P(Stats)]),
        "vzLaunch": (st, [vp, P(LaunchParams), vp, P(Stats)]),
        "vzBatchLaunch": (st, [vp, P(BatchItem), u32, P(Tuning), vp, P(Stats)]),
        "vzAutotune": (st, [vp, P(BatchItem), u32, vp, P(Tuning)]),
        "vzDefaultTuning": (st, [vp, P(Tuning)]),
        "vzServerCreate": (st, [vp, P(ServeConfig), P(vp)]),
        "vzServerSubmit": (st, [vp, vp, P(LaunchParams), P(u64)]),
        "vzServerPoll": (st, [vp, u64, P(RequestResult)]),
        "vzServerDrain": (st, [vp]),
        "vzServerDestroy": (st, [vp]),
        "vzRequestFootprint": (u64, [P(LaunchParams)]),
        "vzGenerateBookScene": (st, [u32, vp, vp, u32, P(u32)]),
        "vzGenerateBoxScene": (st, [u32, u32, vp, vp]),
        "vzCameraLookAt": (st, [vp, vp, f32, f32, f32, f32, P(LaunchParams)]),
        "vzWireSceneSize": (C.c_size_t, [u32]),
        "vzWireEncodeScene": (st, [vp, vp, u32, vp, C.c_size_t, P(C.c_size_t)]),
        "vzWireDecodeScene": (st, [vp, C.c_size_t, vp, vp, u32, P(u32), P(C.c_size_t)]),
        "vzWireEncodeRequest": (st, [P(RequestDesc), vp, vp, u32, vp, C.c_size_t, P(C.c_size_t)]),
        "vzWireDecodeRequest": (st, [vp, C.c_size_t, P(RequestDesc), P(C.c_size_t), P(C.c_size_t)]),
        "vzRequestLaunchParams": (st, [P(RequestDesc), vp, vp, P(LaunchParams)]),
        "vzStatusString": (C.c_char_p, [st]),
        "vzGetLastError": (C.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def lib():
    """Loads libvz.so; raises if it is missing (no fallback path exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VzError(8, f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        _lib = _declare(C.CDLL(LIB_PATH))
    return _lib


def check(status):
    if status != VZ_SUCCESS:
        raise VzError(status, lib().vzGetLastError().decode())
    return status


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- host helpers ------------------------------------------------------------
def book_scene(seed=5489):
    """Book final scene (Scene.h:13-80) as (spheres, materials) arrays."""
    n = C.c_uint32(0)
    cap = 1024
    s = np.zeros(cap, SPHERE_DT)
    m = np.zeros(cap, MATERIAL_DT)
    check(lib().vzGenerateBookScene(seed, _p(s), _p(m), cap, C.byref(n)))
    return s[:n.value].copy(), m[:n.value].copy()


def box_scene(n_small, seed):
    s = np.zeros(n_small + 1, SPHERE_DT)
    m = np.zeros(n_small + 1, MATERIAL_DT)
    check(lib().vzGenerateBoxScene(n_small, seed, _p(s), _p(m)))
    return s, m


def launch_params(width, height, spp, max_depth, lookfrom, lookat, vfov, aperture, focus,
                  subframe=1, flags=0, aspect=None):
    p = LaunchParams()
    p.width, p.height, p.samples_per_pixel, p.max_depth = width, height, spp, max_depth
    p.subframe_index, p.flags = subframe, flags
    f = np.asarray(lookfrom, np.float32)
    a = np.asarray(lookat, np.float32)
    asp = float(np.float32(width) / np.float32(height)) if aspect is None else aspect
    check(lib().vzCameraLookAt(_p(f), _p(a), vfov, asp, aperture, focus, C.byref(p)))
    return p


def footprint(params):
    return lib().vzRequestFootprint(C.byref(params))


# ---- wire format (input of the serving form) -------------------------------------
WIRE_SCENE_HEADER = 24
WIRE_REQUEST_HEADER = 96


def _bytes_ptr(b):
    buf = C.create_string_buffer(bytes(b), len(b)) if len(b) else C.create_string_buffer(1)
    return buf, C.cast(buf, C.c_void_p)


def encode_scene(spheres, materials):
    """Scene record (bytes) of the sphere / material arrays."""
    s = np.ascontiguousarray(spheres, SPHERE_DT)
    m = np.ascontiguousarray(materials, MATERIAL_DT)
    if len(s) != len(m):
        raise ValueError("spheres and materials differ in length")
    size = lib().vzWireSceneSize(len(s))
    out = C.create_string_buffer(size)
    w = C.c_size_t(0)
    check(lib().vzWireEncodeScene(_p(s), _p(m), len(s), C.cast(out, C.c_void_p), size, C.byref(w)))
    return out.raw[:w.value]


def decode_scene(data):
    """(spheres, materials, consumed) of the scene record at the start of data."""
    buf, ptr = _bytes_ptr(data)
    n = C.c_uint32(0)
    used = C.c_size_t(0)
    check(lib().vzWireDecodeScene(ptr, len(data), None, None, 0, C.byref(n), C.byref(used)))
    s = np.zeros(n.value, SPHERE_DT)
    m = np.zeros(n.value, MATERIAL_DT)
    check(lib().vzWireDecodeScene(ptr, len(data), _p(s), _p(m), n.value, C.byref(n), C.byref(used)))
    return s, m, used.value


def request_desc(tag, scene_key, width, height, spp, max_depth, lookfrom, lookat, vfov, aperture=0.0,
                 focus=10.0, flags=0, subframe=1, arrival_ms=0.0):
    d = RequestDesc()
    d.tag, d.scene_key = tag, scene_key
    d.width, d.height, d.samples_per_pixel, d.max_depth = width, height, spp, max_depth
    d.flags, d.subframe_index = flags, subframe
    d.lookfrom[:] = [float(x) for x in lookfrom]
    d.lookat[:] = [float(x) for x in lookat]
    d.vfov_deg, d.aperture, d.focus_dist, d.arrival_ms = vfov, aperture, focus, arrival_ms
    return d


def encode_request(desc, spheres=None, materials=None):
    """Request record (bytes), with an inline scene record when spheres is given."""
    n = 0
    sp = mp = None
    if spheres is not None:
        s = np.ascontiguousarray(spheres, SPHERE_DT)
        m = np.ascontiguousarray(materials, MATERIAL_DT)
        if len(s) != len(m):
            raise ValueError("spheres and materials differ in length")
        n, sp, mp = len(s), _p(s), _p(m)
    size = WIRE_REQUEST_HEADER + (lib().vzWireSceneSize(n) if n else 0)
    out = C.create_string_buffer(size)
    w = C.c_size_t(0)
    check(lib().vzWireEncodeRequest(C.byref(desc), sp, mp, n, C.cast(out, C.c_void_p), size, C.byref(w)))
    return out.raw[:w.value]


def decode_request(data):
    """(RequestDesc, inline scene bytes, consumed) of the request record at the start of data."""
    buf, ptr = _bytes_ptr(data)
    d = RequestDesc()
    sb = C.c_size_t(0)
    used = C.c_size_t(0)
    check(lib().vzWireDecodeRequest(ptr, len(data), C.byref(d), C.byref(sb), C.byref(used)))
    return d, bytes(data[WIRE_REQUEST_HEADER:WIRE_REQUEST_HEADER + sb.value]), used.value


def request_launch_params(desc, image=None, accum=None):
    p = LaunchParams()
    check(lib().vzRequestLaunchParams(C.byref(desc), image, accum, C.byref(p)))
    return p


# ---- device objects --------------------------------------------------------------
class Context:
    """A device context (Renderer::CreateContext, Renderer.h:144-158)."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        self._children = weakref.WeakSet()  # scenes and servers, closed before the context
        check(lib().vzCreate(device, C.byref(self.h)))

    def close(self):
        if self.h:
            # servers first: they finish their requests and release the
            # scenes those name, which can then be destroyed
            children = list(self._children)
            for ch in children:
                if isinstance(ch, Server):
                    ch.close()
            for ch in children:
                if not isinstance(ch, Server):
                    ch.close()
            check(lib().vzDestroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        cus, hbm = C.c_uint32(), C.c_uint64()
        name = C.create_string_buffer(128)
        check(lib().vzDeviceInfo(self.h, C.byref(cus), C.byref(hbm), name, 128))
        return {"num_cus": cus.value, "hbm_bytes": hbm.value, "name": name.value.decode()}

    def default_tuning(self):
        t = Tuning()
        check(lib().vzDefaultTuning(self.h, C.byref(t)))
        return t


class Scene:
    """Resident spheres + LBVH (the GAS and SBT of Renderer.h:160-255, 452-520)."""

    def __init__(self, ctx, spheres, materials, stream=None):
        self.ctx = ctx
        self.spheres = np.ascontiguousarray(spheres, SPHERE_DT)
        self.materials = np.ascontiguousarray(materials, MATERIAL_DT)
        self.n = len(self.spheres)
        self.h = C.c_void_p()
        check(lib().vzSceneCreate(ctx.h, _p(self.spheres), _p(self.materials), self.n,
                                  stream, C.byref(self.h)))
        ctx._children.add(self)

    @classmethod
    def from_device(cls, ctx, spheres_dev_ptr, materials_dev_ptr, n, stream=None):
        """vzSceneCreateDevice: spheres and materials already in device memory."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.spheres = self.materials = None
        self.n = int(n)
        self.h = C.c_void_p()
        check(lib().vzSceneCreateDevice(ctx.h, spheres_dev_ptr, materials_dev_ptr, self.n, stream,
                                        C.byref(self.h)))
        ctx._children.add(self)
        return self

    def close(self):
        """vzSceneDestroy; raises (and keeps the scene) while server
        requests that name it are queued or running."""
        if self.h:
            check(lib().vzSceneDestroy(self.h))
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = AccelInfo()
        check(lib().vzAccelInfoGet(self.h, C.byref(i)))
        return i

    def rebuild(self, stream=None):
        check(lib().vzAccelBuild(self.h, stream))

    def update_spheres(self, spheres, stream=None):
        spheres = np.ascontiguousarray(spheres, SPHERE_DT)
        check(lib().vzSceneUpdateSpheres(self.h, _p(spheres), stream))
        self.spheres = spheres

    def update_spheres_device(self, spheres_dev_ptr, stream=None):
        """vzSceneUpdateSpheresDevice: the new spheres are in device memory."""
        check(lib().vzSceneUpdateSpheresDevice(self.h, spheres_dev_ptr, stream))
        self.spheres = None

    def download(self, stream=None):
        n = self.n
        nodes = np.zeros(max(n - 1, 0), NODE_DT)
        leaf = np.zeros(n, np.uint32)
        morton = np.zeros(n, np.uint32)
        check(lib().vzAccelDownload(self.h, _p(nodes) if n > 1 else None, _p(leaf), _p(morton),
                                    stream))
        return nodes, leaf, morton

    def download_wide(self, stream=None):
        """The 4-wide tree every traversal walks (vzAccelDownloadWide)."""
        nw = self.info().num_wide_nodes
        w = np.zeros(nw, WNODE_DT)
        check(lib().vzAccelDownloadWide(self.h, _p(w) if nw else C.c_void_p(1), stream))
        return w

    def trace(self, rays_dev_ptr, n, hits_dev_ptr, stream=None, stats=False):
        st = Stats()
        check(lib().vzTraceClosest(self.h, rays_dev_ptr, n, hits_dev_ptr, stream,
                                   C.byref(st) if stats else None))
        return st

    def trace_noalloc(self, rays_dev_ptr, n, hits_dev_ptr, stream=None, stats=False):
        """vzTraceClosestNoAlloc: the same query with the library's fixed scratch."""
        st = Stats()
        check(lib().vzTraceClosestNoAlloc(self.h, rays_dev_ptr, n, hits_dev_ptr, stream,
                                          C.byref(st) if stats else None))
        return st

    def launch(self, params, stream=None):
        st = Stats()
        check(lib().vzLaunch(self.h, C.byref(params), stream, C.byref(st)))
        return st


def batch_launch(ctx, scenes_params, tuning=None, stream=None):
    n = len(scenes_params)
    arr = (BatchItem * n)()
    for i, (scene, params) in enumerate(scenes_params):
        arr[i].scene = scene.h
        arr[i].params = params
    st = Stats()
    check(lib().vzBatchLaunch(ctx.h, arr, n, C.byref(tuning) if tuning is not None else None,
                              stream, C.byref(st)))
    return st


def autotune(ctx, scenes_params, stream=None):
    n = len(scenes_params)
    arr = (BatchItem * n)()
    for i, (scene, params) in enumerate(scenes_params):
        arr[i].scene = scene.h
        arr[i].params = params
    t = Tuning()
    check(lib().vzAutotune(ctx.h, arr, n, stream, C.byref(t)))
    return t


class Server:
    """Serving loop: continuous batching over resident scenes."""

    def __init__(self, ctx, hbm_budget_bytes, max_active=256, tuning=None):
        cfg = ServeConfig()
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.max_active = max_active
        if tuning is not None:
            cfg.tuning = tuning
        self.h = C.c_void_p()
        # scenes of requests not yet seen finished: kept alive here so that
        # garbage collection cannot try to destroy a scene a request names
        self._held = {}
        check(lib().vzServerCreate(ctx.h, C.byref(cfg), C.byref(self.h)))
        ctx._children.add(self)

    def submit(self, scene, params):
        rid = C.c_uint64()
        check(lib().vzServerSubmit(self.h, scene.h, C.byref(params), C.byref(rid)))
        self._held[rid.value] = scene
        return rid.value

    def poll(self, rid):
        r = RequestResult()
        s = lib().vzServerPoll(self.h, rid, C.byref(r))
        if s == 4:
            return None
        check(s)
        self._held.pop(rid, None)
        return r

    def drain(self):
        check(lib().vzServerDrain(self.h))
        self._held.clear()

    def close(self):
        if self.h:
            lib().vzServerDestroy(self.h)
            self.h = C.c_void_p()
            self._held.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def declared_symbols():
    """Function names declared in include/vz_api.h."""
    import re
    text = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^(?:vzStatus|uint64_t|size_t|const char\*)\s+(vz\w+)\(", text, re.M)))

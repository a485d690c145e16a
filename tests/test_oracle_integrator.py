"""Pins the oracle's integrator composition with checks that do not reuse
its code.

1. An independent restatement, in plain Python with float32 scalars, of the
   reference's raygen / closest-hit / miss programs for a scene of ONE
   sphere (Core/RayTracer.cu:87-217 raygen and samplers, 272-440 the three
   closest-hit programs, 442-450 miss; Core/random.cuh:31-67 the RNG;
   Core/vec_math.h:483-503, 545-570 the vector helpers).  It predicts every
   segment the oracle traces (origin and direction, bit for bit, read back
   through the oracle's ray recorder) and every pixel's accumulated value.
   The only oracle routine it takes as input is the sphere test's t, which
   test_oracle.py pins by closed forms and against FP64.  It pins:
   tea<4>(y*W + x, frame) seeding, jitter x then y, the u/v formula, the lens
   disk draws, the ray formula, the seed NOT written back after a trace
   (RayTracer.cu:183), each material's scatter and its draw order, albedo
   multiplied for Lambertian and metal but not dielectric, the sky
   gradient, the in-order sample sum, 1/spp and the progressive blend.
2. Closed forms: one Lambertian sphere at max_depth 2 gives blue = a_b
   exactly on every path that hits it and escapes; glass gives blue in
   {0, 1} per path (no attenuation); the Lambertian direction is
   n + a unit vector, so |d' - n| = 1.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

F = np.float32
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M32 = 0xFFFFFFFF


# ---- RNG: TEA (Wheeler & Needham 1994) with 4 rounds and the key of
# random.cuh:31-45; the Numerical Recipes LCG of random.cuh:47-55; a float
# from the low 24 bits (random.cuh:63-67).  Checked against the KAT table.
def tea4(v0, v1):
    s0 = 0
    for _ in range(4):
        s0 = (s0 + 0x9E3779B9) & M32
        v0 = (v0 + ((((v1 << 4) + 0xA341316C) & M32) ^ ((v1 + s0) & M32) ^ (((v1 >> 5) + 0xC8013EA4) & M32))) & M32
        v1 = (v1 + ((((v0 << 4) + 0xAD90777D) & M32) ^ ((v0 + s0) & M32) ^ (((v0 >> 5) + 0x7E95761E) & M32))) & M32
    return v0


class Seed:
    def __init__(self, s):
        self.s = s

    def rnd(self):
        self.s = (1664525 * self.s + 1013904223) & M32
        return F(self.s & 0x00FFFFFF) / F(0x01000000)


def test_python_rng_matches_kat():
    kat = json.load(open(os.path.join(GOLDEN, "rng_kat.json")))
    for row in kat["rows"]:
        assert tea4(row["v0"], row["v1"]) == row["tea4"]
        s = Seed(row["tea4"])
        assert [float(s.rnd()) for _ in range(3)] == row["rnd"] and s.s == row["state"]


# ---- float3 helpers in the reference's operation order (vec_math.h) -------
def v3(x, y, z):
    return (F(x), F(y), F(z))


def add(a, b):
    return (a[0] + b[0], a[1] + b[1], a[2] + b[2])


def sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def mul(a, b):
    return (a[0] * b[0], a[1] * b[1], a[2] * b[2])


def scl(a, s):
    return (a[0] * s, a[1] * s, a[2] * s)


def neg(a):
    return (-a[0], -a[1], -a[2])


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def normalize(v):  # vec_math.h:545-549: v * (1 / sqrt(dot(v, v)))
    return scl(v, F(1.0) / np.sqrt(dot(v, v)))


def reflect(i, n):  # vec_math.h:558-561: i - 2 * n * dot(n, i)
    return sub(i, scl(scl(n, F(2.0)), dot(n, i)))


def refract(uv, n, eta):  # vec_math.h:564-570
    ct = min(dot(neg(uv), n), F(1.0))
    perp = scl(add(uv, scl(n, ct)), eta)
    par = scl(n, -np.sqrt(abs(F(1.0) - dot(perp, perp))))
    return add(perp, par)


def rnd_m11(seed):  # random_float(seed, -1, 1) = -1 + (1 - -1) * rnd
    return F(-1.0) + F(2.0) * seed.rnd()


def in_unit_sphere(seed):  # RayTracer.cu:117-125, components drawn x, y, z
    while True:
        p = (rnd_m11(seed), rnd_m11(seed), rnd_m11(seed))
        if dot(p, p) >= F(1.0):
            continue
        return p


def sky(d):  # RayTracer.cu:442-450: lerp((1,1,1), (0.5,0.7,1), 0.5 (y + 1))
    u = normalize(d)
    t = F(0.5) * (u[1] + F(1.0))
    one = v3(1, 1, 1)
    return add(one, scl(sub(v3(0.5, 0.7, 1.0), one), t))


def camera_ray(cam, W, H, x, y, seed):
    """RayTracer.cu:151-177: jitter x then y, the thin-lens disk, the ray."""
    u = F(2.0) * (F(x) + seed.rnd()) / F(W - 1) - F(1.0)
    v = F(2.0) * (F(y) + seed.rnd()) / F(H - 1) - F(1.0)
    while True:
        p = (rnd_m11(seed), rnd_m11(seed), F(0.0))
        if dot(p, p) >= F(1.0):
            continue
        break
    lr = F(cam.lens_radius)
    rd = scl(p, lr)
    U, V, Wv = tuple(map(F, cam.u)), tuple(map(F, cam.v)), tuple(map(F, cam.w))
    off = add(scl(normalize(U), rd[0]), scl(normalize(V), rd[1]))
    org = add(tuple(map(F, cam.origin)), off)
    d = sub(add(add(Wv, scl(scl(U, u), F(0.5))), scl(scl(V, v), F(0.5))), off)
    return org, d


def scatter(mat, d, n, front, seed):
    """The three closest-hit programs (RayTracer.cu:272-440).  Returns
    (continues, new direction, attenuation)."""
    t = int(mat["type"])
    alb = tuple(map(F, mat["albedo"]))
    if t == O_LAMB:
        sd = add(n, normalize(in_unit_sphere(seed)))
        if all(abs(float(c)) < 1e-8 for c in sd):
            sd = n
        return True, sd, alb
    if t == O_METAL:
        sd = add(reflect(normalize(d), n), scl(in_unit_sphere(seed), F(mat["fuzz"])))
        if dot(sd, n) > F(0.0):
            return True, sd, alb
        return False, None, None
    ir = F(mat["ir"])
    ratio = F(1.0) / ir if front else ir
    u = normalize(d)
    ct = min(dot(neg(u), n), F(1.0))
    st = np.sqrt(F(1.0) - ct * ct)
    refl = ratio * st > F(1.0)
    if not refl:  # Schlick with (1 - c)^5 as multiplies; rnd drawn only here
        r0 = (F(1.0) - ratio) / (F(1.0) + ratio)
        r0 = r0 * r0
        x = F(1.0) - ct
        x2 = x * x
        sch = r0 + (F(1.0) - r0) * ((x2 * x2) * x)
        refl = sch > seed.rnd()
    return True, (reflect(u, n) if refl else refract(u, n, ratio)), v3(1, 1, 1)


O_LAMB, O_METAL, O_DIEL = 0, 1, 2


def _ray_tuple(r):
    return (F(r["ox"]), F(r["oy"]), F(r["oz"])), (F(r["dx"]), F(r["dy"]), F(r["dz"]))


def _hit(sph, o, d):
    """The oracle's sphere test (pinned in test_oracle.py) on one ray."""
    ray = np.zeros(1, O.RAY_DT)
    ray[0] = (o[0], o[1], o[2], 0.001, d[0], d[1], d[2], 1e16)
    h = O.trace_brute(sph, ray)[0]
    return (None if int(h["prim"]) == O.MISS else F(h["t"]))


def restate(sph, mat, cam, W, H, spp, max_depth, frame, rec):
    """Per pixel (y outer, x inner, as the recorder runs) and sample: the
    Python path, checked segment by segment against the recorded rays.
    Returns (accum[H, W, 3] float32 for a zero previous frame, segments)."""
    c = (F(sph[0]["cx"]), F(sph[0]["cy"]), F(sph[0]["cz"]))
    inv_r = F(1.0) / F(sph[0]["r"])
    k = 0
    acc = np.zeros((H, W, 3), np.float32)
    for y in range(H):
        for x in range(W):
            seed = Seed(tea4(y * W + x, frame))
            total = v3(0, 0, 0)
            for _ in range(spp):
                o, d = camera_ray(cam, W, H, x, y, seed)
                path = Seed(seed.s)  # the PRD gets a copy; the raygen seed is not written back
                thr = v3(1, 1, 1)
                depth = max_depth - 1
                while True:
                    ro, rd = _ray_tuple(rec[k])
                    assert ro == o and rd == d, (x, y, k)
                    k += 1
                    t = _hit(sph, o, d)
                    if t is None:
                        L = mul(thr, sky(d))
                        break
                    if depth <= 0:
                        L = v3(0, 0, 0)
                        break
                    p = add(o, scl(d, t))
                    n = scl(sub(p, c), inv_r)
                    front = dot(d, n) < F(0.0)
                    if not front:
                        n = neg(n)
                    ok, nd, att = scatter(mat[0], d, n, front, path)
                    if not ok:
                        L = v3(0, 0, 0)
                        break
                    thr = mul(thr, att)
                    depth -= 1
                    o, d = p, nd
                total = add(total, L)
            avg = scl(total, F(1.0) / F(spp))
            if frame > 1:  # blend 1/frame of this frame into a zero previous frame
                avg = scl(avg, F(1.0) / F(frame))
            acc[y, x] = avg
    return acc, k


def _one_sphere(mtype, albedo, fuzz=0.0, ir=1.5, center=(0.0, 0.0, -3.0), r=1.0):
    s = np.zeros(1, O.SPHERE_DT)
    s[0] = (*center, r)
    m = np.zeros(1, O.MATERIAL_DT)
    m[0]["type"], m[0]["albedo"], m[0]["fuzz"], m[0]["ir"] = mtype, albedo, fuzz, ir
    return s, m


@pytest.mark.parametrize("mtype,albedo,fuzz,depth,frame", [
    (O_LAMB, (0.75, 0.5, 0.25), 0.0, 2, 1),
    (O_LAMB, (0.6, 0.3, 0.9), 0.0, 4, 2),
    (O_METAL, (0.5, 0.75, 0.25), 0.3, 3, 1),
    (O_DIEL, (0.2, 0.4, 0.6), 0.0, 6, 3),   # dielectric: the albedo field must be ignored
])
def test_integrator_matches_independent_restatement(mtype, albedo, fuzz, depth, frame):
    W, H, spp = 14, 9, 3
    s, m = _one_sphere(mtype, albedo, fuzz)
    cam = O.camera_lookat([0.3, 0.4, 0.0], [0.0, 0.0, -3.0], 60.0, float(F(W) / F(H)), 0.4, 3.0)
    p = O.make_params(W, H, spp, depth, cam, subframe=frame)
    acc, _, st = O.render(s, m, p, bvh=None)
    rec = O.make_rayset(s, m, p, int(st.segments), 0, 0)
    want, nseg = restate(s, m, cam, W, H, spp, depth, frame, rec)
    assert nseg == st.segments
    assert 0 < st.hits < st.segments  # the view holds the sphere and sky
    assert np.array_equal(acc[..., :3], want)
    assert np.all(acc[..., 3] == 1.0)


def test_no_write_back_is_what_the_stream_pins():
    """Planted mistake, on the CPU: writing the path's seed back into the
    raygen seed (the correlation fix the reference does NOT make) changes
    the camera rays of later samples, so the restatement would catch an
    oracle that did it."""
    W, H, spp = 6, 4, 4
    s, m = _one_sphere(O_LAMB, (0.5, 0.5, 0.5), center=(0.0, 0.0, -2.0), r=1.5)
    cam = O.camera_lookat([0.0, 0.0, 0.0], [0.0, 0.0, -3.0], 60.0, 1.5, 0.2, 2.0)
    changed = 0
    for y in range(H):
        for x in range(W):
            a, b = Seed(tea4(y * W + x, 1)), Seed(tea4(y * W + x, 1))
            for _ in range(spp):
                ra = camera_ray(cam, W, H, x, y, a)
                rb = camera_ray(cam, W, H, x, y, b)
                changed += ra != rb
                path = Seed(b.s)
                in_unit_sphere(path)  # a bounce's draws ...
                b.s = path.s          # ... written back
    assert changed > 0


# ---- closed forms (RayTracer.cu:272-319, 373-450) ----------------------------
def _covered_pixels(s, cam, W, H, spp):
    """Pixels all of whose camera rays hit the sphere (read from the ray
    recorder at max_depth 1: one segment per sample)."""
    m = np.zeros(1, O.MATERIAL_DT)
    p = O.make_params(W, H, spp, 1, cam)
    rays = O.make_rayset(s, m, p, W * H * spp, 0, 0)
    hit = O.trace_brute(s, rays)["prim"] != O.MISS
    return hit.reshape(H, W, spp).all(axis=2)


def test_one_lambertian_sphere_blue_is_albedo():
    """max_depth 2: a path that hits the sphere scatters once and escapes
    (a bounce off a convex sphere leaves it), so its radiance is
    albedo * sky(d'), and sky's blue is exactly 1 for every direction.  With
    dyadic albedo and spp the pixel average of blue is exactly a_b on every
    fully covered pixel; r and g lie on the sky's line through the albedo."""
    W, H, spp = 24, 16, 8
    a = (0.75, 0.5, 0.25)
    s, m = _one_sphere(O_LAMB, a)
    cam = O.camera_lookat([0.0, 0.0, 0.0], [0.0, 0.0, -3.0], 40.0, 1.5, 0.0, 3.0)
    acc, _, st = O.render(s, m, O.make_params(W, H, spp, 2, cam))
    cov = _covered_pixels(s, cam, W, H, spp)
    assert cov.sum() > 40
    rgb = acc[..., :3][cov].astype(np.float64)
    assert np.all(rgb[:, 2] == 0.25)
    r, g = rgb[:, 0] / a[0], rgb[:, 1] / a[1]
    eps = 1e-6
    assert np.all((r >= 0.5 - eps) & (r <= 1.0 + eps) & (g >= 0.7 - eps) & (g <= 1.0 + eps))
    # r = 1 - 0.5 t and g = 1 - 0.3 t with the same mean t
    assert np.allclose((1 - r) / 0.5, (1 - g) / 0.3, atol=1e-5)
    assert st.hits > 0 and st.segments > W * H * spp


def test_one_glass_sphere_has_no_attenuation():
    """A dielectric path is never attenuated (RayTracer.cu:381-440 has no
    albedo multiply): each escaping sample's blue is exactly 1, a path that
    runs out of depth inside is 0, so blue * spp is an integer on every
    pixel, whatever albedo the material record holds."""
    W, H, spp = 20, 12, 8
    s, m = _one_sphere(O_DIEL, (0.3, 0.6, 0.1))
    cam = O.camera_lookat([0.0, 0.0, 0.0], [0.0, 0.0, -3.0], 40.0, float(F(W) / F(H)), 0.0, 3.0)
    acc, _, st = O.render(s, m, O.make_params(W, H, spp, 8, cam))
    b = acc[..., 2].astype(np.float64) * spp
    assert np.all(b == np.round(b))
    assert st.hits > 0
    cov = _covered_pixels(s, cam, W, H, spp)
    assert cov.sum() > 20 and np.any(b[cov] == spp)  # most glass paths escape unattenuated


def test_lambertian_direction_is_normal_plus_unit_vector():
    """RayTracer.cu:288: normal + random_unit_vector, the unit vector being
    a normalised rejection sample.  So |d' - n| = 1 to a few ulps; the
    unnormalised n + in_unit_sphere would give |d' - n| < 1."""
    rng = np.random.default_rng(5)
    m = np.zeros(1, O.MATERIAL_DT)[0]
    m["type"], m["albedo"] = O_LAMB, (0.5, 0.5, 0.5)
    for seed in range(600):
        n = rng.normal(size=3)
        n = (n / np.linalg.norm(n)).astype(np.float32)
        ok, nd, _ = O.scatter(m, -n, n, True, seed)
        assert ok
        dist = np.linalg.norm(nd.astype(np.float64) - n.astype(np.float64))
        assert abs(dist - 1.0) < 4e-7, (seed, dist)

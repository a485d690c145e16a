"""The C++ Camera's interactive controls (venusaur_amd/core/camera.h) against
the reference's semantics (Core/camera.h:36-85, Core/Camera.cpp:16-37),
restated here in float64 from glm's definitions: glm::rotate(quat(1,0,0,0),
angle, axis) is the unit quaternion of `angle` about the normalised axis, and
glm's `v * q` applies the inverse rotation, i.e. rotates v by -angle.  The
axes are the frame of the last UpdateUVW, so moves made between two frames
all use the same, older frame.  Host-only: the camera is plain C++."""
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include "camera.h"
static void put(const char* tag, const vec3& v) { std::printf("%s %.9g %.9g %.9g\n", tag, v.x, v.y, v.z); }
int main(int argc, char** argv)
{
    Camera cam(vec3(1.0f, 2.0f, 3.0f), 40.0f, 1.5f, 0.2f, 4.0f);
    cam.SetForward(vec3(-0.3f, -0.2f, -1.0f));
    vec3 U, V, W;
    cam.UVWFrame(U, V, W);
    put("U0", U); put("V0", V); put("W0", W);
    for (int i = 1; i + 1 < argc; i += 2) {
        float a = (float)std::atof(argv[i + 1]);
        if (!std::strcmp(argv[i], "pitch")) cam.Pitch(a);
        else if (!std::strcmp(argv[i], "yaw")) cam.Yaw(a);
        else if (!std::strcmp(argv[i], "roll")) cam.Roll(a);
        else if (!std::strcmp(argv[i], "right")) cam.MoveRight(a);
        else if (!std::strcmp(argv[i], "up")) cam.MoveUp(a);
        else if (!std::strcmp(argv[i], "fwd")) cam.MoveForward(a);
        else if (!std::strcmp(argv[i], "frame")) cam.UVWFrame(U, V, W);
    }
    std::printf("changed %d\n", cam.Changed() ? 1 : 0);
    std::printf("changed %d\n", cam.Changed() ? 1 : 0);
    put("F", cam.GetForward());
    put("P", cam.GetPosition());
    cam.UVWFrame(U, V, W);
    put("U", U); put("V", V); put("W", W);
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("cam")
    src = d / "cam.cpp"
    src.write_text(DRIVER)
    exe = d / "cam"
    subprocess.run([cxx, "-std=c++17", "-O1", "-ffp-contract=off", "-I", os.path.join(ROOT, "venusaur_amd", "core"),
                    str(src), "-o", str(exe)], check=True)
    return str(exe)


def run(exe, *moves):
    out = subprocess.run([exe, *[str(m) for m in moves]], capture_output=True, text=True, check=True).stdout
    vals, changed = {}, []
    for line in out.splitlines():
        k, *rest = line.split()
        if k == "changed":
            changed.append(int(rest[0]))
        else:
            vals[k] = np.array([float(x) for x in rest])
    return vals, changed


def quat_rotate_inverse(v, axis, degrees):
    """v * q with q = glm::rotate(identity, radians(degrees), axis): the
    conjugate rotation q^-1 v q, written out as quaternion products."""
    a = math.radians(degrees)
    n = axis / np.linalg.norm(axis)
    w, x, y, z = math.cos(a / 2), *(n * math.sin(a / 2))

    def qmul(p, q):
        pw, px, py, pz = p
        qw, qx, qy, qz = q
        return (pw * qw - px * qx - py * qy - pz * qz, pw * qx + px * qw + py * qz - pz * qy,
                pw * qy - px * qz + py * qw + pz * qx, pw * qz + px * qy - py * qx + pz * qw)
    qc = (w, -x, -y, -z)
    r = qmul(qmul(qc, (0.0, *v)), (w, x, y, z))
    return np.array(r[1:])


def frame(forward, focal=4.0, vfov=40.0, aspect=1.5):
    """Camera.cpp:24-37 in float64."""
    Wv = forward * focal
    U = np.cross(Wv, [0.0, 1.0, 0.0])
    U /= np.linalg.norm(U)
    V = np.cross(U, Wv)
    V /= np.linalg.norm(V)
    h = math.tan(math.radians(vfov) / 2)
    return U * focal * aspect * 2 * h, V * focal * 2 * h, Wv


def test_initial_frame(driver):
    v, _ = run(driver)
    f = np.array([-0.3, -0.2, -1.0])
    f /= np.linalg.norm(f)
    U, V, W = frame(f)
    assert np.allclose(v["U0"], U, atol=2e-6) and np.allclose(v["V0"], V, atol=2e-6)
    assert np.allclose(v["W0"], W, atol=2e-6)


@pytest.mark.parametrize("moves", [
    ("pitch", 7.5), ("yaw", -12.0), ("roll", 30.0),
    ("pitch", 5.0, "yaw", 10.0),                       # both about the same, older frame
    ("yaw", 10.0, "frame", 0, "yaw", 10.0),            # the second yaw about the new frame
    ("pitch", 3.0, "roll", -4.0, "yaw", 6.0, "frame", 0, "pitch", -2.0),
])
def test_rotations_follow_glm_semantics(driver, moves):
    v, changed = run(driver, *moves)
    assert changed == [1, 0]  # Changed() is true once after a move (camera.h:80-85)
    f = np.array([-0.3, -0.2, -1.0])
    f /= np.linalg.norm(f)
    U, V, W = frame(f)
    it = iter(moves)
    for op, arg in zip(it, it):
        if op == "frame":
            U, V, W = frame(f)
            continue
        axis = {"pitch": U, "yaw": V, "roll": W}[op]
        f = quat_rotate_inverse(f, axis, float(arg))
        f /= np.linalg.norm(f)
    assert np.allclose(v["F"], f, atol=3e-6), (v["F"], f)
    U2, V2, W2 = frame(f)
    assert np.allclose(v["U"], U2, atol=5e-6) and np.allclose(v["W"], W2, atol=5e-6)
    # the sign convention is what the check holds: a positive pitch turns
    # the view down (v * q rotates by -angle about U)
    if moves == ("pitch", 7.5):
        assert v["F"][1] < -0.2 / math.sqrt(1.13)


def test_moves_use_the_frame_directions(driver):
    v, _ = run(driver, "right", 2.0, "up", -1.0, "fwd", 0.5)
    f = np.array([-0.3, -0.2, -1.0])
    f /= np.linalg.norm(f)
    U, V, _ = frame(f)
    want = np.array([1.0, 2.0, 3.0]) + U / np.linalg.norm(U) * 2.0 - V / np.linalg.norm(V) + f * 0.5
    assert np.allclose(v["P"], want, atol=5e-6)

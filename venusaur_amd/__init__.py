"""venusaur_amd -- MI355X-native sphere path tracer behind a C ABI.

The product is libvz.so (venusaur_amd/csrc, include/vz_api.h): a GPU LBVH,
short-stack traversal and a wavefront integrator in hand-written HIP for
gfx950, with the reference's Renderer/Scene/Camera API mirrored in C++
(venusaur_amd/core) and, for Python callers, in venusaur_amd.api.
"""
from . import api  # noqa: F401
from .api import (VzError, Context, Scene, Server, batch_launch, autotune, book_scene,  # noqa: F401
                  box_scene, launch_params)

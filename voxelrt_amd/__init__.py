"""voxelrt_amd -- Python host binding of libvoxelrt_hip (MI355X / gfx950).

The product is the C ABI in ``include/voxelrt.h`` implemented by
``voxelrt_amd/libvoxelrt_hip.so`` (HIP kernels + C++ runtime).  This module
is a thin ctypes layer over it, shaped like the reference's host flow
(src/main.js:680-747: one ``frame()`` = pathtrace -> temporal -> a-trous x3,
and ``capture(hdr=True)`` as in main.js:571-661).  There is no CPU fallback:
every call goes to the HIP library, and importing fails loudly when the
library has not been built.
"""
from __future__ import annotations

from .lib import (  # noqa: F401
    ACCUM_RGBA16F, ACCUM_RGBA32F, FACE_MISS, FACE_NONE, FLAG_DENOISE, FLAG_TEMPORAL,
    HIT_DTYPE, NOALLOC_MAX_RAYS, SCENE_RANDOM_BRICKS, SCENE_VOXELRT, SCENE_WORLD, VOXEL_MISS, VOXEL_SIZE,
    Context, DeviceArray, Frame, Scene, SceneDesc, Tuning, Uniforms, VrtError, View, build, camera_matrices,
    device_count, generate_scene, lib_path, load, make_uniforms, scene_spawn,
)
from .renderer import Renderer, orbit_pose  # noqa: F401

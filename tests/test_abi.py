"""The C ABI without a GPU: exports, struct layouts, error paths and scene validation."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADERS = sorted((ROOT / "include").glob("*.h"))


def _declared_functions() -> set[str]:
    names = set()
    for h in HEADERS:
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(vx_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_function(vx):
    declared = _declared_functions()
    assert len(declared) >= 15
    missing = [n for n in sorted(declared) if not hasattr(vx.LIB, n)]
    assert not missing, f"declared in include/*.h but not exported: {missing}"
    assert declared <= set(vx.EXPORTED) | declared  # the Python binding lists them too
    assert not (declared - set(vx.EXPORTED)), sorted(declared - set(vx.EXPORTED))


def test_struct_layouts_compile_time(tmp_path):
    """sizeof/offsetof of the ABI structs against the reference layouts
    (OctreeConfig: libraries/CashSystem/include/VoxelSceneCacher.h:100-133;
    push constant: libraries/RenderGraph/include/Data/CameraData.h:17-27)."""
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stddef.h>
#include "vixen_build.h"
_Static_assert(sizeof(vx_ray) == 32, "ray");
_Static_assert(sizeof(vx_hit) == 32, "hit");
_Static_assert(sizeof(vx_camera) == 64, "camera");
_Static_assert(offsetof(vx_camera, fov_deg) == 28 && offsetof(vx_camera, aspect) == 44 &&
               offsetof(vx_camera, debug_mode) == 60, "camera offsets");
_Static_assert(sizeof(vx_octree_config) == 256, "config");
_Static_assert(offsetof(vx_octree_config, grid_min) == 32 && offsetof(vx_octree_config, grid_max) == 48 &&
               offsetof(vx_octree_config, local_to_world) == 64 &&
               offsetof(vx_octree_config, world_to_local) == 128 &&
               offsetof(vx_octree_config, _padding4) == 192, "config offsets");
_Static_assert(offsetof(vx_hit, rgb) == 20 && offsetof(vx_ray, d) == 16, "record offsets");
_Static_assert(offsetof(vx_scene_desc, config) == 40 && offsetof(vx_scene_desc, dxt_normals) == 48 &&
               offsetof(vx_scene_desc, material_palette) == 56 && sizeof(vx_scene_desc) == 64, "scene descriptor");
_Static_assert(VX_BC1_NONE == 0 && VX_BC1_FROM_RGBA8 == 1 && VX_BC1_FROM_MATERIAL == 2, "BC1 sources");
int main(void) { return 0; }
""")
    r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_numpy_dtypes_match_abi(vx):
    assert vx.RAY_DTYPE.itemsize == C.sizeof(C.c_float) * 8
    assert vx.HIT_DTYPE.fields["rgb"][1] == 20
    assert C.sizeof(vx.VxOctreeConfig) == 256 and C.sizeof(vx.VxCamera) == 64
    assert C.sizeof(vx.VxSceneDesc) == 64 and vx.VxSceneDesc.dxt_normals.offset == 48
    assert vx.VxSceneDesc.material_palette.offset == 56


def test_version_and_no_device_errors(vx):
    assert b"gfx950" in vx.LIB.vx_version()
    h = C.c_void_p()
    assert vx.LIB.vx_ctx_create(-1, C.byref(h)) != 0
    assert vx.LIB.vx_last_error()
    assert vx.LIB.vx_ctx_create(1 << 20, C.byref(h)) != 0
    assert vx.LIB.vx_ctx_create(0, None) == 1


def _desc(vx, built, **over):
    nodes = over.get("nodes", built.nodes)
    cfg = over.get("config", built.config)
    d = vx.VxSceneDesc()
    d.esvo_nodes = nodes.ctypes.data
    d.node_count = over.get("node_count", nodes.shape[0])
    d.brick_data = built.brick_data.ctypes.data
    d.brick_count = over.get("brick_count", built.brick_count)
    d.bc1_colors = None
    d.config = C.pointer(cfg)
    return d, (nodes, cfg)


def _validate(vx, built, **over):
    d, keep = _desc(vx, built, **over)
    st = vx.LIB.vx_scene_validate(C.byref(d))
    return st, vx.LIB.vx_last_error().decode()


def test_validate_accepts_built_scenes(vx, cave, cornell):
    for _, _, built in (cave, cornell):
        assert _validate(vx, built)[0] == 0


def test_validate_rejects_corrupt_trees(vx, cornell):
    built = cornell[2]
    VX_ERR_INVALID_SCENE = 2
    # a leaf bit on the root (an internal level)
    nodes = built.nodes.copy()
    nodes[0, 0] |= ((nodes[0, 0] >> 16) & 0xFF) << 24
    st, msg = _validate(vx, built, nodes=nodes)
    assert st == VX_ERR_INVALID_SCENE and "leaf" in msg
    # a child pointer past the end
    nodes = built.nodes.copy()
    nodes[0, 0] = (nodes[0, 0] & ~np.uint32(0x7FFF)) | np.uint32(0x7FF0)
    st, msg = _validate(vx, built, nodes=nodes)
    assert st == VX_ERR_INVALID_SCENE and "out of range" in msg
    # a far pointer past the end
    nodes = built.nodes.copy()
    nodes[0, 0] |= np.uint32(0x8000)
    nodes[0, 1] = 0xFFFFF0
    assert _validate(vx, built, nodes=nodes)[0] == VX_ERR_INVALID_SCENE
    # a brick id past the pool
    st, msg = _validate(vx, built, brick_count=built.brick_count - 1)
    assert st == VX_ERR_INVALID_SCENE and "brick" in msg


def test_validate_survives_seeded_corruption(vx, cornell):
    """Seeded mutations of a built tree (valid and leaf masks, far bits, child pointers, brick ids, node
    counts): validation answers every one with a status (run with the library built under
    AddressSanitizer this also checks that the walk reads only descriptors that exist), and a tree it
    accepts has every reachable child block inside the array and every reachable brick id valid, which
    is what the traversal kernels rely on."""
    built = cornell[2]
    rng = np.random.default_rng(5)
    outcomes = set()
    for i in range(400):
        nodes = built.nodes.copy()
        nn = nodes.shape[0]
        for _ in range(int(rng.integers(1, 6))):
            k, w = int(rng.integers(0, nn)), int(rng.integers(0, 2))
            if i % 2:
                nodes[k, w] ^= np.uint32(1 << int(rng.integers(0, 32)))
            else:
                nodes[k, w] = np.uint32(rng.integers(0, 2 ** 32))
        count = int(rng.integers(1, nn + 1)) if i % 7 == 0 else nn
        st, _ = _validate(vx, built, nodes=nodes, node_count=count)
        outcomes.add(st)
        assert st in (0, 2), st
        if st == 0:  # the accepted tree, walked from the root as the kernels walk it
            todo = [0]
            while todo:
                idx = todo.pop()
                w0, w1 = int(nodes[idx, 0]), int(nodes[idx, 1])
                valid, leaf = (w0 >> 16) & 0xFF, (w0 >> 24) & 0xFF
                cp = (w1 & 0xFFFFFF) if w0 & 0x8000 else (w0 & 0x7FFF)
                for c in range(bin(valid).count("1")):
                    assert cp + c < count
                    if leaf:
                        b = int(nodes[cp + c, 1]) & 0xFFFFFF
                        assert b == 0xFFFFFF or b < built.brick_count
                    else:
                        todo.append(cp + c)
    assert outcomes == {0, 2}


def test_validate_rejects_bad_configs(vx, cornell):
    built = cornell[2]
    for field, value in (("bricks_per_axis", 3), ("bricks_per_axis", 1), ("user_max_levels", 7),
                         ("brick_size", 4), ("esvo_max_scale", 23)):
        cfg = vx.VxOctreeConfig()
        C.memmove(C.addressof(cfg), C.addressof(built.config), 256)
        setattr(cfg, field, value)
        assert _validate(vx, built, config=cfg)[0] == 2, field
    cfg = vx.VxOctreeConfig()
    C.memmove(C.addressof(cfg), C.addressof(built.config), 256)
    cfg.grid_max[1] = 50.0  # not a cube
    assert _validate(vx, built, config=cfg)[0] == 2
    assert vx.LIB.vx_scene_validate(None) == 1


def test_trace_rejects_null_arguments_without_device(vx):
    assert vx.LIB.vx_trace(None, None, 10, None, None) == 1
    assert vx.LIB.vx_render(None, None, 4, 4, 1.0, None, None, None) == 1
    assert vx.LIB.vx_camera_rays(None, 4, 4, 1.0, None, None) == 1
    # oversized images are refused before anything reaches a device
    cam = vx.VxCamera()
    assert vx.LIB.vx_camera_rays(C.byref(cam), 70000, 4, 1.0, C.c_void_p(16), None) == 1
    assert b"too large" in vx.LIB.vx_last_error()
    out = C.c_void_p()
    assert vx.LIB.vx_scene_create_device(None, None, C.byref(out)) == 1


def test_builder_rejects_bad_sizes(vx):
    out = C.POINTER(vx.VxBuiltScene)()
    g = np.zeros(8 ** 3, np.uint32)
    assert vx.LIB.vx_build_from_dense(g.ctypes.data, 8, 1, 1, C.byref(out)) == 1
    assert b"power of two" in vx.LIB.vx_last_error()
    assert vx.LIB.vx_build_from_dense(None, 16, 1, 1, C.byref(out)) == 1


def test_python_mirror_checks_buffers_before_any_launch(vx):
    # device buffers must be contiguous CUDA tensors on the scene's device, large enough
    import torch
    with pytest.raises(ValueError, match="cuda:0"):
        vx._dev_buf(torch.zeros(64, dtype=torch.uint8), "hits", 0, 32)
    with pytest.raises(ValueError, match="torch tensor"):
        vx._dev_buf(np.zeros(64, np.uint8), "hits", 0, 32)


def test_batched_ray_offsets_are_checked(vx):
    # one entry per scene plus one, non-decreasing, ending inside the rays (the C side reads off[n_scenes]
    # and trusts the rest)
    got = vx._ray_offsets([0, 10, 10, 25], 3, 25)
    assert list(got) == [0, 10, 10, 25]
    with pytest.raises(ValueError, match="entries"):
        vx._ray_offsets([0, 10, 25], 3, 25)            # one entry short
    with pytest.raises(ValueError, match="non-decreasing"):
        vx._ray_offsets([0, 10, 5, 25], 3, 25)
    with pytest.raises(ValueError, match="the rays hold"):
        vx._ray_offsets([0, 10, 10, 26], 3, 25)        # past the last ray

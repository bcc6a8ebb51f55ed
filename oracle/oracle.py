"""ctypes front of the CPU oracle (liboracle_vx.so). TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only to check results (or as the timed CPU baseline), never
as the product path. See vx_oracle.c for what each function restates and how it is
pinned.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_vx.so"

RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("flags", "<u4"), ("t", "<f4"), ("leaf_node", "<u4"), ("brick", "<u4"),
                      ("voxel", "<u4"), ("rgb", "<f4", 3)])


class OrcScene(C.Structure):
    _fields_ = [("nodes", C.c_void_p), ("node_count", C.c_uint64), ("brick_data", C.c_void_p),
                ("brick_count", C.c_uint64), ("bc1", C.c_void_p), ("grid_min", C.c_float * 3),
                ("grid_max", C.c_float * 3), ("max_levels", C.c_int32), ("bricks_per_axis", C.c_int32),
                ("normals", C.c_void_p), ("palette", C.c_void_p)]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    lib = C.CDLL(str(LIB_PATH))
    vp, u64 = C.c_void_p, C.c_uint64
    lib.orc_trace.argtypes = [C.POINTER(OrcScene), vp, u64, vp, vp]
    lib.orc_trace_mt.argtypes = [C.POINTER(OrcScene), vp, u64, vp, vp, C.c_int]
    lib.orc_trace_lod.argtypes = [C.POINTER(OrcScene), vp, u64, C.c_float, C.c_float, vp, C.c_int]
    lib.orc_camera_rays.argtypes = [vp, C.c_uint32, C.c_uint32, C.c_float, vp]
    lib.orc_bc1_decode.argtypes = [C.c_uint64, C.c_int, vp]
    lib.orc_rgba8_decode.argtypes = [C.c_uint32, vp]
    lib.orc_bruteforce.argtypes = [C.POINTER(OrcScene), vp, vp, u64, C.c_double, vp, vp, vp]
    lib.orc_setup_ray.argtypes = [C.POINTER(OrcScene), vp, vp, vp, vp, vp, vp]
    lib.orc_setup_ray.restype = C.c_int
    lib.orc_bc1_decode_shader.argtypes = [C.c_uint64, C.c_int, vp]
    lib.orc_dxt_normal_decode.argtypes = [C.c_uint64, C.c_uint64, C.c_int, vp]
    lib.orc_shade.argtypes = [C.POINTER(OrcScene), vp, vp, C.c_float]
    lib.orc_shade.restype = C.c_uint32
    lib.orc_render.argtypes = [C.POINTER(OrcScene), vp, C.c_uint32, C.c_uint32, C.c_float, vp, C.c_int]
    lib.orc_trace_cpu_reference.argtypes = [C.POINTER(OrcScene), vp, u64, C.c_uint32, vp, vp]
    lib.orc_trace_cpu_reference.restype = C.c_int
    lib.orc_material_color.argtypes = [C.c_uint32, vp]
    lib.orc_material_id_to_color.argtypes = [C.c_uint32, vp]
    for f in ("orc_trace", "orc_trace_mt", "orc_trace_lod", "orc_camera_rays", "orc_bc1_decode", "orc_rgba8_decode",
              "orc_bruteforce", "orc_bc1_decode_shader", "orc_dxt_normal_decode", "orc_render", "orc_material_color",
              "orc_material_id_to_color"):
        getattr(lib, f).restype = None
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


class OracleScene:
    """Holds the arrays an orc_scene points at (nodes, bricks, BC1, bounds)."""

    def __init__(self, nodes, brick_data, bc1, grid_min, grid_max, max_levels, bricks_per_axis, normals=None,
                 palette=None):
        self.nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        self.brick_data = np.ascontiguousarray(brick_data, dtype=np.uint32)
        self.bc1 = None if bc1 is None else np.ascontiguousarray(bc1, dtype=np.uint64)
        self.normals = None if normals is None else np.ascontiguousarray(normals, dtype=np.uint64)
        s = OrcScene()
        s.nodes = self.nodes.ctypes.data
        s.node_count = self.nodes.size // 2
        s.brick_data = self.brick_data.ctypes.data if self.brick_data.size else None
        s.brick_count = self.brick_data.size // 512
        s.bc1 = self.bc1.ctypes.data if (self.bc1 is not None and self.bc1.size) else None
        s.grid_min[:] = [float(x) for x in grid_min]
        s.grid_max[:] = [float(x) for x in grid_max]
        s.max_levels = int(max_levels)
        s.bricks_per_axis = int(bricks_per_axis)
        s.normals = self.normals.ctypes.data if (self.normals is not None and self.normals.size) else None
        self.palette = None if palette is None else np.ascontiguousarray(palette, dtype=np.float32).reshape(256, 3)
        s.palette = self.palette.ctypes.data if self.palette is not None else None
        self.s = s

    @classmethod
    def from_built(cls, built) -> "OracleScene":
        cfg = built.config
        return cls(built.nodes, built.brick_data, built.bc1, list(cfg.grid_min), list(cfg.grid_max),
                   cfg.user_max_levels, cfg.bricks_per_axis, getattr(built, "normals", None),
                   getattr(built, "palette", None))


def trace(scene: OracleScene, rays: np.ndarray, nthreads: int = 1):
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    out = np.zeros(len(rays), HIT_DTYPE)
    counts = np.zeros(5, np.uint64)
    if nthreads > 1:
        lib().orc_trace_mt(C.byref(scene.s), rays.ctypes.data, len(rays), out.ctypes.data, counts.ctypes.data,
                           nthreads)
    else:
        lib().orc_trace(C.byref(scene.s), rays.ctypes.data, len(rays), out.ctypes.data, counts.ctypes.data)
    return out, counts


def trace_lod(scene: OracleScene, rays: np.ndarray, ray_orig_size: float, ray_dir_size: float,
              nthreads: int = 1) -> np.ndarray:
    """castRayWithLOD over a batch (oracle/vx_oracle.c orc_trace_lod)."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    out = np.zeros(len(rays), HIT_DTYPE)
    lib().orc_trace_lod(C.byref(scene.s), rays.ctypes.data, len(rays), ray_orig_size, ray_dir_size,
                        out.ctypes.data, max(1, nthreads))
    return out


CPU_NO_UNIT_TMAX, CPU_OCTANT_EPS, CPU_CLIP_BRICK, CPU_LOOKUP_INSIDE = 1, 2, 4, 8
CPU_ALL_ADJUST = 15


def trace_cpu_reference(scene: OracleScene, rays: np.ndarray, adjust: int = 0):
    """The reference's CPU castRay restated for single-level worlds (vx_cpu_reference.c), with the
    CPU_* adjustments in `adjust`: (records, status per ray: 1 hit, 0 miss, -1 outside the scope)."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    out = np.zeros(len(rays), HIT_DTYPE)
    status = np.zeros(len(rays), np.int32)
    if lib().orc_trace_cpu_reference(C.byref(scene.s), rays.ctypes.data, len(rays), int(adjust), out.ctypes.data,
                                     status.ctypes.data) != 0:
        raise ValueError("trace_cpu_reference: only worlds of at most 2 bricks per axis in voxel units")
    return out, status


def camera_rays(cam_floats: np.ndarray, w: int, h: int, tmax: float) -> np.ndarray:
    cam = np.ascontiguousarray(cam_floats, dtype=np.float32)
    assert cam.size == 16
    out = np.zeros(w * h, RAY_DTYPE)
    lib().orc_camera_rays(cam.ctypes.data, w, h, float(tmax), out.ctypes.data)
    return out


def bc1_decode(block: int, texel: int) -> np.ndarray:
    rgb = np.zeros(3, np.float32)
    lib().orc_bc1_decode(int(block), int(texel), rgb.ctypes.data)
    return rgb


def bc1_decode_shader(block: int, texel: int) -> np.ndarray:
    rgb = np.zeros(3, np.float32)
    lib().orc_bc1_decode_shader(int(block), int(texel), rgb.ctypes.data)
    return rgb


def dxt_normal_decode(block_a: int, block_b: int, texel: int) -> np.ndarray:
    n = np.zeros(3, np.float32)
    lib().orc_dxt_normal_decode(int(block_a), int(block_b), int(texel), n.ctypes.data)
    return n


def render(scene: OracleScene, cam_floats: np.ndarray, w: int, h: int, tmax: float, nthreads: int = 8) -> np.ndarray:
    """vx_render's RGBA8 image (uint32 per pixel, R in bits 0-7), row-major."""
    cam = np.ascontiguousarray(cam_floats, dtype=np.float32)
    out = np.zeros(w * h, np.uint32)
    lib().orc_render(C.byref(scene.s), cam.ctypes.data, w, h, float(tmax), out.ctypes.data, nthreads)
    return out


def rgba8_decode(word: int) -> np.ndarray:
    rgb = np.zeros(3, np.float32)
    lib().orc_rgba8_decode(int(word), rgb.ctypes.data)
    return rgb


def material_color(material_id: int) -> np.ndarray:
    """getMaterialColor (shaders/Materials.glsl:18-62)."""
    rgb = np.zeros(3, np.float32)
    lib().orc_material_color(int(material_id), rgb.ctypes.data)
    return rgb


def material_id_to_color(material_id: int) -> np.ndarray:
    """MaterialIdToColor (libraries/SVO/src/SVORebuild.cpp:68-120)."""
    rgb = np.zeros(3, np.float32)
    lib().orc_material_id_to_color(int(material_id), rgb.ctypes.data)
    return rgb


def material_palette() -> np.ndarray:
    """material_color of ids 0..255, (256, 3) float32."""
    return np.stack([material_color(i) for i in range(256)])


def setup_ray(scene: OracleScene, ray) -> tuple | None:
    r = np.ascontiguousarray(np.asarray(ray, dtype=RAY_DTYPE).reshape(1))
    ol, dl, du = np.zeros(3, np.float32), np.zeros(3, np.float32), np.zeros(3, np.float32)
    ts, te = np.zeros(1, np.float32), np.zeros(1, np.float32)
    ok = lib().orc_setup_ray(C.byref(scene.s), r.ctypes.data, ol.ctypes.data, dl.ctypes.data, du.ctypes.data,
                             ts.ctypes.data, te.ctypes.data)
    if not ok:
        return None
    return ol, dl, du, float(ts[0]), float(te[0])


def bruteforce(scene: OracleScene, grid: np.ndarray, rays: np.ndarray, eps: float = 2e-5):
    """grid[z, y, x] dense words; returns (vox (n,3) int32 xyz or -1, t float64, ambiguous bool)."""
    rays = np.ascontiguousarray(rays, dtype=RAY_DTYPE)
    g = np.ascontiguousarray(grid, dtype=np.uint32)
    n = len(rays)
    vox = np.zeros((n, 3), np.int32)
    t = np.zeros(n, np.float64)
    amb = np.zeros(n, np.int32)
    lib().orc_bruteforce(C.byref(scene.s), g.ctypes.data, rays.ctypes.data, n, float(eps), vox.ctypes.data,
                         t.ctypes.data, amb.ctypes.data)
    return vox, t, amb.astype(bool)

"""vbvs--vixen_amd: MI355X (gfx950) backend for VIXEN's voxel ray-cast hot path.

Python mirror of the reference's ray-cast surface over the C ABI in
``libvixen_hip.so`` (declared in ``include/vixen_hip.h`` and
``include/vixen_build.h``):

* ``build_from_dense`` / ``build_from_bricks`` / ``gen_*`` -- the scene build path
  (``LaineKarrasOctree::rebuild``, libraries/SVO/src/SVORebuild.cpp:228-732, and
  ``VoxelSceneCacher::BuildOctree``, libraries/CashSystem/src/VoxelSceneCacher.cpp:463-628);
  ``read_cache`` / ``write_cache`` -- VIXEN's VSCA v1 scene cache
  (libraries/CashSystem/src/VoxelSceneCacher.cpp:179-364);
  ``bc1_encode_device`` -- the rebuild's BC1 pass on a device brick pool,
  bit-identical to the host builder's pool;
* ``Scene.trace`` -- the batch form of ``ISVOStructure::castRay``
  (libraries/SVO/include/ISVOStructure.h:122-126);
* ``Scene.render`` -- the compute pipeline's per-frame dispatch
  (shaders/VoxelRayMarch_Compressed.comp:493-594);
* ``Scene.trace_lod`` / ``lod_from_camera`` -- ``castRayWithLOD`` and its
  ``LODParameters`` (libraries/SVO/include/SVOLOD.h:50-195);
* ``Batch`` -- every scene of a set in one persistent launch, with its
  autotuner and hipGraph capture;
* ``Scene.albedo_forward`` / ``albedo_backward`` -- the differentiable albedo
  render and its hand-written adjoint (``vx_torch.so`` wraps them in a
  ``torch.autograd.Function``);
* ``LaineKarrasOctreeHip.castRay`` -- one-ray convenience with the reference's
  argument meaning (origin, direction, tMin, tMax) returning a ``RayHit``.

PyTorch is used only for device memory and streams. Every call goes through the
HIP library; there is no CPU fallback: if the library is missing, importing this
package raises.
"""
from __future__ import annotations

import ctypes as C
import math
import threading
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "libvixen_hip.so"

VX_OK = 0
STATUS_NAMES = {0: "VX_OK", 1: "VX_ERR_INVALID_ARG", 2: "VX_ERR_INVALID_SCENE", 3: "VX_ERR_HIP",
                4: "VX_ERR_OUT_OF_MEMORY", 5: "VX_ERR_UNSUPPORTED"}

HIT_FLAG = 1
LOD_FLAG = 1 << 6  # VX_HIT_FLAG_LOD


class VixenError(RuntimeError):
    """A C-ABI call returned a non-zero vx_status."""

    def __init__(self, status: int, where: str, msg: str):
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}: {msg}")
        self.status = status


# --------------------------------------------------------------------------
# ABI structures (must match include/vixen_hip.h; sizes are checked at load)
# --------------------------------------------------------------------------
class VxOctreeConfig(C.Structure):
    _fields_ = [("esvo_max_scale", C.c_int32), ("user_max_levels", C.c_int32),
                ("brick_depth_levels", C.c_int32), ("brick_size", C.c_int32),
                ("min_esvo_scale", C.c_int32), ("brick_esvo_scale", C.c_int32),
                ("bricks_per_axis", C.c_int32), ("_padding1", C.c_int32),
                ("grid_min", C.c_float * 3), ("_padding2", C.c_float),
                ("grid_max", C.c_float * 3), ("_padding3", C.c_float),
                ("local_to_world", C.c_float * 16), ("world_to_local", C.c_float * 16),
                ("_padding4", C.c_float * 16)]


class VxCamera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("time", C.c_float), ("dir", C.c_float * 3), ("fov_deg", C.c_float),
                ("up", C.c_float * 3), ("aspect", C.c_float), ("right", C.c_float * 3),
                ("debug_mode", C.c_int32)]

    def as_floats(self) -> np.ndarray:
        """The 64-byte push constant as 16 float32 words (debug_mode bit-cast)."""
        return np.frombuffer(bytes(self), dtype=np.float32).copy()

    def to_dict(self) -> dict:
        return {k: (list(getattr(self, k)) if k in ("pos", "dir", "up", "right") else getattr(self, k))
                for k, _ in self._fields_}


class VxSceneDesc(C.Structure):
    _fields_ = [("esvo_nodes", C.c_void_p), ("node_count", C.c_uint64),
                ("brick_data", C.c_void_p), ("brick_count", C.c_uint64),
                ("bc1_colors", C.c_void_p), ("config", C.POINTER(VxOctreeConfig)), ("dxt_normals", C.c_void_p),
                ("material_palette", C.c_void_p)]


class VxCacheEntryInfo(C.Structure):
    _fields_ = [("key", C.c_uint64), ("scene_type", C.c_uint32), ("resolution", C.c_uint32), ("density", C.c_float),
                ("seed", C.c_uint32), ("node_count", C.c_uint32), ("brick_count", C.c_uint32),
                ("solid_voxel_count", C.c_uint32), ("materials", C.c_void_p), ("material_count", C.c_uint64),
                ("brick_grid_lookup", C.c_void_p), ("brick_grid_lookup_count", C.c_uint64),
                ("world_grid_size", C.c_float)]


class VxCacheEntryIn(C.Structure):
    _fields_ = [("scene_type", C.c_uint32), ("resolution", C.c_uint32), ("density", C.c_float), ("seed", C.c_uint32),
                ("scene", C.POINTER(VxSceneDesc)), ("materials", C.c_void_p), ("material_count", C.c_uint64),
                ("brick_grid_lookup", C.c_void_p), ("brick_grid_lookup_count", C.c_uint64),
                ("solid_voxel_count", C.c_uint32), ("world_grid_size", C.c_float)]


class VxTuning(C.Structure):
    _fields_ = [("waves_per_group", C.c_uint32), ("chunk", C.c_uint32), ("lds_depth", C.c_uint32),
                ("refill_min", C.c_uint32)]

    def to_dict(self) -> dict:
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class VxBuiltScene(C.Structure):
    _fields_ = [("esvo_nodes", C.POINTER(C.c_uint32)), ("node_count", C.c_uint64),
                ("brick_data", C.POINTER(C.c_uint32)), ("brick_count", C.c_uint64),
                ("bc1_colors", C.POINTER(C.c_uint64)), ("brick_coords", C.POINTER(C.c_uint32)),
                ("internal_count", C.c_uint64), ("config", VxOctreeConfig)]


class VxLod(C.Structure):
    """LODParameters (libraries/SVO/include/SVOLOD.h:50-177)."""
    _fields_ = [("ray_orig_size", C.c_float), ("ray_dir_size", C.c_float)]

    def enabled(self) -> bool:
        return self.ray_dir_size > 0.0


RAY_DTYPE = np.dtype([("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("flags", "<u4"), ("t", "<f4"), ("leaf_node", "<u4"), ("brick", "<u4"),
                      ("voxel", "<u4"), ("rgb", "<f4", 3)])
assert RAY_DTYPE.itemsize == 32 and HIT_DTYPE.itemsize == 32
assert C.sizeof(VxOctreeConfig) == 256 and C.sizeof(VxCamera) == 64


def _load() -> C.CDLL:
    # PyTorch ships its own HIP runtime under the same soname (libamdhip64.so.7).
    # Import it first so this library binds to the runtime torch already uses:
    # one runtime per process, so torch tensors and streams are valid here.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build() or "
                          f"python vbvs--vixen_amd/build.py")
    lib = C.CDLL(str(LIB_PATH))
    vp, u64, u32, i32, st = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int32, C.c_int
    sig = {
        "vx_last_error": ([], C.c_char_p),
        "vx_version": ([], C.c_char_p),
        "vx_ctx_create": ([C.c_int, C.POINTER(vp)], st),
        "vx_ctx_destroy": ([vp], st),
        "vx_scene_create": ([vp, C.POINTER(VxSceneDesc), C.POINTER(vp)], st),
        "vx_scene_validate": ([C.POINTER(VxSceneDesc)], st),
        "vx_scene_retain": ([vp], st),
        "vx_scene_revalidate": ([vp], st),
        "vx_scene_create_device": ([vp, C.POINTER(VxSceneDesc), C.POINTER(vp)], st),
        "vx_scene_destroy": ([vp], st),
        "vx_scene_info": ([vp, C.POINTER(u64), C.POINTER(u64), C.POINTER(i32), C.POINTER(i32)], st),
        "vx_trace": ([vp, vp, u64, vp, vp], st),
        "vx_trace_counters": ([vp, vp, u64, vp, vp, vp], st),
        "vx_trace_lod": ([vp, vp, u64, C.POINTER(VxLod), vp, vp], st),
        "vx_lod_from_camera": ([C.c_float, i32], VxLod),
        "vx_lod_from_camera_near": ([C.c_float, i32, C.c_float], VxLod),
        "vx_lod_with_bias": ([VxLod, C.c_float], VxLod),
        "vx_material_color": ([u32, vp], None),
        "vx_material_palette": ([vp], None),
        "vx_material_id_to_color": ([u32, vp], None),
        "vx_render": ([vp, C.POINTER(VxCamera), u32, u32, C.c_float, vp, vp, vp], st),
        "vx_camera_rays": ([C.POINTER(VxCamera), u32, u32, C.c_float, vp, vp], st),
        "vx_bounce_rays": ([vp, vp, vp, u64, u32, vp, vp], st),
        "vx_build_from_dense": ([vp, u32, C.c_int, C.c_int, C.POINTER(C.POINTER(VxBuiltScene))], st),
        "vx_build_from_bricks": ([vp, vp, u64, u32, C.c_int, C.c_int, C.POINTER(C.POINTER(VxBuiltScene))], st),
        "vx_built_scene_free": ([C.POINTER(VxBuiltScene)], None),
        "vx_bc1_encode_block": ([vp], u64),
        "vx_bc1_encode_device": ([vp, u64, vp, vp, vp], st),
        "vx_encode_normals": ([vp, u64, C.c_int, vp], st),
        "vx_cache_read": ([C.c_char_p, C.POINTER(vp)], st),
        "vx_cache_entry_count": ([vp, C.POINTER(u32)], st),
        "vx_cache_entry": ([vp, u32, C.POINTER(VxCacheEntryInfo), C.POINTER(VxSceneDesc)], st),
        "vx_cache_free": ([vp], st),
        "vx_cache_write": ([C.c_char_p, C.POINTER(VxCacheEntryIn), u32], st),
        "vx_dxt_normal_encode_block": ([vp, vp], None),
        "vx_geometric_normal": ([vp, C.c_int, C.c_int, C.c_int, vp], None),
        "vx_scene_set_lds_depth": ([vp, u32], st),
        "vx_batch_create": ([vp, C.POINTER(vp), u32, C.POINTER(vp)], st),
        "vx_batch_destroy": ([vp], st),
        "vx_trace_batched": ([vp, vp, C.POINTER(u64), vp, C.POINTER(VxTuning), vp], st),
        "vx_batch_autotune": ([vp, vp, C.POINTER(u64), vp, C.POINTER(VxTuning), C.POINTER(C.c_float), vp], st),
        "vx_batch_graph_create": ([vp, vp, C.POINTER(u64), vp, C.POINTER(VxTuning), C.POINTER(vp)], st),
        "vx_graph_launch": ([vp, vp], st),
        "vx_graph_destroy": ([vp], st),
        "vx_scene_voxel_count": ([vp, C.POINTER(u64)], st),
        "vx_albedo_forward": ([vp, C.POINTER(VxCamera), u32, u32, vp, vp, vp, vp, vp, vp], st),
        "vx_albedo_backward": ([vp, C.POINTER(VxCamera), u32, u32, vp, vp, vp, vp, vp, u32, vp], st),
        "vx_albedo_backward_noalloc": ([vp, C.POINTER(VxCamera), u32, u32, vp, vp, vp, vp, vp, u32, vp], st),
        "vx_gen_cornell": ([u32, u64, vp, C.POINTER(VxCamera)], st),
        "vx_gen_cave": ([u32, u64, C.c_double, vp, C.POINTER(VxCamera)], st),
        "vx_gen_urban": ([u32, u64, C.c_double, vp, C.POINTER(VxCamera)], st),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


LIB = _load()

# Every symbol the headers declare (the CPU suite checks the library exports them all).
EXPORTED = ("vx_last_error", "vx_version", "vx_ctx_create", "vx_ctx_destroy", "vx_scene_create", "vx_scene_validate",
            "vx_scene_retain",
            "vx_scene_destroy", "vx_scene_info", "vx_trace", "vx_trace_counters", "vx_render",
            "vx_camera_rays", "vx_build_from_dense", "vx_build_from_bricks", "vx_built_scene_free", "vx_bc1_encode_block",
            "vx_gen_cornell", "vx_gen_cave", "vx_gen_urban", "vx_scene_set_lds_depth", "vx_batch_create",
            "vx_batch_destroy", "vx_trace_batched", "vx_batch_autotune", "vx_batch_graph_create", "vx_graph_launch",
            "vx_graph_destroy", "vx_scene_voxel_count", "vx_albedo_forward", "vx_albedo_backward", "vx_bounce_rays",
            "vx_encode_normals", "vx_dxt_normal_encode_block", "vx_geometric_normal", "vx_cache_read",
            "vx_cache_entry_count", "vx_cache_entry", "vx_cache_free", "vx_cache_write", "vx_trace_lod",
            "vx_lod_from_camera", "vx_lod_from_camera_near", "vx_lod_with_bias", "vx_scene_create_device",
            "vx_albedo_backward_noalloc", "vx_material_color", "vx_material_palette", "vx_material_id_to_color",
            "vx_scene_revalidate", "vx_bc1_encode_device")


def lod_from_camera(fov_y: float, screen_height: int, near_plane: float | None = None) -> VxLod:
    """LODParameters::fromCamera / fromCameraWithNearPlane (fov_y in radians)."""
    if near_plane is None:
        return LIB.vx_lod_from_camera(fov_y, screen_height)
    return LIB.vx_lod_from_camera_near(fov_y, screen_height, near_plane)


def lod_with_bias(lod: VxLod, bias: float) -> VxLod:
    """LODParameters::withBias: both cone sizes times 2^bias."""
    return LIB.vx_lod_with_bias(lod, bias)


def _check(status: int, where: str) -> None:
    if status != VX_OK:
        raise VixenError(status, where, (LIB.vx_last_error() or b"").decode())


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# --------------------------------------------------------------------------
# Host-side scene build and generators
# --------------------------------------------------------------------------
@dataclass
class BuiltScene:
    """Scene buffers in VIXEN's layout (copied out of the C builder)."""
    nodes: np.ndarray        # (node_count, 2) uint32
    brick_data: np.ndarray   # (brick_count, 512) uint32
    bc1: np.ndarray | None   # (brick_count, 32) uint64
    brick_coords: np.ndarray  # (brick_count, 3) int32 grid cell per brick id
    internal_count: int
    config: VxOctreeConfig
    normals: np.ndarray | None = None  # (brick_count, 64) uint64 DXT normal blocks, see encode_normals
    palette: np.ndarray | None = None  # (256, 3) float32 colour per material id (brick word & 0xFF), or None

    @property
    def N(self) -> int:
        return int(self.config.bricks_per_axis) * 8

    @property
    def node_count(self) -> int:
        return int(self.nodes.shape[0])

    @property
    def brick_count(self) -> int:
        return int(self.brick_data.shape[0])

    def device_bytes(self) -> int:
        nb = self.brick_count
        return (self.node_count * 8 + nb * 2048 + nb * 64 + (nb * 256 if self.bc1 is not None else 0) +
                (nb * 512 if self.normals is not None else 0))

    def encode_normals(self, nthreads: int = 0) -> "BuiltScene":
        """Adds the compressed shader's normal pool (vx_encode_normals) and returns self."""
        out = np.zeros((self.brick_count, 64), np.uint64)
        if self.brick_count:
            _check(LIB.vx_encode_normals(_ptr(self.brick_data), self.brick_count, nthreads, _ptr(out)),
                   "vx_encode_normals")
        self.normals = out
        return self


def make_camera(pos, look_dir, fov_deg: float, aspect: float) -> VxCamera:
    """Pinhole camera: right = dir x worldUp (y up, z when looking straight up/down), up = right x dir."""
    d = np.asarray(look_dir, np.float64)
    d = d / np.linalg.norm(d)
    wu = np.array([0.0, 0.0, 1.0]) if abs(d[1]) > 0.99 else np.array([0.0, 1.0, 0.0])
    r = np.cross(d, wu)
    r /= np.linalg.norm(r)
    u = np.cross(r, d)
    cam = VxCamera()
    cam.pos[:] = [float(x) for x in np.asarray(pos, np.float64)]
    cam.dir[:], cam.up[:], cam.right[:] = [float(x) for x in d], [float(x) for x in u], [float(x) for x in r]
    cam.fov_deg, cam.aspect = float(fov_deg), float(aspect)
    return cam


def seeded_views(grid: np.ndarray, seed: int, n: int, fov_deg: float = 70.0, aspect: float = 16 / 9,
                 clearance: int = 2) -> list[VxCamera]:
    """n cameras at seeded empty cells (a (2c+1)^3 empty neighbourhood) with seeded look directions."""
    rng = np.random.default_rng(seed)
    N = grid.shape[0]
    solid = grid != 0
    cams = []
    while len(cams) < n:
        x, y, z = rng.integers(clearance, N - clearance, 3)
        c = clearance
        if solid[z - c:z + c + 1, y - c:y + c + 1, x - c:x + c + 1].any():
            continue
        v = rng.normal(size=3)
        cams.append(make_camera((x + 0.5, y + 0.5, z + 0.5), v, fov_deg, aspect))
    return cams


# what the builders encode BC1 blocks from (VX_BC1_* of vixen_build.h)
BC1_NONE, BC1_FROM_RGBA8, BC1_FROM_MATERIAL = 0, 1, 2


def _bc1_mode(bc1: bool, materials: bool) -> int:
    return (BC1_FROM_MATERIAL if materials else BC1_FROM_RGBA8) if bc1 else BC1_NONE


def build_from_dense(grid: np.ndarray, bc1: bool = True, nthreads: int = 0, materials: bool = False) -> BuiltScene:
    """Dense grid grid[z, y, x] (0 = empty) -> octree buffers. The words are RGBA8 colours, or with
    materials=True VIXEN's material ids: BC1 blocks (bc1=True) then come from MaterialIdToColor, as VIXEN's
    rebuild encodes them; without BC1 the scene carries getMaterialColor's palette (the uncompressed
    pipeline's colour)."""
    grid = np.ascontiguousarray(grid, dtype=np.uint32)
    N = grid.shape[0]
    assert grid.shape == (N, N, N)
    out = C.POINTER(VxBuiltScene)()
    _check(LIB.vx_build_from_dense(_ptr(grid), N, _bc1_mode(bc1, materials), nthreads, C.byref(out)),
           "vx_build_from_dense")
    built = _take_built(out, bc1)
    if materials and not bc1:
        built.palette = material_palette()
    return built


def build_from_bricks(cells: np.ndarray, voxels: np.ndarray, N: int, bc1: bool = True,
                      nthreads: int = 0, materials: bool = False) -> BuiltScene:
    """Sparse world -> octree buffers: cells [n, 3] brick cells (x, y, z) in [0, N/8),
    voxels [n, 8, 8, 8] as voxels[i, z, y, x] (0 = empty), RGBA8 or material ids as in build_from_dense."""
    cells = np.ascontiguousarray(cells, dtype=np.uint32).reshape(-1, 3)
    voxels = np.ascontiguousarray(voxels, dtype=np.uint32).reshape(-1, 512)
    assert len(cells) == len(voxels)
    out = C.POINTER(VxBuiltScene)()
    _check(LIB.vx_build_from_bricks(_ptr(cells), _ptr(voxels), len(cells), N, _bc1_mode(bc1, materials), nthreads,
                                    C.byref(out)), "vx_build_from_bricks")
    built = _take_built(out, bc1)
    if materials and not bc1:
        built.palette = material_palette()
    return built


def material_palette() -> np.ndarray:
    """getMaterialColor (shaders/Materials.glsl:18-62) of ids 0..255, as (256, 3) float32."""
    out = np.zeros((256, 3), np.float32)
    LIB.vx_material_palette(_ptr(out))
    return out


def material_color(material_id: int) -> np.ndarray:
    """getMaterialColor (shaders/Materials.glsl:18-62) of one material id."""
    out = np.zeros(3, np.float32)
    LIB.vx_material_color(int(material_id), _ptr(out))
    return out


def material_id_to_color(material_id: int) -> np.ndarray:
    """MaterialIdToColor (libraries/SVO/src/SVORebuild.cpp:68-120) of one material id."""
    out = np.zeros(3, np.float32)
    LIB.vx_material_id_to_color(int(material_id), _ptr(out))
    return out


def _take_built(out, bc1: bool) -> BuiltScene:
    """Copies a vx_built_scene into numpy arrays and frees it."""
    try:
        s = out.contents
        nc, nb = int(s.node_count), int(s.brick_count)
        nodes = np.ctypeslib.as_array(s.esvo_nodes, shape=(nc * 2,)).reshape(nc, 2).copy()
        bd = (np.ctypeslib.as_array(s.brick_data, shape=(nb * 512,)).reshape(nb, 512).copy() if nb
              else np.zeros((0, 512), np.uint32))
        b1 = None
        if bc1:
            b1 = (np.ctypeslib.as_array(s.bc1_colors, shape=(nb * 32,)).reshape(nb, 32).copy() if nb
                  else np.zeros((0, 32), np.uint64))
        packed = (np.ctypeslib.as_array(s.brick_coords, shape=(nb,)).copy() if nb else np.zeros(0, np.uint32))
        coords = np.stack([packed & 1023, (packed >> 10) & 1023, (packed >> 20) & 1023], axis=1).astype(np.int32)
        cfg = VxOctreeConfig()
        C.memmove(C.addressof(cfg), C.addressof(s.config), C.sizeof(cfg))
        return BuiltScene(nodes, bd, b1, coords, int(s.internal_count), cfg)
    finally:
        LIB.vx_built_scene_free(out)


def bc1_encode_block(colors: np.ndarray) -> int:
    c = np.ascontiguousarray(colors, dtype=np.float32).reshape(16, 3)
    return int(LIB.vx_bc1_encode_block(_ptr(c)))


def material_id_palette() -> np.ndarray:
    """float32 [512, 3]: MaterialIdToColor of ids 0-511, the palette with which bc1_encode_device reproduces the
    builder's material-id BC1 pool (entries 256-511 colour the ids of 256 and above by id mod 256)."""
    return np.stack([material_id_to_color(i) for i in range(512)]).astype(np.float32)


def bc1_encode_device(brick_data, palette=None, out=None, stream=None):
    """The BC1 pool of a device brick pool (torch int32/uint32 tensor of brick_count x 512 words, RGBA8 colours, or
    material ids when `palette` (a device float32 [512, 3] tensor, see material_id_palette) is given): an int64
    tensor [brick_count, 32], bit-identical to the host builder's. vx_bc1_encode_device."""
    torch = _torch()
    if brick_data.dtype not in (torch.int32, torch.uint32) or brick_data.numel() % 512:
        raise ValueError("brick_data must be 32-bit words, 512 per brick")
    nb = brick_data.numel() // 512
    device = brick_data.device.index if brick_data.is_cuda else -1
    _dev_buf(brick_data, "brick_data", device)
    if palette is not None:
        _dev_buf(palette, "palette", device, 512 * 3 * 4)
        if palette.dtype != torch.float32:
            raise ValueError("palette must be float32 [512, 3]")
    if out is None:
        out = torch.empty((nb, 32), dtype=torch.int64, device=brick_data.device)
    _dev_buf(out, "out", device, nb * 32 * 8)
    with torch.cuda.device(device):  # the C call runs on the current device
        _check(LIB.vx_bc1_encode_device(brick_data.data_ptr(), nb,
                                        palette.data_ptr() if palette is not None else None, out.data_ptr(),
                                        _stream_ptr(stream)), "vx_bc1_encode_device")
    return out


def dxt_normal_encode_block(normals: np.ndarray) -> tuple[int, int]:
    """DXTNormalCompressor::encodeBlockTyped of 16 normals: (blockA, blockB)."""
    n = np.ascontiguousarray(normals, dtype=np.float32).reshape(16, 3)
    out = np.zeros(2, np.uint64)
    LIB.vx_dxt_normal_encode_block(_ptr(n), _ptr(out))
    return int(out[0]), int(out[1])


def _scene_desc(built: BuiltScene) -> tuple[VxSceneDesc, VxOctreeConfig]:
    desc = VxSceneDesc()
    desc.esvo_nodes = _ptr(built.nodes)
    desc.node_count = built.node_count
    desc.brick_data = _ptr(built.brick_data) if built.brick_count else None
    desc.brick_count = built.brick_count
    desc.bc1_colors = _ptr(built.bc1) if (built.bc1 is not None and built.brick_count) else None
    desc.dxt_normals = _ptr(built.normals) if (built.normals is not None and built.brick_count) else None
    if built.palette is not None:
        pal = np.ascontiguousarray(built.palette, dtype=np.float32)
        if pal.shape != (256, 3):
            raise ValueError("palette: a (256, 3) float32 array of RGB per material id")
        built.palette = pal
        desc.material_palette = _ptr(pal)
    cfg = built.config
    desc.config = C.pointer(cfg)
    return desc, cfg


def brick_grid_lookup(built: BuiltScene) -> np.ndarray:
    """VIXEN's brickGridLookup: bx + B*by + B*B*bz -> brick id, ~0u where no brick."""
    B = int(built.config.bricks_per_axis)
    out = np.full(B * B * B, 0xFFFFFFFF, np.uint32)
    c = built.brick_coords.astype(np.int64)
    out[c[:, 0] + B * c[:, 1] + B * B * c[:, 2]] = np.arange(built.brick_count, dtype=np.uint32)
    return out


SCENE_TYPES = {"cornell": 0, "noise": 1, "tunnels": 2, "cityscape": 3, "custom": 255}
# VoxelSceneCacher::BuildOctree's WORLD_GRID_SIZE (VoxelSceneCacher.cpp:605-607)
WORLD_GRID_SIZE = 10.0


def shader_frame(built: BuiltScene, world_grid_size: float = WORLD_GRID_SIZE) -> BuiltScene:
    """The scene with its traced world box set to the compute shader's frame, [0, world_grid_size]^3
    (VoxelRayMarch_Compressed.comp:352-354 maps world space to the unit cube with worldToLocal =
    scale(1 / worldGridSize)); the buffers are shared, only the config is copied."""
    cfg = VxOctreeConfig.from_buffer_copy(bytes(built.config))
    for i in range(3):
        cfg.grid_min[i], cfg.grid_max[i] = 0.0, float(world_grid_size)
    return BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count, cfg,
                      built.normals, built.palette)


def write_cache(path, entries) -> None:
    """Writes a VSCA v1 scene cache. entries: dicts with built (BuiltScene), scene_type (int or name),
    resolution, density, seed, and optionally materials (bytes/uint8 array of 32-byte records),
    solid_voxel_count and world_grid_size (default VIXEN's 10, written after the 256-byte config);
    the brick grid lookup is derived from built.brick_coords."""
    keep, ins = [], (VxCacheEntryIn * max(1, len(entries)))()
    for i, e in enumerate(entries):
        built = e["built"]
        desc, cfg = _scene_desc(built)
        lookup = brick_grid_lookup(built)
        mats = np.frombuffer(bytes(e.get("materials", b"")), np.uint8)
        st = e.get("scene_type", 255)
        ins[i].scene_type = SCENE_TYPES[st] if isinstance(st, str) else int(st)
        ins[i].resolution = int(e.get("resolution", built.N))
        ins[i].density = float(e.get("density", 0.5))
        ins[i].seed = int(e.get("seed", 0))
        ins[i].scene = C.pointer(desc)
        ins[i].materials = _ptr(mats) if mats.size else None
        ins[i].material_count = mats.size // 32
        ins[i].brick_grid_lookup = _ptr(lookup)
        ins[i].brick_grid_lookup_count = lookup.size
        ins[i].solid_voxel_count = int(e.get("solid_voxel_count", int((built.brick_data != 0).sum())))
        ins[i].world_grid_size = float(e.get("world_grid_size", WORLD_GRID_SIZE))
        keep.append((desc, cfg, lookup, mats, built))
    _check(LIB.vx_cache_write(str(path).encode(), ins, len(entries)), "vx_cache_write")


def read_cache(path) -> list[tuple[dict, BuiltScene]]:
    """Reads a VSCA v1 scene cache: per entry (info dict, BuiltScene with copies of its buffers)."""
    h = C.c_void_p()
    _check(LIB.vx_cache_read(str(path).encode(), C.byref(h)), "vx_cache_read")
    try:
        n = C.c_uint32()
        _check(LIB.vx_cache_entry_count(h, C.byref(n)), "vx_cache_entry_count")
        out = []
        for i in range(n.value):
            info, desc = VxCacheEntryInfo(), VxSceneDesc()
            _check(LIB.vx_cache_entry(h, i, C.byref(info), C.byref(desc)), "vx_cache_entry")
            nn, nb = int(desc.node_count), int(desc.brick_count)

            def arr(ptr, count, dtype):
                if not ptr or count == 0:
                    return None
                return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(np.ctypeslib.as_ctypes_type(dtype))),
                                             (count,)).copy()
            nodes = arr(desc.esvo_nodes, nn * 2, np.uint32).reshape(nn, 2)
            bricks = arr(desc.brick_data, nb * 512, np.uint32)
            bricks = bricks.reshape(nb, 512) if bricks is not None else np.zeros((0, 512), np.uint32)
            bc1 = arr(desc.bc1_colors, nb * 32, np.uint64)
            nrm = arr(desc.dxt_normals, nb * 64, np.uint64)
            lookup = arr(info.brick_grid_lookup, int(info.brick_grid_lookup_count), np.uint32)
            cfg = VxOctreeConfig.from_buffer_copy(bytes(desc.config.contents))
            B = int(cfg.bricks_per_axis)
            coords = np.zeros((nb, 3), np.int32)
            if lookup is not None and lookup.size == B ** 3:
                cells = np.nonzero(lookup != 0xFFFFFFFF)[0]
                ids = lookup[cells]
                ok = ids < nb
                coords[ids[ok], 0] = cells[ok] % B
                coords[ids[ok], 1] = (cells[ok] // B) % B
                coords[ids[ok], 2] = cells[ok] // (B * B)
            mats = arr(info.materials, int(info.material_count) * 32, np.uint8)
            meta = {"key": int(info.key), "scene_type": int(info.scene_type), "resolution": int(info.resolution),
                    "density": float(info.density), "seed": int(info.seed), "node_count": int(info.node_count),
                    "brick_count": int(info.brick_count), "solid_voxel_count": int(info.solid_voxel_count),
                    "materials": b"" if mats is None else mats.tobytes(),
                    "brick_grid_lookup": lookup, "world_grid_size": float(info.world_grid_size)}
            built = BuiltScene(nodes, bricks, None if bc1 is None else bc1.reshape(nb, 32), coords,
                               max(0, nn - nb), cfg, None if nrm is None else nrm.reshape(nb, 64))
            out.append((meta, built))
        return out
    finally:
        LIB.vx_cache_free(h)


def geometric_normal(brick: np.ndarray, x: int, y: int, z: int) -> np.ndarray:
    """computeGeometricNormal of voxel (x, y, z) in one brick of 512 words."""
    b = np.ascontiguousarray(brick, dtype=np.uint32).reshape(512)
    out = np.zeros(3, np.float32)
    LIB.vx_geometric_normal(_ptr(b), x, y, z, _ptr(out))
    return out


def _gen(fn, where, N, *args):
    grid = np.empty((N, N, N), dtype=np.uint32)
    cam = VxCamera()
    _check(fn(N, *args, _ptr(grid), C.byref(cam)), where)
    return grid, cam


def gen_cornell(N: int = 64, seed: int = 1) -> tuple[np.ndarray, VxCamera]:
    return _gen(LIB.vx_gen_cornell, "vx_gen_cornell", N, seed)


def gen_cave(N: int = 256, seed: int = 2, occupancy: float = 0.5) -> tuple[np.ndarray, VxCamera]:
    return _gen(LIB.vx_gen_cave, "vx_gen_cave", N, seed, float(occupancy))


def gen_urban(N: int = 1024, seed: int = 3, occupancy: float = 1.0) -> tuple[np.ndarray, VxCamera]:
    return _gen(LIB.vx_gen_urban, "vx_gen_urban", N, seed, float(occupancy))


# --------------------------------------------------------------------------
# Device side
# --------------------------------------------------------------------------
def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("vbvs--vixen_amd: no HIP device visible")
    return torch


def _stream_ptr(stream) -> int | None:
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream) or None


def _dev_buf(t, what: str, device: int, min_bytes: int = 0) -> int:
    """The device address of t after checking it is a contiguous CUDA tensor on `device` holding at least
    min_bytes bytes; ValueError otherwise (before anything is launched)."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what} must be a torch tensor, got {type(t).__name__}")
    if not t.is_cuda or t.device.index != device:
        raise ValueError(f"{what} must live on cuda:{device}, it is on {t.device}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    nbytes = t.numel() * t.element_size()
    if nbytes < min_bytes:
        raise ValueError(f"{what} holds {nbytes} bytes, {min_bytes} needed")
    return t.data_ptr()


def _ray_count(rays, what: str, device: int) -> int:
    _dev_buf(rays, what, device)
    nbytes = rays.numel() * rays.element_size()
    if nbytes % 32:
        raise ValueError(f"{what} holds {nbytes} bytes, not a whole number of 32-byte vx_ray records")
    return nbytes // 32


def _ray_offsets(offsets, n_scenes: int, n_rays: int) -> C.Array:
    """A batched call's ray_offsets as the host array the C ABI reads (n_scenes + 1 entries), after checking
    it has that many entries, never decreases and ends within the n_rays rays; ValueError otherwise."""
    o = np.ascontiguousarray(offsets, dtype=np.uint64).reshape(-1)
    if len(o) != n_scenes + 1:
        raise ValueError(f"ray_offsets needs {n_scenes + 1} entries (scenes + 1), got {len(o)}")
    if np.any(o[1:] < o[:-1]):
        raise ValueError("ray_offsets must be non-decreasing")
    if int(o[-1]) > n_rays:
        raise ValueError(f"ray_offsets ends at ray {int(o[-1])}, the rays hold {n_rays}")
    return (C.c_uint64 * len(o))(*[int(x) for x in o])


def _hits_for(hits, n: int, like, device: int, what: str = "hits"):
    """The caller's hit buffer (checked to hold n records and not to overlap the n input records of `like`)
    or a new one."""
    torch = _torch()
    if hits is None:
        return torch.empty(n * 32, dtype=torch.uint8, device=like.device)
    _dev_buf(hits, what, device, 32 * n)
    a, b = hits.data_ptr(), like.data_ptr()
    if n and a < b + 32 * n and b < a + 32 * n:
        raise ValueError(f"{what} overlaps the input rays")
    return hits


class Context:
    """One device (replaces the Vulkan device/queue around ComputeDispatchNode)."""

    def __init__(self, device: int = 0):
        _torch()
        self.device = device
        h = C.c_void_p()
        _check(LIB.vx_ctx_create(device, C.byref(h)), "vx_ctx_create")
        self._h = h

    def close(self) -> None:
        if self._h:
            LIB.vx_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """A scene uploaded to HBM (VoxelSceneCacher::UploadToGPU)."""

    def __init__(self, ctx: Context, built: BuiltScene):
        self.ctx = ctx
        self.built = built
        desc, self._cfg = _scene_desc(built)
        self._keep, self._versions = (), ()
        h = C.c_void_p()
        _check(LIB.vx_scene_create(ctx._h, C.byref(desc), C.byref(h)), "vx_scene_create")
        self._h = h

    @classmethod
    def from_device(cls, ctx: Context, nodes, brick_data, config: VxOctreeConfig, bc1=None, normals=None,
                    palette=None) -> "Scene":
        """A scene over caller-owned device buffers (vx_scene_create_device), e.g. CUDA tensors holding the
        descriptor pairs, brick words, BC1 blocks and DXT normal blocks. They are used in place: the scene keeps
        them referenced, and they must stay unchanged until close() (the kernels read them without range
        checks; the tree was validated once, at creation). Every launch checks that no tensor was modified in
        place since (torch's version counter) and raises otherwise; after an intended change, revalidate()
        re-checks the tree and rebuilds the derived tables. Each must be a contiguous CUDA tensor on the
        context's device; bc1 and normals must hold 256 and 512 bytes per brick. palette (host, (256, 3)
        float32, e.g. material_palette()) is copied."""
        self = cls.__new__(cls)
        self.ctx, self.built = ctx, None
        dev = ctx.device
        _dev_buf(nodes, "nodes", dev)
        _dev_buf(brick_data, "brick_data", dev)
        brick_count = brick_data.numel() * brick_data.element_size() // 2048
        if bc1 is not None:
            _dev_buf(bc1, "bc1", dev, 256 * brick_count)
        if normals is not None:
            _dev_buf(normals, "normals", dev, 512 * brick_count)
        self._cfg = VxOctreeConfig()
        C.memmove(C.addressof(self._cfg), C.addressof(config), C.sizeof(config))
        pal = None
        if palette is not None:
            pal = np.ascontiguousarray(palette, dtype=np.float32)
            if pal.shape != (256, 3):
                raise ValueError("palette: a (256, 3) float32 array of RGB per material id")
        self._keep = tuple(t for t in (nodes, brick_data, bc1, normals) if t is not None)
        self._versions = tuple(t._version for t in self._keep)
        desc = VxSceneDesc()
        desc.material_palette = _ptr(pal) if pal is not None else None
        desc.esvo_nodes = nodes.data_ptr()
        desc.node_count = nodes.numel() * nodes.element_size() // 8
        desc.brick_data = brick_data.data_ptr() if brick_data.numel() else None
        desc.brick_count = brick_count
        desc.bc1_colors = bc1.data_ptr() if bc1 is not None else None
        desc.dxt_normals = normals.data_ptr() if normals is not None else None
        desc.config = C.pointer(self._cfg)
        h = C.c_void_p()
        _check(LIB.vx_scene_create_device(ctx._h, C.byref(desc), C.byref(h)), "vx_scene_create_device")
        self._h = h
        return self

    def close(self) -> None:
        if self._h:
            LIB.vx_scene_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _live(self) -> C.c_void_p:
        """The handle, after checking that no borrowed device buffer changed since creation or revalidate()."""
        if any(t._version != v for t, v in zip(self._keep, self._versions)):
            raise VixenError(2, "Scene", "a borrowed device buffer was modified in place since the scene was "
                                         "validated; call revalidate() first")
        return self._h

    def revalidate(self) -> None:
        """vx_scene_revalidate: re-checks the tree of the (possibly modified) borrowed buffers and rebuilds the
        derived occupancy tables; launches work again once it succeeds."""
        _check(LIB.vx_scene_revalidate(self._h), "vx_scene_revalidate")
        self._versions = tuple(t._version for t in self._keep)

    # rays/hits are contiguous CUDA tensors (any dtype) on the scene's device holding n * 32 bytes
    def trace(self, rays, hits=None, stream=None):
        n = _ray_count(rays, "rays", self.ctx.device)
        hits = _hits_for(hits, n, rays, self.ctx.device)
        _check(LIB.vx_trace(self._live(), rays.data_ptr(), n, hits.data_ptr(), _stream_ptr(stream)), "vx_trace")
        return hits

    def trace_lod(self, rays, lod: VxLod, hits=None, stream=None):
        """castRayWithLOD over a batch; LOD hits carry VX_HIT_FLAG_LOD (bit 6)."""
        n = _ray_count(rays, "rays", self.ctx.device)
        hits = _hits_for(hits, n, rays, self.ctx.device)
        _check(LIB.vx_trace_lod(self._live(), rays.data_ptr(), n, C.byref(lod), hits.data_ptr(), _stream_ptr(stream)),
               "vx_trace_lod")
        return hits

    def trace_counters(self, rays, hits=None, stream=None):
        torch = _torch()
        n = _ray_count(rays, "rays", self.ctx.device)
        hits = _hits_for(hits, n, rays, self.ctx.device)
        counters = torch.zeros(8, dtype=torch.int64, device=rays.device)
        _check(LIB.vx_trace_counters(self._live(), rays.data_ptr(), n, hits.data_ptr(), counters.data_ptr(),
                                     _stream_ptr(stream)), "vx_trace_counters")
        return hits, counters

    def render(self, cam: VxCamera, w: int, h: int, tmax: float = 1e30, hits=True, rgba=False, stream=None):
        torch = _torch()
        dev = torch.device("cuda", self.ctx.device)
        hb = torch.empty(w * h * 32, dtype=torch.uint8, device=dev) if hits else None
        ib = torch.empty(w * h, dtype=torch.int32, device=dev) if rgba else None
        _check(LIB.vx_render(self._live(), C.byref(cam), w, h, float(tmax), hb.data_ptr() if hb is not None else None,
                             ib.data_ptr() if ib is not None else None, _stream_ptr(stream)), "vx_render")
        return hb, ib


    def bounce_rays(self, primary, hits, seed: int, stream=None):
        """Seeded cosine-weighted diffuse bounce rays for the primary hits (device buffers of n x 32 bytes)."""
        torch = _torch()
        n = _ray_count(primary, "primary", self.ctx.device)
        _dev_buf(hits, "hits", self.ctx.device, 32 * n)
        out = torch.empty(n * 32, dtype=torch.uint8, device=primary.device)
        _check(LIB.vx_bounce_rays(self._live(), primary.data_ptr(), hits.data_ptr(), n, seed & 0xFFFFFFFF, out.data_ptr(),
                                  _stream_ptr(stream)), "vx_bounce_rays")
        return out

    def set_lds_depth(self, depth: int) -> None:
        _check(LIB.vx_scene_set_lds_depth(self._h, depth), "vx_scene_set_lds_depth")

    # ---- differentiable albedo rendering (Lambertian, one RGB albedo per occupied voxel) ----
    def voxel_count(self) -> int:
        n = C.c_uint64()
        _check(LIB.vx_scene_voxel_count(self._live(), C.byref(n)), "vx_scene_voxel_count")
        return int(n.value)

    def albedo_forward(self, cam: VxCamera, w: int, h: int, light, albedo, store_hits: bool = True, stream=None):
        """image [h,w,3] (+ per-pixel albedo row and shade when store_hits)."""
        torch = _torch()
        _dev_buf(albedo, "albedo", self.ctx.device)
        if albedo.dtype != torch.float32 or tuple(albedo.shape) != (self.voxel_count(), 3):
            raise ValueError(f"albedo must be float32 [{self.voxel_count()}, 3], got {albedo.dtype} "
                             f"{list(albedo.shape)}")
        lt = np.ascontiguousarray(light, dtype=np.float32)
        image = torch.empty((h, w, 3), dtype=torch.float32, device=albedo.device)
        rows = torch.empty(h * w, dtype=torch.int32, device=albedo.device) if store_hits else None
        shade = torch.empty(h * w, dtype=torch.float32, device=albedo.device) if store_hits else None
        _check(LIB.vx_albedo_forward(self._live(), C.byref(cam), w, h, lt.ctypes.data, albedo.data_ptr(), image.data_ptr(),
                                     rows.data_ptr() if store_hits else None, shade.data_ptr() if store_hits else None,
                                     _stream_ptr(stream)), "vx_albedo_forward")
        return image, rows, shade

    def albedo_backward(self, cam: VxCamera, w: int, h: int, light, grad_image, rows=None, shade=None, grad=None,
                        deterministic: bool = True, stream=None):
        """grad [V,3] += adjoint of albedo_forward for grad_image [h,w,3] (hits re-traced when rows is None)."""
        torch = _torch()
        dev = self.ctx.device
        _dev_buf(grad_image, "grad_image", dev)
        if grad_image.dtype != torch.float32 or grad_image.numel() != h * w * 3:
            raise ValueError(f"grad_image must be float32 with {h * w * 3} elements")
        if (rows is None) != (shade is None):
            raise ValueError("rows and shade go together (both stored hits, or neither to re-trace)")
        if rows is not None:
            _dev_buf(rows, "rows", dev, 4 * w * h)
            _dev_buf(shade, "shade", dev, 4 * w * h)
        V = self.voxel_count()
        if grad is None:
            grad = torch.zeros((V, 3), dtype=torch.float32, device=grad_image.device)
        else:
            _dev_buf(grad, "grad", dev, 12 * V)
        lt = np.ascontiguousarray(light, dtype=np.float32)
        _check(LIB.vx_albedo_backward(self._live(), C.byref(cam), w, h, lt.ctypes.data, grad_image.data_ptr(),
                                      rows.data_ptr() if rows is not None else None,
                                      shade.data_ptr() if shade is not None else None, grad.data_ptr(),
                                      1 if deterministic else 0, _stream_ptr(stream)), "vx_albedo_backward")
        return grad


def _tuning(t) -> VxTuning | None:
    if t is None or isinstance(t, VxTuning):
        return t
    return VxTuning(*[int(t.get(k, 0)) for k in ("waves_per_group", "chunk", "lds_depth", "refill_min")])


class Batch:
    """Many scenes in one persistent launch (concatenated pools, per-scene offsets)."""

    def __init__(self, ctx: Context, scenes: list[Scene]):
        self.ctx = ctx
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(scenes))(*[sc._live().value for sc in scenes])
        h = C.c_void_p()
        _check(LIB.vx_batch_create(ctx._h, arr, len(scenes), C.byref(h)), "vx_batch_create")
        self._h = h

    def close(self) -> None:
        if self._h:
            LIB.vx_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _args(self, rays, offsets, hits, need_hits: bool):
        """Checks a batched call's buffers before anything is launched: ray_offsets has one entry per scene
        plus one, never decreases, and its last entry fits both the rays and the hits."""
        dev = self.ctx.device
        offs = _ray_offsets(offsets, len(self.scenes), _ray_count(rays, "rays", dev))
        if hits is None and need_hits:
            raise ValueError("hits is required")
        if hits is not None:
            _dev_buf(hits, "hits", dev, 32 * int(offs[len(self.scenes)]))
        return offs

    def trace(self, rays, offsets, hits=None, tuning=None, stream=None):
        torch = _torch()
        if hits is None:
            hits = torch.empty(rays.numel() * rays.element_size(), dtype=torch.uint8, device=rays.device)
        offs = self._args(rays, offsets, hits, True)
        t = _tuning(tuning)
        _check(LIB.vx_trace_batched(self._h, rays.data_ptr(), offs, hits.data_ptr(),
                                    C.byref(t) if t is not None else None, _stream_ptr(stream)), "vx_trace_batched")
        return hits

    def autotune(self, rays, offsets, hits, stream=None):
        offs = self._args(rays, offsets, hits, True)
        best = VxTuning()
        ms = C.c_float()
        _check(LIB.vx_batch_autotune(self._h, rays.data_ptr(), offs, hits.data_ptr(),
                                     C.byref(best), C.byref(ms), _stream_ptr(stream)), "vx_batch_autotune")
        return best, float(ms.value)

    def graph(self, rays, offsets, hits, tuning=None) -> "Graph":
        """The captured launch reads rays and writes hits at replay: the Graph keeps both tensors (and the
        batch) referenced until it is closed."""
        offs = self._args(rays, offsets, hits, True)
        t = _tuning(tuning)
        h = C.c_void_p()
        _check(LIB.vx_batch_graph_create(self._h, rays.data_ptr(), offs, hits.data_ptr(),
                                         C.byref(t) if t is not None else None, C.byref(h)), "vx_batch_graph_create")
        return Graph(h, keep=(self, rays, hits))


class Graph:
    """A captured batched launch (hipGraph) replayed per frame."""

    def __init__(self, h: C.c_void_p, keep=()):
        self._h = h
        self._keep = keep  # buffers the captured kernels use at replay

    def launch(self, stream=None) -> None:
        _check(LIB.vx_graph_launch(self._h, _stream_ptr(stream)), "vx_graph_launch")

    def close(self) -> None:
        if self._h:
            LIB.vx_graph_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CapturedFrame:
    """Launches on scenes (vx_trace, vx_trace_lod, vx_render, vx_camera_rays) recorded once into a
    torch.cuda.CUDAGraph and replayed per frame. The graph holds the scenes' raw device pointers, so this
    object takes its own reference to every scene (vx_scene_retain) and keeps the borrowed device buffers of
    device-buffer scenes and the tensors in `keep` alive until close(): a scene may be closed by its creator
    while the frame is still replayed. record() issues the launches and returns the outputs (kept as
    .outputs); warm the launches up once before capturing, as torch.cuda.graph expects."""

    def __init__(self, scenes, record, keep=()):
        torch = _torch()
        self._scenes = []
        for sc in scenes:
            _check(LIB.vx_scene_retain(sc._h), "vx_scene_retain")
            self._scenes.append(C.c_void_p(sc._h.value))
        self._keep = (tuple(getattr(sc, "_keep", ()) for sc in scenes), keep)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.outputs = record()

    def replay(self) -> None:
        self.graph.replay()

    def close(self) -> None:
        if self._scenes:
            self.graph = None
            self.outputs = None
            for h in self._scenes:
                LIB.vx_scene_destroy(h)
            self._scenes = []
            self._keep = ()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_TORCH_EXT = None


def torch_ext():
    """The libtorch autograd binding (vx_torch.so): albedo_render with a hand-written backward."""
    global _TORCH_EXT
    if _TORCH_EXT is None:
        import importlib.machinery
        import importlib.util
        path = PKG_DIR / "vx_torch.so"
        if not path.exists():
            raise ImportError(f"{path} is not built: run __graft_entry__.build()")
        loader = importlib.machinery.ExtensionFileLoader("vx_torch", str(path))
        spec = importlib.util.spec_from_file_location("vx_torch", str(path), loader=loader)
        mod = importlib.util.module_from_spec(spec)
        loader.exec_module(mod)
        _TORCH_EXT = mod
    return _TORCH_EXT


# Bytes per pixel the forward keeps for the backward when it stores its hits:
# the hit voxel's albedo row (u32) and its shade (f32).
HIT_STASH_BYTES_PER_PIXEL = 8


def stores_hits(w: int, h: int, free_bytes: int, memory_fraction: float = 0.25) -> bool:
    """The automatic choice between storing the forward's hits and re-tracing
    them in the backward: store while the w x h stash takes at most
    memory_fraction of the device memory still free, re-trace when memory is tight."""
    if not 0.0 <= memory_fraction <= 1.0:
        raise ValueError(f"memory_fraction must lie in [0, 1], got {memory_fraction}")
    return HIT_STASH_BYTES_PER_PIXEL * int(w) * int(h) <= memory_fraction * int(free_bytes)


def free_device_bytes(device) -> int:
    """Device memory a new allocation can take: what the driver reports free plus
    what torch's caching allocator holds reserved but unused."""
    torch = _torch()
    free, _ = torch.cuda.mem_get_info(device)
    return int(free) + int(torch.cuda.memory_reserved(device)) - int(torch.cuda.memory_allocated(device))


def albedo_render(scene: "Scene", albedo, cam: VxCamera, w: int, h: int, light, store_hits="auto",
                  deterministic: bool = True, stream=None, memory_fraction: float = 0.25):
    """Differentiable render (torch::autograd::Function in vx_torch.so): image [h,w,3]; albedo gets gradients.

    store_hits: True keeps each pixel's albedo row and shade (8 bytes) for the
    backward, False makes the backward re-trace the hits (same gradient, bit for
    bit), "auto" (the default) stores unless that stash would exceed
    memory_fraction of the free device memory (stores_hits)."""
    torch = _torch()
    if isinstance(store_hits, str):
        if store_hits != "auto":
            raise ValueError(f"store_hits must be True, False or 'auto', got {store_hits!r}")
        store_hits = stores_hits(w, h, free_device_bytes(albedo.device), memory_fraction)
    camt = torch.from_numpy(cam.as_floats())
    lt = torch.tensor(np.asarray(light, dtype=np.float32))
    # the binding runs on the current stream (its outputs come from torch's allocator there)
    with torch.cuda.stream(stream if stream is not None else torch.cuda.current_stream()):
        s = torch.cuda.current_stream()
        return torch_ext().albedo_render(albedo, int(scene._live().value), camt, int(w), int(h), lt, bool(store_hits),
                                         bool(deterministic), int(s.cuda_stream))


def camera_rays(cam: VxCamera, w: int, h: int, tmax: float, device: int = 0, stream=None, out=None):
    """The w x h pinhole rays of a camera (vx_ray, 32 bytes each), into `out` when given."""
    torch = _torch()
    if out is not None:
        _dev_buf(out, "out", device, w * h * 32)
    rays = out if out is not None else torch.empty(w * h * 32, dtype=torch.uint8, device=torch.device("cuda", device))
    _check(LIB.vx_camera_rays(C.byref(cam), w, h, float(tmax), rays.data_ptr(), _stream_ptr(stream)),
           "vx_camera_rays")
    return rays


def hits_to_numpy(hits) -> np.ndarray:
    """Device hit buffer -> structured numpy array (HIT_DTYPE)."""
    return hits.cpu().numpy().view(HIT_DTYPE)


def rays_to_device(rays: np.ndarray, device: int = 0):
    torch = _torch()
    raw = np.ascontiguousarray(rays).view(np.uint8).reshape(-1)
    return torch.from_numpy(raw.copy()).to(torch.device("cuda", device))


@dataclass
class RayHit:
    """ISVOStructure::RayHit (libraries/SVO/include/ISVOStructure.h:101-112), voxel-path fields."""
    hit: bool
    tMin: float
    hitPoint: tuple[float, float, float]
    normal: tuple[float, float, float]
    brick: int
    voxel: int
    leaf_node: int
    color: tuple[float, float, float]
    scale: int | None = None  # user scale of an LOD hit (esvoToUserScale), None otherwise


class LaineKarrasOctreeHip:
    """castRay surface of LaineKarrasOctree (libraries/SVO/src/SVOTraversal.cpp:189-241) on the GPU."""

    def __init__(self, scene: Scene):
        self.scene = scene
        self._pinned_in = self._pinned_out = None  # page-locked staging of castRays, grown on demand
        self._staging_lock = threading.Lock()  # one castRays at a time uses the staging buffers

    def _staging(self, nbytes: int):
        torch = _torch()
        if self._pinned_in is None or self._pinned_in.numel() < nbytes:
            self._pinned_in = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
            self._pinned_out = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        return self._pinned_in[:nbytes], self._pinned_out[:nbytes]

    # ---- the structure queries a castRay caller uses (libraries/SVO/include/LaineKarrasOctree.h:97-108) ----
    def getWorldMin(self) -> tuple[float, float, float]:
        return tuple(float(x) for x in self.scene._cfg.grid_min)

    def getWorldMax(self) -> tuple[float, float, float]:
        return tuple(float(x) for x in self.scene._cfg.grid_max)

    def getMaxLevels(self) -> int:
        return int(self.scene._cfg.user_max_levels)

    def getVoxelSize(self, scale: int) -> float:
        """LaineKarrasOctree::getVoxelSize (libraries/SVO/src/LaineKarrasOctree.cpp:327-334): the world's x
        extent over 2^scale in float32, 0 from scale = getMaxLevels() on."""
        if scale >= self.getMaxLevels():
            return 0.0
        ext = np.float32(self.scene._cfg.grid_max[0]) - np.float32(self.scene._cfg.grid_min[0])
        return float(ext / np.float32(2.0 ** scale))

    def getGPUBuffers(self) -> dict:
        """ISVOStructure::GPUBuffers (libraries/SVO/include/ISVOStructure.h:252-269) of a host-built scene:
        hierarchy = the child descriptors, attribute = brickData, and the DXT colour and normal pools (empty
        when absent)."""
        b = self.scene.built
        if b is None:
            raise ValueError("getGPUBuffers: the scene was created over device buffers")
        none = b""
        return {"hierarchyBuffer": b.nodes.tobytes(), "attributeBuffer": b.brick_data.tobytes(), "auxBuffer": none,
                "compressedColorBuffer": b.bc1.tobytes() if b.bc1 is not None else none,
                "compressedNormalBuffer": b.normals.tobytes() if b.normals is not None else none}

    def castRays(self, rays: np.ndarray, lod: VxLod | None = None) -> np.ndarray:
        """Host rays in, host records out (HIT_DTYPE). Both cross PCIe through page-locked staging buffers, filled
        and drained by torch's multi-threaded host copies, so the transfers run at DMA speed."""
        torch = _torch()
        raw = np.ascontiguousarray(np.asarray(rays, dtype=RAY_DTYPE)).view(np.uint8).reshape(-1)
        nbytes = raw.size
        if nbytes == 0:
            return np.zeros(0, HIT_DTYPE)
        out = torch.empty(nbytes, dtype=torch.uint8)
        dev = torch.device("cuda", self.scene.ctx.device)
        with self._staging_lock, torch.cuda.device(dev):
            pin_in, pin_out = self._staging(nbytes)
            pin_in.copy_(torch.from_numpy(raw if raw.flags.writeable else raw.copy()))
            d = pin_in.to(dev, non_blocking=True)
            hits = self.scene.trace(d) if lod is None else self.scene.trace_lod(d, lod)
            pin_out.copy_(hits, non_blocking=True)
            torch.cuda.current_stream().synchronize()
            out.copy_(pin_out)
        return out.numpy().view(HIT_DTYPE)

    def castRay(self, origin, direction, tMin: float = 0.0, tMax: float = 1e30) -> RayHit:
        return self._cast(origin, direction, tMin, tMax, None)

    def castRayLOD(self, origin, direction, lodBias: float, tMin: float = 0.0, tMax: float = 1e30) -> RayHit:
        """The legacy interface: lodBias is the cone growth per unit distance (SVOTraversal.cpp:198-210)."""
        return self._cast(origin, direction, tMin, tMax, VxLod(0.0, lodBias))

    def castRayScreenSpaceLOD(self, origin, direction, fovY: float, screenHeight: int, tMin: float = 0.0,
                              tMax: float = 1e30) -> RayHit:
        return self._cast(origin, direction, tMin, tMax, lod_from_camera(fovY, screenHeight))

    def castRayWithLOD(self, origin, direction, lodParams: VxLod, tMin: float = 0.0, tMax: float = 1e30) -> RayHit:
        return self._cast(origin, direction, tMin, tMax, lodParams)

    def _cast(self, origin, direction, tMin, tMax, lod) -> RayHit:
        r = np.zeros(1, RAY_DTYPE)
        r["o"][0], r["d"][0], r["tmin"][0], r["tmax"][0] = origin, direction, tMin, tMax
        h = self.castRays(r, lod)[0]
        hit = bool(h["flags"] & HIT_FLAG)
        o = np.asarray(origin, np.float64)
        dd = np.asarray(direction, np.float64)
        dd = dd / (np.linalg.norm(dd) or 1.0)
        normal = [0.0, 0.0, 0.0]
        if hit:
            ax = (int(h["flags"]) >> 1) & 3
            normal[ax] = 1.0 if (int(h["flags"]) >> 3) & 1 else -1.0
        t = float(h["t"])
        hp = tuple(o + dd * t) if hit else (math.nan, math.nan, math.nan)
        scale = int(h["flags"] >> 8) & 31 if int(h["flags"]) & LOD_FLAG else None
        return RayHit(hit, t, hp, tuple(normal), int(h["brick"]), int(h["voxel"]), int(h["leaf_node"]),
                      tuple(float(x) for x in h["rgb"]), scale)


def version() -> str:
    return LIB.vx_version().decode()

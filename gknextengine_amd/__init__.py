"""gknextengine_amd - MI355X (gfx950) path-tracing core with a C ABI.

This package is a thin ctypes binding of ``include/gknx_rt.h`` (``libgknx_rt.so``,
built in-tree by ``gknextengine_amd.build``). It exists for tests and benchmarks;
engines bind the C ABI directly. Loading fails loudly when the
library has not been built: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "_lib" / "libgknx_rt.so"

# ---------------------------------------------------------------- constants (gknx_rt.h)
GK_OK = 0
GK_ERR_ARG, GK_ERR_HIP, GK_ERR_OOM, GK_ERR_STATE, GK_ERR_LIMIT, GK_ERR_NO_DEVICE, GK_ERR_IO = -1, -2, -3, -4, -5, -6, -7

UPSCALE_LINEAR, UPSCALE_FSR_EASU = 0, 1
MAT_LAMBERTIAN, MAT_METALLIC, MAT_DIELECTRIC, MAT_ISOTROPIC, MAT_DIFFUSE_LIGHT, MAT_MIXTURE = range(6)
ESTIMATOR_REF_V2, ESTIMATOR_PT_MIS = 0, 1
DENOISER_NONE, DENOISER_BILATERAL, DENOISER_UNET = 0, 1, 2
FRAME_NO_MATERIAL_SORT, FRAME_SKIP_POST, FRAME_ACCUM_FP64, FRAME_LINEAR_OUTPUT, FRAME_MATERIAL_SORT = 1, 2, 4, 8, 16
FRAME_PRIMARY_FROM_VISBUFFER = 32
FRAME_POST_ONLY = 64
FRAME_TLAS_REBUILD = 128
FRAME_FAST_GATHER = 256
FRAME_SINGLE_PIPELINE = 512
TRACE_CLOSEST, TRACE_ANY = 0, 1
TRACE_STATIC_MAX_RAYS = 262144

TARGETS = [
    "SINGLE_DIFFUSE", "SINGLE_SPECULAR", "ALBEDO", "NORMAL", "OBJECTID_0", "OBJECTID_1",
    "MOTIONVECTOR", "MOTIONMOMENT", "PREV_DEPTH", "ACCUM_DIFFUSE", "ACCUM_SPECULAR",
    "ACCUM_ALBEDO", "DENOISED", "MINIGBUFFER", "PRIMARY_HIT", "PRIMARY_INST", "RADIANCE", "RADIANCE64",
]
RT = {name: i for i, name in enumerate(TARGETS)}
# numpy dtype and channel count of each target
TARGET_FORMAT = {
    "SINGLE_DIFFUSE": (np.float16, 4), "SINGLE_SPECULAR": (np.float16, 4), "ALBEDO": (np.float16, 4),
    "NORMAL": (np.float16, 4), "OBJECTID_0": (np.uint32, 1), "OBJECTID_1": (np.uint32, 1),
    "MOTIONVECTOR": (np.float32, 2), "MOTIONMOMENT": (np.uint16, 1), "PREV_DEPTH": (np.float32, 1),
    "ACCUM_DIFFUSE": (np.float16, 4), "ACCUM_SPECULAR": (np.float16, 4), "ACCUM_ALBEDO": (np.float16, 4),
    "DENOISED": (np.float16, 4), "MINIGBUFFER": (np.uint32, 1), "PRIMARY_HIT": (np.float32, 4),
    "PRIMARY_INST": (np.uint32, 1), "RADIANCE": (np.float32, 4), "RADIANCE64": (np.float64, 4),
}


# ---------------------------------------------------------------- structures
class Mesh(C.Structure):
    _fields_ = [("positions", C.c_void_p), ("normals", C.c_void_p), ("texcoords", C.c_void_p),
                ("indices", C.c_void_p), ("tri_material", C.c_void_p),
                ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("tangents", C.c_void_p)]


class Texture(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("srgb", C.c_uint32), ("reserved", C.c_uint32),
                ("rgba8", C.c_void_p)]


class Material(C.Structure):
    _fields_ = [("diffuse", C.c_float * 4), ("diffuse_texture_id", C.c_int32), ("mra_texture_id", C.c_int32),
                ("normal_texture_id", C.c_int32), ("fuzziness", C.c_float), ("refraction_index", C.c_float),
                ("material_model", C.c_uint32), ("metalness", C.c_float), ("refraction_index2", C.c_float),
                ("normal_texture_scale", C.c_float), ("reserved", C.c_float * 3)]


class Light(C.Structure):
    _fields_ = [("p0", C.c_float * 4), ("p1", C.c_float * 4), ("p3", C.c_float * 4),
                ("normal_area", C.c_float * 4), ("light_mat_idx", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class Instance(C.Structure):
    _fields_ = [("world", C.c_float * 16), ("prev_world", C.c_float * 16), ("instance_id", C.c_uint32),
                ("mesh_id", C.c_uint32), ("visible", C.c_uint32), ("nort", C.c_uint32),
                ("mat_id", C.c_uint32 * 16), ("mat_offset", C.c_uint32), ("use_remap", C.c_uint32),
                ("reserved", C.c_uint32 * 2)]


class FrameParams(C.Structure):
    _fields_ = [(n, C.c_float * 16) for n in (
        "model_view", "projection", "model_view_inverse", "projection_inverse", "view_projection",
        "prev_view_projection", "view_projection_unjit", "prev_view_projection_unjit")] + [
        ("sun_direction", C.c_float * 4), ("sun_color", C.c_float * 4), ("sky_color", C.c_float * 4)] + [
        (n, C.c_float) for n in ("aperture", "focus_distance", "paper_white_nit", "bf_sigma", "bf_sigma_lum",
                                 "bf_sigma_normal", "exit_probability", "hit_normal_offset")] + [
        (n, C.c_uint32) for n in ("total_frames", "global_seed", "samples", "bounces", "max_bounces",
                                  "temporal_frames", "bf_size", "estimator", "denoiser", "has_sky", "has_sun",
                                  "taa", "progressive", "hdr", "flags")] + [("sky_rotation", C.c_float)]


class CameraDesc(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_y_degrees", C.c_float), ("aperture", C.c_float), ("focus_distance", C.c_float)]


class DeviceDesc(C.Structure):
    _fields_ = [("device", C.c_int32), ("width", C.c_uint32), ("height", C.c_uint32), ("flags", C.c_uint32)]


class BvhStats(C.Structure):
    _fields_ = [("blas_count", C.c_uint32), ("blas_nodes", C.c_uint32), ("blas_leaves", C.c_uint32),
                ("tlas_nodes", C.c_uint32), ("blas_bytes", C.c_uint64), ("tlas_bytes", C.c_uint64),
                ("sah_cost", C.c_double), ("build_ms", C.c_double), ("tlas_ms", C.c_double),
                ("tlas_sah_cost", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("rays_traced", C.c_uint64), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("path_vertices", C.c_uint64)] + [
        (n, C.c_double) for n in ("ms_tlas", "ms_trace", "ms_shade", "ms_reproject", "ms_denoise", "ms_frame")]


class ProbeGrid(C.Structure):
    _fields_ = [("size_xy", C.c_uint32), ("size_z", C.c_uint32), ("unit", C.c_float), ("offset", C.c_float * 3)]


# gk_ambient_cube (56 bytes) and gk_voxel (16 bytes) as numpy records
CUBE_FIELDS = ("pos_z", "neg_z", "pos_y", "neg_y", "pos_x", "neg_x", "pos_z_d", "neg_z_d", "pos_y_d", "neg_y_d",
               "pos_x_d", "neg_x_d", "sky_vis_pznzpyny", "sky_vis_pxnxs0s1")
CUBE_DTYPE = np.dtype([(n, np.uint32) for n in CUBE_FIELDS])
VOXEL_DTYPE = np.dtype([(n, np.uint32) for n in ("mat_id", "age", "dist_gg_z01", "dist_x01_y01")])
# the reference's grid: 192 x 48 x 192 probes, 0.25 m apart (AmbientCube.slang:10-14)
REFERENCE_PROBE_GRID = (192, 48, 0.25, (-24.0, -0.34375, -24.0))


class GltfView(C.Structure):
    _fields_ = [("meshes", C.c_void_p), ("num_meshes", C.c_uint32), ("materials", C.c_void_p),
                ("num_materials", C.c_uint32), ("instances", C.c_void_p), ("num_instances", C.c_uint32),
                ("lights", C.c_void_p), ("num_lights", C.c_uint32), ("textures", C.c_void_p),
                ("num_textures", C.c_uint32), ("num_cameras", C.c_uint32), ("camera", CameraDesc),
                ("has_sky", C.c_uint32), ("has_sun", C.c_uint32), ("sky_idx", C.c_int32),
                ("sky_intensity", C.c_float), ("sky_rotation", C.c_float), ("sun_intensity", C.c_float),
                ("sun_rotation", C.c_float)]


class UNetLayer(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("cin_a", "cin_b", "cout", "cout_real", "kpad", "up", "relu", "level")] + [
        ("offset", C.c_uint64)]


assert C.sizeof(Material) == 64 and C.sizeof(Instance) == 224 and C.sizeof(FrameParams) == 656
assert C.sizeof(Light) == 80 and C.sizeof(Mesh) == 56 and C.sizeof(Texture) == 24

# every entry point declared in include/gknx_rt.h: (name, restype, argtypes)
_P = C.c_void_p
_U = C.c_uint32
SIGNATURES = [
    ("gk_get_abi_version", C.c_int, []),
    ("gk_status_string", C.c_char_p, [C.c_int]),
    ("gk_create", C.c_int, [C.POINTER(DeviceDesc), C.POINTER(_P)]),
    ("gk_destroy", C.c_int, [_P]),
    ("gk_resize", C.c_int, [_P, _U, _U]),
    ("gk_scene_upload", C.c_int, [_P, C.POINTER(Mesh), _U, C.POINTER(Material), _U, C.POINTER(Light), _U,
                                  C.POINTER(Instance), _U]),
    ("gk_materials_update", C.c_int, [_P, C.POINTER(Material), _U, _U]),
    ("gk_textures_upload", C.c_int, [_P, C.POINTER(Texture), _U]),
    ("gk_sky_upload", C.c_int, [_P, _P, _U, _U]),
    ("gk_instances_update", C.c_int, [_P, _P, _U, _U]),
    ("gk_instances_update_device", C.c_int, [_P, _P, _U, _U]),
    ("gk_get_bvh_stats", C.c_int, [_P, C.POINTER(BvhStats)]),
    ("gk_frame", C.c_int, [_P, C.POINTER(FrameParams)]),
    ("gk_render", C.c_int, [_P]),
    ("gk_sync", C.c_int, [_P]),
    ("gk_reset_accumulation", C.c_int, [_P]),
    ("gk_read_target", C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    ("gk_write_target", C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    ("gk_target_info", C.c_int, [_P, C.c_int, C.POINTER(C.c_size_t), C.POINTER(_P)]),
    ("gk_get_stats", C.c_int, [_P, C.POINTER(Stats)]),
    ("gk_trace", C.c_int, [_P, _P, _U, _P, _P, _U]),
    ("gk_trace_static", C.c_int, [_P, _P, _U, _P, _P, _U]),
    ("gk_make_frame_params", C.c_int, [C.POINTER(CameraDesc), _U, _U, _U, _U, C.POINTER(FrameParams),
                                       C.POINTER(FrameParams)]),
    ("gk_procgen_cornell", C.c_int, [_P, _P, _P, _P, C.POINTER(_U), C.POINTER(_U)]),
    ("gk_procgen_cornell_materials", C.c_int, [C.POINTER(Material), C.POINTER(_U), C.POINTER(Light),
                                               C.POINTER(_U)]),
    ("gk_procgen_city", C.c_int, [_U, _U, _U, _U, C.c_uint64, _P, _P, _P, _P, C.POINTER(_U), C.POINTER(_U),
                                  C.POINTER(Material), C.POINTER(_U)]),
    ("gk_procgen_brick", C.c_int, [_P, _P, _P, _P, C.POINTER(_U), C.POINTER(_U)]),
    ("gk_adjoint", C.c_int, [_P, _P, _P]),
    ("gk_render_radiance", C.c_int, [_P, _P]),
    ("gk_unet_create", C.c_int, [_P, _U, C.c_uint64, C.POINTER(_P)]),
    ("gk_unet_destroy", C.c_int, [_P]),
    ("gk_unet_num_params", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("gk_unet_layer_info", C.c_int, [_P, _U, C.POINTER(UNetLayer)]),
    ("gk_unet_params", C.c_int, [_P, C.POINTER(_P)]),
    ("gk_unet_params_updated", C.c_int, [_P]),
    ("gk_unet_save", C.c_int, [_P, C.c_char_p]),
    ("gk_unet_load", C.c_int, [_P, C.c_char_p]),
    ("gk_unet_grad", C.c_int, [_P, _P, _P, _U, _U, _U, _P, C.POINTER(C.c_float)]),
    ("gk_unet_train_step", C.c_int, [_P, _P, _P, _U, _U, _U, C.c_float, C.POINTER(C.c_float)]),
    ("gk_unet_infer", C.c_int, [_P, _P, _P, _U, _U]),
    ("gk_denoise_input", C.c_int, [_P, _P]),
    ("gk_render_extent", C.c_int, [_U, _U, _U, C.POINTER(_U), C.POINTER(_U)]),
    ("gk_upscale", C.c_int, [_P, C.c_int, _U, _U, _P]),
    ("gk_screenshot", C.c_int, [_P, _P, _U, _U, C.c_char_p]),
    ("gk_gltf_load", C.c_int, [C.c_char_p, C.POINTER(_P)]),
    ("gk_gltf_load_memory", C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    ("gk_gltf_get_view", C.c_int, [_P, C.POINTER(GltfView)]),
    ("gk_gltf_upload", C.c_int, [_P, _P]),
    ("gk_gltf_free", C.c_int, [_P]),
    ("gk_gltf_error", C.c_char_p, []),
    ("gk_probes_init", C.c_int, [_P, C.POINTER(ProbeGrid)]),
    ("gk_probes_voxelize", C.c_int, [_P]),
    ("gk_probes_bake", C.c_int, [_P, _U, _U]),
    ("gk_probes_bake_frame", C.c_int, [_P, _U, _U]),
    ("gk_probes_read", C.c_int, [_P, _P, _P, _U, _U]),
    ("gk_probes_write", C.c_int, [_P, _P, _P, _U, _U]),
]

_lib = None


class GkError(RuntimeError):
    def __init__(self, status: int, what: str):
        super().__init__(f"{what} failed: {status} ({lib().gk_status_string(status).decode()})")
        self.status = status


def lib() -> C.CDLL:
    """Load libgknx_rt.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m gknextengine_amd.build`")
        # PyTorch-ROCm ships its own libamdhip64 (SONAME libamdhip64.so.7, loaded under the
        # unversioned name). Loading it first makes libgknx_rt bind to that same runtime, so
        # the process has one HIP runtime and torch tensors and gk_* calls share the device.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(str(LIB_PATH))
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status: int, what: str) -> None:
    if status != GK_OK:
        raise GkError(status, what)


def ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


# ---------------------------------------------------------------- host helpers
def make_frame_params(eye, target, width, height, fov=40.0, up=(0, 1, 0), frame=0, taa=False,
                      prev: FrameParams | None = None) -> FrameParams:
    cam = CameraDesc()
    cam.eye[:] = eye
    cam.target[:] = target
    cam.up[:] = up
    cam.fov_y_degrees = fov
    cam.aperture = 0.0
    cam.focus_distance = 10.0
    out = FrameParams()
    check(lib().gk_make_frame_params(C.byref(cam), width, height, frame, int(taa),
                                     C.byref(prev) if prev is not None else None, C.byref(out)),
          "gk_make_frame_params")
    return out


class SceneData:
    """Host-side arrays of one scene (meshes, materials, instances)."""

    def __init__(self):
        self.meshes: list[dict] = []
        self.materials: list[Material] = []
        self.instances: list[Instance] = []
        self.lights: list[Light] = []    # LightObject quads (the probe bake uses the first)
        self.textures: list[dict] = []   # {"rgba": (h, w, 4) uint8, "srgb": bool}

    def add_mesh(self, positions, indices, normals=None, tri_material=None, texcoords=None, tangents=None) -> int:
        def arr(a, k):
            return None if a is None else np.ascontiguousarray(a, np.float32).reshape(-1, k)
        m = {"positions": arr(positions, 3),
             "indices": np.ascontiguousarray(indices, np.uint32).reshape(-1, 3),
             "normals": arr(normals, 3),
             "tri_material": None if tri_material is None else np.ascontiguousarray(tri_material, np.uint32),
             "texcoords": arr(texcoords, 2), "tangents": arr(tangents, 4)}
        self.meshes.append(m)
        return len(self.meshes) - 1

    def add_instance(self, mesh_id: int, world=None, prev_world=None, instance_id=None, mat_offset=0,
                     visible=True) -> int:
        inst = Instance()
        w = np.eye(4, dtype=np.float32) if world is None else np.asarray(world, np.float32)
        pw = w if prev_world is None else np.asarray(prev_world, np.float32)
        inst.world[:] = w.T.reshape(-1).tolist()       # column-major
        inst.prev_world[:] = pw.T.reshape(-1).tolist()
        inst.instance_id = len(self.instances) if instance_id is None else instance_id
        inst.mesh_id = mesh_id
        inst.visible = 1 if visible else 0
        inst.nort = 0
        inst.mat_offset = mat_offset
        inst.use_remap = 0
        self.instances.append(inst)
        return len(self.instances) - 1

    def c_meshes(self):
        arr = (Mesh * len(self.meshes))()
        for i, m in enumerate(self.meshes):
            arr[i].positions = ptr(m["positions"])
            arr[i].normals = ptr(m["normals"])
            arr[i].texcoords = ptr(m.get("texcoords"))
            arr[i].tangents = ptr(m.get("tangents"))
            arr[i].indices = ptr(m["indices"])
            arr[i].tri_material = ptr(m["tri_material"])
            arr[i].num_vertices = m["positions"].shape[0]
            arr[i].num_triangles = m["indices"].shape[0]
        return arr

    def c_materials(self):
        arr = (Material * len(self.materials))()
        for i, m in enumerate(self.materials):
            arr[i] = m
        return arr

    def c_instances(self):
        arr = (Instance * len(self.instances))()
        for i, m in enumerate(self.instances):
            arr[i] = m
        return arr

    def add_texture(self, rgba, srgb=True) -> int:
        """An RGBA8 texture ((h, w, 4) uint8, rows top to bottom); returns its id."""
        a = np.ascontiguousarray(rgba, np.uint8)
        assert a.ndim == 3 and a.shape[2] == 4
        self.textures.append({"rgba": a, "srgb": bool(srgb)})
        return len(self.textures) - 1

    def c_textures(self):
        arr = (Texture * max(1, len(self.textures)))()
        for i, t in enumerate(self.textures):
            arr[i].width, arr[i].height = t["rgba"].shape[1], t["rgba"].shape[0]
            arr[i].srgb = 1 if t["srgb"] else 0
            arr[i].rgba8 = t["rgba"].ctypes.data
        return arr

    def c_lights(self):
        if not self.lights:
            return None
        arr = (Light * len(self.lights))()
        for i, m in enumerate(self.lights):
            arr[i] = m
        return arr


class GltfScene:
    """A parsed glTF file (gk_gltf_load): the C-side arrays, copied into ``scene``
    (a SceneData) plus the first camera and the environment extras in ``info``."""

    def __init__(self, source):
        L = lib()
        h = C.c_void_p()
        if isinstance(source, (bytes, bytearray)):
            buf = C.create_string_buffer(bytes(source), len(source))
            st = L.gk_gltf_load_memory(buf, len(source), C.byref(h))
        else:
            st = L.gk_gltf_load(str(source).encode(), C.byref(h))
        if st != 0:
            raise GkError(st, "gk_gltf_load: " + L.gk_gltf_error().decode())
        self.h = h
        v = GltfView()
        check(L.gk_gltf_get_view(h, C.byref(v)), "gk_gltf_get_view")
        s = SceneData()
        meshes = C.cast(v.meshes, C.POINTER(Mesh))
        for i in range(v.num_meshes):
            m = meshes[i]
            nv, nt = m.num_vertices, m.num_triangles

            def grab(p, n, k, dt=np.float32):
                if not p or n == 0:
                    return np.zeros((0, k), dt)
                return np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float if dt == np.float32 else C.c_uint32)),
                                             (n * k,)).reshape(n, k).copy()
            s.add_mesh(grab(m.positions, nv, 3), grab(m.indices, nt, 3, np.uint32), grab(m.normals, nv, 3),
                       grab(m.tri_material, nt, 1, np.uint32).reshape(-1), texcoords=grab(m.texcoords, nv, 2),
                       tangents=grab(m.tangents, nv, 4))
        mats = C.cast(v.materials, C.POINTER(Material))
        s.materials = [Material.from_buffer_copy(mats[i]) for i in range(v.num_materials)]
        insts = C.cast(v.instances, C.POINTER(Instance))
        s.instances = [Instance.from_buffer_copy(insts[i]) for i in range(v.num_instances)]
        lights = C.cast(v.lights, C.POINTER(Light))
        s.lights = [Light.from_buffer_copy(lights[i]) for i in range(v.num_lights)]
        texs = C.cast(v.textures, C.POINTER(Texture))
        for i in range(v.num_textures):
            t = texs[i]
            a = np.ctypeslib.as_array(C.cast(t.rgba8, C.POINTER(C.c_uint8)), (t.height * t.width * 4,))
            s.add_texture(a.reshape(t.height, t.width, 4).copy(), bool(t.srgb))
        self.scene = s
        cam = v.camera
        self.info = {"num_cameras": v.num_cameras, "eye": tuple(cam.eye), "target": tuple(cam.target),
                     "fov": cam.fov_y_degrees, "has_sky": v.has_sky, "has_sun": v.has_sun, "sky_idx": v.sky_idx,
                     "sky_intensity": v.sky_intensity, "sky_rotation": v.sky_rotation,
                     "sun_intensity": v.sun_intensity, "sun_rotation": v.sun_rotation}

    def upload(self, ctx: "Context"):
        """gk_gltf_upload: the loader's arrays straight into the context."""
        check(lib().gk_gltf_upload(ctx.h, self.h), "gk_gltf_upload")

    def close(self):
        if self.h:
            lib().gk_gltf_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def render_extent(out_w: int, out_h: int, super_resolution: int) -> tuple[int, int]:
    w, h = C.c_uint32(0), C.c_uint32(0)
    check(lib().gk_render_extent(out_w, out_h, super_resolution, C.byref(w), C.byref(h)), "gk_render_extent")
    return w.value, h.value


def procgen_cornell() -> SceneData:
    L = lib()
    nv, nt = C.c_uint32(0), C.c_uint32(0)
    check(L.gk_procgen_cornell(None, None, None, None, C.byref(nv), C.byref(nt)), "gk_procgen_cornell")
    pos = np.zeros((nv.value, 3), np.float32)
    nrm = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.uint32)
    mat = np.zeros(nt.value, np.uint32)
    check(L.gk_procgen_cornell(ptr(pos), ptr(nrm), ptr(idx), ptr(mat), C.byref(nv), C.byref(nt)),
          "gk_procgen_cornell")
    nm, nl = C.c_uint32(8), C.c_uint32(4)
    mats = (Material * 8)()
    lights = (Light * 4)()
    check(L.gk_procgen_cornell_materials(mats, C.byref(nm), lights, C.byref(nl)), "gk_procgen_cornell_materials")
    s = SceneData()
    s.add_mesh(pos, idx, nrm, mat)
    s.materials = [mats[i] for i in range(nm.value)]
    s.lights = [lights[i] for i in range(nl.value)]
    s.add_instance(0)
    return s


def procgen_city(grid=64, subdiv=4, ground_res=256, num_lights=64, seed=2024) -> SceneData:
    L = lib()
    nv, nt, nm = C.c_uint32(0), C.c_uint32(0), C.c_uint32(0)
    check(L.gk_procgen_city(grid, subdiv, ground_res, num_lights, seed, None, None, None, None,
                            C.byref(nv), C.byref(nt), None, C.byref(nm)), "gk_procgen_city")
    pos = np.zeros((nv.value, 3), np.float32)
    nrm = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.uint32)
    mat = np.zeros(nt.value, np.uint32)
    mats = (Material * nm.value)()
    check(L.gk_procgen_city(grid, subdiv, ground_res, num_lights, seed, ptr(pos), ptr(nrm), ptr(idx), ptr(mat),
                            C.byref(nv), C.byref(nt), mats, C.byref(nm)), "gk_procgen_city")
    s = SceneData()
    s.add_mesh(pos, idx, nrm, mat)
    s.materials = [mats[i] for i in range(nm.value)]
    s.add_instance(0)
    return s


def procgen_brick():
    L = lib()
    nv, nt = C.c_uint32(0), C.c_uint32(0)
    check(L.gk_procgen_brick(None, None, None, None, C.byref(nv), C.byref(nt)), "gk_procgen_brick")
    pos = np.zeros((nv.value, 3), np.float32)
    nrm = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.uint32)
    mat = np.zeros(nt.value, np.uint32)
    check(L.gk_procgen_brick(ptr(pos), ptr(nrm), ptr(idx), ptr(mat), C.byref(nv), C.byref(nt)), "gk_procgen_brick")
    return pos, nrm, idx, mat


# ---------------------------------------------------------------- device context
class Context:
    """Owns a gk_ctx (one HIP stream, all device state)."""

    def __init__(self, width: int, height: int, device: int = 0):
        d = DeviceDesc(device, width, height, 0)
        h = C.c_void_p()
        check(lib().gk_create(C.byref(d), C.byref(h)), "gk_create")
        self.h = h
        self.width, self.height = width, height
        self._keep = None
        self.unet = None

    def close(self):
        if self.h:
            if getattr(self, "unet", None) is not None:
                self.unet.h = None  # gk_destroy destroys the context's network
                self.unet = None
            lib().gk_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def denoise_input(self, d_out: int):
        """Writes the current frame's 10 denoiser features (width*height*10 FP32) to d_out."""
        check(lib().gk_denoise_input(self.h, d_out), "gk_denoise_input")

    def upload(self, scene: SceneData):
        meshes, mats, insts = scene.c_meshes(), scene.c_materials(), scene.c_instances()
        lights = scene.c_lights()
        textures = scene.c_textures()
        self._keep = (scene, meshes, mats, insts, lights, textures)
        check(lib().gk_scene_upload(self.h, meshes, len(scene.meshes), mats, len(scene.materials), lights,
                                    len(scene.lights), insts, len(scene.instances)), "gk_scene_upload")
        check(lib().gk_textures_upload(self.h, textures, len(scene.textures)), "gk_textures_upload")

    def sky_upload(self, rgba=None):
        """The sky image (equirectangular, H x W x 4 float32, row 0 at +y); None removes it."""
        if rgba is None:
            check(lib().gk_sky_upload(self.h, None, 0, 0), "gk_sky_upload")
            return
        a = np.ascontiguousarray(rgba, np.float32)
        if a.ndim != 3 or a.shape[2] != 4:
            raise ValueError("sky image must be H x W x 4")
        check(lib().gk_sky_upload(self.h, ptr(a), a.shape[1], a.shape[0]), "gk_sky_upload")

    def update_instances(self, worlds: np.ndarray, first: int = 0):
        w = np.ascontiguousarray(np.transpose(np.asarray(worlds, np.float32), (0, 2, 1)))  # column-major
        check(lib().gk_instances_update(self.h, ptr(w), first, w.shape[0]), "gk_instances_update")

    def update_materials(self, materials, first=0):
        arr = (Material * len(materials))(*materials)
        check(lib().gk_materials_update(self.h, arr, first, len(materials)), "gk_materials_update")

    def frame(self, params: FrameParams):
        check(lib().gk_frame(self.h, C.byref(params)), "gk_frame")

    def render(self):
        check(lib().gk_render(self.h), "gk_render")

    def sync(self):
        check(lib().gk_sync(self.h), "gk_sync")

    def reset(self):
        check(lib().gk_reset_accumulation(self.h), "gk_reset_accumulation")

    def read(self, name: str) -> np.ndarray:
        dt, ch = TARGET_FORMAT[name]
        shape = (self.height, self.width, ch) if ch > 1 else (self.height, self.width)
        out = np.zeros(shape, dt)
        check(lib().gk_read_target(self.h, RT[name], ptr(out), out.nbytes), f"gk_read_target({name})")
        return out

    def write(self, name: str, data: np.ndarray):
        dt, ch = TARGET_FORMAT[name]
        a = np.ascontiguousarray(data, dt)
        check(lib().gk_write_target(self.h, RT[name], ptr(a), a.nbytes), f"gk_write_target({name})")

    def target_ptr(self, name: str) -> int:
        n, p = C.c_size_t(0), C.c_void_p()
        check(lib().gk_target_info(self.h, RT[name], C.byref(n), C.byref(p)), "gk_target_info")
        return p.value

    # output stage
    def upscale(self, out_w: int, out_h: int, mode: int = UPSCALE_LINEAR, d_out: int | None = None):
        """GK_RT_DENOISED resolved to out_w x out_h RGBA16F: into the device buffer d_out, or
        read back to the host as a float16 array when d_out is None."""
        import torch
        buf = None
        if d_out is None:
            buf = torch.empty((out_h, out_w, 4), dtype=torch.float16, device="cuda")
            torch.cuda.synchronize()
            d_out = buf.data_ptr()
        check(lib().gk_upscale(self.h, mode, out_w, out_h, d_out), "gk_upscale")
        if buf is not None:
            self.sync()
            return buf.cpu().numpy()
        return None

    def screenshot(self, path, d_img: int | None = None, w: int = 0, h: int = 0):
        check(lib().gk_screenshot(self.h, d_img, w, h, str(path).encode()), "gk_screenshot")

    # ambient-cube probe grid
    def probes_init(self, size_xy=None, size_z=None, unit=None, offset=None):
        """Allocate and clear the probe grid (the reference's grid by default)."""
        d = REFERENCE_PROBE_GRID
        g = ProbeGrid(size_xy or d[0], size_z or d[1], d[2] if unit is None else unit,
                      (C.c_float * 3)(*(d[3] if offset is None else offset)))
        check(lib().gk_probes_init(self.h, C.byref(g)), "gk_probes_init")
        self.probe_count = g.size_xy * g.size_xy * g.size_z

    def probes_voxelize(self):
        check(lib().gk_probes_voxelize(self.h), "gk_probes_voxelize")

    def probes_bake(self, first=0, count=None):
        n = self.probe_count - first if count is None else count
        check(lib().gk_probes_bake(self.h, first, n), "gk_probes_bake")

    def probes_bake_frame(self, frame_count, temporal_frames=120):
        check(lib().gk_probes_bake_frame(self.h, frame_count, temporal_frames), "gk_probes_bake_frame")

    def probes_read(self, first=0, count=None):
        n = self.probe_count - first if count is None else count
        cubes = np.zeros(n, CUBE_DTYPE)
        vox = np.zeros(n, VOXEL_DTYPE)
        check(lib().gk_probes_read(self.h, ptr(cubes), ptr(vox), first, n), "gk_probes_read")
        return cubes, vox

    def probes_write(self, cubes=None, voxels=None, first=0):
        n = len(cubes) if cubes is not None else len(voxels)
        c = None if cubes is None else np.ascontiguousarray(cubes, CUBE_DTYPE)
        v = None if voxels is None else np.ascontiguousarray(voxels, VOXEL_DTYPE)
        check(lib().gk_probes_write(self.h, ptr(c), ptr(v), first, n), "gk_probes_write")

    def bvh_stats(self) -> BvhStats:
        s = BvhStats()
        check(lib().gk_get_bvh_stats(self.h, C.byref(s)), "gk_get_bvh_stats")
        return s

    def stats(self) -> Stats:
        s = Stats()
        check(lib().gk_get_stats(self.h, C.byref(s)), "gk_get_stats")
        return s

    def trace(self, d_rays: int, n: int, d_hits: int, d_inst: int | None = None, any_hit: bool = False):
        """Device-pointer ray cast (gk_ray in, gk_hit out); enqueue only."""
        check(lib().gk_trace(self.h, d_rays, n, d_hits, d_inst, TRACE_ANY if any_hit else TRACE_CLOSEST),
              "gk_trace")

    def trace_static(self, d_rays: int, n: int, d_hits: int, d_inst: int | None = None, any_hit: bool = False):
        """gk_trace without device allocation (the library's static scratch array, at most
        TRACE_STATIC_MAX_RAYS rays); enqueue only."""
        check(lib().gk_trace_static(self.h, d_rays, n, d_hits, d_inst, TRACE_ANY if any_hit else TRACE_CLOSEST),
              "gk_trace_static")


class UNet:
    """The learned denoiser of a Context (gk_unet_*). Arrays are device pointers."""

    IN_CHANNELS = 10
    LAYERS = 12

    def __init__(self, ctx: Context, base_channels: int = 32, seed: int = 1):
        h = C.c_void_p()
        check(lib().gk_unet_create(ctx.h, base_channels, seed, C.byref(h)), "gk_unet_create")
        if getattr(ctx, "unet", None) is not None:
            ctx.unet.h = None  # replaced (and destroyed) by the new network
        self.h = h
        self.ctx = ctx
        self.base = base_channels
        ctx.unet = self

    WEIGHTS = PKG / "weights" / "unet_b32.bin"

    @classmethod
    def pretrained(cls, ctx: Context) -> "UNet":
        """The network trained by gknextengine_amd.denoise_train (weights shipped in the package)."""
        if not cls.WEIGHTS.exists():
            raise RuntimeError(f"{cls.WEIGHTS} is missing")
        net = cls(ctx, 32, seed=0)
        net.load(cls.WEIGHTS)
        return net

    def close(self):
        if self.h:
            lib().gk_unet_destroy(self.h)
            self.h = None
            if self.ctx.unet is self:
                self.ctx.unet = None

    def num_params(self) -> int:
        n = C.c_uint64()
        check(lib().gk_unet_num_params(self.h, C.byref(n)), "gk_unet_num_params")
        return n.value

    def layer(self, i: int) -> UNetLayer:
        L = UNetLayer()
        check(lib().gk_unet_layer_info(self.h, i, C.byref(L)), "gk_unet_layer_info")
        return L

    def params_ptr(self) -> int:
        p = C.c_void_p()
        check(lib().gk_unet_params(self.h, C.byref(p)), "gk_unet_params")
        return p.value

    def params_updated(self):
        check(lib().gk_unet_params_updated(self.h), "gk_unet_params_updated")

    def save(self, path):
        check(lib().gk_unet_save(self.h, str(path).encode()), "gk_unet_save")

    def load(self, path):
        check(lib().gk_unet_load(self.h, str(path).encode()), "gk_unet_load")

    HEADER = 32  # "GKUNET01", u32 base, u32 reserved, u64 count, u64 adam step

    def get_params(self) -> np.ndarray:
        """FP32 parameter vector on the host (through gk_unet_save)."""
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            f = Path(d) / "p.bin"
            self.save(f)
            return np.fromfile(f, np.float32, offset=self.HEADER)

    def set_params(self, params: np.ndarray, step: int = 0):
        """Replaces the parameters (through gk_unet_load; Adam moments restart)."""
        import tempfile
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == self.num_params()
        with tempfile.TemporaryDirectory() as d:
            f = Path(d) / "p.bin"
            with open(f, "wb") as fh:
                fh.write(b"GKUNET01")
                fh.write(np.array([self.base, 0], np.uint32).tobytes())
                fh.write(np.array([p.size, step], np.uint64).tobytes())
                fh.write(p.tobytes())
            self.load(f)

    # The network runs on its context's HIP stream. Inputs produced by torch (on torch's
    # stream) are made visible with a device synchronisation first; grad and train_step
    # return after the stream has drained (they read the loss back), so torch may free or
    # reuse their inputs right after.
    @staticmethod
    def _sync_torch():
        import torch
        torch.cuda.synchronize()

    def grad(self, d_in: int, d_target: int, n: int, h: int, w: int, d_grad: int) -> float:
        self._sync_torch()
        loss = C.c_float()
        check(lib().gk_unet_grad(self.h, d_in, d_target, n, h, w, d_grad, C.byref(loss)), "gk_unet_grad")
        return loss.value

    def train_step(self, d_in: int, d_target: int, n: int, h: int, w: int, lr: float) -> float:
        self._sync_torch()
        loss = C.c_float()
        check(lib().gk_unet_train_step(self.h, d_in, d_target, n, h, w, lr, C.byref(loss)), "gk_unet_train_step")
        return loss.value

    def infer(self, d_in: int, d_out: int, h: int, w: int):
        """Enqueues inference on the context's stream (call ctx.sync() before reading d_out)."""
        self._sync_torch()
        check(lib().gk_unet_infer(self.h, d_in, d_out, h, w), "gk_unet_infer")

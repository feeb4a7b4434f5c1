"""Python binding of the CPU oracle (test infrastructure only).

Loads ``oracle/_build/liboracle.so`` (the C++ restatement of the reference's path
tracing math) and ``oracle/_ref/libtinybvh_ref.so`` (the reference's tinybvh, compiled
from the reference tree). Imported only by tests/, ``__graft_entry__.smoke`` and
bench.py's cpu_baseline leg -- never by the product package.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "_build" / "liboracle.so"
REF_LIB = HERE / "_ref" / "libtinybvh_ref.so"

_P = C.c_void_p
_U = C.c_uint32
_o = None
_r = None


def ptr(a):
    return None if a is None else a.ctypes.data


def oracle():
    global _o
    if _o is None:
        if not ORACLE_LIB.exists():
            raise RuntimeError(f"{ORACLE_LIB} missing: run `make -C oracle`")
        L = C.CDLL(str(ORACLE_LIB))
        sig = {
            "oracle_scene_new": (_P, []),
            "oracle_scene_free": (None, [_P]),
            "oracle_scene_add_mesh": (C.c_int, [_P, _P, _P, _P, _P, _U, _U]),
            "oracle_scene_set_materials": (C.c_int, [_P, _P, _U]),
            "oracle_scene_add_instance": (C.c_int, [_P, _P]),
            "oracle_scene_build": (C.c_int, [_P]),
            "oracle_scene_update_instances": (C.c_int, [_P, _P, _U, _U]),
            "oracle_blas_stats": (None, [_P, _U, C.POINTER(_U), C.POINTER(C.c_float)]),
            "oracle_trace": (None, [_P, _P, _U, _P, _P, C.c_int, _U]),
            "oracle_primary_rays": (None, [_P, _U, _U, _P]),
            "oracle_lens_rays": (None, [_P, _U, _U, _P]),
            "oracle_concentric_disk": (None, [C.c_float, C.c_float, _P]),
            "oracle_vis_pack": (None, [_P, _P, _P, _U, _P]),
            "oracle_vis_decode": (None, [_P, _P, _P, _U, _P]),
            "oracle_render_refv2": (C.c_uint64, [_P, _P, _U, _U, _U, _U, _U, _U, _P, _P, _P, _P, _P, _P, _P, _P, _U]),
            "oracle_render_ptmis": (C.c_uint64, [_P, _P, _U, _U, _U, _U, _U, _U, _U, _U, C.c_int, _P, _P, _P, _U]),
            "oracle_render_ptmis_mode": (C.c_uint64, [_P, _P, _U, _U, _U, _U, _U, _U, _U, _U, C.c_int, _P, _U]),
            "oracle_pcg4d": (None, [_P, _U, _P]),
            "oracle_u2f": (C.c_float, [_U]),
            "oracle_sincos_2pi": (None, [_P, _U, _P, _P]),
            "oracle_sample": (None, [C.c_int, C.c_float, _P, _U, _U, _P]),
            "oracle_ptmis_spec_sample": (None, [C.c_double, _P, _U, _U, _P, _P]),
            "oracle_ptmis_spec_pdf": (C.c_double, [C.c_double, _P, _P]),
            "oracle_num_lights": (_U, [_P]),
            "oracle_diff_render": (None, [_P, _P, _P, _U, _U, _U, _U, _U, _U, _U, _U, _P, _U]),
            "oracle_probe_voxelize": (None, [_P, _P, _P, _U, _U, _U]),
            "oracle_probe_bake": (None, [_P, _P, _P, _P, _P, _P, _U, _U, _U]),
            "oracle_probe_lookup": (None, [_P, _P, _P, _P, _P, _U, C.c_int, _P]),
            "oracle_probe_pack": (_U, [_P]),
            "oracle_probe_unpack": (None, [_U, _P]),
            "oracle_scene_set_probes": (None, [_P, _P, _P, _P]),
            "oracle_scene_set_mesh_attribs": (C.c_int, [_P, _U, _P, _U, _U, _P, _P]),
            "oracle_scene_set_textures": (C.c_int, [_P, _P, _U]),
            "oracle_tex_sample": (None, [_P, _U, _P, _U, _P]),
            "oracle_scene_set_sky": (C.c_int, [_P, _P, _U, _U]),
            "oracle_sky_sh": (None, [_P, _P]),
            "oracle_sky_radiance": (None, [_P, _P, _P, _U, _P]),
            "oracle_sky_rough": (None, [_P, _P, _P, _U, _P]),
            "oracle_atan2_turns": (None, [_P, _P, _U, _P]),
        }
        optional = {
            "oracle_reproject": (None, [_P, _U, _U] + [_P] * 12),
            "oracle_jbf": (None, [_P, _U, _U] + [_P] * 5),
            "oracle_gt_tonemap": (C.c_float, [C.c_float]),
            "oracle_upscale": (None, [C.c_int, _P, _U, _U, _U, _U, _P]),
            "oracle_st2084": (C.c_float, [C.c_float]),
            "oracle_display": (C.c_float, [C.c_float, C.c_float, C.c_int]),
        }
        for name, (res, args) in {**sig, **optional}.items():
            if not hasattr(L, name):
                if name in sig:
                    raise RuntimeError(f"{name} missing from {ORACLE_LIB}")
                continue
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _o = L
    return _o


def ref_available() -> bool:
    return REF_LIB.exists()


def ref():
    global _r
    if _r is None:
        if not REF_LIB.exists():
            raise RuntimeError(f"{REF_LIB} missing: run `make -C oracle ref` where the reference tree exists")
        L = C.CDLL(str(REF_LIB))
        sig = {
            "ref_hardware_concurrency": (C.c_uint, []),
            "ref_blas_build": (_P, [_P, _U, C.POINTER(C.c_double)]),
            "ref_blas_free": (None, [_P]),
            "ref_blas_stats": (None, [_P, C.POINTER(_U), C.POINTER(C.c_double), _P, _P]),
            "ref_intersect": (None, [_P, _P, _U, _P, _U, C.c_int, C.POINTER(C.c_double)]),
            "ref_occluded": (None, [_P, _P, _U, _P, _U, C.POINTER(C.c_double)]),
            "ref_tlas_build": (_P, [_P, _U, _P, _P, _U, C.POINTER(C.c_double)]),
            "ref_tlas_rebuild": (None, [_P, _P, C.POINTER(C.c_double)]),
            "ref_tlas_free": (None, [_P]),
            "ref_tlas_intersect": (None, [_P, _P, _U, _P, _P, _U, C.POINTER(C.c_double)]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _r = L
    return _r


def atan2_turns(y, x):
    """The oracle's atan2(y, x) / (2 pi) (Maclaurin series, bit-identical to the device's)."""
    y = np.ascontiguousarray(y, np.float32).ravel()
    x = np.ascontiguousarray(x, np.float32).ravel()
    out = np.zeros(y.size, np.float32)
    oracle().oracle_atan2_turns(ptr(y), ptr(x), y.size, ptr(out))
    return out


# ---------------------------------------------------------------- oracle scene
class OracleScene:
    """Oracle copy of a gknextengine_amd.SceneData (same arrays and structs)."""

    def __init__(self, scene):
        L = oracle()
        self.h = L.oracle_scene_new()
        self.scene = scene
        for m in scene.meshes:
            r = L.oracle_scene_add_mesh(self.h, ptr(m["positions"]), ptr(m["normals"]), ptr(m["indices"]),
                                        ptr(m["tri_material"]), m["positions"].shape[0], m["indices"].shape[0])
            assert r >= 0
            if m.get("texcoords") is not None or m.get("tangents") is not None:
                assert L.oracle_scene_set_mesh_attribs(self.h, r, ptr(m["indices"]), m["positions"].shape[0],
                                                       m["indices"].shape[0], ptr(m.get("texcoords")),
                                                       ptr(m.get("tangents"))) == 0
        if getattr(scene, "textures", None):
            self._tex = scene.c_textures()
            L.oracle_scene_set_textures(self.h, C.addressof(self._tex), len(scene.textures))
        mats = scene.c_materials()
        L.oracle_scene_set_materials(self.h, C.addressof(mats), len(scene.materials))
        for inst in scene.instances:
            L.oracle_scene_add_instance(self.h, C.addressof(inst))
        assert L.oracle_scene_build(self.h) == 0

    def __del__(self):
        try:
            oracle().oracle_scene_free(self.h)
        except Exception:
            pass

    def update_instances(self, worlds, first=0):
        w = np.ascontiguousarray(np.transpose(np.asarray(worlds, np.float32), (0, 2, 1)))
        assert oracle().oracle_scene_update_instances(self.h, ptr(w), first, w.shape[0]) == 0

    def blas_stats(self, mesh=0):
        n, s = _U(0), C.c_float(0)
        oracle().oracle_blas_stats(self.h, mesh, C.byref(n), C.byref(s))
        return n.value, s.value

    def trace(self, rays, any_hit=False, threads=None):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        inst = np.zeros(n, np.uint32)
        oracle().oracle_trace(self.h, ptr(rays), n, ptr(hits), ptr(inst), int(any_hit), threads or os.cpu_count())
        return hits, inst

    def tex_sample(self, tex_id, uv):
        uv = np.ascontiguousarray(uv, np.float32).reshape(-1, 2)
        out = np.zeros((uv.shape[0], 4), np.float32)
        oracle().oracle_tex_sample(self.h, tex_id, ptr(uv), uv.shape[0], ptr(out))
        return out

    def set_sky(self, rgba=None):
        """The sky image (H x W x 4 float32, row 0 at +y); None removes it."""
        if rgba is None:
            oracle().oracle_scene_set_sky(self.h, None, 0, 0)
            self._sky = None
            return
        self._sky = np.ascontiguousarray(rgba, np.float32)
        assert oracle().oracle_scene_set_sky(self.h, ptr(self._sky), self._sky.shape[1], self._sky.shape[0]) == 0

    def sky_sh(self):
        out = np.zeros(27, np.float32)
        oracle().oracle_sky_sh(self.h, ptr(out))
        return out.reshape(3, 9)

    def sky_radiance(self, params, dirs, rough=False):
        """SampleIBLV2 (rough=False) or SampleIBLRough (rough=True) of the frame's sky at
        (n, 3) directions, times sky_color."""
        d = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
        out = np.zeros((d.shape[0], 4), np.float32)
        fn = oracle().oracle_sky_rough if rough else oracle().oracle_sky_radiance
        fn(self.h, C.byref(params), ptr(d), d.shape[0], ptr(out))
        return out

    def vis_pack(self, hits, inst):
        """The raster visibility pass's ids of traced hits ((n, 4) {t, u, v, prim bits}, (n,)
        instances): (proxy + 1) << 17 | triangle within the 65535-triangle section."""
        h = np.ascontiguousarray(hits, np.float32)
        i = np.ascontiguousarray(inst, np.uint32)
        out = np.zeros(h.shape[0], np.uint32)
        oracle().oracle_vis_pack(self.h, ptr(h), ptr(i), h.shape[0], ptr(out))
        return out

    def vis_decode(self, packed, rays):
        """get_material_dataV2 for packed ids and their rays: (instance, triangle, world-space
        barycentrics (3,), interpolated world position (3,)); instance 0xffffffff for none."""
        pk = np.ascontiguousarray(packed, np.uint32).ravel()
        r = np.ascontiguousarray(rays, np.float32)
        out = np.zeros((pk.size, 8), np.float32)
        oracle().oracle_vis_decode(self.h, ptr(pk), ptr(r), pk.size, ptr(out))
        return out[:, 0].view(np.uint32), out[:, 1].view(np.uint32), out[:, 2:5], out[:, 5:8]

    def num_lights(self):
        return oracle().oracle_num_lights(self.h)

    def render_refv2(self, params, W, H, rect=None, moment=None, threads=None):
        x0, y0, w, h = rect or (0, 0, W, H)
        out = {
            "diffuse": np.zeros((h, w, 4), np.float32), "spec": np.zeros((h, w, 4), np.float32),
            "albedo": np.zeros((h, w, 4), np.float32), "normal": np.zeros((h, w, 4), np.float32),
            "objid": np.zeros((h, w), np.uint32), "motion": np.zeros((h, w, 2), np.float32),
            "depth": np.zeros((h, w), np.float32),
        }
        mom = np.zeros((h, w), np.uint16) if moment is None else moment
        rays = oracle().oracle_render_refv2(self.h, C.addressof(params), W, H, x0, y0, w, h,
                                            *(ptr(out[k]) for k in ("diffuse", "spec", "albedo", "normal", "objid",
                                                                    "motion", "depth")),
                                            ptr(mom), threads or os.cpu_count())
        out["moment"] = mom
        out["rays"] = rays
        return out

    def diff_render(self, params, W, H, eval_materials=None, rect=None, frame0=0, spp=1, threads=None):
        """Detached differentiable PT_MIS in double: sampling from the scene's materials,
        BSDF values from eval_materials (list of Material, same length) when given."""
        x0, y0, w, h = rect or (0, 0, W, H)
        rgb = np.zeros((h, w, 3), np.float64)
        ev = None
        if eval_materials is not None:
            ev = (type(eval_materials[0]) * len(eval_materials))(*eval_materials)
        oracle().oracle_diff_render(self.h, C.addressof(params), None if ev is None else C.addressof(ev), W, H,
                                    x0, y0, w, h, frame0, spp, ptr(rgb), threads or os.cpu_count())
        return rgb

    def render_ptmis(self, params, W, H, rect=None, frame0=0, frames=1, fp64=False, threads=None):
        x0, y0, w, h = rect or (0, 0, W, H)
        rgb = np.zeros((h, w, 3), np.float64)
        d = np.zeros((h, w, 3), np.float64)
        s = np.zeros((h, w, 3), np.float64)
        rays = oracle().oracle_render_ptmis(self.h, C.addressof(params), W, H, x0, y0, w, h, frame0, frames,
                                            int(fp64), ptr(rgb), ptr(d), ptr(s), threads or os.cpu_count())
        return {"rgb": rgb, "diffuse": d, "spec": s, "rays": rays}

    # strategy modes of the PT_MIS oracle (double precision); the PLANTED_* modes are wrong
    # emitter-hit weights, for tests that show their checks fail on them
    MIS, NEE_ONLY, BSDF_ONLY = 0, 1, 2
    PLANTED_NO_COSINE, PLANTED_NO_SELECTION_PDF, PLANTED_BALANCE_HEURISTIC = 3, 4, 5

    def render_ptmis_mode(self, params, W, H, mode, rect=None, frame0=0, frames=1, threads=None):
        """PT_MIS radiance (h, w, 3), double precision, with only some strategies carrying
        light: MIS (the estimator), NEE_ONLY, BSDF_ONLY, or a PLANTED_* mistake. Every mode
        draws the same random numbers, so renders of one seed follow the same paths."""
        x0, y0, w, h = rect or (0, 0, W, H)
        rgb = np.zeros((h, w, 3), np.float64)
        oracle().oracle_render_ptmis_mode(self.h, C.addressof(params), W, H, x0, y0, w, h, frame0, frames, int(mode),
                                          ptr(rgb), threads or os.cpu_count())
        return rgb


# ---------------------------------------------------------------- probes
class ProbeGridO(C.Structure):
    _fields_ = [("size_xy", _U), ("size_z", _U), ("unit", C.c_float), ("offset", C.c_float * 3)]


class OracleProbes:
    """Probe grid arrays (numpy records of the product's CUBE_DTYPE / VOXEL_DTYPE layout)
    with the oracle's voxelize / bake / lookup."""

    def __init__(self, size_xy, size_z, unit, offset, cube_dtype, voxel_dtype):
        self.g = ProbeGridO(size_xy, size_z, unit, (C.c_float * 3)(*offset))
        self.n = size_xy * size_xy * size_z
        self.cubes = np.zeros(self.n, cube_dtype)
        self.vox = np.zeros(self.n, voxel_dtype)

    def voxelize(self, scene: OracleScene, threads=None):
        oracle().oracle_probe_voxelize(scene.h, C.addressof(self.g), ptr(self.vox), 0, self.n,
                                       threads or os.cpu_count())

    def bake(self, scene: OracleScene, params, light=None, first=0, count=None, threads=None):
        n = self.n - first if count is None else count
        oracle().oracle_probe_bake(scene.h, C.addressof(self.g), C.addressof(params),
                                   None if light is None else C.addressof(light), ptr(self.cubes), ptr(self.vox),
                                   first, n, threads or os.cpu_count())

    def lookup(self, pos, nrm, full=True):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        out = np.zeros((pos.shape[0], 4), np.float32)
        oracle().oracle_probe_lookup(C.addressof(self.g), ptr(self.cubes), ptr(self.vox), ptr(pos), ptr(nrm),
                                     pos.shape[0], int(full), ptr(out))
        return out

    def attach(self, scene: OracleScene):
        """REF_V2 renders of scene terminate paths with this grid (keep self alive)."""
        oracle().oracle_scene_set_probes(scene.h, C.addressof(self.g), ptr(self.cubes), ptr(self.vox))


def probe_pack(c):
    a = np.ascontiguousarray(c, np.float32).reshape(4)
    return oracle().oracle_probe_pack(ptr(a))


def probe_unpack(p):
    out = np.zeros(4, np.float32)
    oracle().oracle_probe_unpack(int(p), ptr(out))
    return out


def primary_rays(params, W, H):
    rays = np.zeros((H * W, 8), np.float32)
    oracle().oracle_primary_rays(C.addressof(params), W, H, ptr(rays))
    return rays


def lens_rays(params, W, H):
    """Primary rays with the frame's thin lens (aperture 0: the pinhole rays)."""
    rays = np.zeros((W * H, 8), np.float32)
    oracle().oracle_lens_rays(C.addressof(params), W, H, ptr(rays))
    return rays


def concentric_disk(ua, ub):
    """RandomInUnitDisk's map of one square sample (ua, ub) in [0, 1)^2."""
    out = np.zeros(2, np.float32)
    oracle().oracle_concentric_disk(float(ua), float(ub), ptr(out))
    return out


def pcg4d(seed, n):
    s = np.ascontiguousarray(seed, np.uint32)
    out = np.zeros((n, 4), np.uint32)
    oracle().oracle_pcg4d(ptr(s), n, ptr(out))
    return out


def sincos_2pi(u):
    u = np.ascontiguousarray(u, np.float32)
    s = np.zeros_like(u)
    c = np.zeros_like(u)
    oracle().oracle_sincos_2pi(ptr(u), u.size, ptr(s), ptr(c))
    return s, c


def sample(kind, n, param=0.0, direction=(0, 0, 1), seed=1):
    d = np.asarray(direction, np.float32)
    out = np.zeros((n, 3), np.float32)
    oracle().oracle_sample(kind, param, ptr(d), n, seed, ptr(out))
    return out


# ---------------------------------------------------------------- tinybvh (reference)
def scene_triangles(mesh) -> np.ndarray:
    """(n, 9) float32 triangle corners of a SceneData mesh dict."""
    return np.ascontiguousarray(mesh["positions"][mesh["indices"]].reshape(-1, 9), np.float32)


class TinyBVH:
    """The reference's CPU BVH (tinybvh 1.3.8, binned SAH) over one mesh."""

    def __init__(self, tris9):
        self.tris = np.ascontiguousarray(tris9, np.float32)
        ms = C.c_double(0)
        self.h = ref().ref_blas_build(ptr(self.tris), self.tris.shape[0], C.byref(ms))
        self.build_ms = ms.value

    def __del__(self):
        try:
            ref().ref_blas_free(self.h)
        except Exception:
            pass

    def stats(self):
        n, s = _U(0), C.c_double(0)
        ref().ref_blas_stats(self.h, C.byref(n), C.byref(s), None, None)
        return n.value, s.value

    def intersect(self, rays, threads=1, normalize=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        ms = C.c_double(0)
        ref().ref_intersect(self.h, ptr(rays), rays.shape[0], ptr(hits), threads, int(normalize), C.byref(ms))
        return hits, ms.value

    def occluded(self, rays, threads=1):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros(rays.shape[0], np.uint32)
        ms = C.c_double(0)
        ref().ref_occluded(self.h, ptr(rays), rays.shape[0], ptr(out), threads, C.byref(ms))
        return out, ms.value


class TinyTLAS:
    def __init__(self, blases, worlds_rowmajor, blas_idx):
        self.blases = blases
        arr = (C.c_void_p * len(blases))(*[b.h for b in blases])
        self._arr = arr
        self.xf = np.ascontiguousarray(worlds_rowmajor, np.float32).reshape(-1, 16)
        self.idx = np.ascontiguousarray(blas_idx, np.uint32)
        ms = C.c_double(0)
        self.h = ref().ref_tlas_build(arr, len(blases), ptr(self.xf), ptr(self.idx), self.idx.size, C.byref(ms))
        self.build_ms = ms.value

    def rebuild(self, worlds_rowmajor):
        self.xf = np.ascontiguousarray(worlds_rowmajor, np.float32).reshape(-1, 16)
        ms = C.c_double(0)
        ref().ref_tlas_rebuild(self.h, ptr(self.xf), C.byref(ms))
        return ms.value

    def __del__(self):
        try:
            ref().ref_tlas_free(self.h)
        except Exception:
            pass

    def intersect(self, rays, threads=1):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        inst = np.zeros(rays.shape[0], np.uint32)
        ms = C.c_double(0)
        ref().ref_tlas_intersect(self.h, ptr(rays), rays.shape[0], ptr(hits), ptr(inst), threads, C.byref(ms))
        return hits, inst, ms.value


# ---------------------------------------------------------------- post passes
def _f4(a):
    return np.ascontiguousarray(np.asarray(a, np.float32))


def reproject(src_d, src_s, src_a, normal, prev_d, prev_s, prev_a, obj0, obj1, motion, moment, acc_init,
              temporal_frames=16, progressive=False):
    """Process.ReProject restatement. All colour inputs (H, W, 4) float32 (fp16 values);
    acc_init supplies the alpha kept by the .rgb-only writes. Returns (diffuse, spec, albedo)."""
    H, W = obj0.shape
    params = np.array([int(progressive), temporal_frames], np.uint32)
    acc = np.ascontiguousarray(np.stack([_f4(a) for a in acc_init]), np.float32)
    args = [_f4(x) for x in (src_d, src_s, src_a, normal, prev_d, prev_s, prev_a)]
    o0 = np.ascontiguousarray(obj0, np.uint32)
    o1 = np.ascontiguousarray(obj1, np.uint32)
    mv = _f4(motion)
    mo = np.ascontiguousarray(moment, np.uint16)
    oracle().oracle_reproject(ptr(params), W, H, *(ptr(a) for a in args), ptr(o0), ptr(o1), ptr(mv), ptr(mo), ptr(acc))
    return acc[0], acc[1], acc[2]


def jbf(acc_d, acc_s, acc_a, sigma=2.0, sigma_lum=3.0, paper_white=600.0, bf_size=5, hdr=False, linear=False):
    """Process.DenoiseJBF restatement: display colour, or with linear the composed radiance."""
    H, W = acc_d.shape[:2]
    fp = np.array([sigma, sigma_lum, paper_white, 1.0 if hdr else 0.0, 1.0 if linear else 0.0], np.float32)
    bs = np.array([bf_size], np.uint32)
    out = np.zeros((H, W, 4), np.float32)
    d, s, a = _f4(acc_d), _f4(acc_s), _f4(acc_a)
    oracle().oracle_jbf(ptr(fp), W, H, ptr(d), ptr(s), ptr(a), ptr(bs), ptr(out))
    return out


def upscale(img, out_w, out_h, mode=0):
    """The resolve pass on an (H, W, 4) float image (mode 0 linear blit, 1 FSR EASU)."""
    a = np.ascontiguousarray(img, np.float32)
    H, W = a.shape[:2]
    out = np.zeros((out_h, out_w, 4), np.float32)
    oracle().oracle_upscale(int(mode), ptr(a), W, H, out_w, out_h, ptr(out))
    return out

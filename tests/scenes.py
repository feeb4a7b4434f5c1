"""Benchmark scenes shared by the tests (the benchmark configurations, scaled where noted)."""
from __future__ import annotations

import numpy as np

import gknextengine_amd as gk

# the benchmark cameras and the config-3 scene are the package's own (one definition)
from gknextengine_amd.scenes import BRICKS_CAM, CITY_CAM, CORNELL_CAM, Bricks  # noqa: E402,F401


def glossy_cornell():
    """The Cornell box with glossy materials and a second, large emitter: the ceiling
    (1 x 1, radiance (1, 0.95, 0.9)) besides the small ceiling light. Red wall Mixture
    (roughness 0.6), green wall Lambertian, white walls Mixture (roughness 0.35, metallic
    0.25), short box Mixture (0.3, 0.5), tall box Metallic (0.15). Both emitters make BSDF
    sampling and next-event estimation each carry a large share of the light."""
    from gknextengine_amd.scenes import material
    s = gk.procgen_cornell()
    tm = s.meshes[0]["tri_material"]
    tm[2:4] = 4          # ceiling
    tm[10:22] = 5        # short box
    tm[22:34] = 6        # tall box
    s.materials = [material((0.7, 0.07, 0.07), gk.MAT_MIXTURE, 0.6, 0.0),
                   material((0.07, 0.7, 0.07), gk.MAT_LAMBERTIAN, 1.0, 0.0),
                   material((0.7, 0.7, 0.7), gk.MAT_MIXTURE, 0.35, 0.25),
                   s.materials[3],
                   material((1.0, 0.95, 0.9), gk.MAT_DIFFUSE_LIGHT),
                   material((0.8, 0.6, 0.3), gk.MAT_MIXTURE, 0.3, 0.5),
                   material((0.9, 0.9, 0.9), gk.MAT_METALLIC, 0.15, 1.0)]
    return s


def all_models_cornell():
    """The Cornell box with every material model the estimators shade: walls Mixture,
    Metallic and Lambertian, the short box Dielectric (ior 1.5), the tall box Mixture."""
    from gknextengine_amd.scenes import material
    s = gk.procgen_cornell()
    tm = s.meshes[0]["tri_material"]
    tm[10:22] = 4
    tm[22:34] = 5
    s.materials = [material((0.7, 0.1, 0.1), gk.MAT_MIXTURE, 0.4, 0.3),
                   material((0.1, 0.7, 0.1), gk.MAT_METALLIC, 0.25, 1.0),
                   material((0.7, 0.7, 0.7), gk.MAT_LAMBERTIAN, 1.0, 0.0),
                   s.materials[3],
                   material((0.95, 0.95, 0.95), gk.MAT_DIELECTRIC, 0.0, 0.0, 1.5),
                   material((0.6, 0.6, 0.8), gk.MAT_MIXTURE, 0.2, 0.8)]
    s.materials[4].refraction_index2 = 1.5
    return s


def random_rays(n, lo, hi, seed=0, tmin=0.0, tmax=1e30):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = o
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


# ---------------------------------------------------------------- probe-grid test scenes
def tri_scene(tris, albedo=(0.5, 0.5, 0.5), extra_materials=(), mat=0):
    """One mesh of the given triangles ((n,3,3) vertex positions), flat normals along
    cross(v1-v0, v2-v0), all on material mat; material 0 is Lambertian, extra materials
    follow."""
    from gknextengine_amd.scenes import material
    tris = np.asarray(tris, np.float32).reshape(-1, 3, 3)
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    s = gk.SceneData()
    s.add_mesh(tris.reshape(-1, 3), np.arange(3 * len(tris), dtype=np.uint32).reshape(-1, 3),
               np.repeat(n, 3, axis=0).astype(np.float32), np.full(len(tris), mat, np.uint32))
    s.materials = [material(albedo, model=gk.MAT_LAMBERTIAN), *extra_materials]
    s.add_instance(0)
    return s


def floor_tris(L=100.0, y=0.0):
    """Two triangles of the plane y = const, geometric normal +y."""
    return [[(-L, y, -L), (-L, y, L), (L, y, L)], [(-L, y, -L), (L, y, L), (L, y, -L)]]


def small_tri(center, normal, r=0.2):
    """A triangle of circumradius r around center, facing along normal."""
    c = np.asarray(center, np.float64)
    n = np.asarray(normal, np.float64)
    n /= np.linalg.norm(n)
    a = np.cross(n, (0.0, 1.0, 0.0) if abs(n[1]) < 0.9 else (1.0, 0.0, 0.0))
    a /= np.linalg.norm(a)
    b = np.cross(n, a)
    pts = [c + r * (np.cos(t) * a + np.sin(t) * b) for t in (0.0, 2.0944, 4.1888)]
    return [pts]


def quad_light(center, half, area, mat_idx):
    """A LightObject quad at center (y = const plane) whose p1/p3 midpoint is center."""
    L = gk.Light()
    cx, cy, cz = center
    L.p0[:] = [cx - half, cy, cz - half, 1.0]
    L.p1[:] = [cx - half, cy, cz + half, 1.0]
    L.p3[:] = [cx + half, cy, cz - half, 1.0]
    L.normal_area[:] = [0.0, -1.0, 0.0, area]
    L.light_mat_idx = mat_idx
    return L


# ---------------------------------------------------------------- textured test scenes
def planar_uv(positions, normals, scale=3.0):
    """Texture coordinates by projection along each vertex normal's dominant axis, and the
    matching object-space tangent (the u axis, handedness +1)."""
    p = np.asarray(positions, np.float64)
    n = np.asarray(normals, np.float64)
    ax = np.argmax(np.abs(n), axis=1)
    uv = np.zeros((len(p), 2))
    tan = np.zeros((len(p), 4))
    tan[:, 3] = 1.0
    for a, (iu, iv) in enumerate(((2, 1), (0, 2), (0, 1))):
        sel = ax == a
        uv[sel, 0] = p[sel, iu] * scale
        uv[sel, 1] = p[sel, iv] * scale
        tan[sel, iu] = 1.0
    return uv.astype(np.float32), tan.astype(np.float32)


def checker(n=8, cells=2, seed=0):
    rng = np.random.default_rng(seed)
    cols = rng.integers(40, 255, (cells * cells, 4)).astype(np.uint8)
    cols[:, 3] = 255
    y, x = np.mgrid[0:n, 0:n]
    k = (y * cells // n) * cells + (x * cells // n)
    return cols[k]


def textured_cornell():
    """The Cornell box with planar texture coordinates: a checker base colour and a
    metallic-roughness map on the white material, a normal map on the red one."""
    s = gk.procgen_cornell()
    m = s.meshes[0]
    uv, tan = planar_uv(m["positions"], m["normals"])
    m["texcoords"], m["tangents"] = uv, tan
    rng = np.random.default_rng(3)
    t_col = s.add_texture(checker(8, 4, seed=1), srgb=True)
    mra = rng.integers(0, 256, (4, 4, 4)).astype(np.uint8)
    t_mra = s.add_texture(mra, srgb=False)
    nm = np.zeros((4, 4, 4), np.uint8)
    nm[..., 0] = rng.integers(60, 196, (4, 4))
    nm[..., 1] = rng.integers(60, 196, (4, 4))
    nm[..., 2] = 230
    nm[..., 3] = 255
    t_nrm = s.add_texture(nm, srgb=False)
    s.materials[2].diffuse_texture_id = t_col
    s.materials[2].mra_texture_id = t_mra
    s.materials[0].normal_texture_id = t_nrm
    return s


def face_plane_rays():
    """Rays of the unit Cornell box whose origin lies exactly on a face plane (x, y or z in
    {0, 1}) at grid points just outside, on and inside the box, each fired along the two
    other axes (parallel to that plane) both ways: tinybvh's slab distance to that plane
    is exactly 0 for them, which decides whether a box that ends on the plane is entered
    (the probe voxelize casts such rays from probes that sit on wall planes)."""
    coords = np.array([-0.0625, 0.0, 0.25, 0.5, 1.0, 1.0625], np.float32)
    rays = []
    for ax in range(3):
        others = [a for a in range(3) if a != ax]
        for plane in (0.0, 1.0):
            for c0 in coords:
                for c1 in coords:
                    O = np.zeros(3, np.float32)
                    O[ax] = plane
                    O[others[0]] = c0
                    O[others[1]] = c1
                    for d_ax in others:
                        for sgn in (1.0, -1.0):
                            D = np.zeros(3, np.float32)
                            D[d_ax] = sgn
                            rays.append([*O, 0.0, *D, 4.0])
    return np.asarray(rays, np.float32)


# ---------------------------------------------------------------- outdoor scene under a sky image
SKY_CAM = {"eye": (0.0, 1.4, 4.5), "target": (0.0, 0.6, 0.0), "fov": 60.0}


def sky_image(w=128, h=64, seed=0):
    """An equirectangular RGBA32F test sky (row 0 at +y): a zenith-to-horizon gradient, a
    colour that varies with longitude, a sun disc brighter than the sampler's clamp of 10
    and noise, so that a wrong longitude, latitude or filter shows."""
    rng = np.random.default_rng(seed)
    v = (np.arange(h)[:, None] + 0.5) / h
    u = (np.arange(w)[None, :] + 0.5) / w
    img = np.zeros((h, w, 4), np.float32)
    img[..., 0] = 0.4 + 0.6 * v + 0.3 * np.cos(2 * np.pi * u)
    img[..., 1] = 0.6 + 0.4 * np.sin(np.pi * v) + 0.2 * np.sin(2 * np.pi * u)
    img[..., 2] = 1.2 - 0.8 * v + 0.1 * np.cos(6 * np.pi * u)
    img[..., :3] *= 1.0 + 0.2 * rng.random((h, w, 3))
    sun = (np.abs(v - 0.3) < 0.05) & (np.abs(u - 0.6) < 0.03)
    img[sun, :3] = [40.0, 35.0, 30.0]
    img[..., 3] = 1.0 + 0.5 * rng.random((h, w))   # alpha above 1: the clamp shows
    return img


def sky_scene():
    """A Lambertian floor and four bricks (Lambertian, metal, glass, mixture) in the open."""
    from gknextengine_amd.scenes import material
    tris = np.asarray(floor_tris(L=20.0), np.float32)
    s = gk.SceneData()
    s.add_mesh(tris.reshape(-1, 3), np.arange(6, dtype=np.uint32).reshape(-1, 3),
               np.tile(np.float32([0, 1, 0]), (6, 1)), np.zeros(2, np.uint32))
    bp, bn, bi, bm = gk.procgen_brick()
    s.add_mesh(bp, bi, bn, bm)
    s.materials = [material((0.45, 0.45, 0.4), gk.MAT_LAMBERTIAN),
                   material((0.8, 0.3, 0.2), gk.MAT_LAMBERTIAN),
                   material((0.9, 0.9, 0.9), gk.MAT_METALLIC, rough=0.05),
                   material((1.0, 1.0, 1.0), gk.MAT_DIELECTRIC, rough=0.0, ior=1.5),
                   material((0.3, 0.6, 0.8), gk.MAT_MIXTURE, rough=0.3, metal=0.5)]
    s.add_instance(0)
    for k, (x, z, ang) in enumerate(((-1.6, 0.0, 0.3), (-0.5, -0.6, -0.4), (0.6, 0.4, 0.8), (1.7, -0.3, 0.1))):
        w = np.eye(4, dtype=np.float32)
        c, sn = np.cos(ang), np.sin(ang)
        w[:3, :3] = np.float32([[c, 0, sn], [0, 1, 0], [-sn, 0, c]]) * 1.2
        w[0, 3], w[2, 3] = x, z
        s.add_instance(1, world=w, mat_offset=1 + k)
    return s

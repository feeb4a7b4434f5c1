"""Benchmark scenes of the five measured configurations (host-side scene setup).

    cornell   unit-cube Cornell box (gk_procgen_cornell), 256x256
    city      ~0.92 M-triangle procedural city (gk_procgen_city), 1920x1080
    bricks    4096 instances of a 24-triangle brick (gk_procgen_brick) on a
              16^3 lattice with random rigid motion (angular velocity N(0,0.5) rad/s about
              a random axis, linear velocity N(0,1) m/s), 8 materials
    tiles     the Cornell box with its walls split into 256 material tiles
"""
from __future__ import annotations

import numpy as np

from . import (MAT_DIFFUSE_LIGHT, MAT_LAMBERTIAN, MAT_MIXTURE, Material, SceneData, procgen_brick,
               procgen_cornell)

CORNELL_CAM = dict(eye=(0.5, 0.5, -1.4), target=(0.5, 0.5, 0.0), fov=40.0)
CITY_CAM = dict(eye=(-20.0, 28.0, -44.0), target=(2.0, 0.0, 4.0), fov=50.0)
BRICKS_CAM = dict(eye=(0.0, 10.0, -30.0), target=(0.0, 0.0, 0.0), fov=50.0)


def material(rgb, model=MAT_MIXTURE, rough=1.0, metal=0.0, ior=1.45):
    m = Material()
    m.diffuse[:] = [float(rgb[0]), float(rgb[1]), float(rgb[2]), 1.0]
    m.diffuse_texture_id = m.mra_texture_id = m.normal_texture_id = -1
    m.fuzziness = float(rough)
    m.refraction_index = float(ior)
    m.material_model = int(model)
    m.metalness = float(metal)
    return m


def rotations(axes: np.ndarray, angles: np.ndarray) -> np.ndarray:
    """Rodrigues rotation matrices, (n,3) axes and (n,) angles -> (n,3,3)."""
    a = axes / np.linalg.norm(axes, axis=1, keepdims=True)
    x, y, z = a[:, 0], a[:, 1], a[:, 2]
    c, s = np.cos(angles), np.sin(angles)
    C = 1.0 - c
    R = np.empty((a.shape[0], 3, 3))
    R[:, 0, 0] = c + x * x * C
    R[:, 0, 1] = x * y * C - z * s
    R[:, 0, 2] = x * z * C + y * s
    R[:, 1, 0] = y * x * C + z * s
    R[:, 1, 1] = c + y * y * C
    R[:, 1, 2] = y * z * C - x * s
    R[:, 2, 0] = z * x * C - y * s
    R[:, 2, 1] = z * y * C + x * s
    R[:, 2, 2] = c + z * z * C
    return R


class Bricks:
    """4096 instanced bricks with per-instance rigid motion."""

    def __init__(self, n_side=16, spacing=1.5, n_materials=8, seed=7):
        rng = np.random.default_rng(seed)
        pos, nrm, idx, mat = procgen_brick()
        self.scene = SceneData()
        self.scene.add_mesh(pos, idx, nrm, mat)
        mats = [material(rng.uniform(0.2, 0.9, 3), MAT_MIXTURE, rng.uniform(0.1, 1.0), float(rng.random() < 0.3))
                for _ in range(n_materials)]
        mats.append(material((8.0, 8.0, 8.0), MAT_DIFFUSE_LIGHT, 0.0, 0.0, 0.0))
        self.scene.materials = mats
        n = n_side ** 3
        g = (np.arange(n_side) - (n_side - 1) / 2.0) * spacing
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        self.pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        self.axis = rng.normal(size=(n, 3))
        self.omega = rng.normal(0.0, 0.5, n)
        self.vel = rng.normal(0.0, 1.0, (n, 3))
        self.angle0 = rng.uniform(0, 2 * np.pi, n)
        self.mat_base = rng.integers(0, n_materials - 1, n)
        self.n = n
        for i, w in enumerate(self.worlds(0.0)):
            # slot 0: body, slot 1: studs; one brick in 64 has emissive studs
            off = int(self.mat_base[i]) if i % 64 else n_materials - 1
            self.scene.add_instance(0, world=w, instance_id=i, mat_offset=off)

    def worlds(self, t: float) -> np.ndarray:
        """(n,4,4) row-major world matrices at time t (seconds)."""
        out = np.zeros((self.n, 4, 4), np.float32)
        out[:, :3, :3] = rotations(self.axis, self.angle0 + self.omega * t)
        out[:, :3, 3] = self.pos + self.vel * t
        out[:, 3, 3] = 1.0
        return out


def cornell_tiles(tiles_per_wall=8, seed=3):
    """The differentiable-rendering scene: the Cornell box whose walls and boxes are split into tiles with
    their own Mixture material (256 materials = 5 walls x 8x8 tiles truncated, plus the
    boxes' faces), light unchanged. Returns (SceneData, material count)."""
    base = procgen_cornell()
    pos, idx, nrm, tm = (base.meshes[0][k] for k in ("positions", "indices", "normals", "tri_material"))
    rng = np.random.default_rng(seed)
    s = SceneData()
    P, N, I, M = [], [], [], []
    n_mat = 0

    def add_quad(a, b, c, d, n, m):
        k = len(P)
        P.extend([a, b, c, d])
        N.extend([n] * 4)
        I.append([k, k + 1, k + 2])
        I.append([k, k + 2, k + 3])
        M.extend([m, m])

    # walls: the first 10 triangles are 5 quads (floor, ceiling, back, left, right)
    light_mat = None
    for q in range(int(idx.shape[0] // 2)):
        t0 = idx[2 * q]
        a, b, c = pos[t0[0]], pos[t0[1]], pos[t0[2]]
        d = pos[idx[2 * q + 1][2]]
        n = nrm[t0[0]]
        if tm[2 * q] == 3:  # light quad stays one emitter
            light_mat = "light"
            add_quad(a, b, c, d, n, -1)
            continue
        k = tiles_per_wall if q < 5 else 2
        for i in range(k):
            for j in range(k):
                u0, u1, v0, v1 = i / k, (i + 1) / k, j / k, (j + 1) / k

                def L(u, v):
                    return (a + (b - a) * u) * (1 - v) + (d + (c - d) * u) * v
                add_quad(L(u0, v0), L(u1, v0), L(u1, v1), L(u0, v1), n, n_mat)
                n_mat += 1
    # 5 walls x 8x8 + 12 box faces x 2x2 = 368 tiles share 256 materials (tile % 256)
    n_unique = min(n_mat, 256)
    mats = []
    for i in range(n_unique):
        mats.append(material(rng.uniform(0.2, 0.9, 3), MAT_MIXTURE, rng.uniform(0.2, 0.9), rng.uniform(0.0, 1.0)))
    mats.append(material((15.0, 15.0, 15.0), MAT_DIFFUSE_LIGHT, 0.0, 0.0, 0.0))
    M = [n_unique if m < 0 else m % n_unique for m in M]
    s.add_mesh(np.array(P, np.float32), np.array(I, np.uint32), np.array(N, np.float32), np.array(M, np.uint32))
    s.materials = mats
    s.add_instance(0)
    assert light_mat is not None
    return s, n_unique

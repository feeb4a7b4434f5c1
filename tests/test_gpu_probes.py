"""Parity of the device probe grid (gk_probes_voxelize / gk_probes_bake and REF_V2 path
termination) with the CPU oracle."""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from gknextengine_amd.scenes import material
from oracle import pyoracle as po
from tests import scenes

pytestmark = pytest.mark.gpu

CORNELL_GRID = (20, 20, 0.0625, (-0.125, -0.125, -0.125))


def oracle_grid(g):
    return po.OracleProbes(*g, gk.CUBE_DTYPE, gk.VOXEL_DTYPE)


def bright_params(W=32, H=32):
    p = gk.make_frame_params(scenes.CORNELL_CAM["eye"], scenes.CORNELL_CAM["target"], W, H, fov=40.0)
    p.estimator = gk.ESTIMATOR_REF_V2
    p.has_sky = 1
    p.sky_color[:] = [300.0, 200.0, 100.0, 1.0]
    p.has_sun = 1
    p.sun_direction[:] = [0.3, 0.9, -0.31622776, 0.0]
    p.sun_color[:] = [800.0, 700.0, 600.0, 1.0]
    return p


def bright_cornell():
    s = gk.procgen_cornell()
    s.materials[3].diffuse[:] = [2000.0, 2000.0, 2000.0, 1.0]  # light bright enough to survive RGB10
    return s


def test_voxelize_matches_oracle_cornell_and_floor():
    for s, g in ((bright_cornell(), CORNELL_GRID),
                 (scenes.tri_scene(scenes.floor_tris(), extra_materials=[material((0.3, 0.3, 0.3))], mat=1),
                  (6, 40, 0.25, (0.1, -4.875, 0.2)))):
        G = oracle_grid(g)
        G.voxelize(po.OracleScene(s))
        with gk.Context(32, 32) as ctx:
            ctx.upload(s)
            ctx.probes_init(g[0], g[1], g[2], g[3])
            ctx.probes_voxelize()
            _, vox = ctx.probes_read()
        for f in gk.VOXEL_DTYPE.names:
            assert np.array_equal(vox[f], G.vox[f]), f


def test_bake_passes_match_oracle_cornell():
    s = bright_cornell()
    osc = po.OracleScene(s)
    G = oracle_grid(CORNELL_GRID)
    G.voxelize(osc)
    p = bright_params()
    with gk.Context(32, 32) as ctx:
        ctx.upload(s)
        ctx.frame(p)
        ctx.probes_init(*CORNELL_GRID[:3], CORNELL_GRID[3])
        ctx.probes_voxelize()
        for it in range(4):
            ctx.probes_bake()
            G.bake(osc, p, s.lights[0])
            cubes, vox = ctx.probes_read()
            assert np.array_equal(vox["age"], G.vox["age"])
            diff = np.zeros(len(cubes), bool)
            for f in gk.CUBE_FIELDS:
                diff |= cubes[f] != G.cubes[f]
            print(f"pass {it}: {int(diff.sum())} of {len(cubes)} probes differ", flush=True)
            assert not diff.any()
            # keep both sides on the same state so a difference cannot compound
            G.cubes[:] = cubes
        act = (vox["dist_gg_z01"] & 0xFF) < 8
        assert act.sum() > 1000
        assert (cubes["pos_y"][act] != 0).mean() > 0.2   # bounced light reached the probes


def test_bake_frame_slices_like_the_renderer():
    s = bright_cornell()
    with gk.Context(32, 32) as ctx:
        ctx.upload(s)
        ctx.frame(bright_params())
        ctx.probes_init(32, 16, 0.0625, (-0.5, -0.125, -0.5))
        ctx.probes_voxelize()
        n = 32 * 32 * 16
        per = (n // 64) // 8 * 64
        ctx.probes_bake_frame(11, 8)      # slice 11 % 8 = 3
        _, vox = ctx.probes_read()
        act = (vox["dist_gg_z01"] & 0xFF) < 8
        age = vox["age"]
        sl = np.zeros(n, bool)
        sl[3 * per:4 * per] = True
        assert np.array_equal(age, (sl & act).astype(np.uint32))


@pytest.mark.parametrize("fast_gather", [False, True])
def test_refv2_with_probes_matches_oracle(fast_gather):
    """REF_V2 frames whose terminated paths read a baked grid (the device's grid, read back
    for the oracle); with GK_FRAME_FAST_GATHER every non-dielectric sample ends at its first
    bounce's probe lookup (ExitAfterFirst)."""
    W = H = 64
    s = bright_cornell()
    osc = po.OracleScene(s)
    G = oracle_grid(CORNELL_GRID)
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        p = bright_params(W, H)
        p.bounces = p.max_bounces = 4
        p.samples = 1
        p.has_sun = 0
        ctx.frame(p)
        ctx.probes_init(*CORNELL_GRID[:3], CORNELL_GRID[3])
        ctx.probes_voxelize()
        for _ in range(3):
            ctx.probes_bake()
        G.cubes[:], G.vox[:] = ctx.probes_read()
        G.attach(osc)
        moment = np.zeros((H, W), np.uint16)
        exact, acc_g, acc_c = [], 0.0, 0.0
        for f in range(4):
            q = bright_params(W, H)
            q.total_frames = f
            q.bounces = q.max_bounces = 4
            q.samples = 1
            q.has_sun = 0
            if fast_gather:
                q.flags |= gk.FRAME_FAST_GATHER
            ctx.frame(q)
            ctx.render()
            ctx.sync()
            o = osc.render_refv2(q, W, H, moment=moment)
            gd = ctx.read("SINGLE_DIFFUSE")
            od = o["diffuse"].astype(np.float16)
            exact.append(np.all(gd.view(np.uint16) == od.view(np.uint16), axis=2).mean())
            acc_g = acc_g + gd[..., :3].astype(np.float64)
            acc_c = acc_c + o["diffuse"][..., :3]
        po.oracle().oracle_scene_set_probes(osc.h, None, None, None)
    assert min(exact) > 0.98, exact
    err = float(np.mean((acc_g - acc_c) ** 2 / (acc_c ** 2 + 1e-2)))
    assert err < 1e-4, err
    assert acc_c.max() > 1.0


def test_reference_grid_on_the_city_matches_oracle_samples():
    """The reference's 192 x 48 x 192 grid over the procedural city: the whole grid is
    voxelized and baked on the device; sampled ranges are checked against the oracle."""
    s = gk.procgen_city()
    osc = po.OracleScene(s)
    p = bright_params()
    with gk.Context(32, 32) as ctx:
        ctx.upload(s)
        ctx.frame(p)
        ctx.probes_init()
        ctx.probes_voxelize()
        ctx.probes_bake()
        cubes, vox = ctx.probes_read()
    n = 192 * 48 * 192
    assert len(vox) == n
    act = (vox["dist_gg_z01"] & 0xFF) < 8
    print(f"active probes {int(act.sum())} of {n}", flush=True)
    assert act.sum() > 10000
    G = oracle_grid(gk.REFERENCE_PROBE_GRID)
    rng = np.random.default_rng(5)
    for first in rng.integers(0, n - 4096, 6):
        po.oracle().oracle_probe_voxelize(osc.h, C_addr(G.g), G.vox.ctypes.data, int(first), 4096, 16)
        sl = slice(int(first), int(first) + 4096)
        for f in gk.VOXEL_DTYPE.names:
            if f != "age":
                assert np.array_equal(vox[f][sl], G.vox[f][sl]), (first, f)
    # one bake pass of the oracle over a slice from the voxelized grid (ages 0, cubes 0)
    G.vox[:] = vox
    G.vox["age"] = 0
    first = int(np.nonzero(act)[0][len(np.nonzero(act)[0]) // 2]) // 64 * 64
    G.bake(osc, p, None, first=first, count=4096)
    sl = slice(first, first + 4096)
    for f in gk.CUBE_FIELDS:
        assert np.array_equal(cubes[f][sl], G.cubes[f][sl]), f


def C_addr(struct):
    import ctypes
    return ctypes.addressof(struct)

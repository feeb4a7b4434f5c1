"""Pins the oracle's ambient-cube probe grid (voxelize, bake, lookup, REF_V2 termination)
with closed forms on hand-built scenes and grids. CPU only.

The reference has no tests or stored outputs for its probe grid, so every check here is
derived from the geometry or from the definitions of the packing: a swapped face, byte or
axis, a wrong weight or a missed quirk fails one of them."""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from gknextengine_amd.scenes import material
from oracle import pyoracle as po
from tests import scenes

f32 = np.float32


def grid(size_xy, size_z, unit, offset):
    return po.OracleProbes(size_xy, size_z, unit, offset, gk.CUBE_DTYPE, gk.VOXEL_DTYPE)


def pack_ref(r, g, b):
    """RGB10 word of integer channel steps."""
    return int(r) | (int(g) << 10) | (int(b) << 20)


def bytes_of(w):
    return [(int(w) >> (8 * k)) & 0xFF for k in range(4)]


# ---------------------------------------------------------------- packing
def test_pack_unpack_known_answers():
    assert po.probe_pack([0, 0, 0, 0]) == 0
    assert po.probe_pack([512, 512, 512, 512]) == 0x3FF | (0x3FF << 10) | (0x3FF << 20) | (3 << 30)
    assert po.probe_pack([1e9, -5, 0, 0]) == 0x3FF                      # clamped
    assert po.probe_pack([np.nan, 0, 0, 0]) == 0                        # max(NaN, 0) = 0
    assert po.probe_pack([256, 128, 64, 0]) == pack_ref(511, 255, 127)  # 511.5 -> 511 (truncation)
    assert np.array_equal(po.probe_unpack(0x3FF | (0x200 << 10)), f32([512, 512 * 512 / 1023, 0, 0]))
    # unpack -> pack loses at most one step (the reference's truncation)
    for r in range(0, 1024, 7):
        v = po.probe_unpack(r)
        assert r - po.probe_pack(v) in (0, 1)


# ---------------------------------------------------------------- lookup
def free_voxels(G):
    G.vox["dist_gg_z01"] = 20 | (255 << 8) | (255 << 16) | (255 << 24)
    G.vox["dist_x01_y01"] = 0xFFFFFFFF


def probe_xyz(G, sxy):
    i = np.arange(G.n)
    return i % sxy, i // (sxy * sxy), (i // sxy) % sxy


def test_lookup_face_mapping_and_normalisation():
    G = grid(4, 4, 1.0, (0.0, 0.0, 0.0))
    free_voxels(G)
    faces = ["pos_x", "neg_x", "pos_y", "neg_y", "pos_z", "neg_z"]
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    for k, f in enumerate(faces):
        G.cubes[f + "_d"] = pack_ref(100 + 10 * k, 0, 0)
        G.cubes[f] = pack_ref(0, 300 + 10 * k, 0)
    G.cubes["sky_vis_pznzpyny"] = 255 * 0x01010101
    G.cubes["sky_vis_pxnxs0s1"] = 255 | (255 << 8)
    rng = np.random.default_rng(1)
    pos = rng.uniform(0.0, 2.0, (6, 3))
    for k, d in enumerate(dirs):
        di = G.lookup(pos, np.tile(d, (6, 1)), full=False)
        fu = G.lookup(pos, np.tile(d, (6, 1)), full=True)
        np.testing.assert_allclose(di[:, 0], (100 + 10 * k) * 512 / 1023, rtol=1e-6)
        assert np.all(di[:, 1:] == 0)
        np.testing.assert_allclose(fu[:, 1], (300 + 10 * k) * 512 / 1023, rtol=1e-6)
        np.testing.assert_allclose(fu[:, 3], 1.0, rtol=1e-6)        # sky visibility
    # a normal between +x and +y weights those faces equally
    n = np.tile(np.array([1, 1, 0]) / np.sqrt(2), (6, 1))
    di = G.lookup(pos, n, full=False)
    np.testing.assert_allclose(di[:, 0], (100 + 120) / 2 * 512 / 1023, rtol=1e-5)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_lookup_is_trilinear_on_a_linear_field(axis):
    sxy, sz = 6, 5
    G = grid(sxy, sz, 0.5, (-1.0, 2.0, 3.0))
    free_voxels(G)
    c = probe_xyz(G, sxy)[axis]
    G.cubes["pos_y_d"] = [pack_ref(40 * v, 0, 0) for v in c]
    rng = np.random.default_rng(axis)
    lim = [sxy - 2, sz - 2, sxy - 2]
    u = rng.uniform(0, 1, (32, 3)) * lim
    pos = u * 0.5 + np.array([-1.0, 2.0, 3.0])
    out = G.lookup(pos, np.tile([0, 1, 0], (32, 1)), full=False)
    np.testing.assert_allclose(out[:, 0], 40 * u[:, axis] * 512 / 1023, rtol=1e-5, atol=1e-4)


def test_lookup_outside_inside_and_occlusion():
    G = grid(4, 4, 1.0, (0.0, 0.0, 0.0))
    free_voxels(G)
    G.cubes["pos_y_d"] = pack_ref(200, 0, 0)
    up = [[0, 1, 0]]
    # outside the grid: (0,0,0,1); nearpos may reach size-2
    for p in ([-0.01, 1, 1], [1, 1, 2.01], [1, 2.01, 1]):
        assert np.array_equal(G.lookup([p], up), f32([[0, 0, 0, 1]]))
    assert G.lookup([[2.0, 2.0, 2.0]], up)[0, 0] > 0
    # every probe inside geometry: nothing to blend
    G2 = grid(4, 4, 1.0, (0.0, 0.0, 0.0))
    G2.cubes["pos_y_d"] = pack_ref(200, 0, 0)
    assert np.array_equal(G2.lookup([[0.5, 0.5, 0.5]], up), f32([[0, 0, 0, 0]]))
    # corner (0,0,0) holds another value and has no free distance: it is skipped unless
    # the point lies within 0.05 of it
    G.cubes["pos_y_d"][0] = pack_ref(900, 0, 0)
    G.vox["dist_x01_y01"][0] = 0
    G.vox["dist_gg_z01"][0] = 20 | (255 << 8)
    far = G.lookup([[0.3, 0.2, 0.4]], up)
    np.testing.assert_allclose(far[0, 0], 200 * 512 / 1023, rtol=1e-6)
    near = G.lookup([[0.02, 0.02, 0.02]], up)
    assert near[0, 0] > 205 * 512 / 1023
    # free only towards +x: a point along +x is blended in, one along +y is not
    G.vox["dist_x01_y01"][0] = 255
    along_x = G.lookup([[0.5, 0.0, 0.0]], up)[0, 0]
    along_y = G.lookup([[0.0, 0.5, 0.0]], up)[0, 0]
    np.testing.assert_allclose(along_x, (0.5 * 900 + 0.5 * 200) * 512 / 1023, rtol=1e-6)
    np.testing.assert_allclose(along_y, 200 * 512 / 1023, rtol=1e-6)


# ---------------------------------------------------------------- voxelize
def test_voxelize_above_and_below_a_floor():
    s = scenes.tri_scene(scenes.floor_tris(), extra_materials=[material((0.3, 0.3, 0.3))], mat=1)
    osc = po.OracleScene(s)
    G = grid(4, 40, 0.25, (0.1, -4.875, 0.2))  # probe heights: odd multiples of 1/8
    G.voxelize(osc)
    _, yy, _ = probe_xyz(G, 4)
    y = yy * 0.25 - 4.875
    for i in range(G.n):
        b0, b1 = bytes_of(G.vox["dist_gg_z01"][i]), bytes_of(G.vox["dist_x01_y01"][i])
        h = abs(y[i])
        assert b0[2:] == [255, 255] and b1[:2] == [255, 255]   # z and x rays miss
        assert G.vox["age"][i] == 0
        if y[i] > 0:   # -y ray hits the top of the floor
            d = min(h / 0.25, 1.0)
            assert b1[2] == 255 and b1[3] == int(d * 255)
            assert b0[0] == int(h / 0.25) and b0[1] == int(d * 255)
        else:          # +y ray hits its back: solid within one unit
            if h <= 0.25:
                assert b1[2] == 0 and b0[1] == 0 and b0[0] == 0
            else:
                assert b1[2] == 255 and b0[1] == 255 and b0[0] == int(h / 0.25)
            assert b1[3] == 255
        assert G.vox["mat_id"][i] == (1 if h <= 0.25 else 0)   # material of a hit within a unit


def test_voxelize_far_from_geometry_wraps_the_distance_byte():
    # nothing within 64 units: min distance 255 m -> 1020 units -> byte 252 (& 0xFF)
    s = scenes.tri_scene(scenes.small_tri((500, 500, 500), (1, 0, 0)))
    G = grid(2, 2, 0.25, (0.0, 0.0, 0.0))
    G.voxelize(po.OracleScene(s))
    assert np.all(G.vox["dist_gg_z01"] == (252 | (255 << 8) | (255 << 16) | (255 << 24)))
    assert np.all(G.vox["dist_x01_y01"] == 0xFFFFFFFF)


@pytest.mark.parametrize("diag,seen", [((1, 1, 1), True), ((-1, -1, -1), True), ((1, -1, 1), False)])
def test_voxelize_diagonal_probes(diag, seen):
    # DetectDistance's list repeats (-1,1,1) and (-1,1,-1) and never casts (1,-1,1)
    o = np.array([3.0, 3.0, 3.0])
    d = np.array(diag, np.float64) / np.sqrt(3.0)
    s = scenes.tri_scene(scenes.small_tri(o + 2.1 * d, d))
    G = grid(2, 2, 0.25, tuple(o))
    G.voxelize(po.OracleScene(s))
    b0 = bytes_of(G.vox["dist_gg_z01"][0])
    assert b0[0] == (8 if seen else 252)
    assert b0[1:] == [255, 255, 255]


# ---------------------------------------------------------------- bake
def bake_params(sky=(512, 256, 128, 0), sun=None):
    p = gk.FrameParams()
    p.has_sky = 1
    p.sky_color[:] = list(sky)
    if sun is not None:
        p.has_sun = 1
        p.sun_direction[:] = [0.0, 1.0, 0.0, 0.0]
        p.sun_color[:] = list(sun)
    return p


def test_bake_open_sky_face_closed_forms():
    """Probes 0.125-1.875 m over an infinite floor: the +y face sees only sky and sun, so
    one pass gives lerp(0, sky + sun/4, 1/8); the quad light adds ndotl area / (d^2 pi)."""
    light_mat = material((40.0, 40.0, 40.0), model=gk.MAT_DIFFUSE_LIGHT)
    s = scenes.tri_scene(scenes.floor_tris(), extra_materials=[light_mat])
    osc = po.OracleScene(s)
    G = grid(2, 12, 0.25, (0.0, 0.125, 0.0))
    G.voxelize(osc)
    _, yy, _ = probe_xyz(G, 2)
    y = 0.125 + 0.25 * yy
    active = y < 2.0                       # min distance below 8 units
    assert np.array_equal((G.vox["dist_gg_z01"] & 0xFF) < 8, active)
    light = scenes.quad_light((0.0, 6.0, 0.0), 0.5, 1.0, 1)
    p = bake_params(sun=(1024, 0, 0, 0))
    prev = np.zeros((G.n, 3))
    for it in range(2):
        G.bake(osc, p, light)
        i = np.arange(G.n)
        x = (i % 2) * 0.25
        z = ((i // 2) % 2) * 0.25
        o_y = y + 0.25 * 0.25                       # face origin moved a quarter unit up
        dx, dy, dz = 0.0 - x, 6.0 - o_y, 0.0 - z
        dist = np.sqrt(dx * dx + dy * dy + dz * dz)
        ndotl = dy / dist
        att = ndotl * 1.0 / (dist * dist * 3.14159)
        direct = np.stack([512 + 256 + 40 * att, 256 + 40 * att, 128 + 40 * att], 1)
        new = prev + (direct - prev) * 0.125
        steps = np.floor(np.minimum(new / 512.0, 1.0) * 1023.0)
        frac = np.minimum(new / 512.0, 1.0) * 1023.0 - steps
        assert np.all((frac[active] > 1e-3) & (frac[active] < 1 - 1e-3)), "closed form too close to a step"
        got = G.cubes["pos_y_d"]
        for k in range(3):
            ch = (got >> (10 * k)) & 0x3FF
            assert np.array_equal(ch[active], steps[active, k].astype(np.int64)), (it, k)
        assert np.all(got[~active] == 0)
        prev = np.stack([((got >> (10 * k)) & 0x3FF) * 512.0 / 1023.0 for k in range(3)], 1)
        # +y sky byte: lerp(old, 255 * 16/16, 1/8); sun bit set; age counts passes
        sky = bytes_of_arr(G.cubes["sky_vis_pznzpyny"], 2)
        expect_sky = 31 if it == 0 else int(31 + (255 - 31) * 0.125)
        assert np.all(sky[active] == expect_sky)
        assert np.all(bytes_of_arr(G.cubes["sky_vis_pxnxs0s1"], 2)[active] == 1)
        assert np.all(G.vox["age"][active] == it + 1) and np.all(G.vox["age"][~active] == 0)
        # the -y face sees only floor (bounce off an empty grid): no sky, no light
        assert np.all(G.cubes["neg_y_d"][active] == 0)
        assert np.all(bytes_of_arr(G.cubes["sky_vis_pznzpyny"], 3) == 0)


def bytes_of_arr(w, k):
    return (np.asarray(w, np.uint32) >> (8 * k)) & 0xFF


def test_bake_range_touches_only_its_probes():
    s = scenes.tri_scene(scenes.floor_tris())
    osc = po.OracleScene(s)
    G = grid(4, 4, 0.25, (0.1, 0.125, 0.2))
    G.voxelize(osc)
    G.bake(osc, bake_params(), None, first=16, count=16)
    ages = G.vox["age"]
    assert np.all(ages[16:32] == 1) and np.all(ages[:16] == 0) and np.all(ages[32:] == 0)
    assert np.all(G.cubes["pos_y_d"][:16] == 0) and np.all(G.cubes["pos_y_d"][16:32] != 0)


# ---------------------------------------------------------------- REF_V2 termination
def test_refv2_termination_is_linear_in_the_probe_value():
    """A terminated path is multiplied by the lookup: with a uniform grid whose faces hold
    u, the lookup is 2u (direct + bounced), so frame(u) - frame(none) scales with u."""
    W = H = 24
    s = gk.procgen_cornell()
    osc = po.OracleScene(s)
    p = gk.make_frame_params(scenes.CORNELL_CAM["eye"], scenes.CORNELL_CAM["target"], W, H, fov=40.0)
    p.estimator = gk.ESTIMATOR_REF_V2
    p.bounces = p.max_bounces = 3
    p.samples = 2
    p.exit_probability = 0.5
    base = osc.render_refv2(p, W, H)["diffuse"][..., :3].astype(np.float64)
    outs = []
    for r in (64, 640):
        G = grid(20, 20, 0.0625, (-0.125, -0.125, -0.125))
        free_voxels(G)
        for f in gk.CUBE_FIELDS[:12]:
            G.cubes[f] = pack_ref(r, r, r)
        G.attach(osc)
        outs.append(osc.render_refv2(p, W, H)["diffuse"][..., :3].astype(np.float64) - base)
        # a grid that does not reach the box leaves every frame as without probes
        G2 = grid(4, 4, 0.0625, (50.0, 50.0, 50.0))
        G2.attach(osc)
        assert np.array_equal(osc.render_refv2(p, W, H)["diffuse"][..., :3], base.astype(np.float32))
        G.attach(osc)
        del G2
    po.oracle().oracle_scene_set_probes(osc.h, None, None, None)
    mask = outs[0] > 1e-3
    assert mask.mean() > 0.3
    np.testing.assert_allclose(outs[1][mask] / outs[0][mask], 10.0, rtol=1e-4)


def test_refv2_fast_gather_ends_every_sample_at_its_first_bounce():
    """GK_FRAME_FAST_GATHER (the renderer's ExitAfterFirst, Shading.slang:1118-1124): every
    sample of a non-emitter, non-dielectric primary traces exactly one bounce ray, so a
    frame traces one primary ray per pixel plus one ray per sample of such pixels, with any
    seeds; the first bounce's throughput is multiplied by the probe lookup, so with the sky
    off and a uniform grid the frame is linear in the grid's value, and without a grid
    only samples whose first bounce reaches the light keep any radiance."""
    W = H = 24
    s = gk.procgen_cornell()
    osc = po.OracleScene(s)
    p = gk.make_frame_params(scenes.CORNELL_CAM["eye"], scenes.CORNELL_CAM["target"], W, H, fov=40.0)
    p.estimator = gk.ESTIMATOR_REF_V2
    p.bounces = p.max_bounces = 4
    p.samples = 3
    p.has_sky = p.has_sun = 0
    p.flags |= gk.FRAME_FAST_GATHER
    light = next(i for i, m in enumerate(s.materials) if m.material_model == gk.MAT_DIFFUSE_LIGHT)
    for frame in range(3):
        p.total_frames = frame
        o = osc.render_refv2(p, W, H)
        hit = o["objid"] != 65535
        on_light = hit & np.all(o["albedo"][..., :3] == np.array(s.materials[light].diffuse[:3], np.float32), axis=-1)
        assert hit.mean() > 0.5 and on_light.any()  # misses trace their primary ray only
        assert int(o["rays"]) == W * H + p.samples * int((hit & ~on_light).sum())
    q = gk.FrameParams.from_buffer_copy(p)
    q.flags &= ~gk.FRAME_FAST_GATHER
    assert int(osc.render_refv2(q, W, H)["rays"]) > int(osc.render_refv2(p, W, H)["rays"])
    base = osc.render_refv2(p, W, H)["diffuse"][..., :3].astype(np.float64)
    outs = []
    for r in (64, 640):
        G = grid(20, 20, 0.0625, (-0.125, -0.125, -0.125))
        free_voxels(G)
        for f in gk.CUBE_FIELDS[:12]:
            G.cubes[f] = pack_ref(r, r, r)
        G.attach(osc)
        outs.append(osc.render_refv2(p, W, H)["diffuse"][..., :3].astype(np.float64) - base)
    po.oracle().oracle_scene_set_probes(osc.h, None, None, None)
    mask = outs[0] > 1e-3
    assert mask.mean() > 0.5
    np.testing.assert_allclose(outs[1][mask] / outs[0][mask], 10.0, rtol=1e-4)

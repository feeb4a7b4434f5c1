"""PT_MIS on the GPU (wavefront.hip k_pt_shade) against the oracle on glossy materials.

Same per-pixel pcg4d seeds on both sides, so a path draws the same random numbers: its
lobe choice (GGX VNDF or cosine lobe), the MIS weights of both strategies, next-event
estimation beside a specular lobe, dielectric events and Russian roulette are compared
path by path. Float paths can part where a device and a host result differ in the last
bit near a branch; the tests bound the share of such pixels and the image error.

At the benchmark's own size (the 1080p procedural city, 8 bounces) the frame is checked
against the oracle on crops spread over the image, and through properties that hold at
any size: radiance is linear in the emitters' radiance (bit for bit), and the
material-sorted shading order gives the same frame.
"""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po
from tests import scenes

pytestmark = pytest.mark.gpu


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 1e-2)))


def params(W, H, cam, frame=0, bounces=8, sky=None, seed=0):
    p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"], frame=frame)
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.bounces = p.max_bounces = bounces
    p.samples = 1
    p.global_seed = seed
    p.denoiser = gk.DENOISER_NONE
    if sky is not None:
        p.has_sky = 1
        p.sky_color[:] = [*sky, 1.0]
    return p


def agreement(gpu, cpu, rtol=1e-4):
    close = np.all(np.abs(gpu - cpu) <= rtol * (1 + np.abs(cpu)), axis=-1)
    return float(close.mean())


@pytest.mark.parametrize("which", ["glossy", "all_models"])
def test_pt_mis_glossy_cornell_matches_oracle_same_seeds(which):
    """8 frames x 1 spp at 128x128, 8 bounces: GGX Mixture and Metallic lobes, MIS next-event
    estimation towards two emitters (one of them a large ceiling), dielectric refraction
    and Russian roulette after the third bounce."""
    W = H = 128
    s = scenes.glossy_cornell() if which == "glossy" else scenes.all_models_cornell()
    osc = po.OracleScene(s)
    frames = 8
    cam = scenes.CORNELL_CAM
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        for f in range(frames):
            ctx.frame(params(W, H, cam, frame=f))
            ctx.render()
            if f == 0:
                ctx.sync()
                spec0 = ctx.read("SINGLE_SPECULAR")[..., :3].astype(np.float32)
        ctx.sync()
        rad = ctx.read("RADIANCE")
    gpu = rad[..., :3] / rad[..., 3:4]
    o = osc.render_ptmis(params(W, H, cam), W, H, frame0=0, frames=frames)
    cpu = o["rgb"]
    agree = agreement(gpu, cpu)
    err = relmse(gpu, cpu)
    print(f"{which}: pixels within 1e-4: {agree:.4f}, relMSE {err:.2e}")
    assert agree > 0.95, agree
    assert err < 2e-3, err
    # the primary-lobe split: frame 0's specular target (fp16) against the oracle's
    o0 = osc.render_ptmis(params(W, H, cam), W, H, frame0=0, frames=1)
    same = np.all(spec0 == o0["spec"].astype(np.float16).astype(np.float32), axis=-1) | \
        np.all(np.abs(spec0 - o0["spec"]) <= 2e-3 * (1 + np.abs(o0["spec"])), axis=-1)
    assert same.mean() > 0.95, same.mean()
    assert o0["spec"].max() > 0.1  # the specular lobe carries light in these scenes


@pytest.fixture(scope="module")
def city():
    s = gk.procgen_city(64, 4, 256, 64, 2024)
    return s, po.OracleScene(s)


CROPS = [(128, 96), (512, 400), (900, 500), (1400, 200), (1700, 900), (300, 800), (1100, 1000), (960, 40)]


@pytest.mark.parametrize("sky", [None, (0.6, 0.7, 0.9)])
def test_pt_mis_city_1080p_8_bounces_matches_oracle_on_crops(city, sky):
    """The benchmark frame (the procedural city: 0.92 M triangles, GGX materials with roughness
    U(0.05, 1) and metallic Bernoulli(0.3), 64 emissive roofs, 1920x1080, 8 bounces) and
    the same frame under a sky, against the oracle on eight 64x64 crops."""
    s, osc = city
    W, H = 1920, 1080
    p = params(W, H, scenes.CITY_CAM, frame=5, bounces=8, sky=sky, seed=11)
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        rad = ctx.read("RADIANCE")
    gpu = rad[..., :3] / rad[..., 3:4]
    assert np.isfinite(gpu).all() and (gpu >= 0).all()
    agrees, gs, cs = [], [], []
    for x0, y0 in CROPS:
        o = osc.render_ptmis(p, W, H, rect=(x0, y0, 64, 64), frame0=p.total_frames, frames=1)
        g = gpu[y0:y0 + 64, x0:x0 + 64]
        agrees.append(agreement(g, o["rgb"]))
        gs.append(g)
        cs.append(o["rgb"])
    g, c = np.concatenate(gs), np.concatenate(cs)
    err = relmse(g, c)
    print(f"city sky={sky}: per-crop agreement {np.round(agrees, 4).tolist()}, relMSE {err:.2e}, "
          f"mean {g.mean():.4f} vs {c.mean():.4f}")
    assert c.max() > 0.5  # the crops hold light
    assert min(agrees) > 0.9, agrees
    assert np.mean(agrees) > 0.95, agrees
    assert err < 5e-3, err


def test_pt_mis_city_1080p_radiance_is_linear_in_emission(city):
    """Doubling every emitter's radiance (gk_materials_update, which also refreshes the
    light list next-event estimation samples) doubles the 1080p, 8-bounce frame bit for
    bit: throughputs, Russian roulette, light selection (normalised powers) and MIS
    weights do not depend on the radiance scale, and x2 is exact in floating point."""
    s, _ = city
    W, H = 1920, 1080
    p = params(W, H, scenes.CITY_CAM, frame=3, bounces=8, seed=2)
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        a = ctx.read("RADIANCE").copy()
        mats = [gk.Material.from_buffer_copy(m) for m in s.materials]
        for m in mats:
            if m.material_model == gk.MAT_DIFFUSE_LIGHT:
                for k in range(3):
                    m.diffuse[k] *= 2.0
        ctx.update_materials(mats)
        ctx.reset()
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        b = ctx.read("RADIANCE")
    assert a[..., :3].max() > 1.0
    assert np.array_equal(b[..., :3], 2.0 * a[..., :3])
    assert np.array_equal(b[..., 3], a[..., 3])


def test_materials_update_switches_a_light_off_for_next_event_estimation():
    """A material update that turns the Cornell box's only emitter into a Lambertian
    surface leaves no light in the scene (no sky): the frame is black, so next-event
    estimation no longer samples the old emitter."""
    W = H = 64
    s = gk.procgen_cornell()
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        p = params(W, H, scenes.CORNELL_CAM, bounces=4)
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        assert ctx.read("RADIANCE")[..., :3].mean() > 0.05
        mats = [gk.Material.from_buffer_copy(m) for m in s.materials]
        for m in mats:
            if m.material_model == gk.MAT_DIFFUSE_LIGHT:
                m.material_model = gk.MAT_LAMBERTIAN
                m.diffuse[:] = [0.5, 0.5, 0.5, 1.0]
        ctx.update_materials(mats)
        ctx.reset()
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        rad = ctx.read("RADIANCE")
    assert np.isfinite(rad).all()
    assert np.all(rad[..., :3] == 0.0)


def test_pt_mis_city_1080p_sorted_shading_gives_the_same_frame(city):
    """Material-sorted shading reorders a bounce's queue; every path owns its state, so
    the 1080p frame with materials over four models is bit-identical to path order."""
    s, _ = city
    rng = np.random.default_rng(5)
    models = [gk.MAT_LAMBERTIAN, gk.MAT_METALLIC, gk.MAT_DIELECTRIC, gk.MAT_MIXTURE]
    mixed = gk.SceneData()
    mixed.meshes, mixed.instances, mixed.lights = s.meshes, s.instances, s.lights
    mixed.textures = getattr(s, "textures", [])
    mixed.materials = [gk.Material.from_buffer_copy(m) for m in s.materials]
    for m in mixed.materials:
        if m.material_model != gk.MAT_DIFFUSE_LIGHT:
            m.material_model = int(models[rng.integers(4)])
            m.refraction_index = m.refraction_index2 = 1.5
    W, H = 1920, 1080
    out = []
    for flags in (gk.FRAME_NO_MATERIAL_SORT, gk.FRAME_MATERIAL_SORT):
        with gk.Context(W, H) as ctx:
            ctx.upload(mixed)
            p = params(W, H, scenes.CITY_CAM, frame=1, bounces=8, sky=(0.6, 0.7, 0.9), seed=4)
            p.flags |= flags
            ctx.frame(p)
            ctx.render()
            ctx.sync()
            out.append(ctx.read("RADIANCE").copy())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("W,H", [(520, 520), (515, 515), (1920, 1080)])
def test_pt_mis_two_pipelines_give_the_single_pipeline_frame(city, W, H):
    """Frames of 2^18 paths or more run as two path pipelines on two streams (the even and
    the odd 8x8 tiles; halves of the path range when the size is not a multiple of 8). Three
    frames with temporal reuse and a moving camera, queued back to back: from the second on,
    the pipelines generate and trace their primary rays while the previous frame's
    reprojection and bilateral filter run. Every target bit-identical to
    GK_FRAME_SINGLE_PIPELINE's, with an odd tile count (65 x 65) too, and the same number of
    rays traced (at 1080p the second frame holds the shared-edge tie of
    test_closest_hit_does_not_depend_on_ray_order)."""
    s, _ = city
    names = [t for t in gk.TARGETS if t not in ("MINIGBUFFER", "RADIANCE64")]
    outs = []
    for flags in (gk.FRAME_SINGLE_PIPELINE, 0):
        with gk.Context(W, H) as ctx:
            ctx.upload(s)
            for f in range(3):
                eye = tuple(e + 0.15 * f * d for e, d in zip(scenes.CITY_CAM["eye"], (1.0, 0.0, 0.5)))
                cam = dict(scenes.CITY_CAM, eye=eye)
                p = params(W, H, cam, frame=2 + f, bounces=8, sky=(0.6, 0.7, 0.9), seed=7)
                p.denoiser = gk.DENOISER_BILATERAL
                p.flags |= flags
                ctx.frame(p)
                ctx.render()
            ctx.sync()
            outs.append(({t: ctx.read(t).copy() for t in names}, ctx.stats().rays_traced))
    (a, ra), (b, rb) = outs
    assert ra == rb and ra > W * H, (ra, rb)
    for t in names:
        assert np.array_equal(a[t].view(np.uint8), b[t].view(np.uint8)), t
    assert a["RADIANCE"][..., :3].max() > 0.1


def test_split_frames_queued_with_user_traces(city, device):
    """gk_trace uses the context's ray queues, which a split frame's ray generation also
    writes. Three split frames (520 x 520) queued back to back, each followed by a gk_trace
    of 2^16 random rays queued without a synchronisation: the frames' targets equal those of
    the same frames without the traces, and every trace returns the hits of its rays traced
    alone (a frame after another entry point waits for the context's stream before its
    ray generation)."""
    import torch
    s, _ = city
    W = H = 520
    n = 1 << 16
    rng = np.random.default_rng(11)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.uniform([-30.0, 1.0, -30.0], [30.0, 30.0, 30.0], (n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = 1e30
    r = torch.from_numpy(rays).to(device)
    names = [t for t in gk.TARGETS if t not in ("MINIGBUFFER", "RADIANCE64")]

    def frame_params(f):
        p = params(W, H, scenes.CITY_CAM, frame=f, bounces=8, sky=(0.6, 0.7, 0.9), seed=3)
        p.denoiser = gk.DENOISER_BILATERAL
        return p

    with gk.Context(W, H) as ctx:          # the rays alone
        ctx.upload(s)
        h0 = torch.zeros((n, 4), dtype=torch.float32, device=device)
        i0 = torch.zeros(n, dtype=torch.int32, device=device)
        torch.cuda.synchronize()
        ctx.trace(r.data_ptr(), n, h0.data_ptr(), i0.data_ptr())
        ctx.sync()
        alone = h0.cpu().numpy()
    outs = []
    for with_traces in (False, True):
        with gk.Context(W, H) as ctx:
            ctx.upload(s)
            hs = [torch.zeros((n, 4), dtype=torch.float32, device=device) for _ in range(3)]
            ins = [torch.zeros(n, dtype=torch.int32, device=device) for _ in range(3)]
            torch.cuda.synchronize()
            for f in range(3):
                ctx.frame(frame_params(f))
                ctx.render()
                if with_traces:
                    ctx.trace(r.data_ptr(), n, hs[f].data_ptr(), ins[f].data_ptr())
            ctx.sync()
            outs.append({t: ctx.read(t).copy() for t in names})
            if with_traces:
                for h in hs:
                    assert np.array_equal(h.cpu().numpy().view(np.uint32), alone.view(np.uint32))
    hits = int((alone[:, 0] < 1e30).sum())
    assert hits > n // 4, hits
    for t in names:
        assert np.array_equal(outs[0][t].view(np.uint8), outs[1][t].view(np.uint8)), t


def test_pt_mis_two_pipelines_on_moving_instances_and_visibility_buffer():
    """The split on a two-level scene (4096 brick instances, 640x480 = 307,200 paths): a
    frame, the instances moved and the TLAS rebuilt, a second frame, then a third whose
    primary hits come from the visibility buffer the second wrote; every target bit-identical
    to the single-pipeline frames."""
    from gknextengine_amd import scenes as gs
    b = scenes.Bricks()
    W, H = 640, 480
    names = [t for t in gk.TARGETS if t not in ("RADIANCE64",)]
    outs = []
    for flags in (gk.FRAME_SINGLE_PIPELINE, 0):
        with gk.Context(W, H) as ctx:
            ctx.upload(b.scene)
            for f in range(3):
                if f == 1:
                    ctx.update_instances(b.worlds(1.0 / 60.0))
                p = gk.make_frame_params(gs.BRICKS_CAM["eye"], gs.BRICKS_CAM["target"], W, H,
                                         fov=gs.BRICKS_CAM["fov"], frame=f)
                p.estimator = gk.ESTIMATOR_PT_MIS
                p.bounces = p.max_bounces = 4
                p.denoiser = gk.DENOISER_BILATERAL
                p.flags |= flags | gk.FRAME_TLAS_REBUILD
                if f == 2:
                    p.flags |= gk.FRAME_PRIMARY_FROM_VISBUFFER
                ctx.frame(p)
                ctx.render()
            ctx.sync()
            outs.append(({t: ctx.read(t).copy() for t in names}, ctx.stats().rays_traced))
    (a, ra), (c, rc) = outs
    assert ra == rc, (ra, rc)
    for t in names:
        assert np.array_equal(a[t].view(np.uint8), c[t].view(np.uint8)), t
    assert (a["MINIGBUFFER"] != 0).mean() > 0.3 and a["RADIANCE"][..., :3].max() > 0.1

"""Failure handling of the C ABI on the device: invalid scenes, transforms, ids and frame
constants come back as status codes and leave the context usable; material ids beyond the
table are clamped as the reference's GetMaterial does (BasicTypes.slang:376-379); an
allocation beyond the device-memory budget fails with GK_ERR_OOM and the context recovers."""
from __future__ import annotations

import os
import subprocess
import sys
import textwrap

import numpy as np
import pytest

import gknextengine_amd as gk
from oracle import pyoracle as po
from tests import scenes
from tests.test_oracle import _u32

pytestmark = pytest.mark.gpu


def cornell_frame(W, H, flags=0):
    c = scenes.CORNELL_CAM
    p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"])
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.bounces = p.max_bounces = 2
    p.flags |= flags
    return p


def status_of(fn):
    try:
        fn()
    except gk.GkError as e:
        return e.status
    return gk.GK_OK


def test_invalid_scenes_are_rejected_and_a_valid_upload_recovers():
    W = H = 32
    bad = []
    s = gk.procgen_cornell()
    s.instances[0].world[5] = float("nan")
    bad.append(("NaN world matrix", s))
    s = gk.procgen_cornell()
    s.instances[0].prev_world[12] = float("inf")
    bad.append(("Inf previous matrix", s))
    s = gk.procgen_cornell()
    s.meshes[0]["positions"][7, 1] = np.inf
    bad.append(("Inf vertex", s))
    s = gk.procgen_cornell()
    s.meshes[0]["positions"][3, 0] = np.nan
    bad.append(("NaN vertex", s))
    s = gk.procgen_cornell()
    s.meshes[0]["indices"][3, 2] = 10 ** 6
    bad.append(("vertex index out of range", s))
    s = gk.procgen_cornell()
    s.instances[0].mesh_id = 3
    bad.append(("mesh id out of range", s))
    s = gk.procgen_cornell()
    s.instances = []
    bad.append(("empty scene", s))
    with gk.Context(W, H) as ctx:
        for name, s in bad:
            assert status_of(lambda: ctx.upload(s)) == gk.GK_ERR_ARG, name
        # the first invalid upload that reached the device leaves no scene behind
        assert status_of(lambda: (ctx.frame(cornell_frame(W, H)), ctx.render())) == gk.GK_ERR_STATE
        good = gk.procgen_cornell()
        ctx.upload(good)
        ctx.frame(cornell_frame(W, H))
        ctx.render()
        ctx.sync()
        rad = ctx.read("RADIANCE")
        assert np.isfinite(rad).all() and rad[..., :3].max() > 0.1


def test_bad_updates_and_frames_are_rejected_without_side_effects():
    W = H = 48
    b = scenes.Bricks()
    osc = po.OracleScene(b.scene)
    with gk.Context(W, H) as ctx:
        ctx.upload(b.scene)
        p = cornell_frame(W, H)
        c = scenes.BRICKS_CAM
        p = gk.make_frame_params(c["eye"], c["target"], W, H, fov=c["fov"])
        p.estimator = gk.ESTIMATOR_PT_MIS
        p.bounces = p.max_bounces = 1
        w = b.worlds(0.5)
        ctx.update_instances(w)
        osc.update_instances(w)
        wn = w.copy()
        wn[17, 1, 3] = np.nan
        assert status_of(lambda: ctx.update_instances(wn)) == gk.GK_ERR_ARG
        assert status_of(lambda: ctx.update_instances(w[:5], first=b.n - 2)) == gk.GK_ERR_ARG
        mats = [gk.Material.from_buffer_copy(m) for m in b.scene.materials]
        assert status_of(lambda: ctx.update_materials(mats, first=1)) == gk.GK_ERR_ARG
        q = gk.FrameParams.from_buffer_copy(p)
        q.model_view_inverse[3] = float("inf")
        assert status_of(lambda: ctx.frame(q)) == gk.GK_ERR_ARG
        q = gk.FrameParams.from_buffer_copy(p)
        q.samples = 5000
        assert status_of(lambda: ctx.frame(q)) == gk.GK_ERR_LIMIT
        q = gk.FrameParams.from_buffer_copy(p)
        q.aperture = -1.0
        assert status_of(lambda: ctx.frame(q)) == gk.GK_ERR_ARG
        q.aperture, q.focus_distance = 0.1, 0.0
        assert status_of(lambda: ctx.frame(q)) == gk.GK_ERR_ARG
        q = gk.FrameParams.from_buffer_copy(p)
        q.estimator = 7
        assert status_of(lambda: ctx.frame(q)) == gk.GK_ERR_ARG
        small = np.zeros((4, 4), np.float32)
        assert gk.lib().gk_read_target(ctx.h, gk.RT["RADIANCE"], small.ctypes.data, small.nbytes) == gk.GK_ERR_ARG
        assert gk.lib().gk_read_target(ctx.h, 99, small.ctypes.data, small.nbytes) == gk.GK_ERR_ARG
        # none of the rejected calls changed the scene or the frame
        ctx.frame(p)
        ctx.render()
        ctx.sync()
        hit = ctx.read("PRIMARY_HIT").reshape(-1, 4)
        oh, _ = osc.trace(po.primary_rays(p, W, H))
        assert np.array_equal(_u32(hit), _u32(oh))
        ctx.stats()  # no traversal overflow


def test_material_ids_beyond_the_table_are_clamped_to_the_last_material():
    """GetMaterial clamps the index to the table (BasicTypes.slang:376-379): triangles whose
    local slot or instance offset points past the materials shade with the last one, exactly
    as a scene that names the last material."""
    W = H = 64
    s_far = gk.procgen_cornell()
    last = len(s_far.materials) - 1
    tm = s_far.meshes[0]["tri_material"]
    s_last = gk.procgen_cornell()
    s_last.meshes[0]["tri_material"][10:22] = last
    tm[10:22] = 200                      # the short box: a slot far past the table
    out = []
    for s in (s_far, s_last):
        with gk.Context(W, H) as ctx:
            ctx.upload(s)
            p = cornell_frame(W, H)
            ctx.frame(p)
            ctx.render()
            ctx.sync()
            out.append(ctx.read("RADIANCE").copy())
    assert np.isfinite(out[0]).all()
    assert np.array_equal(out[0], out[1])


def test_device_memory_budget_gives_oom_and_the_context_recovers():
    """GKNX_DEVICE_MEMORY_LIMIT bounds the library's device memory: with 160 MB a 0.92 M-triangle
    scene does not fit (GK_ERR_OOM; the failed upload releases what it had allocated), the
    context then takes a small scene and renders (the traversal's spill area alone is 67 MB), and a
    context whose render targets alone exceed the budget is refused with GK_ERR_OOM. Run in a
    child process (the budget is read once per process)."""
    code = textwrap.dedent("""
        import sys
        sys.path.insert(0, %r)
        import numpy as np
        import gknextengine_amd as gk
        from tests.test_gpu_errors import cornell_frame, status_of
        ctx = gk.Context(64, 64)
        st = status_of(lambda: ctx.upload(gk.procgen_city()))
        assert st == gk.GK_ERR_OOM, st
        ctx.upload(gk.procgen_cornell())
        ctx.frame(cornell_frame(64, 64))
        ctx.render()
        ctx.sync()
        assert ctx.read("RADIANCE")[..., :3].max() > 0.1
        ctx.close()
        st = status_of(lambda: gk.Context(4096, 4096))
        assert st == gk.GK_ERR_OOM, st
        print("budget ok")
    """) % (str(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),)
    env = dict(os.environ, GKNX_DEVICE_MEMORY_LIMIT=str(160 * 1024 * 1024))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0 and "budget ok" in r.stdout

"""The C ABI library: loads, exports every declared symbol, host helpers work without a
GPU, argument errors come back as status codes. CPU only."""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import gknextengine_amd as gk

HEADER = Path(__file__).resolve().parents[1] / "include" / "gknx_rt.h"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gk_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    L = C.CDLL(str(gk.LIB_PATH))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding covers the same set
    assert sorted(n for n, _, _ in gk.SIGNATURES) == names


def test_struct_sizes_match_header_comments():
    assert C.sizeof(gk.Material) == 64
    assert C.sizeof(gk.Light) == 80
    assert C.sizeof(gk.Instance) == 224
    assert C.sizeof(gk.FrameParams) == 656


def test_abi_version_and_status_strings():
    L = gk.lib()
    assert L.gk_get_abi_version() == 1
    for s in (gk.GK_OK, gk.GK_ERR_ARG, gk.GK_ERR_HIP, gk.GK_ERR_OOM, gk.GK_ERR_STATE, gk.GK_ERR_LIMIT,
              gk.GK_ERR_NO_DEVICE, gk.GK_ERR_IO):
        assert L.gk_status_string(s)


def test_null_and_invalid_arguments_return_status():
    L = gk.lib()
    assert L.gk_create(None, None) == gk.GK_ERR_ARG
    d = gk.DeviceDesc(0, 0, 0, 0)
    h = C.c_void_p()
    assert L.gk_create(C.byref(d), C.byref(h)) == gk.GK_ERR_ARG
    d = gk.DeviceDesc(0, 1 << 20, 8, 0)
    assert L.gk_create(C.byref(d), C.byref(h)) == gk.GK_ERR_LIMIT
    for fn in ("gk_destroy", "gk_render", "gk_sync", "gk_reset_accumulation"):
        assert getattr(L, fn)(None) == gk.GK_ERR_ARG
    assert L.gk_frame(None, None) == gk.GK_ERR_ARG
    assert L.gk_trace(None, None, 0, None, None, 0) == gk.GK_ERR_ARG
    assert L.gk_make_frame_params(None, 1, 1, 0, 0, None, None) == gk.GK_ERR_ARG


def test_no_device_is_reported_not_crashed():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    d = gk.DeviceDesc(0, 64, 64, 0)
    h = C.c_void_p()
    assert gk.lib().gk_create(C.byref(d), C.byref(h)) == gk.GK_ERR_NO_DEVICE


def test_frame_params_camera_conventions():
    W, H = 320, 200
    p = gk.make_frame_params((0, 0, 5), (0, 0, 0), W, H, fov=60.0)
    mv = np.array(p.model_view).reshape(4, 4).T     # column-major -> matrix
    pr = np.array(p.projection).reshape(4, 4).T
    mvi = np.array(p.model_view_inverse).reshape(4, 4).T
    pri = np.array(p.projection_inverse).reshape(4, 4).T
    assert np.allclose(mv @ mvi, np.eye(4), atol=1e-5)
    assert np.allclose(pr @ pri, np.eye(4), atol=1e-4)
    # Vulkan-style Y flip and [0,1] depth (GLM_FORCE_DEPTH_ZERO_TO_ONE, Engine.cpp:746-754)
    assert pr[1, 1] < 0 and pr[0, 0] > 0
    t = 1 / np.tan(np.deg2rad(30))
    assert pr[0, 0] == pytest.approx(t / (W / H), rel=1e-6)
    near = pr @ np.array([0, 0, -0.2, 1.0])
    far = pr @ np.array([0, 0, -2000.0, 1.0])
    assert near[2] / near[3] == pytest.approx(0.0, abs=1e-5)
    assert far[2] / far[3] == pytest.approx(1.0, abs=1e-5)
    # TAA: Halton(2,3) jitter of frame % 16 in the third column (Engine.cpp:771-778)
    pj = gk.make_frame_params((0, 0, 5), (0, 0, 0), W, H, fov=60.0, frame=1, taa=True)
    prj = np.array(pj.projection).reshape(4, 4).T
    assert prj[0, 2] == pytest.approx((0.25 - 0.5) / W * 2, rel=1e-5)
    assert prj[1, 2] == pytest.approx((2 / 3 - 0.5) / H * 2, rel=1e-3)


def test_procgen_counts():
    s = gk.procgen_cornell()
    assert s.meshes[0]["indices"].shape == (36, 3)
    assert len(s.materials) == 4
    assert s.materials[3].material_model == gk.MAT_DIFFUSE_LIGHT
    c = gk.procgen_city(grid=8, subdiv=2, ground_res=16, num_lights=5, seed=3)
    assert c.meshes[0]["indices"].shape[0] == 8 * 8 * 12 * 4 + 5 * 2 + 16 * 16 * 2
    assert len(c.materials) == 8 * 8 + 2
    m = c.meshes[0]
    assert m["indices"].max() < m["positions"].shape[0]
    assert np.allclose(np.linalg.norm(m["normals"], axis=1), 1.0)
    pos, nrm, idx, mat = gk.procgen_brick()
    assert idx.shape == (24, 3) and set(mat.tolist()) == {0, 1}
    # same seed -> same city
    c2 = gk.procgen_city(grid=8, subdiv=2, ground_res=16, num_lights=5, seed=3)
    assert np.array_equal(c2.meshes[0]["positions"], m["positions"])


def test_city_full_size_matches_config():
    c = gk.procgen_city()
    nt = c.meshes[0]["indices"].shape[0]
    assert 0.9e6 < nt < 0.95e6  # ~1M triangles (the procedural city)
    metal = np.array([m.metalness for m in c.materials[:4096]])
    rough = np.array([m.fuzziness for m in c.materials[:4096]])
    assert 0.25 < metal.mean() < 0.35 and set(np.unique(metal)) <= {0.0, 1.0}
    assert rough.min() >= 0.05 and rough.max() <= 1.0
    assert sum(m.material_model == gk.MAT_DIFFUSE_LIGHT for m in c.materials) == 1


def test_diff_render_rejects_host_and_misshapen_parameters():
    """diff.render checks params before any pointer reaches the library: a host tensor
    (the kernels read it as device memory) or a wrong shape raises ValueError."""
    import torch

    from gknextengine_amd import diff
    base = [gk.Material() for _ in range(3)]
    with pytest.raises(ValueError, match="CUDA"):
        diff.render(torch.zeros(3, 5), None, base, None)
    if torch.cuda.is_available():
        with pytest.raises(ValueError, match="shape"):
            diff.render(torch.zeros(4, 5, device="cuda"), None, base, None)

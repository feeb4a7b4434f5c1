"""The C++ host adapter (gknextengine_amd/host/gknx_renderer.hpp) drives the library with
no Python in the process: its demo renders the Cornell box and must produce the same FP32
radiance, bit for bit, as the same frames submitted through the Python binding."""
from __future__ import annotations

import json
import subprocess

import numpy as np
import pytest

import gknextengine_amd as gk
from gknextengine_amd import build, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("estimator", [gk.ESTIMATOR_REF_V2, gk.ESTIMATOR_PT_MIS])
def test_cpp_adapter_matches_python_binding(tmp_path, estimator):
    W, H, frames = 96, 64, 4
    out = tmp_path / "rad.f32"
    print("case: cpp demo", W, H, frames, estimator, flush=True)
    res = subprocess.run([str(build.DEMO), str(W), str(H), str(frames), str(estimator), str(out)],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    info = json.loads(res.stdout.strip().splitlines()[-1])
    assert info["rays"] > W * H * frames
    cpp = np.fromfile(out, np.float32).reshape(H, W, 4)
    s = gk.procgen_cornell()
    cam = scenes.CORNELL_CAM
    with gk.Context(W, H) as ctx:
        ctx.upload(s)
        prev = None
        for f in range(frames):
            p = gk.make_frame_params(cam["eye"], cam["target"], W, H, fov=cam["fov"], frame=f, prev=prev)
            p.estimator = estimator
            p.bounces = p.max_bounces = 4
            ctx.frame(p)
            ctx.render()
            prev = p
        ctx.sync()
        py = ctx.read("RADIANCE")
    assert np.array_equal(cpp, py)
    assert cpp[..., :3].mean() > 0 and np.all(cpp[..., 3] > 0)

"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): the C ABI's glTF
loader and procedural generators and the CPU oracle, built with -fsanitize=address,undefined
(tests/sanitize/Makefile) and driven on valid scenes, every pass of the oracle, and glTF files
truncated at many lengths or with corrupted bytes, which must come back as error codes. Any
out-of-bounds access, use after free, leak or undefined behaviour fails the run."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

import pytest

import gknextengine_amd as gk
from tests import scenes
from tests.test_gltf import make_scene

HERE = Path(__file__).resolve().parent / "sanitize"


@pytest.mark.skipif(shutil.which("g++") is None or shutil.which("make") is None, reason="needs g++ and make")
def test_host_code_under_asan_and_ubsan(tmp_path):
    r = subprocess.run(["make", "-C", str(HERE), "-j8"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    c = scenes.CORNELL_CAM
    p = gk.make_frame_params(c["eye"], c["target"], 24, 20, fov=c["fov"])
    p.has_sky = 1
    p.sky_color[:] = [0.5, 0.6, 0.7, 1.0]
    p.sun_direction[:] = [0.3, 0.9, -0.3, 0.0]
    p.sun_color[:] = [2.0, 2.0, 2.0, 1.0]
    (tmp_path / "frame.bin").write_bytes(bytes(p))
    B, _ = make_scene()
    (tmp_path / "scene.glb").write_bytes(B.glb())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([str(HERE / "_build" / "sanitize_driver"), str(tmp_path / "frame.bin"),
                        str(tmp_path / "scene.glb")], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0 and "sanitize ok" in r.stdout, r.stdout[-4000:] + r.stderr[-4000:]

"""The glTF loader (gk_gltf_load): a GLB written here with every feature the reference's
FSceneLoader reads, and the arrays the loader returns checked against what the file says
and the reference's rules (material mapping, texture order, node transforms, area
lights, cameras, section-indexed materials). CPU only."""
from __future__ import annotations

import base64
import json

import numpy as np
import pytest

import gknextengine_amd as gk
from tests.gltf_util import GltfBuilder, png


def quat(axis, deg):
    a = np.asarray(axis, np.float64)
    a /= np.linalg.norm(a)
    h = np.radians(deg) / 2
    return [*(a * np.sin(h)), np.cos(h)]  # x, y, z, w


def trs(t=(0, 0, 0), q=(0, 0, 0, 1), s=(1, 1, 1)):
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    M = np.eye(4)
    M[:3, :3] = R * np.asarray(s, np.float64)[None, :]
    M[:3, 3] = t
    return M


def make_scene():
    rng = np.random.default_rng(7)
    B = GltfBuilder()
    img0 = rng.integers(0, 256, (6, 8, 4)).astype(np.uint8)
    pal = rng.integers(0, 256, (16, 3)).astype(np.uint8)
    trns = np.array([10, 20, 30], np.uint8)
    idx1 = rng.integers(0, 16, (4, 5)).astype(np.uint8)
    img2 = rng.integers(0, 65536, (3, 4, 2)).astype(np.uint16)
    t0 = B.image(png(img0, 6, 8))
    t1 = B.image(png(idx1, 3, 4, palette=pal, trns=trns))
    t2 = B.image(png(img2, 4, 16))
    B.doc["materials"] = [
        {"pbrMetallicRoughness": {"baseColorFactor": [0.25, 0.64, 1.0, 1.0], "roughnessFactor": 0.4,
                                  "metallicFactor": 0.2, "baseColorTexture": {"index": t0}},
         "normalTexture": {"index": t2, "scale": 0.7}},
        {"pbrMetallicRoughness": {"metallicFactor": 0.99, "metallicRoughnessTexture": {"index": t1}}},
        {"emissiveFactor": [1.0, 0.5, 0.25], "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 3.0}}},
        {"pbrMetallicRoughness": {"metallicFactor": 0.0, "roughnessFactor": 0.1},
         "extensions": {"KHR_materials_transmission": {"transmissionFactor": 0.9}, "KHR_materials_ior": {"ior": 1.33}},
         "extras": {"ior2": 1.1}},
        {"pbrMetallicRoughness": {"metallicFactor": 1.0}},
        {"pbrMetallicRoughness": {"metallicFactor": 0.0, "roughnessFactor": 1.0}},
    ]
    # mesh 0: a quad in z = 0 with uv = (x, y), no tangents
    qp = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
    qn = np.tile([0, 0, 1.0], (4, 1))
    quad = {"primitives": [{"attributes": {"POSITION": B.accessor(qp), "NORMAL": B.accessor(qn),
                                           "TEXCOORD_0": B.accessor(qp[:, :2])},
                            "indices": B.accessor(np.array([0, 1, 2, 0, 2, 3]), 5123), "material": 0}]}
    # mesh 1: an indexed triangle with tangents (positions with byteStride 16), a LINES
    # primitive, then a non-indexed pair of triangles without normals or uvs
    tp = np.array([[0, 0, 0], [2, 0, 0], [0, 3, 0]], np.float32)
    tt = np.array([[1, 0, 0, -1], [1, 0, 0, -1], [1, 0, 0, -1]], np.float32)
    pp = np.array([[0, 0, 1], [1, 0, 1], [0, 1, 1], [0, 0, 2], [0, 1, 2], [1, 0, 2]], np.float32)
    parts = {"primitives": [
        {"attributes": {"POSITION": B.accessor(tp, stride=16), "NORMAL": B.accessor(np.tile([0, 0, 1.0], (3, 1))),
                        "TEXCOORD_0": B.accessor(tp[:, :2] * 0.5), "TANGENT": B.accessor(tt)},
         "indices": B.accessor(np.array([0, 1, 2]), 5125), "material": 1},
        {"attributes": {"POSITION": B.accessor(tp)}, "mode": 1, "material": 2},
        {"attributes": {"POSITION": B.accessor(pp)}, "material": 3}]}
    # mesh 2: two primitives (materials 4, 5)
    lp = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], np.float32)
    lmesh = {"primitives": [{"attributes": {"POSITION": B.accessor(lp)}, "material": 4},
                            {"attributes": {"POSITION": B.accessor(lp + 1)}, "material": 5}]}
    B.doc["meshes"] = [quad, parts, lmesh]
    B.doc["cameras"] = [{"type": "perspective", "perspective": {"yfov": 0.5, "znear": 0.1}}]
    q_root, q_parts = quat((0, 1, 0), 90), quat((1, 0, 0), 30)
    B.doc["nodes"] = [
        {"name": "root", "translation": [1, 2, 3], "rotation": q_root, "scale": [2, 2, 2], "children": [1, 2, 4]},
        {"name": "quad", "mesh": 0, "translation": [0, 0, -1]},
        {"name": "parts", "mesh": 1, "rotation": q_parts, "children": [3]},
        {"name": "light", "mesh": 2, "translation": [0, 5, 0], "scale": [0.5, 1, 0.5], "extras": {"arealight": 1}},
        {"name": "cam", "camera": 0, "translation": [0, 1, 10]},
    ]
    B.doc["scenes"] = [{"nodes": [0], "extras": {"SkyIntensity": 2.5, "WithSun": 1, "SunRotation": 0.3}}]
    B.doc["scene"] = 0
    expect = {"img0": img0, "pal": pal, "trns": trns, "idx1": idx1, "img2": img2, "q_root": q_root,
              "q_parts": q_parts, "qp": qp, "tp": tp, "tt": tt, "pp": pp}
    return B, expect


@pytest.fixture(scope="module")
def loaded():
    B, e = make_scene()
    g = gk.GltfScene(B.glb())
    return g, e


def test_materials_follow_the_reference_mapping(loaded):
    g, _ = loaded
    m = g.scene.materials
    assert len(m) == 6
    np.testing.assert_allclose(list(m[0].diffuse), [0.5, 0.8, 1.0, 1.0], rtol=1e-6)
    assert (m[0].material_model, m[0].diffuse_texture_id, m[0].normal_texture_id, m[0].mra_texture_id) == \
        (gk.MAT_MIXTURE, 0, 1, -1)
    assert m[0].fuzziness == np.float32(0.4) and m[0].metalness == np.float32(0.2)
    assert m[0].normal_texture_scale == np.float32(0.7) and m[0].refraction_index == np.float32(1.46)
    # a metallic-roughness texture forces Fuzziness 1 and blocks the Metallic rule
    assert (m[1].material_model, m[1].mra_texture_id, m[1].fuzziness) == (gk.MAT_MIXTURE, 2, 1.0)
    assert m[2].material_model == gk.MAT_DIFFUSE_LIGHT
    np.testing.assert_allclose(list(m[2].diffuse), [300.0, 150.0, 75.0, 1.0], rtol=1e-6)
    assert m[3].material_model == gk.MAT_DIELECTRIC
    assert m[3].refraction_index == np.float32(1.33) and m[3].refraction_index2 == np.float32(1.1)
    assert m[4].material_model == gk.MAT_METALLIC and m[4].fuzziness == 1.0
    assert m[5].material_model == gk.MAT_LAMBERTIAN


def test_textures_decode_in_reference_load_order(loaded):
    g, e = loaded
    t = g.scene.textures
    assert len(t) == 3
    assert [x["srgb"] for x in t] == [True, False, False]  # base colour, normal (mat 0), MR (mat 1)
    assert np.array_equal(t[0]["rgba"], e["img0"])
    hi = (e["img2"] >> 8).astype(np.uint8)
    want2 = np.stack([hi[..., 0], hi[..., 0], hi[..., 0], hi[..., 1]], -1)
    assert np.array_equal(t[1]["rgba"], want2)
    pal = e["pal"][e["idx1"]]
    alpha = np.where(e["idx1"] < 3, e["trns"][np.minimum(e["idx1"], 2)], 255).astype(np.uint8)
    assert np.array_equal(t[2]["rgba"], np.concatenate([pal, alpha[..., None]], -1))


def test_meshes_sections_and_generated_attributes(loaded):
    g, e = loaded
    ms = g.scene.meshes
    assert len(ms) == 3
    q = ms[0]
    assert np.array_equal(q["positions"], e["qp"]) and np.array_equal(q["texcoords"], e["qp"][:, :2])
    assert np.array_equal(q["indices"], [[0, 1, 2], [0, 2, 3]])
    # no TANGENT in the mesh: tangents from the uv derivatives, u along +x, handedness +1
    np.testing.assert_allclose(q["tangents"], np.tile([1, 0, 0, 1.0], (4, 1)), atol=1e-6)
    p = ms[1]
    assert p["positions"].shape == (9, 3)          # 3 (strided) + 6; the LINES primitive is skipped
    assert np.array_equal(p["positions"][:3], e["tp"]) and np.array_equal(p["positions"][3:], e["pp"])
    assert np.array_equal(p["indices"], [[0, 1, 2], [3, 4, 5], [6, 7, 8]])
    assert np.array_equal(p["tri_material"], [0, 1, 1])  # section index counts drawn primitives
    assert np.array_equal(p["tangents"][:3], e["tt"]) and np.all(p["tangents"][3:] == 0)
    # no NORMAL: face normals (triangles of z = 1 and z = 2, opposite windings)
    np.testing.assert_allclose(p["normals"][3:6], np.tile([0, 0, 1.0], (3, 1)), atol=1e-6)
    np.testing.assert_allclose(p["normals"][6:9], np.tile([0, 0, -1.0], (3, 1)), atol=1e-6)
    assert np.all(p["texcoords"][3:] == 0)


def test_nodes_lights_camera_and_environment(loaded):
    g, e = loaded
    ins = g.scene.instances
    assert [i.mesh_id for i in ins] == [0, 1, 2] and [i.instance_id for i in ins] == [1, 2, 3]
    root = trs((1, 2, 3), e["q_root"], (2, 2, 2))
    want = [root @ trs((0, 0, -1)), root @ trs(q=e["q_parts"]),
            root @ trs(q=e["q_parts"]) @ trs((0, 5, 0), s=(0.5, 1, 0.5))]
    for i, w in zip(ins, want):
        M = np.array(list(i.world), np.float64).reshape(4, 4).T  # column-major
        np.testing.assert_allclose(M, w, atol=1e-5)
        assert list(i.world) == list(i.prev_world) and i.visible == 1 and i.use_remap == 1
    # the node's material map indexes all primitives, the LINES one included
    assert list(ins[1].mat_id)[:3] == [1, 2, 3] and list(ins[2].mat_id)[:2] == [4, 5]
    # area light from the node's own transform (its parent is set after the light is made)
    L = g.scene.lights
    assert len(L) == 1
    np.testing.assert_allclose(list(L[0].p0)[:3], [-0.5, 5, -0.5], atol=1e-6)
    np.testing.assert_allclose(list(L[0].p1)[:3], [-0.5, 5, 0.5], atol=1e-6)
    np.testing.assert_allclose(list(L[0].p3)[:3], [0.5, 5, -0.5], atol=1e-6)
    np.testing.assert_allclose(list(L[0].normal_area), [0, 1, 0, 0.5], atol=1e-6)
    assert L[0].light_mat_idx == 5
    # camera from the node's own transform, fov 40
    i = g.info
    assert i["num_cameras"] == 1 and i["fov"] == 40.0
    np.testing.assert_allclose(i["eye"], [0, 1, 10], atol=1e-6)
    np.testing.assert_allclose(i["target"], [0, 1, 9], atol=1e-6)
    assert (i["has_sky"], i["has_sun"]) == (1, 1)
    assert i["sky_intensity"] == np.float32(2.5) and i["sun_rotation"] == np.float32(0.3)


def test_gltf_file_with_external_buffer_and_data_uri_matches_glb(tmp_path):
    B, _ = make_scene()
    glb = gk.GltfScene(B.glb())
    doc = json.loads(json.dumps(B.doc))
    (tmp_path / "scene data.bin").write_bytes(bytes(B.bin))
    doc["buffers"] = [{"byteLength": len(B.bin), "uri": "scene%20data.bin"}]
    # first image as a data URI instead of a buffer view
    v = doc["bufferViews"][doc["images"][0]["bufferView"]]
    raw = bytes(B.bin[v["byteOffset"]:v["byteOffset"] + v["byteLength"]])
    doc["images"][0] = {"uri": "data:image/png;base64," + base64.b64encode(raw).decode()}
    (tmp_path / "scene.gltf").write_text(json.dumps(doc, indent=1))
    g = gk.GltfScene(tmp_path / "scene.gltf")
    for a, b in zip(glb.scene.meshes, g.scene.meshes):
        for k in ("positions", "normals", "texcoords", "tangents", "indices", "tri_material"):
            assert np.array_equal(a[k], b[k]), k
    for a, b in zip(glb.scene.textures, g.scene.textures):
        assert np.array_equal(a["rgba"], b["rgba"])
    assert bytes(glb.scene.c_materials()) == bytes(g.scene.c_materials())
    assert bytes(glb.scene.c_instances()) == bytes(g.scene.c_instances())


def test_loader_errors(tmp_path):
    with pytest.raises(gk.GkError) as ei:
        gk.GltfScene(tmp_path / "missing.glb")
    assert ei.value.status == gk.GK_ERR_IO
    with pytest.raises(gk.GkError) as ei:
        gk.GltfScene(b'{"asset": {"version": "2.0"}, "meshes": [')
    assert ei.value.status == gk.GK_ERR_ARG and "malformed JSON" in str(ei.value)
    B, _ = make_scene()
    B.doc["meshes"][0]["primitives"][0]["indices"] = B.accessor(np.array([0, 1, 9]), 5123)
    with pytest.raises(gk.GkError) as ei:
        gk.GltfScene(B.glb())
    assert "index out of range" in str(ei.value)
    B, _ = make_scene()
    jpeg = b"\xff\xd8\xff\xe0" + bytes(64)
    v = B._view(jpeg)
    B.doc["images"][0] = {"bufferView": v, "mimeType": "image/jpeg"}
    with pytest.raises(gk.GkError) as ei:
        gk.GltfScene(B.glb())
    assert "not PNG" in str(ei.value)

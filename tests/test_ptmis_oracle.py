"""Pins the oracle's PT_MIS estimator (oracle/oracle_render.cpp, PtMis) by what the
mathematics fixes, independently of the formulas it is written with.

* Strategy agreement. Next-event estimation alone, BSDF sampling alone and their
  power-heuristic combination are three unbiased estimators of the same image, so their
  expectations agree. The three renders replay the same paths (the same random numbers,
  only the weights differ), which makes the comparison sharp; on a Cornell box with glossy
  materials and a large emitter, each of three plausible weight mistakes planted in the
  oracle (light pdf without the emitter cosine, without the light-selection probability,
  balance heuristic on one side) fails the same check by a wide margin.
* Furnace. A single planar quad under a uniform sky: a reflected path cannot hit the quad
  again, so a pixel's expected value is L_sky times the BSDF's directional albedo
  E(wo) = integral of f(wo, wi) cos(wi) over the hemisphere. A Lambertian quad of albedo 1
  returns L_sky on every path; for GGX (metallic and Mixture) E(wo) is integrated here by
  quadrature over half-vectors from the published GGX formulas (Walter et al. 2007 normal
  distribution, Heitz 2014 height-correlated Smith masking, Schlick Fresnel), so the
  oracle's bounded-VNDF sampler, its pdf and its throughput update are held to the BSDF
  they sample.
"""
from __future__ import annotations

import numpy as np
import pytest

import gknextengine_amd as gk
from gknextengine_amd.scenes import material
from oracle import pyoracle as po
from tests import scenes


def _params(W, H, bounces, eye=None, target=None, fov=40.0):
    cam = scenes.CORNELL_CAM
    p = gk.make_frame_params(eye or cam["eye"], target or cam["target"], W, H, fov=fov)
    p.estimator = gk.ESTIMATOR_PT_MIS
    p.bounces = p.max_bounces = bounces
    p.samples = 1
    p.global_seed = 17
    return p


# ---------------------------------------------------------------- strategy agreement
def _chunk_means(osc, p, W, mode, chunks, frames):
    return np.array([osc.render_ptmis_mode(p, W, W, mode, frame0=7919 * k, frames=frames).mean()
                     for k in range(chunks)])


def _t(a, b):
    d = a - b
    return float(d.mean() / (d.std(ddof=1) / np.sqrt(d.size)))


@pytest.fixture(scope="module")
def glossy():
    s = scenes.glossy_cornell()
    osc = po.OracleScene(s)
    assert osc.num_lights() == 4  # two ceiling triangles, two small-light triangles
    W, chunks, frames = 32, 16, 256
    p = _params(W, W, 4)
    runs = {m: _chunk_means(osc, p, W, m, chunks, frames)
            for m in (po.OracleScene.MIS, po.OracleScene.NEE_ONLY, po.OracleScene.BSDF_ONLY,
                      po.OracleScene.PLANTED_NO_COSINE, po.OracleScene.PLANTED_NO_SELECTION_PDF,
                      po.OracleScene.PLANTED_BALANCE_HEURISTIC)}
    return runs


def test_mis_agrees_with_nee_only_and_bsdf_only(glossy):
    """t statistics over 16 chunks of 32x32 x 256 paths (4 bounces) with common random
    numbers: the combination is unbiased against both single strategies."""
    mis = glossy[po.OracleScene.MIS]
    t_nee = _t(mis, glossy[po.OracleScene.NEE_ONLY])
    t_bsdf = _t(mis, glossy[po.OracleScene.BSDF_ONLY])
    assert abs(t_nee) < 4.0 and abs(t_bsdf) < 4.0, (t_nee, t_bsdf)
    # and the single strategies agree with each other
    assert abs(_t(glossy[po.OracleScene.NEE_ONLY], glossy[po.OracleScene.BSDF_ONLY])) < 4.0


@pytest.mark.parametrize("mode", ["PLANTED_NO_COSINE", "PLANTED_NO_SELECTION_PDF", "PLANTED_BALANCE_HEURISTIC"])
def test_planted_weight_mistakes_fail_the_agreement_check(glossy, mode):
    """The check above has teeth: each planted mistake moves the image by 1-3 % and the
    t statistic against BSDF-only sampling far past the bound the correct weights meet."""
    bad = glossy[getattr(po.OracleScene, mode)]
    t_bsdf = _t(bad, glossy[po.OracleScene.BSDF_ONLY])
    assert abs(t_bsdf) > 8.0, t_bsdf
    rel = abs(bad.mean() / glossy[po.OracleScene.MIS].mean() - 1.0)
    assert rel > 0.005, rel


# ---------------------------------------------------------------- furnace
def _quad_scene(mat, tilt_deg=50.0):
    """One 1.2 x 1.2 quad centred at (0.5, 0.5, 0.3), tilted about x, facing the camera."""
    a = np.radians(tilt_deg)
    R = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    loc = np.array([[-0.6, -0.6, 0], [0.6, -0.6, 0], [0.6, 0.6, 0], [-0.6, 0.6, 0]])
    c = (loc @ R.T) + np.array([0.5, 0.5, 0.3])
    tris = [[c[0], c[2], c[1]], [c[0], c[3], c[2]]]  # normal towards -z (the camera)
    s = scenes.tri_scene(tris)
    s.materials = [mat]
    n = np.cross(c[2] - c[0], c[1] - c[0])
    return s, n / np.linalg.norm(n)


def _ggx_albedo(cos_o, alpha, metal, albedo=1.0, n_psi=1500, n_phi=720, check_visible=False):
    """Directional albedo of f = (1-metal) albedo/pi + F D G2 / (4 cos_o cos_i), F0 =
    lerp(0.04, albedo, metal), by quadrature over half-vectors h = (theta, phi):
    integral of f cos_i dw_i = (1 - metal) albedo + integral F D G2 (wo.h) / cos_o dw_h
    (dw_i = 4 |wo.h| dw_h), theta = atan(alpha tan psi) to resolve narrow lobes."""
    psi = (np.arange(n_psi) + 0.5) / n_psi * (np.pi / 2)
    th = np.arctan(alpha * np.tan(psi))
    dth = alpha / (np.cos(psi) ** 2 + alpha ** 2 * np.sin(psi) ** 2) * (np.pi / 2 / n_psi)
    ph = (np.arange(n_phi) + 0.5) / n_phi * 2 * np.pi
    dph = 2 * np.pi / n_phi
    T, P = np.meshgrid(th, ph, indexing="ij")
    hx, hy, hz = np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)
    w = np.sin(T) * dth[:, None] * dph
    a2 = alpha * alpha
    D = a2 / (np.pi * (hz * hz * (a2 - 1) + 1) ** 2)
    # the quadrature itself: the projected normal distribution integrates to one
    assert abs(float((D * hz * w).sum()) - 1.0) < 1e-3

    def lam(cz, sz):
        t2 = (sz / np.maximum(cz, 1e-12)) ** 2
        return np.where(cz > 0, (np.sqrt(1 + a2 * t2) - 1) / 2, 0.0)

    f0 = 0.04 * (1 - metal) + albedo * metal
    out = []
    for co in np.atleast_1d(cos_o):
        so = np.sqrt(max(0.0, 1 - co * co))
        wo = np.array([so, 0.0, co])
        oh = wo[0] * hx + wo[2] * hz
        ix, iy, iz = 2 * oh * hx - wo[0], 2 * oh * hy, 2 * oh * hz - wo[2]
        ok = (iz > 0) & (oh > 0)
        si = np.sqrt(np.maximum(0, 1 - iz * iz))
        G2 = 1 / (1 + lam(co, so) + lam(iz, si))
        F = f0 + (1 - f0) * (1 - np.clip(oh, 0, 1)) ** 5
        spec = float((np.where(ok, F * D * G2 * oh / co, 0.0) * w).sum())
        if check_visible:
            # visible normals: G1(wo) max(0, wo.h) D(h) / cos_o integrates to one
            g1 = 1 / (1 + lam(co, so))
            vis = float((np.where(oh > 0, g1 * oh * D / co, 0.0) * w).sum())
            assert abs(vis - 1.0) < 2e-3, (co, alpha, vis)
        out.append((1 - metal) * albedo + spec)
    return np.array(out)


def test_ggx_quadrature_reproduces_known_values():
    """The quadrature of the test against two textbook facts: with the masking term G1 of
    the outgoing direction alone the visible normals integrate to one (Heitz 2014, eq. 38),
    and an F = 1 GGX lobe loses energy to multiple scattering, more so when rough
    (directional albedo < 1, decreasing with roughness)."""
    e = [float(_ggx_albedo(0.7, a, 1.0, check_visible=True)[0]) for a in (0.05, 0.3, 0.8)]
    _ggx_albedo([0.2, 0.95], 0.5, 1.0, check_visible=True)
    assert all(x < 1.0 for x in e) and e[0] > e[1] > e[2] and e[0] > 0.98, e


@pytest.mark.parametrize("kind", ["lambert", "metal_r0.5", "mixture_r0.3"])
def test_furnace_quad_under_uniform_sky(kind):
    sky = 1.5
    if kind == "lambert":
        mat = material((1.0, 1.0, 1.0), gk.MAT_LAMBERTIAN, 1.0, 0.0)
        alpha = metal = None
    elif kind == "metal_r0.5":
        mat = material((1.0, 1.0, 1.0), gk.MAT_METALLIC, 0.5, 1.0)
        alpha, metal = 0.25, 1.0
    else:
        mat = material((1.0, 1.0, 1.0), gk.MAT_MIXTURE, 0.3, 0.5)
        alpha, metal = 0.09, 0.5
    s, n = _quad_scene(mat)
    osc = po.OracleScene(s)
    W = 24
    p = _params(W, W, 3)
    p.has_sky = 1
    p.sky_color[:] = [sky, sky, sky, 1.0]
    rays = po.primary_rays(p, W, W).reshape(W, W, 8)
    hits, _ = osc.trace(rays.reshape(-1, 8))
    on = (hits[:, 3].view(np.uint32) != 0xFFFFFFFF).reshape(W, W)
    assert 0.3 < on.mean() < 0.95
    chunks, frames = 8, 48
    imgs = np.stack([osc.render_ptmis(p, W, W, frame0=1000 * k, frames=frames, fp64=True)["rgb"]
                     for k in range(chunks)])
    img = imgs.mean(0)
    # misses see the sky itself
    assert np.allclose(img[~on], sky, rtol=0, atol=1e-12)
    if kind == "lambert":
        # f cos / pdf = 1 exactly in expectation and per path: every path returns L_sky
        assert np.allclose(imgs[:, on], sky, rtol=1e-12)
        return
    cos_o = np.abs(rays[..., 4:7] @ n)[on]
    table_c = np.linspace(cos_o.min() - 1e-3, min(1.0, cos_o.max() + 1e-3), 24)
    table_e = _ggx_albedo(table_c, alpha, metal)
    expect = sky * np.interp(cos_o, table_c, table_e)
    got = imgs[:, on].mean(axis=2)          # chunks x pixels (grey: rgb equal)
    per_chunk = got.mean(axis=1)
    se = per_chunk.std(ddof=1) / np.sqrt(chunks)
    z = (per_chunk.mean() - expect.mean()) / se
    print(f"{kind}: render {per_chunk.mean():.6f} quadrature {expect.mean():.6f} se {se:.2e} z {z:.2f}")
    assert abs(z) < 4.0, (per_chunk.mean(), expect.mean(), se)
    # a 1 % error in the sampled BSDF weight would be seen
    assert se < 0.002 * expect.mean()

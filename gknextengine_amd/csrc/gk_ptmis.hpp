// gk_ptmis.hpp - BSDF, sampling and light-selection helpers of the PT_MIS estimator,
// shared by the wavefront tracer (wavefront.hip) and the differentiable renderer
// (adjoint.hip).
//
// BSDF: Lambert f_d = (1-metal) albedo / pi plus GGX f_s = F D G2 / (4 cos_o cos_i) with
// alpha = max(roughness, 0.02)^2, Schlick F with F0 = lerp(0.04, albedo, metal); the
// specular lobe is sampled with the bounded VNDF method (Eto & Tokuyoshi 2023), whose
// reflected-direction pdf is D / (2 (k wo.z + t)); lobe probability ps = 0 for Lambertian,
// 0.5 + 0.5 metal otherwise.
#pragma once
#include "gk_scene.hpp"

namespace gk {
namespace ptmis {

constexpr float kPi = 3.14159265358979323846f;

__device__ inline float saturate(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }

struct Bsdf {
    f3 albedo;
    float metal, alpha, ps;
};
__device__ inline Bsdf bsdf_of(const gk_material& m) {
    Bsdf b;
    b.albedo = f3{m.diffuse[0], m.diffuse[1], m.diffuse[2]};
    const float r = fmaxf(m.fuzziness, 0.02f);
    b.alpha = r * r;
    if (m.material_model == GK_MATERIAL_METALLIC) b.metal = 1.f;
    else if (m.material_model == GK_MATERIAL_MIXTURE) b.metal = m.metalness;
    else b.metal = 0.f;
    b.ps = (m.material_model == GK_MATERIAL_LAMBERTIAN || m.material_model == GK_MATERIAL_ISOTROPIC) ? 0.f
                                                                                                       : 0.5f + 0.5f * b.metal;
    return b;
}
__device__ inline float ggx_d(float a2, float cz) {
    const float d = cz * cz * (a2 - 1.f) + 1.f;
    return a2 / (kPi * d * d);
}
__device__ inline float ggx_lambda(float a2, f3 w) {
    const float z2 = w.z * w.z;
    if (z2 <= 0.f) return 0.f;
    const float t2 = (w.x * w.x + w.y * w.y) / z2;
    return (sqrtf(1.f + a2 * t2) - 1.f) * 0.5f;
}
__device__ inline f3 vndf_sample(float alpha, f3 wo, float ux, float uy) {
    const f3 wi = normalize(f3{wo.x * alpha, wo.y * alpha, wo.z});
    float b = wi.z;
    const float a = saturate(alpha);
    const float sl = 1.f + sqrtf(wo.x * wo.x + wo.y * wo.y);
    const float awiz = a * wo.z / sl;
    b = b * ((1.f - a * a) / (1.f + awiz * awiz));
    const float z = (1.f - uy) * (1.f + b) - b;
    float sp, cp;
    sincos_2pi(ux, sp, cp);
    const float r = sqrtf(saturate(1.f - z * z));
    const f3 m = wi + f3{r * cp, r * sp, z};
    return normalize(f3{m.x * alpha, m.y * alpha, m.z});
}
__device__ inline float vndf_pdf(float alpha, f3 wo, f3 m) {
    const float a2 = alpha * alpha;
    const float D = ggx_d(a2, m.z);
    const float len2 = a2 * (wo.x * wo.x + wo.y * wo.y);
    const float t = sqrtf(len2 + wo.z * wo.z);
    const float a = saturate(alpha);
    const float sl = 1.f + sqrtf(wo.x * wo.x + wo.y * wo.y);
    const float s2 = sl * sl, aa = a * a;
    const float k = (1.f - aa) * s2 / (s2 + aa * wo.z * wo.z);
    return D / (2.f * (k * wo.z + t));
}
__device__ inline void bsdf_eval(const Bsdf& b, f3 wo, f3 wi, f3& fd, f3& fs, float& pdd, float& pds) {
    fd = b.albedo * ((1.f - b.metal) / kPi);
    pdd = wi.z / kPi;
    fs = f3{0.f, 0.f, 0.f};
    pds = 0.f;
    if (b.ps <= 0.f) return;
    f3 h = wo + wi;
    const float hl = length(h);
    if (hl <= 0.f) return;
    h = h * (1.f / hl);
    const float a2 = b.alpha * b.alpha;
    const float D = ggx_d(a2, h.z);
    const float G = 1.f / (1.f + ggx_lambda(a2, wo) + ggx_lambda(a2, wi));
    const float c = fmaxf(0.f, dot(wo, h));
    const float x = 1.f - c;
    const float x5 = x * x * x * x * x;
    const f3 f0 = f3{0.04f, 0.04f, 0.04f} * (1.f - b.metal) + b.albedo * b.metal;
    const f3 F = f0 + (f3{1.f, 1.f, 1.f} - f0) * x5;
    fs = F * (D * G / (4.f * wo.z * wi.z));
    pds = vndf_pdf(b.alpha, wo, h);
}
__device__ inline float power_h(float a, float b) {
    const float a2 = a * a, b2 = b * b;
    return a2 + b2 > 0.f ? a2 / (a2 + b2) : 0.f;
}
__device__ inline float light_select_pdf(const DScene& s, uint32_t inst, uint32_t prim) {
    uint32_t lo = 0, hi = s.nlights;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const float4 le = s.lights[mid].le, e2 = s.lights[mid].e2;
        const uint32_t li = __float_as_uint(e2.w), lp = __float_as_uint(le.w);
        if (li < inst || (li == inst && lp < prim)) lo = mid + 1; else hi = mid;
    }
    if (lo >= s.nlights) return 0.f;
    const LightTri& L = s.lights[lo];
    if (__float_as_uint(L.e2.w) != inst || __float_as_uint(L.le.w) != prim) return 0.f;
    return s.light_cdf[lo + 1] - s.light_cdf[lo];
}
__device__ inline uint32_t light_pick(const DScene& s, float u) {
    uint32_t lo = 0, hi = s.nlights - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s.light_cdf[mid + 1] > u) hi = mid; else lo = mid + 1;
    }
    return lo;
}

}  // namespace ptmis
}  // namespace gk

// oracle_render.cpp - CPU restatement of the path-tracing passes (test infrastructure only).
//
//   primary rays   FVisibilityBufferRayCasterV2::TracePrimaryRay ray construction
//                  (Shading.slang:350-356), traced instead of rasterised; with an aperture,
//                  FHardwarePrimaryRayCaster's thin lens (Shading.slang:283-300), whose
//                  RandomInUnitDisk is the path's first draw (none at aperture 0).
//   REF_V2         FPathTracingRendererV2 Init/PrimaryHit/Render/GetRayColor
//                  (Shading.slang:995-1241) with FHardwareRayTracerV2 hit reconstruction
//                  (Shading.slang:784-826), CalculateMotionVectorV2 (Shading.slang:50-72),
//                  FHardwareDirectIlluminatorV2 (Shading.slang:886-907) and the sampling
//                  helpers of Const_Func.slang:8-49,227-340. Path termination
//                  multiplies by interpolateAmbientCubesV2<Full> over the scene's
//                  probe grid ((0,0,0,1) without one).
//   probes         VoxelizeCube (CPUAccelerationStructure.cpp:85-160) and the GPU bake
//                  FGpuProbeGenerator::RenderAmbientCubeOnly/FaceTask (AmbientCube.slang:
//                  472-566,650-690), one probe at a time; a bake pass reads the grid as it
//                  was when the pass started.
//   PT_MIS         the build's GGX/Lambert path tracer with next-event estimation to
//                  emissive triangles and power-heuristic MIS, templated on float/double
//                  (double = the convergence oracle).
#include <algorithm>
#include <atomic>
#include <thread>

#include "oracle_core.hpp"

namespace orc {
float blas_sah_cost(const Mesh& m, uint32_t idx);
uint32_t blas_node_count(const Mesh& m);
}  // namespace orc

using namespace orc;

namespace {

struct InstanceABI {  // gk_instance
    float world[16], prev_world[16];
    uint32_t instance_id, mesh_id, visible, nort, mat_id[16], mat_offset, use_remap, reserved[2];
};
static_assert(sizeof(InstanceABI) == 224, "gk_instance layout");

template <class F>
void parallel_rows(uint32_t n, uint32_t nthreads, F&& f) {
    if (nthreads <= 1) { for (uint32_t i = 0; i < n; ++i) f(i); return; }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < nthreads; ++t)
        pool.emplace_back([&] { for (uint32_t i; (i = next.fetch_add(1)) < n;) f(i); });
    for (auto& th : pool) th.join();
}

constexpr float PI_F = 3.14159265358979323846f;
constexpr float EPS = 1e-3f;                     // PreProcessor.slang:13
constexpr float PT_MAX_TRACE_DISTANCE = 1000.f;  // Shading.slang:15

// ---------------------------------------------------------------- sampling (float)
// Schlick, Const_Func.slang:8-14
float schlick(float cosine, float ior) {
    float r0 = (1.f - ior) / (1.f + ior);
    r0 = r0 * r0;
    const float x = 1.f - cosine;
    return r0 + (1.f - r0) * (x * x * x * x * x);
}
// ONB (Duff et al. 2017), Const_Func.slang:17-22
void onb(F3 n, F3& b1, F3& b2) {
    const float sz = n.z < 0.f ? -1.f : 1.f;
    const float a = -1.0f / (sz + n.z);
    b2 = F3(n.x * n.y * a, sz + n.y * n.y * a, -n.y);
    b1 = F3(1.0f + sz * n.x * n.x * a, sz * b2.x, -sz * n.x);
}
// to_world(v, T, B, N) = v.x*T + v.y*B + v.z*N (PreProcessor.slang:2)
F3 to_world(F3 v, F3 T, F3 B, F3 N) {
    return F3(v.x * T.x + v.y * B.x + v.z * N.x, v.x * T.y + v.y * B.y + v.z * N.y,
              v.x * T.z + v.y * B.z + v.z * N.z);
}
// to_local(v, t, b, n) = (dot(t,v), dot(b,v), dot(n,v)) (PreProcessor.slang:1)
F3 to_local(F3 v, F3 T, F3 B, F3 N) { return F3(dot(T, v), dot(B, v), dot(N, v)); }
// AlignWithNormal, Const_Func.slang:44-49
F3 align_with_normal(F3 ray, F3 normal) {
    F3 T, B;
    onb(normal, T, B);
    return to_world(ray, T, B, normal);
}
// RandomInHemiSphere1 (cosine-weighted), Const_Func.slang:299-305
F3 random_hemisphere1(Rng& s) {
    float ux, uy;
    rnd2(s, ux, uy);
    float sp, cp;
    sincos_2pi(ux, sp, cp);
    const float r = std::sqrt(uy);
    return F3(r * cp, r * sp, std::sqrt(1.0f - uy));
}
// RandomInCone, Const_Func.slang:282-289
F3 random_in_cone(Rng& s, float cos_theta) {
    float ux, uy;
    rnd2(s, ux, uy);
    float sp, cp;
    sincos_2pi(ux, sp, cp);
    const float c = 1.0f + uy * (cos_theta - 1.f);
    const float r = std::sqrt(1.0f - c * c);
    return F3(r * cp, r * sp, c);
}
float saturate(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }
// ggx_sample_vndf (bounded VNDF, Eto & Tokuyoshi 2023), Const_Func.slang:310-328, isotropic
F3 ggx_sample_vndf(float alpha, F3 wi_, float ux, float uy) {
    const F3 wi = normalize(F3(wi_.x * alpha, wi_.y * alpha, wi_.z));
    float b = wi.z;
    if (wi_.z > 0.f) {
        const float a = saturate(alpha);
        const float awiz_s = a * wi_.z / (1.0f + std::sqrt(wi_.x * wi_.x + wi_.y * wi_.y));
        b = b * ((1.0f - a * a) / (1.0f + awiz_s * awiz_s));
    }
    const float z = std::fma(1.0f - uy, 1.0f + b, -b);
    float sp, cp;
    sincos_2pi(ux, sp, cp);
    const float r = std::sqrt(saturate(1.0f - z * z));
    const F3 o(r * cp, r * sp, z);
    const F3 m = wi + o;
    return normalize(F3(m.x * alpha, m.y * alpha, m.z));
}
// ggxSampling, Const_Func.slang:330-340
F3 ggx_sampling(Rng& s, float roughness, F3 normal) {
    F3 T, B;
    onb(normal, T, B);
    const float alpha = roughness * roughness;
    const F3 wl = to_local(normal, T, B, normal);
    float ux, uy;
    rnd2(s, ux, uy);
    return to_world(ggx_sample_vndf(alpha, wl, ux, uy), T, B, normal);
}
// HLSL reflect / refract
F3 reflect(F3 i, F3 n) { return i - n * (2.f * dot(n, i)); }
F3 refract(F3 i, F3 n, float eta) {
    const float d = dot(n, i);
    const float k = 1.f - eta * eta * (1.f - d * d);
    if (k < 0.f) return F3(0.f, 0.f, 0.f);
    return i * eta - n * (eta * d + std::sqrt(k));
}

struct F4 {
    float v[4];
};
F4 mul4(F4 a, F4 b) { return F4{{a.v[0] * b.v[0], a.v[1] * b.v[1], a.v[2] * b.v[2], a.v[3] * b.v[3]}}; }

// ---------------------------------------------------------------- primary rays
// uv = ipos/size*2-1 (no half-pixel offset), origin = MVinv*(0,0,0,1),
// target = Pinv*(uv,1,1), dir = MVinv*(normalize(target.xyz),0), ray = normalize(dir.xyz)
void primary_ray_at(const FrameParams& P, float fx, float fy, uint32_t W, uint32_t H, F3& O, F3& D) {
    const float ux = (fx / float(W)) * 2.0f - 1.0f;
    const float uy = (fy / float(H)) * 2.0f - 1.0f;
    const float o4[4] = {0.f, 0.f, 0.f, 1.f};
    float org[4];
    mat_mul4(P.model_view_inverse, o4, org);
    const float t4[4] = {ux, uy, 1.f, 1.f};
    float tgt[4];
    mat_mul4(P.projection_inverse, t4, tgt);
    const F3 tn = normalize(F3(tgt[0], tgt[1], tgt[2]));
    const float d4[4] = {tn.x, tn.y, tn.z, 0.f};
    float dir[4];
    mat_mul4(P.model_view_inverse, d4, dir);
    O = F3(org[0], org[1], org[2]);
    D = normalize(F3(dir[0], dir[1], dir[2]));
}
void primary_ray(const FrameParams& P, uint32_t x, uint32_t y, uint32_t W, uint32_t H, F3& O, F3& D) {
    primary_ray_at(P, float(x), float(y), W, H, O, D);
}
// FHardwarePrimaryRayCaster's thin lens (Shading.slang:283-300): offset = Aperture / 2 *
// RandomInUnitDisk (the path's first draw), origin MVinv (offset, 0, 1), direction
// MVinv normalize(Pinv (uv, 1, 1) * FocusDistance - (offset, 0)). Aperture 0: the pinhole
// ray, no draw.
void primary_ray_lens(const FrameParams& P, uint32_t x, uint32_t y, uint32_t W, uint32_t H, Rng& seed, F3& O,
                      F3& D) {
    if (!(P.aperture > 0.f)) {
        primary_ray(P, x, y, W, H, O, D);
        return;
    }
    float a, b, dx, dy;
    rnd2(seed, a, b);
    concentric_disk(a, b, dx, dy);
    const float half = P.aperture / 2.f;
    const float lx = half * dx, ly = half * dy;
    const float ux = (float(x) / float(W)) * 2.0f - 1.0f;
    const float uy = (float(y) / float(H)) * 2.0f - 1.0f;
    const float o4[4] = {lx, ly, 0.f, 1.f};
    float org[4];
    mat_mul4(P.model_view_inverse, o4, org);
    const float t4[4] = {ux, uy, 1.f, 1.f};
    float tgt[4];
    mat_mul4(P.projection_inverse, t4, tgt);
    const F3 tl = normalize(F3(tgt[0] * P.focus_distance - lx, tgt[1] * P.focus_distance - ly,
                               tgt[2] * P.focus_distance - 0.f));
    const float d4[4] = {tl.x, tl.y, tl.z, 0.f};
    float dir[4];
    mat_mul4(P.model_view_inverse, d4, dir);
    O = F3(org[0], org[1], org[2]);
    D = normalize(F3(dir[0], dir[1], dir[2]));
}

// hit reconstruction ----------------------------------------------------------
struct Vtx {
    F3 pos, nrm;
    uint32_t mat;
};

// FHardwareRayTracerV2::TraceRay: normal = normalize(WorldToObject^T * Mix(n0,n1,n2,bary)),
// position = origin + dir * t (Shading.slang:803-821).
Vtx hit_vertex(const Scene& sc, const Hit& h, F3 O, F3 D) {
    const Instance& I = sc.inst[h.inst];
    const Mesh& m = sc.meshes[I.mesh_id];
    const F3 n0 = m.n0[h.prim], n1 = m.n1[h.prim], n2 = m.n2[h.prim];
    // Mix(a,b,c,bary) = a + (b-a)*bary.y + (c-a)*bary.z  (PreProcessor.slang:24)
    const F3 mix = (n0 + (n1 - n0) * h.u) + (n2 - n0) * h.v;
    const float* W = I.inv;  // row-major world->object
    F3 nw;
    for (int j = 0; j < 3; ++j) nw[j] = W[0 * 4 + j] * mix.x + W[1 * 4 + j] * mix.y + W[2 * 4 + j] * mix.z;
    Vtx v;
    v.nrm = normalize(nw);
    v.pos = O + D * h.t;
    v.mat = material_of(sc, h.inst, h.prim);
    return v;
}

// primary vertex as get_material_dataV2 (GeneralFunc.slang:20-53): world-space vertex
// normals blended with (1-u-v, u, v): bx*n0 + (by*n1 + bz*n2).
Vtx primary_vertex(const Scene& sc, const Hit& h, F3 O, F3 D) {
    const Instance& I = sc.inst[h.inst];
    const Mesh& m = sc.meshes[I.mesh_id];
    const F3 n0 = xform_vector(I.world, m.n0[h.prim]);
    const F3 n1 = xform_vector(I.world, m.n1[h.prim]);
    const F3 n2 = xform_vector(I.world, m.n2[h.prim]);
    const float by = h.u, bz = h.v, bx = 1.0f - (by + bz);
    const F3 nb = n0 * bx + (n1 * by + n2 * bz);
    Vtx v;
    v.nrm = normalize(nb);
    v.pos = O + D * h.t;
    v.mat = material_of(sc, h.inst, h.prim);
    return v;
}

const Material& mat_at(const Scene& sc, uint32_t i) { return sc.mats[i]; }

// 4x4 column-major product a*b
void mat_mul44(const float* a, const float* b, float* o) {
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            float acc = a[0 * 4 + r] * b[c * 4 + 0];
            acc = acc + a[1 * 4 + r] * b[c * 4 + 1];
            acc = acc + a[2 * 4 + r] * b[c * 4 + 2];
            acc = acc + a[3 * 4 + r] * b[c * 4 + 3];
            o[c * 4 + r] = acc;
        }
}

// combinedPrevTS = prev * inverse(world) (Node::TickVelocity, Node.cpp:82-106), computed
// in double by Gauss-Jordan elimination with partial pivoting, identity when the origin
// moves by at most 0.001.
void combined_prev(const float* world, const float* prev, float* out) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) { a[r][c] = world[c * 4 + r]; a[r][4 + c] = r == c ? 1.0 : 0.0; }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r) if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (p != c) for (int k = 0; k < 8; ++k) std::swap(a[c][k], a[p][k]);
        const double d = a[c][c];
        if (d == 0.0) { for (int i = 0; i < 16; ++i) out[i] = (i % 5 == 0) ? 1.f : 0.f; return; }
        for (int k = 0; k < 8; ++k) a[c][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = a[r][c];
            if (f != 0.0) for (int k = 0; k < 8; ++k) a[r][k] -= f * a[c][k];
        }
    }
    double res[16];  // column-major prev * inv
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += double(prev[k * 4 + r]) * a[k][4 + c];
            res[c * 4 + r] = acc;
        }
    const double tx = res[12], ty = res[13], tz = res[14];
    if (std::sqrt(tx * tx + ty * ty + tz * tz) > 0.001) {
        for (int i = 0; i < 16; ++i) out[i] = float(res[i]);
    } else {
        for (int i = 0; i < 16; ++i) out[i] = (i % 5 == 0) ? 1.f : 0.f;
    }
}

// ---------------------------------------------------------------- textures
// SampleLevel(uv, 0) on the reference's default sampler (Sampler.hpp:9-26: linear
// filter, REPEAT): the coordinate is wrapped into [0, 1), the four texels around
// u * w - 0.5 are decoded to linear (sRGB textures by the sRGB transfer function) and
// blended, first along u, then along v.
F4 texel_lin(const Scene& sc, const TextureO& t, uint32_t x, uint32_t y) {
    const uint32_t p = t.texels[size_t(y) * t.w + x];
    const float* lut = t.srgb ? sc.tex_lut : sc.tex_lut + 256;
    return F4{{lut[p & 0xFF], lut[(p >> 8) & 0xFF], lut[(p >> 16) & 0xFF], sc.tex_lut[256 + (p >> 24)]}};
}
void wrap_axis(float u, uint32_t n, uint32_t& i0, uint32_t& i1, float& a) {
    float w = u - std::floor(u);
    if (!(w >= 0.f && w < 1.f)) w = 0.f;
    const float x = w * float(n) - 0.5f;
    const float fx = std::floor(x);
    a = x - fx;
    const int k = int(fx);
    i0 = k < 0 ? n - 1 : (uint32_t(k) >= n ? uint32_t(k) - n : uint32_t(k));
    i1 = (i0 + 1 == n) ? 0 : i0 + 1;
}
F4 tex_sample(const Scene& sc, uint32_t id, float u, float v) {
    const TextureO& t = sc.tex[id];
    uint32_t x0, x1, y0, y1;
    float ax, ay;
    wrap_axis(u, t.w, x0, x1, ax);
    wrap_axis(v, t.h, y0, y1, ay);
    const F4 a = texel_lin(sc, t, x0, y0), b = texel_lin(sc, t, x1, y0);
    const F4 c = texel_lin(sc, t, x0, y1), d = texel_lin(sc, t, x1, y1);
    F4 o;
    for (int k = 0; k < 4; ++k) {
        const float top = a.v[k] * (1.f - ax) + b.v[k] * ax;
        const float bot = c.v[k] * (1.f - ax) + d.v[k] * ax;
        o.v[k] = top * (1.f - ay) + bot * ay;
    }
    return o;
}
bool has_tex(const Scene& sc, int32_t id) { return id >= 0 && uint32_t(id) < sc.tex.size(); }

// ---------------------------------------------------------------- sky
// SampleIBLV2(SkyIdx, dir, SkyRotation, 0) * SkyIntensity (Shading.slang:194-199, called at
// 1077 and 1193): the direction normalised, longitude u = (atan2(d.x, d.z) + pi rotation) /
// (2 pi), colatitude v = acos(d.y) / pi, here atan2(sqrt((1 - y)(1 + y)), y) / pi; the float
// image sampled with the default sampler (REPEAT on both axes, bilinear as tex_sample), then
// min((10, 10, 10, 1), texel) times the frame's sky_color. Without an image, sky_color.
F4 sky_radiance(const Scene& sc, const FrameParams& P, F3 dir) {
    const F4 col{{P.sky_color[0], P.sky_color[1], P.sky_color[2], P.sky_color[3]}};
    if (sc.sky_w == 0) return col;
    const F3 d = normalize(dir);
    const float u = atan2_turns(d.x, d.z) + P.sky_rotation * 0.5f;
    const float y = std::min(1.f, std::max(-1.f, d.y));
    const float v = 2.f * atan2_turns(std::sqrt((1.f - y) * (1.f + y)), y);
    uint32_t x0, x1, y0, y1;
    float ax, ay;
    wrap_axis(u, sc.sky_w, x0, x1, ax);
    wrap_axis(v, sc.sky_h, y0, y1, ay);
    auto px = [&](uint32_t x, uint32_t yy, int k) { return sc.sky[4 * (size_t(yy) * sc.sky_w + x) + k]; };
    static const float cap[4] = {10.f, 10.f, 10.f, 1.f};
    F4 o;
    for (int k = 0; k < 4; ++k) {
        const float top = px(x0, y0, k) * (1.f - ax) + px(x1, y0, k) * ax;
        const float bot = px(x0, y1, k) * (1.f - ax) + px(x1, y1, k) * ax;
        const float t = top * (1.f - ay) + bot * ay;
        o.v[k] = std::min(cap[k], t) * col.v[k];
    }
    return o;
}

// SampleIBLRough(SkyIdx, dir, SkyRotation) * SkyIntensity (Shading.slang:146-192 as used by
// the probe bake, AmbientCube.slang:515): EvaluateSH with rotate = 1 - SkyRotation, i.e. the
// direction turned about +y by pi (1 - SkyRotation), summed over the nine basis functions in
// order; alpha 1. Times the frame's sky_color; without an image, sky_color.
F4 sky_rough(const Scene& sc, const FrameParams& P, F3 n) {
    const F4 col{{P.sky_color[0], P.sky_color[1], P.sky_color[2], P.sky_color[3]}};
    if (sc.sky_w == 0) return col;
    float turns = (1.f - P.sky_rotation) * 0.5f;
    turns = turns - std::floor(turns);
    float sa, ca;
    sincos_2pi(turns, sa, ca);
    const float rx = n.x * ca + n.z * sa, ry = n.y, rz = -n.x * sa + n.z * ca;
    const float C0 = 0.282095f, C1 = 0.488603f, C2 = 1.092548f, C3 = 0.315392f, C4 = 0.546274f;
    const float basis[9] = {C0, -C1 * ry, C1 * rz, -C1 * rx, C2 * rx * ry, -C2 * ry * rz, C3 * (3.f * ry * ry - 1.0f),
                            -C2 * rx * rz, C4 * (rx * rx - rz * rz)};
    float rgb[3] = {0.f, 0.f, 0.f};
    for (int i = 0; i < 9; ++i)
        for (int ch = 0; ch < 3; ++ch) rgb[ch] += sc.sky_sh[9 * ch + i] * basis[i];
    return F4{{rgb[0] * col.v[0], rgb[1] * col.v[1], rgb[2] * col.v[2], col.v[3]}};
}

// ProjectHdrToSh (Texture.cpp:180-253): for each pixel centre, theta = pi (y + 0.5) / h,
// phi = 2 pi (x + 0.5) / w, direction (sin theta cos phi, cos theta, sin theta sin phi),
// solid-angle weight sin theta (pi / h)(2 pi / w); coefficient i of channel c accumulates
// colour_c * basis_i * weight in float over rows, then columns.
void project_sh(const float* rgba, uint32_t w, uint32_t h, float out[27]) {
    const float PI = 3.14159265358979323846f;
    const float C0 = 0.282095f, C1 = 0.488603f, C2 = 1.092548f, C3 = 0.315392f, C4 = 0.546274f;
    std::fill(out, out + 27, 0.f);
    for (uint32_t y = 0; y < h; ++y) {
        const float theta = ((float(y) + 0.5f) / float(h)) * PI;
        const float sin_t = std::sin(theta), cos_t = std::cos(theta);
        const float weight = sin_t * (PI / float(h)) * (2.0f * PI / float(w));
        for (uint32_t x = 0; x < w; ++x) {
            const float phi = ((float(x) + 0.5f) / float(w)) * 2.0f * PI;
            const float dx = sin_t * std::cos(phi), dy = cos_t, dz = sin_t * std::sin(phi);
            const float basis[9] = {C0, -C1 * dy, C1 * dz, -C1 * dx, C2 * dx * dy, -C2 * dy * dz,
                                    C3 * (3.0f * dy * dy - 1.0f), -C2 * dx * dz, C4 * (dx * dx - dz * dz)};
            const float* c = rgba + 4 * (size_t(y) * w + x);
            for (int ch = 0; ch < 3; ++ch)
                for (int i = 0; i < 9; ++i) out[9 * ch + i] += c[ch] * basis[i] * weight;
        }
    }
}

// Diffuse * base colour texture (GetRayColor, Shading.slang:1047-1052)
F4 mat_albedo(const Scene& sc, const Material& m, float u, float v) {
    F4 a{{m.diffuse[0], m.diffuse[1], m.diffuse[2], m.diffuse[3]}};
    if (has_tex(sc, m.diffuse_texture_id)) a = mul4(a, tex_sample(sc, uint32_t(m.diffuse_texture_id), u, v));
    return a;
}

// texture coordinate at a traced hit: Mix(t0, t1, t2, bary) (Shading.slang:820, PreProcessor.slang:24)
void hit_uv(const Scene& sc, const Hit& h, float& u, float& v) {
    const Mesh& m = sc.meshes[sc.inst[h.inst].mesh_id];
    u = v = 0.f;
    if (m.uv.empty()) return;
    const float* t = &m.uv[6 * size_t(h.prim)];
    u = (t[0] + (t[2] - t[0]) * h.u) + (t[4] - t[0]) * h.v;
    v = (t[1] + (t[3] - t[1]) * h.u) + (t[5] - t[1]) * h.v;
}

// FetchGBufferV2 (Shading.slang:82-112) with get_material_dataV2's interpolation
// bx*a0 + (by*a1 + bz*a2) (GeneralFunc.slang:48-53) and mikkTSpace (Shading.slang:74-80)
void fetch_gbuffer(const Scene& sc, const Hit& h, const Material& mat, F3& nrm, F4& albedo, float& rough) {
    const Instance& I = sc.inst[h.inst];
    const Mesh& m = sc.meshes[I.mesh_id];
    const float by = h.u, bz = h.v, bx = 1.0f - (by + bz);
    float u = 0.f, v = 0.f;
    if (!m.uv.empty()) {
        const float* t = &m.uv[6 * size_t(h.prim)];
        u = bx * t[0] + (by * t[2] + bz * t[4]);
        v = bx * t[1] + (by * t[3] + bz * t[5]);
    }
    albedo = mat_albedo(sc, mat, u, v);
    rough = mat.fuzziness;
    if (has_tex(sc, mat.mra_texture_id)) rough = rough * tex_sample(sc, uint32_t(mat.mra_texture_id), u, v).v[1];
    if (!has_tex(sc, mat.normal_texture_id)) return;
    const F4 nt = tex_sample(sc, uint32_t(mat.normal_texture_id), u, v);
    const F3 ln = m.n0[h.prim] * bx + (m.n1[h.prim] * by + m.n2[h.prim] * bz);
    float lt[4] = {0.f, 0.f, 0.f, 0.f};
    if (!m.tan.empty()) {
        const float* a = &m.tan[12 * size_t(h.prim)];
        for (int k = 0; k < 4; ++k) lt[k] = bx * a[k] + (by * a[4 + k] + bz * a[8 + k]);
    }
    const F3 vnt(nt.v[0] * 2.0f - 1.0f, nt.v[1] * 2.0f - 1.0f, nt.v[2] * 2.0f - 1.0f);
    const F3 tg(lt[0], lt[1], lt[2]);
    const F3 vb = lt[3] * cross(ln, tg);
    const F3 n = normalize(vnt.x * tg + vnt.y * vb + vnt.z * ln);
    nrm = normalize(xform_vector(I.world, n));
}

// ---------------------------------------------------------------- ambient-cube probes
// packRGB10A2 / unpackRGB10A2 (AmbientCube.slang:66-87), MAX_ILLUMINANCE = 512
uint32_t pack_rgb10a2(const F4& c) {
    float q[4];
    for (int i = 0; i < 4; ++i) q[i] = std::fmin(std::fmax(c.v[i] / 512.0f, 0.0f), 1.0f);
    return uint32_t(q[0] * 1023.0f) | (uint32_t(q[1] * 1023.0f) << 10) | (uint32_t(q[2] * 1023.0f) << 20) |
           (uint32_t(q[3] * 3.0f) << 30);
}
F4 unpack_rgb10a2(uint32_t p) {
    const float r = float(p & 0x3FF) / 1023.0f;
    const float g = float((p >> 10) & 0x3FF) / 1023.0f;
    const float b = float((p >> 20) & 0x3FF) / 1023.0f;
    return F4{{r * 512.f, g * 512.f, b * 512.f, 0.0f * 512.f}};
}
// unpack_bytes / pack_bytes (AmbientCube.slang:90-109)
void unpack_bytes(uint32_t p, uint32_t b[4]) {
    b[0] = p & 0xFF; b[1] = (p >> 8) & 0xFF; b[2] = (p >> 16) & 0xFF; b[3] = (p >> 24) & 0xFF;
}
uint32_t pack_bytes(const uint32_t b[4]) {
    return (b[0] & 0xFF) | ((b[1] & 0xFF) << 8) | ((b[2] & 0xFF) << 16) | ((b[3] & 0xFF) << 24);
}
F4 add4(const F4& a, const F4& b) { return F4{{a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2], a.v[3] + b.v[3]}}; }
F4 scale4(float s, const F4& a) { return F4{{s * a.v[0], s * a.v[1], s * a.v[2], s * a.v[3]}}; }
F4 lerp4(const F4& a, const F4& b, float t) {
    F4 o;
    for (int i = 0; i < 4; ++i) o.v[i] = a.v[i] + (b.v[i] - a.v[i]) * t;
    return o;
}

// sampleAmbientCubeHL2_DI and sampleAmbientCubeHL2_Full (AmbientCube.slang:117-232)
F4 sample_cube(const CubeO& cube, F3 n, bool full) {
    const float wx = std::fmax(n.x, 0.f), wnx = std::fmax(-n.x, 0.f);
    const float wy = std::fmax(n.y, 0.f), wny = std::fmax(-n.y, 0.f);
    const float wz = std::fmax(n.z, 0.f), wnz = std::fmax(-n.z, 0.f);
    const float sum = wx + wnx + wy + wny + wz + wnz;
    F4 color{{0.f, 0.f, 0.f, 0.f}};
    float skyvis = 0.f;
    if (full) {
        uint32_t s0[4], s1[4];
        unpack_bytes(cube.sky0, s0);
        unpack_bytes(cube.sky1, s1);
        skyvis += float(s0[0]) * wz;
        skyvis += float(s0[1]) * wnz;
        skyvis += float(s0[2]) * wy;
        skyvis += float(s0[3]) * wny;
        skyvis += float(s1[0]) * wx;
        skyvis += float(s1[1]) * wnx;
    }
    color = add4(color, scale4(wx, unpack_rgb10a2(cube.pos_x_d)));
    color = add4(color, scale4(wnx, unpack_rgb10a2(cube.neg_x_d)));
    color = add4(color, scale4(wy, unpack_rgb10a2(cube.pos_y_d)));
    color = add4(color, scale4(wny, unpack_rgb10a2(cube.neg_y_d)));
    color = add4(color, scale4(wz, unpack_rgb10a2(cube.pos_z_d)));
    color = add4(color, scale4(wnz, unpack_rgb10a2(cube.neg_z_d)));
    const float k = sum > 0.0f ? (1.0f / sum) : 1.0f;
    if (!full) {
        for (int i = 0; i < 4; ++i) color.v[i] *= k;
        return color;
    }
    color = add4(color, scale4(wx, unpack_rgb10a2(cube.pos_x)));
    color = add4(color, scale4(wnx, unpack_rgb10a2(cube.neg_x)));
    color = add4(color, scale4(wy, unpack_rgb10a2(cube.pos_y)));
    color = add4(color, scale4(wny, unpack_rgb10a2(cube.neg_y)));
    color = add4(color, scale4(wz, unpack_rgb10a2(cube.pos_z)));
    color = add4(color, scale4(wnz, unpack_rgb10a2(cube.neg_z)));
    for (int i = 0; i < 3; ++i) color.v[i] *= k;
    skyvis *= k;
    color.v[3] = skyvis / 255.0f;
    return color;
}

// interpolateAmbientCubesV2 (AmbientCube.slang:268-345)
F4 interpolate_cubes(const ProbeRef& pr, F3 in_pos, F3 normal, bool full) {
    if (!pr.cubes) return F4{{0.f, 0.f, 0.f, 1.f}};
    const ProbeGridO& g = pr.grid;
    const F3 nearpos((in_pos.x - g.offset[0]) / g.unit, (in_pos.y - g.offset[1]) / g.unit,
                     (in_pos.z - g.offset[2]) / g.unit);
    const int XY = int(g.size_xy), Z = int(g.size_z);
    // the reference's early out; a NaN position also counts as outside
    if (!(nearpos.x >= 0 && nearpos.y >= 0 && nearpos.z >= 0 && nearpos.x <= float(XY - 2) &&
          nearpos.y <= float(Z - 2) && nearpos.z <= float(XY - 2)))
        return F4{{0.f, 0.f, 0.f, 1.f}};
    const F3 fl(std::floor(nearpos.x), std::floor(nearpos.y), std::floor(nearpos.z));
    const int base[3] = {int(fl.x), int(fl.y), int(fl.z)};
    const F3 frac = nearpos - fl;
    float total_weight = 0.0f;
    F4 result{{0.f, 0.f, 0.f, 0.f}};
    for (int i = 0; i < 8; i++) {
        const int off[3] = {i & 1, (i >> 1) & 1, (i >> 2) & 1};
        const int px = base[0] + off[0], py = base[1] + off[1], pz = base[2] + off[2];
        const size_t idx = size_t(py) * XY * XY + size_t(pz) * XY + size_t(px);
        const CubeO& cube = pr.cubes[idx];
        const VoxelO& voxel = pr.vox[idx];
        uint32_t u0[4], u1[4];
        unpack_bytes(voxel.dist0, u0);
        unpack_bytes(voxel.dist1, u1);
        float d0[4], d1[4];
        for (int k = 0; k < 4; ++k) { d0[k] = float(u0[k]) / 255.0f; d1[k] = float(u1[k]) / 255.0f; }
        if (d0[1] < 0.01f) continue;
        const F3 to_local = frac - F3(float(off[0]), float(off[1]), float(off[2]));
        const float dist_to_probe = length(to_local);
        const F3 dir = normalize(to_local);
        const float dist_pz = d0[2], dist_nz = d0[3], dist_px = d1[0], dist_nx = d1[1], dist_py = d1[2],
                    dist_ny = d1[3];
        const float hit_length = std::sqrt(std::fmax(dir.x, 0.0f)) * dist_px + std::sqrt(std::fmax(-dir.x, 0.0f)) * dist_nx +
                                 std::sqrt(std::fmax(dir.y, 0.0f)) * dist_py + std::sqrt(std::fmax(-dir.y, 0.0f)) * dist_ny +
                                 std::sqrt(std::fmax(dir.z, 0.0f)) * dist_pz + std::sqrt(std::fmax(-dir.z, 0.0f)) * dist_nz;
        if (dist_to_probe > hit_length + 0.05f) continue;
        const float wx = off[0] == 0 ? (1.0f - frac.x) : frac.x;
        const float wy = off[1] == 0 ? (1.0f - frac.y) : frac.y;
        const float wz = off[2] == 0 ? (1.0f - frac.z) : frac.z;
        const float weight = wx * wy * wz;
        const F4 sc = sample_cube(cube, normal, full);
        for (int k = 0; k < 4; ++k) result.v[k] = result.v[k] + sc.v[k] * weight;
        total_weight += weight;
    }
    if (!(total_weight > 0.0f)) return F4{{0.f, 0.f, 0.f, 0.f}};
    for (int k = 0; k < 4; ++k) result.v[k] = result.v[k] / total_weight;
    return result;
}

F3 probe_origin(const ProbeGridO& g, uint32_t idx) {
    const uint32_t y = idx / (g.size_xy * g.size_xy);
    const uint32_t z = (idx - y * g.size_xy * g.size_xy) / g.size_xy;
    const uint32_t x = idx - y * g.size_xy * g.size_xy - z * g.size_xy;
    return F3(float(x) * g.unit + g.offset[0], float(y) * g.unit + g.offset[1], float(z) * g.unit + g.offset[2]);
}

// InsideGeometry (CPUAccelerationStructure.cpp:85-110); the hit normal is the triangle's
// geometric normal taken to world space by the instance transform (:50-58)
bool inside_geometry(const Scene& sc, float unit, F3 origin, F3 ray_dir, VoxelO& cube, float& distance) {
    Hit h;
    if (trace_closest(sc, origin, ray_dir, 0.0f, unit * 64.f, h)) {
        distance = h.t;
        if (distance <= unit) {
            const uint32_t mat = material_of(sc, h.inst, h.prim);
            cube.mat_id = mat;
            const Instance& I = sc.inst[h.inst];
            const Mesh& m = sc.meshes[I.mesh_id];
            const F3 n = xform_vector(I.world, normalize(cross(m.e1[h.prim], m.e2[h.prim])));
            if (dot(n, ray_dir) > 0.0f || sc.mats[mat].material_model == MAT_DIFFUSE_LIGHT) {
                distance = 0;
                return true;
            }
        }
    }
    return false;
}

// DetectDistance (CPUAccelerationStructure.cpp:72-83)
float detect_distance(const Scene& sc, float unit, F3 origin, F3 dir) {
    Hit h;
    if (trace_closest(sc, origin, normalize(dir), 0.0f, unit * 64.f, h)) return h.t;
    return 255;
}

// VoxelizeCube (CPUAccelerationStructure.cpp:112-160)
void voxelize_cube(const Scene& sc, float unit, VoxelO& cube, F3 origin) {
    cube.age = 0;
    cube.mat_id = 0;
    float dist_py = 255.0f, dist_ny = 255.0f, dist_px = 255.0f, dist_nx = 255.0f, dist_pz = 255.0f, dist_nz = 255.0f;
    inside_geometry(sc, unit, origin, F3(0, 1, 0), cube, dist_py);
    inside_geometry(sc, unit, origin, F3(0, -1, 0), cube, dist_ny);
    inside_geometry(sc, unit, origin, F3(1, 0, 0), cube, dist_px);
    inside_geometry(sc, unit, origin, F3(-1, 0, 0), cube, dist_nx);
    inside_geometry(sc, unit, origin, F3(0, 0, 1), cube, dist_pz);
    inside_geometry(sc, unit, origin, F3(0, 0, -1), cube, dist_nz);
    float min_dist = std::min({dist_py, dist_ny, dist_px, dist_nx, dist_pz, dist_nz});
    if (min_dist > 254.0f) {
        min_dist = std::min(min_dist, detect_distance(sc, unit, origin, F3(1, 1, 1)));
        min_dist = std::min(min_dist, detect_distance(sc, unit, origin, F3(-1, 1, 1)));
        min_dist = std::min(min_dist, detect_distance(sc, unit, origin, F3(-1, -1, 1)));
        min_dist = std::min(min_dist, detect_distance(sc, unit, origin, F3(-1, 1, 1)));
        min_dist = std::min(min_dist, detect_distance(sc, unit, origin, F3(1, 1, -1)));
        min_dist = std::min(min_dist, detect_distance(sc, unit, origin, F3(-1, 1, -1)));
        min_dist = std::min(min_dist, detect_distance(sc, unit, origin, F3(-1, -1, -1)));
        min_dist = std::min(min_dist, detect_distance(sc, unit, origin, F3(-1, 1, -1)));
    }
    auto fclamp = [](float x) { return std::fmin(std::fmax(x, 0.0f), 1.0f); };
    dist_py = fclamp(dist_py / unit);
    dist_ny = fclamp(dist_ny / unit);
    dist_px = fclamp(dist_px / unit);
    dist_nx = fclamp(dist_nx / unit);
    dist_pz = fclamp(dist_pz / unit);
    dist_nz = fclamp(dist_nz / unit);
    const float inside = dist_py * dist_ny * dist_px * dist_nx * dist_pz * dist_nz;
    const uint32_t b0[4] = {uint32_t(min_dist / unit), uint32_t(inside * 255.0f), uint32_t(dist_pz * 255.0f),
                            uint32_t(dist_nz * 255.0f)};
    const uint32_t b1[4] = {uint32_t(dist_px * 255.0f), uint32_t(dist_nx * 255.0f), uint32_t(dist_py * 255.0f),
                            uint32_t(dist_ny * 255.0f)};
    cube.dist0 = pack_bytes(b0);
    cube.dist1 = pack_bytes(b1);
}

struct LightABI {  // gk_light / LightObject
    float p0[4], p1[4], p3[4], normal_area[4];
    uint32_t light_mat_idx, reserved[3];
};

// FGpuProbeGenerator (AmbientCube.slang:472-690) over a read-only snapshot of the grid
struct ProbeBaker {
    const Scene& sc;
    const FrameParams& P;
    const LightABI* light;
    ProbeRef old;  // the grid as it was when the pass started

    // FaceTask (AmbientCube.slang:482-566)
    void face_task(F3 origin, F3 basis, uint32_t iterate, uint32_t& direct_light, uint32_t& indirect_light,
                   uint32_t& sky_vis_out, uint32_t& sun_vis_out) const {
        static const float grid3x3[9][2] = {{-0.667f, -0.667f}, {0.0f, -0.667f}, {0.667f, -0.667f},
                                            {-0.667f, 0.0f},    {0.0f, 0.0f},    {0.667f, 0.0f},
                                            {-0.667f, 0.667f},  {0.0f, 0.667f},  {0.667f, 0.667f}};
        static const float grid4x4[16][2] = {
            {-0.75f, -0.75f}, {-0.25f, -0.75f}, {0.25f, -0.75f}, {0.75f, -0.75f}, {-0.75f, -0.25f}, {-0.25f, -0.25f},
            {0.25f, -0.25f},  {0.75f, -0.25f},  {-0.75f, 0.25f}, {-0.25f, 0.25f}, {0.25f, 0.25f},   {0.75f, 0.25f},
            {-0.75f, 0.75f},  {-0.25f, 0.75f},  {0.25f, 0.75f},  {0.75f, 0.75f}};
        const float unit = old.grid.unit;
        origin = origin + (basis * unit) * 0.25f;
        F4 direct_color{{0, 0, 0, 0}}, bounce_color{{0, 0, 0, 0}};
        float sky_visibility = 0.0f;
        const float ox = grid3x3[iterate % 9][0] * 0.25f, oy = grid3x3[iterate % 9][1] * 0.25f;
        for (uint32_t i = 0; i < 16; i++) {
            const F3 hemi = normalize(F3(grid4x4[i][0] + ox, grid4x4[i][1] + oy, 1.0f));
            const F3 ray_dir = align_with_normal(hemi, basis);
            Hit h;
            if (trace_closest(sc, origin, ray_dir, EPS, 20.f, h)) {
                const Vtx v = hit_vertex(sc, h, origin, ray_dir);
                const Material& hm = sc.mats[v.mat];
                float tu, tv;
                hit_uv(sc, h, tu, tv);
                const F4 albedo = mat_albedo(sc, hm, tu, tv);
                const F4 li = interpolate_cubes(old, v.pos, v.nrm, false);
                F4 t = mul4(albedo, li);
                for (int k = 0; k < 4; ++k) t.v[k] = t.v[k] * 1.25f;
                bounce_color = add4(bounce_color, t);
            } else {
                if (P.has_sky) direct_color = add4(direct_color, sky_rough(sc, P, ray_dir));
                sky_visibility += 1.0f;
            }
        }
        for (int k = 0; k < 4; ++k) {
            direct_color.v[k] = direct_color.v[k] / 16.0f;
            bounce_color.v[k] = bounce_color.v[k] / 16.0f;
        }
        if (light) {
            const uint32_t nm = uint32_t(sc.mats.size());
            const Material& lm = sc.mats[light->light_mat_idx > nm - 1 ? nm - 1 : light->light_mat_idx];
            const F3 p1(light->p1[0], light->p1[1], light->p1[2]), p3(light->p3[0], light->p3[1], light->p3[2]);
            const F3 light_pos = p1 + (p3 - p1) * 0.5f;
            // TraceSegment(origin, lightPos, CUBE_UNIT * 0.5f), Shading.slang:759-782
            const float e = unit * 0.5f;
            const F3 seg = light_pos - origin;
            const float len = length(seg);
            const F3 sd(seg.x / len, seg.y / len, seg.z / len);
            const float light_atten = trace_any(sc, origin, sd, e, len - e) ? 0.0f : 1.0f;
            const F3 light_dir = normalize(light_pos - origin);
            const float ndotl = std::fmin(std::fmax(dot(basis, light_dir), 0.0f), 1.0f);
            const float distance = length(light_pos - origin);
            const float attenuation = ndotl * light->normal_area[3] / (distance * distance * 3.14159f);
            for (int k = 0; k < 4; ++k) direct_color.v[k] = direct_color.v[k] + lm.diffuse[k] * attenuation * light_atten;
        }
        if (P.has_sun) {
            const F3 sun_dir(P.sun_direction[0], P.sun_direction[1], P.sun_direction[2]);
            const float sun_atten = trace_any(sc, origin, sun_dir, EPS, PT_MAX_TRACE_DISTANCE) ? 0.0f : 1.0f;
            const float ndotl = std::fmin(std::fmax(dot(basis, sun_dir), 0.0f), 1.0f);
            sun_vis_out = sun_atten > 0.0f ? 1 : 0;
            for (int k = 0; k < 4; ++k)
                direct_color.v[k] = direct_color.v[k] + P.sun_color[k] * sun_atten * ndotl * 1.0f * 0.25f;
        }
        const float curr_weight = 0.125f;
        const float sv_old = float(sky_vis_out);
        sky_vis_out = uint32_t(sv_old + (255.0f * sky_visibility / 16.0f - sv_old) * curr_weight);
        direct_light = pack_rgb10a2(lerp4(unpack_rgb10a2(direct_light), direct_color, curr_weight));
        indirect_light = pack_rgb10a2(lerp4(unpack_rgb10a2(indirect_light), bounce_color, curr_weight));
    }

    // RenderAmbientCubeOnly (AmbientCube.slang:650-690) writing into out (a copy of the
    // probe's old cube) and vox
    void render(uint32_t idx, CubeO& out, VoxelO& vox) const {
        uint32_t up[4];
        unpack_bytes(vox.dist0, up);
        if (up[0] >= 8) return;
        const uint32_t iterate = vox.age;
        vox.age = vox.age + 1;
        uint32_t s0[4], s1[4];
        unpack_bytes(out.sky0, s0);
        unpack_bytes(out.sky1, s1);
        uint32_t sunvis = 0;
        const F3 origin = probe_origin(old.grid, idx);
        face_task(origin, F3(0, 1, 0), iterate, out.pos_y_d, out.pos_y, s0[2], s1[2]);
        face_task(origin, F3(0, -1, 0), iterate, out.neg_y_d, out.neg_y, s0[3], sunvis);
        face_task(origin, F3(1, 0, 0), iterate, out.pos_x_d, out.pos_x, s1[0], sunvis);
        face_task(origin, F3(-1, 0, 0), iterate, out.neg_x_d, out.neg_x, s1[1], sunvis);
        face_task(origin, F3(0, 0, 1), iterate, out.pos_z_d, out.pos_z, s0[0], sunvis);
        face_task(origin, F3(0, 0, -1), iterate, out.neg_z_d, out.neg_z, s0[1], sunvis);
        out.sky0 = pack_bytes(s0);
        out.sky1 = pack_bytes(s1);
    }
};

// ---------------------------------------------------------------- REF_V2 per pixel
struct RefOut {
    float diffuse[4], spec[4], albedo[4], normal[4];
    uint32_t objid;
    float motion[2];
    float depth;
    uint32_t rays;
};

struct RefV2 {
    const Scene& sc;
    const FrameParams& P;
    const std::vector<std::vector<float>>& combined;  // per instance, column-major

    bool trace(F3 O, F3 D, float tmax, Hit& h, uint32_t& rays) const {
        rays++;
        return trace_closest(sc, O, D, EPS, tmax, h);
    }
    bool occluded(F3 O, F3 D, uint32_t& rays) const {
        rays++;
        return trace_any(sc, O, D, EPS, PT_MAX_TRACE_DISTANCE);
    }

    // GetRayColor, Shading.slang:1003-1082
    bool get_ray_color(Vtx& v, F3& ray_dir, F4& color, Rng& seed, bool& hit_reflect, bool& hit_metal,
                       uint32_t& rays) const {
        const Material& mat = mat_at(sc, v.mat);
        const float start_off = mat.material_model == MAT_DIELECTRIC ? 0.0f : 1.0f;
        const float roughness = mat.fuzziness;
        const float dv = dot(ray_dir, v.nrm);
        const bool back = dv > 0.f;
        const F3 outward = back ? -v.nrm : v.nrm;
        const float ni_over_nt = back ? mat.refraction_index2 : (1.f / mat.refraction_index2);
        const float cosine = dv > 0.f ? mat.refraction_index * dv : -dv;
        const float reflect_prob = schlick(cosine, mat.refraction_index);
        const float metal_prob = mat.metalness;
        const bool chance_reflect = rnd1(seed) < reflect_prob;
        const bool chance_metal = rnd1(seed) < metal_prob;
        const bool chance_ggx = chance_reflect || chance_metal;
        const F3 trace_next = chance_ggx ? reflect(ray_dir, outward) : outward;
        F3 trace_dir = chance_ggx ? ggx_sampling(seed, std::sqrt(roughness), trace_next)
                                  : align_with_normal(random_hemisphere1(seed), trace_next);
        hit_reflect = chance_ggx;
        hit_metal = chance_metal;
        if (mat.material_model == MAT_DIELECTRIC && !chance_reflect)
            trace_dir = refract(ray_dir, outward, ni_over_nt);
        ray_dir = trace_dir;
        const F3 org = v.pos + (v.nrm * P.hit_normal_offset) * start_off;
        Hit h;
        if (trace(org, trace_dir, PT_MAX_TRACE_DISTANCE, h, rays)) {
            v = hit_vertex(sc, h, org, trace_dir);
            const Material& hm = mat_at(sc, v.mat);
            float tu, tv;
            hit_uv(sc, h, tu, tv);
            const F4 alb = mat_albedo(sc, hm, tu, tv);
            if (hm.material_model == MAT_DIFFUSE_LIGHT || !chance_reflect) color = mul4(color, alb);
            if (back && mat.material_model != MAT_DIELECTRIC) {
                color = F4{{0.5f, 0.5f, 0.5f, 0.f}};
                return true;
            }
            return hm.material_model == MAT_DIFFUSE_LIGHT;
        }
        const F4 sky = P.has_sky ? sky_radiance(sc, P, trace_dir) : F4{{0.f, 0.f, 0.f, 0.f}};
        color = mul4(color, sky);
        return true;
    }

    void pixel(uint32_t x, uint32_t y, uint32_t W, uint32_t H, uint16_t& moment, RefOut& out) const {
        Rng seed{{x, y, P.total_frames, P.global_seed}};
        out = RefOut{};
        uint32_t rays = 0;
        F3 O, D;
        primary_ray_lens(P, x, y, W, H, seed, O, D);
        Hit h;
        rays++;
        if (!trace_closest(sc, O, D, 0.0f, 1e30f, h)) {
            const F4 sky = P.has_sky ? sky_radiance(sc, P, D) : F4{{0.f, 0.f, 0.f, 0.f}};
            for (int i = 0; i < 4; ++i) out.diffuse[i] = sky.v[i];
            out.albedo[0] = out.albedo[1] = out.albedo[2] = out.albedo[3] = 1.f;
            out.normal[1] = 1.f;
            out.normal[3] = 1.f;
            out.objid = 65535u;
            out.depth = 1e5f;
            out.rays = rays;
            return;
        }
        Vtx pv = primary_vertex(sc, h, O, D);
        const Instance& I = sc.inst[h.inst];
        const Material& mat = mat_at(sc, pv.mat);
        F4 palb;
        float prough;
        fetch_gbuffer(sc, h, mat, pv.nrm, palb, prough);
        // PrimaryHit, Shading.slang:1204-1220
        {
            const float p4[4] = {pv.pos.x, pv.pos.y, pv.pos.z, 1.f};
            float clip[4];
            mat_mul4(P.view_projection, p4, clip);
            out.depth = clip[2] / clip[3];
            float ch[4], ph[4], pp[4], m[16];
            mat_mul4(P.view_projection_unjit, p4, ch);
            const float* comb = combined[h.inst].data();
            mat_mul44(P.prev_view_projection_unjit, comb, m);
            mat_mul4(m, p4, ph);
            const float cfx = ch[0] / ch[3] * 0.5f, cfy = ch[1] / ch[3] * 0.5f;
            const float pfx = ph[0] / ph[3] * 0.5f, pfy = ph[1] / ph[3] * 0.5f;
            mat_mul4(comb, p4, pp);
            const F3 dpos = pv.pos - F3(pp[0], pp[1], pp[2]);
            if (length(dpos) > 0.02f) moment = uint16_t(P.temporal_frames);
            else moment = moment > 0 ? uint16_t(moment - 1) : 0;
            out.motion[0] = (pfx - cfx) * float(W);
            out.motion[1] = (pfy - cfy) * float(H);
            for (int i = 0; i < 4; ++i) out.albedo[i] = palb.v[i];
            out.normal[0] = pv.nrm.x; out.normal[1] = pv.nrm.y; out.normal[2] = pv.nrm.z;
            out.normal[3] = prough;
            out.objid = I.instance_id;
            if (mat.material_model != MAT_DIELECTRIC) pv.pos = pv.pos - D * 0.03f;
        }
        const F4 primary_albedo = palb;
        // Render, Shading.slang:1085-1184
        if (mat.material_model == MAT_DIFFUSE_LIGHT) {
            for (int i = 0; i < 4; ++i) out.diffuse[i] = mat.diffuse[i];
            out.rays = rays;
            return;
        }
        const uint32_t mult = moment > 0 ? 4u : 1u;
        const uint32_t nsamp = std::max(1u, P.samples * mult);
        F4 fc{{0, 0, 0, 0}}, fr{{0, 0, 0, 0}};
        const float o4[4] = {0.f, 0.f, 0.f, 1.f};
        float eye4[4];
        mat_mul4(P.model_view_inverse, o4, eye4);
        const F3 eye(eye4[0], eye4[1], eye4[2]);
        const float cos_cone = float(std::cos(0.25 / 180.0 * 3.14159265358979323846));
        const F3 sun_dir(P.sun_direction[0], P.sun_direction[1], P.sun_direction[2]);
        const F4 sun_col{{P.sun_color[0], P.sun_color[1], P.sun_color[2], P.sun_color[3]}};
        for (uint32_t s = 0; s < nsamp; ++s) {
            F4 color{{1, 1, 1, 1}};
            F3 dir = normalize(pv.pos - eye);
            Vtx vs = pv;
            const uint32_t maxb = mat.material_model == MAT_DIELECTRIC ? P.max_bounces : P.bounces;
            bool hit_reflect = false, hit_metal = false;
            const bool exit_first = get_ray_color(vs, dir, color, seed, hit_reflect, hit_metal, rays);
            if (!exit_first && (P.flags & 256u) && mat.material_model != MAT_DIELECTRIC) {
                // ExitAfterFirst (GK_FRAME_FAST_GATHER), Shading.slang:1118-1124
                color = mul4(color, interpolate_cubes(sc.probes, vs.pos, vs.nrm, true));
            } else if (!exit_first) {
                for (uint32_t b = 1; b < maxb; ++b) {
                    bool dc = false;
                    if (get_ray_color(vs, dir, color, seed, dc, dc, rays)) break;
                    if (P.has_sun && rnd1(seed) < 0.5f) {
                        const F3 cone = align_with_normal(random_in_cone(seed, cos_cone), sun_dir);
                        if (!occluded(vs.pos + vs.nrm * 0.0f, cone, rays)) {
                            color = mul4(color, sun_col);
                            break;
                        }
                    }
                    const bool early = mat.material_model != MAT_DIELECTRIC && rnd1(seed) < P.exit_probability;
                    if (b == maxb - 1 || early) {
                        color = mul4(color, interpolate_cubes(sc.probes, vs.pos, vs.nrm, true));
                        break;
                    }
                }
            }
            if (hit_metal) color = mul4(color, primary_albedo);
            F4& acc = hit_reflect ? fr : fc;
            for (int i = 0; i < 4; ++i) acc.v[i] = acc.v[i] + color.v[i];
        }
        for (int i = 0; i < 4; ++i) { fc.v[i] = fc.v[i] / float(nsamp); fr.v[i] = fr.v[i] / float(nsamp); }
        // FHardwareDirectIlluminatorV2, Shading.slang:888-905
        {
            const float d = std::max(dot(sun_dir, normalize(pv.nrm)), 0.0f) * (1.0f / PI_F);
            float shadow = 0.f;
            if (P.has_sun) {
                const F3 cone = align_with_normal(random_in_cone(seed, cos_cone), sun_dir);
                shadow = 1.f;
                if (occluded(pv.pos, cone, rays)) shadow = 0.f;
            }
            for (int i = 0; i < 4; ++i) fc.v[i] = fc.v[i] + sun_col.v[i] * d * shadow;
        }
        for (int i = 0; i < 4; ++i) { out.diffuse[i] = fc.v[i]; out.spec[i] = fr.v[i]; }
        out.rays = rays;
    }
};

// ---------------------------------------------------------------- PT_MIS (templated)
// mode: which strategies carry light to a vertex (test instrumentation of the oracle):
//   0  MIS: next-event estimation and BSDF sampling, power-heuristic weights (the estimator)
//   1  NEE only: emitters reached by a non-delta BSDF sample add nothing
//   2  BSDF only: shadow rays are drawn but add nothing; emitter hits carry weight 1
//   3-5 planted mistakes, for tests that show their checks fail on them: the emitter-hit
//      weight uses a light pdf without the emitter's cosine (3), without the light
//      selection probability (4), or the balance instead of the power heuristic (5), so
//      the two weights of a direction no longer sum to one
template <class T>
struct PtMis {
    const Scene& sc;
    const FrameParams& P;
    int mode = 0;
    using V = V3<T>;
    static constexpr T PI = T(3.14159265358979323846);

    static V cvt(F3 a) { return V(T(a.x), T(a.y), T(a.z)); }

    T rnd(Rng& s) const { return T(rnd1(s)); }
    void rnd2t(Rng& s, T& a, T& b) const { float x, y; rnd2(s, x, y); a = T(x); b = T(y); }

    static void onb_t(V n, V& b1, V& b2) {
        const T sz = n.z < T(0) ? T(-1) : T(1);
        const T a = T(-1) / (sz + n.z);
        b2 = V(n.x * n.y * a, sz + n.y * n.y * a, -n.y);
        b1 = V(T(1) + sz * n.x * n.x * a, sz * b2.x, -sz * n.x);
    }
    static T sat(T x) { return x < T(0) ? T(0) : (x > T(1) ? T(1) : x); }
    static T ggx_d(T a2, T cz) {
        const T d = cz * cz * (a2 - T(1)) + T(1);
        return a2 / (PI * d * d);
    }
    static T ggx_lambda(T a2, V w) {
        const T z2 = w.z * w.z;
        if (z2 <= T(0)) return T(0);
        const T t2 = (w.x * w.x + w.y * w.y) / z2;
        return (std::sqrt(T(1) + a2 * t2) - T(1)) * T(0.5);
    }
    // bounded VNDF sample (Eto & Tokuyoshi 2023), isotropic alpha, wo.z > 0
    static V vndf_sample(T alpha, V wo, T ux, T uy) {
        const V wi = normalize(V(wo.x * alpha, wo.y * alpha, wo.z));
        T b = wi.z;
        const T a = sat(alpha);
        const T s = T(1) + std::sqrt(wo.x * wo.x + wo.y * wo.y);
        const T awiz = a * wo.z / s;
        b = b * ((T(1) - a * a) / (T(1) + awiz * awiz));
        const T z = (T(1) - uy) * (T(1) + b) - b;
        T sp, cp;
        sincos_2pi<T>(ux, sp, cp);
        const T r = std::sqrt(sat(T(1) - z * z));
        const V m = wi + V(r * cp, r * sp, z);
        return normalize(V(m.x * alpha, m.y * alpha, m.z));
    }
    // pdf of the reflected direction for the bounded VNDF sampler (Eto & Tokuyoshi 2023, eq. 8)
    static T vndf_pdf(T alpha, V wo, V m) {
        const T a2 = alpha * alpha;
        const T D = ggx_d(a2, m.z);
        const T len2 = a2 * (wo.x * wo.x + wo.y * wo.y);
        const T t = std::sqrt(len2 + wo.z * wo.z);
        const T a = sat(alpha);
        const T s = T(1) + std::sqrt(wo.x * wo.x + wo.y * wo.y);
        const T s2 = s * s, aa = a * a;
        const T k = (T(1) - aa) * s2 / (s2 + aa * wo.z * wo.z);
        return D / (T(2) * (k * wo.z + t));
    }

    struct Bsdf {
        V albedo;
        T metal, alpha, ps;
    };
    Bsdf bsdf_of(const Material& m) const {
        Bsdf b;
        b.albedo = V(T(m.diffuse[0]), T(m.diffuse[1]), T(m.diffuse[2]));
        const T r = std::max(T(m.fuzziness), T(0.02));
        b.alpha = r * r;
        if (m.material_model == MAT_METALLIC) b.metal = T(1);
        else if (m.material_model == MAT_MIXTURE) b.metal = T(m.metalness);
        else b.metal = T(0);
        b.ps = m.material_model == MAT_LAMBERTIAN || m.material_model == MAT_ISOTROPIC
                   ? T(0) : T(0.5) + T(0.5) * b.metal;
        return b;
    }
    // f_d and f_s for local wo, wi (z = shading normal); returns cos-free BSDF values
    void eval(const Bsdf& b, V wo, V wi, V& fd, V& fs, T& pdf_d, T& pdf_s) const {
        fd = b.albedo * ((T(1) - b.metal) / PI);
        pdf_d = wi.z / PI;
        fs = V(T(0), T(0), T(0));
        pdf_s = T(0);
        if (b.ps <= T(0)) return;
        V h = wo + wi;
        const T hl = length(h);
        if (hl <= T(0)) return;
        h = h * (T(1) / hl);
        const T a2 = b.alpha * b.alpha;
        const T D = ggx_d(a2, h.z);
        const T G = T(1) / (T(1) + ggx_lambda(a2, wo) + ggx_lambda(a2, wi));
        const T c = std::max(T(0), dot(wo, h));
        const T x = T(1) - c;
        const T x5 = x * x * x * x * x;
        const V f0 = V(T(0.04), T(0.04), T(0.04)) * (T(1) - b.metal) + b.albedo * b.metal;
        const V F = f0 + (V(T(1), T(1), T(1)) - f0) * x5;
        fs = F * (D * G / (T(4) * wo.z * wi.z));
        pdf_s = vndf_pdf(b.alpha, wo, h);
    }
    static T power(T a, T b) { const T a2 = a * a, b2 = b * b; return a2 + b2 > T(0) ? a2 / (a2 + b2) : T(0); }

    // radiance of one path; rgb accumulated into Ld (primary diffuse lobe) and Ls (rest)
    void path(uint32_t x, uint32_t y, uint32_t W, uint32_t H, Rng& seed, V& Ld, V& Ls, uint64_t& rays) const {
        F3 Of, Df;
        primary_ray_lens(P, x, y, W, H, seed, Of, Df);
        V O = cvt(Of), D = cvt(Df);
        V thr(T(1), T(1), T(1));
        int lobe0 = -1;          // 0 diffuse, 1 specular/other at the primary vertex
        T last_pdf = T(0);
        bool last_delta = true;  // camera ray: no MIS
        V last_pos = O;
        const uint32_t max_depth = P.bounces;
        for (uint32_t depth = 0;; ++depth) {
            HitT<T> h;
            rays++;
            const T tmin = T(0);
            if (!trace_closest_t<T>(sc, O, D, tmin, T(1e30), h)) {
                if (P.has_sky) {
                    const F4 s4 = sky_radiance(sc, P, F3(float(D.x), float(D.y), float(D.z)));
                    const V sky(T(s4.v[0]), T(s4.v[1]), T(s4.v[2]));
                    (lobe0 == 0 ? Ld : Ls) = (lobe0 == 0 ? Ld : Ls) + thr * sky;
                }
                break;
            }
            const Instance& I = sc.inst[h.inst];
            const Mesh& m = sc.meshes[I.mesh_id];
            const uint32_t mi = material_of(sc, h.inst, h.prim);
            Material mat = sc.mats[mi];
            // textures: albedo and roughness at every vertex, the normal map at the primary hit
            bool mapped = false;
            F3 mapped_n;
            if (mat.material_model != MAT_DIFFUSE_LIGHT &&
                (has_tex(sc, mat.diffuse_texture_id) || has_tex(sc, mat.mra_texture_id) ||
                 (depth == 0 && has_tex(sc, mat.normal_texture_id)))) {
                const Hit hf{float(h.t), float(h.u), float(h.v), h.prim, h.inst};
                F4 alb;
                float rough;
                if (depth == 0) {
                    fetch_gbuffer(sc, hf, mat, mapped_n, alb, rough);
                    mapped = has_tex(sc, mat.normal_texture_id);
                } else {
                    float tu, tv;
                    hit_uv(sc, hf, tu, tv);
                    alb = mat_albedo(sc, mat, tu, tv);
                    rough = mat.fuzziness;
                    if (has_tex(sc, mat.mra_texture_id)) rough = rough * tex_sample(sc, uint32_t(mat.mra_texture_id), tu, tv).v[1];
                }
                for (int k = 0; k < 4; ++k) mat.diffuse[k] = alb.v[k];
                mat.fuzziness = rough;
            }
            const V P_ = O + D * h.t;
            // world normals: geometric (inverse transpose of e1 x e2) and shading
            const float* Wi = I.inv;
            auto nrm_w = [&](F3 n) {
                V o;
                for (int j = 0; j < 3; ++j)
                    o[j] = T(Wi[0 * 4 + j]) * T(n.x) + T(Wi[1 * 4 + j]) * T(n.y) + T(Wi[2 * 4 + j]) * T(n.z);
                return normalize(o);
            };
            V Ng = nrm_w(cross(m.e1[h.prim], m.e2[h.prim]));
            const F3 n0 = m.n0[h.prim], n1 = m.n1[h.prim], n2 = m.n2[h.prim];
            const T bu = h.u, bv = h.v, bw = T(1) - bu - bv;
            V nsum = cvt(n0) * bw + cvt(n1) * bu + cvt(n2) * bv;
            const float* Wr = I.inv;
            V Ns;
            for (int j = 0; j < 3; ++j)
                Ns[j] = T(Wr[0 * 4 + j]) * nsum.x + T(Wr[1 * 4 + j]) * nsum.y + T(Wr[2 * 4 + j]) * nsum.z;
            Ns = normalize(Ns);
            if (mapped) Ns = cvt(mapped_n);
            if (dot(Ng, D) > T(0)) Ng = -Ng;
            if (dot(Ns, Ng) < T(0)) Ns = -Ns;
            if (mat.material_model == MAT_DIFFUSE_LIGHT) {
                const V Le(T(mat.diffuse[0]), T(mat.diffuse[1]), T(mat.diffuse[2]));
                T w = T(1);
                if (!last_delta && sc.light_total > 0.0) {
                    const T cosl = std::fabs(dot(Ng, D));
                    const T area = T(0.5) * length(nrm_raw(m, h.prim, I));
                    const T psel = light_select_pdf(h.inst, h.prim);
                    const T d2 = h.t * h.t * dot(D, D);
                    const T pl = cosl > T(0) && area > T(0) ? psel * d2 / (cosl * area) : T(0);
                    w = power(last_pdf, pl);
                    if (mode == 1) w = T(0);
                    if (mode == 2) w = T(1);
                    if (mode == 3) w = power(last_pdf, area > T(0) ? psel * d2 / area : T(0));
                    if (mode == 4) w = power(last_pdf, pl / psel);
                    if (mode == 5) w = last_pdf / (last_pdf + pl);
                }
                if (depth == 0) Ls = Ls + Le;  // emitter seen directly: specular target
                else (lobe0 == 0 ? Ld : Ls) = (lobe0 == 0 ? Ld : Ls) + thr * Le * w;
                break;
            }
            if (depth >= max_depth) break;
            const T scale = T(1) + std::max(std::fabs(P_.x), std::max(std::fabs(P_.y), std::fabs(P_.z)));
            const T off = T(1e-4) * scale;
            if (mat.material_model == MAT_DIELECTRIC) {
                const T ior = T(mat.refraction_index);
                const T cosi = -dot(D, Ng);
                const bool entering = dot(D, cvt_face(m, h.prim, I)) < T(0);
                const T eta = entering ? T(1) / ior : ior;
                T r0 = (T(1) - ior) / (T(1) + ior);
                r0 = r0 * r0;
                const T xx = T(1) - cosi;
                const T F = r0 + (T(1) - r0) * xx * xx * xx * xx * xx;
                const T k = T(1) - eta * eta * (T(1) - cosi * cosi);
                const T u = rnd(seed);
                V wi;
                bool refl = k < T(0) || u < F;
                if (refl) wi = D + Ng * (T(2) * cosi);
                else wi = D * eta + Ng * (eta * cosi - std::sqrt(k));
                wi = normalize(wi);
                if (depth == 0) lobe0 = 1;
                thr = thr * V(T(mat.diffuse[0]), T(mat.diffuse[1]), T(mat.diffuse[2]));
                last_delta = true;
                O = P_ + (refl ? Ng : -Ng) * off;
                D = wi;
            } else {
                const Bsdf b = bsdf_of(mat);
                V T1, B1;
                onb_t(Ns, T1, B1);
                auto to_loc = [&](V v) { return V(dot(T1, v), dot(B1, v), dot(Ns, v)); };
                auto to_wld = [&](V v) { return T1 * v.x + B1 * v.y + Ns * v.z; };
                const V wo = to_loc(-D);
                if (wo.z <= T(0)) break;
                // next-event estimation
                if (sc.light_total > 0.0) {
                    const T ul = rnd(seed);
                    T s1, s2;
                    rnd2t(seed, s1, s2);
                    uint32_t li = light_pick(ul);
                    const Light& L = sc.lights[li];
                    const T sq = std::sqrt(s1);
                    const T bu2 = T(1) - sq, bv2 = s2 * sq;
                    const V Pl = cvt(L.p0) + cvt(L.e1) * bu2 + cvt(L.e2) * bv2;
                    const V Po = P_ + Ng * off;
                    V dl = Pl - Po;
                    const T d2 = dot(dl, dl);
                    const T dist = std::sqrt(d2);
                    const V wl = dl * (T(1) / dist);
                    const T cosl = std::fabs(dot(cvt(L.n), wl));
                    const V wil = to_loc(wl);
                    const T psel = T(sc.light_cdf[li + 1] - sc.light_cdf[li]);
                    if (cosl > T(0) && wil.z > T(0) && dot(Ng, wl) > T(0) && psel > T(0)) {
                        const T pl = psel * d2 / (cosl * T(L.area));
                        rays++;
                        if (!trace_any_t<T>(sc, Po, wl, T(0), dist * (T(1) - T(1e-4)) - off)) {
                            V fd, fs;
                            T pdd, pds;
                            eval(b, wo, wil, fd, fs, pdd, pds);
                            const V Le = cvt(L.le);
                            T wd = power(pl, (T(1) - b.ps) * pdd);
                            T ws = power(pl, b.ps * pds);
                            if (mode == 1) wd = ws = T(1);
                            if (mode == 2) wd = ws = T(0);
                            const V cd = thr * Le * fd * (wil.z * wd / pl);
                            const V cs = thr * Le * fs * (wil.z * ws / pl);
                            if (depth == 0) { Ld = Ld + cd; Ls = Ls + cs; }
                            else { V& acc = lobe0 == 0 ? Ld : Ls; acc = acc + cd + cs; }
                        }
                    }
                }
                // BSDF sampling
                const T ulobe = rnd(seed);
                T u1, u2;
                rnd2t(seed, u1, u2);
                const bool spec = ulobe < b.ps;
                V wi;
                if (spec) {
                    const V mh = vndf_sample(b.alpha, wo, u1, u2);
                    wi = mh * (T(2) * dot(wo, mh)) - wo;
                } else {
                    T sp, cp;
                    sincos_2pi<T>(u1, sp, cp);
                    const T r = std::sqrt(u2);
                    wi = V(r * cp, r * sp, std::sqrt(T(1) - u2));
                }
                if (wi.z <= T(0)) break;
                V fd, fs;
                T pdd, pds;
                eval(b, wo, wi, fd, fs, pdd, pds);
                T pdf;
                V f;
                if (spec) { pdf = b.ps * pds; f = fs; }
                else { pdf = (T(1) - b.ps) * pdd; f = fd; }
                if (!(pdf > T(0))) break;
                thr = thr * f * (wi.z / pdf);
                if (depth == 0) lobe0 = spec ? 1 : 0;
                last_pdf = pdf;
                last_delta = false;
                const V wiw = normalize(to_wld(wi));
                O = P_ + (dot(wiw, Ng) > T(0) ? Ng : -Ng) * off;
                D = wiw;
            }
            last_pos = P_;
            if (depth >= 3) {
                const T q = std::min(T(0.95), std::max(thr.x, std::max(thr.y, thr.z)));
                if (rnd(seed) >= q) break;
                thr = thr * (T(1) / q);
            }
        }
        (void)last_pos;
    }

    V nrm_raw(const Mesh& m, uint32_t p, const Instance& I) const {
        // world-space e1 x e2 (for the emitter area)
        const F3 a = xform_vector(I.world, m.e1[p]), c = xform_vector(I.world, m.e2[p]);
        return cvt(cross(a, c));
    }
    V cvt_face(const Mesh& m, uint32_t p, const Instance& I) const { return nrm_raw(m, p, I); }
    T light_select_pdf(uint32_t inst, uint32_t prim) const {
        // lights are sorted by (inst, prim); binary search
        size_t lo = 0, hi = sc.lights.size();
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            const Light& L = sc.lights[mid];
            if (L.inst < inst || (L.inst == inst && L.prim < prim)) lo = mid + 1; else hi = mid;
        }
        if (lo >= sc.lights.size() || sc.lights[lo].inst != inst || sc.lights[lo].prim != prim) return T(0);
        return T(sc.light_cdf[lo + 1] - sc.light_cdf[lo]);
    }
    uint32_t light_pick(T u) const {
        // first i with cdf[i+1] > u
        size_t lo = 0, hi = sc.lights.size() - 1;
        while (lo < hi) {
            const size_t mid = (lo + hi) / 2;
            if (T(sc.light_cdf[mid + 1]) > u) hi = mid; else lo = mid + 1;
        }
        return uint32_t(lo);
    }
};

// Differentiable PT_MIS without Russian roulette, in double, with detached sampling:
// lobe choices, directions, pdfs and MIS weights come from the scene's materials (the
// sampling set); BSDF values come from `eval` (the evaluation set). Central differences
// of this function with respect to `eval` are the detached-sampling gradient that the
// HIP adjoint computes.
struct DiffPath {
    const Scene& sc;
    const FrameParams& P;
    const std::vector<Material>& eval;
    using T = double;
    using V = D3;
    using PT = PtMis<double>;

    V path(uint32_t x, uint32_t y, uint32_t W, uint32_t H, Rng& seed) const {
        PT pt{sc, P};
        F3 Of, Df;
        float jx, jy;  // sub-pixel position of this sample
        rnd2(seed, jx, jy);
        primary_ray_at(P, float(x) + jx, float(y) + jy, W, H, Of, Df);
        V O = PT::cvt(Of), D = PT::cvt(Df);
        V thr(1, 1, 1), L(0, 0, 0);
        T last_pdf = 0;
        bool last_delta = true;
        for (uint32_t depth = 0;; ++depth) {
            HitT<T> h;
            if (!trace_closest_t<T>(sc, O, D, T(0), T(1e30), h)) {
                if (P.has_sky) {
                    const F4 s4 = sky_radiance(sc, P, F3(float(D.x), float(D.y), float(D.z)));
                    L = L + thr * V(s4.v[0], s4.v[1], s4.v[2]);
                }
                break;
            }
            const Instance& I = sc.inst[h.inst];
            const Mesh& m = sc.meshes[I.mesh_id];
            const uint32_t mi = material_of(sc, h.inst, h.prim);
            const Material& ms = sc.mats[mi];
            const Material& me = eval[mi];
            const V Pp = O + D * h.t;
            auto nrm_w = [&](V n) {
                V o;
                for (int j = 0; j < 3; ++j)
                    o[j] = T(I.inv[0 * 4 + j]) * n.x + T(I.inv[1 * 4 + j]) * n.y + T(I.inv[2 * 4 + j]) * n.z;
                return normalize(o);
            };
            V Ng = nrm_w(PT::cvt(cross(m.e1[h.prim], m.e2[h.prim])));
            const T bu = h.u, bv = h.v, bw = T(1) - bu - bv;
            V Ns = nrm_w(PT::cvt(m.n0[h.prim]) * bw + PT::cvt(m.n1[h.prim]) * bu + PT::cvt(m.n2[h.prim]) * bv);
            if (dot(Ng, D) > 0) Ng = -Ng;
            if (dot(Ns, Ng) < 0) Ns = -Ns;
            if (ms.material_model == MAT_DIFFUSE_LIGHT) {
                const V Le(ms.diffuse[0], ms.diffuse[1], ms.diffuse[2]);  // emission is not a parameter
                T w = 1;
                if (!last_delta && sc.light_total > 0.0) {
                    const T cosl = std::fabs(dot(Ng, D));
                    const T area = T(0.5) * length(pt.nrm_raw(m, h.prim, I));
                    const T psel = pt.light_select_pdf(h.inst, h.prim);
                    const T pl = cosl > 0 && area > 0 ? psel * h.t * h.t / (cosl * area) : T(0);
                    w = PT::power(last_pdf, pl);
                }
                L = L + (depth == 0 ? Le : thr * Le * w);
                break;
            }
            if (depth >= P.bounces) break;
            const T off = T(1e-4) * (T(1) + std::max(std::fabs(Pp.x), std::max(std::fabs(Pp.y), std::fabs(Pp.z))));
            const typename PT::Bsdf bs = pt.bsdf_of(ms), be = pt.bsdf_of(me);
            V T1, B1;
            PT::onb_t(Ns, T1, B1);
            auto to_loc = [&](V v) { return V(dot(T1, v), dot(B1, v), dot(Ns, v)); };
            const V wo = to_loc(-D);
            if (wo.z <= 0) break;
            if (sc.light_total > 0.0) {
                const T ul = pt.rnd(seed);
                T s1, s2;
                pt.rnd2t(seed, s1, s2);
                const uint32_t li = pt.light_pick(ul);
                const Light& Lt = sc.lights[li];
                const T sq = std::sqrt(s1);
                const V Pl = PT::cvt(Lt.p0) + PT::cvt(Lt.e1) * (T(1) - sq) + PT::cvt(Lt.e2) * (s2 * sq);
                const V Po = Pp + Ng * off;
                const V dl = Pl - Po;
                const T d2 = dot(dl, dl), dist = std::sqrt(d2);
                const V wl = dl * (T(1) / dist);
                const T cosl = std::fabs(dot(PT::cvt(Lt.n), wl));
                const V wil = to_loc(wl);
                const T psel = T(sc.light_cdf[li + 1] - sc.light_cdf[li]);
                if (cosl > 0 && wil.z > 0 && dot(Ng, wl) > 0 && psel > 0) {
                    const T pl = psel * d2 / (cosl * T(Lt.area));
                    if (!trace_any_t<T>(sc, Po, wl, T(0), dist * (T(1) - T(1e-4)) - off)) {
                        V fd, fs, fd_s, fs_s;
                        T pdd, pds;
                        pt.eval(bs, wo, wil, fd_s, fs_s, pdd, pds);   // pdfs of the sampling set
                        T pdd2, pds2;
                        pt.eval(be, wo, wil, fd, fs, pdd2, pds2);     // values of the evaluation set
                        const T wd = PT::power(pl, (T(1) - bs.ps) * pdd);
                        const T ws = PT::power(pl, bs.ps * pds);
                        L = L + thr * PT::cvt(Lt.le) * (fd * wd + fs * ws) * (wil.z / pl);
                    }
                }
            }
            const T ulobe = pt.rnd(seed);
            T u1, u2;
            pt.rnd2t(seed, u1, u2);
            const bool spec = ulobe < bs.ps;
            V wi;
            if (spec) {
                const V mh = PT::vndf_sample(bs.alpha, wo, u1, u2);
                wi = mh * (T(2) * dot(wo, mh)) - wo;
            } else {
                T sp, cp;
                sincos_2pi<T>(u1, sp, cp);
                const T r = std::sqrt(u2);
                wi = V(r * cp, r * sp, std::sqrt(T(1) - u2));
            }
            if (wi.z <= 0) break;
            V fd, fs, fde, fse;
            T pdd, pds, a, b;
            pt.eval(bs, wo, wi, fd, fs, pdd, pds);
            pt.eval(be, wo, wi, fde, fse, a, b);
            const T pdf = spec ? bs.ps * pds : (T(1) - bs.ps) * pdd;
            if (!(pdf > 0)) break;
            thr = thr * (spec ? fse : fde) * (wi.z / pdf);
            last_pdf = pdf;
            last_delta = false;
            const V wiw = normalize(T1 * wi.x + B1 * wi.y + Ns * wi.z);
            O = Pp + (dot(wiw, Ng) > 0 ? Ng : -Ng) * off;
            D = wiw;
        }
        return L;
    }
};

void build_lights(Scene& sc) {
    sc.lights.clear();
    for (uint32_t ii = 0; ii < sc.inst.size(); ++ii) {
        const Instance& I = sc.inst[ii];
        if (!I.visible || I.nort) continue;
        const Mesh& m = sc.meshes[I.mesh_id];
        for (uint32_t p = 0; p < m.v0.size(); ++p) {
            const uint32_t mi = material_of(sc, ii, p);
            const Material& mat = sc.mats[mi];
            if (mat.material_model != MAT_DIFFUSE_LIGHT) continue;
            Light L;
            L.p0 = xform_point(I.world, m.v0[p]);
            L.e1 = xform_vector(I.world, m.e1[p]);
            L.e2 = xform_vector(I.world, m.e2[p]);
            const F3 c = cross(L.e1, L.e2);
            L.area = 0.5f * length(c);
            L.n = normalize(c);
            L.le = F3(mat.diffuse[0], mat.diffuse[1], mat.diffuse[2]);
            L.inst = ii;
            L.prim = p;
            if (L.area > 0.f) sc.lights.push_back(L);
        }
    }
    sc.light_cdf.assign(sc.lights.size() + 1, 0.0);
    double acc = 0.0;
    for (size_t i = 0; i < sc.lights.size(); ++i) {
        const Light& L = sc.lights[i];
        const double lum = 0.2126 * L.le.x + 0.7152 * L.le.y + 0.0722 * L.le.z;
        acc += double(L.area) * lum;
        sc.light_cdf[i + 1] = acc;
    }
    sc.light_total = acc;
    if (acc > 0.0) for (auto& c : sc.light_cdf) c /= acc;
}

}  // namespace

// =================================================================== C API
extern "C" {

void* oracle_scene_new() { return new Scene(); }
void oracle_scene_free(void* s) { delete static_cast<Scene*>(s); }

// positions/normals: 3*nv floats (normals may be NULL: face normals); idx: 3*nt; tri_mat: nt or NULL
int oracle_scene_add_mesh(void* s, const float* pos, const float* nrm, const uint32_t* idx,
                          const uint32_t* tri_mat, uint32_t nv, uint32_t nt) {
    Scene& sc = *static_cast<Scene*>(s);
    Mesh m;
    m.v0.resize(nt); m.e1.resize(nt); m.e2.resize(nt);
    m.n0.resize(nt); m.n1.resize(nt); m.n2.resize(nt); m.mat.resize(nt);
    for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
        if (a >= nv || b >= nv || c >= nv) return -1;
        const F3 pa(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]);
        const F3 pb(pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]);
        const F3 pc(pos[3 * c], pos[3 * c + 1], pos[3 * c + 2]);
        m.v0[t] = pa;
        m.e1[t] = pb - pa;
        m.e2[t] = pc - pa;
        if (nrm) {
            m.n0[t] = F3(nrm[3 * a], nrm[3 * a + 1], nrm[3 * a + 2]);
            m.n1[t] = F3(nrm[3 * b], nrm[3 * b + 1], nrm[3 * b + 2]);
            m.n2[t] = F3(nrm[3 * c], nrm[3 * c + 1], nrm[3 * c + 2]);
        } else {
            const F3 fn = normalize(cross(m.e1[t], m.e2[t]));
            m.n0[t] = m.n1[t] = m.n2[t] = fn;
        }
        m.mat[t] = tri_mat ? tri_mat[t] : 0u;
    }
    sc.meshes.push_back(std::move(m));
    return int(sc.meshes.size() - 1);
}

// per-vertex texture coordinates (2 floats) and tangents (4 floats) of mesh `mesh`
// (either may be NULL), gathered per triangle through idx
int oracle_scene_set_mesh_attribs(void* s, uint32_t mesh, const uint32_t* idx, uint32_t nv, uint32_t nt,
                                  const float* uv, const float* tan) {
    Scene& sc = *static_cast<Scene*>(s);
    if (mesh >= sc.meshes.size()) return -1;
    Mesh& m = sc.meshes[mesh];
    m.uv.assign(uv ? 6 * size_t(nt) : 0, 0.f);
    m.tan.assign(tan ? 12 * size_t(nt) : 0, 0.f);
    for (uint32_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k) {
            const uint32_t a = idx[3 * t + k];
            if (a >= nv) return -1;
            if (uv) { m.uv[6 * t + 2 * k] = uv[2 * a]; m.uv[6 * t + 2 * k + 1] = uv[2 * a + 1]; }
            if (tan)
                for (int j = 0; j < 4; ++j) m.tan[12 * t + 4 * k + j] = tan[4 * a + j];
        }
    return 0;
}

// tex_sample at n coordinates (uv: 2 floats each, out: 4 floats each)
void oracle_tex_sample(void* s, uint32_t id, const float* uv, uint32_t n, float* out) {
    const Scene& sc = *static_cast<Scene*>(s);
    for (uint32_t i = 0; i < n; ++i) {
        const F4 r = tex_sample(sc, id, uv[2 * i], uv[2 * i + 1]);
        std::memcpy(out + 4 * i, r.v, 16);
    }
}

// the sky image (rgba: w * h * 4 floats; w or h 0 removes it) and its SH coefficients
int oracle_scene_set_sky(void* s, const float* rgba, uint32_t w, uint32_t h) {
    Scene& sc = *static_cast<Scene*>(s);
    if (w == 0 || h == 0) {
        sc.sky.clear();
        sc.sky_w = sc.sky_h = 0;
        std::fill(sc.sky_sh, sc.sky_sh + 27, 0.f);
        return 0;
    }
    sc.sky.assign(rgba, rgba + 4 * size_t(w) * h);
    sc.sky_w = w;
    sc.sky_h = h;
    project_sh(rgba, w, h, sc.sky_sh);
    return 0;
}
void oracle_sky_sh(void* s, float* out27) { std::memcpy(out27, static_cast<Scene*>(s)->sky_sh, 27 * 4); }
// sky_radiance / sky_rough of the frame P at n directions (3 floats each; out 4 each)
void oracle_sky_radiance(void* s, const FrameParams* P, const float* dirs, uint32_t n, float* out) {
    const Scene& sc = *static_cast<Scene*>(s);
    for (uint32_t i = 0; i < n; ++i) {
        const F4 r = sky_radiance(sc, *P, F3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        std::memcpy(out + 4 * i, r.v, 16);
    }
}
void oracle_sky_rough(void* s, const FrameParams* P, const float* dirs, uint32_t n, float* out) {
    const Scene& sc = *static_cast<Scene*>(s);
    for (uint32_t i = 0; i < n; ++i) {
        const F4 r = sky_rough(sc, *P, F3(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
        std::memcpy(out + 4 * i, r.v, 16);
    }
}
void oracle_atan2_turns(const float* y, const float* x, uint32_t n, float* out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = atan2_turns(y[i], x[i]);
}

struct TexABI {  // gk_texture
    uint32_t w, h, srgb, reserved;
    const uint8_t* rgba8;
};

// the scene's textures; the byte -> linear tables as the sRGB transfer function and
// byte / 255, evaluated in double
int oracle_scene_set_textures(void* s, const TexABI* t, uint32_t n) {
    Scene& sc = *static_cast<Scene*>(s);
    sc.tex.assign(n, TextureO{});
    for (uint32_t i = 0; i < n; ++i) {
        TextureO& o = sc.tex[i];
        o.w = t[i].w; o.h = t[i].h; o.srgb = t[i].srgb ? 1 : 0;
        o.texels.resize(size_t(o.w) * o.h);
        std::memcpy(o.texels.data(), t[i].rgba8, 4 * o.texels.size());
    }
    for (int b = 0; b < 256; ++b) {
        const double x = b / 255.0;
        sc.tex_lut[b] = float(x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4));
        sc.tex_lut[256 + b] = float(x);
    }
    return 0;
}

int oracle_scene_set_materials(void* s, const Material* mats, uint32_t n) {
    Scene& sc = *static_cast<Scene*>(s);
    sc.mats.assign(mats, mats + n);
    return 0;
}

int oracle_scene_add_instance(void* s, const InstanceABI* in) {
    Scene& sc = *static_cast<Scene*>(s);
    Instance I{};
    std::memcpy(I.world, in->world, 64);
    std::memcpy(I.prev_world, in->prev_world, 64);
    I.instance_id = in->instance_id;
    I.mesh_id = in->mesh_id;
    I.visible = in->visible;
    I.nort = in->nort;
    std::memcpy(I.mat_id, in->mat_id, 64);
    I.mat_offset = in->mat_offset;
    I.use_remap = in->use_remap;
    sc.inst.push_back(I);
    return int(sc.inst.size() - 1);
}

int oracle_scene_build(void* s) {
    Scene& sc = *static_cast<Scene*>(s);
    for (const Instance& I : sc.inst) if (I.mesh_id >= sc.meshes.size()) return -1;
    if (sc.mats.empty()) return -1;
    scene_build(sc);
    build_lights(sc);
    return 0;
}

// New world matrices (column-major, n*16) for instances [first, first+n); prev <- old world.
int oracle_scene_update_instances(void* s, const float* world16, uint32_t first, uint32_t n) {
    Scene& sc = *static_cast<Scene*>(s);
    if (size_t(first) + n > sc.inst.size()) return -1;
    for (uint32_t i = 0; i < n; ++i) {
        Instance& I = sc.inst[first + i];
        std::memcpy(I.prev_world, I.world, 64);
        std::memcpy(I.world, world16 + size_t(i) * 16, 64);
    }
    scene_update_instances(sc);
    build_lights(sc);
    return 0;
}

void oracle_blas_stats(void* s, uint32_t mesh, uint32_t* nodes, float* sah) {
    Scene& sc = *static_cast<Scene*>(s);
    *nodes = blas_node_count(sc.meshes[mesh]);
    *sah = sc.meshes[mesh].nodes.empty() ? 0.f : blas_sah_cost(sc.meshes[mesh], 0);
}

// rays: n gk_ray; hits: n*{t,u,v,prim}; inst may be NULL. any != 0: occlusion (hits[i].prim = 0/1)
void oracle_trace(void* s, const float* rays, uint32_t n, float* hits, uint32_t* inst, int any,
                  uint32_t nthreads) {
    const Scene& sc = *static_cast<Scene*>(s);
    const uint32_t chunk = 4096;
    parallel_rows((n + chunk - 1) / chunk, nthreads, [&](uint32_t c) {
        for (uint32_t i = c * chunk; i < std::min(n, (c + 1) * chunk); ++i) {
            const float* r = rays + size_t(i) * 8;
            const F3 O(r[0], r[1], r[2]), D(r[4], r[5], r[6]);
            float* o = hits + size_t(i) * 4;
            if (any) {
                const uint32_t occ = trace_any(sc, O, D, r[3], r[7]) ? 1u : 0u;
                o[0] = o[1] = o[2] = 0.f;
                std::memcpy(&o[3], &occ, 4);
                continue;
            }
            Hit h;
            if (trace_closest(sc, O, D, r[3], r[7], h)) {
                o[0] = h.t; o[1] = h.u; o[2] = h.v;
                std::memcpy(&o[3], &h.prim, 4);
                if (inst) inst[i] = h.inst;
            } else {
                const float inf = std::numeric_limits<float>::infinity();
                const uint32_t miss = 0xffffffffu;
                o[0] = inf; o[1] = 0.f; o[2] = 0.f;
                std::memcpy(&o[3], &miss, 4);
                if (inst) inst[i] = miss;
            }
        }
    });
}

// The frame's primary rays with the thin lens of P (each pixel's first draw from its
// pcg4d seed (x, y, total_frames, global_seed)), as gk_ray records (tmin 0, tmax 1e30).
void oracle_lens_rays(const FrameParams* P, uint32_t W, uint32_t H, float* rays) {
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            Rng seed{{x, y, P->total_frames, P->global_seed}};
            F3 O, D;
            primary_ray_lens(*P, x, y, W, H, seed, O, D);
            float* r = rays + (size_t(y) * W + x) * 8;
            r[0] = O.x; r[1] = O.y; r[2] = O.z; r[3] = 0.f;
            r[4] = D.x; r[5] = D.y; r[6] = D.z; r[7] = 1e30f;
        }
}

// Primary rays of a W x H frame as gk_ray records (tmin 0, tmax 1e30).
void oracle_primary_rays(const FrameParams* P, uint32_t W, uint32_t H, float* rays) {
    for (uint32_t y = 0; y < H; ++y)
        for (uint32_t x = 0; x < W; ++x) {
            F3 O, D;
            primary_ray(*P, x, y, W, H, O, D);
            float* r = rays + (size_t(y) * W + x) * 8;
            r[0] = O.x; r[1] = O.y; r[2] = O.z; r[3] = 0.f;
            r[4] = D.x; r[5] = D.y; r[6] = D.z; r[7] = 1e30f;
        }
}

// ---------------------------------------------------------------- probe exports
// Voxelizes probes [first, first+count) of grid g into vox (indexed by probe).
void oracle_probe_voxelize(void* s, const ProbeGridO* g, VoxelO* vox, uint32_t first, uint32_t count,
                           uint32_t nthreads) {
    const Scene& sc = *static_cast<Scene*>(s);
    parallel_rows(count, nthreads, [&](uint32_t i) {
        const uint32_t idx = first + i;
        voxelize_cube(sc, g->unit, vox[idx], probe_origin(*g, idx));
    });
}

// One bake pass over probes [first, first+count): cubes/vox are the whole grid, updated in
// place; light may be NULL (no LightObject).
void oracle_probe_bake(void* s, const ProbeGridO* g, const FrameParams* P, const LightABI* light, CubeO* cubes,
                       VoxelO* vox, uint32_t first, uint32_t count, uint32_t nthreads) {
    const Scene& sc = *static_cast<Scene*>(s);
    const size_t n = size_t(g->size_xy) * g->size_xy * g->size_z;
    std::vector<CubeO> snapshot(cubes, cubes + n);
    std::vector<VoxelO> vsnap(vox, vox + n);
    ProbeBaker B{sc, *P, light, ProbeRef{*g, snapshot.data(), vsnap.data()}};
    parallel_rows(count, nthreads, [&](uint32_t i) {
        const uint32_t idx = first + i;
        B.render(idx, cubes[idx], vox[idx]);
    });
}

// interpolateAmbientCubesV2 at n points (pos, nrm: 3 floats each; out: 4 floats each)
void oracle_probe_lookup(const ProbeGridO* g, const CubeO* cubes, const VoxelO* vox, const float* pos,
                         const float* nrm, uint32_t n, int full, float* out) {
    const ProbeRef pr{*g, cubes, vox};
    for (uint32_t i = 0; i < n; ++i) {
        const F4 r = interpolate_cubes(pr, F3(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                                       F3(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]), full != 0);
        std::memcpy(out + 4 * i, r.v, 16);
    }
}

uint32_t oracle_probe_pack(const float* c) { return pack_rgb10a2(F4{{c[0], c[1], c[2], c[3]}}); }
void oracle_probe_unpack(uint32_t p, float* c) { const F4 r = unpack_rgb10a2(p); std::memcpy(c, r.v, 16); }

// REF_V2 termination reads this grid (arrays must outlive the renders); cubes NULL detaches.
void oracle_scene_set_probes(void* s, const ProbeGridO* g, const CubeO* cubes, const VoxelO* vox) {
    Scene& sc = *static_cast<Scene*>(s);
    sc.probes = cubes ? ProbeRef{*g, cubes, vox} : ProbeRef{};
}

// REF_V2 frame over the pixel rectangle [x0,x0+w) x [y0,y0+h) of a W x H image.
// Outputs are w*h arrays (row-major within the rectangle). moment is in/out (u16).
// Returns the number of rays traced.
uint64_t oracle_render_refv2(void* s, const FrameParams* P, uint32_t W, uint32_t H, uint32_t x0,
                             uint32_t y0, uint32_t w, uint32_t h, float* diffuse, float* spec,
                             float* albedo, float* normal, uint32_t* objid, float* motion,
                             float* depth, uint16_t* moment, uint32_t nthreads) {
    const Scene& sc = *static_cast<Scene*>(s);
    std::vector<std::vector<float>> comb(sc.inst.size(), std::vector<float>(16));
    for (size_t i = 0; i < sc.inst.size(); ++i) combined_prev(sc.inst[i].world, sc.inst[i].prev_world, comb[i].data());
    RefV2 r{sc, *P, comb};
    std::atomic<uint64_t> rays{0};
    parallel_rows(h, nthreads, [&](uint32_t yy) {
        uint64_t local = 0;
        for (uint32_t xx = 0; xx < w; ++xx) {
            const size_t o = size_t(yy) * w + xx;
            RefOut out;
            r.pixel(x0 + xx, y0 + yy, W, H, moment[o], out);
            std::memcpy(diffuse + 4 * o, out.diffuse, 16);
            std::memcpy(spec + 4 * o, out.spec, 16);
            std::memcpy(albedo + 4 * o, out.albedo, 16);
            std::memcpy(normal + 4 * o, out.normal, 16);
            objid[o] = out.objid;
            motion[2 * o] = out.motion[0];
            motion[2 * o + 1] = out.motion[1];
            depth[o] = out.depth;
            local += out.rays;
        }
        rays += local;
    });
    return rays.load();
}

}  // extern "C"

// PT_MIS over a rectangle: radiance (rgb) averaged over spp samples drawn from frames
// [frame0, frame0+frames) (seed z = frame), spp_per_frame = P->samples paths each.
// diffuse/spec receive the primary-lobe split (w*h*3 each, may be NULL).
template <class T>
uint64_t render_ptmis(const Scene& sc, const FrameParams* P, uint32_t W, uint32_t H, uint32_t x0,
                      uint32_t y0, uint32_t w, uint32_t h, uint32_t frame0, uint32_t frames,
                      double* rgb, double* diffuse, double* spec, uint32_t nthreads, int mode = 0) {
    PtMis<T> pt{sc, *P};
    pt.mode = mode;
    std::atomic<uint64_t> rays{0};
    const uint32_t spp = std::max(1u, P->samples);
    parallel_rows(h, nthreads, [&](uint32_t yy) {
        uint64_t local = 0;
        for (uint32_t xx = 0; xx < w; ++xx) {
            V3<T> sd(T(0), T(0), T(0)), ss(T(0), T(0), T(0));
            for (uint32_t f = 0; f < frames; ++f) {
                Rng seed{{x0 + xx, y0 + yy, frame0 + f, P->global_seed}};
                for (uint32_t s = 0; s < spp; ++s) {
                    V3<T> Ld(T(0), T(0), T(0)), Ls(T(0), T(0), T(0));
                    pt.path(x0 + xx, y0 + yy, W, H, seed, Ld, Ls, local);
                    sd = sd + Ld;
                    ss = ss + Ls;
                }
            }
            const T inv = T(1) / T(frames * spp);
            const size_t o = size_t(yy) * w + xx;
            for (int c = 0; c < 3; ++c) {
                rgb[3 * o + c] = double((sd[c] + ss[c]) * inv);
                if (diffuse) diffuse[3 * o + c] = double(sd[c] * inv);
                if (spec) spec[3 * o + c] = double(ss[c] * inv);
            }
        }
        rays += local;
    });
    return rays.load();
}

extern "C" {

uint64_t oracle_render_ptmis(void* s, const FrameParams* P, uint32_t W, uint32_t H, uint32_t x0,
                             uint32_t y0, uint32_t w, uint32_t h, uint32_t frame0, uint32_t frames,
                             int fp64, double* rgb, double* diffuse, double* spec, uint32_t nthreads) {
    const Scene& sc = *static_cast<Scene*>(s);
    if (fp64) return render_ptmis<double>(sc, P, W, H, x0, y0, w, h, frame0, frames, rgb, diffuse, spec, nthreads);
    return render_ptmis<float>(sc, P, W, H, x0, y0, w, h, frame0, frames, rgb, diffuse, spec, nthreads);
}

// The same render with a strategy mode of PtMis (0 MIS, 1 NEE only, 2 BSDF only, 3-5 planted
// weight mistakes); double precision.
uint64_t oracle_render_ptmis_mode(void* s, const FrameParams* P, uint32_t W, uint32_t H, uint32_t x0, uint32_t y0,
                                  uint32_t w, uint32_t h, uint32_t frame0, uint32_t frames, int mode, double* rgb,
                                  uint32_t nthreads) {
    const Scene& sc = *static_cast<Scene*>(s);
    return render_ptmis<double>(sc, P, W, H, x0, y0, w, h, frame0, frames, rgb, nullptr, nullptr, nthreads, mode);
}

// Detached differentiable PT_MIS (DiffPath) over a rectangle: mean of spp paths seeded
// (x, y, frame0 + s, global_seed); eval: num_materials materials or NULL (= scene's).
void oracle_diff_render(void* s, const FrameParams* P, const Material* eval, uint32_t W, uint32_t H, uint32_t x0,
                        uint32_t y0, uint32_t w, uint32_t h, uint32_t frame0, uint32_t spp, double* rgb,
                        uint32_t nthreads) {
    const Scene& sc = *static_cast<Scene*>(s);
    std::vector<Material> ev = eval ? std::vector<Material>(eval, eval + sc.mats.size()) : sc.mats;
    DiffPath dp{sc, *P, ev};
    parallel_rows(h, nthreads, [&](uint32_t yy) {
        for (uint32_t xx = 0; xx < w; ++xx) {
            D3 acc(0, 0, 0);
            for (uint32_t k = 0; k < spp; ++k) {
                Rng seed{{x0 + xx, y0 + yy, frame0 + k, P->global_seed}};
                acc = acc + dp.path(x0 + xx, y0 + yy, W, H, seed);
            }
            const size_t o = size_t(yy) * w + xx;
            for (int c = 0; c < 3; ++c) rgb[3 * o + c] = acc[c] / double(spp);
        }
    });
}

// pcg4d known answers: n steps from seed, writes 4*n words
void oracle_pcg4d(const uint32_t* seed, uint32_t n, uint32_t* out) {
    Rng s{{seed[0], seed[1], seed[2], seed[3]}};
    for (uint32_t i = 0; i < n; ++i) {
        pcg4d(s);
        std::memcpy(out + 4 * i, s.v, 16);
    }
}
float oracle_u2f(uint32_t x) { return u2f(x); }
void oracle_sincos_2pi(const float* u, uint32_t n, float* s, float* c) {
    for (uint32_t i = 0; i < n; ++i) sincos_2pi<float>(u[i], s[i], c[i]);
}
// sampler draws for distribution tests: kind 0 = RandomInHemiSphere1 (around +z),
// 1 = RandomInCone(cos_theta = param), 2 = ggx_sample_vndf(alpha = param, wi = dir)
void oracle_sample(int kind, float param, const float* dir, uint32_t n, uint32_t seed, float* out) {
    Rng s{{seed, 0u, 0u, 0u}};
    for (uint32_t i = 0; i < n; ++i) {
        F3 v;
        if (kind == 0) v = random_hemisphere1(s);
        else if (kind == 1) v = random_in_cone(s, param);
        else if (kind == 3) {  // RandomInUnitDisk
            float a, b, dx, dy;
            rnd2(s, a, b);
            concentric_disk(a, b, dx, dy);
            v = F3(dx, dy, 0.f);
        } else {
            float a, b;
            rnd2(s, a, b);
            v = ggx_sample_vndf(param, F3(dir[0], dir[1], dir[2]), a, b);
        }
        out[3 * i] = v.x; out[3 * i + 1] = v.y; out[3 * i + 2] = v.z;
    }
}
// bounded-VNDF reflected-direction sampling and its pdf in double (PT_MIS):
// writes n directions (3 floats) and pdfs.
void oracle_ptmis_spec_sample(double alpha, const double* wo, uint32_t n, uint32_t seed, double* dirs,
                              double* pdfs) {
    Rng s{{seed, 1u, 2u, 3u}};
    const D3 w(wo[0], wo[1], wo[2]);
    for (uint32_t i = 0; i < n; ++i) {
        float a, b;
        rnd2(s, a, b);
        const D3 m = PtMis<double>::vndf_sample(alpha, w, a, b);
        const D3 wi = m * (2.0 * dot(w, m)) - w;
        dirs[3 * i] = wi.x; dirs[3 * i + 1] = wi.y; dirs[3 * i + 2] = wi.z;
        pdfs[i] = PtMis<double>::vndf_pdf(alpha, w, m);
    }
}
double oracle_ptmis_spec_pdf(double alpha, const double* wo, const double* wi) {
    const D3 o(wo[0], wo[1], wo[2]), i(wi[0], wi[1], wi[2]);
    D3 h = o + i;
    h = normalize(h);
    return PtMis<double>::vndf_pdf(alpha, o, h);
}
void oracle_concentric_disk(float ua, float ub, float* out) { concentric_disk(ua, ub, out[0], out[1]); }

// ---------------------------------------------------------------- visibility buffer
// The raster pass's ids (Rast.VisibilityPass.vert.slang:18-22) for the NodeProxy list of
// UpdateNodesGpuDriven (Scene.cpp:472-478): one proxy per instance and 65535-triangle
// section of its mesh (RebuildMeshBuffer, Scene.cpp:202-256), in instance order.
static std::vector<uint32_t> proxy_bases(const Scene& sc) {
    std::vector<uint32_t> base(sc.inst.size() + 1, 0u);
    for (size_t i = 0; i < sc.inst.size(); ++i) {
        const uint32_t tris = uint32_t(sc.meshes[sc.inst[i].mesh_id].v0.size());
        base[i + 1] = base[i] + (tris + 65534u) / 65535u;
    }
    return base;
}
// hits: n x {t, u, v, prim bits}, inst: n instance indices -> packed ids (0 for a miss or a
// proxy the 15-bit field cannot name)
void oracle_vis_pack(void* s, const float* hits, const uint32_t* inst, uint32_t n, uint32_t* out) {
    const Scene& sc = *static_cast<Scene*>(s);
    const std::vector<uint32_t> base = proxy_bases(sc);
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t prim;
        std::memcpy(&prim, hits + 4 * i + 3, 4);
        out[i] = 0u;
        if (prim == 0xffffffffu) continue;
        const uint32_t proxy = base[inst[i]] + prim / 65535u;
        if (proxy + 1u <= 0x7fffu) out[i] = ((proxy + 1u) << 17) | (prim % 65535u);
    }
}
// get_material_dataV2 (GeneralFunc.slang:6-56) for the packed id of each ray: the named
// triangle's corners moved to world space by the instance's worldTS, barycentrics from the
// world-space ray by its Moeller-Trumbore form without rejection (:31-45), and the
// interpolated position. out: n x {instance, triangle (uint bits), b.x, b.y, b.z, pos.xyz};
// instance 0xffffffff where the id names no triangle. Corners in the mesh's own order (the
// reference's provoking-vertex reorder relabels them; the position is the same point).
void oracle_vis_decode(void* s, const uint32_t* packed, const float* rays, uint32_t n, float* out) {
    const Scene& sc = *static_cast<Scene*>(s);
    const std::vector<uint32_t> base = proxy_bases(sc);
    const uint32_t nproxy = base.back();
    for (uint32_t i = 0; i < n; ++i) {
        float* o = out + 8 * size_t(i);
        const uint32_t miss = 0xffffffffu;
        std::memcpy(&o[0], &miss, 4);
        std::memcpy(&o[1], &miss, 4);
        for (int k = 2; k < 8; ++k) o[k] = 0.f;
        const uint32_t pk = packed[i];
        const uint32_t ip1 = (pk >> 17) & 0x7fffu, stri = pk & 0x1ffffu;
        if (ip1 == 0u || ip1 - 1u >= nproxy || stri >= 65535u) continue;
        const uint32_t proxy = ip1 - 1u;
        const uint32_t ii = uint32_t(std::upper_bound(base.begin(), base.end(), proxy) - base.begin()) - 1u;
        const Instance& I = sc.inst[ii];
        const Mesh& m = sc.meshes[I.mesh_id];
        const uint32_t tri = (proxy - base[ii]) * 65535u + stri;
        if (tri >= m.v0.size()) continue;
        const F3 lp[3] = {m.v0[tri], m.v0[tri] + m.e1[tri], m.v0[tri] + m.e2[tri]};
        F3 p[3];
        for (int k = 0; k < 3; ++k) {
            const float v4[4] = {lp[k].x, lp[k].y, lp[k].z, 1.f};
            float w4[4];
            mat_mul4(I.world, v4, w4);
            p[k] = F3(w4[0], w4[1], w4[2]);
        }
        const F3 O(rays[8 * size_t(i)], rays[8 * size_t(i) + 1], rays[8 * size_t(i) + 2]);
        const F3 D(rays[8 * size_t(i) + 4], rays[8 * size_t(i) + 5], rays[8 * size_t(i) + 6]);
        const F3 e0 = p[1] - p[0], e1 = p[2] - p[0];
        const F3 rce1 = cross(D, e1);
        const float rcp = 1.0f / dot(e0, rce1);
        const F3 r0 = O - p[0];
        const float by = rcp * dot(r0, rce1);
        const F3 e0c0 = cross(e0, r0);
        const float bz = -rcp * dot(D, e0c0);
        const float bx = 1.0f - (by + bz);
        const F3 pos = p[0] * bx + (p[1] * by + p[2] * bz);
        std::memcpy(&o[0], &ii, 4);
        std::memcpy(&o[1], &tri, 4);
        o[2] = bx; o[3] = by; o[4] = bz;
        o[5] = pos.x; o[6] = pos.y; o[7] = pos.z;
    }
}
uint32_t oracle_num_lights(void* s) { return uint32_t(static_cast<Scene*>(s)->lights.size()); }

}  // extern "C"

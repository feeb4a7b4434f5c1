// gk_scene.hpp - device view of the scene and per-hit reconstruction shared by the
// wavefront kernels (wavefront.hip, adjoint.hip).
#pragma once
#include <hip/hip_fp16.h>

#include "gk_ctx.hpp"

namespace gk {

struct DScene {
    const Bvh4Node* blas;
    const TriRec* tris;
    const Bvh4Node* tlas;
    const uint32_t* tlas_prims;
    const InstDev* inst;
    const float4* tri_nrm;
    const uint32_t* tri_mat;
    const gk_material* mats;
    const LightTri* lights;
    const float* light_cdf;
    uint32_t nmat;
    uint32_t nlights;
    uint32_t blas_root;
    uint32_t two_level;
    // textures
    const float2* tri_uv;     // null: no texture coordinates
    const float4* tri_tan;    // null: no tangents
    const TexInfo* tex;
    const uint32_t* texels;
    const float* tex_lut;
    uint32_t ntex;
    // sky image (null: constant sky)
    const float4* sky;
    uint32_t sky_w, sky_h;
};

inline DScene make_dscene(const Ctx& c) {
    DScene s;
    s.blas = c.blas_nodes.ptr;
    s.tris = c.tris.ptr;
    s.tlas = c.tlas_nodes.ptr;
    s.tlas_prims = c.tlas_prims.ptr;
    s.inst = c.inst.ptr;
    s.tri_nrm = c.tri_nrm.ptr;
    s.tri_mat = c.tri_mat.ptr;
    s.mats = c.materials.ptr;
    s.lights = c.lights.ptr;
    s.light_cdf = c.light_cdf.ptr;
    s.nmat = uint32_t(c.h_materials.size());
    s.nlights = c.num_lights;
    s.blas_root = c.h_meshes.empty() ? 0u : c.h_meshes[0].root;
    s.two_level = c.blas_only ? 0u : 1u;
    s.tri_uv = c.tri_uv.count ? c.tri_uv.ptr : nullptr;
    s.tri_tan = c.tri_tan.count ? c.tri_tan.ptr : nullptr;
    s.tex = c.tex.ptr;
    s.texels = c.texels.ptr;
    s.tex_lut = c.tex_lut.ptr;
    s.ntex = c.num_tex;
    s.sky = c.sky_w ? c.sky.ptr : nullptr;
    s.sky_w = c.sky_w;
    s.sky_h = c.sky_h;
    return s;
}

// FetchMaterialId (Const_Func.slang:370-424) or offset form, clamped as GetMaterial
// (BasicTypes.slang:376-379)
__device__ inline uint32_t material_of(const DScene& s, uint32_t inst, uint32_t prim) {
    const InstDev& I = s.inst[inst];
    const uint32_t local = s.tri_mat[I.tri_base + prim];
    const uint32_t m = I.use_remap ? I.mat_id[local & 15u] : I.mat_offset + local;
    return m > s.nmat - 1u ? s.nmat - 1u : m;
}

// row-major world->object inverse applied as the inverse transpose (normals)
__device__ inline f3 inv_transpose(const float* W, f3 v) {
    f3 o;
    o.x = W[0] * v.x + W[4] * v.y + W[8] * v.z;
    o.y = W[1] * v.x + W[5] * v.y + W[9] * v.z;
    o.z = W[2] * v.x + W[6] * v.y + W[10] * v.z;
    return o;
}

// branchless orthonormal basis around n (Duff et al. 2017) and AlignWithNormal
// (Const_Func.slang:8-49): r expressed in the frame (T, B, n)
__device__ inline void onb(f3 n, f3& b1, f3& b2) {
    const float sz = n.z < 0.f ? -1.f : 1.f;
    const float a = -1.0f / (sz + n.z);
    b2 = f3{n.x * n.y * a, sz + n.y * n.y * a, -n.y};
    b1 = f3{1.0f + sz * n.x * n.x * a, sz * b2.x, -sz * n.x};
}
__device__ inline f3 to_world(f3 v, f3 T, f3 B, f3 N) {
    return f3{v.x * T.x + v.y * B.x + v.z * N.x, v.x * T.y + v.y * B.y + v.z * N.y, v.x * T.z + v.y * B.z + v.z * N.z};
}
__device__ inline f3 align_with_normal(f3 r, f3 n) {
    f3 T, B;
    onb(n, T, B);
    return to_world(r, T, B, n);
}

// FHardwareRayTracerV2 hit vertex (Shading.slang:803-821)
__device__ inline void hit_vertex(const DScene& s, uint32_t inst, uint32_t prim, float t, float u, float v, f3 O,
                                  f3 D, f3& pos, f3& nrm) {
    const InstDev& I = s.inst[inst];
    const size_t b = 3 * size_t(I.tri_base + prim);
    const f3 n0 = ld3(s.tri_nrm[b]), n1 = ld3(s.tri_nrm[b + 1]), n2 = ld3(s.tri_nrm[b + 2]);
    const f3 mix = (n0 + (n1 - n0) * u) + (n2 - n0) * v;
    nrm = normalize(inv_transpose(I.inv, mix));
    pos = O + D * t;
}

// ---------------------------------------------------------------- textures
// SampleLevel(uv, 0) with the default sampler (Sampler.hpp:9-26): REPEAT addressing,
// bilinear between the four texels around u * w - 0.5, v * h - 0.5, each decoded to
// linear first (sRGB textures through the sRGB transfer function).
__device__ inline float4 tex_texel(const DScene& s, const TexInfo& t, uint32_t x, uint32_t y) {
    const uint32_t p = s.texels[t.offset + size_t(y) * t.w + x];
    const float* lut = s.tex_lut + (t.srgb ? 0 : 256);
    return make_float4(lut[p & 0xffu], lut[(p >> 8) & 0xffu], lut[(p >> 16) & 0xffu], s.tex_lut[256 + (p >> 24)]);
}
__device__ inline void tex_axis(float u, uint32_t n, uint32_t& i0, uint32_t& i1, float& a) {
    float w = u - floorf(u);
    if (!(w >= 0.f && w < 1.f)) w = 0.f;  // NaN / infinite coordinates
    const float x = w * float(n) - 0.5f;
    const float fx = floorf(x);
    a = x - fx;
    const int k = int(fx);
    i0 = k < 0 ? n - 1u : (uint32_t(k) >= n ? uint32_t(k) - n : uint32_t(k));
    i1 = i0 + 1u == n ? 0u : i0 + 1u;
}
__device__ inline float4 tex_sample(const DScene& s, uint32_t id, float2 uv) {
    const TexInfo t = s.tex[id];
    uint32_t x0, x1, y0, y1;
    float ax, ay;
    tex_axis(uv.x, t.w, x0, x1, ax);
    tex_axis(uv.y, t.h, y0, y1, ay);
    const float4 a = tex_texel(s, t, x0, y0), b = tex_texel(s, t, x1, y0);
    const float4 c = tex_texel(s, t, x0, y1), d = tex_texel(s, t, x1, y1);
    const float bx = 1.f - ax, by = 1.f - ay;
    const float4 top = make_float4(a.x * bx + b.x * ax, a.y * bx + b.y * ax, a.z * bx + b.z * ax, a.w * bx + b.w * ax);
    const float4 bot = make_float4(c.x * bx + d.x * ax, c.y * bx + d.y * ax, c.z * bx + d.z * ax, c.w * bx + d.w * ax);
    return make_float4(top.x * by + bot.x * ay, top.y * by + bot.y * ay, top.z * by + bot.z * ay, top.w * by + bot.w * ay);
}
__device__ inline bool has_tex(const DScene& s, int32_t id) { return id >= 0 && uint32_t(id) < s.ntex; }

// ---------------------------------------------------------------- sky
// SampleIBLV2(SkyIdx, dir, SkyRotation, 0) * SkyIntensity (Shading.slang:194-199 at its
// callers 1077 and 1193): u = (atan2(d.x, d.z) + pi rotation) / (2 pi), v = acos(d.y) / pi
// (as atan2(sqrt(1 - y^2), y) / pi) on the equirectangular image, bilinear with REPEAT on
// both axes like tex_sample, clamped to (10, 10, 10, 1), times col (the frame's sky_color).
// Without an image the sky is col.
__device__ inline float4 sky_radiance(const DScene& s, const float* col, float rot, f3 dir) {
    if (!s.sky) return make_float4(col[0], col[1], col[2], col[3]);
    const f3 d = normalize(dir);
    const float u = atan2_turns(d.x, d.z) + rot * 0.5f;
    const float y = d.y < -1.f ? -1.f : (d.y > 1.f ? 1.f : d.y);
    const float v = 2.f * atan2_turns(sqrtf((1.f - y) * (1.f + y)), y);
    uint32_t x0, x1, y0, y1;
    float ax, ay;
    tex_axis(u, s.sky_w, x0, x1, ax);
    tex_axis(v, s.sky_h, y0, y1, ay);
    const float4 a = s.sky[size_t(y0) * s.sky_w + x0], b = s.sky[size_t(y0) * s.sky_w + x1];
    const float4 c = s.sky[size_t(y1) * s.sky_w + x0], e = s.sky[size_t(y1) * s.sky_w + x1];
    const float bx = 1.f - ax, by = 1.f - ay;
    const float4 top = make_float4(a.x * bx + b.x * ax, a.y * bx + b.y * ax, a.z * bx + b.z * ax, a.w * bx + b.w * ax);
    const float4 bot = make_float4(c.x * bx + e.x * ax, c.y * bx + e.y * ax, c.z * bx + e.z * ax, c.w * bx + e.w * ax);
    const float4 t = make_float4(top.x * by + bot.x * ay, top.y * by + bot.y * ay, top.z * by + bot.z * ay,
                                 top.w * by + bot.w * ay);
    return make_float4(fminf(t.x, 10.f) * col[0], fminf(t.y, 10.f) * col[1], fminf(t.z, 10.f) * col[2],
                       fminf(t.w, 1.f) * col[3]);
}

// SampleIBLRough(SkyIdx, dir, SkyRotation) (Shading.slang:146-192): EvaluateSH of the image's
// coefficients at dir turned about +y by pi * (1 - rotation), alpha 1, times col; without an
// image col. sh: [channel][9].
__device__ inline float4 sky_rough(const float* sh, bool has_image, const float* col, float rot, f3 n) {
    if (!has_image) return make_float4(col[0], col[1], col[2], col[3]);
    float turns = (1.f - rot) * 0.5f;
    turns = turns - floorf(turns);
    float sa, ca;
    sincos_2pi(turns, sa, ca);
    const float rx = n.x * ca + n.z * sa, ry = n.y, rz = -n.x * sa + n.z * ca;
    const float b[9] = {0.282095f, -0.488603f * ry, 0.488603f * rz, -0.488603f * rx, 1.092548f * rx * ry,
                        -1.092548f * ry * rz, 0.315392f * (3.f * ry * ry - 1.0f), -1.092548f * rx * rz,
                        0.546274f * (rx * rx - rz * rz)};
    float r = 0.f, g = 0.f, bl = 0.f;
    for (int i = 0; i < 9; ++i) {
        r += sh[i] * b[i];
        g += sh[9 + i] * b[i];
        bl += sh[18 + i] * b[i];
    }
    return make_float4(r * col[0], g * col[1], bl * col[2], col[3]);
}

// texture coordinate at a traced hit: Mix(t0, t1, t2, bary) = t0 + (t1-t0) u + (t2-t0) v
// (FHardwareRayTracerV2::TraceRay, Shading.slang:820)
__device__ inline float2 hit_uv(const DScene& s, uint32_t inst, uint32_t prim, float u, float v) {
    if (!s.tri_uv) return make_float2(0.f, 0.f);
    const size_t b = 3 * size_t(s.inst[inst].tri_base + prim);
    const float2 t0 = s.tri_uv[b], t1 = s.tri_uv[b + 1], t2 = s.tri_uv[b + 2];
    return make_float2((t0.x + (t1.x - t0.x) * u) + (t2.x - t0.x) * v, (t0.y + (t1.y - t0.y) * u) + (t2.y - t0.y) * v);
}

// albedo = Diffuse * base colour texture (FetchGBufferV2 / GetRayColor, Shading.slang:84-90,1047-1052)
__device__ inline float4 mat_albedo(const DScene& s, const gk_material& m, float2 uv) {
    float4 a = make_float4(m.diffuse[0], m.diffuse[1], m.diffuse[2], m.diffuse[3]);
    if (has_tex(s, m.diffuse_texture_id)) {
        const float4 t = tex_sample(s, uint32_t(m.diffuse_texture_id), uv);
        a = make_float4(a.x * t.x, a.y * t.y, a.z * t.z, a.w * t.w);
    }
    return a;
}

// FetchGBufferV2 at the primary hit (Shading.slang:82-112) with get_material_dataV2's
// attributes (GeneralFunc.slang:48-53: bx*a0 + (by*a1 + bz*a2)): albedo, roughness and
// the shading normal, which a normal map replaces by
// normalize(world * normalize(mikkTSpace(tex, local normal, local tangent))).
__device__ inline void fetch_gbuffer(const DScene& s, uint32_t inst, uint32_t prim, float u, float v,
                                     const gk_material& m, f3& nrm, float4& albedo, float& rough) {
    const float by = u, bz = v, bx = 1.0f - (by + bz);
    const size_t b = 3 * size_t(s.inst[inst].tri_base + prim);
    float2 uv = make_float2(0.f, 0.f);
    if (s.tri_uv) {
        const float2 t0 = s.tri_uv[b], t1 = s.tri_uv[b + 1], t2 = s.tri_uv[b + 2];
        uv = make_float2(bx * t0.x + (by * t1.x + bz * t2.x), bx * t0.y + (by * t1.y + bz * t2.y));
    }
    albedo = mat_albedo(s, m, uv);
    rough = m.fuzziness;
    if (has_tex(s, m.mra_texture_id)) rough = rough * tex_sample(s, uint32_t(m.mra_texture_id), uv).y;
    if (has_tex(s, m.normal_texture_id)) {
        const float4 nt = tex_sample(s, uint32_t(m.normal_texture_id), uv);
        const f3 n0 = ld3(s.tri_nrm[b]), n1 = ld3(s.tri_nrm[b + 1]), n2 = ld3(s.tri_nrm[b + 2]);
        const f3 ln = n0 * bx + (n1 * by + n2 * bz);
        float4 lt = make_float4(0.f, 0.f, 0.f, 0.f);
        if (s.tri_tan) {
            const float4 a0 = s.tri_tan[b], a1 = s.tri_tan[b + 1], a2 = s.tri_tan[b + 2];
            lt = make_float4(bx * a0.x + (by * a1.x + bz * a2.x), bx * a0.y + (by * a1.y + bz * a2.y),
                             bx * a0.z + (by * a1.z + bz * a2.z), bx * a0.w + (by * a1.w + bz * a2.w));
        }
        const f3 vn = f3{nt.x * 2.0f - 1.0f, nt.y * 2.0f - 1.0f, nt.z * 2.0f - 1.0f};
        const f3 t3 = f3{lt.x, lt.y, lt.z};
        const f3 vb = cross(ln, t3) * lt.w;
        const f3 mapped = normalize((t3 * vn.x + vb * vn.y) + ln * vn.z);
        float w4[4];
        mul44v(s.inst[inst].world, mapped.x, mapped.y, mapped.z, 0.f, w4);
        nrm = normalize(f3{w4[0], w4[1], w4[2]});
    }
}

__device__ inline ushort4 to_h4(float a, float b, float c, float d) {
    return make_ushort4(__half_as_ushort(__float2half_rn(a)), __half_as_ushort(__float2half_rn(b)),
                        __half_as_ushort(__float2half_rn(c)), __half_as_ushort(__float2half_rn(d)));
}
__device__ inline float4 from_h4(ushort4 v) {
    return make_float4(__half2float(__ushort_as_half(v.x)), __half2float(__ushort_as_half(v.y)),
                       __half2float(__ushort_as_half(v.z)), __half2float(__ushort_as_half(v.w)));
}

// ---------------------------------------------------------------- single-ray closest hit
// Per-thread traversal with a private stack for per-pixel kernels (the differentiable
// renderer). Same triangle test, widened box test and (t, instance, primitive)
// tie-break as k_trace, so it returns the same hit for the same ray.
struct SHit {
    float t, u, v;
    uint32_t prim, inst;
};

__device__ inline bool trace_closest_single(const DScene& s, f3 O, f3 D, float tmin, float tmax, SHit& best) {
    uint32_t stack[96];
    int sp = 0;
    const f3 wO = O, wD = D;
    f3 rD = f3{safercp(D.x), safercp(D.y), safercp(D.z)};
    const f3 wrD = rD;
    uint32_t level = s.two_level ? 0u : 1u;
    uint32_t node = s.two_level ? 0u : s.blas_root;
    uint32_t cur_inst = 0;
    bool found = false;
    best.t = tmax;
    best.prim = best.inst = 0xffffffffu;
    best.u = best.v = 0.f;
    for (;;) {
        if (node == 0xfffffffeu) {
            if (sp == 0) return found;
            const uint32_t e = stack[--sp];
            if (e == kStackMark) { O = wO; D = wD; rD = wrD; level = 0; continue; }
            if (e & kInstFlag) {
                cur_inst = e & ~kInstFlag;
                const InstDev& I = s.inst[cur_inst];
                const float* M = I.inv;
                f3 r = f3{M[0] * wO.x + M[1] * wO.y + M[2] * wO.z + M[3], M[4] * wO.x + M[5] * wO.y + M[6] * wO.z + M[7],
                          M[8] * wO.x + M[9] * wO.y + M[10] * wO.z + M[11]};
                const float w = M[12] * wO.x + M[13] * wO.y + M[14] * wO.z + M[15];
                O = w == 1.0f ? r : r * (1.f / w);
                D = f3{M[0] * wD.x + M[1] * wD.y + M[2] * wD.z, M[4] * wD.x + M[5] * wD.y + M[6] * wD.z,
                       M[8] * wD.x + M[9] * wD.y + M[10] * wD.z};
                rD = f3{safercp(D.x), safercp(D.y), safercp(D.z)};
                if (sp < 96) stack[sp++] = kStackMark;
                level = 1;
                node = I.root;
                continue;
            }
            node = e;
            continue;
        }
        const Bvh4Node n = (level == 0) ? s.tlas[node] : s.blas[node];
        const float lx[4] = {n.lox.x, n.lox.y, n.lox.z, n.lox.w}, hx[4] = {n.hix.x, n.hix.y, n.hix.z, n.hix.w};
        const float ly[4] = {n.loy.x, n.loy.y, n.loy.z, n.loy.w}, hy[4] = {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w};
        const float lz[4] = {n.loz.x, n.loz.y, n.loz.z, n.loz.w}, hz[4] = {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w};
        const uint32_t ch[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
        const uint32_t me[4] = {n.meta.x, n.meta.y, n.meta.z, n.meta.w};
        node = 0xfffffffeu;
        // internal children are visited near to far (pushed far first), so that a near
        // hit culls the far subtrees; leaves are tested as they are met
        float kt[4] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff()};
        uint32_t kc[4] = {0u, 0u, 0u, 0u};
        for (int i = 0; i < 4; ++i) {
            float a = (lx[i] - O.x) * rD.x, b = (hx[i] - O.x) * rD.x;
            float lo = fminf(a, b), hi = fmaxf(a, b);
            a = (ly[i] - O.y) * rD.y; b = (hy[i] - O.y) * rD.y;
            lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
            a = (lz[i] - O.z) * rD.z; b = (hz[i] - O.z) * rD.z;
            lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
            // widened by 2^-16 relative, as k_trace's box test
            lo = lo > 0.f ? lo * 0.9999847412109375f : lo * 1.0000152587890625f;
            hi = hi > 0.f ? hi * 1.0000152587890625f : hi * 0.9999847412109375f;
            if (!(hi >= lo && hi >= tmin && lo <= best.t)) continue;
            if (me[i] == 0u) { kt[i] = lo; kc[i] = ch[i]; continue; }
            if (level == 0) {
                for (uint32_t k = 0; k < me[i]; ++k) if (sp < 96) stack[sp++] = kInstFlag | s.tlas_prims[ch[i] + k];
                continue;
            }
            for (uint32_t k = 0; k < me[i]; ++k) {
                const TriRec tr = s.tris[(ch[i] & kLeafFirstMask) + k];
                const f3 v0 = ld3(tr.v0), e1 = ld3(tr.e1), e2 = ld3(tr.e2);
                const f3 h = cross(D, e2);
                const float aa = dot(e1, h);
                if (fabsf(aa) < 0.0000001f) continue;
                const float f = 1 / aa;
                const f3 sv = O - v0;
                const float u = f * dot(sv, h);
                if (u < 0 || u > 1) continue;
                const f3 q = cross(sv, e1);
                const float v = f * dot(D, q);
                if (v < 0 || u + v > 1) continue;
                const float t = f * dot(e2, q);
                if (!(t > tmin && t < tmax)) continue;
                const uint32_t prim = __float_as_uint(tr.v0.w);
                const uint32_t ci = s.two_level ? cur_inst : 0u;
                if (!found || t < best.t || (t == best.t && (ci < best.inst || (ci == best.inst && prim < best.prim)))) {
                    best.t = t; best.u = u; best.v = v; best.prim = prim; best.inst = ci;
                    found = true;
                }
            }
        }
        // a fixed 5-exchange sorting network (static indices); the closest hit does not
        // depend on the order of equal distances
        auto cx = [&](int a, int b) {
            const bool sw = kt[b] < kt[a];
            const float ta = kt[a], tb = kt[b];
            const uint32_t ca = kc[a], cb = kc[b];
            kt[a] = sw ? tb : ta; kt[b] = sw ? ta : tb;
            kc[a] = sw ? cb : ca; kc[b] = sw ? ca : cb;
        };
        cx(0, 1); cx(2, 3); cx(0, 2); cx(1, 3); cx(1, 2);
#pragma unroll
        for (int k = 3; k >= 0; --k)
            if (kt[k] < __builtin_inff() && sp < 96) stack[sp++] = kc[k];
    }
}

// ---------------------------------------------------------------- single-ray any-hit
// Occlusion query for rare inline uses (REF_V2 sun cone); private stack.
__device__ inline bool occluded_single(const DScene& s, f3 O, f3 D, float tmin, float tmax) {
    uint32_t stack[96];
    int sp = 0;
    f3 wO = O, wD = D;
    f3 rD = f3{safercp(D.x), safercp(D.y), safercp(D.z)};
    f3 wrD = rD;
    uint32_t level = s.two_level ? 0u : 1u;
    uint32_t node = s.two_level ? 0u : s.blas_root;
    for (;;) {
        if (node == 0xfffffffeu) {
            if (sp == 0) return false;
            const uint32_t e = stack[--sp];
            if (e == kStackMark) { O = wO; D = wD; rD = wrD; level = 0; continue; }
            if (e & kInstFlag) {
                const InstDev& I = s.inst[e & ~kInstFlag];
                const float* M = I.inv;
                f3 r = f3{M[0] * wO.x + M[1] * wO.y + M[2] * wO.z + M[3], M[4] * wO.x + M[5] * wO.y + M[6] * wO.z + M[7],
                          M[8] * wO.x + M[9] * wO.y + M[10] * wO.z + M[11]};
                const float w = M[12] * wO.x + M[13] * wO.y + M[14] * wO.z + M[15];
                O = w == 1.0f ? r : r * (1.f / w);
                D = f3{M[0] * wD.x + M[1] * wD.y + M[2] * wD.z, M[4] * wD.x + M[5] * wD.y + M[6] * wD.z,
                       M[8] * wD.x + M[9] * wD.y + M[10] * wD.z};
                rD = f3{safercp(D.x), safercp(D.y), safercp(D.z)};
                if (sp < 96) stack[sp++] = kStackMark;
                level = 1;
                node = I.root;
                continue;
            }
            node = e;
            continue;
        }
        const Bvh4Node n = (level == 0) ? s.tlas[node] : s.blas[node];
        const float lx[4] = {n.lox.x, n.lox.y, n.lox.z, n.lox.w}, hx[4] = {n.hix.x, n.hix.y, n.hix.z, n.hix.w};
        const float ly[4] = {n.loy.x, n.loy.y, n.loy.z, n.loy.w}, hy[4] = {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w};
        const float lz[4] = {n.loz.x, n.loz.y, n.loz.z, n.loz.w}, hz[4] = {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w};
        const uint32_t ch[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
        const uint32_t me[4] = {n.meta.x, n.meta.y, n.meta.z, n.meta.w};
        node = 0xfffffffeu;
        for (int i = 0; i < 4; ++i) {
            float a = (lx[i] - O.x) * rD.x, b = (hx[i] - O.x) * rD.x;
            float lo = fminf(a, b), hi = fmaxf(a, b);
            a = (ly[i] - O.y) * rD.y; b = (hy[i] - O.y) * rD.y;
            lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
            a = (lz[i] - O.z) * rD.z; b = (hz[i] - O.z) * rD.z;
            lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
            // widened by 2^-16 relative, as k_trace's box test
            lo = lo > 0.f ? lo * 0.9999847412109375f : lo * 1.0000152587890625f;
            hi = hi > 0.f ? hi * 1.0000152587890625f : hi * 0.9999847412109375f;
            if (!(hi >= lo && hi >= tmin && lo <= tmax)) continue;
            if (me[i] == 0u) { if (sp < 96) stack[sp++] = ch[i]; continue; }
            if (level == 0) {
                for (uint32_t k = 0; k < me[i]; ++k) if (sp < 96) stack[sp++] = kInstFlag | s.tlas_prims[ch[i] + k];
                continue;
            }
            for (uint32_t k = 0; k < me[i]; ++k) {
                const TriRec tr = s.tris[(ch[i] & kLeafFirstMask) + k];
                const f3 v0 = ld3(tr.v0), e1 = ld3(tr.e1), e2 = ld3(tr.e2);
                const f3 h = cross(D, e2);
                const float aa = dot(e1, h);
                if (fabsf(aa) < 0.0000001f) continue;
                const float f = 1 / aa;
                const f3 sv = O - v0;
                const float u = f * dot(sv, h);
                if (u < 0 || u > 1) continue;
                const f3 q = cross(sv, e1);
                const float v = f * dot(D, q);
                if (v < 0 || u + v > 1) continue;
                const float t = f * dot(e2, q);
                if (t > tmin && t < tmax) return true;
            }
        }
    }
}

}  // namespace gk

// gk_probe.hpp - ambient-cube probe grid on the device: colour packing and the
// trilinear, occlusion-tested lookup interpolateAmbientCubesV2 (AmbientCube.slang:66-100,
// 135-232, 268-345). Used by the bake (probe.hip) and by REF_V2 path termination
// (wavefront.hip).
//
// Probe (x, y, z) sits at (x, y, z) * unit + offset and is element
// y * sxy * sxy + z * sxy + x of both arrays.
#pragma once
#include "gk_ctx.hpp"

namespace gk {

struct ProbeDev {
    const gk_ambient_cube* cubes;  // nullptr: no grid (every lookup is "outside")
    const gk_voxel* voxels;
    uint32_t sxy, sz;
    float unit;
    float off[3];
};

constexpr float kProbeMaxIlluminance = 512.f;  // MAX_ILLUMINANCE (AmbientCube.slang:4)

__host__ __device__ inline float clamp01(float x) { return fminf(fmaxf(x, 0.f), 1.f); }

// packRGB10A2 / unpackRGB10A2 (AmbientCube.slang:66-87)
__device__ inline uint32_t probe_pack(float4 c) {
    const float r = clamp01(c.x / kProbeMaxIlluminance), g = clamp01(c.y / kProbeMaxIlluminance);
    const float b = clamp01(c.z / kProbeMaxIlluminance), a = clamp01(c.w / kProbeMaxIlluminance);
    return uint32_t(r * 1023.0f) | (uint32_t(g * 1023.0f) << 10) | (uint32_t(b * 1023.0f) << 20) |
           (uint32_t(a * 3.0f) << 30);
}
__device__ inline float4 probe_unpack(uint32_t p) {
    const float r = float(p & 0x3ffu) / 1023.0f;
    const float g = float((p >> 10) & 0x3ffu) / 1023.0f;
    const float b = float((p >> 20) & 0x3ffu) / 1023.0f;
    return make_float4(r * kProbeMaxIlluminance, g * kProbeMaxIlluminance, b * kProbeMaxIlluminance,
                       0.f * kProbeMaxIlluminance);
}

__device__ inline void acc_scaled(float4& acc, float w, float4 v) {
    acc.x = acc.x + w * v.x;
    acc.y = acc.y + w * v.y;
    acc.z = acc.z + w * v.z;
    acc.w = acc.w + w * v.w;
}

// sampleAmbientCubeHL2_DI (direct faces only) and sampleAmbientCubeHL2_Full (direct +
// bounced faces, sky visibility in w), AmbientCube.slang:117-232
template <bool FULL>
__device__ inline float4 probe_sample(const gk_ambient_cube& c, f3 n) {
    const float wx = fmaxf(n.x, 0.f), wnx = fmaxf(-n.x, 0.f);
    const float wy = fmaxf(n.y, 0.f), wny = fmaxf(-n.y, 0.f);
    const float wz = fmaxf(n.z, 0.f), wnz = fmaxf(-n.z, 0.f);
    const float sum = wx + wnx + wy + wny + wz + wnz;
    float4 col = make_float4(0.f, 0.f, 0.f, 0.f);
    float skyvis = 0.f;
    if (FULL) {
        const uint32_t s0 = c.sky_vis_pznzpyny, s1 = c.sky_vis_pxnxs0s1;
        skyvis = skyvis + float(s0 & 0xffu) * wz;
        skyvis = skyvis + float((s0 >> 8) & 0xffu) * wnz;
        skyvis = skyvis + float((s0 >> 16) & 0xffu) * wy;
        skyvis = skyvis + float(s0 >> 24) * wny;
        skyvis = skyvis + float(s1 & 0xffu) * wx;
        skyvis = skyvis + float((s1 >> 8) & 0xffu) * wnx;
    }
    acc_scaled(col, wx, probe_unpack(c.pos_x_d));
    acc_scaled(col, wnx, probe_unpack(c.neg_x_d));
    acc_scaled(col, wy, probe_unpack(c.pos_y_d));
    acc_scaled(col, wny, probe_unpack(c.neg_y_d));
    acc_scaled(col, wz, probe_unpack(c.pos_z_d));
    acc_scaled(col, wnz, probe_unpack(c.neg_z_d));
    const float norm = sum > 0.f ? 1.0f / sum : 1.0f;
    if (FULL) {
        acc_scaled(col, wx, probe_unpack(c.pos_x));
        acc_scaled(col, wnx, probe_unpack(c.neg_x));
        acc_scaled(col, wy, probe_unpack(c.pos_y));
        acc_scaled(col, wny, probe_unpack(c.neg_y));
        acc_scaled(col, wz, probe_unpack(c.pos_z));
        acc_scaled(col, wnz, probe_unpack(c.neg_z));
        col.x = col.x * norm;
        col.y = col.y * norm;
        col.z = col.z * norm;
        skyvis = skyvis * norm;
        col.w = skyvis / 255.0f;
    } else {
        col.x = col.x * norm;
        col.y = col.y * norm;
        col.z = col.z * norm;
        col.w = col.w * norm;
    }
    return col;
}

// interpolateAmbientCubesV2 (AmbientCube.slang:268-345): the eight probes around pos,
// trilinear weights, a probe skipped when its voxel is inside geometry or when the
// direction-blended free distance of the voxel is shorter than the way to pos.
template <bool FULL>
__device__ inline float4 probe_lookup(const ProbeDev& P, f3 pos, f3 normal) {
    if (!P.cubes) return make_float4(0.f, 0.f, 0.f, 1.f);
    const f3 np = f3{(pos.x - P.off[0]) / P.unit, (pos.y - P.off[1]) / P.unit, (pos.z - P.off[2]) / P.unit};
    const float lim_xy = float(int(P.sxy) - 2), lim_z = float(int(P.sz) - 2);
    // written so that a NaN position also counts as outside
    if (!(np.x >= 0.f && np.y >= 0.f && np.z >= 0.f && np.x <= lim_xy && np.y <= lim_z && np.z <= lim_xy))
        return make_float4(0.f, 0.f, 0.f, 1.f);
    const f3 fl = f3{floorf(np.x), floorf(np.y), floorf(np.z)};
    const int bx = int(fl.x), by = int(fl.y), bz = int(fl.z);
    const f3 fr = np - fl;
    float total = 0.f;
    float4 res = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int i = 0; i < 8; ++i) {
        const int ox = i & 1, oy = (i >> 1) & 1, oz = (i >> 2) & 1;
        const size_t idx = (size_t(by + oy) * P.sxy + size_t(bz + oz)) * P.sxy + size_t(bx + ox);
        const gk_voxel v = P.voxels[idx];
        const uint32_t u0 = v.dist_gg_z01, u1 = v.dist_x01_y01;
        const float inside = float((u0 >> 8) & 0xffu) / 255.0f;
        if (inside < 0.01f) continue;
        const float dpz = float((u0 >> 16) & 0xffu) / 255.0f, dnz = float(u0 >> 24) / 255.0f;
        const float dpx = float(u1 & 0xffu) / 255.0f, dnx = float((u1 >> 8) & 0xffu) / 255.0f;
        const float dpy = float((u1 >> 16) & 0xffu) / 255.0f, dny = float(u1 >> 24) / 255.0f;
        const f3 to_local = fr - f3{float(ox), float(oy), float(oz)};
        const float dist = length(to_local);
        const f3 dir = normalize(to_local);
        const float hit_len = sqrtf(fmaxf(dir.x, 0.f)) * dpx + sqrtf(fmaxf(-dir.x, 0.f)) * dnx +
                              sqrtf(fmaxf(dir.y, 0.f)) * dpy + sqrtf(fmaxf(-dir.y, 0.f)) * dny +
                              sqrtf(fmaxf(dir.z, 0.f)) * dpz + sqrtf(fmaxf(-dir.z, 0.f)) * dnz;
        if (dist > hit_len + 0.05f) continue;
        const float wx = ox == 0 ? (1.0f - fr.x) : fr.x;
        const float wy = oy == 0 ? (1.0f - fr.y) : fr.y;
        const float wz = oz == 0 ? (1.0f - fr.z) : fr.z;
        const float w = wx * wy * wz;
        const float4 s = probe_sample<FULL>(P.cubes[idx], normal);
        res.x = res.x + s.x * w;
        res.y = res.y + s.y * w;
        res.z = res.z + s.z * w;
        res.w = res.w + s.w * w;
        total = total + w;
    }
    if (!(total > 0.f)) return make_float4(0.f, 0.f, 0.f, 0.f);
    return make_float4(res.x / total, res.y / total, res.z / total, res.w / total);
}

}  // namespace gk

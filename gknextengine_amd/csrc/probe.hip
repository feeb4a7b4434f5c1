// probe.hip - the ambient-cube probe grid: voxelization and bake.
//
//   voxelize  VoxelizeCube (CPUAccelerationStructure.cpp:85-160), which the reference runs
//             on CPU threads over 16x16-probe columns (AsyncProcessFull/AsyncProcessGroup,
//             :340-431), here one wavefront pass over the grid: 6 axis rays and 8 diagonal
//             rays per probe, all closest-hit, traced by the persistent traversal kernel.
//   bake      Bake.HwAmbientCube (Bake.HwAmbientCube.comp.slang:1-19) ->
//             FGpuProbeGenerator::RenderAmbientCubeOnly / FaceTask (AmbientCube.slang:
//             472-566,650-690): per probe within 8 units of a surface and per face, 16
//             hemisphere rays on a jittered 4x4 grid, a segment to the first quad light and
//             a sun ray; hits add albedo x the direct-light lookup x 1.25, misses add the
//             sky (its spherical harmonics when a sky image is set).
//             The new face value is lerp(old, new, 1/8) in RGB10.
//
// Work is done in chunks of kChunk probes: rays of a chunk are written to queues, traced,
// then reduced per (probe, face). A pass writes its results to a second array that is
// copied over the grid when the pass ends, so every lookup of the pass reads the grid as
// it was when the pass started.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "gk_probe.hpp"
#include "gk_scene.hpp"

namespace gk {

int collect_timing(Ctx& c);

namespace {

constexpr uint32_t kChunk = 262144;         // probes per chunk (bake: 25 M rays)
constexpr uint32_t kVoxRays = 14;            // 6 axis + 8 diagonal rays per probe
constexpr uint32_t kFaceRays = 16;           // FACE_TRACING (AmbientCube.slang:27,57)
constexpr uint32_t kProbeRays = 6 * kFaceRays;
constexpr float kEps = 1e-3f;                // EPS (PreProcessor.slang:13)
constexpr float kMaxTrace = 1000.f;          // PT_MAX_TRACE_DISTANCE (Shading.slang:15)
constexpr float kFastMaxTrace = 20.f;        // FAST_MAX_TRACE_DISTANCE (Shading.slang:16)
constexpr uint32_t kMiss = 0xffffffffu;

// grid4x4 and grid3x3 sample offsets (AmbientCube.slang:37-49)
__constant__ float kGrid4[16][2] = {
    {-0.75f, -0.75f}, {-0.25f, -0.75f}, {0.25f, -0.75f}, {0.75f, -0.75f},
    {-0.75f, -0.25f}, {-0.25f, -0.25f}, {0.25f, -0.25f}, {0.75f, -0.25f},
    {-0.75f, 0.25f},  {-0.25f, 0.25f},  {0.25f, 0.25f},  {0.75f, 0.25f},
    {-0.75f, 0.75f},  {-0.25f, 0.75f},  {0.25f, 0.75f},  {0.75f, 0.75f}};
__constant__ float kGrid3[9][2] = {{-0.667f, -0.667f}, {0.0f, -0.667f}, {0.667f, -0.667f},
                                   {-0.667f, 0.0f},    {0.0f, 0.0f},    {0.667f, 0.0f},
                                   {-0.667f, 0.667f},  {0.0f, 0.667f},  {0.667f, 0.667f}};

// faces in FaceTask call order (AmbientCube.slang:674-679): +y, -y, +x, -x, +z, -z
__device__ inline f3 face_basis(uint32_t f) {
    switch (f) {
        case 0: return f3{0.f, 1.f, 0.f};
        case 1: return f3{0.f, -1.f, 0.f};
        case 2: return f3{1.f, 0.f, 0.f};
        case 3: return f3{-1.f, 0.f, 0.f};
        case 4: return f3{0.f, 0.f, 1.f};
        default: return f3{0.f, 0.f, -1.f};
    }
}

// VoxelizeCube ray k: axis rays +y, -y, +x, -x, +z, -z, then DetectDistance's eight
// diagonals (CPUAccelerationStructure.cpp:126-144, directions as written there,
// normalised by tinybvh::Ray)
__device__ inline f3 vox_dir(uint32_t k) {
    if (k < 6) return face_basis(k);
    // (1,1,1) (-1,1,1) (-1,-1,1) (-1,1,1) (1,1,-1) (-1,1,-1) (-1,-1,-1) (-1,1,-1)
    const uint32_t j = k - 6u;
    const float x = (j == 0u || j == 4u) ? 1.f : -1.f;
    const float y = (j == 2u || j == 6u) ? -1.f : 1.f;
    const float z = j < 4u ? 1.f : -1.f;
    return normalize(f3{x, y, z});
}

struct GridK {
    uint32_t sxy, sz, count;
    float unit, off[3];
};

__device__ inline f3 probe_pos(const GridK& g, uint32_t idx) {
    const uint32_t plane = g.sxy * g.sxy;
    const uint32_t y = idx / plane;
    const uint32_t rem = idx - y * plane;
    const uint32_t z = rem / g.sxy;
    const uint32_t x = rem - z * g.sxy;
    return f3{float(x) * g.unit + g.off[0], float(y) * g.unit + g.off[1], float(z) * g.unit + g.off[2]};
}

// ---------------------------------------------------------------- voxelize
__global__ void k_vox_rays(GridK g, uint32_t first, uint32_t n, float4* __restrict__ ro, float4* __restrict__ rd) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n * kVoxRays) return;
    const uint32_t v = r / kVoxRays, k = r - v * kVoxRays;
    const f3 o = probe_pos(g, first + v);
    const f3 d = vox_dir(k);
    ro[r] = make_float4(o.x, o.y, o.z, 0.f);
    rd[r] = make_float4(d.x, d.y, d.z, g.unit * 64.f);
}

// InsideGeometry + VoxelizeCube's encoding (CPUAccelerationStructure.cpp:85-160). The
// solid test uses the hit triangle's geometric normal in world space.
__global__ void k_vox_finish(DScene s, GridK g, uint32_t first, uint32_t n, const float4* __restrict__ rd,
                             const float4* __restrict__ hit, const uint32_t* __restrict__ hit_inst,
                             const uint32_t* __restrict__ tri_rec, gk_voxel* __restrict__ vox) {
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    uint32_t mat_id = 0;
    float dist[6];
    for (uint32_t k = 0; k < 6; ++k) {
        const size_t r = size_t(v) * kVoxRays + k;
        const float4 h = hit[r];
        dist[k] = 255.0f;
        if (__float_as_uint(h.w) == kMiss) continue;
        dist[k] = h.x;
        if (h.x <= g.unit) {
            const uint32_t inst = s.two_level ? hit_inst[r] : 0u;
            const uint32_t prim = __float_as_uint(h.w);
            const uint32_t mi = material_of(s, inst, prim);
            mat_id = mi;
            const InstDev& I = s.inst[inst];
            const TriRec& tr = s.tris[I.tri_base + tri_rec[I.tri_base + prim]];
            const f3 nl = normalize(cross(ld3(tr.e1), ld3(tr.e2)));
            float nw[4];
            mul44v(I.world, nl.x, nl.y, nl.z, 0.f, nw);
            const f3 d = ld3(rd[r]);
            if (nw[0] * d.x + nw[1] * d.y + nw[2] * d.z > 0.0f ||
                s.mats[mi].material_model == GK_MATERIAL_DIFFUSE_LIGHT)
                dist[k] = 0.f;
        }
    }
    // order of the reference's distances: +y, -y, +x, -x, +z, -z
    float mn = dist[0];
    for (int k = 1; k < 6; ++k) mn = dist[k] < mn ? dist[k] : mn;
    if (mn > 254.0f) {
        for (uint32_t k = 6; k < kVoxRays; ++k) {
            const float4 h = hit[size_t(v) * kVoxRays + k];
            const float dd = __float_as_uint(h.w) == kMiss ? 255.0f : h.x;
            mn = dd < mn ? dd : mn;
        }
    }
    const float py = clamp01(dist[0] / g.unit), ny = clamp01(dist[1] / g.unit);
    const float px = clamp01(dist[2] / g.unit), nx = clamp01(dist[3] / g.unit);
    const float pz = clamp01(dist[4] / g.unit), nz = clamp01(dist[5] / g.unit);
    const float inside = py * ny * px * nx * pz * nz;
    gk_voxel out;
    out.mat_id = mat_id;
    out.age = 0u;
    out.dist_gg_z01 = (uint32_t(mn / g.unit) & 0xffu) | ((uint32_t(inside * 255.0f) & 0xffu) << 8) |
                      ((uint32_t(pz * 255.0f) & 0xffu) << 16) | ((uint32_t(nz * 255.0f) & 0xffu) << 24);
    out.dist_x01_y01 = (uint32_t(px * 255.0f) & 0xffu) | ((uint32_t(nx * 255.0f) & 0xffu) << 8) |
                       ((uint32_t(py * 255.0f) & 0xffu) << 16) | ((uint32_t(ny * 255.0f) & 0xffu) << 24);
    vox[first + v] = out;
}

// ---------------------------------------------------------------- bake
struct BakeK {
    GridK g;
    float sky[4], sun_dir[4], sun_col[4];
    float sky_sh[27], sky_rot;  // SH of the sky image (sky_image != 0)
    uint32_t has_sky, sky_image, has_sun, has_light, nany;  // nany: occlusion rays per face (light, sun)
    float light_pos[3], light_area;
    uint32_t light_mat;
};

__device__ inline bool probe_active(const gk_voxel& v) { return (v.dist_gg_z01 & 0xffu) < 8u; }

// probes of [first, first+n) that get baked (RenderAmbientCubeOnly's test), in order
// within each work-group
__global__ __launch_bounds__(256) void k_bake_select(const gk_voxel* __restrict__ vox, uint32_t first, uint32_t n,
                                                     uint32_t* __restrict__ list, uint32_t* __restrict__ counters) {
    __shared__ uint32_t cnt[5];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool want = i < n && probe_active(vox[first + i]);
    const uint64_t m = __ballot(want);
    const uint32_t wv = threadIdx.x >> 6;
    if (__lane_id() == 0) cnt[wv] = uint32_t(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        cnt[4] = t ? atomicAdd(counters, t) : 0u;
    }
    __syncthreads();
    uint32_t b = cnt[4];
    for (uint32_t w = 0; w < wv; ++w) b += cnt[w];
    if (want) list[b + uint32_t(__popcll(m & ((1ull << __lane_id()) - 1ull)))] = first + i;
}

__global__ void k_bake_counts(uint32_t* counters, uint32_t nany) {
    counters[1] = counters[0] * kProbeRays;
    counters[2] = counters[0] * 6u * nany;
}

__device__ inline f3 face_origin(const GridK& g, uint32_t idx, f3 basis) {
    return probe_pos(g, idx) + (basis * g.unit) * 0.25f;
}

// FaceTask's rays (AmbientCube.slang:488-549)
__global__ void k_bake_rays(BakeK K, const gk_voxel* __restrict__ vox, const uint32_t* __restrict__ list,
                            const uint32_t* __restrict__ counters, float4* __restrict__ ro, float4* __restrict__ rd,
                            float4* __restrict__ so, float4* __restrict__ sd) {
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t np = counters[0];
    if (r < np * kProbeRays) {
        const uint32_t s = r / kProbeRays, rem = r - s * kProbeRays;
        const uint32_t f = rem / kFaceRays, i = rem - f * kFaceRays;
        const uint32_t idx = list[s];
        const uint32_t it = vox[idx].age % 9u;
        const f3 basis = face_basis(f);
        const f3 o = face_origin(K.g, idx, basis);
        const float jx = kGrid3[it][0] * 0.25f, jy = kGrid3[it][1] * 0.25f;
        const f3 hemi = normalize(f3{kGrid4[i][0] + jx, kGrid4[i][1] + jy, 1.0f});
        const f3 d = align_with_normal(hemi, basis);
        ro[r] = make_float4(o.x, o.y, o.z, kEps);
        rd[r] = make_float4(d.x, d.y, d.z, kFastMaxTrace);
    }
    if (K.nany && r < np * 6u * K.nany) {
        const uint32_t sf = r / K.nany, j = r - sf * K.nany;
        const uint32_t s = sf / 6u, f = sf - s * 6u;
        const uint32_t idx = list[s];
        const f3 o = face_origin(K.g, idx, face_basis(f));
        if (K.has_light && j == 0u) {
            // TraceSegment(origin, lightPos, CUBE_UNIT * 0.5) (Shading.slang:759-782)
            const float e = K.g.unit * 0.5f;
            const f3 dir = f3{K.light_pos[0], K.light_pos[1], K.light_pos[2]} - o;
            const float len = length(dir);
            so[r] = make_float4(o.x, o.y, o.z, e);
            sd[r] = make_float4(dir.x / len, dir.y / len, dir.z / len, len - e);
        } else {
            // TraceOcclusion(origin, SunDirection) (Shading.slang:737-757)
            so[r] = make_float4(o.x, o.y, o.z, kEps);
            sd[r] = make_float4(K.sun_dir[0], K.sun_dir[1], K.sun_dir[2], kMaxTrace);
        }
    }
}

__device__ inline float4 lerp4(float4 a, float4 b, float t) {
    return make_float4(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t, a.w + (b.w - a.w) * t);
}

__device__ inline uint32_t& face_direct(gk_ambient_cube& c, uint32_t f) {
    switch (f) {
        case 0: return c.pos_y_d;
        case 1: return c.neg_y_d;
        case 2: return c.pos_x_d;
        case 3: return c.neg_x_d;
        case 4: return c.pos_z_d;
        default: return c.neg_z_d;
    }
}
__device__ inline uint32_t& face_bounce(gk_ambient_cube& c, uint32_t f) {
    switch (f) {
        case 0: return c.pos_y;
        case 1: return c.neg_y;
        case 2: return c.pos_x;
        case 3: return c.neg_x;
        case 4: return c.pos_z;
        default: return c.neg_z;
    }
}
// sky-visibility byte of face f: (word, shift)
__device__ inline void face_sky_slot(uint32_t f, uint32_t& word, uint32_t& shift) {
    switch (f) {
        case 0: word = 0; shift = 16; break;  // +y -> pznzpyny.z
        case 1: word = 0; shift = 24; break;  // -y -> pznzpyny.w
        case 2: word = 1; shift = 0; break;   // +x -> pxnxs0s1.x
        case 3: word = 1; shift = 8; break;   // -x -> pxnxs0s1.y
        case 4: word = 0; shift = 0; break;   // +z -> pznzpyny.x
        default: word = 0; shift = 8; break;  // -z -> pznzpyny.y
    }
}

// FaceTask's shading and accumulation (AmbientCube.slang:497-566). One thread per ray
// shades its hit (albedo x direct-light lookup x 1.25) or miss (sky); the first thread of
// each 16-ray face then sums the face's terms in ray order, as the reference's loop does,
// adds the light and sun terms and blends the face into the new grid.
__global__ __launch_bounds__(256) void k_bake_face(DScene s, BakeK K, ProbeDev P, const gk_voxel* __restrict__ vox,
                                                   const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ counters,
                                                   const float4* __restrict__ ro, const float4* __restrict__ rd,
                                                   const float4* __restrict__ hit,
                                                   const uint32_t* __restrict__ hit_inst,
                                                   const float4* __restrict__ shit, gk_ambient_cube* __restrict__ next,
                                                   uint32_t* __restrict__ face_sky) {
    __shared__ float4 s_bounce[256], s_direct[256];
    __shared__ float s_miss[256];
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = r < counters[0] * kProbeRays;
    float4 b = make_float4(0.f, 0.f, 0.f, 0.f), d = make_float4(0.f, 0.f, 0.f, 0.f);
    float miss = 0.f;
    if (live) {
        const float4 h = hit[r];
        if (__float_as_uint(h.w) != kMiss) {
            const uint32_t inst = s.two_level ? hit_inst[r] : 0u;
            const uint32_t prim = __float_as_uint(h.w);
            f3 pos, nrm;
            hit_vertex(s, inst, prim, h.x, h.y, h.z, ld3(ro[r]), ld3(rd[r]), pos, nrm);
            const gk_material& m = s.mats[material_of(s, inst, prim)];
            const float4 alb = mat_albedo(s, m, hit_uv(s, inst, prim, h.y, h.z));
            const float4 li = probe_lookup<false>(P, pos, nrm);
            b = make_float4((alb.x * li.x) * 1.25f, (alb.y * li.y) * 1.25f, (alb.z * li.z) * 1.25f,
                            (alb.w * li.w) * 1.25f);
        } else {
            if (K.has_sky) d = sky_rough(K.sky_sh, K.sky_image != 0u, K.sky, K.sky_rot, ld3(rd[r]));
            miss = 1.f;
        }
    }
    s_bounce[threadIdx.x] = b;
    s_direct[threadIdx.x] = d;
    s_miss[threadIdx.x] = miss;
    __syncthreads();
    if (!live || (threadIdx.x & (kFaceRays - 1u)) != 0u) return;
    // x + 0 = x: adding the other kind's zero keeps the reference's conditional sums
    float4 direct = make_float4(0.f, 0.f, 0.f, 0.f), bounce = make_float4(0.f, 0.f, 0.f, 0.f);
    float skyv = 0.f;
    for (uint32_t i = 0; i < kFaceRays; ++i) {
        const float4 bi = s_bounce[threadIdx.x + i], di = s_direct[threadIdx.x + i];
        bounce = make_float4(bounce.x + bi.x, bounce.y + bi.y, bounce.z + bi.z, bounce.w + bi.w);
        direct = make_float4(direct.x + di.x, direct.y + di.y, direct.z + di.z, direct.w + di.w);
        skyv = skyv + s_miss[threadIdx.x + i];
    }
    const uint32_t t = r / kFaceRays;  // (probe, face)
    const uint32_t sidx = t / 6u, f = t - sidx * 6u;
    const uint32_t idx = list[sidx];
    const f3 basis = face_basis(f);
    const f3 o = face_origin(K.g, idx, basis);
    const float nf = float(kFaceRays);
    direct = make_float4(direct.x / nf, direct.y / nf, direct.z / nf, direct.w / nf);
    bounce = make_float4(bounce.x / nf, bounce.y / nf, bounce.z / nf, bounce.w / nf);
    uint32_t j = 0;
    if (K.has_light) {
        const bool occ = shit[(size_t(sidx) * 6u + f) * K.nany + j].w != 0.f;
        ++j;
        const float atten_l = occ ? 0.0f : 1.0f;
        const f3 lp = f3{K.light_pos[0], K.light_pos[1], K.light_pos[2]};
        const f3 ldir = normalize(lp - o);
        const float ndotl = clamp01(dot(basis, ldir));
        const float dist = length(lp - o);
        const float att = ndotl * K.light_area / (dist * dist * 3.14159f);
        const gk_material& lm = s.mats[K.light_mat];
        direct.x = direct.x + (lm.diffuse[0] * att) * atten_l;
        direct.y = direct.y + (lm.diffuse[1] * att) * atten_l;
        direct.z = direct.z + (lm.diffuse[2] * att) * atten_l;
        direct.w = direct.w + (lm.diffuse[3] * att) * atten_l;
    }
    uint32_t sun_bits = 0;
    if (K.has_sun) {
        const bool occ = shit[(size_t(sidx) * 6u + f) * K.nany + j].w != 0.f;
        const float atten_s = occ ? 0.0f : 1.0f;
        const float ndotl = clamp01(dot(basis, f3{K.sun_dir[0], K.sun_dir[1], K.sun_dir[2]}));
        sun_bits = 0x100u | (atten_s > 0.0f ? 0x200u : 0u);
        direct.x = direct.x + (((K.sun_col[0] * atten_s) * ndotl) * 1.0f) * 0.25f;
        direct.y = direct.y + (((K.sun_col[1] * atten_s) * ndotl) * 1.0f) * 0.25f;
        direct.z = direct.z + (((K.sun_col[2] * atten_s) * ndotl) * 1.0f) * 0.25f;
        direct.w = direct.w + (((K.sun_col[3] * atten_s) * ndotl) * 1.0f) * 0.25f;
    }
    // lerp towards this pass with weight 1/8 (AmbientCube.slang:561-565)
    const float w = 0.125f;
    gk_ambient_cube old = P.cubes[idx];
    uint32_t word, shift;
    face_sky_slot(f, word, shift);
    const uint32_t sky_old = ((word ? old.sky_vis_pxnxs0s1 : old.sky_vis_pznzpyny) >> shift) & 0xffu;
    const float sky_new = 255.0f * skyv / nf;
    const float sky_l = float(sky_old) + (sky_new - float(sky_old)) * w;
    const uint32_t d_new = probe_pack(lerp4(probe_unpack(face_direct(old, f)), direct, w));
    const uint32_t b_new = probe_pack(lerp4(probe_unpack(face_bounce(old, f)), bounce, w));
    gk_ambient_cube& out = next[idx];
    face_direct(out, f) = d_new;
    face_bounce(out, f) = b_new;
    face_sky[t] = (uint32_t(sky_l) & 0xffu) | (f == 0u ? sun_bits : 0u);
}

// sky bytes of the six faces, the +y face's sun bit, and the probe's age
__global__ void k_bake_commit(const uint32_t* __restrict__ list, const uint32_t* __restrict__ counters,
                              const gk_ambient_cube* __restrict__ cubes, const uint32_t* __restrict__ face_sky,
                              gk_ambient_cube* __restrict__ next, gk_voxel* __restrict__ vox) {
    const uint32_t sidx = blockIdx.x * blockDim.x + threadIdx.x;
    if (sidx >= counters[0]) return;
    const uint32_t idx = list[sidx];
    uint32_t w[2] = {cubes[idx].sky_vis_pznzpyny, cubes[idx].sky_vis_pxnxs0s1};
    for (uint32_t f = 0; f < 6; ++f) {
        const uint32_t v = face_sky[sidx * 6u + f];
        uint32_t word, shift;
        face_sky_slot(f, word, shift);
        w[word] = (w[word] & ~(0xffu << shift)) | ((v & 0xffu) << shift);
        if (v & 0x100u) w[1] = (w[1] & ~(0xffu << 16)) | (((v >> 9) & 1u) << 16);
    }
    next[idx].sky_vis_pznzpyny = w[0];
    next[idx].sky_vis_pxnxs0s1 = w[1];
    vox[idx].age = vox[idx].age + 1u;
}

__global__ void k_bake_apply(const gk_voxel* __restrict__ vox, uint32_t first, uint32_t n,
                             const gk_ambient_cube* __restrict__ next, gk_ambient_cube* __restrict__ cubes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (probe_active(vox[first + i])) cubes[first + i] = next[first + i];
}

__global__ void k_zero_u32(uint32_t* p, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

inline uint32_t nblk(size_t n, uint32_t b = 256) { return uint32_t((n + b - 1) / b); }

GridK grid_k(const Ctx& c) {
    const auto& g = c.probes.grid;
    GridK k;
    k.sxy = g.size_xy;
    k.sz = g.size_z;
    k.count = c.probes.count;
    k.unit = g.unit;
    memcpy(k.off, g.offset, 12);
    return k;
}

// keep room for the trace launches' timing events (folded into the frame statistics)
int reserve_trace_events(Ctx& c) {
    if (c.tev_used + 64u > c.tev.size() && c.tev_used > 0) {
        // the recorded launches may be on any slot's stream
        GK_HIP(hipStreamSynchronize(c.stream));
        for (auto& ps : c.pstream) if (ps) GK_HIP(hipStreamSynchronize(ps));
        return collect_timing(c);
    }
    return GK_OK;
}

}  // namespace

ProbeDev probe_dev(const Ctx& c) {
    ProbeDev P{};
    if (c.probes.count == 0) return P;
    P.cubes = c.probes.cubes.ptr;
    P.voxels = c.probes.voxels.ptr;
    P.sxy = c.probes.grid.size_xy;
    P.sz = c.probes.grid.size_z;
    P.unit = c.probes.grid.unit;
    memcpy(P.off, c.probes.grid.offset, 12);
    return P;
}

int probes_init(Ctx& c, const gk_probe_grid& g) {
    if (g.size_xy < 2 || g.size_z < 2 || !(g.unit > 0.f) || !std::isfinite(g.unit)) return GK_ERR_ARG;
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(g.offset[i])) return GK_ERR_ARG;
    const uint64_t n = uint64_t(g.size_xy) * g.size_xy * g.size_z;
    if (n > (1ull << 26)) return GK_ERR_LIMIT;
    auto& P = c.probes;
    GK_HIP(hipStreamSynchronize(c.stream));
    P.count = 0;
    GK_HIP(P.cubes.resize(n));
    GK_HIP(P.next.resize(n));
    GK_HIP(P.voxels.resize(n));
    k_zero_u32<<<nblk(n * 14), 256, 0, c.stream>>>(reinterpret_cast<uint32_t*>(P.cubes.ptr), n * 14);
    k_zero_u32<<<nblk(n * 4), 256, 0, c.stream>>>(reinterpret_cast<uint32_t*>(P.voxels.ptr), n * 4);
    GK_HIP(hipGetLastError());
    P.grid = g;
    P.count = uint32_t(n);
    return GK_OK;
}

int probes_voxelize(Ctx& c) {
    auto& P = c.probes;
    if (P.count == 0 || !c.scene_ready) return GK_ERR_STATE;
    int st = reserve_trace_events(c);
    if (st != GK_OK) return st;
    const uint32_t chunk = std::min(kChunk, P.count);
    auto& W = P.w[0];
    const size_t nr = size_t(chunk) * kVoxRays;
    GK_HIP(W.ro.resize(std::max(W.ro.count, nr)));
    GK_HIP(W.rd.resize(std::max(W.rd.count, nr)));
    GK_HIP(W.hit.resize(std::max(W.hit.count, nr)));
    GK_HIP(W.hit_inst.resize(std::max(W.hit_inst.count, nr)));
    const GridK g = grid_k(c);
    const DScene s = make_dscene(c);
    for (uint32_t first = 0; first < P.count; first += chunk) {
        const uint32_t n = std::min(chunk, P.count - first);
        k_vox_rays<<<nblk(size_t(n) * kVoxRays), 256, 0, c.stream>>>(g, first, n, W.ro.ptr, W.rd.ptr);
        GK_HIP(hipGetLastError());
        st = launch_trace(c, W.ro.ptr, W.rd.ptr, nullptr, n * kVoxRays, W.hit.ptr, W.hit_inst.ptr, false, nullptr);
        if (st != GK_OK) return st;
        k_vox_finish<<<nblk(n), 256, 0, c.stream>>>(s, g, first, n, W.rd.ptr, W.hit.ptr, W.hit_inst.ptr,
                                                    c.tri_rec.ptr, P.voxels.ptr);
        GK_HIP(hipGetLastError());
        st = reserve_trace_events(c);
        if (st != GK_OK) return st;
    }
    return GK_OK;
}

int probes_bake(Ctx& c, uint32_t first, uint32_t count) {
    auto& P = c.probes;
    if (P.count == 0 || !c.scene_ready || !c.frame_ready) return GK_ERR_STATE;
    if (first > P.count || count > P.count - first) return GK_ERR_ARG;
    if (count == 0) return GK_OK;
    int st = reserve_trace_events(c);
    if (st != GK_OK) return st;
    const gk_frame_params& fp = c.params;
    BakeK K{};
    K.g = grid_k(c);
    memcpy(K.sky, fp.sky_color, 16);
    memcpy(K.sky_sh, c.sky_sh, sizeof(K.sky_sh));
    K.sky_rot = fp.sky_rotation;
    K.sky_image = c.sky_w ? 1u : 0u;
    memcpy(K.sun_dir, fp.sun_direction, 16);
    memcpy(K.sun_col, fp.sun_color, 16);
    K.has_sky = fp.has_sky ? 1u : 0u;
    K.has_sun = fp.has_sun ? 1u : 0u;
    K.has_light = c.h_lights.empty() ? 0u : 1u;
    if (K.has_light) {
        const gk_light& L = c.h_lights[0];
        // lerp(p1, p3, 0.5)
        for (int i = 0; i < 3; ++i) K.light_pos[i] = L.p1[i] + (L.p3[i] - L.p1[i]) * 0.5f;
        K.light_area = L.normal_area[3];
        const uint32_t nm = uint32_t(c.h_materials.size());
        K.light_mat = L.light_mat_idx > nm - 1u ? nm - 1u : L.light_mat_idx;
    }
    K.nany = K.has_light + K.has_sun;
    const uint32_t chunk = std::min(kChunk, count);
    // Consecutive chunks alternate between two work sets, trace slots and streams, so one
    // chunk's probe lookups and its trace's tail overlap the next chunk's trace. Chunks touch
    // disjoint probes (their list, next and face results); k_bake_apply waits for both.
    const bool two = count > chunk;
    const int nw = two ? 2 : 1;
    const size_t nr = size_t(chunk) * kProbeRays;
    const size_t ns = size_t(chunk) * 6u * std::max(1u, K.nany);
    for (int h = 0; h < nw; ++h) {
        auto& W = P.w[h];
        GK_HIP(W.list.resize(std::max<size_t>(W.list.count, chunk)));
        GK_HIP(W.counters.resize(4));
        GK_HIP(W.face_sky.resize(std::max<size_t>(W.face_sky.count, size_t(chunk) * 6)));
        GK_HIP(W.ro.resize(std::max(W.ro.count, nr)));
        GK_HIP(W.rd.resize(std::max(W.rd.count, nr)));
        GK_HIP(W.hit.resize(std::max(W.hit.count, nr)));
        GK_HIP(W.hit_inst.resize(std::max(W.hit_inst.count, nr)));
        GK_HIP(W.so.resize(std::max(W.so.count, ns)));
        GK_HIP(W.sd.resize(std::max(W.sd.count, ns)));
        GK_HIP(W.shit.resize(std::max(W.shit.count, ns)));
        if (!trace_counter(c, h)) return GK_ERR_OOM;  // a slot's first use clears its counters on `stream`
    }
    hipStream_t ss[2] = {c.stream, c.stream};
    if (two) {
        if (!c.fork_ev) GK_HIP(hipEventCreateWithFlags(&c.fork_ev, hipEventDisableTiming));
        if (!c.pstream[0]) {
            GK_HIP(hipStreamCreateWithFlags(&c.pstream[0], hipStreamNonBlocking));
            GK_HIP(hipEventCreateWithFlags(&c.join_ev[0], hipEventDisableTiming));
        }
        ss[1] = c.pstream[0];
        GK_HIP(hipEventRecord(c.fork_ev, c.stream));
        GK_HIP(hipStreamWaitEvent(ss[1], c.fork_ev, 0));
    }
    c.trace_split = two;
    const DScene s = make_dscene(c);
    const ProbeDev PD = probe_dev(c);
    int h = 0;
    for (uint32_t c0 = first; c0 < first + count; c0 += chunk) {
        const uint32_t n = std::min(chunk, first + count - c0);
        auto& W = P.w[h];
        hipStream_t st_h = ss[h];
        k_zero_u32<<<1, 64, 0, st_h>>>(W.counters.ptr, 4);
        k_bake_select<<<nblk(n), 256, 0, st_h>>>(P.voxels.ptr, c0, n, W.list.ptr, W.counters.ptr);
        k_bake_counts<<<1, 1, 0, st_h>>>(W.counters.ptr, K.nany);
        k_bake_rays<<<nblk(size_t(n) * kProbeRays), 256, 0, st_h>>>(K, P.voxels.ptr, W.list.ptr, W.counters.ptr,
                                                                     W.ro.ptr, W.rd.ptr, W.so.ptr, W.sd.ptr);
        GK_HIP(hipGetLastError());
        if (K.nany)
            st = launch_trace_pair(c, W.ro.ptr, W.rd.ptr, W.counters.ptr + 1, n * kProbeRays, W.hit.ptr,
                                   W.hit_inst.ptr, W.so.ptr, W.sd.ptr, W.counters.ptr + 2, n * 6u * K.nany,
                                   W.shit.ptr, h);
        else
            st = launch_trace(c, W.ro.ptr, W.rd.ptr, W.counters.ptr + 1, n * kProbeRays, W.hit.ptr, W.hit_inst.ptr,
                              false, nullptr, h);
        if (st != GK_OK) return st;
        k_bake_face<<<nblk(size_t(n) * kProbeRays), 256, 0, st_h>>>(s, K, PD, P.voxels.ptr, W.list.ptr, W.counters.ptr,
                                                             W.ro.ptr, W.rd.ptr, W.hit.ptr, W.hit_inst.ptr,
                                                             W.shit.ptr, P.next.ptr, W.face_sky.ptr);
        k_bake_commit<<<nblk(n), 256, 0, st_h>>>(W.list.ptr, W.counters.ptr, P.cubes.ptr, W.face_sky.ptr,
                                                 P.next.ptr, P.voxels.ptr);
        GK_HIP(hipGetLastError());
        st = reserve_trace_events(c);
        if (st != GK_OK) return st;
        h = two ? h ^ 1 : 0;
    }
    c.trace_split = false;
    if (two) {
        GK_HIP(hipEventRecord(c.join_ev[0], ss[1]));
        GK_HIP(hipStreamWaitEvent(c.stream, c.join_ev[0], 0));
    }
    k_bake_apply<<<nblk(count), 256, 0, c.stream>>>(P.voxels.ptr, first, count, P.next.ptr, P.cubes.ptr);
    GK_HIP(hipGetLastError());
    return GK_OK;
}

}  // namespace gk

extern "C" {

int gk_probes_init(gk_ctx* ctx, const gk_probe_grid* grid) {
    GK_CHECK_CTX(ctx);
    gk::Ctx& c = ctx->c;
    GK_HIP(hipSetDevice(c.device));
    const gk_probe_grid def{192u, 48u, 0.25f, {-24.0f, -0.34375f, -24.0f}};
    return gk::probes_init(c, grid ? *grid : def);
}

int gk_probes_voxelize(gk_ctx* ctx) {
    GK_CHECK_CTX(ctx);
    GK_HIP(hipSetDevice(ctx->c.device));
    return gk::probes_voxelize(ctx->c);
}

int gk_probes_bake(gk_ctx* ctx, uint32_t first, uint32_t count) {
    GK_CHECK_CTX(ctx);
    GK_HIP(hipSetDevice(ctx->c.device));
    return gk::probes_bake(ctx->c, first, count);
}

int gk_probes_bake_frame(gk_ctx* ctx, uint32_t frame_count, uint32_t temporal_frames) {
    GK_CHECK_CTX(ctx);
    if (temporal_frames == 0) return GK_ERR_ARG;
    const uint32_t n = ctx->c.probes.count;
    if (n == 0) return GK_ERR_STATE;
    const uint32_t groups = n / 64u;
    const uint32_t per_frame = groups / temporal_frames;
    const uint32_t first = (frame_count % temporal_frames) * per_frame * 64u;
    return gk_probes_bake(ctx, first, per_frame * 64u);
}

int gk_probes_read(gk_ctx* ctx, gk_ambient_cube* cubes, gk_voxel* voxels, uint32_t first, uint32_t count) {
    GK_CHECK_CTX(ctx);
    gk::Ctx& c = ctx->c;
    if (c.probes.count == 0) return GK_ERR_STATE;
    if (first > c.probes.count || count > c.probes.count - first) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(c.device));
    if (cubes)
        GK_HIP(hipMemcpyAsync(cubes, c.probes.cubes.ptr + first, sizeof(gk_ambient_cube) * size_t(count),
                              hipMemcpyDeviceToHost, c.stream));
    if (voxels)
        GK_HIP(hipMemcpyAsync(voxels, c.probes.voxels.ptr + first, sizeof(gk_voxel) * size_t(count),
                              hipMemcpyDeviceToHost, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

int gk_probes_write(gk_ctx* ctx, const gk_ambient_cube* cubes, const gk_voxel* voxels, uint32_t first,
                    uint32_t count) {
    GK_CHECK_CTX(ctx);
    gk::Ctx& c = ctx->c;
    if (c.probes.count == 0) return GK_ERR_STATE;
    if (first > c.probes.count || count > c.probes.count - first) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(c.device));
    if (cubes)
        GK_HIP(hipMemcpyAsync(c.probes.cubes.ptr + first, cubes, sizeof(gk_ambient_cube) * size_t(count),
                              hipMemcpyHostToDevice, c.stream));
    if (voxels)
        GK_HIP(hipMemcpyAsync(c.probes.voxels.ptr + first, voxels, sizeof(gk_voxel) * size_t(count),
                              hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

}  // extern "C"

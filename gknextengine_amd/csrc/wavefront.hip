// wavefront.hip - wavefront path tracer (the "rt pass" of PathTracingRenderer::Render,
// PathTracingRenderer.cpp:104-218, whose megakernel is Core.PathTracing.comp.slang:5-31).
//
// One path per pixel lives in SoA state buffers; each bounce is
//     trace(extension queue) -> [sort by material] -> shade -> trace(shadow queue) -> resolve
// with queues compacted by wave-level ballots (one atomic per wave), so that no host
// synchronisation happens inside a frame. Two estimators:
//   REF_V2  FPathTracingRendererV2 (Shading.slang:995-1241) step by step: its
//           GetRayColor is split at the trace call into "begin" (material, RNG draws,
//           new direction) and "finish" (hit albedo, back-face and emitter rules); the
//           sample loop, sun cone, early exit and DirectIlluminate follow Render.
//           A terminated path is multiplied by the ambient-cube probe lookup
//           (gk_probe.hpp); without a probe grid that lookup is (0,0,0,1).
//   PT_MIS  GGX (bounded VNDF) + Lambert with next-event estimation to emissive
//           triangles, power-heuristic MIS, Russian roulette after 3 bounces.
// Every RNG draw happens in the same order as in the per-pixel formulation, so a path's
// random numbers do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "gk_probe.hpp"
#include "gk_ptmis.hpp"

namespace gk {

using ptmis::Bsdf;
using ptmis::bsdf_eval;
using ptmis::bsdf_of;
using ptmis::light_pick;
using ptmis::light_select_pdf;
using ptmis::power_h;
using ptmis::vndf_sample;

int post_process(Ctx& c);  // post.hip
ProbeDev probe_dev(const Ctx& c);  // probe.hip

namespace {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kEps = 1e-3f;                 // PreProcessor.slang:13
constexpr float kMaxTrace = 1000.f;           // Shading.slang:15
constexpr uint32_t kSplitPaths = 1u << 18;    // PT_MIS frames of at least this many paths are split
// path pipelines of a split frame (trace slots allow 4; 1080p city, ms per frame: 2 -> 2.98,
// 3 -> 3.08, 4 -> 4.43)
constexpr int kPathParts = 2;
constexpr uint32_t kMiss = 0xffffffffu;

// per-frame constants
struct FrameK {
    float mvi[16], pi[16], vp[16], vpu[16], pvpu[16];
    float sun_dir[4], sun_col[4], sky[4];
    float exit_prob, hit_off;
    float aperture, focus;  // thin lens (aperture 0: pinhole, no random numbers drawn)
    float sky_rot;          // SkyRotation
    uint32_t W, H, frame, seed, samples, bounces, max_bounces, temporal, has_sky, has_sun, tiled;
    uint32_t tparts;      // tiled: the path ids fall into tparts parts, part h covering the
    uint32_t tstart[4];   // 8x8 tiles t = h mod tparts from path tile tstart[h] on
    uint32_t vis_primary;  // primary hits come from GK_RT_MINIGBUFFER (left unwritten)
    uint32_t fast_gather;  // REF_V2 ExitAfterFirst
    ProbeDev probes;       // REF_V2 path termination
};

struct Targets {
    ushort4* diffuse;
    ushort4* spec;
    ushort4* albedo;
    ushort4* normal;
    uint32_t* objid;
    float2* motion;
    uint16_t* moment;
    float* depth;
    uint32_t* minigbuffer;
    float4* phit;
    uint32_t* pinst;
    float4* radiance;
    double4* radiance64;  // written when fp64 != 0
    uint32_t fp64;
};

struct Queues {
    float4* o[2];
    float4* d[2];
    uint32_t* path[2];
    float4* hit;
    uint32_t* hit_inst;
    float4* so;
    float4* sd;
    float4* sc_d;   // shadow contribution to Ld (rgb)
    float4* sc_s;   // shadow contribution to Ls (rgb)
    uint32_t* spath;
    float4* shit;   // occlusion result per shadow entry (w = 1 when occluded)
    // queue sizes: extension queue k at word 2k, the shadow queue of the bounce that fills
    // extension queue k at 2k + 1, so a bounce's two appends share one 8-byte word
    uint32_t* cnt;
    uint32_t* sort_keys;    // material class of each queue entry
    uint32_t* sorted;       // queue entries in material order (nullptr: path order)
    uint32_t* sort_cnt;     // 16 class counts, then 16 cursors
    // PT_MIS path state stored with each queue entry (read where the hit is read, written
    // where the extension ray is appended), so shading reads and writes it coalesced
    float4* qthr[2];        // throughput rgb, pdf of the last bsdf sample (w)
    uint4* qrng[2];
    uint4* qmisc[2];        // x pixel, y flags
};

struct PathState {
    float4* thr;    // REF_V2: RayColor; PT_MIS: throughput + last pdf (w)
    float4* ld;     // PT_MIS: primary-diffuse radiance; REF_V2: FinalColor
    float4* ls;     // PT_MIS: other radiance; REF_V2: FinalReflection
    uint4* rng;
    uint4* misc;    // x pixel, y flags, z counters, w material of the current vertex
    float4* pos;    // current vertex position (w: primary material bits for REF_V2)
    float4* nrm;    // current vertex normal
    float4* dir;    // current incoming direction
    float4* ppos;   // REF_V2 primary vertex (pulled back), w = primary material bits
    float4* pnrm;   // REF_V2 primary normal
    float4* palb;   // REF_V2 primary albedo (FetchGBufferV2: textured)
};

// Path p -> pixel. Tiled frames walk 8x8 tiles, 64 paths each, in parts: part h of the path
// ids covers the tiles whose row-major index is h modulo the part count, so the pipelines of
// a split frame (one per part) see the same mix of the image.
__device__ inline uint32_t pixel_of(const FrameK& K, uint32_t p) {
    if (!K.tiled) return p;
    const uint32_t t = p >> 6, within = p & 63u;
    uint32_t h = 0;
    for (uint32_t k = 1; k < K.tparts; ++k) h = t >= K.tstart[k] ? k : h;
    const uint32_t tile = (t - K.tstart[h]) * K.tparts + h;
    const uint32_t tx_n = K.W >> 3;
    const uint32_t tx = tile % tx_n, ty = tile / tx_n;
    const uint32_t x = (tx << 3) | (within & 7u), y = (ty << 3) | (within >> 3);
    return y * K.W + x;
}

// FVisibilityBufferRayCasterV2 ray (Shading.slang:350-356): uv = ipos/size*2-1
__device__ inline void camera_ray(const FrameK& K, uint32_t x, uint32_t y, f3& O, f3& D) {
    const float ux = (float(x) / float(K.W)) * 2.0f - 1.0f;
    const float uy = (float(y) / float(K.H)) * 2.0f - 1.0f;
    float o[4], t[4], d[4];
    mul44v(K.mvi, 0.f, 0.f, 0.f, 1.f, o);
    mul44v(K.pi, ux, uy, 1.f, 1.f, t);
    const f3 tn = normalize(f3{t[0], t[1], t[2]});
    mul44v(K.mvi, tn.x, tn.y, tn.z, 0.f, d);
    O = f3{o[0], o[1], o[2]};
    D = normalize(f3{d[0], d[1], d[2]});
}

// The thin-lens ray of FHardwarePrimaryRayCaster::TracePrimaryRay (Shading.slang:283-300):
// a lens offset Aperture / 2 * RandomInUnitDisk, origin MVinv (offset, 0, 1), direction
// MVinv normalize(Pinv (uv, 1, 1) * FocusDistance - (offset, 0)). The disk draw is the
// path's first random number. With aperture 0 the pinhole ray is taken and nothing is drawn,
// as the live visibility-buffer caster draws only for a circle of confusion (:372-379).
__device__ inline void camera_ray_lens(const FrameK& K, uint32_t x, uint32_t y, rng4& seed, f3& O, f3& D) {
    if (!(K.aperture > 0.f)) {
        camera_ray(K, x, y, O, D);
        return;
    }
    float a, b, dx, dy;
    rnd2(seed, a, b);
    concentric_disk(a, b, dx, dy);
    const float half = K.aperture / 2.f;
    const float lx = half * dx, ly = half * dy;
    const float ux = (float(x) / float(K.W)) * 2.0f - 1.0f;
    const float uy = (float(y) / float(K.H)) * 2.0f - 1.0f;
    float o[4], t[4], d[4];
    mul44v(K.mvi, lx, ly, 0.f, 1.f, o);
    mul44v(K.pi, ux, uy, 1.f, 1.f, t);
    const f3 tl = normalize(f3{t[0] * K.focus - lx, t[1] * K.focus - ly, t[2] * K.focus - 0.f});
    mul44v(K.mvi, tl.x, tl.y, tl.z, 0.f, d);
    O = f3{o[0], o[1], o[2]};
    D = normalize(f3{d[0], d[1], d[2]});
}

// Queue appends. All queue counters share one cache line, and per-wave atomics on it
// serialised the shading kernels (1080p city: 1.2 ms per frame), so a 256-thread
// work-group combines its wave counts in LDS and issues one atomic per counter; waves
// take consecutive ranges. Every thread of the group must call these.
//
// Two appends (extension and shadow queue) at once:
// Appends to two queues from a 256-thread work-group with one 64-bit atomic on the pair of
// 32-bit sizes at pair[0] (a) and pair[1] (b) (8-byte aligned; a's size stays below 2^32,
// so the low half never carries): sa, sb = this thread's slots. Every thread must call it.
__device__ inline void block_append_pair(uint32_t* pair, bool wa, uint32_t& sa, bool wb, uint32_t& sb) {
    __shared__ uint32_t cnt[10];
    const uint64_t ma = __ballot(wa), mb = __ballot(wb);
    const uint32_t wv = threadIdx.x >> 6;
    if (__lane_id() == 0) { cnt[wv] = uint32_t(__popcll(ma)); cnt[4 + wv] = uint32_t(__popcll(mb)); }
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t ta = cnt[0] + cnt[1] + cnt[2] + cnt[3], tb = cnt[4] + cnt[5] + cnt[6] + cnt[7];
        uint64_t old = 0u;
        if (ta | tb)
            old = atomicAdd(reinterpret_cast<unsigned long long*>(pair),
                            (static_cast<unsigned long long>(tb) << 32) | ta);
        cnt[8] = uint32_t(old);
        cnt[9] = uint32_t(old >> 32);
    }
    __syncthreads();
    uint32_t ba = cnt[8], bb = cnt[9];
    for (uint32_t w = 0; w < wv; ++w) { ba += cnt[w]; bb += cnt[4 + w]; }
    const uint64_t lt = (1ull << __lane_id()) - 1ull;
    sa = ba + uint32_t(__popcll(ma & lt));
    sb = bb + uint32_t(__popcll(mb & lt));
}

// One append for a 256-thread work-group (one atomic); every thread must call it.
__device__ inline uint32_t block_append(uint32_t* counter, bool want) {
    __shared__ uint32_t cnt[5];
    const uint64_t m = __ballot(want);
    const uint32_t wv = threadIdx.x >> 6;
    if (__lane_id() == 0) cnt[wv] = uint32_t(__popcll(m));
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t t = cnt[0] + cnt[1] + cnt[2] + cnt[3];
        cnt[4] = t ? atomicAdd(counter, t) : 0u;
    }
    __syncthreads();
    uint32_t b = cnt[4];
    for (uint32_t w = 0; w < wv; ++w) b += cnt[w];
    return b + uint32_t(__popcll(m & ((1ull << __lane_id()) - 1ull)));
}

// ---------------------------------------------------------------- sampling (REF_V2, float)
__device__ inline float schlick(float cosine, float ior) {
    float r0 = (1.f - ior) / (1.f + ior);
    r0 = r0 * r0;
    const float x = 1.f - cosine;
    return r0 + (1.f - r0) * (x * x * x * x * x);
}
__device__ inline f3 random_hemisphere1(rng4& s) {
    float ux, uy;
    rnd2(s, ux, uy);
    float sp, cp;
    sincos_2pi(ux, sp, cp);
    const float r = sqrtf(uy);
    return f3{r * cp, r * sp, sqrtf(1.0f - uy)};
}
__device__ inline f3 random_in_cone(rng4& s, float cos_theta) {
    float ux, uy;
    rnd2(s, ux, uy);
    float sp, cp;
    sincos_2pi(ux, sp, cp);
    const float c = 1.0f + uy * (cos_theta - 1.f);
    const float r = sqrtf(1.0f - c * c);
    return f3{r * cp, r * sp, c};
}
__device__ inline float saturate(float x) { return x < 0.f ? 0.f : (x > 1.f ? 1.f : x); }
__device__ inline f3 ggx_sample_vndf(float alpha, f3 wi_, float ux, float uy) {
    const f3 wi = normalize(f3{wi_.x * alpha, wi_.y * alpha, wi_.z});
    float b = wi.z;
    if (wi_.z > 0.f) {
        const float a = saturate(alpha);
        const float awiz_s = a * wi_.z / (1.0f + sqrtf(wi_.x * wi_.x + wi_.y * wi_.y));
        b = b * ((1.0f - a * a) / (1.0f + awiz_s * awiz_s));
    }
    const float z = fmaf(1.0f - uy, 1.0f + b, -b);
    float sp, cp;
    sincos_2pi(ux, sp, cp);
    const float r = sqrtf(saturate(1.0f - z * z));
    const f3 m = wi + f3{r * cp, r * sp, z};
    return normalize(f3{m.x * alpha, m.y * alpha, m.z});
}
__device__ inline f3 ggx_sampling(rng4& s, float roughness, f3 normal) {
    f3 T, B;
    onb(normal, T, B);
    const float alpha = roughness * roughness;
    const f3 wl = f3{dot(T, normal), dot(B, normal), dot(normal, normal)};
    float ux, uy;
    rnd2(s, ux, uy);
    return to_world(ggx_sample_vndf(alpha, wl, ux, uy), T, B, normal);
}
__device__ inline f3 reflect(f3 i, f3 n) { return i - n * (2.f * dot(n, i)); }
__device__ inline f3 refract(f3 i, f3 n, float eta) {
    const float d = dot(n, i);
    const float k = 1.f - eta * eta * (1.f - d * d);
    if (k < 0.f) return f3{0.f, 0.f, 0.f};
    return i * eta - n * (eta * d + sqrtf(k));
}

// get_material_dataV2 normal (GeneralFunc.slang:20-53): bx*n0 + (by*n1 + bz*n2), world normals
__device__ inline f3 primary_normal(const DScene& s, uint32_t inst, uint32_t prim, float u, float v) {
    const InstDev& I = s.inst[inst];
    const size_t b = 3 * size_t(I.tri_base + prim);
    float a[4], bb[4], cc[4];
    const float4 n0 = s.tri_nrm[b], n1 = s.tri_nrm[b + 1], n2 = s.tri_nrm[b + 2];
    mul44v(I.world, n0.x, n0.y, n0.z, 0.f, a);
    mul44v(I.world, n1.x, n1.y, n1.z, 0.f, bb);
    mul44v(I.world, n2.x, n2.y, n2.z, 0.f, cc);
    const float by = u, bz = v, bx = 1.0f - (by + bz);
    const f3 nb = f3{a[0], a[1], a[2]} * bx + (f3{bb[0], bb[1], bb[2]} * by + f3{cc[0], cc[1], cc[2]} * bz);
    return normalize(nb);
}

__device__ inline void mul44m(const float* a, const float* b, float* o) {
    for (int cc = 0; cc < 4; ++cc)
        for (int r = 0; r < 4; ++r) {
            float acc = a[r] * b[cc * 4 + 0];
            acc = acc + a[4 + r] * b[cc * 4 + 1];
            acc = acc + a[8 + r] * b[cc * 4 + 2];
            acc = acc + a[12 + r] * b[cc * 4 + 3];
            o[cc * 4 + r] = acc;
        }
}

// PrimaryHit G-buffer (Shading.slang:1204-1215) + CalculateMotionVectorV2 (:50-72)
__device__ inline void write_gbuffer(const DScene& s, const FrameK& K, const Targets& T, uint32_t px, uint32_t inst,
                                     uint32_t prim, float4 hit, f3 pos, f3 nrm, float4 albedo, float rough,
                                     uint32_t& moment_out) {
    float clip[4];
    mul44v(K.vp, pos.x, pos.y, pos.z, 1.f, clip);
    T.depth[px] = clip[2] / clip[3];
    const InstDev& I = s.inst[inst];
    float ch[4], ph[4], pp[4], mm[16];
    mul44v(K.vpu, pos.x, pos.y, pos.z, 1.f, ch);
    mul44m(K.pvpu, I.combined, mm);
    mul44v(mm, pos.x, pos.y, pos.z, 1.f, ph);
    const float cfx = ch[0] / ch[3] * 0.5f, cfy = ch[1] / ch[3] * 0.5f;
    const float pfx = ph[0] / ph[3] * 0.5f, pfy = ph[1] / ph[3] * 0.5f;
    mul44v(I.combined, pos.x, pos.y, pos.z, 1.f, pp);
    const f3 dpos = pos - f3{pp[0], pp[1], pp[2]};
    uint32_t mo = T.moment[px];
    if (length(dpos) > 0.02f) mo = K.temporal;
    else mo = mo > 0u ? mo - 1u : 0u;
    T.moment[px] = uint16_t(mo);
    moment_out = mo;
    T.motion[px] = make_float2((pfx - cfx) * float(K.W), (pfy - cfy) * float(K.H));
    T.albedo[px] = to_h4(albedo.x, albedo.y, albedo.z, albedo.w);
    T.normal[px] = to_h4(nrm.x, nrm.y, nrm.z, rough);
    T.objid[px] = I.instance_id;
    if (!K.vis_primary) {
        // the reference's packing of (NodeProxy + 1, triangle of the section); 0 (no geometry)
        // for a proxy the 15-bit field cannot name
        const uint32_t proxy = I.proxy_base + prim / kSectionTris;
        T.minigbuffer[px] = proxy < kMaxVisProxies ? ((proxy + 1u) << 17) | (prim % kSectionTris) : 0u;
    }
    T.phit[px] = hit;
    T.pinst[px] = inst;
}

__device__ inline void write_miss_gbuffer(const FrameK& K, const Targets& T, uint32_t px, float4 hit) {
    T.motion[px] = make_float2(0.f, 0.f);
    T.albedo[px] = to_h4(1.f, 1.f, 1.f, 1.f);
    T.normal[px] = to_h4(0.f, 1.f, 0.f, 1.f);
    T.objid[px] = 65535u;
    T.depth[px] = 1e5f;
    if (!K.vis_primary) T.minigbuffer[px] = 0u;
    T.phit[px] = hit;
    T.pinst[px] = kMiss;
}

// ---------------------------------------------------------------- raygen
// queue_state: PT_MIS keeps its per-path state with the queue entries (Queues::q*),
// REF_V2 in PathState
// camera rays of paths p_base .. p_base + n - 1 into queue entries 0 .. n - 1; also zeroes the
// queue counters and the ray counter of the trace launch that follows (trace_next)
__global__ void k_raygen(FrameK K, Queues Q, PathState S, uint32_t n, uint32_t queue_state, uint32_t p_base,
                         uint32_t* __restrict__ trace_next) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    // queue 0 (word 0) holds the n primary rays; queue 1 (word 2) and both shadow sizes are 0
    if (i == 0) { Q.cnt[0] = n; Q.cnt[1] = 0u; Q.cnt[2] = 0u; Q.cnt[3] = 0u; *trace_next = 0u; }
    if (i >= n) return;
    const uint32_t p = p_base + i;
    const uint32_t px = pixel_of(K, p);
    const uint32_t x = px % K.W, y = px / K.W;
    f3 O, D;
    rng4 seed{x, y, K.frame, K.seed};
    camera_ray_lens(K, x, y, seed, O, D);
    Q.o[0][i] = make_float4(O.x, O.y, O.z, 0.f);
    Q.d[0][i] = make_float4(D.x, D.y, D.z, 1e30f);
    Q.path[0][i] = p;
    if (queue_state) {
        Q.qrng[0][i] = make_uint4(seed.x, seed.y, seed.z, seed.w);
        Q.qmisc[0][i] = make_uint4(px, 0u, 0u, 0u);
        Q.qthr[0][i] = make_float4(1.f, 1.f, 1.f, 0.f);
    } else {
        S.rng[p] = make_uint4(seed.x, seed.y, seed.z, seed.w);
        S.misc[p] = make_uint4(px, 0u, 0u, 0u);
        S.thr[p] = make_float4(1.f, 1.f, 1.f, 0.f);
    }
    S.ld[p] = make_float4(0.f, 0.f, 0.f, 0.f);
    S.ls[p] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Primary hits from a visibility buffer (GK_FRAME_PRIMARY_FROM_VISBUFFER): the camera ray of
// path p hits the triangle GK_RT_MINIGBUFFER names for its pixel ((proxy+1) << 17 | tri, 0 = no
// geometry; Rast.VisibilityPass.vert.slang:18-22, where proxy numbers the (instance, section)
// pairs and tri counts within the section's 65535 triangles), as in FVisibilityBufferRayCasterV2::
// TracePrimaryRay (Shading.slang:336-370). The barycentrics are get_material_dataV2's
// Moeller-Trumbore form without rejection (GeneralFunc.slang:31-45); t follows the same
// form. They are evaluated in the triangle's object space with the tracer's arithmetic
// (the ray moved by the instance inverse as tinybvh does), so a buffer written by a traced
// frame of this core reproduces that frame's primary hits bit for bit.
__global__ void k_vis_primary(DScene s, FrameK K, Queues Q, const uint32_t* __restrict__ vis,
                              const uint32_t* __restrict__ tri_rec, const MeshInfo* __restrict__ meshes,
                              const uint32_t* __restrict__ proxy_inst, uint32_t nproxy, uint32_t n,
                              uint32_t p_base) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t px = pixel_of(K, p_base + i);
    const uint32_t pk = vis[px];
    const uint32_t ip1 = (pk >> 17) & 0x7fffu, stri = pk & 0x1ffffu;
    float4 hit = make_float4(__builtin_inff(), 0.f, 0.f, __uint_as_float(kMiss));
    uint32_t hinst = kMiss;
    if (ip1 != 0u && ip1 - 1u < nproxy && stri < kSectionTris) {
        const uint32_t inst = proxy_inst[ip1 - 1u];
        const InstDev& I = s.inst[inst];
        const uint32_t tri = (ip1 - 1u - I.proxy_base) * kSectionTris + stri;  // within the mesh
        if (tri < meshes[I.mesh].tri_count) {
            f3 O = ld3(Q.o[0][i]), D = ld3(Q.d[0][i]);
            if (s.two_level) {
                const float* M = I.inv;
                const f3 r = f3{M[0] * O.x + M[1] * O.y + M[2] * O.z + M[3], M[4] * O.x + M[5] * O.y + M[6] * O.z + M[7],
                                M[8] * O.x + M[9] * O.y + M[10] * O.z + M[11]};
                const float w = M[12] * O.x + M[13] * O.y + M[14] * O.z + M[15];
                const f3 d = f3{M[0] * D.x + M[1] * D.y + M[2] * D.z, M[4] * D.x + M[5] * D.y + M[6] * D.z,
                                M[8] * D.x + M[9] * D.y + M[10] * D.z};
                O = w == 1.0f ? r : r * (1.f / w);
                D = d;
            }
            const TriRec tr = s.tris[I.tri_base + tri_rec[I.tri_base + tri]];
            const f3 v0 = ld3(tr.v0), e1 = ld3(tr.e1), e2 = ld3(tr.e2);
            const f3 h = cross(D, e2);
            const float f = 1 / dot(e1, h);
            const f3 sv = O - v0;
            const float u = f * dot(sv, h);
            const f3 q = cross(sv, e1);
            const float v = f * dot(D, q);
            const float t = f * dot(e2, q);
            hit = make_float4(t, u, v, __uint_as_float(__float_as_uint(tr.v0.w)));
            hinst = s.two_level ? inst : 0u;
        }
    }
    Q.hit[i] = hit;
    Q.hit_inst[i] = hinst;
}

// ================================================================ REF_V2
// flags (misc.y): bit0 chanceReflect of the pending GetRayColor, bit1 BackFace,
// bit2 origin material is dielectric, bit3 hitReflect (first call), bit4 hitMetal (first call)
// counters (misc.z): bits 0-7 bounce b of the pending call, 8-19 sample, 20-31 sample count - 1
// (gk_frame allows 1024 samples, x4 on moving pixels: 4096 samples, indices 0..4095)

// GetRayColor up to the trace (Shading.slang:1007-1036); emits the extension ray.
__device__ inline void ref_begin(const DScene& s, const FrameK& K, const Queues& Q, const PathState& S, uint32_t p,
                                 f3 pos, f3 nrm, uint32_t mi, f3 dir, rng4& seed, uint32_t& flags, uint32_t b, int nq,
                                 bool& emit, float4& ro, float4& rd) {
    const gk_material& mat = s.mats[mi];
    const bool diel = mat.material_model == GK_MATERIAL_DIELECTRIC;
    const float start_off = diel ? 0.0f : 1.0f;
    const float roughness = mat.fuzziness;
    const float dv = dot(dir, nrm);
    const bool back = dv > 0.f;
    const f3 outward = back ? -nrm : nrm;
    const float ni_over_nt = back ? mat.refraction_index2 : (1.f / mat.refraction_index2);
    const float cosine = dv > 0.f ? mat.refraction_index * dv : -dv;
    const float reflect_prob = schlick(cosine, mat.refraction_index);
    const float metal_prob = mat.metalness;
    const bool chance_reflect = rnd1(seed) < reflect_prob;
    const bool chance_metal = rnd1(seed) < metal_prob;
    const bool chance_ggx = chance_reflect || chance_metal;
    const f3 trace_next = chance_ggx ? reflect(dir, outward) : outward;
    f3 trace_dir = chance_ggx ? ggx_sampling(seed, sqrtf(roughness), trace_next)
                              : align_with_normal(random_hemisphere1(seed), trace_next);
    if (diel && !chance_reflect) trace_dir = refract(dir, outward, ni_over_nt);
    flags &= ~7u;
    flags |= (chance_reflect ? 1u : 0u) | (back ? 2u : 0u) | (diel ? 4u : 0u);
    if (b == 0u) {
        flags &= ~(8u | 16u);
        flags |= (chance_ggx ? 8u : 0u) | (chance_metal ? 16u : 0u);
    }
    const f3 org = pos + (nrm * K.hit_off) * start_off;
    S.dir[p] = make_float4(trace_dir.x, trace_dir.y, trace_dir.z, 0.f);
    ro = make_float4(org.x, org.y, org.z, kEps);
    rd = make_float4(trace_dir.x, trace_dir.y, trace_dir.z, kMaxTrace);
    emit = true;
    (void)nq;
    (void)Q;
}

// end of a sample (Shading.slang:1159-1171); restarts the next sample or finalizes
__device__ inline void ref_end_sample(const DScene& s, const FrameK& K, const Targets& T, const PathState& S,
                                      uint32_t p, uint32_t px, float4 color, uint32_t& flags, uint32_t& counters,
                                      rng4& seed, bool& emit, float4& ro, float4& rd) {
    const float4 pp = S.ppos[p];
    const uint32_t pmat = __float_as_uint(pp.w);
    if (flags & 16u) {
        const float4 pa = S.palb[p];
        color.x = color.x * pa.x; color.y = color.y * pa.y;
        color.z = color.z * pa.z; color.w = color.w * pa.w;
    }
    float4 acc = (flags & 8u) ? S.ls[p] : S.ld[p];
    acc.x = acc.x + color.x; acc.y = acc.y + color.y; acc.z = acc.z + color.z; acc.w = acc.w + color.w;
    if (flags & 8u) S.ls[p] = acc; else S.ld[p] = acc;
    const uint32_t sample = ((counters >> 8) & 0xfffu) + 1u;
    const uint32_t nsamp = (counters >> 20) + 1u;
    const f3 ppos = ld3(pp), pnrm = ld3(S.pnrm[p]);
    if (sample < nsamp) {
        counters = (counters & 0xfff00000u) | (sample << 8);
        float e4[4];
        mul44v(K.mvi, 0.f, 0.f, 0.f, 1.f, e4);
        const f3 dir = normalize(ppos - f3{e4[0], e4[1], e4[2]});
        S.thr[p] = make_float4(1.f, 1.f, 1.f, 1.f);
        ref_begin(s, K, Queues{}, S, p, ppos, pnrm, pmat, dir, seed, flags, 0u, 0, emit, ro, rd);
        return;
    }
    // finalize: average, DirectIlluminate (Shading.slang:1175-1183, 888-905)
    float4 fc = S.ld[p], fr = S.ls[p];
    const float ns = float(nsamp);
    fc.x = fc.x / ns; fc.y = fc.y / ns; fc.z = fc.z / ns; fc.w = fc.w / ns;
    fr.x = fr.x / ns; fr.y = fr.y / ns; fr.z = fr.z / ns; fr.w = fr.w / ns;
    const f3 sun = f3{K.sun_dir[0], K.sun_dir[1], K.sun_dir[2]};
    const float d = fmaxf(dot(sun, normalize(pnrm)), 0.0f) * (1.0f / kPi);
    float shadow = 0.f;
    if (K.has_sun) {
        const float cos_cone = 0.99999048072578f;  // cos(0.25 deg)
        const f3 cone = align_with_normal(random_in_cone(seed, cos_cone), sun);
        shadow = occluded_single(s, ppos, cone, kEps, kMaxTrace) ? 0.f : 1.f;
    }
    fc.x = fc.x + K.sun_col[0] * d * shadow; fc.y = fc.y + K.sun_col[1] * d * shadow;
    fc.z = fc.z + K.sun_col[2] * d * shadow; fc.w = fc.w + K.sun_col[3] * d * shadow;
    T.diffuse[px] = to_h4(fc.x, fc.y, fc.z, fc.w);
    T.spec[px] = to_h4(fr.x, fr.y, fr.z, fr.w);
    float4 r = T.radiance[px];
    r.x += fc.x + fr.x; r.y += fc.y + fr.y; r.z += fc.z + fr.z; r.w += 1.f;
    T.radiance[px] = r;
    if (T.fp64) {
        double4 q = T.radiance64[px];
        q.x += double(fc.x) + double(fr.x); q.y += double(fc.y) + double(fr.y); q.z += double(fc.z) + double(fr.z);
        q.w += 1.0;
        T.radiance64[px] = q;
    }
    emit = false;
}

// trace_next: the ray counter of the trace launch that follows, zeroed here
__global__ __launch_bounds__(256) void k_ref_primary(DScene s, FrameK K, Targets T, Queues Q, PathState S,
                                                     uint32_t* __restrict__ trace_next) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) *trace_next = 0u;
    const uint32_t n = Q.cnt[0];
    const bool live = i < n;
    bool emit = false;
    float4 ro{}, rd{};
    uint32_t p = 0;
    if (live) {
        p = Q.path[0][i];
        const uint32_t px = S.misc[p].x;
        const float4 h = Q.hit[i];
        const uint32_t inst = Q.hit_inst[i];
        rng4 seed;
        const uint4 sr = S.rng[p];
        seed = rng4{sr.x, sr.y, sr.z, sr.w};
        if (__float_as_uint(h.w) == kMiss) {
            write_miss_gbuffer(K, T, px, h);
            const float4 sky = K.has_sky ? sky_radiance(s, K.sky, K.sky_rot, ld3(Q.d[0][i])) : make_float4(0, 0, 0, 0);
            T.diffuse[px] = to_h4(sky.x, sky.y, sky.z, sky.w);
            T.spec[px] = to_h4(0.f, 0.f, 0.f, 0.f);
            float4 r = T.radiance[px];
            r.x += sky.x; r.y += sky.y; r.z += sky.z; r.w += 1.f;
            T.radiance[px] = r;
        } else {
            const float4 o4 = Q.o[0][i], d4 = Q.d[0][i];
            const f3 O = ld3(o4), D = ld3(d4);
            const uint32_t prim = __float_as_uint(h.w);
            const uint32_t mi = material_of(s, inst, prim);
            const gk_material& m = s.mats[mi];
            f3 pos = O + D * h.x;
            f3 nrm = primary_normal(s, inst, prim, h.y, h.z);
            float4 alb;
            float rough;
            fetch_gbuffer(s, inst, prim, h.y, h.z, m, nrm, alb, rough);
            uint32_t moment = 0;
            write_gbuffer(s, K, T, px, inst, prim, h, pos, nrm, alb, rough, moment);
            if (m.material_model != GK_MATERIAL_DIELECTRIC) pos = pos - D * 0.03f;
            if (m.material_model == GK_MATERIAL_DIFFUSE_LIGHT) {
                T.diffuse[px] = to_h4(m.diffuse[0], m.diffuse[1], m.diffuse[2], m.diffuse[3]);
                T.spec[px] = to_h4(0.f, 0.f, 0.f, 0.f);
                float4 r = T.radiance[px];
                r.x += m.diffuse[0]; r.y += m.diffuse[1]; r.z += m.diffuse[2]; r.w += 1.f;
                T.radiance[px] = r;
            } else {
                const uint32_t mult = moment > 0u ? 4u : 1u;
                uint32_t nsamp = K.samples * mult;
                if (nsamp < 1u) nsamp = 1u;
                S.ppos[p] = make_float4(pos.x, pos.y, pos.z, __uint_as_float(mi));
                S.pnrm[p] = make_float4(nrm.x, nrm.y, nrm.z, 0.f);
                S.palb[p] = alb;
                S.ld[p] = make_float4(0.f, 0.f, 0.f, 0.f);
                S.ls[p] = make_float4(0.f, 0.f, 0.f, 0.f);
                S.thr[p] = make_float4(1.f, 1.f, 1.f, 1.f);
                float e4[4];
                mul44v(K.mvi, 0.f, 0.f, 0.f, 1.f, e4);
                const f3 dir = normalize(pos - f3{e4[0], e4[1], e4[2]});
                uint32_t flags = 0u;
                const uint32_t counters = (nsamp - 1u) << 20;
                ref_begin(s, K, Q, S, p, pos, nrm, mi, dir, seed, flags, 0u, 1, emit, ro, rd);
                S.misc[p] = make_uint4(px, flags, counters, mi);
                S.pos[p] = make_float4(pos.x, pos.y, pos.z, 0.f);
                S.nrm[p] = make_float4(nrm.x, nrm.y, nrm.z, 0.f);
            }
        }
        S.rng[p] = make_uint4(seed.x, seed.y, seed.z, seed.w);
    }
    const uint32_t slot = block_append(&Q.cnt[2], emit);
    if (emit) { Q.o[1][slot] = ro; Q.d[1][slot] = rd; Q.path[1][slot] = p; }
}

// GetRayColor after the trace (Shading.slang:1039-1081) and the Render loop body
// (Shading.slang:1117-1171); cur = queue the hits belong to; trace_next as k_ref_primary.
__global__ __launch_bounds__(256) void k_ref_shade(DScene s, FrameK K, Targets T, Queues Q, PathState S, int cur,
                                                   uint32_t* __restrict__ trace_next) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0u) *trace_next = 0u;
    const uint32_t n = Q.cnt[2 * cur];
    const int nxt = cur ^ 1;
    const bool live = i < n;
    bool emit = false;
    float4 ro{}, rd{};
    uint32_t p = 0;
    if (live) {
        p = Q.path[cur][i];
        uint4 misc = S.misc[p];
        const uint32_t px = misc.x;
        uint32_t flags = misc.y, counters = misc.z, mi = misc.w;
        const uint4 sr = S.rng[p];
        rng4 seed{sr.x, sr.y, sr.z, sr.w};
        float4 color = S.thr[p];
        const float4 h = Q.hit[i];
        const uint32_t inst = Q.hit_inst[i];
        const uint32_t b = counters & 0xffu;
        const bool chance_reflect = flags & 1u, back = flags & 2u, origin_diel = flags & 4u;
        bool ret;
        f3 pos = ld3(S.pos[p]), nrm = ld3(S.nrm[p]);
        const f3 dir = ld3(S.dir[p]);
        if (__float_as_uint(h.w) != kMiss) {
            const f3 O = ld3(Q.o[cur][i]);
            const uint32_t prim = __float_as_uint(h.w);
            hit_vertex(s, inst, prim, h.x, h.y, h.z, O, dir, pos, nrm);
            mi = material_of(s, inst, prim);
            const gk_material& hm = s.mats[mi];
            const bool hit_light = hm.material_model == GK_MATERIAL_DIFFUSE_LIGHT;
            if (hit_light || !chance_reflect) {
                const float4 ha = mat_albedo(s, hm, hit_uv(s, inst, prim, h.y, h.z));
                color.x = color.x * ha.x; color.y = color.y * ha.y;
                color.z = color.z * ha.z; color.w = color.w * ha.w;
            }
            if (back && !origin_diel) {
                color = make_float4(0.5f, 0.5f, 0.5f, 0.f);
                ret = true;
            } else {
                ret = hit_light;
            }
        } else {
            const float4 sky = K.has_sky ? sky_radiance(s, K.sky, K.sky_rot, dir) : make_float4(0, 0, 0, 0);
            color.x = color.x * sky.x; color.y = color.y * sky.y; color.z = color.z * sky.z; color.w = color.w * sky.w;
            ret = true;
        }
        const float4 pp = S.ppos[p];
        const uint32_t pmat = __float_as_uint(pp.w);
        const bool prim_diel = s.mats[pmat].material_model == GK_MATERIAL_DIELECTRIC;
        const uint32_t maxb = prim_diel ? K.max_bounces : K.bounces;
        bool end_sample = false;
        if (b == 0u) {
            if (ret || !(1u < maxb)) {
                end_sample = true;
            } else if (K.fast_gather && !prim_diel) {
                // ExitAfterFirst (Shading.slang:1118-1124): the probe lookup at the first
                // bounce's vertex ends the sample
                const float4 pr = probe_lookup<true>(K.probes, pos, nrm);
                color.x = color.x * pr.x; color.y = color.y * pr.y;
                color.z = color.z * pr.z; color.w = color.w * pr.w;
                end_sample = true;
            }
        } else {
            if (ret) {
                end_sample = true;
            } else {
                bool done = false;
                if (K.has_sun && rnd1(seed) < 0.5f) {
                    const f3 sun = f3{K.sun_dir[0], K.sun_dir[1], K.sun_dir[2]};
                    const f3 cone = align_with_normal(random_in_cone(seed, 0.99999048072578f), sun);
                    if (!occluded_single(s, pos + nrm * 0.0f, cone, kEps, kMaxTrace)) {
                        color.x = color.x * K.sun_col[0]; color.y = color.y * K.sun_col[1];
                        color.z = color.z * K.sun_col[2]; color.w = color.w * K.sun_col[3];
                        done = true;
                    }
                }
                if (!done) {
                    const bool early = !prim_diel && rnd1(seed) < K.exit_prob;
                    if (b == maxb - 1u || early) {
                        // RayColor *= interpolateAmbientCubesV2<FullAmbientCubeSampler>(vertex)
                        const float4 pr = probe_lookup<true>(K.probes, pos, nrm);
                        color.x = color.x * pr.x; color.y = color.y * pr.y;
                        color.z = color.z * pr.z; color.w = color.w * pr.w;
                        done = true;
                    }
                }
                end_sample = done;
            }
        }
        if (end_sample) {
            ref_end_sample(s, K, T, S, p, px, color, flags, counters, seed, emit, ro, rd);
            if (emit) mi = pmat;
        } else {
            const uint32_t nb = b + 1u;
            counters = (counters & ~0xffu) | nb;
            S.thr[p] = color;
            ref_begin(s, K, Q, S, p, pos, nrm, mi, dir, seed, flags, nb, nxt, emit, ro, rd);
        }
        if (emit && end_sample) {
            // restarted sample: vertex = primary
            pos = ld3(pp);
            nrm = ld3(S.pnrm[p]);
            counters = counters & ~0xffu;
        } else if (!end_sample) {
            // keep RayColor
        }
        if (!end_sample) S.thr[p] = color;
        S.pos[p] = make_float4(pos.x, pos.y, pos.z, 0.f);
        S.nrm[p] = make_float4(nrm.x, nrm.y, nrm.z, 0.f);
        S.misc[p] = make_uint4(px, flags, counters, mi);
        S.rng[p] = make_uint4(seed.x, seed.y, seed.z, seed.w);
    }
    const uint32_t slot = block_append(&Q.cnt[2 * nxt], emit);
    if (emit) { Q.o[nxt][slot] = ro; Q.d[nxt][slot] = rd; Q.path[nxt][slot] = p; }
}


// PT_MIS flags (misc.y): bits 0-7 depth, bit8 lobe0 decided, bit9 lobe0 = diffuse,
// bit10 last vertex was delta (or camera), bits 16-27 sample index
// sh: the shadow queue's counter word, 2 nxt + 1 (3 and 1 alternate between bounces)
__global__ __launch_bounds__(256) void k_pt_shade(DScene s, FrameK K, Targets T, Queues Q, PathState S, int cur,
                                                  int sh, uint32_t* __restrict__ trace_next) {
    const uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 == 0u) *trace_next = 0u;  // the ray counter of the trace launch that follows
    const uint32_t n = Q.cnt[2 * cur];
    const int nxt = cur ^ 1;
    bool emit = false, emit_sh = false;
    float4 ro{}, rd{}, so{}, sd{}, scd{}, scs{}, nthr{};
    uint4 nrng{}, nmisc{};
    uint32_t p = 0;
    if (i0 < n) {
        const uint32_t i = Q.sorted ? Q.sorted[i0] : i0;  // material order
        p = Q.path[cur][i];
        const uint4 misc = Q.qmisc[cur][i];
        const uint32_t px = misc.x;
        uint32_t flags = misc.y;
        const uint32_t depth = flags & 0xffu;
        const uint4 sr = Q.qrng[cur][i];
        rng4 seed{sr.x, sr.y, sr.z, sr.w};
        const float4 thr4 = Q.qthr[cur][i];
        f3 thr = ld3(thr4);
        float last_pdf = thr4.w;
        const bool last_delta = depth == 0u || (flags & 1024u);
        // radiance found at this vertex; added to the path's sums only when non-zero
        f3 dLd = f3{0.f, 0.f, 0.f}, dLs = f3{0.f, 0.f, 0.f};
        bool added = false;
        const float4 h = Q.hit[i];
        const uint32_t inst = Q.hit_inst[i];
        const f3 O = ld3(Q.o[cur][i]), D = ld3(Q.d[cur][i]);
        const bool lobe0_diffuse = (flags & 512u) != 0u;
        auto add = [&](f3 v, bool to_diffuse) {
            if (to_diffuse) dLd = dLd + v;
            else dLs = dLs + v;
            added = true;
        };
        bool alive = false;
        if (depth == 0u && ((flags >> 16) & 0xfffu) == 0u) {
            // primary hit: G-buffer
            if (__float_as_uint(h.w) == kMiss) {
                write_miss_gbuffer(K, T, px, h);
            } else {
                const uint32_t prim = __float_as_uint(h.w);
                const gk_material& m = s.mats[material_of(s, inst, prim)];
                f3 nrm = primary_normal(s, inst, prim, h.y, h.z);
                float4 alb;
                float rough;
                fetch_gbuffer(s, inst, prim, h.y, h.z, m, nrm, alb, rough);
                uint32_t moment = 0;
                write_gbuffer(s, K, T, px, inst, prim, h, O + D * h.x, nrm, alb, rough, moment);
            }
        }
        if (__float_as_uint(h.w) == kMiss) {
            if (K.has_sky) {
                const float4 sky = sky_radiance(s, K.sky, K.sky_rot, D);
                add(thr * f3{sky.x, sky.y, sky.z}, lobe0_diffuse);
            }
        } else {
            const uint32_t prim = __float_as_uint(h.w);
            const InstDev& I = s.inst[inst];
            const uint32_t mi = material_of(s, inst, prim);
            gk_material mat = s.mats[mi];
            f3 mapped_n = f3{0.f, 0.f, 0.f};
            bool mapped = false;
            if (mat.material_model != GK_MATERIAL_DIFFUSE_LIGHT &&
                (has_tex(s, mat.diffuse_texture_id) || has_tex(s, mat.mra_texture_id) ||
                 (depth == 0u && has_tex(s, mat.normal_texture_id)))) {
                // textured albedo and roughness at every vertex, the normal map at the primary hit
                float4 alb;
                float rough;
                if (depth == 0u) {
                    fetch_gbuffer(s, inst, prim, h.y, h.z, mat, mapped_n, alb, rough);
                    mapped = has_tex(s, mat.normal_texture_id);
                } else {
                    const float2 uv = hit_uv(s, inst, prim, h.y, h.z);
                    alb = mat_albedo(s, mat, uv);
                    rough = mat.fuzziness;
                    if (has_tex(s, mat.mra_texture_id)) rough = rough * tex_sample(s, uint32_t(mat.mra_texture_id), uv).y;
                }
                mat.diffuse[0] = alb.x; mat.diffuse[1] = alb.y; mat.diffuse[2] = alb.z; mat.diffuse[3] = alb.w;
                mat.fuzziness = rough;
            }
            const f3 P = O + D * h.x;
            // shading normal: interpolated vertex normals; geometric normal: the object-space
            // face normal e1 x e2 kept in the w lanes of tri_nrm; both by the inverse transpose
            const size_t tb = 3 * size_t(I.tri_base + prim);
            const float4 n0 = s.tri_nrm[tb], n1 = s.tri_nrm[tb + 1], n2 = s.tri_nrm[tb + 2];
            const float bu = h.y, bv = h.z, bw = 1.f - bu - bv;
            f3 Ns = normalize(inv_transpose(I.inv, (ld3(n0) * bw + ld3(n1) * bu) + ld3(n2) * bv));
            const f3 face = f3{n0.w, n1.w, n2.w};
            f3 Ng = normalize(inv_transpose(I.inv, face));
            if (mapped) Ns = mapped_n;
            if (dot(Ng, D) > 0.f) Ng = -Ng;
            if (dot(Ns, Ng) < 0.f) Ns = -Ns;
            if (mat.material_model == GK_MATERIAL_DIFFUSE_LIGHT) {
                const f3 Le = f3{mat.diffuse[0], mat.diffuse[1], mat.diffuse[2]};
                float w = 1.f;
                if (!last_delta && s.nlights > 0u) {
                    const float cosl = fabsf(dot(Ng, D));
                    // selection pdf and world-space area of the hit emitter (its light record)
                    const float psel = light_select_pdf(s, inst, prim);
                    float area = 0.f;
                    {
                        uint32_t lo = 0, hi = s.nlights;
                        while (lo < hi) {
                            const uint32_t mid = (lo + hi) >> 1;
                            const uint32_t li = __float_as_uint(s.lights[mid].e2.w), lp = __float_as_uint(s.lights[mid].le.w);
                            if (li < inst || (li == inst && lp < prim)) lo = mid + 1; else hi = mid;
                        }
                        if (lo < s.nlights) area = s.lights[lo].p0.w;
                    }
                    const float d2 = h.x * h.x * dot(D, D);
                    const float pl = (cosl > 0.f && area > 0.f) ? psel * d2 / (cosl * area) : 0.f;
                    w = power_h(last_pdf, pl);
                }
                if (depth == 0u) add(Le, false);
                else add(thr * Le * w, lobe0_diffuse);
            } else if (depth < K.bounces) {
                const float scale = 1.f + fmaxf(fabsf(P.x), fmaxf(fabsf(P.y), fabsf(P.z)));
                const float off = 1e-4f * scale;
                if (mat.material_model == GK_MATERIAL_DIELECTRIC) {
                    const float ior = mat.refraction_index;
                    const float cosi = -dot(D, Ng);
                    const bool entering = dot(D, inv_transpose(I.inv, face)) < 0.f;
                    const float eta = entering ? 1.f / ior : ior;
                    float r0 = (1.f - ior) / (1.f + ior);
                    r0 = r0 * r0;
                    const float xx = 1.f - cosi;
                    const float F = r0 + (1.f - r0) * xx * xx * xx * xx * xx;
                    const float k = 1.f - eta * eta * (1.f - cosi * cosi);
                    const float u = rnd1(seed);
                    const bool refl = k < 0.f || u < F;
                    f3 wi = refl ? D + Ng * (2.f * cosi) : D * eta + Ng * (eta * cosi - sqrtf(fmaxf(k, 0.f)));
                    wi = normalize(wi);
                    if (depth == 0u) flags = (flags & ~768u) | 256u;
                    thr = thr * f3{mat.diffuse[0], mat.diffuse[1], mat.diffuse[2]};
                    flags |= 1024u;
                    const f3 no = P + (refl ? Ng : -Ng) * off;
                    ro = make_float4(no.x, no.y, no.z, 0.f);
                    rd = make_float4(wi.x, wi.y, wi.z, 1e30f);
                    alive = true;
                } else {
                    const Bsdf b = bsdf_of(mat);
                    f3 T1, B1;
                    onb(Ns, T1, B1);
                    const f3 wo = f3{dot(T1, -D), dot(B1, -D), dot(Ns, -D)};
                    if (wo.z > 0.f) {
                        if (s.nlights > 0u) {
                            const float ul = rnd1(seed);
                            float s1, s2;
                            rnd2(seed, s1, s2);
                            const uint32_t li = light_pick(s, ul);
                            const LightTri L = s.lights[li];
                            const float sq = sqrtf(s1);
                            const float bu2 = 1.f - sq, bv2 = s2 * sq;
                            const f3 Pl = (ld3(L.p0) + ld3(L.e1) * bu2) + ld3(L.e2) * bv2;
                            const f3 Po = P + Ng * off;
                            const f3 dl = Pl - Po;
                            const float d2 = dot(dl, dl);
                            const float dist = sqrtf(d2);
                            const f3 wl = dl * (1.f / dist);
                            const f3 ln = normalize(cross(ld3(L.e1), ld3(L.e2)));
                            const float cosl = fabsf(dot(ln, wl));
                            const f3 wil = f3{dot(T1, wl), dot(B1, wl), dot(Ns, wl)};
                            // psel 0: a light that a material update switched off
                            const float psel = s.light_cdf[li + 1] - s.light_cdf[li];
                            if (cosl > 0.f && wil.z > 0.f && dot(Ng, wl) > 0.f && psel > 0.f) {
                                const float pl = psel * d2 / (cosl * L.p0.w);
                                f3 fd, fs;
                                float pdd, pds;
                                bsdf_eval(b, wo, wil, fd, fs, pdd, pds);
                                const f3 Le = ld3(L.le);
                                const float wd = power_h(pl, (1.f - b.ps) * pdd);
                                const float ws = power_h(pl, b.ps * pds);
                                const f3 cd = thr * Le * fd * (wil.z * wd / pl);
                                const f3 cs = thr * Le * fs * (wil.z * ws / pl);
                                so = make_float4(Po.x, Po.y, Po.z, 0.f);
                                sd = make_float4(wl.x, wl.y, wl.z, dist * (1.f - 1e-4f) - off);
                                if (depth == 0u) {
                                    scd = make_float4(cd.x, cd.y, cd.z, 0.f);
                                    scs = make_float4(cs.x, cs.y, cs.z, 0.f);
                                } else if (lobe0_diffuse) {
                                    scd = make_float4(cd.x + cs.x, cd.y + cs.y, cd.z + cs.z, 0.f);
                                    scs = make_float4(0.f, 0.f, 0.f, 0.f);
                                } else {
                                    scd = make_float4(0.f, 0.f, 0.f, 0.f);
                                    scs = make_float4(cd.x + cs.x, cd.y + cs.y, cd.z + cs.z, 0.f);
                                }
                                emit_sh = true;
                            }
                        }
                        const float ulobe = rnd1(seed);
                        float u1, u2;
                        rnd2(seed, u1, u2);
                        const bool spec = ulobe < b.ps;
                        f3 wi;
                        if (spec) {
                            const f3 mh = vndf_sample(b.alpha, wo, u1, u2);
                            wi = mh * (2.f * dot(wo, mh)) - wo;
                        } else {
                            float sp, cp;
                            sincos_2pi(u1, sp, cp);
                            const float r = sqrtf(u2);
                            wi = f3{r * cp, r * sp, sqrtf(1.f - u2)};
                        }
                        if (wi.z > 0.f) {
                            f3 fd, fs;
                            float pdd, pds;
                            bsdf_eval(b, wo, wi, fd, fs, pdd, pds);
                            const float pdf = spec ? b.ps * pds : (1.f - b.ps) * pdd;
                            // per-component selects of values: a select of the two f3 objects
                            // copies from a selected address, which keeps both in scratch memory
                            const f3 f = f3{spec ? +fs.x : +fd.x, spec ? +fs.y : +fd.y, spec ? +fs.z : +fd.z};
                            if (pdf > 0.f) {
                                thr = thr * f * (wi.z / pdf);
                                if (depth == 0u) flags = (flags & ~768u) | 256u | (spec ? 0u : 512u);
                                last_pdf = pdf;
                                flags &= ~1024u;
                                const f3 wiw = normalize((T1 * wi.x + B1 * wi.y) + Ns * wi.z);
                                const f3 no = P + (dot(wiw, Ng) > 0.f ? Ng : -Ng) * off;
                                ro = make_float4(no.x, no.y, no.z, 0.f);
                                rd = make_float4(wiw.x, wiw.y, wiw.z, 1e30f);
                                alive = true;
                            }
                        }
                    }
                }
                if (alive && depth >= 3u) {
                    const float q = fminf(0.95f, fmaxf(thr.x, fmaxf(thr.y, thr.z)));
                    if (rnd1(seed) >= q) alive = false;
                    else thr = thr * (1.f / q);
                }
            }
        }
        if (added) {
            float4 Ld = S.ld[p], Ls = S.ls[p];
            Ld.x += dLd.x; Ld.y += dLd.y; Ld.z += dLd.z;
            Ls.x += dLs.x; Ls.y += dLs.y; Ls.z += dLs.z;
            S.ld[p] = Ld;
            S.ls[p] = Ls;
        }
        if (alive) {
            flags = (flags & ~0xffu) | (depth + 1u);
            nthr = make_float4(thr.x, thr.y, thr.z, last_pdf);
            emit = true;
        } else {
            // path finished: restart for the next sample of this frame, if any
            const uint32_t sample = ((flags >> 16) & 0xfffu) + 1u;
            if (sample < K.samples) {
                const uint32_t x = px % K.W, y = px / K.W;
                f3 O2, D2;
                camera_ray_lens(K, x, y, seed, O2, D2);
                ro = make_float4(O2.x, O2.y, O2.z, 0.f);
                rd = make_float4(D2.x, D2.y, D2.z, 1e30f);
                flags = (flags & 0x300u) | (sample << 16) | 1024u;  // keep the primary lobe, depth 0 -> 1 below
                flags = (flags & ~0xffu);
                nthr = make_float4(1.f, 1.f, 1.f, 0.f);
                emit = true;
            }
        }
        nmisc = make_uint4(px, flags, misc.z, misc.w);
        nrng = make_uint4(seed.x, seed.y, seed.z, seed.w);
    }
    uint32_t slot, ss;
    block_append_pair(&Q.cnt[2 * nxt], emit, slot, emit_sh, ss);  // sh == 2 * nxt + 1
    if (emit) {
        Q.o[nxt][slot] = ro;
        Q.d[nxt][slot] = rd;
        Q.path[nxt][slot] = p;
        Q.qthr[nxt][slot] = nthr;
        Q.qrng[nxt][slot] = nrng;
        Q.qmisc[nxt][slot] = nmisc;
    }
    if (emit_sh) { Q.so[ss] = so; Q.sd[ss] = sd; Q.sc_d[ss] = scd; Q.sc_s[ss] = scs; Q.spath[ss] = p; }
}

// ---------------------------------------------------------------- material sort
// Counting sort of the hit queue by material class (model x metal, misses last), so
// that a wave shades one BSDF type: per-block LDS histograms, one scan, per-block
// range reservation. The order inside a class is arbitrary; results are not affected
// because every path owns its state.
__device__ inline uint32_t sort_key(const DScene& s, float4 h, uint32_t inst) {
    const uint32_t prim = __float_as_uint(h.w);
    if (prim == kMiss) return 15u;
    const gk_material& m = s.mats[material_of(s, inst, prim)];
    return (m.material_model * 2u + (m.metalness >= 0.5f ? 1u : 0u)) & 15u;
}

__global__ __launch_bounds__(256) void k_sort_count(DScene s, Queues Q, int cur) {
    __shared__ uint32_t hist[16];
    if (threadIdx.x < 16) hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < Q.cnt[2 * cur]) {
        const uint32_t k = sort_key(s, Q.hit[i], Q.hit_inst[i]);
        Q.sort_keys[i] = k;
        atomicAdd(&hist[k], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 16 && hist[threadIdx.x]) atomicAdd(&Q.sort_cnt[threadIdx.x], hist[threadIdx.x]);
}

__global__ void k_sort_scan(uint32_t* cnt) {
    if (threadIdx.x != 0) return;
    uint32_t acc = 0;
    for (int k = 0; k < 16; ++k) {
        cnt[16 + k] = acc;
        acc += cnt[k];
        cnt[k] = 0u;  // ready for the next sort
    }
}

__global__ __launch_bounds__(256) void k_sort_scatter(Queues Q, int cur) {
    __shared__ uint32_t hist[16], base[16];
    if (threadIdx.x < 16) hist[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = i < Q.cnt[2 * cur];
    uint32_t k = 0, rank = 0;
    if (live) {
        k = Q.sort_keys[i];
        rank = atomicAdd(&hist[k], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 16 && hist[threadIdx.x]) base[threadIdx.x] = atomicAdd(&Q.sort_cnt[16 + threadIdx.x], hist[threadIdx.x]);
    __syncthreads();
    if (live) Q.sorted[base[k] + rank] = i;
}

// shadow rays: add the NEE contribution of unoccluded rays
// Also clears the counters the next bounce appends to: the queue this bounce shaded (cur)
// and the other shadow counter; neither is read after this bounce's trace.
__global__ void k_pt_shadow_resolve(Queues Q, PathState S, int cur, int sh) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        Q.cnt[2 * cur] = 0u;
        Q.cnt[sh ^ 2] = 0u;
    }
    if (i >= Q.cnt[sh]) return;
    if (__float_as_uint(Q.shit[i].w) != 0u) return;  // occluded
    const uint32_t p = Q.spath[i];
    const float4 a = Q.sc_d[i], b = Q.sc_s[i];
    float4 ld = S.ld[p], ls = S.ls[p];
    ld.x += a.x; ld.y += a.y; ld.z += a.z;
    ls.x += b.x; ls.y += b.y; ls.z += b.z;
    S.ld[p] = ld;
    S.ls[p] = ls;
}

// PT_MIS targets: diffuse demodulated by the primary albedo, specular, FP32 radiance sum
// paths p_base .. p_base + n - 1
__global__ void k_pt_finalize(FrameK K, Targets T, PathState S, uint32_t n, uint32_t p_base) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t p = p_base + i;
    const uint32_t px = pixel_of(K, p);
    const float inv = 1.f / float(K.samples > 0u ? K.samples : 1u);
    const float4 ld = S.ld[p], ls = S.ls[p];
    const float4 alb = from_h4(T.albedo[px]);
    const float dx = ld.x * inv, dy = ld.y * inv, dz = ld.z * inv;
    const float sx = ls.x * inv, sy = ls.y * inv, sz = ls.z * inv;
    T.diffuse[px] = to_h4(alb.x > 0.f ? dx / alb.x : 0.f, alb.y > 0.f ? dy / alb.y : 0.f, alb.z > 0.f ? dz / alb.z : 0.f, 1.f);
    T.spec[px] = to_h4(sx, sy, sz, 1.f);
    float4 r = T.radiance[px];
    r.x += dx + sx; r.y += dy + sy; r.z += dz + sz; r.w += 1.f;
    T.radiance[px] = r;
    if (T.fp64) {
        double4 q = T.radiance64[px];
        q.x += double(dx) + double(sx); q.y += double(dy) + double(sy); q.z += double(dz) + double(sz); q.w += 1.0;
        T.radiance64[px] = q;
    }
}

__global__ void k_copy_objid(const uint32_t* a, uint32_t* b, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

inline uint32_t nblk(uint32_t n, uint32_t b = 256) { return (n + b - 1) / b; }

}  // namespace

// ---------------------------------------------------------------- frame driver
int ensure_wavefront(Ctx& c) {
    auto& w = c.wf;
    const uint32_t n = c.width * c.height;
    if (w.capacity >= n && w.counters.ptr) return GK_OK;
    for (int k = 0; k < 2; ++k) {
        GK_HIP(w.ray_o[k].resize(n)); GK_HIP(w.ray_d[k].resize(n)); GK_HIP(w.ray_path[k].resize(n));
    }
    GK_HIP(w.hit.resize(n)); GK_HIP(w.hit_inst.resize(n));
    GK_HIP(w.shadow_o.resize(n)); GK_HIP(w.shadow_d.resize(n));
    GK_HIP(w.shadow_c.resize(2 * size_t(n))); GK_HIP(w.shadow_path.resize(n)); GK_HIP(w.shadow_hit.resize(n));
    GK_HIP(w.counters.resize(32 * kPathParts));  // 32 words per path pipeline
    GK_HIP(w.sort_keys.resize(n));
    GK_HIP(w.sorted_idx.resize(n));
    GK_HIP(w.sort_cnt.resize(32 * 4));  // 16 class counts and 16 cursors per path pipeline
    GK_HIP(hipMemsetAsync(w.sort_cnt.ptr, 0, 32 * 4 * 4, c.stream));
    for (int k = 0; k < 2; ++k) {
        GK_HIP(w.q_thr[k].resize(n)); GK_HIP(w.q_rng[k].resize(n)); GK_HIP(w.q_misc[k].resize(n));
    }
    GK_HIP(w.s_thr.resize(n)); GK_HIP(w.s_ld.resize(n)); GK_HIP(w.s_ls.resize(n));
    GK_HIP(w.s_rng.resize(n)); GK_HIP(w.s_misc.resize(n));
    GK_HIP(w.s_pos.resize(n)); GK_HIP(w.s_nrm.resize(n)); GK_HIP(w.s_dir.resize(n));
    GK_HIP(w.s_ppos.resize(n)); GK_HIP(w.s_pnrm.resize(n)); GK_HIP(w.s_palb.resize(n));
    w.capacity = n;
    return GK_OK;
}

// whether a frame runs as kPathParts path pipelines (render_frame)
static bool frame_splits(const Ctx& c) {
    return c.width * c.height >= kSplitPaths && !(c.params.flags & GK_FRAME_SINGLE_PIPELINE);
}

static FrameK make_k(const Ctx& c) {
    const gk_frame_params& P = c.params;
    FrameK K;
    memcpy(K.mvi, P.model_view_inverse, 64);
    memcpy(K.pi, P.projection_inverse, 64);
    memcpy(K.vp, P.view_projection, 64);
    memcpy(K.vpu, P.view_projection_unjit, 64);
    memcpy(K.pvpu, P.prev_view_projection_unjit, 64);
    memcpy(K.sun_dir, P.sun_direction, 16);
    memcpy(K.sun_col, P.sun_color, 16);
    memcpy(K.sky, P.sky_color, 16);
    K.sky_rot = P.sky_rotation;
    K.exit_prob = P.exit_probability;
    K.hit_off = P.hit_normal_offset;
    K.aperture = P.aperture;
    K.focus = P.focus_distance;
    K.W = c.width;
    K.H = c.height;
    K.frame = P.total_frames;
    K.seed = P.global_seed;
    K.samples = P.samples > 0 ? P.samples : 1;
    K.bounces = P.bounces;
    K.max_bounces = P.max_bounces;
    K.temporal = P.temporal_frames;
    K.has_sky = P.has_sky;
    K.has_sun = P.has_sun;
    K.tiled = (c.width % 8 == 0 && c.height % 8 == 0) ? 1u : 0u;
    {
        // one part (plain row-major tile order, the best cache locality) unless the frame
        // will be split into path pipelines
        const uint32_t tiles = (c.width / 8u) * (c.height / 8u);
        K.tparts = frame_splits(c) ? uint32_t(kPathParts) : 1u;
        uint32_t at = 0;
        for (uint32_t h = 0; h < 4; ++h) {
            K.tstart[h] = at;
            if (h < K.tparts) at += (tiles + K.tparts - 1u - h) / K.tparts;  // tiles = h mod tparts
        }
    }
    K.vis_primary = (P.flags & GK_FRAME_PRIMARY_FROM_VISBUFFER) ? 1u : 0u;
    K.fast_gather = (P.flags & GK_FRAME_FAST_GATHER) ? 1u : 0u;
    K.probes = probe_dev(c);
    return K;
}

int render_frame(Ctx& c) {
    c.trace_split = false;
    const bool early_ok = c.early_ok && c.early_w == c.width && c.early_h == c.height;
    c.early_ok = false;
    int st = ensure_wavefront(c);
    if (st != GK_OK) return st;
    auto& w = c.wf;
    hipStream_t s = c.stream;
    const uint32_t n = c.width * c.height;
    if (c.tev_used + 256u > c.tev.size() && c.tev_used > 0) {
        // keep room for this frame's trace events: fold the recorded ones into the totals
        GK_HIP(hipStreamSynchronize(s));
        st = collect_timing(c);
        if (st != GK_OK) return st;
    }
    hipEventRecord(c.ev[0], s);
    if (c.params.flags & GK_FRAME_POST_ONLY) {  // the caller wrote the rt pass's targets
        hipEventRecord(c.ev[1], s);
        hipEventRecord(c.ev[2], s);
        st = post_process(c);
        if (st != GK_OK) return st;
        hipEventRecord(c.ev[3], s);
        return GK_OK;
    }
    const bool tlas_work = c.tlas_dirty;
    if (c.tlas_dirty) {
        st = (c.params.flags & GK_FRAME_TLAS_REBUILD) ? rebuild_tlas(c) : refit_tlas(c);
        if (st != GK_OK) return st;
    }
    hipEventRecord(c.ev[1], s);
    const FrameK K = make_k(c);
    const DScene S0 = make_dscene(c);
    Targets T;
    T.diffuse = static_cast<ushort4*>(c.target_ptr[GK_RT_SINGLE_DIFFUSE]);
    T.spec = static_cast<ushort4*>(c.target_ptr[GK_RT_SINGLE_SPECULAR]);
    T.albedo = static_cast<ushort4*>(c.target_ptr[GK_RT_ALBEDO]);
    T.normal = static_cast<ushort4*>(c.target_ptr[GK_RT_NORMAL]);
    T.objid = static_cast<uint32_t*>(c.target_ptr[GK_RT_OBJECTID_0]);
    T.motion = static_cast<float2*>(c.target_ptr[GK_RT_MOTIONVECTOR]);
    T.moment = static_cast<uint16_t*>(c.target_ptr[GK_RT_MOTIONMOMENT]);
    T.depth = static_cast<float*>(c.target_ptr[GK_RT_PREV_DEPTH]);
    T.minigbuffer = static_cast<uint32_t*>(c.target_ptr[GK_RT_MINIGBUFFER]);
    T.phit = static_cast<float4*>(c.target_ptr[GK_RT_PRIMARY_HIT]);
    T.pinst = static_cast<uint32_t*>(c.target_ptr[GK_RT_PRIMARY_INST]);
    T.radiance = static_cast<float4*>(c.target_ptr[GK_RT_RADIANCE]);
    T.radiance64 = static_cast<double4*>(c.target_ptr[GK_RT_RADIANCE64]);
    T.fp64 = (c.params.flags & GK_FRAME_ACCUM_FP64) ? 1u : 0u;
    Queues Q;
    for (int k = 0; k < 2; ++k) {
        Q.o[k] = w.ray_o[k].ptr; Q.d[k] = w.ray_d[k].ptr; Q.path[k] = w.ray_path[k].ptr;
        Q.qthr[k] = w.q_thr[k].ptr; Q.qrng[k] = w.q_rng[k].ptr; Q.qmisc[k] = w.q_misc[k].ptr;
    }
    Q.hit = w.hit.ptr;
    Q.hit_inst = w.hit_inst.ptr;
    Q.so = w.shadow_o.ptr;
    Q.sd = w.shadow_d.ptr;
    Q.sc_d = w.shadow_c.ptr;
    Q.sc_s = w.shadow_c.ptr + n;
    Q.spath = w.shadow_path.ptr;
    Q.shit = w.shadow_hit.ptr;
    Q.cnt = w.counters.ptr;
    Q.sort_keys = w.sort_keys.ptr;
    Q.sort_cnt = w.sort_cnt.ptr;
    // Material-ordered shading is opt-in: PT_MIS shades every non-dielectric model with one
    // BSDF code path, so the sort (two passes over the queue, 0.57 ms per 1080p city frame)
    // costs more than the divergence it removes, also with four material models in the
    // scene (1080p city: 5.50 ms per frame sorted, 4.92 ms unsorted)
    const bool sort = (c.params.flags & GK_FRAME_MATERIAL_SORT) && !(c.params.flags & GK_FRAME_NO_MATERIAL_SORT);
    Q.sorted = sort ? w.sorted_idx.ptr : nullptr;
    PathState P;
    P.thr = w.s_thr.ptr; P.ld = w.s_ld.ptr; P.ls = w.s_ls.ptr; P.rng = w.s_rng.ptr; P.misc = w.s_misc.ptr;
    P.pos = w.s_pos.ptr; P.nrm = w.s_nrm.ptr; P.dir = w.s_dir.ptr; P.ppos = w.s_ppos.ptr; P.pnrm = w.s_pnrm.ptr;
    P.palb = w.s_palb.ptr;

    // previous object ids for reprojection (copy pass, PathTracingRenderer.cpp:160-216)
    k_copy_objid<<<nblk(n), 256, 0, s>>>(T.objid, static_cast<uint32_t*>(c.target_ptr[GK_RT_OBJECTID_1]), n);
    GK_HIP(hipGetLastError());
    const bool ref = c.params.estimator == GK_ESTIMATOR_REF_V2;
    // Frames of many paths run as kPathParts path pipelines (part h: the 8x8 tiles h mod
    // kPathParts), each with its own queue range, counters and trace slot, on its own stream:
    // a persistent trace launch spends a long tail finishing its last rays with most lanes
    // idle, and one pipeline's launches overlap the other's. Each trace grid then holds 3
    // work-groups per CU (kSplitBlocksPerCU), so both fit at once (1080p city, PT_MIS: 3.33
    // -> 2.97 ms per frame; 4096 bricks: 4.36 -> 4.10 ms).
    const bool split = frame_splits(c);
    const int np = split ? kPathParts : 1;
    uint32_t pbase[4] = {0u, 0u, 0u, 0u}, pcount[4] = {n, 0u, 0u, 0u};
    if (split) {
        for (int h = 0; h < np; ++h) {
            pbase[h] = K.tiled ? std::min(n, K.tstart[h] * 64u) : uint32_t(uint64_t(n) * uint32_t(h) / uint32_t(np));
            const uint32_t end = h + 1 < np ? (K.tiled ? std::min(n, K.tstart[h + 1] * 64u)
                                                      : uint32_t(uint64_t(n) * uint32_t(h + 1) / uint32_t(np)))
                                            : n;
            pcount[h] = end - pbase[h];
        }
    }
    Queues QH[4] = {Q, Q, Q, Q};
    // pipeline h of a split frame runs on pstream[h] and traces in slot h + 1; an unsplit
    // frame runs on `s` and traces in slot 0
    hipStream_t ss[4] = {s, s, s, s};
    int sl[4] = {0, 0, 0, 0};
    uint32_t* tn[4] = {nullptr, nullptr, nullptr, nullptr};
    static_assert(kPathParts < 4, "a split frame's pipelines use trace slots 1..kPathParts");
    for (int h = 0; split && h < np; ++h) {
        if (!c.fork_ev) GK_HIP(hipEventCreateWithFlags(&c.fork_ev, hipEventDisableTiming));
        if (!c.pstream[h]) {
            GK_HIP(hipStreamCreateWithFlags(&c.pstream[h], hipStreamNonBlocking));
            GK_HIP(hipEventCreateWithFlags(&c.join_ev[h], hipEventDisableTiming));
        }
        ss[h] = c.pstream[h];
        sl[h] = h + 1;
    }
    for (int h = 1; h < np; ++h) {
        // pipeline h: queue entries from its first path id on, counters from word 32 h
        const uint32_t o = pbase[h];
        Queues& B = QH[h];
        for (int k = 0; k < 2; ++k) {
            B.o[k] += o; B.d[k] += o; B.path[k] += o;
            B.qthr[k] += o; B.qrng[k] += o; B.qmisc[k] += o;
        }
        B.hit += o; B.hit_inst += o;
        B.so += o; B.sd += o; B.sc_d += o; B.sc_s += o; B.spath += o; B.shit += o;
        B.cnt += 32 * h;  // a 128-byte line of its own: the pipelines' appends do not share one
        B.sort_keys += o;
        if (B.sorted) B.sorted += o;
        B.sort_cnt += 32 * h;
    }
    // the slots' ray counters (a slot's first use allocates and clears them on `s`, so
    // before the fork)
    for (int h = 0; h < np; ++h) {
        tn[h] = trace_counter(c, sl[h]);
        if (!tn[h]) return GK_ERR_OOM;
    }
    // Early start: ray generation and the primary trace read the queues, path state and
    // slots of their own pipeline (each stream's previous frame is behind them in stream
    // order) and the scene's BVH, and write no target. When no work that they depend on was
    // queued on `s` since the last split frame of this size (no TLAS update, no other entry
    // point), they start without waiting for `s`, so the last frame's post passes run beside
    // them; the pipelines wait for `s` (this frame's object-id copy after those passes)
    // before their first kernel that writes a target. Under stream capture every stream
    // forks from the captured one.
    hipStreamCaptureStatus cap0 = hipStreamCaptureStatusNone;
    GK_HIP(hipStreamIsCapturing(s, &cap0));
    const bool early = split && early_ok && !tlas_work && !K.vis_primary && cap0 == hipStreamCaptureStatusNone;
    if (split) {
        GK_HIP(hipEventRecord(c.fork_ev, s));
        if (!early)
            for (int h = 0; h < np; ++h) GK_HIP(hipStreamWaitEvent(ss[h], c.fork_ev, 0));
    }
    c.trace_split = split;
    for (int h = 0; h < np; ++h) {
        k_raygen<<<nblk(pcount[h]), 256, 0, ss[h]>>>(K, QH[h], P, pcount[h], ref ? 0u : 1u, pbase[h], tn[h]);
        GK_HIP(hipGetLastError());
        // primary rays: traced, or their hits taken from the caller's visibility buffer
        if (K.vis_primary) {
            k_vis_primary<<<nblk(pcount[h]), 256, 0, ss[h]>>>(S0, K, QH[h], T.minigbuffer, c.tri_rec.ptr, c.meshes.ptr,
                                                              c.proxy_inst.ptr, c.num_proxies, pcount[h], pbase[h]);
            GK_HIP(hipGetLastError());
        } else {
            c.trace_counter_zeroed[sl[h]] = true;
            st = launch_trace(c, QH[h].o[0], QH[h].d[0], QH[h].cnt + 0, pcount[h], QH[h].hit, QH[h].hit_inst, false,
                              nullptr, sl[h]);
            if (st != GK_OK) return st;
        }
    }
    if (early)
        for (int h = 0; h < np; ++h) GK_HIP(hipStreamWaitEvent(ss[h], c.fork_ev, 0));
    int cur = 0;
    if (ref) {
        for (int h = 0; h < np; ++h) {
            k_ref_primary<<<nblk(pcount[h]), 256, 0, ss[h]>>>(S0, K, T, QH[h], P, tn[h]);
            GK_HIP(hipGetLastError());
        }
        cur = 1;
        // worst case: every sample uses all bounces (samples * 4 for moving pixels); a sample
        // traces at least once, also with no bounces
        const uint32_t maxb = std::max(1u, std::max(K.bounces, K.max_bounces));
        const uint32_t iters = maxb * K.samples * 4u + 4u;
        // A pipeline's loop ends early once the queue it is about to trace is empty: every
        // sample has finished. Reading the counters back costs a synchronisation of the
        // streams, so they are read after the iterations a pixel without motion can take,
        // then every maxb iterations; a stream under capture runs the worst case.
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        GK_HIP(hipStreamIsCapturing(s, &cap));
        if (cap == hipStreamCaptureStatusNone && !c.h_word) GK_HIP(hipHostMalloc(&c.h_word, 4 * sizeof(uint32_t)));
        uint32_t check = cap == hipStreamCaptureStatusNone ? maxb * std::max(1u, K.samples) + 1u : iters;
        bool live[4] = {true, true, true, true};
        for (uint32_t it = 0; it < iters; ++it) {
            if (it == check) {
                for (int h = 0; h < np; ++h)
                    if (live[h])
                        GK_HIP(hipMemcpyAsync(c.h_word + h, QH[h].cnt + 2 * cur, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                              ss[h]));
                bool any = false;
                for (int h = 0; h < np; ++h) {
                    if (!live[h]) continue;
                    GK_HIP(hipStreamSynchronize(ss[h]));
                    live[h] = c.h_word[h] != 0u;
                    any = any || live[h];
                }
                if (!any) break;
                check += maxb;
            }
            for (int h = 0; h < np; ++h) {
                if (!live[h]) continue;
                c.trace_counter_zeroed[sl[h]] = true;
                // the trace zeroes the counter of the queue k_ref_shade appends to
                const Queues& q = QH[h];
                st = launch_trace(c, q.o[cur], q.d[cur], q.cnt + 2 * cur, pcount[h], q.hit, q.hit_inst, false,
                                  q.cnt + 2 * (cur ^ 1), sl[h]);
                if (st != GK_OK) return st;
            }
            for (int h = 0; h < np; ++h) {
                if (!live[h]) continue;
                k_ref_shade<<<nblk(pcount[h]), 256, 0, ss[h]>>>(S0, K, T, QH[h], P, cur, tn[h]);
                GK_HIP(hipGetLastError());
            }
            cur ^= 1;
        }
        c.trace_split = false;
        for (int h = 0; split && h < np; ++h) {
            GK_HIP(hipEventRecord(c.join_ev[h], ss[h]));
            GK_HIP(hipStreamWaitEvent(s, c.join_ev[h], 0));
        }
    } else {
        // Each iteration shades the hits of queue cur, then one launch traces both the
        // extension rays it produced (closest hit, for the next iteration) and its shadow
        // rays (occlusion), whose contributions are resolved before the next shade. The two
        // pipelines' launches are issued alternately.
        const uint32_t iters = (K.bounces + 1u) * K.samples;
        for (uint32_t it = 0; it < iters; ++it) {
            // queue counters: 2 cur (size of this bounce's hits), 2 nxt (next queue, zero),
            // sh = 2 nxt + 1 (this bounce's shadow rays, zero; next to the next queue's size,
            // so k_pt_shade appends to both with one atomic); k_raygen and the previous
            // resolve cleared them
            const int nxt = cur ^ 1;
            const int sh = 2 * nxt + 1;
            for (int h = 0; h < np; ++h) {
                const Queues& q = QH[h];
                const uint32_t m = pcount[h];
                if (sort) {
                    k_sort_count<<<nblk(m), 256, 0, ss[h]>>>(S0, q, cur);
                    k_sort_scan<<<1, 64, 0, ss[h]>>>(q.sort_cnt);
                    k_sort_scatter<<<nblk(m), 256, 0, ss[h]>>>(q, cur);
                }
                k_pt_shade<<<nblk(m), 256, 0, ss[h]>>>(S0, K, T, q, P, cur, sh, tn[h]);
                GK_HIP(hipGetLastError());
                c.trace_counter_zeroed[sl[h]] = true;
            }
            for (int h = 0; h < np; ++h) {
                const Queues& q = QH[h];
                const uint32_t m = pcount[h];
                if (it + 1 < iters)
                    st = launch_trace_pair(c, q.o[nxt], q.d[nxt], q.cnt + 2 * nxt, m, q.hit, q.hit_inst, q.so, q.sd,
                                           q.cnt + sh, m, q.shit, sl[h]);
                else
                    st = launch_trace(c, q.so, q.sd, q.cnt + sh, m, q.shit, nullptr, true, nullptr, sl[h]);
                if (st != GK_OK) return st;
            }
            for (int h = 0; h < np; ++h) {
                k_pt_shadow_resolve<<<nblk(pcount[h]), 256, 0, ss[h]>>>(QH[h], P, cur, sh);
                GK_HIP(hipGetLastError());
            }
            cur = nxt;
        }
        c.trace_split = false;
        // each pipeline writes its own pixels' targets before joining
        for (int h = 0; h < np; ++h) {
            k_pt_finalize<<<nblk(pcount[h]), 256, 0, ss[h]>>>(K, T, P, pcount[h], pbase[h]);
            GK_HIP(hipGetLastError());
        }
        for (int h = 0; split && h < np; ++h) {
            GK_HIP(hipEventRecord(c.join_ev[h], ss[h]));
            GK_HIP(hipStreamWaitEvent(s, c.join_ev[h], 0));
        }
    }
    hipEventRecord(c.ev[2], s);
    if (!(c.params.flags & GK_FRAME_SKIP_POST)) {
        st = post_process(c);
        if (st != GK_OK) return st;
    }
    hipEventRecord(c.ev[3], s);
    c.early_ok = split && cap0 == hipStreamCaptureStatusNone;
    c.early_w = c.width;
    c.early_h = c.height;
    return GK_OK;
}

}  // namespace gk

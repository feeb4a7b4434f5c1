// gk_ctx.hpp - host-side context of the path-tracing core (one per gk_create).
//
// The context owns every device allocation and one HIP stream; the layout of the
// device data is described next to each member.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../include/gknx_rt.h"
#include "gk_math.hpp"

namespace gk {

// ------------------------------------------------------------------ device memory budget
// Device bytes held by the library's buffers (all contexts of the process), and an optional
// budget: GKNX_DEVICE_MEMORY_LIMIT=<bytes> in the environment when the library first
// allocates. An allocation that would exceed it fails as hipMalloc does when the device is
// full (hipErrorOutOfMemory -> GK_ERR_OOM), so callers can bound the library's footprint and
// exercise their out-of-memory handling.
inline std::atomic<uint64_t>& device_bytes() {
    static std::atomic<uint64_t> n{0};
    return n;
}
inline uint64_t device_budget() {
    static const uint64_t limit = [] {
        const char* v = std::getenv("GKNX_DEVICE_MEMORY_LIMIT");
        return v && *v ? uint64_t(std::strtoull(v, nullptr, 10)) : UINT64_MAX;
    }();
    return limit;
}

// ------------------------------------------------------------------ device buffer
template <class T>
struct DevBuf {
    T* ptr = nullptr;
    size_t count = 0;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (ptr) {
            (void)hipFree(ptr);
            device_bytes() -= uint64_t(cap) * sizeof(T);
        }
        ptr = nullptr;
        count = cap = 0;
    }
    // (re)allocate for n elements; keeps the allocation when it is large enough
    hipError_t resize(size_t n) {
        if (n <= cap) { count = n; return hipSuccess; }
        release();
        if (n == 0) return hipSuccess;
        const uint64_t bytes = uint64_t(n) * sizeof(T);
        if (device_bytes() + bytes > device_budget()) return hipErrorOutOfMemory;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&ptr), n * sizeof(T));
        if (e != hipSuccess) { ptr = nullptr; return e; }
        device_bytes() += bytes;
        count = cap = n;
        return hipSuccess;
    }
    size_t bytes() const { return count * sizeof(T); }
};

// ------------------------------------------------------------------ scene records
struct MeshInfo {        // per mesh
    uint32_t tri_base;   // first triangle of the mesh in tri attribute arrays (original order)
    uint32_t tri_count;
    uint32_t root;       // root node (absolute index into blas_nodes)
    uint32_t depth;      // wide-tree depth
    float bmin[3];
    float bmax[3];
};

struct InstDev {         // per instance, 256 bytes
    float world[16];     // column-major
    float inv[16];       // row-major world->object (tinybvh InvertTransform layout)
    float combined[16];  // column-major prev * inverse(world), identity when static
    uint32_t mesh;
    uint32_t instance_id;
    uint32_t active;     // visible && !nort
    uint32_t use_remap;
    uint32_t mat_offset;
    uint32_t tri_base;   // = meshes[mesh].tri_base
    uint32_t root;       // = meshes[mesh].root
    uint32_t proxy_base; // visibility-buffer proxy of the mesh's section 0 (kSectionTris each)
    uint32_t mat_id[16];
};
static_assert(sizeof(InstDev) == 288, "InstDev size");

// Visibility-buffer numbering (Scene::RebuildMeshBuffer slices a model into sections of at
// most 65535 triangles, Scene.cpp:202-256, and UpdateNodesGpuDriven emits one NodeProxy per
// node and section, Scene.cpp:472-478): proxy = instance's proxy_base + prim / kSectionTris,
// and the packed id is (proxy + 1) << 17 | prim % kSectionTris.
constexpr uint32_t kSectionTris = 65535u;
constexpr uint32_t kMaxVisProxies = 0x7fffu;  // proxies whose proxy + 1 fits the 15-bit field

struct TexInfo {         // one texture of the pool
    uint64_t offset;     // first texel in Ctx::texels
    uint32_t w, h;
    uint32_t srgb, pad;
};

struct LightTri {        // emissive triangle in world space, 64 bytes
    float4 p0;           // w = area
    float4 e1;           // w = luminance-weighted selection pdf
    float4 e2;           // w = inst (bits)
    float4 le;           // w = prim (bits)
};

// ------------------------------------------------------------------ build scratch
struct BuildScratch {
    DevBuf<uint64_t> keys, keys_alt;
    DevBuf<uint32_t> vals, vals_alt;
    DevBuf<uint8_t> sort_tmp;
    DevBuf<float4> prim_lo, prim_hi;      // primitive boxes (w unused)
    DevBuf<uint32_t> bparent;             // binary node parents
    DevBuf<uint2> bchild;                 // binary internal children (leaf: bit 31)
    DevBuf<uint2> brange;                 // [first, last] of each internal node
    DevBuf<float4> bnode_lo, bnode_hi;    // internal node boxes (w = cost / prim count)
    DevBuf<uint32_t> bflag;               // bottom-up arrival counters
    DevBuf<uint32_t> queue_a, queue_b;    // top-down collapse work queues (pairs)
    DevBuf<uint32_t> counters;            // [0] wide node count [1] queue size [2] next queue size
    DevBuf<float> bounds;                 // scene centroid bounds (6 floats as ordered ints)
    DevBuf<uint32_t> node_depth;
};

// ------------------------------------------------------------------ context
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // streams of trace slots 1..3: a frame split into independent path pipelines runs
    // pipeline h on pstream[h] and traces in slot h + 1 (created on first use; default
    // priority); the events that fork them from and join them to `stream`
    hipStream_t pstream[3] = {nullptr, nullptr, nullptr};
    hipEvent_t fork_ev = nullptr, join_ev[3] = {nullptr, nullptr, nullptr};
    // The next split frame's pipelines may generate and trace their primary rays without
    // waiting for the work queued on `stream` before the frame (the last frame's post
    // passes): set by a split frame of early_w x early_h pixels; cleared by every entry point
    // but frame set-up, render, sync and read-only queries (GK_CHECK_CTX)
    bool early_ok = false;
    uint32_t early_w = 0, early_h = 0;
    uint32_t width = 0, height = 0;
    bool scene_ready = false;
    bool frame_ready = false;
    bool tlas_dirty = false;
    bool blas_only = false;   // one identity instance: trace the BLAS directly

    // scene (host copies)
    std::vector<gk_material> h_materials;
    std::vector<MeshInfo> h_meshes;
    std::vector<gk_instance> h_instances;
    uint32_t num_tris = 0;

    // scene (device)
    DevBuf<gk_material> materials;
    DevBuf<TriRec> tris;            // all meshes, leaf order (absolute record index)
    DevBuf<float4> tri_nrm;         // 3 per triangle, original order: n0, n1, n2 (object space)
    DevBuf<uint32_t> tri_mat;       // local material slot per triangle, original order
    DevBuf<uint32_t> tri_rec;       // original order -> leaf-order record index (within the mesh)
    DevBuf<float2> tri_uv;          // 3 per triangle, original order (empty: no mesh has texcoords)
    DevBuf<float4> tri_tan;         // 3 per triangle, original order (empty: no mesh has tangents)
    // textures (gk_textures_upload)
    DevBuf<uint32_t> texels;        // RGBA8, all textures back to back
    DevBuf<TexInfo> tex;
    DevBuf<float> tex_lut;          // [0,256): sRGB byte -> linear, [256,512): byte / 255
    uint32_t num_tex = 0;
    // sky image (gk_sky_upload): equirectangular RGBA32F, row 0 at +y; none while sky_w == 0
    DevBuf<float4> sky;
    uint32_t sky_w = 0, sky_h = 0;
    float sky_sh[27] = {};          // ProjectHdrToSh of the image: [channel r, g, b][9]
    uint32_t max_mesh_tris = 0;     // largest mesh
    uint32_t num_proxies = 0;       // visibility-buffer proxies (instance x section)
    DevBuf<uint32_t> proxy_inst;    // proxy -> instance
    DevBuf<Bvh4Node> blas_nodes;
    DevBuf<MeshInfo> meshes;
    DevBuf<InstDev> inst;
    DevBuf<float> inst_world_in;    // staging for instance updates (16 floats each)
    DevBuf<float4> inst_lo, inst_hi;  // world-space instance boxes
    DevBuf<Bvh4Node> tlas_nodes;
    DevBuf<uint32_t> tlas_prims;    // TLAS leaf -> instance index
    DevBuf<uint32_t> tlas_refit_order;  // nodes, deepest level first
    DevBuf<uint32_t> tlas_level_info;   // [levels, start of each level in tlas_refit_order ...]
    std::vector<uint32_t> tlas_level_offsets;
    uint32_t tlas_node_count = 0;
    uint32_t tlas_depth = 0;
    DevBuf<LightTri> lights;
    uint32_t num_lights = 0;
    DevBuf<float> light_cdf;        // num_lights + 1
    DevBuf<float4> light_obj;       // 3 per light: object-space v0, e1, e2
    std::vector<uint32_t> h_light_mat;  // material of each light (refreshed by gk_materials_update)
    std::vector<float> h_light_area;    // world-space area at upload (rigid motion keeps it)
    DevBuf<float4> light_power;     // staging for gk_materials_update: per light (le rgb, pdf)
    BuildScratch scratch;
    gk_bvh_stats bvh_stats{};
    DevBuf<float> grad_partial;     // adjoint: per-work-group gradient rows
    ::gk_unet* unet = nullptr;      // learned denoiser (GK_DENOISER_UNET), owned by the context
    std::vector<gk_light> h_lights; // LightObject quads (the probe bake lights with the first)

    // ambient-cube probe grid (probe.hip)
    struct Probes {
        gk_probe_grid grid{};
        uint32_t count = 0;             // probes; 0 = no grid
        DevBuf<gk_ambient_cube> cubes;
        DevBuf<gk_ambient_cube> next;   // bake results, applied when the pass ends
        DevBuf<gk_voxel> voxels;
        struct Work {                   // one chunk in flight; consecutive bake chunks alternate
                                        // between two (trace slots 0 and 1, two streams)
            DevBuf<uint32_t> list;      // probes of the chunk being baked
            DevBuf<uint32_t> counters;  // [0] probes in list, [1] closest rays, [2] occlusion rays
            DevBuf<float4> ro, rd, hit;
            DevBuf<uint32_t> hit_inst;
            DevBuf<float4> so, sd, shit;
            DevBuf<uint32_t> face_sky;  // per (probe, face): sky byte, sun flag and bit
        } w[2];
    } probes;

    // frame
    gk_frame_params params{};
    gk_frame_params prev_params{};
    uint32_t frames_rendered = 0;

    // render targets (indexed by gk_target)
    void* target_ptr[GK_RT_COUNT] = {};
    size_t target_bytes[GK_RT_COUNT] = {};
    std::vector<void*> target_allocs;

    // wavefront buffers
    struct Wavefront {
        uint32_t capacity = 0;
        DevBuf<float4> ray_o[2], ray_d[2];    // queue rays (o,tmin) (d,tmax)
        DevBuf<uint32_t> ray_path[2];         // queue entry -> path index
        DevBuf<float4> hit;                   // per queue entry: t,u,v,prim
        DevBuf<uint32_t> hit_inst;
        DevBuf<uint32_t> sorted_idx;          // material-sorted permutation of queue entries
        DevBuf<uint32_t> sort_keys;           // material class per queue entry
        DevBuf<uint32_t> sort_cnt;            // 16 class counts + 16 cursors
        DevBuf<float4> shadow_o, shadow_d;    // shadow queue
        DevBuf<float4> shadow_c;              // contribution (rgb) + target (w)
        DevBuf<float4> shadow_hit;            // occlusion results (w = 1 when occluded)
        DevBuf<uint32_t> shadow_path;
        DevBuf<uint32_t> counters;            // [0..1] queue sizes [2] shadow size [3..] sort
        // path state (SoA)
        DevBuf<float4> q_thr[2];              // PT_MIS state travelling with queue entries:
        DevBuf<uint4> q_rng[2], q_misc[2];    //   throughput + pdf, rng, (pixel, flags)
        DevBuf<float4> s_thr;                 // throughput rgb, last bsdf pdf
        DevBuf<float4> s_ld;                  // primary-diffuse-lobe radiance, w = flags
        DevBuf<float4> s_ls;                  // other radiance
        DevBuf<uint4> s_rng;
        DevBuf<uint4> s_misc;                 // pixel, depth/flags, sample counters, lobe
        // REF_V2 extra state
        DevBuf<float4> s_pos, s_nrm;          // current vertex (w: material bits / misc)
        DevBuf<float4> s_dir;                 // current ray direction
        DevBuf<float4> s_ppos, s_pnrm, s_palb;  // primary vertex, its (textured) albedo
        DevBuf<ushort4> hist_d, hist_s, hist_a;  // previous accumulation targets (RGBA16F)
        DevBuf<uint32_t> spill;               // traversal stack overflow, then the launch counters
        DevBuf<uint32_t> pspill[3];           // the same for trace slots 1..3
        uint32_t spill_threads = 0;
    } wf;

    // timing: per-launch trace events (collected after a stream synchronisation)
    hipEvent_t ev[16] = {};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> tev;
    uint32_t tev_used = 0;
    uint64_t trace_launches = 0;
    bool stack_overflow = false;        // a traversal stack overflowed since gk_get_stats last reported
    bool trace_counter_zeroed[4] = {};  // per slot: the next launch's counter is already 0
    bool trace_split = false;           // a frame's two path pipelines are being enqueued
    uint32_t* h_word = nullptr;         // pinned host words (one per path pipeline) for counter read-backs
    gk_stats stats{};
};

// ------------------------------------------------------------------ helpers
#define GK_HIP(call)                                            \
    do {                                                        \
        hipError_t e__ = (call);                                \
        if (e__ != hipSuccess) {                                \
            gk::set_error(#call, e__);                          \
            return e__ == hipErrorOutOfMemory ? GK_ERR_OOM : GK_ERR_HIP; \
        }                                                       \
    } while (0)

void set_error(const char* what, hipError_t e);

// bvh_build.hip
int build_blas(Ctx& c, const gk_mesh* meshes, uint32_t num_meshes);
int build_tlas(Ctx& c);
int update_instances_device(Ctx& c, uint32_t first, uint32_t count);  // from inst_world_in
int refit_tlas(Ctx& c);
int rebuild_tlas(Ctx& c);  // device-only TLAS build over the current instance boxes
int tlas_sah(Ctx& c, double& out);  // SAH cost of the TLAS as last refitted or rebuilt (synchronises)
int update_lights(Ctx& c);  // world-space emissive triangles from the instance matrices
// ProjectHdrToSh (Texture.cpp:180-253): 9 spherical-harmonic coefficients per channel of an
// equirectangular RGBA32F image, out[channel * 9 + i]
void project_sky_sh(const float* rgba, uint32_t w, uint32_t h, float out[27]);

// trace.hip. A trace launch runs in a slot (0..3): 0 = `stream` with wf.spill, h > 0 =
// pstream[h - 1] with wf.pspill[h - 1]; launches in different slots may run at the same time.
int launch_trace(Ctx& c, const float4* ray_o, const float4* ray_d, const uint32_t* count_dev,
                 uint32_t max_count, float4* hit, uint32_t* hit_inst, bool any_hit,
                 uint32_t* clear, int slot = 0);
// a closest-hit queue (o, d -> hit, hit_inst) and an occlusion queue (so, sd -> shit) in one launch
int launch_trace_pair(Ctx& c, const float4* o, const float4* d, const uint32_t* count_dev, uint32_t max_count,
                      float4* hit, uint32_t* hit_inst, const float4* so, const float4* sd,
                      const uint32_t* scount_dev, uint32_t smax_count, float4* shit, int slot = 0);
int launch_trace_user(Ctx& c, const gk_ray* rays, uint32_t n, gk_hit* hits, uint32_t* inst, uint32_t flags);
// the ray counter a trace launch in `slot` starts from (allocates the slot's spill area); a
// caller whose kernel zeroed it sets Ctx::trace_counter_zeroed[slot], and the next launch in
// that slot skips its reset
uint32_t* trace_counter(Ctx& c, int slot = 0);
int ensure_spill(Ctx& c, uint32_t threads, int slot = 0);
int collect_timing(Ctx& c);
// gk_trace_static's device array: resolved per GPU by gk_create; the overflow flag of the
// last call on c's GPU, read and cleared
int static_scratch_init(int device);
int static_scratch_overflow(Ctx& c, bool& overflow);

}  // namespace gk

struct gk_ctx {
    gk::Ctx c;
};
// Entry points that may queue work on the context's stream or change what a frame's path
// pipelines read (scene, instances, queues, slots) end the pipelines' early start of the
// next frame; frame set-up, render, sync and read-only queries use GK_CHECK_CTX_KEEP.
#define GK_CHECK_CTX(ctx) do { if (!(ctx)) return GK_ERR_ARG; (ctx)->c.early_ok = false; } while (0)
#define GK_CHECK_CTX_KEEP(ctx) do { if (!(ctx)) return GK_ERR_ARG; } while (0)

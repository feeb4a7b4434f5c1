// api.hip - the C ABI (include/gknx_rt.h): context, scene upload, frame constants,
// render targets, ray casting and host helpers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>

#include "gk_ctx.hpp"

namespace gk {

// layouts the ABI fixes (the reference's byte-identical records, and the ctypes binding)
static_assert(sizeof(gk_mesh) == 56 && sizeof(gk_texture) == 24, "mesh / texture records");
static_assert(sizeof(gk_material) == 64 && sizeof(gk_light) == 80 && sizeof(gk_instance) == 224, "scene records");
static_assert(sizeof(gk_ambient_cube) == 56 && sizeof(gk_voxel) == 16, "probe records");

namespace {
thread_local char g_last_error[256] = {0};
}

void set_error(const char* what, hipError_t e) {
    std::snprintf(g_last_error, sizeof(g_last_error), "%s: %s", what, hipGetErrorString(e));
}

int render_frame(Ctx& c);          // wavefront.hip
int post_process(Ctx& c);          // post.hip

namespace {

// bytes per pixel of each target
size_t target_bpp(int t) {
    switch (t) {
        case GK_RT_SINGLE_DIFFUSE: case GK_RT_SINGLE_SPECULAR: case GK_RT_ALBEDO: case GK_RT_NORMAL:
        case GK_RT_ACCUM_DIFFUSE: case GK_RT_ACCUM_SPECULAR: case GK_RT_ACCUM_ALBEDO: case GK_RT_DENOISED:
            return 8;  // RGBA16F
        case GK_RT_OBJECTID_0: case GK_RT_OBJECTID_1: case GK_RT_MINIGBUFFER: case GK_RT_PRIMARY_INST:
        case GK_RT_PREV_DEPTH:
            return 4;
        case GK_RT_MOTIONVECTOR: return 8;
        case GK_RT_MOTIONMOMENT: return 2;
        case GK_RT_PRIMARY_HIT: case GK_RT_RADIANCE: return 16;
        case GK_RT_RADIANCE64: return 32;
        default: return 0;
    }
}

void free_targets(Ctx& c) {
    for (void* p : c.target_allocs) (void)hipFree(p);
    c.target_allocs.clear();
    for (int t = 0; t < GK_RT_COUNT; ++t) {
        if (c.target_ptr[t]) device_bytes() -= c.target_bytes[t];
        c.target_ptr[t] = nullptr;
        c.target_bytes[t] = 0;
    }
}

__global__ void k_fill_u16(uint16_t* p, size_t n, uint16_t v) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int alloc_targets(Ctx& c) {
    free_targets(c);
    const size_t px = size_t(c.width) * c.height;
    for (int t = 0; t < GK_RT_COUNT; ++t) {
        const size_t b = px * target_bpp(t);
        void* p = nullptr;
        if (device_bytes() + b > device_budget()) return GK_ERR_OOM;
        GK_HIP(hipMalloc(&p, b));
        device_bytes() += b;
        c.target_allocs.push_back(p);
        c.target_ptr[t] = p;
        c.target_bytes[t] = b;
        // zero-fill with a kernel (safe under graph capture)
        const size_t n16 = b / 2;
        k_fill_u16<<<uint32_t((n16 + 255) / 256), 256, 0, c.stream>>>(static_cast<uint16_t*>(p), n16, 0);
    }
    GK_HIP(hipGetLastError());
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

// After a failed gk_scene_upload: the scene's device buffers and the build scratch are
// released, so that a failed (for instance out-of-memory) upload leaves no memory behind.
void release_scene(Ctx& c) {
    c.scene_ready = false;
    c.materials.release(); c.tris.release(); c.tri_nrm.release(); c.tri_mat.release(); c.tri_rec.release();
    c.tri_uv.release(); c.tri_tan.release(); c.proxy_inst.release(); c.blas_nodes.release(); c.meshes.release();
    c.inst.release(); c.inst_world_in.release(); c.inst_lo.release(); c.inst_hi.release();
    c.tlas_nodes.release(); c.tlas_prims.release(); c.tlas_refit_order.release(); c.tlas_level_info.release();
    c.lights.release(); c.light_cdf.release(); c.light_obj.release(); c.light_power.release();
    BuildScratch& b = c.scratch;
    b.keys.release(); b.keys_alt.release(); b.vals.release(); b.vals_alt.release(); b.sort_tmp.release();
    b.prim_lo.release(); b.prim_hi.release(); b.bparent.release(); b.bchild.release(); b.brange.release();
    b.bnode_lo.release(); b.bnode_hi.release(); b.bflag.release(); b.queue_a.release(); b.queue_b.release();
    b.counters.release(); b.bounds.release(); b.node_depth.release();
    c.h_instances.clear();
    c.h_meshes.clear();
    c.num_lights = 0;
    c.num_proxies = 0;
    c.tlas_node_count = 0;
}

bool is_identity(const float* m) {
    for (int i = 0; i < 16; ++i) if (m[i] != ((i % 5 == 0) ? 1.f : 0.f)) return false;
    return true;
}

// ---------------------------------------------------------------- light list (host)
int upload_lights(Ctx& c, const gk_mesh* meshes) {
    struct L { uint32_t inst, prim, mat; float v0[3], e1[3], e2[3]; float le[3]; };
    std::vector<L> ls;
    const uint32_t nmat = uint32_t(c.h_materials.size());
    for (uint32_t ii = 0; ii < c.h_instances.size(); ++ii) {
        const gk_instance& I = c.h_instances[ii];
        if (!I.visible || I.nort) continue;
        const gk_mesh& M = meshes[I.mesh_id];
        for (uint32_t p = 0; p < M.num_triangles; ++p) {
            const uint32_t local = M.tri_material ? M.tri_material[p] : 0u;
            uint32_t mi = I.use_remap ? I.mat_id[local & 15u] : I.mat_offset + local;
            if (mi > nmat - 1) mi = nmat - 1;
            const gk_material& m = c.h_materials[mi];
            if (m.material_model != GK_MATERIAL_DIFFUSE_LIGHT) continue;
            L l;
            l.inst = ii;
            l.prim = p;
            l.mat = mi;
            const uint32_t a = M.indices[3 * p], b = M.indices[3 * p + 1], d = M.indices[3 * p + 2];
            for (int k = 0; k < 3; ++k) {
                l.v0[k] = M.positions[3 * a + k];
                l.e1[k] = M.positions[3 * b + k] - M.positions[3 * a + k];
                l.e2[k] = M.positions[3 * d + k] - M.positions[3 * a + k];
                l.le[k] = m.diffuse[k];
            }
            ls.push_back(l);
        }
    }
    c.num_lights = uint32_t(ls.size());
    c.h_light_mat.resize(ls.size());
    c.h_light_area.resize(ls.size());
    GK_HIP(c.lights.resize(std::max<size_t>(ls.size(), 1)));
    GK_HIP(c.light_obj.resize(3 * std::max<size_t>(ls.size(), 1)));
    GK_HIP(c.light_cdf.resize(ls.size() + 1));
    if (ls.empty()) return GK_OK;
    // selection weight: area * luminance(Le), area from the upload-time world matrix
    std::vector<double> w(ls.size());
    std::vector<LightTri> lt(ls.size());
    std::vector<float4> obj(3 * ls.size());
    double acc = 0.0;
    std::vector<float> cdf(ls.size() + 1, 0.f);
    std::vector<double> cdfd(ls.size() + 1, 0.0);
    for (size_t i = 0; i < ls.size(); ++i) {
        const L& l = ls[i];
        const float* W = c.h_instances[l.inst].world;
        float e1[4], e2[4];
        mul44v(W, l.e1[0], l.e1[1], l.e1[2], 0.f, e1);
        mul44v(W, l.e2[0], l.e2[1], l.e2[2], 0.f, e2);
        const f3 cr = cross(f3{e1[0], e1[1], e1[2]}, f3{e2[0], e2[1], e2[2]});
        const float area = 0.5f * length(cr);
        c.h_light_mat[i] = l.mat;
        c.h_light_area[i] = area;
        const double lum = 0.2126 * l.le[0] + 0.7152 * l.le[1] + 0.0722 * l.le[2];
        acc += double(area) * lum;
        cdfd[i + 1] = acc;
        obj[3 * i] = make_float4(l.v0[0], l.v0[1], l.v0[2], 0.f);
        obj[3 * i + 1] = make_float4(l.e1[0], l.e1[1], l.e1[2], 0.f);
        obj[3 * i + 2] = make_float4(l.e2[0], l.e2[1], l.e2[2], 0.f);
        LightTri& t = lt[i];
        memset(&t, 0, sizeof(t));
        uint32_t inst_bits = l.inst, prim_bits = l.prim;
        float fi, fp;
        memcpy(&fi, &inst_bits, 4);
        memcpy(&fp, &prim_bits, 4);
        t.e2.w = fi;
        t.le = make_float4(l.le[0], l.le[1], l.le[2], fp);
    }
    for (size_t i = 0; i <= ls.size(); ++i) cdf[i] = acc > 0.0 ? float(cdfd[i] / acc) : 0.f;
    for (size_t i = 0; i < ls.size(); ++i) lt[i].e1.w = acc > 0.0 ? float((cdfd[i + 1] - cdfd[i]) / acc) : 0.f;
    GK_HIP(hipMemcpyAsync(c.lights.ptr, lt.data(), sizeof(LightTri) * lt.size(), hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipMemcpyAsync(c.light_obj.ptr, obj.data(), sizeof(float4) * obj.size(), hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipMemcpyAsync(c.light_cdf.ptr, cdf.data(), 4 * cdf.size(), hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

// Radiance and selection pdfs of the light list from the current materials: a light whose
// material is no longer an emitter gets zero radiance and zero selection weight. Emitters
// made by a material update are not in the list: their light arrives by BSDF sampling
// alone (the emitter-hit weight is 1 where the light-selection pdf is 0) until the next
// gk_scene_upload rebuilds the list.
__global__ void k_set_light_power(LightTri* __restrict__ lights, const float4* __restrict__ power, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 q = power[i];
    LightTri L = lights[i];
    L.le.x = q.x; L.le.y = q.y; L.le.z = q.z;  // le.w keeps the primitive
    L.e1.w = q.w;
    lights[i] = L;
}

int refresh_light_power(Ctx& c) {
    const size_t n = c.num_lights;
    if (n == 0) return GK_OK;
    std::vector<float4> q(n);
    std::vector<double> cdfd(n + 1, 0.0);
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) {
        const gk_material& m = c.h_materials[c.h_light_mat[i]];
        const bool emits = m.material_model == GK_MATERIAL_DIFFUSE_LIGHT;
        const float le[3] = {emits ? m.diffuse[0] : 0.f, emits ? m.diffuse[1] : 0.f, emits ? m.diffuse[2] : 0.f};
        const double lum = 0.2126 * le[0] + 0.7152 * le[1] + 0.0722 * le[2];
        acc += double(c.h_light_area[i]) * lum;
        cdfd[i + 1] = acc;
        q[i] = make_float4(le[0], le[1], le[2], 0.f);
    }
    std::vector<float> cdf(n + 1, 0.f);
    for (size_t i = 0; i <= n; ++i) cdf[i] = acc > 0.0 ? float(cdfd[i] / acc) : 0.f;
    for (size_t i = 0; i < n; ++i) q[i].w = acc > 0.0 ? float((cdfd[i + 1] - cdfd[i]) / acc) : 0.f;
    GK_HIP(c.light_power.resize(n));
    GK_HIP(hipMemcpyAsync(c.light_power.ptr, q.data(), sizeof(float4) * n, hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipMemcpyAsync(c.light_cdf.ptr, cdf.data(), 4 * cdf.size(), hipMemcpyHostToDevice, c.stream));
    k_set_light_power<<<uint32_t((n + 127) / 128), 128, 0, c.stream>>>(c.lights.ptr, c.light_power.ptr, uint32_t(n));
    GK_HIP(hipGetLastError());
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

// ---------------------------------------------------------------- small matrix helpers (host)
void mat_identity(double* m) { for (int i = 0; i < 16; ++i) m[i] = (i % 5 == 0) ? 1.0 : 0.0; }
void mat_mul(const double* a, const double* b, double* o) {  // column-major o = a*b
    double r[16];
    for (int cc = 0; cc < 4; ++cc)
        for (int rr = 0; rr < 4; ++rr) {
            double s = 0.0;
            for (int k = 0; k < 4; ++k) s += a[k * 4 + rr] * b[cc * 4 + k];
            r[cc * 4 + rr] = s;
        }
    memcpy(o, r, sizeof(r));
}
bool mat_inverse(const double* m, double* out) {  // Gauss-Jordan, column-major
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int cc = 0; cc < 4; ++cc) { a[r][cc] = m[cc * 4 + r]; a[r][4 + cc] = r == cc ? 1.0 : 0.0; }
    for (int cc = 0; cc < 4; ++cc) {
        int p = cc;
        for (int r = cc + 1; r < 4; ++r) if (std::fabs(a[r][cc]) > std::fabs(a[p][cc])) p = r;
        if (a[p][cc] == 0.0) return false;
        if (p != cc) for (int k = 0; k < 8; ++k) std::swap(a[cc][k], a[p][k]);
        const double d = a[cc][cc];
        for (int k = 0; k < 8; ++k) a[cc][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == cc) continue;
            const double f = a[r][cc];
            for (int k = 0; k < 8; ++k) a[r][k] -= f * a[cc][k];
        }
    }
    for (int r = 0; r < 4; ++r)
        for (int cc = 0; cc < 4; ++cc) out[cc * 4 + r] = a[r][4 + cc];
    return true;
}
void to_float(const double* d, float* f) { for (int i = 0; i < 16; ++i) f[i] = float(d[i]); }

// HaltonSequence (Engine.cpp:642-651)
float halton(int index, int base) {
    float f = 1.0f, r = 0.0f;
    while (index > 0) {
        f = f / base;
        r = r + f * (index % base);
        index = index / base;
    }
    return r;
}

}  // namespace

// Texture.cpp:180-253: the solid-angle weighted sum over pixel centres of each colour times
// the nine real SH basis functions (the reference's signs), in float, rows then columns.
void project_sky_sh(const float* rgba, uint32_t w, uint32_t h, float out[27]) {
    constexpr float kPiF = 3.14159265358979323846f;
    constexpr float c0 = 0.282095f, c1 = 0.488603f, c2 = 1.092548f, c3 = 0.315392f, c4 = 0.546274f;
    for (int i = 0; i < 27; ++i) out[i] = 0.f;
    for (uint32_t y = 0; y < h; ++y) {
        const float v = (float(y) + 0.5f) / float(h);
        const float theta = v * kPiF;
        const float st = std::sin(theta), ct = std::cos(theta);
        const float weight = st * (kPiF / float(h)) * (2.0f * kPiF / float(w));
        for (uint32_t x = 0; x < w; ++x) {
            const float u = (float(x) + 0.5f) / float(w);
            const float phi = u * 2.0f * kPiF;
            const float sp = std::sin(phi), cp = std::cos(phi);
            const float dx = st * cp, dy = ct, dz = st * sp;
            const float b[9] = {c0, -c1 * dy, c1 * dz, -c1 * dx, c2 * dx * dy, -c2 * dy * dz,
                                c3 * (3.0f * dy * dy - 1.0f), -c2 * dx * dz, c4 * (dx * dx - dz * dz)};
            const float* px = rgba + 4 * (size_t(y) * w + x);
            for (int i = 0; i < 9; ++i) {
                out[i] += px[0] * b[i] * weight;
                out[9 + i] += px[1] * b[i] * weight;
                out[18 + i] += px[2] * b[i] * weight;
            }
        }
    }
}
}  // namespace gk

using namespace gk;


extern "C" {

int gk_get_abi_version(void) { return GKNX_ABI_VERSION; }

const char* gk_status_string(int s) {
    switch (s) {
        case GK_OK: return "ok";
        case GK_ERR_ARG: return "invalid argument";
        case GK_ERR_HIP: return g_last_error[0] ? g_last_error : "HIP runtime error";
        case GK_ERR_OOM: return "device out of memory";
        case GK_ERR_STATE: return "invalid call order";
        case GK_ERR_LIMIT: return "capacity exceeded";
        case GK_ERR_NO_DEVICE: return "no usable HIP device";
        case GK_ERR_IO: return "file i/o failed";
        default: return "unknown status";
    }
}

int gk_create(const gk_device_desc* desc, gk_ctx** out) {
    if (!desc || !out || desc->width == 0 || desc->height == 0) return GK_ERR_ARG;
    if (desc->width > 16384 || desc->height > 16384) return GK_ERR_LIMIT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= desc->device || desc->device < 0) return GK_ERR_NO_DEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, desc->device) != hipSuccess) return GK_ERR_NO_DEVICE;
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GK_ERR_NO_DEVICE;
    gk_ctx* ctx = new gk_ctx();
    Ctx& c = ctx->c;
    c.device = desc->device;
    c.width = desc->width;
    c.height = desc->height;
    if (hipSetDevice(c.device) != hipSuccess || hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return GK_ERR_HIP;
    }
    for (auto& e : c.ev) (void)hipEventCreate(&e);
    // gk_trace_static's array on this GPU; a failure is reported by gk_trace_static itself
    (void)static_scratch_init(c.device);
    const int s = alloc_targets(c);
    if (s != GK_OK) { gk_destroy(ctx); return s; }
    *out = ctx;
    return GK_OK;
}

int gk_destroy(gk_ctx* ctx) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    for (auto& ps : c.pstream) if (ps) (void)hipStreamSynchronize(ps);
    if (c.unet) (void)gk_unet_destroy(c.unet);  // detaches itself from c
    free_targets(c);
    for (auto& e : c.ev) if (e) (void)hipEventDestroy(e);
    if (c.h_word) (void)hipHostFree(c.h_word);
    if (c.fork_ev) (void)hipEventDestroy(c.fork_ev);
    for (auto& e : c.join_ev) if (e) (void)hipEventDestroy(e);
    for (auto& ps : c.pstream) if (ps) (void)hipStreamDestroy(ps);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
    return GK_OK;
}

int gk_resize(gk_ctx* ctx, uint32_t w, uint32_t h) {
    GK_CHECK_CTX(ctx);
    if (w == 0 || h == 0) return GK_ERR_ARG;
    if (w > 16384 || h > 16384) return GK_ERR_LIMIT;
    Ctx& c = ctx->c;
    GK_HIP(hipSetDevice(c.device));
    GK_HIP(hipStreamSynchronize(c.stream));
    c.width = w;
    c.height = h;
    c.frames_rendered = 0;
    return alloc_targets(c);
}

int gk_scene_upload(gk_ctx* ctx, const gk_mesh* meshes, uint32_t num_meshes, const gk_material* materials,
                    uint32_t num_materials, const gk_light* lights, uint32_t num_lights,
                    const gk_instance* instances, uint32_t num_instances) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    // path tracing finds emitters as DIFFUSE_LIGHT triangles; the LightObject quads are kept
    // for the probe bake, which lights probes with the first one
    if (num_lights && !lights) return GK_ERR_ARG;
    if (num_meshes == 0 || !meshes || num_materials == 0 || !materials || num_instances == 0 || !instances)
        return GK_ERR_ARG;
    if (num_meshes > GK_MAX_MESHES || num_materials > GK_MAX_MATERIALS || num_instances > GK_MAX_INSTANCES)
        return GK_ERR_LIMIT;
    uint64_t total_tris = 0;
    for (uint32_t m = 0; m < num_meshes; ++m) {
        const gk_mesh& M = meshes[m];
        if (M.num_triangles && (!M.positions || !M.indices || M.num_vertices == 0)) return GK_ERR_ARG;
        total_tris += M.num_triangles;
    }
    if (total_tris >= 0x7fffffffull) return GK_ERR_LIMIT;
    for (uint32_t i = 0; i < num_instances; ++i) {
        if (instances[i].mesh_id >= num_meshes) return GK_ERR_ARG;
        for (int k = 0; k < 16; ++k)
            if (!std::isfinite(instances[i].world[k]) || !std::isfinite(instances[i].prev_world[k])) return GK_ERR_ARG;
    }
    GK_HIP(hipSetDevice(c.device));
    GK_HIP(hipStreamSynchronize(c.stream));
    c.scene_ready = false;
    c.h_materials.assign(materials, materials + num_materials);
    c.h_lights.assign(lights, lights + num_lights);
    GK_HIP(c.materials.resize(num_materials));
    GK_HIP(hipMemcpyAsync(c.materials.ptr, materials, sizeof(gk_material) * num_materials, hipMemcpyHostToDevice, c.stream));
    int s = build_blas(c, meshes, num_meshes);
    if (s == GK_OK) {
        c.h_instances.assign(instances, instances + num_instances);
        s = build_tlas(c);
    }
    if (s == GK_OK) s = upload_lights(c, meshes);
    if (s == GK_OK) s = update_lights(c);
    if (s != GK_OK) {
        (void)hipStreamSynchronize(c.stream);
        release_scene(c);
        return s;
    }
    c.blas_only = num_instances == 1 && is_identity(instances[0].world) && instances[0].visible && !instances[0].nort &&
                  instances[0].mesh_id == 0;
    c.frames_rendered = 0;
    c.scene_ready = true;
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

int gk_materials_update(gk_ctx* ctx, const gk_material* materials, uint32_t first, uint32_t count) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    if (!c.scene_ready) return GK_ERR_STATE;
    if (!materials || uint64_t(first) + count > c.h_materials.size()) return GK_ERR_ARG;
    std::copy(materials, materials + count, c.h_materials.begin() + first);
    GK_HIP(hipMemcpyAsync(c.materials.ptr + first, materials, sizeof(gk_material) * count, hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    // next-event estimation reads emitter radiance from the light list
    for (uint32_t m : c.h_light_mat)
        if (m >= first && m - first < count) return refresh_light_power(c);
    return GK_OK;
}

int gk_textures_upload(gk_ctx* ctx, const gk_texture* textures, uint32_t count) {
    GK_CHECK_CTX(ctx);
    if (count && !textures) return GK_ERR_ARG;
    uint64_t texels = 0;
    for (uint32_t i = 0; i < count; ++i) {
        const gk_texture& t = textures[i];
        if (t.width == 0 || t.height == 0 || !t.rgba8) return GK_ERR_ARG;
        if (t.width > 16384 || t.height > 16384) return GK_ERR_LIMIT;
        texels += uint64_t(t.width) * t.height;
    }
    if (texels > (uint64_t(1) << 31)) return GK_ERR_LIMIT;
    Ctx& c = ctx->c;
    GK_HIP(hipSetDevice(c.device));
    GK_HIP(hipStreamSynchronize(c.stream));
    c.num_tex = 0;
    if (count == 0) return GK_OK;
    std::vector<TexInfo> info(count);
    uint64_t off = 0;
    for (uint32_t i = 0; i < count; ++i) {
        info[i] = TexInfo{off, textures[i].width, textures[i].height, textures[i].srgb ? 1u : 0u, 0u};
        off += uint64_t(textures[i].width) * textures[i].height;
    }
    GK_HIP(c.texels.resize(std::max<uint64_t>(texels, 1)));
    GK_HIP(c.tex.resize(count));
    for (uint32_t i = 0; i < count; ++i)
        GK_HIP(hipMemcpyAsync(c.texels.ptr + info[i].offset, textures[i].rgba8, 4 * uint64_t(info[i].w) * info[i].h,
                              hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipMemcpyAsync(c.tex.ptr, info.data(), sizeof(TexInfo) * count, hipMemcpyHostToDevice, c.stream));
    // byte -> linear: the sRGB transfer function (as an R8G8B8A8_SRGB view decodes), and
    // byte / 255 for UNORM; evaluated in double on the host so every device reads the same floats
    float lut[512];
    for (int b = 0; b < 256; ++b) {
        const double x = b / 255.0;
        lut[b] = float(x <= 0.04045 ? x / 12.92 : std::pow((x + 0.055) / 1.055, 2.4));
        lut[256 + b] = float(x);
    }
    GK_HIP(c.tex_lut.resize(512));
    GK_HIP(hipMemcpyAsync(c.tex_lut.ptr, lut, sizeof(lut), hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    c.num_tex = count;
    return GK_OK;
}

int gk_sky_upload(gk_ctx* ctx, const float* rgba32f, uint32_t width, uint32_t height) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    if (width == 0 || height == 0) {
        c.sky_w = c.sky_h = 0;
        return GK_OK;
    }
    if (!rgba32f) return GK_ERR_ARG;
    if (width > 16384 || height > 16384) return GK_ERR_LIMIT;
    const size_t n = size_t(width) * height;
    for (size_t k = 0; k < 4 * n; ++k)
        if (!std::isfinite(rgba32f[k])) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(c.device));
    GK_HIP(hipStreamSynchronize(c.stream));
    c.sky_w = c.sky_h = 0;
    GK_HIP(c.sky.resize(n));
    GK_HIP(hipMemcpyAsync(c.sky.ptr, rgba32f, 16 * n, hipMemcpyHostToDevice, c.stream));
    project_sky_sh(rgba32f, width, height, c.sky_sh);
    GK_HIP(hipStreamSynchronize(c.stream));
    c.sky_w = width;
    c.sky_h = height;
    return GK_OK;
}

int gk_instances_update(gk_ctx* ctx, const float* world16, uint32_t first, uint32_t count) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    if (!c.scene_ready) return GK_ERR_STATE;
    if (!world16 || uint64_t(first) + count > c.h_instances.size()) return GK_ERR_ARG;
    for (size_t k = 0; k < size_t(count) * 16; ++k) if (!std::isfinite(world16[k])) return GK_ERR_ARG;
    for (uint32_t i = 0; i < count; ++i) {
        memcpy(c.h_instances[first + i].prev_world, c.h_instances[first + i].world, 64);
        memcpy(c.h_instances[first + i].world, world16 + 16 * size_t(i), 64);
    }
    if (count == 0) return GK_OK;
    GK_HIP(hipMemcpyAsync(c.inst_world_in.ptr, world16, 64 * size_t(count), hipMemcpyHostToDevice, c.stream));
    c.blas_only = false;
    return update_instances_device(c, first, count);
}

int gk_instances_update_device(gk_ctx* ctx, const float* d_world16, uint32_t first, uint32_t count) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    if (!c.scene_ready) return GK_ERR_STATE;
    if (!d_world16 || uint64_t(first) + count > c.h_instances.size()) return GK_ERR_ARG;
    if (count == 0) return GK_OK;
    GK_HIP(hipMemcpyAsync(c.inst_world_in.ptr, d_world16, 64 * size_t(count), hipMemcpyDeviceToDevice, c.stream));
    c.blas_only = false;
    return update_instances_device(c, first, count);
}

int gk_get_bvh_stats(gk_ctx* ctx, gk_bvh_stats* out) {
    GK_CHECK_CTX_KEEP(ctx);
    if (!out) return GK_ERR_ARG;
    Ctx& c = ctx->c;
    if (c.scene_ready) {
        GK_HIP(hipSetDevice(c.device));
        const int s = tlas_sah(c, c.bvh_stats.tlas_sah_cost);
        if (s != GK_OK) return s;
    }
    *out = c.bvh_stats;
    return GK_OK;
}

int gk_frame(gk_ctx* ctx, const gk_frame_params* p) {
    GK_CHECK_CTX_KEEP(ctx);
    if (!p) return GK_ERR_ARG;
    if (p->estimator > GK_ESTIMATOR_PT_MIS || p->denoiser > GK_DENOISER_UNET) return GK_ERR_ARG;
    if (p->samples > 1024 || p->bounces > 64 || p->max_bounces > 64) return GK_ERR_LIMIT;
    for (int k = 0; k < 16; ++k)
        if (!std::isfinite(p->model_view_inverse[k]) || !std::isfinite(p->projection_inverse[k])) return GK_ERR_ARG;
    // thin lens: a finite aperture >= 0, and a positive focus distance when it is open; the
    // visibility-buffer input is a pinhole's (its DOF resample is not built)
    if (!std::isfinite(p->aperture) || p->aperture < 0.f) return GK_ERR_ARG;
    if (p->aperture > 0.f && (!std::isfinite(p->focus_distance) || p->focus_distance <= 0.f)) return GK_ERR_ARG;
    if (p->aperture > 0.f && (p->flags & GK_FRAME_PRIMARY_FROM_VISBUFFER)) return GK_ERR_ARG;
    Ctx& c = ctx->c;
    // the visibility-buffer packing holds 15 bits of proxy+1 (one proxy per instance and
    // 65535-triangle section of its mesh) and the triangle within the section
    if ((p->flags & GK_FRAME_PRIMARY_FROM_VISBUFFER) && c.num_proxies > kMaxVisProxies) return GK_ERR_LIMIT;
    c.prev_params = c.params;
    c.params = *p;
    c.frame_ready = true;
    return GK_OK;
}

int gk_render(gk_ctx* ctx) {
    GK_CHECK_CTX_KEEP(ctx);
    Ctx& c = ctx->c;
    if (!c.scene_ready || !c.frame_ready) return GK_ERR_STATE;
    GK_HIP(hipSetDevice(c.device));
    int s = render_frame(c);
    if (s != GK_OK) return s;
    c.frames_rendered++;
    return GK_OK;
}

int gk_sync(gk_ctx* ctx) {
    GK_CHECK_CTX_KEEP(ctx);
    GK_HIP(hipStreamSynchronize(ctx->c.stream));
    // a split frame joins its other streams into the first; this also covers a frame whose
    // enqueue stopped on an error between the fork and the join
    for (auto& ps : ctx->c.pstream) if (ps) GK_HIP(hipStreamSynchronize(ps));
    return GK_OK;
}

int gk_reset_accumulation(gk_ctx* ctx) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    GK_HIP(hipStreamSynchronize(c.stream));
    for (int t : {GK_RT_RADIANCE, GK_RT_RADIANCE64, GK_RT_ACCUM_DIFFUSE, GK_RT_ACCUM_SPECULAR, GK_RT_ACCUM_ALBEDO,
                  GK_RT_MOTIONMOMENT, GK_RT_OBJECTID_1}) {
        const size_t n16 = c.target_bytes[t] / 2;
        k_fill_u16<<<uint32_t((n16 + 255) / 256), 256, 0, c.stream>>>(static_cast<uint16_t*>(c.target_ptr[t]), n16, 0);
    }
    GK_HIP(hipGetLastError());
    c.frames_rendered = 0;
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

int gk_read_target(gk_ctx* ctx, int target, void* host, size_t bytes) {
    GK_CHECK_CTX_KEEP(ctx);
    Ctx& c = ctx->c;
    if (target < 0 || target >= GK_RT_COUNT || !host) return GK_ERR_ARG;
    if (bytes != c.target_bytes[target]) return GK_ERR_ARG;
    GK_HIP(hipMemcpyAsync(host, c.target_ptr[target], bytes, hipMemcpyDeviceToHost, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

int gk_write_target(gk_ctx* ctx, int target, const void* host, size_t bytes) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    if (target < 0 || target >= GK_RT_COUNT || !host) return GK_ERR_ARG;
    if (bytes != c.target_bytes[target]) return GK_ERR_ARG;
    GK_HIP(hipMemcpyAsync(c.target_ptr[target], host, bytes, hipMemcpyHostToDevice, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

int gk_target_info(gk_ctx* ctx, int target, size_t* bytes, void** device_ptr) {
    GK_CHECK_CTX_KEEP(ctx);
    Ctx& c = ctx->c;
    if (target < 0 || target >= GK_RT_COUNT) return GK_ERR_ARG;
    if (bytes) *bytes = c.target_bytes[target];
    if (device_ptr) *device_ptr = c.target_ptr[target];
    return GK_OK;
}

// Statistics since the context was created: ray counts from device counters, trace
// time summed over the HIP events bracketing every trace launch, and the stage times of
// the last frame (events ev[0..6] of render_frame / post_process).
int gk_get_stats(gk_ctx* ctx, gk_stats* out) {
    GK_CHECK_CTX_KEEP(ctx);
    if (!out) return GK_ERR_ARG;
    Ctx& c = ctx->c;
    GK_HIP(hipStreamSynchronize(c.stream));
    const int s = collect_timing(c);
    if (s != GK_OK) return s;
    if (c.frames_rendered > 0) {
        float a = 0.f, b = 0.f, r = 0.f, d = 0.f;
        if (hipEventElapsedTime(&a, c.ev[0], c.ev[1]) == hipSuccess) c.stats.ms_tlas = a;
        if (hipEventElapsedTime(&b, c.ev[0], c.ev[3]) == hipSuccess) c.stats.ms_frame = b;
        if (hipEventElapsedTime(&r, c.ev[4], c.ev[5]) == hipSuccess) c.stats.ms_reproject = r;
        if (hipEventElapsedTime(&d, c.ev[5], c.ev[6]) == hipSuccess) c.stats.ms_denoise = d;
    }
    c.stats.extension_rays = c.stats.rays_traced - c.stats.shadow_rays;
    *out = c.stats;
    bool static_ovf = false;
    const int so = static_scratch_overflow(c, static_ovf);
    if (so != GK_OK) return so;
    // a traversal stack overflowed since the last report (gk_trace, gk_render, or the last
    // gk_trace_static on this GPU): those results are invalid; reported once
    const bool ovf = c.stack_overflow || static_ovf;
    c.stack_overflow = false;
    return ovf ? GK_ERR_LIMIT : GK_OK;
}

int gk_trace(gk_ctx* ctx, const gk_ray* d_rays, uint32_t n, gk_hit* d_hits, uint32_t* d_inst, uint32_t flags) {
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    if (!c.scene_ready) return GK_ERR_STATE;
    if ((!d_rays || !d_hits) && n) return GK_ERR_ARG;
    if (flags > GK_TRACE_ANY) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(c.device));
    if (c.tlas_dirty) {
        const int s = refit_tlas(c);
        if (s != GK_OK) return s;
    }
    return launch_trace_user(c, d_rays, n, d_hits, d_inst, flags);
}

int gk_make_frame_params(const gk_camera_desc* cam, uint32_t width, uint32_t height, uint32_t frame_index,
                         uint32_t taa, const gk_frame_params* prev, gk_frame_params* out) {
    if (!cam || !out || width == 0 || height == 0) return GK_ERR_ARG;
    gk_frame_params p{};
    // lookAt (right-handed), glm layout
    double e[3] = {cam->eye[0], cam->eye[1], cam->eye[2]};
    double f[3] = {double(cam->target[0]) - e[0], double(cam->target[1]) - e[1], double(cam->target[2]) - e[2]};
    double fl = std::sqrt(f[0] * f[0] + f[1] * f[1] + f[2] * f[2]);
    if (fl == 0.0) return GK_ERR_ARG;
    for (double& v : f) v /= fl;
    const double up[3] = {cam->up[0], cam->up[1], cam->up[2]};
    double s[3] = {f[1] * up[2] - f[2] * up[1], f[2] * up[0] - f[0] * up[2], f[0] * up[1] - f[1] * up[0]};
    double sl = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    if (sl == 0.0) return GK_ERR_ARG;
    for (double& v : s) v /= sl;
    const double u[3] = {s[1] * f[2] - s[2] * f[1], s[2] * f[0] - s[0] * f[2], s[0] * f[1] - s[1] * f[0]};
    double mv[16];
    mat_identity(mv);
    mv[0] = s[0]; mv[4] = s[1]; mv[8] = s[2];
    mv[1] = u[0]; mv[5] = u[1]; mv[9] = u[2];
    mv[2] = -f[0]; mv[6] = -f[1]; mv[10] = -f[2];
    mv[12] = -(s[0] * e[0] + s[1] * e[1] + s[2] * e[2]);
    mv[13] = -(u[0] * e[0] + u[1] * e[1] + u[2] * e[2]);
    mv[14] = (f[0] * e[0] + f[1] * e[1] + f[2] * e[2]);
    // perspectiveRH_ZO(fov, aspect, 0.2, 2000), then Projection[1][1] *= -1 (Engine.cpp:746-754)
    const double zn = 0.2, zf = 2000.0;
    const double th = std::tan(double(cam->fov_y_degrees) * 3.14159265358979323846 / 360.0);
    const double aspect = double(width) / double(height);
    double pr[16] = {0};
    pr[0] = 1.0 / (aspect * th);
    pr[5] = 1.0 / th;
    pr[10] = zf / (zn - zf);
    pr[11] = -1.0;
    pr[14] = -(zf * zn) / (zf - zn);
    pr[5] *= -1.0;
    double unjit[16];
    memcpy(unjit, pr, sizeof(pr));
    p.temporal_frames = 16;
    if (taa) {
        // Halton(2,3) jitter of frame % TemporalFrames (Engine.cpp:771-778)
        const int idx = int(frame_index % p.temporal_frames);
        const float jx = halton(idx + 1, 2) - 0.5f, jy = halton(idx + 1, 3) - 0.5f;
        pr[8] = double(jx / float(width) * 2.0f);
        pr[9] = double(jy / float(height) * 2.0f);
    }
    double mvi[16], pi[16], vp[16], vpu[16];
    if (!mat_inverse(mv, mvi) || !mat_inverse(pr, pi)) return GK_ERR_ARG;
    mat_mul(pr, mv, vp);
    mat_mul(unjit, mv, vpu);
    to_float(mv, p.model_view);
    to_float(pr, p.projection);
    to_float(mvi, p.model_view_inverse);
    to_float(pi, p.projection_inverse);
    to_float(vp, p.view_projection);
    to_float(vpu, p.view_projection_unjit);
    if (prev && prev->total_frames != 0) {
        memcpy(p.prev_view_projection, prev->view_projection, 64);
        memcpy(p.prev_view_projection_unjit, prev->view_projection_unjit, 64);
    } else {
        memcpy(p.prev_view_projection, p.view_projection, 64);
        memcpy(p.prev_view_projection_unjit, p.view_projection_unjit, 64);
    }
    p.aperture = cam->aperture;
    p.focus_distance = cam->focus_distance;
    // CreateUserSettings defaults (Engine.cpp:111-163) and Core.PathTracing constants
    p.paper_white_nit = 600.f;
    p.bf_sigma = 2.0f;
    p.bf_sigma_lum = 3.0f;
    p.bf_sigma_normal = 0.005f;
    p.bf_size = 5;
    p.exit_probability = 0.5f;
    p.hit_normal_offset = 0.001f;
    p.total_frames = frame_index;
    p.samples = 1;
    p.bounces = 4;
    p.max_bounces = 10;
    p.taa = taa;
    p.estimator = GK_ESTIMATOR_REF_V2;
    p.denoiser = GK_DENOISER_NONE;
    *out = p;
    return GK_OK;
}

}  // extern "C"

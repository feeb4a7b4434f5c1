/*
 * gknx_rt.h - C ABI of the MI355X (gfx950) path-tracing core.
 *
 * This is the drop-in boundary for the ray-traced renderer of gameknife/gkNextEngine.
 * Every entry point below replaces one piece of the reference's renderer/scene
 * interface; the replaced interface is cited (file:line, relative to the reference
 * repository root) on each declaration.
 *
 * Conventions
 *   - Plain C, no exceptions cross the ABI. Every call returns an int status
 *     (GK_OK == 0, negative on error; gk_status_string() names it).
 *     The reference throws std::runtime_error from Check(VkResult)
 *     (src/Vulkan/Vulkan.hpp:39); status codes replace that.
 *   - A gk_ctx owns all device memory (scene, acceleration structures, render
 *     targets, wavefront queues) and one HIP stream. Calls on one context must be
 *     serialized by the caller (the reference renders on the SDL main thread,
 *     src/DesktopMain.cpp:21-28). Host pointers are borrowed for the duration of
 *     the call only.
 *   - Pointers documented as "device" are HIP device pointers (for example the
 *     data_ptr() of a torch tensor on cuda:0); all others are host pointers.
 *   - Matrices are 4x4 float, column-major (glm layout, m[col*4+row]), exactly as
 *     the reference uploads UniformBufferObject / NodeProxy (BasicTypes.slang:22-95).
 */
#ifndef GKNX_RT_H_
#define GKNX_RT_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GKNX_ABI_VERSION 1

/* ---------------------------------------------------------------- status */
enum gk_status {
    GK_OK = 0,
    GK_ERR_ARG = -1,        /* null pointer, bad size, out-of-range index     */
    GK_ERR_HIP = -2,        /* a HIP runtime call failed                       */
    GK_ERR_OOM = -3,        /* device allocation failed                        */
    GK_ERR_STATE = -4,      /* call order violated (e.g. render before upload) */
    GK_ERR_LIMIT = -5,      /* a fixed capacity is exceeded                    */
    GK_ERR_NO_DEVICE = -6,  /* no HIP device / wrong architecture              */
    GK_ERR_IO = -7          /* file read/write failed (denoiser weights)       */
};

/* ---------------------------------------------------------------- limits
 * Mirrors MAX_NODES / MAX_MATERIALS of BasicTypes.slang:10-11.           */
#define GK_MAX_INSTANCES 65535u
#define GK_MAX_MATERIALS 8192u
#define GK_MAX_MESHES 4096u

/* Material models, BasicTypes.slang:183-188. */
enum gk_material_model {
    GK_MATERIAL_LAMBERTIAN = 0,
    GK_MATERIAL_METALLIC = 1,
    GK_MATERIAL_DIELECTRIC = 2,
    GK_MATERIAL_ISOTROPIC = 3,
    GK_MATERIAL_DIFFUSE_LIGHT = 4,
    GK_MATERIAL_MIXTURE = 5
};

/* Estimators. REF_V2 restates FPathTracingRendererV2 (Shading.slang:995-1261);
 * PT_MIS is the GGX-VNDF / next-event-estimation / MIS path tracer. */
enum gk_estimator { GK_ESTIMATOR_REF_V2 = 0, GK_ESTIMATOR_PT_MIS = 1 };

/* Post-process denoiser selection. */
enum gk_denoiser { GK_DENOISER_NONE = 0, GK_DENOISER_BILATERAL = 1, GK_DENOISER_UNET = 2 };

/* Render targets: the RT_* bindless storage images of
 * VulkanBaseRenderer::CreateRenderImages (VulkanBaseRenderer.cpp:429-463),
 * same formats, plus three targets that only this core has. */
enum gk_target {
    GK_RT_SINGLE_DIFFUSE = 0,   /* RGBA16F  this frame's diffuse radiance        */
    GK_RT_SINGLE_SPECULAR = 1,  /* RGBA16F  this frame's specular radiance       */
    GK_RT_ALBEDO = 2,           /* RGBA16F  primary albedo                       */
    GK_RT_NORMAL = 3,           /* RGBA16F  primary normal (xyz) + roughness (w) */
    GK_RT_OBJECTID_0 = 4,       /* R32UI    instance id, 65535 on miss           */
    GK_RT_OBJECTID_1 = 5,       /* R32UI    previous frame's instance id         */
    GK_RT_MOTIONVECTOR = 6,     /* RG32F    motion in pixels (prev - curr)       */
    GK_RT_MOTIONMOMENT = 7,     /* R16UI    frames since the object last moved   */
    GK_RT_PREV_DEPTH = 8,       /* R32F     ndc depth                            */
    GK_RT_ACCUM_DIFFUSE = 9,    /* RGBA16F  reprojected/accumulated diffuse      */
    GK_RT_ACCUM_SPECULAR = 10,  /* RGBA16F                                        */
    GK_RT_ACCUM_ALBEDO = 11,    /* RGBA16F                                        */
    GK_RT_DENOISED = 12,        /* RGBA16F  tonemapped output (linear radiance with GK_FRAME_LINEAR_OUTPUT) */
    GK_RT_MINIGBUFFER = 13,     /* R32UI    (proxy+1)<<17 | tri, visibility buffer (below) */
    GK_RT_PRIMARY_HIT = 14,     /* gk_hit   primary-ray hit {t,u,v,prim}         */
    GK_RT_PRIMARY_INST = 15,    /* R32UI    primary-ray hit instance             */
    GK_RT_RADIANCE = 16,        /* RGBA32F  running FP32 sum of diffuse+specular  */
    GK_RT_RADIANCE64 = 17,      /* RGBA64F  FP64 sum, written with GK_FRAME_ACCUM_FP64 */
    GK_RT_COUNT = 18
};

/* ---------------------------------------------------------------- data types */
typedef struct gk_ctx gk_ctx;

typedef struct gk_device_desc {
    int32_t device;     /* HIP device ordinal                          */
    uint32_t width;     /* render extent                               */
    uint32_t height;
    uint32_t flags;     /* reserved, 0                                 */
} gk_device_desc;

/* One mesh (a reference Model / BLAS). FP32 positions; the reference packs FP16
 * (Vertex.hpp:83-96) but its CPU BVH is FP32 (CPUAccelerationStructure.cpp:199-201). */
typedef struct gk_mesh {
    const float* positions;        /* 3 * num_vertices                            */
    const float* normals;          /* 3 * num_vertices, or NULL (face normals)    */
    const float* texcoords;        /* 2 * num_vertices (Vertex::TexCoord), or NULL (0, 0) */
    const uint32_t* indices;       /* 3 * num_triangles                           */
    const uint32_t* tri_material;  /* num_triangles local material slots, or NULL */
    uint32_t num_vertices;
    uint32_t num_triangles;
    const float* tangents;         /* 4 * num_vertices (Vertex::Tangent: xyz, handedness w),
                                      or NULL (zero: normal maps then keep the vertex normal) */
} gk_mesh;

/* Material, byte-identical to Material (BasicTypes.slang:218-236). 64 bytes. */
typedef struct gk_material {
    float diffuse[4];               /* albedo, or radiance for DIFFUSE_LIGHT */
    int32_t diffuse_texture_id;     /* -1: none; else an index of gk_textures_upload */
    int32_t mra_texture_id;
    int32_t normal_texture_id;
    float fuzziness;                /* roughness                             */
    float refraction_index;
    uint32_t material_model;        /* gk_material_model                     */
    float metalness;
    float refraction_index2;
    float normal_texture_scale;
    float reserved[3];
} gk_material;

/* An RGBA8 texture (GlobalTexturePool's LDR path, Texture.cpp:600-612). srgb: the texels
 * are sRGB-encoded (base colour; VK_FORMAT_R8G8B8A8_SRGB) and are decoded to linear before
 * filtering; otherwise UNORM (metallic-roughness, normal maps). Sampled as the reference's
 * SampleLevel(uv, 0) with its default sampler (Sampler.hpp:9-26): bilinear at level 0,
 * REPEAT in u and v. */
typedef struct gk_texture {
    uint32_t width, height;
    uint32_t srgb;
    uint32_t reserved;
    const uint8_t* rgba8;           /* width * height * 4 bytes, rows top to bottom */
} gk_texture;

/* Quad light, byte-identical to LightObject (BasicTypes.slang:171-182). */
typedef struct gk_light {
    float p0[4];
    float p1[4];
    float p3[4];
    float normal_area[4];
    uint32_t light_mat_idx;
    uint32_t reserved[3];
} gk_light;

/* Instance (scene node), the fields of NodeProxy (BasicTypes.slang:86-95) that
 * the host supplies. combinedPrevTS = prev_world * inverse(world) is derived on
 * the device (Node::TickVelocity, Node.cpp:82-106).
 * Material of a triangle = use_remap ? mat_id[local & 15] : mat_offset + local,
 * local being gk_mesh.tri_material (FetchMaterialId, Const_Func.slang:370-424;
 * the offset form lifts the reference's 16-slot limit). 224 bytes. */
typedef struct gk_instance {
    float world[16];
    float prev_world[16];
    uint32_t instance_id;   /* written to GK_RT_OBJECTID_0 */
    uint32_t mesh_id;
    uint32_t visible;
    uint32_t nort;          /* 1: not ray traced (masked out of the TLAS) */
    uint32_t mat_id[16];
    uint32_t mat_offset;
    uint32_t use_remap;
    uint32_t reserved[2];
} gk_instance;

/* Per-frame constants: the subset of UniformBufferObject (BasicTypes.slang:22-84)
 * that the path-tracing passes read, plus the estimator and denoiser switches.
 * gk_make_frame_params() fills it the way NextEngine::GetUniformBufferObject does
 * (Engine.cpp:736-849). 656 bytes. */
typedef struct gk_frame_params {
    float model_view[16];
    float projection[16];
    float model_view_inverse[16];
    float projection_inverse[16];
    float view_projection[16];
    float prev_view_projection[16];
    float view_projection_unjit[16];
    float prev_view_projection_unjit[16];
    float sun_direction[4];
    float sun_color[4];
    float sky_color[4];        /* sky radiance without a sky image; with one (gk_sky_upload)
                                  the image's value times this, component-wise: all four
                                  set to SkyIntensity is the reference's sky        */
    float aperture;            /* thin-lens diameter (FHardwarePrimaryRayCaster, Shading.slang:
                                  283-300); 0: pinhole. > 0 needs focus_distance > 0; not
                                  with GK_FRAME_PRIMARY_FROM_VISBUFFER, gk_render_radiance
                                  or gk_adjoint (GK_ERR_ARG) */
    float focus_distance;
    float paper_white_nit;
    float bf_sigma;
    float bf_sigma_lum;
    float bf_sigma_normal;
    float exit_probability;    /* REF_V2: 0.5 (Core.PathTracing.comp.slang:16)  */
    float hit_normal_offset;   /* REF_V2: 0.001 (Core.PathTracing.comp.slang:18) */
    uint32_t total_frames;     /* pcg4d seed z (Shading.slang:1232)             */
    uint32_t global_seed;      /* pcg4d seed w                                  */
    uint32_t samples;          /* NumberOfSamples                               */
    uint32_t bounces;          /* NumberOfBounces                               */
    uint32_t max_bounces;      /* MaxNumberOfBounces (dielectric primaries)     */
    uint32_t temporal_frames;
    uint32_t bf_size;          /* 0 disables the bilateral filter               */
    uint32_t estimator;        /* gk_estimator                                  */
    uint32_t denoiser;         /* gk_denoiser                                   */
    uint32_t has_sky;
    uint32_t has_sun;
    uint32_t taa;
    uint32_t progressive;
    uint32_t hdr;
    uint32_t flags;            /* GK_FRAME_* bits                               */
    float sky_rotation;        /* SkyRotation: turns the sky image by pi * this about +y
                                  (Shading.slang:194-199); 0 with no sky image       */
} gk_frame_params;

/* gk_frame_params.flags */
#define GK_FRAME_NO_MATERIAL_SORT 1u   /* shade in path order (the default; overrides 16u) */
#define GK_FRAME_MATERIAL_SORT 16u     /* sort each hit queue by material class (model x
                                          metal, misses last) before shading               */
/* Visibility-buffer ids (GK_RT_MINIGBUFFER), numbered as the reference's raster pass numbers
 * them: Scene::RebuildMeshBuffer slices each mesh into sections of at most 65535 triangles
 * (Scene.cpp:202-256) and UpdateNodesGpuDriven emits one NodeProxy per node and section
 * (Scene.cpp:472-478), so a pixel holds (proxy+1) << 17 | triangle-within-section
 * (Rast.VisibilityPass.vert.slang:18-22), 0 where no geometry covers it. Here the proxies
 * are the instances in upload order, each taking ceil(triangles / 65535) consecutive
 * numbers: triangle t of instance i is proxy proxy_base(i) + t / 65535, id t % 65535. Traced
 * frames write 0 for a proxy above 0x7ffe (the 15-bit field cannot name it). The reference
 * keeps at most 10 sections per model (it warns and drops the rest); all sections are
 * numbered here. Its provoking-vertex reorder (meshopt_generateProvokingIndexBuffer)
 * changes only the labelling of a triangle's corners, not these ids.
 * Primary hits from the caller's visibility buffer instead of traced camera rays: before
 * gk_render the caller writes GK_RT_MINIGBUFFER (device pointer from gk_target_info) with
 * those ids (FVisibilityBufferRayCasterV2::TracePrimaryRay, Shading.slang:336-370); the
 * frame then leaves that target unwritten. GK_ERR_LIMIT from gk_frame when the scene has
 * more than 0x7fff proxies. */
#define GK_FRAME_PRIMARY_FROM_VISBUFFER 32u
#define GK_FRAME_SKIP_POST 2u          /* stop after the tracer (no reproject/denoise)  */
/* Run only the post passes (reprojection, denoiser, history copy) on the targets as they
 * stand: the caller writes SINGLE_DIFFUSE/SPECULAR, ALBEDO, NORMAL, OBJECTID_0/1,
 * MOTIONVECTOR and MOTIONMOMENT (gk_write_target) as another tracer's rt pass would
 * (PathTracingRenderer::Render's reproject and denoise dispatches,
 * PathTracingRenderer.cpp:104-218, without its rt pass). No ray is traced. */
#define GK_FRAME_POST_ONLY 64u
/* Rebuild the TLAS from the instances' current boxes when instances moved, instead of
 * refitting the existing tree (the reference rebuilds every frame,
 * TopLevelAccelerationStructure.cpp:73-81): a device-only LBVH + 4-wide collapse, no host
 * round trip. A refit keeps the tree of the last build, whose boxes grow as instances drift
 * apart. */
#define GK_FRAME_TLAS_REBUILD 128u
/* REF_V2: FastGather (UserSettings::FastGather -> UniformBufferObject::FastGather, the
 * renderer's ExitAfterFirst, Core.PathTracing.comp.slang:17): a sample whose first bounce
 * neither ends on an emitter/sky nor starts at a dielectric primary ends there, multiplied
 * by the probe lookup at that bounce's vertex (Shading.slang:1118-1124). */
#define GK_FRAME_FAST_GATHER 256u
/* Frames of at least 2^18 paths run as two independent path pipelines (even and odd 8x8
 * pixel tiles) on two streams, so that one pipeline's launches overlap the other's;
 * the image is the same either way. This flag keeps one pipeline. */
#define GK_FRAME_SINGLE_PIPELINE 512u
#define GK_FRAME_ACCUM_FP64 4u         /* also add the frame's radiance to GK_RT_RADIANCE64 */
#define GK_FRAME_LINEAR_OUTPUT 8u      /* GK_RT_DENOISED holds linear radiance, not the tonemapped image */

/* Camera description for gk_make_frame_params (Assets::Camera, Scene.hpp). */
typedef struct gk_camera_desc {
    float eye[3];
    float target[3];
    float up[3];
    float fov_y_degrees;
    float aperture;
    float focus_distance;
} gk_camera_desc;

/* Ray and hit records of gk_trace. Ray: 32 bytes, hit: 16 bytes.
 * t/u/v follow tinybvh's Moeller-Trumbore (tiny_bvh.h:6815-6844): the hit point is
 * (1-u-v)*v0 + u*v1 + v*v2; prim = triangle index within its mesh;
 * t = +inf and prim = 0xffffffff on a miss. */
typedef struct gk_ray {
    float origin[3];
    float tmin;     /* hits need t > tmin */
    float dir[3];
    float tmax;     /* hits need t < tmax */
} gk_ray;

typedef struct gk_hit {
    float t;
    float u;
    float v;
    uint32_t prim;
} gk_hit;

#define GK_TRACE_CLOSEST 0u
#define GK_TRACE_ANY 1u     /* occlusion: first hit ends the search (t/u/v unspecified) */

/* Acceleration-structure statistics. */
typedef struct gk_bvh_stats {
    uint32_t blas_count;
    uint32_t blas_nodes;       /* total wide nodes over all meshes           */
    uint32_t blas_leaves;
    uint32_t tlas_nodes;
    uint64_t blas_bytes;       /* device bytes of nodes + triangle records   */
    uint64_t tlas_bytes;
    double sah_cost;           /* BVH2-equivalent SAH cost of mesh 0 (C_TRAV=C_INT=1) */
    double build_ms;           /* last BLAS build, device time               */
    double tlas_ms;            /* last TLAS build/refit, device time         */
    double tlas_sah_cost;      /* SAH cost of the TLAS as it stands (after refits or
                                  rebuilds), C_TRAV = C_INT = 1, relative to its root box */
} gk_bvh_stats;

/* Per-frame statistics (the reference's SCOPED_GPU_TIMER passes). */
typedef struct gk_stats {
    uint64_t rays_traced;      /* primary + extension + shadow rays           */
    uint64_t extension_rays;
    uint64_t shadow_rays;
    uint64_t path_vertices;    /* shaded hits                                  */
    double ms_tlas;
    double ms_trace;           /* all trace launches                           */
    double ms_shade;           /* raygen + shade + queue compaction/sort       */
    double ms_reproject;
    double ms_denoise;
    double ms_frame;           /* gk_render, end to end                        */
} gk_stats;

/* ---------------------------------------------------------------- lifecycle */

int gk_get_abi_version(void);
const char* gk_status_string(int status);

/* Replaces LogicRendererBase::OnDeviceSet + CreateSwapChain
 * (VulkanBaseRenderer.hpp:213-243; PathTracingRenderer.cpp:66-90). */
int gk_create(const gk_device_desc* desc, gk_ctx** out_ctx);
/* Replaces DeleteSwapChain + device teardown. */
int gk_destroy(gk_ctx* ctx);
/* Replaces VulkanBaseRenderer::CreateRenderImages on a swap-chain resize
 * (VulkanBaseRenderer.cpp:429-463). Clears accumulation history. */
int gk_resize(gk_ctx* ctx, uint32_t width, uint32_t height);

/* ---------------------------------------------------------------- scene */

/* Replaces Scene::Reload + RebuildMeshBuffer + OnPostLoadScene's
 * CreateAccelerationStructures (Scene.cpp:92-289; RayTraceBaseRenderer.cpp:106-124,
 * 276-386): uploads geometry, builds one BLAS per mesh on the device and the TLAS.
 * GK_ERR_LIMIT when the meshes hold 2^27 triangles or more in total; GK_ERR_ARG for no
 * instance, a mesh id, vertex index or transform that is out of range or not finite, or a
 * non-finite vertex position. An error found before the device work leaves the previous
 * scene in place; one found during it (a bad vertex, GK_ERR_OOM) leaves no scene and frees
 * the scene's device memory (gk_render: GK_ERR_STATE until an upload succeeds).
 * Material ids past the table are clamped to the last material, as GetMaterial does
 * (BasicTypes.slang:376-379). GKNX_DEVICE_MEMORY_LIMIT=<bytes> in the environment bounds the
 * device memory of the library's buffers in the process (GK_ERR_OOM beyond it). */
int gk_scene_upload(gk_ctx* ctx,
                    const gk_mesh* meshes, uint32_t num_meshes,
                    const gk_material* materials, uint32_t num_materials,
                    const gk_light* lights, uint32_t num_lights,
                    const gk_instance* instances, uint32_t num_instances);

/* Replaces Scene::UpdateAllMaterials (Scene.cpp:393-408). Next-event estimation follows:
 * the emissive triangles of the last gk_scene_upload take their radiance and selection
 * probability from the updated materials (zero for one that is no longer DIFFUSE_LIGHT).
 * A material made DIFFUSE_LIGHT by an update is lit by BSDF sampling alone until the
 * next gk_scene_upload adds its triangles to the light list (still unbiased). */
int gk_materials_update(gk_ctx* ctx, const gk_material* materials, uint32_t first, uint32_t count);

/* Replaces the context's textures (GlobalTexturePool::LoadTexture for each image of a
 * scene, Texture.cpp:260-273,369-700; FSceneLoader.cpp:253-291): texture i is what
 * gk_material texture ids i refer to. Materials sample them as FetchGBufferV2 and
 * GetRayColor do (Shading.slang:82-112,1046-1052): albedo = Diffuse * base colour texture
 * at every hit; at the primary hit also roughness = Fuzziness * MRA.g (the NORMAL
 * target's w) and the normal map (mikkTSpace with the interpolated object-space normal
 * and tangent, then the instance transform). A texture id outside [0, count) counts as
 * none. count 0 removes all textures. Texels are stored uncompressed (the reference
 * transcodes LDR images to BC7). */
int gk_textures_upload(gk_ctx* ctx, const gk_texture* textures, uint32_t count);

/* Replaces the context's sky image (an HDRI of GlobalTexturePool's hdr branch,
 * Texture.cpp:436-457, lit as Camera.SkyIdx): an equirectangular RGBA32F image, row 0 the
 * +y pole, width * height * 4 floats. Every ray that leaves the scene (REF_V2 primary and
 * GetRayColor misses, PT_MIS and gk_adjoint misses) sees SampleIBLV2(dir, sky_rotation, 0)
 * (Shading.slang:194-199): u = (atan2(d.x, d.z) + pi * sky_rotation) / (2 pi),
 * v = acos(d.y) / pi, bilinear with REPEAT on both axes, clamped to (10, 10, 10, 1), times
 * sky_color; the probe bake's misses see SampleIBLRough (Shading.slang:146-192,
 * AmbientCube.slang:515), the 9-coefficient spherical harmonics that ProjectHdrToSh
 * (Texture.cpp:180-253) computes here from the image at upload, times sky_color. Both only
 * where has_sky is set. width or height 0 (pixels may be null) removes the image: the sky
 * is then sky_color everywhere. GK_ERR_ARG for a non-finite pixel, GK_ERR_LIMIT for a side
 * above 16384; the image survives gk_scene_upload. */
int gk_sky_upload(gk_ctx* ctx, const float* rgba32f, uint32_t width, uint32_t height);

/* Replaces Scene::UpdateNodesGpuDriven + RayTraceBaseRenderer::AfterUpdateScene /
 * PreRender TLAS update (Scene.cpp:443-492; RayTraceBaseRenderer.cpp:165-218).
 * world16: count column-major 4x4 matrices (host). The previous world matrix of
 * each updated instance becomes its prev_world. The TLAS is refit on the device
 * during the next gk_render. */
int gk_instances_update(gk_ctx* ctx, const float* world16, uint32_t first, uint32_t count);

/* Same as gk_instances_update with device-resident matrices (count*16 floats). */
int gk_instances_update_device(gk_ctx* ctx, const float* d_world16, uint32_t first, uint32_t count);

int gk_get_bvh_stats(gk_ctx* ctx, gk_bvh_stats* out);

/* ---------------------------------------------------------------- frame */

/* Replaces VulkanBaseRenderer::UpdateUniformBuffer (GetUniformBufferObject,
 * Engine.cpp:736-849): sets the constants of the next gk_render. */
int gk_frame(gk_ctx* ctx, const gk_frame_params* params);

/* Replaces PathTracingRenderer::Render (PathTracingRenderer.cpp:104-218) together
 * with the PreRender TLAS update: TLAS refit -> wavefront trace/shade ->
 * reproject -> denoise -> copy accum to history. Enqueues on the context stream. */
int gk_render(gk_ctx* ctx);

/* Waits for the context stream (the reference's fence wait, VulkanBaseRenderer.cpp:832). */
int gk_sync(gk_ctx* ctx);

/* Clears temporal history and the FP32 radiance sum (progressive restart). */
int gk_reset_accumulation(gk_ctx* ctx);

/* Copies a render target to host memory. bytes must equal gk_target_bytes(). */
int gk_read_target(gk_ctx* ctx, int target, void* host, size_t bytes);
/* Copies host memory into a render target (same size rule); for inputs such as the
 * visibility buffer of GK_FRAME_PRIMARY_FROM_VISBUFFER. */
int gk_write_target(gk_ctx* ctx, int target, const void* host, size_t bytes);
/* Size in bytes and device address of a render target (for zero-copy readers). */
int gk_target_info(gk_ctx* ctx, int target, size_t* bytes, void** device_ptr);

/* Statistics since the context was created. GK_ERR_LIMIT (with *out filled) when a
 * traversal stack overflowed since the last call (gk_trace, gk_render, or the last
 * gk_trace_static on the context's GPU): those hits are invalid. Reported once. */
int gk_get_stats(gk_ctx* ctx, gk_stats* out);

/* ---------------------------------------------------------------- ray casting */

/* Batched closest-hit / any-hit query against the current scene; replaces
 * NextEngine::RayCastGPU -> FCPUAccelerationStructure::RayCastInCPU
 * (Engine.cpp:684-690; CPUAccelerationStructure.cpp:286-313) and the inline
 * RayQuery of FHardwareRayTracerV2 (Shading.slang:735-827).
 * d_rays, d_hits: device arrays of n records; d_inst: device array of n instance
 * indices, or NULL. Enqueued on the context stream; call gk_sync before reading.
 * Ties in t are broken towards the lower (instance, prim), so the result does
 * not depend on the traversal order. */
int gk_trace(gk_ctx* ctx, const gk_ray* d_rays, uint32_t n, gk_hit* d_hits,
             uint32_t* d_inst, uint32_t flags);

/* gk_trace without any device allocation. No counterpart in the reference: it is for
 * callers that cannot allocate device memory during the call (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same arguments and
 * results as gk_trace (bit-identical hits); the ray staging, traversal spill stacks, ray
 * counter and overflow flag live in one static device array of the library, about 24 MiB,
 * that every GPU the library runs on holds until the process ends. At most
 * GK_TRACE_STATIC_MAX_RAYS rays per call (GK_ERR_LIMIT above, nothing enqueued). All calls
 * on one GPU share that array, from any context: calls whose execution overlaps on one GPU
 * (two streams, two contexts, two threads) race and must be serialised by the caller.
 * Enqueued on the context stream; a traversal stack overflow of the last such call on the
 * context's GPU is reported by gk_get_stats (GK_ERR_LIMIT). The scene must be uploaded; a
 * pending instance update is refitted first, as gk_trace does. */
#define GK_TRACE_STATIC_MAX_RAYS 262144u
int gk_trace_static(gk_ctx* ctx, const gk_ray* d_rays, uint32_t n, gk_hit* d_hits,
                    uint32_t* d_inst, uint32_t flags);

/* ---------------------------------------------------------------- host helpers
 * No device needed. */

/* Fills camera matrices and defaults like GetUniformBufferObject
 * (Engine.cpp:736-849): perspective(fov, w/h, 0.2, 2000) with Y flipped, Halton
 * jitter when taa != 0, inverses, and PrevViewProjection from prev (or the current
 * view-projection when prev is NULL). */
int gk_make_frame_params(const gk_camera_desc* camera, uint32_t width, uint32_t height,
                         uint32_t frame_index, uint32_t taa,
                         const gk_frame_params* prev, gk_frame_params* out);

/* Procedural scenes. Each generator has a sizing call (pointers NULL) that
 * reports counts, then a fill call. Vertices are per-face (flat normals).
 *   cornell:  unit-cube Cornell box, 5 walls + 2 rotated boxes + ceiling light.
 *   city:     grid x grid boxes, lognormal(1, 0.6) heights, faces subdivided
 *             subdiv x subdiv, ground plane, num_lights emissive roof quads.
 *   brick:    one 24-triangle brick mesh (box with two studs).
 * Material arrays are written by gk_procgen_*_materials. */
int gk_procgen_cornell(float* positions, float* normals, uint32_t* indices, uint32_t* tri_material,
                       uint32_t* num_vertices, uint32_t* num_triangles);
int gk_procgen_cornell_materials(gk_material* materials, uint32_t* num_materials,
                                 gk_light* lights, uint32_t* num_lights);
int gk_procgen_city(uint32_t grid, uint32_t subdiv, uint32_t ground_res, uint32_t num_lights,
                    uint64_t seed,
                    float* positions, float* normals, uint32_t* indices, uint32_t* tri_material,
                    uint32_t* num_vertices, uint32_t* num_triangles,
                    gk_material* materials, uint32_t* num_materials);
int gk_procgen_brick(float* positions, float* normals, uint32_t* indices, uint32_t* tri_material,
                     uint32_t* num_vertices, uint32_t* num_triangles);

/* ---------------------------------------------------------------- differentiation
 * Path-replay backpropagation of an image loss with respect to per-material
 * parameters (albedo rgb, roughness, metalness): 5 floats per material, in
 * material order. No counterpart in the reference (it has no differentiable
 * path). The primal is the PT_MIS estimator at params->samples spp; paths are
 * replayed from their seeds, never stored.
 *   d_dl_dimage: device float4 per pixel (dL/d radiance rgb, w ignored)
 *   d_grad:      device 5*num_materials floats, overwritten                    */
int gk_adjoint(gk_ctx* ctx, const float* d_dl_dimage, float* d_grad);

/* Renders the PT_MIS primal at params->samples spp into a device float4 image
 * (the same estimator gk_adjoint differentiates). */
int gk_render_radiance(gk_ctx* ctx, float* d_image);

/* ---------------------------------------------------------------- learned denoiser
 * No counterpart in the reference renderer, whose learned denoiser is a prebuilt
 * library called from src/Rendering/PathTracing/PathTracingRenderer.cpp:244-268.
 *
 * A 5-level convolutional U-Net. Channels per level: base * {1, 1.5, 2, 3, 4}
 * rounded up to multiples of 16 (base 32: 32/48/64/96/128). Twelve 3x3 "same"
 * convolutions (zero padding), ReLU after all but the last, 2x2 average pooling
 * down, nearest-neighbour upsampling with skip concatenation up:
 *   L0 in->c0  L1 c0->c0 | L2 c0->c1 | L3 c1->c2 | L4 c2->c3 | L5 c3->c4 (1/16 res)
 *   L6 up(c4)+c3->c3 | L7 up(c3)+c2->c2 | L8 up(c2)+c1->c1 | L9 up(c1)+c0->c0
 *   L10 c0->c0  L11 c0->3 (linear)
 * Input features per pixel (10 floats, NHWC): linear radiance rgb, albedo rgb,
 * normal xyz, depth feature log(1+min(t,1e4))/10 of the primary hit distance t.
 * The network sees log(1+radiance) and predicts z = log(1+radiance); outputs are
 * max(exp(z)-1, 0). Training loss: mean over pixels and rgb of
 * (z - log(1+max(target,0)))^2.
 *
 * Parameters: one FP32 vector (Adam master copy, moments beside it); layer l is
 * a row-major [cout_pad][kpad] block at its offset, column tap*cin_pad + ci with
 * tap = 3*(dy+1) + (dx+1), the bias in column 9*cin_pad, zero elsewhere in the
 * padding. Kernels use a BF16 copy refreshed after every update.
 * Training sizes: h and w multiples of 16. Inference: any size (zero padded). */
typedef struct gk_unet gk_unet;

#define GK_UNET_IN_CHANNELS 10u
#define GK_UNET_LAYERS 12u

typedef struct gk_unet_layer {
    uint32_t cin_a;      /* channels of the first input (upsampled when up = 1) */
    uint32_t cin_b;      /* channels of the skip input (0: none)                 */
    uint32_t cout;       /* output channels, padded to a multiple of 16          */
    uint32_t cout_real;  /* 3 for the last layer, else cout                      */
    uint32_t kpad;       /* row length: 9*(cin_a+cin_b)+1 rounded up to 32       */
    uint32_t up;         /* first input read at half resolution                  */
    uint32_t relu;
    uint32_t level;      /* output resolution = input resolution >> level        */
    uint64_t offset;     /* first parameter of the layer                         */
} gk_unet_layer;

/* Creates a network on ctx's device and stream with He-normal weights drawn from
 * seed, and makes it ctx's GK_DENOISER_UNET. A context owns at most one network:
 * creating another destroys the previous one, and gk_destroy destroys it. The forward
 * pass of base_channels 32 runs kernels built for its layer shapes; with
 * GKNX_UNET_GENERIC=1 in the environment at creation it runs the runtime-shape kernels
 * instead (same results, for comparison). */
int gk_unet_create(gk_ctx* ctx, uint32_t base_channels, uint64_t seed, gk_unet** out);
int gk_unet_destroy(gk_unet* net);
int gk_unet_num_params(gk_unet* net, uint64_t* out);
int gk_unet_layer_info(gk_unet* net, uint32_t layer, gk_unet_layer* out);
/* Device pointer to the FP32 parameter vector; after writing it call
 * gk_unet_params_updated. */
int gk_unet_params(gk_unet* net, float** d_params);
int gk_unet_params_updated(gk_unet* net);
/* File: "GKUNET01", u32 base_channels, u32 reserved, u64 num_params, u64 adam step,
 * then num_params FP32 values. */
int gk_unet_save(gk_unet* net, const char* path);
int gk_unet_load(gk_unet* net, const char* path);
/* Gradient of the training loss on a batch (no update). d_input n*h*w*10 and
 * d_target n*h*w*3 (FP32 NHWC, device); d_grad num_params floats (overwritten);
 * *loss receives the loss (host). */
int gk_unet_grad(gk_unet* net, const float* d_input, const float* d_target, uint32_t n, uint32_t h,
                 uint32_t w, float* d_grad, float* loss);
/* gk_unet_grad followed by one Adam step (beta 0.9/0.999, eps 1e-8) with rate lr. */
int gk_unet_train_step(gk_unet* net, const float* d_input, const float* d_target, uint32_t n, uint32_t h,
                       uint32_t w, float lr, float* loss);
/* Inference: d_input h*w*10 -> d_output h*w*3 linear radiance (device, FP32). */
int gk_unet_infer(gk_unet* net, const float* d_input, float* d_output, uint32_t h, uint32_t w);
/* The 10 input features of ctx's current frame (accumulation targets after
 * reprojection, normal, primary hit distance) as d_input, width*height*10 FP32. */
int gk_denoise_input(gk_ctx* ctx, float* d_input);

/* ---------------------------------------------------------------- ambient-cube probes
 * A regular grid of probes (AmbientCube.slang:10-14: 192 x 48 x 192 probes, 0.25 m
 * apart, first probe at (-24, -0.34375, -24)); probe (x, y, z) is element
 * y*size_xy*size_xy + z*size_xy + x. Each probe has a voxel record (scene distance
 * field, FCPUProbeBaker) and an ambient cube (six faces of direct and of bounced light,
 * RGB10 at 512/1023 per step, plus sky visibility bytes). REF_V2 multiplies a path
 * ended by early exit or the bounce limit by the probes' trilinear, occlusion-tested
 * lookup (interpolateAmbientCubesV2<FullAmbientCubeSampler>, AmbientCube.slang:268-345;
 * Shading.slang:1147-1153). Without gk_probes_init a terminated path gets (0,0,0,1), the
 * value of a lookup outside the grid. */
typedef struct gk_probe_grid {
    uint32_t size_xy;     /* probes along x and z (CUBE_SIZE_XY) */
    uint32_t size_z;      /* probes along y (CUBE_SIZE_Z) */
    float unit;           /* spacing in metres (CUBE_UNIT) */
    float offset[3];      /* position of probe (0,0,0) (CUBE_OFFSET) */
} gk_probe_grid;

/* AmbientCube, byte-identical (BasicTypes.slang:116-138), 56 bytes. Colours are packed
 * r | g << 10 | b << 20 | a << 30 of clamp(c / 512, 0, 1) * 1023 (a: * 3). Sky
 * visibility bytes: sky_vis_pznzpyny = +z, -z, +y, -y (low byte first),
 * sky_vis_pxnxs0s1 = +x, -x, sun visibility of the +y face, unused. */
typedef struct gk_ambient_cube {
    uint32_t pos_z, neg_z, pos_y, neg_y, pos_x, neg_x;              /* bounced light */
    uint32_t pos_z_d, neg_z_d, pos_y_d, neg_y_d, pos_x_d, neg_x_d;  /* direct light  */
    uint32_t sky_vis_pznzpyny, sky_vis_pxnxs0s1;
} gk_ambient_cube;

/* VoxelData, byte-identical (BasicTypes.slang:145-151), 16 bytes. dist_gg_z01 bytes:
 * min(hit distance)/unit, 255 * product of the six clamped axis distances ("not inside"),
 * +z, -z; dist_x01_y01 bytes: +x, -x, +y, -y axis distances clamped to one unit, * 255. */
typedef struct gk_voxel {
    uint32_t mat_id;
    uint32_t age;         /* bake passes this probe has received */
    uint32_t dist_gg_z01;
    uint32_t dist_x01_y01;
} gk_voxel;

/* Allocates the grid of ctx and clears it, as FCPUProbeBaker::Init + ClearAmbientCubes
 * (CPUAccelerationStructure.cpp:165-170,491-497). grid NULL: the reference's grid.
 * At most 2^26 probes (GK_ERR_LIMIT). */
int gk_probes_init(gk_ctx* ctx, const gk_probe_grid* grid);
/* Voxelizes every probe against the current scene: VoxelizeCube
 * (CPUAccelerationStructure.cpp:85-160), run on the device instead of CPU threads
 * (AsyncProcessFull / AsyncProcessGroup, :340-431). Resets every probe's age. */
int gk_probes_voxelize(gk_ctx* ctx);
/* One bake pass over probes [first, first+count) of the grid: Bake.HwAmbientCube
 * (Bake.HwAmbientCube.comp.slang:1-19 -> FGpuProbeGenerator::RenderAmbientCubeOnly,
 * AmbientCube.slang:472-690) with sky and sun of the last gk_frame parameters and the
 * first gk_light of the scene. Only probes whose voxel lies within 8 units of a surface
 * are baked. Reads of the grid inside one pass see it as it was when the pass started
 * (the reference reads probes that other threads of the same dispatch are writing). */
int gk_probes_bake(gk_ctx* ctx, uint32_t first, uint32_t count);
/* The per-frame slice RayTraceBaseRenderer::PostRender bakes (RayTraceBaseRenderer.cpp:
 * 225-274): groups of 64 probes, grid/64/temporal_frames groups per frame, slice
 * frame_count % temporal_frames (temporal_frames 30, 120 or 300 for BakeSpeedLevel 0-2). */
int gk_probes_bake_frame(gk_ctx* ctx, uint32_t frame_count, uint32_t temporal_frames);
/* Copies probes [first, first+count) to / from host memory (either pointer may be NULL);
 * gk_probes_write is FCPUProbeBaker::UploadGPU (CPUAccelerationStructure.cpp:333-338). */
int gk_probes_read(gk_ctx* ctx, gk_ambient_cube* cubes, gk_voxel* voxels, uint32_t first, uint32_t count);
int gk_probes_write(gk_ctx* ctx, const gk_ambient_cube* cubes, const gk_voxel* voxels, uint32_t first,
                    uint32_t count);

/* ---------------------------------------------------------------- output stage
 * The compose pass writes GK_RT_DENOISED at the render extent: GT-tonemapped, or with
 * gk_frame_params.hdr ST 2084 (PQ) of the tonemapped radiance scaled by PaperWhiteNit / 230
 * (Process.DenoiseJBF.comp.slang:170-181, Const_Func.slang:51-67). */
enum gk_upscale_mode { GK_UPSCALE_LINEAR = 0, GK_UPSCALE_FSR_EASU = 1 };
/* The render extent of SuperResolution 0 / 1 / 2 (50 / 66 / 100 %): out * 2 / (4 - sr)
 * (VulkanBaseRenderer.cpp:474-476). */
int gk_render_extent(uint32_t out_w, uint32_t out_h, uint32_t super_resolution, uint32_t* w, uint32_t* h);
/* The resolve pass (VulkanBaseRenderer.cpp:1041-1067): GK_RT_DENOISED to a device RGBA16F
 * image of out_w x out_h. GK_UPSCALE_LINEAR is the live vkCmdBlitImage with
 * VK_FILTER_LINEAR (pixel centres mapped into the source, bilinear, clamp to edge);
 * GK_UPSCALE_FSR_EASU is Process.UpScaleFSR.comp.slang (FSR 1 EASU; a copy at equal
 * sizes), the compute path the reference keeps compiled out. Enqueued on ctx's stream. */
int gk_upscale(gk_ctx* ctx, int mode, uint32_t out_w, uint32_t out_h, void* d_out_rgba16f);
/* ScreenShot::SaveSwapChainToFile (src/Runtime/ScreenShot.cpp:99-283): an RGBA16F device
 * image of w x h (NULL: GK_RT_DENOISED at the render extent) stored as the UNORM swap chain
 * stores it (round(clamp(x, 0, 1) * 255)) and written as an 8-bit RGB PNG (the reference
 * writes JPEG at quality 91). Synchronises ctx's stream. */
int gk_screenshot(gk_ctx* ctx, const void* d_rgba16f, uint32_t w, uint32_t h, const char* path);

/* ---------------------------------------------------------------- glTF scenes
 * FSceneLoader::LoadGLTFScene (src/Assets/FSceneLoader.cpp:209-700) for .glb and .gltf
 * files (buffers in the GLB chunk, in files next to the .gltf, or base64 data URIs;
 * images PNG). The result is the host arrays gk_scene_upload / gk_textures_upload take:
 *   meshes     one per glTF mesh: its TRIANGLES primitives concatenated, tri_material =
 *              the primitive's section index (FSceneLoader.cpp:398-510); tangents from
 *              TANGENT, else generated from the texture coordinates
 *   materials  the reference's PBR mapping (FSceneLoader.cpp:308-388): Diffuse =
 *              (sqrt(baseColorFactor.rgb), 1), Fuzziness = roughnessFactor, Metalness =
 *              metallicFactor, Mixture by default, Metallic / Lambertian by the same
 *              thresholds, KHR_materials_ior, KHR_materials_transmission > 0.8 ->
 *              Dielectric, extras.ior2, KHR_materials_emissive_strength -> DiffuseLight
 *              of emissive * strength * 100
 *   textures   base colour (sRGB), metallic-roughness and normal images, numbered in
 *              the reference's load order (per material: base colour, MR, normal)
 *   instances  one per node with a mesh, depth-first from scene 0's roots, world =
 *              parent * T * R * S, the node's primitive materials in mat_id[] (use_remap)
 *   lights     nodes with extras.arealight (FSceneLoader.cpp:152-172)
 * plus the first camera node and scene 0's environment extras. Node matrices (as
 * opposed to TRS), animations and sparse accessors are not read. */
typedef struct gk_gltf gk_gltf;

typedef struct gk_gltf_view {
    const gk_mesh* meshes;           uint32_t num_meshes;
    const gk_material* materials;    uint32_t num_materials;
    const gk_instance* instances;    uint32_t num_instances;
    const gk_light* lights;          uint32_t num_lights;
    const gk_texture* textures;      uint32_t num_textures;
    uint32_t num_cameras;            /* camera nodes; the first one is in camera        */
    gk_camera_desc camera;           /* eye, eye + forward, up (0,1,0), fov 40          */
    uint32_t has_sky, has_sun;       /* scene 0 extras (FSceneLoader.cpp:512-550)       */
    int32_t sky_idx;
    float sky_intensity, sky_rotation, sun_intensity, sun_rotation;
} gk_gltf_view;

/* Parse a .glb / .gltf file, or .glb bytes in memory. GK_ERR_IO: unreadable file;
 * GK_ERR_ARG: malformed or unsupported content (gk_gltf_error() says what). */
int gk_gltf_load(const char* path, gk_gltf** out);
int gk_gltf_load_memory(const void* glb, size_t bytes, gk_gltf** out);
/* Pointers stay valid until gk_gltf_free. */
int gk_gltf_get_view(gk_gltf* scene, gk_gltf_view* out);
/* gk_scene_upload + gk_textures_upload of the loaded arrays. */
int gk_gltf_upload(gk_ctx* ctx, gk_gltf* scene);
int gk_gltf_free(gk_gltf* scene);
/* The calling thread's last glTF loader message ("" if none). */
const char* gk_gltf_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GKNX_RT_H_ */

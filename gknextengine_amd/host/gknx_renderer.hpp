// gknx_renderer.hpp - C++ host adapter over the C ABI (include/gknx_rt.h).
//
// Mirrors the hooks through which gkNextEngine drives its ray-traced renderer, so an
// engine-side logic renderer can forward them one to one:
//   CreateSwapChain / DeleteSwapChain   LogicRendererBase (src/Rendering/VulkanBaseRenderer.hpp:213-243),
//                                       PathTracingRenderer::CreateSwapChain (PathTracingRenderer.cpp:66-90)
//   OnPostLoadScene                     RayTraceBaseRenderer::OnPostLoadScene -> CreateAccelerationStructures
//                                       (src/Rendering/RayTraceBaseRenderer.cpp:106-124,276-386)
//   UpdateMaterials                     Scene::UpdateAllMaterials (src/Assets/Scene.cpp:393-408)
//   UpdateNodes                         Scene::UpdateNodesGpuDriven + the per-frame TLAS update
//                                       (Scene.cpp:443-492, RayTraceBaseRenderer.cpp:165-218)
//   UpdateUniformBuffer                 NextEngine::GetUniformBufferObject (src/Runtime/Engine.cpp:736-849)
//   Render                              PathTracingRenderer::Render (PathTracingRenderer.cpp:104-218)
//   BeforeNextFrame                     the frame fence wait (VulkanBaseRenderer.cpp:832)
//   VoxelizeProbes                      FCPUAccelerationStructure::AsyncProcessFull + UploadGPU
//                                       (src/Assets/CPUAccelerationStructure.cpp:333-431)
//   BakeProbes                          RayTraceBaseRenderer::PostRender's hw-lightbake slice
//                                       (src/Rendering/RayTraceBaseRenderer.cpp:225-274)
//   RayCast                             NextEngine::RayCastGPU / RayCastInCPU (Engine.cpp:684-690)
// Errors are thrown as std::runtime_error, as the reference's Check(VkResult) does
// (src/Vulkan/Vulkan.hpp:39). Header-only; link libgknx_rt.so.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gknx_rt.h"

namespace gknx {

inline void check(int status, const char* what) {
    if (status != GK_OK) throw std::runtime_error(std::string(what) + " failed: " + gk_status_string(status));
}

struct SceneArrays {
    std::vector<gk_mesh> meshes;
    std::vector<gk_material> materials;
    std::vector<gk_instance> instances;
};

class PathTracingRenderer {
public:
    explicit PathTracingRenderer(int device = 0) : device_(device) {}
    PathTracingRenderer(const PathTracingRenderer&) = delete;
    PathTracingRenderer& operator=(const PathTracingRenderer&) = delete;
    ~PathTracingRenderer() { DeleteSwapChain(); }

    void CreateSwapChain(uint32_t width, uint32_t height) {
        if (!ctx_) {
            gk_device_desc d{device_, width, height, 0u};
            check(gk_create(&d, &ctx_), "gk_create");
        } else {
            check(gk_resize(ctx_, width, height), "gk_resize");
        }
        width_ = width;
        height_ = height;
        has_prev_ = false;
    }
    void DeleteSwapChain() {
        if (ctx_) gk_destroy(ctx_);
        ctx_ = nullptr;
    }

    // Host arrays are borrowed for the duration of the call.
    void OnPostLoadScene(const SceneArrays& s, const std::vector<gk_light>& lights = {}) {
        check(gk_scene_upload(ctx(), s.meshes.data(), uint32_t(s.meshes.size()), s.materials.data(),
                              uint32_t(s.materials.size()), lights.empty() ? nullptr : lights.data(),
                              uint32_t(lights.size()), s.instances.data(), uint32_t(s.instances.size())),
              "gk_scene_upload");
        has_prev_ = false;
    }
    void UpdateMaterials(const std::vector<gk_material>& m, uint32_t first = 0) {
        check(gk_materials_update(ctx(), m.data(), first, uint32_t(m.size())), "gk_materials_update");
    }
    // worlds: count column-major 4x4 matrices (the nodes' current WorldTransform)
    void UpdateNodes(const float* worlds, uint32_t first, uint32_t count) {
        check(gk_instances_update(ctx(), worlds, first, count), "gk_instances_update");
    }

    // Frame constants from a camera; keeps last frame's matrices for motion vectors.
    gk_frame_params& UpdateUniformBuffer(const gk_camera_desc& cam, uint32_t total_frames, bool taa = false) {
        gk_frame_params p{};
        check(gk_make_frame_params(&cam, width_, height_, total_frames, taa ? 1u : 0u, has_prev_ ? &params_ : nullptr,
                                   &p),
              "gk_make_frame_params");
        params_ = p;
        has_prev_ = true;
        return params_;
    }
    gk_frame_params& Params() { return params_; }

    void Render() {
        check(gk_frame(ctx(), &params_), "gk_frame");
        check(gk_render(ctx()), "gk_render");
    }
    void BeforeNextFrame() { check(gk_sync(ctx()), "gk_sync"); }

    // Ambient-cube probes: grid nullptr = the reference's 192 x 48 x 192 grid.
    void VoxelizeProbes(const gk_probe_grid* grid = nullptr) {
        check(gk_probes_init(ctx(), grid), "gk_probes_init");
        check(gk_probes_voxelize(ctx()), "gk_probes_voxelize");
    }
    // BakeSpeedLevel 0/1/2 -> temporal_frames 30/120/300; level 2 does not bake
    void BakeProbes(uint32_t frame_count, int bake_speed_level = 1) {
        if (bake_speed_level == 2) return;
        check(gk_probes_bake_frame(ctx(), frame_count, bake_speed_level == 0 ? 30u : 120u), "gk_probes_bake_frame");
    }
    void ClearAccumulation() { check(gk_reset_accumulation(ctx()), "gk_reset_accumulation"); }

    template <class T>
    std::vector<T> ReadTarget(gk_target t) {
        size_t bytes = 0;
        void* dptr = nullptr;
        check(gk_target_info(ctx(), int(t), &bytes, &dptr), "gk_target_info");
        std::vector<T> out(bytes / sizeof(T));
        check(gk_read_target(ctx(), int(t), out.data(), out.size() * sizeof(T)), "gk_read_target");
        return out;
    }
    gk_stats Stats() {
        gk_stats s{};
        check(gk_get_stats(ctx(), &s), "gk_get_stats");
        return s;
    }
    // device rays -> device hits (enqueue only)
    void RayCast(const gk_ray* d_rays, uint32_t n, gk_hit* d_hits, uint32_t* d_inst, bool any_hit = false) {
        check(gk_trace(ctx(), d_rays, n, d_hits, d_inst, any_hit ? GK_TRACE_ANY : GK_TRACE_CLOSEST), "gk_trace");
    }
    gk_ctx* Handle() const { return ctx_; }

private:
    gk_ctx* ctx() const {
        if (!ctx_) throw std::runtime_error("CreateSwapChain has not been called");
        return ctx_;
    }
    int device_;
    gk_ctx* ctx_ = nullptr;
    uint32_t width_ = 0, height_ = 0;
    gk_frame_params params_{};
    bool has_prev_ = false;
};

// The procedural Cornell box (gk_procgen_cornell) as engine-side arrays; the vectors
// own the vertex data the gk_mesh points into.
struct CornellScene {
    std::vector<float> positions, normals;
    std::vector<uint32_t> indices, tri_material;
    SceneArrays arrays;
    CornellScene() {
        uint32_t nv = 0, nt = 0;
        check(gk_procgen_cornell(nullptr, nullptr, nullptr, nullptr, &nv, &nt), "gk_procgen_cornell");
        positions.resize(3 * nv);
        normals.resize(3 * nv);
        indices.resize(3 * nt);
        tri_material.resize(nt);
        check(gk_procgen_cornell(positions.data(), normals.data(), indices.data(), tri_material.data(), &nv, &nt),
              "gk_procgen_cornell");
        uint32_t nm = 8;
        arrays.materials.resize(nm);
        check(gk_procgen_cornell_materials(arrays.materials.data(), &nm, nullptr, nullptr),
              "gk_procgen_cornell_materials");
        arrays.materials.resize(nm);
        gk_mesh m{};
        m.positions = positions.data();
        m.normals = normals.data();
        m.indices = indices.data();
        m.tri_material = tri_material.data();
        m.num_vertices = nv;
        m.num_triangles = nt;
        arrays.meshes.push_back(m);
        gk_instance in{};
        for (int i = 0; i < 4; ++i) in.world[i * 5] = in.prev_world[i * 5] = 1.f;
        in.visible = 1;
        arrays.instances.push_back(in);
    }
};

}  // namespace gknx

// ref_tinybvh.cpp - C wrapper around the reference's vendored tinybvh 1.3.8.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile straight from the header in the
// reference tree (src/ThirdParty/tinybvh/tiny_bvh.h, included by path, never copied)
// into oracle/_ref/libtinybvh_ref.so. Used by tests/ to produce golden vectors and by
// bench.py's cpu_baseline leg; never linked into or called by the product library.
//
// Reference call sites mirrored here:
//   BLAS build      FCPUAccelerationStructure::InitBVH   CPUAccelerationStructure.cpp:172-235
//                   -> tinybvh::BVH::Build(bvhvec4*, n)   tiny_bvh.h:1510-1531,1674-1767
//   TLAS build      FCPUAccelerationStructure::UpdateBVH CPUAccelerationStructure.cpp:237-284
//                   -> BVH::Build(BLASInstance*, ...)     tiny_bvh.h:1565-1603
//   ray cast        RayCastInCPU                         CPUAccelerationStructure.cpp:286-313
//                   -> BVH::Intersect / IntersectTLAS     tiny_bvh.h:2245-2353
//   multithreading  GenShadowMap tiles of 256x256 rays   CPUAccelerationStructure.cpp:589-625
#define TINYBVH_IMPLEMENTATION
#include "tiny_bvh.h"

#include <atomic>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct Blas {
    std::vector<tinybvh::bvhvec4> verts;  // tinybvh keeps a pointer to these
    tinybvh::BVH bvh;
};

struct Tlas {
    std::vector<tinybvh::BLASInstance> inst;
    std::vector<tinybvh::BVHBase*> blas;
    tinybvh::BVH bvh;
};

double now_ms() {
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Ray setup without renormalisation: O, D exactly as given, rD = safercp(D)
// (the Ray constructor at tiny_bvh.h:562-567 additionally normalises D).
inline void make_ray(tinybvh::Ray& r, const float* rec, int normalize) {
    const tinybvh::bvhvec3 O(rec[0], rec[1], rec[2]);
    const tinybvh::bvhvec3 D(rec[4], rec[5], rec[6]);
    if (normalize) {
        r = tinybvh::Ray(O, D, rec[7]);
    } else {
        memset(&r, 0, sizeof(r));
        r.O = O;
        r.D = D;
        r.rD = tinybvh::tinybvh_safercp(D);
        r.hit.t = rec[7];
    }
}

template <class F>
void parallel_chunks(uint32_t n, uint32_t chunk, uint32_t nthreads, F&& f) {
    if (nthreads <= 1) {
        for (uint32_t b = 0; b < n; b += chunk) f(b, b + chunk < n ? b + chunk : n);
        return;
    }
    std::atomic<uint32_t> next{0};
    std::vector<std::thread> pool;
    for (uint32_t t = 0; t < nthreads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                const uint32_t b = next.fetch_add(chunk);
                if (b >= n) break;
                f(b, b + chunk < n ? b + chunk : n);
            }
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

unsigned ref_hardware_concurrency() { return std::thread::hardware_concurrency(); }

// tris: ntri * 9 floats (v0.xyz, v1.xyz, v2.xyz).
void* ref_blas_build(const float* tris, uint32_t ntri, double* build_ms) {
    Blas* b = new Blas();
    b->verts.resize(size_t(ntri) * 3);
    for (size_t i = 0; i < size_t(ntri) * 3; ++i)
        b->verts[i] = tinybvh::bvhvec4(tris[i * 3 + 0], tris[i * 3 + 1], tris[i * 3 + 2], 0.f);
    const double t0 = now_ms();
    b->bvh.Build(b->verts.data(), ntri);
    if (build_ms) *build_ms = now_ms() - t0;
    return b;
}

void ref_blas_free(void* h) { delete static_cast<Blas*>(h); }

// Node count and SAH cost (tiny_bvh.h:706-707), C_TRAV = C_INT = 1.
void ref_blas_stats(void* h, uint32_t* nodes, double* sah, float* root_min, float* root_max) {
    Blas* b = static_cast<Blas*>(h);
    if (nodes) *nodes = uint32_t(b->bvh.NodeCount());
    if (sah) *sah = double(b->bvh.SAHCost(0));
    if (root_min) { root_min[0] = b->bvh.aabbMin.x; root_min[1] = b->bvh.aabbMin.y; root_min[2] = b->bvh.aabbMin.z; }
    if (root_max) { root_max[0] = b->bvh.aabbMax.x; root_max[1] = b->bvh.aabbMax.y; root_max[2] = b->bvh.aabbMax.z; }
}

// rays: n gk_ray records (8 floats); hits: n * {t, u, v, prim(bits)}.
// tmin is ignored (tinybvh accepts t > 0). Miss: t = tmax, prim = 0xffffffff.
void ref_intersect(void* h, const float* rays, uint32_t n, float* hits, uint32_t nthreads,
                   int normalize, double* ms) {
    Blas* b = static_cast<Blas*>(h);
    const double t0 = now_ms();
    parallel_chunks(n, 65536, nthreads, [&](uint32_t lo, uint32_t hi) {
        tinybvh::Ray r;
        for (uint32_t i = lo; i < hi; ++i) {
            make_ray(r, rays + size_t(i) * 8, normalize);
            r.hit.prim = 0xffffffffu;
            b->bvh.Intersect(r);
            float* o = hits + size_t(i) * 4;
            o[0] = r.hit.t;
            o[1] = r.hit.u;
            o[2] = r.hit.v;
            memcpy(&o[3], &r.hit.prim, 4);
        }
    });
    if (ms) *ms = now_ms() - t0;
}

// Occlusion query (IsOccluded, tiny_bvh.h:2355-2397): out[i] = 1 if any hit in (0, tmax).
void ref_occluded(void* h, const float* rays, uint32_t n, uint32_t* out, uint32_t nthreads, double* ms) {
    Blas* b = static_cast<Blas*>(h);
    const double t0 = now_ms();
    parallel_chunks(n, 65536, nthreads, [&](uint32_t lo, uint32_t hi) {
        tinybvh::Ray r;
        for (uint32_t i = lo; i < hi; ++i) {
            make_ray(r, rays + size_t(i) * 8, 0);
            out[i] = b->bvh.IsOccluded(r) ? 1u : 0u;
        }
    });
    if (ms) *ms = now_ms() - t0;
}

// TLAS over instances. transforms: ninst row-major 4x4 (tinybvh BLASInstance layout,
// the transpose of the reference's glm worldTS, CPUAccelerationStructure.cpp:253-258).
void* ref_tlas_build(void* const* blas, uint32_t nblas, const float* transforms,
                     const uint32_t* blas_idx, uint32_t ninst, double* build_ms) {
    Tlas* t = new Tlas();
    t->blas.resize(nblas);
    for (uint32_t i = 0; i < nblas; ++i) t->blas[i] = &static_cast<Blas*>(blas[i])->bvh;
    t->inst.resize(ninst);
    for (uint32_t i = 0; i < ninst; ++i) {
        t->inst[i].blasIdx = blas_idx[i];
        memcpy(t->inst[i].transform, transforms + size_t(i) * 16, 64);
    }
    const double t0 = now_ms();
    t->bvh.Build(t->inst.data(), ninst, t->blas.data(), nblas);
    if (build_ms) *build_ms = now_ms() - t0;
    return t;
}

// Rebuild the TLAS after new transforms (UpdateBVH rebuilds fully every call).
void ref_tlas_rebuild(void* h, const float* transforms, double* build_ms) {
    Tlas* t = static_cast<Tlas*>(h);
    for (size_t i = 0; i < t->inst.size(); ++i) memcpy(t->inst[i].transform, transforms + i * 16, 64);
    const double t0 = now_ms();
    t->bvh.Build(t->inst.data(), uint32_t(t->inst.size()), t->blas.data(), uint32_t(t->blas.size()));
    if (build_ms) *build_ms = now_ms() - t0;
}

void ref_tlas_free(void* h) { delete static_cast<Tlas*>(h); }

void ref_tlas_intersect(void* h, const float* rays, uint32_t n, float* hits, uint32_t* inst,
                        uint32_t nthreads, double* ms) {
    Tlas* t = static_cast<Tlas*>(h);
    const double t0 = now_ms();
    parallel_chunks(n, 65536, nthreads, [&](uint32_t lo, uint32_t hi) {
        tinybvh::Ray r;
        for (uint32_t i = lo; i < hi; ++i) {
            make_ray(r, rays + size_t(i) * 8, 0);
            r.hit.prim = 0xffffffffu;
            r.hit.inst = 0xffffffffu;
            t->bvh.Intersect(r);
            float* o = hits + size_t(i) * 4;
            o[0] = r.hit.t;
            o[1] = r.hit.u;
            o[2] = r.hit.v;
            memcpy(&o[3], &r.hit.prim, 4);
            if (inst) inst[i] = r.hit.inst;
        }
    });
    if (ms) *ms = now_ms() - t0;
}

}  // extern "C"

// Host-code sanitizer run (AddressSanitizer + UndefinedBehaviorSanitizer, test
// infrastructure only): the glTF loader and the procedural generators of the C ABI
// (gknextengine_amd/csrc/gltf.cpp, procgen.cpp) and the CPU oracle (oracle/*.cpp), built with
// -fsanitize=address,undefined and driven through their entry points on valid and corrupted
// inputs. gk_scene_upload / gk_textures_upload, which the loader's gk_gltf_upload calls, are
// replaced here by functions that read every element of every array they receive, so an
// out-of-bounds view handed over by the loader is reported. Prints "sanitize ok" on success.
//
// usage: sanitize_driver <frame_params.bin> <file.glb> [<file.glb> ...]
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <random>
#include <vector>

#include "gknx_rt.h"

// ---------------------------------------------------------------- oracle exports (oracle/*.cpp)
extern "C" {
void* oracle_scene_new();
void oracle_scene_free(void* s);
int oracle_scene_add_mesh(void* s, const float* pos, const float* nrm, const uint32_t* idx, const uint32_t* tri_mat,
                          uint32_t nv, uint32_t nt);
int oracle_scene_set_materials(void* s, const gk_material* mats, uint32_t n);
int oracle_scene_add_instance(void* s, const gk_instance* in);
int oracle_scene_build(void* s);
int oracle_scene_update_instances(void* s, const float* world16, uint32_t first, uint32_t n);
void oracle_trace(void* s, const float* rays, uint32_t n, float* hits, uint32_t* inst, int any, uint32_t nthreads);
void oracle_primary_rays(const gk_frame_params* P, uint32_t W, uint32_t H, float* rays);
void oracle_lens_rays(const gk_frame_params* P, uint32_t W, uint32_t H, float* rays);
uint64_t oracle_render_refv2(void* s, const gk_frame_params* P, uint32_t W, uint32_t H, uint32_t x0, uint32_t y0,
                             uint32_t w, uint32_t h, float* diffuse, float* spec, float* albedo, float* normal,
                             uint32_t* objid, float* motion, float* depth, uint16_t* moment, uint32_t nthreads);
uint64_t oracle_render_ptmis(void* s, const gk_frame_params* P, uint32_t W, uint32_t H, uint32_t x0, uint32_t y0,
                             uint32_t w, uint32_t h, uint32_t frame0, uint32_t frames, int fp64, double* rgb,
                             double* diffuse, double* spec, uint32_t nthreads);
uint64_t oracle_render_ptmis_mode(void* s, const gk_frame_params* P, uint32_t W, uint32_t H, uint32_t x0,
                                  uint32_t y0, uint32_t w, uint32_t h, uint32_t frame0, uint32_t frames, int mode,
                                  double* rgb, uint32_t nthreads);
void oracle_diff_render(void* s, const gk_frame_params* P, const gk_material* eval, uint32_t W, uint32_t H,
                        uint32_t x0, uint32_t y0, uint32_t w, uint32_t h, uint32_t frame0, uint32_t spp, double* rgb,
                        uint32_t nthreads);
void oracle_vis_pack(void* s, const float* hits, const uint32_t* inst, uint32_t n, uint32_t* out);
void oracle_vis_decode(void* s, const uint32_t* packed, const float* rays, uint32_t n, float* out);
void oracle_probe_voxelize(void* s, const gk_probe_grid* g, gk_voxel* vox, uint32_t first, uint32_t count,
                           uint32_t nthreads);
void oracle_probe_bake(void* s, const gk_probe_grid* g, const gk_frame_params* P, const gk_light* light,
                       gk_ambient_cube* cubes, gk_voxel* vox, uint32_t first, uint32_t count, uint32_t nthreads);
void oracle_scene_set_probes(void* s, const gk_probe_grid* g, const gk_ambient_cube* cubes, const gk_voxel* vox);
void oracle_reproject(const uint32_t* params, uint32_t W, uint32_t H, const float* src_d, const float* src_s,
                      const float* src_a, const float* normal, const float* prev_d, const float* prev_s,
                      const float* prev_a, const uint32_t* obj0, const uint32_t* obj1, const float* motion,
                      const uint16_t* moment, float* acc);
void oracle_jbf(const float* fparams, uint32_t W, uint32_t H, const float* acc_d, const float* acc_s,
                const float* acc_a, const uint32_t* bf_size, float* out);
void oracle_upscale(int mode, const float* src, uint32_t sw, uint32_t sh, uint32_t dw, uint32_t dh, float* out);
}

// ---------------------------------------------------------------- upload shims
static double g_sum = 0.0;  // every value read, so that no read is optimised away
static uint32_t g_uploads = 0;

template <class T>
static void touch(const T* p, size_t n) {
    for (size_t i = 0; i < n; ++i) g_sum += double(p[i]) * 1e-30;
}

extern "C" int gk_scene_upload(gk_ctx*, const gk_mesh* meshes, uint32_t num_meshes, const gk_material* materials,
                               uint32_t num_materials, const gk_light* lights, uint32_t num_lights,
                               const gk_instance* instances, uint32_t num_instances) {
    for (uint32_t m = 0; m < num_meshes; ++m) {
        const gk_mesh& M = meshes[m];
        touch(M.positions, 3 * size_t(M.num_vertices));
        if (M.normals) touch(M.normals, 3 * size_t(M.num_vertices));
        if (M.texcoords) touch(M.texcoords, 2 * size_t(M.num_vertices));
        if (M.tangents) touch(M.tangents, 4 * size_t(M.num_vertices));
        touch(M.indices, 3 * size_t(M.num_triangles));
        if (M.tri_material) touch(M.tri_material, M.num_triangles);
        for (size_t i = 0; i < 3 * size_t(M.num_triangles); ++i)
            if (M.indices[i] >= M.num_vertices) return GK_ERR_ARG;
    }
    for (uint32_t i = 0; i < num_materials; ++i) touch(materials[i].diffuse, 4);
    for (uint32_t i = 0; i < num_lights; ++i) touch(lights[i].p0, 4);
    for (uint32_t i = 0; i < num_instances; ++i) {
        touch(instances[i].world, 16);
        if (instances[i].mesh_id >= num_meshes) return GK_ERR_ARG;
    }
    ++g_uploads;
    return GK_OK;
}

extern "C" int gk_textures_upload(gk_ctx*, const gk_texture* t, uint32_t count) {
    for (uint32_t i = 0; i < count; ++i) touch(t[i].rgba8, 4 * size_t(t[i].width) * t[i].height);
    return GK_OK;
}

static int fails = 0;
#define CHECK(cond)                                                        \
    do {                                                                   \
        if (!(cond)) { std::printf("check failed: %s (line %d)\n", #cond, __LINE__); ++fails; } \
    } while (0)

// ---------------------------------------------------------------- glTF loader
static void load_and_use(const std::vector<uint8_t>& bytes, bool must_load) {
    gk_gltf* g = nullptr;
    const int st = gk_gltf_load_memory(bytes.data(), bytes.size(), &g);
    if (must_load) CHECK(st == GK_OK);
    if (st != GK_OK) return;
    gk_gltf_view v;
    CHECK(gk_gltf_get_view(g, &v) == GK_OK);
    (void)gk_gltf_upload(reinterpret_cast<gk_ctx*>(&g_sum), g);  // any non-null handle: the shims ignore it
    CHECK(gk_gltf_free(g) == GK_OK);
}

static void gltf_files(int argc, char** argv) {
    std::mt19937 rng(11);
    for (int a = 2; a < argc; ++a) {
        std::ifstream f(argv[a], std::ios::binary);
        const std::vector<uint8_t> bytes((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        CHECK(!bytes.empty());
        gk_gltf* g = nullptr;
        CHECK(gk_gltf_load(argv[a], &g) == GK_OK);
        if (g) CHECK(gk_gltf_free(g) == GK_OK);
        load_and_use(bytes, true);
        // every truncation length in steps, then single-byte corruptions
        for (size_t n = 0; n < bytes.size(); n += 1 + bytes.size() / 400) {
            std::vector<uint8_t> cut(bytes.begin(), bytes.begin() + n);
            load_and_use(cut, false);
        }
        for (int k = 0; k < 1500; ++k) {
            std::vector<uint8_t> bad = bytes;
            const size_t at = rng() % bad.size();
            bad[at] = uint8_t(rng());
            if (k % 3 == 0) bad[rng() % bad.size()] ^= 0x80;
            load_and_use(bad, false);
        }
    }
}

// ---------------------------------------------------------------- procedural scenes + oracle
static void oracle_passes(const gk_frame_params& P0) {
    uint32_t nv = 0, nt = 0;
    CHECK(gk_procgen_cornell(nullptr, nullptr, nullptr, nullptr, &nv, &nt) == GK_OK);
    std::vector<float> pos(3 * nv), nrm(3 * nv);
    std::vector<uint32_t> idx(3 * nt), tm(nt);
    CHECK(gk_procgen_cornell(pos.data(), nrm.data(), idx.data(), tm.data(), &nv, &nt) == GK_OK);
    gk_material mats[8];
    gk_light lights[4];
    uint32_t nm = 8, nl = 4;
    CHECK(gk_procgen_cornell_materials(mats, &nm, lights, &nl) == GK_OK);
    // a small city and the brick, sized then filled
    uint32_t cv = 0, ct = 0, cm = 0;
    CHECK(gk_procgen_city(8, 2, 16, 4, 5, nullptr, nullptr, nullptr, nullptr, &cv, &ct, nullptr, &cm) == GK_OK);
    std::vector<float> cpos(3 * cv), cnrm(3 * cv);
    std::vector<uint32_t> cidx(3 * ct), ctm(ct);
    std::vector<gk_material> cmats(cm);
    CHECK(gk_procgen_city(8, 2, 16, 4, 5, cpos.data(), cnrm.data(), cidx.data(), ctm.data(), &cv, &ct, cmats.data(),
                          &cm) == GK_OK);
    uint32_t bv = 0, bt = 0;
    CHECK(gk_procgen_brick(nullptr, nullptr, nullptr, nullptr, &bv, &bt) == GK_OK);
    std::vector<float> bpos(3 * bv), bnrm(3 * bv);
    std::vector<uint32_t> bidx(3 * bt), btm(bt);
    CHECK(gk_procgen_brick(bpos.data(), bnrm.data(), bidx.data(), btm.data(), &bv, &bt) == GK_OK);

    void* s = oracle_scene_new();
    CHECK(oracle_scene_add_mesh(s, pos.data(), nrm.data(), idx.data(), tm.data(), nv, nt) == 0);
    CHECK(oracle_scene_add_mesh(s, bpos.data(), bnrm.data(), bidx.data(), btm.data(), bv, bt) == 1);
    CHECK(oracle_scene_set_materials(s, mats, nm) == 0);
    gk_instance in{};
    for (int k = 0; k < 16; ++k) in.world[k] = in.prev_world[k] = (k % 5 == 0) ? 1.f : 0.f;
    in.visible = 1;
    CHECK(oracle_scene_add_instance(s, &in) == 0);
    gk_instance br = in;
    br.mesh_id = 1;
    br.instance_id = 1;
    for (int k = 0; k < 16; ++k) br.world[k] = br.prev_world[k] = (k % 5 == 0) ? 0.05f : 0.f;
    br.world[15] = br.prev_world[15] = 1.f;
    br.world[12] = 0.5f; br.world[13] = 0.1f; br.world[14] = 0.5f;
    CHECK(oracle_scene_add_instance(s, &br) >= 0);
    CHECK(oracle_scene_build(s) == 0);
    br.world[12] = 0.55f;
    CHECK(oracle_scene_update_instances(s, br.world, 1, 1) == 0);

    const uint32_t W = 24, H = 20, n = W * H;
    gk_frame_params P = P0;
    std::vector<float> rays(8 * n), hits(4 * n);
    std::vector<uint32_t> inst(n);
    oracle_primary_rays(&P, W, H, rays.data());
    oracle_trace(s, rays.data(), n, hits.data(), inst.data(), 0, 2);
    oracle_trace(s, rays.data(), n, hits.data(), inst.data(), 1, 2);
    oracle_trace(s, rays.data(), n, hits.data(), inst.data(), 0, 2);
    P.aperture = 0.05f;
    P.focus_distance = 1.8f;
    oracle_lens_rays(&P, W, H, rays.data());
    P.aperture = 0.f;
    std::vector<uint32_t> packed(n);
    std::vector<float> dec(8 * n);
    oracle_primary_rays(&P, W, H, rays.data());
    oracle_vis_pack(s, hits.data(), inst.data(), n, packed.data());
    oracle_vis_decode(s, packed.data(), rays.data(), n, dec.data());

    std::vector<float> d(4 * n), sp(4 * n), al(4 * n), nr(4 * n), mo(2 * n), dp(n);
    std::vector<uint32_t> oid(n);
    std::vector<uint16_t> mom(n, 0);
    for (uint32_t flags : {0u, 256u}) {
        P.flags = flags;
        P.estimator = GK_ESTIMATOR_REF_V2;
        P.samples = 2;
        P.bounces = P.max_bounces = 3;
        oracle_render_refv2(s, &P, W, H, 0, 0, W, H, d.data(), sp.data(), al.data(), nr.data(), oid.data(), mo.data(),
                            dp.data(), mom.data(), 2);
    }
    P.flags = 0;
    std::vector<double> rgb(3 * n), rd(3 * n), rs(3 * n);
    oracle_render_ptmis(s, &P, W, H, 0, 0, W, H, 0, 2, 0, rgb.data(), rd.data(), rs.data(), 2);
    oracle_render_ptmis(s, &P, W, H, 2, 3, W - 4, H - 5, 0, 1, 1, rgb.data(), rd.data(), rs.data(), 2);
    for (int mode = 0; mode < 6; ++mode)
        oracle_render_ptmis_mode(s, &P, W, H, 0, 0, W, H, 0, 1, mode, rgb.data(), 2);
    std::vector<gk_material> eval(mats, mats + nm);
    oracle_diff_render(s, &P, eval.data(), W, H, 0, 0, W, H, 0, 2, rgb.data(), 2);

    // probe grid: voxelize, two bake passes, REF_V2 with the grid attached
    const gk_probe_grid g{6u, 6u, 0.2f, {-0.1f, -0.1f, -0.1f}};
    const uint32_t np = 6 * 6 * 6;
    std::vector<gk_voxel> vox(np);
    std::vector<gk_ambient_cube> cubes(np);
    std::memset(vox.data(), 0, sizeof(gk_voxel) * np);
    std::memset(cubes.data(), 0, sizeof(gk_ambient_cube) * np);
    oracle_probe_voxelize(s, &g, vox.data(), 0, np, 2);
    P.has_sun = 1;
    oracle_probe_bake(s, &g, &P, &lights[0], cubes.data(), vox.data(), 0, np, 2);
    oracle_probe_bake(s, &g, &P, nullptr, cubes.data(), vox.data(), 0, np, 2);
    oracle_scene_set_probes(s, &g, cubes.data(), vox.data());
    P.estimator = GK_ESTIMATOR_REF_V2;
    oracle_render_refv2(s, &P, W, H, 0, 0, W, H, d.data(), sp.data(), al.data(), nr.data(), oid.data(), mo.data(),
                        dp.data(), mom.data(), 2);
    oracle_scene_set_probes(s, nullptr, nullptr, nullptr);

    // post passes and the resolve
    const uint32_t rp[2] = {0u, 16u};
    std::vector<float> acc(12 * n, 0.f), out(4 * n), up(4 * 31 * 27);
    for (uint32_t i = 0; i < n; ++i) { mo[2 * i] = float(i % 7) - 3.5f; mo[2 * i + 1] = float(i % 5) * 0.75f - 1.f; }
    oracle_reproject(rp, W, H, d.data(), sp.data(), al.data(), nr.data(), d.data(), sp.data(), al.data(), oid.data(),
                     oid.data(), mo.data(), mom.data(), acc.data());
    const float fp[5] = {2.f, 3.f, 600.f, 0.f, 0.f};
    const uint32_t bf = 5;
    oracle_jbf(fp, W, H, acc.data(), acc.data() + 4 * n, acc.data() + 8 * n, &bf, out.data());
    const float fph[5] = {2.f, 3.f, 600.f, 1.f, 1.f};
    oracle_jbf(fph, W, H, acc.data(), acc.data() + 4 * n, acc.data() + 8 * n, &bf, out.data());
    oracle_upscale(0, out.data(), W, H, 31, 27, up.data());
    oracle_upscale(1, out.data(), W, H, 31, 27, up.data());
    oracle_scene_free(s);

    // the city through the oracle's BVH builder
    void* c = oracle_scene_new();
    CHECK(oracle_scene_add_mesh(c, cpos.data(), cnrm.data(), cidx.data(), ctm.data(), cv, ct) == 0);
    CHECK(oracle_scene_set_materials(c, cmats.data(), cm) == 0);
    CHECK(oracle_scene_add_instance(c, &in) == 0);
    CHECK(oracle_scene_build(c) == 0);
    oracle_trace(c, rays.data(), n, hits.data(), inst.data(), 0, 2);
    oracle_scene_free(c);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        std::printf("usage: %s frame_params.bin file.glb ...\n", argv[0]);
        return 2;
    }
    gk_frame_params P{};
    {
        std::ifstream f(argv[1], std::ios::binary);
        f.read(reinterpret_cast<char*>(&P), sizeof(P));
        CHECK(size_t(f.gcount()) == sizeof(P));
    }
    oracle_passes(P);
    gltf_files(argc, argv);
    CHECK(g_uploads > 0);
    if (fails) {
        std::printf("%d checks failed\n", fails);
        return 1;
    }
    std::printf("sanitize ok (%u uploads, %g)\n", g_uploads, g_sum);
    return 0;
}

// procgen.cpp - procedural scenes (host code; no device needed).
//
// Counterparts of FProcModel::CreateCornellBox / CreateBox (FProcModel.cpp:16-176) and
// SceneList::CornellBox (SceneList.cpp:180-226), sized to the benchmark scenes:
//   cornell  unit cube [0,1]^3, open at z = 0: floor/ceiling/back white, left red,
//            right green (albedo 0.7), two boxes rotated about y, a 0.25 x 0.25 ceiling
//            light of radiance 15 -> 36 triangles.
//   city     grid x grid boxes (footprint 0.8, pitch 1) with lognormal(1, 0.6) heights,
//            each face split subdiv x subdiv (12 * subdiv^2 triangles per box), a ground
//            of ground_res x ground_res quads, num_lights emissive roof quads (radiance 20);
//            one Mixture material per box: roughness U(0.05,1), metallic Bernoulli(0.3),
//            albedo U(0.1,0.9)^3; seeded PCG32 (O'Neill 2014).
//   brick    a 24-triangle brick: a 1 x 0.4 x 0.5 box and two 6-triangle studs.
// Vertices are per face (flat normals), materials are referenced by local slot.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/gknx_rt.h"

namespace {

struct V { float x, y, z; };

struct Builder {
    std::vector<float> pos, nrm;
    std::vector<uint32_t> idx, mat;
    void quad(V a, V b, V c, V d, V n, uint32_t m) {  // a,b,c,d counter-clockwise around n
        const uint32_t base = uint32_t(pos.size() / 3);
        for (const V& p : {a, b, c, d}) {
            pos.insert(pos.end(), {p.x, p.y, p.z});
            nrm.insert(nrm.end(), {n.x, n.y, n.z});
        }
        idx.insert(idx.end(), {base, base + 1, base + 2, base, base + 2, base + 3});
        mat.insert(mat.end(), {m, m});
    }
    // quad split into s x s sub-quads
    void quad_sub(V a, V b, V c, V d, V n, uint32_t m, uint32_t s) {
        auto lerp = [](V p, V q, float t) { return V{p.x + (q.x - p.x) * t, p.y + (q.y - p.y) * t, p.z + (q.z - p.z) * t}; };
        for (uint32_t i = 0; i < s; ++i)
            for (uint32_t j = 0; j < s; ++j) {
                const float u0 = float(i) / s, u1 = float(i + 1) / s, v0 = float(j) / s, v1 = float(j + 1) / s;
                auto P = [&](float u, float v) { return lerp(lerp(a, b, u), lerp(d, c, u), v); };
                quad(P(u0, v0), P(u1, v0), P(u1, v1), P(u0, v1), n, m);
            }
    }
    // box rotated by angle (radians) about y around its centre; outward normals
    void box(V c, V half, float angle, uint32_t m, uint32_t s = 1, bool bottom = true) {
        const float ca = std::cos(angle), sa = std::sin(angle);
        auto R = [&](V p) { return V{c.x + ca * p.x + sa * p.z, c.y + p.y, c.z - sa * p.x + ca * p.z}; };
        auto Rn = [&](V p) { return V{ca * p.x + sa * p.z, p.y, -sa * p.x + ca * p.z}; };
        const float hx = half.x, hy = half.y, hz = half.z;
        // +x
        quad_sub(R({hx, -hy, -hz}), R({hx, hy, -hz}), R({hx, hy, hz}), R({hx, -hy, hz}), Rn({1, 0, 0}), m, s);
        // -x
        quad_sub(R({-hx, -hy, hz}), R({-hx, hy, hz}), R({-hx, hy, -hz}), R({-hx, -hy, -hz}), Rn({-1, 0, 0}), m, s);
        // +y
        quad_sub(R({-hx, hy, -hz}), R({-hx, hy, hz}), R({hx, hy, hz}), R({hx, hy, -hz}), Rn({0, 1, 0}), m, s);
        // -y
        if (bottom) quad_sub(R({-hx, -hy, hz}), R({-hx, -hy, -hz}), R({hx, -hy, -hz}), R({hx, -hy, hz}), Rn({0, -1, 0}), m, s);
        // +z
        quad_sub(R({hx, -hy, hz}), R({hx, hy, hz}), R({-hx, hy, hz}), R({-hx, -hy, hz}), Rn({0, 0, 1}), m, s);
        // -z
        quad_sub(R({-hx, -hy, -hz}), R({-hx, hy, -hz}), R({hx, hy, -hz}), R({hx, -hy, -hz}), Rn({0, 0, -1}), m, s);
    }
    int emit(float* p, float* n, uint32_t* i, uint32_t* tm, uint32_t* nv, uint32_t* nt) const {
        if (!nv || !nt) return GK_ERR_ARG;
        const uint32_t v = uint32_t(pos.size() / 3), t = uint32_t(idx.size() / 3);
        if (p || n || i || tm) {
            if (*nv < v || *nt < t) return GK_ERR_LIMIT;
            if (p) std::memcpy(p, pos.data(), pos.size() * 4);
            if (n) std::memcpy(n, nrm.data(), nrm.size() * 4);
            if (i) std::memcpy(i, idx.data(), idx.size() * 4);
            if (tm) std::memcpy(tm, mat.data(), mat.size() * 4);
        }
        *nv = v;
        *nt = t;
        return GK_OK;
    }
};

gk_material lambertian(float r, float g, float b) {
    gk_material m{};
    m.diffuse[0] = r; m.diffuse[1] = g; m.diffuse[2] = b; m.diffuse[3] = 1.f;
    m.diffuse_texture_id = m.mra_texture_id = m.normal_texture_id = -1;
    m.fuzziness = 1.0f;  // Material::Lambertian (Material.hpp:10-13)
    m.refraction_index = 1.0f;
    m.material_model = GK_MATERIAL_LAMBERTIAN;
    return m;
}
gk_material diffuse_light(float r) {
    gk_material m{};
    m.diffuse[0] = m.diffuse[1] = m.diffuse[2] = r; m.diffuse[3] = 1.f;
    m.diffuse_texture_id = m.mra_texture_id = m.normal_texture_id = -1;
    m.material_model = GK_MATERIAL_DIFFUSE_LIGHT;  // Material::DiffuseLight (Material.hpp:35-38)
    return m;
}
gk_material mixture(float r, float g, float b, float rough, float metal) {
    gk_material m{};
    m.diffuse[0] = r; m.diffuse[1] = g; m.diffuse[2] = b; m.diffuse[3] = 1.f;
    m.diffuse_texture_id = m.mra_texture_id = m.normal_texture_id = -1;
    m.fuzziness = rough;
    m.refraction_index = 1.45f;  // Material::Mixture (Material.hpp:20-23)
    m.material_model = GK_MATERIAL_MIXTURE;
    m.metalness = metal;
    return m;
}

// PCG32 (O'Neill 2014, pcg32_random_r)
struct Pcg32 {
    uint64_t state, inc;
    Pcg32(uint64_t seed, uint64_t seq) : state(0), inc((seq << 1u) | 1u) { next(); state += seed; next(); }
    uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ull + inc;
        const uint32_t xs = uint32_t(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = uint32_t(old >> 59u);
        return (xs >> rot) | (xs << ((32u - rot) & 31u));
    }
    double uniform() { return (next() >> 8) * (1.0 / 16777216.0); }  // [0,1)
    double normal() {  // Box-Muller
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-12) u1 = 1e-12;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

}  // namespace

extern "C" {

int gk_procgen_cornell(float* positions, float* normals, uint32_t* indices, uint32_t* tri_material,
                       uint32_t* num_vertices, uint32_t* num_triangles) {
    Builder b;
    const uint32_t RED = 0, GREEN = 1, WHITE = 2, LIGHT = 3;
    // walls (inward normals)
    b.quad({0, 0, 0}, {0, 0, 1}, {1, 0, 1}, {1, 0, 0}, {0, 1, 0}, WHITE);     // floor
    b.quad({0, 1, 0}, {1, 1, 0}, {1, 1, 1}, {0, 1, 1}, {0, -1, 0}, WHITE);    // ceiling
    b.quad({0, 0, 1}, {0, 1, 1}, {1, 1, 1}, {1, 0, 1}, {0, 0, -1}, WHITE);    // back
    b.quad({0, 0, 0}, {0, 1, 0}, {0, 1, 1}, {0, 0, 1}, {1, 0, 0}, RED);       // left
    b.quad({1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}, {-1, 0, 0}, GREEN);    // right
    // boxes
    b.box({0.66f, 0.15f, 0.35f}, {0.15f, 0.15f, 0.15f}, -0.314159f, WHITE);  // short, -18 deg
    b.box({0.34f, 0.30f, 0.66f}, {0.15f, 0.30f, 0.15f}, 0.296706f, WHITE);   // tall, +17 deg
    // ceiling light, facing down
    const float y = 1.0f - 1e-3f, a = 0.375f, c = 0.625f;
    b.quad({a, y, a}, {c, y, a}, {c, y, c}, {a, y, c}, {0, -1, 0}, LIGHT);
    return b.emit(positions, normals, indices, tri_material, num_vertices, num_triangles);
}

int gk_procgen_cornell_materials(gk_material* materials, uint32_t* num_materials, gk_light* lights,
                                 uint32_t* num_lights) {
    if (!num_materials) return GK_ERR_ARG;
    if (materials) {
        if (*num_materials < 4) return GK_ERR_LIMIT;
        materials[0] = lambertian(0.7f, 0.07f, 0.07f);
        materials[1] = lambertian(0.07f, 0.7f, 0.07f);
        materials[2] = lambertian(0.7f, 0.7f, 0.7f);
        materials[3] = diffuse_light(15.0f);
    }
    *num_materials = 4;
    if (num_lights) {
        if (lights) {
            if (*num_lights < 1) return GK_ERR_LIMIT;
            // LightObject of the ceiling quad (FProcModel.cpp:96-119 layout)
            gk_light L{};
            const float y = 1.0f - 1e-3f, a = 0.375f, c = 0.625f;
            L.p0[0] = a; L.p0[1] = y; L.p0[2] = c; L.p0[3] = 1;
            L.p1[0] = a; L.p1[1] = y; L.p1[2] = a; L.p1[3] = 1;
            L.p3[0] = c; L.p3[1] = y; L.p3[2] = c; L.p3[3] = 1;
            L.normal_area[1] = -1; L.normal_area[3] = (c - a) * (c - a);
            L.light_mat_idx = 3;
            lights[0] = L;
        }
        *num_lights = 1;
    }
    return GK_OK;
}

int gk_procgen_city(uint32_t grid, uint32_t subdiv, uint32_t ground_res, uint32_t num_lights, uint64_t seed,
                    float* positions, float* normals, uint32_t* indices, uint32_t* tri_material,
                    uint32_t* num_vertices, uint32_t* num_triangles, gk_material* materials,
                    uint32_t* num_materials) {
    if (grid == 0 || subdiv == 0 || !num_materials) return GK_ERR_ARG;
    const uint32_t nbox = grid * grid;
    const uint32_t nmat = nbox + 2;  // boxes, ground, light
    if (nmat > GK_MAX_MATERIALS) return GK_ERR_LIMIT;
    const bool sizing = !positions && !normals && !indices && !tri_material && !materials;
    Pcg32 rng(seed, 1u);
    std::vector<float> heights(nbox);
    std::vector<gk_material> mats(nmat);
    for (uint32_t i = 0; i < nbox; ++i) {
        heights[i] = float(std::exp(1.0 + 0.6 * rng.normal()));
        const float rough = float(0.05 + 0.95 * rng.uniform());
        const float metal = rng.uniform() < 0.3 ? 1.f : 0.f;
        const float r = float(0.1 + 0.8 * rng.uniform()), g = float(0.1 + 0.8 * rng.uniform()),
                    bb = float(0.1 + 0.8 * rng.uniform());
        mats[i] = mixture(r, g, bb, rough, metal);
    }
    const uint32_t GROUND = nbox, LIGHT = nbox + 1;
    mats[GROUND] = mixture(0.5f, 0.5f, 0.5f, 0.8f, 0.f);
    mats[LIGHT] = diffuse_light(20.0f);
    // emissive roofs: distinct random boxes
    std::vector<uint8_t> lit(nbox, 0);
    const uint32_t nl = num_lights < nbox ? num_lights : nbox;
    for (uint32_t k = 0; k < nl;) {
        const uint32_t i = rng.next() % nbox;
        if (!lit[i]) { lit[i] = 1; ++k; }
    }
    Builder b;
    if (!sizing) {
        b.pos.reserve(size_t(nbox) * 6 * subdiv * subdiv * 4 * 3);
        b.idx.reserve(size_t(nbox) * 12 * subdiv * subdiv * 3);
    }
    const float half = 0.5f * float(grid);
    uint64_t nv = 0, nt = 0;
    for (uint32_t gz = 0; gz < grid; ++gz)
        for (uint32_t gx = 0; gx < grid; ++gx) {
            const uint32_t i = gz * grid + gx;
            const float h = heights[i];
            const V c{float(gx) + 0.5f - half, 0.5f * h, float(gz) + 0.5f - half};
            if (sizing) { nv += 24ull * subdiv * subdiv; nt += 12ull * subdiv * subdiv; }
            else b.box(c, {0.4f, 0.5f * h, 0.4f}, 0.f, i, subdiv);
            if (lit[i]) {
                const float y = h + 0.02f, e = 0.3f;
                if (sizing) { nv += 4; nt += 2; }
                else b.quad({c.x - e, y, c.z - e}, {c.x - e, y, c.z + e}, {c.x + e, y, c.z + e}, {c.x + e, y, c.z - e},
                            {0, 1, 0}, LIGHT);
            }
        }
    // ground
    const float cell = float(grid) / float(ground_res);
    for (uint32_t z = 0; z < ground_res; ++z)
        for (uint32_t x = 0; x < ground_res; ++x) {
            if (sizing) { nv += 4; nt += 2; continue; }
            const float x0 = -half + cell * x, x1 = x0 + cell, z0 = -half + cell * z, z1 = z0 + cell;
            b.quad({x0, 0, z0}, {x0, 0, z1}, {x1, 0, z1}, {x1, 0, z0}, {0, 1, 0}, GROUND);
        }
    if (materials) {
        if (*num_materials < nmat) return GK_ERR_LIMIT;
        std::memcpy(materials, mats.data(), sizeof(gk_material) * nmat);
    }
    *num_materials = nmat;
    if (sizing) {
        if (!num_vertices || !num_triangles) return GK_ERR_ARG;
        if (nv > 0xffffffffull || nt > 0xffffffffull) return GK_ERR_LIMIT;
        *num_vertices = uint32_t(nv);
        *num_triangles = uint32_t(nt);
        return GK_OK;
    }
    return b.emit(positions, normals, indices, tri_material, num_vertices, num_triangles);
}

int gk_procgen_brick(float* positions, float* normals, uint32_t* indices, uint32_t* tri_material,
                     uint32_t* num_vertices, uint32_t* num_triangles) {
    Builder b;
    b.box({0.f, 0.2f, 0.f}, {0.5f, 0.2f, 0.25f}, 0.f, 0);  // 12 triangles
    // studs: top + two side walls (6 triangles each)
    for (int s = 0; s < 2; ++s) {
        const float cx = s == 0 ? -0.25f : 0.25f, y0 = 0.4f, y1 = 0.5f, e = 0.12f;
        b.quad({cx - e, y1, -e}, {cx - e, y1, e}, {cx + e, y1, e}, {cx + e, y1, -e}, {0, 1, 0}, 1);
        b.quad({cx + e, y0, -e}, {cx + e, y1, -e}, {cx + e, y1, e}, {cx + e, y0, e}, {1, 0, 0}, 1);
        b.quad({cx - e, y0, e}, {cx - e, y1, e}, {cx - e, y1, -e}, {cx - e, y0, -e}, {-1, 0, 0}, 1);
    }
    return b.emit(positions, normals, indices, tri_material, num_vertices, num_triangles);
}

}  // extern "C"

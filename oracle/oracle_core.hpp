// oracle_core.hpp - shared definitions of the CPU oracle.
//
// TEST INFRASTRUCTURE ONLY: the oracle is a plain-C++ restatement of the reference's
// path-tracing math, used by tests/ (and bench.py's cpu_baseline) as the checker of the
// HIP product. Nothing in the product library includes, links or calls it.
//
// Every arithmetic sequence that the parity tests expect to be bit-exact is written
// out operation by operation and the library is compiled with -ffp-contract=off, so
// that no multiply-add is fused behind our back.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace orc {

// ------------------------------------------------------------------ ABI mirrors
// Byte layouts of include/gknx_rt.h (the oracle does not include the product header;
// tests/ check sizeof on both sides).
struct Material {  // BasicTypes.slang:218-236
    float diffuse[4];
    int32_t diffuse_texture_id, mra_texture_id, normal_texture_id;
    float fuzziness, refraction_index;
    uint32_t material_model;
    float metalness, refraction_index2, normal_texture_scale, reserved[3];
};
static_assert(sizeof(Material) == 64, "Material layout");

struct FrameParams {  // gk_frame_params (UniformBufferObject subset, BasicTypes.slang:22-84)
    float model_view[16], projection[16], model_view_inverse[16], projection_inverse[16];
    float view_projection[16], prev_view_projection[16], view_projection_unjit[16],
        prev_view_projection_unjit[16];
    float sun_direction[4], sun_color[4], sky_color[4];
    float aperture, focus_distance, paper_white_nit, bf_sigma, bf_sigma_lum, bf_sigma_normal,
        exit_probability, hit_normal_offset;
    uint32_t total_frames, global_seed, samples, bounces, max_bounces, temporal_frames, bf_size,
        estimator, denoiser, has_sky, has_sun, taa, progressive, hdr, flags;
    float sky_rotation;  // SkyRotation
};
static_assert(sizeof(FrameParams) == 656, "FrameParams layout");

enum { MAT_LAMBERTIAN = 0, MAT_METALLIC = 1, MAT_DIELECTRIC = 2, MAT_ISOTROPIC = 3,
       MAT_DIFFUSE_LIGHT = 4, MAT_MIXTURE = 5 };

// ------------------------------------------------------------------ vectors
template <class T>
struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    T& operator[](int i) { return (&x)[i]; }
    T operator[](int i) const { return (&x)[i]; }
};
template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> inline V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
// dot(a,b) = (a.x*b.x + a.y*b.y) + a.z*b.z
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
// normalize(v) = v * (1 / sqrt(dot(v, v)))  (IEEE sqrt and divide; also tinybvh_normalize
// for non-zero v, tiny_bvh.h:391-395)
template <class T> inline V3<T> normalize(V3<T> a) {
    const T l = std::sqrt(dot(a, a));
    const T rl = l == T(0) ? T(0) : T(1) / l;
    return a * rl;
}
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) {
    return {a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z};
}
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) {
    return {a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z};
}
using F3 = V3<float>;
using D3 = V3<double>;

// column-major 4x4 (glm layout) times (x, y, z, w): ((m0*x + m1*y) + m2*z) + m3*w
inline void mat_mul4(const float* m, const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r) {
        float acc = m[0 * 4 + r] * v[0];
        acc = acc + m[1 * 4 + r] * v[1];
        acc = acc + m[2 * 4 + r] * v[2];
        acc = acc + m[3 * 4 + r] * v[3];
        out[r] = acc;
    }
}

// ------------------------------------------------------------------ RNG
// pcg4d, Const_Func.slang:227-233 (Jarzynski & Olano 2020).
struct Rng {
    uint32_t v[4];
};
inline void pcg4d(Rng& s) {
    uint32_t* v = s.v;
    for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
}
// uint_to_float, Const_Func.slang:235
inline float u2f(uint32_t x) {
    const uint32_t b = 0x3f800000u | (x >> 9);
    float f;
    std::memcpy(&f, &b, 4);
    return f - 1.0f;
}
inline float rnd1(Rng& s) { pcg4d(s); return u2f(s.v[0]); }                        // RandomFloat
inline void rnd2(Rng& s, float& a, float& b) { pcg4d(s); a = u2f(s.v[0]); b = u2f(s.v[1]); }  // RandomFloat2

// ------------------------------------------------------------------ sin/cos of 2*pi*u
// Written from the Taylor series of sin/cos about 0 on |x| <= pi/4 and the quarter-turn
// symmetries; used for every azimuth phi = 2*pi*u of the samplers so that the oracle
// and the device produce identical bits. Input u in [0, 1).
template <class T>
inline void sincos_2pi(T u, T& s, T& c) {
    const T q4 = u * T(4);
    int q = int(q4);
    if (q > 3) q = 3;
    const T r = q4 - T(q);                       // [0,1): angle r*pi/2 within quadrant
    const T x = (r - T(0.5)) * T(1.5707963267948966);  // [-pi/4, pi/4)
    const T x2 = x * x;
    // sin x = x - x^3/3! + x^5/5! - x^7/7! + x^9/9!
    T sp = T(1.0 / 362880.0);
    sp = sp * x2 - T(1.0 / 5040.0);
    sp = sp * x2 + T(1.0 / 120.0);
    sp = sp * x2 - T(1.0 / 6.0);
    sp = sp * x2 + T(1);
    const T sx = sp * x;
    // cos x = 1 - x^2/2! + x^4/4! - x^6/6! + x^8/8! - x^10/10!
    T cp = T(-1.0 / 3628800.0);
    cp = cp * x2 + T(1.0 / 40320.0);
    cp = cp * x2 - T(1.0 / 720.0);
    cp = cp * x2 + T(1.0 / 24.0);
    cp = cp * x2 - T(0.5);
    cp = cp * x2 + T(1);
    const T cx = cp;
    // angle within quadrant = pi/4 + x
    const T k = T(0.70710678118654752);
    const T sq = (sx + cx) * k;  // sin(pi/4 + x)
    const T cq = (cx - sx) * k;  // cos(pi/4 + x)
    switch (q) {
        case 0: s = sq; c = cq; break;
        case 1: s = cq; c = -sq; break;
        case 2: s = -sq; c = -cq; break;
        default: s = -cq; c = sq; break;
    }
}

// ------------------------------------------------------------------ atan2 in turns
// atan2(y, x) / (2 pi) in (-1/2, 1/2], 0 at the origin, from the Maclaurin series of atan:
// the ratio t = min/max of |x|, |y| in [0, 1] is moved to |s| <= tan(pi/8) with
// atan t = pi/4 + atan((t - 1)/(t + 1)) above tan(pi/8); atan s = s (1 - s^2/3 + ... - s^18/19)
// (remainder below s^21/21 < 1e-8) by Horner's rule; then atan(1/t) = pi/2 - atan t and the
// reflections into the other octants. Used for the sky image's longitude and colatitude,
// so that the oracle and the device produce identical bits.
inline float atan2_turns(float y, float x) {
    const float ax = std::fabs(x), ay = std::fabs(y);
    const bool steep = ay > ax;
    const float hi = steep ? ay : ax, lo = steep ? ax : ay;
    if (hi == 0.f) return 0.f;
    float t = lo / hi;
    float quarter_eighths = 0.f;  // 1/8 turn = pi/4
    if (t > 0.41421356f) {
        t = (t - 1.f) / (t + 1.f);
        quarter_eighths = 0.125f;
    }
    const float t2 = t * t;
    static const float inv_odd[10] = {1.f / 19.f, 1.f / 17.f, 1.f / 15.f, 1.f / 13.f, 1.f / 11.f,
                                      1.f / 9.f,  1.f / 7.f,  1.f / 5.f,  1.f / 3.f,  1.f};
    // q = -(1 - t2/3 + t2^2/5 - ...): coefficient of t2^k is (-1)^(k+1) / (2k + 1)
    float q = inv_odd[0];
    for (int k = 1; k < 10; ++k) q = (k % 2) ? q * t2 - inv_odd[k] : q * t2 + inv_odd[k];
    float a = quarter_eighths - (q * t) * 0.15915494f;
    if (steep) a = 0.25f - a;
    if (x < 0.f) a = 0.5f - a;
    return y < 0.f ? -a : a;
}

// ------------------------------------------------------------------ RandomInUnitDisk
// concentric_sample_disk (Const_Func.slang:260-275): offset 2u - 1, zero (|x| < 1e-35 on
// both axes) maps to zero, the larger coordinate is the radius and the angle is
// pi/4 * smaller / larger; the x-major branch returns r (cos, sin), the other r (sin,
// sqrt(1 - sin^2)) as the reference writes it. sin/cos of the angle come from sincos_2pi of
// |ratio| / 8 (exact), sign restored.
template <class T>
inline void concentric_disk(T ua, T ub, T& dx, T& dy) {
    const T ox = ua + (ua - T(1)), oy = ub + (ub - T(1));
    const T nz = T(1e-35);
    if (ox > -nz && ox < nz && oy > -nz && oy < nz) { dx = T(0); dy = T(0); return; }
    const bool xmaj = std::fabs(ox) > std::fabs(oy);
    const T r = xmaj ? oy / ox : ox / oy;
    T s, c;
    sincos_2pi(std::fabs(r) * T(0.125), s, c);
    if (r < T(0)) s = -s;
    if (xmaj) { dx = ox * c; dy = ox * s; }
    else { dx = oy * s; dy = oy * std::sqrt(T(1) - s * s); }
}

// ------------------------------------------------------------------ triangle test
// Moeller-Trumbore exactly as tinybvh IntersectTri (tiny_bvh.h:6815-6844), with the
// t-window (tmin, tmax) of a RayDesc. Returns true and writes t,u,v on a hit.
template <class T>
inline bool intersect_tri(const V3<T>& O, const V3<T>& D, const V3<T>& v0, const V3<T>& e1,
                          const V3<T>& e2, T tmin, T tmax, T& t_out, T& u_out, T& v_out) {
    const V3<T> h = cross(D, e2);
    const T a = dot(e1, h);
    if (std::fabs(a) < T(0.0000001f)) return false;
    const T f = T(1) / a;
    const V3<T> s = O - v0;
    const T u = f * dot(s, h);
    if (u < T(0) || u > T(1)) return false;
    const V3<T> q = cross(s, e1);
    const T v = f * dot(D, q);
    if (v < T(0) || u + v > T(1)) return false;
    const T t = f * dot(e2, q);
    if (!(t > tmin && t < tmax)) return false;
    t_out = t; u_out = u; v_out = v;
    return true;
}

// ------------------------------------------------------------------ scene
struct Mesh {
    std::vector<F3> v0, e1, e2;     // per triangle: v0, v1-v0, v2-v0 (object space)
    std::vector<F3> n0, n1, n2;     // per-triangle vertex normals (object space)
    std::vector<uint32_t> mat;      // local material slot per triangle
    std::vector<float> uv;          // 6 per triangle (vertex texture coordinates), or empty
    std::vector<float> tan;         // 12 per triangle (vertex tangents xyzw), or empty
    // BVH2 (binned SAH, restating tinybvh BVH::Build)
    struct Node { F3 bmin; uint32_t left_first; F3 bmax; uint32_t count; };
    std::vector<Node> nodes;
    std::vector<uint32_t> prim_idx;
    F3 bmin, bmax;
};

struct Instance {
    float world[16];        // column-major
    float prev_world[16];
    float inv[16];          // row-major inverse (tinybvh InvertTransform layout)
    float xf[16];           // row-major world (tinybvh BLASInstance::transform)
    uint32_t instance_id, mesh_id, visible, nort, mat_id[16], mat_offset, use_remap;
    F3 bmin, bmax;
};

struct Light {              // one emissive triangle (world space)
    F3 p0, e1, e2, n;       // n = unit geometric normal
    float area;
    F3 le;                  // emitted radiance
    uint32_t inst, prim;
};

// ambient-cube probe grid (gk_probe_grid / AmbientCube / VoxelData layouts)
struct ProbeGridO {
    uint32_t size_xy, size_z;
    float unit, offset[3];
};
struct CubeO {            // AmbientCube, BasicTypes.slang:116-138
    uint32_t pos_z, neg_z, pos_y, neg_y, pos_x, neg_x;
    uint32_t pos_z_d, neg_z_d, pos_y_d, neg_y_d, pos_x_d, neg_x_d;
    uint32_t sky0, sky1;  // skyVisibility_pznzpyny, skyVisibility_pxnxs0s1
};
struct VoxelO {           // VoxelData, BasicTypes.slang:145-151
    uint32_t mat_id, age, dist0, dist1;  // dist0: distanceToSolid_gg_z01, dist1: _x01_y01
};
struct ProbeRef {
    ProbeGridO grid{};
    const CubeO* cubes = nullptr;  // nullptr: no grid
    const VoxelO* vox = nullptr;
};

struct TextureO {          // an RGBA8 texture (gk_texture)
    uint32_t w = 0, h = 0, srgb = 0;
    std::vector<uint32_t> texels;
};

struct Scene {
    ProbeRef probes;        // REF_V2 termination lookup (arrays owned by the caller)
    std::vector<TextureO> tex;
    // sky image: equirectangular RGBA32F, row 0 at +y (empty: constant sky), and its
    // ProjectHdrToSh coefficients [channel][9]
    std::vector<float> sky;
    uint32_t sky_w = 0, sky_h = 0;
    float sky_sh[27] = {};
    float tex_lut[512] = {};  // sRGB byte -> linear, then byte / 255
    std::vector<Mesh> meshes;
    std::vector<Instance> inst;
    std::vector<Material> mats;
    // TLAS BVH2 over instances
    std::vector<Mesh::Node> tnodes;
    std::vector<uint32_t> tprim;
    std::vector<Light> lights;
    std::vector<double> light_cdf;
    double light_total = 0.0;
};

struct Hit {
    float t, u, v;
    uint32_t prim, inst;
};

template <class T>
struct HitT {
    T t, u, v;
    uint32_t prim, inst;
};

// Build (or rebuild) all acceleration structures and the light list.
void scene_build(Scene& sc);
void scene_update_instances(Scene& sc);
// Closest hit with ties broken towards the lower (inst, prim). tmin < t < tmax.
bool trace_closest(const Scene& sc, F3 O, F3 D, float tmin, float tmax, Hit& h);
bool trace_any(const Scene& sc, F3 O, F3 D, float tmin, float tmax);
template <class T>
bool trace_closest_t(const Scene& sc, V3<T> O, V3<T> D, T tmin, T tmax, HitT<T>& h);
template <class T>
bool trace_any_t(const Scene& sc, V3<T> O, V3<T> D, T tmin, T tmax);

// material index of (inst, prim): FetchMaterialId semantics (Const_Func.slang:370-424)
// plus the offset form, clamped like GetMaterial (BasicTypes.slang:376-379).
inline uint32_t material_of(const Scene& sc, uint32_t inst, uint32_t prim) {
    const Instance& I = sc.inst[inst];
    const uint32_t local = sc.meshes[I.mesh_id].mat[prim];
    uint32_t m = I.use_remap ? I.mat_id[local & 15u] : I.mat_offset + local;
    const uint32_t maxm = uint32_t(sc.mats.size()) - 1u;
    return m > maxm ? maxm : m;
}

// world-space transform helpers (column-major world matrix)
inline F3 xform_point(const float* m, F3 p) {
    float v[4] = {p.x, p.y, p.z, 1.f}, o[4];
    mat_mul4(m, v, o);
    return {o[0], o[1], o[2]};
}
inline F3 xform_vector(const float* m, F3 p) {
    float v[4] = {p.x, p.y, p.z, 0.f}, o[4];
    mat_mul4(m, v, o);
    return {o[0], o[1], o[2]};
}

}  // namespace orc

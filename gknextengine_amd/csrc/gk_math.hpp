// gk_math.hpp - device math shared by the HIP kernels of the path-tracing core.
//
// All of the library is compiled with -ffp-contract=off: a multiply-add is fused only
// where fmaf() is written out. The ray/triangle test, the camera ray and the samplers
// therefore produce the same bits on the GPU as the IEEE-754 single-precision
// arithmetic they are specified with.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gk {

struct f3 {
    float x, y, z;
};
__host__ __device__ inline f3 mk(float x, float y, float z) { return f3{x, y, z}; }
__host__ __device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
__host__ __device__ inline f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__host__ __device__ inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__host__ __device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline f3 cross(f3 a, f3 b) {
    return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__host__ __device__ inline float length(f3 a) { return sqrtf(dot(a, a)); }
// v * (1/sqrt(dot(v,v))), IEEE sqrt and divide (not rsqrt)
__host__ __device__ inline f3 normalize(f3 a) {
    const float l = sqrtf(dot(a, a));
    const float rl = l == 0.f ? 0.f : 1.0f / l;
    return a * rl;
}
__host__ __device__ inline f3 fmin3(f3 a, f3 b) {
    return f3{a.x < b.x ? a.x : b.x, a.y < b.y ? a.y : b.y, a.z < b.z ? a.z : b.z};
}
__host__ __device__ inline f3 fmax3(f3 a, f3 b) {
    return f3{a.x > b.x ? a.x : b.x, a.y > b.y ? a.y : b.y, a.z > b.z ? a.z : b.z};
}
__device__ inline f3 ld3(const float4& v) { return f3{v.x, v.y, v.z}; }

// column-major 4x4 times (x,y,z,w): ((m0*x + m1*y) + m2*z) + m3*w
__host__ __device__ inline void mul44v(const float* m, float x, float y, float z, float w, float* o) {
    for (int r = 0; r < 4; ++r) {
        float a = m[r] * x;
        a = a + m[4 + r] * y;
        a = a + m[8 + r] * z;
        a = a + m[12 + r] * w;
        o[r] = a;
    }
}

// ---------------------------------------------------------------- RNG
// pcg4d hash (Jarzynski & Olano 2020), the per-pixel generator of Const_Func.slang:227-258.
struct rng4 {
    uint32_t x, y, z, w;
};
__device__ inline void pcg4d(rng4& v) {
    v.x = v.x * 1664525u + 1013904223u;
    v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u;
    v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
    v.x ^= v.x >> 16u; v.y ^= v.y >> 16u; v.z ^= v.z >> 16u; v.w ^= v.w >> 16u;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
}
__device__ inline float u2f(uint32_t x) { return __uint_as_float(0x3f800000u | (x >> 9)) - 1.0f; }
__device__ inline float rnd1(rng4& s) { pcg4d(s); return u2f(s.x); }
__device__ inline void rnd2(rng4& s, float& a, float& b) { pcg4d(s); a = u2f(s.x); b = u2f(s.y); }

// sin/cos of 2*pi*u for u in [0,1): quarter-turn reduction and the Taylor series of
// sin (degree 9) and cos (degree 10) on |x| <= pi/4, evaluated by Horner's rule.
template <class T>
__device__ inline void sincos_2pi(T u, T& s, T& c) {
    const T q4 = u * T(4);
    int q = int(q4);
    if (q > 3) q = 3;
    const T r = q4 - T(q);
    const T x = (r - T(0.5)) * T(1.5707963267948966);
    const T x2 = x * x;
    T sp = T(1.0 / 362880.0);
    sp = sp * x2 - T(1.0 / 5040.0);
    sp = sp * x2 + T(1.0 / 120.0);
    sp = sp * x2 - T(1.0 / 6.0);
    sp = sp * x2 + T(1);
    const T sx = sp * x;
    T cp = T(-1.0 / 3628800.0);
    cp = cp * x2 + T(1.0 / 40320.0);
    cp = cp * x2 - T(1.0 / 720.0);
    cp = cp * x2 + T(1.0 / 24.0);
    cp = cp * x2 - T(0.5);
    cp = cp * x2 + T(1);
    const T k = T(0.70710678118654752);
    const T sq = (sx + cp) * k;
    const T cq = (cp - sx) * k;
    switch (q) {
        case 0: s = sq; c = cq; break;
        case 1: s = cq; c = -sq; break;
        case 2: s = -sq; c = -cq; break;
        default: s = -cq; c = sq; break;
    }
}

// atan2(y, x) / (2 pi), in (-1/2, 1/2]; 0 for (0, 0). The ratio t = min/max of |x|, |y|
// is reduced with atan t = pi/4 + atan((t - 1)/(t + 1)) when t > tan(pi/8), leaving
// |s| <= tan(pi/8), where the Maclaurin series of atan (terms to s^19, truncation below
// 1e-8) is summed by Horner's rule in s^2; the octant and quadrant follow from
// atan(1/t) = pi/2 - atan t and atan2's symmetries.
__device__ inline float atan2_turns(float y, float x) {
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = ax > ay ? ax : ay, mn = ax > ay ? ay : ax;
    if (mx == 0.f) return 0.f;
    float t = mn / mx;
    float base = 0.f;
    if (t > 0.41421356f) { t = (t - 1.f) / (t + 1.f); base = 0.125f; }
    const float t2 = t * t;
    float p = 1.f / 19.f;
    p = p * t2 - 1.f / 17.f;
    p = p * t2 + 1.f / 15.f;
    p = p * t2 - 1.f / 13.f;
    p = p * t2 + 1.f / 11.f;
    p = p * t2 - 1.f / 9.f;
    p = p * t2 + 1.f / 7.f;
    p = p * t2 - 1.f / 5.f;
    p = p * t2 + 1.f / 3.f;
    p = p * t2 - 1.f;
    float a = base - (p * t) * 0.15915494f;   // -p t = atan t; 0.15915494 = 1 / (2 pi)
    if (ay > ax) a = 0.25f - a;
    if (x < 0.f) a = 0.5f - a;
    return y < 0.f ? -a : a;
}

// concentric_sample_disk (Const_Func.slang:260-275) of RandomInUnitDisk (:277-280): offset
// 2u - 1; zero maps to zero (isZERO, |x| < 1e-35); the larger coordinate is the radius and
// the angle pi/4 * (smaller / larger), whose sin/cos come from sincos_2pi of |ratio| / 8
// (exact) with the sign restored; the second branch keeps the reference's form
// (sin of the angle, then sqrt(1 - s^2)).
template <class T>
__device__ inline void concentric_disk(T ua, T ub, T& dx, T& dy) {
    const T ox = ua + (ua - T(1)), oy = ub + (ub - T(1));
    const T nz = T(1e-35);
    if (ox > -nz && ox < nz && oy > -nz && oy < nz) { dx = T(0); dy = T(0); return; }
    const bool xmaj = fabs(ox) > fabs(oy);
    const T r = xmaj ? oy / ox : ox / oy;
    T s, c;
    sincos_2pi(fabs(r) * T(0.125), s, c);
    if (r < T(0)) s = -s;
    if (xmaj) { dx = ox * c; dy = ox * s; }
    else { dx = oy * s; dy = oy * sqrt(T(1) - s * s); }
}

// ---------------------------------------------------------------- acceleration structures
// 4-wide BVH node, one 128-byte cache line. Child i is
//   internal: meta[i] == 0, child[i] = node index
//   leaf:     meta[i] = primitive count (> 0); child[i] = the first primitive record in a
//             TLAS node, and in a BLAS node the traversal stack entry of the leaf,
//             kBlasLeaf | count << 27 | first triangle record (blas_leaf_entry)
//   empty:    a point box far outside the scene (never hit), child and meta 0
constexpr uint32_t kBlasLeaf = 0x40000000u;
constexpr uint32_t kMaxLeafTris = 4;  // BLAS leaf size cap of the builder (<= 7: the count has 3 bits)
constexpr uint32_t kLeafFirstMask = 0x07ffffffu;
__host__ __device__ inline uint32_t blas_leaf_entry(uint32_t first, uint32_t count) {
    return kBlasLeaf | (count << 27) | first;
}
struct alignas(128) Bvh4Node {
    float4 lox, hix, loy, hiy, loz, hiz;
    uint4 child;
    uint4 meta;
};
static_assert(sizeof(Bvh4Node) == 128, "node size");

// Triangle record, 48 bytes: v0 (w = primitive index within its mesh), e1 = v1 - v0,
// e2 = v2 - v0 (w of e1 = local material slot).
struct TriRec {
    float4 v0;
    float4 e1;
    float4 e2;
};

constexpr uint32_t kStackMark = 0xffffffffu;  // return from an instance to the TLAS
constexpr uint32_t kInstFlag = 0x80000000u;   // stack entry: enter instance (low bits)
constexpr float kFar = 1e30f;

__host__ __device__ inline float safercp(float x) {
    return x > 1e-12f ? (1.0f / x) : (x < -1e-12f ? (1.0f / x) : kFar);
}

}  // namespace gk

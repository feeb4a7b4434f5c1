// trace.hip - software ray traversal for gfx950 (no RT cores on CDNA4).
//
// Replaces the hardware RayQuery of FHardwareRayTracerV2::TraceRay / TraceOcclusion
// (Shading.slang:735-827) and the CPU ray cast BVH::Intersect / IntersectTLAS /
// IsOccluded (tiny_bvh.h:2245-2397).
//
// Design
//   * 4-wide BVH, one 128-byte node per visit (gk_math.hpp Bvh4Node), 48-byte triangle
//     records stored in leaf order.
//   * Persistent waves with dynamic ray fetch: lanes whose ray has finished are refilled
//     together -- one ballot, one atomicAdd per wave, ranks from mbcnt -- once kRefill of
//     the 64 lanes are idle (wave-level compaction).
//   * Traversal stack: a 16-entry ring per lane in LDS; when a push would overwrite the
//     oldest entry that entry is spilled to a per-thread global array (coalesced
//     [entry][thread] layout) and read back when the ring runs empty.
//   * One step tests either one node (4 boxes) or one leaf (<= 4 triangles): BLAS leaf
//     children become stack entries (0x40000000 | count << 27 | first triangle) ordered
//     near to far with the internal children, which bounds the divergence of a step and
//     lets near hits cull far subtrees (measured on the 1080p city: trace 4.46 -> 3.94 ms).
//     In single-level scenes a lane postpones the leaf it reaches and goes on with nodes;
//     the wave tests the postponed leaves together (kLeafBatch).
//   * Launches of at most kQuadRays rays per quad of the grid (a frame's last bounces;
//     kQuadRaysTwo in two-level scenes) run trace_quads instead: four lanes per ray, one
//     child box or triangle each.
//   * Two-level scenes: a TLAS leaf holds one instance, whose entry is ordered near to far
//     with its siblings; entering an instance pushes a return mark and continues in the
//     instance's BLAS with the ray transformed by the same inverse-matrix arithmetic
//     tinybvh uses (tiny_bvh.h:396-409,2309-2316).
//   * Triangle test: Moeller-Trumbore with exactly tinybvh's operations and rejects
//     (tiny_bvh.h:6815-6844). Closest hit keeps the smallest t, ties broken towards the
//     lower (instance, primitive), so results do not depend on traversal order. Box
//     tests are widened by 2^-16 relative so they do not cull an accepted triangle (a
//     triangle test accepts points a few tens of ulps outside the triangle, near its
//     edges; 2^-20 missed one such shared-edge hit, and the result depended on ray order).
#include <hip/hip_runtime.h>

#include <mutex>

#include "gk_ctx.hpp"

namespace gk {

namespace {

constexpr int kBlock = 256;
constexpr int kRing = 16;            // LDS stack entries per lane
constexpr int kSpill = 64;           // global spill entries per thread
// Resident work-groups per CU of a trace launch (the spill area is sized for it): at most the
// kernel's occupancy, capped at 4, i.e. 4 waves per SIMD. Fewer waves share each CU's caches;
// measured on the 1080p city (trace ms per frame): 6 -> 3.13, 5 -> 3.03, 4 -> 3.00,
// 3 -> 3.09, 2 -> 3.57.
constexpr int kMaxBlocksPerCU = 4;
// The same while a frame's two path pipelines trace at once: a CU holds 5 of these
// work-groups (LDS and registers), so grids of 3 per CU let both pipelines' launches be
// resident together. 1080p city, ms per frame with two pipelines: 4 -> 3.18, 3 -> 2.96-2.99,
// 2 -> 3.17, 3 and 2 -> 3.12-3.19; 4096 bricks: 4 -> 4.37, 3 -> 4.10, 2 -> 4.28.
constexpr int kSplitBlocksPerCU = 3;
constexpr uint32_t kRefill = 56;     // idle lanes that trigger a wave refill (<= 64)
// postponed leaves that start a leaf phase (single level; measured on the 1080p city, trace
// ms per frame: testing leaves as met 2.40, batches of 16 or 32 -> 2.36, 48 -> 2.37, 64 -> 2.41)
constexpr uint32_t kLeafBatch = 32;

struct Lane {
    f3 O, D, rD;
    f3 wO, wD, wrD;                  // world-space ray (two-level)
    float tmin, tmax;
    float bt, bu, bv;
    uint32_t bprim, binst;
    uint32_t node;                   // next node to visit, or kNone
    uint32_t leaf;                   // single-level: a postponed leaf entry, or kNone
    uint32_t cur_inst;
    uint32_t level;                  // 0 = TLAS, 1 = BLAS
    uint32_t head, count, gsp;       // ring head, entries in ring, spilled entries
    uint32_t ray;
    uint32_t sx, sy, sz;             // byte offset (0 or 16) of the near plane of each axis
                                     // within a node's lo/hi pair, for the current rD
    bool active, found, overflow;
    bool any;                        // occlusion ray: the first hit ends it
    bool job1;                       // ray of the second job (two-job launches)
};

// One queue of rays for a trace launch. count_dev (device, may be null) caps max_count.
struct TraceJob {
    const float4* o;                 // origin, tmin (w)
    const float4* d;                 // direction, tmax (w)
    const uint32_t* count_dev;
    uint32_t max_count;
    float4* hit;                     // closest: (t, u, v, prim); any: w = 1 when occluded
    uint32_t* inst;                  // closest: instance of the hit (may be null)
    uint32_t* clear;                 // a counter other than count_dev the launch sets to 0 (may be
                                     // null): the queue the shading after the trace appends to
};

// MODE 0: one closest-hit job; 1: one any-hit job; 2: a closest-hit job then an any-hit
// job drained by the same persistent waves (their latency tails overlap).
template <int MODE>
__device__ inline bool lane_any(const bool any_flag) { return MODE == 1 ? true : (MODE == 0 ? false : any_flag); }

constexpr uint32_t kNone = 0xfffffffeu;

__device__ inline void push(Lane& L, uint32_t (*ring)[kBlock], uint32_t* spill, uint32_t gtid,
                            uint32_t stride, uint32_t v) {
    const uint32_t lane = threadIdx.x;
    if (L.count == kRing) {
        // ring full: the slot at head holds the oldest entry; move it to global memory
        if (L.gsp < kSpill) spill[size_t(L.gsp) * stride + gtid] = ring[L.head][lane];
        else L.overflow = true;
        L.gsp++;
        ring[L.head][lane] = v;
        L.head = (L.head + 1) & (kRing - 1);
    } else {
        ring[L.head][lane] = v;
        L.head = (L.head + 1) & (kRing - 1);
        L.count++;
    }
}

__device__ inline bool pop(Lane& L, uint32_t (*ring)[kBlock], const uint32_t* spill, uint32_t gtid,
                           uint32_t stride, uint32_t& v) {
    if (L.count > 0) {
        L.head = (L.head - 1) & (kRing - 1);
        L.count--;
        v = ring[L.head][threadIdx.x];
        return true;
    }
    if (L.gsp > 0) {
        L.gsp--;
        v = L.gsp < kSpill ? spill[size_t(L.gsp) * stride + gtid] : 0u;
        return true;
    }
    return false;
}

// tinybvh_transform_point / _vector with a row-major 4x4 (tiny_bvh.h:396-409)
__device__ inline f3 tpoint(f3 v, const float* M) {
    const f3 r = f3{M[0] * v.x + M[1] * v.y + M[2] * v.z + M[3],
                    M[4] * v.x + M[5] * v.y + M[6] * v.z + M[7],
                    M[8] * v.x + M[9] * v.y + M[10] * v.z + M[11]};
    const float w = M[12] * v.x + M[13] * v.y + M[14] * v.z + M[15];
    if (w == 1.0f) return r;
    return r * (1.f / w);
}
__device__ inline f3 tvector(f3 v, const float* M) {
    return f3{M[0] * v.x + M[1] * v.y + M[2] * v.z, M[4] * v.x + M[5] * v.y + M[6] * v.z,
              M[8] * v.x + M[9] * v.y + M[10] * v.z};
}

// widen an interval end by 2^-16 relative: x +- |x| * 2^-16 is exactly x * (1 +- 2^-16), and
// the fma rounds it once, so this equals the rounded product; one instruction (|x| is a
// source modifier). An end at +-infinity becomes NaN and the box is culled: a box whose
// entry or exit lies at infinity holds no hit at a finite distance.
__device__ inline float widen_hi(float x) { return __builtin_fmaf(__builtin_fabsf(x), 0x1p-16f, x); }
__device__ inline float widen_lo(float x) { return __builtin_fmaf(-__builtin_fabsf(x), 0x1p-16f, x); }

// slab test of one child box; returns its entry distance, or +inf when missed
__device__ inline float slab1(float lx, float hx, float ly, float hy, float lz, float hz, f3 O, f3 rD,
                              float tmin, float limit) {
    // tinybvh's (plane - O) * rD, not one fma with a precomputed O * rD: for a ray whose
    // origin lies on a box face and whose direction is parallel to it, the difference is
    // exactly 0, so the box is culled as the reference culls it (its voxelize misses such
    // grazing hits); the fma form is off by the rounding of O * rD there and enters the box.
    float a = (lx - O.x) * rD.x, b = (hx - O.x) * rD.x;
    float lo = fminf(a, b), hi = fmaxf(a, b);
    a = (ly - O.y) * rD.y; b = (hy - O.y) * rD.y;
    lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
    a = (lz - O.z) * rD.z; b = (hz - O.z) * rD.z;
    lo = fmaxf(lo, fminf(a, b)); hi = fminf(hi, fmaxf(a, b));
    lo = widen_lo(lo);
    hi = widen_hi(hi);
    return (hi >= lo && hi >= tmin && lo <= limit) ? lo : __builtin_inff();
}

// slab test of the 4 children; returns entry distances (or +inf when missed)
__device__ inline void slab4(const Bvh4Node& n, const Lane& L, float limit, float tl[4]) {
    const float lx[4] = {n.lox.x, n.lox.y, n.lox.z, n.lox.w};
    const float hx[4] = {n.hix.x, n.hix.y, n.hix.z, n.hix.w};
    const float ly[4] = {n.loy.x, n.loy.y, n.loy.z, n.loy.w};
    const float hy[4] = {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w};
    const float lz[4] = {n.loz.x, n.loz.y, n.loz.z, n.loz.w};
    const float hz[4] = {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) tl[i] = slab1(lx[i], hx[i], ly[i], hy[i], lz[i], hz[i], L.O, L.rD, L.tmin, limit);
}

// near_far_offsets: for a direction whose reciprocal component is negative the near plane
// of that axis is the box's hi plane (offset 16 in the node's lo/hi float4 pair).
__device__ inline void near_far_offsets(Lane& L) {
    L.sx = (__float_as_uint(L.rD.x) >> 27) & 16u;
    L.sy = (__float_as_uint(L.rD.y) >> 27) & 16u;
    L.sz = (__float_as_uint(L.rD.z) >> 27) & 16u;
}

// slab4 with the near and far planes fetched by address: per axis, (near - O) * rD and
// (far - O) * rD are the min and the max of the two plane distances, bit for bit (the
// rounded difference and the product by a nonzero reciprocal are monotonic, and no term
// is NaN), so the entry is max3 of the near terms and the exit min3 of the far terms and
// the per-axis min/max pairs of slab1 go away. nb: the node array; node: its index.
__device__ inline void slab4_ordered(const char* __restrict__ nb, uint32_t node, const Lane& L, float limit,
                                     float tl[4], uint4& child, uint4& meta) {
    const uint32_t base = node * uint32_t(sizeof(Bvh4Node));
    const uint32_t ax = base | L.sx, ay = base | L.sy, az = base | L.sz;
    auto ld = [&](uint32_t off) { return *reinterpret_cast<const float4*>(nb + off); };
    const float4 nx = ld(ax), fx = ld(ax ^ 16u);
    const float4 ny = ld(ay + 32u), fy = ld((ay ^ 16u) + 32u);
    const float4 nz = ld(az + 64u), fz = ld((az ^ 16u) + 64u);
    child = *reinterpret_cast<const uint4*>(nb + base + 96u);
    meta = *reinterpret_cast<const uint4*>(nb + base + 112u);
    const float nxs[4] = {nx.x, nx.y, nx.z, nx.w}, fxs[4] = {fx.x, fx.y, fx.z, fx.w};
    const float nys[4] = {ny.x, ny.y, ny.z, ny.w}, fys[4] = {fy.x, fy.y, fy.z, fy.w};
    const float nzs[4] = {nz.x, nz.y, nz.z, nz.w}, fzs[4] = {fz.x, fz.y, fz.z, fz.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float lo = widen_lo(fmaxf(fmaxf((nxs[i] - L.O.x) * L.rD.x, (nys[i] - L.O.y) * L.rD.y),
                                        (nzs[i] - L.O.z) * L.rD.z));
        const float hi = widen_hi(fminf(fminf((fxs[i] - L.O.x) * L.rD.x, (fys[i] - L.O.y) * L.rD.y),
                                        (fzs[i] - L.O.z) * L.rD.z));
        tl[i] = (hi >= lo && hi >= L.tmin && lo <= limit) ? lo : __builtin_inff();
    }
}

// compare-exchange of two (distance, entry) pairs, nearer first
__device__ inline void cswap(float (&t)[4], uint32_t (&c)[4], int a, int b) {
    const bool s = t[b] < t[a];
    const float ta = t[a], tb = t[b];
    const uint32_t ca = c[a], cb = c[b];
    t[a] = s ? tb : ta;
    t[b] = s ? ta : tb;
    c[a] = s ? cb : ca;
    c[b] = s ? ca : cb;
}

// Moeller-Trumbore, tinybvh IntersectTri operation order (tiny_bvh.h:6818-6833)
__device__ inline bool mt(const TriRec& tr, f3 O, f3 D, float& t, float& u, float& v) {
    const f3 v0 = ld3(tr.v0), e1 = ld3(tr.e1), e2 = ld3(tr.e2);
    const f3 h = cross(D, e2);
    const float a = dot(e1, h);
    if (fabsf(a) < 0.0000001f) return false;
    const float f = 1 / a;
    const f3 s = O - v0;
    u = f * dot(s, h);
    if (u < 0 || u > 1) return false;
    const f3 q = cross(s, e1);
    v = f * dot(D, q);
    if (v < 0 || u + v > 1) return false;
    t = f * dot(e2, q);
    return true;
}

// ---- quad traversal: four lanes per ray, for launches with few rays. The last bounces of
// a frame trace a few thousand rays, so each lane of the per-lane traversal holds one ray
// and a step runs the node and the leaf code in turn for a few waves: the launch lasts as
// long as its slowest ray's chain of dependent steps. Here lane q of a quad tests child q
// of a node, or triangle q of a leaf, and the quad shares one stack in LDS, which shortens
// each step about fourfold. The per-child box test and the triangle test are the same
// functions as in the per-lane traversal, and the closest hit is the smallest
// (t, primitive), so the results are the same. Every branch below is uniform across a
// quad, so all four lanes of a quad are active at each quad exchange.
// launches of at most this many rays per quad of the grid (measured on the 1080p city, trace
// ms per frame, quad traversal from: 1 ray per quad 2.43, 8 -> 2.40, 16 -> 2.56, 64 -> 3.29;
// the one-lane traversal alone 2.79; again once the one-lane traversal postponed leaves:
// 4 -> 2.38, 6 -> 2.35, 8 -> 2.36)
constexpr uint32_t kQuadRays = 8;
// the same for a two-level scene, whose quads also enter and leave instances (4096 bricks at
// 1080p, ms per frame: one-lane traversal alone 5.54, 1 ray per quad 5.48, 2 -> 5.47,
// 3 -> 5.47, 4 -> 5.65, 8 -> 5.61)
constexpr uint32_t kQuadRaysTwo = 2;
constexpr int kQuads = kBlock / 4;                   // quads per work-group
constexpr int kQStack = 128;                         // stack entries per quad (deeper than a lane's
                                                     // 16 in LDS + kSpill in memory)
static_assert(kQStack >= kRing + kSpill, "a quad holds at least a lane's stack depth");
// LDS words of a work-group: the lanes' rings or the quads' stacks
constexpr int kLdsWords = kRing * kBlock > kQStack * kQuads ? kRing * kBlock : kQStack * kQuads;

// the value of lane j of this lane's quad (DPP quad permute, a VALU move)
template <int J>
__device__ inline uint32_t quad_get(uint32_t v) {
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), J * 0x55, 0xf, 0xf, false));
}
template <int J>
__device__ inline float quad_getf(float v) { return __uint_as_float(quad_get<J>(__float_as_uint(v))); }
// the value of lane q ^ M of this lane's quad
template <int M>
__device__ inline uint32_t quad_xor(uint32_t v) {
    constexpr int ctrl = (0 ^ M) | ((1 ^ M) << 2) | ((2 ^ M) << 4) | ((3 ^ M) << 6);
    return uint32_t(__builtin_amdgcn_mov_dpp(int(v), ctrl, 0xf, 0xf, false));
}

template <int M>
__device__ inline void quad_min_hit(float& t, float& u, float& v, uint32_t& prim) {
    const float ot = __uint_as_float(quad_xor<M>(__float_as_uint(t)));
    const float ou = __uint_as_float(quad_xor<M>(__float_as_uint(u)));
    const float ov = __uint_as_float(quad_xor<M>(__float_as_uint(v)));
    const uint32_t op = quad_xor<M>(prim);
    if (ot < t || (ot == t && op < prim)) { t = ot; u = ou; v = ov; prim = op; }
}

template <int MODE, bool TWO>
__device__ void trace_quads(const Bvh4Node* __restrict__ blas, const TriRec* __restrict__ tris,
                            const Bvh4Node* __restrict__ tlas, const uint32_t* __restrict__ tlas_prims,
                            const InstDev* __restrict__ inst, uint32_t blas_root,
                            const TraceJob& j0, const TraceJob& j1, uint32_t c0, uint32_t count,
                            uint32_t (*stack)[kQuads], uint32_t* __restrict__ overflow_flag) {
    const uint32_t q = __lane_id() & 3u;
    const uint32_t qi = threadIdx.x >> 2;
    const uint32_t nquads = gridDim.x * kQuads;
    for (uint32_t r = blockIdx.x * kQuads + qi; r < count; r += nquads) {
        const bool second = MODE == 2 && r >= c0;
        const uint32_t ri = second ? r - c0 : r;
        const bool any = MODE == 1 || second;
        const float4 o = second ? j1.o[ri] : j0.o[ri];
        const float4 d = second ? j1.d[ri] : j0.d[ri];
        const f3 wO = ld3(o), wD = ld3(d);
        const f3 wrD = f3{safercp(wD.x), safercp(wD.y), safercp(wD.z)};
        f3 O = wO, D = wD, rD = wrD;
        const float tmin = o.w, tmax = d.w;
        float bt = tmax, bu = 0.f, bv = 0.f;
        uint32_t bprim = 0xffffffffu, binst = 0xffffffffu;
        uint32_t cur_inst = 0u, level = TWO ? 0u : 1u;
        bool found = false, overflow = false;
        uint32_t node = TWO ? 0u : blas_root;
        int sp = 0;
        // enter instance ii: the ray goes to its object space, and a mark beneath its BLAS
        // entries restores the world ray when they are done. False when the stack is full:
        // the ray then ends (a dropped mark would leave TLAS entries read as BLAS nodes).
        auto enter = [&](uint32_t ii) {
            if (sp >= kQStack) return false;
            stack[sp++][qi] = kStackMark;
            const InstDev& I = inst[ii];
            O = tpoint(wO, I.inv);
            D = tvector(wD, I.inv);
            rD = f3{safercp(D.x), safercp(D.y), safercp(D.z)};
            cur_inst = ii;
            level = 1u;
            node = I.root;
            return true;
        };
        for (;;) {
            if (node == kNone) {
                if (sp == 0) break;
                node = stack[--sp][qi];
                if (TWO) {
                    if (node == kStackMark) {  // leave the instance
                        O = wO; D = wD; rD = wrD;
                        level = 0u;
                        node = kNone;
                        continue;
                    }
                    if (node & kInstFlag) {
                        if (!enter(node & ~kInstFlag)) { overflow = true; break; }
                    }
                }
            }
            if (level != 0u && (node & 0xc0000000u) == kBlasLeaf) {
                const uint32_t first = node & kLeafFirstMask, cnt = (node >> 27) & 7u;
                float ht = __builtin_inff(), hu = 0.f, hv = 0.f;
                uint32_t hp = 0xffffffffu;
                for (uint32_t k = q; k < cnt; k += 4u) {
                    const TriRec tr = tris[first + k];
                    float t, u, v;
                    if (!mt(tr, O, D, t, u, v)) continue;
                    if (!(t > tmin && t < tmax)) continue;
                    const uint32_t prim = __float_as_uint(tr.v0.w);
                    if (t < ht || (t == ht && prim < hp)) { ht = t; hu = u; hv = v; hp = prim; }
                }
                quad_min_hit<1>(ht, hu, hv, hp);
                quad_min_hit<2>(ht, hu, hv, hp);
                node = kNone;
                if (hp != 0xffffffffu) {
                    if (any) { found = true; break; }
                    // the leaf's hits share one instance: its smallest (t, primitive) against
                    // the best so far in (t, instance, primitive) order
                    if (!found || ht < bt ||
                        (ht == bt && (cur_inst < binst || (cur_inst == binst && hp < bprim)))) {
                        bt = ht; bu = hu; bv = hv; bprim = hp; binst = cur_inst; found = true;
                    }
                }
                continue;
            }
            const bool tl_level = TWO && level == 0u;
            const Bvh4Node& N = tl_level ? tlas[node] : blas[node];
            const float* f = reinterpret_cast<const float*>(&N);
            float tq = slab1(f[q], f[4 + q], f[8 + q], f[12 + q], f[16 + q], f[20 + q], O, rD, tmin, bt);
            uint32_t cq = reinterpret_cast<const uint32_t*>(&N.child)[q];
            if (TWO && tl_level) {
                // a TLAS leaf child: one instance joins the near-to-far order as an entry;
                // the instances of a larger leaf are pushed beneath the other children
                const uint32_t mq = reinterpret_cast<const uint32_t*>(&N.meta)[q];
                const bool hitq = tq < __builtin_inff();
                const uint32_t many = (hitq && mq > 1u) ? mq : 0u;
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const uint32_t m = j == 0 ? quad_get<0>(many) : j == 1 ? quad_get<1>(many)
                                     : j == 2 ? quad_get<2>(many) : quad_get<3>(many);
                    const uint32_t c = j == 0 ? quad_get<0>(cq) : j == 1 ? quad_get<1>(cq)
                                     : j == 2 ? quad_get<2>(cq) : quad_get<3>(cq);
                    for (uint32_t k = 0; k < m; ++k) {
                        if (sp < kQStack) stack[sp++][qi] = kInstFlag | tlas_prims[c + k];
                        else overflow = true;
                    }
                }
                if (many) tq = __builtin_inff();
                else if (hitq && mq == 1u) cq = kInstFlag | tlas_prims[cq];
            }
            float kt[4] = {quad_getf<0>(tq), quad_getf<1>(tq), quad_getf<2>(tq), quad_getf<3>(tq)};
            uint32_t kc[4] = {quad_get<0>(cq), quad_get<1>(cq), quad_get<2>(cq), quad_get<3>(cq)};
            cswap(kt, kc, 0, 1);
            cswap(kt, kc, 2, 3);
            cswap(kt, kc, 0, 2);
            cswap(kt, kc, 1, 3);
            cswap(kt, kc, 1, 2);
#pragma unroll
            for (int k = 3; k >= 1; --k) {
                if (kt[k] < __builtin_inff()) {
                    if (sp < kQStack) stack[sp++][qi] = kc[k];
                    else overflow = true;
                }
            }
            node = kt[0] < __builtin_inff() ? kc[0] : kNone;
            if (TWO && node != kNone && (node & kInstFlag)) {
                if (!enter(node & ~kInstFlag)) { overflow = true; break; }
            }
        }
        if (q == 0u) {
            if (any) {
                float4* out = second ? j1.hit : j0.hit;
                out[ri] = make_float4(0.f, 0.f, 0.f, __uint_as_float(found ? 1u : 0u));
            } else {
                j0.hit[ri] = found ? make_float4(bt, bu, bv, __uint_as_float(bprim))
                                   : make_float4(__builtin_inff(), 0.f, 0.f, __uint_as_float(0xffffffffu));
                if (j0.inst) j0.inst[ri] = found ? binst : 0xffffffffu;
            }
            if (overflow) atomicOr(overflow_flag, 1u);
        }
    }
}

template <int MODE, bool TWO>
__global__ __launch_bounds__(kBlock) void k_trace(
    const Bvh4Node* __restrict__ blas, const TriRec* __restrict__ tris,
    const Bvh4Node* __restrict__ tlas, const uint32_t* __restrict__ tlas_prims,
    const InstDev* __restrict__ inst, uint32_t blas_root, TraceJob j0, TraceJob j1,
    uint32_t* __restrict__ next_ray, uint32_t* __restrict__ spill,
    uint32_t* __restrict__ overflow_flag, unsigned long long* __restrict__ ray_stats) {
    __shared__ uint32_t lds[kLdsWords];
    auto ring = reinterpret_cast<uint32_t (*)[kBlock]>(lds);
    const uint32_t stride = gridDim.x * kBlock;
    const uint32_t gtid = blockIdx.x * kBlock + threadIdx.x;
    uint32_t c0 = j0.max_count, c1 = 0;
    if (j0.count_dev) { const uint32_t c = *j0.count_dev; c0 = c < c0 ? c : c0; }
    if (MODE == 2) {
        c1 = j1.max_count;
        if (j1.count_dev) { const uint32_t c = *j1.count_dev; c1 = c < c1 ? c : c1; }
    }
    const uint32_t count = c0 + c1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (j0.clear) *j0.clear = 0u;
        atomicAdd(&ray_stats[MODE == 1 ? 1 : 0], (unsigned long long)c0);
        if (MODE == 2) atomicAdd(&ray_stats[1], (unsigned long long)c1);
    }
    if (count <= (TWO ? kQuadRaysTwo : kQuadRays) * gridDim.x * kQuads) {
        trace_quads<MODE, TWO>(blas, tris, tlas, tlas_prims, inst, blas_root, j0, j1, c0, count,
                               reinterpret_cast<uint32_t (*)[kQuads]>(lds), overflow_flag);
        return;
    }
    // rays a wave holds at once: its share of the launch. The launch holds more rays than the
    // quad traversal takes, so the share is at least 1 (a wave learns that the queue is
    // drained only from a fetch, so a share of 0 would never end).
    static_assert(kQuadRays >= 1 && kQuadRaysTwo >= 1, "an empty launch takes the quad traversal");
    const uint32_t waves = gridDim.x * (kBlock / 64u);
    const uint32_t quota = (count + waves - 1u) / waves;

    Lane L;
    L.active = false;
    L.overflow = false;
    L.any = MODE == 1;
    L.job1 = false;
    bool exhausted = false;
    for (;;) {
        // ---- refill idle lanes: one atomic per wave, ranks from the ballot. A refill
        // stalls the whole wave on the ray loads, so it waits until kRefill lanes are idle
        // (measured on the 1080p city, trace ms per frame at 4 waves per SIMD: 32 -> 3.26,
        // 40 -> 3.09, 48 -> 3.00, 52 -> 2.96, 56 -> 2.95, 60 -> 2.95, 64 -> 3.00).
        if (!exhausted) {
            const uint64_t need = __ballot(!L.active);
            const uint32_t n_need = __popcll(need);
            const uint32_t n_act = 64u - n_need;
            if (n_need >= kRefill && n_act < quota) {
                const uint32_t room = quota - n_act;
                const uint32_t n_take = n_need < room ? n_need : room;
                uint32_t base = 0;
                if (__lane_id() == __ffsll((unsigned long long)need) - 1) base = atomicAdd(next_ray, n_take);
                base = __shfl(base, __ffsll((unsigned long long)need) - 1);
                if (base >= count) exhausted = true;
                if (!L.active) {
                    const uint32_t rank = __popcll(need & ((1ull << __lane_id()) - 1ull));
                    const uint32_t r = base + rank;
                    if (rank < n_take && r < count) {
                        const bool second = MODE == 2 && r >= c0;
                        const uint32_t ri = second ? r - c0 : r;
                        const float4 o = second ? j1.o[ri] : j0.o[ri];
                        const float4 d = second ? j1.d[ri] : j0.d[ri];
                        L.ray = ri;
                        L.job1 = second;
                        L.any = MODE == 1 || second;
                        L.O = ld3(o);
                        L.D = ld3(d);
                        L.tmin = o.w;
                        L.tmax = d.w;
                        L.rD = f3{safercp(L.D.x), safercp(L.D.y), safercp(L.D.z)};
                        near_far_offsets(L);
                        L.wO = L.O; L.wD = L.D; L.wrD = L.rD;
                        L.bt = L.tmax; L.bu = 0.f; L.bv = 0.f;
                        L.bprim = 0xffffffffu; L.binst = 0xffffffffu;
                        L.found = false;
                        L.head = 0; L.count = 0; L.gsp = 0;
                        L.level = TWO ? 0u : 1u;
                        L.cur_inst = 0;
                        L.node = TWO ? 0u : blas_root;
                        L.leaf = kNone;
                        L.active = true;
                    }
                }
            }
        }
        if (__ballot(L.active) == 0ull) {
            if (exhausted) break;
            continue;
        }
        // ---- ray finished
        auto finish = [&]() {
            if (lane_any<MODE>(L.any)) {
                float4* out = (MODE == 2 && L.job1) ? j1.hit : j0.hit;
                out[L.ray] = make_float4(0.f, 0.f, 0.f, __uint_as_float(L.found ? 1u : 0u));
            } else {
                j0.hit[L.ray] = L.found ? make_float4(L.bt, L.bu, L.bv, __uint_as_float(L.bprim))
                                        : make_float4(__builtin_inff(), 0.f, 0.f, __uint_as_float(0xffffffffu));
                if (j0.inst) j0.inst[L.ray] = L.found ? L.binst : 0xffffffffu;
            }
            if (L.overflow) atomicOr(overflow_flag, 1u);
            L.active = false;
        };
        // ---- the triangles of a BLAS leaf entry; an any-hit ray that hits drops its stack
        auto test_leaf = [&](uint32_t e) {
            const uint32_t first = e & kLeafFirstMask, cnt = (e >> 27) & 7u;
            // the next record is loaded while the current one is tested
            TriRec nxt = tris[first];
            bool term = false;
#pragma unroll 1
            for (uint32_t k = 0; k < cnt; ++k) {
                const TriRec tr = nxt;
                if (k + 1 < cnt) nxt = tris[first + k + 1];
                float t, u, v;
                if (!mt(tr, L.O, L.D, t, u, v)) continue;
                if (!(t > L.tmin && t < L.tmax)) continue;
                const uint32_t prim = __float_as_uint(tr.v0.w);
                if (lane_any<MODE>(L.any)) { L.found = true; term = true; break; }
                const uint32_t ci = TWO ? L.cur_inst : 0u;
                if (!L.found || t < L.bt ||
                    (t == L.bt && (ci < L.binst || (ci == L.binst && prim < L.bprim)))) {
                    L.bt = t; L.bu = u; L.bv = v; L.bprim = prim; L.binst = ci; L.found = true;
                }
            }
            if (MODE != 0 && term) { L.count = 0; L.gsp = 0; L.node = kNone; }
        };
        auto is_leaf = [](uint32_t e) { return (e & 0xc0000000u) == kBlasLeaf; };

        if (!TWO) {
            // ---- single level: leaves are postponed. A lane that reaches a leaf keeps it
            // in L.leaf and goes on with its next nodes; the wave tests leaves together
            // once kLeafBatch lanes hold one, or when no lane has a node to visit. Testing
            // each step's leaves as they come ran the leaf code for about 2 of 40 busy
            // lanes in more than half of the steps. The closest hit does not depend on
            // the order in which leaves are tested.
            if (L.active) {
                if (L.leaf == kNone && L.node != kNone && is_leaf(L.node)) { L.leaf = L.node; L.node = kNone; }
                if (L.leaf == kNone && L.node == kNone && L.count == 0u && L.gsp == 0u) finish();
            }
            // every active lane now has a node to visit or to pop, or a postponed leaf
            const bool nodeable = L.active && (L.node != kNone ? !is_leaf(L.node) : (L.count | L.gsp) != 0u);
            const uint64_t bn = __ballot(nodeable);
            const uint64_t bl = __ballot(L.active && L.leaf != kNone);
            if ((bn | bl) == 0ull) continue;
            if (bn == 0ull || uint32_t(__popcll(bl)) >= kLeafBatch) {
                if (L.active && L.leaf != kNone) {
                    test_leaf(L.leaf);
                    L.leaf = kNone;
                }
                continue;
            }
            if (!nodeable) continue;
            if (L.node == kNone) {
                uint32_t e = kNone;
                pop(L, ring, spill, gtid, stride, e);  // the stack is not empty
                if (is_leaf(e)) {
                    if (L.leaf == kNone) L.leaf = e;
                    else L.node = e;                    // waits for the leaf phase
                    continue;
                }
                L.node = e;
            }
        } else {
            if (!L.active) continue;
            // ---- one traversal step
            if (L.node == kNone) {
                uint32_t e;
                bool done = false;
                for (;;) {
                    if (!pop(L, ring, spill, gtid, stride, e)) { done = true; break; }
                    if (e == kStackMark) {  // leave the instance
                        L.O = L.wO; L.D = L.wD; L.rD = L.wrD;
                        near_far_offsets(L);
                        L.level = 0;
                        continue;
                    }
                    if (e & kInstFlag) {  // enter an instance
                        const uint32_t ii = e & ~kInstFlag;
                        const InstDev& I = inst[ii];
                        L.O = tpoint(L.wO, I.inv);
                        L.D = tvector(L.wD, I.inv);
                        L.rD = f3{safercp(L.D.x), safercp(L.D.y), safercp(L.D.z)};
                        near_far_offsets(L);
                        L.cur_inst = ii;
                        L.level = 1;
                        push(L, ring, spill, gtid, stride, kStackMark);
                        L.node = I.root;
                        break;
                    }
                    L.node = e;
                    break;
                }
                if (done) {
                    finish();
                    continue;
                }
            }
            if (L.level != 0u && is_leaf(L.node)) {
                test_leaf(L.node);
                L.node = kNone;
                continue;
            }
        }
        const bool tl_level = TWO && L.level == 0;
        float tl[4];
        uint4 nch, nmeta;
        slab4_ordered(reinterpret_cast<const char*>(tl_level ? tlas : blas), L.node, L, L.bt, tl, nch, nmeta);
        const uint32_t ch[4] = {nch.x, nch.y, nch.z, nch.w};
        const uint32_t mt4[4] = {nmeta.x, nmeta.y, nmeta.z, nmeta.w};
        // TLAS leaves push their instances; BLAS leaves join the internal children as
        // entries (a BLAS node stores a leaf child as its stack entry, kBlasLeaf | count
        // << 27 | first triangle), all near to far. The
        // four (distance, entry) pairs go through a fixed 5-exchange sorting network with
        // rejected children at +inf, so every index is static (a variable-length sort
        // compiles to chains of selects on a dynamic index); the accepted children are
        // then the finite prefix of the sorted four (slab4 returns +inf for a child it
        // rejects, and accepts only entries <= L.bt). The order of equal distances does
        // not matter: the closest hit is canonical.
        float kt[4];
        uint32_t kc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            kt[i] = tl[i];
            kc[i] = ch[i];
            if (tl_level && mt4[i] != 0u && tl[i] <= L.bt) {
                if (mt4[i] == 1u) {  // a one-instance leaf is ordered with the other children
                    kc[i] = kInstFlag | tlas_prims[ch[i]];
                } else {
                    for (uint32_t k = 0; k < mt4[i]; ++k)
                        push(L, ring, spill, gtid, stride, kInstFlag | tlas_prims[ch[i] + k]);
                    kt[i] = __builtin_inff();
                }
            }
        }
        cswap(kt, kc, 0, 1);
        cswap(kt, kc, 2, 3);
        cswap(kt, kc, 0, 2);
        cswap(kt, kc, 1, 3);
        cswap(kt, kc, 1, 2);
        if (kt[3] < __builtin_inff()) push(L, ring, spill, gtid, stride, kc[3]);
        if (kt[2] < __builtin_inff()) push(L, ring, spill, gtid, stride, kc[2]);
        if (kt[1] < __builtin_inff()) push(L, ring, spill, gtid, stride, kc[1]);
        L.node = kt[0] < __builtin_inff() ? kc[0] : kNone;
        if (TWO && L.node != kNone && (L.node & kInstFlag)) {  // instances are entered by the pop path
            push(L, ring, spill, gtid, stride, L.node);
            L.node = kNone;
        }
    }
}

__global__ void k_reset_counter(uint32_t* p) { *p = 0u; }

// user gk_ray/gk_hit records <-> queue layout
__global__ void k_unpack_rays(const gk_ray* __restrict__ r, uint32_t n, float4* o, float4* d) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const gk_ray x = r[i];
    o[i] = make_float4(x.origin[0], x.origin[1], x.origin[2], x.tmin);
    d[i] = make_float4(x.dir[0], x.dir[1], x.dir[2], x.tmax);
}

uint32_t cu_count(int device) {
    static int cus = 0;
    if (!cus) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, device) == hipSuccess) cus = p.multiProcessorCount;
        if (cus <= 0) cus = 256;
    }
    return uint32_t(cus);
}

// Persistent grid: as many blocks as the kernel keeps resident per CU (its VGPR and LDS
// use decide: 5 for the single-level kernels at 92 VGPRs, 4 for the two-level ones).
template <class K>
uint32_t persistent_blocks(int device, K kern, int cap = kMaxBlocksPerCU) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, kBlock, 0) != hipSuccess || nb < 1) nb = 4;
    if (nb > cap) nb = cap;
    return cu_count(device) * uint32_t(nb);
}

}  // namespace

int ensure_spill(Ctx& c, uint32_t threads, int slot) {
    const size_t need = size_t(threads) * kSpill;
    DevBuf<uint32_t>& sp = slot ? c.wf.pspill[slot - 1] : c.wf.spill;
    if (sp.count < need + 8) {
        GK_HIP(sp.resize(need + 8));
        GK_HIP(hipMemsetAsync(sp.ptr + need, 0, 32, c.stream));
    }
    c.wf.spill_threads = threads;
    if (c.tev.empty()) {
        c.tev.resize(1024);
        for (auto& e : c.tev) {
            GK_HIP(hipEventCreate(&e.first));
            GK_HIP(hipEventCreate(&e.second));
        }
    }
    return GK_OK;
}

uint32_t* trace_counter(Ctx& c, int slot) {
    // the spill area is laid out for the largest grid, so its tail (counters) never moves.
    // A slot's area is allocated (and its counters cleared on `stream`) on first use.
    const uint32_t cap_threads = cu_count(c.device) * kMaxBlocksPerCU * kBlock;
    if (ensure_spill(c, cap_threads, slot) != GK_OK) return nullptr;
    const DevBuf<uint32_t>& sp = slot ? c.wf.pspill[slot - 1] : c.wf.spill;
    return sp.ptr + size_t(cap_threads) * kSpill;  // [next ray, overflow, stats u64 x2]
}

// Launches one persistent trace over up to two jobs (mode as k_trace) in a slot (its stream
// and spill area). Each launch is bracketed by HIP events on that stream (collected by
// collect_timing).
static int launch_jobs(Ctx& c, int mode, const TraceJob& j0, const TraceJob& j1, int slot) {
    const bool zeroed = c.trace_counter_zeroed[slot];  // applies to this launch only
    c.trace_counter_zeroed[slot] = false;
    uint32_t* tail = trace_counter(c, slot);
    if (!tail) return GK_ERR_OOM;
    hipStream_t st = slot ? c.pstream[slot - 1] : c.stream;
    uint32_t* spill = slot ? c.wf.pspill[slot - 1].ptr : c.wf.spill.ptr;
    uint32_t* next = tail;
    uint32_t* ovf = tail + 1;
    unsigned long long* rs = reinterpret_cast<unsigned long long*>(tail + 2);
    if (!zeroed) k_reset_counter<<<1, 1, 0, st>>>(next);
    const bool two = !c.blas_only;
    const uint32_t root = c.h_meshes.empty() ? 0u : c.h_meshes[0].root;
    const Bvh4Node* tl = c.tlas_nodes.ptr;
    auto* kern = mode == 0 ? (two ? k_trace<0, true> : k_trace<0, false>)
               : mode == 1 ? (two ? k_trace<1, true> : k_trace<1, false>)
                           : (two ? k_trace<2, true> : k_trace<2, false>);
    const uint32_t blocks = persistent_blocks(c.device, kern, c.trace_split ? kSplitBlocksPerCU : kMaxBlocksPerCU);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (c.tev_used < c.tev.size()) { e0 = c.tev[c.tev_used].first; e1 = c.tev[c.tev_used].second; c.tev_used++; }
    if (e0) GK_HIP(hipEventRecord(e0, st));
    kern<<<blocks, kBlock, 0, st>>>(c.blas_nodes.ptr, c.tris.ptr, tl, c.tlas_prims.ptr, c.inst.ptr, root, j0, j1,
                                   next, spill, ovf, rs);
    GK_HIP(hipGetLastError());
    if (e1) GK_HIP(hipEventRecord(e1, st));
    return GK_OK;
}

// Trace a queue of rays. count_dev (device) caps the number of rays when non-null.
// hit: float4 (t,u,v,prim) per ray (any-hit: w = 1 when occluded). clear (device, may be
// null): a counter the launch zeroes, in place of a launch of its own.
int launch_trace(Ctx& c, const float4* ray_o, const float4* ray_d, const uint32_t* count_dev,
                 uint32_t max_count, float4* hit, uint32_t* hit_inst, bool any_hit, uint32_t* clear, int slot) {
    if (max_count == 0) return GK_OK;
    const TraceJob j0{ray_o, ray_d, count_dev, max_count, hit, any_hit ? nullptr : hit_inst, clear};
    return launch_jobs(c, any_hit ? 1 : 0, j0, TraceJob{}, slot);
}

// A closest-hit queue and an occlusion queue in one launch.
int launch_trace_pair(Ctx& c, const float4* o, const float4* d, const uint32_t* count_dev, uint32_t max_count,
                      float4* hit, uint32_t* hit_inst, const float4* so, const float4* sd,
                      const uint32_t* scount_dev, uint32_t smax_count, float4* shit, int slot) {
    const TraceJob j0{o, d, count_dev, max_count, hit, hit_inst, nullptr};
    const TraceJob j1{so, sd, scount_dev, smax_count, shit, nullptr, nullptr};
    return launch_jobs(c, 2, j0, j1, slot);
}

// Sums the elapsed time of the trace launches recorded since the last call (the stream
// must have been synchronised) and reads the device ray counters.
int collect_timing(Ctx& c) {
    for (uint32_t i = 0; i < c.tev_used; ++i) {
        float ms = 0.f;
        GK_HIP(hipEventElapsedTime(&ms, c.tev[i].first, c.tev[i].second));
        c.stats.ms_trace += ms;
        c.trace_launches++;
    }
    c.tev_used = 0;
    if (c.wf.spill_threads) {
        // every slot's counters: rays since the spill areas were allocated, overflow flags
        c.stats.rays_traced = 0;
        c.stats.shadow_rays = 0;
        for (int slot = 0; slot < 4; ++slot) {
            const DevBuf<uint32_t>& sp = slot ? c.wf.pspill[slot - 1] : c.wf.spill;
            if (!sp.ptr) continue;
            unsigned long long v[2] = {0, 0};
            uint32_t ovf = 0;
            const uint32_t* tail = sp.ptr + size_t(c.wf.spill_threads) * kSpill;
            GK_HIP(hipMemcpy(v, tail + 2, 16, hipMemcpyDeviceToHost));
            GK_HIP(hipMemcpy(&ovf, tail + 1, 4, hipMemcpyDeviceToHost));
            c.stats.rays_traced += v[0] + v[1];
            c.stats.shadow_rays += v[1];
            if (ovf) {  // kept until gk_get_stats reports it; the device flag starts over
                c.stack_overflow = true;
                const uint32_t zero = 0;
                GK_HIP(hipMemcpy(const_cast<uint32_t*>(tail + 1), &zero, 4, hipMemcpyHostToDevice));
            }
        }
    }
    return GK_OK;
}

int launch_trace_user(Ctx& c, const gk_ray* rays, uint32_t n, gk_hit* hits, uint32_t* inst, uint32_t flags) {
    if (n == 0) return GK_OK;
    // staging in the wavefront queue 0 (grown on demand)
    auto& w = c.wf;
    if (w.ray_o[0].count < n) { GK_HIP(w.ray_o[0].resize(n)); GK_HIP(w.ray_d[0].resize(n)); }
    k_unpack_rays<<<(n + 255) / 256, 256, 0, c.stream>>>(rays, n, w.ray_o[0].ptr, w.ray_d[0].ptr);
    GK_HIP(hipGetLastError());
    static_assert(sizeof(gk_hit) == sizeof(float4), "gk_hit is a float4");
    return launch_trace(c, w.ray_o[0].ptr, w.ray_d[0].ptr, nullptr, n, reinterpret_cast<float4*>(hits), inst,
                        (flags & GK_TRACE_ANY) != 0, nullptr);
}

// ---------------------------------------------------------------- allocation-free ray cast
// gk_trace_static: the same launch as gk_trace with everything beyond the caller's rays
// and hits in `scratch`, a static device array that each GPU holds from the library's load
// to the process's end. Layout (32-bit words): a 64-word head (next ray, overflow flag,
// two 64-bit ray counts), ray staging (origin/tmin then direction/tmax, float4 each, for
// kStaticMaxRays rays), then the spill stacks of a kStaticBlocks-work-group launch
// ([entry][thread], kSpill entries per thread). 24 MiB + 256 bytes.
// Host code never passes `scratch` itself to a kernel (on the host it names a shadow
// variable, not device memory): gk_create resolves its device address once per GPU with
// hipGetSymbolAddress, with that GPU current, and every use below goes through it.
constexpr uint32_t kStaticMaxRays = GK_TRACE_STATIC_MAX_RAYS;
constexpr uint32_t kStaticBlocks = 256;  // work-groups of a launch (the spill area is sized for it)
constexpr size_t kStaticHead = 64;
constexpr size_t kStaticWords = kStaticHead + size_t(kStaticMaxRays) * 8 + size_t(kStaticBlocks) * kBlock * kSpill;
static __device__ __align__(256) uint32_t scratch[kStaticWords];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch[kMaxDevices] = {};

__global__ void k_clear_words(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}
}  // namespace

// Resolves this GPU's copy of `scratch` (the caller has made `device` current).
int static_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return GK_ERR_NO_DEVICE;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return GK_OK;
    void* p = nullptr;
    GK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = static_cast<uint32_t*>(p);
    return GK_OK;
}

// Reads and clears the overflow flag of the last gk_trace_static call on c's GPU.
int static_scratch_overflow(Ctx& c, bool& overflow) {
    overflow = false;
    uint32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c.device >= 0 && c.device < kMaxDevices) base = g_scratch[c.device];
    }
    if (!base) return GK_OK;
    uint32_t v = 0;
    GK_HIP(hipMemcpy(&v, base + 1, 4, hipMemcpyDeviceToHost));
    if (v) {
        const uint32_t zero = 0;
        GK_HIP(hipMemcpy(base + 1, &zero, 4, hipMemcpyHostToDevice));
        overflow = true;
    }
    return GK_OK;
}

}  // namespace gk

extern "C" int gk_trace_static(gk_ctx* ctx, const gk_ray* d_rays, uint32_t n, gk_hit* d_hits, uint32_t* d_inst,
                               uint32_t flags) {
    using namespace gk;
    GK_CHECK_CTX(ctx);
    Ctx& c = ctx->c;
    if (!c.scene_ready) return GK_ERR_STATE;
    if ((!d_rays || !d_hits) && n) return GK_ERR_ARG;
    if (flags > GK_TRACE_ANY) return GK_ERR_ARG;
    if (n > kStaticMaxRays) return GK_ERR_LIMIT;
    uint32_t* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c.device >= 0 && c.device < kMaxDevices) base = g_scratch[c.device];
    }
    if (!base) return GK_ERR_HIP;  // gk_create could not resolve this GPU's array
    GK_HIP(hipSetDevice(c.device));
    if (c.tlas_dirty) {
        const int s = refit_tlas(c);  // in place, no allocation
        if (s != GK_OK) return s;
    }
    if (n == 0) return GK_OK;
    uint32_t* head = base;                                   // next ray, overflow, ray counts
    float4* so = reinterpret_cast<float4*>(base + kStaticHead);
    float4* sd = so + kStaticMaxRays;
    uint32_t* spill = reinterpret_cast<uint32_t*>(sd + kStaticMaxRays);
    // the previous call's counters are still there: start from zero
    k_clear_words<<<1, uint32_t(kStaticHead), 0, c.stream>>>(head, uint32_t(kStaticHead));
    k_unpack_rays<<<(n + 255) / 256, 256, 0, c.stream>>>(d_rays, n, so, sd);
    GK_HIP(hipGetLastError());
    const bool any = (flags & GK_TRACE_ANY) != 0;
    const bool two = !c.blas_only;
    const uint32_t root = c.h_meshes.empty() ? 0u : c.h_meshes[0].root;
    auto* kern = any ? (two ? k_trace<1, true> : k_trace<1, false>) : (two ? k_trace<0, true> : k_trace<0, false>);
    uint32_t blocks = persistent_blocks(c.device, kern);
    if (blocks > kStaticBlocks) blocks = kStaticBlocks;
    const TraceJob j0{so, sd, nullptr, n, reinterpret_cast<float4*>(d_hits), any ? nullptr : d_inst, nullptr};
    kern<<<blocks, kBlock, 0, c.stream>>>(c.blas_nodes.ptr, c.tris.ptr, c.tlas_nodes.ptr, c.tlas_prims.ptr, c.inst.ptr,
                                         root, j0, TraceJob{}, head, spill, head + 1,
                                         reinterpret_cast<unsigned long long*>(head + 2));
    GK_HIP(hipGetLastError());
    return GK_OK;
}

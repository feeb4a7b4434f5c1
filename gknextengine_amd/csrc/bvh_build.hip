// bvh_build.hip - acceleration-structure construction on the device.
//
// Replaces the driver-built BLAS/TLAS of RayTraceBaseRenderer
// (CreateBottomLevelStructures / CreateTopLevelStructures / PreRender TLAS update,
// RayTraceBaseRenderer.cpp:205-218,276-386) and, for the CPU path, tinybvh
// BVH::Build / Build(BLASInstance*) (tiny_bvh.h:1565-1603,1674-1767).
//
// Pipeline per primitive set (a mesh's triangles, or the instances of the TLAS):
//   1. primitive boxes; centroid bounds (block reduction + ordered-int atomics)
//   2. 63-bit Morton codes of the centroids, radix sort (hipCUB)
//   3. binary radix tree over the sorted codes (Karras 2012), ties in the code broken
//      by the primitive's sorted position
//   4. bottom-up pass (one thread per leaf, second arrival continues): node boxes and
//      the SAH cost min(C_INT*n*SA, C_TRAV*SA + C(l) + C(r)) with C_TRAV = C_INT = 1
//      (tinybvh's constants, tiny_bvh.h:118-123); a subtree whose leaf cost wins and
//      that holds <= max_leaf primitives becomes one leaf (its primitives are
//      contiguous in Morton order)
//   5. top-down collapse into 4-wide nodes, one launch per level: a wide node starts
//      from the two children of its binary node and repeatedly opens the child with
//      the largest surface area until it has 4 children
//   6. triangle records reordered into leaf order
// TLAS per frame: instance matrices -> inverse (tinybvh's InvertTransform expansion),
// combinedPrevTS (Node::TickVelocity), world boxes of the BLAS root (BLASInstance::Update,
// tiny_bvh.h:6718-6731), then a bottom-up refit of the wide TLAS by levels.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "gk_ctx.hpp"

namespace gk {

namespace {

constexpr float kEmpty = 3e37f;  // empty child slot: a point box that no ray reaches
constexpr uint32_t kLeafBit = 0x80000000u;

__device__ inline uint32_t fenc(float f) {
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__host__ __device__ inline float fdec(uint32_t u) {
    const uint32_t b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &b, 4);
    return f;
}
__device__ inline float half_area(float3 lo, float3 hi) {
    const float ex = hi.x - lo.x, ey = hi.y - lo.y, ez = hi.z - lo.z;
    return ex * ey + ey * ez + ez * ex;
}

// ---------------------------------------------------------------- triangle prep
__global__ void k_prep_tris(const float* __restrict__ pos, const uint32_t* __restrict__ idx,
                            const float* __restrict__ nrm, const uint32_t* __restrict__ tmat,
                            uint32_t nv, uint32_t nt, uint32_t tri_base, TriRec* __restrict__ tmp,
                            float4* __restrict__ lo, float4* __restrict__ hi, float4* __restrict__ tri_nrm,
                            uint32_t* __restrict__ tri_mat, uint32_t* __restrict__ err) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    uint32_t a = idx[3 * t], b = idx[3 * t + 1], c = idx[3 * t + 2];
    if (a >= nv || b >= nv || c >= nv) { atomicOr(err, 1u); a = b = c = 0; }
    f3 pa = f3{pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]};
    f3 pb = f3{pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]};
    f3 pc = f3{pos[3 * c], pos[3 * c + 1], pos[3 * c + 2]};
    if (!(isfinite(pa.x) && isfinite(pa.y) && isfinite(pa.z) && isfinite(pb.x) && isfinite(pb.y) && isfinite(pb.z) &&
          isfinite(pc.x) && isfinite(pc.y) && isfinite(pc.z))) {  // a NaN/Inf vertex: the upload fails
        atomicOr(err, 2u);
        pa = pb = pc = f3{0.f, 0.f, 0.f};
    }
    const f3 e1 = pb - pa, e2 = pc - pa;
    const uint32_t m = tmat ? tmat[t] : 0u;
    tmp[t].v0 = make_float4(pa.x, pa.y, pa.z, __uint_as_float(t));
    tmp[t].e1 = make_float4(e1.x, e1.y, e1.z, __uint_as_float(m));
    tmp[t].e2 = make_float4(e2.x, e2.y, e2.z, 0.f);
    const f3 l = fmin3(fmin3(pa, pb), pc), h = fmax3(fmax3(pa, pb), pc);
    lo[t] = make_float4(l.x, l.y, l.z, 0.f);
    hi[t] = make_float4(h.x, h.y, h.z, 0.f);
    f3 n0, n1, n2;
    if (nrm) {
        n0 = f3{nrm[3 * a], nrm[3 * a + 1], nrm[3 * a + 2]};
        n1 = f3{nrm[3 * b], nrm[3 * b + 1], nrm[3 * b + 2]};
        n2 = f3{nrm[3 * c], nrm[3 * c + 1], nrm[3 * c + 2]};
    } else {
        n0 = n1 = n2 = normalize(cross(e1, e2));
    }
    // vertex normals in xyz, the unnormalised face normal e1 x e2 in the w lanes
    const f3 fn = cross(e1, e2);
    const size_t o = 3 * size_t(tri_base + t);
    tri_nrm[o] = make_float4(n0.x, n0.y, n0.z, fn.x);
    tri_nrm[o + 1] = make_float4(n1.x, n1.y, n1.z, fn.y);
    tri_nrm[o + 2] = make_float4(n2.x, n2.y, n2.z, fn.z);
    tri_mat[tri_base + t] = m;
}

// texture coordinates and tangents of the triangle's vertices (original order); a mesh
// without them gets zeros
__global__ void k_prep_tex(const uint32_t* __restrict__ idx, const float* __restrict__ uv,
                           const float* __restrict__ tan, uint32_t nv, uint32_t nt, uint32_t tri_base,
                           float2* __restrict__ tri_uv, float4* __restrict__ tri_tan) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nt) return;
    for (int k = 0; k < 3; ++k) {
        uint32_t a = idx[3 * t + k];
        if (a >= nv) a = 0;  // reported by k_prep_tris
        const size_t o = 3 * size_t(tri_base + t) + k;
        if (tri_uv) tri_uv[o] = uv ? make_float2(uv[2 * a], uv[2 * a + 1]) : make_float2(0.f, 0.f);
        if (tri_tan)
            tri_tan[o] = tan ? make_float4(tan[4 * a], tan[4 * a + 1], tan[4 * a + 2], tan[4 * a + 3])
                             : make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// ---------------------------------------------------------------- centroid bounds
__global__ void k_init_bounds(uint32_t* enc) {
    if (threadIdx.x < 3) enc[threadIdx.x] = 0xffffffffu;       // min
    else if (threadIdx.x < 6) enc[threadIdx.x] = 0u;           // max
}

__global__ void k_bounds(const float4* __restrict__ lo, const float4* __restrict__ hi, uint32_t n,
                         uint32_t* __restrict__ enc) {
    __shared__ uint32_t s[6][256];
    const uint32_t tid = threadIdx.x;
    uint32_t mn[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, mx[3] = {0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * blockDim.x + tid; i < n; i += gridDim.x * blockDim.x) {
        const float4 a = lo[i], b = hi[i];
        const float c[3] = {(a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f};
        for (int k = 0; k < 3; ++k) {
            const uint32_t e = fenc(c[k]);
            mn[k] = e < mn[k] ? e : mn[k];
            mx[k] = e > mx[k] ? e : mx[k];
        }
    }
    for (int k = 0; k < 3; ++k) { s[k][tid] = mn[k]; s[3 + k][tid] = mx[k]; }
    __syncthreads();
    for (uint32_t w = blockDim.x / 2; w > 0; w >>= 1) {
        if (tid < w)
            for (int k = 0; k < 3; ++k) {
                s[k][tid] = min(s[k][tid], s[k][tid + w]);
                s[3 + k][tid] = max(s[3 + k][tid], s[3 + k][tid + w]);
            }
        __syncthreads();
    }
    if (tid == 0)
        for (int k = 0; k < 3; ++k) { atomicMin(&enc[k], s[k][0]); atomicMax(&enc[3 + k], s[3 + k][0]); }
}

__device__ inline uint64_t spread21(uint64_t x) {
    x &= 0x1fffffull;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ void k_morton(const float4* __restrict__ lo, const float4* __restrict__ hi, uint32_t n,
                         const uint32_t* __restrict__ enc, uint64_t* __restrict__ keys,
                         uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = lo[i], b = hi[i];
    const float c[3] = {(a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f, (a.z + b.z) * 0.5f};
    uint64_t code = 0;
    for (int k = 0; k < 3; ++k) {
        const float mn = fdec(enc[k]), mx = fdec(enc[3 + k]);
        const float ext = mx - mn;
        float u = ext > 0.f ? (c[k] - mn) / ext : 0.5f;
        u = fminf(fmaxf(u, 0.f), 1.f);
        const uint64_t q = uint64_t(fminf(u * 2097152.f, 2097151.f));
        code |= spread21(q) << (2 - k);
    }
    keys[i] = code;
    vals[i] = i;
}

// ---------------------------------------------------------------- radix tree (Karras 2012)
__device__ inline int delta(const uint64_t* __restrict__ k, int n, int i, int j) {
    if (j < 0 || j > n - 1) return -1;
    const uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz(uint32_t(i ^ j));
    return __clzll(a ^ b);
}

__global__ void k_karras(const uint64_t* __restrict__ keys, int n, uint2* __restrict__ bchild,
                         uint2* __restrict__ brange, uint32_t* __restrict__ bparent) {
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n - 1) return;
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0;
    int div = 2;
    int t = (l + div - 1) / div;
    while (t >= 1) {
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
        if (t == 1) break;
        div <<= 1;
        t = (l + div - 1) / div;
    }
    const int gamma = i + s * d + (d < 0 ? d : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t left, right;
    if (lo == gamma) { left = kLeafBit | uint32_t(gamma); bparent[n - 1 + gamma] = uint32_t(i); }
    else { left = uint32_t(gamma); bparent[gamma] = uint32_t(i); }
    if (hi == gamma + 1) { right = kLeafBit | uint32_t(gamma + 1); bparent[n - 1 + gamma + 1] = uint32_t(i); }
    else { right = uint32_t(gamma + 1); bparent[gamma + 1] = uint32_t(i); }
    bchild[i] = make_uint2(left, right);
    brange[i] = make_uint2(uint32_t(lo), uint32_t(hi));
}

__global__ void k_zero_u32(uint32_t* p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__device__ inline float ld_coherent(const float* p) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT));
}

// box + cost of a binary child (leaf: primitive box, cost = its half area)
__device__ inline void child_info(uint32_t c, const float4* __restrict__ plo, const float4* __restrict__ phi,
                                  const uint32_t* __restrict__ vals, const float4* blo, const float4* bhi,
                                  float3& lo, float3& hi, float& cost) {
    if (c & kLeafBit) {
        const uint32_t p = vals[c & ~kLeafBit];
        const float4 a = plo[p], b = phi[p];
        lo = make_float3(a.x, a.y, a.z);
        hi = make_float3(b.x, b.y, b.z);
        cost = half_area(lo, hi);
    } else {
        const float* a = reinterpret_cast<const float*>(blo + c);
        const float* b = reinterpret_cast<const float*>(bhi + c);
        lo = make_float3(ld_coherent(a), ld_coherent(a + 1), ld_coherent(a + 2));
        hi = make_float3(ld_coherent(b), ld_coherent(b + 1), ld_coherent(b + 2));
        cost = ld_coherent(a + 3);
    }
}

__global__ void k_bottom_up(const float4* __restrict__ plo, const float4* __restrict__ phi,
                            const uint32_t* __restrict__ vals, int n, const uint2* __restrict__ bchild,
                            const uint2* __restrict__ brange, const uint32_t* __restrict__ bparent,
                            uint32_t* __restrict__ bflag, float4* blo, float4* bhi, uint32_t max_leaf) {
    const int k = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    uint32_t node = bparent[n - 1 + k];
    for (;;) {
        __threadfence();
        const uint32_t old = atomicAdd(&bflag[node], 1u);
        if (old == 0u) return;
        __threadfence();
        const uint2 ch = bchild[node];
        float3 la, ha, lb, hb;
        float ca, cb;
        child_info(ch.x, plo, phi, vals, blo, bhi, la, ha, ca);
        child_info(ch.y, plo, phi, vals, blo, bhi, lb, hb, cb);
        const float3 lo = make_float3(fminf(la.x, lb.x), fminf(la.y, lb.y), fminf(la.z, lb.z));
        const float3 hi = make_float3(fmaxf(ha.x, hb.x), fmaxf(ha.y, hb.y), fmaxf(ha.z, hb.z));
        const float sa = half_area(lo, hi);
        const uint2 r = brange[node];
        const uint32_t cnt = r.y - r.x + 1u;
        const float c_split = sa + ca + cb;
        const float c_leaf = cnt <= max_leaf ? sa * float(cnt) : __builtin_inff();
        const bool leaf = c_leaf <= c_split;
        const float cost = leaf ? c_leaf : c_split;
        blo[node] = make_float4(lo.x, lo.y, lo.z, cost);
        bhi[node] = make_float4(hi.x, hi.y, hi.z, leaf ? 1.f : 0.f);
        if (node == 0u) return;
        node = bparent[node];
    }
}

// ---------------------------------------------------------------- wide collapse
struct CollapseArgs {
    const uint2* bchild;
    const uint2* brange;
    const float4* blo;
    const float4* bhi;
    const float4* plo;
    const float4* phi;
    const uint32_t* vals;
    Bvh4Node* out;
    uint32_t node_base;
    uint32_t leaf_base;
    uint32_t blas;    // 1: leaf children hold their traversal stack entry (blas_leaf_entry)
    uint32_t* wide_count;
    uint32_t* depth;  // per wide node (may be null)
    uint32_t level;
};

__device__ inline bool is_collapsed(const CollapseArgs& a, uint32_t e) {
    return (e & kLeafBit) || a.bhi[e].w != 0.f;
}
__device__ inline void elem_box(const CollapseArgs& a, uint32_t e, float3& lo, float3& hi) {
    if (e & kLeafBit) {
        const uint32_t p = a.vals[e & ~kLeafBit];
        const float4 l = a.plo[p], h = a.phi[p];
        lo = make_float3(l.x, l.y, l.z);
        hi = make_float3(h.x, h.y, h.z);
    } else {
        const float4 l = a.blo[e], h = a.bhi[e];
        lo = make_float3(l.x, l.y, l.z);
        hi = make_float3(h.x, h.y, h.z);
    }
}

// One queue item of the top-down collapse: binary subtree root_elem becomes wide node w, its
// internal children are queued for the next level with fresh wide-node indices.
__device__ void collapse_item(const CollapseArgs& a, const uint2 item, uint2* __restrict__ qout,
                              uint32_t* __restrict__ qout_count) {
    const uint32_t root_elem = item.x, w = item.y;
    uint32_t set[4];
    int ns;
    if (is_collapsed(a, root_elem)) {
        set[0] = root_elem;
        ns = 1;
    } else {
        const uint2 ch = a.bchild[root_elem];
        set[0] = ch.x;
        set[1] = ch.y;
        ns = 2;
        while (ns < 4) {
            int best = -1;
            float best_sa = -1.f;
            for (int j = 0; j < ns; ++j) {
                if (is_collapsed(a, set[j])) continue;
                float3 lo, hi;
                elem_box(a, set[j], lo, hi);
                const float sa = half_area(lo, hi);
                if (sa > best_sa) { best_sa = sa; best = j; }
            }
            if (best < 0) break;
            const uint2 ch2 = a.bchild[set[best]];
            set[best] = ch2.x;
            set[ns++] = ch2.y;
        }
    }
    float lo4[3][4], hi4[3][4];
    uint32_t child[4], meta[4];
    for (int j = 0; j < 4; ++j) {
        if (j >= ns) {
            for (int k = 0; k < 3; ++k) { lo4[k][j] = kEmpty; hi4[k][j] = kEmpty; }
            child[j] = 0u;
            meta[j] = 0u;
            continue;
        }
        const uint32_t e = set[j];
        float3 lo, hi;
        elem_box(a, e, lo, hi);
        lo4[0][j] = lo.x; lo4[1][j] = lo.y; lo4[2][j] = lo.z;
        hi4[0][j] = hi.x; hi4[1][j] = hi.y; hi4[2][j] = hi.z;
        if (e & kLeafBit) {
            child[j] = a.leaf_base + (e & ~kLeafBit);
            meta[j] = 1u;
            if (a.blas) child[j] = blas_leaf_entry(child[j], 1u);
        } else if (a.bhi[e].w != 0.f) {
            const uint2 r = a.brange[e];
            child[j] = a.leaf_base + r.x;
            meta[j] = r.y - r.x + 1u;
            if (a.blas) child[j] = blas_leaf_entry(child[j], meta[j]);
        } else {
            const uint32_t idx = atomicAdd(a.wide_count, 1u);
            child[j] = a.node_base + idx;
            meta[j] = 0u;
            const uint32_t q = atomicAdd(qout_count, 1u);
            qout[q] = make_uint2(e, idx);
        }
    }
    Bvh4Node nd;
    nd.lox = make_float4(lo4[0][0], lo4[0][1], lo4[0][2], lo4[0][3]);
    nd.hix = make_float4(hi4[0][0], hi4[0][1], hi4[0][2], hi4[0][3]);
    nd.loy = make_float4(lo4[1][0], lo4[1][1], lo4[1][2], lo4[1][3]);
    nd.hiy = make_float4(hi4[1][0], hi4[1][1], hi4[1][2], hi4[1][3]);
    nd.loz = make_float4(lo4[2][0], lo4[2][1], lo4[2][2], lo4[2][3]);
    nd.hiz = make_float4(hi4[2][0], hi4[2][1], hi4[2][2], hi4[2][3]);
    nd.child = make_uint4(child[0], child[1], child[2], child[3]);
    nd.meta = make_uint4(meta[0], meta[1], meta[2], meta[3]);
    a.out[a.node_base + w] = nd;
    if (a.depth) a.depth[w] = a.level;
}

__global__ void k_collapse(CollapseArgs a, const uint2* __restrict__ qin, const uint32_t* __restrict__ qin_count,
                           uint2* __restrict__ qout, uint32_t* __restrict__ qout_count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *qin_count) return;
    collapse_item(a, qin[i], qout, qout_count);
}

// The whole top-down collapse in one work-group, level after level with barriers in between,
// so that no level waits for the host (a TLAS rebuild per frame). Wide nodes are numbered
// level by level (a level's children take consecutive indices), so level lv holds nodes
// [end[lv-1], end[lv]); the kernel then writes the refit order of k_refit_tlas, deepest level
// first: level_info[0] = number of levels, level_info[1 + k] = start of the k-th group in
// order[]. counters as k_set_root left them: [0] wide nodes, [1] queue a size, [2] queue b size.
constexpr uint32_t kCollapseMaxLevels = 255;
__global__ __launch_bounds__(1024) void k_collapse_block(CollapseArgs a, uint2* __restrict__ qa,
                                                         uint2* __restrict__ qb, uint32_t* __restrict__ counters,
                                                         uint32_t* __restrict__ order,
                                                         uint32_t* __restrict__ level_info) {
    __shared__ uint32_t s_count;
    __shared__ uint32_t s_end[kCollapseMaxLevels + 1];
    uint32_t L = 0;
    for (;; ++L) {
        uint32_t* cin = counters + 1 + (L & 1u);
        uint32_t* cout = counters + 1 + ((L + 1u) & 1u);
        if (threadIdx.x == 0) {
            s_count = L < kCollapseMaxLevels ? __hip_atomic_load(cin, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
            s_end[L] = __hip_atomic_load(counters, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);  // levels 0..L
            __hip_atomic_store(cout, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
        const uint32_t count = s_count;
        if (count == 0) break;
        CollapseArgs al = a;
        al.level = L;
        const uint2* qin = (L & 1u) ? qb : qa;
        uint2* qout = (L & 1u) ? qa : qb;
        for (uint32_t i = threadIdx.x; i < count; i += blockDim.x) collapse_item(al, qin[i], qout, cout);
        __threadfence();
        __syncthreads();
        // the queues alternate: drop any L1 copy of the entries this level wrote
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    // L levels: 0 .. L-1 (L == kCollapseMaxLevels: the tree is deeper than the kernel
    // follows; 0 levels tell the refit there is nothing valid to refit)
    const uint32_t levels = L < kCollapseMaxLevels ? L : 0u;
    uint32_t pos = 0;
    for (uint32_t k = 0; k < levels; ++k) {
        const uint32_t lv = levels - 1u - k;
        const uint32_t b = lv ? s_end[lv - 1u] : 0u, e = s_end[lv];
        if (threadIdx.x == 0) level_info[1 + k] = pos;
        for (uint32_t i = b + threadIdx.x; i < e; i += blockDim.x) order[pos + (i - b)] = i;
        pos += e - b;
    }
    if (threadIdx.x == 0) {
        level_info[0] = levels;
        level_info[1 + levels] = pos;
    }
}

__global__ void k_set_root(uint2* q, uint32_t* counters, uint32_t root_elem) {
    q[0] = make_uint2(root_elem, 0u);
    counters[0] = 1u;  // wide nodes allocated (root)
    counters[1] = 1u;  // queue a size
    counters[2] = 0u;  // queue b size
    counters[3] = 0u;
}
__global__ void k_zero_word(uint32_t* p) { *p = 0u; }

__global__ void k_reorder_tris(const TriRec* __restrict__ tmp, const uint32_t* __restrict__ vals, uint32_t n,
                               TriRec* __restrict__ out, uint32_t* __restrict__ rec_of) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = tmp[vals[k]];
    rec_of[vals[k]] = k;
}

__global__ void k_copy_u32(const uint32_t* __restrict__ a, uint32_t* __restrict__ b, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) b[k] = a[k];
}

inline uint32_t nblk(uint32_t n, uint32_t b = 256) { return (n + b - 1) / b; }

struct WideResult {
    uint32_t nodes = 0;
    uint32_t depth = 0;
    double sah = 0.0;
    const uint32_t* vals = nullptr;  // sorted primitive order (device)
};

// LBVH + SAH collapse + 4-wide collapse of n primitive boxes. Wide nodes are written
// to out[node_base ...], leaf children point at leaf_base + sorted position.
int build_wide(Ctx& c, const float4* plo, const float4* phi, uint32_t n, uint32_t max_leaf, Bvh4Node* out,
               uint32_t node_base, uint32_t leaf_base, uint32_t* depth_out, WideResult& res, bool blas,
               uint32_t* refit_order, uint32_t* level_info) {
    BuildScratch& s = c.scratch;
    hipStream_t st = c.stream;
    GK_HIP(s.keys.resize(n)); GK_HIP(s.keys_alt.resize(n));
    GK_HIP(s.vals.resize(n)); GK_HIP(s.vals_alt.resize(n));
    GK_HIP(s.bounds.resize(8));
    GK_HIP(s.counters.resize(8));
    uint32_t* enc = reinterpret_cast<uint32_t*>(s.bounds.ptr);
    k_init_bounds<<<1, 64, 0, st>>>(enc);
    k_bounds<<<std::min<uint32_t>(nblk(n), 1024), 256, 0, st>>>(plo, phi, n, enc);
    k_morton<<<nblk(n), 256, 0, st>>>(plo, phi, n, enc, s.keys.ptr, s.vals.ptr);
    GK_HIP(hipGetLastError());
    hipcub::DoubleBuffer<uint64_t> dk(s.keys.ptr, s.keys_alt.ptr);
    hipcub::DoubleBuffer<uint32_t> dv(s.vals.ptr, s.vals_alt.ptr);
    size_t tmp_bytes = 0;
    GK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, dk, dv, int(n), 0, 63, st));
    GK_HIP(s.sort_tmp.resize(tmp_bytes + 16));
    GK_HIP(hipcub::DeviceRadixSort::SortPairs(s.sort_tmp.ptr, tmp_bytes, dk, dv, int(n), 0, 63, st));
    const uint64_t* keys = dk.Current();
    const uint32_t* vals = dv.Current();
    res.vals = vals;
    const uint32_t ni = n > 1 ? n - 1 : 1;
    GK_HIP(s.bchild.resize(ni)); GK_HIP(s.brange.resize(ni));
    GK_HIP(s.bparent.resize(2 * n));
    GK_HIP(s.bnode_lo.resize(ni)); GK_HIP(s.bnode_hi.resize(ni));
    GK_HIP(s.bflag.resize(ni));
    uint32_t root_elem = kLeafBit | 0u;
    if (n > 1) {
        k_karras<<<nblk(n - 1), 256, 0, st>>>(keys, int(n), s.bchild.ptr, s.brange.ptr, s.bparent.ptr);
        k_zero_u32<<<nblk(n - 1), 256, 0, st>>>(s.bflag.ptr, n - 1);
        k_bottom_up<<<nblk(n), 256, 0, st>>>(plo, phi, vals, int(n), s.bchild.ptr, s.brange.ptr, s.bparent.ptr,
                                            s.bflag.ptr, s.bnode_lo.ptr, s.bnode_hi.ptr, max_leaf);
        GK_HIP(hipGetLastError());
        root_elem = 0u;
    }
    // top-down collapse
    GK_HIP(s.queue_a.resize(2 * size_t(n) + 2));
    GK_HIP(s.queue_b.resize(2 * size_t(n) + 2));
    uint2* qa = reinterpret_cast<uint2*>(s.queue_a.ptr);
    uint2* qb = reinterpret_cast<uint2*>(s.queue_b.ptr);
    k_set_root<<<1, 1, 0, st>>>(qa, s.counters.ptr, root_elem);
    CollapseArgs a{s.bchild.ptr, s.brange.ptr, s.bnode_lo.ptr, s.bnode_hi.ptr, plo, phi, vals,
                   out, node_base, leaf_base, blas ? 1u : 0u, s.counters.ptr + 0, depth_out, 0};
    if (refit_order) {
        // device only: one work-group runs every level and writes the refit order; nothing
        // is read back (res keeps only the sorted primitive order)
        k_collapse_block<<<1, 1024, 0, st>>>(a, qa, qb, s.counters.ptr, refit_order, level_info);
        GK_HIP(hipGetLastError());
        return GK_OK;
    }
    uint32_t qcount = 1, level = 0;
    uint32_t* cin = s.counters.ptr + 1;
    uint32_t* cout = s.counters.ptr + 2;
    while (qcount > 0) {
        a.level = level;
        k_zero_word<<<1, 1, 0, st>>>(cout);
        k_collapse<<<nblk(qcount), 256, 0, st>>>(a, qa, cin, qb, cout);
        GK_HIP(hipGetLastError());
        GK_HIP(hipMemcpyAsync(&qcount, cout, 4, hipMemcpyDeviceToHost, st));
        GK_HIP(hipStreamSynchronize(st));
        std::swap(qa, qb);
        std::swap(cin, cout);
        level++;
        if (level > 200) return GK_ERR_LIMIT;
    }
    uint32_t wide = 0;
    GK_HIP(hipMemcpyAsync(&wide, s.counters.ptr, 4, hipMemcpyDeviceToHost, st));
    float4 rl{}, rh{};
    if (n > 1) {
        GK_HIP(hipMemcpyAsync(&rl, s.bnode_lo.ptr, 16, hipMemcpyDeviceToHost, st));
        GK_HIP(hipMemcpyAsync(&rh, s.bnode_hi.ptr, 16, hipMemcpyDeviceToHost, st));
    }
    GK_HIP(hipStreamSynchronize(st));
    res.nodes = wide;
    res.depth = level;
    if (n > 1) {
        const float sa = (rh.x - rl.x) * (rh.y - rl.y) + (rh.y - rl.y) * (rh.z - rl.z) + (rh.z - rl.z) * (rh.x - rl.x);
        res.sah = sa > 0.f ? double(rl.w) / double(sa) : 0.0;
    } else {
        res.sah = 1.0;
    }
    return GK_OK;
}

// ---------------------------------------------------------------- instances
// tinybvh BLASInstance::InvertTransform expansion (tiny_bvh.h:6734-6756), row-major
__device__ void invert_rowmajor(const float* T, float* inv) {
    inv[0] = T[5] * T[10] * T[15] - T[5] * T[11] * T[14] - T[9] * T[6] * T[15] + T[9] * T[7] * T[14] + T[13] * T[6] * T[11] - T[13] * T[7] * T[10];
    inv[1] = -T[1] * T[10] * T[15] + T[1] * T[11] * T[14] + T[9] * T[2] * T[15] - T[9] * T[3] * T[14] - T[13] * T[2] * T[11] + T[13] * T[3] * T[10];
    inv[2] = T[1] * T[6] * T[15] - T[1] * T[7] * T[14] - T[5] * T[2] * T[15] + T[5] * T[3] * T[14] + T[13] * T[2] * T[7] - T[13] * T[3] * T[6];
    inv[3] = -T[1] * T[6] * T[11] + T[1] * T[7] * T[10] + T[5] * T[2] * T[11] - T[5] * T[3] * T[10] - T[9] * T[2] * T[7] + T[9] * T[3] * T[6];
    inv[4] = -T[4] * T[10] * T[15] + T[4] * T[11] * T[14] + T[8] * T[6] * T[15] - T[8] * T[7] * T[14] - T[12] * T[6] * T[11] + T[12] * T[7] * T[10];
    inv[5] = T[0] * T[10] * T[15] - T[0] * T[11] * T[14] - T[8] * T[2] * T[15] + T[8] * T[3] * T[14] + T[12] * T[2] * T[11] - T[12] * T[3] * T[10];
    inv[6] = -T[0] * T[6] * T[15] + T[0] * T[7] * T[14] + T[4] * T[2] * T[15] - T[4] * T[3] * T[14] - T[12] * T[2] * T[7] + T[12] * T[3] * T[6];
    inv[7] = T[0] * T[6] * T[11] - T[0] * T[7] * T[10] - T[4] * T[2] * T[11] + T[4] * T[3] * T[10] + T[8] * T[2] * T[7] - T[8] * T[3] * T[6];
    inv[8] = T[4] * T[9] * T[15] - T[4] * T[11] * T[13] - T[8] * T[5] * T[15] + T[8] * T[7] * T[13] + T[12] * T[5] * T[11] - T[12] * T[7] * T[9];
    inv[9] = -T[0] * T[9] * T[15] + T[0] * T[11] * T[13] + T[8] * T[1] * T[15] - T[8] * T[3] * T[13] - T[12] * T[1] * T[11] + T[12] * T[3] * T[9];
    inv[10] = T[0] * T[5] * T[15] - T[0] * T[7] * T[13] - T[4] * T[1] * T[15] + T[4] * T[3] * T[13] + T[12] * T[1] * T[7] - T[12] * T[3] * T[5];
    inv[11] = -T[0] * T[5] * T[11] + T[0] * T[7] * T[9] + T[4] * T[1] * T[11] - T[4] * T[3] * T[9] - T[8] * T[1] * T[7] + T[8] * T[3] * T[5];
    inv[12] = -T[4] * T[9] * T[14] + T[4] * T[10] * T[13] + T[8] * T[5] * T[14] - T[8] * T[6] * T[13] - T[12] * T[5] * T[10] + T[12] * T[6] * T[9];
    inv[13] = T[0] * T[9] * T[14] - T[0] * T[10] * T[13] - T[8] * T[1] * T[14] + T[8] * T[2] * T[13] + T[12] * T[1] * T[10] - T[12] * T[2] * T[9];
    inv[14] = -T[0] * T[5] * T[14] + T[0] * T[6] * T[13] + T[4] * T[1] * T[14] - T[4] * T[2] * T[13] - T[12] * T[1] * T[6] + T[12] * T[2] * T[5];
    inv[15] = T[0] * T[5] * T[10] - T[0] * T[6] * T[9] - T[4] * T[1] * T[10] + T[4] * T[2] * T[9] + T[8] * T[1] * T[6] - T[8] * T[2] * T[5];
    const float det = T[0] * inv[0] + T[1] * inv[4] + T[2] * inv[8] + T[3] * inv[12];
    if (det == 0) return;
    const float invdet = 1.0f / det;
    for (int i = 0; i < 16; i++) inv[i] *= invdet;
}

// combinedPrevTS = prev * inverse(world) in double (Gauss-Jordan, partial pivoting);
// identity when the origin moves by at most 0.001 (Node::TickVelocity, Node.cpp:82-106)
__device__ void combined_prev(const float* world, const float* prev, float* out) {
    double a[4][8];
    for (int r = 0; r < 4; ++r)
        for (int cc = 0; cc < 4; ++cc) { a[r][cc] = world[cc * 4 + r]; a[r][4 + cc] = r == cc ? 1.0 : 0.0; }
    bool singular = false;
    for (int cc = 0; cc < 4 && !singular; ++cc) {
        int p = cc;
        for (int r = cc + 1; r < 4; ++r) if (fabs(a[r][cc]) > fabs(a[p][cc])) p = r;
        if (p != cc) for (int k = 0; k < 8; ++k) { const double t = a[cc][k]; a[cc][k] = a[p][k]; a[p][k] = t; }
        const double d = a[cc][cc];
        if (d == 0.0) { singular = true; break; }
        for (int k = 0; k < 8; ++k) a[cc][k] /= d;
        for (int r = 0; r < 4; ++r) {
            if (r == cc) continue;
            const double f = a[r][cc];
            if (f != 0.0) for (int k = 0; k < 8; ++k) a[r][k] -= f * a[cc][k];
        }
    }
    if (singular) { for (int i = 0; i < 16; ++i) out[i] = (i % 5 == 0) ? 1.f : 0.f; return; }
    double res[16];
    for (int cc = 0; cc < 4; ++cc)
        for (int r = 0; r < 4; ++r) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += double(prev[k * 4 + r]) * a[k][4 + cc];
            res[cc * 4 + r] = acc;
        }
    const double tx = res[12], ty = res[13], tz = res[14];
    if (sqrt(tx * tx + ty * ty + tz * tz) > 0.001) for (int i = 0; i < 16; ++i) out[i] = float(res[i]);
    else for (int i = 0; i < 16; ++i) out[i] = (i % 5 == 0) ? 1.f : 0.f;
}

__device__ inline f3 tpoint_rm(f3 v, const float* M) {
    const f3 r = f3{M[0] * v.x + M[1] * v.y + M[2] * v.z + M[3],
                    M[4] * v.x + M[5] * v.y + M[6] * v.z + M[7],
                    M[8] * v.x + M[9] * v.y + M[10] * v.z + M[11]};
    const float w = M[12] * v.x + M[13] * v.y + M[14] * v.z + M[15];
    if (w == 1.0f) return r;
    return r * (1.f / w);
}

// Per instance: new world (column-major) from `world_in` (or keep), prev from `prev_in`
// (or the old world), then inverse, combinedPrevTS and world box.
__global__ void k_update_instances(InstDev* __restrict__ inst, const float* __restrict__ world_in,
                                   const float* __restrict__ prev_in, const MeshInfo* __restrict__ meshes,
                                   uint32_t first, uint32_t count, float4* __restrict__ ilo,
                                   float4* __restrict__ ihi) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    InstDev& I = inst[first + k];
    float world[16], prev[16];
    for (int i = 0; i < 16; ++i) {
        world[i] = world_in ? world_in[16 * k + i] : I.world[i];
        prev[i] = prev_in ? prev_in[16 * k + i] : I.world[i];
    }
    float xf[16], inv[16];
    for (int r = 0; r < 4; ++r)
        for (int cc = 0; cc < 4; ++cc) xf[r * 4 + cc] = world[cc * 4 + r];
    for (int i = 0; i < 16; ++i) inv[i] = xf[i];
    invert_rowmajor(xf, inv);
    float comb[16];
    combined_prev(world, prev, comb);
    for (int i = 0; i < 16; ++i) { I.world[i] = world[i]; I.inv[i] = inv[i]; I.combined[i] = comb[i]; }
    const MeshInfo m = meshes[I.mesh];
    f3 lo = f3{kFar, kFar, kFar}, hi = f3{-kFar, -kFar, -kFar};
    for (int j = 0; j < 8; ++j) {
        const f3 p = f3{j & 1 ? m.bmax[0] : m.bmin[0], j & 2 ? m.bmax[1] : m.bmin[1], j & 4 ? m.bmax[2] : m.bmin[2]};
        const f3 t = tpoint_rm(p, xf);
        lo = fmin3(lo, t);
        hi = fmax3(hi, t);
    }
    if (!I.active || m.tri_count == 0) { lo = f3{kEmpty, kEmpty, kEmpty}; hi = lo; }
    ilo[first + k] = make_float4(lo.x, lo.y, lo.z, 0.f);
    ihi[first + k] = make_float4(hi.x, hi.y, hi.z, 0.f);
}

// Bottom-up refit of the wide TLAS by levels (deepest first), one work-group.
// level_info: [0] number of levels, [1 ..] start of each level's nodes in order[] (deepest
// level first), as build_tlas or k_collapse_block wrote them
__global__ void k_refit_tlas(Bvh4Node* __restrict__ nodes, const uint32_t* __restrict__ order,
                             const uint32_t* __restrict__ level_info,
                             const uint32_t* __restrict__ prims, const float4* __restrict__ ilo,
                             const float4* __restrict__ ihi) {
    const uint32_t levels = level_info[0];
    const uint32_t* level_off = level_info + 1;
    for (uint32_t L = 0; L < levels; ++L) {
        for (uint32_t i = level_off[L] + threadIdx.x; i < level_off[L + 1]; i += blockDim.x) {
            Bvh4Node& n = nodes[order[i]];
            float lo[3][4], hi[3][4];
            const uint32_t ch[4] = {n.child.x, n.child.y, n.child.z, n.child.w};
            const uint32_t me[4] = {n.meta.x, n.meta.y, n.meta.z, n.meta.w};
            for (int j = 0; j < 4; ++j) {
                float l[3] = {kEmpty, kEmpty, kEmpty}, h[3] = {kEmpty, kEmpty, kEmpty};
                bool any = false;
                if (me[j] > 0u) {
                    for (uint32_t k = 0; k < me[j]; ++k) {
                        const uint32_t p = prims[ch[j] + k];
                        const float4 a = ilo[p], b = ihi[p];
                        if (a.x == kEmpty) continue;
                        if (!any) { l[0] = a.x; l[1] = a.y; l[2] = a.z; h[0] = b.x; h[1] = b.y; h[2] = b.z; any = true; }
                        else {
                            l[0] = fminf(l[0], a.x); l[1] = fminf(l[1], a.y); l[2] = fminf(l[2], a.z);
                            h[0] = fmaxf(h[0], b.x); h[1] = fmaxf(h[1], b.y); h[2] = fmaxf(h[2], b.z);
                        }
                    }
                } else if (ch[j] != 0u) {  // internal child (child 0 with meta 0 marks an empty slot)
                    const Bvh4Node& c = nodes[ch[j]];
                    const float clx[4] = {c.lox.x, c.lox.y, c.lox.z, c.lox.w}, chx[4] = {c.hix.x, c.hix.y, c.hix.z, c.hix.w};
                    const float cly[4] = {c.loy.x, c.loy.y, c.loy.z, c.loy.w}, chy[4] = {c.hiy.x, c.hiy.y, c.hiy.z, c.hiy.w};
                    const float clz[4] = {c.loz.x, c.loz.y, c.loz.z, c.loz.w}, chz[4] = {c.hiz.x, c.hiz.y, c.hiz.z, c.hiz.w};
                    for (int k = 0; k < 4; ++k) {
                        if (clx[k] == kEmpty) continue;
                        if (!any) { l[0] = clx[k]; l[1] = cly[k]; l[2] = clz[k]; h[0] = chx[k]; h[1] = chy[k]; h[2] = chz[k]; any = true; }
                        else {
                            l[0] = fminf(l[0], clx[k]); l[1] = fminf(l[1], cly[k]); l[2] = fminf(l[2], clz[k]);
                            h[0] = fmaxf(h[0], chx[k]); h[1] = fmaxf(h[1], chy[k]); h[2] = fmaxf(h[2], chz[k]);
                        }
                    }
                }
                for (int a = 0; a < 3; ++a) { lo[a][j] = l[a]; hi[a][j] = h[a]; }
            }
            n.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
            n.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
            n.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
            n.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
            n.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
            n.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
        }
        __threadfence_block();
        __syncthreads();
    }
}

// world-space emissive triangles: p = world * (v0,1), e = world * (e,0) (column-major, as
// the oracle), area and selection pdf kept from the upload
__global__ void k_update_lights(LightTri* __restrict__ lights, const float4* __restrict__ obj, uint32_t n,
                                const InstDev* __restrict__ inst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    LightTri L = lights[i];
    const InstDev& I = inst[__float_as_uint(L.e2.w)];
    const float4 a = obj[3 * i], b = obj[3 * i + 1], c = obj[3 * i + 2];
    float o[4];
    mul44v(I.world, a.x, a.y, a.z, 1.f, o);
    L.p0.x = o[0]; L.p0.y = o[1]; L.p0.z = o[2];
    mul44v(I.world, b.x, b.y, b.z, 0.f, o);
    L.e1.x = o[0]; L.e1.y = o[1]; L.e1.z = o[2];
    mul44v(I.world, c.x, c.y, c.z, 0.f, o);
    L.e2.x = o[0]; L.e2.y = o[1]; L.e2.z = o[2];
    const f3 cr = cross(f3{L.e1.x, L.e1.y, L.e1.z}, f3{L.e2.x, L.e2.y, L.e2.z});
    L.p0.w = 0.5f * length(cr);
    lights[i] = L;
}

}  // namespace

// ---------------------------------------------------------------- BLAS
int build_blas(Ctx& c, const gk_mesh* meshes, uint32_t num_meshes) {
    hipStream_t st = c.stream;
    uint64_t total64 = 0;
    for (uint32_t m = 0; m < num_meshes; ++m) total64 += meshes[m].num_triangles;
    // traversal stack entries address a leaf's first triangle with 27 bits
    if (total64 >= (uint64_t(1) << 27)) return GK_ERR_LIMIT;
    const uint32_t total = uint32_t(total64);
    c.num_tris = total;
    GK_HIP(c.tris.resize(std::max<uint32_t>(total, 1)));
    GK_HIP(c.tri_nrm.resize(3 * size_t(std::max<uint32_t>(total, 1))));
    GK_HIP(c.tri_mat.resize(std::max<uint32_t>(total, 1)));
    GK_HIP(c.tri_rec.resize(std::max<uint32_t>(total, 1)));
    bool any_uv = false, any_tan = false;
    for (uint32_t m = 0; m < num_meshes; ++m) {
        any_uv |= meshes[m].texcoords != nullptr;
        any_tan |= meshes[m].tangents != nullptr;
    }
    if (any_uv) GK_HIP(c.tri_uv.resize(3 * size_t(std::max<uint32_t>(total, 1))));
    else c.tri_uv.release();
    if (any_tan) GK_HIP(c.tri_tan.resize(3 * size_t(std::max<uint32_t>(total, 1))));
    else c.tri_tan.release();
    c.max_mesh_tris = 0;
    for (uint32_t m = 0; m < num_meshes; ++m) c.max_mesh_tris = std::max(c.max_mesh_tris, meshes[m].num_triangles);
    size_t node_cap = 0;
    for (uint32_t m = 0; m < num_meshes; ++m) node_cap += meshes[m].num_triangles + 1;
    GK_HIP(c.blas_nodes.resize(std::max<size_t>(node_cap, 1)));
    c.h_meshes.assign(num_meshes, MeshInfo{});
    DevBuf<float> d_pos, d_nrm, d_uv, d_tan;
    DevBuf<uint32_t> d_idx, d_mat, d_err;
    DevBuf<TriRec> tmp;
    GK_HIP(d_err.resize(1));
    GK_HIP(hipMemsetAsync(d_err.ptr, 0, 4, st));
    uint32_t tri_base = 0, node_base = 0;
    double build_ms = 0.0;
    hipEvent_t e0, e1;
    GK_HIP(hipEventCreate(&e0));
    GK_HIP(hipEventCreate(&e1));
    GK_HIP(hipEventRecord(e0, st));
    for (uint32_t m = 0; m < num_meshes; ++m) {
        const gk_mesh& M = meshes[m];
        MeshInfo& info = c.h_meshes[m];
        info.tri_base = tri_base;
        info.tri_count = M.num_triangles;
        info.root = node_base;
        const uint32_t nt = M.num_triangles, nv = M.num_vertices;
        if (nt == 0) {
            // empty mesh: a root whose four slots are empty
            Bvh4Node empty{};
            empty.lox = empty.hix = empty.loy = empty.hiy = empty.loz = empty.hiz =
                make_float4(kEmpty, kEmpty, kEmpty, kEmpty);
            GK_HIP(hipMemcpyAsync(c.blas_nodes.ptr + node_base, &empty, sizeof(empty), hipMemcpyHostToDevice, st));
            GK_HIP(hipStreamSynchronize(st));
            for (int k = 0; k < 3; ++k) { info.bmin[k] = kEmpty; info.bmax[k] = kEmpty; }
            node_base += 1;
            continue;
        }
        GK_HIP(d_pos.resize(3 * size_t(nv)));
        GK_HIP(d_idx.resize(3 * size_t(nt)));
        GK_HIP(hipMemcpyAsync(d_pos.ptr, M.positions, 12 * size_t(nv), hipMemcpyHostToDevice, st));
        GK_HIP(hipMemcpyAsync(d_idx.ptr, M.indices, 12 * size_t(nt), hipMemcpyHostToDevice, st));
        if (M.normals) {
            GK_HIP(d_nrm.resize(3 * size_t(nv)));
            GK_HIP(hipMemcpyAsync(d_nrm.ptr, M.normals, 12 * size_t(nv), hipMemcpyHostToDevice, st));
        }
        if (M.tri_material) {
            GK_HIP(d_mat.resize(nt));
            GK_HIP(hipMemcpyAsync(d_mat.ptr, M.tri_material, 4 * size_t(nt), hipMemcpyHostToDevice, st));
        }
        if (M.texcoords) {
            GK_HIP(d_uv.resize(2 * size_t(nv)));
            GK_HIP(hipMemcpyAsync(d_uv.ptr, M.texcoords, 8 * size_t(nv), hipMemcpyHostToDevice, st));
        }
        if (M.tangents) {
            GK_HIP(d_tan.resize(4 * size_t(nv)));
            GK_HIP(hipMemcpyAsync(d_tan.ptr, M.tangents, 16 * size_t(nv), hipMemcpyHostToDevice, st));
        }
        if (any_uv || any_tan) {
            k_prep_tex<<<nblk(nt), 256, 0, st>>>(d_idx.ptr, M.texcoords ? d_uv.ptr : nullptr,
                                                M.tangents ? d_tan.ptr : nullptr, nv, nt, tri_base,
                                                any_uv ? c.tri_uv.ptr : nullptr, any_tan ? c.tri_tan.ptr : nullptr);
            GK_HIP(hipGetLastError());
        }
        GK_HIP(tmp.resize(nt));
        GK_HIP(c.scratch.prim_lo.resize(nt));
        GK_HIP(c.scratch.prim_hi.resize(nt));
        k_prep_tris<<<nblk(nt), 256, 0, st>>>(d_pos.ptr, d_idx.ptr, M.normals ? d_nrm.ptr : nullptr,
                                             M.tri_material ? d_mat.ptr : nullptr, nv, nt, tri_base, tmp.ptr,
                                             c.scratch.prim_lo.ptr, c.scratch.prim_hi.ptr, c.tri_nrm.ptr,
                                             c.tri_mat.ptr, d_err.ptr);
        GK_HIP(hipGetLastError());
        WideResult r;
        int s = build_wide(c, c.scratch.prim_lo.ptr, c.scratch.prim_hi.ptr, nt, kMaxLeafTris, c.blas_nodes.ptr, node_base,
                           tri_base, nullptr, r, true, nullptr, nullptr);
        if (s != GK_OK) return s;
        k_reorder_tris<<<nblk(nt), 256, 0, st>>>(tmp.ptr, r.vals, nt, c.tris.ptr + tri_base, c.tri_rec.ptr + tri_base);
        GK_HIP(hipGetLastError());
        // mesh bounds = union of the root's children
        Bvh4Node root;
        GK_HIP(hipMemcpyAsync(&root, c.blas_nodes.ptr + node_base, sizeof(root), hipMemcpyDeviceToHost, st));
        GK_HIP(hipStreamSynchronize(st));
        const float lx[4] = {root.lox.x, root.lox.y, root.lox.z, root.lox.w}, hx[4] = {root.hix.x, root.hix.y, root.hix.z, root.hix.w};
        const float ly[4] = {root.loy.x, root.loy.y, root.loy.z, root.loy.w}, hy[4] = {root.hiy.x, root.hiy.y, root.hiy.z, root.hiy.w};
        const float lz[4] = {root.loz.x, root.loz.y, root.loz.z, root.loz.w}, hz[4] = {root.hiz.x, root.hiz.y, root.hiz.z, root.hiz.w};
        float bmin[3] = {kFar, kFar, kFar}, bmax[3] = {-kFar, -kFar, -kFar};
        for (int j = 0; j < 4; ++j) {
            if (lx[j] == kEmpty) continue;
            bmin[0] = std::min(bmin[0], lx[j]); bmin[1] = std::min(bmin[1], ly[j]); bmin[2] = std::min(bmin[2], lz[j]);
            bmax[0] = std::max(bmax[0], hx[j]); bmax[1] = std::max(bmax[1], hy[j]); bmax[2] = std::max(bmax[2], hz[j]);
        }
        for (int k = 0; k < 3; ++k) { info.bmin[k] = bmin[k]; info.bmax[k] = bmax[k]; }
        info.depth = r.depth;
        if (m == 0) c.bvh_stats.sah_cost = r.sah;
        node_base += r.nodes;
        tri_base += nt;
    }
    GK_HIP(hipEventRecord(e1, st));
    GK_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GK_HIP(hipEventElapsedTime(&ms, e0, e1));
    build_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    uint32_t err = 0;
    GK_HIP(hipMemcpy(&err, d_err.ptr, 4, hipMemcpyDeviceToHost));
    if (err) return GK_ERR_ARG;
    GK_HIP(c.meshes.resize(std::max<uint32_t>(num_meshes, 1)));
    GK_HIP(hipMemcpyAsync(c.meshes.ptr, c.h_meshes.data(), sizeof(MeshInfo) * num_meshes, hipMemcpyHostToDevice, st));
    GK_HIP(hipStreamSynchronize(st));
    c.bvh_stats.blas_count = num_meshes;
    c.bvh_stats.blas_nodes = node_base;
    c.bvh_stats.blas_bytes = uint64_t(node_base) * sizeof(Bvh4Node) + uint64_t(total) * sizeof(TriRec);
    c.bvh_stats.build_ms = build_ms;
    return GK_OK;
}

// ---------------------------------------------------------------- TLAS
int build_tlas(Ctx& c) {
    hipStream_t st = c.stream;
    const uint32_t n = uint32_t(c.h_instances.size());
    // instance records (world, prev from the host structs)
    std::vector<InstDev> h(n);
    std::vector<float> worlds(16 * size_t(n)), prevs(16 * size_t(n));
    std::vector<uint32_t> proxy_inst;
    for (uint32_t i = 0; i < n; ++i) {
        const gk_instance& g = c.h_instances[i];
        InstDev& d = h[i];
        memset(&d, 0, sizeof(d));
        memcpy(d.world, g.world, 64);
        d.mesh = g.mesh_id;
        d.instance_id = g.instance_id;
        d.active = (g.visible && !g.nort) ? 1u : 0u;
        d.use_remap = g.use_remap;
        d.mat_offset = g.mat_offset;
        d.tri_base = c.h_meshes[g.mesh_id].tri_base;
        d.root = c.h_meshes[g.mesh_id].root;
        // one visibility-buffer proxy per section of the mesh, in instance order
        d.proxy_base = uint32_t(proxy_inst.size());
        const uint32_t sections = (c.h_meshes[g.mesh_id].tri_count + kSectionTris - 1) / kSectionTris;
        proxy_inst.insert(proxy_inst.end(), sections, i);
        memcpy(d.mat_id, g.mat_id, 64);
        memcpy(&worlds[16 * size_t(i)], g.world, 64);
        memcpy(&prevs[16 * size_t(i)], g.prev_world, 64);
    }
    GK_HIP(c.inst.resize(std::max<uint32_t>(n, 1)));
    GK_HIP(c.inst_lo.resize(std::max<uint32_t>(n, 1)));
    GK_HIP(c.inst_hi.resize(std::max<uint32_t>(n, 1)));
    GK_HIP(c.inst_world_in.resize(32 * size_t(std::max<uint32_t>(n, 1))));
    c.num_proxies = uint32_t(proxy_inst.size());
    GK_HIP(c.proxy_inst.resize(std::max<size_t>(proxy_inst.size(), 1)));
    if (!proxy_inst.empty())
        GK_HIP(hipMemcpyAsync(c.proxy_inst.ptr, proxy_inst.data(), 4 * proxy_inst.size(), hipMemcpyHostToDevice, st));
    if (n == 0) return GK_OK;
    GK_HIP(hipMemcpyAsync(c.inst.ptr, h.data(), sizeof(InstDev) * n, hipMemcpyHostToDevice, st));
    GK_HIP(hipMemcpyAsync(c.inst_world_in.ptr, worlds.data(), 64 * size_t(n), hipMemcpyHostToDevice, st));
    GK_HIP(hipMemcpyAsync(c.inst_world_in.ptr + 16 * size_t(n), prevs.data(), 64 * size_t(n), hipMemcpyHostToDevice, st));
    k_update_instances<<<nblk(n, 128), 128, 0, st>>>(c.inst.ptr, c.inst_world_in.ptr, c.inst_world_in.ptr + 16 * size_t(n),
                                               c.meshes.ptr, 0, n, c.inst_lo.ptr, c.inst_hi.ptr);
    GK_HIP(hipGetLastError());
    GK_HIP(c.tlas_nodes.resize(n + 1));
    GK_HIP(c.tlas_prims.resize(n));
    DevBuf<uint32_t> depth;
    GK_HIP(depth.resize(n + 1));
    WideResult r;
    hipEvent_t e0, e1;
    GK_HIP(hipEventCreate(&e0));
    GK_HIP(hipEventCreate(&e1));
    GK_HIP(hipEventRecord(e0, st));
    // one instance per TLAS leaf: each instance's box is tested before the ray enters it, which
    // costs less than entering (4096 bricks at 1080p: 5.69 -> 5.54 ms per frame with 2 per leaf)
    int s = build_wide(c, c.inst_lo.ptr, c.inst_hi.ptr, n, 1, c.tlas_nodes.ptr, 0, 0, depth.ptr, r, false, nullptr,
                       nullptr);
    if (s != GK_OK) return s;
    k_copy_u32<<<nblk(n), 256, 0, st>>>(r.vals, c.tlas_prims.ptr, n);
    GK_HIP(hipEventRecord(e1, st));
    GK_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    GK_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c.tlas_node_count = r.nodes;
    c.tlas_depth = r.depth;
    // refit order: nodes grouped by depth, deepest level first
    std::vector<uint32_t> hd(r.nodes);
    GK_HIP(hipMemcpy(hd.data(), depth.ptr, 4 * size_t(r.nodes), hipMemcpyDeviceToHost));
    uint32_t maxd = 0;
    for (uint32_t d : hd) maxd = std::max(maxd, d);
    std::vector<uint32_t> order;
    order.reserve(r.nodes);
    c.tlas_level_offsets.assign(1, 0u);
    for (int lv = int(maxd); lv >= 0; --lv) {
        for (uint32_t i = 0; i < r.nodes; ++i) if (hd[i] == uint32_t(lv)) order.push_back(i);
        c.tlas_level_offsets.push_back(uint32_t(order.size()));
    }
    // sized for any rebuild of the same instances (at most n wide nodes, kCollapseMaxLevels levels)
    GK_HIP(c.tlas_refit_order.resize(std::max<size_t>(order.size(), n)));
    GK_HIP(c.tlas_level_info.resize(kCollapseMaxLevels + 2));
    GK_HIP(hipMemcpy(c.tlas_refit_order.ptr, order.data(), 4 * order.size(), hipMemcpyHostToDevice));
    std::vector<uint32_t> info(1, uint32_t(c.tlas_level_offsets.size()) - 1u);
    info.insert(info.end(), c.tlas_level_offsets.begin(), c.tlas_level_offsets.end());
    GK_HIP(hipMemcpy(c.tlas_level_info.ptr, info.data(), 4 * info.size(), hipMemcpyHostToDevice));
    c.bvh_stats.tlas_nodes = r.nodes;
    c.bvh_stats.tlas_bytes = uint64_t(r.nodes) * sizeof(Bvh4Node);
    c.bvh_stats.tlas_ms = ms;
    c.tlas_dirty = false;
    return GK_OK;
}

// SAH cost of the TLAS as it stands: sum over wide nodes of the node's box area plus, for
// leaf children, instance count x child box area, over the root box area (C_TRAV = C_INT = 1).
// One work-group; result in *out.
__global__ __launch_bounds__(256) void k_tlas_sah(const Bvh4Node* __restrict__ nodes, const uint32_t* __restrict__ order,
                                                   const uint32_t* __restrict__ level_info, double* __restrict__ out) {
    __shared__ double s_sum[256];
    const uint32_t levels = level_info[0];
    const uint32_t nn = level_info[1 + levels];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < nn; i += blockDim.x) {
        const Bvh4Node& n = nodes[order[i]];
        const float lx[4] = {n.lox.x, n.lox.y, n.lox.z, n.lox.w}, hx[4] = {n.hix.x, n.hix.y, n.hix.z, n.hix.w};
        const float ly[4] = {n.loy.x, n.loy.y, n.loy.z, n.loy.w}, hy[4] = {n.hiy.x, n.hiy.y, n.hiy.z, n.hiy.w};
        const float lz[4] = {n.loz.x, n.loz.y, n.loz.z, n.loz.w}, hz[4] = {n.hiz.x, n.hiz.y, n.hiz.z, n.hiz.w};
        const uint32_t me[4] = {n.meta.x, n.meta.y, n.meta.z, n.meta.w};
        float3 ulo = make_float3(3e38f, 3e38f, 3e38f), uhi = make_float3(-3e38f, -3e38f, -3e38f);
        for (int j = 0; j < 4; ++j) {
            if (lx[j] == kEmpty) continue;
            ulo = make_float3(fminf(ulo.x, lx[j]), fminf(ulo.y, ly[j]), fminf(ulo.z, lz[j]));
            uhi = make_float3(fmaxf(uhi.x, hx[j]), fmaxf(uhi.y, hy[j]), fmaxf(uhi.z, hz[j]));
            if (me[j] > 0u) {
                const double dx = hx[j] - lx[j], dy = hy[j] - ly[j], dz = hz[j] - lz[j];
                acc += double(me[j]) * 2.0 * (dx * dy + dy * dz + dz * dx);
            }
        }
        if (uhi.x >= ulo.x) {
            const double dx = uhi.x - ulo.x, dy = uhi.y - ulo.y, dz = uhi.z - ulo.z;
            const double a = 2.0 * (dx * dy + dy * dz + dz * dx);
            acc += a;
        }
    }
    s_sum[threadIdx.x] = acc;
    __syncthreads();
    for (uint32_t w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_sum[threadIdx.x] += s_sum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const Bvh4Node& r = nodes[0];
        const float lx[4] = {r.lox.x, r.lox.y, r.lox.z, r.lox.w}, hx[4] = {r.hix.x, r.hix.y, r.hix.z, r.hix.w};
        const float ly[4] = {r.loy.x, r.loy.y, r.loy.z, r.loy.w}, hy[4] = {r.hiy.x, r.hiy.y, r.hiy.z, r.hiy.w};
        const float lz[4] = {r.loz.x, r.loz.y, r.loz.z, r.loz.w}, hz[4] = {r.hiz.x, r.hiz.y, r.hiz.z, r.hiz.w};
        double l[3] = {3e38, 3e38, 3e38}, h[3] = {-3e38, -3e38, -3e38};
        for (int j = 0; j < 4; ++j) {
            if (lx[j] == kEmpty) continue;
            l[0] = fmin(l[0], lx[j]); l[1] = fmin(l[1], ly[j]); l[2] = fmin(l[2], lz[j]);
            h[0] = fmax(h[0], hx[j]); h[1] = fmax(h[1], hy[j]); h[2] = fmax(h[2], hz[j]);
        }
        const double dx = h[0] - l[0], dy = h[1] - l[1], dz = h[2] - l[2];
        const double root = 2.0 * (dx * dy + dy * dz + dz * dx);
        *out = root > 0.0 ? s_sum[0] / root : 0.0;
    }
}

int tlas_sah(Ctx& c, double& out) {
    out = 0.0;
    if (c.h_instances.empty() || c.tlas_node_count == 0) return GK_OK;
    DevBuf<double> d;
    GK_HIP(d.resize(1));
    k_tlas_sah<<<1, 256, 0, c.stream>>>(c.tlas_nodes.ptr, c.tlas_refit_order.ptr, c.tlas_level_info.ptr, d.ptr);
    GK_HIP(hipGetLastError());
    GK_HIP(hipMemcpyAsync(&out, d.ptr, sizeof(double), hipMemcpyDeviceToHost, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    return GK_OK;
}

// A new TLAS over the instances' current world boxes, built on the device without waiting
// for the host (LBVH, SAH leaf collapse, then the single-work-group 4-wide collapse, which
// also writes the refit order), as the reference rebuilds its TLAS every frame
// (TopLevelAccelerationStructure.cpp:73-81). Buffers keep the sizes of the scene's build.
int rebuild_tlas(Ctx& c) {
    const uint32_t n = uint32_t(c.h_instances.size());
    if (n == 0 || c.tlas_node_count == 0) return GK_OK;
    WideResult r;
    const int s = build_wide(c, c.inst_lo.ptr, c.inst_hi.ptr, n, 1, c.tlas_nodes.ptr, 0, 0, nullptr, r, false,
                             c.tlas_refit_order.ptr, c.tlas_level_info.ptr);
    if (s != GK_OK) return s;
    k_copy_u32<<<nblk(n), 256, 0, c.stream>>>(r.vals, c.tlas_prims.ptr, n);
    GK_HIP(hipGetLastError());
    if (c.num_lights) {
        k_update_lights<<<nblk(c.num_lights, 128), 128, 0, c.stream>>>(c.lights.ptr, c.light_obj.ptr, c.num_lights,
                                                                  c.inst.ptr);
        GK_HIP(hipGetLastError());
    }
    c.tlas_dirty = false;
    return GK_OK;
}

// new world matrices are in c.inst_world_in[0 .. 16*count)
int update_instances_device(Ctx& c, uint32_t first, uint32_t count) {
    if (count == 0) return GK_OK;
    k_update_instances<<<nblk(count, 128), 128, 0, c.stream>>>(c.inst.ptr, c.inst_world_in.ptr, nullptr, c.meshes.ptr,
                                                         first, count, c.inst_lo.ptr, c.inst_hi.ptr);
    GK_HIP(hipGetLastError());
    c.tlas_dirty = true;
    return GK_OK;
}

int update_lights(Ctx& c) {
    if (c.num_lights == 0) return GK_OK;
    k_update_lights<<<nblk(c.num_lights, 128), 128, 0, c.stream>>>(c.lights.ptr, c.light_obj.ptr, c.num_lights,
                                                                      c.inst.ptr);
    GK_HIP(hipGetLastError());
    return GK_OK;
}

int refit_tlas(Ctx& c) {
    if (c.h_instances.empty() || c.tlas_node_count == 0) return GK_OK;
    k_refit_tlas<<<1, 1024, 0, c.stream>>>(c.tlas_nodes.ptr, c.tlas_refit_order.ptr, c.tlas_level_info.ptr,
                                          c.tlas_prims.ptr, c.inst_lo.ptr, c.inst_hi.ptr);
    GK_HIP(hipGetLastError());
    if (c.num_lights) {
        k_update_lights<<<nblk(c.num_lights, 128), 128, 0, c.stream>>>(c.lights.ptr, c.light_obj.ptr, c.num_lights,
                                                                  c.inst.ptr);
        GK_HIP(hipGetLastError());
    }
    c.tlas_dirty = false;
    return GK_OK;
}

}  // namespace gk

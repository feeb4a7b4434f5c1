// oracle_bvh.cpp - CPU acceleration structures of the oracle (test infrastructure only).
//
// BLAS: binned SAH with 8 bins, C_TRAV = C_INT = 1, restating tinybvh BVH::Build
//       (tiny_bvh.h:1674-1767) as used by FCPUAccelerationStructure::InitBVH
//       (CPUAccelerationStructure.cpp:172-235). tests/ check node count and SAH cost
//       against the compiled tinybvh.
// TLAS: the same builder over world-space instance boxes (BLASInstance::Update,
//       tiny_bvh.h:6718-6731; BVH::Build(BLASInstance*), tiny_bvh.h:1565-1603).
// Traversal returns the closest hit with ties broken towards the lower (inst, prim):
//       the answer equals a brute-force loop over all triangles, independent of the
//       tree. Box tests are widened by a relative 2^-16 so they do not cull a triangle
//       that the Moeller-Trumbore test accepts.
#include <algorithm>
#include <limits>

#include "oracle_core.hpp"

namespace orc {

namespace {

constexpr int BINS = 8;
constexpr float FAR = 1e30f;

struct Frag { F3 bmin, bmax; };

float half_area(F3 e) { return e.x < -FAR ? 0.f : (e.x * e.y + e.y * e.z + e.z * e.x); }

int bin_of(float c, float nmin, float rpd) {
    const float f = (c - nmin) * rpd;
    if (!(f >= 0.f)) return 0;  // NaN (zero extent) and negatives clamp to bin 0
    if (f >= float(BINS - 1)) return BINS - 1;
    return int(f);
}

// Binned SAH build over fragments (restating tiny_bvh.h:1674-1767).
void build_bvh2(const std::vector<Frag>& frag, std::vector<Mesh::Node>& nodes,
                std::vector<uint32_t>& prim) {
    const uint32_t n = uint32_t(frag.size());
    nodes.assign(std::max<uint32_t>(2 * n, 2), Mesh::Node{});
    prim.resize(n);
    Mesh::Node& root = nodes[0];
    root.left_first = 0;
    root.count = n;
    root.bmin = F3(FAR, FAR, FAR);
    root.bmax = F3(-FAR, -FAR, -FAR);
    for (uint32_t i = 0; i < n; ++i) {
        prim[i] = i;
        root.bmin = vmin(root.bmin, frag[i].bmin);
        root.bmax = vmax(root.bmax, frag[i].bmax);
    }
    if (n == 0) return;
    uint32_t next = 2;
    const F3 min_dim = (root.bmax - root.bmin) * 1e-20f;
    std::vector<uint32_t> task;
    uint32_t node_idx = 0;
    for (;;) {
        for (;;) {
            Mesh::Node& node = nodes[node_idx];
            F3 bin_min[3][BINS], bin_max[3][BINS];
            uint32_t cnt[3][BINS] = {};
            for (int a = 0; a < 3; ++a)
                for (int i = 0; i < BINS; ++i) {
                    bin_min[a][i] = F3(FAR, FAR, FAR);
                    bin_max[a][i] = F3(-FAR, -FAR, -FAR);
                }
            const F3 ext = node.bmax - node.bmin;
            const F3 rpd(float(BINS) / ext.x, float(BINS) / ext.y, float(BINS) / ext.z);
            const F3 nmin = node.bmin;
            for (uint32_t i = 0; i < node.count; ++i) {
                const Frag& f = frag[prim[node.left_first + i]];
                const F3 c = (f.bmin + f.bmax) * 0.5f;
                for (int a = 0; a < 3; ++a) {
                    const int b = bin_of(c[a], nmin[a], rpd[a]);
                    bin_min[a][b] = vmin(bin_min[a][b], f.bmin);
                    bin_max[a][b] = vmax(bin_max[a][b], f.bmax);
                    cnt[a][b]++;
                }
            }
            const float rsav = 1.0f / half_area(ext);
            float split_cost = FAR;
            int best_axis = 0, best_pos = 0;
            F3 bl_min, bl_max, br_min, br_max;
            for (int a = 0; a < 3; ++a) {
                if (!(ext[a] > min_dim[a])) continue;
                F3 lmin[BINS - 1], lmax[BINS - 1], rmin[BINS - 1], rmax[BINS - 1];
                float anl[BINS - 1], anr[BINS - 1];
                F3 l1(FAR, FAR, FAR), l2(-FAR, -FAR, -FAR), r1(FAR, FAR, FAR), r2(-FAR, -FAR, -FAR);
                uint32_t ln = 0, rn = 0;
                for (int i = 0; i < BINS - 1; ++i) {
                    lmin[i] = l1 = vmin(l1, bin_min[a][i]);
                    rmin[BINS - 2 - i] = r1 = vmin(r1, bin_min[a][BINS - 1 - i]);
                    lmax[i] = l2 = vmax(l2, bin_max[a][i]);
                    rmax[BINS - 2 - i] = r2 = vmax(r2, bin_max[a][BINS - 1 - i]);
                    ln += cnt[a][i];
                    rn += cnt[a][BINS - 1 - i];
                    anl[i] = ln == 0 ? FAR : half_area(l2 - l1) * float(ln);
                    anr[BINS - 2 - i] = rn == 0 ? FAR : half_area(r2 - r1) * float(rn);
                }
                for (int i = 0; i < BINS - 1; ++i) {
                    const float C = 1.0f + rsav * 1.0f * (anl[i] + anr[i]);
                    if (C < split_cost) {
                        split_cost = C; best_axis = a; best_pos = i;
                        bl_min = lmin[i]; bl_max = lmax[i]; br_min = rmin[i]; br_max = rmax[i];
                    }
                }
            }
            if (split_cost >= float(node.count)) break;
            uint32_t j = node.left_first + node.count, src = node.left_first;
            for (uint32_t i = 0; i < node.count; ++i) {
                const Frag& f = frag[prim[src]];
                const float c = (f.bmin[best_axis] + f.bmax[best_axis]) * 0.5f;
                if (bin_of(c, nmin[best_axis], rpd[best_axis]) <= best_pos) src++;
                else std::swap(prim[src], prim[--j]);
            }
            const uint32_t lc = src - node.left_first, rc = node.count - lc;
            if (lc == 0 || rc == 0) break;
            const uint32_t li = next++, ri = next++;
            nodes[li] = Mesh::Node{bl_min, node.left_first, bl_max, lc};
            nodes[ri] = Mesh::Node{br_min, j, br_max, rc};
            Mesh::Node& nn = nodes[node_idx];
            nn.left_first = li;
            nn.count = 0;
            task.push_back(ri);
            node_idx = li;
        }
        if (task.empty()) break;
        node_idx = task.back();
        task.pop_back();
    }
    nodes.resize(next);
}

// Row-major inverse of a row-major 4x4 by cofactors; same expansion and order as
// BLASInstance::InvertTransform (tiny_bvh.h:6734-6756).
void invert_rowmajor(const float* T, float* inv) {
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

// tinybvh_transform_point / _vector (tiny_bvh.h:396-409), row-major T
template <class T>
V3<T> tpoint(V3<T> v, const float* M) {
    const V3<T> r(T(M[0]) * v.x + T(M[1]) * v.y + T(M[2]) * v.z + T(M[3]),
                  T(M[4]) * v.x + T(M[5]) * v.y + T(M[6]) * v.z + T(M[7]),
                  T(M[8]) * v.x + T(M[9]) * v.y + T(M[10]) * v.z + T(M[11]));
    const T w = T(M[12]) * v.x + T(M[13]) * v.y + T(M[14]) * v.z + T(M[15]);
    if (w == T(1)) return r;
    return r * (T(1) / w);
}
template <class T>
V3<T> tvector(V3<T> v, const float* M) {
    return V3<T>(T(M[0]) * v.x + T(M[1]) * v.y + T(M[2]) * v.z,
                 T(M[4]) * v.x + T(M[5]) * v.y + T(M[6]) * v.z,
                 T(M[8]) * v.x + T(M[9]) * v.y + T(M[10]) * v.z);
}

template <class T>
T safercp(T x) { return x > T(1e-12f) ? T(1) / x : (x < T(-1e-12f) ? T(1) / x : T(FAR)); }

// conservative slab test: returns entry distance or +inf
template <class T>
T slab(const V3<T>& O, const V3<T>& rD, const F3& bmin, const F3& bmax, T tmin, T tmax) {
    T t1 = (T(bmin.x) - O.x) * rD.x, t2 = (T(bmax.x) - O.x) * rD.x;
    T lo = std::min(t1, t2), hi = std::max(t1, t2);
    t1 = (T(bmin.y) - O.y) * rD.y; t2 = (T(bmax.y) - O.y) * rD.y;
    lo = std::max(lo, std::min(t1, t2)); hi = std::min(hi, std::max(t1, t2));
    t1 = (T(bmin.z) - O.z) * rD.z; t2 = (T(bmax.z) - O.z) * rD.z;
    lo = std::max(lo, std::min(t1, t2)); hi = std::min(hi, std::max(t1, t2));
    hi = hi + std::fabs(hi) * T(1.52587890625e-05);  // widen by 2^-16 relative
    lo = lo - std::fabs(lo) * T(1.52587890625e-05);
    if (hi >= lo && lo <= tmax && hi >= tmin) return lo;
    return std::numeric_limits<T>::infinity();
}

inline bool key_less(uint32_t ia, uint32_t pa, uint32_t ib, uint32_t pb) {
    return ia < ib || (ia == ib && pa < pb);
}

template <class T>
void blas_closest(const Mesh& m, V3<T> O, V3<T> D, T tmin, T tmax, uint32_t inst, HitT<T>& best,
                  bool& found) {
    if (m.nodes.empty()) return;
    const V3<T> rD(safercp(D.x), safercp(D.y), safercp(D.z));
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Mesh::Node& nd = m.nodes[stack[--sp]];
        const T limit = found ? best.t : tmax;
        if (slab(O, rD, nd.bmin, nd.bmax, tmin, limit) == std::numeric_limits<T>::infinity()) continue;
        if (nd.count) {
            for (uint32_t i = 0; i < nd.count; ++i) {
                const uint32_t p = m.prim_idx[nd.left_first + i];
                V3<T> v0(T(m.v0[p].x), T(m.v0[p].y), T(m.v0[p].z));
                V3<T> e1(T(m.e1[p].x), T(m.e1[p].y), T(m.e1[p].z));
                V3<T> e2(T(m.e2[p].x), T(m.e2[p].y), T(m.e2[p].z));
                T t, u, v;
                if (!intersect_tri(O, D, v0, e1, e2, tmin, tmax, t, u, v)) continue;
                if (!found || t < best.t || (t == best.t && key_less(inst, p, best.inst, best.prim))) {
                    best.t = t; best.u = u; best.v = v; best.prim = p; best.inst = inst;
                    found = true;
                }
            }
        } else {
            stack[sp++] = nd.left_first;
            stack[sp++] = nd.left_first + 1;
        }
    }
}

template <class T>
bool blas_any(const Mesh& m, V3<T> O, V3<T> D, T tmin, T tmax) {
    if (m.nodes.empty()) return false;
    const V3<T> rD(safercp(D.x), safercp(D.y), safercp(D.z));
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Mesh::Node& nd = m.nodes[stack[--sp]];
        if (slab(O, rD, nd.bmin, nd.bmax, tmin, tmax) == std::numeric_limits<T>::infinity()) continue;
        if (nd.count) {
            for (uint32_t i = 0; i < nd.count; ++i) {
                const uint32_t p = m.prim_idx[nd.left_first + i];
                V3<T> v0(T(m.v0[p].x), T(m.v0[p].y), T(m.v0[p].z));
                V3<T> e1(T(m.e1[p].x), T(m.e1[p].y), T(m.e1[p].z));
                V3<T> e2(T(m.e2[p].x), T(m.e2[p].y), T(m.e2[p].z));
                T t, u, v;
                if (intersect_tri(O, D, v0, e1, e2, tmin, tmax, t, u, v)) return true;
            }
        } else {
            stack[sp++] = nd.left_first;
            stack[sp++] = nd.left_first + 1;
        }
    }
    return false;
}

bool inst_active(const Instance& I) { return I.visible && !I.nort; }

}  // namespace

void scene_update_instances(Scene& sc) {
    std::vector<Frag> frag(sc.inst.size());
    for (size_t i = 0; i < sc.inst.size(); ++i) {
        Instance& I = sc.inst[i];
        // row-major copy of the column-major world matrix
        for (int r = 0; r < 4; ++r)
            for (int c = 0; c < 4; ++c) I.xf[r * 4 + c] = I.world[c * 4 + r];
        std::memcpy(I.inv, I.xf, sizeof(I.inv));
        invert_rowmajor(I.xf, I.inv);
        const Mesh& m = sc.meshes[I.mesh_id];
        F3 lo(FAR, FAR, FAR), hi(-FAR, -FAR, -FAR);
        for (int j = 0; j < 8; ++j) {
            const F3 p(j & 1 ? m.bmax.x : m.bmin.x, j & 2 ? m.bmax.y : m.bmin.y, j & 4 ? m.bmax.z : m.bmin.z);
            const F3 t = tpoint<float>(p, I.xf);
            lo = vmin(lo, t);
            hi = vmax(hi, t);
        }
        if (!inst_active(I) || m.v0.empty()) {  // inactive instances: empty box far away
            lo = F3(FAR, FAR, FAR);
            hi = F3(FAR, FAR, FAR);
        }
        I.bmin = lo;
        I.bmax = hi;
        frag[i] = Frag{lo, hi};
    }
    build_bvh2(frag, sc.tnodes, sc.tprim);
}

void scene_build(Scene& sc) {
    for (Mesh& m : sc.meshes) {
        const size_t n = m.v0.size();
        std::vector<Frag> frag(n);
        for (size_t i = 0; i < n; ++i) {
            const F3 a = m.v0[i], b = m.v0[i] + m.e1[i], c = m.v0[i] + m.e2[i];
            frag[i] = Frag{vmin(vmin(a, b), c), vmax(vmax(a, b), c)};
        }
        build_bvh2(frag, m.nodes, m.prim_idx);
        if (n) { m.bmin = m.nodes[0].bmin; m.bmax = m.nodes[0].bmax; }
    }
    scene_update_instances(sc);
}

template <class T>
bool trace_closest_t(const Scene& sc, V3<T> O, V3<T> D, T tmin, T tmax, HitT<T>& h) {
    bool found = false;
    if (sc.tnodes.empty()) return false;
    const V3<T> rD(safercp(D.x), safercp(D.y), safercp(D.z));
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Mesh::Node& nd = sc.tnodes[stack[--sp]];
        const T limit = found ? h.t : tmax;
        if (slab(O, rD, nd.bmin, nd.bmax, tmin, limit) == std::numeric_limits<T>::infinity()) continue;
        if (nd.count) {
            for (uint32_t i = 0; i < nd.count; ++i) {
                const uint32_t ii = sc.tprim[nd.left_first + i];
                const Instance& I = sc.inst[ii];
                if (!inst_active(I)) continue;
                const V3<T> Oo = tpoint<T>(O, I.inv);
                const V3<T> Do = tvector<T>(D, I.inv);
                blas_closest(sc.meshes[I.mesh_id], Oo, Do, tmin, tmax, ii, h, found);
            }
        } else {
            stack[sp++] = nd.left_first;
            stack[sp++] = nd.left_first + 1;
        }
    }
    return found;
}

template <class T>
bool trace_any_t(const Scene& sc, V3<T> O, V3<T> D, T tmin, T tmax) {
    if (sc.tnodes.empty()) return false;
    const V3<T> rD(safercp(D.x), safercp(D.y), safercp(D.z));
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Mesh::Node& nd = sc.tnodes[stack[--sp]];
        if (slab(O, rD, nd.bmin, nd.bmax, tmin, tmax) == std::numeric_limits<T>::infinity()) continue;
        if (nd.count) {
            for (uint32_t i = 0; i < nd.count; ++i) {
                const Instance& I = sc.inst[sc.tprim[nd.left_first + i]];
                if (!inst_active(I)) continue;
                if (blas_any(sc.meshes[I.mesh_id], tpoint<T>(O, I.inv), tvector<T>(D, I.inv), tmin, tmax))
                    return true;
            }
        } else {
            stack[sp++] = nd.left_first;
            stack[sp++] = nd.left_first + 1;
        }
    }
    return false;
}

template bool trace_closest_t<float>(const Scene&, F3, F3, float, float, HitT<float>&);
template bool trace_closest_t<double>(const Scene&, D3, D3, double, double, HitT<double>&);
template bool trace_any_t<float>(const Scene&, F3, F3, float, float);
template bool trace_any_t<double>(const Scene&, D3, D3, double, double);

bool trace_closest(const Scene& sc, F3 O, F3 D, float tmin, float tmax, Hit& h) {
    HitT<float> ht{};
    if (!trace_closest_t<float>(sc, O, D, tmin, tmax, ht)) return false;
    h.t = ht.t; h.u = ht.u; h.v = ht.v; h.prim = ht.prim; h.inst = ht.inst;
    return true;
}

bool trace_any(const Scene& sc, F3 O, F3 D, float tmin, float tmax) {
    return trace_any_t<float>(sc, O, D, tmin, tmax);
}

// float SAH cost of a BLAS exactly as BVH::SAHCost (tiny_bvh.h:1396-1404)
float blas_sah_cost(const Mesh& m, uint32_t idx) {
    const Mesh::Node& n = m.nodes[idx];
    const float sa = half_area(n.bmax - n.bmin);
    if (n.count) return sa * float(n.count);
    const float cost = sa + blas_sah_cost(m, n.left_first) + blas_sah_cost(m, n.left_first + 1);
    return idx == 0 ? cost / sa : cost;
}

uint32_t blas_node_count(const Mesh& m) {
    if (m.nodes.empty()) return 0;
    uint32_t cnt = 0, stack[128];
    int sp = 0;
    stack[sp++] = 0;
    while (sp) {
        const Mesh::Node& n = m.nodes[stack[--sp]];
        cnt++;
        if (!n.count) { stack[sp++] = n.left_first; stack[sp++] = n.left_first + 1; }
    }
    return cnt;
}

}  // namespace orc

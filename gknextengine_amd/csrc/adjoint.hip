// adjoint.hip - differentiable PT_MIS and its path-replay adjoint.
//
// No counterpart in the reference. The primal is the PT_MIS estimator of wavefront.hip
// (bounded-VNDF GGX + Lambert, NEE to emissive triangles, power-heuristic MIS) without
// Russian roulette, one thread per pixel, `samples` paths per pixel seeded with pcg4d
// (x, y, total_frames + s, global_seed); each path starts at a uniformly drawn position
// inside its pixel (box filter), so image edges and grid-aligned geometry edges are
// antialiased rather than sampled at one fixed point.
//
// Gradients use detached sampling: sampled directions, lobe choices, pdfs and MIS weights
// are held fixed, and the derivative of a path contribution is taken through the BSDF
// values it multiplies (f/p with p detached). For a path whose contributions are C_j
// (emission found at vertex j, or NEE at vertex j) this gives
//     dC/dtheta = sum_k d log f_k / dtheta * (sum_{j > k} C_j) + sum_j dC_nee_j / dtheta
// where f_k is the lobe sampled at vertex k. The adjoint replays each path from its seed
// twice: pass 1 computes the path's total radiance Ltot, pass 2 accumulates the prefix P
// and uses Ltot - P as the suffix radiance at every vertex (path-replay backpropagation,
// Vicini, Speierer & Jakob 2021). Nothing is stored per vertex.
//
// Parameters per material (5 floats): albedo r, g, b, roughness (fuzziness), metallic.
// Gradient contributions are summed per work-group in LDS and the per-group partials are
// added in a fixed order by a second kernel.
#include <hip/hip_runtime.h>

#include <cstring>

#include "gk_ptmis.hpp"

namespace gk {

namespace {

using namespace ptmis;

struct DiffK {
    float mvi[16], pi[16];
    float sky[4], sky_rot;
    uint32_t W, H, frame0, seed, spp, bounces, has_sky;
};

// camera ray through the image position (fx, fy) in pixels
__device__ inline void camera_ray(const DiffK& K, float fx, float fy, f3& O, f3& D) {
    const float ux = (fx / float(K.W)) * 2.0f - 1.0f;
    const float uy = (fy / float(K.H)) * 2.0f - 1.0f;
    float o[4], t[4], d[4];
    mul44v(K.mvi, 0.f, 0.f, 0.f, 1.f, o);
    mul44v(K.pi, ux, uy, 1.f, 1.f, t);
    const f3 tn = normalize(f3{t[0], t[1], t[2]});
    mul44v(K.mvi, tn.x, tn.y, tn.z, 0.f, d);
    O = f3{o[0], o[1], o[2]};
    D = normalize(f3{d[0], d[1], d[2]});
}

// pieces of the specular lobe needed for its derivatives
struct SpecParts {
    f3 F;        // Fresnel per channel
    float K;     // D * G / (4 cos_o cos_i)
    float x5;    // (1 - wo.h)^5
    float dlogK_dr;  // d log K / d roughness
};

__device__ inline SpecParts spec_parts(const Bsdf& b, const gk_material& m, f3 wo, f3 wi) {
    SpecParts P{f3{0.f, 0.f, 0.f}, 0.f, 0.f, 0.f};
    if (b.ps <= 0.f) return P;
    f3 h = wo + wi;
    const float hl = length(h);
    if (hl <= 0.f) return P;
    h = h * (1.f / hl);
    const float al = b.alpha, a2 = al * al;
    const float d = h.z * h.z * (a2 - 1.f) + 1.f;
    const float D = a2 / (kPi * d * d);
    const float lo = ggx_lambda(a2, wo), li = ggx_lambda(a2, wi);
    const float G = 1.f / (1.f + lo + li);
    const float c = fmaxf(0.f, dot(wo, h));
    const float x = 1.f - c;
    P.x5 = x * x * x * x * x;
    const f3 f0 = f3{0.04f, 0.04f, 0.04f} * (1.f - b.metal) + b.albedo * b.metal;
    P.F = f0 + (f3{1.f, 1.f, 1.f} - f0) * P.x5;
    P.K = D * G / (4.f * wo.z * wi.z);
    // d log D / d alpha = 2/alpha - 4 alpha cz^2 / d ; d Lambda / d alpha = alpha t^2 / (2 sqrt(1 + alpha^2 t^2))
    auto dlam = [&](f3 w) {
        const float z2 = w.z * w.z;
        if (z2 <= 0.f) return 0.f;
        const float t2 = (w.x * w.x + w.y * w.y) / z2;
        return al * t2 / (2.f * sqrtf(1.f + a2 * t2));
    };
    const float dlogD = 2.f / al - 4.f * al * h.z * h.z / d;
    const float dlogG = -G * (dlam(wo) + dlam(wi));
    const float r = m.fuzziness;
    const float dalpha_dr = r > 0.02f ? 2.f * r : 0.f;
    P.dlogK_dr = (dlogD + dlogG) * dalpha_dr;
    return P;
}

__device__ inline bool metal_live(const gk_material& m) { return m.material_model == GK_MATERIAL_MIXTURE; }

template <bool ADJ>
__device__ inline f3 dpath(const DScene& s, const DiffK& K, uint32_t x, uint32_t y, rng4& seed, f3 dL, f3 Ltot,
                           float* g /* LDS, 5 per material */) {
    f3 O, D;
    float jx, jy;  // sub-pixel position of this sample
    rnd2(seed, jx, jy);
    camera_ray(K, float(x) + jx, float(y) + jy, O, D);
    f3 T = f3{1.f, 1.f, 1.f};
    f3 Lsum = f3{0.f, 0.f, 0.f};
    float last_pdf = 0.f;
    bool last_delta = true;
    auto add = [&](f3 C) { Lsum = Lsum + C; };
    for (uint32_t depth = 0;; ++depth) {
        SHit h;
        if (!trace_closest_single(s, O, D, 0.f, 1e30f, h)) {
            if (K.has_sky) {
                const float4 sky = sky_radiance(s, K.sky, K.sky_rot, D);
                add(T * f3{sky.x, sky.y, sky.z});
            }
            break;
        }
        const InstDev& I = s.inst[h.inst];
        const uint32_t mi = material_of(s, h.inst, h.prim);
        const gk_material& mat = s.mats[mi];
        const f3 P_ = O + D * h.t;
        const size_t tb = 3 * size_t(I.tri_base + h.prim);
        const float4 n0 = s.tri_nrm[tb], n1 = s.tri_nrm[tb + 1], n2 = s.tri_nrm[tb + 2];
        const float bu = h.u, bv = h.v, bw = 1.f - bu - bv;
        f3 Ns = normalize(inv_transpose(I.inv, (ld3(n0) * bw + ld3(n1) * bu) + ld3(n2) * bv));
        f3 Ng = normalize(inv_transpose(I.inv, f3{n0.w, n1.w, n2.w}));
        if (dot(Ng, D) > 0.f) Ng = -Ng;
        if (dot(Ns, Ng) < 0.f) Ns = -Ns;
        if (mat.material_model == GK_MATERIAL_DIFFUSE_LIGHT) {
            const f3 Le = f3{mat.diffuse[0], mat.diffuse[1], mat.diffuse[2]};
            float w = 1.f;
            if (!last_delta && s.nlights > 0u) {
                const float cosl = fabsf(dot(Ng, D));
                const float psel = light_select_pdf(s, h.inst, h.prim);
                float area = 0.f;
                uint32_t lo = 0, hi = s.nlights;
                while (lo < hi) {
                    const uint32_t mid = (lo + hi) >> 1;
                    const uint32_t li = __float_as_uint(s.lights[mid].e2.w), lp = __float_as_uint(s.lights[mid].le.w);
                    if (li < h.inst || (li == h.inst && lp < h.prim)) lo = mid + 1; else hi = mid;
                }
                if (lo < s.nlights) area = s.lights[lo].p0.w;
                const float d2 = h.t * h.t * dot(D, D);
                const float pl = (cosl > 0.f && area > 0.f) ? psel * d2 / (cosl * area) : 0.f;
                w = power_h(last_pdf, pl);
            }
            add(depth == 0u ? Le : T * Le * w);
            break;
        }
        if (depth >= K.bounces) break;
        const float scale = 1.f + fmaxf(fabsf(P_.x), fmaxf(fabsf(P_.y), fabsf(P_.z)));
        const float off = 1e-4f * scale;
        if (mat.material_model == GK_MATERIAL_DIELECTRIC) {
            const float ior = mat.refraction_index;
            const float cosi = -dot(D, Ng);
            const bool entering = dot(D, inv_transpose(I.inv, f3{n0.w, n1.w, n2.w})) < 0.f;
            const float eta = entering ? 1.f / ior : ior;
            float r0 = (1.f - ior) / (1.f + ior);
            r0 = r0 * r0;
            const float xx = 1.f - cosi;
            const float F = r0 + (1.f - r0) * xx * xx * xx * xx * xx;
            const float k = 1.f - eta * eta * (1.f - cosi * cosi);
            const float u = rnd1(seed);
            const bool refl = k < 0.f || u < F;
            f3 wi = refl ? D + Ng * (2.f * cosi) : D * eta + Ng * (eta * cosi - sqrtf(fmaxf(k, 0.f)));
            wi = normalize(wi);
            T = T * f3{mat.diffuse[0], mat.diffuse[1], mat.diffuse[2]};
            last_delta = true;
            O = P_ + (refl ? Ng : -Ng) * off;
            D = wi;
            continue;
        }
        const Bsdf b = bsdf_of(mat);
        f3 T1, B1;
        onb(Ns, T1, B1);
        const f3 wo = f3{dot(T1, -D), dot(B1, -D), dot(Ns, -D)};
        if (wo.z <= 0.f) break;
        if (s.nlights > 0u) {
            const float ul = rnd1(seed);
            float s1, s2;
            rnd2(seed, s1, s2);
            const uint32_t li = light_pick(s, ul);
            const LightTri L = s.lights[li];
            const float sq = sqrtf(s1);
            const f3 Pl = (ld3(L.p0) + ld3(L.e1) * (1.f - sq)) + ld3(L.e2) * (s2 * sq);
            const f3 Po = P_ + Ng * off;
            const f3 dl = Pl - Po;
            const float d2 = dot(dl, dl);
            const float dist = sqrtf(d2);
            const f3 wl = dl * (1.f / dist);
            const float cosl = fabsf(dot(normalize(cross(ld3(L.e1), ld3(L.e2))), wl));
            const f3 wil = f3{dot(T1, wl), dot(B1, wl), dot(Ns, wl)};
            const float psel = s.light_cdf[li + 1] - s.light_cdf[li];  // 0: a light switched off
            if (cosl > 0.f && wil.z > 0.f && dot(Ng, wl) > 0.f && psel > 0.f) {
                const float pl = psel * d2 / (cosl * L.p0.w);
                if (!occluded_single(s, Po, wl, 0.f, dist * (1.f - 1e-4f) - off)) {
                    f3 fd, fs;
                    float pdd, pds;
                    bsdf_eval(b, wo, wil, fd, fs, pdd, pds);
                    const f3 Le = ld3(L.le);
                    const float wd = power_h(pl, (1.f - b.ps) * pdd);
                    const float ws = power_h(pl, b.ps * pds);
                    const f3 A = T * Le * (wil.z / pl);
                    const f3 C = A * (fd * wd + fs * ws);
                    add(C);
                    if (ADJ) {
                        // d C_nee / d theta at this vertex's material
                        const SpecParts sp = spec_parts(b, mat, wo, wil);
                        const float om = 1.f - b.metal;
                        const f3 w3 = dL * A;  // dLoss/dC per channel times A
                        float* gm = g + 5 * mi;
                        const float dfd_da = om / kPi;
                        const float dfs_da = b.metal * (1.f - sp.x5) * sp.K;
                        atomicAdd(&gm[0], w3.x * (wd * dfd_da + ws * dfs_da));
                        atomicAdd(&gm[1], w3.y * (wd * dfd_da + ws * dfs_da));
                        atomicAdd(&gm[2], w3.z * (wd * dfd_da + ws * dfs_da));
                        if (b.ps > 0.f)
                            atomicAdd(&gm[3], ws * sp.K * sp.dlogK_dr * (w3.x * sp.F.x + w3.y * sp.F.y + w3.z * sp.F.z));
                        if (metal_live(mat)) {
                            const float k1 = (1.f - sp.x5) * sp.K;
                            const float dm = w3.x * (wd * (-b.albedo.x / kPi) + ws * (b.albedo.x - 0.04f) * k1) +
                                             w3.y * (wd * (-b.albedo.y / kPi) + ws * (b.albedo.y - 0.04f) * k1) +
                                             w3.z * (wd * (-b.albedo.z / kPi) + ws * (b.albedo.z - 0.04f) * k1);
                            atomicAdd(&gm[4], dm);
                        }
                    }
                }
            }
        }
        const float ulobe = rnd1(seed);
        float u1, u2;
        rnd2(seed, u1, u2);
        const bool spec = ulobe < b.ps;
        f3 wi;
        if (spec) {
            const f3 mh = vndf_sample(b.alpha, wo, u1, u2);
            wi = mh * (2.f * dot(wo, mh)) - wo;
        } else {
            float spn, cpn;
            sincos_2pi(u1, spn, cpn);
            const float r = sqrtf(u2);
            wi = f3{r * cpn, r * spn, sqrtf(1.f - u2)};
        }
        if (wi.z <= 0.f) break;
        f3 fd, fs;
        float pdd, pds;
        bsdf_eval(b, wo, wi, fd, fs, pdd, pds);
        const float pdf = spec ? b.ps * pds : (1.f - b.ps) * pdd;
        // per-component selects of values (a select of the two f3 objects keeps them in scratch)
        const f3 f = f3{spec ? +fs.x : +fd.x, spec ? +fs.y : +fd.y, spec ? +fs.z : +fd.z};
        if (!(pdf > 0.f)) break;
        if (ADJ) {
            // d log f_lobe / d theta times the suffix radiance Ltot - P
            const f3 S = (Ltot - Lsum) * dL;
            float* gm = g + 5 * mi;
            if (spec) {
                const SpecParts sp = spec_parts(b, mat, wo, wi);
                const float k1 = 1.f - sp.x5;
                if (sp.F.x > 0.f) atomicAdd(&gm[0], S.x * b.metal * k1 / sp.F.x);
                if (sp.F.y > 0.f) atomicAdd(&gm[1], S.y * b.metal * k1 / sp.F.y);
                if (sp.F.z > 0.f) atomicAdd(&gm[2], S.z * b.metal * k1 / sp.F.z);
                atomicAdd(&gm[3], (S.x + S.y + S.z) * sp.dlogK_dr);
                if (metal_live(mat)) {
                    float dm = 0.f;
                    if (sp.F.x > 0.f) dm += S.x * (b.albedo.x - 0.04f) * k1 / sp.F.x;
                    if (sp.F.y > 0.f) dm += S.y * (b.albedo.y - 0.04f) * k1 / sp.F.y;
                    if (sp.F.z > 0.f) dm += S.z * (b.albedo.z - 0.04f) * k1 / sp.F.z;
                    atomicAdd(&gm[4], dm);
                }
            } else {
                if (b.albedo.x > 0.f) atomicAdd(&gm[0], S.x / b.albedo.x);
                if (b.albedo.y > 0.f) atomicAdd(&gm[1], S.y / b.albedo.y);
                if (b.albedo.z > 0.f) atomicAdd(&gm[2], S.z / b.albedo.z);
                if (metal_live(mat) && b.metal < 1.f) atomicAdd(&gm[4], -(S.x + S.y + S.z) / (1.f - b.metal));
            }
        }
        T = T * f * (wi.z / pdf);
        last_pdf = pdf;
        last_delta = false;
        const f3 wiw = normalize((T1 * wi.x + B1 * wi.y) + Ns * wi.z);
        O = P_ + (dot(wiw, Ng) > 0.f ? Ng : -Ng) * off;
        D = wiw;
    }
    return Lsum;
}

__global__ __launch_bounds__(256) void k_diff_primal(DScene s, DiffK K, float4* __restrict__ img) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= K.W * K.H) return;
    const uint32_t x = i % K.W, y = i / K.W;
    f3 acc = f3{0.f, 0.f, 0.f};
    for (uint32_t smp = 0; smp < K.spp; ++smp) {
        rng4 seed{x, y, K.frame0 + smp, K.seed};
        acc = acc + dpath<false>(s, K, x, y, seed, f3{0, 0, 0}, f3{0, 0, 0}, nullptr);
    }
    const float inv = 1.f / float(K.spp);
    img[i] = make_float4(acc.x * inv, acc.y * inv, acc.z * inv, 1.f);
}

// grad partials: one row of 5*nmat floats per work-group
__global__ __launch_bounds__(256) void k_diff_adjoint(DScene s, DiffK K, const float4* __restrict__ dimg,
                                                      float* __restrict__ partial, uint32_t nparam) {
    extern __shared__ float g[];
    for (uint32_t k = threadIdx.x; k < nparam; k += blockDim.x) g[k] = 0.f;
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < K.W * K.H) {
        const uint32_t x = i % K.W, y = i / K.W;
        const float inv = 1.f / float(K.spp);
        const float4 dl4 = dimg[i];
        const f3 dL = f3{dl4.x * inv, dl4.y * inv, dl4.z * inv};  // image = mean over samples
        if (dL.x != 0.f || dL.y != 0.f || dL.z != 0.f) {
            for (uint32_t smp = 0; smp < K.spp; ++smp) {
                rng4 seed{x, y, K.frame0 + smp, K.seed};
                const f3 Ltot = dpath<false>(s, K, x, y, seed, f3{0, 0, 0}, f3{0, 0, 0}, nullptr);
                rng4 seed2{x, y, K.frame0 + smp, K.seed};
                dpath<true>(s, K, x, y, seed2, dL, Ltot, g);
            }
        }
    }
    __syncthreads();
    float* row = partial + size_t(blockIdx.x) * nparam;
    for (uint32_t k = threadIdx.x; k < nparam; k += blockDim.x) row[k] = g[k];
}

__global__ void k_reduce_partials(const float* __restrict__ partial, uint32_t rows, uint32_t nparam,
                                  float* __restrict__ out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nparam) return;
    double acc = 0.0;
    for (uint32_t r = 0; r < rows; ++r) acc += double(partial[size_t(r) * nparam + k]);
    out[k] = float(acc);
}

DiffK make_diffk(const Ctx& c) {
    DiffK K;
    memcpy(K.mvi, c.params.model_view_inverse, 64);
    memcpy(K.pi, c.params.projection_inverse, 64);
    memcpy(K.sky, c.params.sky_color, 16);
    K.sky_rot = c.params.sky_rotation;
    K.W = c.width;
    K.H = c.height;
    K.frame0 = c.params.total_frames;
    K.seed = c.params.global_seed;
    K.spp = c.params.samples > 0 ? c.params.samples : 1;
    K.bounces = c.params.bounces;
    K.has_sky = c.params.has_sky;
    return K;
}

}  // namespace

}  // namespace gk

using namespace gk;

extern "C" {

int gk_render_radiance(gk_ctx* ctx, float* d_image) {
    if (!ctx || !d_image) return GK_ERR_ARG;
    ctx->c.early_ok = false;  // queues work on the context's stream
    Ctx& c = ctx->c;
    if (!c.scene_ready || !c.frame_ready) return GK_ERR_STATE;
    if (c.params.aperture > 0.f) return GK_ERR_ARG;  // the differentiable camera is a pinhole
    GK_HIP(hipSetDevice(c.device));
    if (c.tlas_dirty) {
        const int st = refit_tlas(c);
        if (st != GK_OK) return st;
    }
    const uint32_t n = c.width * c.height;
    k_diff_primal<<<(n + 255) / 256, 256, 0, c.stream>>>(make_dscene(c), make_diffk(c),
                                                          reinterpret_cast<float4*>(d_image));
    GK_HIP(hipGetLastError());
    return GK_OK;
}

int gk_adjoint(gk_ctx* ctx, const float* d_dl_dimage, float* d_grad) {
    if (!ctx || !d_dl_dimage || !d_grad) return GK_ERR_ARG;
    ctx->c.early_ok = false;  // queues work on the context's stream
    Ctx& c = ctx->c;
    if (!c.scene_ready || !c.frame_ready) return GK_ERR_STATE;
    if (c.params.aperture > 0.f) return GK_ERR_ARG;  // the differentiable camera is a pinhole
    const uint32_t nparam = 5u * uint32_t(c.h_materials.size());
    if (nparam * 4u > 64u * 1024u) return GK_ERR_LIMIT;  // LDS accumulation: <= 3276 materials
    GK_HIP(hipSetDevice(c.device));
    if (c.tlas_dirty) {
        const int st = refit_tlas(c);
        if (st != GK_OK) return st;
    }
    const uint32_t n = c.width * c.height;
    const uint32_t blocks = (n + 255) / 256;
    if (c.grad_partial.count < size_t(blocks) * nparam) GK_HIP(c.grad_partial.resize(size_t(blocks) * nparam));
    k_diff_adjoint<<<blocks, 256, nparam * 4, c.stream>>>(make_dscene(c), make_diffk(c),
                                                          reinterpret_cast<const float4*>(d_dl_dimage),
                                                          c.grad_partial.ptr, nparam);
    GK_HIP(hipGetLastError());
    k_reduce_partials<<<(nparam + 255) / 256, 256, 0, c.stream>>>(c.grad_partial.ptr, blocks, nparam, d_grad);
    GK_HIP(hipGetLastError());
    return GK_OK;
}

}  // extern "C"

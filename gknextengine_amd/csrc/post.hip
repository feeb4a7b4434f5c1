// post.hip - temporal reprojection, joint bilateral filter + tonemap, history copy.
//
//   k_reproject   Process.ReProject.comp.slang:71-223 (16x16 groups, one thread per pixel)
//   k_jbf         Process.DenoiseJBF.comp.slang:75-181: 16x16 pixels per work-group over
//                 a 26x26 LDS tile of {diffuse, spec}, 6x6 taps at stride 2 without the
//                 centre, weights exp(-d^2/2s^2) * exp(-dlum^2/2sL^2) (the spatial factor
//                 takes four values; the luminance factor uses the hardware exp2, as a
//                 shader's exp does); GT tonemap (Const_Func.slang:84-126) of
//                 Total * PaperWhiteNit / 40000
//   history copy  the "copy pass" accum -> prev of PathTracingRenderer.cpp:160-216
// Storage images read out of bounds return zero, as robust Vulkan image loads do.
#include <hip/hip_runtime.h>

#include "gk_post.hpp"
#include "gk_scene.hpp"

namespace gk {

int unet_denoise_frame(Ctx& c);  // unet.hip

namespace {

struct PostK {
    uint32_t W, H, progressive, temporal, bf_size, hdr, linear;
    float bf_sigma, bf_sigma_lum, paper_white;
};

__device__ inline bool inb(int x, int y, const PostK& K) { return x >= 0 && y >= 0 && x < int(K.W) && y < int(K.H); }
__device__ inline float4 ld_h4(const ushort4* p, int x, int y, const PostK& K) {
    return inb(x, y, K) ? from_h4(p[size_t(y) * K.W + x]) : make_float4(0.f, 0.f, 0.f, 0.f);
}
__device__ inline float lerpf(float a, float b, float t) { return a * (1.f - t) + b * t; }

struct PostT {
    const ushort4 *src_d, *src_s, *src_a, *normal;
    const ushort4 *prev_d, *prev_s, *prev_a;
    ushort4 *acc_d, *acc_s, *acc_a;
    const uint32_t *obj0, *obj1;
    const float2* motion;
    const uint16_t* moment;
    ushort4* out;
};

// The previous frame's 2x2 footprint at (px, py), with px, py inside the image: element
// offsets of the right and lower neighbours (0 where they fall outside, so every address is
// inside) and whether they are inside. Outside elements read as zero.
struct Quad {
    uint32_t b0, o1, oW;
    bool x1, y1;
};
__device__ inline Quad quad_at(int px, int py, const PostK& K) {
    Quad q;
    q.x1 = uint32_t(px + 1) < K.W;
    q.y1 = uint32_t(py + 1) < K.H;
    q.b0 = uint32_t(py) * K.W + uint32_t(px);
    q.o1 = q.x1 ? 1u : 0u;
    q.oW = q.y1 ? K.W : 0u;
    return q;
}

// FilterHistoryColor, Process.ReProject.comp.slang:49-68. The four loads are issued
// unconditionally (from in-image addresses) and the selects pick history, zero or the
// current colour.
__device__ inline f3 filter_history(const ushort4* prev, float sx, float sy, const Quad& q, uint32_t cur,
                                    uint32_t p0, uint32_t p1, uint32_t p2, uint32_t p3, float4 curr) {
    const f3 c = ld3(curr);
    // outside elements are cleared bit by bit (+0 in half precision): every load's value is
    // used, so none is moved under a branch of its own
    const uint16_t m1 = q.x1 ? 0xFFFFu : 0u, m2 = q.y1 ? 0xFFFFu : 0u, m3 = m1 & m2;
    const ushort4 r0 = prev[q.b0], r1 = prev[q.b0 + q.o1], r2 = prev[q.b0 + q.oW], r3 = prev[q.b0 + q.oW + q.o1];
    const f3 t0 = ld3(from_h4(r0));
    const f3 t1 = ld3(from_h4(make_ushort4(r1.x & m1, r1.y & m1, r1.z & m1, 0)));
    const f3 t2 = ld3(from_h4(make_ushort4(r2.x & m2, r2.y & m2, r2.z & m2, 0)));
    const f3 t3 = ld3(from_h4(make_ushort4(r3.x & m3, r3.y & m3, r3.z & m3, 0)));
    const f3 h0 = cur == p0 ? t0 : c;
    const f3 h1 = cur == p1 ? t1 : c;
    const f3 h2 = cur == p2 ? t2 : c;
    const f3 h3 = cur == p3 ? t3 : c;
    f3 a = f3{lerpf(h0.x, h1.x, sx), lerpf(h0.y, h1.y, sx), lerpf(h0.z, h1.z, sx)};
    f3 b = f3{lerpf(h2.x, h3.x, sx), lerpf(h2.y, h3.y, sx), lerpf(h2.z, h3.z, sx)};
    f3 h = f3{lerpf(a.x, b.x, sy), lerpf(a.y, b.y, sy), lerpf(a.z, b.z, sy)};
    h.x = fminf(fmaxf(h.x, 0.f), 1600.f);
    h.y = fminf(fmaxf(h.y, 0.f), 1600.f);
    h.z = fminf(fmaxf(h.z, 0.f), 1600.f);
    return h;
}

// The 5x5 neighbourhood (object ids, normals, colours) and the stride-2 motion samples of
// an 8x8 block are staged in LDS first (zero outside the image, as the loads there return).
// kRB x kRB pixels per work-group; tiles of the 5x5 neighbourhood (kTileP) and of the
// motion samples two pixels apart (kTileM)
constexpr int kRB = 16, kRP = 2, kTileP = kRB + 2 * kRP, kTileM = kRB + 4 * kRP;

__global__ __launch_bounds__(kRB * kRB) void k_reproject(PostK K, PostT T) {
    // the neighbourhood is converted to float once per tile element, and the motion samples
    // are kept as their lengths, instead of once per tap
    __shared__ uint32_t s_obj[kTileP][kTileP];
    __shared__ float4 s_nrm[kTileP][kTileP], s_d[kTileP][kTileP], s_s[kTileP][kTileP];
    __shared__ float s_ml[kTileM][kTileM];
    __shared__ float s_kt[25];
    const int tx = int(threadIdx.x), ty = int(threadIdx.y);
    const int gx = int(blockIdx.x) * kRB, gy = int(blockIdx.y) * kRB;
    const int x = gx + tx, y = gy + ty;
    if (K.progressive) {
        if (x >= int(K.W) || y >= int(K.H)) return;
        const size_t i = size_t(y) * K.W + x;
        const float4 cd = from_h4(T.src_d[i]), cs = from_h4(T.src_s[i]), ca = from_h4(T.src_a[i]);
        const float t = 1.0f / 512.0f;
        const float4 pd = from_h4(T.prev_d[i]), ps = from_h4(T.prev_s[i]), pa = from_h4(T.prev_a[i]);
        T.acc_d[i] = to_h4(lerpf(pd.x, cd.x, t), lerpf(pd.y, cd.y, t), lerpf(pd.z, cd.z, t), lerpf(pd.w, cd.w, t));
        T.acc_s[i] = to_h4(lerpf(ps.x, cs.x, t), lerpf(ps.y, cs.y, t), lerpf(ps.z, cs.z, t), lerpf(ps.w, cs.w, t));
        T.acc_a[i] = to_h4(lerpf(pa.x, ca.x, t), lerpf(pa.y, ca.y, t), lerpf(pa.z, ca.z, t), lerpf(pa.w, ca.w, t));
        return;
    }
    const ushort4 z4 = make_ushort4(0, 0, 0, 0);
    for (int k = ty * kRB + tx; k < kTileP * kTileP; k += kRB * kRB) {
        const int r = k / kTileP, cc = k % kTileP;
        const int sx = gx + cc - kRP, sy = gy + r - kRP;
        const bool in = inb(sx, sy, K);
        // loads under a branch: a select of two lvalues (a global element or the local zero)
        // selects their addresses and reads through a flat pointer into scratch memory
        uint32_t ov = 0u;
        ushort4 nv = z4, dv = z4, sv = z4;
        if (in) {
            const uint32_t j = uint32_t(sy) * K.W + uint32_t(sx);
            ov = T.obj0[j];
            nv = T.normal[j];
            dv = T.src_d[j];
            sv = T.src_s[j];
        }
        s_obj[r][cc] = ov;
        s_nrm[r][cc] = from_h4(nv);
        s_d[r][cc] = from_h4(dv);
        s_s[r][cc] = from_h4(sv);
    }
    for (int k = ty * kRB + tx; k < kTileM * kTileM; k += kRB * kRB) {
        const int r = k / kTileM, cc = k % kTileM;
        const int mx = gx + cc - 2 * kRP, my = gy + r - 2 * kRP;
        const float2 m2 = inb(mx, my, K) ? T.motion[uint32_t(my) * K.W + uint32_t(mx)] : make_float2(0.f, 0.f);
        s_ml[r][cc] = sqrtf(m2.x * m2.x + m2.y * m2.y);
    }
    // calculateWeight (Process.ReProject.comp.slang:32-47) off the centre: for the same
    // object and a clamped normal cosine nd >= 0.98, (nd - 0.98) / (1 - 0.98) * 2 / (1.5 |o| + 4);
    // the table holds its constant factor per tap, 0 at the centre (weighted apart, 0.4)
    const float th = 0.98f;
    if (const int t = ty * kRB + tx; t < 25) {
        const int ox = t % 5 - 2, oy = t / 5 - 2, n = ox * ox + oy * oy;
        s_kt[t] = n == 0 ? 0.f : 2.0f / ((1.0f - th) * (sqrtf(float(n)) * 1.5f + 4.0f));
    }
    __syncthreads();
    if (x >= int(K.W) || y >= int(K.H)) return;
    const uint32_t i = uint32_t(y) * K.W + uint32_t(x);
    const float2 mv = T.motion[i];
    const float fx = float(x) + mv.x, fy = float(y) + mv.y;
    const int ppx = int(floorf(fx)), ppy = int(floorf(fy));
    // all(previpos < isize) && all(previpos >= -1) with int2 previpos and uint2 isize
    // (Process.ReProject.comp.slang:109-110): the mixed-sign comparison converts previpos to
    // unsigned, so -1 fails it and history is used only for 0 <= previpos < size
    const bool inside = uint32_t(ppx) < K.W && uint32_t(ppy) < K.H;
    // the centre of the tile is this pixel's own object id, normal and colours
    const uint32_t cur = s_obj[ty + kRP][tx + kRP];
    const bool miss = cur == 65535u;
    const uint32_t moment = T.moment[i];
    if (miss || !inside || moment > 0u) {
        T.acc_d[i] = T.src_d[i];
        T.acc_s[i] = T.src_s[i];
        T.acc_a[i] = T.src_a[i];
        return;
    }
    const float4 cd = s_d[ty + kRP][tx + kRP], cs = s_s[ty + kRP][tx + kRP], ca = from_h4(T.src_a[i]);
    const Quad q = quad_at(ppx, ppy, K);
    const uint32_t o0 = T.obj1[q.b0], o1 = T.obj1[q.b0 + q.o1], o2 = T.obj1[q.b0 + q.oW],
                   o3 = T.obj1[q.b0 + q.oW + q.o1];
    const float4 cn4 = s_nrm[ty + kRP][tx + kRP];
    const f3 cn = f3{cn4.x, cn4.y, cn4.z};
    float4 sd = make_float4(0.f, 0.f, 0.f, 0.f), ss = make_float4(0.f, 0.f, 0.f, 0.f);
    float tw = 0.f, motion_in_range = 0.f;
    // rows one at a time, the five taps of a row unrolled: fully unrolled, all 25 taps' tile
    // reads are hoisted ahead of the arithmetic and exhaust the registers
#pragma unroll 1
    for (int oy = -2; oy <= 2; ++oy) {
        const int r = ty + kRP + oy;
#pragma unroll
        for (int ox = -2; ox <= 2; ++ox) {
            const int cc = tx + kRP + ox;
            const float kt = s_kt[(oy + 2) * 5 + ox + 2];
            const float4 n4 = s_nrm[r][cc];
            const float nd = fminf(fmaxf(dot(f3{n4.x, n4.y, n4.z}, cn), 0.f), 1.f);
            const float w = (s_obj[r][cc] == cur && nd >= th) ? (nd - th) * kt : 0.f;
            const float4 a = s_d[r][cc], b = s_s[r][cc];
            sd.x += a.x * w; sd.y += a.y * w; sd.z += a.z * w;
            ss.x += b.x * w; ss.y += b.y * w; ss.z += b.z * w;
            tw += w;
            // out-of-image samples are zero in the tile and add sqrt(0) = 0
            motion_in_range += s_ml[ty + 2 * kRP + 2 * oy][tx + 2 * kRP + 2 * ox];
        }
    }
    // the centre, weight 0.4, last (the shader adds it in raster order; the sums differ by
    // rounding only)
    sd.x += cd.x * 0.4f; sd.y += cd.y * 0.4f; sd.z += cd.z * 0.4f;
    ss.x += cs.x * 0.4f; ss.y += cs.y * 0.4f; ss.z += cs.z * 0.4f;
    tw += 0.4f;
    sd.x /= tw; sd.y /= tw; sd.z /= tw;
    ss.x /= tw; ss.y /= tw; ss.z /= tw;
    // object ids of the footprint, zero outside the image; all the current id when the
    // neighbourhood barely moves
    const bool still = fabsf(motion_in_range) < 0.1f;
    const uint32_t p0 = still ? cur : o0;
    const uint32_t p1 = still ? cur : (q.x1 ? o1 : 0u);
    const uint32_t p2 = still ? cur : (q.y1 ? o2 : 0u);
    const uint32_t p3 = still ? cur : (q.x1 && q.y1 ? o3 : 0u);
    const float subx = fx - floorf(fx), suby = fy - floorf(fy);
    const f3 hd = filter_history(T.prev_d, subx, suby, q, cur, p0, p1, p2, p3, sd);
    const f3 hs = filter_history(T.prev_s, subx, suby, q, cur, p0, p1, p2, p3, ss);
    const f3 ha = filter_history(T.prev_a, subx, suby, q, cur, p0, p1, p2, p3, ca);
    float keep = 1.0f / float(K.temporal > 1u ? K.temporal : 1u);
    keep = fminf(fmaxf(keep, 0.f), 1.f);
    // .rgb written, alpha keeps the accumulation target's previous value
    const float4 od = from_h4(T.acc_d[i]), os = from_h4(T.acc_s[i]), oa = from_h4(T.acc_a[i]);
    T.acc_d[i] = to_h4(lerpf(hd.x, cd.x, keep), lerpf(hd.y, cd.y, keep), lerpf(hd.z, cd.z, keep), od.w);
    T.acc_s[i] = to_h4(lerpf(hs.x, cs.x, keep), lerpf(hs.y, cs.y, keep), lerpf(hs.z, cs.z, keep), os.w);
    T.acc_a[i] = to_h4(lerpf(ha.x, ca.x, keep), lerpf(ha.y, ca.y, keep), lerpf(ha.z, ca.z, keep), oa.w);
}

constexpr int kR = 5, kJB = 16, kTile = kJB + 2 * kR;  // kJB x kJB pixels per work-group

__global__ __launch_bounds__(kJB * kJB) void k_jbf(PostK K, PostT T) {
    // each tile element holds c = colour + 0.001 and its luminance dot(c, lc): computed once
    // per element here instead of once per tap below
    __shared__ float4 sd[kTile][kTile];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int gx = blockIdx.x * kJB, gy = blockIdx.y * kJB;
    const int x = gx + tx, y = gy + ty;
    const f3 lc = f3{0.212671f, 0.715160f, 0.072169f};
    f3 total;
    if (K.bf_size > 0u) {
        for (int idx = ty * kJB + tx; idx < kTile * kTile; idx += kJB * kJB) {
            const int row = idx / kTile, col = idx % kTile;
            const int cx = gx + col - kR, cy = gy + row - kR;
            const float4 d = ld_h4(T.acc_d, cx, cy, K);
            const f3 c = f3{d.x + 0.001f, d.y + 0.001f, d.z + 0.001f};
            sd[row][col] = make_float4(c.x, c.y, c.z, dot(c, lc));
        }
        __syncthreads();
        if (x >= int(K.W) || y >= int(K.H)) return;
        const float clum = sd[ty + kR][tx + kR].w;
        // the specular term is read at the centre only
        const float4 s4 = ld_h4(T.acc_s, x, y, K);
        const f3 cspec = f3{s4.x + 0.001f, s4.y + 0.001f, s4.z + 0.001f};
        const float sigma = K.bf_sigma, sigma_l = K.bf_sigma_lum * 100.0f;
        // the spatial weight depends on i*i + j*j only, which is 2, 10, 18 or (clamped to
        // distance 1) at least 26: four values, computed once with the per-tap expression
        auto spatial = [&](int n) {
            float dist = float(n) / float(kR * kR);
            dist = fminf(fmaxf(dist, 0.f), 1.f);
            return expf(-dist * dist / (2.0f * sigma * sigma));
        };
        const float F2 = spatial(2), F10 = spatial(10), F18 = spatial(18), Ffar = spatial(2 * kR * kR);
        // exp(-dl^2 / (2 sL^2)) as the hardware exp2 of dl^2 * (-log2(e) / (2 sL^2))
        const float lk = -1.44269504f / (2.0f * sigma_l * sigma_l);
        f3 acc = f3{0.f, 0.f, 0.f};
        float wsum = 0.f;
#pragma unroll
        for (int i = -kR; i <= kR; i += 2)
#pragma unroll
            for (int j = -kR; j <= kR; j += 2) {
                const float4 q = sd[ty + kR + i][tx + kR + j];
                const int n = i * i + j * j;
                const float dl = (clum - q.w) * (clum - q.w);
                const float Fi = n == 2 ? F2 : n == 10 ? F10 : n == 18 ? F18 : Ffar;
                const float w = Fi * __builtin_amdgcn_exp2f(dl * dl * lk);
                acc.x = fmaf(q.x, w, acc.x);
                acc.y = fmaf(q.y, w, acc.y);
                acc.z = fmaf(q.z, w, acc.z);
                wsum += w;
            }
        total = acc * (1.f / wsum);
        const f3 alb = ld3(from_h4(T.acc_a[size_t(y) * K.W + x]));
        total = total * alb + cspec;
    } else {
        if (x >= int(K.W) || y >= int(K.H)) return;
        const size_t i = size_t(y) * K.W + x;
        total = ld3(from_h4(T.acc_d[i])) * ld3(from_h4(T.acc_a[i])) + ld3(from_h4(T.acc_s[i]));
    }
    const bool hdr = K.hdr != 0u;
    T.out[size_t(y) * K.W + x] =
        K.linear ? to_h4(total.x, total.y, total.z, 1.f)
                 : to_h4(display_encode(total.x, K.paper_white, hdr), display_encode(total.y, K.paper_white, hdr),
                         display_encode(total.z, K.paper_white, hdr), 1.f);
}

__global__ void k_copy_history(const ushort4* a, const ushort4* b, const ushort4* c, ushort4* pa, ushort4* pb,
                               ushort4* pc, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    pa[i] = a[i];
    pb[i] = b[i];
    pc[i] = c[i];
}

}  // namespace

int post_process(Ctx& c) {
    const uint32_t n = c.width * c.height;
    // history images (RT_SINGLE_PREV_*): last frame's accumulation targets
    auto& w = c.wf;
    if (w.hist_d.count < n) {
        GK_HIP(w.hist_d.resize(n));
        GK_HIP(w.hist_s.resize(n));
        GK_HIP(w.hist_a.resize(n));
        GK_HIP(hipMemsetAsync(w.hist_d.ptr, 0, n * 8, c.stream));
        GK_HIP(hipMemsetAsync(w.hist_s.ptr, 0, n * 8, c.stream));
        GK_HIP(hipMemsetAsync(w.hist_a.ptr, 0, n * 8, c.stream));
    }
    ushort4* prev_d = w.hist_d.ptr;
    ushort4* prev_s = w.hist_s.ptr;
    ushort4* prev_a = w.hist_a.ptr;
    PostK K;
    K.W = c.width;
    K.H = c.height;
    K.progressive = c.params.progressive;
    K.temporal = c.params.temporal_frames;
    K.bf_size = c.params.denoiser == GK_DENOISER_BILATERAL ? c.params.bf_size : 0u;
    K.hdr = c.params.hdr;
    K.linear = (c.params.flags & GK_FRAME_LINEAR_OUTPUT) ? 1u : 0u;
    K.bf_sigma = c.params.bf_sigma;
    K.bf_sigma_lum = c.params.bf_sigma_lum;
    K.paper_white = c.params.paper_white_nit;
    PostT T;
    T.src_d = static_cast<const ushort4*>(c.target_ptr[GK_RT_SINGLE_DIFFUSE]);
    T.src_s = static_cast<const ushort4*>(c.target_ptr[GK_RT_SINGLE_SPECULAR]);
    T.src_a = static_cast<const ushort4*>(c.target_ptr[GK_RT_ALBEDO]);
    T.normal = static_cast<const ushort4*>(c.target_ptr[GK_RT_NORMAL]);
    T.prev_d = prev_d;
    T.prev_s = prev_s;
    T.prev_a = prev_a;
    T.acc_d = static_cast<ushort4*>(c.target_ptr[GK_RT_ACCUM_DIFFUSE]);
    T.acc_s = static_cast<ushort4*>(c.target_ptr[GK_RT_ACCUM_SPECULAR]);
    T.acc_a = static_cast<ushort4*>(c.target_ptr[GK_RT_ACCUM_ALBEDO]);
    T.obj0 = static_cast<const uint32_t*>(c.target_ptr[GK_RT_OBJECTID_0]);
    T.obj1 = static_cast<const uint32_t*>(c.target_ptr[GK_RT_OBJECTID_1]);
    T.motion = static_cast<const float2*>(c.target_ptr[GK_RT_MOTIONVECTOR]);
    T.moment = static_cast<const uint16_t*>(c.target_ptr[GK_RT_MOTIONMOMENT]);
    T.out = static_cast<ushort4*>(c.target_ptr[GK_RT_DENOISED]);
    const dim3 grid((c.width + kRB - 1) / kRB, (c.height + kRB - 1) / kRB), blk(kRB, kRB);
    hipEventRecord(c.ev[4], c.stream);
    k_reproject<<<grid, blk, 0, c.stream>>>(K, T);
    GK_HIP(hipGetLastError());
    hipEventRecord(c.ev[5], c.stream);
    if (c.params.denoiser == GK_DENOISER_UNET) {
        const int s = unet_denoise_frame(c);
        if (s != GK_OK) return s;
    } else {
        const dim3 jgrid((c.width + kJB - 1) / kJB, (c.height + kJB - 1) / kJB), jblk(kJB, kJB);
        k_jbf<<<jgrid, jblk, 0, c.stream>>>(K, T);
        GK_HIP(hipGetLastError());
    }
    hipEventRecord(c.ev[6], c.stream);
    k_copy_history<<<(n + 255) / 256, 256, 0, c.stream>>>(T.acc_d, T.acc_s, T.acc_a, prev_d, prev_s, prev_a, n);
    GK_HIP(hipGetLastError());
    return GK_OK;
}

}  // namespace gk

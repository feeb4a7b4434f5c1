// output.hip - the output stage after the compose pass: the resolve to the display
// extent (VulkanBaseRenderer.cpp:1041-1067) and screenshots (ScreenShot.cpp:99-283).
//
//   GK_UPSCALE_LINEAR    the live resolve: vkCmdBlitImage with VK_FILTER_LINEAR, i.e. the
//                        destination pixel centre mapped into the source, bilinear with
//                        clamp-to-edge
//   GK_UPSCALE_FSR_EASU  Process.UpScaleFSR.comp.slang: FidelityFX FSR 1 EASU (12-tap
//                        edge-adaptive Lanczos-like kernel with deringing), equal sizes copy
#include <hip/hip_runtime.h>
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "gk_ctx.hpp"
#include "gk_scene.hpp"

namespace gk {
namespace {

__device__ inline float4 ld_px(const ushort4* img, int x, int y, int w) { return from_h4(img[size_t(y) * w + x]); }

__global__ void k_blit_linear(const ushort4* __restrict__ src, int sw, int sh, ushort4* __restrict__ dst, int dw, int dh) {
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x), y = int(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= dw || y >= dh) return;
    const float u = (float(x) + 0.5f) * (float(sw) / float(dw)) - 0.5f;
    const float v = (float(y) + 0.5f) * (float(sh) / float(dh)) - 0.5f;
    const float fu = floorf(u), fv = floorf(v);
    const float a = u - fu, b = v - fv;
    const int x0 = min(max(int(fu), 0), sw - 1), x1 = min(max(int(fu) + 1, 0), sw - 1);
    const int y0 = min(max(int(fv), 0), sh - 1), y1 = min(max(int(fv) + 1, 0), sh - 1);
    const float4 p = ld_px(src, x0, y0, sw), q = ld_px(src, x1, y0, sw);
    const float4 r = ld_px(src, x0, y1, sw), s = ld_px(src, x1, y1, sw);
    const float ia = 1.f - a, ib = 1.f - b;
    const float4 top = make_float4(p.x * ia + q.x * a, p.y * ia + q.y * a, p.z * ia + q.z * a, p.w * ia + q.w * a);
    const float4 bot = make_float4(r.x * ia + s.x * a, r.y * ia + s.y * a, r.z * ia + s.z * a, r.w * ia + s.w * a);
    dst[size_t(y) * dw + x] = to_h4(top.x * ib + bot.x * b, top.y * ib + bot.y * b, top.z * ib + bot.z * b,
                                    top.w * ib + bot.w * b);
}

// FsrEasuCF: the texel holding normalised position p (point sampling, clamped)
__device__ inline float4 easu_tap(const ushort4* src, int sw, int sh, float px, float py) {
    const int cx = min(max(int(px * float(sw)), 0), sw - 1);
    const int cy = min(max(int(py * float(sh)), 0), sh - 1);
    return ld_px(src, cx, cy, sw);
}
__device__ inline float luma2(float4 c) { return c.y + 0.5f * (c.x + c.z); }

// FsrEasuSetF: direction and length from the '+' pattern around c
__device__ inline void easu_set(float& dx, float& dy, float& len, float w, float lA, float lB, float lC, float lD,
                                float lE) {
    float lenX = fmaxf(fabsf(lD - lC), fabsf(lC - lB));
    const float dirX = lD - lB;
    dx = dx + dirX * w;
    lenX = fminf(fmaxf(fabsf(dirX) / lenX, 0.f), 1.f);
    lenX = lenX * lenX;
    len = len + lenX * w;
    float lenY = fmaxf(fabsf(lE - lC), fabsf(lC - lA));
    const float dirY = lE - lA;
    dy = dy + dirY * w;
    lenY = fminf(fmaxf(fabsf(dirY) / lenY, 0.f), 1.f);
    lenY = lenY * lenY;
    len = len + lenY * w;
}

// FsrEasuTapF: one tap of the kernel, offset (ox, oy) from the resolve position
__device__ inline void easu_acc(float4& aC, float& aW, float ox, float oy, float dx, float dy, float lx, float ly,
                                float lob, float clp, float4 c) {
    float vx = ox * dx + oy * dy;
    float vy = ox * (-dy) + oy * dx;
    vx = vx * lx;
    vy = vy * ly;
    const float d2 = fminf(vx * vx + vy * vy, clp);
    float wB = 0.4f * d2 - 1.f;
    float wA = lob * d2 - 1.f;
    wB = wB * wB;
    wA = wA * wA;
    wB = 1.5625f * wB - 0.5625f;
    const float w = wB * wA;
    aC = make_float4(aC.x + c.x * w, aC.y + c.y * w, aC.z + c.z * w, aC.w + c.w * w);
    aW = aW + w;
}

__global__ void k_fsr_easu(const ushort4* __restrict__ src, int sw, int sh, ushort4* __restrict__ dst, int dw, int dh) {
    const int x = int(blockIdx.x * blockDim.x + threadIdx.x), y = int(blockIdx.y * blockDim.y + threadIdx.y);
    if (x >= dw || y >= dh) return;
    if (sw == dw && sh == dh) {
        dst[size_t(y) * dw + x] = src[size_t(y) * sw + x];
        return;
    }
    // FsrEasuCon (input viewport = input size = the render extent)
    const float c0x = float(sw) / float(dw), c0y = float(sh) / float(dh);
    const float c0z = 0.5f * float(sw) / float(dw) - 0.5f, c0w = 0.5f * float(sh) / float(dh) - 0.5f;
    const float c1x = 1.f / float(sw), c1y = 1.f / float(sh), c1z = 1.f / float(sw), c1w = -1.f / float(sh);
    const float c2x = -1.f / float(sw), c2y = 2.f / float(sh), c2z = 1.f / float(sw), c2w = 2.f / float(sh);
    const float c3y = 4.f / float(sh);
    // the shader's uv * output size
    const float ipx = (float(x) / float(dw)) * float(dw), ipy = (float(y) / float(dh)) * float(dh);
    float ppx = ipx * c0x + c0z, ppy = ipy * c0y + c0w;
    const float fpx = floorf(ppx), fpy = floorf(ppy);
    ppx = ppx - fpx;
    ppy = ppy - fpy;
    const float p0x = fpx * c1x + c1z, p0y = fpy * c1y + c1w;
    const float p1x = p0x + c2x, p1y = p0y + c2y;
    const float p2x = p0x + c2z, p2y = p0y + c2w;
    const float p3x = p0x + 0.f, p3y = p0y + c3y;
    const float ox = -0.5f * c1x, oy = 0.5f * c1x, oz = -0.5f * c1y, ow = 0.5f * c1y;
    const float4 bC = easu_tap(src, sw, sh, p0x + ox, p0y + ow);
    const float4 cC = easu_tap(src, sw, sh, p0x + oy, p0y + ow);
    const float4 iC = easu_tap(src, sw, sh, p1x + ox, p1y + ow);
    const float4 jC = easu_tap(src, sw, sh, p1x + oy, p1y + ow);
    const float4 fC = easu_tap(src, sw, sh, p1x + oy, p1y + oz);
    const float4 eC = easu_tap(src, sw, sh, p1x + ox, p1y + oz);
    const float4 kC = easu_tap(src, sw, sh, p2x + ox, p2y + ow);
    const float4 lC = easu_tap(src, sw, sh, p2x + oy, p2y + ow);
    const float4 hC = easu_tap(src, sw, sh, p2x + oy, p2y + oz);
    const float4 gC = easu_tap(src, sw, sh, p2x + ox, p2y + oz);
    const float4 oC = easu_tap(src, sw, sh, p3x + oy, p3y + oz);
    const float4 nC = easu_tap(src, sw, sh, p3x + ox, p3y + oz);
    const float bL = luma2(bC), cL = luma2(cC), iL = luma2(iC), jL = luma2(jC), fL = luma2(fC), eL = luma2(eC);
    const float kL = luma2(kC), lL = luma2(lC), hL = luma2(hC), gL = luma2(gC), oL = luma2(oC), nL = luma2(nC);
    float dx = 0.f, dy = 0.f, len = 0.f;
    easu_set(dx, dy, len, (1.f - ppx) * (1.f - ppy), bL, eL, fL, gL, jL);
    easu_set(dx, dy, len, ppx * (1.f - ppy), cL, fL, gL, hL, kL);
    easu_set(dx, dy, len, (1.f - ppx) * ppy, fL, iL, jL, kL, nL);
    easu_set(dx, dy, len, ppx * ppy, gL, jL, kL, lL, oL);
    float dirR = dx * dx + dy * dy;
    const bool zro = dirR < (1.0f / 32768.0f);
    dirR = 1.0f / sqrtf(dirR);
    dirR = zro ? 1.0f : dirR;
    dx = zro ? 1.0f : dx;
    dx = dx * dirR;
    dy = dy * dirR;
    len = len * 0.5f;
    len = len * len;
    const float stretch = (dx * dx + dy * dy) / fmaxf(fabsf(dx), fabsf(dy));
    const float lx = 1.f + (stretch - 1.0f) * len, ly = 1.f - 0.5f * len;
    const float lob = 0.5f - 0.29f * len;
    const float clp = 1.f / lob;
    const float4 mn = make_float4(fminf(fminf(fC.x, gC.x), fminf(jC.x, kC.x)), fminf(fminf(fC.y, gC.y), fminf(jC.y, kC.y)),
                                  fminf(fminf(fC.z, gC.z), fminf(jC.z, kC.z)), fminf(fminf(fC.w, gC.w), fminf(jC.w, kC.w)));
    const float4 mx = make_float4(fmaxf(fmaxf(fC.x, gC.x), fmaxf(jC.x, kC.x)), fmaxf(fmaxf(fC.y, gC.y), fmaxf(jC.y, kC.y)),
                                  fmaxf(fmaxf(fC.z, gC.z), fmaxf(jC.z, kC.z)), fmaxf(fmaxf(fC.w, gC.w), fmaxf(jC.w, kC.w)));
    float4 aC = make_float4(0.f, 0.f, 0.f, 0.f);
    float aW = 0.f;
    easu_acc(aC, aW, 0.f - ppx, -1.f - ppy, dx, dy, lx, ly, lob, clp, bC);
    easu_acc(aC, aW, 1.f - ppx, -1.f - ppy, dx, dy, lx, ly, lob, clp, cC);
    easu_acc(aC, aW, -1.f - ppx, 1.f - ppy, dx, dy, lx, ly, lob, clp, iC);
    easu_acc(aC, aW, 0.f - ppx, 1.f - ppy, dx, dy, lx, ly, lob, clp, jC);
    easu_acc(aC, aW, 0.f - ppx, 0.f - ppy, dx, dy, lx, ly, lob, clp, fC);
    easu_acc(aC, aW, -1.f - ppx, 0.f - ppy, dx, dy, lx, ly, lob, clp, eC);
    easu_acc(aC, aW, 1.f - ppx, 1.f - ppy, dx, dy, lx, ly, lob, clp, kC);
    easu_acc(aC, aW, 2.f - ppx, 1.f - ppy, dx, dy, lx, ly, lob, clp, lC);
    easu_acc(aC, aW, 2.f - ppx, 0.f - ppy, dx, dy, lx, ly, lob, clp, hC);
    easu_acc(aC, aW, 1.f - ppx, 0.f - ppy, dx, dy, lx, ly, lob, clp, gC);
    easu_acc(aC, aW, 1.f - ppx, 2.f - ppy, dx, dy, lx, ly, lob, clp, oC);
    easu_acc(aC, aW, 0.f - ppx, 2.f - ppy, dx, dy, lx, ly, lob, clp, nC);
    const float4 r = make_float4(aC.x / aW, aC.y / aW, aC.z / aW, aC.w / aW);
    dst[size_t(y) * dw + x] = to_h4(fminf(mx.x, fmaxf(mn.x, r.x)), fminf(mx.y, fmaxf(mn.y, r.y)),
                                    fminf(mx.z, fmaxf(mn.z, r.z)), fminf(mx.w, fmaxf(mn.w, r.w)));
}

float half_to_float(uint16_t h) {
    const uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 31u, m = h & 1023u;
    float v;
    if (e == 0) v = std::ldexp(float(m), -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp(float(m | 1024u), int(e) - 25);
    return s ? -v : v;
}

void put32(std::vector<uint8_t>& o, uint32_t v) {
    for (int k = 3; k >= 0; --k) o.push_back(uint8_t(v >> (8 * k)));
}
void chunk(std::vector<uint8_t>& o, const char* type, const std::vector<uint8_t>& data) {
    put32(o, uint32_t(data.size()));
    const size_t start = o.size();
    o.insert(o.end(), type, type + 4);
    o.insert(o.end(), data.begin(), data.end());
    put32(o, uint32_t(crc32(0L, o.data() + start, uInt(o.size() - start))));
}

}  // namespace
}  // namespace gk

extern "C" {

int gk_render_extent(uint32_t out_w, uint32_t out_h, uint32_t super_resolution, uint32_t* w, uint32_t* h) {
    if (!w || !h || super_resolution > 2) return GK_ERR_ARG;
    const uint32_t divider = 4u - super_resolution;  // VulkanBaseRenderer.cpp:474-476
    *w = out_w * 2u / divider;
    *h = out_h * 2u / divider;
    return GK_OK;
}

int gk_upscale(gk_ctx* ctx, int mode, uint32_t out_w, uint32_t out_h, void* d_out) {
    GK_CHECK_CTX_KEEP(ctx);  // reads the display target only
    if (!d_out || out_w == 0 || out_h == 0 || (mode != GK_UPSCALE_LINEAR && mode != GK_UPSCALE_FSR_EASU))
        return GK_ERR_ARG;
    if (out_w > 16384 || out_h > 16384) return GK_ERR_LIMIT;
    gk::Ctx& c = ctx->c;
    GK_HIP(hipSetDevice(c.device));
    const ushort4* src = static_cast<const ushort4*>(c.target_ptr[GK_RT_DENOISED]);
    ushort4* dst = static_cast<ushort4*>(d_out);
    const dim3 blk(16, 16), grd((out_w + 15) / 16, (out_h + 15) / 16);
    if (mode == GK_UPSCALE_LINEAR)
        gk::k_blit_linear<<<grd, blk, 0, c.stream>>>(src, int(c.width), int(c.height), dst, int(out_w), int(out_h));
    else
        gk::k_fsr_easu<<<grd, blk, 0, c.stream>>>(src, int(c.width), int(c.height), dst, int(out_w), int(out_h));
    GK_HIP(hipGetLastError());
    return GK_OK;
}

int gk_screenshot(gk_ctx* ctx, const void* d_rgba16f, uint32_t w, uint32_t h, const char* path) {
    GK_CHECK_CTX_KEEP(ctx);  // reads the display target only
    if (!path) return GK_ERR_ARG;
    gk::Ctx& c = ctx->c;
    if (!d_rgba16f) {
        d_rgba16f = c.target_ptr[GK_RT_DENOISED];
        w = c.width;
        h = c.height;
    }
    if (w == 0 || h == 0) return GK_ERR_ARG;
    GK_HIP(hipSetDevice(c.device));
    std::vector<uint16_t> px(size_t(w) * h * 4);
    GK_HIP(hipMemcpyAsync(px.data(), d_rgba16f, px.size() * 2, hipMemcpyDeviceToHost, c.stream));
    GK_HIP(hipStreamSynchronize(c.stream));
    // float -> 8-bit UNORM as the swap chain stores it (round to nearest of clamp(x) * 255)
    std::vector<uint8_t> raw;
    raw.reserve((size_t(w) * 3 + 1) * h);
    for (uint32_t y = 0; y < h; ++y) {
        raw.push_back(0);  // filter: none
        for (uint32_t x = 0; x < w; ++x)
            for (int k = 0; k < 3; ++k) {
                float v = gk::half_to_float(px[(size_t(y) * w + x) * 4 + k]);
                v = std::isnan(v) ? 0.f : std::fmin(std::fmax(v, 0.f), 1.f);
                raw.push_back(uint8_t(std::lround(v * 255.f)));
            }
    }
    uLongf zlen = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zlen);
    if (compress2(z.data(), &zlen, raw.data(), uLong(raw.size()), 6) != Z_OK) return GK_ERR_IO;
    z.resize(zlen);
    std::vector<uint8_t> png = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    gk::put32(ihdr, w);
    gk::put32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit RGB
    gk::chunk(png, "IHDR", ihdr);
    gk::chunk(png, "IDAT", z);
    gk::chunk(png, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f) return GK_ERR_IO;
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    return (std::fclose(f) == 0 && ok) ? GK_OK : GK_ERR_IO;
}

}  // extern "C"

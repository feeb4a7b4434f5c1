// oracle_post.cpp - CPU restatement of the post passes (test infrastructure only).
//
//   oracle_reproject  Process.ReProject.comp.slang:71-223 (FastReproject = false)
//   oracle_jbf        Process.DenoiseJBF.comp.slang:75-181 (HDR off, DebugDraw off)
//   oracle_gt_tonemap GranTurismoTonemapper, Const_Func.slang:84-121
// Images are row-major float4 arrays holding the values of the RGBA16F targets
// (callers pass fp16-rounded inputs). Reads outside the image return 0, as robust
// storage-image loads do.
#include <cmath>
#include <cstdint>

namespace {

struct Img4 {
    const float* p;
    int W, H;
    void at(int x, int y, float o[4]) const {
        if (x < 0 || y < 0 || x >= W || y >= H) { o[0] = o[1] = o[2] = o[3] = 0.f; return; }
        const float* q = p + 4 * (size_t(y) * W + x);
        o[0] = q[0]; o[1] = q[1]; o[2] = q[2]; o[3] = q[3];
    }
};
struct ImgU {
    const uint32_t* p;
    int W, H;
    uint32_t at(int x, int y) const { return (x < 0 || y < 0 || x >= W || y >= H) ? 0u : p[size_t(y) * W + x]; }
};

float lerpf(float a, float b, float t) { return a * (1.f - t) + b * t; }

// calculateWeight, Process.ReProject.comp.slang:32-47
float calc_weight(float center_dist, bool same, bool center, const float* n, const float* cn) {
    if (center) return 0.4f;
    if (!same) return 0.f;
    float nd = n[0] * cn[0] + n[1] * cn[1] + n[2] * cn[2];
    nd = nd < 0.f ? 0.f : (nd > 1.f ? 1.f : nd);
    const float th = 0.98f;
    if (nd < th) return 0.f;
    const float nw = (nd - th) / (1.0f - th);
    return nw * 2.0f / (center_dist * 1.5f + 4.0f);
}

// FilterHistoryColor, Process.ReProject.comp.slang:49-68
void filter_history(const Img4& prev, float sx, float sy, int px, int py, uint32_t cur, const uint32_t p[4],
                    const float* curr, float out[3]) {
    float h[4][4];
    const int ox[4] = {0, 1, 0, 1}, oy[4] = {0, 0, 1, 1};
    for (int k = 0; k < 4; ++k) {
        if (cur == p[k]) prev.at(px + ox[k], py + oy[k], h[k]);
        else for (int c = 0; c < 3; ++c) h[k][c] = curr[c];
    }
    for (int c = 0; c < 3; ++c) {
        const float a = lerpf(h[0][c], h[1][c], sx), b = lerpf(h[2][c], h[3][c], sx);
        float v = lerpf(a, b, sy);
        v = v < 0.f ? 0.f : (v > 1600.f ? 1600.f : v);
        out[c] = v;
    }
}

float w_f(float x, float e0, float e1) {
    if (x <= e0) return 0.f;
    if (x >= e1) return 1.f;
    const float a = (x - e0) / (e1 - e0);
    return a * a * (3.f - 2.f * a);
}
float h_f(float x, float e0, float e1) {
    if (x <= e0) return 0.f;
    if (x >= e1) return 1.f;
    return (x - e0) / (e1 - e0);
}
float gt(float x) {
    const float P = 1.f, a = 0.7f, m = 0.22f, l = 0.4f, c = 1.33f, b = 0.f;
    const float l0 = (P - m) * l / a;
    const float L_x = m + a * (x - m);
    const float T_x = m * std::pow(x / m, c) + b;
    const float S0 = m + l0;
    const float S1 = m + a * l0;
    const float C2 = a * P / (P - S1);
    const float S_x = P - (P - S1) * std::pow(2.71828f, -(C2 * (x - S0) / P));
    const float w0 = 1.f - w_f(x, 0.f, m);
    const float w2 = h_f(x, m + l0, m + l0);
    const float w1 = 1.f - w0 - w2;
    return T_x * w0 + L_x * w1 + S_x * w2;
}

}  // namespace

extern "C" {

float oracle_gt_tonemap(float x) { return gt(x); }

// params: {progressive, temporal_frames}. Accumulation targets are in/out: their alpha
// channel is kept where the reference writes .rgb only.
void oracle_reproject(const uint32_t* params, uint32_t W, uint32_t H, const float* src_d, const float* src_s,
                      const float* src_a, const float* normal, const float* prev_d, const float* prev_s,
                      const float* prev_a, const uint32_t* obj0, const uint32_t* obj1, const float* motion,
                      const uint16_t* moment, float* acc /* 3 x W*H*4: diffuse, spec, albedo */) {
    const int w = int(W), h = int(H);
    const Img4 SD{src_d, w, h}, SS{src_s, w, h}, NB{normal, w, h}, PD{prev_d, w, h}, PS{prev_s, w, h},
        PA{prev_a, w, h};
    const ImgU O0{obj0, w, h}, O1{obj1, w, h};
    const size_t n = size_t(W) * H;
    float* ad = acc;
    float* as = acc + 4 * n;
    float* aa = acc + 8 * n;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = size_t(y) * W + x;
            const float* cd = src_d + 4 * i;
            const float* cs = src_s + 4 * i;
            const float* ca = src_a + 4 * i;
            if (params[0]) {
                const float t = 1.0f / 512.0f;
                for (int c = 0; c < 4; ++c) {
                    ad[4 * i + c] = lerpf(prev_d[4 * i + c], cd[c], t);
                    as[4 * i + c] = lerpf(prev_s[4 * i + c], cs[c], t);
                    aa[4 * i + c] = lerpf(prev_a[4 * i + c], ca[c], t);
                }
                continue;
            }
            const float mx = motion[2 * i], my = motion[2 * i + 1];
            const float fx = float(x) + mx, fy = float(y) + my;
            const int ppx = int(std::floor(fx)), ppy = int(std::floor(fy));
            // int2 previpos < uint2 isize compares as unsigned (Process.ReProject.comp.slang:
            // 109-110): -1 is rejected, so the >= -1 clause never decides
            const bool inside = uint32_t(ppx) < W && uint32_t(ppy) < H;
            const uint32_t cur = obj0[i];
            uint32_t p[4];
            p[0] = O1.at(ppx, ppy);
            if (cur == 65535u || !inside || moment[i] > 0) {
                for (int c = 0; c < 4; ++c) { ad[4 * i + c] = cd[c]; as[4 * i + c] = cs[c]; aa[4 * i + c] = ca[c]; }
                continue;
            }
            const float* cn = normal + 4 * i;
            float sd[4] = {0, 0, 0, 0}, ss[4] = {0, 0, 0, 0};
            float tw = 0.f, mir = 0.f;
            for (int oy = -2; oy <= 2; ++oy)
                for (int ox = -2; ox <= 2; ++ox) {
                    const int sx = x + ox, sy = y + oy;
                    float wn[4], a[4], b[4];
                    NB.at(sx, sy, wn);
                    const float wgt = calc_weight(std::sqrt(float(ox * ox + oy * oy)), O0.at(sx, sy) == cur,
                                                  ox == 0 && oy == 0, wn, cn);
                    SD.at(sx, sy, a);
                    SS.at(sx, sy, b);
                    for (int c = 0; c < 4; ++c) { sd[c] += a[c] * wgt; ss[c] += b[c] * wgt; }
                    tw += wgt;
                    const int qx = x + 2 * ox, qy = y + 2 * oy;
                    if (qx >= 0 && qy >= 0 && qx < w && qy < h) {
                        const float* m2 = motion + 2 * (size_t(qy) * W + qx);
                        mir += std::sqrt(m2[0] * m2[0] + m2[1] * m2[1]);
                    }
                }
            for (int c = 0; c < 4; ++c) { sd[c] /= tw; ss[c] /= tw; }
            p[1] = O1.at(ppx + 1, ppy);
            p[2] = O1.at(ppx, ppy + 1);
            p[3] = O1.at(ppx + 1, ppy + 1);
            if (std::fabs(mir) < 0.1f) p[0] = p[1] = p[2] = p[3] = cur;
            const float sx = fx - std::floor(fx), sy = fy - std::floor(fy);
            float hd[3], hs[3], ha[3];
            filter_history(PD, sx, sy, ppx, ppy, cur, p, sd, hd);
            filter_history(PS, sx, sy, ppx, ppy, cur, p, ss, hs);
            filter_history(PA, sx, sy, ppx, ppy, cur, p, ca, ha);
            float keep = 1.0f / float(params[1] > 1u ? params[1] : 1u);
            keep = keep < 0.f ? 0.f : (keep > 1.f ? 1.f : keep);
            for (int c = 0; c < 3; ++c) {
                ad[4 * i + c] = lerpf(hd[c], cd[c], keep);
                as[4 * i + c] = lerpf(hs[c], cs[c], keep);
                aa[4 * i + c] = lerpf(ha[c], ca[c], keep);
            }
        }
}

// LinearToST2084UE, Const_Func.slang:51-67
float st2084(float lin) {
    const float m1 = 0.1593017578125f, m2 = 78.84375f, c1 = 0.8359375f, c2 = 18.8515625f, c3 = 18.6875f;
    const float L = lin / 10000.f;
    const float Lm = std::pow(L, m1);
    const float N1 = c1 + c2 * Lm, N2 = 1.0f + c3 * Lm;
    return std::pow(N1 * (1.0f / N2), m2);
}

// the compose pass's output transform (Process.DenoiseJBF.comp.slang:170-181)
float display(float v, float paper_white, bool hdr) {
    if (!hdr) return gt(v * (paper_white / 40000.0f));
    const float t = gt(v / 2000.f) * 2000.f;
    return st2084(t * paper_white / 230.0f);
}

// fparams: {bf_sigma, bf_sigma_lum, paper_white_nit, hdr (0/1), linear (0/1)}, bf_size: 0
// disables the filter. out: W*H*4 display colour (alpha 1), or with linear the composed
// radiance before the display transform (the kernel's GK_FRAME_LINEAR_OUTPUT).
void oracle_jbf(const float* fparams, uint32_t W, uint32_t H, const float* acc_d, const float* acc_s,
                const float* acc_a, const uint32_t* bf_size, float* out) {
    const int w = int(W), h = int(H);
    const Img4 D{acc_d, w, h}, S{acc_s, w, h};
    const float sigma = fparams[0], sigma_l = fparams[1] * 100.0f, pw = fparams[2];
    const bool hdr = fparams[3] != 0.f;
    const bool linear = fparams[4] != 0.f;
    const float lc[3] = {0.212671f, 0.715160f, 0.072169f};
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = size_t(y) * W + x;
            float total[3];
            if (*bf_size > 0) {
                float c4[4], s4[4];
                D.at(x, y, c4);
                S.at(x, y, s4);
                float cc[3], cspec[3];
                for (int c = 0; c < 3; ++c) { cc[c] = c4[c] + 0.001f; cspec[c] = s4[c] + 0.001f; }
                const float clum = cc[0] * lc[0] + cc[1] * lc[1] + cc[2] * lc[2];
                float acc[3] = {0, 0, 0}, wsum = 0.f;
                for (int a = -5; a <= 5; a += 2)
                    for (int b = -5; b <= 5; b += 2) {
                        float q[4];
                        D.at(x + b, y + a, q);
                        float ci[3];
                        for (int c = 0; c < 3; ++c) ci[c] = q[c] + 0.001f;
                        const float lumi = ci[0] * lc[0] + ci[1] * lc[1] + ci[2] * lc[2];
                        float dist = float(a * a + b * b) / 25.f;
                        dist = dist < 0.f ? 0.f : (dist > 1.f ? 1.f : dist);
                        const float dl = (clum - lumi) * (clum - lumi);
                        const float Fi = std::exp(-dist * dist / (2.0f * sigma * sigma));
                        const float Li = std::exp(-dl * dl / (2.0f * sigma_l * sigma_l));
                        for (int c = 0; c < 3; ++c) acc[c] += ci[c] * (Fi * Li);
                        wsum += Fi * Li;
                    }
                for (int c = 0; c < 3; ++c) total[c] = acc[c] * (1.f / wsum) * acc_a[4 * i + c] + cspec[c];
            } else {
                for (int c = 0; c < 3; ++c) total[c] = acc_d[4 * i + c] * acc_a[4 * i + c] + acc_s[4 * i + c];
            }
            for (int c = 0; c < 3; ++c) out[4 * i + c] = linear ? total[c] : display(total[c], pw, hdr);
            out[4 * i + 3] = 1.f;
        }
}

// ---------------------------------------------------------------- resolve to the display
// src, out: float RGBA images (row-major); mode 0: vkCmdBlitImage with VK_FILTER_LINEAR
// (VulkanBaseRenderer.cpp:1050-1064), mode 1: Process.UpScaleFSR.comp.slang (FSR 1 EASU).
void oracle_upscale(int mode, const float* src, uint32_t sw_, uint32_t sh_, uint32_t dw_, uint32_t dh_, float* out) {
    const int sw = int(sw_), sh = int(sh_), dw = int(dw_), dh = int(dh_);
    auto texel = [&](int x, int y, float c[4]) {
        x = std::min(std::max(x, 0), sw - 1);
        y = std::min(std::max(y, 0), sh - 1);
        for (int k = 0; k < 4; ++k) c[k] = src[(size_t(y) * sw + x) * 4 + k];
    };
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x) {
            float* o = out + (size_t(y) * dw + x) * 4;
            if (mode == 0) {
                // destination centre in source texels, minus the texel-centre half
                const float u = (float(x) + 0.5f) * (float(sw) / float(dw)) - 0.5f;
                const float v = (float(y) + 0.5f) * (float(sh) / float(dh)) - 0.5f;
                const float fu = std::floor(u), fv = std::floor(v);
                const float a = u - fu, b = v - fv;
                float p[4], q[4], r[4], t[4];
                texel(int(fu), int(fv), p);
                texel(int(fu) + 1, int(fv), q);
                texel(int(fu), int(fv) + 1, r);
                texel(int(fu) + 1, int(fv) + 1, t);
                for (int k = 0; k < 4; ++k) {
                    const float top = p[k] * (1.f - a) + q[k] * a;
                    const float bot = r[k] * (1.f - a) + t[k] * a;
                    o[k] = top * (1.f - b) + bot * b;
                }
                continue;
            }
            if (sw == dw && sh == dh) {  // the shader's copy branch
                texel(x, y, o);
                continue;
            }
            // FsrEasuCon with input viewport = input size
            const float con0[4] = {float(sw) / float(dw), float(sh) / float(dh), 0.5f * float(sw) / float(dw) - 0.5f,
                                   0.5f * float(sh) / float(dh) - 0.5f};
            const float con1[4] = {1.f / float(sw), 1.f / float(sh), 1.f / float(sw), -1.f / float(sh)};
            const float con2[4] = {-1.f / float(sw), 2.f / float(sh), 1.f / float(sw), 2.f / float(sh)};
            const float con3[2] = {0.f, 4.f / float(sh)};
            const float ip[2] = {(float(x) / float(dw)) * float(dw), (float(y) / float(dh)) * float(dh)};
            float pp[2] = {ip[0] * con0[0] + con0[2], ip[1] * con0[1] + con0[3]};
            const float fp[2] = {std::floor(pp[0]), std::floor(pp[1])};
            pp[0] -= fp[0];
            pp[1] -= fp[1];
            const float p0[2] = {fp[0] * con1[0] + con1[2], fp[1] * con1[1] + con1[3]};
            const float base[4][2] = {{p0[0], p0[1]}, {p0[0] + con2[0], p0[1] + con2[1]},
                                      {p0[0] + con2[2], p0[1] + con2[3]}, {p0[0] + con3[0], p0[1] + con3[1]}};
            const float offx[2] = {-.5f * con1[0], .5f * con1[0]}, offy[2] = {-.5f * con1[1], .5f * con1[1]};
            // taps b c i j f e k l h g o n: (base, x offset -/+, y offset -/+)
            static const int tb[12][3] = {{0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}, {1, 0, 0},
                                          {2, 0, 1}, {2, 1, 1}, {2, 1, 0}, {2, 0, 0}, {3, 1, 0}, {3, 0, 0}};
            float C[12][4], Lm[12];
            for (int t = 0; t < 12; ++t) {
                const float px = base[tb[t][0]][0] + offx[tb[t][1]], py = base[tb[t][0]][1] + offy[tb[t][2]];
                texel(int(px * float(sw)), int(py * float(sh)), C[t]);
                Lm[t] = C[t][1] + 0.5f * (C[t][0] + C[t][2]);
            }
            enum { B, Cc, I, J, F, E, K, L, Hh, G, O, N };
            float dir[2] = {0.f, 0.f}, len = 0.f;
            auto setf = [&](float w, float lA, float lB, float lC, float lD, float lE) {
                float lenX = std::max(std::fabs(lD - lC), std::fabs(lC - lB));
                const float dirX = lD - lB;
                dir[0] += dirX * w;
                // clamp as the GPU does: a flat neighbourhood's 0 / 0 becomes 0 (maxNum)
                lenX = std::fmin(std::fmax(std::fabs(dirX) / lenX, 0.f), 1.f);
                lenX *= lenX;
                len += lenX * w;
                float lenY = std::max(std::fabs(lE - lC), std::fabs(lC - lA));
                const float dirY = lE - lA;
                dir[1] += dirY * w;
                lenY = std::fmin(std::fmax(std::fabs(dirY) / lenY, 0.f), 1.f);
                lenY *= lenY;
                len += lenY * w;
            };
            setf((1.f - pp[0]) * (1.f - pp[1]), Lm[B], Lm[E], Lm[F], Lm[G], Lm[J]);
            setf(pp[0] * (1.f - pp[1]), Lm[Cc], Lm[F], Lm[G], Lm[Hh], Lm[K]);
            setf((1.f - pp[0]) * pp[1], Lm[F], Lm[I], Lm[J], Lm[K], Lm[N]);
            setf(pp[0] * pp[1], Lm[G], Lm[J], Lm[K], Lm[L], Lm[O]);
            float dirR = dir[0] * dir[0] + dir[1] * dir[1];
            const bool zro = dirR < (1.0f / 32768.0f);
            dirR = 1.0f / std::sqrt(dirR);
            if (zro) { dirR = 1.0f; dir[0] = 1.0f; }
            dir[0] *= dirR;
            dir[1] *= dirR;
            len = len * 0.5f;
            len *= len;
            const float stretch = (dir[0] * dir[0] + dir[1] * dir[1]) / std::max(std::fabs(dir[0]), std::fabs(dir[1]));
            const float len2[2] = {1.f + (stretch - 1.0f) * len, 1.f - .5f * len};
            const float lob = .5f - .29f * len, clp = 1.f / lob;
            static const float toff[12][2] = {{0, -1}, {1, -1}, {-1, 1}, {0, 1}, {0, 0}, {-1, 0},
                                              {1, 1}, {2, 1}, {2, 0}, {1, 0}, {1, 2}, {0, 2}};
            float aC[4] = {0, 0, 0, 0}, aW = 0.f;
            for (int t = 0; t < 12; ++t) {
                const float ox = toff[t][0] - pp[0], oy = toff[t][1] - pp[1];
                float vx = ox * dir[0] + oy * dir[1];
                float vy = ox * -dir[1] + oy * dir[0];
                vx *= len2[0];
                vy *= len2[1];
                const float d2 = std::min(vx * vx + vy * vy, clp);
                float wB = .4f * d2 - 1.f, wA = lob * d2 - 1.f;
                wB *= wB;
                wA *= wA;
                wB = 1.5625f * wB - .5625f;
                const float w = wB * wA;
                for (int k = 0; k < 4; ++k) aC[k] += C[t][k] * w;
                aW += w;
            }
            for (int k = 0; k < 4; ++k) {
                const float mn = std::min(std::min(C[F][k], C[G][k]), std::min(C[J][k], C[K][k]));
                const float mx = std::max(std::max(C[F][k], C[G][k]), std::max(C[J][k], C[K][k]));
                o[k] = std::min(mx, std::max(mn, aC[k] / aW));
            }
        }
}

float oracle_st2084(float x) { return st2084(x); }
float oracle_display(float v, float paper_white, int hdr) { return display(v, paper_white, hdr != 0); }

}  // extern "C"

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
            const bool inside = ppx < w && ppy < h && ppx >= -1 && ppy >= -1;
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

// fparams: {bf_sigma, bf_sigma_lum, paper_white_nit}, bf_size: 0 disables the filter.
// out: W*H*4 tonemapped colour (alpha 1).
void oracle_jbf(const float* fparams, uint32_t W, uint32_t H, const float* acc_d, const float* acc_s,
                const float* acc_a, const uint32_t* bf_size, float* out) {
    const int w = int(W), h = int(H);
    const Img4 D{acc_d, w, h}, S{acc_s, w, h};
    const float sigma = fparams[0], sigma_l = fparams[1] * 100.0f, sc = fparams[2] / 40000.0f;
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
            for (int c = 0; c < 3; ++c) out[4 * i + c] = gt(total[c] * sc);
            out[4 * i + 3] = 1.f;
        }
}

}  // extern "C"

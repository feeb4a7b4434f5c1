// gk_post.hpp - display transform shared by the bilateral and learned denoisers.
#pragma once
#include <hip/hip_runtime.h>

namespace gk {

// W_f / H_f / GranTurismoTonemapper, Const_Func.slang:84-121
__device__ inline float w_f(float x, float e0, float e1) {
    if (x <= e0) return 0.f;
    if (x >= e1) return 1.f;
    const float a = (x - e0) / (e1 - e0);
    return a * a * (3.f - 2.f * a);
}
__device__ inline float h_f(float x, float e0, float e1) {
    if (x <= e0) return 0.f;
    if (x >= e1) return 1.f;
    return (x - e0) / (e1 - e0);
}
// x^y for x >= 0 as the hardware exp2(y * log2(x)) (relative error a few ulp, as a
// shader's pow): 0^y = 0 and inf^y = inf for y > 0, NaN for x < 0
__device__ inline float pow_hw(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }

__device__ inline float gt_tonemap(float x) {
    const float P = 1.f, a = 0.7f, m = 0.22f, l = 0.4f, c = 1.33f, b = 0.f;
    const float l0 = (P - m) * l / a;
    const float L_x = m + a * (x - m);
    const float T_x = m * pow_hw(x / m, c) + b;
    const float S0 = m + l0;
    const float S1 = m + a * l0;
    const float C2 = a * P / (P - S1);
    // 2.71828^y = exp2(y * log2(2.71828f)), log2(2.71828f) = 1.4426941f
    const float S_x = P - (P - S1) * __builtin_amdgcn_exp2f(-(C2 * (x - S0) / P) * 1.4426941f);
    const float w0 = 1.f - w_f(x, 0.f, m);
    const float w2 = h_f(x, m + l0, m + l0);
    const float w1 = 1.f - w0 - w2;
    return T_x * w0 + L_x * w1 + S_x * w2;
}

// LinearToST2084UE (Const_Func.slang:51-67): SMPTE ST 2084 (PQ) of luminance in nits
__device__ inline float st2084(float lin) {
    const float m1 = 0.1593017578125f, m2 = 78.84375f, c1 = 0.8359375f, c2 = 18.8515625f, c3 = 18.6875f;
    const float L = lin / 10000.f;
    const float Lm = pow_hw(L, m1);
    const float N1 = c1 + c2 * Lm, N2 = 1.0f + c3 * Lm;
    return pow_hw(N1 * (1.0f / N2), m2);
}

// the final transform of the compose pass (Process.DenoiseJBF.comp.slang:170-181): SDR
// GT-tonemaps radiance * PaperWhiteNit / 40000; HDR tonemaps radiance / 2000, scales back
// by 2000 and encodes radiance * PaperWhiteNit / 230 with ST 2084
__device__ inline float display_encode(float v, float paper_white, bool hdr) {
    if (!hdr) return gt_tonemap(v * (paper_white / 40000.0f));
    const float t = gt_tonemap(v / 2000.f) * 2000.f;
    return st2084(t * paper_white / 230.0f);
}

}  // namespace gk

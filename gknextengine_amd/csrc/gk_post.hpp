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
__device__ inline float gt_tonemap(float x) {
    const float P = 1.f, a = 0.7f, m = 0.22f, l = 0.4f, c = 1.33f, b = 0.f;
    const float l0 = (P - m) * l / a;
    const float L_x = m + a * (x - m);
    const float T_x = m * powf(x / m, c) + b;
    const float S0 = m + l0;
    const float S1 = m + a * l0;
    const float C2 = a * P / (P - S1);
    const float S_x = P - (P - S1) * powf(2.71828f, -(C2 * (x - S0) / P));
    const float w0 = 1.f - w_f(x, 0.f, m);
    const float w2 = h_f(x, m + l0, m + l0);
    const float w1 = 1.f - w0 - w2;
    return T_x * w0 + L_x * w1 + S_x * w2;
}

}  // namespace gk

// cornell_demo.cpp - renders the Cornell box through the C++ adapter (no Python).
//
//   gk_cornell_demo <width> <height> <frames> <estimator 0|1> <out.f32>
//
// Writes the FP32 radiance sum target (width*height RGBA floats, alpha = sample count)
// to out.f32 and one JSON line with the frame time and the ray count.
#include <chrono>
#include <cstdio>
#include <cstdlib>

#include "gknx_renderer.hpp"

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s width height frames estimator out.f32\n", argv[0]);
        return 2;
    }
    const uint32_t W = uint32_t(std::atoi(argv[1])), H = uint32_t(std::atoi(argv[2]));
    const uint32_t frames = uint32_t(std::atoi(argv[3]));
    const uint32_t estimator = uint32_t(std::atoi(argv[4]));
    try {
        gknx::PathTracingRenderer r;
        r.CreateSwapChain(W, H);
        gknx::CornellScene scene;
        r.OnPostLoadScene(scene.arrays);
        gk_camera_desc cam{{0.5f, 0.5f, -1.4f}, {0.5f, 0.5f, 0.f}, {0.f, 1.f, 0.f}, 40.f, 0.f, 10.f};
        const auto t0 = std::chrono::steady_clock::now();
        for (uint32_t f = 0; f < frames; ++f) {
            gk_frame_params& p = r.UpdateUniformBuffer(cam, f);
            p.estimator = estimator;
            p.bounces = p.max_bounces = 4;
            r.Render();
        }
        r.BeforeNextFrame();
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        const std::vector<float> rad = r.ReadTarget<float>(GK_RT_RADIANCE);
        FILE* f = std::fopen(argv[5], "wb");
        if (!f || std::fwrite(rad.data(), sizeof(float), rad.size(), f) != rad.size()) {
            std::fprintf(stderr, "cannot write %s\n", argv[5]);
            return 1;
        }
        std::fclose(f);
        const gk_stats s = r.Stats();
        std::printf("{\"width\": %u, \"height\": %u, \"frames\": %u, \"ms_per_frame\": %.4f, \"rays\": %llu}\n", W, H,
                    frames, ms / frames, (unsigned long long)s.rays_traced);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}

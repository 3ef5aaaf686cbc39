// `eleven <scene_dir> <spp> <output> [--mode plain|wavefront] [--depth N]
//         [--seed S] [--bvh 4|8] [--device D]`
//
// Headless replacement of the reference CLI (main.cpp:111-194): loadScene ->
// renderSetup -> renderCuda -> getBuffers -> clampPixels -> gamma -> write.
// The reference saves fastPow(x, 1/2.2) * 255 (truncated) of the clamped
// beauty pass for all four channels (main.cpp:150-160) with stbi_write_png,
// whatever the file name; here the container follows the extension (.bmp
// writes a BMP, anything else a PNG) and the process exits when done.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/eleven.h"
#include "scene.hpp"

static int usage() {
    std::fprintf(stderr,
                 "usage: eleven <scene_dir> <spp> <output.bmp|.png> [--mode plain|wavefront] [--depth N]\n"
                 "              [--seed S] [--bvh 4|8] [--device D]\n");
    return 2;
}

int main(int argc, char** argv) {
    if (argc < 4) return usage();
    std::string sceneDir = argv[1], out = argv[3];
    long spp = std::strtol(argv[2], nullptr, 10);
    if (spp < 0) {
        std::fprintf(stderr, "spp must be >= 0\n");
        return 2;
    }
    el_render_params p{};
    p.spp = (uint32_t)spp;
    p.max_depth = 5;
    p.mode = EL_MODE_WAVEFRONT;
    int device = 0;
    for (int i = 4; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char* { return i + 1 < argc ? argv[++i] : nullptr; };
        const char* v = nullptr;
        if (a == "--mode" && (v = next())) {
            p.mode = std::strcmp(v, "plain") == 0 ? EL_MODE_PLAIN : EL_MODE_WAVEFRONT;
        } else if (a == "--depth" && (v = next())) {
            p.max_depth = (uint32_t)std::atoi(v);
        } else if (a == "--seed" && (v = next())) {
            p.seed = std::strtoull(v, nullptr, 10);
        } else if (a == "--bvh" && (v = next())) {
            p.bvh_width = (uint32_t)std::atoi(v);
        } else if (a == "--device" && (v = next())) {
            device = std::atoi(v);
        } else {
            return usage();
        }
    }

    auto t0 = std::chrono::steady_clock::now();
    el_scene* scene = nullptr;
    int rc = el_scene_load(sceneDir.c_str(), &scene);
    if (rc) {
        std::fprintf(stderr, "loading %s failed: %s\n", sceneDir.c_str(), el_scene_last_error());
        return 1;
    }
    el_scene_view view;
    el_scene_view_get(scene, &view);
    std::printf("scene %s: %u tris, %u materials, %u textures, %u lights, %ux%u\n", sceneDir.c_str(), view.tri_count,
                view.material_count, view.texture_count, view.light_count, view.camera.x_res, view.camera.y_res);

    el_ctx* ctx = nullptr;
    rc = el_create(&ctx, device);
    if (rc) {
        std::fprintf(stderr, "no usable HIP device (status %d)\n", rc);
        el_scene_free(scene);
        return 1;
    }
    auto t1 = std::chrono::steady_clock::now();
    rc = el_upload_scene(ctx, &view);
    if (rc) {
        std::fprintf(stderr, "upload failed: %s\n", el_last_error(ctx));
        el_destroy(ctx);
        el_scene_free(scene);
        return 1;
    }
    auto t2 = std::chrono::steady_clock::now();
    rc = el_render(ctx, &p);
    if (rc) {
        std::fprintf(stderr, "render failed: %s\n", el_last_error(ctx));
        el_destroy(ctx);
        el_scene_free(scene);
        return 1;
    }
    el_stats st;
    el_get_stats(ctx, &st);
    size_t W = view.camera.x_res, H = view.camera.y_res;
    std::vector<float> rgba(W * H * 4);
    uint64_t rays = 0, verts = 0;
    el_read_passes(ctx, rgba.data(), nullptr, &rays, &verts);

    std::vector<unsigned char> img(W * H * 4);
    for (size_t i = 0; i < W * H * 4; i++) {
        float x = rgba[i];
        x = x < 0 ? 0 : (x > 1 ? 1 : x);   // clampPixels (PostProcessing.cpp:30-33)
        img[i] = (unsigned char)(eleven::fastPow(x, 1.0 / 2.2) * 255);
    }
    bool bmp = out.size() >= 4 && (out.compare(out.size() - 4, 4, ".bmp") == 0 || out.compare(out.size() - 4, 4, ".BMP") == 0);
    bool ok = bmp ? eleven::writeBMP(out, (int)W, (int)H, img.data()) : eleven::writePNG(out, (int)W, (int)H, img.data());
    double loadMs = std::chrono::duration<double, std::milli>(t1 - t0).count();
    double setupMs = std::chrono::duration<double, std::milli>(t2 - t1).count();
    std::printf("load %.1f ms, setup (BVH + upload) %.1f ms, render %.1f ms (%s, depth %u, %u spp)\n", loadMs, setupMs,
                st.render_ms, p.mode == EL_MODE_PLAIN ? "plain" : "wavefront", p.max_depth ? p.max_depth : 5, p.spp);
    std::printf("kPaths/s: %.1f  Mrays/s: %.1f  (%llu vertices, %llu rays)\n", verts / st.render_ms,
                rays / st.render_ms / 1000.0, (unsigned long long)verts, (unsigned long long)rays);
    el_destroy(ctx);
    el_scene_free(scene);
    if (!ok) {
        std::fprintf(stderr, "cannot write %s\n", out.c_str());
        return 1;
    }
    std::printf("saved %s\n", out.c_str());
    return 0;
}

// Host code under AddressSanitizer + UndefinedBehaviorSanitizer: the scene
// loader (.json/.obj/.mtl/.bmp/.hdr parsing, scene_loader.cpp), the
// MikkTSpace tangents it runs per mesh (tangent_space.cpp), the BVH build,
// wide collapse and quantization (bvh.cpp), the HDRI sampler tables
// (hdri_table.cpp) and the scene C ABI (scene_capi.cpp).
//
//   host_sanitize <scene dir>... -- <scratch dir>
//
// Every good scene is loaded through the C ABI, built into BVH2 / BVH4 /
// BVH8, and its HDRI tables are built and checked against the reference
// search on a stride of random numbers.  Then each file of the first scene is
// damaged -- truncated at many lengths, bytes overwritten, header fields made
// huge or negative -- and the damaged scene is loaded: the loader must
// either succeed or report an error (LoadError / a non-zero status), and the
// sanitizers must stay silent.  Exits non-zero on any mismatch; a sanitizer
// finding aborts the process (-fno-sanitize-recover=all).
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "../../include/eleven.h"
#include "bvh.hpp"
#include "hdri_table.hpp"
#include "scene.hpp"

namespace fs = std::filesystem;
using namespace eleven;

static int bad = 0;
#define CHECK(c, ...)                          \
    do {                                       \
        if (!(c)) {                            \
            if (bad < 20) printf(__VA_ARGS__); \
            bad++;                             \
        }                                      \
    } while (0)

static std::vector<char> readAll(const fs::path& p) {
    std::ifstream f(p, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}

static void writeAll(const fs::path& p, const std::vector<char>& b) {
    std::ofstream f(p, std::ios::binary | std::ios::trunc);
    f.write(b.data(), (std::streamsize)b.size());
}

// A scene loads through the C ABI; returns the status.
static int loadViaAbi(const std::string& dir, bool expectOk) {
    el_scene* s = nullptr;
    const int rc = el_scene_load(dir.c_str(), &s);
    if (rc == EL_OK) {
        el_scene_view v{};
        CHECK(el_scene_view_get(s, &v) == EL_OK, "view of %s\n", dir.c_str());
        // touch every array the view exposes
        double acc = 0;
        for (uint32_t i = 0; i < v.tri_count * 9; i++) acc += v.vertices[i] + v.normals[i] + v.tangents[i];
        for (uint32_t i = 0; i < v.tri_count * 6; i++) acc += v.uvs[i];
        for (uint32_t i = 0; i < v.tri_count; i++) {
            acc += v.tangent_sign[i];
            CHECK(v.object_id[i] >= 0 && (uint32_t)v.object_id[i] < v.mesh_count, "object id\n");
        }
        for (uint32_t m = 0; m < v.mesh_count; m++)
            CHECK(v.mesh_material[m] >= 0 && (uint32_t)v.mesh_material[m] < v.material_count, "mesh material\n");
        for (uint32_t t = 0; t < v.texture_count; t++) {
            const el_texture& tx = v.textures[t];
            CHECK(tx.width > 0 && tx.height > 0, "texture size\n");
            acc += tx.data[0] + tx.data[(size_t)tx.width * tx.height * 3 - 1];
        }
        acc += v.hdri.data[(size_t)v.hdri.width * v.hdri.height * 3 - 1];
        (void)acc;
        el_scene_free(s);
    } else {
        CHECK(s == nullptr, "failed load left a scene\n");
    }
    if (expectOk) CHECK(rc == EL_OK, "%s did not load (%d)\n", dir.c_str(), rc);
    return rc;
}

static void checkScene(const std::string& dir) {
    Scene sc = loadScene(dir);
    const int n = sc.triCount();
    std::vector<float> v((size_t)n * 9);
    for (int i = 0; i < n; i++)
        for (int k = 0; k < 3; k++) {
            v[i * 9 + k * 3 + 0] = sc.tris[i].vertices[k].x;
            v[i * 9 + k * 3 + 1] = sc.tris[i].vertices[k].y;
            v[i * 9 + k * 3 + 2] = sc.tris[i].vertices[k].z;
        }
    if (n > 0) {
        Bvh2 b = buildBvh2(v.data(), n, 4);
        for (int w : {4, 8}) {
            WideBvh wb = collapseWide(b, v.data(), n, w);
            CHECK((int)wb.order.size() == n, "order size\n");
            CHECK(wb.triStream.size() == (size_t)n * 12, "tri stream size\n");
            std::vector<char> seen(n, 0);
            for (int i : wb.order) {
                CHECK(i >= 0 && i < n && !seen[i], "order is not a permutation\n");
                if (i >= 0 && i < n) seen[i] = 1;
            }
        }
    }
    const Texture& t = sc.hdri.texture;
    HdriSampler hs = buildHdriSampler(t.data.data(), t.width, t.height);
    CHECK(hs.length == t.width * t.height, "hdri length\n");
    std::vector<uint32_t> bucket = hdriBadBuckets(hs, t.data.data(), t.width, t.height, t.xTile, t.yTile,
                                                  t.xOffset, t.yOffset);
    CHECK(bucket.size() == (size_t)(1 << 16) / 32, "bucket bitmap size %zu\n", bucket.size());
    for (uint32_t k = 0; k < (1u << 24); k += 4093) {
        const float r = (float)k * (1.0f / 16777216.0f);
        const int a = hdriLookup(hs.cdf.data(), hs.guide.data(), hs.map2.data(), hs.log2G, r);
        const int e = hdriBinarySearch(hs.cdf.data(), r, hs.length);
        CHECK(a == e, "hdri lookup %d vs search %d at r=%g\n", a, e, r);
    }
    loadViaAbi(dir, true);
}

// Load a damaged copy of `src` where `file` is replaced by `bytes`.
static int tried = 0, failed = 0;
static void tryDamaged(const fs::path& src, const fs::path& work, const std::string& file,
                       const std::vector<char>& bytes) {
    fs::remove_all(work);
    fs::create_directories(work);
    for (const auto& e : fs::directory_iterator(src)) fs::copy(e.path(), work / e.path().filename());
    writeAll(work / file, bytes);
    tried++;
    try {
        Scene s = loadScene(work.string());
        (void)s;
    } catch (const LoadError&) {
        failed++;
    }
    loadViaAbi(work.string(), false);
}

static void damage(const fs::path& src, const fs::path& scratch) {
    std::mt19937 g(7);
    const fs::path work = scratch / "damaged";
    for (const auto& e : fs::directory_iterator(src)) {
        const std::string name = e.path().filename().string();
        const std::vector<char> good = readAll(e.path());
        if (good.empty()) continue;
        // truncations: a spread of lengths, denser near the start (headers)
        std::vector<size_t> cuts = {0, 1, 2, 8, 14, 18, 26, 34, 54, 60, 100};
        for (int k = 1; k < 12; k++) cuts.push_back(good.size() * k / 12);
        for (size_t c : cuts)
            if (c < good.size()) tryDamaged(src, work, name, std::vector<char>(good.begin(), good.begin() + c));
        // overwritten bytes (headers first, then anywhere)
        for (int it = 0; it < 24; it++) {
            std::vector<char> b = good;
            const size_t span = it < 12 ? std::min<size_t>(b.size(), 64) : b.size();
            for (int j = 0; j < 1 + (int)(g() % 4); j++) b[g() % span] = (char)(g() & 0xff);
            tryDamaged(src, work, name, b);
        }
        // text files: numbers made huge, negative or malformed
        if (name.size() > 4 && (name.substr(name.size() - 4) == ".obj" || name.substr(name.size() - 4) == ".mtl" ||
                                name.substr(name.size() - 5) == ".json")) {
            const char* subs[] = {"-99999999", "1e39", "nan", "-1", "2147483647", "", "0x", "--"};
            for (const char* sub : subs) {
                std::string s(good.begin(), good.end());
                size_t at = 0;
                for (int k = 0; k < 6; k++) {
                    at = s.find_first_of("0123456789", at + (g() % 200));
                    if (at == std::string::npos) break;
                    size_t end = s.find_first_not_of("0123456789.", at);
                    if (end == std::string::npos) end = s.size();
                    s.replace(at, end - at, sub);
                }
                tryDamaged(src, work, name, std::vector<char>(s.begin(), s.end()));
            }
        }
    }
    fs::remove_all(work);
}

int main(int argc, char** argv) {
    std::vector<std::string> scenes;
    std::string scratch;
    for (int i = 1; i < argc; i++) {
        if (std::string(argv[i]) == "--" && i + 1 < argc) {
            scratch = argv[i + 1];
            break;
        }
        scenes.push_back(argv[i]);
    }
    if (scenes.empty() || scratch.empty()) {
        fprintf(stderr, "usage: host_sanitize <scene dir>... -- <scratch dir>\n");
        return 2;
    }
    for (const auto& s : scenes) {
        checkScene(s);
        printf("scene %s checked\n", s.c_str());
    }
    damage(scenes[0], scratch);
    printf("damaged loads: %d tried, %d reported an error\n", tried, failed);
    // the loader must notice damage at least sometimes, or the cases are vacuous
    CHECK(failed > 0 && failed < tried, "damage detection %d of %d\n", failed, tried);
    printf("bad=%d\n", bad);
    return bad ? 1 : 0;
}

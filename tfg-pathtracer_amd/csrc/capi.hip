// C-ABI device context: scene upload (renderSetup), rendering (renderCuda),
// readback (getBuffers / getSamples) and the ray-query hooks.
//
// Reference host side (src/tfg-pathtracer/kernel.cu:485-724) deep-copies the
// AoS scene with one cudaMemcpy per array plus "pointer binding", keeps the
// accumulation in fixed 1920x1080 device globals and reads it back on a
// fresh, unsynchronised stream.  Here one context owns one stream; the scene
// becomes a handful of SoA buffers; accumulation is sized to the camera; and
// every readback is ordered after the last render on the same stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/eleven.h"
#include "bvh.hpp"
#include "hdri_table.hpp"
#include "host_parallel.hpp"
#include "kernels.hpp"
#include "pt_device.hpp"
#include "scene.hpp"

using namespace eleven;
using eleven::dev::DevScene;

namespace {

struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t ensure(size_t bytes) {
        if (bytes <= n && p) return hipSuccess;
        release();
        if (bytes == 0) bytes = 16;
        hipError_t e = hipMalloc(&p, bytes);
        if (e == hipSuccess) n = bytes;
        return e;
    }
    template <class T>
    T* as() const {
        return (T*)p;
    }
};

constexpr int kCounterWords = kXcds * kStatSlots;   // per-XCD copies of the tallies (kernels.hpp)

// Sum of the per-XCD copies of tally `slot`.
inline unsigned long long tally(const unsigned long long* h, int slot) {
    unsigned long long s = 0;
    for (int x = 0; x < kXcds; x++) s += h[x * kStatSlots + slot];
    return s;
}

}  // namespace

// ---------------------------------------------------- static device scratch --
// The scratch of el_trace_closest_noalloc: the tallies (kCounterWords 64-bit
// words) and the traversal stack's per-thread spill area, one array per GPU
// that lives until the process ends.  Host code never passes `scratch`
// itself to a kernel (on the host it is only a stand-in for the symbol): each
// device's address comes from hipGetSymbolAddress, looked up once per device
// by el_create with that device current.
constexpr size_t kScratchCounterBytes = 512;                  // >= kCounterWords * 8
constexpr size_t kScratchSpillInts = size_t(4) << 20;         // 16 MiB of spill slots
constexpr size_t kScratchBytes = kScratchCounterBytes + kScratchSpillInts * sizeof(int);
static_assert(kScratchCounterBytes >= sizeof(unsigned long long) * kXcds * kStatSlots, "tallies fit");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratchMu;
static std::vector<void*> g_scratchPtr;   // per device, null until looked up

// Device address of `scratch` on the current device `device`.
static hipError_t scratchAddress(int device, void** out) {
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if ((size_t)device >= g_scratchPtr.size()) g_scratchPtr.resize(device + 1, nullptr);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[device] = p;
    }
    *out = g_scratchPtr[device];
    return hipSuccess;
}

struct el_ctx {
    int device = 0;
    void* scratch = nullptr;       // this device's static scratch (scratchAddress)
    hipStream_t stream = nullptr;
    std::string err;
    int cus = 0;
    uint32_t maxThreads = 0;
    uint32_t block = kDefaultBlock;
    uint32_t refill = kDefaultRefill;

    bool hasScene = false;
    uint32_t ntri = 0;
    std::vector<float> verts;
    Bvh2 bvh2;
    WideBvh wide;
    int width = 4;
    int gcap = 1;

    DevScene ds{};
    el_camera cam{};
    std::vector<dev::DevMaterial> mats;
    std::vector<dev::DevTex> texs;
    std::vector<uint64_t> texF32;   // FP32-pool offset of each texture (el_set_texture)

    DevBuf dNodes, dTris, dShade, dMats, dTexs, dTexels, dLights, dEnvCdf, dEnvGuide, dEnvMap, dEnvBad, dEnvTab, dTexels8, dTexLut, dTangents;
    DevBuf dGstack, dAccum, dCounts, dAov, dCounters;
    DevBuf wRay0, wRay1, wHits, wShH, wShP, wT, wL, wPix, wSample, wQuiet, wAov, wCounters, wPend;
    size_t wfCap = 0, wfQcap = 0;

    DevBuf tRays, tTmax, tTri, tT, tUV, tOcc;   // scratch for host-pointer ray queries

    uint32_t sampleBase = 0;
    uint64_t accSeed = 0;          // seed / depth / RNG scene word of the renders in the accumulation
    uint32_t accDepth = 0, accScene = 0;
    bool accMixed = false;         // renders since el_clear used different seeds or depths
    DevBuf dRec;                   // per (sample, pixel) path radiance of a recorded render
    DevBuf dRecHits, dRecVis, wVis;  // its vertices (wavefront recordings: RenderArgs::recHits / recVis)
    bool recHits = false;          // with recValid: dRecHits / dRecVis hold the recorded render's vertices
    bool adjSumValid = false;      // aMat / aTex hold the sums of the el_adjoint calls so far
    int32_t adjSumTex = -1;        // their grad_texture
    bool recValid = false;         // dRec covers the whole accumulation
    uint64_t totalRays = 0, totalVerts = 0, totalPaths = 0;
    DevBuf aImage, aMat, aTex;     // adjoint image, FP64 material and FP32 texel gradients
    DevBuf aCache;                 // el_adjoint: each grid thread's vertices of the path it replays
    DevBuf aTarget, aLoss;         // el_residual_l2: target image, FP64 loss
    size_t residualPx = 0;         // aImage holds el_residual_l2's residual for this many pixels (0: none)
    el_stats stats{};
    bool aovOn = false;

    std::vector<hipEvent_t> evPool;
    // Wavefront renders: the shadow launches of bounce b run on `shadowStream`
    // (HDRI) and `shadowStreamP` (point lights) while the extension launch of
    // bounce b + 1 runs on `stream`; syncPool holds the events that order
    // them (no timing).
    hipStream_t shadowStream = nullptr, shadowStreamP = nullptr;
    std::vector<hipEvent_t> syncPool;
};

static int fail(el_ctx* c, int code, const std::string& what) {
    if (c) c->err = what;
    return code;
}
static int failHip(el_ctx* c, const char* what, hipError_t e) {
    return fail(c, EL_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define EL_HIP(ctx, x)                                   \
    do {                                                 \
        hipError_t e_ = (x);                             \
        if (e_ != hipSuccess) return failHip(ctx, #x, e_); \
    } while (0)

static hipError_t upload(DevBuf& b, const void* src, size_t bytes, hipStream_t st) {
    hipError_t e = b.ensure(bytes);
    if (e != hipSuccess) return e;
    if (bytes) e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st);
    return e;
}

static LaunchCfg cfgOf(el_ctx* c) {
    LaunchCfg l;
    l.stream = c->stream;
    l.block = c->block;
    l.shadeBlock = 256;
    l.refill = c->refill;
    l.maxThreads = c->maxThreads;
    l.gstack = c->dGstack.as<int>();
    return l;
}

extern "C" {

int el_device_count(int* out) {
    if (!out) return EL_ERR_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *out = n;
    return EL_OK;
}

int el_create(el_ctx** out, int device) {
    if (!out) return EL_ERR_ARG;
    *out = nullptr;
    std::unique_ptr<el_ctx> c(new el_ctx());
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return EL_ERR_HIP;
    if (device < 0 || device >= n) return EL_ERR_ARG;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess) return EL_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return EL_ERR_HIP;
    c->cus = prop.multiProcessorCount;
    // resident-thread capacity: 32 waves of 64 lanes per CU
    c->maxThreads = (uint32_t)c->cus * 2048u;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) return EL_ERR_HIP;
    if (hipStreamCreateWithFlags(&c->shadowStream, hipStreamNonBlocking) != hipSuccess) return EL_ERR_HIP;
    if (hipStreamCreateWithFlags(&c->shadowStreamP, hipStreamNonBlocking) != hipSuccess) return EL_ERR_HIP;
    if (c->dCounters.ensure(sizeof(unsigned long long) * kCounterWords) != hipSuccess) return EL_ERR_HIP;
    if (scratchAddress(device, &c->scratch) != hipSuccess) return EL_ERR_HIP;
    *out = c.release();
    return EL_OK;
}

void el_destroy(el_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->shadowStream) (void)hipStreamSynchronize(c->shadowStream);
    if (c->shadowStreamP) (void)hipStreamSynchronize(c->shadowStreamP);
    for (auto e : c->evPool) (void)hipEventDestroy(e);
    for (auto e : c->syncPool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    if (c->shadowStream) (void)hipStreamDestroy(c->shadowStream);
    if (c->shadowStreamP) (void)hipStreamDestroy(c->shadowStreamP);
    delete c;
}

const char* el_last_error(const el_ctx* c) { return c ? c->err.c_str() : "null context"; }

}  // extern "C"

// Pack and upload the wide BVH of the current width.
// 8-bit copy of an RGB FP32 texture: R | G << 8 | B << 16 per texel, each
// byte the index of the value in `table` (bit-exact); false if some value is
// not in the table.  The table is non-decreasing (the loader's fastPow of
// i / 255), so each value is found by binary search.
static bool packLdr(const float* rgb, size_t n, const float (&table)[256], uint32_t* out) {
    std::atomic<bool> ok{true};
    parallelFor(n, [&](size_t i0, size_t i1) {
        for (size_t i = i0; i < i1 && ok.load(std::memory_order_relaxed); i++) {
            uint32_t w = 0;
            for (int ch = 0; ch < 3; ch++) {
                const float v = rgb[3 * i + ch];
                const float* p = std::lower_bound(table, table + 256, v);
                if (p == table + 256 || std::memcmp(p, &v, 4) != 0) {
                    ok = false;
                    return;
                }
                w |= (uint32_t)(p - table) << (8 * ch);
            }
            out[i] = w;
        }
    });
    return ok;
}

// The device camera: the host fields plus the six trig values the reference
// evaluates per ray (kernel.cu:304-306) after rotation *= (PI / 180.0) as
// Vector3 *= float (kernel.cu:301).
static dev::DevCamera devCamera(const el_camera& cam) {
    dev::DevCamera dc{};
    dc.xRes = cam.x_res;
    dc.yRes = cam.y_res;
    dc.focalLength = cam.focal_length;
    dc.sensorWidth = cam.sensor_width;
    dc.sensorHeight = cam.sensor_height;
    dc.aperture = cam.aperture;
    dc.focusDistance = cam.focus_distance;
    const float toRad = (float)(3.14159265358979323846f / 180.0);
    const float rx = cam.rotation.x * toRad, ry = cam.rotation.y * toRad, rz = cam.rotation.z * toRad;
    dc.cosRot[0] = std::cos(rx);
    dc.sinRot[0] = std::sin(rx);
    dc.cosRot[1] = std::cos(ry);
    dc.sinRot[1] = std::sin(ry);
    dc.cosRot[2] = std::cos(rz);
    dc.sinRot[2] = std::sin(rz);
    dc.pos[0] = cam.position.x;
    dc.pos[1] = cam.position.y;
    dc.pos[2] = cam.position.z;
    dc.bokeh = cam.bokeh;
    return dc;
}

// Material values as the kernels read them (texture ids are set by the
// caller); for untextured roughness and metallic also generateHitData's
// powf(x, 2.2f) (kernel.cu:113-114), once per material.
static void setMaterialValues(dev::DevMaterial& d, const el_material& m) {
    d.albedo[0] = m.albedo.x;
    d.albedo[1] = m.albedo.y;
    d.albedo[2] = m.albedo.z;
    d.emission[0] = m.emission.x;
    d.emission[1] = m.emission.y;
    d.emission[2] = m.emission.z;
    d.roughness = m.roughness;
    d.metallic = m.metallic;
    d.roughnessEff = powf(m.roughness, 2.2f);
    d.metallicEff = powf(m.metallic, 2.2f);
    d.clearcoatGloss = m.clearcoat_gloss;
    d.clearcoat = m.clearcoat;
    d.anisotropic = m.anisotropic;
    d.eta = m.eta;
    d.transmission = m.transmission;
    d.specular = m.specular;
    d.specularTint = m.specular_tint;
    d.sheenTint = m.sheen_tint;
    d.subsurface = m.subsurface;
    d.sheen = m.sheen;
}

static int uploadBvh(el_ctx* c, int width) {
    try {
        c->wide = collapseWide(c->bvh2, c->verts.data(), (int)c->ntri, width);
    } catch (const std::exception& ex) {
        return fail(c, EL_ERR_RANGE, ex.what());
    }
    c->width = width;
    EL_HIP(c, upload(c->dNodes, c->wide.nodes.data(), c->wide.nodes.size(), c->stream));
    EL_HIP(c, upload(c->dTris, c->wide.triStream.data(), c->wide.triStream.size() * sizeof(float), c->stream));
    c->gcap = std::max(1, c->wide.stackBound + 2 - dev::kShortStack);
    EL_HIP(c, c->dGstack.ensure((size_t)kSpillSets * c->gcap * c->maxThreads * sizeof(int)));
    c->ds.nodes = c->dNodes.p;
    c->ds.tris = c->dTris.as<float4>();
    c->ds.width = width;
    c->ds.gcap = c->gcap;
    c->stats.bvh_width = (uint32_t)width;
    c->stats.stack_bound = (uint32_t)c->wide.stackBound;
    c->stats.node_bytes = c->wide.nodes.size();
    c->stats.tri_bytes = c->wide.triStream.size() * sizeof(float);
    c->stats.node_stride = (uint32_t)c->wide.nodeStride;
    c->stats.tri_stride = 12 * sizeof(float);
    EL_HIP(c, hipStreamSynchronize(c->stream));
    return EL_OK;
}

// Releases the buffers of a recording render (EL_FLAG_RECORD_PATHS), which
// cover only the accumulation they were recorded into.
static void releaseRecord(el_ctx* c) {
    c->dRec.release();
    c->dRecHits.release();
    c->dRecVis.release();
    c->wVis.release();
    c->recValid = false;
    c->recHits = false;
}

static int clearAccum(el_ctx* c) {
    size_t npx = (size_t)c->cam.x_res * c->cam.y_res;
    EL_HIP(c, hipMemsetAsync(c->dAccum.p, 0, npx * sizeof(float4), c->stream));
    EL_HIP(c, hipMemsetAsync(c->dCounts.p, 0, npx * sizeof(uint32_t), c->stream));
    if (c->dAov.p) EL_HIP(c, hipMemsetAsync(c->dAov.p, 0, c->dAov.n, c->stream));
    c->sampleBase = 0;
    c->accMixed = false;
    releaseRecord(c);
    c->totalRays = c->totalVerts = c->totalPaths = 0;
    EL_HIP(c, hipStreamSynchronize(c->stream));
    return EL_OK;
}

extern "C" {

int el_upload_scene(el_ctx* c, const el_scene_view* v) {
    if (!c || !v) return fail(c, EL_ERR_ARG, "null argument");
    if (v->tri_count && (!v->vertices || !v->object_id)) return fail(c, EL_ERR_ARG, "missing triangle arrays");
    if (v->camera.x_res == 0 || v->camera.y_res == 0) return fail(c, EL_ERR_ARG, "camera resolution is zero");
    if ((uint64_t)v->camera.x_res * v->camera.y_res > (1ull << 28))
        return fail(c, EL_ERR_RANGE, "resolution above 2^28 pixels");
    if (v->tri_count >= (uint32_t)kLeafFirstLimit) return fail(c, EL_ERR_RANGE, "more than 2^27 triangles");
    if (v->material_count == 0 || !v->materials) return fail(c, EL_ERR_ARG, "at least one material is required");
    if (v->mesh_count == 0 || !v->mesh_material) return fail(c, EL_ERR_ARG, "mesh table missing");
    if (!v->hdri.data || v->hdri.width <= 0 || v->hdri.height <= 0) return fail(c, EL_ERR_ARG, "hdri missing");
    if (v->light_count && !v->lights) return fail(c, EL_ERR_ARG, "lights missing");
    for (uint32_t i = 0; i < v->texture_count; i++)
        if (!v->textures[i].data || v->textures[i].width <= 0 || v->textures[i].height <= 0)
            return fail(c, EL_ERR_ARG, "texture " + std::to_string(i) + " is empty");
    EL_HIP(c, hipSetDevice(c->device));
    c->hasScene = false;
    c->adjSumValid = false;
    c->residualPx = 0;
    c->ntri = v->tri_count;
    c->cam = v->camera;

    // geometry + BVH
    c->verts.assign(v->vertices, v->vertices + (size_t)v->tri_count * 9);
    c->bvh2 = buildBvh2(c->verts.data(), (int)c->ntri);
    int rc = uploadBvh(c, c->width);
    if (rc) return rc;

    // shading records and corner tangents, indexed by original triangle
    std::vector<float> shade((size_t)v->tri_count * dev::kShadeFloats, 0.f);
    std::vector<float> tans((size_t)v->tri_count * dev::kTanFloats, 0.f);
    parallelFor(v->tri_count, [&](size_t i0, size_t i1) {
        for (uint32_t i = (uint32_t)i0; i < (uint32_t)i1; i++) {
            float* S = &shade[(size_t)i * dev::kShadeFloats];
            std::memcpy(S, v->vertices + (size_t)i * 9, 9 * sizeof(float));
            if (v->normals) std::memcpy(S + 9, v->normals + (size_t)i * 9, 9 * sizeof(float));
            if (v->uvs) std::memcpy(S + 18, v->uvs + (size_t)i * 6, 6 * sizeof(float));
            int32_t obj = v->object_id[i];
            int32_t mat = (obj >= 0 && (uint32_t)obj < v->mesh_count) ? v->mesh_material[obj] : 0;
            if (mat < 0 || (uint32_t)mat >= v->material_count) mat = 0;
            std::memcpy(&S[24], &mat, 4);
            S[25] = v->tangent_sign ? v->tangent_sign[i] : 0.f;
            if (v->tangents) std::memcpy(&tans[(size_t)i * dev::kTanFloats], v->tangents + (size_t)i * 9, 9 * sizeof(float));
        }
    });
    EL_HIP(c, upload(c->dShade, shade.data(), shade.size() * sizeof(float), c->stream));
    EL_HIP(c, upload(c->dTangents, tans.data(), tans.size() * sizeof(float), c->stream));

    // textures: one FP32 pool, the environment map last.  A texture whose
    // every value is an entry of one of the loader's byte -> value tables
    // (an 8-bit file) also gets an 8-bit copy, which the kernels read.
    std::vector<float> pool;
    std::vector<uint32_t> pool8;
    float tables[2][256];
    ldrTable(sRGB, tables[0]);
    ldrTable(LINEAR, tables[1]);
    c->texs.clear();
    c->texF32.clear();
    auto addTex = [&](const el_texture& t, bool mayPack) {
        dev::DevTex d{};
        d.w = t.width;
        d.h = t.height;
        d.xTile = t.x_tile;
        d.yTile = t.y_tile;
        d.xOffset = t.x_offset;
        d.yOffset = t.y_offset;
        d.filter = t.filter;
        d.format = dev::kTexF32;
        d.offset = pool.size();
        const size_t n = (size_t)t.width * t.height;
        pool.insert(pool.end(), t.data, t.data + n * 3);
        for (int k = 0; mayPack && k < 2; k++) {
            std::vector<uint32_t> words(n);
            if (!packLdr(t.data, n, tables[k], words.data())) continue;
            d.format = dev::kTexLdr + k;
            d.offset = pool8.size();
            pool8.insert(pool8.end(), words.begin(), words.end());
            break;
        }
        return d;
    };
    for (uint32_t i = 0; i < v->texture_count; i++) {
        c->texF32.push_back(pool.size());
        c->texs.push_back(addTex(v->textures[i], true));
    }
    dev::DevTex env = addTex(v->hdri, false);
    EL_HIP(c, upload(c->dTexels, pool.data(), pool.size() * sizeof(float), c->stream));
    EL_HIP(c, upload(c->dTexels8, pool8.data(), pool8.size() * sizeof(uint32_t), c->stream));
    EL_HIP(c, upload(c->dTexLut, tables, sizeof tables, c->stream));
    c->stats.texels_8bit = pool8.size();
    EL_HIP(c, upload(c->dTexs, c->texs.data(), c->texs.size() * sizeof(dev::DevTex), c->stream));

    // materials
    c->mats.assign(v->material_count, dev::DevMaterial{});
    for (uint32_t i = 0; i < v->material_count; i++) {
        const el_material& m = v->materials[i];
        dev::DevMaterial& d = c->mats[i];
        auto texOk = [&](int32_t id) { return (id >= 0 && (uint32_t)id < v->texture_count) ? id : -1; };
        d.albedoTex = texOk(m.albedo_tex);
        d.emissionTex = texOk(m.emission_tex);
        d.roughnessTex = texOk(m.roughness_tex);
        d.metallicTex = texOk(m.metallic_tex);
        d.normalTex = texOk(m.normal_tex);
        setMaterialValues(d, m);
    }
    EL_HIP(c, upload(c->dMats, c->mats.data(), c->mats.size() * sizeof(dev::DevMaterial), c->stream));

    // lights
    std::vector<float> lights((size_t)v->light_count * 6);
    for (uint32_t i = 0; i < v->light_count; i++) {
        const el_point_light& l = v->lights[i];
        float* L = &lights[(size_t)i * 6];
        L[0] = l.position.x;
        L[1] = l.position.y;
        L[2] = l.position.z;
        L[3] = l.radiance.x;
        L[4] = l.radiance.y;
        L[5] = l.radiance.z;
    }
    EL_HIP(c, upload(c->dLights, lights.data(), lights.size() * sizeof(float), c->stream));

    // HDRI CDF exactly as HDRI::generateCDF at construction (offsets still 0),
    // then the alias table of the search's effective distribution.
    const el_texture& h = v->hdri;
    HdriSampler hs = buildHdriSampler(h.data, h.width, h.height);
    EL_HIP(c, upload(c->dEnvCdf, hs.cdf.data(), hs.cdf.size() * sizeof(float), c->stream));
    EL_HIP(c, upload(c->dEnvGuide, hs.guide.data(), hs.guide.size() * sizeof(int32_t), c->stream));
    EL_HIP(c, upload(c->dEnvMap, hs.map2.data(), hs.map2.size() * sizeof(int32_t), c->stream));
    const std::vector<uint32_t> badBits =
        hdriBadBuckets(hs, h.data, h.width, h.height, h.x_tile, h.y_tile, h.x_offset, h.y_offset);
    EL_HIP(c, upload(c->dEnvBad, badBits.data(), badBits.size() * sizeof(uint32_t), c->stream));

    // accumulation
    size_t npx = (size_t)v->camera.x_res * v->camera.y_res;
    EL_HIP(c, c->dAccum.ensure(npx * sizeof(float4)));
    EL_HIP(c, c->dCounts.ensure(npx * sizeof(uint32_t)));
    c->dAov.release();

    DevScene& s = c->ds;
    s.shade = c->dShade.as<float4>();
    s.tangents = c->dTangents.as<float4>();
    s.mats = c->dMats.as<dev::DevMaterial>();
    s.texs = c->dTexs.as<dev::DevTex>();
    s.texels = c->dTexels.as<float>();
    s.texels8 = c->dTexels8.as<uint32_t>();
    s.texLut = c->dTexLut.as<float>();
    s.lights = c->dLights.as<float>();
    s.envCdf = c->dEnvCdf.as<float>();
    s.envGuide = c->dEnvGuide.as<int32_t>();
    s.envMap = c->dEnvMap.as<int32_t>();
    s.envBad = c->dEnvBad.as<uint32_t>();
    s.env = env;
    s.envRadianceSum = hs.radianceSum;
    EL_HIP(c, c->dEnvTab.ensure((size_t)h.width * h.height * 2 * sizeof(float4)));
    s.envTab = c->dEnvTab.as<float4>();
    s.envLog2G = hs.log2G;
    s.nLights = (int32_t)v->light_count;
    s.nTris = (int32_t)v->tri_count;
    s.cam = devCamera(v->camera);
    EL_HIP(c, launchEnvTable(c->stream, s, c->dEnvTab.as<float4>()));
    rc = clearAccum(c);
    if (rc) return rc;
    c->hasScene = true;
    return EL_OK;
}

int el_set_camera(el_ctx* c, const el_camera* cam) {
    if (!c || !cam) return EL_ERR_ARG;
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (cam->x_res == 0 || cam->y_res == 0) return fail(c, EL_ERR_ARG, "camera resolution is zero");
    if ((uint64_t)cam->x_res * cam->y_res > (1ull << 28)) return fail(c, EL_ERR_RANGE, "resolution above 2^28 pixels");
    EL_HIP(c, hipSetDevice(c->device));
    const size_t npx = (size_t)cam->x_res * cam->y_res;
    EL_HIP(c, c->dAccum.ensure(npx * sizeof(float4)));
    EL_HIP(c, c->dCounts.ensure(npx * sizeof(uint32_t)));
    if (cam->x_res != c->cam.x_res || cam->y_res != c->cam.y_res) c->dAov.release();
    c->cam = *cam;
    c->ds.cam = devCamera(*cam);
    c->residualPx = 0;   // the residual of el_residual_l2 belongs to the previous view
    return clearAccum(c);
}

int el_clear(el_ctx* c) {
    if (!c) return EL_ERR_ARG;
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    EL_HIP(c, hipSetDevice(c->device));
    return clearAccum(c);
}

}  // extern "C"

// Generations with fewer paths overlap every bounce's shadow launches with
// the next extension; larger ones from the third bounce on.
constexpr uint32_t kOverlapPaths = 16u << 20;

static hipEvent_t syncEvent(el_ctx* c, size_t i) {
    while (c->syncPool.size() <= i) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        c->syncPool.push_back(e);
    }
    return c->syncPool[i];
}

static hipEvent_t poolEvent(el_ctx* c, size_t i) {
    while (c->evPool.size() <= i) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->evPool.push_back(e);
    }
    return c->evPool[i];
}

// Path slots: `cap`; queue entries: max(cap, kXcds * segSize) (the camera
// queue is indexed by slot, the compacted queues by XCD segment).  Every
// buffer keeps the largest size asked of it (DevBuf::ensure); the per-path
// AOV buffer is sized by this call's capacity whatever the others hold, since
// renders without AOV may have grown them past its size.
static int ensureWf(el_ctx* c, size_t cap, size_t qcap, bool aov) {
    if (aov) EL_HIP(c, c->wAov.ensure(cap * 48));
    if (cap <= c->wfCap && qcap <= c->wfQcap) return EL_OK;
    EL_HIP(c, c->wRay0.ensure(qcap * 32));
    EL_HIP(c, c->wRay1.ensure(qcap * 32));
    EL_HIP(c, c->wHits.ensure(qcap * 16));
    EL_HIP(c, c->wShH.ensure(qcap * 48));
    EL_HIP(c, c->wShP.ensure(qcap * 48));
    EL_HIP(c, c->wT.ensure(cap * 16));
    EL_HIP(c, c->wL.ensure(cap * 16));
    EL_HIP(c, c->wPix.ensure(cap * 4));
    EL_HIP(c, c->wSample.ensure(cap * 4));
    EL_HIP(c, c->wQuiet.ensure(cap * 8));
    EL_HIP(c, c->wPend.ensure(cap * 16));
    EL_HIP(c, c->wCounters.ensure(kQCounters * sizeof(uint32_t)));
    c->wfCap = std::max(c->wfCap, cap);
    c->wfQcap = std::max(c->wfQcap, qcap);
    return EL_OK;
}

// Generation size and wavefront state for `total` work items of a frame
// with `npad` padded pixels: Q paths per generation (queueCap, or by default
// as many as a quarter of the free device memory holds), rounded to whole
// sample passes or 64-path groups; q points at the context's buffers.
static int wfPlan(el_ctx* c, const LaunchCfg& lc, uint64_t total, uint64_t npad, uint64_t queueCap, bool aov,
                  WfState& q, uint64_t& Q) {
    // Default: as many paths in flight as a quarter of the free device
    // memory holds (~216 B of queue and path state per path), so a
    // 1080p x 64 spp frame runs as a single generation.
    Q = queueCap;
    if (!Q) {
        size_t freeB = 0, totalB = 0;
        EL_HIP(c, hipMemGetInfo(&freeB, &totalB));
        const uint64_t perPath = 6 * 32 + 60 + (aov ? 48 : 0);
        Q = std::max<uint64_t>(1u << 20, freeB / 4 / perPath);
    }
    if (Q >= npad)
        Q = (Q / npad) * npad;
    else
        Q = std::max<uint64_t>(64, (Q / 64) * 64);
    if (Q > total) Q = total;
    if (Q > 0x7fffffffull) return fail(c, EL_ERR_RANGE, "queue capacity above 2^31");
    // Entries one XCD's blocks can append in one launch: at most
    // ceil(Q/8) + maxThreads/8 + block (grid-stride over <= maxThreads,
    // block <= kMaxBlock).
    const uint64_t segSize = (Q + kXcds - 1) / kXcds + lc.maxThreads / kXcds + kMaxBlock;
    const uint64_t qcap = std::max<uint64_t>(Q, kXcds * segSize);
    int rc = ensureWf(c, Q, qcap, aov);
    if (rc) return rc;
    // the launches index Q path slots and qcap queue entries in every buffer
    if (c->wRay0.n < qcap * 32 || c->wRay1.n < qcap * 32 || c->wHits.n < qcap * 16 || c->wShH.n < qcap * 48 ||
        c->wShP.n < qcap * 48 || c->wT.n < Q * 16 || c->wL.n < Q * 16 || c->wPix.n < Q * 4 || c->wSample.n < Q * 4 ||
        c->wQuiet.n < Q * 8 || c->wPend.n < Q * 16 || (aov && c->wAov.n < Q * 48))
        return fail(c, EL_ERR_STATE, "wavefront buffers smaller than the generation (internal error)");
    q.ray[0] = c->wRay0.as<float4>();
    q.ray[1] = c->wRay1.as<float4>();
    q.hits = c->wHits.as<float4>();
    q.shadowH = c->wShH.as<float4>();
    q.shadowP = c->wShP.as<float4>();
    q.pathT = c->wT.as<float4>();
    q.pathL = c->wL.as<float4>();
    q.pathPix = c->wPix.as<uint32_t>();
    q.pathSample = c->wSample.as<uint32_t>();
    q.pathQuiet = c->wQuiet.as<uint64_t>();
    q.pathAov = aov ? c->wAov.as<float4>() : nullptr;
    q.pendP = c->wPend.as<float4>();
    q.counters = c->wCounters.as<uint32_t>();
    q.segSize = (uint32_t)segSize;
    q.gstack = lc.gstack;
    q.gthreads = lc.maxThreads;
    return EL_OK;
}

extern "C" int el_render(el_ctx* c, const el_render_params* p) {
    if (!c || !p) return fail(c, EL_ERR_ARG, "null argument");
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (p->mode != EL_MODE_PLAIN && p->mode != EL_MODE_WAVEFRONT) return fail(c, EL_ERR_ARG, "unknown mode");
    if (p->bvh_width != 0 && p->bvh_width != 4 && p->bvh_width != 8) return fail(c, EL_ERR_ARG, "bvh_width must be 4 or 8");
    if (p->block != 0 && (p->block % 64 != 0 || p->block > (uint32_t)kMaxBlock))
        return fail(c, EL_ERR_ARG, "block must be a multiple of 64, at most 256");
    uint32_t depth = p->max_depth ? p->max_depth : 5;
    if (depth > 64) return fail(c, EL_ERR_RANGE, "max_depth above 64");
    EL_HIP(c, hipSetDevice(c->device));
    if (p->spp == 0) return EL_OK;
    if (p->bvh_width && (int)p->bvh_width != c->width) {
        int rc = uploadBvh(c, (int)p->bvh_width);
        if (rc) return rc;
    }
    if (p->refill > 64) return fail(c, EL_ERR_ARG, "refill must be 0..64");
    c->block = p->block ? p->block : kDefaultBlock;
    c->refill = p->refill ? p->refill : kDefaultRefill;
    bool aov = (p->flags & EL_FLAG_AOV) != 0;
    size_t npx = (size_t)c->cam.x_res * c->cam.y_res;
    if (aov && !c->dAov.p) {
        EL_HIP(c, c->dAov.ensure(npx * 3 * sizeof(float4)));
        EL_HIP(c, hipMemsetAsync(c->dAov.p, 0, c->dAov.n, c->stream));
    }
    c->aovOn = c->aovOn || aov;
    float4* aovBuf = aov ? c->dAov.as<float4>() : nullptr;
    unsigned long long* ctr = c->dCounters.as<unsigned long long>();
    EL_HIP(c, hipMemsetAsync(ctr, 0, sizeof(unsigned long long) * kCounterWords, c->stream));
    LaunchCfg lc = cfgOf(c);
    RenderArgs a{};
    a.k0 = (uint32_t)p->seed;
    a.k1 = (uint32_t)(p->seed >> 32);
    a.scene = p->rng_scene;
    a.maxDepth = depth;
    a.sample = c->sampleBase;
    a.cull = (p->flags & EL_FLAG_NO_CULL) ? 0u : 1u;
    bool record = (p->flags & EL_FLAG_RECORD_PATHS) != 0;
    bool recordHits = false;
    if (record) {
        if (c->sampleBase != 0) return fail(c, EL_ERR_STATE, "EL_FLAG_RECORD_PATHS needs a cleared accumulation");
        // the radiance record (16 B per path) is kept only when it fits half
        // of the free device memory; otherwise el_adjoint replays without it
        size_t freeB = 0, totalB = 0;
        EL_HIP(c, hipMemGetInfo(&freeB, &totalB));
        const size_t recBytes = (size_t)p->spp * npx * sizeof(float4);
        record = recBytes <= c->dRec.n || recBytes <= freeB / 2;
    }
    if (record) {
        EL_HIP(c, c->dRec.ensure((size_t)p->spp * npx * sizeof(float4)));
        a.rec = c->dRec.as<float4>();
        // Wavefront renders also keep every vertex's hit and shadow answers
        // (16 B per vertex + 4 B per path) when they fit a quarter of the
        // free memory, so that el_adjoint replays without tracing.
        if (p->mode != EL_MODE_PLAIN && depth <= (uint32_t)dev::kAdjMaxDepthHost) {
            size_t freeB = 0, totalB = 0;
            EL_HIP(c, hipMemGetInfo(&freeB, &totalB));
            const size_t hitBytes = (size_t)p->spp * npx * depth * sizeof(float4);
            recordHits = hitBytes <= c->dRecHits.n || hitBytes <= freeB / 4;
            if (recordHits) {
                EL_HIP(c, c->dRecHits.ensure(hitBytes));
                EL_HIP(c, c->dRecVis.ensure((size_t)p->spp * npx * sizeof(uint32_t)));
                a.recHits = c->dRecHits.as<float4>();
                a.recVis = c->dRecVis.as<uint32_t>();
                a.recStride = (uint64_t)p->spp * npx;
            }
        }
    }
    c->recValid = false;

    hipEvent_t e0 = poolEvent(c, 0), e1 = poolEvent(c, 1);
    if (!e0 || !e1) return fail(c, EL_ERR_HIP, "event creation failed");
    double extendMs = 0, shadeMs = 0, shadowMs = 0;
    uint64_t extendLaunches = 0;
    EL_HIP(c, hipEventRecord(e0, c->stream));
    const bool countTrav = (p->flags & 2u) != 0;
    unsigned long long h[kCounterWords];   // the launches' tallies, read back with the last sync

    if (p->mode == EL_MODE_PLAIN) {
        for (uint32_t s = 0; s < p->spp; s++) {
            a.sample = c->sampleBase + s;
            EL_HIP(c, launchPlain(lc, c->ds, a, c->dAccum.as<float4>(), c->dCounts.as<uint32_t>(), aovBuf, ctr));
        }
    } else {
        const uint32_t tilesX = (c->cam.x_res + 7) / 8, tilesY = (c->cam.y_res + 7) / 8;
        const uint64_t npad = (uint64_t)tilesX * tilesY * 64;
        uint64_t total = (uint64_t)p->spp * npad;
        if (npad >= (1ull << 31)) return fail(c, EL_ERR_RANGE, "frame too large for the wavefront renderer");
        uint64_t Q = 0;
        WfState q{};
        int rc = wfPlan(c, lc, total, npad, p->queue_cap, aov, q, Q);
        if (rc) return rc;
        if (recordHits) {
            EL_HIP(c, c->wVis.ensure(Q * sizeof(uint32_t)));
            q.pathVis = c->wVis.as<uint32_t>();
            // the shadow launches set visibility bits of one word per path:
            // they stay in order, and the point-light terms are added at once
            q.pendP = nullptr;
        }
        a.npad = (uint32_t)npad;
        size_t ev = 2;
        std::vector<std::pair<size_t, int>> timed;   // (event index, kind) kind 0 extend 1 shade 2 shadow
        for (uint64_t start = 0; start < total; start += Q) {
            a.genPixel = (uint32_t)(start % npad);
            a.genSample = (uint32_t)(start / npad);
            a.genCount = (uint32_t)std::min<uint64_t>(Q, total - start);
            EL_HIP(c, launchWfCamera(lc, c->ds, a, q));
            // Per bounce b: extend (stream) -> shade (stream) -> both shadow
            // launches (shadowStream, shadowStreamP).  The shadows of bounce
            // b need only what shade b wrote and update nothing the extension
            // of bounce b + 1 reads (their counters are set b & 1, the
            // extension clears set (b + 1) & 1), so the three overlap; shade
            // b + 1 and finalize wait for them (they add to the path
            // radiance).  The point-light launch writes its terms to pendP,
            // which shade b + 1 or finalize adds after the HDRI term, so it
            // also runs beside the HDRI launch (a recording render keeps the
            // two in order on one stream).
            LaunchCfg lcs = lc, lcp = lc;
            lcs.stream = c->shadowStream;
            lcp.stream = c->shadowStreamP;
            const size_t spillSet = (size_t)c->gcap * lc.maxThreads;
            hipEvent_t doneH = nullptr, doneP = nullptr;
            // Large launches keep the GPU busy on their own, and overlapping
            // two of them costs cache hits (the extension's L2 hit rate fell
            // from 0.70 to 0.62 next to the first bounce's shadows); small
            // ones last one ray's chain of dependent fetches, which the
            // overlap hides.
            const bool serial = (p->flags & EL_FLAG_SERIAL_SHADOWS) != 0;
            const bool kernelTimes = (p->flags & EL_FLAG_KERNEL_TIMES) != 0;
            for (uint32_t b = 0; b < depth; b++) {
                const bool overlap = !serial && (a.genCount < kOverlapPaths || b >= 2);
                a.bounce = b;
                q.parity = b & 1u;
                // point-light terms deferred where the two shadow launches run
                // at once (elsewhere the P launch adds them itself)
                q.deferP = overlap && q.pendP != nullptr ? 1u : 0u;
                const int cur = b & 1, nxt = cur ^ 1;
                uint32_t* cntCur = q.counters + (cur ? kQRay1 : kQRay0);
                uint32_t* cntNxt = q.counters + (nxt ? kQRay1 : kQRay0);
                hipEvent_t x0 = poolEvent(c, ev), x1 = poolEvent(c, ev + 1);
                if (kernelTimes) EL_HIP(c, hipEventRecord(x0, c->stream));
                EL_HIP(c, launchWfExtend(lc, c->ds, q, q.ray[cur], cntCur, cntNxt, a.genCount, countTrav, ctr));
                if (kernelTimes) EL_HIP(c, hipEventRecord(x1, c->stream));
                if (kernelTimes) timed.push_back({ev, 0});
                ev += 2;
                if (doneH) EL_HIP(c, hipStreamWaitEvent(c->stream, doneH, 0));
                if (doneP) EL_HIP(c, hipStreamWaitEvent(c->stream, doneP, 0));
                hipEvent_t y0 = poolEvent(c, ev), y1 = poolEvent(c, ev + 1);
                if (kernelTimes) EL_HIP(c, hipEventRecord(y0, c->stream));
                EL_HIP(c, launchWfShade(lc, c->ds, a, q, q.ray[cur], cntCur, q.ray[nxt], cntNxt, a.genCount, ctr));
                if (kernelTimes) EL_HIP(c, hipEventRecord(y1, c->stream));
                if (kernelTimes) timed.push_back({ev, 1});
                ev += 2;
                // beside the extension, each shadow launch spills to its own area
                const bool split = q.deferP != 0u;   // H and P on their own streams
                const LaunchCfg& lsH = overlap ? lcs : lc;
                const LaunchCfg& lsP = split ? lcp : lsH;
                WfState qh = q, qp = q;
                if (overlap) qh.gstack = q.gstack + spillSet;
                qp.gstack = split ? q.gstack + 2 * spillSet : qh.gstack;
                if (recordHits) {
                    qh.visBit = 1u << (2 * b);
                    qp.visBit = 1u << (2 * b + 1);
                }
                if (overlap) {
                    hipEvent_t shaded = syncEvent(c, 3 * b);
                    if (!shaded) return fail(c, EL_ERR_HIP, "event creation failed");
                    EL_HIP(c, hipEventRecord(shaded, c->stream));
                    EL_HIP(c, hipStreamWaitEvent(c->shadowStream, shaded, 0));
                    if (split) EL_HIP(c, hipStreamWaitEvent(c->shadowStreamP, shaded, 0));
                }
                hipEvent_t z0 = poolEvent(c, ev), z1 = poolEvent(c, ev + 1);
                if (kernelTimes) EL_HIP(c, hipEventRecord(z0, lsH.stream));
                EL_HIP(c, launchWfShadow(lsH, c->ds, qh, q.shadowH, q.counters + qShadow(q.parity, 0), a.genCount,
                                         ctr));
                if (kernelTimes) EL_HIP(c, hipEventRecord(z1, lsH.stream));
                if (kernelTimes) timed.push_back({ev, 2});
                ev += 2;
                hipEvent_t w0 = poolEvent(c, ev), w1 = poolEvent(c, ev + 1);
                if (kernelTimes) EL_HIP(c, hipEventRecord(w0, lsP.stream));
                EL_HIP(c, launchWfShadow(lsP, c->ds, qp, q.shadowP, q.counters + qShadow(q.parity, 1), a.genCount,
                                         ctr));
                if (kernelTimes) EL_HIP(c, hipEventRecord(w1, lsP.stream));
                if (kernelTimes) timed.push_back({ev, 2});
                ev += 2;
                doneH = doneP = nullptr;
                if (overlap) {
                    doneH = syncEvent(c, 3 * b + 1);
                    if (!doneH) return fail(c, EL_ERR_HIP, "event creation failed");
                    EL_HIP(c, hipEventRecord(doneH, c->shadowStream));
                }
                if (split) {
                    doneP = syncEvent(c, 3 * b + 2);
                    if (!doneP) return fail(c, EL_ERR_HIP, "event creation failed");
                    EL_HIP(c, hipEventRecord(doneP, c->shadowStreamP));
                }
                extendLaunches++;
            }
            if (doneH) EL_HIP(c, hipStreamWaitEvent(c->stream, doneH, 0));
            if (doneP) EL_HIP(c, hipStreamWaitEvent(c->stream, doneP, 0));
            EL_HIP(c, launchWfFinalize(lc, c->ds, a, q, c->dAccum.as<float4>(), c->dCounts.as<uint32_t>(), aovBuf,
                                       ctr));
        }
        EL_HIP(c, hipEventRecord(e1, c->stream));
        EL_HIP(c, hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
        EL_HIP(c, hipStreamSynchronize(c->stream));   // the tallies and every event
        for (auto& t : timed) {
            float ms = 0;
            EL_HIP(c, hipEventElapsedTime(&ms, c->evPool[t.first], c->evPool[t.first + 1]));
            (t.second == 0 ? extendMs : t.second == 1 ? shadeMs : shadowMs) += ms;
        }
    }
    if (p->mode == EL_MODE_PLAIN) {
        EL_HIP(c, hipEventRecord(e1, c->stream));
        EL_HIP(c, hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
        EL_HIP(c, hipStreamSynchronize(c->stream));
    }
    float totalMs = 0;
    EL_HIP(c, hipEventElapsedTime(&totalMs, e0, e1));
    if (c->sampleBase == 0) {
        c->accSeed = p->seed;
        c->accDepth = depth;
        c->accScene = p->rng_scene;
    } else if (c->accSeed != p->seed || c->accDepth != depth || c->accScene != p->rng_scene) {
        c->accMixed = true;
    }
    c->sampleBase += p->spp;
    c->recValid = record;
    c->recHits = recordHits;
    c->stats.record_bytes = record ? c->dRec.n + c->dRecHits.n + c->dRecVis.n + c->wVis.n : 0;
    c->totalRays += tally(h, kStatRays);
    c->totalVerts += tally(h, kStatVerts);
    c->totalPaths += tally(h, kStatPaths);
    c->stats.rays = tally(h, kStatRays);
    c->stats.vertices = tally(h, kStatVerts);
    c->stats.paths = tally(h, kStatPaths);
    c->stats.render_ms = totalMs;
    c->stats.extend_ms = extendMs;
    c->stats.shade_ms = shadeMs;
    c->stats.shadow_ms = shadowMs;
    c->stats.extend_launches = extendLaunches;
    c->stats.extend_rays = tally(h, kStatExtRays);
    c->stats.node_fetches = tally(h, kStatNodes);
    c->stats.tri_tests = tally(h, kStatTris);
    if (tally(h, kStatOverflow)) return fail(c, EL_ERR_RANGE, "traversal stack or queue overflow (results incomplete)");
    return EL_OK;
}

extern "C" int el_residual_l2(el_ctx* c, const float* target, double* loss) {
    if (!c || !target) return fail(c, EL_ERR_ARG, "null argument");
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (c->sampleBase == 0) return fail(c, EL_ERR_STATE, "nothing rendered since el_clear");
    EL_HIP(c, hipSetDevice(c->device));
    const size_t npx = (size_t)c->cam.x_res * c->cam.y_res;
    EL_HIP(c, c->aImage.ensure(npx * 3 * sizeof(float)));
    EL_HIP(c, c->aTarget.ensure(npx * 3 * sizeof(float)));
    EL_HIP(c, c->aLoss.ensure(sizeof(double)));
    EL_HIP(c, hipMemcpyAsync(c->aTarget.p, target, npx * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    EL_HIP(c, hipMemsetAsync(c->aLoss.p, 0, sizeof(double), c->stream));
    EL_HIP(c, launchResidualL2(c->stream, c->dAccum.as<float4>(), c->dCounts.as<uint32_t>(), c->aTarget.as<float>(),
                               (uint32_t)npx, c->aImage.as<float>(), c->aLoss.as<double>()));
    double h = 0.0;
    EL_HIP(c, hipMemcpyAsync(&h, c->aLoss.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EL_HIP(c, hipStreamSynchronize(c->stream));
    if (loss) *loss = h;
    c->residualPx = npx;
    return EL_OK;
}

extern "C" int el_adjoint(el_ctx* c, const el_adjoint_params* p, const float* dimage, float* dmat, float* dtex,
                          int devp) {
    const bool accumulate = (p && (p->flags & EL_ADJ_ACCUMULATE) != 0);
    const bool devResidual = (p && (p->flags & EL_ADJ_DEVICE_RESIDUAL) != 0);
    if (!c || !p || (!dimage && !devResidual)) return fail(c, EL_ERR_ARG, "null argument");
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (c->sampleBase == 0) return fail(c, EL_ERR_STATE, "nothing rendered since el_clear: no image to differentiate");
    if (c->accMixed) return fail(c, EL_ERR_STATE, "the accumulation mixes renders of different seeds or depths");
    if (c->accDepth > (uint32_t)dev::kAdjMaxDepthHost) return fail(c, EL_ERR_RANGE, "adjoint supports max_depth <= 16");
    if (p->block != 0 && (p->block % 64 != 0 || p->block > (uint32_t)kMaxBlock))
        return fail(c, EL_ERR_ARG, "block must be a multiple of 64, at most 256");
    const int gtex = p->grad_texture;
    if (gtex >= (int)c->texs.size()) return fail(c, EL_ERR_ARG, "grad_texture out of range");
    if (accumulate && (!c->adjSumValid || c->adjSumTex != gtex))
        return fail(c, EL_ERR_STATE, "EL_ADJ_ACCUMULATE without a previous el_adjoint with the same grad_texture");
    EL_HIP(c, hipSetDevice(c->device));
    const size_t npx = (size_t)c->cam.x_res * c->cam.y_res;
    const size_t M = c->mats.size();
    const size_t ntex = gtex >= 0 ? (size_t)c->texs[gtex].w * c->texs[gtex].h * 3 : 0;
    const hipMemcpyKind h2d = devp ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (devResidual) {
        if (c->residualPx != npx)
            return fail(c, EL_ERR_STATE, "EL_ADJ_DEVICE_RESIDUAL without an el_residual_l2 at this resolution");
    } else {
        EL_HIP(c, c->aImage.ensure(npx * 3 * sizeof(float)));
        EL_HIP(c, hipMemcpyAsync(c->aImage.p, dimage, npx * 3 * sizeof(float), h2d, c->stream));
        c->residualPx = 0;
    }
    if (!accumulate) {
        EL_HIP(c, c->aMat.ensure(std::max<size_t>(1, M) * 6 * sizeof(double)));
        EL_HIP(c, hipMemsetAsync(c->aMat.p, 0, M * 6 * sizeof(double), c->stream));
        if (ntex) {
            EL_HIP(c, c->aTex.ensure(ntex * sizeof(float)));
            EL_HIP(c, hipMemsetAsync(c->aTex.p, 0, ntex * sizeof(float), c->stream));
        }
        c->adjSumValid = false;
    }
    unsigned long long* ctr = c->dCounters.as<unsigned long long>();
    EL_HIP(c, hipMemsetAsync(ctr, 0, sizeof(unsigned long long) * kCounterWords, c->stream));
    LaunchCfg lc = cfgOf(c);
    if (p->block) lc.block = p->block;
    RenderArgs a{};
    a.k0 = (uint32_t)c->accSeed;
    a.k1 = (uint32_t)(c->accSeed >> 32);
    a.scene = c->accScene;
    a.maxDepth = c->accDepth;
    a.sample = 0;
    a.cull = 1;   // exact: a zero-throughput path adds neither radiance nor gradient
    if (c->recValid && c->recHits) {
        // the recording stopped no later than this replay does (it culled
        // alike or not at all), so every vertex the replay reads is recorded
        a.recHits = c->dRecHits.as<float4>();
        a.recVis = c->dRecVis.as<uint32_t>();
        a.recStride = (uint64_t)c->sampleBase * npx;
    }
    float4* cache = nullptr;
    if (!c->recValid) {   // the primal walk keeps its vertices for the replay, per grid thread
        EL_HIP(c, c->aCache.ensure((size_t)lc.maxThreads * c->accDepth * sizeof(float4)));
        cache = c->aCache.as<float4>();
    }
    hipEvent_t e0 = poolEvent(c, 0), e1 = poolEvent(c, 1);
    if (!e0 || !e1) return fail(c, EL_ERR_HIP, "event creation failed");
    EL_HIP(c, hipEventRecord(e0, c->stream));
    EL_HIP(c, launchAdjoint(lc, c->ds, a, c->sampleBase, c->aImage.as<float>(), c->dCounts.as<uint32_t>(),
                            c->recValid ? c->dRec.as<float4>() : nullptr, cache, gtex, c->aMat.as<double>(),
                            ntex ? c->aTex.as<float>() : nullptr, ctr));
    EL_HIP(c, hipEventRecord(e1, c->stream));
    std::vector<double> hm(M * 6);
    EL_HIP(c, hipMemcpyAsync(hm.data(), c->aMat.p, M * 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    unsigned long long h[kCounterWords];
    EL_HIP(c, hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
    // the texel gradients with the same sync
    if (ntex && dtex)
        EL_HIP(c, hipMemcpyAsync(dtex, c->aTex.p, ntex * sizeof(float),
                                 devp ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, c->stream));
    EL_HIP(c, hipStreamSynchronize(c->stream));
    float ms = 0;
    EL_HIP(c, hipEventElapsedTime(&ms, e0, e1));
    std::vector<float> fm(M * 6);
    for (size_t i = 0; i < M * 6; i++) fm[i] = (float)hm[i];
    if (dmat) {
        if (devp)
            EL_HIP(c, hipMemcpy(dmat, fm.data(), fm.size() * sizeof(float), hipMemcpyHostToDevice));
        else
            std::memcpy(dmat, fm.data(), fm.size() * sizeof(float));
    }
    c->adjSumValid = true;
    c->adjSumTex = gtex;
    c->stats.adjoint_ms = ms;
    c->stats.adjoint_rays = tally(h, kStatRays);
    c->stats.adjoint_bytes = c->aCache.n + c->aImage.n + c->aMat.n + c->aTex.n;
    if (tally(h, kStatOverflow)) return fail(c, EL_ERR_RANGE, "traversal stack overflow (gradients incomplete)");
    return EL_OK;
}

extern "C" {

int el_read_passes(el_ctx* c, float* rgba, uint32_t* samples, uint64_t* rays, uint64_t* vertices) {
    if (!c) return EL_ERR_ARG;
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    EL_HIP(c, hipSetDevice(c->device));
    size_t npx = (size_t)c->cam.x_res * c->cam.y_res;
    std::vector<float> acc(npx * 4);
    std::vector<uint32_t> cnt(npx);
    EL_HIP(c, hipMemcpyAsync(acc.data(), c->dAccum.p, npx * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    EL_HIP(c, hipMemcpyAsync(cnt.data(), c->dCounts.p, npx * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    EL_HIP(c, hipStreamSynchronize(c->stream));
    if (rgba)
        for (size_t i = 0; i < npx; i++) {
            float n = (float)cnt[i];
            for (int k = 0; k < 3; k++) rgba[4 * i + k] = cnt[i] ? acc[4 * i + k] / n : 0.f;
            rgba[4 * i + 3] = 1.f;
        }
    if (samples) std::memcpy(samples, cnt.data(), npx * sizeof(uint32_t));
    if (rays) *rays = c->totalRays;
    if (vertices) *vertices = c->totalVerts;
    return EL_OK;
}

int el_read_aov(el_ctx* c, int pass, float* rgba) {
    if (!c || !rgba) return EL_ERR_ARG;
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (pass != EL_PASS_NORMAL && pass != EL_PASS_TANGENT && pass != EL_PASS_BITANGENT)
        return fail(c, EL_ERR_ARG, "pass must be NORMAL, TANGENT or BITANGENT");
    if (!c->dAov.p) return fail(c, EL_ERR_STATE, "no AOV rendered (EL_FLAG_AOV)");
    EL_HIP(c, hipSetDevice(c->device));
    size_t npx = (size_t)c->cam.x_res * c->cam.y_res;
    std::vector<float> a(npx * 12);
    std::vector<uint32_t> cnt(npx);
    EL_HIP(c, hipMemcpyAsync(a.data(), c->dAov.p, npx * 48, hipMemcpyDeviceToHost, c->stream));
    EL_HIP(c, hipMemcpyAsync(cnt.data(), c->dCounts.p, npx * 4, hipMemcpyDeviceToHost, c->stream));
    EL_HIP(c, hipStreamSynchronize(c->stream));
    int k = pass - EL_PASS_NORMAL;
    for (size_t i = 0; i < npx; i++) {
        for (int j = 0; j < 3; j++) rgba[4 * i + j] = cnt[i] ? a[12 * i + 4 * k + j] / (float)cnt[i] : 0.f;
        rgba[4 * i + 3] = 1.f;
    }
    return EL_OK;
}

int el_get_stats(el_ctx* c, el_stats* out) {
    if (!c || !out) return EL_ERR_ARG;
    *out = c->stats;
    return EL_OK;
}

int el_set_materials(el_ctx* c, const el_material* m, uint32_t count) {
    if (!c || !m) return EL_ERR_ARG;
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (count != c->mats.size()) return fail(c, EL_ERR_ARG, "material count differs from the uploaded scene");
    for (uint32_t i = 0; i < count; i++) {
        setMaterialValues(c->mats[i], m[i]);
    }
    EL_HIP(c, hipSetDevice(c->device));
    EL_HIP(c, hipMemcpyAsync(c->dMats.p, c->mats.data(), count * sizeof(dev::DevMaterial), hipMemcpyHostToDevice,
                             c->stream));
    EL_HIP(c, hipStreamSynchronize(c->stream));
    return EL_OK;
}

int el_set_texture(el_ctx* c, int32_t id, const float* rgb) {
    if (!c || !rgb) return EL_ERR_ARG;
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (id < 0 || (size_t)id >= c->texs.size()) return fail(c, EL_ERR_ARG, "texture id out of range");
    // new values go to the FP32 copy, which the texture reads from then on
    dev::DevTex& t = c->texs[id];
    EL_HIP(c, hipSetDevice(c->device));
    EL_HIP(c, hipMemcpyAsync(c->dTexels.as<float>() + c->texF32[id], rgb, (size_t)t.w * t.h * 3 * sizeof(float),
                             hipMemcpyHostToDevice, c->stream));
    if (t.format != dev::kTexF32) {
        c->stats.texels_8bit -= (uint64_t)t.w * t.h;
        t.format = dev::kTexF32;
        t.offset = c->texF32[id];
        EL_HIP(c, hipMemcpyAsync(c->dTexs.as<dev::DevTex>() + id, &t, sizeof t, hipMemcpyHostToDevice, c->stream));
    }
    EL_HIP(c, hipStreamSynchronize(c->stream));
    return EL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- queries --
static int traceCommon(el_ctx* c, const el_ray* rays, const float* tmax, uint32_t n, int32_t* tri, float* t,
                       float* uv, uint8_t* occ, int devp, bool any, bool count, uint64_t* nodes, uint64_t* tris) {
    if (!c || !rays) return fail(c, EL_ERR_ARG, "null argument");
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (any && (!tmax || !occ)) return fail(c, EL_ERR_ARG, "null argument");
    if (!any && !count && (!tri || !t)) return fail(c, EL_ERR_ARG, "null argument");
    if (n == 0) return EL_OK;
    EL_HIP(c, hipSetDevice(c->device));
    const float* dR = (const float*)rays;
    const float* dTm = tmax;
    int32_t* dTri = tri;
    float* dT = t;
    float* dUV = uv;
    uint8_t* dOcc = occ;
    if (!devp) {
        EL_HIP(c, upload(c->tRays, rays, (size_t)n * sizeof(el_ray), c->stream));
        dR = c->tRays.as<float>();
        if (any) {
            EL_HIP(c, upload(c->tTmax, tmax, (size_t)n * 4, c->stream));
            dTm = c->tTmax.as<float>();
            EL_HIP(c, c->tOcc.ensure(n));
            dOcc = c->tOcc.as<uint8_t>();
        } else {
            EL_HIP(c, c->tTri.ensure((size_t)n * 4));
            EL_HIP(c, c->tT.ensure((size_t)n * 4));
            dTri = c->tTri.as<int32_t>();
            dT = c->tT.as<float>();
            if (uv) {
                EL_HIP(c, c->tUV.ensure((size_t)n * 8));
                dUV = c->tUV.as<float>();
            }
        }
    } else if (count) {
        EL_HIP(c, c->tTri.ensure((size_t)n * 4));
        EL_HIP(c, c->tT.ensure((size_t)n * 4));
        if (!dTri) dTri = c->tTri.as<int32_t>();
        if (!dT) dT = c->tT.as<float>();
    }
    if (!devp && count) {
        EL_HIP(c, c->tTri.ensure((size_t)n * 4));
        EL_HIP(c, c->tT.ensure((size_t)n * 4));
        dTri = c->tTri.as<int32_t>();
        dT = c->tT.as<float>();
    }
    unsigned long long* ctr = c->dCounters.as<unsigned long long>();
    EL_HIP(c, hipMemsetAsync(ctr, 0, sizeof(unsigned long long) * kCounterWords, c->stream));
    EL_HIP(c, launchTrace(cfgOf(c), c->ds, dR, dTm, n, dTri, dT, dUV, dOcc, any, count, ctr));
    unsigned long long h[kCounterWords];
    EL_HIP(c, hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
    if (!devp) {
        if (any) {
            EL_HIP(c, hipMemcpyAsync(occ, dOcc, n, hipMemcpyDeviceToHost, c->stream));
        } else if (!count) {
            EL_HIP(c, hipMemcpyAsync(tri, dTri, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
            EL_HIP(c, hipMemcpyAsync(t, dT, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream));
            if (uv) EL_HIP(c, hipMemcpyAsync(uv, dUV, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
        }
    }
    EL_HIP(c, hipStreamSynchronize(c->stream));
    if (nodes) *nodes = tally(h, kStatNodes);
    if (tris) *tris = tally(h, kStatTris);
    if (tally(h, kStatOverflow)) return fail(c, EL_ERR_RANGE, "traversal stack overflow (results incomplete)");
    return EL_OK;
}

extern "C" {

int el_trace_closest(el_ctx* c, const el_ray* rays, uint32_t n, int32_t* tri, float* t, float* uv, int devp) {
    return traceCommon(c, rays, nullptr, n, tri, t, uv, nullptr, devp, false, false, nullptr, nullptr);
}

int el_trace_any(el_ctx* c, const el_ray* rays, const float* tmax, uint32_t n, uint8_t* occ, int devp) {
    return traceCommon(c, rays, tmax, n, nullptr, nullptr, nullptr, occ, devp, true, false, nullptr, nullptr);
}

int el_trace_count(el_ctx* c, const el_ray* rays, uint32_t n, uint64_t* nodes, uint64_t* tris, int devp) {
    if (!nodes || !tris) return fail(c, EL_ERR_ARG, "null argument");
    return traceCommon(c, rays, nullptr, n, nullptr, nullptr, nullptr, nullptr, devp, false, true, nodes, tris);
}

int el_trace_closest_noalloc(el_ctx* c, const el_ray* d_rays, uint32_t n, int32_t* d_tri, float* d_t,
                             float* d_uv) {
    if (!c || !d_rays || !d_tri || !d_t) return fail(c, EL_ERR_ARG, "null argument");
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    if (!c->scratch) return fail(c, EL_ERR_STATE, "no device scratch");
    if (n == 0) return EL_OK;
    EL_HIP(c, hipSetDevice(c->device));
    // as many resident threads as the spill area holds at the scene's depth
    LaunchCfg lc = cfgOf(c);
    const uint64_t fit = kScratchSpillInts / (uint64_t)c->gcap;
    uint32_t threads = (uint32_t)std::min<uint64_t>(c->maxThreads, fit);
    threads -= threads % lc.block;
    if (threads < lc.block)
        return fail(c, EL_ERR_RANGE, "the scene's traversal stack (" + std::to_string(c->gcap) +
                                         " spill slots per thread) exceeds the static scratch for one block");
    lc.maxThreads = threads;
    lc.gstack = (int*)((unsigned char*)c->scratch + kScratchCounterBytes);
    unsigned long long* ctr = (unsigned long long*)c->scratch;
    EL_HIP(c, hipMemsetAsync(ctr, 0, sizeof(unsigned long long) * kCounterWords, c->stream));
    EL_HIP(c, launchTrace(lc, c->ds, (const float*)d_rays, nullptr, n, d_tri, d_t, d_uv, nullptr, false, false, ctr));
    unsigned long long h[kCounterWords];
    EL_HIP(c, hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
    EL_HIP(c, hipStreamSynchronize(c->stream));
    if (tally(h, kStatOverflow)) return fail(c, EL_ERR_RANGE, "traversal stack overflow (results incomplete)");
    return EL_OK;
}

int el_trace_bench(el_ctx* c, const el_ray* d_rays, uint32_t n, int32_t* d_tri, float* d_t, int iters,
                   double* mean_ms) {
    if (!c || !d_rays || !d_tri || !d_t || !mean_ms || iters <= 0) return fail(c, EL_ERR_ARG, "null argument");
    if (!c->hasScene) return fail(c, EL_ERR_STATE, "no scene uploaded");
    EL_HIP(c, hipSetDevice(c->device));
    unsigned long long* ctr = c->dCounters.as<unsigned long long>();
    hipEvent_t e0 = poolEvent(c, 0), e1 = poolEvent(c, 1);
    double total = 0;
    for (int i = 0; i < iters; i++) {
        EL_HIP(c, hipEventRecord(e0, c->stream));
        EL_HIP(c, launchTrace(cfgOf(c), c->ds, (const float*)d_rays, nullptr, n, d_tri, d_t, nullptr, nullptr, false,
                              false, ctr));
        EL_HIP(c, hipEventRecord(e1, c->stream));
        EL_HIP(c, hipEventSynchronize(e1));
        float ms = 0;
        EL_HIP(c, hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    *mean_ms = total / iters;
    return EL_OK;
}

int el_device_alloc(el_ctx* c, size_t bytes, void** out) {
    if (!c || !out) return EL_ERR_ARG;
    EL_HIP(c, hipSetDevice(c->device));
    EL_HIP(c, hipMalloc(out, bytes ? bytes : 16));
    return EL_OK;
}

int el_device_free(el_ctx* c, void* p) {
    if (!c) return EL_ERR_ARG;
    EL_HIP(c, hipSetDevice(c->device));
    EL_HIP(c, hipFree(p));
    return EL_OK;
}

int el_memcpy_htod(el_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src) return EL_ERR_ARG;
    EL_HIP(c, hipSetDevice(c->device));
    EL_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    EL_HIP(c, hipStreamSynchronize(c->stream));
    return EL_OK;
}

int el_memcpy_dtoh(el_ctx* c, void* dst, const void* src, size_t bytes) {
    if (!c || !dst || !src) return EL_ERR_ARG;
    EL_HIP(c, hipSetDevice(c->device));
    EL_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    EL_HIP(c, hipStreamSynchronize(c->stream));
    return EL_OK;
}

}  // extern "C"

// ---------------------------------------------------------------- batches --
// Many scenes rendered by one wavefront loop.  Each scene is uploaded into
// its own context (geometry, textures, HDRI tables and its accumulation stay
// in that context's buffers); the batch owns the queues, the traversal
// stack, a device array of the scenes' DevScene records and the generation
// plan.  Work items are enumerated scene after scene, sample-major inside a
// scene; a generation takes the next queue_cap of them, whatever scenes they
// belong to, and a segment table tells the camera and finalize kernels which
// scene each path slot belongs to.  With EL_BATCH_GRAPH the whole launch
// sequence is captured once into a hipGraph and replayed.
struct el_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err, log;
    uint32_t maxThreads = 0;
    std::vector<el_ctx*> scenes;
    std::vector<uint32_t> spp;
    std::vector<uint64_t> seed;
    int width = 4;
    DevBuf dScenes, dSegs, dGstack, dCounters;
    DevBuf wRay0, wRay1, wHits, wShH, wShP, wT, wL, wPix, wSample, wQuiet, wCounters;
    size_t wfCap = 0, wfQcap = 0;
    hipGraphExec_t graph = nullptr;
    std::string graphKey;
    el_stats stats{};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipStream_t shadowStream = nullptr;   // shadow launches of small bounces (as el_render)
    std::vector<hipEvent_t> syncPool;
};

static hipEvent_t batchSyncEvent(el_batch* b, size_t i) {
    while (b->syncPool.size() <= i) {
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        b->syncPool.push_back(e);
    }
    return b->syncPool[i];
}

static int bfail(el_batch* b, int code, const std::string& what) {
    if (b) b->err = what;
    return code;
}

#define EL_BHIP(b, x)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return bfail(b, EL_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct BatchGen {
    size_t segOffset;
    uint32_t nseg, count, finTotal;
};

extern "C" {

int el_batch_create(el_batch** out, int device) {
    if (!out) return EL_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return EL_ERR_HIP;
    if (device < 0 || device >= n) return EL_ERR_ARG;
    std::unique_ptr<el_batch> b(new el_batch());
    b->device = device;
    if (hipSetDevice(device) != hipSuccess) return EL_ERR_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return EL_ERR_HIP;
    b->maxThreads = (uint32_t)prop.multiProcessorCount * 2048u;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return EL_ERR_HIP;
    if (hipStreamCreateWithFlags(&b->shadowStream, hipStreamNonBlocking) != hipSuccess) return EL_ERR_HIP;
    if (hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) return EL_ERR_HIP;
    if (b->dCounters.ensure(sizeof(unsigned long long) * kCounterWords) != hipSuccess) return EL_ERR_HIP;
    *out = b.release();
    return EL_OK;
}

void el_batch_destroy(el_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    for (auto* c : b->scenes) el_destroy(c);
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    if (b->shadowStream) (void)hipStreamSynchronize(b->shadowStream);
    for (auto e : b->syncPool) (void)hipEventDestroy(e);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->shadowStream) (void)hipStreamDestroy(b->shadowStream);
    delete b;
}

const char* el_batch_last_error(const el_batch* b) { return b ? b->err.c_str() : "null batch"; }
const char* el_batch_log(const el_batch* b) { return b ? b->log.c_str() : ""; }

int el_batch_add(el_batch* b, const el_scene_view* view, uint32_t spp, uint64_t seed, int32_t* index) {
    if (!b || !view) return bfail(b, EL_ERR_ARG, "null argument");
    if (spp == 0) return bfail(b, EL_ERR_ARG, "spp must be positive");
    if (b->scenes.size() >= 65536) return bfail(b, EL_ERR_RANGE, "at most 65536 scenes per batch");
    el_ctx* c = nullptr;
    int rc = el_create(&c, b->device);
    if (rc) return bfail(b, rc, "context creation failed");
    rc = el_upload_scene(c, view);
    if (rc) {
        std::string m = el_last_error(c);
        el_destroy(c);
        return bfail(b, rc, m);
    }
    if (c->width != b->width) {
        rc = uploadBvh(c, b->width);
        if (rc) {
            std::string m = el_last_error(c);
            el_destroy(c);
            return bfail(b, rc, m);
        }
    }
    b->scenes.push_back(c);
    b->spp.push_back(spp);
    b->seed.push_back(seed);
    if (b->graph) {
        (void)hipGraphExecDestroy(b->graph);
        b->graph = nullptr;
    }
    if (index) *index = (int32_t)b->scenes.size() - 1;
    return EL_OK;
}

int el_batch_read(el_batch* b, int32_t scene, float* rgba, uint32_t* samples) {
    if (!b) return EL_ERR_ARG;
    if (scene < 0 || (size_t)scene >= b->scenes.size()) return bfail(b, EL_ERR_ARG, "scene index out of range");
    int rc = el_read_passes(b->scenes[scene], rgba, samples, nullptr, nullptr);
    if (rc) return bfail(b, rc, el_last_error(b->scenes[scene]));
    return EL_OK;
}

int el_batch_stats(el_batch* b, el_stats* out) {
    if (!b || !out) return EL_ERR_ARG;
    *out = b->stats;
    return EL_OK;
}

}  // extern "C"

static int batchEnsureWf(el_batch* b, size_t cap, size_t qcap) {
    if (cap <= b->wfCap && qcap <= b->wfQcap) return EL_OK;
    EL_BHIP(b, b->wRay0.ensure(qcap * 32));
    EL_BHIP(b, b->wRay1.ensure(qcap * 32));
    EL_BHIP(b, b->wHits.ensure(qcap * 16));
    EL_BHIP(b, b->wShH.ensure(qcap * 48));
    EL_BHIP(b, b->wShP.ensure(qcap * 48));
    EL_BHIP(b, b->wT.ensure(cap * 16));
    EL_BHIP(b, b->wL.ensure(cap * 16));
    EL_BHIP(b, b->wPix.ensure(cap * 4));
    EL_BHIP(b, b->wSample.ensure(cap * 4));
    EL_BHIP(b, b->wQuiet.ensure(cap * 8));
    EL_BHIP(b, b->wCounters.ensure(kQCounters * sizeof(uint32_t)));
    b->wfCap = std::max(b->wfCap, cap);
    b->wfQcap = std::max(b->wfQcap, qcap);
    return EL_OK;
}

// Renders every scene of the batch from a cleared accumulation with spp[i]
// (capped at sppCap if nonzero).  Returns the device time in *ms.
static int batchRender(el_batch* b, const el_batch_params* p, uint32_t sppCap, double* ms) {
    const size_t S = b->scenes.size();
    if (S == 0) return bfail(b, EL_ERR_STATE, "empty batch");
    const uint32_t depth = p->max_depth ? p->max_depth : 5;
    if (depth > 64) return bfail(b, EL_ERR_RANGE, "max_depth above 64");
    if (p->bvh_width != 0 && p->bvh_width != 4 && p->bvh_width != 8) return bfail(b, EL_ERR_ARG, "bvh_width must be 4 or 8");
    if (p->block != 0 && (p->block % 64 != 0 || p->block > (uint32_t)kMaxBlock))
        return bfail(b, EL_ERR_ARG, "block must be a multiple of 64, at most 256");
    if (p->refill > 64) return bfail(b, EL_ERR_ARG, "refill must be 0..64");
    EL_BHIP(b, hipSetDevice(b->device));
    const int width = p->bvh_width ? (int)p->bvh_width : b->width;
    if (width != b->width) {
        for (auto* c : b->scenes) {
            int rc = uploadBvh(c, width);
            if (rc) return bfail(b, rc, el_last_error(c));
        }
        b->width = width;
    }
    // scene records: shared spill capacity, per-scene key and accumulation
    int gcap = 1;
    for (auto* c : b->scenes) gcap = std::max(gcap, c->gcap);
    std::vector<DevScene> recs(S);
    std::vector<uint64_t> work(S), npadv(S);
    uint64_t total = 0;
    for (size_t i = 0; i < S; i++) {
        el_ctx* c = b->scenes[i];
        recs[i] = c->ds;
        recs[i].gcap = gcap;
        recs[i].seedLo = (uint32_t)b->seed[i];
        recs[i].seedHi = (uint32_t)(b->seed[i] >> 32);
        recs[i].accum = c->dAccum.as<float4>();
        recs[i].counts = c->dCounts.as<uint32_t>();
        const uint64_t tx = (c->cam.x_res + 7) / 8, ty = (c->cam.y_res + 7) / 8;
        npadv[i] = tx * ty * 64;
        const uint32_t spp = sppCap ? std::min(sppCap, b->spp[i]) : b->spp[i];
        work[i] = (uint64_t)spp * npadv[i];
        total += work[i];
    }
    EL_BHIP(b, upload(b->dScenes, recs.data(), S * sizeof(DevScene), b->stream));
    EL_BHIP(b, b->dGstack.ensure((size_t)kSpillSets * gcap * b->maxThreads * sizeof(int)));
    // generation plan
    uint64_t Q = p->queue_cap;
    if (!Q) {
        size_t freeB = 0, totalB = 0;
        EL_BHIP(b, hipMemGetInfo(&freeB, &totalB));
        Q = std::max<uint64_t>(1u << 20, freeB / 4 / (6 * 32 + 44));
    }
    Q = std::max<uint64_t>(64, (Q / 64) * 64);
    if (Q > total) Q = total;
    if (Q > 0x7fffffffull) return bfail(b, EL_ERR_RANGE, "queue capacity above 2^31");
    std::vector<BatchSeg> segs;
    std::vector<BatchGen> gens;
    {
        size_t sc = 0;
        uint64_t sceneBase = 0;
        for (uint64_t gs = 0; gs < total; gs += Q) {
            const uint64_t ge = std::min(total, gs + Q);
            BatchGen g{segs.size(), 0, (uint32_t)(ge - gs), 0};
            uint64_t w = gs;
            while (w < ge) {
                while (w >= sceneBase + work[sc]) {
                    sceneBase += work[sc];
                    sc++;
                }
                const uint64_t segEnd = std::min(ge, sceneBase + work[sc]);
                BatchSeg s{};
                s.slotStart = (uint32_t)(w - gs);
                s.count = (uint32_t)(segEnd - w);
                s.scene = (uint32_t)sc;
                s.npad = (uint32_t)npadv[sc];
                s.wStart = w - sceneBase;
                s.finStart = g.finTotal;
                g.finTotal += std::min<uint32_t>(s.count, s.npad);
                segs.push_back(s);
                g.nseg++;
                w = segEnd;
            }
            gens.push_back(g);
        }
    }
    EL_BHIP(b, upload(b->dSegs, segs.data(), segs.size() * sizeof(BatchSeg), b->stream));
    LaunchCfg lc;
    lc.stream = b->stream;
    lc.block = p->block ? p->block : kDefaultBlock;
    lc.shadeBlock = 256;
    lc.maxThreads = b->maxThreads;
    lc.refill = p->refill ? p->refill : kDefaultRefill;
    lc.gstack = b->dGstack.as<int>();
    const uint64_t segSize = (Q + kXcds - 1) / kXcds + lc.maxThreads / kXcds + kMaxBlock;
    const uint64_t qcap = std::max<uint64_t>(Q, kXcds * segSize);
    int rc = batchEnsureWf(b, Q, qcap);
    if (rc) return rc;
    WfState q{};
    q.ray[0] = b->wRay0.as<float4>();
    q.ray[1] = b->wRay1.as<float4>();
    q.hits = b->wHits.as<float4>();
    q.shadowH = b->wShH.as<float4>();
    q.shadowP = b->wShP.as<float4>();
    q.pathT = b->wT.as<float4>();
    q.pathL = b->wL.as<float4>();
    q.pathPix = b->wPix.as<uint32_t>();
    q.pathSample = b->wSample.as<uint32_t>();
    q.pathQuiet = b->wQuiet.as<uint64_t>();
    q.pathAov = nullptr;
    q.counters = b->wCounters.as<uint32_t>();
    q.segSize = (uint32_t)segSize;
    q.gstack = lc.gstack;
    q.gthreads = lc.maxThreads;
    // clear the accumulations and tallies (outside any captured graph)
    for (size_t i = 0; i < S; i++) {
        el_ctx* c = b->scenes[i];
        const size_t npx = (size_t)c->cam.x_res * c->cam.y_res;
        EL_BHIP(b, hipMemsetAsync(c->dAccum.p, 0, npx * sizeof(float4), b->stream));
        EL_BHIP(b, hipMemsetAsync(c->dCounts.p, 0, npx * sizeof(uint32_t), b->stream));
    }
    unsigned long long* ctr = b->dCounters.as<unsigned long long>();
    EL_BHIP(b, hipMemsetAsync(ctr, 0, sizeof(unsigned long long) * kCounterWords, b->stream));
    const DevScene* dsc = b->dScenes.as<DevScene>();
    // Shadow launches of small bounces overlap the next extension on a
    // second stream, as in el_render (inside a captured graph the events
    // become the graph's fork and join edges).
    LaunchCfg lcs = lc;
    lcs.stream = b->shadowStream;
    const bool serialShadows = (p->flags & EL_BATCH_SERIAL_SHADOWS) != 0;
    auto launchAll = [&]() -> hipError_t {
        hipError_t e;
        size_t ev = 0;
        for (const BatchGen& g : gens) {
            hipEvent_t shadowsDone = nullptr;
            const BatchSeg* gseg = b->dSegs.as<BatchSeg>() + g.segOffset;
            if ((e = launchBtCamera(lc, dsc, gseg, g.nseg, g.count, depth, !(p->flags & EL_BATCH_NO_CULL), q)) !=
                hipSuccess)
                return e;
            for (uint32_t bo = 0; bo < depth; bo++) {
                RenderArgs a{};
                a.maxDepth = depth;
                a.bounce = bo;
                a.cull = (p->flags & EL_BATCH_NO_CULL) ? 0u : 1u;
                const int cur = bo & 1, nxt = cur ^ 1;
                uint32_t* cntCur = q.counters + (cur ? kQRay1 : kQRay0);
                uint32_t* cntNxt = q.counters + (nxt ? kQRay1 : kQRay0);
                q.parity = bo & 1u;
                const bool overlap = !serialShadows && (g.count < kOverlapPaths || bo >= 2);
                if ((e = launchBtExtend(lc, dsc, width, q, q.ray[cur], cntCur, cntNxt, g.count, ctr)) != hipSuccess)
                    return e;
                if (shadowsDone && (e = hipStreamWaitEvent(b->stream, shadowsDone, 0)) != hipSuccess) return e;
                if ((e = launchBtShade(lc, dsc, a, q, q.ray[cur], cntCur, q.ray[nxt], cntNxt, g.count, ctr)) !=
                    hipSuccess)
                    return e;
                shadowsDone = nullptr;
                if (overlap) {
                    hipEvent_t shaded = batchSyncEvent(b, ev++), done = batchSyncEvent(b, ev++);
                    if (!shaded || !done) return hipErrorOutOfMemory;
                    if ((e = hipEventRecord(shaded, b->stream)) != hipSuccess) return e;
                    if ((e = hipStreamWaitEvent(b->shadowStream, shaded, 0)) != hipSuccess) return e;
                    WfState qs = q;   // the shadow stream's spill area
                    qs.gstack = q.gstack + (size_t)gcap * lc.maxThreads;
                    if ((e = launchBtShadow(lcs, dsc, width, qs, q.shadowH, q.counters + qShadow(q.parity, 0),
                                            g.count, ctr)) != hipSuccess)
                        return e;
                    if ((e = launchBtShadow(lcs, dsc, width, qs, q.shadowP, q.counters + qShadow(q.parity, 1),
                                            g.count, ctr)) != hipSuccess)
                        return e;
                    if ((e = hipEventRecord(done, b->shadowStream)) != hipSuccess) return e;
                    shadowsDone = done;
                } else {
                    if ((e = launchBtShadow(lc, dsc, width, q, q.shadowH, q.counters + qShadow(q.parity, 0), g.count,
                                            ctr)) != hipSuccess)
                        return e;
                    if ((e = launchBtShadow(lc, dsc, width, q, q.shadowP, q.counters + qShadow(q.parity, 1), g.count,
                                            ctr)) != hipSuccess)
                        return e;
                }
            }
            if (shadowsDone && (e = hipStreamWaitEvent(b->stream, shadowsDone, 0)) != hipSuccess) return e;
            if ((e = launchBtFinalize(lc, dsc, gseg, g.nseg, g.finTotal, q, ctr)) != hipSuccess) return e;
        }
        return hipSuccess;
    };
    const bool useGraph = (p->flags & EL_BATCH_GRAPH) != 0;
    if (useGraph) {
        // the captured sequence depends on the plan and the launch shape
        // FNV-1a of the segment table and the generations: kernel arguments
        // (segment counts, slot totals) are baked into the graph
        uint64_t plan = 1469598103934665603ull;
        auto mix = [&](const void* p, size_t n) {
            for (size_t i = 0; i < n; i++) plan = (plan ^ ((const unsigned char*)p)[i]) * 1099511628211ull;
        };
        mix(segs.data(), segs.size() * sizeof(BatchSeg));
        for (const BatchGen& g : gens) {
            const uint64_t f[4] = {g.segOffset, g.nseg, g.count, g.finTotal};
            mix(f, sizeof f);
        }
        std::string key = std::to_string(S) + "/" + std::to_string(depth) + "/" + std::to_string(width) + "/" +
                          std::to_string(p->flags & ~EL_BATCH_GRAPH) + "/" + std::to_string(plan) + "/" +
                          std::to_string(lc.block) + "/" + std::to_string(lc.refill) + "/" + std::to_string(Q) +
                          "/" + std::to_string(sppCap) + "/" + std::to_string(total) + "/" +
                          std::to_string((uintptr_t)b->dSegs.p) + "/" + std::to_string((uintptr_t)b->wRay0.p) +
                          "/" + std::to_string((uintptr_t)b->dScenes.p) + "/" + std::to_string((uintptr_t)b->dGstack.p);
        if (!b->graph || key != b->graphKey) {
            if (b->graph) {
                EL_BHIP(b, hipGraphExecDestroy(b->graph));
                b->graph = nullptr;
            }
            EL_BHIP(b, hipStreamSynchronize(b->stream));
            hipGraph_t gph = nullptr;
            EL_BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t le = launchAll();
            hipError_t ce = hipStreamEndCapture(b->stream, &gph);
            if (le != hipSuccess) return bfail(b, EL_ERR_HIP, std::string("capture: ") + hipGetErrorString(le));
            if (ce != hipSuccess) return bfail(b, EL_ERR_HIP, std::string("end capture: ") + hipGetErrorString(ce));
            hipError_t ie = hipGraphInstantiate(&b->graph, gph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(gph);
            if (ie != hipSuccess) {
                b->graph = nullptr;
                return bfail(b, EL_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
            }
            b->graphKey = key;
        }
        EL_BHIP(b, hipEventRecord(b->e0, b->stream));
        EL_BHIP(b, hipGraphLaunch(b->graph, b->stream));
        EL_BHIP(b, hipEventRecord(b->e1, b->stream));
    } else {
        EL_BHIP(b, hipEventRecord(b->e0, b->stream));
        EL_BHIP(b, launchAll());
        EL_BHIP(b, hipEventRecord(b->e1, b->stream));
    }
    unsigned long long h[kCounterWords];
    EL_BHIP(b, hipMemcpyAsync(h, ctr, sizeof h, hipMemcpyDeviceToHost, b->stream));
    EL_BHIP(b, hipStreamSynchronize(b->stream));
    float t = 0;
    EL_BHIP(b, hipEventElapsedTime(&t, b->e0, b->e1));
    *ms = t;
    for (size_t i = 0; i < S; i++) {
        el_ctx* c = b->scenes[i];
        c->sampleBase = sppCap ? std::min(sppCap, b->spp[i]) : b->spp[i];
        c->accSeed = b->seed[i];
        c->accDepth = depth;
        c->accMixed = true;   // batch RNG streams carry the scene index: no el_adjoint on them
        c->recValid = false;
    }
    b->stats = el_stats{};
    b->stats.rays = tally(h, kStatRays);
    b->stats.vertices = tally(h, kStatVerts);
    b->stats.paths = tally(h, kStatPaths);
    b->stats.render_ms = t;
    b->stats.bvh_width = (uint32_t)width;
    b->stats.extend_launches = gens.size() * depth;
    if (tally(h, kStatOverflow)) return bfail(b, EL_ERR_RANGE, "traversal stack or queue overflow (results incomplete)");
    return EL_OK;
}

extern "C" int el_batch_render(el_batch* b, const el_batch_params* p) {
    if (!b || !p) return bfail(b, EL_ERR_ARG, "null argument");
    double ms = 0;
    return batchRender(b, p, 0, &ms);
}

// Coordinate descent over BVH width, queue capacity, traversal block size
// and refill threshold, each trial rendering every scene at min(spp,
// trial_spp) samples; keeps the fastest setting of each knob in turn.
// A setting replaces the current best only if its trial is this much faster.
constexpr double kTuneMargin = 0.98;

extern "C" int el_batch_tune(el_batch* b, el_batch_params* io, uint32_t trial_spp) {
    if (!b || !io) return bfail(b, EL_ERR_ARG, "null argument");
    if (trial_spp == 0) trial_spp = 2;
    el_batch_params best = *io;
    best.flags &= ~EL_BATCH_GRAPH;
    // explicit defaults, so a sweep value equal to the default is not retried
    if (!best.bvh_width) best.bvh_width = (uint32_t)b->width;
    if (!best.block) best.block = kDefaultBlock;
    if (!best.refill) best.refill = kDefaultRefill;
    std::ostringstream log;
    auto trial = [&](const el_batch_params& p, double* ms) -> int {
        double t = 0;
        int rc = batchRender(b, &p, trial_spp, &t);   // warm (BVH repack, buffers)
        if (rc) return rc;
        rc = batchRender(b, &p, trial_spp, &t);
        if (rc) return rc;
        *ms = t;
        log << "width " << (p.bvh_width ? p.bvh_width : 4) << " queue " << p.queue_cap << " block "
            << (p.block ? p.block : kDefaultBlock) << " refill " << (p.refill ? p.refill : kDefaultRefill) << ": "
            << t << " ms\n";
        return EL_OK;
    };
    double bestMs = 0;
    int rc = trial(best, &bestMs);
    if (rc) return rc;
    auto sweep = [&](uint32_t el_batch_params::*field, std::vector<uint32_t> values) -> int {
        for (uint32_t v : values) {
            el_batch_params p = best;
            p.*field = v;
            if (p.*field == best.*field) continue;
            double ms = 0;
            int r = trial(p, &ms);
            if (r) return r;
            if (ms < bestMs * kTuneMargin) {   // trials closer than the margin are noise
                bestMs = ms;
                best = p;
            }
        }
        return EL_OK;
    };
    if ((rc = sweep(&el_batch_params::bvh_width, {4, 8}))) return rc;
    if ((rc = sweep(&el_batch_params::queue_cap, {1u << 22, 1u << 24, 1u << 26, 0}))) return rc;
    if ((rc = sweep(&el_batch_params::block, {64, 128, 256}))) return rc;
    if ((rc = sweep(&el_batch_params::refill, {8, 16, 32, 64}))) return rc;
    log << "chosen: width " << (best.bvh_width ? best.bvh_width : 4) << " queue " << best.queue_cap << " block "
        << (best.block ? best.block : kDefaultBlock) << " refill " << (best.refill ? best.refill : kDefaultRefill)
        << " (" << bestMs << " ms per trial)\n";
    b->log = log.str();
    best.flags = io->flags;
    *io = best;
    return EL_OK;
}

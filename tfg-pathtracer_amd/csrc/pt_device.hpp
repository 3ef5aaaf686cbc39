// Device-side core of the MI355X path tracer (gfx950, wave64).
//
// Estimator semantics follow the reference kernel (src/tfg-pathtracer/):
//   camera ray          kernel.cu:260-337      hit attributes   Tri.hpp:38-160
//   material fetch      kernel.cu:54-119       Disney           Disney.hpp:41-253
//   sampling            Sampling.hpp:10-54     point light NEE  kernel.cu:175-205
//   HDRI NEE            kernel.cu:210-258, HDRI.hpp:130-162, Texture.hpp:95-207
//   MIS combine         kernel.cu:339-358      path loop        kernel.cu:369-481
// Defined behaviour where the reference is undefined or replaced:
//   * RNG: Philox4x32-10 (Salmon et al., SC'11) keyed by (seed, scene) with
//     counter (pixel, sample, block, 0); uniforms in [0,1) from the top 24 bits.
//   * The HDRI NEE pdf is computed whether or not the shadow ray is blocked
//     (the reference leaves it uninitialised when blocked, kernel.cu:248/254).
//   * Point-light visibility: blocked iff some triangle is hit at
//     t < dist - 0.001 from the offset origin (the reference compares the
//     closest hit's smooth-shading position, kernel.cu:193-197; the two differ
//     only for an occluder within the shading offset of the light).
//   * The HDRI pixel is binarySearch(cdf, r1) bit for bit, found through a
//     guide table (hdri_table.hpp); the point light uses the same r1.
//   * acos arguments are clamped to [-1, 1] for the environment lookup.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hdri_table.hpp"

namespace eleven {
namespace dev {

#define EL_PI 3.14159265358979323846f

// Traversal variants (compile-time).
#ifndef EL_LEAF_PIPE   // leaf triangles: next record loaded before the current one is tested
#define EL_LEAF_PIPE 1
#endif
#ifndef EL_SLAB_AB     // slab distances as A + q B per axis (2 VALU per plane) instead of (b - o) / d
#define EL_SLAB_AB 1
#endif
#ifndef EL_PUSH_BULK   // one stack-room check per node visit instead of one per push
#define EL_PUSH_BULK 1
#endif

// ---------------------------------------------------------------- vectors --
struct v3 {
    float x, y, z;
};
__device__ __forceinline__ v3 mk(float x, float y, float z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 operator*(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ v3 operator*(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 operator*(float s, v3 a) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ v3 operator/(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ v3 operator/(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// Vector.hpp:229-231 form (y component negated difference).
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, -(a.x * b.z - a.z * b.x), a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(v3 a) { return sqrtf(a.x * a.x + a.y * a.y + a.z * a.z); }
// Vector.hpp:173-176: zero vectors pass through unchanged.
__device__ __forceinline__ v3 normalized(v3 a) {
    float l = length(a);
    if (l == 0.f) return a;
    return a / l;
}
// Hardware reciprocal and square roots (v_rcp_f32, v_sqrt_f32, v_rsq_f32:
// 1 ulp) for value arithmetic only -- BRDF and pdf values, MIS weights,
// throughput.  Every quantity that decides a direction, an index or a branch
// of the path (sampling, bases, ray and shadow directions, texel and HDRI
// index math, intersection) keeps IEEE division and square root.
__device__ __forceinline__ float frcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// normalized() for value arithmetic: zero vectors pass through unchanged.
__device__ __forceinline__ v3 fnormalized(v3 a) {
    float l2 = dot(a, a);
    return l2 == 0.f ? a : a * __builtin_amdgcn_rsqf(l2);
}
__device__ __forceinline__ float lerpf(float a, float b, float c) { return a + c * (b - a); }  // FAST_LERP
__device__ __forceinline__ v3 lerp3(v3 a, v3 b, float c) {
    return mk(lerpf(a.x, b.x, c), lerpf(a.y, b.y, c), lerpf(a.z, b.z, c));
}
__device__ __forceinline__ float clampf(float a, float b, float c) { return a < b ? b : (a > c ? c : a); }
__device__ __forceinline__ float minf(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float maxf(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ v3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

// ------------------------------------------------------------------- RNG --
// Philox4x32-10, Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC'11.  Multipliers and Weyl key increments as published.
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        // one 32 x 32 -> 64-bit multiply per product (v_mad_u64_u32) instead
        // of separate high and low multiplies
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x, p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}
__device__ __forceinline__ float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }

// Random numbers of one path, one Philox block per use:
//   block 0      camera r1..r4          block 1 .x   camera r5
//   block 2 + b  bounce b: DisneySample r1, r2, r3 (.x .y .z), shade r1 (.w)
// (shade's r2 and r3 are never read on the reference's HDRI-IS path,
// kernel.cu:234-257, so they are not drawn).
struct PathRng {
    uint32_t k0, k1, pixel, sample, scene;
    __device__ __forceinline__ uint4 block(uint32_t b) const {
        return philox4x32_10(make_uint4(pixel, sample, b, scene), k0, k1);
    }
};

// ----------------------------------------------------------- scene layout --
// Texel storage: kTexF32 = RGB FP32 in the float pool (offset = first
// float); kTexLdr + k = one 32-bit word per texel in the 8-bit pool (offset =
// first texel), bytes R, G, B indexing the loader's byte -> value table k
// (0 sRGB, 1 linear): the same FP32 values in a third of the bytes.
constexpr int32_t kTexF32 = 0, kTexLdr = 1;
struct DevTex {
    int32_t w, h;
    float xTile, yTile, xOffset, yOffset;
    int32_t filter;
    int32_t format;
    uint64_t offset;
};

struct DevMaterial {
    int32_t albedoTex, emissionTex, roughnessTex, metallicTex, normalTex, pad0;
    float albedo[3], emission[3];
    float roughness, metallic, clearcoatGloss, clearcoat, anisotropic, eta, transmission, specular,
        specularTint, sheenTint, subsurface, sheen;
    float roughnessEff, metallicEff;   // powf(roughness, 2.2f), powf(metallic, 2.2f) (kernel.cu:113-114)
    float pad1[2];
};

struct DevCamera {
    uint32_t xRes, yRes;
    float focalLength, sensorWidth, sensorHeight, aperture, focusDistance;
    float cosRot[3], sinRot[3];   // cos/sin of the Euler angles in radians, per camera
    float pos[3];
    int32_t bokeh;
};

// Shading record per original triangle: 8 float4 = 128 B, one cache line
// per hit: [0..8] v0 v1 v2 | [9..17] n0 n1 n2 | [18..23] uv0 uv1 uv2 (x,y) |
// [24] material id (int bits) | [25] tangentsSign | [26..31] unused.
// The corner tangents, needed only for normal maps and the tangent AOVs,
// are a separate stream of 3 float4 per triangle: t0 t1 t2 | unused.
constexpr int kShadeWords = 8, kShadeFloats = 4 * kShadeWords;
constexpr int kTanWords = 3, kTanFloats = 4 * kTanWords;

struct DevScene {
    const void* nodes;             // WideNode<W>[]
    const float4* tris;            // 3 float4 per leaf-order triangle
    const float4* shade;           // kShadeWords per original triangle
    const float4* tangents;        // kTanWords per original triangle
    const DevMaterial* mats;
    const DevTex* texs;
    const float* texels;           // pool of all textures (RGB FP32)
    const uint32_t* texels8;       // 8-bit copies of the LDR textures (DevTex::format)
    const float* texLut;           // 2 x 256 byte -> value tables (sRGB, linear)
    const float* lights;           // 6 floats per light: position, radiance
    const float* envCdf;           // HDRI CDF (L + 1)
    const int32_t* envGuide;       // guide table (2^envLog2G + 1 pairs)
    const int32_t* envMap;         // per interval: open-interval and tie results
    const uint32_t* envBad;        // 2^16-bit map: random-number buckets whose HDRI pdf may be 0 or inf
    const float4* envTab;          // 2 per HDRI pixel k: envSampleAt(k) (direction + pdf, value)
    DevTex env;                    // env texture (texels in the pool)
    float envRadianceSum;
    int32_t envLog2G;
    int32_t nLights;
    int32_t nTris;
    int32_t width;                 // BVH width (4 or 8)
    int32_t gcap;                  // global stack entries per thread
    DevCamera cam;
    // batches only: the scene's RNG key and accumulation buffers
    uint32_t seedLo, seedHi;
    float4* accum;
    uint32_t* counts;
};

// --------------------------------------------------------------- textures --
struct TexPools {
    const float* f32;
    const uint32_t* u8;
    const float* lut;   // the byte -> value tables (global or an LDS copy)
};
__device__ __forceinline__ TexPools texPools(const DevScene& s, const float* lut = nullptr) {
    return TexPools{s.texels, s.texels8, lut ? lut : s.texLut};
}
// Texture::getValueFromCoordinates (Texture.hpp:95-106).
__device__ __forceinline__ v3 texel(const TexPools& P, const DevTex& t, int x, int y) {
    x = (int)(t.xTile * (x + t.xOffset * t.w)) % t.w;
    y = (int)(t.yTile * (y + t.yOffset * t.h)) % t.h;
    // Out-of-range coordinates (negative uv) are undefined in the reference;
    // wrap them into the texture instead of reading out of bounds.
    if (x < 0) x += t.w;
    if (y < 0) y += t.h;
    const uint64_t i = (uint64_t)y * t.w + x;
    if (t.format == kTexF32) return ld3(P.f32 + t.offset + 3 * i);
    const uint32_t v = P.u8[t.offset + i];
    const float* L = P.lut + 256 * (t.format - kTexLdr);
    return mk(L[v & 255u], L[(v >> 8) & 255u], L[(v >> 16) & 255u]);
}
// Index (texel units, within the texture) that texel() reads for (x, y).
__device__ __forceinline__ uint32_t texelIndex(const DevTex& t, int x, int y) {
    x = (int)(t.xTile * (x + t.xOffset * t.w)) % t.w;
    y = (int)(t.yTile * (y + t.yOffset * t.h)) % t.h;
    if (x < 0) x += t.w;
    if (y < 0) y += t.h;
    return (uint32_t)y * (uint32_t)t.w + (uint32_t)x;
}
__device__ __forceinline__ v3 texUV(const TexPools& pool, const DevTex& t, float u, float v) {
    return texel(pool, t, (int)(u * t.w), (int)(v * t.h));
}
// Texture.hpp:114-135
__device__ __forceinline__ v3 texBilinear(const TexPools& pool, const DevTex& t, float u, float v) {
    float x = u * t.w, y = v * t.h;
    float t1x = floorf(x), t1y = floorf(y);
    float t2x = t1x + 1, t2y = t1y + 1;
    float a = (x - t1x) / (t2x - t1x), b = (y - t1y) / (t2y - t1y);
    v3 v1 = texel(pool, t, (int)t1x, (int)t1y), v2 = texel(pool, t, (int)t2x, (int)t1y);
    v3 v3_ = texel(pool, t, (int)t1x, (int)t2y), v4 = texel(pool, t, (int)t2x, (int)t2y);
    return lerp3(lerp3(v1, v2, a), lerp3(v3_, v4, a), b);
}
__device__ __forceinline__ v3 texFiltered(const TexPools& pool, const DevTex& t, float u, float v) {
    return t.filter == 1 ? texBilinear(pool, t, u, v) : texUV(pool, t, u, v);
}

// Texture::sphericalMapping (Texture.hpp:144-156) for a unit sphere at origin.
__device__ __forceinline__ void sphericalMapping(v3 p, float& u, float& v) {
    float theta = acosf(clampf(-p.y, -1.f, 1.f));
    float phi = atan2f(-p.z, p.x) + EL_PI;
    u = phi / (2 * EL_PI);
    v = theta / EL_PI;
    // limitUV (Math.hpp:38-41)
    u += -(float)(u > 1) + -(float)(u < 0);
    v += -(float)(v > 1) + -(float)(v < 0);
}

// -------------------------------------------------------------- sampling --
__device__ __forceinline__ v3 cosineSampleHemisphere(float u1, float u2) {   // Sampling.hpp:31-41
    float r = sqrtf(u1);
    float phi = 2.0f * EL_PI * u2;
    v3 d;
    d.x = r * cosf(phi);
    d.y = r * sinf(phi);
    d.z = sqrtf(maxf(0.0f, 1.0f - d.x * d.x - d.y * d.y));
    return d;
}
__device__ __forceinline__ v3 importanceSampleGGX(float rgh, float r1, float r2) {   // Sampling.hpp:43-54
    float a = maxf(0.001f, rgh);
    float phi = r1 * EL_PI * 2;
    float cosTheta = sqrtf((1.0f - r2) / (1.0f + (a * a - 1.0f) * r2));
    float sinTheta = clampf(sqrtf(1.0f - (cosTheta * cosTheta)), 0.0f, 1.0f);
    return mk(sinTheta * cosf(phi), sinTheta * sinf(phi), cosTheta);
}
__device__ __forceinline__ void uniformCircleSampling(float u1, float u2, float u3, float& x, float& y) {
    float t = 2 * EL_PI * u1;
    float u = u2 + u3;
    float r = u > 1 ? 2 - u : u;
    x = r * cosf(t);
    y = r * sinf(t);
}

// ---------------------------------------------------------------- Disney --
struct HitData {
    float metallic, roughness, clearcoatGloss, clearcoat, anisotropic, eta, transmission, specular,
        specularTint, sheenTint, subsurface, sheen;
    v3 emission, albedo, normal;
};

__device__ __forceinline__ void createBasis(v3 n, v3& t, v3& b) {   // Disney.hpp:41-45
    t = normalized(cross(mk(0, 1, 0), n));
    b = normalized(cross(n, t));
}
__device__ __forceinline__ float schlickFresnel(float u) {
    float m = clampf(1.0f - u, 0.0f, 1.0f);
    float m2 = m * m;
    return m2 * m2 * m;
}
// The microfacet terms below are value arithmetic (frcp, fsqrt).
__device__ __forceinline__ float GTR1(float NDotH, float a) {
    if (a >= 1.0f) return 1.0f / EL_PI;
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * NDotH * NDotH;
    return (a2 - 1.0f) * frcp(EL_PI * logf(a2) * t);
}
// 1 / (pi ax ay c^2) with rax = 1 / ax, ray = 1 / ay, piAxAy = pi * ax * ay.
__device__ __forceinline__ float GTR2_aniso(float NDotH, float HDotX, float HDotY, float rax, float ray,
                                            float piAxAy) {
    float a = HDotX * rax, b = HDotY * ray;
    float c = a * a + b * b + NDotH * NDotH;
    return frcp(piAxAy * c * c);
}
__device__ __forceinline__ float smithG_GGX(float NDotV, float alphaG) {
    float a = alphaG * alphaG, b = NDotV * NDotV;
    return frcp(NDotV + fsqrt(a + b - a * b));
}
__device__ __forceinline__ float smithG_GGX_aniso(float NDotV, float VDotX, float VDotY, float ax, float ay) {
    float a = VDotX * ax, b = VDotY * ay, c = NDotV;
    return frcp(NDotV + fsqrt(a * a + b * b + c * c));
}

// Anisotropic GGX widths of Disney.hpp:128-130 / 213-215.
struct GgxWidths {
    float ax, ay, rax, ray, piAxAy;
};
__device__ __forceinline__ GgxWidths ggxWidths(const HitData& h) {
    GgxWidths g;
    float aspect = fsqrt(1.0f - h.anisotropic * 0.9f);
    float ra = frcp(aspect);
    g.ax = maxf(0.001f, h.roughness * ra);
    g.ay = maxf(0.001f, h.roughness * aspect);
    g.rax = frcp(g.ax);
    g.ray = frcp(g.ay);
    g.piAxAy = EL_PI * g.ax * g.ay;
    return g;
}

// Disney.hpp:108-147 given the half vector H = normalize(L + V) and
// Ds = GTR2_aniso(N.H, H.T, H.B).  `dir` is the incoming ray direction
// (V = -dir).  With clearcoat == 0 the GTR1/GTR2 mix weight is exactly 1.
__device__ __forceinline__ float disneyPdfH(const HitData& h, v3 L, v3 H, float NDotH, float Ds) {
    v3 N = h.normal;
    if (dot(N, L) <= 0.0f) return 1.0f;
    float diffuseRatio = 0.5f * (1.0f - h.metallic);
    float specularRatio = 1.0f - diffuseRatio;
    float pdfGTR2 = Ds * NDotH;
    float mix = pdfGTR2;
    if (h.clearcoat != 0.0f) {
        float clearcoatAlpha = lerpf(0.1f, 0.001f, h.clearcoatGloss);
        float pdfGTR1 = GTR1(NDotH, clearcoatAlpha) * NDotH;
        mix = lerpf(pdfGTR1, pdfGTR2, frcp(1.0f + h.clearcoat));
    }
    float pdfSpec = mix * frcp(4.0f * fabsf(dot(L, H)));
    float pdfDiff = fabsf(dot(L, N)) * (1.0f / EL_PI);
    return diffuseRatio * pdfDiff + specularRatio * pdfSpec;
}
__device__ __forceinline__ float disneyPdf(v3 dir, const HitData& h, v3 L) {
    v3 V = dir * -1.0f;
    v3 H = fnormalized(L + V);
    v3 T, B;
    createBasis(h.normal, T, B);
    float NDotH = fabsf(dot(h.normal, H));
    GgxWidths g = ggxWidths(h);
    return disneyPdfH(h, L, H, NDotH, GTR2_aniso(NDotH, dot(H, T), dot(H, B), g.rax, g.ray, g.piAxAy));
}

// Disney.hpp:150-177
__device__ __forceinline__ v3 disneySample(v3 dir, const HitData& h, float r1, float r2, float r3) {
    v3 N = h.normal;
    v3 V = dir * -1.0f;
    v3 T, B;
    createBasis(h.normal, T, B);
    float diffuseRatio = 0.5f * (1.0f - h.metallic);
    if (r3 < diffuseRatio) {
        v3 H = cosineSampleHemisphere(r1, r2);
        return T * H.x + B * H.y + N * H.z;
    }
    v3 H = importanceSampleGGX(h.roughness, r1, r2);
    H = T * H.x + B * H.y + N * H.z;
    v3 v1 = V * -1.0f;
    return v1 - 2 * dot(v1, H) * H;   // reflect (Vector.hpp:218-220)
}

// Disney.hpp:179-253; with PDF also disneyPdf(dir, h, L) into `pdf`, which
// shares the half vector and the GTR2 term.  The clearcoat lobe is skipped
// when its weight is 0 (it is finite, so it adds exactly 0).
template <bool PDF>
__device__ __forceinline__ v3 disneyEvalT(v3 dir, const HitData& h, v3 L, float& pdf) {
    v3 V = dir * -1.0f;
    v3 T, B;
    createBasis(h.normal, T, B);
    v3 H = fnormalized(L + V);
    float NDotL = fabsf(dot(h.normal, L));
    float NDotV = fabsf(dot(h.normal, V));
    float NDotH = fabsf(dot(h.normal, H));
    float LDotH = fabsf(dot(L, H));
    GgxWidths g = ggxWidths(h);
    float Ds = GTR2_aniso(NDotH, dot(H, T), dot(H, B), g.rax, g.ray, g.piAxAy);
    if (PDF) pdf = disneyPdfH(h, L, H, NDotH, Ds);
    v3 brdf = mk(0, 0, 0);
    if (h.transmission < 1.0f && dot(h.normal, L) > 0.0f && dot(h.normal, V) > 0.0f) {
        v3 Cdlin = h.albedo;
        float Cdlum = 0.3f * Cdlin.x + 0.6f * Cdlin.y + 0.1f * Cdlin.z;
        v3 Ctint = Cdlum > 0.0f ? Cdlin * frcp(Cdlum) : mk(1, 1, 1);
        v3 Cspec0 = lerp3(h.specular * 0.08f * lerp3(mk(1, 1, 1), Ctint, h.specularTint), Cdlin, h.metallic);
        v3 Csheen = lerp3(mk(1, 1, 1), Ctint, h.sheenTint);
        float FL = schlickFresnel(NDotL), FV = schlickFresnel(NDotV);
        float Fd90 = 0.5f + 2.0f * LDotH * LDotH * h.roughness;
        float Fd = lerpf(1.0f, Fd90, FL) * lerpf(1.0f, Fd90, FV);
        float Fss90 = LDotH * LDotH * h.roughness;
        float Fss = lerpf(1.0f, Fss90, FL) * lerpf(1.0f, Fss90, FV);
        float ss = 1.25f * (Fss * (frcp(NDotL + NDotV) - 0.5f) + 0.5f);
        float FH = schlickFresnel(LDotH);
        v3 Fs = lerp3(Cspec0, mk(1, 1, 1), FH);
        float Gs = smithG_GGX_aniso(NDotL, dot(L, T), dot(L, B), g.ax, g.ay);
        Gs *= smithG_GGX_aniso(NDotV, dot(V, T), dot(V, B), g.ax, g.ay);
        v3 Fsheen = FH * h.sheen * Csheen;
        brdf = ((1.0f / EL_PI) * lerpf(Fd, ss, h.subsurface) * Cdlin + Fsheen) * (1.0f - h.metallic) + Gs * Fs * Ds;
        if (h.clearcoat != 0.0f) {
            float Dr = GTR1(NDotH, lerpf(0.1f, 0.001f, h.clearcoatGloss));
            float Fr = lerpf(0.04f, 1.0f, FH);
            float Gr = smithG_GGX(NDotL, 0.25f) * smithG_GGX(NDotV, 0.25f);
            brdf = brdf + mk(1, 1, 1) * (0.25f * h.clearcoat * Gr * Fr * Dr);
        }
    }
    return brdf;
}
__device__ __forceinline__ v3 disneyEval(v3 dir, const HitData& h, v3 L) {
    float unused;
    return disneyEvalT<false>(dir, h, L, unused);
}
__device__ __forceinline__ v3 disneyEvalPdf(v3 dir, const HitData& h, v3 L, float& pdf) {
    return disneyEvalT<true>(dir, h, L, pdf);
}

// d disneyEval / d (albedo.x, albedo.y, albedo.z, roughness, metallic,
// specular) of HitData (effective values, after generateHitData's pow 2.2),
// written out analytically from the expression above; used by the adjoint.
__device__ __forceinline__ void disneyEvalGrad(v3 dir, const HitData& h, v3 L, v3 (&df)[6]) {
#pragma unroll
    for (int k = 0; k < 6; k++) df[k] = mk(0, 0, 0);
    v3 V = dir * -1.0f;
    v3 T, B;
    createBasis(h.normal, T, B);
    v3 H = normalized(L + V);
    float NDotL = fabsf(dot(h.normal, L));
    float NDotV = fabsf(dot(h.normal, V));
    float NDotH = fabsf(dot(h.normal, H));
    float LDotH = fabsf(dot(L, H));
    if (!(h.transmission < 1.0f && dot(h.normal, L) > 0.0f && dot(h.normal, V) > 0.0f)) return;
    const v3 C = h.albedo;
    const float wl[3] = {0.3f, 0.6f, 0.1f};
    const float lum = wl[0] * C.x + wl[1] * C.y + wl[2] * C.z;
    const v3 Ctint = lum > 0.0f ? C / lum : mk(1, 1, 1);
    const float mu = h.metallic, rho = h.roughness;
    const v3 tintMix = mk(1, 1, 1) + (Ctint - mk(1, 1, 1)) * h.specularTint;
    const float S0 = h.specular * 0.08f;
    const v3 Cspec0 = S0 * tintMix + (C - S0 * tintMix) * mu;
    const v3 Csheen = mk(1, 1, 1) + (Ctint - mk(1, 1, 1)) * h.sheenTint;
    const float FL = schlickFresnel(NDotL), FV = schlickFresnel(NDotV), FH = schlickFresnel(LDotH);
    const float Fd90 = 0.5f + 2.0f * LDotH * LDotH * rho;
    const float A = 1 + FL * (Fd90 - 1), Bd = 1 + FV * (Fd90 - 1);
    const float Fd = A * Bd, dFd = 2 * LDotH * LDotH * (FL * Bd + FV * A);
    const float Fss90 = LDotH * LDotH * rho;
    const float A2 = 1 + FL * (Fss90 - 1), B2 = 1 + FV * (Fss90 - 1);
    const float Fss = A2 * B2, dFss = LDotH * LDotH * (FL * B2 + FV * A2);
    const float k1 = 1.0f / (NDotL + NDotV) - 0.5f;
    const float ss = 1.25f * (Fss * k1 + 0.5f), dss = 1.25f * k1 * dFss;
    const float D0 = Fd + h.subsurface * (ss - Fd), dD0 = dFd + h.subsurface * (dss - dFd);
    const float aspect = sqrtf(1.0f - h.anisotropic * 0.9f);
    const float ax = maxf(0.001f, rho / aspect), ay = maxf(0.001f, rho * aspect);
    const float dax = rho / aspect > 0.001f ? 1.0f / aspect : 0.0f, day = rho * aspect > 0.001f ? aspect : 0.0f;
    const float HX = dot(H, T), HY = dot(H, B);
    const float a = HX / ax, b = HY / ay, c = a * a + b * b + NDotH * NDotH;
    const float Ds = 1.0f / (EL_PI * ax * ay * c * c);
    const float dDs = Ds * (-1.0f / ax + 4.0f * a * a / (c * ax)) * dax + Ds * (-1.0f / ay + 4.0f * b * b / (c * ay)) * day;
    float GL, dGL, GV, dGV;
    {
        float X = dot(L, T), Y = dot(L, B);
        float q = sqrtf((X * ax) * (X * ax) + (Y * ay) * (Y * ay) + NDotL * NDotL);
        GL = 1.0f / (NDotL + q);
        dGL = -GL * GL * ((X * X * ax) * dax + (Y * Y * ay) * day) / q;
    }
    {
        float X = dot(V, T), Y = dot(V, B);
        float q = sqrtf((X * ax) * (X * ax) + (Y * ay) * (Y * ay) + NDotV * NDotV);
        GV = 1.0f / (NDotV + q);
        dGV = -GV * GV * ((X * X * ax) * dax + (Y * Y * ay) * day) / q;
    }
    const float Gs = GL * GV, dGs = dGL * GV + GL * dGV;
    const v3 Fs = Cspec0 + (mk(1, 1, 1) - Cspec0) * FH;
    const v3 Fsheen = Csheen * (FH * h.sheen);
    const v3 diffusePart = C * ((1.0f / EL_PI) * D0) + Fsheen;
    const float spec = Gs * Ds * (1.0f - FH);
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const v3 e = mk(k == 0 ? 1.f : 0.f, k == 1 ? 1.f : 0.f, k == 2 ? 1.f : 0.f);
        const v3 dCtint = lum > 0.0f ? e / lum - C * (wl[k] / (lum * lum)) : mk(0, 0, 0);
        const v3 dCspec0 = (S0 * h.specularTint * (1.0f - mu)) * dCtint + e * mu;
        const v3 dFsheen = dCtint * (FH * h.sheen * h.sheenTint);
        df[k] = (e * ((1.0f / EL_PI) * D0) + dFsheen) * (1.0f - mu) + dCspec0 * spec;
    }
    df[3] = C * ((1.0f / EL_PI) * dD0 * (1.0f - mu)) + Fs * (dGs * Ds + Gs * dDs);
    df[4] = diffusePart * -1.0f + (C - S0 * tintMix) * spec;
    df[5] = tintMix * (0.08f * (1.0f - mu) * spec);
}

// ----------------------------------------------------------------- camera --
// kernel.cu:260-337 (double-promoted terms kept as in the reference).
__device__ __forceinline__ void cameraRay(const DevCamera& c, int x, int y, float r1, float r2, float r3, float r4,
                                          float r5, v3& o, v3& d) {
    float dx = c.pos[0] + ((float)x) / ((float)c.xRes) * c.sensorWidth;
    float dy = c.pos[1] + ((float)y) / ((float)c.yRes) * c.sensorHeight;
    float odx = (float)((-c.sensorWidth / 2.0) + dx);
    float ody = (float)((-c.sensorHeight / 2.0) + dy);
    float rx = (float)((1.0 / (float)c.xRes) * (r1 - 0.5) * c.sensorWidth);
    float ry = (float)((1.0 / (float)c.yRes) * (r2 - 0.5) * c.sensorHeight);
    float SPx = odx + rx, SPy = ody + ry;
    float SPz = c.pos[2] + c.focalLength;
    // XYZ Euler rotation; the per-camera cos/sin are evaluated once on the host
    // (the reference evaluates the same six values for every ray)
    const float cx = c.cosRot[0], sx = c.sinRot[0], cy = c.cosRot[1], sy = c.sinRot[1];
    const float cz = c.cosRot[2], sz = c.sinRot[2];
    v3 dir = mk(SPx, SPy, SPz) - mk(c.pos[0], c.pos[1], c.pos[2]);
    v3 dX = mk(dir.x, dir.y * cx - dir.z * sx, dir.y * sx + dir.z * cx);
    v3 dY = mk(dX.x * cy + dX.z * sy, dX.y, dX.z * cy - dX.x * sy);
    v3 dZ = mk(dY.x * cz - dY.y * sz, dY.x * sz + dY.y * cz, dY.z);
    o = mk(c.pos[0], c.pos[1], c.pos[2]);
    d = normalized(dZ);
    if (c.bokeh) {
        float diameter = c.focalLength / c.aperture;
        float l = c.focusDistance + c.focalLength;
        v3 focus = o + d * l;
        float px, py;
        uniformCircleSampling(r3, r4, r5, px, py);
        px *= diameter * 0.5f;
        py *= diameter * 0.5f;
        v3 orig = o + mk(px, py, 0);
        o = orig;
        d = normalized(focus - orig);
    }
}

// --------------------------------------------------------------- hit data --
struct SurfaceHit {
    v3 position, normal;
    float tu, tv;
    int32_t material, tri;
    float u, v, sign;   // barycentrics and tangentsSign, for surfaceTangents
};

// Tri::hit's attribute part (Tri.hpp:70-157) from the traversal's (t, u, v).
__device__ __forceinline__ SurfaceHit surfaceAt(const DevScene& s, int tri, float t, float u, float v, v3 o, v3 d) {
    const float4* R = s.shade + (uint64_t)tri * kShadeWords;
    const float4 A = R[0], B = R[1], C = R[2], D = R[3], E = R[4], F = R[5], G = R[6];
    SurfaceHit h;
    v3 v0 = mk(A.x, A.y, A.z), v1 = mk(A.w, B.x, B.y), v2 = mk(B.z, B.w, C.x);
    v3 n0 = mk(C.y, C.z, C.w), n1 = mk(D.x, D.y, D.z), n2 = mk(D.w, E.x, E.y);
    float uv0x = E.z, uv0y = E.w, uv1x = F.x, uv1y = F.y, uv2x = F.z, uv2y = F.w;
    h.material = __float_as_int(G.x);
    h.sign = G.y;
    h.tri = tri;
    h.u = u;
    h.v = v;
    h.tu = uv0x + (uv1x - uv0x) * u + (uv2x - uv0x) * v;
    h.tv = uv0y + (uv1y - uv0y) * u + (uv2y - uv0y) * v;
    v3 geom = o + d * t;
    v3 Ns = n0 + (n1 - n0) * u + (n2 - n0) * v;
    v3 p0 = geom - dot(geom - v0, n0) * n0;
    v3 p1 = geom - dot(geom - v1, n1) * n1;
    v3 p2 = geom - dot(geom - v2, n2) * n2;
    v3 sp = p0 + (p1 - p0) * u + (p2 - p0) * v;
    bool convex = dot(sp - geom, Ns) > 0.0f;
    h.position = convex ? sp : geom;
    h.normal = Ns;
    return h;
}

// Tangent frame at a hit (Tri.hpp:78-79, 153): the interpolated corner
// tangents (not renormalised) and bitangent = tangentsSign * N x T.
__device__ __forceinline__ void surfaceTangents(const DevScene& s, const SurfaceHit& h, v3& T, v3& B) {
    const float4* R = s.tangents + (uint64_t)h.tri * kTanWords;
    const float4 a = R[0], b = R[1], c = R[2];
    const v3 t0 = mk(a.x, a.y, a.z), t1 = mk(a.w, b.x, b.y), t2 = mk(b.z, b.w, c.x);
    T = t0 + (t1 - t0) * h.u + (t2 - t0) * h.v;
    B = h.sign * cross(h.normal, T);
}

// generateHitData (kernel.cu:54-119)
// `lut`: an LDS copy of s.texLut, or null for the global one.
__device__ __forceinline__ HitData hitData(const DevScene& s, const SurfaceHit& h, const float* lut = nullptr) {
    const DevMaterial& m = s.mats[h.material];
    const TexPools P = texPools(s, lut);
    HitData d;
    d.albedo = m.albedoTex < 0 ? ld3(m.albedo) : texFiltered(P, s.texs[m.albedoTex], h.tu, h.tv);
    d.emission = m.emissionTex < 0 ? ld3(m.emission) : texFiltered(P, s.texs[m.emissionTex], h.tu, h.tv);
    // powf(x, 2.2f) of kernel.cu:113-114: for untextured parameters computed
    // once per material at upload
    d.roughness = m.roughnessTex < 0 ? m.roughnessEff
                                     : powf(texFiltered(P, s.texs[m.roughnessTex], h.tu, h.tv).x, 2.2f);
    d.metallic = m.metallicTex < 0 ? m.metallicEff
                                   : powf(texFiltered(P, s.texs[m.metallicTex], h.tu, h.tv).x, 2.2f);
    if (m.normalTex < 0) {
        d.normal = h.normal;
    } else {
        v3 c = texUV(P, s.texs[m.normalTex], h.tu, h.tv);
        v3 ln = c * 2.0f - mk(1, 1, 1);
        v3 T, B;
        surfaceTangents(s, h, T, B);
        d.normal = normalized(ln.x * T - ln.y * B + ln.z * h.normal);
    }
    d.clearcoatGloss = m.clearcoatGloss;
    d.clearcoat = m.clearcoat;
    d.anisotropic = m.anisotropic;
    d.eta = m.eta;
    d.transmission = m.transmission;
    d.specular = m.specular;
    d.specularTint = m.specularTint;
    d.sheenTint = m.sheenTint;
    d.subsurface = m.subsurface;
    d.sheen = m.sheen;
    return d;
}

// Escaped ray: env(-dir), nearest/filtered lookup (kernel.cu:414-419).
__device__ __forceinline__ v3 envLookup(const DevScene& s, v3 d) {
    float u, v;
    sphericalMapping(d * -1.0f, u, v);
    return texFiltered(texPools(s), s.env, u, v);
}

// HDRI NEE sample (kernel.cu:236-254 with HDRI.hpp:145-162, Texture.hpp:177-207).
struct EnvSample {
    v3 dir, value;
    float pdf;
    v3 shadowDir;   // normalized(dir), the shadow ray's direction
};
// Everything the NEE sample derives from the environment pixel k alone;
// tabulated per pixel at upload (envTab) by this same code.
__device__ __forceinline__ EnvSample envSampleAt(const DevScene& s, int k) {
    const DevTex& t = s.env;
    int x = k % t.w, y = k / t.w;
    float nu = x / (float)t.w, nv = y / (float)t.h;
    // inverseTransformUV
    int ix = (int)(nu * t.w), iy = (int)(nv * t.h);
    ix = (int)(t.xTile * (ix - t.xOffset * t.w)) % t.w;
    iy = (int)(t.yTile * (iy - t.yOffset * t.h)) % t.h;
    float iu = (float)ix / (float)t.w, iv = (float)iy / (float)t.h;
    iu += -(float)(iu > 1) + -(float)(iu < 0);
    iv += -(float)(iv > 1) + -(float)(iv < 0);
    // reverseSphericalMapping, negated and normalised
    float phi = iu * 2 * EL_PI, theta = iv * EL_PI;
    float px = cosf(phi - EL_PI), py = -cosf(theta), pz = -sinf(phi - EL_PI);
    float a = sqrtf(1 - py * py);
    EnvSample e;
    e.dir = normalized(mk(a * px, py, a * pz)) * -1.0f;
    e.value = texUV(texPools(s), t, iu, iv);
    // HDRI::pdf((int)(iu*W), (int)(iv*H))
    int qx = (int)(iu * t.w), qy = (int)(iv * t.h);
    v3 dv = texel(texPools(s), t, qx, qy);
    float th = (((float)qy / (float)t.h)) * EL_PI;
    e.pdf = ((dv.x + dv.y + dv.z) / s.envRadianceSum) * (float)(t.w * t.h) / (2.0f * EL_PI * sinf(th));
    e.shadowDir = normalized(e.dir);
    return e;
}
__device__ __forceinline__ EnvSample envSample(const DevScene& s, float r1) {
    // HDRI::sample(r1): exactly binarySearch(cdf, r1, W*H)
    const int k = hdriLookup(s.envCdf, s.envGuide, s.envMap, s.envLog2G, r1);
    const float4 a = s.envTab[2 * k], b = s.envTab[2 * k + 1];
    EnvSample e;
    e.dir = mk(a.x, a.y, a.z);
    e.pdf = a.w;
    e.value = mk(b.x, b.y, b.z);
    e.shadowDir = normalized(e.dir);   // as envSampleAt
    return e;
}

// ------------------------------------------------------------- traversal --
// Moller-Trumbore exactly as Tri::hit (Tri.hpp:38-68): FP32, no contraction,
// IEEE division for 1/det, inclusive barycentric bounds, t >= 0.
__device__ __forceinline__ bool mtHit(float4 A, float4 B, float4 C, float ox, float oy, float oz, float dx, float dy,
                                      float dz, float& t, float& u, float& v) {
#pragma clang fp contract(off)
    const float EPS = 0.0000001f;
    float e1x = A.w, e1y = B.x, e1z = B.y;
    float e2x = B.z, e2y = B.w, e2z = C.x;
    float px = dy * e2z - dz * e2y;
    float py = -(dx * e2z - dz * e2x);
    float pz = dx * e2y - dy * e2x;
    float det = e1x * px + e1y * py + e1z * pz;
    if (det > -EPS && det < EPS) return false;
    float inv = 1.0f / det;
    float tx = ox - A.x, ty = oy - A.y, tz = oz - A.z;
    u = (tx * px + ty * py + tz * pz) * inv;
    if (u < 0.0f || u > 1.0f) return false;
    float qx = ty * e1z - tz * e1y;
    float qy = -(tx * e1z - tz * e1x);
    float qz = tx * e1y - ty * e1x;
    v = (dx * qx + dy * qy + dz * qz) * inv;
    if (v < 0.0f || (u + v) > 1.0f) return false;
    t = (e2x * qx + e2y * qy + e2z * qz) * inv;
    if (t < 0.0f) return false;
    return true;
}

constexpr int kShortStack = 16;   // LDS entries per lane (power of two)

// Per-lane traversal stack: the top kShortStack entries live in LDS
// (lane-interleaved, lds[k*blockDim + tid]); older entries spill to a global
// per-thread area (gstack[k*gstride]).  The host sizes the global part from
// the tree's worst-case depth so it never overflows; if it would, the entry is
// dropped and `overflow` is set (no out-of-bounds access).
struct TravStack {
    int* lds;
    int ldsStride;
    int* g;
    int gstride;
    int gcap;
    int sp, lo;
    bool overflow;
    __device__ __forceinline__ void push(int e) {
        if (sp - lo == kShortStack) {
            int old = lds[(lo & (kShortStack - 1)) * ldsStride];
            if (lo < gcap)
                g[(uint64_t)lo * gstride] = old;
            else
                overflow = true;
            lo++;
        }
        lds[(sp & (kShortStack - 1)) * ldsStride] = e;
        sp++;
    }
    // Room for k more pushes (k <= kShortStack): the oldest LDS entries go
    // to the global area first, as push would spill them one at a time.
    __device__ __forceinline__ void reserve(int k) {
        while (sp - lo + k > kShortStack) {
            int old = lds[(lo & (kShortStack - 1)) * ldsStride];
            if (lo < gcap)
                g[(uint64_t)lo * gstride] = old;
            else
                overflow = true;
            lo++;
        }
    }
    // push after reserve: no spill check
    __device__ __forceinline__ void pushReserved(int e) {
        lds[(sp & (kShortStack - 1)) * ldsStride] = e;
        sp++;
    }
    __device__ __forceinline__ int pop() {
        if (sp == 0) return -1;
        sp--;
        if (sp >= lo) return lds[(sp & (kShortStack - 1)) * ldsStride];
        lo = sp;
        return sp < gcap ? g[(uint64_t)sp * gstride] : -1;
    }
};

// Quantized wide node as stored (eleven::QNode<W>, bvh.hpp): 16-byte words
// per node, W = 4: 4 (header, q0-3, q4-5 + child0-1, child2-3 + pad);
// W = 8: 6 (header, q0-11 in three words, child0-7 in two).
template <int W>
struct QNodeWords {
    static constexpr int kWords = W == 4 ? 4 : 6;
};

struct TraceResult {
    int32_t tri;   // original index, -1 if none
    float t, u, v;
};

struct TraceCounters {
    uint32_t nodes, tris;
};

// A ray prepared for slab tests: origin, direction and safe reciprocal.
struct TRay {
    float ox, oy, oz, dx, dy, dz, ix, iy, iz;
};

__device__ __forceinline__ TRay makeTRay(float ox, float oy, float oz, float dx, float dy, float dz) {
    const float tiny = 1e-30f;
    TRay r;
    r.ox = ox;
    r.oy = oy;
    r.oz = oz;
    r.dx = dx;
    r.dy = dy;
    r.dz = dz;
    r.ix = 1.0f / (fabsf(dx) > tiny ? dx : copysignf(tiny, dx));
    r.iy = 1.0f / (fabsf(dy) > tiny ? dy : copysignf(tiny, dy));
    r.iz = 1.0f / (fabsf(dz) > tiny ? dz : copysignf(tiny, dz));
    return r;
}

// Visits inner node `cur`: slab-tests its W children against [0, tcull],
// pushes the hit children far-to-near and returns the nearest (or the next
// stack entry if none is hit; -1 when the stack is empty).  With UNSORTED
// (any-hit queries, whose result does not depend on the visiting order) the
// hit children are taken in slot order instead.  Child encoding:
// >= 0 inner node, <= -2 leaf, -1 empty slot (its box is inverted, so it is
// never hit).
//
// Planes decode exactly (p + q * step is an FP32 value by construction), so
// the computed slab distances (b - o) * (1/d) carry only relative rounding
// error (< 2^-21 with the reciprocal's); widening tfar by 2^-20 of its
// magnitude makes the test conservative: no box the ray touches is culled.
// visitWords tests a node whose words are already loaded (the root, kept in
// scalar registers by the wavefront traversal); visitNode loads them.
template <int W, bool UNSORTED = false>
__device__ __forceinline__ int visitWords(const uint4* np, const TRay& ray, float tcull, TravStack& st) {
    constexpr int D = W / 4;   // dwords of bytes per (axis, side)
    const uint4 h = np[0];
    uint32_t q[6 * D];
    int32_t child[W];
    if constexpr (W == 4) {
        const uint4 a = np[1], b = np[2], c = np[3];
        q[0] = a.x, q[1] = a.y, q[2] = a.z, q[3] = a.w, q[4] = b.x, q[5] = b.y;
        child[0] = (int32_t)b.z, child[1] = (int32_t)b.w, child[2] = (int32_t)c.x, child[3] = (int32_t)c.y;
    } else {
        const uint4 a = np[1], b = np[2], c = np[3], d = np[4], e = np[5];
        q[0] = a.x, q[1] = a.y, q[2] = a.z, q[3] = a.w, q[4] = b.x, q[5] = b.y;
        q[6] = b.z, q[7] = b.w, q[8] = c.x, q[9] = c.y, q[10] = c.z, q[11] = c.w;
        child[0] = (int32_t)d.x, child[1] = (int32_t)d.y, child[2] = (int32_t)d.z, child[3] = (int32_t)d.w;
        child[4] = (int32_t)e.x, child[5] = (int32_t)e.y, child[6] = (int32_t)e.z, child[7] = (int32_t)e.w;
    }
    const float px = __uint_as_float(h.x), py = __uint_as_float(h.y), pz = __uint_as_float(h.z);
    const float sx = __uint_as_float((h.w & 0xffu) << 23);
    const float sy = __uint_as_float(((h.w >> 8) & 0xffu) << 23);
    const float sz = __uint_as_float(((h.w >> 16) & 0xffu) << 23);
    // near / far plane bytes per axis by the direction's sign (q index: axis*2*D + side*D + dword)
    const bool nx = ray.ix < 0, ny = ray.iy < 0, nz = ray.iz < 0;
    uint32_t nearX[D], farX[D], nearY[D], farY[D], nearZ[D], farZ[D];
#pragma unroll
    for (int d = 0; d < D; d++) {
        nearX[d] = nx ? q[1 * D + d] : q[0 * D + d];
        farX[d] = nx ? q[0 * D + d] : q[1 * D + d];
        nearY[d] = ny ? q[3 * D + d] : q[2 * D + d];
        farY[d] = ny ? q[2 * D + d] : q[3 * D + d];
        nearZ[d] = nz ? q[5 * D + d] : q[4 * D + d];
        farZ[d] = nz ? q[4 * D + d] : q[5 * D + d];
    }
    uint32_t keys[W];
#if EL_SLAB_AB
    // Plane b = p + q s, so (b - o) / d = A + q B per axis with
    // A = (p - o) inv and B = s inv (exact: s is a power of two), and
    // fma(q, B, A) rounds once; A carries two roundings, so a distance is off
    // by at most 2^-23 |A| + 2^-24 |t|.  Near planes use A - E and far planes
    // A + E with E = 2^-19 |A| of their own axis (an axis whose direction is
    // near zero has a huge A, but also huge distances of its own), and the
    // 2^-20 |tf| widening below covers the |t| term: the test stays
    // conservative.  An empty slot (child -1) has an inverted box, 255 grid
    // steps from near to far plane; where 255 steps are below the rounding
    // of A (a node smaller than ~4e-6 of its distance from the origin on
    // every axis) the box passes the widened test, so the slot's code is
    // checked as well: entering it would end the traversal (-1 is the empty
    // stack's code).
    const float Ax = (px - ray.ox) * ray.ix, Ay = (py - ray.oy) * ray.iy, Az = (pz - ray.oz) * ray.iz;
    const float Bx = sx * ray.ix, By = sy * ray.iy, Bz = sz * ray.iz;
    const float Ex = fabsf(Ax) * (1.0f / 524288.0f), Ey = fabsf(Ay) * (1.0f / 524288.0f),
                Ez = fabsf(Az) * (1.0f / 524288.0f);
    const float Anx = Ax - Ex, Afx = Ax + Ex, Any = Ay - Ey, Afy = Ay + Ey, Anz = Az - Ez, Afz = Az + Ez;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const int d = i >> 2, sh = 8 * (i & 3);
        const float tx0 = fmaf((float)((nearX[d] >> sh) & 0xffu), Bx, Anx);
        const float tx1 = fmaf((float)((farX[d] >> sh) & 0xffu), Bx, Afx);
        const float ty0 = fmaf((float)((nearY[d] >> sh) & 0xffu), By, Any);
        const float ty1 = fmaf((float)((farY[d] >> sh) & 0xffu), By, Afy);
        const float tz0 = fmaf((float)((nearZ[d] >> sh) & 0xffu), Bz, Anz);
        const float tz1 = fmaf((float)((farZ[d] >> sh) & 0xffu), Bz, Afz);
        const float tn = fmaxf(fmaxf(tx0, ty0), fmaxf(tz0, 0.0f));
        float tf = fminf(fminf(tx1, ty1), fminf(tz1, tcull));
        tf = fmaf(fabsf(tf), 1.0f / 1048576.0f, tf);
        keys[i] = (tn <= tf && child[i] != -1) ? ((__float_as_uint(tn) & ~7u) | (uint32_t)i) : 0xFFFFFFFFu;
    }
#else
#pragma unroll
    for (int i = 0; i < W; i++) {
        const int d = i >> 2, sh = 8 * (i & 3);
        const float bx0 = fmaf((float)((nearX[d] >> sh) & 0xffu), sx, px);
        const float bx1 = fmaf((float)((farX[d] >> sh) & 0xffu), sx, px);
        const float by0 = fmaf((float)((nearY[d] >> sh) & 0xffu), sy, py);
        const float by1 = fmaf((float)((farY[d] >> sh) & 0xffu), sy, py);
        const float bz0 = fmaf((float)((nearZ[d] >> sh) & 0xffu), sz, pz);
        const float bz1 = fmaf((float)((farZ[d] >> sh) & 0xffu), sz, pz);
        const float tn = fmaxf(fmaxf((bx0 - ray.ox) * ray.ix, (by0 - ray.oy) * ray.iy),
                               fmaxf((bz0 - ray.oz) * ray.iz, 0.0f));
        float tf = fminf(fminf((bx1 - ray.ox) * ray.ix, (by1 - ray.oy) * ray.iy), fminf((bz1 - ray.oz) * ray.iz, tcull));
        tf = fmaf(fabsf(tf), 1.0f / 1048576.0f, tf);
        keys[i] = (tn <= tf && child[i] != -1) ? ((__float_as_uint(tn) & ~7u) | (uint32_t)i) : 0xFFFFFFFFu;
    }
#endif
    if constexpr (UNSORTED) {
        int next = -2;   // the lowest hit slot is visited next, the others are pushed
#if EL_PUSH_BULK
        int nhit = 0;
#pragma unroll
        for (int i = 0; i < W; i++) nhit += keys[i] != 0xFFFFFFFFu;
        if (nhit > 1) st.reserve(nhit - 1);
#endif
#pragma unroll
        for (int i = W - 1; i >= 0; i--) {
            if (keys[i] == 0xFFFFFFFFu) continue;
#if EL_PUSH_BULK
            if (next != -2) st.pushReserved(child[next]);
#else
            if (next != -2) st.push(child[next]);
#endif
            next = i;
        }
        return next == -2 ? st.pop() : child[next];
    }
    // ascending sort (tn >= 0, so float bits order as unsigned)
#define EL_CSWAP(a, b)                        \
    {                                         \
        uint32_t lo_ = min(keys[a], keys[b]); \
        uint32_t hi_ = max(keys[a], keys[b]); \
        keys[a] = lo_;                        \
        keys[b] = hi_;                        \
    }
    if (W == 4) {
        EL_CSWAP(0, 1) EL_CSWAP(2, 3) EL_CSWAP(0, 2) EL_CSWAP(1, 3) EL_CSWAP(1, 2)
    } else {
#pragma unroll
        for (int i = 1; i < W; i++)
#pragma unroll
            for (int j = i; j > 0; j--) EL_CSWAP(j - 1, j)
    }
#undef EL_CSWAP
    int nh = 0;
#pragma unroll
    for (int i = 0; i < W; i++) nh += keys[i] != 0xFFFFFFFFu;
    if (nh == 0) return st.pop();
#if EL_PUSH_BULK
    if (nh > 1) st.reserve(nh - 1);
#pragma unroll
    for (int i = W - 1; i >= 1; i--)
        if (i < nh) st.pushReserved(child[keys[i] & 7u]);
#else
#pragma unroll
    for (int i = W - 1; i >= 1; i--)
        if (i < nh) st.push(child[keys[i] & 7u]);
#endif
    return child[keys[0] & 7u];
}

template <int W, bool UNSORTED = false>
__device__ __forceinline__ int visitNode(const uint4* __restrict__ nodes, int cur, const TRay& ray, float tcull,
                                         TravStack& st) {
    constexpr int K = QNodeWords<W>::kWords;
    const uint4* np = nodes + (uint64_t)cur * K;
    uint4 w[K];
#pragma unroll
    for (int k = 0; k < K; k++) w[k] = np[k];
    return visitWords<W, UNSORTED>(w, ray, tcull, st);
}

// Tests one triangle record and updates the closest hit: smaller t wins,
// equal t goes to the lower original index.  For ANY the first hit below
// r.t is taken and true is returned (the query ends).
template <bool ANY>
__device__ __forceinline__ bool hitTriangle(const float4& A, const float4& B, const float4& C, const TRay& ray,
                                            TraceResult& r) {
    float t, u, v;
    if (mtHit(A, B, C, ray.ox, ray.oy, ray.oz, ray.dx, ray.dy, ray.dz, t, u, v)) {
        const int id = __float_as_int(C.y);
        if (ANY) {
            if (t < r.t) {
                r.tri = id;
                r.t = t;
                r.u = u;
                r.v = v;
                return true;
            }
        } else if (t < r.t || (t == r.t && (unsigned)id < (unsigned)r.tri)) {
            r.tri = id;
            r.t = t;
            r.u = u;
            r.v = v;
        }
    }
    return false;
}

// Tests the triangles of leaf `code` (<= -2), as hitTriangle.  The leaf's
// triangles are contiguous records; the three words of the next
// triangle are loaded before the current one is tested, and all three
// words of a triangle together (the vertex word is otherwise fetched only
// after the determinant test), so a leaf costs about one memory round trip
// instead of two per triangle.
template <bool ANY, bool COUNT>
__device__ __forceinline__ bool testLeaf(const float4* __restrict__ tris, int code, const TRay& ray, TraceResult& r,
                                         TraceCounters& cnt) {
    int k = -code - 2;
    int first = k >> 4, count = (k & 15) + 1;
    const float4* p = tris + (uint64_t)first * 3;
#if EL_LEAF_PIPE
    float4 nA = p[0], nB = p[1], nC = p[2];
#endif
    for (int i = 0; i < count; i++) {
#if EL_LEAF_PIPE
        const float4 A = nA, B = nB, C = nC;
        if (i + 1 < count) {
            nA = p[3 * (i + 1)];
            nB = p[3 * (i + 1) + 1];
            nC = p[3 * (i + 1) + 2];
        }
#else
        float4 A = p[3 * i], B = p[3 * i + 1], C = p[3 * i + 2];
#endif
        if (COUNT) cnt.tris++;
        if (hitTriangle<ANY>(A, B, C, ray, r)) return true;
    }
    return false;
}

// Culling bound for child boxes: closest-hit keeps boxes slightly beyond the
// current hit so an equal-t hit with a lower index is still found.
template <bool ANY>
__device__ __forceinline__ float cullT(float t) {
    return ANY ? t : t * (1.0f + 1.0f / 1048576.0f);
}

// One ray, one thread: depth-first traversal with nearest-child-first order.
template <int W, bool ANY, bool COUNT>
__device__ __forceinline__ TraceResult traceRay(const DevScene& s, float ox, float oy, float oz, float dx, float dy,
                                                float dz, float tmax, TravStack& st, TraceCounters& cnt) {
    const uint4* __restrict__ nodes = (const uint4*)s.nodes;
    const TRay ray = makeTRay(ox, oy, oz, dx, dy, dz);
    TraceResult r;
    r.tri = -1;
    r.t = tmax;
    r.u = r.v = 0.f;
    st.sp = 0;
    st.lo = 0;
    int cur = 0;
    while (cur != -1) {
        if (cur >= 0) {
            if (COUNT) cnt.nodes++;
            cur = visitNode<W, ANY>(nodes, cur, ray, cullT<ANY>(r.t), st);
        } else {
            bool done = testLeaf<ANY, COUNT>(s.tris, cur, ray, r, cnt);
            cur = done ? -1 : st.pop();
        }
    }
    return r;
}

// Persistent traversal of the items [begin, end) owned by one wave, with
// in-wave refill: a lane whose ray is finished takes the wave's next item, so
// the wave keeps its lanes busy instead of waiting for its slowest ray (the
// refill happens once at least `refill` lanes are idle, or all are).  Inner
// nodes and leaves are separated (Aila & Laine, "Understanding the Efficiency
// of Ray Traversal on GPUs", HPG 2009, speculative while-while): a lane that
// reaches a leaf postpones it and keeps descending until every live lane of
// the wave holds a leaf, then the wave tests leaves together.  With one BVH
// per wave a lane may postpone a second leaf and go on descending instead
// of idling while the other lanes look for their first (measured: a third
// does not pay, nor does a second with per-ray BVHs, whose lanes hold more
// registers).  The
// closest hit is order-independent (t, then lower index), so results are
// identical to traceRay.
//
//   load(i, ox, oy, oz, dx, dy, dz, tmax, bvh) -> false if item i has no ray;
//                                   with PER_RAY_BVH it also sets the BVH
//                                   (nodes, triangles) the ray traverses
//   store(i, result, traced)        once per item
//
// Every lane of the wave must call with the same begin/end.
struct BvhPtr {
    const uint4* nodes;
    const float4* tris;
};

// The items a wave traces: a fixed range [begin, end), or (ctr set) chunks
// of `chunk` items handed out at run time.  Chunk c belongs to share
// c % parts; share x is counted by ctr[x * stride].  A wave draws from its
// home share first, then from the others, so waves that meet cheap rays take
// more chunks and a launch ends when the work does, not when its slowest
// fixed range does.  One scene uses 8 shares (one per XCD: the waves of an
// XCD draw neighbouring chunks, which keeps their rays coherent in its L2),
// a batch 64 (measured faster there).  Every counter starts the launch at 0.  A share is skipped without an atomic once a coherent read shows it
// drawn out (a read never runs ahead of the counter, so no chunk is lost).
// A queue with no more chunks than the grid has waves (the late bounces, small
// frames) is handed out without counters: wave w takes chunk w -- the chunk
// the dynamic scheme's share order would also give its XCD -- and waves past
// the last chunk return at once.  Thousands of waves contending for eight
// counters cost such a launch ~100 us whatever its size.
constexpr uint32_t kWorkShares = 64;   // shares of a wide WorkSource (one per lane of a wave)
struct WorkSource {
    uint32_t begin, end;
    uint32_t* ctr;
    uint32_t n, chunk, stride, parts, home;
    uint32_t chunks;   // ceil(n / chunk)
    bool fixed;        // chunks <= waves of the grid: wave w takes chunk w, no counter atomics
};

template <int W, bool ANY, bool COUNT, bool PER_RAY_BVH = false, class Load, class Store>
__device__ __forceinline__ void traceWaveRange(const DevScene& s, TravStack& st, const WorkSource& ws,
                                               uint32_t refill, Load&& load, Store&& store, TraceCounters& cnt,
                                               uint32_t& traced) {
    const uint32_t waveId = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (ws.ctr && ws.fixed && waveId >= ws.chunks) return;   // no chunk for this wave
    bool fixedDrawn = false;
    BvhPtr bvh{(const uint4*)s.nodes, s.tris};
    // One BVH for the whole wave: its root node (visited first by every ray)
    // is read once into scalar registers, and a new ray is tested against it
    // as it is loaded, so no ray fetches the root.
    constexpr int K = QNodeWords<W>::kWords;
    uint4 root[K];
    if constexpr (!PER_RAY_BVH) {
        const uint4* rn = (const uint4*)s.nodes;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const uint4 v = rn[k];
            root[k] = make_uint4(__builtin_amdgcn_readfirstlane(v.x), __builtin_amdgcn_readfirstlane(v.y),
                                 __builtin_amdgcn_readfirstlane(v.z), __builtin_amdgcn_readfirstlane(v.w));
        }
    }
    const uint32_t lane = __lane_id();
    const unsigned long long ltMask = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t cursor = ws.begin, end = ws.end;   // wave-uniform
    uint32_t tried = 0;                         // narrow source: shares found drawn out
    bool more = true;                           // wave-uniform: items may remain
    bool live = false;
    uint32_t item = 0;
    TRay ray{};
    TraceResult r{};
    int cur = -1, leaf = -1, leaf2 = -1;
    while (true) {
        // ---- refill idle lanes from the wave's range or the next chunk
        const unsigned long long idle = __ballot(!live);
        const uint32_t nIdle = (uint32_t)__popcll(idle);
        const bool want = nIdle >= refill || nIdle == 64;
        if (want && cursor >= end) {
            more = false;
            if (ws.ctr && ws.fixed) {   // the wave's one chunk
                if (!fixedDrawn) {
                    fixedDrawn = true;
                    cursor = waveId * ws.chunk;
                    end = cursor + ws.chunk < ws.n ? cursor + ws.chunk : ws.n;
                    more = cursor < end;
                }
            }
            // Wide sources (batches, parts == 64): lane l looks at share
            // (home + l) % 64 at once (one load per lane instead of 64
            // dependent loads; the whole wave runs here), and the wave draws
            // from the first share with chunks left.  Share x holds the
            // chunks x, x + 64, ...: (chunks - x + 63) / 64 of them.
            while (PER_RAY_BVH && ws.ctr && !ws.fixed) {
                const uint32_t x = (ws.home + lane) & (kWorkShares - 1);
                const uint32_t seen =
                    __hip_atomic_load(ws.ctr + x * ws.stride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const bool left = x < ws.chunks && seen < ((ws.chunks - x + kWorkShares - 1) >> 6);
                const unsigned long long have = __ballot(left);
                if (!have) break;
                const uint32_t pick = (ws.home + (uint32_t)(__ffsll((long long)have) - 1)) & (kWorkShares - 1);
                uint32_t k = 0;
                if (lane == 0) k = atomicAdd(ws.ctr + pick * ws.stride, 1u);
                k = __builtin_amdgcn_readfirstlane(k);
                const uint64_t first = ((uint64_t)pick + (uint64_t)kWorkShares * k) * ws.chunk;
                if (first < ws.n) {
                    cursor = (uint32_t)first;
                    end = (uint32_t)(first + ws.chunk < ws.n ? first + ws.chunk : ws.n);
                    more = true;
                    break;
                }
            }
            // Narrow sources (one scene, parts == 8): one share at a time,
            // from the wave's home share on; a share found drawn out is not
            // looked at again.
            while (!PER_RAY_BVH && ws.ctr && !ws.fixed && tried < ws.parts) {
                const uint32_t x = (ws.home + tried) % ws.parts;
                uint32_t* const c = ws.ctr + x * ws.stride;
                const uint32_t seen = __hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (((uint64_t)x + (uint64_t)ws.parts * seen) * ws.chunk >= ws.n) {
                    tried++;
                    continue;
                }
                uint32_t k = 0;
                if (lane == (uint32_t)(__ffsll((long long)__ballot(true)) - 1)) k = atomicAdd(c, 1u);
                k = __builtin_amdgcn_readfirstlane(k);
                const uint64_t first = ((uint64_t)x + (uint64_t)ws.parts * k) * ws.chunk;
                if (first < ws.n) {
                    cursor = (uint32_t)first;
                    end = (uint32_t)(first + ws.chunk < ws.n ? first + ws.chunk : ws.n);
                    more = true;
                    break;
                }
                tried++;
            }
        }
        if (want && more) {
            if (!live) {
                uint32_t i = cursor + (uint32_t)__popcll(idle & ltMask);
                if (i < end) {
                    float ox, oy, oz, dx, dy, dz, tmax;
                    BvhPtr rb = bvh;
                    if (load(i, ox, oy, oz, dx, dy, dz, tmax, rb)) {
                        if (PER_RAY_BVH) bvh = rb;
                        ray = makeTRay(ox, oy, oz, dx, dy, dz);
                        r.tri = -1;
                        r.t = tmax;
                        r.u = r.v = 0.f;
                        st.sp = 0;
                        st.lo = 0;
                        leaf = leaf2 = -1;
                        cur = 0;
                        if constexpr (!PER_RAY_BVH) {
                            cur = visitWords<W, ANY>(root, ray, cullT<ANY>(r.t), st);
                            if (cur <= -2) {
                                leaf = cur;
                                cur = st.pop();
                            }
                        }
                        item = i;
                        live = true;
                    } else {
                        TraceResult dead;
                        dead.tri = -2;
                        dead.t = dead.u = dead.v = 0.f;
                        store(i, dead, false);
                    }
                }
            }
            cursor += nIdle;
        } else if (nIdle == 64 && !more) {
            break;   // no items left and every lane finished
        }
        // ---- inner nodes: descend until every live lane holds a leaf
        while (true) {
            const bool act = live && cur >= 0;
            if (__all(!act || leaf != -1)) break;   // also when no lane can descend
            // closest hit, one BVH per wave: with enough lanes finished and
            // items left, test the leaves held so far and refill instead of
            // idling until the descent ends (any-hit queries and per-ray
            // BVHs measured slower with it)
            if (!ANY && !PER_RAY_BVH && more &&
                (uint32_t)__popcll(__ballot(!live || (cur == -1 && leaf == -1 && leaf2 == -1))) >= refill)
                break;
            if (act) {
                if (COUNT) cnt.nodes++;
                cur = visitNode<W, ANY>(PER_RAY_BVH ? bvh.nodes : (const uint4*)s.nodes, cur, ray, cullT<ANY>(r.t),
                                        st);
                if (cur <= -2 && leaf == -1) {
                    leaf = cur;
                    cur = st.pop();
                } else if (!PER_RAY_BVH && cur <= -2 && leaf2 == -1) {   // a second leaf: postponed as well
                    leaf2 = cur;
                    cur = st.pop();
                }
            }
        }
        // ---- leaves: the postponed one, then any leaves that follow on the stack
        while (true) {
            const bool has = live && (leaf != -1 || leaf2 != -1 || cur <= -2);
            if (!__any(has)) break;
            if (has) {
                int code;
                if (leaf != -1) {
                    code = leaf;
                    leaf = -1;
                } else if (leaf2 != -1) {
                    code = leaf2;
                    leaf2 = -1;
                } else {
                    code = cur;
                    cur = st.pop();
                }
                if (testLeaf<ANY, COUNT>(PER_RAY_BVH ? bvh.tris : s.tris, code, ray, r, cnt)) {
                    cur = -1;   // any-hit: done
                    leaf = leaf2 = -1;
                }
            }
        }
        // ---- finished rays
        if (live && cur == -1 && leaf == -1 && leaf2 == -1) {
            store(item, r, true);
            traced++;
            live = false;
        }
    }
}

}  // namespace dev
}  // namespace eleven

// CPU oracle: restatement of the reference estimator (test infrastructure only;
// see oracle.h).  Written from the reference's text (src/tfg-pathtracer/),
// each piece citing the lines it follows.  Compiled with -ffp-contract=off so
// FP32 expressions round exactly as written.
//
// R = float reproduces the reference's types, including its double-literal
// promotions (e.g. `0.5 * (1.0 - metallic)` is evaluated in double and
// stored to float, as in the reference).  R = double is the FP64 mode.
#include "oracle.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

const float PI_F = 3.14159265358979323846f;   // Math.hpp:7
template <class R>
inline R piR() {
    return (R)PI_F;
}
template <>
inline double piR<double>() {
    return 3.14159265358979323846;
}

// ----------------------------------------------------------- Vector3 -----
template <class R>
struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R a) : x(a), y(a), z(a) {}
    R length() const { return std::sqrt(x * x + y * y + z * z); }   // Vector.hpp:161-163
    V3 normalized() const {                                          // Vector.hpp:242-245
        if (length() == 0) return *this;
        R l = length();
        return V3(x / l, y / l, z / l);
    }
};
template <class R> V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> V3<R> operator*(V3<R> a, V3<R> b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> V3<R> operator*(V3<R> a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> V3<R> operator*(R s, V3<R> a) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> V3<R> operator/(V3<R> a, R s) { return V3<R>(a.x / s, a.y / s, a.z / s); }
template <class R> V3<R> operator/(V3<R> a, V3<R> b) { return V3<R>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> V3<R> cross(V3<R> a, V3<R> b) {   // Vector.hpp:229-231
    return V3<R>((a.y * b.z - a.z * b.y), -(a.x * b.z - a.z * b.x), (a.x * b.y - a.y * b.x));
}
template <class R> V3<R> neg(V3<R> a) { return a * (R)-1; }
template <class R> V3<R> reflect(V3<R> v1, V3<R> v2) { return v1 - (R)(2 * dot(v1, v2)) * v2; }   // Vector.hpp:218-220

template <class R> R clampR(R a, R b, R c) { return a < b ? b : a > c ? c : a; }   // Math.hpp:30-32
template <class R> R lerpR(R a, R b, R c) { return a + c * (b - a); }              // Math.hpp:43-50 (FAST_LERP)
template <class R> V3<R> lerpV(V3<R> a, V3<R> b, R c) {
    return V3<R>(lerpR(a.x, b.x, c), lerpR(a.y, b.y, c), lerpR(a.z, b.z, c));
}
template <class R> R minR(R a, R b) { return a < b ? a : b; }   // Math.hpp:60-66
template <class R> R maxR(R a, R b) { return a > b ? a : b; }

// ------------------------------------------------------------ Tri::hit -----
// Tri.hpp:38-68 (the geometric part).  FP32, 1/det in double.
inline bool triHit(const float* v, const float o[3], const float d[3], float& t, float& u, float& vv) {
    float EPSILON = 0.0000001;
    float e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
    float e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
    // pvec = cross(direction, edge2)
    float px = d[1] * e2z - d[2] * e2y;
    float py = -(d[0] * e2z - d[2] * e2x);
    float pz = d[0] * e2y - d[1] * e2x;
    float det = e1x * px + e1y * py + e1z * pz;
    float inv_det = 1.0 / det;
    if (det > -EPSILON && det < EPSILON) return false;
    float tx = o[0] - v[0], ty = o[1] - v[1], tz = o[2] - v[2];
    u = (tx * px + ty * py + tz * pz) * inv_det;
    if (u < 0.0 || u > 1.0) return false;
    // qvec = cross(tvec, edge1)
    float qx = ty * e1z - tz * e1y;
    float qy = -(tx * e1z - tz * e1x);
    float qz = tx * e1y - ty * e1x;
    vv = (d[0] * qx + d[1] * qy + d[2] * qz) * inv_det;
    if (vv < 0.0 || (u + vv) > 1.0) return false;
    t = (e2x * qx + e2y * qy + e2z * qz) * inv_det;
    if (t < 0) return false;
    return true;
}

struct Hit {
    int tri = -1;
    float t = std::numeric_limits<float>::infinity(), u = 0, v = 0;
    void offer(int id, float tt, float uu, float vv) {
        if (tri < 0 || tt < t || (tt == t && id < tri)) {
            tri = id;
            t = tt;
            u = uu;
            v = vv;
        }
    }
};

template <class F>
void parallelFor(int n, int threads, F f) {
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::max(1, std::min(threads, n));
    if (threads == 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> ts;
    for (int k = 0; k < threads; k++)
        ts.emplace_back([&] {
            int i;
            while ((i = next.fetch_add(64)) < n)
                for (int j = i; j < std::min(n, i + 64); j++) f(j);
        });
    for (auto& t : ts) t.join();
}

// ------------------------------------------------------ reference BVH -----
struct RNode {   // BVH.hpp:29-48 (44 bytes of payload)
    float b1[3] = {0, 0, 0}, b2[3] = {0, 0, 0};
    int from = 0, to = 0, idx = 0, depth = 0;
    bool valid = false;
};

struct RefBVH {
    int D = 18, bins = 14;
    std::vector<RNode> nodes;
    std::vector<int> triIndices;
    std::vector<float> verts;
    int n = 0, nodeIdx = 0, triIdx = 0;
};

inline float mapf(float a, float b, float c, float d, float e) { return d + ((a - b) / (c - b)) * (e - d); }   // Math.hpp:22-24
// float -> int as x86 cvttss2si (out-of-range and NaN give INT_MIN), which
// is what the reference's host build does for map() of a zero extent.
inline int toIntX86(float f) {
    if (!(f > -2147483904.0f && f < 2147483648.0f)) return INT_MIN;
    return (int)f;
}

struct Box {
    float b1[3] = {0, 0, 0}, b2[3] = {0, 0, 0};
};
inline float boundsArea(const Box& b) {   // BVH.hpp:519-527
    float x = b.b2[0] - b.b1[0], y = b.b2[1] - b.b1[1], z = b.b2[2] - b.b1[2];
    return 2 * (x * y + x * z + y * z);
}
inline Box boundsUnion(Box a, Box b) {   // BVH.hpp:492-517
    Box r = a;   // reference leaves b5/b6 as passed when neither branch applies
    if (boundsArea(a) <= 0 || boundsArea(b) <= 0) {
        if (boundsArea(a) <= 0) r = b;
        if (boundsArea(b) <= 0) r = a;
    } else {
        for (int k = 0; k < 3; k++) {
            r.b1[k] = minR(a.b1[k], minR(a.b2[k], minR(b.b1[k], b.b2[k])));
            r.b2[k] = maxR(a.b1[k], maxR(a.b2[k], maxR(b.b1[k], b.b2[k])));
        }
    }
    return r;
}
inline Box triBounds(const float* v) {   // BVH.hpp:529-538
    Box b;
    for (int k = 0; k < 3; k++) {
        b.b1[k] = minR(v[k], minR(v[3 + k], v[6 + k]));
        b.b2[k] = maxR(v[k], maxR(v[3 + k], v[6 + k]));
    }
    return b;
}
// BVH.hpp:540-578: an empty list leaves the outputs untouched.
inline void listBounds(const RefBVH& h, const std::vector<int>& l, Box& out) {
    if (l.empty()) return;
    const float* v0 = &h.verts[(size_t)l[0] * 9];
    for (int k = 0; k < 3; k++) out.b1[k] = out.b2[k] = v0[k];
    for (int i : l) {
        const float* v = &h.verts[(size_t)i * 9];
        for (int c = 0; c < 3; c++)
            for (int k = 0; k < 3; k++) {
                out.b1[k] = minR(v[3 * c + k], out.b1[k]);
                out.b2[k] = maxR(v[3 * c + k], out.b2[k]);
            }
    }
}
inline float centroidAxis(const float* v, int axis) {   // Tri.hpp:27-32
    return (v[axis] + v[3 + axis] + v[6 + axis]) / (float)3.0;
}

// BVH.hpp:373-460
void divideSAH(const RefBVH& h, const std::vector<int>& tris, std::vector<int>& L, std::vector<int>& Rr) {
    if (tris.empty()) return;
    const int B = h.bins;
    int bestBin = 0, bestAxis = 0;
    float bestH = std::numeric_limits<float>::max();
    Box tot;
    listBounds(h, tris, tot);
    std::vector<Box> bb(B);
    std::vector<int> cnt(B);
    for (int axis = 0; axis < 3; axis++) {
        for (int i = 0; i < B; i++) {
            cnt[i] = 0;
            bb[i] = Box();
        }
        for (int t : tris) {
            const float* v = &h.verts[(size_t)t * 9];
            int bin = 0;
            if (tot.b1[axis] != tot.b2[axis]) bin = toIntX86(mapf(centroidAxis(v, axis), tot.b1[axis], tot.b2[axis], 0, B - 1));
            cnt[bin]++;
            bb[bin] = boundsUnion(bb[bin], triBounds(v));
        }
        for (int i = 0; i < B; i++) {
            int c1 = 0, c2 = 0;
            Box l, r;
            for (int j = 0; j < i; j++) {
                c1 += cnt[j];
                l = boundsUnion(l, bb[j]);
            }
            for (int k = i; k < B; k++) {
                c2 += cnt[k];
                r = boundsUnion(r, bb[k]);
            }
            float heur = boundsArea(l) * (float)c1 + boundsArea(r) * (float)c2;
            if (heur < bestH) {
                bestH = heur;
                bestBin = i;
                bestAxis = axis;
            }
        }
    }
    for (int t : tris) {
        const float* v = &h.verts[(size_t)t * 9];
        int bin = toIntX86(mapf(centroidAxis(v, bestAxis), tot.b1[bestAxis], tot.b2[bestAxis], 0, B - 1));
        (bin < bestBin ? L : Rr).push_back(t);
    }
}

// BVH.hpp:291-330
void buildAux(RefBVH& h, int depth, const std::vector<int>& tris) {
    Box b;
    listBounds(h, tris, b);
    RNode nd;
    std::memcpy(nd.b1, b.b1, sizeof nd.b1);
    std::memcpy(nd.b2, b.b2, sizeof nd.b2);
    nd.idx = h.nodeIdx;
    nd.depth = depth;
    nd.valid = true;
    if (depth == h.D) {
        nd.from = h.triIdx;
        nd.to = h.triIdx + (int)tris.size();
        h.nodes[h.nodeIdx++] = nd;
        for (int t : tris) h.triIndices[h.triIdx++] = t;
    } else {
        h.nodes[h.nodeIdx++] = nd;
        std::vector<int> L, Rr;
        divideSAH(h, tris, L, Rr);
        buildAux(h, depth + 1, L);
        buildAux(h, depth + 1, Rr);
    }
}

// BVH.hpp:69-101
inline bool slab(const float o[3], const float d[3], const float* b1, const float* b2) {
    float fx = 1.0f / d[0], fy = 1.0f / d[1], fz = 1.0f / d[2];
    float t1 = (b1[0] - o[0]) * fx, t2 = (b2[0] - o[0]) * fx;
    float t3 = (b1[1] - o[1]) * fy, t4 = (b2[1] - o[1]) * fy;
    float t5 = (b1[2] - o[2]) * fz, t6 = (b2[2] - o[2]) * fz;
    float tmin = maxR(maxR(minR(t1, t2), minR(t3, t4)), minR(t5, t6));
    float tmax = minR(minR(maxR(t1, t2), maxR(t3, t4)), maxR(t5, t6));
    if (tmax < 0) return false;
    if (tmin > tmax) return false;
    return true;
}

inline void intersectNode(const RefBVH& h, const RNode& nd, const float o[3], const float d[3], Hit& best) {
    for (int i = nd.from; i < nd.to; i++) {   // BVH.hpp:159-175
        int ti = h.triIndices[i];
        float t, u, v;
        if (triHit(&h.verts[(size_t)ti * 9], o, d, t, u, v)) best.offer(ti, t, u, v);
    }
}

// BVH.hpp:120-157 and 177-185
Hit refTraverse(const RefBVH& h, const float o[3], const float d[3]) {
    Hit best;
    RNode stack[64];
    int sp = 0;
    stack[sp++].valid = false;
    RNode node = h.nodes[0];
    const RNode invalid;
    do {
        const RNode& l = node.depth == h.D ? invalid : h.nodes[node.idx + 1];
        const RNode& r = node.depth == h.D ? invalid : h.nodes[node.idx + (2 << (h.D - node.depth - 1))];
        bool lO = slab(o, d, l.b1, l.b2);
        bool rO = slab(o, d, r.b1, r.b2);
        if (node.depth == h.D - 1 && rO) intersectNode(h, r, o, d, best);
        if (node.depth == h.D - 1 && lO) intersectNode(h, l, o, d, best);
        bool travL = lO && node.depth != h.D - 1;
        bool travR = rO && node.depth != h.D - 1;
        if (!travL && !travR) {
            node = stack[--sp];
        } else {
            RNode nx = travL ? l : r;
            if (travL && travR) stack[sp++] = r;
            node = nx;
        }
    } while (node.valid);
    return best;
}

// --------------------------------------------------------------- RNG -----
inline void philox(const uint32_t c_[4], const uint32_t k_[2], uint32_t out[4]) {
    uint32_t c[4] = {c_[0], c_[1], c_[2], c_[3]};
    uint32_t k0 = k_[0], k1 = k_[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    std::memcpy(out, c, sizeof c);
}
inline float u01(uint32_t u) { return (float)(u >> 8) * (1.0f / 16777216.0f); }
// Random block b of path (pixel, sample): Philox counter (pixel, sample, b, 0).
// Block 0: camera r1..r4; block 1 word 0: camera r5; block 2+i: bounce i
// (DisneySample r1, r2, r3, shade r1).  Shade's r2/r3 are never read on the
// reference's HDRI-IS path (kernel.cu:234-257).
inline void pathBlock(uint32_t k0, uint32_t k1, uint32_t pixel, uint32_t sample, uint32_t b, uint32_t* out) {
    uint32_t key[2] = {k0, k1};
    uint32_t c[4] = {pixel, sample, b, 0};
    philox(c, key, out);
}

// ------------------------------------------------------------ Disney -----
template <class R>
struct HitData {   // kernel.h:47-69
    R metallic, roughness, clearcoatGloss, clearcoat, anisotropic, eta, transmission, specular, specularTint,
        sheenTint, subsurface, sheen;
    V3<R> emission, albedo, normal, tangent, bitangent;
};

template <class R> void createBasis(V3<R> normal, V3<R>& tangent, V3<R>& bitangent) {   // Disney.hpp:41-45
    V3<R> up((R)0, (R)1, (R)0);
    tangent = cross(up, normal).normalized();
    bitangent = cross(normal, tangent).normalized();
}
template <class R> R SchlickFresnel(R u) {   // Disney.hpp:47-51
    R m = clampR<R>(1.0 - u, 0.0, 1.0);
    R m2 = m * m;
    return m2 * m2 * m;
}
template <class R> R GTR1(R NDotH, R a) {   // Disney.hpp:68-74
    if (a >= 1.0) return (1.0 / piR<R>());
    R a2 = a * a;
    R t = 1.0 + (a2 - 1.0) * NDotH * NDotH;
    return (a2 - 1.0) / (piR<R>() * std::log(a2) * t);
}
template <class R> R GTR2_aniso(R NDotH, R HDotX, R HDotY, R ax, R ay) {   // Disney.hpp:82-87
    R a = HDotX / ax;
    R b = HDotY / ay;
    R c = a * a + b * b + NDotH * NDotH;
    return 1.0 / (piR<R>() * ax * ay * c * c);
}
template <class R> R SmithG_GGX(R NDotV, R alphaG) {   // Disney.hpp:89-93
    R a = alphaG * alphaG;
    R b = NDotV * NDotV;
    return 1.0 / (NDotV + std::sqrt(a + b - a * b));
}
template <class R> R SmithG_GGX_aniso(R NDotV, R VDotX, R VDotY, R ax, R ay) {   // Disney.hpp:95-100
    R a = VDotX * ax;
    R b = VDotY * ay;
    R c = NDotV;
    return 1.0 / (NDotV + std::sqrt(a * a + b * b + c * c));
}

template <class R> R DisneyPdf(V3<R> dir, const HitData<R>& h, V3<R> L) {   // Disney.hpp:108-147
    V3<R> N = h.normal;
    V3<R> V = neg(dir);
    V3<R> H = (L + V).normalized();
    V3<R> T, B;
    createBasis(h.normal, T, B);
    R NDotH = std::fabs(dot(N, H));
    if (dot(N, L) <= 0.0) return 1.0;
    R specularAlpha = maxR<R>(0.001, h.roughness);
    (void)specularAlpha;
    R clearcoatAlpha = lerpR<R>(0.1, 0.001, h.clearcoatGloss);
    R diffuseRatio = 0.5 * (1.0 - h.metallic);
    R specularRatio = 1.0 - diffuseRatio;
    R aspect = std::sqrt((R)(1.0 - h.anisotropic * 0.9));
    R ax = maxR<R>(0.001, h.roughness / aspect);
    R ay = maxR<R>(0.001, h.roughness * aspect);
    R pdfGTR2_aniso = GTR2_aniso<R>(NDotH, dot(H, T), dot(H, B), ax, ay) * NDotH;
    R pdfGTR1 = GTR1<R>(NDotH, clearcoatAlpha) * NDotH;
    R ratio = 1.0 / (1.0 + h.clearcoat);
    R pdfSpec = lerpR<R>(pdfGTR1, pdfGTR2_aniso, ratio) / (4.0 * std::fabs(dot(L, H)));
    R pdfDiff = std::fabs(dot(L, N)) * (1.0 / piR<R>());
    return diffuseRatio * pdfDiff + specularRatio * pdfSpec;
}

template <class R> V3<R> CosineSampleHemisphere(R u1, R u2) {   // Sampling.hpp:31-41
    V3<R> dir;
    R r = std::sqrt(u1);
    R phi = 2.0 * piR<R>() * u2;
    dir.x = r * std::cos(phi);
    dir.y = r * std::sin(phi);
    dir.z = std::sqrt(maxR<R>(0.0, 1.0 - dir.x * dir.x - dir.y * dir.y));
    return dir;
}
template <class R> V3<R> ImportanceSampleGGX(R rgh, R r1, R r2) {   // Sampling.hpp:43-54
    R a = maxR<R>(0.001, rgh);
    R phi = r1 * piR<R>() * 2;
    R cosTheta = std::sqrt((R)((1.0 - r2) / (1.0 + (a * a - 1.0) * r2)));
    R sinTheta = clampR<R>(std::sqrt((R)(1.0 - (cosTheta * cosTheta))), 0.0, 1.0);
    R sinPhi = std::sin(phi);
    R cosPhi = std::cos(phi);
    return V3<R>(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}
template <class R> void uniformCircleSampling(R u1, R u2, R u3, R& x, R& y) {   // Sampling.hpp:21-29
    R t = 2 * piR<R>() * u1;
    R u = u2 + u3;
    R r = u > 1 ? 2 - u : u;
    x = r * std::cos(t);
    y = r * std::sin(t);
}

template <class R> V3<R> DisneySample(V3<R> dir, const HitData<R>& h, R r1, R r2, R r3) {   // Disney.hpp:150-177
    V3<R> N = h.normal;
    V3<R> V = neg(dir);
    V3<R> T, B;
    createBasis(h.normal, T, B);
    R diffuseRatio = 0.5 * (1.0 - h.metallic);
    if (r3 < diffuseRatio) {
        V3<R> H = CosineSampleHemisphere<R>(r1, r2);
        H = T * H.x + B * H.y + N * H.z;
        return H;
    }
    V3<R> H = ImportanceSampleGGX<R>(h.roughness, r1, r2);
    H = T * H.x + B * H.y + N * H.z;
    return reflect(neg(V), H);
}

template <class R> V3<R> DisneyEval(V3<R> dir, const HitData<R>& h, V3<R> L) {   // Disney.hpp:179-253
    V3<R> V = neg(dir);
    V3<R> T, B;
    createBasis(h.normal, T, B);
    V3<R> H = (L + V).normalized();
    R NDotL = std::fabs(dot(h.normal, L));
    R NDotV = std::fabs(dot(h.normal, V));
    R NDotH = std::fabs(dot(h.normal, H));
    R LDotH = std::fabs(dot(L, H));
    V3<R> brdf((R)0.0);
    if (h.transmission < 1.0 && dot(h.normal, L) > 0.0 && dot(h.normal, V) > 0.0) {
        V3<R> Cdlin = h.albedo;
        R Cdlum = 0.3 * Cdlin.x + 0.6 * Cdlin.y + 0.1 * Cdlin.z;
        V3<R> Ctint = Cdlum > 0.0 ? Cdlin / Cdlum : V3<R>((R)1.0);
        V3<R> Cspec0 = lerpV<R>((R)(h.specular * 0.08) * lerpV<R>(V3<R>((R)1.0), Ctint, h.specularTint), Cdlin,
                                h.metallic);
        V3<R> Csheen = lerpV<R>(V3<R>((R)1.0), Ctint, h.sheenTint);
        R FL = SchlickFresnel<R>(NDotL), FV = SchlickFresnel<R>(NDotV);
        R Fd90 = 0.5 + 2.0 * LDotH * LDotH * h.roughness;
        R Fd = lerpR<R>(1.0, Fd90, FL) * lerpR<R>(1.0, Fd90, FV);
        R Fss90 = LDotH * LDotH * h.roughness;
        R Fss = lerpR<R>(1.0, Fss90, FL) * lerpR<R>(1.0, Fss90, FV);
        R ss = 1.25 * (Fss * (1.0 / (NDotL + NDotV) - 0.5) + 0.5);
        R aspect = std::sqrt((R)(1.0 - h.anisotropic * 0.9));
        R ax = maxR<R>(0.001, h.roughness / aspect);
        R ay = maxR<R>(0.001, h.roughness * aspect);
        R Ds = GTR2_aniso<R>(NDotH, dot(H, T), dot(H, B), ax, ay);
        R FH = SchlickFresnel<R>(LDotH);
        V3<R> Fs = lerpV<R>(Cspec0, V3<R>((R)1.0), FH);
        R Gs = SmithG_GGX_aniso<R>(NDotL, dot(L, T), dot(L, B), ax, ay);
        Gs *= SmithG_GGX_aniso<R>(NDotV, dot(V, T), dot(V, B), ax, ay);
        V3<R> Fsheen = (R)(FH * h.sheen) * Csheen;
        R Dr = GTR1<R>(NDotH, lerpR<R>(0.1, 0.001, h.clearcoatGloss));
        R Fr = lerpR<R>(0.04, 1.0, FH);
        R Gr = SmithG_GGX<R>(NDotL, 0.25) * SmithG_GGX<R>(NDotV, 0.25);
        brdf = ((R)((1.0 / piR<R>()) * lerpR<R>(Fd, ss, h.subsurface)) * Cdlin + Fsheen) * (R)(1.0 - h.metallic) +
               (Gs * Fs) * Ds + V3<R>((R)(0.25 * h.clearcoat * Gr * Fr * Dr));
    }
    return brdf;
}

// Derivatives of DisneyEval (above) with respect to k = albedo.x, albedo.y,
// albedo.z, roughness, metallic, specular of HitData (the effective values,
// after generateHitData's pow 2.2), written out analytically from the same
// expression; df[k] is the RGB derivative.  Not part of the reference (it has
// no adjoint); pinned against central differences of DisneyEval.
template <class R> void DisneyEvalGrad(V3<R> dir, const HitData<R>& h, V3<R> L, V3<R>* df) {
    for (int k = 0; k < 6; k++) df[k] = V3<R>((R)0);
    V3<R> V = neg(dir);
    V3<R> T, B;
    createBasis(h.normal, T, B);
    V3<R> H = (L + V).normalized();
    R NDotL = std::fabs(dot(h.normal, L));
    R NDotV = std::fabs(dot(h.normal, V));
    R NDotH = std::fabs(dot(h.normal, H));
    R LDotH = std::fabs(dot(L, H));
    if (!(h.transmission < 1.0 && dot(h.normal, L) > 0.0 && dot(h.normal, V) > 0.0)) return;
    const R pi = piR<R>();
    V3<R> C = h.albedo;
    const R wl[3] = {(R)0.3, (R)0.6, (R)0.1};
    R lum = wl[0] * C.x + wl[1] * C.y + wl[2] * C.z;
    V3<R> Ctint = lum > 0.0 ? C / lum : V3<R>((R)1.0);
    V3<R> dCtint[3];
    for (int k = 0; k < 3; k++) {
        V3<R> e((R)(k == 0), (R)(k == 1), (R)(k == 2));
        dCtint[k] = lum > 0.0 ? e / lum - C * (wl[k] / (lum * lum)) : V3<R>((R)0);
    }
    const R mu = h.metallic, rho = h.roughness, sig = h.specular;
    V3<R> tintMix = V3<R>((R)1) + (Ctint - V3<R>((R)1)) * h.specularTint;
    R S0 = sig * (R)0.08;
    V3<R> Cspec0 = S0 * tintMix + (C - S0 * tintMix) * mu;
    V3<R> Csheen = V3<R>((R)1) + (Ctint - V3<R>((R)1)) * h.sheenTint;
    R FL = SchlickFresnel<R>(NDotL), FV = SchlickFresnel<R>(NDotV), FH = SchlickFresnel<R>(LDotH);
    R Fd90 = 0.5 + 2.0 * LDotH * LDotH * rho;
    R A = 1 + FL * (Fd90 - 1), Bd = 1 + FV * (Fd90 - 1);
    R Fd = A * Bd, dFd = 2 * LDotH * LDotH * (FL * Bd + FV * A);
    R Fss90 = LDotH * LDotH * rho;
    R A2 = 1 + FL * (Fss90 - 1), B2 = 1 + FV * (Fss90 - 1);
    R Fss = A2 * B2, dFss = LDotH * LDotH * (FL * B2 + FV * A2);
    R k1 = 1.0 / (NDotL + NDotV) - 0.5;
    R ss = 1.25 * (Fss * k1 + 0.5), dss = 1.25 * k1 * dFss;
    R D0 = Fd + h.subsurface * (ss - Fd), dD0 = dFd + h.subsurface * (dss - dFd);
    R aspect = std::sqrt((R)(1.0 - h.anisotropic * 0.9));
    R ax = maxR<R>(0.001, rho / aspect), ay = maxR<R>(0.001, rho * aspect);
    R dax = rho / aspect > 0.001 ? 1 / aspect : 0, day = rho * aspect > 0.001 ? aspect : 0;
    R HX = dot(H, T), HY = dot(H, B);
    R a = HX / ax, b = HY / ay, c = a * a + b * b + NDotH * NDotH;
    R Ds = 1.0 / (pi * ax * ay * c * c);
    R dDs = Ds * (-1 / ax + 4 * a * a / (c * ax)) * dax + Ds * (-1 / ay + 4 * b * b / (c * ay)) * day;
    auto G = [&](R N, R X, R Y, R& g, R& dg) {
        R q = std::sqrt((X * ax) * (X * ax) + (Y * ay) * (Y * ay) + N * N);
        g = 1.0 / (N + q);
        dg = -g * g * ((X * X * ax) * dax + (Y * Y * ay) * day) / q;
    };
    R GL, dGL, GV, dGV;
    G(NDotL, dot(L, T), dot(L, B), GL, dGL);
    G(NDotV, dot(V, T), dot(V, B), GV, dGV);
    R Gs = GL * GV, dGs = dGL * GV + GL * dGV;
    V3<R> Fs = Cspec0 + (V3<R>((R)1) - Cspec0) * FH;
    V3<R> Fsheen = Csheen * (FH * h.sheen);
    V3<R> diffusePart = C * (R)((1.0 / pi) * D0) + Fsheen;
    R spec = Gs * Ds * (1 - FH);
    for (int k = 0; k < 3; k++) {
        V3<R> e((R)(k == 0), (R)(k == 1), (R)(k == 2));
        V3<R> dCspec0 = S0 * h.specularTint * (1 - mu) * dCtint[k] + e * mu;
        V3<R> dFsheen = dCtint[k] * (FH * h.sheen * h.sheenTint);
        df[k] = (e * (R)((1.0 / pi) * D0) + dFsheen) * (R)(1 - mu) + dCspec0 * spec;
    }
    df[3] = C * (R)((1.0 / pi) * dD0 * (1 - mu)) + Fs * (R)(dGs * Ds + Gs * dDs);
    df[4] = neg(diffusePart) + (C - S0 * tintMix) * spec;
    df[5] = tintMix * ((R)0.08 * (1 - mu) * spec);
}

// ----------------------------------------------------------- texture -----
struct Tex {
    int w = 1, h = 1;
    const float* data = nullptr;
    float xTile = 1, yTile = 1, xOffset = 0, yOffset = 0;
    int filter = 0;
};
inline Tex texOf(const el_texture& t) {
    Tex x;
    x.w = t.width;
    x.h = t.height;
    x.data = t.data;
    x.xTile = t.x_tile;
    x.yTile = t.y_tile;
    x.xOffset = t.x_offset;
    x.yOffset = t.y_offset;
    x.filter = t.filter;
    return x;
}
// Texture.hpp:95-106; negative coordinates (undefined in the reference) wrap.
template <class R> V3<R> texel(const Tex& t, int x, int y) {
    x = (int)(t.xTile * (x + t.xOffset * t.w)) % t.w;
    y = (int)(t.yTile * (y + t.yOffset * t.h)) % t.h;
    if (x < 0) x += t.w;
    if (y < 0) y += t.h;
    const float* p = t.data + 3 * ((size_t)y * t.w + x);
    return V3<R>((R)p[0], (R)p[1], (R)p[2]);
}
// Linear index of the texel texel() reads for (x, y).
inline size_t texelIndex(const Tex& t, int x, int y) {
    x = (int)(t.xTile * (x + t.xOffset * t.w)) % t.w;
    y = (int)(t.yTile * (y + t.yOffset * t.h)) % t.h;
    if (x < 0) x += t.w;
    if (y < 0) y += t.h;
    return (size_t)y * t.w + x;
}
template <class R> V3<R> texUV(const Tex& t, R u, R v) { return texel<R>(t, (int)(u * t.w), (int)(v * t.h)); }
template <class R> V3<R> texBilinear(const Tex& t, R u, R v) {   // Texture.hpp:114-135
    R x = u * t.w, y = v * t.h;
    R t1x = std::floor(x), t1y = std::floor(y);
    R t2x = t1x + 1, t2y = t1y + 1;
    R a = (x - t1x) / (t2x - t1x), b = (y - t1y) / (t2y - t1y);
    V3<R> v1 = texel<R>(t, (int)t1x, (int)t1y), v2 = texel<R>(t, (int)t2x, (int)t1y);
    V3<R> v3 = texel<R>(t, (int)t1x, (int)t2y), v4 = texel<R>(t, (int)t2x, (int)t2y);
    return lerpV(lerpV(v1, v2, a), lerpV(v3, v4, a), b);
}
template <class R> V3<R> texFiltered(const Tex& t, R u, R v) {
    return t.filter == 1 ? texBilinear<R>(t, u, v) : texUV<R>(t, u, v);
}
template <class R> void limitUV(R& u, R& v) {   // Math.hpp:38-41
    u += -(R)(u > 1) + -(R)(u < 0);
    v += -(R)(v > 1) + -(R)(v < 0);
}
template <class R> void sphericalMapping(V3<R> p, R& u, R& v) {   // Texture.hpp:144-156 (unit sphere)
    R theta = std::acos(clampR<R>(-p.y, -1, 1));
    R phi = std::atan2(-p.z, p.x) + piR<R>();
    u = phi / (2 * piR<R>());
    v = theta / piR<R>();
    limitUV(u, v);
}

// ---------------------------------------------------------- estimator -----
struct SceneC {
    const el_scene_view* v;
    const RefBVH* bvh;
    Tex env;
    std::vector<Tex> texs;
    std::vector<float> cdf;
    float radianceSum = 0;
    const el_material* evalMats = nullptr;   // or_render_params::eval_materials
};

template <class R>
struct Surf {
    V3<R> position, normal, tangent, bitangent;
    R tu, tv;
    int material;
};

// Tri.hpp:70-157 (attributes for a hit found by the traversal)
template <class R>
Surf<R> surface(const SceneC& s, int tri, float t, float u_, float v_, V3<R> o, V3<R> d) {
    const el_scene_view& v = *s.v;
    auto vtx = [&](const float* a, int k) { return V3<R>((R)a[3 * k], (R)a[3 * k + 1], (R)a[3 * k + 2]); };
    const float* P = v.vertices + (size_t)tri * 9;
    const float* N = v.normals ? v.normals + (size_t)tri * 9 : nullptr;
    const float* T = v.tangents ? v.tangents + (size_t)tri * 9 : nullptr;
    const float* UV = v.uvs ? v.uvs + (size_t)tri * 6 : nullptr;
    R u = u_, w = v_;
    Surf<R> h;
    V3<R> zero;
    V3<R> n0 = N ? vtx(N, 0) : zero, n1 = N ? vtx(N, 1) : zero, n2 = N ? vtx(N, 2) : zero;
    V3<R> t0 = T ? vtx(T, 0) : zero, t1 = T ? vtx(T, 1) : zero, t2 = T ? vtx(T, 2) : zero;
    R uv0x = UV ? UV[0] : 0, uv0y = UV ? UV[1] : 0, uv1x = UV ? UV[2] : 0, uv1y = UV ? UV[3] : 0;
    R uv2x = UV ? UV[4] : 0, uv2y = UV ? UV[5] : 0;
    h.tu = uv0x + (uv1x - uv0x) * u + (uv2x - uv0x) * w;
    h.tv = uv0y + (uv1y - uv0y) * u + (uv2y - uv0y) * w;
    V3<R> geom = o + d * (R)t;
    V3<R> Ns = n0 + (n1 - n0) * u + (n2 - n0) * w;
    V3<R> Ts = t0 + (t1 - t0) * u + (t2 - t0) * w;
    V3<R> v0 = vtx(P, 0), v1 = vtx(P, 1), v2 = vtx(P, 2);
    V3<R> p0 = geom - dot(geom - v0, n0) * n0;   // projectOnPlane (Tri.hpp:34-36)
    V3<R> p1 = geom - dot(geom - v1, n1) * n1;
    V3<R> p2 = geom - dot(geom - v2, n2) * n2;
    V3<R> sp = p0 + (p1 - p0) * u + (p2 - p0) * w;
    bool convex = dot(sp - geom, Ns) > 0.0f;
    h.position = convex ? sp : geom;
    h.normal = Ns;
    h.tangent = Ts;
    R sign = v.tangent_sign ? (R)v.tangent_sign[tri] : 0;
    h.bitangent = sign * cross(h.normal, h.tangent);
    int obj = v.object_id[tri];
    int mat = (obj >= 0 && (uint32_t)obj < v.mesh_count) ? v.mesh_material[obj] : 0;
    if (mat < 0 || (uint32_t)mat >= v.material_count) mat = 0;
    h.material = mat;
    return h;
}

// generateHitData (kernel.cu:54-119)
template <class R>
HitData<R> hitData(const SceneC& s, const Surf<R>& h, const el_material* mats = nullptr) {
    const el_material& m = (mats ? mats : s.v->materials)[h.material];
    auto tex = [&](int id) -> const Tex* {
        return (id >= 0 && (size_t)id < s.texs.size()) ? &s.texs[id] : nullptr;
    };
    HitData<R> d;
    d.albedo = tex(m.albedo_tex) ? texFiltered<R>(*tex(m.albedo_tex), h.tu, h.tv)
                                 : V3<R>((R)m.albedo.x, (R)m.albedo.y, (R)m.albedo.z);
    d.emission = tex(m.emission_tex) ? texFiltered<R>(*tex(m.emission_tex), h.tu, h.tv)
                                     : V3<R>((R)m.emission.x, (R)m.emission.y, (R)m.emission.z);
    d.roughness = tex(m.roughness_tex) ? texFiltered<R>(*tex(m.roughness_tex), h.tu, h.tv).x : (R)m.roughness;
    d.metallic = tex(m.metallic_tex) ? texFiltered<R>(*tex(m.metallic_tex), h.tu, h.tv).x : (R)m.metallic;
    if (!tex(m.normal_tex)) {
        d.normal = h.normal;
    } else {
        V3<R> c = texUV<R>(*tex(m.normal_tex), h.tu, h.tv);
        V3<R> ln = c * (R)2 - V3<R>((R)1);
        d.normal = (ln.x * h.tangent - ln.y * h.bitangent + ln.z * h.normal).normalized();
    }
    d.roughness = std::pow(d.roughness, (R)2.2f);
    d.metallic = std::pow(d.metallic, (R)2.2f);
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
    d.tangent = h.tangent;
    d.bitangent = h.bitangent;
    return d;
}

// HDRI.hpp:130-142
inline int binarySearch(const float* arr, float value, int length) {
    int from = 0, to = length - 1;
    while (to - from > 0) {
        int m = from + (to - from) / 2;
        if (value == arr[m]) return m;
        if (value < arr[m]) to = m - 1;
        if (value > arr[m]) from = m + 1;
    }
    return to;
}

template <class R>
Hit closest(const SceneC& s, V3<R> o, V3<R> d) {
    float of[3] = {(float)o.x, (float)o.y, (float)o.z}, df[3] = {(float)d.x, (float)d.y, (float)d.z};
    return refTraverse(*s.bvh, of, df);
}

// Everything one path vertex computed, for the adjoint's replay.
template <class R>
struct VertexRec {
    int depth;
    V3<R> d, T;                 // incoming direction, throughput before the vertex
    Surf<R> sh;
    HitData<R> hd;
    V3<R> bdir, f;              // sampled direction and DisneyEval there
    R bpdf, cosS;
    V3<R> hdir, hval;           // HDRI sample direction and texel
    R hpdf, cosH;
    bool visH;
    bool hasP, visP;
    V3<R> ldir, lrad;           // light direction, radiance / dist^2
    R ppdf, cosL;
    R w1, w2, w3;
    V3<R> Lincl;                // radiance gathered up to and including this vertex
};

// One path (kernel.cu:369-481), radiance before the clamp; `hook` sees every
// vertex.
template <class R, class Hook>
void walkPath(const SceneC& s, const or_render_params& p, int x, int y, uint32_t sample, V3<R>& Lraw,
              uint64_t& rays, uint64_t& verts, Hook&& hook) {
    const el_scene_view& v = *s.v;
    const el_camera& cam = v.camera;
    uint32_t W = cam.x_res, Hh = cam.y_res;
    uint32_t idx = W * (Hh - y - 1) + x;
    uint32_t k0 = (uint32_t)p.seed, k1 = (uint32_t)(p.seed >> 32);
    uint32_t c[5], c1[4];
    pathBlock(k0, k1, idx, sample, 0, c);
    pathBlock(k0, k1, idx, sample, 1, c1);
    c[4] = c1[0];

    // calculateCameraRay (kernel.cu:260-337)
    V3<R> o, d;
    {
        R r1 = u01(c[0]), r2 = u01(c[1]), r3 = u01(c[2]), r4 = u01(c[3]), r5 = u01(c[4]);
        R dx = cam.position.x + ((R)x) / ((R)cam.x_res) * (R)cam.sensor_width;
        R dy = cam.position.y + ((R)y) / ((R)cam.y_res) * (R)cam.sensor_height;
        R odx = (-cam.sensor_width / 2.0) + dx;
        R ody = (-cam.sensor_height / 2.0) + dy;
        R rx = (1.0 / (R)cam.x_res) * (r1 - 0.5) * cam.sensor_width;
        R ry = (1.0 / (R)cam.y_res) * (r2 - 0.5) * cam.sensor_height;
        R SPx = odx + rx, SPy = ody + ry;
        R SPz = (R)cam.position.z + (R)cam.focal_length;
        V3<R> rot((R)cam.rotation.x, (R)cam.rotation.y, (R)cam.rotation.z);
        rot = rot * (R)(piR<R>() / 180.0);
        V3<R> pos((R)cam.position.x, (R)cam.position.y, (R)cam.position.z);
        V3<R> dir = V3<R>(SPx, SPy, SPz) - pos;
        V3<R> dX(dir.x, dir.y * std::cos(rot.x) - dir.z * std::sin(rot.x), dir.y * std::sin(rot.x) + dir.z * std::cos(rot.x));
        V3<R> dY(dX.x * std::cos(rot.y) + dX.z * std::sin(rot.y), dX.y, dX.z * std::cos(rot.y) - dX.x * std::sin(rot.y));
        V3<R> dZ(dY.x * std::cos(rot.z) - dY.y * std::sin(rot.z), dY.x * std::sin(rot.z) + dY.y * std::cos(rot.z), dY.z);
        o = pos;
        d = dZ.normalized();
        if (cam.bokeh) {
            R diameter = (R)cam.focal_length / (R)cam.aperture;
            R l = (R)cam.focus_distance + (R)cam.focal_length;
            V3<R> focus = o + d * l;
            R px, py;
            uniformCircleSampling<R>(r3, r4, r5, px, py);
            px *= (R)(diameter * 0.5);
            py *= (R)(diameter * 0.5);
            V3<R> orig = pos + V3<R>(px, py, (R)0);
            o = orig;
            d = (focus - orig).normalized();
        }
    }
    V3<R> L((R)0), T((R)1);
    int i;
    const int nl = (int)v.light_count;
    for (i = 0; i < (int)p.depth; i++) {
        rays++;
        Hit hit = closest(s, o, d);
        if (hit.tri < 0) {
            R u, w;
            sphericalMapping<R>(neg(d), u, w);
            L = L + texFiltered<R>(s.env, u, w) * T;
            break;
        }
        Surf<R> sh = surface<R>(s, hit.tri, hit.t, hit.u, hit.v, o, d);
        HitData<R> hd = hitData<R>(s, sh);
        // BRDF values from the evaluation materials when given (sampling,
        // pdfs and MIS weights keep the scene's): the detached-sampling
        // estimator whose derivative the adjoint computes
        const HitData<R> he = s.evalMats ? hitData<R>(s, sh, s.evalMats) : hd;
        uint32_t r[4];
        pathBlock(k0, k1, idx, sample, 2 + i, r);
        V3<R> bdir = DisneySample<R>(d, hd, u01(r[0]), u01(r[1]), u01(r[2]));

        // shade (kernel.cu:339-358)
        V3<R> f = DisneyEval<R>(d, he, bdir);
        R bpdf = DisneyPdf<R>(d, hd, bdir);
        R r1 = u01(r[3]);
        // hdriLight (kernel.cu:234-257)
        R hpdf;
        V3<R> Hl, vHdir, vHval, vLdir, vLrad;
        bool vHvis = false, vPvis = false;
        {
            const Tex& e = s.env;
            int k = binarySearch(s.cdf.data(), (float)r1, e.w * e.h);   // HDRI::sample
            int tx = k % e.w, ty = k / e.w;
            R nu = tx / (R)e.w, nv = ty / (R)e.h;
            int ix = (int)(nu * e.w), iy = (int)(nv * e.h);   // inverseTransformUV (Texture.hpp:177-193)
            ix = (int)(e.xTile * (ix - e.xOffset * e.w)) % e.w;
            iy = (int)(e.yTile * (iy - e.yOffset * e.h)) % e.h;
            R iu = (R)ix / (R)e.w, iv = (R)iy / (R)e.h;
            limitUV(iu, iv);
            R phi = iu * 2 * piR<R>(), theta = iv * piR<R>();   // reverseSphericalMapping (Texture.hpp:195-207)
            R px = std::cos(phi - piR<R>()), py = -std::cos(theta), pz = -std::sin(phi - piR<R>());
            R a = std::sqrt(1 - py * py);
            V3<R> nd = neg(V3<R>(a * px, py, a * pz).normalized());
            V3<R> so = sh.position + nd * (R)0.001;
            V3<R> sd = nd.normalized();
            rays++;
            bool occ = closest(s, so, sd).tri >= 0;
            V3<R> val = texUV<R>(e, iu, iv);
            V3<R> fe = DisneyEval<R>(d, he, nd);
            // HDRI::pdf((int)(iu*W), (int)(iv*H)) (HDRI.hpp:145-152), evaluated
            // whether or not the shadow ray is blocked
            int qx = (int)(iu * e.w), qy = (int)(iv * e.h);
            V3<R> dv = texel<R>(e, qx, qy);
            R th = (((R)qy / (R)e.h)) * piR<R>();
            hpdf = ((dv.x + dv.y + dv.z) / (R)s.radianceSum) * (R)e.w * (R)e.h / (2.0 * piR<R>() * std::sin(th));
            Hl = occ ? V3<R>() : fe * std::fabs(dot(nd, hd.normal)) * val / hpdf;
            vHdir = nd;
            vHval = val;
            vHvis = !occ;
        }
        // pointLight (kernel.cu:175-205)
        R ppdf = 0;
        V3<R> Pl;
        if (nl > 0) {
            ppdf = ((R)nl) / (2.0 * piR<R>());
            int li = std::min((int)(nl * r1), nl - 1);
            const el_point_light& lt = v.lights[li];
            V3<R> lp((R)lt.position.x, (R)lt.position.y, (R)lt.position.z);
            V3<R> lr((R)lt.radiance.x, (R)lt.radiance.y, (R)lt.radiance.z);
            V3<R> nd = (lp - sh.position).normalized();
            R dist = (lp - sh.position).length();
            V3<R> so = sh.position + nd * (R)0.001;
            V3<R> sd = nd.normalized();
            rays++;
            Hit sh2 = closest(s, so, sd);
            bool occ;
            if (p.shadow_mode == 0) {
                if (sh2.tri >= 0) {
                    Surf<R> ss = surface<R>(s, sh2.tri, sh2.t, sh2.u, sh2.v, so, sd);
                    occ = (ss.position - sh.position).length() < dist;
                } else {
                    occ = false;
                }
            } else {
                occ = sh2.tri >= 0 && sh2.t < (float)(dist - (R)0.001);
            }
            Pl = occ ? V3<R>() : (lr / (dist * dist)) * DisneyEval<R>(d, he, nd) * std::fabs(dot(nd, hd.normal)) / ppdf;
            vLdir = nd;
            vLrad = lr / (dist * dist);
            vPvis = !occ;
        }
        V3<R> B = hd.emission * (f * std::fabs(dot(bdir, hd.normal))) / bpdf;
        R w1 = hpdf / (hpdf + ppdf + bpdf);
        R w2 = ppdf / (hpdf + ppdf + bpdf);
        R w3 = bpdf / (hpdf + ppdf + bpdf);
        VertexRec<R> rec;
        rec.depth = i;
        rec.d = d;
        rec.T = T;
        L = L + T * (w1 * Hl + w2 * Pl + w3 * B);
        T = T * ((f * std::fabs(dot(bdir, hd.normal))) / bpdf);
        verts++;
        rec.sh = sh;
        rec.hd = hd;
        rec.bdir = bdir;
        rec.f = f;
        rec.bpdf = bpdf;
        rec.cosS = std::fabs(dot(bdir, hd.normal));
        rec.hdir = vHdir;
        rec.hval = vHval;
        rec.hpdf = hpdf;
        rec.cosH = std::fabs(dot(vHdir, hd.normal));
        rec.visH = vHvis;
        rec.hasP = nl > 0;
        rec.visP = vPvis;
        rec.ldir = vLdir;
        rec.lrad = vLrad;
        rec.ppdf = ppdf;
        rec.cosL = std::fabs(dot(vLdir, hd.normal));
        rec.w1 = w1;
        rec.w2 = w2;
        rec.w3 = w3;
        rec.Lincl = L;
        hook(rec);
        o = sh.position + bdir * (R)0.001;
        d = bdir.normalized();
    }
    (void)i;
    Lraw = L;
}

struct NoHook {
    template <class T>
    void operator()(const T&) const {}
};

// One path (kernel.cu:369-481).  Returns true if the sample is accepted.
// aov9 (optional): the first hit's normal, tangent and bitangent
// (kernel.cu:435-440; zero if the camera ray escapes).
template <class R>
bool tracePath(const SceneC& s, const or_render_params& p, int x, int y, uint32_t sample, V3<R>& Lout,
               uint64_t& rays, uint64_t& verts, double* aov9 = nullptr) {
    V3<R> L;
    if (aov9)
        for (int k = 0; k < 9; k++) aov9[k] = 0;
    walkPath<R>(s, p, x, y, sample, L, rays, verts, [&](const VertexRec<R>& rec) {
        if (aov9 && rec.depth == 0) {
            const V3<R> a[3] = {rec.sh.normal, rec.sh.tangent, rec.sh.bitangent};
            for (int k = 0; k < 3; k++) {
                aov9[3 * k] = (double)a[k].x;
                aov9[3 * k + 1] = (double)a[k].y;
                aov9[3 * k + 2] = (double)a[k].z;
            }
        }
    });
    L = V3<R>(clampR<R>(L.x, 0, 10), clampR<R>(L.y, 0, 10), clampR<R>(L.z, 0, 10));
    if (std::isnan(L.x) || std::isnan(L.y) || std::isnan(L.z)) return false;
    Lout = L;
    return true;
}

template <class R>
void renderT(const SceneC& s, const or_render_params& p, double* sum, double* sumsq, uint32_t* count,
             std::atomic<uint64_t>& rays, std::atomic<uint64_t>& verts) {
    const el_camera& cam = s.v->camera;
    int nx = p.x1 - p.x0, ny = p.y1 - p.y0;
    parallelFor(nx * ny, p.threads, [&](int k) {
        int x = p.x0 + k % nx, y = p.y0 + k / nx;
        uint32_t idx = cam.x_res * (cam.y_res - y - 1) + x;
        uint64_t lr = 0, lv = 0;
        double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, ao[9] = {0}, a9[9];
        uint32_t n = 0;
        for (uint32_t sidx = 0; sidx < p.spp; sidx++) {
            V3<R> L;
            if (tracePath<R>(s, p, x, y, p.s0 + sidx, L, lr, lv, p.aov ? a9 : nullptr)) {
                double c3[3] = {(double)L.x, (double)L.y, (double)L.z};
                for (int j = 0; j < 3; j++) {
                    a[j] += c3[j];
                    b[j] += c3[j] * c3[j];
                }
                if (p.aov)
                    for (int j = 0; j < 9; j++) ao[j] += a9[j];
                n++;
            }
        }
        for (int j = 0; j < 3; j++) {
            sum[3 * (size_t)idx + j] += a[j];
            sumsq[3 * (size_t)idx + j] += b[j];
        }
        if (p.aov)
            for (int j = 0; j < 9; j++) p.aov[9 * (size_t)idx + j] += ao[j];
        count[idx] += n;
        rays += lr;
        verts += lv;
    });
}

// Path-replay adjoint (Vicini, Speierer, Jakob 2021) of the estimator above
// with detached sampling: sampled directions, lobe choices, pdfs and MIS
// weights are held; the derivative flows through DisneyEval at the sampled,
// HDRI and light directions.  Per accepted sample with adjoint weight a
// (dLoss/dI_pixel / accepted samples, zeroed on clamped channels):
//   g += a . [ (T w3 Le cosS/bpdf + L_after / f) dF_s
//            + T w1 env cosH/hpdf dF_h [visible] + T w2 rad cosL/ppdf dF_l [visible] ]
// with L_after = radiance gathered after the vertex (L_total - L_incl).
template <class R>
void adjointT(const SceneC& s, const or_render_params& rp, const or_adjoint_params& p, const float* dimage,
              double* dmat, double* dtex, double* image) {
    const el_scene_view& v = *s.v;
    const el_camera& cam = v.camera;
    const int W = (int)cam.x_res, Hh = (int)cam.y_res, npx = W * Hh;
    const int nthreads = std::max(1, rp.threads > 0 ? rp.threads : (int)std::thread::hardware_concurrency());
    const int M = (int)v.material_count;
    const Tex* gtex = (p.grad_texture >= 0 && (size_t)p.grad_texture < s.texs.size()) ? &s.texs[p.grad_texture] : nullptr;
    const size_t ntex = gtex ? (size_t)gtex->w * gtex->h * 3 : 0;
    std::vector<std::vector<double>> gm(nthreads, std::vector<double>((size_t)M * 6, 0.0));
    std::vector<std::vector<double>> gt(nthreads, std::vector<double>(ntex, 0.0));
    std::atomic<int> next(0);
    auto work = [&](int tid) {
        std::vector<V3<R>> Ls(p.spp);
        std::vector<char> ok(p.spp);
        for (;;) {
            int k = next.fetch_add(1);
            if (k >= npx) break;
            int x = k % W, y = k / W;
            uint32_t idx = (uint32_t)(W * (Hh - y - 1) + x);
            uint64_t rays = 0, verts = 0;
            int n = 0;
            double mean[3] = {0, 0, 0};
            for (uint32_t j = 0; j < p.spp; j++) {
                walkPath<R>(s, rp, x, y, p.s0 + j, Ls[j], rays, verts, NoHook());
                V3<R> c(clampR<R>(Ls[j].x, 0, 10), clampR<R>(Ls[j].y, 0, 10), clampR<R>(Ls[j].z, 0, 10));
                ok[j] = !(std::isnan(c.x) || std::isnan(c.y) || std::isnan(c.z));
                if (ok[j]) {
                    n++;
                    mean[0] += (double)c.x;
                    mean[1] += (double)c.y;
                    mean[2] += (double)c.z;
                }
            }
            if (image)
                for (int c = 0; c < 3; c++) image[3 * (size_t)idx + c] = n ? mean[c] / n : 0.0;
            if (!n) continue;
            const float* di = dimage + 3 * (size_t)idx;
            for (uint32_t j = 0; j < p.spp; j++) {
                if (!ok[j]) continue;
                const V3<R> Lt = Ls[j];
                R a[3];
                const R Lc[3] = {Lt.x, Lt.y, Lt.z};
                for (int c = 0; c < 3; c++) a[c] = (Lc[c] >= 0 && Lc[c] <= 10) ? (R)di[c] / (R)n : (R)0;
                if (a[0] == 0 && a[1] == 0 && a[2] == 0) continue;
                auto hook = [&](const VertexRec<R>& rec) {
                    V3<R> dfs[6], dfh[6], dfl[6];
                    DisneyEvalGrad<R>(rec.d, rec.hd, rec.bdir, dfs);
                    DisneyEvalGrad<R>(rec.d, rec.hd, rec.hdir, dfh);
                    for (int q = 0; q < 6; q++) dfl[q] = V3<R>((R)0);
                    if (rec.hasP) DisneyEvalGrad<R>(rec.d, rec.hd, rec.ldir, dfl);
                    V3<R> La = Lt - rec.Lincl;
                    V3<R> cS = rec.T * rec.hd.emission * (rec.w3 * rec.cosS / rec.bpdf);
                    V3<R> cH = rec.visH ? rec.T * rec.hval * (rec.w1 * rec.cosH / rec.hpdf) : V3<R>((R)0);
                    V3<R> cP = (rec.hasP && rec.visP) ? rec.T * rec.lrad * (rec.w2 * rec.cosL / rec.ppdf) : V3<R>((R)0);
                    const R fa[3] = {rec.f.x, rec.f.y, rec.f.z}, la[3] = {La.x, La.y, La.z};
                    R ratio[3];
                    for (int c = 0; c < 3; c++) ratio[c] = fa[c] != 0 ? la[c] / fa[c] : (R)0;
                    const R cs[3] = {cS.x, cS.y, cS.z}, ch[3] = {cH.x, cH.y, cH.z}, cp[3] = {cP.x, cP.y, cP.z};
                    R g[6];
                    for (int q = 0; q < 6; q++) {
                        const R ds[3] = {dfs[q].x, dfs[q].y, dfs[q].z}, dh[3] = {dfh[q].x, dfh[q].y, dfh[q].z},
                                dl[3] = {dfl[q].x, dfl[q].y, dfl[q].z};
                        R acc = 0;
                        for (int c = 0; c < 3; c++) acc += a[c] * ((cs[c] + ratio[c]) * ds[c] + ch[c] * dh[c] + cp[c] * dl[c]);
                        g[q] = acc;
                    }
                    const int mat = rec.sh.material;
                    const el_material& m = v.materials[mat];
                    double* G = gm[tid].data() + (size_t)mat * 6;
                    const R e = (R)2.2f;
                    if (m.albedo_tex < 0) {
                        for (int q = 0; q < 3; q++) G[q] += (double)g[q];
                    } else if (gtex && m.albedo_tex == p.grad_texture) {
                        const Tex& t = *gtex;
                        double* D = gt[tid].data();
                        if (t.filter == 1) {
                            R xx = rec.sh.tu * t.w, yy = rec.sh.tv * t.h;
                            R t1x = std::floor(xx), t1y = std::floor(yy);
                            R fx = xx - t1x, fy = yy - t1y;
                            const int cx[4] = {(int)t1x, (int)t1x + 1, (int)t1x, (int)t1x + 1};
                            const int cy[4] = {(int)t1y, (int)t1y, (int)t1y + 1, (int)t1y + 1};
                            const R wgt[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
                            for (int c4 = 0; c4 < 4; c4++) {
                                size_t ti = texelIndex(t, cx[c4], cy[c4]);
                                for (int q = 0; q < 3; q++) D[3 * ti + q] += (double)(g[q] * wgt[c4]);
                            }
                        } else {
                            size_t ti = texelIndex(t, (int)(rec.sh.tu * t.w), (int)(rec.sh.tv * t.h));
                            for (int q = 0; q < 3; q++) D[3 * ti + q] += (double)g[q];
                        }
                    }
                    if (m.roughness_tex < 0) G[3] += (double)(g[3] * e * std::pow((R)m.roughness, e - 1));
                    if (m.metallic_tex < 0) G[4] += (double)(g[4] * e * std::pow((R)m.metallic, e - 1));
                    G[5] += (double)g[5];
                };
                V3<R> Lr;
                walkPath<R>(s, rp, x, y, p.s0 + j, Lr, rays, verts, hook);
            }
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; t++) pool.emplace_back(work, t);
    for (auto& t : pool) t.join();
    for (int t = 0; t < nthreads; t++) {
        for (size_t i = 0; i < (size_t)M * 6; i++) dmat[i] += gm[t][i];
        if (dtex)
            for (size_t i = 0; i < ntex; i++) dtex[i] += gt[t][i];
    }
}

}  // namespace

extern "C" {

int or_tri_hit(const float* v9, const float* o3, const float* d3, float* t, float* u, float* v) {
    float tt = 0, uu = 0, vv = 0;
    bool h = triHit(v9, o3, d3, tt, uu, vv);
    if (h) {
        *t = tt;
        *u = uu;
        *v = vv;
    }
    return h ? 1 : 0;
}

void or_closest_brute(const float* verts, int n, const float* rays, int nrays, int32_t* tri, float* t, float* u,
                      float* v, int threads) {
    parallelFor(nrays, threads, [&](int i) {
        const float* r = rays + (size_t)i * 6;
        Hit best;
        for (int k = 0; k < n; k++) {
            float tt, uu, vv;
            if (triHit(verts + (size_t)k * 9, r, r + 3, tt, uu, vv)) best.offer(k, tt, uu, vv);
        }
        tri[i] = best.tri;
        t[i] = best.tri >= 0 ? best.t : std::numeric_limits<float>::infinity();
        if (u) u[i] = best.u;
        if (v) v[i] = best.v;
    });
}

void* or_refbvh_build(const float* verts, int n, int depth, int bins) {
    if (depth < 1 || depth > 24 || bins < 2) return nullptr;
    RefBVH* h = new RefBVH();
    h->D = depth;
    h->bins = bins;
    h->n = n;
    h->verts.assign(verts, verts + (size_t)n * 9);
    h->nodes.assign((size_t)2 << depth, RNode());
    h->triIndices.assign((size_t)n, 0);
    std::vector<int> all((size_t)n);
    for (int i = 0; i < n; i++) all[i] = i;
    buildAux(*h, 0, all);
    return h;
}

void or_refbvh_free(void* h) { delete (RefBVH*)h; }

int or_refbvh_node_count(void* h) { return ((RefBVH*)h)->nodeIdx; }

void or_refbvh_closest(void* hv, const float* rays, int nrays, int32_t* tri, float* t, int threads) {
    const RefBVH& h = *(const RefBVH*)hv;
    parallelFor(nrays, threads, [&](int i) {
        const float* r = rays + (size_t)i * 6;
        Hit b = refTraverse(h, r, r + 3);
        tri[i] = b.tri;
        t[i] = b.tri >= 0 ? b.t : std::numeric_limits<float>::infinity();
    });
}

void or_refbvh_any(void* hv, const float* rays, const float* tmax, int nrays, uint8_t* occ, int threads) {
    const RefBVH& h = *(const RefBVH*)hv;
    parallelFor(nrays, threads, [&](int i) {
        const float* r = rays + (size_t)i * 6;
        Hit b = refTraverse(h, r, r + 3);
        occ[i] = (b.tri >= 0 && b.t < tmax[i]) ? 1 : 0;
    });
}

void or_philox(const uint32_t* c, const uint32_t* k, uint32_t* out) { philox(c, k, out); }

static HitData<double> hdD(const double* a) {
    HitData<double> h;
    h.metallic = a[0];
    h.roughness = a[1];
    h.clearcoatGloss = a[2];
    h.clearcoat = a[3];
    h.anisotropic = a[4];
    h.eta = a[5];
    h.transmission = a[6];
    h.specular = a[7];
    h.specularTint = a[8];
    h.sheenTint = a[9];
    h.subsurface = a[10];
    h.sheen = a[11];
    h.emission = V3<double>(a[12], a[13], a[14]);
    h.albedo = V3<double>(a[15], a[16], a[17]);
    h.normal = V3<double>(a[18], a[19], a[20]);
    return h;
}
static HitData<float> hdF(const double* a) {
    HitData<double> d = hdD(a);
    HitData<float> h;
    h.metallic = (float)d.metallic;
    h.roughness = (float)d.roughness;
    h.clearcoatGloss = (float)d.clearcoatGloss;
    h.clearcoat = (float)d.clearcoat;
    h.anisotropic = (float)d.anisotropic;
    h.eta = (float)d.eta;
    h.transmission = (float)d.transmission;
    h.specular = (float)d.specular;
    h.specularTint = (float)d.specularTint;
    h.sheenTint = (float)d.sheenTint;
    h.subsurface = (float)d.subsurface;
    h.sheen = (float)d.sheen;
    h.emission = V3<float>((float)a[12], (float)a[13], (float)a[14]);
    h.albedo = V3<float>((float)a[15], (float)a[16], (float)a[17]);
    h.normal = V3<float>((float)a[18], (float)a[19], (float)a[20]);
    return h;
}

void or_disney_eval(const double* hd, const double* dir, const double* L, int precision, double* out) {
    if (precision == 1) {
        V3<double> r = DisneyEval<double>(V3<double>(dir[0], dir[1], dir[2]), hdD(hd), V3<double>(L[0], L[1], L[2]));
        out[0] = r.x;
        out[1] = r.y;
        out[2] = r.z;
    } else {
        V3<float> r = DisneyEval<float>(V3<float>((float)dir[0], (float)dir[1], (float)dir[2]), hdF(hd),
                                        V3<float>((float)L[0], (float)L[1], (float)L[2]));
        out[0] = r.x;
        out[1] = r.y;
        out[2] = r.z;
    }
}

double or_disney_pdf(const double* hd, const double* dir, const double* L, int precision) {
    if (precision == 1) return DisneyPdf<double>(V3<double>(dir[0], dir[1], dir[2]), hdD(hd), V3<double>(L[0], L[1], L[2]));
    return DisneyPdf<float>(V3<float>((float)dir[0], (float)dir[1], (float)dir[2]), hdF(hd),
                            V3<float>((float)L[0], (float)L[1], (float)L[2]));
}

void or_disney_sample(const double* hd, const double* dir, double r1, double r2, double r3, int precision, double* out) {
    if (precision == 1) {
        V3<double> r = DisneySample<double>(V3<double>(dir[0], dir[1], dir[2]), hdD(hd), r1, r2, r3);
        out[0] = r.x;
        out[1] = r.y;
        out[2] = r.z;
    } else {
        V3<float> r = DisneySample<float>(V3<float>((float)dir[0], (float)dir[1], (float)dir[2]), hdF(hd), (float)r1,
                                          (float)r2, (float)r3);
        out[0] = r.x;
        out[1] = r.y;
        out[2] = r.z;
    }
}

void or_cosine_sample_hemisphere(double u1, double u2, double* out) {
    V3<double> r = CosineSampleHemisphere<double>(u1, u2);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
void or_importance_sample_ggx(double rgh, double r1, double r2, double* out) {
    V3<double> r = ImportanceSampleGGX<double>(rgh, r1, r2);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
void or_uniform_circle(double u1, double u2, double u3, double* x, double* y) { uniformCircleSampling<double>(u1, u2, u3, *x, *y); }

void or_hdri_cdf(const float* rgb, int w, int h, float* cdf, float* radianceSum) {   // HDRI.hpp:107-128
    float sum = 0;
    size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; i++) sum += rgb[3 * i] + rgb[3 * i + 1] + rgb[3 * i + 2];
    cdf[0] = 0;
    for (size_t c = 0; c < n; c++) cdf[c + 1] = cdf[c] + (rgb[3 * c] + rgb[3 * c + 1] + rgb[3 * c + 2]) / sum;
    *radianceSum = sum;
}

int or_hdri_search(const float* cdf, float r, int length) { return binarySearch(cdf, r, length); }

float or_hdri_pdf(const float* rgb, int w, int h, float xOffset, float radianceSum, int x, int y) {
    Tex t;
    t.w = w;
    t.h = h;
    t.data = rgb;
    t.xOffset = xOffset;
    V3<float> dv = texel<float>(t, x, y);
    float theta = (((float)y / (float)h)) * PI_F;
    return ((dv.x + dv.y + dv.z) / radianceSum) * w * h / (2.0 * PI_F * std::sin(theta));
}

void or_camera_ray(const el_camera* cam, int x, int y, const float* r5, float* o3, float* d3) {
    const el_camera& c = *cam;   // kernel.cu:260-337 in FP32
    float dx = c.position.x + ((float)x) / ((float)c.x_res) * c.sensor_width;
    float dy = c.position.y + ((float)y) / ((float)c.y_res) * c.sensor_height;
    float odx = (-c.sensor_width / 2.0) + dx;
    float ody = (-c.sensor_height / 2.0) + dy;
    float rx = (1.0 / (float)c.x_res) * (r5[0] - 0.5) * c.sensor_width;
    float ry = (1.0 / (float)c.y_res) * (r5[1] - 0.5) * c.sensor_height;
    float SPx = odx + rx, SPy = ody + ry, SPz = c.position.z + c.focal_length;
    V3<float> rot(c.rotation.x, c.rotation.y, c.rotation.z);
    rot = rot * (float)(PI_F / 180.0);
    V3<float> pos(c.position.x, c.position.y, c.position.z);
    V3<float> dir = V3<float>(SPx, SPy, SPz) - pos;
    V3<float> dX(dir.x, dir.y * std::cos(rot.x) - dir.z * std::sin(rot.x), dir.y * std::sin(rot.x) + dir.z * std::cos(rot.x));
    V3<float> dY(dX.x * std::cos(rot.y) + dX.z * std::sin(rot.y), dX.y, dX.z * std::cos(rot.y) - dX.x * std::sin(rot.y));
    V3<float> dZ(dY.x * std::cos(rot.z) - dY.y * std::sin(rot.z), dY.x * std::sin(rot.z) + dY.y * std::cos(rot.z), dY.z);
    V3<float> o = pos, d = dZ.normalized();
    if (c.bokeh) {
        float diameter = c.focal_length / c.aperture;
        float l = c.focus_distance + c.focal_length;
        V3<float> focus = o + d * l;
        float px, py;
        uniformCircleSampling<float>(r5[2], r5[3], r5[4], px, py);
        px *= diameter * 0.5;
        py *= diameter * 0.5;
        o = pos + V3<float>(px, py, 0.f);
        d = (focus - o).normalized();
    }
    o3[0] = o.x;
    o3[1] = o.y;
    o3[2] = o.z;
    d3[0] = d.x;
    d3[1] = d.y;
    d3[2] = d.z;
}

int or_render(const el_scene_view* v, void* refbvh, const or_render_params* p, double* sum, double* sumsq,
              uint32_t* count, uint64_t* rays, uint64_t* vertices) {
    if (!v || !refbvh || !p || !sum || !sumsq || !count) return -1;
    if (p->x0 < 0 || p->y0 < 0 || p->x1 > (int)v->camera.x_res || p->y1 > (int)v->camera.y_res || p->x0 >= p->x1 ||
        p->y0 >= p->y1)
        return -1;
    SceneC s;
    s.v = v;
    s.bvh = (const RefBVH*)refbvh;
    s.env = texOf(v->hdri);
    for (uint32_t i = 0; i < v->texture_count; i++) s.texs.push_back(texOf(v->textures[i]));
    // HDRI::generateCDF at construction: offsets not yet applied
    size_t L = (size_t)v->hdri.width * v->hdri.height;
    s.cdf.assign(L + 1, 0.f);
    or_hdri_cdf(v->hdri.data, v->hdri.width, v->hdri.height, s.cdf.data(), &s.radianceSum);
    s.evalMats = p->eval_materials;
    std::atomic<uint64_t> r{0}, vv{0};
    if (p->precision == 1)
        renderT<double>(s, *p, sum, sumsq, count, r, vv);
    else
        renderT<float>(s, *p, sum, sumsq, count, r, vv);
    if (rays) *rays = r.load();
    if (vertices) *vertices = vv.load();
    return 0;
}

}  // extern "C"

extern "C" {

void or_disney_eval_grad(const double* hd, const double* dir, const double* L, double* out18) {
    V3<double> df[6];
    DisneyEvalGrad<double>(V3<double>(dir[0], dir[1], dir[2]), hdD(hd), V3<double>(L[0], L[1], L[2]), df);
    for (int k = 0; k < 6; k++) {
        out18[3 * k] = df[k].x;
        out18[3 * k + 1] = df[k].y;
        out18[3 * k + 2] = df[k].z;
    }
}

int or_adjoint(const el_scene_view* v, void* refbvh, const or_adjoint_params* p, const float* dimage, double* dmat,
               double* dtex, double* image) {
    if (!v || !refbvh || !p || !dimage || !dmat) return -1;
    SceneC s;
    s.v = v;
    s.bvh = (const RefBVH*)refbvh;
    s.env = texOf(v->hdri);
    for (uint32_t i = 0; i < v->texture_count; i++) s.texs.push_back(texOf(v->textures[i]));
    size_t L = (size_t)v->hdri.width * v->hdri.height;
    s.cdf.assign(L + 1, 0.f);
    or_hdri_cdf(v->hdri.data, v->hdri.width, v->hdri.height, s.cdf.data(), &s.radianceSum);
    or_render_params rp{};
    rp.spp = p->spp;
    rp.s0 = p->s0;
    rp.depth = p->depth;
    rp.seed = p->seed;
    rp.precision = p->precision;
    rp.shadow_mode = p->shadow_mode;
    rp.threads = p->threads;
    rp.x0 = 0;
    rp.x1 = (int)v->camera.x_res;
    rp.y0 = 0;
    rp.y1 = (int)v->camera.y_res;
    if (p->precision == 1)
        adjointT<double>(s, rp, *p, dimage, dmat, dtex, image);
    else
        adjointT<float>(s, rp, *p, dimage, dmat, dtex, image);
    return 0;
}

}  // extern "C"

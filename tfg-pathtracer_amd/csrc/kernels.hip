// HIP kernels of the path-tracing core for gfx950 (MI355X, wave64).
//
//   k_trace        closest / any hit for a ray batch (parity + traversal bench)
//   k_plain        one thread per pixel, full path per thread: the plain
//                  megakernel, same structure as renderingKernel
//                  (kernel.cu:369-481) with per-sample launches
//                  (kernel.cu:675-683)
//   k_wf_*         wavefront: camera -> extend -> shade -> shadow x2 ->
//                  finalize over per-bounce ray queues
//
// Queues and tallies avoid same-address atomics across the chip: MI355X has 8
// XCDs with private L2s, and workgroups are dispatched to them round-robin,
// so blockIdx % 8 names the XCD.  Every queue is split into 8 segments, one
// per XCD; a block appends all its entries of one loop iteration with a single
// atomic on its XCD's counter (ballot + LDS prefix), and consumers address the
// 8 segments as one dense index range.  Tallies are reduced per block and
// added to per-XCD copies of the counters.
//
// All traversal kernels are grid-stride loops over a grid no larger than the
// device's resident-thread capacity, so the per-thread global spill area of
// the traversal stack is sized once per context.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "pt_device.hpp"

namespace eleven {
namespace dev {

constexpr int kMaxWaves = kMaxBlock / 64;

// Minimum waves per SIMD the register allocator must leave room for, per
// kernel family (1 = no constraint).  Traversal is latency-bound and gains
// from occupancy; the plain megakernel trades spills against occupancy.
#ifndef EL_WPE_TRAV
#define EL_WPE_TRAV 8
#endif
#ifndef EL_WPE_SHADE
#define EL_WPE_SHADE 3
#endif
#ifndef EL_WPE_PLAIN
#define EL_WPE_PLAIN 4
#endif
#ifndef EL_WPE_ADJ
#define EL_WPE_ADJ 4
#endif
#ifndef EL_WPE_ADJ_REPLAY   // the adjoint over recorded vertices (no traversal)
#define EL_WPE_ADJ_REPLAY 3
#endif
#define EL_KERNEL(wpe) __global__ __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(wpe)))

__device__ __forceinline__ uint32_t waveSumU32(uint32_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ uint32_t xcdOf() { return blockIdx.x % kXcds; }

// Block-wide sum of N per-thread tallies, one atomic per tally per block on
// the block's XCD copy of the counters.  Every thread of the block must call.
template <int N>
__device__ __forceinline__ void blockFlush(unsigned long long* counters, const int (&slot)[N],
                                          const uint32_t (&v)[N]) {
    __shared__ uint32_t red[kMaxWaves][N];
    const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
    for (int k = 0; k < N; k++) {
        uint32_t s = waveSumU32(v[k]);
        if (__lane_id() == 0) red[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        unsigned long long s = 0;
        for (int w = 0; w < nw; w++) s += red[w][threadIdx.x];
        if (s) atomicAdd(&counters[xcdOf() * kStatSlots + slot[threadIdx.x]], s);
    }
}

// Chunked work of a traversal launch over n queue items (counter set `kind`):
// chunks of kChunkRays, smaller (down to 64) when the queue would not give
// every wave of the grid eight of them.
template <bool WIDE>
__device__ __forceinline__ WorkSource chunkSource(uint32_t* qcounters, int kind, uint32_t n) {
    WorkSource ws;
    ws.begin = ws.end = 0;
    ws.ctr = qcounters + kQChunk + kind * kChunkShares * kChunkStride;
    ws.n = n;
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t fit = (uint32_t)(((uint64_t)n / ((uint64_t)waves * 8)) & ~63ull);
    ws.chunk = fit < 64 ? 64u : (fit > kChunkRays ? kChunkRays : fit);
    ws.stride = kChunkStride;
    ws.chunks = (uint32_t)(((uint64_t)n + ws.chunk - 1) / ws.chunk);
    static_assert(kChunkShares == (int)kWorkShares, "one chunk share per lane");
    ws.parts = WIDE ? kChunkShares : kXcds;
    ws.home = blockIdx.x % ws.parts;
    ws.fixed = ws.chunks <= waves;
    return ws;
}
// Zeroes the chunk counters of `kind` (threads 0..kChunkShares-1 of block 0).
__device__ __forceinline__ void resetChunks(uint32_t* qcounters, int kind) {
    if (blockIdx.x == 0 && threadIdx.x < kChunkShares)
        qcounters[kQChunk + (kind * kChunkShares + threadIdx.x) * kChunkStride] = 0;
}

// Dense index i over an XCD-segmented queue -> entry index.
struct QueueMap {
    uint32_t pre[kXcds + 1];
    __device__ __forceinline__ void load(const uint32_t* cnt) {
        pre[0] = 0;
#pragma unroll
        for (int k = 0; k < kXcds; k++) pre[k + 1] = pre[k] + cnt[k];
    }
    __device__ __forceinline__ uint32_t total() const { return pre[kXcds]; }
    __device__ __forceinline__ uint64_t entry(uint32_t i, uint32_t segSize) const {
        // selects only: a dynamic index into pre[] would put it in scratch
        uint32_t s = 0, base = 0;
#pragma unroll
        for (int k = 1; k < kXcds; k++) {
            bool past = i >= pre[k];
            s = past ? (uint32_t)k : s;
            base = past ? pre[k] : base;
        }
        return (uint64_t)s * segSize + (i - base);
    }
};

// Appends up to three entries per thread (one per queue) with one atomic per
// queue per block.  Returns entry indices (segment-relative index checked
// against segSize; ~0 if not wanted or if the segment would overflow).
// Every thread of the block must call.
__device__ __forceinline__ void blockAlloc3(const bool (&want)[3], uint32_t* const (&cnt)[3], uint32_t segSize,
                                           uint64_t (&out)[3], uint32_t& overflow) {
    __shared__ uint32_t s_cnt[kMaxWaves][3];
    const int wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6, lane = __lane_id();
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    uint32_t pos[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        unsigned long long m = __ballot(want[k]);
        pos[k] = (uint32_t)__popcll(m & lt);
        if (lane == 0) s_cnt[wave][k] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int k = threadIdx.x;
        uint32_t total = 0;
        for (int w = 0; w < nw; w++) total += s_cnt[w][k];
        uint32_t base = total ? atomicAdd(cnt[k] + xcdOf(), total) : 0u;
        for (int w = 0; w < nw; w++) {
            uint32_t c = s_cnt[w][k];
            s_cnt[w][k] = base;
            base += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; k++) {
        uint32_t idx = s_cnt[wave][k] + pos[k];
        bool ok = want[k] && idx < segSize;
        if (want[k] && !ok) overflow = 1;
        out[k] = ok ? (uint64_t)xcdOf() * segSize + idx : ~0ull;
    }
    __syncthreads();   // s_cnt is reused by the next call
}

__device__ __forceinline__ TravStack makeStack(const DevScene& s, int* gstack, uint32_t gthreads) {
    extern __shared__ int lds[];
    TravStack st;
    st.lds = lds + threadIdx.x;
    st.ldsStride = blockDim.x;
    uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    st.g = gstack + gtid;
    st.gstride = (int)gthreads;
    st.gcap = s.gcap;
    st.sp = st.lo = 0;
    st.overflow = false;
    return st;
}

// ------------------------------------------------------------------ trace --
// Items [begin, end) of the calling wave: n items split evenly over the
// grid's waves (block sizes are multiples of 64).
// Items [begin, end) of the calling wave: n items in chunks of a multiple of
// 64 over the grid's waves (block sizes are multiples of 64), so a small
// batch fills few waves completely rather than many waves with one lane.
__device__ __forceinline__ void waveRange(uint32_t n, uint32_t& begin, uint32_t& end) {
    const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
    const uint32_t w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t chunk = (((uint64_t)n + waves - 1) / waves + 63) & ~63ull;
    const uint64_t b = chunk * w, e = b + chunk;
    begin = (uint32_t)(b < n ? b : n);
    end = (uint32_t)(e < n ? e : n);
}

template <int W, bool ANY, bool COUNT>
EL_KERNEL(EL_WPE_TRAV) void k_trace(DevScene s, const float* __restrict__ rays, const float* __restrict__ tmax,
                                    uint32_t n, int32_t* __restrict__ outTri, float* __restrict__ outT,
                                    float* __restrict__ outUV, uint8_t* __restrict__ outOcc, int* gstack,
                                    uint32_t gthreads, uint32_t refill, unsigned long long* counters) {
    TravStack st = makeStack(s, gstack, gthreads);
    TraceCounters cnt{0, 0};
    uint32_t traced = 0;
    WorkSource ws{};
    waveRange(n, ws.begin, ws.end);
    auto load = [&](uint32_t i, float& ox, float& oy, float& oz, float& dx, float& dy, float& dz, float& tm,
                    BvhPtr&) {
        const float* r = rays + (uint64_t)i * 6;
        ox = r[0];
        oy = r[1];
        oz = r[2];
        dx = r[3];
        dy = r[4];
        dz = r[5];
        tm = ANY ? tmax[i] : __int_as_float(0x7f800000);
        return true;
    };
    auto store = [&](uint32_t i, const TraceResult& h, bool) {
        if (ANY) {
            outOcc[i] = h.tri >= 0 ? 1 : 0;
        } else {
            outTri[i] = h.tri;
            outT[i] = h.tri >= 0 ? h.t : __int_as_float(0x7f800000);
            if (outUV) {
                outUV[2 * (uint64_t)i] = h.u;
                outUV[2 * (uint64_t)i + 1] = h.v;
            }
        }
    };
    traceWaveRange<W, ANY, COUNT>(s, st, ws, refill, load, store, cnt, traced);
    const int slots[3] = {kStatNodes, kStatTris, kStatOverflow};
    const uint32_t vals[3] = {cnt.nodes, cnt.tris, (uint32_t)st.overflow};
    blockFlush<3>(counters, slots, vals);
}

// --------------------------------------------------------- shared shading --
struct VertexOut {
    v3 base;        // T * (w1*0 + w2*0 + w3*B): added unconditionally
    v3 contribH;    // T * w1 * H_unoccluded: added if the HDRI shadow ray is clear
    v3 contribP;    // T * w2 * P_unoccluded: added if the point shadow ray is clear
    v3 hOrig, hDir; // HDRI shadow ray
    v3 pOrig, pDir; // point-light shadow ray
    float pTmax;
    bool hasP;
    v3 newT;        // throughput after the bounce
    v3 nextO, nextD;
    // primal terms the adjoint replay reuses
    v3 bdir, f;     // sampled direction, DisneyEval there
    float bpdf, w1, w2, w3;
    v3 hdir, hval;  // HDRI sample direction and texel
    float hpdf;
    v3 ldir, lrad;  // light direction, radiance / dist^2
    float ppdf;
};

// One path vertex: DisneySample, DisneyEval/Pdf, HDRI and point-light NEE
// terms and the balance-heuristic combine of shade() (kernel.cu:339-358),
// with the shadow queries left to the caller.  r = the bounce's Philox block.
__device__ __forceinline__ VertexOut shadeVertex(const DevScene& s, v3 d, v3 T, const SurfaceHit& sh,
                                                 const HitData& hd, const uint32_t* r) {
    VertexOut o;
    v3 bdir = disneySample(d, hd, u01(r[0]), u01(r[1]), u01(r[2]));
    float bpdf;
    v3 f = disneyEvalPdf(d, hd, bdir, bpdf);
    float r1 = u01(r[3]);

    EnvSample es = envSample(s, r1);
    v3 H = disneyEval(d, hd, es.dir) * fabsf(dot(es.dir, hd.normal)) * es.value * frcp(es.pdf);
    float hpdf = es.pdf;
    o.hOrig = sh.position + es.dir * 0.001f;
    o.hDir = es.shadowDir;

    float ppdf = 0.f;
    v3 P = mk(0, 0, 0), ldir = mk(0, 0, 1), lrad = mk(0, 0, 0);
    o.hasP = s.nLights > 0;
    if (o.hasP) {
        ppdf = ((float)s.nLights) / (float)(2.0 * EL_PI);
        int li = (int)(s.nLights * r1);
        li = li < s.nLights ? li : s.nLights - 1;
        v3 lp = ld3(s.lights + 6 * li), lr = ld3(s.lights + 6 * li + 3);
        v3 nd = normalized(lp - sh.position);
        float dist = length(lp - sh.position);
        o.pOrig = sh.position + nd * 0.001f;
        o.pDir = normalized(nd);
        o.pTmax = dist - 0.001f;
        lrad = lr * frcp(dist * dist);
        P = lrad * disneyEval(d, hd, nd) * fabsf(dot(nd, hd.normal)) * frcp(ppdf);
        ldir = nd;
    }
    o.ldir = ldir;
    o.lrad = lrad;
    v3 fc = f * fabsf(dot(bdir, hd.normal));
    float rb = frcp(bpdf);
    v3 B = hd.emission * fc * rb;
    float rs = frcp(hpdf + ppdf + bpdf);
    float w1 = hpdf * rs, w2 = ppdf * rs, w3 = bpdf * rs;
    o.base = T * (w1 * mk(0, 0, 0) + w2 * mk(0, 0, 0) + w3 * B);
    o.contribH = T * (w1 * H);
    o.contribP = T * (w2 * P);
    o.newT = T * (fc * rb);
    o.nextO = sh.position + bdir * 0.001f;
    o.nextD = normalized(bdir);
    o.bdir = bdir;
    o.f = f;
    o.bpdf = bpdf;
    o.w1 = w1;
    o.w2 = w2;
    o.w3 = w3;
    o.hdir = es.dir;
    o.hval = es.value;
    o.hpdf = hpdf;
    o.ppdf = ppdf;
    return o;
}

__device__ __forceinline__ bool anyNan(v3 a) { return isnan(a.x) || isnan(a.y) || isnan(a.z); }

// A path whose throughput is exactly zero adds nothing more to its pixel --
// unless a later vertex makes its radiance NaN, which rejects the whole sample
// (kernel.cu:449).  With zero throughput the only such term (outside events
// of measure zero) is the HDRI one: 0 * f * env / pdf is NaN when the sampled
// environment pixel has an infinite pdf (pole row) or a zero one, and which
// pixel a bounce samples depends on its random number alone.  True if no
// remaining bounce's number falls in a bucket that can reach such a pixel
// (s.envBad, conservative): stopping the path then leaves the estimate and
// the accepted-sample count exactly as if it were traced.
__device__ __forceinline__ bool quietDeath(const DevScene& s, const PathRng& rng, int fromBounce, int maxDepth) {
    for (int b = fromBounce; b < maxDepth; b++) {
        const uint32_t bucket = rng.block(2 + b).w >> 16;
        if ((s.envBad[bucket >> 5] >> (bucket & 31)) & 1u) return false;
    }
    return true;
}

// Bit b set: bounce b's environment random number falls in a bucket that can
// reach a pdf-0 / pdf-inf pixel (see quietDeath).  A path that dies at
// bounce k < maxDepth - 1 stops quietly iff (mask >> (k + 1)) == 0; maxDepth
// <= 64, so every bounce has its bit and the shift stays below 64.  Bit 0
// is never read (a death is checked from the next bounce on): no Philox
// block is drawn for it.
__device__ __forceinline__ uint64_t nanProneBounces(const DevScene& s, const PathRng& rng, int maxDepth) {
    uint64_t m = 0;
    for (int b = 1; b < maxDepth && b < 64; b++) {
        const uint32_t bucket = rng.block(2 + b).w >> 16;
        m |= (uint64_t)((s.envBad[bucket >> 5] >> (bucket & 31)) & 1u) << b;
    }
    return m;
}

__device__ __forceinline__ bool isZero(v3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
__device__ __forceinline__ v3 clamp3(v3 a) {
    return mk(clampf(a.x, 0.f, 10.f), clampf(a.y, 0.f, 10.f), clampf(a.z, 0.f, 10.f));
}

// ------------------------------------------------------------------ plain --
// What one path vertex of walkPath computed, for hooks (AOV capture, adjoint).
struct PathVertex {
    int depth;
    v3 d, T;             // incoming direction, throughput before the vertex
    const SurfaceHit* sh;
    const HitData* hd;
    const VertexOut* vo;
    bool visH, visP;
    v3 Lincl;            // radiance gathered up to and including this vertex
};

// Where a walk's closest hits and shadow answers come from.
//   kWalkTrace      traced.
//   kWalkTraceKeep  traced; vertex i's closest hit (tri, t, u, v; tri -1:
//                   escaped) is also written to io.keep[i * io.keepStride]
//                   and the shadow answers to io.visOut (bit 2i HDRI, bit
//                   2i + 1 point light, set when the light is visible).
//   kWalkReplay     read from io.hits[i * io.stride] and io.vis, the record of
//                   an earlier walk of the same path (a recording wavefront
//                   render, or a kWalkTraceKeep walk); nothing is traced.  The
//                   replay stops where the recorded walk did (both cull alike),
//                   so every record read was written (a triangle index out of
//                   range reads as an escape).
constexpr int kWalkTrace = 0, kWalkTraceKeep = 1, kWalkReplay = 2;
struct WalkIO {
    const float4* hits;   // kWalkReplay
    uint64_t stride;
    uint32_t vis;
    float4* keep;         // kWalkTraceKeep
    uint64_t keepStride;
    uint32_t visOut;
};

// One path of renderingKernel (kernel.cu:369-481) for pixel (x, y), sample
// `sample`: camera ray, up to maxDepth vertices with DisneySample / shade and
// both shadow queries, escape to the environment.  Returns the radiance
// before the clamp; `hook` sees every shaded vertex; `verts` counts them.
template <int W, int MODE = kWalkTrace, class Hook>
__device__ __forceinline__ v3 walkPath(const DevScene& s, const RenderArgs& a, uint32_t x, uint32_t y,
                                       uint32_t idx, uint32_t sample, TravStack& st, TraceCounters& cnt,
                                       uint32_t& rays, uint32_t& verts, Hook&& hook, WalkIO& io) {
    constexpr bool REPLAY = MODE == kWalkReplay, KEEP = MODE == kWalkTraceKeep;
    PathRng rng{a.k0, a.k1, idx, sample, a.scene};
    const uint32_t vis = REPLAY ? io.vis : 0u;
    uint32_t visOut = 0;
    uint4 c = rng.block(0);
    uint32_t c5 = rng.block(1).x;
    v3 o, d;
    cameraRay(s.cam, (int)x, (int)y, u01(c.x), u01(c.y), u01(c.z), u01(c.w), u01(c5), o, d);
    v3 L = mk(0, 0, 0), T = mk(1, 1, 1);
    int i;
    for (i = 0; i < (int)a.maxDepth; i++) {
        TraceResult h;
        if constexpr (REPLAY) {
            const float4 r = io.hits[(uint64_t)i * io.stride];
            h.tri = __float_as_int(r.x);
            h.t = r.y;
            h.u = r.z;
            h.v = r.w;
            if (h.tri >= s.nTris) h.tri = -1;
        } else {
            rays++;
            h = traceRay<W, false, false>(s, o.x, o.y, o.z, d.x, d.y, d.z, __int_as_float(0x7f800000), st, cnt);
            if constexpr (KEEP) io.keep[(uint64_t)i * io.keepStride] = make_float4(__int_as_float(h.tri), h.t, h.u, h.v);
        }
        if (h.tri < 0) {
            L = L + envLookup(s, d) * T;
            break;
        }
        SurfaceHit sh = surfaceAt(s, h.tri, h.t, h.u, h.v, o, d);
        HitData hd = hitData(s, sh);
        uint4 rb = rng.block(2 + i);
        uint32_t r[4] = {rb.x, rb.y, rb.z, rb.w};
        VertexOut vo = shadeVertex(s, d, T, sh, hd, r);
        bool visH;
        if constexpr (REPLAY) {
            visH = (vis >> (2 * i)) & 1u;
        } else {
            rays++;
            visH = traceRay<W, true, false>(s, vo.hOrig.x, vo.hOrig.y, vo.hOrig.z, vo.hDir.x, vo.hDir.y, vo.hDir.z,
                                            __int_as_float(0x7f800000), st, cnt).tri < 0;
        }
        v3 add = vo.base;
        if (visH) add = add + vo.contribH;
        bool visP = false;
        if (vo.hasP) {
            if constexpr (REPLAY) {
                visP = (vis >> (2 * i + 1)) & 1u;
            } else {
                rays++;
                visP = traceRay<W, true, false>(s, vo.pOrig.x, vo.pOrig.y, vo.pOrig.z, vo.pDir.x, vo.pDir.y,
                                                vo.pDir.z, vo.pTmax, st, cnt).tri < 0;
            }
            if (visP) add = add + vo.contribP;
        }
        if (KEEP) visOut |= (visH ? 1u : 0u) << (2 * i) | (visP ? 1u : 0u) << (2 * i + 1);
        L = L + add;
        hook(PathVertex{i, d, T, &sh, &hd, &vo, visH, visP, L});
        T = vo.newT;
        o = vo.nextO;
        d = vo.nextD;
        if (a.cull && isZero(T) && quietDeath(s, rng, i + 1, (int)a.maxDepth)) {
            i++;
            break;
        }
    }
    if (KEEP) io.visOut = visOut;
    verts += (uint32_t)i;
    return L;
}
template <int W, class Hook>
__device__ __forceinline__ v3 walkPath(const DevScene& s, const RenderArgs& a, uint32_t x, uint32_t y,
                                       uint32_t idx, uint32_t sample, TravStack& st, TraceCounters& cnt,
                                       uint32_t& rays, uint32_t& verts, Hook&& hook) {
    WalkIO io{};
    return walkPath<W, kWalkTrace>(s, a, x, y, idx, sample, st, cnt, rays, verts, hook, io);
}

template <int W>
EL_KERNEL(EL_WPE_PLAIN) void k_plain(DevScene s, RenderArgs a, float4* __restrict__ accum, uint32_t* __restrict__ counts,
                        float4* __restrict__ aov, int* gstack, uint32_t gthreads,
                        unsigned long long* counters) {
    TravStack st = makeStack(s, gstack, gthreads);
    TraceCounters cnt{0, 0};
    const uint32_t W_ = s.cam.xRes, H_ = s.cam.yRes;
    const uint32_t tilesX = (W_ + 15) / 16;
    uint32_t rays = 0, verts = 0, paths = 0, overflow = 0;
    uint32_t stride = gridDim.x * blockDim.x;
    uint32_t total = tilesX * ((H_ + 15) / 16) * 256;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < total; w += stride) {
        uint32_t tile = w >> 8, in = w & 255;
        uint32_t x = (tile % tilesX) * 16 + (in & 15), y = (tile / tilesX) * 16 + (in >> 4);
        if (x >= W_ || y >= H_) continue;
        uint32_t idx = W_ * (H_ - y - 1) + x;
        v3 an = mk(0, 0, 0), at = mk(0, 0, 0), ab = mk(0, 0, 0);
        paths++;
        v3 L = walkPath<W>(s, a, x, y, idx, a.sample, st, cnt, rays, verts, [&](const PathVertex& pv) {
            if (pv.depth == 0 && aov) {
                an = pv.sh->normal;
                surfaceTangents(s, *pv.sh, at, ab);
            }
        });
        overflow |= st.overflow;
        if (a.rec) a.rec[(uint64_t)a.sample * W_ * H_ + idx] = make_float4(L.x, L.y, L.z, 1.f);
        L = clamp3(L);
        if (!anyNan(L)) {
            float4 acc = accum[idx];
            accum[idx] = make_float4(acc.x + L.x, acc.y + L.y, acc.z + L.z, acc.w);
            counts[idx] += 1;
            if (aov) {
                float4* A = aov + 3 * (uint64_t)idx;
                A[0] = make_float4(A[0].x + an.x, A[0].y + an.y, A[0].z + an.z, 0);
                A[1] = make_float4(A[1].x + at.x, A[1].y + at.y, A[1].z + at.z, 0);
                A[2] = make_float4(A[2].x + ab.x, A[2].y + ab.y, A[2].z + ab.z, 0);
            }
        }
    }
    const int slots[4] = {kStatRays, kStatVerts, kStatPaths, kStatOverflow};
    const uint32_t vals[4] = {rays, verts, paths, overflow};
    blockFlush<4>(counters, slots, vals);
}

// ---------------------------------------------------------------- adjoint --
// Gradient of one path vertex (the bracket in the comment of k_adjoint
// below): the six parameters of the vertex's material in g (albedo zeroed
// when it comes from a texture: those go to the texels of gradTex with FP32
// atomics, through the bilinear weights when the texture filters).  d: incoming direction, T: throughput
// before the vertex, La: radiance the path gathers after the vertex.
// Returns the material index.
__device__ __forceinline__ int vertexGradient(const DevScene& s, v3 d, v3 T, const SurfaceHit& sh,
                                              const HitData& hd, const VertexOut& vo, bool visH, bool visP, v3 La,
                                              const float ad[3], int gradTex, float* __restrict__ dtex, float g[6]) {
    const DevTex* gt = gradTex >= 0 ? &s.texs[gradTex] : nullptr;
    v3 dfs[6], dfh[6], dfl[6];
    disneyEvalGrad(d, hd, vo.bdir, dfs);
    disneyEvalGrad(d, hd, vo.hdir, dfh);
    if (vo.hasP)
        disneyEvalGrad(d, hd, vo.ldir, dfl);
    else
#pragma unroll
        for (int q = 0; q < 6; q++) dfl[q] = mk(0, 0, 0);
    const float cosS = fabsf(dot(vo.bdir, hd.normal));
    const v3 cS = T * hd.emission * (vo.w3 * cosS / vo.bpdf);
    const v3 cH = visH ? T * vo.hval * (vo.w1 * fabsf(dot(vo.hdir, hd.normal)) / vo.hpdf) : mk(0, 0, 0);
    const v3 cP = (vo.hasP && visP) ? T * vo.lrad * (vo.w2 * fabsf(dot(vo.ldir, hd.normal)) / vo.ppdf)
                                    : mk(0, 0, 0);
    const v3 ratio = mk(vo.f.x != 0.f ? La.x / vo.f.x : 0.f, vo.f.y != 0.f ? La.y / vo.f.y : 0.f,
                        vo.f.z != 0.f ? La.z / vo.f.z : 0.f);
    const v3 kS = (cS + ratio) * mk(ad[0], ad[1], ad[2]);
    const v3 kH = cH * mk(ad[0], ad[1], ad[2]);
    const v3 kP = cP * mk(ad[0], ad[1], ad[2]);
#pragma unroll
    for (int q = 0; q < 6; q++) g[q] = dot(kS, dfs[q]) + dot(kH, dfh[q]) + dot(kP, dfl[q]);
    const int mat = sh.material;
    const DevMaterial& m = s.mats[mat];
    if (m.albedoTex >= 0) {
        if (gt && m.albedoTex == gradTex) {
            const float tu = sh.tu, tv = sh.tv;
            if (gt->filter == 1) {
                const float xx = tu * gt->w, yy = tv * gt->h;
                const float t1x = floorf(xx), t1y = floorf(yy);
                const float fx = xx - t1x, fy = yy - t1y;
                const int cx[4] = {(int)t1x, (int)t1x + 1, (int)t1x, (int)t1x + 1};
                const int cy[4] = {(int)t1y, (int)t1y, (int)t1y + 1, (int)t1y + 1};
                const float wg[4] = {(1 - fx) * (1 - fy), fx * (1 - fy), (1 - fx) * fy, fx * fy};
                for (int k4 = 0; k4 < 4; k4++) {
                    const uint64_t ti = 3 * (uint64_t)texelIndex(*gt, cx[k4], cy[k4]);
                    for (int q = 0; q < 3; q++) atomicAdd(dtex + ti + q, g[q] * wg[k4]);
                }
            } else {
                const uint64_t ti = 3 * (uint64_t)texelIndex(*gt, (int)(tu * gt->w), (int)(tv * gt->h));
                for (int q = 0; q < 3; q++) atomicAdd(dtex + ti + q, g[q]);
            }
        }
        g[0] = g[1] = g[2] = 0.f;
    }
    const float e = 2.2f;
    g[3] = m.roughnessTex < 0 ? g[3] * e * powf(m.roughness, e - 1.0f) : 0.f;
    g[4] = m.metallicTex < 0 ? g[4] * e * powf(m.metallic, e - 1.0f) : 0.f;
    return mat;
}

// Adjoint weight of one path sample: dLoss/dI of its pixel over the pixel's
// accepted samples, per channel, 0 where the [0, 10] clamp is active or the
// sample was rejected (NaN) or the pixel has none.
__device__ __forceinline__ void adjointWeight(v3 Lt, const float* __restrict__ dimage, uint32_t idx, uint32_t n,
                                              float ad[3]) {
    ad[0] = ad[1] = ad[2] = 0.f;
    if (!n || anyNan(clamp3(Lt))) return;
    const float* di = dimage + 3 * (uint64_t)idx;
    const float Lr[3] = {Lt.x, Lt.y, Lt.z};
#pragma unroll
    for (int ch = 0; ch < 3; ch++) ad[ch] = (Lr[ch] >= 0.f && Lr[ch] <= 10.f) ? di[ch] / (float)n : 0.f;
}

// Sums (material, g[6]) records across the wave, one FP64 atomic per
// material and parameter (lanes with m < 0 hold none).  Every lane of the
// wave must call it.
__device__ __forceinline__ void waveAddGradients(int m, const float g[6], double* __restrict__ dmat) {
    const uint32_t lane = __lane_id();
    unsigned long long pending = __ballot(m >= 0);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int m0 = __shfl(m, leader);
        const bool mine = m == m0;
        const unsigned long long grp = __ballot(mine);
#pragma unroll
        for (int q = 0; q < 6; q++) {
            float v = mine ? g[q] : 0.f;
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            if ((int)lane == leader && v != 0.f) atomicAdd(dmat + (uint64_t)m0 * 6 + q, (double)v);
        }
        pending &= ~grp;
    }
}

// Path-replay backpropagation (Vicini, Speierer, Jakob, "Path Replay
// Backpropagation", SIGGRAPH 2021) of the estimator above, with detached
// sampling: the sampled directions, lobe choices, pdfs and MIS weights of a
// path are held; the derivative flows through DisneyEval at the sampled, HDRI
// and light directions.  Each thread owns one (pixel, sample): it walks the
// path once for its radiance L (the primal), then replays it from the same
// Philox counters and, per vertex, adds
//   a . [ (T w3 Le cosS/bpdf + L_after / f) dF_s
//       + T w1 env cosH/hpdf dF_h [HDRI visible] + T w2 rad cosL/ppdf dF_l [light visible] ]
// with a = dLoss/dI_pixel / accepted samples (0 on clamped channels) and
// L_after = L - (radiance up to and including the vertex).  Nothing of the
// path outlives its thread: the primal walk keeps the closest hit of each of
// its vertices and the shadow answers in the thread's own slots of `cache`
// (vertex i at cache[i * gthreads + thread], coalesced across a wave; sized
// by the resident threads and the depth, not by the pixels or samples), so
// the replay recomputes every vertex from the RNG counters without tracing
// again.  Memory is the adjoint image, the counts, the gradients and that
// per-thread cache.
// Per-vertex records (material, 6 gradients) are kept per lane (private
// memory: they are indexed by the vertex number) until the path ends; the
// wave then reduces them per material (ballot + shuffle sums) and one lane
// per material adds them with an FP64 atomic.  Texel
// gradients of the selected texture go straight to FP32 atomics.
constexpr int kAdjMaxDepth = kAdjMaxDepthHost;

// REPLAY: the paths' vertices come from a recording render (a.recHits /
// a.recVis, EL_FLAG_RECORD_PATHS), so the kernel traces nothing.  Otherwise,
// with `rec` (the recorded radiance of each path) one traced walk suffices;
// without it the primal walk keeps its vertices in `cache` for the replay.
template <int W, bool REPLAY>
EL_KERNEL(REPLAY ? EL_WPE_ADJ_REPLAY : EL_WPE_ADJ) void k_adjoint(DevScene s, RenderArgs a, uint32_t spp, const float* __restrict__ dimage,
                                       const uint32_t* __restrict__ counts, const float4* __restrict__ rec,
                                       float4* __restrict__ cache, int gradTex, double* __restrict__ dmat,
                                       float* __restrict__ dtex, int* gstack, uint32_t gthreads,
                                       unsigned long long* counters) {
    TravStack st{};
    if constexpr (!REPLAY) st = makeStack(s, gstack, gthreads);
    TraceCounters cnt{0, 0};
    const uint32_t W_ = s.cam.xRes, H_ = s.cam.yRes;
    const uint32_t tilesX = (W_ + 15) / 16;
    const uint32_t perSample = tilesX * ((H_ + 15) / 16) * 256;
    const uint64_t total = (uint64_t)perSample * spp;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint32_t gtid = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t rays = 0, verts = 0, paths = 0, overflow = 0;
    const uint32_t lane = __lane_id();
    // uniform trip count: every lane of a wave runs the same iterations
    const uint64_t base0 = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
    for (uint64_t wb = base0; wb < total; wb += stride) {
        const uint64_t w = wb + lane;
        int nrec = 0;
        int recMat[kAdjMaxDepth];
        float recG[kAdjMaxDepth][6];
        if (w < total) {
            const uint32_t sample = a.sample + (uint32_t)(w / perSample);
            const uint32_t pw = (uint32_t)(w % perSample);
            const uint32_t tile = pw >> 8, in = pw & 255;
            const uint32_t x = (tile % tilesX) * 16 + (in & 15), y = (tile / tilesX) * 16 + (in >> 4);
            if (x < W_ && y < H_) {
                const uint32_t idx = W_ * (H_ - y - 1) + x;
                const uint32_t n = counts[idx];
                uint32_t vdummy = 0;
                v3 Lt;
                WalkIO io{};
                if (REPLAY || rec) {
                    const float4 r = rec[(uint64_t)sample * W_ * H_ + idx];
                    Lt = mk(r.x, r.y, r.z);
                } else if constexpr (!REPLAY) {
                    io.keep = cache + gtid;
                    io.keepStride = gthreads;
                    Lt = walkPath<W, kWalkTraceKeep>(s, a, x, y, idx, sample, st, cnt, rays, vdummy,
                                                     [](const PathVertex&) {}, io);
                }
                paths++;
                float ad[3];
                adjointWeight(Lt, dimage, idx, n, ad);
                if (ad[0] != 0.f || ad[1] != 0.f || ad[2] != 0.f) {
                    auto grad = [&](const PathVertex& pv) {
                        float g[6];
                        const int mat = vertexGradient(s, pv.d, pv.T, *pv.sh, *pv.hd, *pv.vo, pv.visH, pv.visP,
                                                       Lt - pv.Lincl, ad, gradTex, dtex, g);
                        if (nrec < kAdjMaxDepth) {
                            recMat[nrec] = mat;
#pragma unroll
                            for (int q = 0; q < 6; q++) recG[nrec][q] = g[q];
                            nrec++;
                        }
                    };
                    if constexpr (REPLAY) {
                        const uint64_t at = (uint64_t)sample * W_ * H_ + idx;
                        io.hits = a.recHits + at;
                        io.stride = a.recStride;
                        io.vis = a.recVis[at];
                        walkPath<W, kWalkReplay>(s, a, x, y, idx, sample, st, cnt, rays, verts, grad, io);
                    } else if (rec) {
                        walkPath<W, kWalkTrace>(s, a, x, y, idx, sample, st, cnt, rays, verts, grad, io);
                    } else {
                        io.hits = cache + gtid;
                        io.stride = gthreads;
                        io.vis = io.visOut;
                        walkPath<W, kWalkReplay>(s, a, x, y, idx, sample, st, cnt, rays, verts, grad, io);
                    }
                }
            }
        }
        overflow |= st.overflow;
        // converged: reduce the records per material across the wave
        for (int r = 0; r < (int)a.maxDepth && r < kAdjMaxDepth; r++)
            waveAddGradients(r < nrec ? recMat[r] : -1, recG[r < nrec ? r : 0], dmat);
    }
    const int slots[4] = {kStatRays, kStatVerts, kStatPaths, kStatOverflow};
    const uint32_t vals[4] = {rays, verts, paths, overflow};
    blockFlush<4>(counters, slots, vals);
}

// -------------------------------------------------------------- wavefront --
// Work item w (0 <= w < spp * npad): sample w / npad, padded pixel w % npad,
// padded pixels in 8x8 tiles.
__device__ __forceinline__ bool workPixel(const DevScene& s, uint32_t p, uint32_t& x, uint32_t& y) {
    uint32_t tilesX = (s.cam.xRes + 7) / 8;
    uint32_t tile = p >> 6, in = p & 63;
    x = (tile % tilesX) * 8 + (in & 7);
    y = (tile / tilesX) * 8 + (in >> 3);
    return x < s.cam.xRes && y < s.cam.yRes;
}

// Path slot of a single-scene generation -> output pixel index
// (0xFFFFFFFF: padding) and sample index; (x, y) the kernel pixel.  The
// camera, shade and finalize kernels recompute it instead of storing it.
// genPixel + slot < 2^32: both are below 2^31 (checked on the host).
__device__ __forceinline__ uint32_t slotPixel(const DevScene& s, const RenderArgs& a, uint32_t slot,
                                              uint32_t& sample, uint32_t& x, uint32_t& y) {
    const uint32_t u = a.genPixel + slot;
    sample = a.sample + a.genSample + u / a.npad;
    return workPixel(s, u % a.npad, x, y) ? s.cam.xRes * (s.cam.yRes - y - 1) + x : 0xFFFFFFFFu;
}

// Camera rays go to queue 0 at index = path slot (no compaction); padding
// slots get a zero direction, which the extend kernel reports as dead.
EL_KERNEL(1) void k_wf_camera(DevScene s, RenderArgs a, WfState q) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < a.genCount; slot += stride) {
        uint32_t sample, x, y;
        const uint32_t idx = slotPixel(s, a, slot, sample, x, y);
        const bool valid = idx != 0xFFFFFFFFu;
        // throughput and radiance start at (1, 0): the first shade assumes them
        if (q.pathAov) {
            q.pathAov[3 * (uint64_t)slot + 0] = make_float4(0, 0, 0, 0);
            q.pathAov[3 * (uint64_t)slot + 1] = make_float4(0, 0, 0, 0);
            q.pathAov[3 * (uint64_t)slot + 2] = make_float4(0, 0, 0, 0);
        }
        v3 o = mk(0, 0, 0), d = mk(0, 0, 0);
        uint64_t quiet = ~0ull;
        if (valid) {
            PathRng rng{a.k0, a.k1, idx, sample, a.scene};
            uint4 c = rng.block(0);
            uint32_t c5 = rng.block(1).x;
            cameraRay(s.cam, (int)x, (int)y, u01(c.x), u01(c.y), u01(c.z), u01(c.w), u01(c5), o, d);
            if (a.cull) quiet = nanProneBounces(s, rng, (int)a.maxDepth);
        }
        q.pathQuiet[slot] = quiet;
        if (q.pathVis) q.pathVis[slot] = 0u;
        float4* Q = q.ray[0] + 2 * (uint64_t)slot;
        Q[0] = make_float4(o.x, o.y, o.z, __uint_as_float(slot));
        Q[1] = make_float4(d.x, d.y, d.z, __uint_as_float(a.scene));
    }
    if (blockIdx.x == 0 && threadIdx.x < kXcds) q.counters[kQRay0 + threadIdx.x] = threadIdx.x == 0 ? a.genCount : 0u;
    resetChunks(q.counters, kChunkExt);
}

// Scene addressing of the wavefront kernels.  OneScene: the kernel argument
// is the scene.  ManyScenes: a device array of scenes (a batch), indexed by
// the scene id every queue entry carries (ray queue: bits of d.w; shadow
// queue: bits of the contribution's w).  All scenes of a batch share the BVH
// width and the traversal stack's spill capacity.
struct OneScene {
    DevScene s;
    static constexpr bool kMany = false;
    __device__ __forceinline__ const DevScene& get(uint32_t) const { return s; }
    __device__ __forceinline__ const DevScene& any() const { return s; }
};
struct ManyScenes {
    const DevScene* p;
    static constexpr bool kMany = true;
    __device__ __forceinline__ const DevScene& get(uint32_t i) const { return p[i]; }
    __device__ __forceinline__ const DevScene& any() const { return p[0]; }
};

// Also clears the counters the following shade kernel appends to: the next
// ray queue and both shadow queues.
template <int W, bool COUNT, class P>
EL_KERNEL(EL_WPE_TRAV) void k_wf_extend(P sc, WfState q, const float4* __restrict__ rayIn,
                                        const uint32_t* countIn, uint32_t* resetNext, uint32_t refill,
                                        unsigned long long* counters) {
    const DevScene& s0 = sc.any();
    TravStack st = makeStack(s0, q.gstack, q.gthreads);
    TraceCounters cnt{0, 0};
    QueueMap qm;
    qm.load(countIn);
    const uint32_t n = qm.total();
    // the counters the shade and shadow launches of this bounce use (set
    // q.parity; the shadow launches of the previous bounce, which may still
    // run, use the other set)
    if (blockIdx.x == 0 && threadIdx.x < kXcds) {
        resetNext[threadIdx.x] = 0;
        q.counters[qShadow(q.parity, 0) + threadIdx.x] = 0;
        q.counters[qShadow(q.parity, 1) + threadIdx.x] = 0;
    }
    resetChunks(q.counters, chunkShadow(q.parity, 0));
    resetChunks(q.counters, chunkShadow(q.parity, 1));
    uint32_t rays = 0;
    const WorkSource ws = chunkSource<P::kMany>(q.counters, kChunkExt, n);
    const uint32_t segSize = q.segSize;
    auto load = [&](uint32_t i, float& ox, float& oy, float& oz, float& dx, float& dy, float& dz, float& tm,
                    BvhPtr& bvh) {
        const uint64_t e = qm.entry(i, segSize);
        const float4 A = rayIn[2 * e], B = rayIn[2 * e + 1];
        ox = A.x;
        oy = A.y;
        oz = A.z;
        dx = B.x;
        dy = B.y;
        dz = B.z;
        tm = __int_as_float(0x7f800000);
        if (P::kMany) {
            const DevScene& sx = sc.get(__float_as_uint(B.w));
            bvh = BvhPtr{(const uint4*)sx.nodes, sx.tris};
        }
        return !(B.x == 0.f && B.y == 0.f && B.z == 0.f);   // zero direction: camera-queue padding
    };
    auto store = [&](uint32_t i, const TraceResult& h, bool) {
        q.hits[qm.entry(i, segSize)] = make_float4(__int_as_float(h.tri), h.t, h.u, h.v);
    };
    traceWaveRange<W, false, COUNT, P::kMany>(s0, st, ws, refill, load, store, cnt, rays);
    const int slots[5] = {kStatRays, kStatOverflow, kStatNodes, kStatTris, kStatExtRays};
    const uint32_t vals[5] = {rays, (uint32_t)st.overflow, cnt.nodes, cnt.tris, rays};
    blockFlush<5>(counters, slots, vals);
}

// Adds `c` to a path's radiance.  The first bounce writes it (the radiance
// starts at zero there and is not stored before); later bounces skip the
// read-modify-write when every component is zero, which leaves L bit for bit
// as the addition would (x + 0 == x, x + -0 == x; a NaN is still added).
// `pre` (later bounces): the previous bounce's deferred point-light term,
// added first, where its shadow launch used to add it.  It is zero when the
// light was blocked (a visible term is never zero: its query is traced only
// then).
__device__ __forceinline__ bool isZero4(float4 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
__device__ __forceinline__ void addRadiance(float4* __restrict__ pathL, uint32_t slot, int bounce, v3 c,
                                            float4 pre = make_float4(0.f, 0.f, 0.f, 0.f)) {
    if (bounce == 0) {
        pathL[slot] = make_float4(0.f + c.x, 0.f + c.y, 0.f + c.z, 1.f);   // 0 + -0 == +0, as a sum
        return;
    }
    const bool addPre = !isZero4(pre), addC = !(c.x == 0.f && c.y == 0.f && c.z == 0.f);
    if (!addPre && !addC) return;
    float4 L = pathL[slot];
    if (addPre) L = make_float4(L.x + pre.x, L.y + pre.y, L.z + pre.z, L.w);
    if (addC) L = make_float4(L.x + c.x, L.y + c.y, L.z + c.z, L.w);
    pathL[slot] = L;
}

// Marks on pathL.w / pathT.w that a path's last shade emitted a point-light
// query whose term is deferred (pendP): 2 on the radiance of a path that
// ends there (finalize adds the term), 1 on the throughput of one that
// continues (the next shade adds it).
constexpr float kPendEnd = 2.f, kPendNext = 1.f;

// Escaped rays (a texture lookup) and hits (full shading) arrive mixed in
// the queue.  Each block classifies a chunk of blockDim entries, adds the
// environment term of its escaped rays at once and pushes its hits onto an
// LDS pool; hits are shaded blockDim at a time from the top of the pool, so
// the long shading code runs on full waves.  Results do not depend on the
// order: every path has one entry per queue.
template <class P>
EL_KERNEL(EL_WPE_SHADE) void k_wf_shade(P sc, RenderArgs a, WfState q, const float4* __restrict__ rayIn,
                           const uint32_t* countIn, float4* __restrict__ rayOut, uint32_t* countOut,
                           unsigned long long* counters) {
    __shared__ uint32_t s_pool[2 * kMaxBlock];
    __shared__ uint32_t s_wcnt[kMaxWaves];
    __shared__ float s_lut[512];   // the byte -> value tables (the same for every scene)
    for (uint32_t i = threadIdx.x; i < 512; i += blockDim.x) s_lut[i] = sc.any().texLut[i];
    __syncthreads();
    QueueMap qm;
    qm.load(countIn);
    const uint32_t n = qm.total();
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t rays = 0, verts = 0, overflow = 0;
    const int bounce = (int)a.bounce;
    const bool last = bounce + 1 >= (int)a.maxDepth;
    uint32_t* const outCnt[3] = {countOut, q.counters + qShadow(q.parity, 0), q.counters + qShadow(q.parity, 1)};
    const int wave = threadIdx.x >> 6, nw = blockDim.x >> 6, lane = __lane_id();
    const unsigned long long lt = lane ? (~0ull >> (64 - lane)) : 0ull;
    resetChunks(q.counters, kChunkExt);   // for the next bounce's extend launch
    uint32_t pending = 0;   // hits in the pool (block-uniform)
    for (uint32_t b = blockIdx.x * blockDim.x;; b += stride) {
        const bool more = b < n;
        if (more) {
            const uint32_t i = b + threadIdx.x;
            uint32_t e = 0;
            bool hit = false;
            if (i < n) {
                e = (uint32_t)qm.entry(i, q.segSize);
                const int tri = __float_as_int(q.hits[e].x);
                if (tri == -1) {
                    const float4 A = rayIn[2 * (uint64_t)e], B = rayIn[2 * (uint64_t)e + 1];
                    const uint32_t slot = __float_as_uint(A.w);
                    const DevScene& s = sc.get(__float_as_uint(B.w));
                    const float4 Tf = bounce == 0 ? make_float4(1, 1, 1, 0) : q.pathT[slot];
                    const v3 ev = envLookup(s, mk(B.x, B.y, B.z)) * mk(Tf.x, Tf.y, Tf.z);
                    addRadiance(q.pathL, slot, bounce, ev,
                                Tf.w == kPendNext ? q.pendP[slot] : make_float4(0.f, 0.f, 0.f, 0.f));
                }
                hit = tri >= 0;
                if constexpr (!P::kMany) {
                    // recording render: vertex `bounce` of this path, hit or
                    // escape (tri -2: camera-queue padding, no path)
                    if (a.recHits && tri >= -1) {
                        const DevScene& s = sc.get(0);
                        uint32_t smp, px, py;
                        const uint32_t pix = slotPixel(s, a, __float_as_uint(rayIn[2 * (uint64_t)e].w), smp, px, py);
                        if (pix != 0xFFFFFFFFu)
                            a.recHits[(uint64_t)bounce * a.recStride + (uint64_t)smp * s.cam.xRes * s.cam.yRes +
                                      pix] = q.hits[e];
                    }
                }
            }
            const unsigned long long m = __ballot(hit);
            if (lane == 0) s_wcnt[wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t off = pending, total = 0;
            for (int w = 0; w < nw; w++) {
                const uint32_t c = s_wcnt[w];
                off += w < wave ? c : 0u;
                total += c;
            }
            if (hit) s_pool[off + (uint32_t)__popcll(m & lt)] = e;
            pending += total;
            __syncthreads();   // pool entries visible, s_wcnt free
            if (pending < blockDim.x) continue;
        } else if (pending == 0) {
            break;
        }
        const uint32_t take = pending < blockDim.x ? pending : blockDim.x;
        const bool has = threadIdx.x < take;
        const uint32_t e = has ? s_pool[pending - take + threadIdx.x] : 0u;
        pending -= take;
        bool cont = false, wantH = false, wantP = false;
        VertexOut vo;
        uint32_t slot = 0, sid = 0;
        if (has) {
            const float4 A = rayIn[2 * (uint64_t)e], B = rayIn[2 * (uint64_t)e + 1];
            const float4 hr = q.hits[e];
            slot = __float_as_uint(A.w);
            sid = __float_as_uint(B.w);
            const DevScene& s = sc.get(sid);
            const int tri = __float_as_int(hr.x);
            v3 o = mk(A.x, A.y, A.z), d = mk(B.x, B.y, B.z);
            const float4 Tf = bounce == 0 ? make_float4(1, 1, 1, 0) : q.pathT[slot];
            SurfaceHit sh = surfaceAt(s, tri, hr.y, hr.z, hr.w, o, d);
            HitData hd = hitData(s, sh, s_lut);
            uint32_t k0 = a.k0, k1 = a.k1;
            if (P::kMany) {
                k0 = s.seedLo;
                k1 = s.seedHi;
            }
            uint32_t pix, smp;
            if constexpr (P::kMany) {
                pix = q.pathPix[slot];
                smp = q.pathSample[slot];
            } else {
                uint32_t px, py;
                pix = slotPixel(s, a, slot, smp, px, py);
            }
            PathRng rng{k0, k1, pix, smp, sid};
            uint4 rb = rng.block(2 + bounce);
            uint32_t r[4] = {rb.x, rb.y, rb.z, rb.w};
            vo = shadeVertex(s, d, mk(Tf.x, Tf.y, Tf.z), sh, hd, r);
            addRadiance(q.pathL, slot, bounce, vo.base,
                        Tf.w == kPendNext ? q.pendP[slot] : make_float4(0.f, 0.f, 0.f, 0.f));
            if (bounce == 0 && q.pathAov) {
                v3 at, ab;
                surfaceTangents(s, sh, at, ab);
                q.pathAov[3 * (uint64_t)slot + 0] = make_float4(sh.normal.x, sh.normal.y, sh.normal.z, 0);
                q.pathAov[3 * (uint64_t)slot + 1] = make_float4(at.x, at.y, at.z, 0);
                q.pathAov[3 * (uint64_t)slot + 2] = make_float4(ab.x, ab.y, ab.z, 0);
            }
            verts++;
            rays += 1 + (vo.hasP ? 1 : 0);
            // A shadow ray whose contribution is exactly zero cannot change L.
            // A recording render also traces such a query when the direction
            // lies in the BRDF's support, where the adjoint's derivative
            // term may be nonzero although the value is zero (a black
            // albedo); outside it disneyEvalGrad is zero and the answer is
            // not needed (the 1e-6 margins absorb rounding differences
            // between this test and the one the adjoint makes).
            const bool recV = !P::kMany && q.pathVis != nullptr;
            const bool front = recV && hd.transmission < 1.0f && dot(hd.normal, d) < 1e-6f;
            wantH = vo.contribH.x != 0.f || vo.contribH.y != 0.f || vo.contribH.z != 0.f ||
                    (front && dot(hd.normal, vo.hdir) > -1e-6f);
            wantP = vo.hasP && (vo.contribP.x != 0.f || vo.contribP.y != 0.f || vo.contribP.z != 0.f ||
                                (front && dot(hd.normal, vo.ldir) > -1e-6f));
            cont = !last && !(a.cull && isZero(vo.newT) && (q.pathQuiet[slot] >> (bounce + 1)) == 0ull);
            // only an extended path reads its throughput again
            const bool pend = q.deferP != 0u && wantP;
            if (cont) q.pathT[slot] = make_float4(vo.newT.x, vo.newT.y, vo.newT.z, pend ? kPendNext : 0.f);
            if (pend && !cont) reinterpret_cast<float*>(q.pathL)[4 * (uint64_t)slot + 3] = kPendEnd;
        }
        const bool want[3] = {cont, wantH, wantP};
        uint64_t at[3];
        blockAlloc3(want, outCnt, q.segSize, at, overflow);   // ends with a barrier: the pool may be refilled
        if (at[0] != ~0ull) {
            float4* Q = rayOut + 2 * at[0];
            Q[0] = make_float4(vo.nextO.x, vo.nextO.y, vo.nextO.z, __uint_as_float(slot));
            Q[1] = make_float4(vo.nextD.x, vo.nextD.y, vo.nextD.z, __uint_as_float(sid));
        }
        if (at[1] != ~0ull) {
            float4* S = q.shadowH + 3 * at[1];
            S[0] = make_float4(vo.hOrig.x, vo.hOrig.y, vo.hOrig.z, __uint_as_float(slot));
            S[1] = make_float4(vo.hDir.x, vo.hDir.y, vo.hDir.z, __int_as_float(0x7f800000));
            S[2] = make_float4(vo.contribH.x, vo.contribH.y, vo.contribH.z, __uint_as_float(sid));
        }
        if (at[2] != ~0ull) {
            float4* S = q.shadowP + 3 * at[2];
            S[0] = make_float4(vo.pOrig.x, vo.pOrig.y, vo.pOrig.z, __uint_as_float(slot));
            S[1] = make_float4(vo.pDir.x, vo.pDir.y, vo.pDir.z, vo.pTmax);
            S[2] = make_float4(vo.contribP.x, vo.contribP.y, vo.contribP.z, __uint_as_float(sid));
        }
    }
    const int slots[3] = {kStatRays, kStatVerts, kStatOverflow};
    const uint32_t vals[3] = {rays, verts, overflow};
    blockFlush<3>(counters, slots, vals);
}

// Any-hit test of a shadow queue; an unoccluded entry adds its contribution
// to its path's radiance (each path has at most one entry per queue, and the
// two shadow queues are separate launches, so the update needs no atomics).
// REC: a recording render's launch; visible queries set q.visBit in the
// path's visibility word, and queries traced only for the record add nothing.
template <int W, class P, bool REC = false>
EL_KERNEL(EL_WPE_TRAV) void k_wf_shadow(P sc, WfState q, const float4* __restrict__ sh,
                                        const uint32_t* countIn, uint32_t refill,
                                        unsigned long long* counters) {
    const DevScene& s0 = sc.any();
    TravStack st = makeStack(s0, q.gstack, q.gthreads);
    TraceCounters cnt{0, 0};
    QueueMap qm;
    qm.load(countIn);
    const uint32_t n = qm.total();
    uint32_t traced = 0;
    const WorkSource ws = chunkSource<P::kMany>(q.counters, chunkShadow(q.parity, sh == q.shadowH ? 0 : 1), n);
    const uint32_t segSize = q.segSize;
    auto load = [&](uint32_t i, float& ox, float& oy, float& oz, float& dx, float& dy, float& dz, float& tm,
                    BvhPtr& bvh) {
        const uint64_t e = qm.entry(i, segSize);
        const float4 A = sh[3 * e], B = sh[3 * e + 1];
        ox = A.x;
        oy = A.y;
        oz = A.z;
        dx = B.x;
        dy = B.y;
        dz = B.z;
        tm = B.w;
        if (P::kMany) {
            const DevScene& sx = sc.get(__float_as_uint(sh[3 * e + 2].w));
            bvh = BvhPtr{(const uint4*)sx.nodes, sx.tris};
        }
        return true;
    };
    // point-light queries of a render with deferred terms: every query
    // writes its term (zero if blocked) for the next shade or finalize
    const bool defer = !REC && !P::kMany && q.deferP != 0u && sh == q.shadowP;
    auto store = [&](uint32_t i, const TraceResult& h, bool) {
        if (h.tri >= 0 && !defer) return;
        const uint64_t e = qm.entry(i, segSize);
        const uint32_t slot = __float_as_uint(sh[3 * e].w);
        if (defer) {
            const float4 C = sh[3 * e + 2];
            q.pendP[slot] = h.tri >= 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : make_float4(C.x, C.y, C.z, 0.f);
            return;
        }
        const float4 C = sh[3 * e + 2];
        if constexpr (REC) {
            q.pathVis[slot] |= q.visBit;
            if (C.x == 0.f && C.y == 0.f && C.z == 0.f) return;   // traced only for the record
        }
        const float4 L = q.pathL[slot];
        q.pathL[slot] = make_float4(L.x + C.x, L.y + C.y, L.z + C.z, L.w);
    };
    traceWaveRange<W, true, false, P::kMany>(s0, st, ws, refill, load, store, cnt, traced);
    const int slots[1] = {kStatOverflow};
    const uint32_t vals[1] = {(uint32_t)st.overflow};
    blockFlush<1>(counters, slots, vals);
}

// One thread per distinct pixel of the generation; slots j, j+npad, ... hold
// the same pixel, so accumulation needs no atomics.
EL_KERNEL(1) void k_wf_finalize(DevScene s, RenderArgs a, WfState q, float4* __restrict__ accum,
                              uint32_t* __restrict__ counts, float4* __restrict__ aov,
                              unsigned long long* counters) {
    uint32_t stride = gridDim.x * blockDim.x;
    uint32_t lim = a.genCount < a.npad ? a.genCount : a.npad;
    uint32_t paths = 0;
    for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < lim; j += stride) {
        // every slot of this thread is the same pixel: its sums stay in
        // registers (added in sample order) and are stored once
        uint32_t sample, x, y;
        const uint32_t idx = slotPixel(s, a, j, sample, x, y);
        if (idx == 0xFFFFFFFFu) continue;   // padding pixel
        float4 acc = accum[idx];
        uint32_t n = counts[idx];
        float4 A[3];
        if (aov)
            for (int k = 0; k < 3; k++) A[k] = aov[3 * (uint64_t)idx + k];
        for (uint32_t slot = j; slot < a.genCount; slot += a.npad, sample++) {
            float4 Lf = q.pathL[slot];
            if (Lf.w == kPendEnd) {   // the last bounce's deferred point-light term, added last
                const float4 Pp = q.pendP[slot];
                if (!isZero4(Pp)) Lf = make_float4(Lf.x + Pp.x, Lf.y + Pp.y, Lf.z + Pp.z, Lf.w);
            }
            paths++;
            if (a.rec) a.rec[(uint64_t)sample * s.cam.xRes * s.cam.yRes + idx] = make_float4(Lf.x, Lf.y, Lf.z, 1.f);
            if (a.recVis) a.recVis[(uint64_t)sample * s.cam.xRes * s.cam.yRes + idx] = q.pathVis[slot];
            const v3 L = clamp3(mk(Lf.x, Lf.y, Lf.z));
            if (anyNan(L)) continue;
            acc = make_float4(acc.x + L.x, acc.y + L.y, acc.z + L.z, acc.w);
            n += 1;
            if (aov) {
                const float4* P = q.pathAov + 3 * (uint64_t)slot;
                for (int k = 0; k < 3; k++) A[k] = make_float4(A[k].x + P[k].x, A[k].y + P[k].y, A[k].z + P[k].z, 0);
            }
        }
        accum[idx] = acc;
        counts[idx] = n;
        if (aov)
            for (int k = 0; k < 3; k++) aov[3 * (uint64_t)idx + k] = A[k];
    }
    const int slots[1] = {kStatPaths};
    const uint32_t vals[1] = {paths};
    blockFlush<1>(counters, slots, vals);
}

// ------------------------------------------------------------------ batch --
// Segment of a generation slot (or finalize thread): the last segment whose
// start is <= t.  nseg is small (at most the number of scenes + 1).
__device__ __forceinline__ uint32_t findSeg(const BatchSeg* segs, uint32_t nseg, uint32_t t, bool fin) {
    uint32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if ((fin ? segs[mid].finStart : segs[mid].slotStart) <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Camera rays of a batch generation: slot -> (scene, sample, pixel) through
// the segment table; the scene id rides in the ray's d.w.
EL_KERNEL(1) void k_bt_camera(ManyScenes sc, const BatchSeg* __restrict__ segs, uint32_t nseg, uint32_t genCount,
                              uint32_t maxDepth, uint32_t cull, WfState q) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x; slot < genCount; slot += stride) {
        const BatchSeg g = segs[findSeg(segs, nseg, slot, false)];
        const DevScene& s = sc.get(g.scene);
        const uint64_t w = g.wStart + (slot - g.slotStart);
        const uint32_t p = (uint32_t)(w % g.npad);
        const uint32_t sample = (uint32_t)(w / g.npad);
        uint32_t x, y;
        const bool valid = workPixel(s, p, x, y);
        const uint32_t idx = valid ? s.cam.xRes * (s.cam.yRes - y - 1) + x : 0xFFFFFFFFu;
        q.pathPix[slot] = idx;   // 0xFFFFFFFF: padding; throughput and radiance: set by the first shade
        q.pathSample[slot] = sample;
        v3 o = mk(0, 0, 0), d = mk(0, 0, 0);
        uint64_t quiet = ~0ull;
        if (valid) {
            PathRng rng{s.seedLo, s.seedHi, idx, sample, g.scene};
            uint4 c = rng.block(0);
            uint32_t c5 = rng.block(1).x;
            cameraRay(s.cam, (int)x, (int)y, u01(c.x), u01(c.y), u01(c.z), u01(c.w), u01(c5), o, d);
            if (cull) quiet = nanProneBounces(s, rng, (int)maxDepth);
        }
        q.pathQuiet[slot] = quiet;
        float4* Q = q.ray[0] + 2 * (uint64_t)slot;
        Q[0] = make_float4(o.x, o.y, o.z, __uint_as_float(slot));
        Q[1] = make_float4(d.x, d.y, d.z, __uint_as_float(g.scene));
    }
    if (blockIdx.x == 0 && threadIdx.x < kXcds) q.counters[kQRay0 + threadIdx.x] = threadIdx.x == 0 ? genCount : 0u;
    resetChunks(q.counters, kChunkExt);
}

// Accumulation of a batch generation: one thread per distinct (scene, pixel)
// of the generation (slots j, j + npad, ... of a segment share the pixel),
// so no two threads touch one pixel and no atomics are needed.
EL_KERNEL(1) void k_bt_finalize(ManyScenes sc, const BatchSeg* __restrict__ segs, uint32_t nseg, uint32_t finTotal,
                                WfState q, unsigned long long* counters) {
    const uint32_t stride = gridDim.x * blockDim.x;
    uint32_t paths = 0;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < finTotal; t += stride) {
        const BatchSeg g = segs[findSeg(segs, nseg, t, true)];
        const DevScene& s = sc.get(g.scene);
        const uint32_t end = g.slotStart + g.count;
        const uint32_t first = g.slotStart + (t - g.finStart);
        // every slot of this thread is the same pixel of the scene: its sums
        // stay in registers (added in sample order) and are stored once
        const uint32_t idx = q.pathPix[first];
        if (idx == 0xFFFFFFFFu) continue;   // padding pixel
        float4 acc = s.accum[idx];
        uint32_t n = s.counts[idx];
        for (uint32_t slot = first; slot < end; slot += g.npad) {
            const float4 Lf = q.pathL[slot];
            paths++;
            const v3 L = clamp3(mk(Lf.x, Lf.y, Lf.z));
            if (anyNan(L)) continue;
            acc = make_float4(acc.x + L.x, acc.y + L.y, acc.z + L.z, acc.w);
            n += 1;
        }
        s.accum[idx] = acc;
        s.counts[idx] = n;
    }
    const int slots[1] = {kStatPaths};
    const uint32_t vals[1] = {paths};
    blockFlush<1>(counters, slots, vals);
}

// envTab[2k], envTab[2k+1] = envSampleAt(k) for every environment pixel k < n
// (its shadow direction is recomputed from the direction at sampling time).
__global__ void __launch_bounds__(256) k_env_table(DevScene s, float4* tab, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const EnvSample e = envSampleAt(s, (int)k);
    tab[2 * (uint64_t)k] = make_float4(e.dir.x, e.dir.y, e.dir.z, e.pdf);
    tab[2 * (uint64_t)k + 1] = make_float4(e.value.x, e.value.y, e.value.z, 0.f);
}

// The L2 residual of the mean image (accum / counts, as el_read_passes
// divides) against a target, and the block sums of its squares.
__global__ void __launch_bounds__(256) k_residual_l2(const float4* __restrict__ accum,
                                                     const uint32_t* __restrict__ counts,
                                                     const float* __restrict__ target, uint32_t npx,
                                                     float* __restrict__ r, double* loss) {
    __shared__ double s_part[256 / 64];
    double part = 0.0;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < npx; i += gridDim.x * blockDim.x) {
        const float4 acc = accum[i];
        const uint32_t n = counts[i];
        const float fn = (float)n;
        const float I[3] = {n ? acc.x / fn : 0.f, n ? acc.y / fn : 0.f, n ? acc.z / fn : 0.f};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double d = (double)I[k] - (double)target[3 * (uint64_t)i + k];
            r[3 * (uint64_t)i + k] = (float)(2.0 * d);
            part += d * d;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) part += __shfl_xor(part, o);
    if (__lane_id() == 0) s_part[threadIdx.x >> 6] = part;
    __syncthreads();
    if (threadIdx.x == 0) {
        double b = 0.0;
        for (uint32_t w = 0; w < (blockDim.x >> 6); w++) b += s_part[w];
        atomicAdd(loss, b);
    }
}

}  // namespace dev

// ------------------------------------------------------------ host launch --
using namespace dev;

static uint32_t gridFor(uint64_t n, uint32_t block, uint32_t maxBlocks) {
    uint64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (uint32_t)(g < maxBlocks ? g : maxBlocks);
}

hipError_t launchTrace(const LaunchCfg& c, const DevScene& s, const float* rays, const float* tmax, uint32_t n,
                       int32_t* tri, float* t, float* uv, uint8_t* occ, bool any, bool count,
                       unsigned long long* counters) {
    uint32_t grid = gridFor(n, c.block, c.maxThreads / c.block);
    size_t lds = sizeof(int) * kShortStack * c.block;
#define EL_TRACE(WW, A, C)                                                                                   \
    hipLaunchKernelGGL((k_trace<WW, A, C>), dim3(grid), dim3(c.block), lds, c.stream, s, rays, tmax, n, tri, t, \
                       uv, occ, c.gstack, c.maxThreads, c.refill, counters)
    if (s.width == 4) {
        if (any)
            EL_TRACE(4, true, false);
        else if (count)
            EL_TRACE(4, false, true);
        else
            EL_TRACE(4, false, false);
    } else {
        if (any)
            EL_TRACE(8, true, false);
        else if (count)
            EL_TRACE(8, false, true);
        else
            EL_TRACE(8, false, false);
    }
#undef EL_TRACE
    return hipGetLastError();
}

hipError_t launchPlain(const LaunchCfg& c, const DevScene& s, const RenderArgs& a, float4* accum, uint32_t* counts,
                       float4* aov, unsigned long long* counters) {
    uint64_t total = (uint64_t)((s.cam.xRes + 15) / 16) * ((s.cam.yRes + 15) / 16) * 256;
    uint32_t grid = gridFor(total, c.block, c.maxThreads / c.block);
    size_t lds = sizeof(int) * kShortStack * c.block;
    if (s.width == 4)
        hipLaunchKernelGGL(k_plain<4>, dim3(grid), dim3(c.block), lds, c.stream, s, a, accum, counts, aov, c.gstack,
                           c.maxThreads, counters);
    else
        hipLaunchKernelGGL(k_plain<8>, dim3(grid), dim3(c.block), lds, c.stream, s, a, accum, counts, aov, c.gstack,
                           c.maxThreads, counters);
    return hipGetLastError();
}

hipError_t launchAdjoint(const LaunchCfg& c, const DevScene& s, const RenderArgs& a, uint32_t spp,
                         const float* dimage, const uint32_t* counts, const float4* rec, float4* cache, int gradTex,
                         double* dmat, float* dtex, unsigned long long* counters) {
    uint64_t total = (uint64_t)((s.cam.xRes + 15) / 16) * ((s.cam.yRes + 15) / 16) * 256 * spp;
    uint32_t grid = gridFor(total, c.block, c.maxThreads / c.block);
    size_t lds = sizeof(int) * kShortStack * c.block;
    if (a.recHits)   // replays recorded vertices: no traversal, no stack
        hipLaunchKernelGGL((k_adjoint<4, true>), dim3(grid), dim3(c.block), 0, c.stream, s, a, spp, dimage, counts,
                           rec, cache, gradTex, dmat, dtex, c.gstack, c.maxThreads, counters);
    else if (s.width == 4)
        hipLaunchKernelGGL((k_adjoint<4, false>), dim3(grid), dim3(c.block), lds, c.stream, s, a, spp, dimage, counts,
                           rec, cache, gradTex, dmat, dtex, c.gstack, c.maxThreads, counters);
    else
        hipLaunchKernelGGL((k_adjoint<8, false>), dim3(grid), dim3(c.block), lds, c.stream, s, a, spp, dimage, counts,
                           rec, cache, gradTex, dmat, dtex, c.gstack, c.maxThreads, counters);
    return hipGetLastError();
}

hipError_t launchResidualL2(hipStream_t stream, const float4* accum, const uint32_t* counts, const float* target,
                            uint32_t npx, float* r, double* loss) {
    const uint32_t grid = gridFor(npx, 256, 4096);
    hipLaunchKernelGGL(k_residual_l2, dim3(grid), dim3(256), 0, stream, accum, counts, target, npx, r, loss);
    return hipGetLastError();
}

hipError_t launchEnvTable(hipStream_t stream, const DevScene& s, float4* tab) {
    const uint32_t n = (uint32_t)s.env.w * (uint32_t)s.env.h;
    hipLaunchKernelGGL(k_env_table, dim3((n + 255) / 256), dim3(256), 0, stream, s, tab, n);
    return hipGetLastError();
}

hipError_t launchWfCamera(const LaunchCfg& c, const DevScene& s, const RenderArgs& a, const WfState& q) {
    uint32_t grid = gridFor(a.genCount, c.shadeBlock, c.maxThreads / c.shadeBlock);
    hipLaunchKernelGGL(k_wf_camera, dim3(grid), dim3(c.shadeBlock), 0, c.stream, s, a, q);
    return hipGetLastError();
}

hipError_t launchWfExtend(const LaunchCfg& c, const DevScene& s, const WfState& q, const float4* rayIn,
                          const uint32_t* countIn, uint32_t* resetNext, uint32_t maxCount, bool countTraversal,
                          unsigned long long* counters) {
    uint32_t grid = gridFor(maxCount, c.block, c.maxThreads / c.block);
    size_t lds = sizeof(int) * kShortStack * c.block;
#define EL_EXT(WW, C)                                                                                       \
    hipLaunchKernelGGL((k_wf_extend<WW, C, OneScene>), dim3(grid), dim3(c.block), lds, c.stream, OneScene{s}, q, \
                       rayIn, countIn, resetNext, c.refill, counters)
    if (s.width == 4) {
        if (countTraversal)
            EL_EXT(4, true);
        else
            EL_EXT(4, false);
    } else {
        if (countTraversal)
            EL_EXT(8, true);
        else
            EL_EXT(8, false);
    }
#undef EL_EXT
    return hipGetLastError();
}

hipError_t launchWfShade(const LaunchCfg& c, const DevScene& s, const RenderArgs& a, const WfState& q,
                         const float4* rayIn, const uint32_t* countIn, float4* rayOut, uint32_t* countOut,
                         uint32_t maxCount, unsigned long long* counters) {
    uint32_t grid = gridFor(maxCount, c.shadeBlock, c.maxThreads / c.shadeBlock);
    hipLaunchKernelGGL(k_wf_shade<OneScene>, dim3(grid), dim3(c.shadeBlock), 0, c.stream, OneScene{s}, a, q, rayIn,
                       countIn, rayOut, countOut, counters);
    return hipGetLastError();
}

hipError_t launchWfShadow(const LaunchCfg& c, const DevScene& s, const WfState& q, const float4* shadow,
                          const uint32_t* countIn, uint32_t maxCount, unsigned long long* counters) {
    uint32_t grid = gridFor(maxCount, c.block, c.maxThreads / c.block);
    size_t lds = sizeof(int) * kShortStack * c.block;
    const bool rec = q.visBit != 0;
    if (s.width == 4 && rec)
        hipLaunchKernelGGL((k_wf_shadow<4, OneScene, true>), dim3(grid), dim3(c.block), lds, c.stream, OneScene{s},
                           q, shadow, countIn, c.refill, counters);
    else if (s.width == 4)
        hipLaunchKernelGGL((k_wf_shadow<4, OneScene>), dim3(grid), dim3(c.block), lds, c.stream, OneScene{s}, q,
                           shadow, countIn, c.refill, counters);
    else if (rec)
        hipLaunchKernelGGL((k_wf_shadow<8, OneScene, true>), dim3(grid), dim3(c.block), lds, c.stream, OneScene{s},
                           q, shadow, countIn, c.refill, counters);
    else
        hipLaunchKernelGGL((k_wf_shadow<8, OneScene>), dim3(grid), dim3(c.block), lds, c.stream, OneScene{s}, q,
                           shadow, countIn, c.refill, counters);
    return hipGetLastError();
}

hipError_t launchWfFinalize(const LaunchCfg& c, const DevScene& s, const RenderArgs& a, const WfState& q,
                            float4* accum, uint32_t* counts, float4* aov, unsigned long long* counters) {
    uint32_t lim = a.genCount < a.npad ? a.genCount : a.npad;
    uint32_t grid = gridFor(lim, c.shadeBlock, c.maxThreads / c.shadeBlock);
    hipLaunchKernelGGL(k_wf_finalize, dim3(grid), dim3(c.shadeBlock), 0, c.stream, s, a, q, accum, counts, aov,
                       counters);
    return hipGetLastError();
}

// ---- batch launchers
hipError_t launchBtCamera(const LaunchCfg& c, const DevScene* scenes, const BatchSeg* segs, uint32_t nseg,
                          uint32_t genCount, uint32_t maxDepth, bool cull, const WfState& q) {
    uint32_t grid = gridFor(genCount, c.shadeBlock, c.maxThreads / c.shadeBlock);
    hipLaunchKernelGGL(k_bt_camera, dim3(grid), dim3(c.shadeBlock), 0, c.stream, ManyScenes{scenes}, segs, nseg,
                       genCount, maxDepth, cull ? 1u : 0u, q);
    return hipGetLastError();
}

hipError_t launchBtExtend(const LaunchCfg& c, const DevScene* scenes, int width, const WfState& q,
                          const float4* rayIn, const uint32_t* countIn, uint32_t* resetNext, uint32_t maxCount,
                          unsigned long long* counters) {
    uint32_t grid = gridFor(maxCount, c.block, c.maxThreads / c.block);
    size_t lds = sizeof(int) * kShortStack * c.block;
    if (width == 4)
        hipLaunchKernelGGL((k_wf_extend<4, false, ManyScenes>), dim3(grid), dim3(c.block), lds, c.stream,
                           ManyScenes{scenes}, q, rayIn, countIn, resetNext, c.refill, counters);
    else
        hipLaunchKernelGGL((k_wf_extend<8, false, ManyScenes>), dim3(grid), dim3(c.block), lds, c.stream,
                           ManyScenes{scenes}, q, rayIn, countIn, resetNext, c.refill, counters);
    return hipGetLastError();
}

hipError_t launchBtShade(const LaunchCfg& c, const DevScene* scenes, const RenderArgs& a, const WfState& q,
                         const float4* rayIn, const uint32_t* countIn, float4* rayOut, uint32_t* countOut,
                         uint32_t maxCount, unsigned long long* counters) {
    uint32_t grid = gridFor(maxCount, c.shadeBlock, c.maxThreads / c.shadeBlock);
    hipLaunchKernelGGL(k_wf_shade<ManyScenes>, dim3(grid), dim3(c.shadeBlock), 0, c.stream, ManyScenes{scenes}, a, q,
                       rayIn, countIn, rayOut, countOut, counters);
    return hipGetLastError();
}

hipError_t launchBtShadow(const LaunchCfg& c, const DevScene* scenes, int width, const WfState& q,
                          const float4* shadow, const uint32_t* countIn, uint32_t maxCount,
                          unsigned long long* counters) {
    uint32_t grid = gridFor(maxCount, c.block, c.maxThreads / c.block);
    size_t lds = sizeof(int) * kShortStack * c.block;
    if (width == 4)
        hipLaunchKernelGGL((k_wf_shadow<4, ManyScenes>), dim3(grid), dim3(c.block), lds, c.stream,
                           ManyScenes{scenes}, q, shadow, countIn, c.refill, counters);
    else
        hipLaunchKernelGGL((k_wf_shadow<8, ManyScenes>), dim3(grid), dim3(c.block), lds, c.stream,
                           ManyScenes{scenes}, q, shadow, countIn, c.refill, counters);
    return hipGetLastError();
}

hipError_t launchBtFinalize(const LaunchCfg& c, const DevScene* scenes, const BatchSeg* segs, uint32_t nseg,
                            uint32_t finTotal, const WfState& q, unsigned long long* counters) {
    uint32_t grid = gridFor(finTotal, c.shadeBlock, c.maxThreads / c.shadeBlock);
    hipLaunchKernelGGL(k_bt_finalize, dim3(grid), dim3(c.shadeBlock), 0, c.stream, ManyScenes{scenes}, segs, nseg,
                       finTotal, q, counters);
    return hipGetLastError();
}

}  // namespace eleven

// Host-visible kernel arguments and launchers (implemented in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace eleven {
namespace dev {
struct DevScene;
}

// MI355X dispatches workgroups to its 8 XCDs round-robin; queues and tallies
// keep one segment / copy per XCD (see kernels.hip).
constexpr int kXcds = 8;

// Largest workgroup the kernels are compiled for (launch bounds).
constexpr int kMaxBlock = 256;

// Tuned defaults (heightfield 1M tris at 1080p on MI355X): traversal block
// size and the idle-lane count that triggers an in-wave ray refill.
constexpr uint32_t kDefaultBlock = 64;
constexpr uint32_t kDefaultRefill = 16;

namespace dev {
constexpr int kAdjMaxDepthHost = 16;   // deepest path the adjoint records (k_adjoint's kAdjMaxDepth)
}

// Tally slots (per XCD copy): the host sums the kXcds copies.
constexpr int kStatSlots = 8;
constexpr int kStatRays = 0, kStatVerts = 1, kStatPaths = 2, kStatOverflow = 3, kStatNodes = 4, kStatTris = 5,
              kStatExtRays = 6;   // rays of the closest-hit (extension) launches

// Queue counters: kXcds per queue.  The shadow queues' counts come in two
// sets by bounce parity (bounce b uses set b & 1): the shadow launches of
// bounce b may run while the extension launch of bounce b + 1, which clears
// the set bounce b + 1 fills, is running.
constexpr int kQRay0 = 0, kQRay1 = kXcds, kQShadow = 2 * kXcds;
__host__ __device__ constexpr int qShadow(int parity, int light) { return kQShadow + (2 * parity + light) * kXcds; }
// Work-chunk counters of the traversal launches (extend, and per parity the
// HDRI and the light shadow), each on its own 128-byte line: a batch launch
// draws from kChunkShares shares (share x first by the blocks b with
// b % 64 == x), a single-scene launch from the first kXcds (one per XCD).
constexpr int kChunkShares = 64;
constexpr int kChunkStride = 32, kChunkExt = 0;
__host__ __device__ constexpr int chunkShadow(int parity, int light) { return 1 + 2 * parity + light; }
constexpr int kQChunk = 6 * kXcds, kQCounters = kQChunk + 5 * kChunkShares * kChunkStride;
constexpr uint32_t kChunkRays = 256;   // rays per work chunk (a multiple of 64)

struct RenderArgs {
    uint32_t k0, k1;        // RNG key (seed)
    uint32_t sample;        // first sample index of this launch / generation
    uint32_t scene;         // scene id (counter word 3)
    uint32_t maxDepth;
    uint32_t bounce;        // wavefront: current bounce
    uint32_t genPixel;      // wavefront: first work item of the generation, as padded pixel
    uint32_t genSample;     //   and sample offset (item = genSample * npad + genPixel)
    uint32_t genCount;      // wavefront: work items in the generation
    uint32_t npad;          // wavefront: padded pixel count (8x8 tiles)
    uint32_t cull;          // stop zero-throughput paths that cannot turn NaN
    float4* rec;            // per (sample, pixel): the path's radiance before the clamp, or null
    // Wavefront renders that record, and adjoints that replay, path vertices
    // (one scene; all null otherwise): the closest hit (tri, t, u, v; tri -1:
    // escaped) of vertex b of path (sample, pixel) at recHits[b * recStride +
    // sample * pixels + pixel], and per path the shadow queries that found
    // the light visible: bit 2b HDRI, bit 2b + 1 point light of vertex b.
    float4* recHits;
    uint32_t* recVis;
    uint64_t recStride;
};

struct WfState {
    float4* ray[2];         // ray queues: (o, slot) (d, -) per entry
    float4* hits;           // (tri, t, u, v) per entry of the current queue
    float4* shadowH;        // (o, slot) (d, tmax) (contribution) per entry
    float4* shadowP;
    float4* pathT;          // throughput per path slot (from the first shade on)
    float4* pathL;          // radiance per path slot (from the first shade on)
    uint32_t* pathPix;      // batches: output pixel index per path slot, 0xFFFFFFFF: padding
    uint32_t* pathSample;   // batches: sample index per path slot (one scene: slotPixel)
    uint64_t* pathQuiet;    // per path slot: bounces whose HDRI number is NaN-prone (bit b, b < 64)
    float4* pathAov;        // 3 per path slot (normal, tangent, bitangent) or null
    uint32_t* pathVis;      // recording renders: visibility bits per path slot (RenderArgs::recVis), else null
    float4* pendP;          // per path slot: the point-light term of the last shadow P launch (0: blocked), or
                            // null; with deferP the launch writes it here instead of adding it to pathL, so
                            // it can run beside the HDRI shadow launch, and shade / finalize add it in order
    uint32_t deferP;        // this bounce's point-light terms go to pendP (shade marks, shadow P writes)
    uint32_t visBit;        // shadow launches of a recording render: the bit a visible query sets
    uint32_t parity;        // bounce & 1: the shadow counter set of this bounce's launches
    uint32_t* counters;     // kQCounters: per-XCD segment counts of the queues, work-chunk counters
    uint32_t segSize;       // entries per XCD segment of a compacted queue
    int* gstack;
    uint32_t gthreads;
};

// One scene's share of a batch generation: path slots [slotStart,
// slotStart + count) hold the scene's work items wStart .. (sample-major:
// w = sample * npad + padded pixel).  finStart: prefix over the generation's
// segments of min(count, npad), the finalize threads of the segment.
struct BatchSeg {
    uint32_t slotStart, count, scene, npad;
    uint64_t wStart;
    uint32_t finStart, pad;
};

// Global traversal-stack spill areas per context (gcap x maxThreads ints
// each): launches on the main stream use the first; the shadow launches
// that run beside the next extension use the second (HDRI) and third (point
// light).  A spill slot is addressed by the thread's grid index, which
// kernels running at once share.
constexpr int kSpillSets = 3;

struct LaunchCfg {
    hipStream_t stream;
    uint32_t block;         // traversal kernels (trace, extend, shadow) and the plain megakernel
    uint32_t shadeBlock;    // camera, shade, finalize
    uint32_t maxThreads;    // grid * block upper bound (size of the spill area)
    uint32_t refill;        // persistent traversal: idle lanes that trigger a refill
    int* gstack;
};

// The environment-sample table of s (2 float4 per HDRI pixel), from the
// scene's textures; s.envTab is not read.
hipError_t launchEnvTable(hipStream_t stream, const dev::DevScene& s, float4* tab);
hipError_t launchTrace(const LaunchCfg& c, const dev::DevScene& s, const float* rays, const float* tmax, uint32_t n,
                       int32_t* tri, float* t, float* uv, uint8_t* occ, bool any, bool count,
                       unsigned long long* counters);
hipError_t launchPlain(const LaunchCfg& c, const dev::DevScene& s, const RenderArgs& a, float4* accum,
                       uint32_t* counts, float4* aov, unsigned long long* counters);
// r = 2 (mean image - target) per pixel and channel (FP64 arithmetic, FP32
// store); *loss += sum (mean image - target)^2 (the caller zeroes it).
hipError_t launchResidualL2(hipStream_t stream, const float4* accum, const uint32_t* counts, const float* target,
                            uint32_t npx, float* r, double* loss);
// rec: the forward render's per-path radiance (EL_FLAG_RECORD_PATHS) or null:
// then every path is walked once for its radiance, keeping its vertices in
// `cache` (c.maxThreads * a.maxDepth float4, vertex i of grid thread g at
// i * c.maxThreads + g), and replayed from there for the gradient.
hipError_t launchAdjoint(const LaunchCfg& c, const dev::DevScene& s, const RenderArgs& a, uint32_t spp,
                         const float* dimage, const uint32_t* counts, const float4* rec, float4* cache, int gradTex,
                         double* dmat, float* dtex, unsigned long long* counters);
// Batch (many scenes, scene id per queue entry): scenes = device array of
// DevScene; segs = the generation's segment table (device).
hipError_t launchBtCamera(const LaunchCfg& c, const dev::DevScene* scenes, const BatchSeg* segs, uint32_t nseg,
                          uint32_t genCount, uint32_t maxDepth, bool cull, const WfState& q);
hipError_t launchBtExtend(const LaunchCfg& c, const dev::DevScene* scenes, int width, const WfState& q,
                          const float4* rayIn, const uint32_t* countIn, uint32_t* resetNext, uint32_t maxCount,
                          unsigned long long* counters);
hipError_t launchBtShade(const LaunchCfg& c, const dev::DevScene* scenes, const RenderArgs& a, const WfState& q,
                         const float4* rayIn, const uint32_t* countIn, float4* rayOut, uint32_t* countOut,
                         uint32_t maxCount, unsigned long long* counters);
hipError_t launchBtShadow(const LaunchCfg& c, const dev::DevScene* scenes, int width, const WfState& q,
                          const float4* shadow, const uint32_t* countIn, uint32_t maxCount,
                          unsigned long long* counters);
hipError_t launchBtFinalize(const LaunchCfg& c, const dev::DevScene* scenes, const BatchSeg* segs, uint32_t nseg,
                            uint32_t finTotal, const WfState& q, unsigned long long* counters);
hipError_t launchWfCamera(const LaunchCfg& c, const dev::DevScene& s, const RenderArgs& a, const WfState& q);
hipError_t launchWfExtend(const LaunchCfg& c, const dev::DevScene& s, const WfState& q, const float4* rayIn,
                          const uint32_t* countIn, uint32_t* resetNext, uint32_t maxCount, bool countTraversal,
                          unsigned long long* counters);
hipError_t launchWfShade(const LaunchCfg& c, const dev::DevScene& s, const RenderArgs& a, const WfState& q,
                         const float4* rayIn, const uint32_t* countIn, float4* rayOut, uint32_t* countOut,
                         uint32_t maxCount, unsigned long long* counters);
hipError_t launchWfShadow(const LaunchCfg& c, const dev::DevScene& s, const WfState& q, const float4* shadow,
                          const uint32_t* countIn, uint32_t maxCount, unsigned long long* counters);
hipError_t launchWfFinalize(const LaunchCfg& c, const dev::DevScene& s, const RenderArgs& a, const WfState& q,
                            float4* accum, uint32_t* counts, float4* aov, unsigned long long* counters);

}  // namespace eleven

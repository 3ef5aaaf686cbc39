// HDRI importance sampling that reproduces the reference's sampler exactly.
//
// The reference draws an environment pixel with HDRI::sample(r1)
// (HDRI.hpp:154-162) = binarySearch(cdf, r1, W*H) (HDRI.hpp:130-142), and
// draws the point light from the same r1 (kernel.cu:347-348, 185).  Because
// each path's radiance is clamped to [0, 10] before accumulation
// (kernel.cu:447), the estimator's expectation depends on the joint
// distribution of (pixel, light), so a sampler with the same pixel marginal
// but a different map r1 -> pixel (an alias table) changes the image.  This
// sampler returns binarySearch(cdf, r1, L) bit for bit, in O(1) expected
// memory accesses:
//
//   * that search is not a lower bound.  For r strictly inside
//     (cdf[j], cdf[j+1]) every comparison `r < cdf[m]` is `m > j` (the CDF is
//     non-decreasing), so the result is a function open[j] of the interval
//     index; for r == cdf[j] exactly the result is tie[j], the search run on
//     that value.  Both tables are precomputed per interval.
//   * the interval j(r) = max{ j < L : cdf[j] <= r } is found with a guide
//     table over G = 2^k >= L equal cells of [0, 1) (the cutpoint method,
//     Chen & Asau 1974): guide[c] = j(c / G), then a binary search inside
//     [guide[c], guide[c+1]], which is usually empty or one step.
#pragma once

#include <cstdint>
#include <vector>

#ifdef __HIPCC__
#define EL_HD __host__ __device__
#else
#define EL_HD
#endif

namespace eleven {

// open[j]: result of the reference search for any r in (cdf[j], cdf[j+1]).
int hdriSearchIndex(int j, int length);
// The reference loop itself (HDRI.hpp:130-142), used for the tie table.
int hdriBinarySearch(const float* cdf, float value, int length);

// Pixel (qx, qy) at which HDRI::pdf (HDRI.hpp:145-152) is evaluated for the
// environment sample of index k: inverseTransformUV (Texture.hpp:177-193) of
// k's pixel corner, as the renderer's envSample computes it.
EL_HD inline void hdriPdfPixel(int k, int w, int h, float xTile, float yTile, float xOffset, float yOffset, int& qx,
                               int& qy) {
    int x = k % w, y = k / w;
    float nu = x / (float)w, nv = y / (float)h;
    int ix = (int)(nu * w), iy = (int)(nv * h);
    ix = (int)(xTile * (ix - xOffset * w)) % w;
    iy = (int)(yTile * (iy - yOffset * h)) % h;
    float iu = (float)ix / (float)w, iv = (float)iy / (float)h;
    iu += -(float)(iu > 1) + -(float)(iu < 0);
    iv += -(float)(iv > 1) + -(float)(iv < 0);
    qx = (int)(iu * w);
    qy = (int)(iv * h);
}

// Buckets of the environment random number whose sample can have a
// non-finite or zero pdf: bit b (of 2^16) covers r = v * 2^-24 for
// v in [256 b, 256 b + 255] (v = Philox word >> 8), set if any environment
// index the search can return for such an r lands on the pole row (pdf
// infinite) or on a black pixel (pdf zero).  Conservative: every bucket is
// set when the texture uses tiling or offsets, or is black overall.
std::vector<uint32_t> hdriBadBuckets(const struct HdriSampler& s, const float* rgb, int w, int h, float xTile,
                                     float yTile, float xOffset, float yOffset);

struct HdriSampler {
    std::vector<float> cdf;       // L + 1 entries, HDRI::generateCDF layout
    std::vector<int32_t> guide;   // G + 1 pairs: [2c] = interval j(c / G),
                                  // [2c + 1] = the result for every r of cell c
                                  // when it is one value (no CDF boundary in
                                  // the cell), else -1
    std::vector<int32_t> map2;    // [2j] = open[j], [2j+1] = tie[j]
    float radianceSum = 0;
    int32_t length = 0;           // L = W * H
    int32_t log2G = 0;
};

// From the RGB texels: the CDF and radiance sum exactly as HDRI::generateCDF
// computes them at construction (sequential FP32 sums in row-major order,
// before any offset is applied, HDRI.hpp:107-128), then the tables.
HdriSampler buildHdriSampler(const float* rgb, int width, int height);

// The lookup shared by host and device: binarySearch(cdf, r, L) for r in [0, 1).
// Samples concentrate on heavy CDF intervals, which span many guide cells
// with no boundary inside: those cells answer from the guide entry alone.
EL_HD inline int hdriLookup(const float* cdf, const int32_t* guide, const int32_t* map2, int log2G, float r) {
    int c = (int)(r * (float)(1 << log2G));
    const int direct = guide[2 * c + 1];
    if (direct >= 0) return direct;
    int lo = guide[2 * c], hi = guide[2 * c + 2];
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (cdf[mid] <= r)
            lo = mid;
        else
            hi = mid - 1;
    }
    return cdf[lo] == r ? map2[2 * lo + 1] : map2[2 * lo];
}

}  // namespace eleven

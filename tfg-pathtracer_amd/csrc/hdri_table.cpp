#include "hdri_table.hpp"
#include "host_parallel.hpp"

#include <algorithm>

namespace eleven {

int hdriSearchIndex(int j, int length) {
    // The reference loop with its comparisons resolved for a value strictly
    // between cdf[j] and cdf[j+1]: `value < arr[m]` <=> m > j.
    int from = 0, to = length - 1;
    while (to - from > 0) {
        int m = from + (to - from) / 2;
        if (m > j)
            to = m - 1;
        else
            from = m + 1;
    }
    return to;
}

int hdriBinarySearch(const float* arr, float value, int length) {
    int from = 0, to = length - 1;
    while (to - from > 0) {
        int m = from + (to - from) / 2;
        if (value == arr[m]) return m;
        if (value < arr[m]) to = m - 1;
        if (value > arr[m]) from = m + 1;
    }
    return to;
}

HdriSampler buildHdriSampler(const float* rgb, int w, int h) {
    HdriSampler s;
    const int L = w * h;
    s.length = L;
    s.cdf.assign((size_t)L + 1, 0.f);
    float sum = 0;
    for (int i = 0; i < L; i++) sum += rgb[3 * (size_t)i] + rgb[3 * (size_t)i + 1] + rgb[3 * (size_t)i + 2];
    for (int i = 0; i < L; i++)
        s.cdf[i + 1] = s.cdf[i] + (rgb[3 * (size_t)i] + rgb[3 * (size_t)i + 1] + rgb[3 * (size_t)i + 2]) / sum;
    s.radianceSum = sum;

    s.map2.resize(2 * (size_t)L);
    parallelFor((size_t)L, [&](size_t j0, size_t j1) {
        for (int j = (int)j0; j < (int)j1; j++) {
            s.map2[2 * (size_t)j] = hdriSearchIndex(j, L);
            s.map2[2 * (size_t)j + 1] = hdriBinarySearch(s.cdf.data(), s.cdf[j], L);
        }
    });

    int k = 0;
    while ((1 << k) < L) k++;
    s.log2G = k;
    const int G = 1 << k;
    s.guide.assign(2 * ((size_t)G + 1), -1);
    // guide[2c] = max{ j < L : cdf[j] <= c / G }; c / G is exact in FP32.
    int j = 0;
    for (int c = 0; c <= G; c++) {
        float edge = (float)c / (float)G;
        while (j + 1 < L && s.cdf[j + 1] <= edge) j++;
        s.guide[2 * (size_t)c] = j;
    }
    // Cell c = [c/G, (c+1)/G) holds one result when every r in it lies in
    // the interval j = guide[2c] strictly above cdf[j] (or the tie result is
    // the same): no later boundary below (c+1)/G and cdf[j] < c/G.
    for (int c = 0; c < G; c++) {
        const int jc = s.guide[2 * (size_t)c];
        const float edge = (float)c / (float)G;
        const bool oneInterval = jc + 1 >= L || s.cdf[jc + 1] >= (float)(c + 1) / (float)G;
        const bool noTie = s.cdf[jc] < edge || s.map2[2 * (size_t)jc] == s.map2[2 * (size_t)jc + 1];
        if (oneInterval && noTie) s.guide[2 * (size_t)c + 1] = s.map2[2 * (size_t)jc];
    }
    return s;
}

std::vector<uint32_t> hdriBadBuckets(const HdriSampler& s, const float* rgb, int w, int h, float xTile, float yTile,
                                     float xOffset, float yOffset) {
    constexpr int kBuckets = 1 << 16;
    std::vector<uint32_t> bits(kBuckets / 32, 0xFFFFFFFFu);
    const int L = s.length;
    if (L <= 0 || !(s.radianceSum > 0) || xTile != 1.f || yTile != 1.f || xOffset != 0.f || yOffset != 0.f)
        return bits;
    // bad[k]: the pdf of environment index k is infinite (row 0: sin 0) or zero
    std::vector<uint8_t> bad(L);
    for (int k = 0; k < L; k++) {
        int qx, qy;
        hdriPdfPixel(k, w, h, xTile, yTile, xOffset, yOffset, qx, qy);
        const float* p = rgb + 3 * ((size_t)qy * w + qx);
        bad[k] = qy == 0 || (p[0] + p[1] + p[2]) == 0.f;
    }
    std::fill(bits.begin(), bits.end(), 0u);
    const float* cdf = s.cdf.data();
    for (int b = 0; b < kBuckets; b++) {
        const float rlo = (float)(b * 256) * (1.0f / 16777216.0f);
        const float rhi = (float)(b * 256 + 255) * (1.0f / 16777216.0f);
        // intervals j = max{ j < L : cdf[j] <= r } over the bucket
        int jlo = (int)(std::upper_bound(cdf, cdf + L, rlo) - cdf) - 1;
        int jhi = (int)(std::upper_bound(cdf, cdf + L, rhi) - cdf) - 1;
        jlo = std::max(0, jlo);
        jhi = std::max(jlo, jhi);
        bool any = false;
        for (int j = jlo; j <= jhi && !any; j++) {
            const int a = s.map2[2 * (size_t)j], t = s.map2[2 * (size_t)j + 1];
            any = (a >= 0 && a < L && bad[a]) || (t >= 0 && t < L && bad[t]) || a < 0 || a >= L || t < 0 || t >= L;
        }
        if (any) bits[b >> 5] |= 1u << (b & 31);
    }
    return bits;
}

}  // namespace eleven

#include "hdri_table.hpp"

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
    for (int j = 0; j < L; j++) {
        s.map2[2 * (size_t)j] = hdriSearchIndex(j, L);
        s.map2[2 * (size_t)j + 1] = hdriBinarySearch(s.cdf.data(), s.cdf[j], L);
    }

    int k = 0;
    while ((1 << k) < L) k++;
    s.log2G = k;
    const int G = 1 << k;
    s.guide.resize((size_t)G + 1);
    // guide[c] = max{ j < L : cdf[j] <= c / G }; c / G is exact in FP32.
    int j = 0;
    for (int c = 0; c <= G; c++) {
        float edge = (float)c / (float)G;
        while (j + 1 < L && s.cdf[j + 1] <= edge) j++;
        s.guide[c] = j;
    }
    return s;
}

}  // namespace eleven

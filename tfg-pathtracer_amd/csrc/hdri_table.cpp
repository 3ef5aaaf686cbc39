#include "hdri_table.hpp"

#include <algorithm>

namespace eleven {

int hdriSearchIndex(int j, int length) {
    // The reference loop (HDRI.hpp:130-142) with its comparisons resolved for
    // a value strictly between cdf[j] and cdf[j+1].
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

static double clamp01(double x) { return x < 0 ? 0 : (x > 1 ? 1 : x); }

HdriTable buildHdriTable(const float* cdf, int L) {
    HdriTable t;
    t.pmf.assign((size_t)L, 0.0);
    for (int j = 0; j < L; j++) {
        double a = clamp01((double)cdf[j]);
        double b = (j == L - 1) ? 1.0 : clamp01((double)cdf[j + 1]);
        double m = b > a ? b - a : 0.0;
        if (m > 0) t.pmf[(size_t)hdriSearchIndex(j, L)] += m;
    }
    double total = 0;
    for (double p : t.pmf) total += p;
    if (total <= 0) {   // degenerate (all-black) map: uniform, matching g over equal intervals
        for (auto& p : t.pmf) p = 1.0 / L;
        total = 1.0;
    }
    for (auto& p : t.pmf) p /= total;

    // Vose's alias method.
    t.prob.assign((size_t)L, 1.f);
    t.alias.resize((size_t)L);
    std::vector<double> scaled((size_t)L);
    std::vector<int32_t> small, large;
    small.reserve(L);
    large.reserve(L);
    for (int i = 0; i < L; i++) {
        scaled[i] = t.pmf[i] * L;
        t.alias[i] = i;
        (scaled[i] < 1.0 ? small : large).push_back(i);
    }
    while (!small.empty() && !large.empty()) {
        int s = small.back();
        small.pop_back();
        int l = large.back();
        t.prob[s] = (float)scaled[s];
        t.alias[s] = l;
        scaled[l] = (scaled[l] + scaled[s]) - 1.0;
        if (scaled[l] < 1.0) {
            large.pop_back();
            small.push_back(l);
        }
    }
    for (int i : large) t.prob[i] = 1.f;
    for (int i : small) t.prob[i] = 1.f;
    return t;
}

}  // namespace eleven

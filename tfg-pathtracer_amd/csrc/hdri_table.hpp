// HDRI importance sampling as an alias table over the *effective* pixel
// distribution of the reference's sampler.
//
// The reference draws a pixel with HDRI::sample(r1) (HDRI.hpp:154-162), i.e.
// binarySearch(cdf, r1, W*H) (HDRI.hpp:130-142).  That search is not a lower
// bound: for r strictly inside (cdf[j], cdf[j+1]) every comparison
// `r < cdf[m]` is equivalent to `m > j` (the CDF is non-decreasing), so the
// returned pixel is a function g(j) of the interval index alone.  The
// effective probability of pixel k is therefore the sum of the interval
// masses cdf[j+1]-cdf[j] over all j with g(j) = k, with the last interval
// extending to 1.  The table reproduces that distribution exactly; drawing
// the bin uses the full 32-bit uniform and the acceptance test a second one.
#pragma once

#include <cstdint>
#include <vector>

namespace eleven {

// g(j): pixel returned by the reference search for any r in (cdf[j], cdf[j+1]).
int hdriSearchIndex(int j, int length);

struct HdriTable {
    std::vector<float> prob;    // acceptance probability per bin
    std::vector<int32_t> alias; // alias pixel per bin
    std::vector<double> pmf;    // effective pixel pmf (sums to 1)
};

// cdf has length+1 entries (HDRI::generateCDF layout).
HdriTable buildHdriTable(const float* cdf, int length);

}  // namespace eleven

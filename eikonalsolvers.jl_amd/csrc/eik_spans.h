// eik_spans.h -- total length of a set of time intervals (the round launches of one call, on one
// stream or on the two streams of a split call): the time in which at least one of them ran.
// Plain host code with no HIP dependency.
#pragma once
#include <algorithm>
#include <utility>
#include <vector>

namespace eik {

// length of the union of [start, end] intervals, in their own unit; an interval whose end lies
// before its start counts as empty
inline double span_union(std::vector<std::pair<float, float>> v) {
    std::sort(v.begin(), v.end());
    double tot = 0.0, lo = 0.0, hi = 0.0;
    bool open = false;
    for (const auto& x : v) {
        const double a = x.first, b = std::max(x.first, x.second);
        if (!open || a > hi) {
            if (open) tot += hi - lo;
            lo = a;
            hi = b;
            open = true;
        } else if (b > hi) {
            hi = b;
        }
    }
    if (open) tot += hi - lo;
    return tot;
}

}  // namespace eik

// Extremum graph simplification (host).
#pragma once
#include <stdint.h>

#include <vector>

namespace msceer {

struct MinNode {
    int64_t cell;  // critical vertex cell id
    float value;   // value at that vertex
};

struct SaddleArc {
    int64_t saddle;      // critical 1-cell id
    int64_t min1, min2;  // critical vertex cells its two V-paths end at
    float value;         // value of the saddle's highest vertex
};

// Representative node index per node after cancelling every arc whose
// persistence does not exceed `threshold`.
std::vector<int> simplify_minima(const std::vector<MinNode>& nodes, const std::vector<SaddleArc>& saddles,
                                 float threshold);

}  // namespace msceer

// Extremum graph simplification (host).
#pragma once
#include <stdint.h>

#include <vector>

namespace msceer {

struct MinNode {
    int64_t cell;    // critical cell id (a vertex cell; a hex for the max graph)
    float value;     // value of its highest vertex
    int64_t vertex;  // that vertex (order ties: value, then vertex, then cell)
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
// The max graph (UFMergeGraph MAXIMAL): the higher extremum of a merge stays
// representative; arcs join maxima (min1/min2 hold their cells).
std::vector<int> simplify_maxima(const std::vector<MinNode>& nodes, const std::vector<SaddleArc>& saddles,
                                 float threshold);

}  // namespace msceer

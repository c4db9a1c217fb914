// Host side of the extremum graph: persistence simplification of the minima
// joined by 1-saddles, or of the maxima joined by 2-saddles (UFMergeGraph
// MINIMAL / MAXIMAL + SimplifyToThreshold,
// gi_extrema_region_builder.h:28-268).  Inputs are small (one node per
// critical vertex, one arc per 1-saddle) and the cancellation order is a
// global priority order, so this runs on the host between GPU stages.
#include <algorithm>
#include <cmath>
#include <queue>
#include <unordered_map>
#include <vector>

#include "extgraph.hpp"

namespace msceer {

namespace {

struct Arc {
    int64_t saddle;
    int e1, e2;
    float pers, sval;
};

// lowest persistence first, ties by the lower saddle cell id
struct ArcLater {
    bool operator()(const Arc& l, const Arc& r) const {
        if (l.pers < r.pers) return false;
        if (l.pers > r.pers) return true;
        return l.saddle > r.saddle;
    }
};

struct Graph {
    const std::vector<MinNode>* nodes;
    std::vector<int> rep;
    std::priority_queue<Arc, std::vector<Arc>, ArcLater> arcs;
    float thr = 0.f;
    bool minimal = true;  // MINIMAL: the lower extremum of a merge stays representative; MAXIMAL: the higher

    // lessThan (gi_topological_max_vertex_mesh_function.h:45-62): value of
    // the highest vertex, then that vertex's id, then the cell id
    bool less(int a, int b) const {
        const MinNode& x = (*nodes)[a];
        const MinNode& y = (*nodes)[b];
        if (x.value < y.value) return true;
        if (y.value < x.value) return false;
        if (x.vertex != y.vertex) return x.vertex < y.vertex;
        return x.cell < y.cell;
    }
    int find(int l) {
        int r = l;
        while (rep[r] != r) r = rep[r];
        while (rep[l] != r) { const int n = rep[l]; rep[l] = r; l = n; }
        return r;
    }
    // MergeByVal (gi_extrema_region_builder.h:84-94): the later extremum
    // joins the earlier one's set (lower for minima, higher for maxima)
    void merge(int a, int b) {
        if (less(a, b) ^ minimal) rep[a] = rep[b];
        else rep[b] = rep[a];
    }
    void cancel(Arc m) {
        if (m.pers > thr) return;
        const int r1 = find(m.e1), r2 = find(m.e2);
        if (r1 == r2) return;
        // current representatives: the reference reads them crossed
        const float v1 = (*nodes)[r2].value, v2 = (*nodes)[r1].value;
        if (m.e1 != r1 || m.e2 != r2) {
            const float d1 = std::fabs(m.sval - v1), d2 = std::fabs(m.sval - v2);
            if (d1 < d2) {
                if (d1 > thr) return;
                m.pers = d1; m.e1 = r2; m.e2 = r1;
            } else {
                if (d2 > thr) return;
                m.pers = d2; m.e1 = r1; m.e2 = r2;
            }
            arcs.push(m);
            return;
        }
        merge(r1, r2);
    }
};

}  // namespace

namespace {

std::vector<int> simplify(const std::vector<MinNode>& nodes, const std::vector<SaddleArc>& saddles, float threshold,
                          bool minimal) {
    Graph g;
    g.nodes = &nodes;
    g.minimal = minimal;
    g.rep.resize(nodes.size());
    for (size_t i = 0; i < nodes.size(); i++) g.rep[i] = (int)i;
    // node index of a cell: nodes come in ascending cell order
    const bool sorted = std::is_sorted(nodes.begin(), nodes.end(),
                                       [](const MinNode& a, const MinNode& b) { return a.cell < b.cell; });
    std::unordered_map<int64_t, int> id;
    if (!sorted) {
        id.reserve(nodes.size() * 2);
        for (size_t i = 0; i < nodes.size(); i++) id[nodes[i].cell] = (int)i;
    }
    auto index_of = [&](int64_t cell) -> int {
        if (!sorted) {
            auto it = id.find(cell);
            return it == id.end() ? -1 : it->second;
        }
        auto it = std::lower_bound(nodes.begin(), nodes.end(), cell,
                                   [](const MinNode& a, int64_t c) { return a.cell < c; });
        return (it != nodes.end() && it->cell == cell) ? (int)(it - nodes.begin()) : -1;
    };
    std::vector<Arc> start;
    for (const SaddleArc& s : saddles) {
        if (s.min1 == s.min2 || s.min1 < 0 || s.min2 < 0) continue;
        const int i1 = index_of(s.min1), i2 = index_of(s.min2);
        if (i1 < 0 || i2 < 0) continue;
        Arc a;
        a.saddle = s.saddle;
        a.e1 = i1;
        a.e2 = i2;
        a.sval = s.value;
        float d1 = nodes[a.e1].value - a.sval, d2 = nodes[a.e2].value - a.sval;
        if (d1 < 0) d1 *= -1;
        if (d2 < 0) d2 *= -1;
        a.pers = d1 < d2 ? d1 : d2;
        // an arc above the threshold is popped after every cancellable one and
        // then dropped unprocessed, so it never needs to enter the queue
        if (a.pers > threshold) continue;
        start.push_back(a);
    }
    g.arcs = std::priority_queue<Arc, std::vector<Arc>, ArcLater>(ArcLater(), std::move(start));
    g.thr = threshold;
    while (!g.arcs.empty()) {
        const Arc a = g.arcs.top();
        g.arcs.pop();
        g.cancel(a);
    }
    std::vector<int> out(nodes.size());
    for (size_t i = 0; i < nodes.size(); i++) out[i] = g.find((int)i);
    return out;
}

}  // namespace

std::vector<int> simplify_minima(const std::vector<MinNode>& nodes, const std::vector<SaddleArc>& saddles,
                                 float threshold) {
    return simplify(nodes, saddles, threshold, true);
}

std::vector<int> simplify_maxima(const std::vector<MinNode>& nodes, const std::vector<SaddleArc>& saddles,
                                 float threshold) {
    return simplify(nodes, saddles, threshold, false);
}

}  // namespace msceer

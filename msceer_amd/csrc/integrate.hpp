// Launchers of the accurate-labelling kernels (integrate.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msceer {

struct MList;

// Vertex grid geometry for kernels.
struct VGrid {
    int X, Y, Z;
    int64_t sxy;  // X*Y
    __host__ __device__ int64_t verts() const { return sxy * Z; }
};

inline VGrid make_vgrid(int64_t x, int64_t y, int64_t z) {
    VGrid g;
    g.X = (int)x; g.Y = (int)y; g.Z = (int)z;
    g.sxy = x * y;
    return g;
}

// Device buffers of one labelling run (all sized to the vertex count V unless noted).
struct LabelBuffers {
    const float* values;
    float4* gradf;       // FP32 gradient field
    double4* gradd;      // FP64 gradient field (pass 3)
    int32_t* label;      // region label per vertex (-1 unassigned)
    int32_t* label2;     // scratch labels (pass 3: end vertices of streamlines traced ahead)
    uint8_t* dest;       // DestType per vertex
    uint8_t* seedmask;   // 1 at region seeds (their gradient is zero)
    uint8_t* redo;       // 1: re-trace in FP64 (pass 1 ran out of budget)
    uint32_t* stamp;     // work-list dedupe stamps
    const int32_t* snap; // labels when pass 3 begins (read for re-integration results)
    int32_t* rep;        // component representatives (cleanup)
    int32_t* rep2;
    uint32_t* stats;     // device counters (see IntegStat)
    int32_t desc;        // cleanup order: 0 ascending labels, 1 descending (greater vertex first)
};

enum IntegStat : int {
    kStatTraced1 = 0,    // streamlines traced in pass 1
    kStatOver1 = 1,      // pass-1 streamlines that ran out of step budget
    kStatTraced3 = 2,    // (unused)
    kStatOver3 = 3,      // pass-3 FP64 streamlines that ran out of step budget
    kStatChanged = 4,    // sweep / round change flag
    kStatMasked = 5,     // unassigned vertices inside the integration mask
    kStatAhead3 = 6,     // pass-3 streamlines traced ahead of a chain
    kStatTraced3f = 7,   // re-integrations in pass 3
    kStatDoubt = 12,     // FP32-first: re-integrations traced again in FP64 (a near-tie or the step budget)
    kStatSteps1 = 8,     // uint64 (two words): pass-1 steps, accepted and rejected
    kStatSteps3 = 10,    // uint64: pass-3 FP64 steps
    kStatSteps3f = 14,   // uint64: pass-3 FP32 steps
    kStatNum = 16
};

hipError_t launch_grad_field(const VGrid& g, const float* values, float4* gradf, hipStream_t s);
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s);
// dst[i] |= src[i] for n bytes.
hipError_t launch_or_u8(uint8_t* dst, const uint8_t* src, int64_t n, hipStream_t s);
// b.gradf = 0 where b.seedmask is set (the region extrema).
hipError_t launch_zero_seed_gradient(const VGrid& g, const LabelBuffers& b, hipStream_t s);
hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
// Seeds: label[seed_vert[i]] = seed_label[i]; zero gradient; seedmask = 1.
hipError_t launch_seed_scatter(const int32_t* seed_vert, const int32_t* seed_label, int n, const LabelBuffers& b,
                               hipStream_t s);
// dest = UNASSIGNED where label == -1 else CERTAIN_TERMINAL; redo = 0; stamp = no claim.
hipError_t launch_dest_init(const VGrid& g, const LabelBuffers& b, hipStream_t s);
// Pass 1 at one level of the coarse-to-fine order (sign -1: flow downhill to
// minima, +1: uphill to maxima; the same for pass 3).
hipError_t launch_pass1(const VGrid& g, const LabelBuffers& b, int stride, bool first_level, float gthr, float ethr,
                        int max_iters, int sign, hipStream_t s);
// End of a pass-1 level: its seeds and claimed path vertices become assigned
// (claims in b.stamp, all kNoClaim on entry; reset to zero after the last
// level for passes 2 and 3).
hipError_t launch_pass1_finish(const VGrid& g, const LabelBuffers& b, int stride, bool first_level, bool last,
                               hipStream_t s);
// Pass 2: every ASSIGNED vertex with a differently labelled All6 neighbour
// (or flagged for re-tracing) goes to `out`.
// Re-integration work queue: `items` (capacity entries, 0xFFFFFFFF until
// written) and its claim/reserve counters.
struct WorkQueue {
    uint32_t* items;
    uint32_t capacity;
    uint32_t* head;         // next slot of the items queued during pass 3
    uint32_t* tail;         // slots reserved
    uint32_t* overflow;
    uint32_t* ahead_on;     // set once pass 3 traces vertices ahead
    uint32_t* chunk_head;   // next chunk of the items pass 2 queued
    uint32_t* first_count;  // items pass 2 queued
};
hipError_t launch_pass2(const VGrid& g, const LabelBuffers& b, const WorkQueue& q, hipStream_t s);
// Certain regions in dependency order (label[v] becomes the label shared by
// all earlier All6 neighbours of v when they all carry one; seeds are
// labelled beforehand; cnt: one uint32 per vertex; lists of 64-bit entries,
// capacity entries each).  init lists the vertices
// with no earlier non-seed neighbour; each round resolves the list `in` and
// appends the vertices it releases to `out`, and zeroes *clear_count
// (expected: about how many vertices `in` holds, which sizes the launch).
hipError_t launch_certain_init(const VGrid& g, const float* values, const uint8_t* seedmask, uint32_t* cnt,
                               uint64_t* out, uint32_t* out_count, uint32_t capacity, uint32_t* overflow,
                               hipStream_t s, bool same_bdry = false, bool desc = false);
// NumericIntegratorExpandingRegionStop (minima, downhill): extrema (no
// SameBdry6 neighbour before the vertex) labelled by their own ids, the rest
// -1, counts[0] += extrema; stop states from labels (certain: terminal),
// counts[1] += certain; every unassigned vertex traced alone in FP64 with
// `budget` iterations, lab_out[v] = lab_in[where it stops], counts[slot] +=
// streamlines; unambiguous vertices (all SameBdry6 neighbours share the
// label) made terminals, counts[3] += them.
hipError_t launch_ers_seeds(const VGrid& g, const float* values, uint8_t* seedmask, int32_t* label, uint32_t* counts,
                            hipStream_t s);
hipError_t launch_ers_state(const VGrid& g, const int32_t* label, uint8_t* dest, uint32_t* counts, hipStream_t s);
hipError_t launch_ers_trace(const VGrid& g, const double4* gradd, const uint8_t* dest, const int32_t* lab_in,
                            int32_t* lab_out, float gthr, float ethr, int budget, uint32_t* counts, int slot,
                            hipStream_t s);
hipError_t launch_ers_mark(const VGrid& g, const int32_t* label, uint8_t* dest, uint32_t* counts, hipStream_t s);
hipError_t launch_certain_round(const VGrid& g, int32_t* label, uint32_t* cnt, const uint64_t* in,
                                const uint32_t* in_count, uint64_t* out, uint32_t* out_count, uint32_t* clear_count,
                                uint32_t capacity, uint32_t* overflow, uint64_t expected, hipStream_t s);
hipError_t launch_grad_field64(const VGrid& g, const LabelBuffers& b, hipStream_t s);
// Pass 3 over the queue filled by pass 2 (zero head, tail and chunk_head
// before pass 2): one persistent launch, FP64 re-integration, or with
// fp32_first FP32 re-integration and FP64 where the FP32 trace is in doubt.
hipError_t launch_pass3(const VGrid& g, const LabelBuffers& b, const WorkQueue& q, bool fp32_first, float gthr,
                        float ethr, int max_iters, int sign, hipStream_t s);
// Cleanup of isolated components.
hipError_t launch_cc_init(const VGrid& g, const LabelBuffers& b, hipStream_t s);
// Same-label components (rep[] = earliest vertex of each component).
hipError_t launch_cc_union(const VGrid& g, const LabelBuffers& b, hipStream_t s);
// ints per collected candidate root: root, label, then for six earlier
// neighbours their component roots, those roots' candidate slots, their labels
constexpr int kCollectStride = 20;
// Flags the component roots that may merge (cand, one word per vertex) and
// counts them; maps every root to itself in rep2.
hipError_t launch_cc_candidates(const VGrid& g, const LabelBuffers& b, uint32_t* cand, uint32_t* count,
                                hipStream_t s);
hipError_t launch_cc_collect(const VGrid& g, const LabelBuffers& b, const uint32_t* cand, const uint32_t* slot_of,
                             int32_t* out, hipStream_t s);
// Exclusive prefix sum of n words (hipCUB); temp == nullptr queries *temp_bytes.
hipError_t launch_exclusive_sum(const uint32_t* in, uint32_t* out, int64_t n, void* temp, size_t* temp_bytes,
                                hipStream_t s);
hipError_t launch_cc_scatter(const int32_t* pairs, uint32_t n, int32_t* rep2, hipStream_t s);
hipError_t launch_cc_apply(const VGrid& g, const LabelBuffers& b, hipStream_t s);
// Restriction labels from vertex labels (MIN boundary): 1 on every cell
// with an edge whose two vertices carry valid, different labels.
hipError_t launch_min_boundary(const VGrid& g, const int32_t* label, int8_t* restriction, hipStream_t s);
// MAX boundary (descending manifolds) added onto `restriction`: +2 on interior
// quads between hexes whose highest vertices carry different valid labels,
// then on interior edges under such a quad, then on interior vertices under
// such an edge.
hipError_t launch_max_boundary(const VGrid& g, const float* values, const int32_t* label, int8_t* restriction,
                               hipStream_t s);
// Region boundary kinds (regions.hip): per cell, 0..3 by how many regions
// of `label` meet there (ComputeRegionBoundaryKind).
hipError_t launch_region_kind(const VGrid& g, const int32_t* label, int8_t* out, hipStream_t s);
// One arc of the extremum graph: a critical 1-cell (saddle), the critical
// vertices its two V-paths end at, the value of its highest vertex and the
// values at the two minima.
struct ArcRecord {
    int64_t saddle, min1, min2;
    float sval, v1, v2, pad;
};
// Follow the V-paths below both vertices of every critical 1-cell of `list`
// to critical vertices and append the arcs that join two different minima
// with persistence min(|v1 - sval|, |v2 - sval|) not above `threshold` to
// out[0..cap) (count: appended arcs, also those beyond cap; order arbitrary).
hipError_t launch_saddle_arcs(const VGrid& g, const float* values, const uint8_t* grad, const MList& list,
                              float threshold, ArcRecord* out, uint32_t cap, uint32_t* count, hipStream_t s);
// The same for the max extremum graph: critical quads off the domain
// boundary traced up through their two hexes to critical hexes (maxima);
// h1, h2: the maxima's highest vertices.
struct MaxArcRecord {
    int64_t saddle, max1, max2, h1, h2;
    float sval, v1, v2, pad;
};
hipError_t launch_saddle2_arcs(const VGrid& g, const float* values, const uint8_t* grad, const MList& list,
                               float threshold, MaxArcRecord* out, uint32_t cap, uint32_t* count, hipStream_t s);
// The highest vertex of every critical hex, appended to `out`.
hipError_t launch_collect_maxima(const VGrid& g, const float* values, const uint8_t* grad, const MList& out,
                                 uint32_t* overflow, hipStream_t s);
// Critical vertices of a .grad (pair code 7 at even cells) appended to `out`.
hipError_t launch_collect_minima(const VGrid& g, const uint8_t* grad, const MList& out, uint32_t* overflow,
                                 hipStream_t s);
// Every vertex of `minima` becomes a seed labelled by itself (zero gradient,
// seed mask set).
hipError_t launch_seed_minima(const MList& minima, const LabelBuffers& b, hipStream_t s);
// Outside the integration mask, unassigned vertices take their
// combinatorial label (b.label2) and count as assigned; counts the rest.
hipError_t launch_apply_mask(const VGrid& g, const LabelBuffers& b, const uint8_t* mask, hipStream_t s);
// Combinatorial labels: every vertex follows its V-path down to a critical
// vertex and takes that vertex's region label (for vertices not integrated).
hipError_t launch_combinatorial_labels(const VGrid& g, const uint8_t* grad, const int32_t* min_label, int32_t* out,
                                       const uint8_t* only_where_zero, hipStream_t s);

}  // namespace msceer

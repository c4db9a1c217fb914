// Accurate ascending-manifold labelling on gfx950: gradient field, certain
// regions, the adaptive-Euler streamline integrator, region cleanup and the
// restriction labels.
//
// Semantics restated (reference file:line):
//   gradient:   gi_regular_grid_trilinear_function.h:285-341 (central
//               differences, zero on the boundary of an axis)
//   advection:  gi_adaptive_euler_advector.h:51-305 (one voxel per call, step
//               doubling, step cap 0.5 voxel, x1.8 growth, halving on reject)
//   checkers:   gi_advection_checker.h:212-262
//   integrator: gi_numeric_integrator_path_compressing.h:75-433 (pass 1 with
//               path compression, pass 2 boundary collection, pass 3
//               re-integration to the original certain regions)
//   regions:    gi_extrema_region_builder.h:806-1042
//   cleanup:    gi_isolated_region_remover.h:155-192
//   boundary:   gi_labeling_to_bounary_labeling.h:48-116
//
// Precision: pass 1 traces in FP32 with an FP32 gradient field; pass 3
// (vertices next to a basin boundary, and streamlines that ran out of step
// budget in pass 1) re-traces in FP64 with the gradient recomputed in FP64
// from the values, which is how the reference computes it.  Products and sums
// are not fused into FMAs here: each rounds as in the reference's scalar
// arithmetic, so a streamline's positions follow the reference's operation by
// operation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <hipcub/hipcub.hpp>

#include "integrate.hpp"
#include "mlist.hpp"
#include "tracer.hpp"

#pragma clang fp contract(off)

namespace msceer {

namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ bool before(const float* f, int64_t a, int64_t b) {
    const float fa = f[a], fb = f[b];
    if (fb > fa) return true;
    if (fa > fb) return false;
    return b > a;
}

// The order of a descending labelling (IndexCompareGreaterThan,
// gi_numeric_integrator_expanding_region_stop.h:41-59): a vertex comes first
// when it is greater.  desc false: before().
__device__ __forceinline__ bool ord_before(const float* f, int64_t a, int64_t b, int desc) {
    return desc ? before(f, b, a) : before(f, a, b);
}

// ---- gradient field ---------------------------------------------------------

// Coordinates of vertex v in 32-bit arithmetic: every grid here has fewer
// than 2^32 vertices (its cells number fewer than 2^32), and 64-bit integer
// division is a long instruction sequence on the GPU.
__device__ __forceinline__ void vertex_xyz(const VGrid& g, int64_t v, int& x, int& y, int& z) {
    const uint32_t u = (uint32_t)v, X = (uint32_t)g.X, Y = (uint32_t)g.Y;
    const uint32_t q = u / X, r = q / Y;
    x = (int)(u - q * X);
    y = (int)(q - r * Y);
    z = (int)r;
}

// Coordinates of cell c of an Mx x My x Mz cell grid (fewer than 2^32 cells)
// in 32-bit arithmetic.
__device__ __forceinline__ void cell_xyz(int64_t c, int64_t Mx, int64_t My, int64_t& x, int64_t& y, int64_t& z) {
    const uint32_t u = (uint32_t)c, mx = (uint32_t)Mx, my = (uint32_t)My;
    const uint32_t q = u / mx, r = q / my;
    x = (int64_t)(u - q * mx);
    y = (int64_t)(q - r * my);
    z = (int64_t)r;
}

__global__ void __launch_bounds__(kThreads)
grad_field_kernel(const float* __restrict__ f, VGrid g, float4* __restrict__ out) {
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        int x, y, z;
        vertex_xyz(g, i, x, y, z);
        float4 r;
        r.x = (x > 0 && x < g.X - 1) ? (float)(0.5 * ((double)f[i + 1] - (double)f[i - 1])) : 0.f;
        r.y = (y > 0 && y < g.Y - 1) ? (float)(0.5 * ((double)f[i + g.X] - (double)f[i - g.X])) : 0.f;
        r.z = (z > 0 && z < g.Z - 1) ? (float)(0.5 * ((double)f[i + g.sxy] - (double)f[i - g.sxy])) : 0.f;
        r.w = 0.f;
        out[i] = r;
    }
}

__global__ void seed_scatter_kernel(const int32_t* __restrict__ seed_vert, const int32_t* __restrict__ seed_label,
                                    int n, int32_t* __restrict__ label, float4* __restrict__ gradf,
                                    uint8_t* __restrict__ seedmask) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t v = seed_vert[i];
    label[v] = seed_label[i];
    if (gradf) gradf[v] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (seedmask) seedmask[v] = 1;
}

__global__ void zero_seed_gradient_kernel(VGrid g, const uint8_t* __restrict__ seedmask, float4* __restrict__ gradf) {
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (seedmask[i]) gradf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ void fill_i32_kernel(int32_t* p, int64_t n, int32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void or_u8_kernel(uint8_t* dst, const uint8_t* src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] |= src[i];
}
__global__ void fill_u8_kernel(uint8_t* p, int64_t n, uint8_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

// ---- certain regions --------------------------------------------------------

// (The certain-region kernels follow the work-list helpers below.)

// ---- streamline integration -------------------------------------------------

// FP64 gradient field, as the reference computes it: FP32 coefficients times
// the values (exact), summed in FP64; zero on boundary axes and at region
// extrema (seedmask).
__global__ void __launch_bounds__(kThreads)
grad_field64_kernel(const float* __restrict__ f, VGrid g, const uint8_t* __restrict__ seedmask,
                    double4* __restrict__ out) {
    const int64_t n = g.verts();
    for (int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x; v < n; v += (int64_t)gridDim.x * kThreads) {
        double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
        if (!seedmask || !seedmask[v]) {
            int x, y, z;
            vertex_xyz(g, v, x, y, z);
            r.x = (x > 0 && x < g.X - 1) ? (0.0 + (double)(-0.5f * f[v - 1])) + (double)(0.5f * f[v + 1]) : 0.0;
            r.y = (y > 0 && y < g.Y - 1) ? (0.0 + (double)(-0.5f * f[v - g.X])) + (double)(0.5f * f[v + g.X]) : 0.0;
            r.z = (z > 0 && z < g.Z - 1) ? (0.0 + (double)(-0.5f * f[v - g.sxy])) + (double)(0.5f * f[v + g.sxy])
                                         : 0.0;
        }
        out[v] = r;
    }
}

// Pass 1 at one level of the coarse-to-fine order: the seeds are the
// unassigned vertices whose coordinates are multiples of `stride` (and not
// all multiples of 2*stride unless this is the first level).
//
// Levels are synchronous, so that the labels do not depend on the order in
// which streamlines run: a streamline stops on a certain terminal or on a
// vertex assigned by an EARLIER level (dest becomes kAssigned only in
// pass1_finish_kernel, between levels), never on a path of its own level.
// The seed takes the label its streamline reaches, written at once (no
// streamline of the level reads it).  Every other vertex among the first
// kPath of the path claims it with seed id + 1 (atomic minimum on the claim
// array), and pass1_finish_kernel gives it the label of its lowest claiming
// seed.  The labels a level reads were all written by earlier launches.
// Streamlines that run out of step budget are flagged for FP64 re-tracing.
// (The reference labels paths as each streamline ends and stops on any
// assigned vertex, gi_numeric_integrator_path_compressing.h:176-247, which
// depends on the schedule of its threads.)
constexpr int kPath = 32;
constexpr uint32_t kNoClaim = 0xFFFFFFFFu;

__device__ __forceinline__ bool pass1_seed(int x, int y, int z, int stride, int first_level) {
    if (x % stride || y % stride || z % stride) return false;
    return first_level || (x % (2 * stride)) || (y % (2 * stride)) || (z % (2 * stride));
}

__global__ void __launch_bounds__(kThreads)
pass1_kernel(GradSrc<float> src, VGrid g, LabelBuffers b, int stride, int first_level, float dir, float gthr,
             float ethr, int max_iters) {
    __shared__ int32_t path_mem[kPath * kThreads];
    int32_t* path = path_mem + threadIdx.x;
    uint32_t* const claim = b.stamp;
    const int nx = (g.X + stride - 1) / stride, ny = (g.Y + stride - 1) / stride, nz = (g.Z + stride - 1) / stride;
    const int64_t n = (int64_t)nx * ny * nz;
    uint32_t traced = 0, over = 0;
    uint64_t steps = 0;
    int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t t_step = (int64_t)gridDim.x * kThreads;
    Tracer<float> tr;
    bool active = false;
    int np = 0;
    int64_t v0 = 0;
    // a lane whose streamline ended takes its next seed at once: no
    // streamline waits for another of its level
    while (true) {
        while (!active && t < n) {
            const int x = (int)(t % nx) * stride, y = (int)((t / nx) % ny) * stride;
            const int z = (int)(t / ((int64_t)nx * ny)) * stride;
            t += t_step;
            v0 = (int64_t)x + (int64_t)g.X * ((int64_t)y + (int64_t)g.Y * z);
            if (pass1_seed(x, y, z, stride, first_level) && b.dest[v0] == kUnassigned) {
                traced++;
                np = 0;
                path[(np++) * kThreads] = (int32_t)v0;
                tracer_start(src, g, tr, x, y, z, max_iters);
                active = true;
            }
        }
        if (!__any(active)) break;
        if (!active) continue;
        int64_t next = 0;
        bool moved = false;
        int ev = tracer_iterate(src, g, b.dest, true, dir, gthr, ethr, tr, next, moved);
        if (ev == kOutOfVoxel) {
            next = tracer_reenter(src, g, tr);
            moved = true;
            ev = kNone;
        }
        if (moved && np < kPath) path[(np++) * kThreads] = (int32_t)next;
        if (ev == kNone) continue;
        active = false;
        if (ev != kHitPreassigned) next = nearest(g, tr.cur, tr.near);
        if (np < kPath) path[(np++) * kThreads] = (int32_t)next;
        // the label reached: a certain terminal's or an earlier level's; a
        // stop elsewhere (low gradient, step budget) at a vertex without one
        // gives -1, the reference's label of an unassigned vertex
        const uint8_t dn = b.dest[next];
        b.label[v0] = (dn == kAssigned || dn == kCertainTerminal) ? b.label[next] : -1;
        for (int j = 1; j < np; j++) {
            const int32_t p = path[j * kThreads];
            if (p != (int32_t)v0 && b.dest[p] == kUnassigned)
                __hip_atomic_fetch_min(&claim[p], (uint32_t)v0 + 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        steps += (uint64_t)(max_iters - (tr.iters > 0 ? tr.iters : 0));
        if (ev == kOverMax) {
            over++;
            b.redo[v0] = 1;
        }
    }
    if (traced) atomicAdd(&b.stats[kStatTraced1], traced);
    if (over) atomicAdd(&b.stats[kStatOver1], over);
    if (steps) atomicAdd(reinterpret_cast<unsigned long long*>(&b.stats[kStatSteps1]), (unsigned long long)steps);
}

// End of a pass-1 level: its seeds (labelled by their own streamlines) and
// every claimed vertex become assigned; a claimed vertex takes the label of
// its lowest claiming seed.  Claims are reset (to zero after the last level:
// the array then serves as pass 2/3's queue stamps).
__global__ void __launch_bounds__(kThreads)
pass1_finish_kernel(VGrid g, LabelBuffers b, int stride, int first_level, int last) {
    uint32_t* const claim = b.stamp;
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const uint32_t c = claim[i];
        if (b.dest[i] == kUnassigned) {
            int x, y, z;
            vertex_xyz(g, i, x, y, z);
            if (pass1_seed(x, y, z, stride, first_level)) {
                b.dest[i] = kAssigned;
            } else if (c != kNoClaim) {
                b.label[i] = b.label[c - 1u];
                b.dest[i] = kAssigned;
            }
        }
        claim[i] = last ? 0u : kNoClaim;
    }
}

// Block-staged appends to a sub-list (as in the ldir rounds).
constexpr int kStage = 2048;
struct Outbox {
    uint32_t* stage;
    uint32_t* n;
    MList out;
    int sub;
    uint32_t* overflow;
    __device__ __forceinline__ void push(uint32_t item) const {
        const uint32_t slot = atomicAdd(n, 1u);
        if (slot < (uint32_t)kStage) {
            stage[slot] = item;
        } else {
            if (!mlist_push(out, sub, item)) atomicOr(overflow, 1u);
        }
    }
};

__device__ __forceinline__ int all6(const VGrid& g, int64_t i, int64_t nb[6]) {
    int x, y, z;
    vertex_xyz(g, i, x, y, z);
    int k = 0;
    if (x < g.X - 1) nb[k++] = i + 1;
    if (x > 0) nb[k++] = i - 1;
    if (y < g.Y - 1) nb[k++] = i + g.X;
    if (y > 0) nb[k++] = i - g.X;
    if (z < g.Z - 1) nb[k++] = i + g.sxy;
    if (z > 0) nb[k++] = i - g.sxy;
    return k;
}

constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;             // queue slot not written yet
constexpr uint64_t kIdleTicks = 400000000ull;  // 4 s of the 100 MHz real-time counter

__device__ __forceinline__ uint32_t wave_sum(uint32_t x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return x;
}

__device__ __forceinline__ uint32_t wave_exclusive_scan(uint32_t x, int lane) {
    uint32_t incl = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o);
        if (lane >= o) incl += y;
    }
    return incl - x;
}

// ---- NumericIntegratorExpandingRegionStop --------------------------------
// (gi_numeric_integrator_expanding_region_stop.h:63-440; minima, downhill)

// SameBdry6 neighbours of (x, y, z): the axes along which it is interior.
__device__ __forceinline__ int same_bdry6(const VGrid& g, int64_t v, int x, int y, int z, int64_t out[6]) {
    int k = 0;
    if (x > 0 && x < g.X - 1) { out[k++] = v + 1; out[k++] = v - 1; }
    if (y > 0 && y < g.Y - 1) { out[k++] = v + g.X; out[k++] = v - g.X; }
    if (z > 0 && z < g.Z - 1) { out[k++] = v + g.sxy; out[k++] = v - g.sxy; }
    return k;
}

// Extrema (:94-107, :218-233): no SameBdry6 neighbour before the vertex; each
// labelled by its own id, the rest -1.  counts[0]: extrema.
__global__ void __launch_bounds__(kThreads)
ers_seed_kernel(const float* __restrict__ f, VGrid g, uint8_t* __restrict__ seedmask, int32_t* __restrict__ label,
                uint32_t* __restrict__ counts) {
    uint32_t found = 0;
    for (int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x; v < g.verts(); v += (int64_t)gridDim.x * kThreads) {
        int x, y, z;
        vertex_xyz(g, v, x, y, z);
        int64_t nb[6];
        const int k = same_bdry6(g, v, x, y, z, nb);
        bool ext = true;
        for (int j = 0; j < k; j++)
            if (before(f, nb[j], v)) ext = false;
        seedmask[v] = ext ? 1 : 0;
        label[v] = ext ? (int32_t)v : -1;
        found += ext ? 1u : 0u;
    }
    if (found) atomicAdd(&counts[0], found);
}

// Stop states of the streamlines: certain vertices are terminals
// (TerminateNearCertain, gi_advection_checker.h:183-208).  counts[1]: certain.
__global__ void __launch_bounds__(kThreads)
ers_state_kernel(VGrid g, const int32_t* __restrict__ label, uint8_t* __restrict__ dest, uint32_t* __restrict__ counts) {
    uint32_t n = 0;
    for (int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x; v < g.verts(); v += (int64_t)gridDim.x * kThreads) {
        const bool c = label[v] != -1;
        dest[v] = c ? kCertainTerminal : kUnassigned;
        n += c ? 1u : 0u;
    }
    if (n) atomicAdd(&counts[1], n);
}

// Every vertex still unassigned traces its own streamline in FP64 (budget
// `budget` iterations) to a certain terminal, or wherever it stops, and takes
// lab_in of that vertex.  counts[slot]: streamlines.
__global__ void __launch_bounds__(kThreads)
ers_trace_kernel(GradSrc<double> src, VGrid g, const uint8_t* __restrict__ dest, const int32_t* __restrict__ lab_in,
                 int32_t* __restrict__ lab_out, double gthr, double ethr, int budget, uint32_t* __restrict__ counts,
                 int slot) {
    uint32_t traced = 0;
    int64_t t = (int64_t)blockIdx.x * kThreads + threadIdx.x;
    const int64_t t_step = (int64_t)gridDim.x * kThreads;
    Tracer<double> tr;
    bool active = false;
    int64_t v0 = 0;
    // a lane whose streamline ended takes its next vertex at once
    while (true) {
        while (!active && t < g.verts()) {
            v0 = t;
            t += t_step;
            if (dest[v0] == kUnassigned) {
                int x, y, z;
                vertex_xyz(g, v0, x, y, z);
                tracer_start(src, g, tr, x, y, z, budget);
                traced++;
                active = true;
            }
        }
        if (!__any(active)) break;
        if (!active) continue;
        int64_t next = 0;
        bool moved = false;
        int ev = tracer_iterate(src, g, dest, false, -1.0, gthr, ethr, tr, next, moved);
        if (ev == kOutOfVoxel) {
            next = tracer_reenter(src, g, tr);
            ev = kNone;
        }
        if (ev == kNone) continue;
        active = false;
        if (ev != kHitPreassigned) next = nearest(g, tr.cur, tr.near);
        lab_out[v0] = lab_in[next];
    }
    if (traced) atomicAdd(&counts[slot], traced);
}

// Unambiguous vertices (:348-373): every SameBdry6 neighbour has the vertex's
// label; they become terminals.  counts[3]: marked.
__global__ void __launch_bounds__(kThreads)
ers_mark_kernel(VGrid g, const int32_t* __restrict__ label, uint8_t* __restrict__ dest, uint32_t* __restrict__ counts) {
    uint32_t n = 0;
    for (int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x; v < g.verts(); v += (int64_t)gridDim.x * kThreads) {
        if (dest[v] != kUnassigned) continue;
        int x, y, z;
        vertex_xyz(g, v, x, y, z);
        int64_t nb[6];
        const int k = same_bdry6(g, v, x, y, z, nb);
        const int32_t l = label[v];
        bool same = true;
        for (int j = 0; j < k; j++)
            if (label[nb[j]] != l) same = false;
        if (same) {
            dest[v] = kCertainTerminal;
            n++;
        }
    }
    if (n) atomicAdd(&counts[3], n);
}

// ---- certain regions (GridSimplifiedExtremalRegionBuilder::BeginIntegration) ----
//
// A vertex takes the label shared by all of its earlier All6 neighbours if
// they all carry one (gi_extrema_region_builder.h:806-857); the region seeds
// are labelled from the start.  On the DAG "earlier neighbour -> vertex" this
// has one solution, computed here in dependency order: a vertex is resolved
// once its last unresolved earlier neighbour is, so every vertex is visited
// once instead of once per sweep of a fixed-point iteration.  cnt[v] counts
// the unresolved earlier neighbours of a non-seed vertex.
//
// Rounds in dependency order: round r resolves the vertices released in
// round r - 1 and appends those it releases to a flat list for round r + 1
// (one atomic per wave).  Lists alternate between two buffers; three count
// words rotate, and each round clears the one round r + 2 will append to.
//
// cnt[v] packs what a round needs of v, so that a round reads no values:
// bits 0-7 count the unresolved earlier non-seed neighbours, bits 8-13 flag
// the earlier neighbours and bits 16-21 the later non-seed neighbours, bit j
// for the direction j of +x, -x, +y, -y, +z, -z.
constexpr uint32_t kCntMask = 0xFFu;
constexpr int kEarlierShift = 8, kLaterShift = 16;
// SameBdry6 (not symmetric: a vertex on a face has no neighbour across it,
// while the interior vertex beside it counts the face vertex): bits 22-27
// flag the earlier vertices that count this one as their neighbour, and bit
// 28 says the test below applies -- the priority flood reaches a vertex only
// from such a vertex of its region (Enqueue_Later_Neighbors,
// gi_numeric_integrator_expanding_region_stop.h:108-123).
constexpr int kReachShift = 22;
constexpr uint32_t kAsym = 1u << 28;

__device__ __forceinline__ int64_t dir_offset(const VGrid& g, int j) {
    const int64_t o = j < 2 ? 1 : (j < 4 ? (int64_t)g.X : g.sxy);
    return (j & 1) ? -o : o;
}

// A list entry carries the vertex and its packed word (the count bits then
// zero), so a round reads no counter of its own vertices.
__device__ __forceinline__ uint64_t list_entry(uint32_t v, uint32_t meta) {
    return (uint64_t)(meta & ~kCntMask) << 32 | v;
}

// The vertices with no earlier non-seed neighbour start the first list.
// same_bdry: SameBdry6 neighbours (both along an axis, and only where the
// vertex is interior along it: gi_regular_grid.h:232-265) instead of All6.
__global__ void __launch_bounds__(kThreads)
certain_init_kernel(const float* __restrict__ f, VGrid g, const uint8_t* __restrict__ seedmask,
                    uint32_t* __restrict__ cnt, uint64_t* __restrict__ out, uint32_t* __restrict__ out_count,
                    uint32_t capacity, uint32_t* __restrict__ overflow, int same_bdry, int desc) {
    __shared__ uint64_t stage[kStage];
    __shared__ uint32_t s_n, s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t n = g.verts();
    // each block covers a contiguous range of at most kStage vertices
    const int64_t lo = (int64_t)blockIdx.x * kStage, hi = lo + kStage < n ? lo + kStage : n;
    for (int64_t v = lo + threadIdx.x; v < hi; v += kThreads) {
        if (seedmask[v]) {
            cnt[v] = 0u;
            continue;
        }
        int x, y, z;
        vertex_xyz(g, v, x, y, z);
        const bool exists[6] = {x < g.X - 1, x > 0, y < g.Y - 1, y > 0, z < g.Z - 1, z > 0};
        // along axis a: the vertex's own SameBdry6 neighbours (interior along a),
        // and whether the neighbour in direction j counts this vertex
        const int c3[3] = {x, y, z}, n3[3] = {g.X, g.Y, g.Z};
        uint32_t c = 0, earlier = 0, later = 0, reach = 0;
        for (int j = 0; j < 6; j++) {
            if (!exists[j]) continue;
            const int a = j >> 1, cu = c3[a] + ((j & 1) ? -1 : 1);
            const bool mine = !same_bdry || (c3[a] > 0 && c3[a] < n3[a] - 1);   // u is my neighbour
            const bool theirs = !same_bdry || (cu > 0 && cu < n3[a] - 1);       // I am u's neighbour
            const int64_t u = v + dir_offset(g, j);
            const bool seed = seedmask[u] != 0;
            if (ord_before(f, u, v, desc)) {
                if (mine) earlier |= 1u << j;
                if (theirs) reach |= 1u << j;
                if (!seed && (mine || theirs)) c++;
            } else if (!seed && (mine || theirs)) {
                later |= 1u << j;
            }
        }
        uint32_t meta = c | earlier << kEarlierShift | later << kLaterShift;
        if (same_bdry) meta |= reach << kReachShift | kAsym;
        cnt[v] = meta;
        if (c == 0) stage[atomicAdd(&s_n, 1u)] = list_entry((uint32_t)v, meta);
    }
    __syncthreads();
    const uint32_t m = s_n;
    if (threadIdx.x == 0 && m) s_base = atomicAdd(out_count, m);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < m; j += kThreads) {
        const uint32_t at = s_base + j;
        if (at < capacity) out[at] = stage[j];
        else atomicOr(overflow, 1u);
    }
}

// One round: resolve the listed vertices, release their later neighbours.
__global__ void __launch_bounds__(kThreads)
certain_round_kernel(VGrid g, int32_t* __restrict__ label, uint32_t* __restrict__ cnt,
                     const uint64_t* __restrict__ in, const uint32_t* __restrict__ in_count,
                     uint64_t* __restrict__ out, uint32_t* __restrict__ out_count, uint32_t* __restrict__ clear_count,
                     uint32_t capacity, uint32_t* __restrict__ overflow) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *clear_count = 0u;
    const int lane = (int)(threadIdx.x & 63u);
    const uint32_t total = *in_count < capacity ? *in_count : capacity;
    const uint32_t stride = gridDim.x * kThreads;
    // every lane of a wave runs the same number of iterations (wave-wide appends)
    const uint32_t first = blockIdx.x * kThreads + (threadIdx.x & ~63u);
    for (uint32_t base = first; base < total; base += stride) {
        const uint32_t it = base + (uint32_t)lane;
        uint64_t kids[6];
        int nk = 0;
        if (it < total) {
            const uint64_t e = in[it];
            const int64_t v = (int64_t)(uint32_t)e;
            const uint32_t meta = (uint32_t)(e >> 32);
            int32_t l[6];
            for (int j = 0; j < 6; j++)
                l[j] = (meta >> (kEarlierShift + j)) & 1u ? label[v + dir_offset(g, j)] : -2;
            int32_t val = -1;
            bool has = false, ok = true;
            for (int j = 0; j < 6; j++) {
                if (l[j] == -2) continue;
                if (l[j] == -1) ok = false;
                else if (!has) { val = l[j]; has = true; }
                else if (val != l[j]) ok = false;
            }
            if (ok && has && (meta & kAsym)) {
                bool reached = false;
                for (int j = 0; j < 6; j++)
                    if ((meta >> (kReachShift + j)) & 1u) reached = reached || label[v + dir_offset(g, j)] == val;
                ok = reached;
            }
            if (ok && has) label[v] = val;
            for (int j = 0; j < 6; j++) {
                if (!((meta >> (kLaterShift + j)) & 1u)) continue;
                const int64_t u = v + dir_offset(g, j);
                const uint32_t old = atomicSub(&cnt[u], 1u);
                if ((old & kCntMask) == 1u) kids[nk++] = list_entry((uint32_t)u, old);
            }
        }
        if (__any(nk > 0)) {
            const uint32_t off = wave_exclusive_scan((uint32_t)nk, lane);
            const uint32_t n_all = __shfl(off + (uint32_t)nk, 63);
            uint32_t at0 = 0;
            if (lane == 0) at0 = atomicAdd(out_count, n_all);
            at0 = __shfl(at0, 0);
            for (int j = 0; j < nk; j++) {
                const uint32_t at = at0 + off + (uint32_t)j;
                if (at < capacity) out[at] = kids[j];
                else atomicOr(overflow, 1u);
            }
        }
    }
}

// Pass 2: every assigned vertex next to a different label, or flagged for
// re-tracing, is queued (and stamped as queued).  One reservation per block.
__global__ void __launch_bounds__(kThreads)
pass2_kernel(VGrid g, LabelBuffers b, WorkQueue q) {
    __shared__ uint32_t stage[kStage];
    __shared__ uint32_t s_n, s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int64_t n = g.verts();
    // each block covers a contiguous range of at most kStage vertices
    const int64_t lo = (int64_t)blockIdx.x * kStage, hi = lo + kStage < n ? lo + kStage : n;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) {
        if (b.dest[i] != kAssigned) continue;
        bool push = b.redo[i] != 0;
        if (!push) {
            int64_t nb[6];
            const int k = all6(g, i, nb);
            const int32_t l = b.label[i];
            for (int j = 0; j < k; j++)
                if (b.label[nb[j]] != l) { push = true; break; }
        }
        if (push) {
            b.stamp[i] = 1u;
            stage[atomicAdd(&s_n, 1u)] = (uint32_t)i;
        }
    }
    __syncthreads();
    const uint32_t cnt = s_n;
    if (threadIdx.x == 0 && cnt) s_base = atomicAdd(q.tail, cnt);
    __syncthreads();
    uint32_t lost = 0;
    for (uint32_t j = threadIdx.x; j < cnt; j += kThreads) {
        const uint32_t slot = s_base + j;
        if (slot < q.capacity) q.items[slot] = stage[j];
        else lost++;
    }
    if (lost) atomicOr(q.overflow, 1u);
}

// Pass 3: re-integration to the original certain regions only, until the
// queue drains.  A re-integrated vertex whose label changed queues its
// assigned All6 neighbours that carry another label (each vertex is queued at
// most once: after its re-integration it is no longer assigned).  A
// re-integration takes the label its end vertex carried when pass 3 began
// (b.snap): a certain terminal's, which never changes, or, for a streamline
// stopped elsewhere by the gradient threshold or the step budget, that
// vertex's pass-1 label.  (The reference reads the label current at that
// moment, gi_numeric_integrator_path_compressing.h:370, which depends on its
// threads' schedule.)  So every re-integration's result depends on its vertex
// alone; a vertex is re-integrated iff pass 2 queued it or a re-integrated
// neighbour's new label differs from both that neighbour's and its own pass-1
// labels, a closure that does not depend on the processing order either, and
// the labels are the same from run to run.
//
// One launch of resident waves.  Each wave keeps a ring of queued vertices in
// LDS: it refills an empty ring with a chunk of the items pass 2 queued (one
// atomic per chunk) and keeps the children it queues there while the ring has
// room.  Children beyond that go to the global part of the queue: a producer
// reserves slots (tail) and writes them at once; idle lanes claim only
// reserved slots (head < tail, by compare-and-swap), so a claimed slot is
// always about to be written.  A lane whose streamline ends takes new work
// while the others keep advecting; the wave advects in bursts of up to kBurst
// steps that end with the first streamline to end, so a wave left with one
// long streamline does not visit the queue at every step.  A wave that holds
// no streamline, no ring item and no claimed slot leaves once it has found
// no unclaimed reserved slot for kLeaveTicks: every item queued later is
// queued by a wave still running, which drains it before it leaves itself.
// Waves thus do not wait through the tail of a few long dependent streamlines,
// and kernels of other streams get their compute units.
constexpr uint32_t kRing = 256;                // per-wave ring (a power of two)
constexpr uint32_t kRingKeep = 192;            // children stay in the ring up to this fill
constexpr uint32_t kChunk = 64;                // pass-2 items claimed at a time
constexpr int kBurst = 32;                     // advection steps between visits to the queue
constexpr uint64_t kLeaveTicks = 20000;        // 200 us idle (100 MHz real-time counter) before a wave leaves

// Between pass 2 and pass 3: the pass-2 items are [0, tail); global slots
// are claimed from tail on.
__global__ void pass3_prepare_kernel(WorkQueue q) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const uint32_t n = *q.tail;
        *q.first_count = n;
        *q.head = n;
        *q.chunk_head = 0u;
    }
}

__device__ __forceinline__ uint32_t load_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Tracing ahead.  The tail of the pass is a chain of long streamlines run
// one after another: a re-integration whose label changed queues its
// neighbours, which re-integrate and queue theirs.  A streamline depends on no
// label (it stops only at certain terminals), so where such a chain runs, idle
// lanes trace the vertices two to five All6 steps further ahead that carry
// another label than the one spreading, and keep only the vertex each
// streamline ends at.  When the chain reaches such a
// vertex, its re-integration takes that end vertex's label at that time, as a
// trace of its own would.  spec[v] (the scratch labels): kSpecNone,
// kSpecBusy (a trace ahead runs), kSpecWanted (v was reached meanwhile: the
// lane tracing ahead finishes it), else the end vertex.
constexpr int32_t kSpecNone = -1, kSpecBusy = -2, kSpecWanted = -3;
// flags kept beside the end vertex of a trace ahead (vertex ids < 2^29), for
// the counts of the vertex that takes it: re-traced in FP64, out of budget
constexpr int32_t kSpecFP64 = 0x40000000, kSpecOver = 0x20000000;
constexpr uint32_t kSpecItem = 0x80000000u;  // queue item flag: trace ahead (vertex ids < 2^31)
constexpr int kSpecSteps = 512;              // only chains of streamlines this long are traced ahead
constexpr int kShellR = 5;                   // traced ahead: All6 distance 2..kShellR (224 offsets)
constexpr int kShell = 224;
// (dx + 5) | (dy + 5) << 4 | (dz + 5) << 8, nearer distances first
__constant__ uint16_t kShellOffsets[kShell] = {
    853, 1093, 1108, 1110, 1125, 1333, 1348, 1350, 1363, 1367, 1380, 1382, 1397, 1605, 1620, 1622, 1637, 1877, 597,
    837, 852, 854, 869, 1077, 1092, 1094, 1107, 1111, 1124, 1126, 1141, 1317, 1332, 1334, 1347, 1351, 1362, 1368,
    1379, 1383, 1396, 1398, 1413, 1589, 1604, 1606, 1619, 1623, 1636, 1638, 1653, 1861, 1876, 1878, 1893, 2133, 341,
    581, 596, 598, 613, 821, 836, 838, 851, 855, 868, 870, 885, 1061, 1076, 1078, 1091, 1095, 1106, 1112, 1123,
    1127, 1140, 1142, 1157, 1301, 1316, 1318, 1331, 1335, 1346, 1352, 1361, 1369, 1378, 1384, 1395, 1399, 1412,
    1414, 1429, 1573, 1588, 1590, 1603, 1607, 1618, 1624, 1635, 1639, 1652, 1654, 1669, 1845, 1860, 1862, 1875,
    1879, 1892, 1894, 1909, 2117, 2132, 2134, 2149, 2389, 85, 325, 340, 342, 357, 565, 580, 582, 595, 599, 612, 614,
    629, 805, 820, 822, 835, 839, 850, 856, 867, 871, 884, 886, 901, 1045, 1060, 1062, 1075, 1079, 1090, 1096, 1105,
    1113, 1122, 1128, 1139, 1143, 1156, 1158, 1173, 1285, 1300, 1302, 1315, 1319, 1330, 1336, 1345, 1353, 1360,
    1370, 1377, 1385, 1394, 1400, 1411, 1415, 1428, 1430, 1445, 1557, 1572, 1574, 1587, 1591, 1602, 1608, 1617,
    1625, 1634, 1640, 1651, 1655, 1668, 1670, 1685, 1829, 1844, 1846, 1859, 1863, 1874, 1880, 1891, 1895, 1908,
    1910, 1925, 2101, 2116, 2118, 2131, 2135, 2148, 2150, 2165, 2373, 2388, 2390, 2405, 2645};

// Precision of pass 3.  By default every re-integration is traced in FP64
// with the gradient recomputed in FP64 from the values (the reference's
// arithmetic).  With kFP32First a re-integration is traced in FP32 first
// (the voxel-local tracer, TracerL); when that trace ends with a near-tie
// decision or out of step budget (doubt), the same lane traces the same
// vertex again in FP64 and only that result is used.  Either way a
// re-integration's result is a function of its vertex alone, which keeps the
// closure argument above, and the escalation happens inside the one launch,
// so the chains of dependent re-integrations and the tracing ahead run on as
// before.  With both states the corner gradients live in LDS (72 KB per
// work-group), so that they fit the registers of two waves per SIMD.
template <bool kFP32First>
struct Pass3State;
template <>
struct Pass3State<false> {
    Tracer<double> d;
};
template <>
struct Pass3State<true> {
    TracerL<LdsCorners<float, kThreads>> f;
    Tracer<double, LdsCorners<double, kThreads>> d;
};

template <bool kFP32First>
__global__ void __launch_bounds__(kThreads, 2)
pass3_kernel(GradSrc<float> srcf, GradSrc<double> srcd, VGrid g, LabelBuffers b, WorkQueue q, float dir, float gthr,
             float ethr, int max_iters) {
    __shared__ uint32_t ring_mem[kThreads / 64][kRing];
    __shared__ double corner_mem[kFP32First ? 24 * kThreads : 1];
    __shared__ float corner_mem32[kFP32First ? 24 * kThreads : 1];
    uint32_t* ring = ring_mem[threadIdx.x >> 6];
    int32_t* const spec = b.label2;
    const int lane = (int)(threadIdx.x & 63u);
    const uint64_t below = (1ull << lane) - 1ull;
    const uint32_t first = *q.first_count;
    uint32_t traced = 0, over = 0, ahead = 0, escalated = 0;
    uint64_t steps32 = 0, steps64 = 0;
    uint32_t rhead = 0, rtail = 0;  // ring cursors (the same in every lane)
    bool chunks = true;             // pass-2 chunks may remain (the same in every lane)
    bool active = false, claimed = false, speculative = false, fp64 = false;
    bool ahead_seen = false;        // the wave has seen tracing ahead (the same in every lane)
    uint32_t slot = 0;
    uint64_t wait_since = 0, idle_since = 0;
    int64_t v = 0;
    Pass3State<kFP32First> tr;
    if constexpr (kFP32First) {
        tr.d.c.p = corner_mem + threadIdx.x;
        tr.f.c.p = corner_mem32 + threadIdx.x;
    }
    // queue n items of every lane: in the ring while it has room, globally beyond
    auto push = [&](const uint32_t* items, int n) {
        if (!__any(n > 0)) return;
        const uint32_t off = wave_exclusive_scan((uint32_t)n, lane);
        const uint32_t total = __shfl(off + (uint32_t)n, 63);
        const uint32_t fill = rtail - rhead;
        const uint32_t room = fill < kRingKeep ? kRingKeep - fill : 0u;
        const uint32_t nloc = total < room ? total : room;
        const uint32_t nglob = total - nloc;
        uint32_t gbase = 0;
        if (nglob) {
            if (lane == 0) gbase = atomicAdd(q.tail, nglob);
            gbase = __shfl(gbase, 0);
        }
        uint32_t lost = 0;
        for (int j = 0; j < n; j++) {
            const uint32_t k = off + (uint32_t)j;
            if (k < nloc) {
                ring[(rtail + k) & (kRing - 1)] = items[j];
            } else {
                const uint32_t at = gbase + (k - nloc);
                if (at < q.capacity)
                    __hip_atomic_store(&q.items[at], items[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else
                    lost++;
            }
        }
        rtail += nloc;
        __builtin_amdgcn_wave_barrier();
        if (__any(lost != 0u) && lane == 0) atomicOr(q.overflow, 1u);
    };
    while (true) {
        // a lane without a streamline or a claimed slot takes ring items first
        bool need = !active && !claimed;
        uint64_t want = __ballot(need);
        if (want && rhead == rtail && chunks) {
            const int leader = __ffsll((unsigned long long)want) - 1;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(q.chunk_head, kChunk);
            base = __shfl(base, leader);
            if (base >= first) {
                chunks = false;
            } else {
                const uint32_t n = first - base < kChunk ? first - base : kChunk;
                if ((uint32_t)lane < n) ring[(rtail + (uint32_t)lane) & (kRing - 1)] = q.items[base + lane];
                rtail += n;
                __builtin_amdgcn_wave_barrier();
            }
        }
        uint32_t item = kEmptySlot;
        if (want && rhead != rtail) {
            const uint32_t avail = rtail - rhead, nw = (uint32_t)__popcll(want);
            const uint32_t rank = (uint32_t)__popcll(want & below);
            if (need && rank < avail) {
                item = ring[(rhead + rank) & (kRing - 1)];
                need = false;
            }
            rhead += avail < nw ? avail : nw;
        }
        // then reserved global slots, one per lane
        want = __ballot(need);
        if (want && rhead == rtail && !chunks) {
            const int leader = __ffsll((unsigned long long)want) - 1;
            uint32_t base = 0, n = 0, on = 0;
            if (lane == leader) {
                uint32_t h = load_agent(q.head);
                uint32_t t = load_agent(q.tail);
                on = load_agent(q.ahead_on);
                t = t < q.capacity ? t : q.capacity;
                if (h < t) {
                    const uint32_t nw = (uint32_t)__popcll(want), k = t - h < nw ? t - h : nw;
                    if (__hip_atomic_compare_exchange_strong(q.head, &h, h + k, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT)) {
                        base = h;
                        n = k;
                    }
                }
            }
            base = __shfl(base, leader);
            n = __shfl(n, leader);
            ahead_seen = ahead_seen || __shfl(on, leader) != 0u;
            const uint32_t rank = (uint32_t)__popcll(want & below);
            if (need && rank < n) {
                slot = base + rank;
                claimed = true;
            }
        }
        if (claimed) {
            const uint32_t it = load_agent(&q.items[slot]);
            if (it != kEmptySlot) {
                claimed = false;
                item = it;
            }
        }
        // fin: re-integration of v done, ending at vertex `end`
        bool fin = false, chain = false;
        int64_t end = 0;
        if (item != kEmptySlot) {
            const bool ahead_item = (item & kSpecItem) != 0u;
            v = (int64_t)(item & ~kSpecItem);
            // spec is read only once this wave has seen tracing ahead (a
            // kSpecNone read where a trace ahead exists only costs a second trace)
            const int32_t seen = ahead_seen ? spec[v] : kSpecNone;
            if (b.dest[v] == kAssigned) {
                bool trace = true;
                if (!ahead_item) {
                    traced++;
                    const int32_t sp = seen == kSpecNone ? kSpecNone : atomicCAS(&spec[v], kSpecBusy, kSpecWanted);
                    if (sp == kSpecBusy) {
                        trace = false;  // the lane tracing v ahead finishes it
                    } else if (sp >= 0) {
                        trace = false;  // traced ahead
                        fin = chain = true;
                        end = sp & ~(kSpecFP64 | kSpecOver);
                        escalated += (sp & kSpecFP64) ? 1u : 0u;
                        over += (sp & kSpecOver) ? 1u : 0u;
                    }
                }
                if (trace) {
                    active = true;
                    fp64 = !kFP32First;
                    speculative = ahead_item;
                    if (ahead_item) ahead++;
                    int sx, sy, sz;
                    vertex_xyz(g, v, sx, sy, sz);
                    if constexpr (kFP32First) tracer_start(srcf, g, tr.f, sx, sy, sz, max_iters);
                    else tracer_start(srcd, g, tr.d, sx, sy, sz, max_iters);
                }
            }
        }
        if (!__any(active || fin)) {
            if (rhead == rtail && !chunks) {
                const uint64_t waiting = __ballot(claimed);
                if (!waiting) {
                    // nothing held: leave after kLeaveTicks without a reserved
                    // slot left unclaimed
                    uint32_t left = 0;
                    if (lane == 0) {
                        const uint32_t t = load_agent(q.tail);
                        left = load_agent(q.head) < (t < q.capacity ? t : q.capacity);
                    }
                    if (!__shfl(left, 0)) {
                        const uint64_t now = __builtin_amdgcn_s_memrealtime();
                        if (idle_since == 0) idle_since = now;
                        else if (now - idle_since > kLeaveTicks) break;
                        __builtin_amdgcn_s_sleep(16);
                    } else {
                        idle_since = 0;
                    }
                } else {
                    // a claimed slot is written right after its reservation; a
                    // wave that waited kIdleTicks reports it and leaves, so the
                    // kernel ends even if that were ever broken
                    const uint64_t now = __builtin_amdgcn_s_memrealtime();
                    if (wait_since == 0) {
                        wait_since = now;
                    } else if (now - wait_since > kIdleTicks) {
                        if (lane == __ffsll((unsigned long long)waiting) - 1) atomicOr(q.overflow, 2u);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
            }
            continue;
        }
        wait_since = idle_since = 0;
        // advect in bursts of up to kBurst steps (none while a vertex traced
        // ahead waits to be finished); a burst ends when a streamline of the
        // wave ends (the lanes that ran the burst agree)
        const int burst = __any(fin) ? 0 : kBurst;
        if (active) {
            int64_t next = 0;
            int ev = kNone;
            const double dird = (double)dir, gthrd = (double)gthr, ethrd = (double)ethr;
            for (int k = 0; k < burst; k++) {
                bool moved = false;
                if constexpr (kFP32First) {
                    if (!fp64) {
                        ev = tracer_iterate(srcf, g, b.dest, false, dir, gthr, ethr, tr.f, next, moved);
                        if (ev == kOutOfVoxel) {
                            tracer_reenter(srcf, g, tr.f);
                            ev = kNone;
                        }
                        if (__any(ev != kNone)) break;
                        continue;
                    }
                }
                ev = tracer_iterate(srcd, g, b.dest, false, dird, gthrd, ethrd, tr.d, next, moved);
                if (ev == kOutOfVoxel) {
                    tracer_reenter(srcd, g, tr.d);
                    ev = kNone;
                }
                if (__any(ev != kNone)) break;
            }
            if (ev != kNone) {
                if (!fp64) {
                    if constexpr (kFP32First) {
                        const int n_steps = max_iters - (tr.f.iters > 0 ? tr.f.iters : 0);
                        steps32 += (uint64_t)n_steps;
                        if (tr.f.doubt) {
                            // trace v again in FP64 and use only that result
                            fp64 = true;
                            int sx, sy, sz;
                            vertex_xyz(g, v, sx, sy, sz);
                            tracer_start(srcd, g, tr.d, sx, sy, sz, max_iters);
                            ev = kNone;
                        } else {
                            if (ev != kHitPreassigned) next = tracer_nearest(g, tr.f);
                            chain = n_steps >= kSpecSteps;
                        }
                    }
                } else {
                    const int n_steps = max_iters - (tr.d.iters > 0 ? tr.d.iters : 0);
                    steps64 += (uint64_t)n_steps;
                    if (ev != kHitPreassigned) next = tracer_nearest(g, tr.d);
                    chain = n_steps >= kSpecSteps;
                }
            }
            if (ev != kNone) {
                active = false;
                end = next;
                const bool ob = fp64 && ev == kOverMax, esc = kFP32First && fp64;
                // a vertex traced ahead is finished here only if it was reached meanwhile
                fin = !speculative || atomicExch(&spec[v], (int32_t)next | (esc ? kSpecFP64 : 0) |
                                                               (ob ? kSpecOver : 0)) == kSpecWanted;
                if (fin) {
                    escalated += esc ? 1u : 0u;
                    over += ob ? 1u : 0u;
                }
            }
        }
        uint32_t kids[6];
        int nk = 0;
        bool spawn = false;
        int32_t spawn_label = 0;
        if (fin) {
            const int32_t init = b.label[v];
            const int32_t nl = b.snap[end];
            b.label[v] = nl;
            b.dest[v] = kCertainNonterminal;
            if (nl != init) {
                int64_t nb[6];
                const int k = all6(g, v, nb);
                for (int j = 0; j < k; j++) {
                    if (b.dest[nb[j]] == kAssigned && b.label[nb[j]] != nl &&
                        atomicExch(&b.stamp[nb[j]], 1u) == 0u)
                        kids[nk++] = (uint32_t)nb[j];
                }
                spawn = chain && !chunks;
                spawn_label = nl;
            }
        }
        push(kids, nk);
        // trace ahead around each spawning lane's vertex, one shell vertex per
        // lane, through the global part of the queue (any idle lane takes them)
        for (uint64_t sp = __ballot(spawn); sp; sp &= sp - 1) {
            ahead_seen = true;
            const int from = __ffsll((unsigned long long)sp) - 1;
            const int64_t c = (int64_t)__shfl((uint32_t)v, from);
            const int32_t cl = __shfl(spawn_label, from);
            for (int o0 = 0; o0 < kShell; o0 += 64) {
                uint32_t it = 0;
                int n = 0;
                if (o0 + lane < kShell) {
                    const uint32_t o = kShellOffsets[o0 + lane];
                    int cx, cy, cz;
                    vertex_xyz(g, c, cx, cy, cz);
                    const int ux = cx + (int)(o & 15u) - kShellR;
                    const int uy = cy + (int)((o >> 4) & 15u) - kShellR;
                    const int uz = cz + (int)(o >> 8) - kShellR;
                    if (ux >= 0 && uy >= 0 && uz >= 0 && ux < g.X && uy < g.Y && uz < g.Z) {
                        const int64_t u = (int64_t)ux + (int64_t)g.X * ((int64_t)uy + (int64_t)g.Y * uz);
                        if (b.dest[u] == kAssigned && b.stamp[u] == 0u && spec[u] == kSpecNone && b.label[u] != cl &&
                            atomicCAS(&spec[u], kSpecNone, kSpecBusy) == kSpecNone) {
                            it = (uint32_t)u | kSpecItem;
                            n = 1;
                        }
                    }
                }
                if (__any(n > 0) && lane == 0)
                    __hip_atomic_store(q.ahead_on, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                push(&it, n);
            }
        }
    }
    if (traced) atomicAdd(&b.stats[kStatTraced3f], traced);
    if (over) atomicAdd(&b.stats[kStatOver3], over);
    if (ahead) atomicAdd(&b.stats[kStatAhead3], ahead);
    if (escalated) atomicAdd(&b.stats[kStatDoubt], escalated);
    if (steps32) atomicAdd(reinterpret_cast<unsigned long long*>(&b.stats[kStatSteps3f]), (unsigned long long)steps32);
    if (steps64) atomicAdd(reinterpret_cast<unsigned long long*>(&b.stats[kStatSteps3]), (unsigned long long)steps64);
}

// ---- cleanup of isolated components ------------------------------------------

__global__ void cc_init_kernel(VGrid g, LabelBuffers b) {
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        b.rep[i] = (int32_t)i;
        b.label2[i] = b.label[i];
    }
}

// Same-label All6 components by concurrent union-find.  Roots are always
// hooked under the earlier of the two roots, so a component's root is its
// earliest vertex, as in the reference's union-find (whose merges point the
// later representative at the earlier one).  Finds halve paths with plain
// stores: a racing store only ever points a vertex at one of its ancestors.
__device__ __forceinline__ int32_t uf_parent(int32_t* rep, int32_t x) {
    return __hip_atomic_load(&rep[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ int32_t uf_find(int32_t* rep, int32_t x) {
    int32_t p = uf_parent(rep, x);
    while (p != x) {
        const int32_t gp = uf_parent(rep, p);
        if (gp != p) __hip_atomic_store(&rep[x], gp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = p;
        p = gp;
    }
    return x;
}

__device__ void uf_union(const float* f, int32_t* rep, int32_t a, int32_t b, int desc) {
    while (true) {
        a = uf_find(rep, a);
        b = uf_find(rep, b);
        if (a == b) return;
        if (ord_before(f, b, a, desc)) {  // hook the later root a under b
            const int32_t t = a; a = b; b = t;
        }
        // b is the later root now
        int32_t expected = b;
        if (__hip_atomic_compare_exchange_strong(&rep[b], &expected, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

// Unions in two steps.  A work-group first joins the same-label edges inside
// its 8x8x4 tile in LDS (hooking the later root under the earlier, as the
// global unions do) and sets rep of each vertex to its tile-local root; then
// only the edges that cross tile faces take the global unions, whose parent
// reads and hooks are device-scope atomics performed at the memory.
constexpr int kTx = 8, kTy = 8, kTz = 4;  // kTx * kTy * kTz == kThreads

__device__ __forceinline__ int lds_find(int32_t* parent, int x) {
    int p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    while (p != x) {
        x = p;
        p = __hip_atomic_load(&parent[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return x;
}

__global__ void __launch_bounds__(kThreads)
cc_union_tile_kernel(const float* __restrict__ f, VGrid g, LabelBuffers b) {
    __shared__ int32_t parent[kThreads];
    __shared__ int32_t lab[kThreads];
    __shared__ float val[kThreads];
    const int tx = (g.X + kTx - 1) / kTx, ty = (g.Y + kTy - 1) / kTy;
    const int64_t tile = blockIdx.x;
    const int ox = (int)(tile % tx) * kTx, oy = (int)((tile / tx) % ty) * kTy, oz = (int)(tile / ((int64_t)tx * ty)) * kTz;
    const int t = (int)threadIdx.x, lx = t % kTx, ly = (t / kTx) % kTy, lz = t / (kTx * kTy);
    const int x = ox + lx, y = oy + ly, z = oz + lz;
    const bool inside = x < g.X && y < g.Y && z < g.Z;
    const int64_t i = (int64_t)x + (int64_t)g.X * ((int64_t)y + (int64_t)g.Y * z);
    parent[t] = t;
    lab[t] = inside ? b.label2[i] : INT32_MIN;
    val[t] = inside ? f[i] : 0.0f;
    __syncthreads();
    // the earlier of two tile vertices in the total order (value, vertex id);
    // the local index order is the vertex id order within a tile
    auto earlier = [&](int a, int c) {
        return b.desc ? (val[a] > val[c] || (val[a] == val[c] && a > c))
                      : (val[a] < val[c] || (val[a] == val[c] && a < c));
    };
    auto join = [&](int a, int c) {
        while (true) {
            a = lds_find(parent, a);
            c = lds_find(parent, c);
            if (a == c) return;
            if (earlier(c, a)) {
                const int s2 = a; a = c; c = s2;
            }
            if (atomicCAS(&parent[c], c, a) == c) return;  // hook the later root c under a
        }
    };
    if (inside) {
        const int32_t l = lab[t];
        if (lx + 1 < kTx && x + 1 < g.X && lab[t + 1] == l) join(t, t + 1);
        if (ly + 1 < kTy && y + 1 < g.Y && lab[t + kTx] == l) join(t, t + kTx);
        if (lz + 1 < kTz && z + 1 < g.Z && lab[t + kTx * kTy] == l) join(t, t + kTx * kTy);
    }
    __syncthreads();
    if (inside) {
        const int r = lds_find(parent, t);
        const int64_t ri = (int64_t)(ox + r % kTx) + (int64_t)g.X * ((int64_t)(oy + (r / kTx) % kTy) +
                                                                       (int64_t)g.Y * (oz + r / (kTx * kTy)));
        b.rep[i] = (int32_t)ri;
    }
}

// The same-label edges across tile faces (+x, +y, +z).
__global__ void __launch_bounds__(kThreads)
cc_union_kernel(const float* __restrict__ f, VGrid g, LabelBuffers b) {
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        int x, y, z;
        vertex_xyz(g, i, x, y, z);
        const bool fx = (x + 1) % kTx == 0 && x < g.X - 1, fy = (y + 1) % kTy == 0 && y < g.Y - 1;
        const bool fz = (z + 1) % kTz == 0 && z < g.Z - 1;
        if (!(fx || fy || fz)) continue;
        const int32_t l = b.label2[i];
        if (fx && b.label2[i + 1] == l) uf_union(f, b.rep, (int32_t)i, (int32_t)(i + 1), b.desc);
        if (fy && b.label2[i + g.X] == l) uf_union(f, b.rep, (int32_t)i, (int32_t)(i + g.X), b.desc);
        if (fz && b.label2[i + g.sxy] == l) uf_union(f, b.rep, (int32_t)i, (int32_t)(i + g.sxy), b.desc);
    }
}

// After the unions: rep[i] = root.  No path halving here: a halving store
// racing with thread i's own store could leave rep[i] at a non-root ancestor.
// Only thread i writes rep[i]; other threads' walks see either the old
// ancestor or the root, both on i's path.
__global__ void __launch_bounds__(kThreads)
cc_flatten_kernel(VGrid g, LabelBuffers b) {
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        int32_t x = (int32_t)i, p = uf_parent(b.rep, x);
        while (p != x) {
            x = p;
            p = uf_parent(b.rep, x);
        }
        __hip_atomic_store(&b.rep[i], x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Component roots that may merge: roots with an All6 neighbour before them
// carrying another label (gi_isolated_region_remover.h:172-186).  Any other
// root keeps its region: its earlier neighbours lie in its own component.
// Every root starts mapped to itself (rep2); candidates are flagged in cand.
__global__ void __launch_bounds__(kThreads)
cc_candidates_kernel(const float* __restrict__ f, VGrid g, LabelBuffers b, uint32_t* __restrict__ cand,
                     uint32_t* __restrict__ count) {
    const int64_t n = g.verts();
    uint32_t mine = 0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        cand[i] = 0u;
        if (b.rep[i] != (int32_t)i) continue;
        b.rep2[i] = (int32_t)i;
        int64_t nb[6];
        const int k = all6(g, i, nb);
        const int32_t l = b.label2[i];
        bool c = false;
        for (int j = 0; j < k; j++)
            if (ord_before(f, nb[j], i, b.desc) && b.label2[nb[j]] != l) c = true;
        if (c) {
            cand[i] = 1u;
            mine++;
        }
    }
    if (mine) atomicAdd(count, mine);
}

// Candidate roots in vertex-id order (slot = exclusive prefix count of the
// candidate flags): root, its label, then for its All6 neighbours before it,
// earliest first, their component roots, those roots' candidate slots (-1
// for a root that is no candidate) and their labels (-1 padded).  The merge
// itself is sequential in root order (a root sees the merges of every
// earlier root), so it runs on the host over this list.
__global__ void __launch_bounds__(kThreads)
cc_collect_kernel(const float* __restrict__ f, VGrid g, LabelBuffers b, const uint32_t* __restrict__ cand,
                  const uint32_t* __restrict__ slot_of, int32_t* __restrict__ out) {
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        if (!cand[i]) continue;
        int64_t nb[6];
        const int k = all6(g, i, nb);
        int64_t lower[6];
        int m = 0;
        for (int j = 0; j < k; j++) {
            if (!ord_before(f, nb[j], i, b.desc)) continue;
            int p = m++;
            while (p > 0 && ord_before(f, nb[j], lower[p - 1], b.desc)) {
                lower[p] = lower[p - 1];
                p--;
            }
            lower[p] = nb[j];
        }
        int32_t* o = out + (size_t)slot_of[i] * kCollectStride;
        o[0] = (int32_t)i;
        o[1] = b.label2[i];
        for (int j = 0; j < 6; j++) {
            const int32_t r = j < m ? b.rep[lower[j]] : -1;
            o[2 + j] = r;
            o[8 + j] = r >= 0 && cand[r] ? (int32_t)slot_of[r] : -1;
            o[14 + j] = j < m ? b.label2[lower[j]] : -1;
        }
    }
}

// rep2[root] = the root it finally joins, from the host's merge.
__global__ void __launch_bounds__(kThreads)
cc_scatter_kernel(const int32_t* __restrict__ pairs, uint32_t n, int32_t* __restrict__ rep2) {
    for (uint32_t k = blockIdx.x * kThreads + threadIdx.x; k < n; k += gridDim.x * kThreads)
        rep2[pairs[2 * k]] = pairs[2 * k + 1];
}

__global__ void __launch_bounds__(kThreads)
cc_apply_kernel(VGrid g, LabelBuffers b) {
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        if (b.label2[i] == -1) continue;
        b.label[i] = b.label2[b.rep2[b.rep[i]]];
    }
}

// ---- restriction labels -------------------------------------------------------

// VertexLabelingToBoundaryLabeling::ComputeMAXBoundary
// (gi_labeling_to_bounary_labeling.h:118-195), three passes that add 2:
// (stage 0) every interior quad whose two hexes have different highest
// vertices with different valid labels; (stage 1) every interior edge with
// such a quad (label > 1) among its cofacets; (stage 2) every interior vertex
// with such an edge.  "Interior" is boundaryValue 0: no coordinate on a face.
__device__ __forceinline__ int64_t hex_highest_vertex(const VGrid& g, const float* __restrict__ f, int64_t hx,
                                                      int64_t hy, int64_t hz) {
    // vertices of the hex at odd cell coordinates (hx, hy, hz)
    const int64_t x0 = (hx - 1) / 2, y0 = (hy - 1) / 2, z0 = (hz - 1) / 2;
    int64_t best = x0 + (int64_t)g.X * (y0 + (int64_t)g.Y * z0);
#pragma unroll
    for (int k = 1; k < 8; k++) {
        const int64_t v = (x0 + (k & 1)) + (int64_t)g.X * ((y0 + ((k >> 1) & 1)) + (int64_t)g.Y * (z0 + (k >> 2)));
        if (before(f, best, v)) best = v;  // highest by (value, vertex id)
    }
    return best;
}

template <int kStage>
__global__ void __launch_bounds__(kThreads)
max_boundary_kernel(VGrid g, const float* __restrict__ f, const int32_t* __restrict__ label, int8_t* __restrict__ r) {
    const int64_t Mx = 2 * (int64_t)g.X - 1, My = 2 * (int64_t)g.Y - 1, Mz = 2 * (int64_t)g.Z - 1;
    const int64_t sxy = Mx * My, C = sxy * Mz;
    for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < C; c += (int64_t)gridDim.x * kThreads) {
        int64_t cx, cy, cz;
        cell_xyz(c, Mx, My, cx, cy, cz);
        if (cx == 0 || cx == Mx - 1 || cy == 0 || cy == My - 1 || cz == 0 || cz == Mz - 1) continue;
        const int ox = (int)(cx & 1), oy = (int)(cy & 1), oz = (int)(cz & 1);
        const int dim = ox + oy + oz;
        if (kStage == 0) {
            if (dim != 2) continue;
            // the even axis carries the two hexes
            const int64_t dx = ox ? 0 : 1, dy = oy ? 0 : 1, dz = oz ? 0 : 1;
            const int64_t v1 = hex_highest_vertex(g, f, cx + dx, cy + dy, cz + dz);
            const int64_t v2 = hex_highest_vertex(g, f, cx - dx, cy - dy, cz - dz);
            if (v1 == v2) continue;  // both hexes in one lower star
            const int32_t l1 = label[v1], l2 = label[v2];
            if (l1 == -1 || l2 == -1 || l1 == l2) continue;
            r[c] = (int8_t)(r[c] + 2);
        } else {
            // edges (stage 1) look at their quads, vertices (stage 2) at their edges
            if (dim != 2 - kStage) continue;
            const int64_t off[3] = {1, Mx, sxy};
            const int odd[3] = {ox, oy, oz};
            bool mark = false;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (odd[a]) continue;
                if (r[c + off[a]] > 1 || r[c - off[a]] > 1) mark = true;
            }
            if (mark) r[c] = (int8_t)(r[c] + 2);
        }
    }
}

__global__ void __launch_bounds__(kThreads)
min_boundary_kernel(VGrid g, const int32_t* __restrict__ label, int8_t* __restrict__ r) {
    const int64_t Mx = 2 * (int64_t)g.X - 1, My = 2 * (int64_t)g.Y - 1, Mz = 2 * (int64_t)g.Z - 1;
    const int64_t C = Mx * My * Mz;
    for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < C; c += (int64_t)gridDim.x * kThreads) {
        int64_t cx, cy, cz;
        cell_xyz(c, Mx, My, cx, cy, cz);
        const int ox = (int)(cx & 1), oy = (int)(cy & 1), oz = (int)(cz & 1);
        int8_t res = 0;
        if (ox + oy + oz > 0) {
            // vertices of the cell: corners (cx-ox+2a, ...)/2 for a in {0,1}
            const int64_t vx = (cx - ox) / 2, vy = (cy - oy) / 2, vz = (cz - oz) / 2;
            int32_t l[8];
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const int a = k & 1, bb = (k >> 1) & 1, d = k >> 2;
                if ((a && !ox) || (bb && !oy) || (d && !oz)) { l[k] = -2; continue; }
                l[k] = label[(vx + a) + (int64_t)g.X * ((vy + bb) + (int64_t)g.Y * (vz + d))];
            }
            // an edge of the cell joins corners that differ in one odd axis
#pragma unroll
            for (int k = 0; k < 8; k++) {
#pragma unroll
                for (int ax = 0; ax < 3; ax++) {
                    const int k2 = k | (1 << ax);
                    if (k2 == k) continue;
                    if (l[k] < -1 || l[k2] < -1) continue;
                    if (l[k] == -1 || l[k2] == -1) continue;
                    if (l[k] != l[k2]) res = 1;
                }
            }
        }
        r[c] = res;
    }
}

// ---- combinatorial structure --------------------------------------------------

struct CellG {
    int64_t Mx, My, sxy;
};

__device__ __forceinline__ int64_t grad_pair(const CellG& cg, const uint8_t* grad, int64_t c) {
    const uint32_t d = (grad[c] >> 2) & 7u;
    switch (d) {
        case 0: return c + 1;
        case 1: return c - 1;
        case 2: return c + cg.Mx;
        case 3: return c - cg.Mx;
        case 4: return c + cg.sxy;
        case 5: return c - cg.sxy;
    }
    return c;
}

// From a vertex cell, follow vertex -> paired edge -> other vertex until a
// critical vertex (rec_td, gi_extrema_region_builder.h:286-301).
__device__ int64_t trace_down_vertex(const CellG& cg, const uint8_t* grad, int64_t c, int64_t limit, int64_t cells) {
    for (int64_t s = 0; s < limit; s++) {
        const uint32_t d = (grad[c] >> 2) & 7u;
        if (d == 7u) return c;
        const int64_t e = grad_pair(cg, grad, c);
        // the other vertex of edge e: e + (e - c)
        const int64_t nx = 2 * e - c;
        if (e < 0 || e >= cells || nx < 0 || nx >= cells) return -1;  // only a corrupt byte leads out
        c = nx;
    }
    return -1;
}

__device__ inline int64_t vertex_of_cell(const CellG& cg, int64_t c) {
    int64_t x, y, z;
    cell_xyz(c, cg.Mx, cg.My, x, y, z);
    return x / 2 + ((cg.Mx + 1) / 2) * (y / 2 + ((cg.My + 1) / 2) * (z / 2));
}

// Arcs of the extremum graph below the simplification threshold.  The
// persistence is computed as the host simplification computes it (float
// differences, then the smaller magnitude); an arc above the threshold is
// never cancelled, so only the others leave the device.
__global__ void __launch_bounds__(kThreads)
saddle_arcs_kernel(VGrid g, const float* __restrict__ f, const uint8_t* __restrict__ grad, MList list, float thr,
                   ArcRecord* __restrict__ out, uint32_t cap, uint32_t* __restrict__ count) {
    __shared__ uint32_t prefix[kSubs + 2];
    const uint32_t total = mlist_prefix(list, prefix);
    const CellG cg{2 * (int64_t)g.X - 1, 2 * (int64_t)g.Y - 1, (2 * (int64_t)g.X - 1) * (2 * (int64_t)g.Y - 1)};
    const int64_t limit = g.verts() + 1, cells = cg.sxy * (2 * (int64_t)g.Z - 1);
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < total; i += gridDim.x * kThreads) {
        const int64_t s = mlist_at(list, prefix, i);
        int64_t cx, cy, cz;
        cell_xyz(s, cg.Mx, cg.My, cx, cy, cz);
        const int64_t step = (cx & 1) ? 1 : ((cy & 1) ? cg.Mx : cg.sxy);
        // facets of the edge in +axis, -axis order
        const int64_t m1 = trace_down_vertex(cg, grad, s + step, limit, cells);
        const int64_t m2 = trace_down_vertex(cg, grad, s - step, limit, cells);
        if (m1 == m2 || m1 < 0 || m2 < 0) continue;
        // value of the saddle's highest vertex (value, then vertex id)
        const int64_t va = vertex_of_cell(cg, s + step), vb = vertex_of_cell(cg, s - step);
        const float fa = f[va], fb = f[vb];
        const bool a_high = fa > fb || (!(fb > fa) && va > vb);
        const float sval = a_high ? fa : fb;
        const float v1 = f[vertex_of_cell(cg, m1)], v2 = f[vertex_of_cell(cg, m2)];
        float d1 = v1 - sval, d2 = v2 - sval;
        if (d1 < 0) d1 *= -1;
        if (d2 < 0) d2 *= -1;
        const float pers = d1 < d2 ? d1 : d2;
        if (pers > thr) continue;
        const uint32_t slot = atomicAdd(count, 1u);
        if (slot < cap) out[slot] = ArcRecord{s, m1, m2, sval, v1, v2, 0.f};
    }
}

// From a hex, follow hex -> paired quad -> the quad's other hex until a
// critical hex (rec_tu, gi_extrema_region_builder.h:330-352); -1 where a hex
// is paired with a quad on the domain boundary (no other hex).
__device__ int64_t trace_up_hex(const CellG& cg, int64_t Mz, const uint8_t* grad, int64_t c, int64_t limit) {
    for (int64_t s = 0; s < limit; s++) {
        const uint32_t d = (grad[c] >> 2) & 7u;
        if (d == 7u) return c;
        if (d > 5u) return -1;  // only a corrupt byte
        const int64_t q = grad_pair(cg, grad, c);
        if (((grad[q] >> 2) & 7u) == 7u) return q;
        int64_t x, y, z;
        cell_xyz(q, cg.Mx, cg.My, x, y, z);
        const int a = (int)(d >> 1);
        const int64_t along = a == 0 ? x : (a == 1 ? y : z), ext = a == 0 ? cg.Mx : (a == 1 ? cg.My : Mz);
        if (along == 0 || along == ext - 1) return -1;
        c = 2 * q - c;
    }
    return -1;
}

// Highest vertex (value, then vertex id) of the cell at (x, y, z): its
// corners are the vertices (x +- (x odd)) / 2 ... (the max vertex labelling,
// gi_max_vertex_labeling.h).
__device__ int64_t highest_vertex(const VGrid& g, const float* __restrict__ f, int64_t x, int64_t y, int64_t z) {
    int64_t best = -1;
    for (int k = 0; k < 8; k++) {
        const int64_t vx = (x + ((x & 1) ? ((k & 1) ? 1 : -1) : 0)) / 2;
        const int64_t vy = (y + ((y & 1) ? ((k & 2) ? 1 : -1) : 0)) / 2;
        const int64_t vz = (z + ((z & 1) ? ((k & 4) ? 1 : -1) : 0)) / 2;
        const int64_t v = vx + (int64_t)g.X * (vy + (int64_t)g.Y * vz);
        if (best < 0 || before(f, best, v)) best = v;
    }
    return best;
}

// Arcs of the max extremum graph (trace_up_2saddle, gi_extrema_region_builder.h:354-372):
// every critical quad off the domain boundary traced up through its two
// hexes (+ side first) to maxima; arcs joining two different maxima with
// persistence not above `thr` are appended.
__global__ void __launch_bounds__(kThreads)
saddle2_arcs_kernel(VGrid g, const float* __restrict__ f, const uint8_t* __restrict__ grad, MList list, float thr,
                    MaxArcRecord* __restrict__ out, uint32_t cap, uint32_t* __restrict__ count) {
    __shared__ uint32_t prefix[kSubs + 2];
    const uint32_t total = mlist_prefix(list, prefix);
    const CellG cg{2 * (int64_t)g.X - 1, 2 * (int64_t)g.Y - 1, (2 * (int64_t)g.X - 1) * (2 * (int64_t)g.Y - 1)};
    const int64_t Mz = 2 * (int64_t)g.Z - 1, limit = g.verts() + 1;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < total; i += gridDim.x * kThreads) {
        const int64_t s = mlist_at(list, prefix, i);
        int64_t cx, cy, cz;
        cell_xyz(s, cg.Mx, cg.My, cx, cy, cz);
        // the quad's even axis: its hexes lie along it; on the boundary it has one
        const int a = !(cx & 1) ? 0 : (!(cy & 1) ? 1 : 2);
        const int64_t along = a == 0 ? cx : (a == 1 ? cy : cz), ext = a == 0 ? cg.Mx : (a == 1 ? cg.My : Mz);
        if (along == 0 || along == ext - 1) continue;
        const int64_t step = a == 0 ? 1 : (a == 1 ? cg.Mx : cg.sxy);
        const int64_t m1 = trace_up_hex(cg, Mz, grad, s + step, limit);
        const int64_t m2 = trace_up_hex(cg, Mz, grad, s - step, limit);
        if (m1 == m2 || m1 < 0 || m2 < 0) continue;
        const float sval = f[highest_vertex(g, f, cx, cy, cz)];
        int64_t x, y, z;
        cell_xyz(m1, cg.Mx, cg.My, x, y, z);
        const int64_t h1 = highest_vertex(g, f, x, y, z);
        cell_xyz(m2, cg.Mx, cg.My, x, y, z);
        const int64_t h2 = highest_vertex(g, f, x, y, z);
        const float v1 = f[h1], v2 = f[h2];
        float d1 = v1 - sval, d2 = v2 - sval;
        if (d1 < 0) d1 *= -1;
        if (d2 < 0) d2 *= -1;
        const float pers = d1 < d2 ? d1 : d2;
        if (pers > thr) continue;
        const uint32_t slot = atomicAdd(count, 1u);
        if (slot < cap) out[slot] = MaxArcRecord{s, m1, m2, h1, h2, sval, v1, v2, 0.f};
    }
}

// The highest vertex of every critical hex (the maxima of the max graph).
__global__ void __launch_bounds__(kThreads)
collect_maxima_kernel(VGrid g, const float* __restrict__ f, const uint8_t* __restrict__ grad, MList out,
                      uint32_t* __restrict__ overflow) {
    __shared__ uint32_t stage[kStage];
    __shared__ uint32_t s_n;
    __shared__ MSlot s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const Outbox box{stage, &s_n, out, (int)(blockIdx.x % kSubs), overflow};
    const int64_t Mx = 2 * (int64_t)g.X - 1, My = 2 * (int64_t)g.Y - 1;
    const int64_t HX = g.X - 1, HY = g.Y - 1, n = HX * HY * (int64_t)(g.Z - 1);
    for (int64_t h = (int64_t)blockIdx.x * kThreads + threadIdx.x; h < n; h += (int64_t)gridDim.x * kThreads) {
        const int64_t x = 2 * (h % HX) + 1, y = 2 * ((h / HX) % HY) + 1, z = 2 * (h / (HX * HY)) + 1;
        const int64_t c = x + Mx * (y + My * z);
        if (((grad[c] >> 2) & 7u) == 7u) box.push((uint32_t)highest_vertex(g, f, x, y, z));
    }
    __syncthreads();
    mlist_flush_block(out, box.sub, stage, s_n < (uint32_t)kStage ? s_n : (uint32_t)kStage, &s_base, overflow);
}

__global__ void __launch_bounds__(kThreads)
seed_minima_kernel(MList minima, int32_t* __restrict__ label, float4* __restrict__ gradf,
                   uint8_t* __restrict__ seedmask) {
    __shared__ uint32_t prefix[kSubs + 2];
    const uint32_t total = mlist_prefix(minima, prefix);
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < total; i += gridDim.x * kThreads) {
        const int64_t v = mlist_at(minima, prefix, i);
        label[v] = (int32_t)v;
        gradf[v] = make_float4(0.f, 0.f, 0.f, 0.f);
        seedmask[v] = 1;
    }
}

__global__ void __launch_bounds__(kThreads)
collect_minima_kernel(VGrid g, const uint8_t* __restrict__ grad, MList out, uint32_t* __restrict__ overflow) {
    __shared__ uint32_t stage[kStage];
    __shared__ uint32_t s_n;
    __shared__ MSlot s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const Outbox box{stage, &s_n, out, (int)(blockIdx.x % kSubs), overflow};
    const int64_t Mx = 2 * (int64_t)g.X - 1, My = 2 * (int64_t)g.Y - 1;
    const int64_t n = g.verts();
    for (int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x; v < n; v += (int64_t)gridDim.x * kThreads) {
        int xi, yi, zi;
        vertex_xyz(g, v, xi, yi, zi);
        const int64_t x = xi, y = yi, z = zi;
        const int64_t c = 2 * x + Mx * (2 * y + My * 2 * z);
        if (((grad[c] >> 2) & 7u) == 7u) box.push((uint32_t)v);
    }
    __syncthreads();
    mlist_flush_block(out, box.sub, stage, s_n < (uint32_t)kStage ? s_n : (uint32_t)kStage, &s_base, overflow);
}

__global__ void __launch_bounds__(kThreads)
combinatorial_labels_kernel(VGrid g, const uint8_t* __restrict__ grad, const int32_t* __restrict__ min_label,
                            int32_t* __restrict__ out, const uint8_t* __restrict__ only_where_zero) {
    const CellG cg{2 * (int64_t)g.X - 1, 2 * (int64_t)g.Y - 1, (2 * (int64_t)g.X - 1) * (2 * (int64_t)g.Y - 1)};
    const int64_t n = g.verts(), limit = n + 1, cells = cg.sxy * (2 * (int64_t)g.Z - 1);
    for (int64_t v = (int64_t)blockIdx.x * kThreads + threadIdx.x; v < n; v += (int64_t)gridDim.x * kThreads) {
        if (only_where_zero && only_where_zero[v]) continue;
        int xi, yi, zi;
        vertex_xyz(g, v, xi, yi, zi);
        const int64_t x = xi, y = yi, z = zi;
        const int64_t c = 2 * x + cg.Mx * (2 * y + cg.My * 2 * z);
        const int64_t m = trace_down_vertex(cg, grad, c, limit, cells);
        if (m < 0) { out[v] = -1; continue; }
        int64_t mx, my, mz;
        cell_xyz(m, cg.Mx, cg.My, mx, my, mz);
        mx /= 2;
        my /= 2;
        mz /= 2;
        out[v] = min_label[mx + (int64_t)g.X * (my + (int64_t)g.Y * mz)];
    }
}

__global__ void dest_init_kernel(VGrid g, LabelBuffers b) {
    const int64_t n = g.verts();
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        b.dest[i] = b.label[i] == -1 ? kUnassigned : kCertainTerminal;
        b.redo[i] = 0;
        b.stamp[i] = kNoClaim;  // pass-1 claims
    }
}

__global__ void __launch_bounds__(kThreads)
apply_mask_kernel(VGrid g, LabelBuffers b, const uint8_t* __restrict__ mask) {
    const int64_t n = g.verts();
    uint32_t inside = 0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        if (b.dest[i] != kUnassigned) continue;
        if (mask[i]) {
            inside++;
        } else {
            // trusted combinatorial label: a terminal for every streamline
            b.label[i] = b.label2[i];
            b.dest[i] = kCertainTerminal;
        }
    }
    if (inside) atomicAdd(&b.stats[kStatMasked], inside);
}

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

inline uint32_t grid_for(int64_t n, uint32_t cap = 16384) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    return (uint32_t)(b > cap ? cap : b);
}

}  // namespace

hipError_t launch_grad_field(const VGrid& g, const float* values, float4* gradf, hipStream_t s) {
    hipLaunchKernelGGL(grad_field_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, values, g, gradf);
    return hipGetLastError();
}
hipError_t launch_zero_seed_gradient(const VGrid& g, const LabelBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(zero_seed_gradient_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, b.seedmask,
                       b.gradf);
    return hipGetLastError();
}
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t s) {
    hipLaunchKernelGGL(fill_i32_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t s) {
    hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_or_u8(uint8_t* dst, const uint8_t* src, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(or_u8_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, dst, src, n);
    return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_seed_scatter(const int32_t* seed_vert, const int32_t* seed_label, int n, const LabelBuffers& b,
                               hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(seed_scatter_kernel, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, s, seed_vert,
                       seed_label, n, b.label, b.gradf, b.seedmask);
    return hipGetLastError();
}
namespace {
uint32_t compute_units() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return 256;
    return (uint32_t)cus;
}
// Work-groups of a persistent kernel that are resident at once on the device.
template <typename K>
uint32_t resident_blocks(K kernel) {
    int dev = 0, cus = 0, per = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 1024;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return 1024;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kThreads, 0) != hipSuccess || per <= 0) per = 1;
    return (uint32_t)(cus * per);
}
}  // namespace
hipError_t launch_certain_init(const VGrid& g, const float* values, const uint8_t* seedmask, uint32_t* cnt,
                               uint64_t* out, uint32_t* out_count, uint32_t capacity, uint32_t* overflow,
                               hipStream_t s, bool same_bdry, bool desc) {
    const int64_t blocks = (g.verts() + kStage - 1) / kStage;
    hipLaunchKernelGGL(certain_init_kernel, dim3((uint32_t)blocks), dim3(kThreads), 0, s, values, g, seedmask, cnt,
                       out, out_count, capacity, overflow, same_bdry ? 1 : 0, desc ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_certain_round(const VGrid& g, int32_t* label, uint32_t* cnt, const uint64_t* in,
                                const uint32_t* in_count, uint64_t* out, uint32_t* out_count, uint32_t* clear_count,
                                uint32_t capacity, uint32_t* overflow, uint64_t expected, hipStream_t s) {
    // about two listed vertices per lane: one work-group per compute unit for
    // the thousands of a smooth field's rounds, up to 32 for the millions of a
    // noisy one's (the rounds are latency-bound)
    static const uint32_t cus = compute_units();
    const uint64_t want = (expected + 2 * kThreads - 1) / (2 * kThreads);
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(want, cus), 32ull * cus);
    hipLaunchKernelGGL(certain_round_kernel, dim3(blocks), dim3(kThreads), 0, s, g, label, cnt, in, in_count, out,
                       out_count, clear_count, capacity, overflow);
    return hipGetLastError();
}
hipError_t launch_ers_seeds(const VGrid& g, const float* values, uint8_t* seedmask, int32_t* label, uint32_t* counts,
                            hipStream_t s) {
    hipLaunchKernelGGL(ers_seed_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, values, g, seedmask, label,
                       counts);
    return hipGetLastError();
}
hipError_t launch_ers_state(const VGrid& g, const int32_t* label, uint8_t* dest, uint32_t* counts, hipStream_t s) {
    hipLaunchKernelGGL(ers_state_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, label, dest, counts);
    return hipGetLastError();
}
hipError_t launch_ers_trace(const VGrid& g, const double4* gradd, const uint8_t* dest, const int32_t* lab_in,
                            int32_t* lab_out, float gthr, float ethr, int budget, uint32_t* counts, int slot,
                            hipStream_t s) {
    static const uint32_t blocks = resident_blocks(ers_trace_kernel);
    GradSrc<double> src{gradd};
    hipLaunchKernelGGL(ers_trace_kernel, dim3(blocks), dim3(kThreads), 0, s, src, g, dest, lab_in, lab_out,
                       (double)gthr, (double)ethr, budget, counts, slot);
    return hipGetLastError();
}
hipError_t launch_ers_mark(const VGrid& g, const int32_t* label, uint8_t* dest, uint32_t* counts, hipStream_t s) {
    hipLaunchKernelGGL(ers_mark_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, label, dest, counts);
    return hipGetLastError();
}
hipError_t launch_dest_init(const VGrid& g, const LabelBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(dest_init_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, b);
    return hipGetLastError();
}
hipError_t launch_pass1(const VGrid& g, const LabelBuffers& b, int stride, bool first_level, float gthr, float ethr,
                        int max_iters, int sign, hipStream_t s) {
    GradSrc<float> src{b.gradf, b.values, b.seedmask, g};
    const int64_t n = (int64_t)((g.X + stride - 1) / stride) * ((g.Y + stride - 1) / stride) * ((g.Z + stride - 1) / stride);
    hipLaunchKernelGGL(pass1_kernel, dim3(grid_for(n, 8192)), dim3(kThreads), 0, s, src, g, b, stride,
                       first_level ? 1 : 0, (float)sign, gthr, ethr, max_iters);
    return hipGetLastError();
}
hipError_t launch_pass1_finish(const VGrid& g, const LabelBuffers& b, int stride, bool first_level, bool last,
                               hipStream_t s) {
    hipLaunchKernelGGL(pass1_finish_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, b, stride,
                       first_level ? 1 : 0, last ? 1 : 0);
    return hipGetLastError();
}
hipError_t launch_pass2(const VGrid& g, const LabelBuffers& b, const WorkQueue& q, hipStream_t s) {
    const int64_t blocks = (g.verts() + kStage - 1) / kStage;
    hipLaunchKernelGGL(pass2_kernel, dim3((uint32_t)blocks), dim3(kThreads), 0, s, g, b, q);
    return hipGetLastError();
}
hipError_t launch_grad_field64(const VGrid& g, const LabelBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(grad_field64_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, b.values, g, b.seedmask,
                       b.gradd);
    return hipGetLastError();
}
hipError_t launch_pass3(const VGrid& g, const LabelBuffers& b, const WorkQueue& q, bool fp32_first, float gthr,
                        float ethr, int max_iters, int sign, hipStream_t s) {
    GradSrc<float> srcf{b.gradf, b.values, b.seedmask, g};
    GradSrc<double> srcd{b.gradd};
    hipLaunchKernelGGL(pass3_prepare_kernel, dim3(1), dim3(64), 0, s, q);
    if (fp32_first) {
        static uint32_t blocks = resident_blocks(pass3_kernel<true>);
        hipLaunchKernelGGL(pass3_kernel<true>, dim3(blocks), dim3(kThreads), 0, s, srcf, srcd, g, b, q, (float)sign,
                           gthr, ethr, max_iters);
    } else {
        static uint32_t blocks = resident_blocks(pass3_kernel<false>);
        hipLaunchKernelGGL(pass3_kernel<false>, dim3(blocks), dim3(kThreads), 0, s, srcf, srcd, g, b, q, (float)sign,
                           gthr, ethr, max_iters);
    }
    return hipGetLastError();
}
hipError_t launch_cc_init(const VGrid& g, const LabelBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(cc_init_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, b);
    return hipGetLastError();
}
hipError_t launch_cc_union(const VGrid& g, const LabelBuffers& b, hipStream_t s) {
    static_assert(kTx * kTy * kTz == kThreads, "one tile vertex per thread");
    const int64_t tiles = (int64_t)((g.X + kTx - 1) / kTx) * ((g.Y + kTy - 1) / kTy) * ((g.Z + kTz - 1) / kTz);
    if (tiles > 0x7FFFFFFF) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cc_union_tile_kernel, dim3((uint32_t)tiles), dim3(kThreads), 0, s, b.values, g, b);
    if (hipGetLastError() != hipSuccess) return hipGetLastError();
    hipLaunchKernelGGL(cc_union_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, b.values, g, b);
    if (hipGetLastError() != hipSuccess) return hipGetLastError();
    hipLaunchKernelGGL(cc_flatten_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, b);
    return hipGetLastError();
}
hipError_t launch_cc_candidates(const VGrid& g, const LabelBuffers& b, uint32_t* cand, uint32_t* count,
                                hipStream_t s) {
    hipLaunchKernelGGL(cc_candidates_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, b.values, g, b, cand,
                       count);
    return hipGetLastError();
}
hipError_t launch_cc_collect(const VGrid& g, const LabelBuffers& b, const uint32_t* cand, const uint32_t* slot_of,
                             int32_t* out, hipStream_t s) {
    hipLaunchKernelGGL(cc_collect_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, b.values, g, b, cand,
                       slot_of, out);
    return hipGetLastError();
}
hipError_t launch_exclusive_sum(const uint32_t* in, uint32_t* out, int64_t n, void* temp, size_t* temp_bytes,
                                hipStream_t s) {
    return hipcub::DeviceScan::ExclusiveSum(temp, *temp_bytes, in, out, (int)n, s);
}
hipError_t launch_cc_scatter(const int32_t* pairs, uint32_t n, int32_t* rep2, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(cc_scatter_kernel, dim3(grid_for(n)), dim3(kThreads), 0, s, pairs, n, rep2);
    return hipGetLastError();
}
hipError_t launch_cc_apply(const VGrid& g, const LabelBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(cc_apply_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, b);
    return hipGetLastError();
}
hipError_t launch_min_boundary(const VGrid& g, const int32_t* label, int8_t* restriction, hipStream_t s) {
    const int64_t C = (2 * (int64_t)g.X - 1) * (2 * (int64_t)g.Y - 1) * (2 * (int64_t)g.Z - 1);
    hipLaunchKernelGGL(min_boundary_kernel, dim3(grid_for(C)), dim3(kThreads), 0, s, g, label, restriction);
    return hipGetLastError();
}
hipError_t launch_max_boundary(const VGrid& g, const float* values, const int32_t* label, int8_t* restriction,
                               hipStream_t s) {
    const int64_t C = (2 * (int64_t)g.X - 1) * (2 * (int64_t)g.Y - 1) * (2 * (int64_t)g.Z - 1);
    hipLaunchKernelGGL(max_boundary_kernel<0>, dim3(grid_for(C)), dim3(kThreads), 0, s, g, values, label, restriction);
    hipLaunchKernelGGL(max_boundary_kernel<1>, dim3(grid_for(C)), dim3(kThreads), 0, s, g, values, label, restriction);
    hipLaunchKernelGGL(max_boundary_kernel<2>, dim3(grid_for(C)), dim3(kThreads), 0, s, g, values, label, restriction);
    return hipGetLastError();
}
hipError_t launch_saddle_arcs(const VGrid& g, const float* values, const uint8_t* grad, const MList& list,
                              float threshold, ArcRecord* out, uint32_t cap, uint32_t* count, hipStream_t s) {
    hipLaunchKernelGGL(saddle_arcs_kernel, dim3(1024), dim3(kThreads), 0, s, g, values, grad, list, threshold, out,
                       cap, count);
    return hipGetLastError();
}
hipError_t launch_saddle2_arcs(const VGrid& g, const float* values, const uint8_t* grad, const MList& list,
                               float threshold, MaxArcRecord* out, uint32_t cap, uint32_t* count, hipStream_t s) {
    hipLaunchKernelGGL(saddle2_arcs_kernel, dim3(1024), dim3(kThreads), 0, s, g, values, grad, list, threshold, out,
                       cap, count);
    return hipGetLastError();
}
hipError_t launch_collect_maxima(const VGrid& g, const float* values, const uint8_t* grad, const MList& out,
                                 uint32_t* overflow, hipStream_t s) {
    const int64_t n = (int64_t)(g.X - 1) * (g.Y - 1) * (g.Z - 1);
    hipLaunchKernelGGL(collect_maxima_kernel, dim3(grid_for(n > 0 ? n : 1, 8192)), dim3(kThreads), 0, s, g, values,
                       grad, out, overflow);
    return hipGetLastError();
}
hipError_t launch_seed_minima(const MList& minima, const LabelBuffers& b, hipStream_t s) {
    hipLaunchKernelGGL(seed_minima_kernel, dim3(256), dim3(kThreads), 0, s, minima, b.label, b.gradf, b.seedmask);
    return hipGetLastError();
}
hipError_t launch_collect_minima(const VGrid& g, const uint8_t* grad, const MList& out, uint32_t* overflow,
                                 hipStream_t s) {
    hipLaunchKernelGGL(collect_minima_kernel, dim3(grid_for(g.verts(), 8192)), dim3(kThreads), 0, s, g, grad, out,
                       overflow);
    return hipGetLastError();
}
hipError_t launch_apply_mask(const VGrid& g, const LabelBuffers& b, const uint8_t* mask, hipStream_t s) {
    hipLaunchKernelGGL(apply_mask_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, b, mask);
    return hipGetLastError();
}
hipError_t launch_combinatorial_labels(const VGrid& g, const uint8_t* grad, const int32_t* min_label, int32_t* out,
                                       const uint8_t* only_where_zero, hipStream_t s) {
    hipLaunchKernelGGL(combinatorial_labels_kernel, dim3(grid_for(g.verts())), dim3(kThreads), 0, s, g, grad,
                       min_label, out, only_where_zero);
    return hipGetLastError();
}

}  // namespace msceer

// Accurate 1-manifolds: streamlines from the saddles, digitized into the
// restriction labels of the conforming gradient.
//
// Semantics restated (reference file:line):
//   driver:     ondemandaccurate.cxx:254-500 (ReIntegrateUpFrom2Saddles,
//               ReIntegrateDownFrom1Saddles), 1075-1091
//   streamline: gi_numeric_streamline_integrator.h:734-788 (ASC), 1018-1072 (DSC)
//   digitizing: gi_numeric_streamline_integrator.h:566-700 (ASC), 824-985 (DSC)
//   checkers:   gi_advection_checker.h:59-100 (TerminateNearAssignedHex / Vert)
//   advection:  gi_adaptive_euler_advector.h:51-305 (sign +1 ASC, -1 DSC)
//
// ASC: the critical 2-cells (quads) of the steepest gradient, in decreasing
// (value of the highest vertex, cell id) order; DSC: the critical 1-cells
// (edges), in increasing value, decreasing id.  Each saddle's seeds are the
// cells a short V-path search reaches (hexes ASC, vertices DSC, at most six);
// from each seed, in order, one FP64 streamline (uphill ASC, downhill DSC)
// labels the cells it crosses -- for each crossed face, the cell it leaves and
// the face -- with the target value (4 ASC, 5 DSC), stops as soon as it steps
// into a cell (the voxel ASC, the nearest vertex DSC) that already carries the
// target, or near the boundary; after its streamlines the saddle itself takes
// the target.  One reference thread runs this in order, so later streamlines
// stop on earlier ones: the result is defined by that order.
//
// On the GPU the order is kept exactly by speculation over windows of
// streamlines ("lines"; the saddle's own label is a line without a trace):
//  1. every line of the window is traced at once against the labels committed
//     before the window (plus its own), recording its footprint: the cells it
//     read (tested for the target) and the cells it labelled;
//  2. lines are joined (union-find) when one labels a cell another's footprint
//     holds; a line alone in its component saw exactly the labels it would see
//     in order, so its labels stand;
//  3. each component of several lines is replayed by one lane, its lines in
//     order, re-tracing each against the committed labels and those of its
//     earlier members.  A replayed line that reads a cell outside its own
//     footprint, or labels one it did not label before, may depend on or
//     affect lines of other components: the window is cut before it;
//  4. the labels of the lines before the cut are committed; the next window
//     starts at the cut.  The first line of a window never fails (nothing
//     earlier in the window), so every window commits at least one line.
// A line whose footprint outgrows its record is traced alone, in order, when
// it is first in its window.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <hipcub/hipcub.hpp>

#include "ctx.hpp"
#include "integrate.hpp"
#include "lines.hpp"
#include "tracer.hpp"

#pragma clang fp contract(off)

namespace msceer {
namespace {

constexpr int kThreads = 256;
constexpr uint32_t kNoLine = 0xFFFFFFFFu;
constexpr int kFoot = 128;           // footprint entries recorded per line
constexpr uint8_t kRead = 1, kWrote = 2;
constexpr int kAbort = 100;          // trace event: the footprint rule stopped it

__device__ __forceinline__ uint32_t wave_append(bool want, uint32_t* counter) {
    const uint64_t m = __ballot(want);
    if (!m) return 0;
    const int lane = (int)(threadIdx.x & 63u);
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    return base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
}

struct CellGeom {
    int64_t Mx, My, Mz, sxy;
    __device__ uint32_t id(int64_t x, int64_t y, int64_t z) const { return (uint32_t)(x + Mx * (y + My * z)); }
    __device__ void xyz(uint32_t c, int64_t& x, int64_t& y, int64_t& z) const {
        x = (int64_t)c % Mx;
        y = ((int64_t)c / Mx) % My;
        z = (int64_t)c / sxy;
    }
    __device__ int64_t off(int a) const { return a == 0 ? 1 : (a == 1 ? Mx : sxy); }
    __device__ int64_t ext(int a) const { return a == 0 ? Mx : (a == 1 ? My : Mz); }
    // Cell coordinates inside the complex.  A streamline away from the
    // boundary stays inside; this only keeps a step that a huge error
    // threshold let run far from addressing outside the arrays.
    __device__ bool inside(int64_t x, int64_t y, int64_t z) const {
        return x >= 0 && y >= 0 && z >= 0 && x < Mx && y < My && z < Mz;
    }
};

// Order key of a float: monotone, -0 and +0 equal (as float comparison has them).
__device__ __forceinline__ uint32_t value_key(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) == 0u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ bool critical(const uint8_t* grad, uint32_t c) { return ((grad[c] >> 2) & 7u) == 7u; }

__device__ __forceinline__ uint32_t pair_of(const CellGeom& m, const uint8_t* grad, uint32_t c) {
    const uint32_t code = (grad[c] >> 2) & 7u;
    const int64_t o = m.off((int)(code >> 1));
    return (uint32_t)((int64_t)c + ((code & 1u) ? -o : o));
}

__device__ __forceinline__ double comp(const V3<double>& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// ---- the saddles and their seeds -------------------------------------------

// Critical cells of dimension sdim, keyed (float key of the value of the
// highest vertex -- negated for DSC -- << 32 | cell id): sorted descending,
// this is the reference's processing order (pairs (value, id) sorted
// ascending and taken from the end).
__global__ void __launch_bounds__(kThreads)
line_saddles_kernel(const uint8_t* __restrict__ grad, CellGeom m, int X, int Y, const float* __restrict__ f,
                    int sdim, int asc, uint64_t* __restrict__ keys, uint32_t* count, uint32_t cap) {
    const uint64_t C = (uint64_t)m.sxy * (uint64_t)m.Mz;
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t base = (uint64_t)blockIdx.x * kThreads; base < C; base += stride) {
        const uint64_t c = base + threadIdx.x;
        bool want = false;
        uint64_t key = 0;
        if (c < C && critical(grad, (uint32_t)c)) {
            int64_t x, y, z;
            m.xyz((uint32_t)c, x, y, z);
            if ((int)((x & 1) + (y & 1) + (z & 1)) == sdim) {
                // highest vertex by (value, id) among the cell's vertices
                int64_t best = -1;
                for (int k = 0; k < 8; k++) {
                    if (((k & 1) && !(x & 1)) || ((k & 2) && !(y & 1)) || ((k & 4) && !(z & 1))) continue;
                    const int64_t vx = (x >> 1) + (k & 1), vy = (y >> 1) + ((k >> 1) & 1), vz = (z >> 1) + (k >> 2);
                    const int64_t v = vx + (int64_t)X * (vy + (int64_t)Y * vz);
                    if (best < 0 || f[v] > f[best] || (!(f[best] > f[v]) && v > best)) best = v;
                }
                const float val = f[best];
                key = ((uint64_t)value_key(asc ? val : -1.0f * val) << 32) | (uint64_t)(uint32_t)c;
                want = true;
            }
        }
        const uint32_t at = wave_append(want, count);
        if (want && at < cap) keys[at] = key;
    }
}

// Per saddle (rank k in processing order): the seed cells of the short
// V-path search (ondemandaccurate.cxx:319-351 ASC, 448-480 DSC), at most six,
// in slots 7k .. 7k+5 (kNoLine where absent), and the saddle in slot 7k+6.
// The search pops at most five cells: the saddle, whose two cofacets (ASC) /
// facets (DSC) are new, then cells that each add one new neighbour.
__global__ void __launch_bounds__(kThreads)
line_seeds_kernel(const uint8_t* __restrict__ grad, CellGeom m, const uint64_t* __restrict__ keys, uint32_t n,
                  int asc, uint32_t* __restrict__ slots) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= n) return;
    const uint32_t s = (uint32_t)keys[k];
    uint32_t result[6], queue[6], seen[16];
    int nres = 0, qh = 0, qt = 0, ns = 0;
    auto is_seen = [&](uint32_t c) {
        for (int i = 0; i < ns; i++)
            if (seen[i] == c) return true;
        return false;
    };
    queue[qt++] = s;
    for (int counter = 4; qh < qt && counter >= 0; counter--) {
        const uint32_t cur = queue[qh++];
        if (!is_seen(cur) && ns < 16) seen[ns++] = cur;
        int64_t p[3];
        m.xyz(cur, p[0], p[1], p[2]);
        for (int a = 0; a < 3; a++) {
            if (asc ? (p[a] & 1) != 0 : (p[a] & 1) == 0) continue;
            for (int sg = 0; sg < 2; sg++) {
                if (sg == 0 ? p[a] >= m.ext(a) - 1 : p[a] <= 0) continue;
                const uint32_t t = (uint32_t)((int64_t)cur + (sg == 0 ? m.off(a) : -m.off(a)));
                if (is_seen(t)) continue;
                if (nres < 6) result[nres++] = t;
                if (ns < 16) seen[ns++] = t;
                if (!critical(grad, t)) {
                    const uint32_t pr = pair_of(m, grad, t);
                    if (ns < 16) seen[ns++] = pr;
                    if (qt < 6) queue[qt++] = pr;
                }
            }
        }
    }
    uint32_t* out = slots + 7 * (uint64_t)k;
    for (int j = 0; j < 6; j++) out[j] = j < nres ? result[j] : kNoLine;
    out[6] = s;
}

// ---- one streamline ----------------------------------------------------------

// Label policies of a trace: test(c) -- does cell c carry the target (a
// checker or digitizing test); write(c) -- give it the target; ok() -- false
// once the trace must stop for a footprint rule.

// Alone and in order: the restriction labels themselves.
struct DirectLabels {
    int8_t* restr;
    int8_t target;
    __device__ bool test(uint32_t c) { return restr[c] == target; }
    __device__ void write(uint32_t c) { restr[c] = target; }
    __device__ bool ok() const { return true; }
};

// Speculation: committed labels plus the line's own; every cell tested or
// labelled goes to the footprint (cells, flags), labelled ones also claim
// wstamp (lowest line of the window that labels the cell).
struct SpecLabels {
    const int8_t* restr;
    int8_t target;
    uint32_t* wstamp;
    uint32_t line;
    uint32_t* cells;
    uint8_t* flags;
    int n;
    bool over;
    __device__ int find(uint32_t c) const {
        for (int i = 0; i < n; i++)
            if (cells[i] == c) return i;
        return -1;
    }
    __device__ int add(uint32_t c) {
        int i = find(c);
        if (i >= 0) return i;
        if (n >= kFoot) {
            over = true;
            return -1;
        }
        cells[n] = c;
        flags[n] = 0;
        return n++;
    }
    __device__ bool test(uint32_t c) {
        const int i = add(c);
        if (i < 0) return false;
        flags[i] |= kRead;
        return (flags[i] & kWrote) || restr[c] == target;
    }
    __device__ void write(uint32_t c) {
        const int i = add(c);
        if (i < 0) return;
        flags[i] |= kWrote;
        atomicMin(&wstamp[c], line);
    }
    __device__ bool ok() const { return !over; }
};

// Replay: committed labels plus those of the component's earlier lines
// (lstamp); every access must stay inside the line's speculative footprint
// and every label inside its speculative labels, else the window is cut here.
struct ReplayLabels {
    const int8_t* restr;
    int8_t target;
    uint32_t* lstamp;
    uint32_t line;
    const uint32_t* cells;
    const uint8_t* flags;
    int n;
    uint32_t* fail;
    bool bad;
    __device__ int find(uint32_t c) const {
        for (int i = 0; i < n; i++)
            if (cells[i] == c) return i;
        return -1;
    }
    __device__ bool test(uint32_t c) {
        if (find(c) < 0) {
            bad = true;
            return false;
        }
        return restr[c] == target || lstamp[c] != kNoLine;
    }
    __device__ void write(uint32_t c) {
        const int i = find(c);
        if (i < 0 || !(flags[i] & kWrote)) {
            bad = true;
            return;
        }
        lstamp[c] = line;
    }
    __device__ bool ok() const { return !bad; }
};

// DigitizeSegmentInternal: the faces the segment a -> b crosses, by their
// crossing parameter; for each face that bounds the current cell, label the
// current cell and the face (stop at the first already labelled) and move to
// the cell beyond.  ASC: voxel faces, cells = hexes; DSC: the dual grid
// (positions shifted by half a voxel), faces = edges, cells = vertices.
template <class L>
__device__ void digitize(const CellGeom& m, V3<double> a, V3<double> b, bool asc, L& lab) {
    if (!asc) {
        a = {a.x + 0.5, a.y + 0.5, a.z + 0.5};
        b = {b.x + 0.5, b.y + 0.5, b.z + 0.5};
    }
    const V3<double> ab{b.x - a.x, b.y - a.y, b.z - a.z};
    uint32_t face[3];
    double tc[3];
    int n = 0;
    for (int d = 0; d < 3; d++) {
        const double av = comp(a, d), bv = comp(b, d);
        const int va = (int)floor(av), vb = (int)floor(bv);
        if (va == vb) continue;
        const int plane = va > vb ? va : vb;
        const double t = ((double)plane - av) / (bv - av);
        tc[n] = t;
        if (t < 0 || t > 1) return;
        int64_t fc[3];
        for (int k = 0; k < 3; k++)
            fc[k] = 2 * (int64_t)floor(comp(ab, k) * t + comp(a, k)) + (k == d ? 0 : 1) - (asc ? 0 : 1);
        if (!m.inside(fc[0], fc[1], fc[2])) return;
        face[n++] = m.id(fc[0], fc[1], fc[2]);
    }
    if (n == 0) return;
    for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++)
            if (tc[i] > tc[j]) {
                const uint32_t tf = face[i]; face[i] = face[j]; face[j] = tf;
                const double tt = tc[i]; tc[i] = tc[j]; tc[j] = tt;
            }
    const int64_t s0 = 2 * (int64_t)floor(a.x) + (asc ? 1 : 0), s1 = 2 * (int64_t)floor(a.y) + (asc ? 1 : 0),
                  s2 = 2 * (int64_t)floor(a.z) + (asc ? 1 : 0);
    if (!m.inside(s0, s1, s2)) return;
    uint32_t cur = m.id(s0, s1, s2);
    for (int i = 0; i < n; i++) {
        int64_t p[3];
        m.xyz(face[i], p[0], p[1], p[2]);
        bool has = false;
        uint32_t next = cur;
        for (int d = 0; d < 3; d++) {
            if (asc ? (p[d] & 1) != 0 : (p[d] & 1) == 0) continue;
            if (p[d] < m.ext(d) - 1) {
                const uint32_t c = (uint32_t)((int64_t)face[i] + m.off(d));
                if (c == cur) has = true; else next = c;
            }
            if (p[d] > 0) {
                const uint32_t c = (uint32_t)((int64_t)face[i] - m.off(d));
                if (c == cur) has = true; else next = c;
            }
        }
        if (!has) continue;
        if (lab.test(cur)) return;
        lab.write(cur);
        if (!lab.ok()) return;
        if (lab.test(face[i])) return;
        lab.write(face[i]);
        if (!lab.ok()) return;
        cur = next;
    }
}

// AdvectThroughVoxel{NoCheck,NearBoundary} in FP64 with the line checker:
// near the boundary the first accepted step ends the line; elsewhere a step
// into a cell carrying the target does.
template <class L>
__device__ int advect_voxel(const VGrid& g, const CellGeom& m, Tracer<double>& t, double dir, double gthr,
                            double ethr, bool asc, L& lab) {
    while (true) {
        if (!(t.iters-- > 0)) return kOverMax;
        if (!(t.old.x == t.cur.x && t.old.y == t.cur.y && t.old.z == t.cur.z)) {
            if ((int)t.cur.x != (int)t.old.x || (int)t.cur.y != (int)t.old.y || (int)t.cur.z != (int)t.old.z)
                return kOutOfVoxel;
            t.gr = trilin(t.c, t.cur.x - (double)t.bx, t.cur.y - (double)t.by, t.cur.z - (double)t.bz);
            t.gm = sqrt(t.gr.x * t.gr.x + t.gr.y * t.gr.y + t.gr.z * t.gr.z);
            if (t.gm * t.step > kMaxStep) t.step = kMaxStep / t.gm;
            t.old = t.cur;
        }
        if (t.gm < gthr) return kLowGradient;
        const double s = dir * t.step, hs = (dir * 0.5) * t.step;
        const V3<double> full{t.gr.x * s, t.gr.y * s, t.gr.z * s};
        const V3<double> half{t.gr.x * hs, t.gr.y * hs, t.gr.z * hs};
        const V3<double> hg = trilin(t.c, t.cur.x + half.x - (double)t.bx, t.cur.y + half.y - (double)t.by,
                                     t.cur.z + half.z - (double)t.bz);
        const V3<double> two{hg.x * hs + half.x, hg.y * hs + half.y, hg.z * hs + half.z};
        const double dx = two.x - full.x, dy = two.y - full.y, dz = two.z - full.z;
        const double err = dx * dx + dy * dy + dz * dz;
        if (!(err < ethr)) {
            t.step = t.step * 0.5;
            continue;
        }
        t.cur = {t.cur.x + two.x, t.cur.y + two.y, t.cur.z + two.z};
        if (t.near) { t.cur.x = clampc(t.cur.x, g.X); t.cur.y = clampc(t.cur.y, g.Y); t.cur.z = clampc(t.cur.z, g.Z); }
        const double grow = t.step * 1.8, cap = kMaxStep / t.gm;
        t.step = grow < cap ? grow : cap;
        if (t.near) return kHitPreassigned;
        const int64_t cx = asc ? 2 * (int64_t)floor(t.cur.x) + 1 : 2 * (int64_t)floor(t.cur.x + 0.5);
        const int64_t cy = asc ? 2 * (int64_t)floor(t.cur.y) + 1 : 2 * (int64_t)floor(t.cur.y + 0.5);
        const int64_t cz = asc ? 2 * (int64_t)floor(t.cur.z) + 1 : 2 * (int64_t)floor(t.cur.z + 0.5);
        if (!m.inside(cx, cy, cz)) continue;
        const bool hit = lab.test(m.id(cx, cy, cz));
        if (!lab.ok()) return kAbort;
        if (hit) return kHitPreassigned;
    }
}

// IntegrateStreamline from the centre of seed cell `seed` (cell coordinates / 2).
template <class L>
__device__ int trace_line(const GradSrc<double>& src, const VGrid& g, const CellGeom& m, uint32_t seed, bool asc,
                          double gthr, double ethr, int max_iters, L& lab, uint32_t* voxels) {
    int64_t sx, sy, sz;
    m.xyz(seed, sx, sy, sz);
    Tracer<double> t;
    t.cur = {(double)sx * 0.5, (double)sy * 0.5, (double)sz * 0.5};
    t.iters = max_iters;
    tracer_load(src, g, t, (int)(t.cur.x + 0.5), (int)(t.cur.y + 0.5), (int)(t.cur.z + 0.5));
    tracer_begin(t);
    const double dir = asc ? 1.0 : -1.0;
    while (true) {
        const V3<double> start = t.cur;
        const int ev = advect_voxel(g, m, t, dir, gthr, ethr, asc, lab);
        (*voxels)++;
        if (ev == kAbort) return kAbort;
        if (!t.near) {
            digitize(m, start, t.cur, asc, lab);
            if (!lab.ok()) return kAbort;
        }
        if (ev != kOutOfVoxel) return ev;
        int x, y, z;
        nearest_xyz(g, t.cur, t.near, x, y, z);
        tracer_load(src, g, t, x, y, z);
        tracer_begin(t);
    }
}

struct LineArgs {
    GradSrc<double> src;
    VGrid g;
    CellGeom m;
    const uint32_t* slots;  // 7 per saddle
    int8_t* restr;
    int8_t target;
    bool asc;
    double gthr, ethr;
    int max_iters;
    uint32_t* wstamp;  // per cell, kNoLine outside a window
    uint32_t* lstamp;  // per cell, kNoLine outside a window
    uint32_t* foot;    // kFoot cells per line of the window
    uint8_t* fflags;   // kFoot flags per line of the window
    uint32_t* fcount;  // entries per line of the window
    uint32_t* ctl;     // kLineCtl words
};

enum LineCtl : int {
    kLcOverflow = 0,  // lowest line whose footprint outgrew its record
    kLcFail = 1,      // lowest line whose replay left its footprint
    kLcComps = 2,     // components of the window
    kLcTraced = 3,    // streamlines traced (all phases)
    kLcReplayed = 4,  // streamlines re-traced in replays
    kLcVoxels = 5,    // voxel traversals (all phases)
    kLcWords = 8
};

// 1. Speculation over lines [w0, w1).
__global__ void __launch_bounds__(kThreads) line_spec_kernel(LineArgs a, uint32_t w0, uint32_t w1) {
    const uint32_t i = w0 + blockIdx.x * kThreads + threadIdx.x;
    if (i >= w1) return;
    const uint64_t r = i - w0;
    SpecLabels lab{a.restr, a.target, a.wstamp, i, a.foot + r * kFoot, a.fflags + r * kFoot, 0, false};
    const uint32_t c = a.slots[i];
    if (c != kNoLine) {
        if (i % 7 == 6) {
            lab.write(c);  // the saddle's own label
        } else {
            uint32_t vox = 0;
            trace_line(a.src, a.g, a.m, c, a.asc, a.gthr, a.ethr, a.max_iters, lab, &vox);
            atomicAdd(&a.ctl[kLcTraced], 1u);
            atomicAdd(&a.ctl[kLcVoxels], vox);
        }
    }
    a.fcount[r] = (uint32_t)lab.n;
    if (lab.over) atomicMin(&a.ctl[kLcOverflow], i);
}

__device__ __forceinline__ uint32_t lf_parent(uint32_t* p, uint32_t x) {
    return __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ uint32_t lf_find(uint32_t* p, uint32_t x) {
    uint32_t q = lf_parent(p, x);
    while (q != x) {
        const uint32_t gq = lf_parent(p, q);
        if (gq != q) __hip_atomic_store(&p[x], gq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = q;
        q = gq;
    }
    return x;
}
// Hook the later root under the earlier one, so a component's root is its
// first line.
__device__ void lf_union(uint32_t* p, uint32_t a, uint32_t b) {
    while (true) {
        a = lf_find(p, a);
        b = lf_find(p, b);
        if (a == b) return;
        if (a > b) { const uint32_t t = a; a = b; b = t; }
        uint32_t expected = b;
        if (__hip_atomic_compare_exchange_strong(&p[b], &expected, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}

// 2. Components: a line joins the lowest line of [w0, w1) that labels any
// cell of its footprint (every line that labels a cell also tests it, so all
// lines touching a labelled cell meet in one component).
__global__ void __launch_bounds__(kThreads) line_init_parent_kernel(uint32_t* parent, uint32_t n) {
    const uint32_t r = blockIdx.x * kThreads + threadIdx.x;
    if (r < n) parent[r] = r;
}
__global__ void __launch_bounds__(kThreads) line_union_kernel(LineArgs a, uint32_t w0, uint32_t w1, uint32_t* parent) {
    const uint32_t i = w0 + blockIdx.x * kThreads + threadIdx.x;
    if (i >= w1) return;
    const uint64_t r = i - w0;
    const uint32_t n = a.fcount[r];
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t w = a.wstamp[a.foot[r * kFoot + j]];
        if (w != i && w >= w0 && w < w1) lf_union(parent, (uint32_t)r, w - w0);
    }
}
__global__ void __launch_bounds__(kThreads) line_keys_kernel(uint32_t* parent, uint32_t n, uint64_t* keys) {
    const uint32_t r = blockIdx.x * kThreads + threadIdx.x;
    if (r < n) keys[r] = ((uint64_t)lf_find(parent, r) << 32) | r;
}
// First entry of every component in the sorted keys.
__global__ void __launch_bounds__(kThreads)
line_starts_kernel(const uint64_t* keys, uint32_t n, uint32_t* starts, uint32_t* count) {
    const uint32_t stride = gridDim.x * kThreads;
    for (uint32_t base = blockIdx.x * kThreads; base < n; base += stride) {
        const uint32_t r = base + threadIdx.x;
        const bool want = r < n && (r == 0 || (keys[r] >> 32) != (keys[r - 1] >> 32));
        const uint32_t at = wave_append(want, count);
        if (want) starts[at] = r;
    }
}

// 3. Replay: one lane per component.  A lone line keeps its speculative
// labels; the lines of a larger component are re-traced in order.
__global__ void __launch_bounds__(kThreads)
line_replay_kernel(LineArgs a, uint32_t w0, uint32_t n, const uint64_t* keys, const uint32_t* starts) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= a.ctl[kLcComps]) return;
    const uint32_t s0 = starts[k];
    const uint32_t root = (uint32_t)(keys[s0] >> 32);
    uint32_t e = s0 + 1;
    while (e < n && (uint32_t)(keys[e] >> 32) == root) e++;
    if (e - s0 == 1) {
        const uint64_t r = (uint32_t)keys[s0];
        const uint32_t cnt = a.fcount[r];
        for (uint32_t j = 0; j < cnt; j++)
            if (a.fflags[r * kFoot + j] & kWrote) a.lstamp[a.foot[r * kFoot + j]] = w0 + (uint32_t)r;
        return;
    }
    uint32_t replayed = 0, vox = 0;
    for (uint32_t q = s0; q < e; q++) {
        const uint64_t r = (uint32_t)keys[q];
        const uint32_t i = w0 + (uint32_t)r;
        ReplayLabels lab{a.restr, a.target, a.lstamp, i, a.foot + r * kFoot, a.fflags + r * kFoot,
                         (int)a.fcount[r], &a.ctl[kLcFail], false};
        const uint32_t c = a.slots[i];
        if (c != kNoLine) {
            if (i % 7 == 6) {
                lab.write(c);
            } else {
                trace_line(a.src, a.g, a.m, c, a.asc, a.gthr, a.ethr, a.max_iters, lab, &vox);
                replayed++;
            }
        }
        if (lab.bad) {
            atomicMin(&a.ctl[kLcFail], i);
            break;
        }
    }
    atomicAdd(&a.ctl[kLcReplayed], replayed);
    atomicAdd(&a.ctl[kLcVoxels], vox);
}

// 4. Commit the labels of the lines before `cut`, then clear the stamps of
// every footprint of the window.
__global__ void __launch_bounds__(kThreads) line_commit_kernel(LineArgs a, uint32_t w0, uint32_t w1, uint32_t cut) {
    const uint32_t i = w0 + blockIdx.x * kThreads + threadIdx.x;
    if (i >= w1) return;
    const uint64_t r = i - w0;
    const uint32_t n = a.fcount[r];
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t c = a.foot[r * kFoot + j];
        if (a.lstamp[c] < cut) a.restr[c] = a.target;
    }
}
__global__ void __launch_bounds__(kThreads) line_clear_kernel(LineArgs a, uint32_t w0, uint32_t w1) {
    const uint32_t i = w0 + blockIdx.x * kThreads + threadIdx.x;
    if (i >= w1) return;
    const uint64_t r = i - w0;
    const uint32_t n = a.fcount[r];
    for (uint32_t j = 0; j < n; j++) {
        const uint32_t c = a.foot[r * kFoot + j];
        a.wstamp[c] = kNoLine;
        a.lstamp[c] = kNoLine;
    }
}

// A line traced alone, in order (first of its window, footprint too large).
__global__ void line_direct_kernel(LineArgs a, uint32_t i) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    DirectLabels lab{a.restr, a.target};
    const uint32_t c = a.slots[i];
    if (c == kNoLine) return;
    if (i % 7 == 6) {
        lab.write(c);
        return;
    }
    uint32_t vox = 0;
    trace_line(a.src, a.g, a.m, c, a.asc, a.gthr, a.ethr, a.max_iters, lab, &vox);
    a.ctl[kLcTraced] += 1;
    a.ctl[kLcVoxels] += vox;
}

__global__ void line_ctl_init_kernel(uint32_t* ctl, bool all) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl[kLcOverflow] = kNoLine;
        ctl[kLcFail] = kNoLine;
        ctl[kLcComps] = 0;
        if (all) {
            ctl[kLcTraced] = 0;
            ctl[kLcReplayed] = 0;
            ctl[kLcVoxels] = 0;
        }
    }
}

__global__ void __launch_bounds__(kThreads) line_fill_kernel(uint32_t* p, uint64_t n, uint32_t v) {
    const uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) p[i] = v;
}

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)std::max<uint64_t>(1, (n + kThreads - 1) / kThreads); }

}  // namespace

// Window of lines speculated at once (bounds the footprint scratch).  The
// environment variable MSCEER_LINE_WINDOW (8 .. 2^20) sets a smaller one,
// which tests use to exercise many windows on small volumes.
constexpr uint32_t kWindow = 1u << 20;

uint32_t line_window() {
    if (const char* e = getenv("MSCEER_LINE_WINDOW")) {
        const long v = atol(e);
        if (v >= 8 && v <= (long)kWindow) return (uint32_t)v;
    }
    return kWindow;
}

int run_lines(msceer_ctx* ctx, const Dims& d, const float* values, const uint8_t* grad, const double4* gradd,
              int8_t* restr, bool asc, float gthr, float ethr, int max_iters, LineStats* out) {
    hipStream_t s = ctx->stream;
    const VGrid g = make_vgrid(d.X, d.Y, d.Z);
    CellGeom m{(int64_t)d.Mx, (int64_t)d.My, (int64_t)d.Mz, (int64_t)d.Mx * (int64_t)d.My};
    if (d.cells >= 0xFFFFFFFFull) return MSCEER_ETOOBIG;
    uint32_t* ctl = ctl_dev(ctx);
    uint32_t* count = ctl + kCtlCount;

    // stamps: one word each per cell, kNoLine between windows
    CHECK_HIP(ctx->line_stamps.ensure(8 * d.cells));
    uint32_t* wstamp = ctx->line_stamps.as<uint32_t>();
    uint32_t* lstamp = wstamp + d.cells;
    if (!ctx->line_stamps_clear || ctx->line_stamps_cells != d.cells) {
        hipLaunchKernelGGL(line_fill_kernel, dim3(2048), dim3(kThreads), 0, s, wstamp, 2 * d.cells, kNoLine);
        CHECK_HIP(hipGetLastError());
        ctx->line_stamps_cells = d.cells;
    }
    ctx->line_stamps_clear = false;  // until every window has cleared its footprints

    // saddles: count first, then keys sorted into processing order
    CHECK_HIP(launch_zero(count, 1, s));
    hipLaunchKernelGGL(line_saddles_kernel, dim3(2048), dim3(kThreads), 0, s, grad, m, d.X, d.Y, values, asc ? 2 : 1,
                       asc ? 1 : 0, nullptr, count, 0u);
    CHECK_HIP(hipGetLastError());
    int rc = read_ctl(ctx, kCtlCount, 1);
    if (rc) return rc;
    const uint32_t S = ctx->h_ctl[kCtlCount];
    out->saddles = S;
    if (S == 0) {
        ctx->line_stamps_clear = true;
        return MSCEER_OK;
    }
    if ((uint64_t)S * 7 >= 0xFFFFFFFFull) return MSCEER_ETOOBIG;
    const uint64_t nslots = 7ull * S;
    const uint32_t W = (uint32_t)std::min<uint64_t>(nslots, line_window());

    // scratch layout
    size_t sort_tmp = 0, sort_tmp2 = 0;
    CHECK_HIP(hipcub::DeviceRadixSort::SortKeysDescending(nullptr, sort_tmp, (const uint64_t*)nullptr,
                                                          (uint64_t*)nullptr, (int)S, 0, 64, s));
    CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, sort_tmp2, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                                (int)W, 0, 64, s));
    const size_t tmp_bytes = std::max(sort_tmp, sort_tmp2);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t off = 0;
    const size_t o_keys = off; off += al(8 * (size_t)S);
    const size_t o_sorted = off; off += al(8 * (size_t)S);
    const size_t o_slots = off; off += al(4 * nslots);
    const size_t o_foot = off; off += al((size_t)W * 4 * kFoot);
    const size_t o_fflags = off; off += al((size_t)W * kFoot);
    const size_t o_fcount = off; off += al((size_t)W * 4);
    const size_t o_parent = off; off += al((size_t)W * 4);
    const size_t o_wkeys = off; off += al((size_t)W * 8);
    const size_t o_wsorted = off; off += al((size_t)W * 8);
    const size_t o_starts = off; off += al((size_t)W * 4);
    const size_t o_lctl = off; off += al(4 * kLcWords);
    const size_t o_tmp = off; off += al(tmp_bytes);
    CHECK_HIP(ctx->lines.ensure(off));
    char* base = ctx->lines.as<char>();
    uint64_t* keys = (uint64_t*)(base + o_keys);
    uint64_t* sorted = (uint64_t*)(base + o_sorted);
    uint32_t* slots = (uint32_t*)(base + o_slots);
    uint32_t* parent = (uint32_t*)(base + o_parent);
    uint64_t* wkeys = (uint64_t*)(base + o_wkeys);
    uint64_t* wsorted = (uint64_t*)(base + o_wsorted);
    uint32_t* starts = (uint32_t*)(base + o_starts);
    uint32_t* lctl = (uint32_t*)(base + o_lctl);
    void* tmp = base + o_tmp;

    CHECK_HIP(launch_zero(count, 1, s));
    hipLaunchKernelGGL(line_saddles_kernel, dim3(2048), dim3(kThreads), 0, s, grad, m, d.X, d.Y, values, asc ? 2 : 1,
                       asc ? 1 : 0, keys, count, S);
    CHECK_HIP(hipGetLastError());
    size_t tb = tmp_bytes;
    CHECK_HIP(hipcub::DeviceRadixSort::SortKeysDescending(tmp, tb, keys, sorted, (int)S, 0, 64, s));
    hipLaunchKernelGGL(line_seeds_kernel, dim3(blocks_for(S)), dim3(kThreads), 0, s, grad, m, sorted, S,
                       asc ? 1 : 0, slots);
    CHECK_HIP(hipGetLastError());

    LineArgs a{};
    a.src = GradSrc<double>{gradd};
    a.g = g;
    a.m = m;
    a.slots = slots;
    a.restr = restr;
    a.target = asc ? 4 : 5;
    a.asc = asc;
    a.gthr = (double)gthr;
    a.ethr = (double)ethr;
    a.max_iters = max_iters;
    a.wstamp = wstamp;
    a.lstamp = lstamp;
    a.foot = (uint32_t*)(base + o_foot);
    a.fflags = (uint8_t*)(base + o_fflags);
    a.fcount = (uint32_t*)(base + o_fcount);
    a.ctl = lctl;
    hipLaunchKernelGGL(line_ctl_init_kernel, dim3(1), dim3(64), 0, s, lctl, true);
    CHECK_HIP(hipGetLastError());

    uint32_t h[kLcWords];
    uint64_t w0 = 0;
    while (w0 < nslots) {
        const uint32_t a0 = (uint32_t)w0;
        const uint32_t w1 = (uint32_t)std::min<uint64_t>(nslots, w0 + W);
        const uint32_t n = w1 - a0;
        hipLaunchKernelGGL(line_ctl_init_kernel, dim3(1), dim3(64), 0, s, lctl, false);
        hipLaunchKernelGGL(line_spec_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, s, a, a0, w1);
        CHECK_HIP(hipGetLastError());
        CHECK_HIP(hipMemcpyAsync(h, lctl, sizeof(h), hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipStreamSynchronize(s));
        out->windows++;
        if (h[kLcOverflow] == a0) {
            // the first line's footprint outgrew its record: it alone, in order
            hipLaunchKernelGGL(line_clear_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, s, a, a0, w1);
            hipLaunchKernelGGL(line_direct_kernel, dim3(1), dim3(64), 0, s, a, a0);
            CHECK_HIP(hipGetLastError());
            out->direct++;
            w0 = a0 + 1;
            continue;
        }
        const uint32_t wc = h[kLcOverflow] < w1 ? h[kLcOverflow] : w1;  // lines of this window
        const uint32_t nc = wc - a0;
        hipLaunchKernelGGL(line_init_parent_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, parent, nc);
        hipLaunchKernelGGL(line_union_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, a, a0, wc, parent);
        hipLaunchKernelGGL(line_keys_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, parent, nc, wkeys);
        CHECK_HIP(hipGetLastError());
        tb = tmp_bytes;
        CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, tb, wkeys, wsorted, (int)nc, 0, 64, s));
        hipLaunchKernelGGL(line_starts_kernel, dim3(std::min<uint32_t>(blocks_for(nc), 1024)), dim3(kThreads), 0, s,
                           wsorted, nc, starts, lctl + kLcComps);
        hipLaunchKernelGGL(line_replay_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, a, a0, nc, wsorted,
                           starts);
        CHECK_HIP(hipGetLastError());
        CHECK_HIP(hipMemcpyAsync(h, lctl, sizeof(h), hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipStreamSynchronize(s));
        const uint32_t cut = h[kLcFail] < wc ? h[kLcFail] : wc;
        if (cut == a0) return MSCEER_EHIP;  // the first line of a window cannot fail
        hipLaunchKernelGGL(line_commit_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, a, a0, wc, cut);
        hipLaunchKernelGGL(line_clear_kernel, dim3(blocks_for(n)), dim3(kThreads), 0, s, a, a0, w1);
        CHECK_HIP(hipGetLastError());
        out->components += h[kLcComps];
        if (cut < wc) out->cuts++;
        w0 = cut;
    }
    CHECK_HIP(hipMemcpyAsync(h, lctl, sizeof(h), hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    ctx->line_stamps_clear = true;
    out->traced = h[kLcTraced];
    out->replayed = h[kLcReplayed];
    out->voxels = h[kLcVoxels];
    return MSCEER_OK;
}

}  // namespace msceer

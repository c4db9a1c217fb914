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
//     before the window plus its own, recording its footprint: the cells it
//     tested for the target and the cells it labelled (a cell labelled by
//     lines of the window belongs to the lowest of them, wstamp);
//  2. lines are joined (union-find) when one labels a cell in another's
//     footprint.  A line whose footprint no earlier line of the window
//     labels saw exactly the labels it would see in order: its labels stand;
//  3. each component is replayed by one lane, its lines in order: those
//     depending on an earlier line are re-traced against the committed labels
//     and the labels of their component's earlier lines.  A replayed line
//     that tests a cell of another component's labels, or labels a cell its
//     component did not label in speculation, went beyond what the
//     speculation saw: the window is cut before it;
//  4. the labels of the lines before the cut are committed; the next window
//     starts at the cut.  The first line of a window depends on nothing, so
//     every window commits at least one line.
// A line whose footprint does not fit the pool ends its window; first in a
// window, it is traced alone, in order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <hipcub/hipcub.hpp>

#include "ctx.hpp"
#include "integrate.hpp"
#include "lines.hpp"
#include "lines_core.hpp"

#pragma clang fp contract(off)

namespace msceer {
namespace {

constexpr int kThreads = 256;

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

// Per saddle (rank k in processing order): its seeds in slots 7k .. 7k+5 and
// the saddle in slot 7k+6.
__global__ void __launch_bounds__(kThreads)
line_seeds_kernel(const uint8_t* __restrict__ grad, CellGeom m, const uint64_t* __restrict__ keys, uint32_t n,
                  int asc, uint32_t* __restrict__ slots) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k < n) saddle_seeds(grad, m, (uint32_t)keys[k], asc != 0, slots + 7 * (uint64_t)k);
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
    uint32_t* pstamp;  // per cell, kNoLine outside a window
    uint32_t* lstamp;  // per cell, kNoLine outside a window
    uint32_t* own;     // per cell, kNoLine outside a window
    Pool pool;
    uint32_t* fhead;   // first footprint chunk per line of the window (kNoLine: none)
    uint32_t* whead;   // first chunk of the cells each line labelled
    uint32_t* mhead;   // first chunk of the cells each line's stops could label
    uint32_t* touched; // cells labelled speculatively in the window
    uint32_t* ctl;     // kLcWords words
};

enum LineCtl : int {
    kLcOverflow = 0,  // lowest line whose footprint did not fit the pool
    kLcFail = 1,      // lowest line whose replay left its component
    kLcComps = 2,     // components of the window
    kLcTraced = 3,    // streamlines traced (all phases)
    kLcReplayed = 4,  // streamlines re-traced in replays
    kLcVoxels = 5,    // voxel traversals (all phases)
    kLcTouched = 6,   // cells labelled speculatively in the window
    kLcPoolTop = 7,   // next free word of the footprint pool
    kLcWhyTest = 8,   // replays stopped by a test of another component's cell
    kLcWhyLabel = 9,  // replays stopped by a label the speculation did not give
    kLcWords = 12
};

// 1. Speculation over lines [w0, w1).
__global__ void __launch_bounds__(kThreads) line_spec_kernel(LineArgs a, uint32_t w0, uint32_t w1) {
    const uint32_t i = w0 + blockIdx.x * kThreads + threadIdx.x;
    if (i >= w1) return;
    SpecLabels lab{a.restr, a.target, a.wstamp, a.pstamp, a.own, i, a.pool, a.touched, &a.ctl[kLcTouched]};
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
    lab.finish();
    a.fhead[i - w0] = lab.over ? kNoLine : lab.all.head;
    a.whead[i - w0] = lab.over ? kNoLine : lab.wrote.head;
    a.mhead[i - w0] = lab.over ? kNoLine : lab.maybe.head;
    if (lab.over) atomicMin(&a.ctl[kLcOverflow], i);
}

// 2. Components: a line joins the lowest line of [w0, wc) that labels any
// cell of its footprint (every line that labels a cell also has it in its
// footprint, so all lines touching a labelled cell meet in one component).
__global__ void __launch_bounds__(kThreads) line_init_parent_kernel(uint32_t* parent, uint32_t n) {
    const uint32_t r = blockIdx.x * kThreads + threadIdx.x;
    if (r < n) parent[r] = r;
}
// Also flags the lines that depend on an earlier line of the window (an
// earlier line labels a cell of their footprint): only those are re-traced.
__global__ void __launch_bounds__(kThreads)
line_union_kernel(LineArgs a, uint32_t w0, uint32_t wc, uint32_t* parent, uint8_t* dep) {
    const uint32_t i = w0 + blockIdx.x * kThreads + threadIdx.x;
    if (i >= wc) return;
    bool d = false;
    for (uint32_t ch = a.fhead[i - w0]; ch != kNoLine; ch = a.pool.cells[ch]) {
        for (uint32_t j = 1; j < kChunk; j++) {
            const uint32_t c = a.pool.cells[ch + j];
            if (c == kNoLine) break;
            const uint32_t owners[2] = {a.wstamp[c], a.pstamp[c]};
            for (uint32_t w : owners)
                if (w != i && w >= w0 && w < wc) {
                    lf_union(parent, i - w0, w - w0);
                    d = d || w < i;
                }
        }
    }
    dep[i - w0] = d ? 1 : 0;
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

// 3. Replay: one lane per component, its lines in order.  A line no earlier
// line of the window depends on keeps its speculative labels (the cells of
// its footprint it claimed in wstamp: no earlier line labels any cell it
// reads, so it saw what it would see in order); the others are re-traced.
__device__ void keep_speculation(const LineArgs& a, uint32_t w0, uint32_t i) {
    for (uint32_t ch = a.fhead[i - w0]; ch != kNoLine; ch = a.pool.cells[ch])
        for (uint32_t j = 1; j < kChunk; j++) {
            const uint32_t c = a.pool.cells[ch + j];
            if (c == kNoLine) break;
            if (a.wstamp[c] == i) a.lstamp[c] = i;
        }
}

__global__ void __launch_bounds__(kThreads)
line_replay_kernel(LineArgs a, uint32_t w0, uint32_t wc, const uint64_t* keys, const uint32_t* starts,
                   uint32_t* parent, const uint8_t* dep) {
    const uint32_t k = blockIdx.x * kThreads + threadIdx.x;
    if (k >= a.ctl[kLcComps]) return;
    const uint32_t n = wc - w0;
    const uint32_t s0 = starts[k];
    const uint32_t root = (uint32_t)(keys[s0] >> 32);
    uint32_t replayed = 0, vox = 0;
    for (uint32_t q = s0; q < n && (uint32_t)(keys[q] >> 32) == root; q++) {
        const uint32_t i = w0 + (uint32_t)keys[q];
        if (!dep[i - w0]) {
            keep_speculation(a, w0, i);
            continue;
        }
        ReplayLabels lab{a.restr, a.target,      a.wstamp,       a.pstamp, a.lstamp, a.own,
                         a.pool.cells, a.whead[i - w0], a.mhead[i - w0], parent, w0, wc, root, i,
                         &a.ctl[kLcWhyTest], false};
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
// every cell labelled speculatively in the window.
__global__ void __launch_bounds__(kThreads) line_commit_kernel(LineArgs a, uint32_t cut) {
    const uint32_t n = min(a.ctl[kLcTouched], a.pool.capacity);
    for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < n; t += gridDim.x * kThreads) {
        const uint32_t c = a.touched[t];
        if (a.lstamp[c] < cut) a.restr[c] = a.target;
    }
}
__global__ void __launch_bounds__(kThreads) line_clear_kernel(LineArgs a) {
    const uint32_t n = min(a.ctl[kLcTouched], a.pool.capacity);
    for (uint32_t t = blockIdx.x * kThreads + threadIdx.x; t < n; t += gridDim.x * kThreads) {
        const uint32_t c = a.touched[t];
        a.wstamp[c] = kNoLine;
        a.pstamp[c] = kNoLine;
        a.lstamp[c] = kNoLine;
        a.own[c] = kNoLine;
    }
}

// A line traced alone, in order (first of its window, footprint beyond the pool).
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
        ctl[kLcTouched] = 0;
        ctl[kLcPoolTop] = 0;
        if (all) {
            ctl[kLcTraced] = 0;
            ctl[kLcReplayed] = 0;
            ctl[kLcVoxels] = 0;
            ctl[kLcWhyTest] = 0;
            ctl[kLcWhyLabel] = 0;
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

    // per-cell stamps (lowest speculative labeller, lowest possible labeller,
    // replay labeller, last speculative labeller), kNoLine between windows
    CHECK_HIP(ctx->line_stamps.ensure(16 * d.cells));
    uint32_t* wstamp = ctx->line_stamps.as<uint32_t>();
    uint32_t* pstamp = wstamp + d.cells;
    uint32_t* lstamp = pstamp + d.cells;
    uint32_t* own = lstamp + d.cells;
    if (!ctx->line_stamps_clear || ctx->line_stamps_cells != d.cells) {
        hipLaunchKernelGGL(line_fill_kernel, dim3(2048), dim3(kThreads), 0, s, wstamp, 4 * d.cells, kNoLine);
        CHECK_HIP(hipGetLastError());
        ctx->line_stamps_cells = d.cells;
    }
    ctx->line_stamps_clear = false;  // until every window has cleared its cells

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
    // footprint pool: 256 cells per line of a window on average, at least 2^22
    // words (a speculative trace runs until committed or its own labels stop
    // it, so early in a volume it runs to the extremum)
    const uint32_t pool_cap = (uint32_t)std::min<uint64_t>(
        std::max<uint64_t>((uint64_t)W * 256, 1u << 22) / kChunk * kChunk, 0x7FFFFFFFull / kChunk * kChunk);

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
    const size_t o_pool = off; off += al(4 * (size_t)pool_cap);
    const size_t o_touched = off; off += al(4 * (size_t)pool_cap);
    const size_t o_fhead = off; off += al((size_t)W * 4);
    const size_t o_whead = off; off += al((size_t)W * 4);
    const size_t o_mhead = off; off += al((size_t)W * 4);
    const size_t o_parent = off; off += al((size_t)W * 4);
    const size_t o_dep = off; off += al((size_t)W);
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
    uint8_t* dep = (uint8_t*)(base + o_dep);
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
    a.pstamp = pstamp;
    a.lstamp = lstamp;
    a.own = own;
    a.pool = Pool{(uint32_t*)(base + o_pool), pool_cap, lctl + kLcPoolTop};
    a.fhead = (uint32_t*)(base + o_fhead);
    a.whead = (uint32_t*)(base + o_whead);
    a.mhead = (uint32_t*)(base + o_mhead);
    a.touched = (uint32_t*)(base + o_touched);
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
            // the first line's footprint did not fit the pool: it alone, in order
            hipLaunchKernelGGL(line_clear_kernel, dim3(256), dim3(kThreads), 0, s, a);
            hipLaunchKernelGGL(line_direct_kernel, dim3(1), dim3(64), 0, s, a, a0);
            CHECK_HIP(hipGetLastError());
            out->direct++;
            w0 = a0 + 1;
            continue;
        }
        const uint32_t wc = h[kLcOverflow] < w1 ? h[kLcOverflow] : w1;  // lines of this window
        if (wc < w1) out->overflows++;
        const uint32_t nc = wc - a0;
        hipLaunchKernelGGL(line_init_parent_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, parent, nc);
        hipLaunchKernelGGL(line_union_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, a, a0, wc, parent, dep);
        hipLaunchKernelGGL(line_keys_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, parent, nc, wkeys);
        CHECK_HIP(hipGetLastError());
        tb = tmp_bytes;
        CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(tmp, tb, wkeys, wsorted, (int)nc, 0, 64, s));
        hipLaunchKernelGGL(line_starts_kernel, dim3(std::min<uint32_t>(blocks_for(nc), 1024)), dim3(kThreads), 0, s,
                           wsorted, nc, starts, lctl + kLcComps);
        hipLaunchKernelGGL(line_replay_kernel, dim3(blocks_for(nc)), dim3(kThreads), 0, s, a, a0, wc, wsorted, starts,
                           parent, dep);
        CHECK_HIP(hipGetLastError());
        CHECK_HIP(hipMemcpyAsync(h, lctl, sizeof(h), hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipStreamSynchronize(s));
        const uint32_t cut = h[kLcFail] < wc ? h[kLcFail] : wc;
        if (cut == a0) {
            // (the first line's speculation is exact, so its replay is too;
            // should it fail all the same, it is traced alone, in order)
            hipLaunchKernelGGL(line_clear_kernel, dim3(256), dim3(kThreads), 0, s, a);
            hipLaunchKernelGGL(line_direct_kernel, dim3(1), dim3(64), 0, s, a, a0);
            CHECK_HIP(hipGetLastError());
            out->direct++;
            w0 = a0 + 1;
            continue;
        }
        const uint32_t tb_blocks = std::min<uint32_t>(blocks_for(h[kLcTouched]), 4096);
        hipLaunchKernelGGL(line_commit_kernel, dim3(tb_blocks), dim3(kThreads), 0, s, a, cut);
        hipLaunchKernelGGL(line_clear_kernel, dim3(tb_blocks), dim3(kThreads), 0, s, a);
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
    out->fail_tests = h[kLcWhyTest];
    out->fail_labels = h[kLcWhyLabel];
    return MSCEER_OK;
}

}  // namespace msceer

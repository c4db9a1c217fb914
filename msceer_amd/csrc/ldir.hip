// Ascending-manifold dimension (ldir bits) of a discrete gradient.
//
// Restates TopologicalGradientUsingAlgorithms::setAscendingManifoldDimensions
// and rec_man_trace_up_marking (gi_topological_gradient_using_algorithms.h:395-478):
// pass 1 starts from every critical 1-cell with value 2, pass 2 from every
// critical 2-cell with value 1.  Processing an item: if its ldir already
// equals the value, stop; else set it, and for every cofacet t that is
// assigned, not critical, not paired with the item and whose ldir differs
// from the value, set ldir(t) and continue from pair(t).  The reached set is
// independent of the visiting order, so the GPU walks it with many lanes at
// once, from an LDS pool per work-group; what does not finish in a round is
// handed to the next frontier.  A lane marks a cofacet with a plain byte
// store when it reads it unmarked, without an atomic claim: two lanes that
// read the same cofacet unmarked both continue from its pair, which only
// repeats work, because every writer in a pass writes the same ldir and
// leaves the other bits alone, and the walk follows V-paths backwards, which
// are acyclic, so every repeated walk ends too.  (Device-scope atomics are
// performed at the memory, not in the L2; dropping them took the 512^3 passes
// from 14.2 to 12.0 ms.)
//
// The lower-star kernel already preset ldir to 3 (0 for critical 3-cells,
// which both passes skip), so the passes touch only the reached cells.
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "mlist.hpp"

namespace msceer {

namespace {

constexpr int kThreads = 256;
constexpr int kPool = 4096;         // per-block LDS pool of pending items
constexpr int kStage = 4096;        // per-block LDS staging of outgoing items (sweep)
constexpr int kMaxSteps = 96;       // block steps per round before the rest is handed on
constexpr int kPerLane = 2;         // items a lane works on per step
constexpr int kRoundBlocks = 2048;  // fixed grid of a round

struct CellGeom {
    uint32_t Mx, My, Mz;
    uint32_t sxy;    // Mx*My
    uint32_t cells;  // Mx*My*Mz (< 2^32)
};

// The cell count (2X-1)(2Y-1)(2Z-1) is odd, so the last 32-bit word of the
// byte array would reach past its end; its bytes are claimed in a shadow
// word instead (tail_start = cells & ~3) and copied back after the passes.
struct GradView {
    uint8_t* grad;
    uint32_t tail_start;
    uint32_t* tail;
};

__device__ __forceinline__ uint32_t ldir_of(const GradView& v, uint32_t c) {
    if (c >= v.tail_start) return (*v.tail >> ((c - v.tail_start) * 8u + 5u)) & 7u;
    return (uint32_t)v.grad[c] >> 5;
}

__device__ __forceinline__ void cell_coords(const CellGeom& g, uint32_t c, uint32_t& x, uint32_t& y, uint32_t& z) {
    z = c / g.sxy;
    const uint32_t r = c - z * g.sxy;
    y = r / g.Mx;
    x = r - y * g.Mx;
}

__device__ __forceinline__ int64_t dir_offset(const CellGeom& g, uint32_t dir) {
    const int64_t a = dir < 2 ? 1 : (dir < 4 ? (int64_t)g.Mx : (int64_t)g.sxy);
    return (dir & 1u) ? -a : a;
}

// Set the ldir field of byte c to value unless it already holds it; true if
// this call set it.  Pass 1 (value 2) only ever moves ldir 3 -> 2, one
// cleared bit: a single atomic AND decides.  Pass 2 (value 1) moves 3 or 2
// -> 1 and takes a compare-and-swap.
__device__ __forceinline__ bool claim(const GradView& v, uint32_t c, uint32_t value) {
    const bool in_tail = c >= v.tail_start;
    uint32_t* w = in_tail ? v.tail : reinterpret_cast<uint32_t*>(v.grad) + (c >> 2);
    const uint32_t sh = (in_tail ? (c - v.tail_start) : (c & 3u)) * 8u + 5u;
    if (value == 2u) {
        const uint32_t old = atomicAnd(w, ~(1u << sh));
        return ((old >> sh) & 7u) == 3u;
    }
    uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        if (((old >> sh) & 7u) == value) return false;
        const uint32_t nw = (old & ~(7u << sh)) | (value << sh);
        const uint32_t prev = atomicCAS(w, old, nw);
        if (prev == old) return true;
        old = prev;
    }
}

// Set the ldir of an item (3 -> 2 in pass 1; 3 or 2 -> 1 in pass 2) without
// waiting for the result.
__device__ __forceinline__ void set_ldir(const GradView& v, uint32_t c, uint32_t value) {
    const bool in_tail = c >= v.tail_start;
    uint32_t* w = in_tail ? v.tail : reinterpret_cast<uint32_t*>(v.grad) + (c >> 2);
    const uint32_t sh = (in_tail ? (c - v.tail_start) : (c & 3u)) * 8u + 5u;
    if (value == 2u) {
        atomicAnd(w, ~(1u << sh));
    } else {
        atomicAnd(w, ~(2u << sh));
        atomicOr(w, 1u << sh);
    }
}

// Outgoing items are staged per block in LDS and flushed with one
// reservation; when the stage is full they go straight to the sub-list.
struct Outbox {
    uint32_t* stage;  // LDS, kStage
    uint32_t* n;      // LDS counter
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

// One round.  A block takes an even share of the incoming frontier into an
// LDS pool and works it breadth-first: each step, every lane pops one item,
// and the items found behind it go back into the pool, so a large manifold
// reached from one start item is spread over the whole block.  After
// kMaxSteps steps whatever is left goes to the outgoing frontier, which the
// next round shares out again.  Counts are read on the device, so rounds are
// enqueued back to back; `clear` (the counts the next round appends to) is
// zeroed here.
//
// A non-start item is the partner of a mark; it is pushed once per lane that
// read the mark's byte unmarked (usually one), and start items (critical
// cells) appear once in the start list.  Only the tail word's bytes, which
// live in the shadow word, are claimed atomically.
__global__ void __launch_bounds__(kThreads)
ldir_round_kernel(GradView v, CellGeom g, MList in, MList out, uint32_t* __restrict__ clear,
                  uint32_t* __restrict__ overflow, uint32_t value) {
    __shared__ uint32_t pool[kPool];
    __shared__ uint32_t prefix[kSubs + 2];
    __shared__ uint32_t s_n, s_next;
    __shared__ MSlot s_base;
    const uint32_t tid = threadIdx.x;
    if (blockIdx.x == 0)
        for (uint32_t j = tid; j < (uint32_t)kCountWords; j += kThreads) clear[j] = 0u;
    const uint32_t total = mlist_prefix(in, prefix);
    const uint32_t chunk = (total + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = min(total, blockIdx.x * chunk), hi = min(total, lo + chunk);
    const int sub = (int)(blockIdx.x % kSubs);
    if (tid == 0) { s_n = 0; s_next = lo; }
    __syncthreads();
    for (int step = 0; step < kMaxSteps; step++) {
        uint32_t n = min(s_n, (uint32_t)kPool);
        const uint32_t nx = s_next;
        if (n < (uint32_t)(kPerLane * kThreads) && nx < hi) {
            const uint32_t take = min(hi - nx, min((uint32_t)kPool - n, 4u * kThreads));
            for (uint32_t j = tid; j < take; j += kThreads) pool[n + j] = mlist_at(in, prefix, nx + j);
            __syncthreads();
            if (tid == 0) { s_n = n + take; s_next = nx + take; }
            __syncthreads();
            n += take;
        }
        if (n == 0) break;
        // every lane takes up to kPerLane items and keeps their cofacet
        // loads in flight together
        const uint32_t m = min(n, (uint32_t)(kPerLane * kThreads));
        uint32_t items[kPerLane];
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
            const uint32_t j = tid + (uint32_t)k * kThreads;
            items[k] = j < m ? pool[n - 1 - j] : 0xFFFFFFFFu;
        }
        __syncthreads();
        if (tid == 0) s_n = n - m;
        __syncthreads();
        uint32_t tc[kPerLane][4], tb[kPerLane][4];
        int nt[kPerLane];
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
            nt[k] = 0;
            const uint32_t item = items[k];
            if (item == 0xFFFFFFFFu) continue;
            uint32_t x, y, z;
            cell_coords(g, item, x, y, z);
            const uint32_t pos[3] = {x, y, z};
            const uint32_t ext[3] = {g.Mx, g.My, g.Mz};
            const int64_t step_off[3] = {1, (int64_t)g.Mx, (int64_t)g.sxy};
            // the item itself: its ldir is set below by a byte store (the
            // tail word's bytes live in the shadow word and take atomics)
            if (item >= v.tail_start) set_ldir(v, item, value);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (pos[a] & 1u) continue;  // cofacets lie along even axes
#pragma unroll
                for (int sgn = 0; sgn < 2; sgn++) {
                    if (sgn == 0 && pos[a] == ext[a] - 1) continue;
                    if (sgn == 1 && pos[a] == 0) continue;
                    if (nt[k] < 4) tc[k][nt[k]++] = (uint32_t)((int64_t)item + (sgn ? -step_off[a] : step_off[a]));
                }
            }
        }
        uint32_t ib[kPerLane];
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
            ib[k] = items[k] < v.tail_start ? v.grad[items[k]] : 0u;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (j < nt[k]) tb[k][j] = v.grad[tc[k][j]];
        }
        // the items' ldir by plain byte stores (see the top of the file)
#pragma unroll
        for (int k = 0; k < kPerLane; k++)
            if (items[k] < v.tail_start) v.grad[items[k]] = (uint8_t)((ib[k] & 0x1Fu) | (value << 5));
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
            const uint32_t item = items[k];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (j >= nt[k]) break;
                const uint32_t t = tc[k][j];
                const uint32_t b = tb[k][j];
                const uint32_t dir = (b >> 2) & 7u;
                if (dir == 7u || !(b & 1u)) continue;
                if (t < v.tail_start && (b >> 5) == value) continue;  // already marked
                const int64_t p64 = (int64_t)t + dir_offset(g, dir);
                if (p64 < 0 || p64 >= (int64_t)g.cells) continue;  // only a corrupt byte points out
                const uint32_t p = (uint32_t)p64;
                if (p == item) continue;
                if (t >= v.tail_start) {
                    if (!claim(v, t, value)) continue;
                } else {
                    v.grad[t] = (uint8_t)((b & 0x1Fu) | (value << 5));
                }
                const uint32_t slot = atomicAdd(&s_n, 1u);
                if (slot < (uint32_t)kPool) {
                    pool[slot] = p;
                } else {
                    if (!mlist_push(out, sub, p)) atomicOr(overflow, 1u);
                }
            }
        }
        __syncthreads();
    }
    // hand on what is left: the pool, then the unread part of the share
    __syncthreads();
    const uint32_t n = min(s_n, (uint32_t)kPool);
    mlist_flush_block(out, sub, pool, n, &s_base, overflow);
    for (uint32_t nx = s_next; nx < hi; nx += kPool) {
        const uint32_t take = min(hi - nx, (uint32_t)kPool);
        __syncthreads();
        for (uint32_t j = tid; j < take; j += kThreads) pool[j] = mlist_at(in, prefix, nx + j);
        __syncthreads();
        mlist_flush_block(out, sub, pool, take, &s_base, overflow);
    }
}

// Frontier from the bytes alone: unclaimed critical item-dimension cells, and
// the partner of every claimed mark whose partner is not yet claimed.
__global__ void __launch_bounds__(kThreads)
ldir_sweep_kernel(GradView v, CellGeom g, uint64_t cells, uint32_t item_dim, uint32_t value, MList out,
                  uint32_t* __restrict__ overflow) {
    __shared__ uint32_t stage[kStage];
    __shared__ uint32_t s_n;
    __shared__ MSlot s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const Outbox box{stage, &s_n, out, (int)(blockIdx.x % kSubs), overflow};
    const uint32_t mark_dim = item_dim + 1;
    for (uint64_t c = (uint64_t)blockIdx.x * kThreads + threadIdx.x; c < cells;
         c += (uint64_t)gridDim.x * kThreads) {
        const uint32_t b = v.grad[c];
        uint32_t x, y, z;
        cell_coords(g, (uint32_t)c, x, y, z);
        const uint32_t dim = (x & 1u) + (y & 1u) + (z & 1u);
        const uint32_t dir = (b >> 2) & 7u;
        const uint32_t ldir = ldir_of(v, (uint32_t)c);
        if (dim == item_dim && dir == 7u && ldir != value) {
            box.push((uint32_t)c);
        } else if (dim == mark_dim && dir != 7u && (b & 1u) && ldir == value) {
            const int64_t p64 = (int64_t)c + dir_offset(g, dir);
            if (p64 < 0 || p64 >= (int64_t)g.cells) continue;
            const uint32_t p = (uint32_t)p64;
            if (ldir_of(v, p) != value) box.push(p);
        }
    }
    __syncthreads();
    const uint32_t n = s_n < (uint32_t)kStage ? s_n : (uint32_t)kStage;
    mlist_flush_block(out, box.sub, stage, n, &s_base, overflow);
}

__global__ void zero_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// Shadow of the tail bytes: packed before the passes, unpacked after.
__global__ void ldir_tail_kernel(uint8_t* grad, uint32_t tail_start, uint32_t cells, uint32_t* tail,
                                 int unpack) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (!unpack) {
        uint32_t w = 0;
        for (uint32_t c = tail_start; c < cells; c++) w |= (uint32_t)grad[c] << ((c - tail_start) * 8u);
        *tail = w;
    } else {
        const uint32_t w = *tail;
        for (uint32_t c = tail_start; c < cells; c++) grad[c] = (uint8_t)(w >> ((c - tail_start) * 8u));
    }
}

GradView make_view(const Dims& d, uint8_t* grad, uint32_t* tail) {
    GradView v;
    v.grad = grad;
    v.tail_start = (uint32_t)(d.cells & ~3ull);
    v.tail = tail;
    return v;
}

}  // namespace

hipError_t launch_ldir_tail(const Dims& d, uint8_t* grad, uint32_t* tail, bool unpack, hipStream_t s) {
    const uint32_t tail_start = (uint32_t)(d.cells & ~3ull);
    if (tail_start == d.cells) return hipSuccess;
    hipLaunchKernelGGL(ldir_tail_kernel, dim3(1), dim3(64), 0, s, grad, tail_start, (uint32_t)d.cells,
                       tail, unpack ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_ldir_round(const Dims& d, uint8_t* grad, uint32_t* tail, const MList& in, const MList& out,
                             uint32_t* clear_counts, uint32_t* overflow, uint32_t value, hipStream_t s) {
    CellGeom g{d.Mx, d.My, d.Mz, d.Mx * d.My, (uint32_t)d.cells};
    hipLaunchKernelGGL(ldir_round_kernel, dim3(kRoundBlocks), dim3(kThreads), 0, s, make_view(d, grad, tail),
                       g, in, out, clear_counts, overflow, value);
    return hipGetLastError();
}

hipError_t launch_ldir_sweep(const Dims& d, uint8_t* grad, uint32_t* tail, uint32_t item_dim,
                             uint32_t value, const MList& out, uint32_t* overflow, hipStream_t s) {
    CellGeom g{d.Mx, d.My, d.Mz, d.Mx * d.My, (uint32_t)d.cells};
    uint64_t blocks = (d.cells + kThreads - 1) / kThreads;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(ldir_sweep_kernel, dim3((uint32_t)blocks), dim3(kThreads), 0, s,
                       make_view(d, grad, tail), g, d.cells, item_dim, value, out, overflow);
    return hipGetLastError();
}

hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    uint32_t blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(zero_kernel, dim3(blocks), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

}  // namespace msceer

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
// once: a lane claims a byte by an atomic compare-and-swap of the ldir field
// in the 32-bit word holding it, and follows its chain depth-first with a
// small LDS stack; chains that do not fit are handed to the next frontier.
//
// The lower-star kernel already preset ldir to 3 (0 for critical 3-cells,
// which both passes skip), so the passes touch only the reached cells.
#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace msceer {

namespace {

constexpr int kThreads = 256;
constexpr int kStack = 24;        // per-lane depth-first stack (LDS)
constexpr int kBudget = 1 << 12;  // items per lane before the rest is handed on

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

// Set the ldir field of byte c to value unless it already holds it.
__device__ __forceinline__ bool claim(const GradView& v, uint32_t c, uint32_t value) {
    const bool in_tail = c >= v.tail_start;
    uint32_t* w = in_tail ? v.tail : reinterpret_cast<uint32_t*>(v.grad) + (c >> 2);
    const uint32_t sh = (in_tail ? (c - v.tail_start) : (c & 3u)) * 8u + 5u;
    uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        if (((old >> sh) & 7u) == value) return false;
        const uint32_t nw = (old & ~(7u << sh)) | (value << sh);
        const uint32_t prev = atomicCAS(w, old, nw);
        if (prev == old) return true;
        old = prev;
    }
}

__device__ __forceinline__ void push_out(uint32_t* out, uint32_t cap, uint32_t* counters, uint32_t item) {
    const uint32_t slot = atomicAdd(&counters[kFrontCount], 1u);
    if (slot < cap) {
        out[slot] = item;
    } else {
        atomicOr(&counters[kFrontOverflow], 1u);
    }
}

__global__ void __launch_bounds__(kThreads)
ldir_expand_kernel(GradView v, CellGeom g, const uint32_t* __restrict__ in,
                   uint32_t n_in, uint32_t* __restrict__ out, uint32_t cap,
                   uint32_t* __restrict__ counters, uint32_t value) {
    __shared__ uint32_t stack_mem[kThreads * kStack];
    uint32_t* stack = stack_mem + threadIdx.x;  // strided: entry k at stack[k*kThreads]
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n_in) return;
    int sp = 0;
    stack[0] = in[i];
    sp = 1;
    int work = 0;
    while (sp > 0) {
        const uint32_t item = stack[(--sp) * kThreads];
        if (work >= kBudget) {
            push_out(out, cap, counters, item);
            continue;
        }
        work++;
        if (!claim(v, item, value)) continue;
        uint32_t x, y, z;
        cell_coords(g, item, x, y, z);
        const uint32_t pos[3] = {x, y, z};
        const uint32_t ext[3] = {g.Mx, g.My, g.Mz};
        const int64_t step[3] = {1, (int64_t)g.Mx, (int64_t)g.sxy};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (pos[a] & 1u) continue;  // cofacets lie along even axes
#pragma unroll
            for (int sgn = 0; sgn < 2; sgn++) {
                if (sgn == 0 && pos[a] == ext[a] - 1) continue;
                if (sgn == 1 && pos[a] == 0) continue;
                const uint32_t t = (uint32_t)((int64_t)item + (sgn ? -step[a] : step[a]));
                const uint32_t b = v.grad[t];
                const uint32_t dir = (b >> 2) & 7u;
                if (dir == 7u || !(b & 1u)) continue;
                const int64_t p64 = (int64_t)t + dir_offset(g, dir);
                if (p64 < 0 || p64 >= (int64_t)g.cells) continue;  // only a corrupt byte points out
                const uint32_t p = (uint32_t)p64;
                if (p == item) continue;
                if (!claim(v, t, value)) continue;
                if (sp < kStack) {
                    stack[(sp++) * kThreads] = p;
                } else {
                    push_out(out, cap, counters, p);
                }
            }
        }
    }
}

// Frontier from the bytes alone: unclaimed critical item-dimension cells, and
// the partner of every claimed mark whose partner is not yet claimed.
__global__ void __launch_bounds__(kThreads)
ldir_sweep_kernel(GradView v, CellGeom g, uint64_t cells, uint32_t item_dim,
                  uint32_t value, uint32_t* __restrict__ out, uint32_t cap,
                  uint32_t* __restrict__ counters) {
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
            push_out(out, cap, counters, (uint32_t)c);
        } else if (dim == mark_dim && dir != 7u && (b & 1u) && ldir == value) {
            const int64_t p64 = (int64_t)c + dir_offset(g, dir);
            if (p64 < 0 || p64 >= (int64_t)g.cells) continue;
            const uint32_t p = (uint32_t)p64;
            if (ldir_of(v, p) != value) push_out(out, cap, counters, p);
        }
    }
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

hipError_t launch_ldir_expand(const Dims& d, uint8_t* grad, uint32_t* tail, const uint32_t* in,
                              uint32_t n_in, uint32_t* out, uint32_t out_cap, uint32_t* counters,
                              uint32_t value, hipStream_t s) {
    if (n_in == 0) return hipSuccess;
    CellGeom g{d.Mx, d.My, d.Mz, d.Mx * d.My, (uint32_t)d.cells};
    const uint32_t blocks = (n_in + kThreads - 1) / kThreads;
    hipLaunchKernelGGL(ldir_expand_kernel, dim3(blocks), dim3(kThreads), 0, s, make_view(d, grad, tail),
                       g, in, n_in, out, out_cap, counters, value);
    return hipGetLastError();
}

hipError_t launch_ldir_sweep(const Dims& d, uint8_t* grad, uint32_t* tail, uint32_t item_dim,
                             uint32_t value, uint32_t* out, uint32_t out_cap, uint32_t* counters,
                             hipStream_t s) {
    CellGeom g{d.Mx, d.My, d.Mz, d.Mx * d.My, (uint32_t)d.cells};
    uint64_t blocks = (d.cells + kThreads - 1) / kThreads;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(ldir_sweep_kernel, dim3((uint32_t)blocks), dim3(kThreads), 0, s,
                       make_view(d, grad, tail), g, d.cells, item_dim, value, out, out_cap, counters);
    return hipGetLastError();
}

}  // namespace msceer

// Ascending-manifold dimension (ldir bits) of a discrete gradient.
//
// Restates TopologicalGradientUsingAlgorithms::setAscendingManifoldDimensions
// and rec_man_trace_up_marking (gi_topological_gradient_using_algorithms.h:395-478):
// ldir = 3 everywhere; pass 1 walks from every critical 1-cell with value 2,
// pass 2 from every critical 2-cell with value 1; critical 3-cells get 0.
// Walking an item: stop if its ldir already equals the value, else set it, and
// for every cofacet t that is assigned, not critical, not paired with the item
// and whose ldir differs from the value, set ldir(t) and walk pair(t).
//
// In pass 1 the items are edges (and the hexes some quads pair up with, which
// have no cofacets) and the marked cofacets are quads; in pass 2 the items are
// quads and the marked cofacets hexes.  So the marking has a closed form in
// terms of the items each pass reaches (visited1: edges, visited2: quads):
//   edge e: 2 if visited1(e), else 3;
//   quad t: 1 if visited2(t); else 2 if t is assigned, not critical and has a
//           facet edge e != pair(t) with visited1(e) ("marked1"); else 3;
//   hex h:  0 if critical; else 1 if h is assigned and has a facet quad
//           q != pair(h) with visited2(q); else 2 if pair(h) is a quad t with
//           pair(t) = h and marked1(t); else 3;
//   vertex: 3.
// Reachability does not depend on the visiting order.
//
// Three phases on the GPU:
//  1. ldir_records_kernel, one streaming pass over the .grad: for every edge
//     the pair directions of its cofacet quads that lead on to another edge
//     (rec1, 4 bits per cofacet slot), for every quad those of its cofacet
//     hexes that lead on to another quad (rec2).  A walk step then reads one
//     record instead of four scattered .grad bytes.
//  2. Frontier rounds (ldir_round_kernel) walk pass 1 over rec1 and pass 2
//     over rec2; a reached item is claimed in a bitmap (visited1 / visited2,
//     one bit per record, atomic OR), so each item is expanded once.  The
//     .grad bytes are not written during the walks.
//  3. ldir_finish_kernel writes every cell's ldir from the closed form above,
//     one streaming pass over the .grad (whole rows per wave).
//
// Records and bits are indexed by 3 * (vertex id of the cell's lowest corner)
// + axis, axis = the odd axis of an edge or the even axis of a quad.
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

struct LGeom {
    uint32_t Mx, My, Mz;  // cell grid
    uint32_t sxy;         // Mx * My
    uint32_t X, Y;        // vertex grid (record indices)
    uint32_t V;           // vertices
};

__device__ __forceinline__ void cell_coords(const LGeom& g, uint32_t c, uint32_t& x, uint32_t& y, uint32_t& z) {
    z = c / g.sxy;
    const uint32_t r = c - z * g.sxy;
    y = r / g.Mx;
    x = r - y * g.Mx;
}

__device__ __forceinline__ uint32_t cell_id(const LGeom& g, uint32_t x, uint32_t y, uint32_t z) {
    return x + g.Mx * y + g.sxy * z;
}

// Record / bit index of the edge or quad at cell coordinates (x, y, z).
__device__ __forceinline__ uint32_t rec_index(const LGeom& g, uint32_t x, uint32_t y, uint32_t z, uint32_t axis) {
    return 3u * ((x >> 1) + g.X * ((y >> 1) + g.Y * (z >> 1))) + axis;
}

__device__ __forceinline__ bool test_bit(const uint32_t* bits, uint32_t i) { return (bits[i >> 5] >> (i & 31u)) & 1u; }

// Claim bit i; true if this call set it.
__device__ __forceinline__ bool claim_bit(uint32_t* bits, uint32_t i) {
    const uint32_t m = 1u << (i & 31u);
    if (__hip_atomic_load(&bits[i >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & m) return false;
    return !(atomicOr(&bits[i >> 5], m) & m);
}

// Pair direction code d (0..5 = +x,-x,+y,-y,+z,-z) applied to coordinates.
__device__ __forceinline__ void apply_dir(uint32_t d, uint32_t& x, uint32_t& y, uint32_t& z) {
    const uint32_t a = d >> 1;
    const uint32_t s = (d & 1u) ? 0xFFFFFFFFu : 1u;  // -1 or +1 (wrapping add)
    if (a == 0) x += s;
    else if (a == 1) y += s;
    else z += s;
}

__device__ __forceinline__ void axis_step(uint32_t a, int sgn, uint32_t& x, uint32_t& y, uint32_t& z) {
    const uint32_t s = sgn > 0 ? 1u : 0xFFFFFFFFu;
    if (a == 0) x += s;
    else if (a == 1) y += s;
    else z += s;
}

__device__ __forceinline__ uint32_t coord(uint32_t a, uint32_t x, uint32_t y, uint32_t z) {
    return a == 0 ? x : (a == 1 ? y : z);
}
__device__ __forceinline__ uint32_t extent(const LGeom& g, uint32_t a) { return a == 0 ? g.Mx : (a == 1 ? g.My : g.Mz); }

// The two axes other than a, in increasing order.
__device__ __forceinline__ void other_axes(uint32_t a, uint32_t& a1, uint32_t& a2) {
    a1 = a == 0 ? 1u : 0u;
    a2 = a == 2 ? 1u : 2u;
}

// Does pair direction d of a cell reached from `from` by one step (sign sgn
// along axis a) point straight back to `from`?
__device__ __forceinline__ bool points_back(uint32_t d, uint32_t a, int sgn) {
    return (d >> 1) == a && ((d & 1u) != 0u) == (sgn > 0);
}

// ---- phase 1: records ------------------------------------------------------

// One thread per vertex (i, j, k): its three "+" edges and three "+" quads.
// rec1 slot 2*m + (sgn < 0) is the cofacet of the edge along its m-th even
// axis; rec2 slot (sgn < 0) the cofacet hex of the quad along its even axis.
// A slot holds 8 | pair direction of the cofacet when the walk goes on
// through it, else 0.
__global__ void __launch_bounds__(kThreads)
ldir_records_kernel(const uint8_t* __restrict__ grad, LGeom g, uint16_t* __restrict__ rec1,
                    uint8_t* __restrict__ rec2, uint32_t* __restrict__ bits, uint32_t nbits_words) {
    for (uint32_t w = blockIdx.x * kThreads + threadIdx.x; w < nbits_words; w += gridDim.x * kThreads) bits[w] = 0u;
    for (uint32_t v = blockIdx.x * kThreads + threadIdx.x; v < g.V; v += gridDim.x * kThreads) {
        const uint32_t q = v / g.X, k = q / g.Y;
        const uint32_t i = v - q * g.X, j = q - k * g.Y;
        const uint32_t vx = 2 * i, vy = 2 * j, vz = 2 * k;
        // edges: odd axis o
        for (uint32_t o = 0; o < 3; o++) {
            uint32_t ex = vx, ey = vy, ez = vz;
            axis_step(o, 1, ex, ey, ez);
            if (coord(o, ex, ey, ez) >= extent(g, o)) continue;
            uint32_t a1, a2;
            other_axes(o, a1, a2);
            uint32_t r = 0;
#pragma unroll
            for (int slot = 0; slot < 4; slot++) {
                const uint32_t a = slot < 2 ? a1 : a2;
                const int sgn = (slot & 1) ? -1 : 1;
                const uint32_t ca = coord(a, ex, ey, ez);
                if (sgn > 0 ? ca + 1 >= extent(g, a) : ca == 0) continue;
                uint32_t tx = ex, ty = ey, tz = ez;
                axis_step(a, sgn, tx, ty, tz);
                const uint32_t b = grad[cell_id(g, tx, ty, tz)];
                const uint32_t d = (b >> 2) & 7u;
                // on to an edge: the pair is a facet of t (along o or a), not the item
                if (!(b & 1u) || d == 7u) continue;
                if ((d >> 1) != o && (d >> 1) != a) continue;
                if (points_back(d, a, sgn)) continue;
                r |= (8u | d) << (4 * slot);
            }
            rec1[rec_index(g, ex, ey, ez, o)] = (uint16_t)r;
        }
        // quads: even axis a
        for (uint32_t a = 0; a < 3; a++) {
            uint32_t qx = vx, qy = vy, qz = vz;
            uint32_t a1, a2;
            other_axes(a, a1, a2);
            axis_step(a1, 1, qx, qy, qz);
            axis_step(a2, 1, qx, qy, qz);
            if (coord(a1, qx, qy, qz) >= extent(g, a1) || coord(a2, qx, qy, qz) >= extent(g, a2)) continue;
            uint32_t r = 0;
#pragma unroll
            for (int slot = 0; slot < 2; slot++) {
                const int sgn = slot ? -1 : 1;
                const uint32_t ca = coord(a, qx, qy, qz);
                if (sgn > 0 ? ca + 1 >= extent(g, a) : ca == 0) continue;
                uint32_t hx = qx, hy = qy, hz = qz;
                axis_step(a, sgn, hx, hy, hz);
                const uint32_t b = grad[cell_id(g, hx, hy, hz)];
                const uint32_t d = (b >> 2) & 7u;
                if (!(b & 1u) || d == 7u) continue;
                if (points_back(d, a, sgn)) continue;
                r |= (8u | d) << (4 * slot);
            }
            rec2[rec_index(g, qx, qy, qz, a)] = (uint8_t)r;
        }
    }
}

// ---- phase 2: walks ----------------------------------------------------------

struct Walk {
    const uint16_t* rec1;
    const uint8_t* rec2;
    uint32_t* bits;  // visited1 (pass 1) or visited2 (pass 2)
    int pass;        // 1: edges over rec1, 2: quads over rec2
};

// The items reached in one step from `item` (cell id): up to 4, as cell ids
// whose bits this call claimed.  `start`: the item is a critical cell of the
// start list, claimed here first (skipped if already reached).
__device__ __forceinline__ int expand(const LGeom& g, const Walk& w, uint32_t item, bool start, uint32_t out[4]) {
    uint32_t x, y, z;
    cell_coords(g, item, x, y, z);
    int n = 0;
    if (w.pass == 1) {
        const uint32_t o = (x & 1u) ? 0u : ((y & 1u) ? 1u : 2u);
        const uint32_t ri = rec_index(g, x, y, z, o);
        if (start && !claim_bit(w.bits, ri)) return 0;
        const uint32_t r = w.rec1[ri];
        if (!r) return 0;
        uint32_t a1, a2;
        other_axes(o, a1, a2);
#pragma unroll
        for (int slot = 0; slot < 4; slot++) {
            const uint32_t f = (r >> (4 * slot)) & 15u;
            if (!(f & 8u)) continue;
            uint32_t px = x, py = y, pz = z;
            axis_step(slot < 2 ? a1 : a2, (slot & 1) ? -1 : 1, px, py, pz);
            apply_dir(f & 7u, px, py, pz);
            // p is an edge: its odd axis
            const uint32_t po = (px & 1u) ? 0u : ((py & 1u) ? 1u : 2u);
            if (claim_bit(w.bits, rec_index(g, px, py, pz, po))) out[n++] = cell_id(g, px, py, pz);
        }
    } else {
        const uint32_t a = !(x & 1u) ? 0u : (!(y & 1u) ? 1u : 2u);
        const uint32_t ri = rec_index(g, x, y, z, a);
        if (start && !claim_bit(w.bits, ri)) return 0;
        const uint32_t r = w.rec2[ri];
        if (!r) return 0;
#pragma unroll
        for (int slot = 0; slot < 2; slot++) {
            const uint32_t f = (r >> (4 * slot)) & 15u;
            if (!(f & 8u)) continue;
            uint32_t px = x, py = y, pz = z;
            axis_step(a, slot ? -1 : 1, px, py, pz);
            apply_dir(f & 7u, px, py, pz);
            const uint32_t pa = !(px & 1u) ? 0u : (!(py & 1u) ? 1u : 2u);
            if (claim_bit(w.bits, rec_index(g, px, py, pz, pa))) out[n++] = cell_id(g, px, py, pz);
        }
    }
    return n;
}

// One round.  A block takes an even share of the incoming frontier into an
// LDS pool and works it breadth-first: each step, every lane pops up to
// kPerLane items, and the items found behind them go back into the pool, so
// a large manifold reached from one start item is spread over the whole
// block.  After kMaxSteps steps whatever is left goes to the outgoing
// frontier, which the next round shares out again.  Counts are read on the
// device, so rounds are enqueued back to back; `clear` (the counts the next
// round appends to) is zeroed here.
__global__ void __launch_bounds__(kThreads)
ldir_round_kernel(LGeom g, Walk w, MList in, MList out, uint32_t* __restrict__ clear, uint32_t* __restrict__ overflow,
                  int start) {
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
    // start items come from the share only (never from the pool, which holds
    // claimed items); they are claimed when expanded
    for (int step = 0; step < kMaxSteps; step++) {
        uint32_t n = min(s_n, (uint32_t)kPool);
        const uint32_t nx = s_next;
        uint32_t items[kPerLane];
        bool from_share[kPerLane];
        uint32_t m = 0;
        if (start) {
            // the pool is worked first; share items fill the rest of the step
            const uint32_t mp = min(n, (uint32_t)(kPerLane * kThreads));
            const uint32_t ms = min(hi - nx, (uint32_t)(kPerLane * kThreads) - mp);
#pragma unroll
            for (int k = 0; k < kPerLane; k++) {
                const uint32_t j = tid + (uint32_t)k * kThreads;
                from_share[k] = j >= mp && j < mp + ms;
                items[k] = j < mp ? pool[n - 1 - j] : (from_share[k] ? mlist_at(in, prefix, nx + (j - mp)) : 0xFFFFFFFFu);
            }
            m = mp + ms;
            __syncthreads();
            if (tid == 0) { s_n = n - mp; s_next = nx + ms; }
            __syncthreads();
        } else {
            if (n < (uint32_t)(kPerLane * kThreads) && nx < hi) {
                const uint32_t take = min(hi - nx, min((uint32_t)kPool - n, 4u * kThreads));
                for (uint32_t j = tid; j < take; j += kThreads) pool[n + j] = mlist_at(in, prefix, nx + j);
                __syncthreads();
                if (tid == 0) { s_n = n + take; s_next = nx + take; }
                __syncthreads();
                n += take;
            }
            m = min(n, (uint32_t)(kPerLane * kThreads));
#pragma unroll
            for (int k = 0; k < kPerLane; k++) {
                const uint32_t j = tid + (uint32_t)k * kThreads;
                items[k] = j < m ? pool[n - 1 - j] : 0xFFFFFFFFu;
                from_share[k] = false;
            }
            __syncthreads();
            if (tid == 0) s_n = n - m;
            __syncthreads();
        }
        if (m == 0) break;
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
            if (items[k] == 0xFFFFFFFFu) continue;
            uint32_t nb[4];
            const int c = expand(g, w, items[k], from_share[k], nb);
            for (int j = 0; j < c; j++) {
                const uint32_t slot = atomicAdd(&s_n, 1u);
                if (slot < (uint32_t)kPool) {
                    pool[slot] = nb[j];
                } else {
                    if (!mlist_push(out, sub, nb[j])) atomicOr(overflow, 1u);
                }
            }
        }
        __syncthreads();
    }
    // hand on what is left: the pool (claimed items), then the unread part of
    // the share (start items stay start items only in the first round, so a
    // start round hands on nothing unread: see launch_ldir_round)
    __syncthreads();
    const uint32_t n = min(s_n, (uint32_t)kPool);
    mlist_flush_block(out, sub, pool, n, &s_base, overflow);
    if (!start) {
        for (uint32_t nx = s_next; nx < hi; nx += kPool) {
            const uint32_t take = min(hi - nx, (uint32_t)kPool);
            __syncthreads();
            for (uint32_t j = tid; j < take; j += kThreads) pool[j] = mlist_at(in, prefix, nx + j);
            __syncthreads();
            mlist_flush_block(out, sub, pool, take, &s_base, overflow);
        }
    } else if (s_next < hi && tid == 0) {
        atomicOr(overflow, 4u);  // start items left unread: the caller sweeps
    }
}

// Frontier from the records and bits alone (after any overflow): the
// unreached critical cells of the pass (claimed here), and every unreached
// cell one step from a reached one (claimed here).
__global__ void __launch_bounds__(kThreads)
ldir_sweep_kernel(const uint8_t* __restrict__ grad, LGeom g, Walk w, MList out, uint32_t* __restrict__ overflow) {
    __shared__ uint32_t stage[kStage];
    __shared__ uint32_t s_n;
    __shared__ MSlot s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int sub = (int)(blockIdx.x % kSubs);
    auto push = [&](uint32_t item) {
        const uint32_t slot = atomicAdd(&s_n, 1u);
        if (slot < (uint32_t)kStage) stage[slot] = item;
        else if (!mlist_push(out, sub, item)) atomicOr(overflow, 1u);
    };
    for (uint32_t v = blockIdx.x * kThreads + threadIdx.x; v < g.V; v += gridDim.x * kThreads) {
        const uint32_t q = v / g.X, k = q / g.Y;
        const uint32_t i = v - q * g.X, j = q - k * g.Y;
        for (uint32_t a = 0; a < 3; a++) {
            uint32_t cx = 2 * i, cy = 2 * j, cz = 2 * k;
            if (w.pass == 1) {
                axis_step(a, 1, cx, cy, cz);  // edge with odd axis a
                if (coord(a, cx, cy, cz) >= extent(g, a)) continue;
            } else {
                uint32_t a1, a2;
                other_axes(a, a1, a2);  // quad with even axis a
                axis_step(a1, 1, cx, cy, cz);
                axis_step(a2, 1, cx, cy, cz);
                if (coord(a1, cx, cy, cz) >= extent(g, a1) || coord(a2, cx, cy, cz) >= extent(g, a2)) continue;
            }
            const uint32_t c = cell_id(g, cx, cy, cz);
            const uint32_t ri = rec_index(g, cx, cy, cz, a);
            if (test_bit(w.bits, ri)) {
                uint32_t nb[4];
                const int n = expand(g, w, c, false, nb);
                for (int t = 0; t < n; t++) push(nb[t]);
            } else if (((grad[c] >> 2) & 7u) == 7u && (grad[c] & 1u)) {
                if (claim_bit(w.bits, ri)) push(c);
            }
        }
    }
    __syncthreads();
    const uint32_t n = s_n < (uint32_t)kStage ? s_n : (uint32_t)kStage;
    mlist_flush_block(out, sub, stage, n, &s_base, overflow);
}

// ---- phase 3: the ldir bits -------------------------------------------------

// marked1 of quad t at (x, y, z) with even axis a and pair direction d (7 if
// critical): a facet edge other than its pair was reached in pass 1.
__device__ __forceinline__ bool marked1(const LGeom& g, const uint32_t* v1, uint32_t x, uint32_t y, uint32_t z,
                                        uint32_t a, uint32_t d) {
    uint32_t u1, u2;
    other_axes(a, u1, u2);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const uint32_t u = k < 2 ? u1 : u2, other = k < 2 ? u2 : u1;
        const int sgn = (k & 1) ? -1 : 1;
        if (d != 7u && (d >> 1) == u && ((d & 1u) != 0u) == (sgn < 0)) continue;  // the pair itself
        uint32_t ex = x, ey = y, ez = z;
        axis_step(u, sgn, ex, ey, ez);
        if (test_bit(v1, rec_index(g, ex, ey, ez, other))) return true;
    }
    return false;
}

__device__ __forceinline__ uint32_t final_ldir(const LGeom& g, const uint8_t* __restrict__ grad, const uint32_t* v1,
                                               const uint32_t* v2, uint32_t x, uint32_t y, uint32_t z, uint32_t b) {
    const uint32_t ox = x & 1u, oy = y & 1u, oz = z & 1u;
    const uint32_t dim = ox + oy + oz;
    const uint32_t d = (b >> 2) & 7u;
    const bool assigned = b & 1u;
    if (dim == 1) {
        const uint32_t o = ox ? 0u : (oy ? 1u : 2u);
        return test_bit(v1, rec_index(g, x, y, z, o)) ? 2u : 3u;
    }
    if (dim == 2) {
        const uint32_t a = !ox ? 0u : (!oy ? 1u : 2u);
        if (test_bit(v2, rec_index(g, x, y, z, a))) return 1u;
        if (assigned && d != 7u && marked1(g, v1, x, y, z, a, d)) return 2u;
        return 3u;
    }
    if (dim == 3) {
        if (d == 7u) return (b >> 5) & 7u;  // critical (0, preset by the lower-star pass), or unassigned
        if (assigned) {
#pragma unroll
            for (uint32_t u = 0; u < 3; u++) {
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const int sgn = s ? -1 : 1;
                    if ((d >> 1) == u && ((d & 1u) != 0u) == (sgn < 0)) continue;  // the pair itself
                    uint32_t qx = x, qy = y, qz = z;
                    axis_step(u, sgn, qx, qy, qz);
                    if (test_bit(v2, rec_index(g, qx, qy, qz, u))) return 1u;
                }
            }
            // walked in pass 1 as the partner of a marked quad
            uint32_t tx = x, ty = y, tz = z;
            apply_dir(d, tx, ty, tz);
            const uint32_t bt = grad[cell_id(g, tx, ty, tz)];
            const uint32_t dt = (bt >> 2) & 7u;
            if ((bt & 1u) && dt == (d ^ 1u) && marked1(g, v1, tx, ty, tz, d >> 1, dt)) return 2u;
        }
        return 3u;
    }
    return 3u;
}

// One thread per vertex: the cells of its "+" octant (vertex bytes are
// rewritten unchanged, so a wave writes whole row segments).
__global__ void __launch_bounds__(kThreads)
ldir_finish_kernel(uint8_t* __restrict__ grad, LGeom g, const uint32_t* __restrict__ v1,
                   const uint32_t* __restrict__ v2) {
    for (uint32_t v = blockIdx.x * kThreads + threadIdx.x; v < g.V; v += gridDim.x * kThreads) {
        const uint32_t q = v / g.X, k = q / g.Y;
        const uint32_t i = v - q * g.X, j = q - k * g.Y;
#pragma unroll
        for (uint32_t o = 0; o < 8; o++) {
            const uint32_t x = 2 * i + (o & 1u), y = 2 * j + ((o >> 1) & 1u), z = 2 * k + (o >> 2);
            if (x >= g.Mx || y >= g.My || z >= g.Mz) continue;
            const uint32_t c = cell_id(g, x, y, z);
            const uint32_t b = grad[c];
            const uint32_t l = o == 0 ? (b >> 5) : final_ldir(g, grad, v1, v2, x, y, z, b);
            grad[c] = (uint8_t)((b & 0x1Fu) | (l << 5));
        }
    }
}

__global__ void zero_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

LGeom make_geom(const Dims& d) {
    return LGeom{d.Mx, d.My, d.Mz, d.Mx * d.My, (uint32_t)d.X, (uint32_t)d.Y, (uint32_t)d.verts};
}

uint32_t grid_of(uint64_t n, uint32_t cap) {
    uint64_t b = (n + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    return (uint32_t)(b > cap ? cap : b);
}

}  // namespace

size_t ldir_scratch_bytes(const Dims& d, LdirScratch* s) {
    const uint64_t n = 3ull * d.verts;
    const uint64_t words = (n + 31) / 32;
    const size_t r1 = (2 * n + 255) & ~(size_t)255, r2 = (n + 255) & ~(size_t)255;
    const size_t bw = ((words * 4) + 255) & ~(size_t)255;
    if (s) {
        s->rec1_off = 0;
        s->rec2_off = r1;
        s->bits1_off = r1 + r2;
        s->bits2_off = r1 + r2 + bw;
        s->bit_words = (uint32_t)words;
    }
    return r1 + r2 + 2 * bw;
}

hipError_t launch_ldir_records(const Dims& d, const uint8_t* grad, void* scratch, const LdirScratch& s,
                               hipStream_t st) {
    uint8_t* base = static_cast<uint8_t*>(scratch);
    uint32_t* bits1 = reinterpret_cast<uint32_t*>(base + s.bits1_off);
    // visited1 and visited2 are adjacent: one clear covers both
    const uint32_t words = (uint32_t)((s.bits2_off - s.bits1_off) / 4) + s.bit_words;
    hipLaunchKernelGGL(ldir_records_kernel, dim3(grid_of(d.verts, 16384)), dim3(kThreads), 0, st, grad, make_geom(d),
                       reinterpret_cast<uint16_t*>(base + s.rec1_off), base + s.rec2_off, bits1, words);
    return hipGetLastError();
}

static Walk make_walk(void* scratch, const LdirScratch& s, int pass) {
    uint8_t* base = static_cast<uint8_t*>(scratch);
    return Walk{reinterpret_cast<const uint16_t*>(base + s.rec1_off), base + s.rec2_off,
                reinterpret_cast<uint32_t*>(base + (pass == 1 ? s.bits1_off : s.bits2_off)), pass};
}

hipError_t launch_ldir_round(const Dims& d, void* scratch, const LdirScratch& s, int pass, bool start, const MList& in,
                             const MList& out, uint32_t* clear_counts, uint32_t* overflow, hipStream_t st) {
    hipLaunchKernelGGL(ldir_round_kernel, dim3(kRoundBlocks), dim3(kThreads), 0, st, make_geom(d),
                       make_walk(scratch, s, pass), in, out, clear_counts, overflow, start ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_ldir_sweep(const Dims& d, const uint8_t* grad, void* scratch, const LdirScratch& s, int pass,
                             const MList& out, uint32_t* overflow, hipStream_t st) {
    hipLaunchKernelGGL(ldir_sweep_kernel, dim3(grid_of(d.verts, 8192)), dim3(kThreads), 0, st, grad, make_geom(d),
                       make_walk(scratch, s, pass), out, overflow);
    return hipGetLastError();
}

hipError_t launch_ldir_finish(const Dims& d, uint8_t* grad, void* scratch, const LdirScratch& s, hipStream_t st) {
    uint8_t* base = static_cast<uint8_t*>(scratch);
    hipLaunchKernelGGL(ldir_finish_kernel, dim3(grid_of(d.verts, 16384)), dim3(kThreads), 0, st, grad, make_geom(d),
                       reinterpret_cast<const uint32_t*>(base + s.bits1_off),
                       reinterpret_cast<const uint32_t*>(base + s.bits2_off));
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

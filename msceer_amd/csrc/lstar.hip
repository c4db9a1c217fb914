// Lower-star discrete gradient on gfx950.
//
// One lane owns one vertex.  A work-group stages a halo'd brick of vertex
// values in LDS as order-preserving 32-bit keys, every lane builds its lower
// star as a 27-bit mask over its star (see star.hpp) and runs the relaxed
// Robins homotopy expansion with mask algebra, then writes the bytes of the
// cells it owns.  Cell ownership is by highest vertex, so each .grad byte is
// written by exactly one lane and no pre-clear of the output is needed.
//
// Semantics restated (reference file:line):
//   order:      gi_regular_grid_trilinear_function.h:309-314 (value, then id)
//   lower star: gi_fast_robins_noalloc.h:353-428, gi_modified_robins.h:1281-1343
//               (subsets keyed restriction*3 + boundaryValue, 20 or 24 slots)
//   expansion:  gi_fast_robins_noalloc.h:147-321, gi_modified_robins.h:999-1199
//   byte:       gi_discrete_gradient_labeling.h:11-19,66-171,200-226
//   ldir init:  gi_topological_gradient_using_algorithms.h:438-477 (3, and 0
//               for critical 3-cells, which the ldir passes never touch)
#include <hip/hip_runtime.h>

#include "internal.hpp"
#include "star.hpp"

namespace msceer {

namespace {

constexpr int kBX = 64, kBY = 2, kBZ = 2;           // vertices per brick
constexpr int kTX = kBX + 2, kTY = kBY + 2, kTZ = kBZ + 2;
constexpr int kTileN = kTX * kTY * kTZ;
constexpr int kThreads = kBX * kBY * kBZ;

// Order-preserving key of a float: +0 and -0 map to the same key.
__device__ __forceinline__ uint32_t float_key(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) == 0u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ int tid_x(int t) { return t - 3 * ((t * 11) >> 5) - 1; }
__device__ __forceinline__ int tid_y(int t) { return ((t * 11) >> 5) - 3 * ((t * 57) >> 9) - 1; }
__device__ __forceinline__ int tid_z(int t) { return ((t * 57) >> 9) - 1; }

// LDS offset of the neighbour vertex at star bit t.
__device__ __forceinline__ int tile_off(int t) {
    return tid_x(t) + kTX * tid_y(t) + kTX * kTY * tid_z(t);
}

// (key, star bit) of a neighbour vertex: comparing these as 64-bit integers is
// the reference order restricted to one neighbourhood, because star-bit order
// is global-id order for the vertices of one 3x3x3 block.
__device__ __forceinline__ uint64_t vkey(const uint32_t* tile, int base, int t) {
    return ((uint64_t)tile[base + tile_off(t)] << 5) | (uint64_t)t;
}

// Lowest vertex of the cell at star bit t.
__device__ uint64_t cell_lowest(const uint32_t* tile, int base, int t) {
    const int dx = tid_x(t), dy = tid_y(t), dz = tid_z(t);
    uint64_t best = ~0ull;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int sx = (s & 1) ? dx : 0, sy = (s & 2) ? dy : 0, sz = (s & 4) ? dz : 0;
        const int q = (sx + 1) + 3 * (sy + 1) + 9 * (sz + 1);
        const uint64_t k = ((uint64_t)tile[base + sx + kTX * sy + kTX * kTY * sz] << 5) | (uint64_t)q;
        best = k < best ? k : best;
    }
    return best;
}

struct StarResult {
    uint32_t c0 = 0, c1 = 0, c2 = 0;  // bits of each cell's 3-bit pair code
    uint32_t crit = 0;                // critical cells
    __device__ __forceinline__ void set_code(int t, uint32_t code) {
        c0 |= (code & 1u) << t;
        c1 |= ((code >> 1) & 1u) << t;
        c2 |= (code >> 2) << t;
    }
    __device__ __forceinline__ void set_critical(uint32_t m) {
        c0 |= m; c1 |= m; c2 |= m; crit |= m;
    }
};

__device__ __forceinline__ bool axis_zero(int t, int axis) {
    return axis == 0 ? tid_x(t) == 0 : (axis == 1 ? tid_y(t) == 0 : tid_z(t) == 0);
}

// Relaxed Robins expansion of one subset S of the lower star.
__device__ void expand_subset(uint32_t S, const uint32_t* tile, int base, StarResult& r) {
    uint32_t P = 0;  // processed (paired or critical)
    if (S & kCenter) {
        const uint32_t E = S & kDim1;
        if (!E) {
            r.set_critical(kCenter);
            P |= kCenter;
        } else {
            // steepest edge: the edge whose other vertex comes first
            int best = -1;
            uint64_t bk = ~0ull;
            uint32_t m = E;
            while (m) {
                const int t = __builtin_ctz(m);
                m &= m - 1;
                const uint64_t k = vkey(tile, base, t);
                if (k < bk) { bk = k; best = t; }
            }
            r.set_code(13, star_dir_code(13, best));
            r.set_code(best, star_dir_code(best, 13));
            P |= kCenter | (1u << best);
        }
    }
#pragma unroll
    for (int d = 1; d <= 2; d++) {
        const uint32_t dim_d = d == 1 ? kDim1 : kDim2;
        const uint32_t dim_n = d == 1 ? kDim2 : kDim3;
        while (true) {
            const uint32_t U = S & ~P;
            const uint32_t Ud = U & dim_d;
            if (!Ud) break;
            // cells of dimension d+1 with exactly one unprocessed facet
            const uint32_t fx = star_facet_x(U), fy = star_facet_y(U), fz = star_facet_z(U);
            const uint32_t one = (fx ^ fy ^ fz) & ~(fx & fy & fz);
            const uint32_t cand = S & dim_n & one;
            const uint32_t ready = Ud & star_has_cofacet_in(cand);
            if (ready) {
                // first ready cell in star order; its candidates in cofacet
                // order +x,-x,+y,-y,+z,-z; first strictly lowest wins
                const int a = __builtin_ctz(ready);
                int best = -1;
                uint64_t bk = ~0ull;
#pragma unroll
                for (int i = 0; i < 6; i++) {
                    const int step = (i < 2 ? 1 : (i < 4 ? 3 : 9)) * ((i & 1) ? -1 : 1);
                    if (!axis_zero(a, i >> 1)) continue;
                    const int c = a + step;
                    if (!((cand >> c) & 1u)) continue;
                    const uint64_t k = cell_lowest(tile, base, c);
                    if (k < bk) { bk = k; best = c; }
                }
                r.set_code(a, star_dir_code(a, best));
                r.set_code(best, star_dir_code(best, a));
                P |= (1u << a) | (1u << best);
            } else {
                // no pairing possible: the first cell with the lowest lowest
                // vertex becomes critical
                int best = -1;
                uint64_t bk = ~0ull;
                uint32_t m = Ud;
                while (m) {
                    const int t = __builtin_ctz(m);
                    m &= m - 1;
                    const uint64_t k = d == 1 ? vkey(tile, base, t) : cell_lowest(tile, base, t);
                    if (k < bk) { bk = k; best = t; }
                }
                r.set_critical(1u << best);
                P |= 1u << best;
            }
        }
    }
    // 3-cells left over have no cofacets: all critical
    r.set_critical(S & kDim3 & ~P);
}

template <bool kConstrained>
__global__ void __launch_bounds__(kThreads)
lstar_kernel(const float* __restrict__ values, uint8_t* __restrict__ grad,
             const int8_t* __restrict__ restriction, int X, int Y, int Z, uint32_t Mx, uint32_t My,
             uint32_t* __restrict__ list1, uint32_t* __restrict__ list2, uint32_t cap,
             uint32_t* __restrict__ counters) {
    __shared__ uint32_t tile[kTileN];
    const int x0 = blockIdx.x * kBX, y0 = blockIdx.y * kBY, z0 = blockIdx.z * kBZ;
    const int flat = threadIdx.x + kBX * (threadIdx.y + kBY * threadIdx.z);
    for (int i = flat; i < kTileN; i += kThreads) {
        const int ix = i % kTX, iy = (i / kTX) % kTY, iz = i / (kTX * kTY);
        int gx = x0 - 1 + ix, gy = y0 - 1 + iy, gz = z0 - 1 + iz;
        gx = gx < 0 ? 0 : (gx >= X ? X - 1 : gx);
        gy = gy < 0 ? 0 : (gy >= Y ? Y - 1 : gy);
        gz = gz < 0 ? 0 : (gz >= Z ? Z - 1 : gz);
        tile[i] = float_key(values[(size_t)gx + (size_t)X * ((size_t)gy + (size_t)Y * gz)]);
    }
    __syncthreads();

    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y, z = z0 + threadIdx.z;
    if (x >= X || y >= Y || z >= Z) return;
    const int base = (threadIdx.x + 1) + kTX * ((threadIdx.y + 1) + kTY * (threadIdx.z + 1));

    // neighbours strictly before this vertex
    const uint32_t kc = tile[base];
    uint32_t lower = 0;
#pragma unroll
    for (int t = 0; t < 27; t++) {
        if (t == 13) continue;
        const uint32_t kq = tile[base + tile_off(t)];
        const bool lo = t < 13 ? (kq <= kc) : (kq < kc);
        lower |= (uint32_t)lo << t;
    }
    const uint32_t invalid = (x == 0 ? kXM : 0u) | (x == X - 1 ? kXP : 0u) | (y == 0 ? kYM : 0u) |
                             (y == Y - 1 ? kYP : 0u) | (z == 0 ? kZM : 0u) | (z == Z - 1 ? kZP : 0u);
    const uint32_t L = ~star_dilate(((~lower) & kAll & ~kCenter) | invalid) & kAll;

    // boundary strata of the star cells (gi_topological_regular_grid.h:1291-1303)
    const uint32_t bx = (x == 0 || x == X - 1) ? kX0 : 0u;
    const uint32_t by = (y == 0 || y == Y - 1) ? kY0 : 0u;
    const uint32_t bz = (z == 0 || z == Z - 1) ? kZ0 : 0u;
    const uint32_t onb = bx | by | bz;

    const uint64_t vcell = 2ull * x + (uint64_t)Mx * (2ull * y + (uint64_t)My * 2ull * z);
    StarResult r;
    if (!kConstrained) {
        // steepest pass <5,4>: subsets keyed by boundaryValue 0..3; interior
        // vertices have a single subset
        const uint32_t odd = bx ^ by ^ bz, two = (bx & by) | (bx & bz) | (by & bz);
        const uint32_t sub[4] = {~onb, odd & ~two, two & ~odd, odd & two};
#pragma unroll 1
        for (int k = 0; k < 4; k++) {
            const uint32_t S = L & sub[k];
            if (S) expand_subset(S, tile, base, r);
        }
    } else {
        // constrained pass <4,6>: key = label*3 + boundaryValue, clamped to 23
        uint32_t keyed[24];
        for (int k = 0; k < 24; k++) keyed[k] = 0;
        for (uint32_t m = L; m; m &= m - 1) {
            const int t = __builtin_ctz(m);
            const int bvt = (int)((bx >> t) & 1u) + (int)((by >> t) & 1u) + (int)((bz >> t) & 1u);
            const int64_t off = (int64_t)tid_x(t) + (int64_t)Mx * ((int64_t)tid_y(t) + (int64_t)My * tid_z(t));
            int key = (int)restriction[(int64_t)vcell + off] * 3 + bvt;
            key = key > 23 ? 23 : (key < 0 ? 0 : key);
            keyed[key] |= 1u << t;
        }
#pragma unroll 1
        for (int k = 0; k < 24; k++) {
            if (keyed[k]) expand_subset(keyed[k], tile, base, r);
        }
    }

    // write the owned cells
#pragma unroll
    for (int t = 0; t < 27; t++) {
        if (!((L >> t) & 1u)) continue;
        const uint32_t code = ((r.c0 >> t) & 1u) | (((r.c1 >> t) & 1u) << 1) | (((r.c2 >> t) & 1u) << 2);
        const uint32_t flag = (onb >> t) & 1u;
        const uint32_t ldir = ((kDim3 >> t) & (r.crit >> t) & 1u) ? 0u : 3u;
        const int64_t off = (int64_t)tid_x(t) + (int64_t)Mx * ((int64_t)tid_y(t) + (int64_t)My * tid_z(t));
        grad[(int64_t)vcell + off] = (uint8_t)(1u | (flag << 1) | (code << 2) | (ldir << 5));
    }

    // critical cells: statistics and the ldir start lists
    const uint32_t crit = r.crit;
    if (crit) {
        atomicAdd(&counters[kCritDim0], (crit & kDim0) ? 1u : 0u);
        atomicAdd(&counters[kCritDim1], (uint32_t)__builtin_popcount(crit & kDim1));
        atomicAdd(&counters[kCritDim2], (uint32_t)__builtin_popcount(crit & kDim2));
        atomicAdd(&counters[kCritDim3], (uint32_t)__builtin_popcount(crit & kDim3));
        for (uint32_t m = crit & (kDim1 | kDim2); m; m &= m - 1) {
            const int t = __builtin_ctz(m);
            const int64_t off = (int64_t)tid_x(t) + (int64_t)Mx * ((int64_t)tid_y(t) + (int64_t)My * tid_z(t));
            const uint32_t cell = (uint32_t)((int64_t)vcell + off);
            const bool one = (kDim1 >> t) & 1u;
            const uint32_t slot = atomicAdd(&counters[one ? kCritCount1 : kCritCount2], 1u);
            if (slot < cap) {
                (one ? list1 : list2)[slot] = cell;
            } else {
                atomicOr(&counters[kCritOverflow], 1u);
            }
        }
    }
}

}  // namespace

hipError_t launch_lstar(const Dims& d, const LstarBuffers& b, hipStream_t s) {
    dim3 block(kBX, kBY, kBZ);
    dim3 grid((d.X + kBX - 1) / kBX, (d.Y + kBY - 1) / kBY, (d.Z + kBZ - 1) / kBZ);
    if (b.restriction) {
        hipLaunchKernelGGL(lstar_kernel<true>, grid, block, 0, s, b.values, b.grad, b.restriction, d.X,
                           d.Y, d.Z, d.Mx, d.My, b.crit_list1, b.crit_list2, b.crit_cap, b.counters);
    } else {
        hipLaunchKernelGGL(lstar_kernel<false>, grid, block, 0, s, b.values, b.grad, b.restriction,
                           d.X, d.Y, d.Z, d.Mx, d.My, b.crit_list1, b.crit_list2, b.crit_cap,
                           b.counters);
    }
    return hipGetLastError();
}

}  // namespace msceer

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
#include "lstar_core.hpp"
#include "mlist.hpp"

namespace msceer {

namespace {

constexpr int kThreads = 256;  // vertices per brick, any of the shapes below

// Brick shapes (vertices along x, y, z); the autotuner picks one per volume
// shape: wide bricks coalesce best, narrow ones waste fewer lanes when the
// x extent is small or not a multiple of the brick width.
template <int kBX, int kBY, int kBZ>
struct Brick {
    static_assert(kBX * kBY * kBZ == kThreads, "256 lanes");
    static constexpr int kTX = kBX + 2, kTY = kBY + 2, kTZ = kBZ + 2;
    static constexpr int kTileN = kTX * kTY * kTZ;
};

// Work-groups are dealt to the XCDs round robin (work-group i runs on XCD
// i % 8), each XCD with its own L2.  The brick of work-group i is chosen so
// that every XCD takes a contiguous run of bricks in x-fastest order: bricks
// that share halo values and partly written cache lines of the .grad then
// meet in the same L2.  A bijection of [0, n).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_brick(uint32_t b, uint32_t n) {
    const uint32_t per = n / kXcds, rem = n % kXcds, xcd = b % kXcds, k = b / kXcds;
    return xcd < rem ? xcd * (per + 1) + k : rem * (per + 1) + (xcd - rem) * per + k;
}

template <bool kConstrained, int kBX, int kBY, int kBZ>
__global__ void __launch_bounds__(kThreads)
lstar_kernel(const float* __restrict__ values, uint8_t* __restrict__ grad,
             const int8_t* __restrict__ restriction, int X, int Y, int Z, uint32_t Mx, uint32_t My,
             MList list1, MList list2, uint32_t* __restrict__ stats, uint32_t* __restrict__ overflow, int nbx,
             int nby, int nbz) {
    using B = Brick<kBX, kBY, kBZ>;
    constexpr int kTX = B::kTX, kTY = B::kTY, kTileN = B::kTileN;
    __shared__ uint64_t tile[kTileN];  // (order key << 32 | vertex id) of the halo'd brick
    __shared__ int s_off[27];  // tile offset of star vertex t
    __shared__ int s_cofs[27];  // .grad offset of star cell t from the centre's cell
    __shared__ uint32_t s_cnt[4];  // per block: critical 1-cells, 2-cells, 0-cells, 3-cells
    __shared__ MSlot s_slot[2];
    __shared__ uint32_t s_mask[kThreads];     // lower star of each vertex of the brick
    __shared__ uint16_t s_order[kThreads];    // brick vertices by lower-star size
    __shared__ uint32_t s_bucket[32];
    __shared__ uint32_t s_res[4][kThreads];   // expansion result by brick vertex: pair-code bits, critical
    const uint32_t brick = xcd_brick(blockIdx.x, (uint32_t)nbx * nby * nbz);
    const int x0 = (int)(brick % nbx) * kBX, y0 = (int)((brick / nbx) % nby) * kBY;
    const int z0 = (int)(brick / ((uint32_t)nbx * nby)) * kBZ;
    const int flat = threadIdx.x;
    const int tx = flat % kBX, ty = (flat / kBX) % kBY, tz = flat / (kBX * kBY);
    for (int i = flat; i < kTileN; i += kThreads) {
        const int ix = i % kTX, iy = (i / kTX) % kTY, iz = i / (kTX * kTY);
        int gx = x0 - 1 + ix, gy = y0 - 1 + iy, gz = z0 - 1 + iz;
        gx = gx < 0 ? 0 : (gx >= X ? X - 1 : gx);
        gy = gy < 0 ? 0 : (gy >= Y ? Y - 1 : gy);
        gz = gz < 0 ? 0 : (gz >= Z ? Z - 1 : gz);
        const uint32_t id = (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz);
        tile[i] = ((uint64_t)float_key(values[id]) << 32) | id;
    }
    if (flat < 4) s_cnt[flat] = 0;
    if (flat < 27) {
        s_off[flat] = tid_x(flat) + kTX * (tid_y(flat) + kTY * tid_z(flat));
        s_cofs[flat] = tid_x(flat) + (int)Mx * (tid_y(flat) + (int)My * tid_z(flat));
    }
    if (flat < 32) s_bucket[flat] = 0;
    __syncthreads();

    // Lower star of the lane's own vertex (0: outside the volume; a lower star
    // always holds its vertex).  The expansion loops of a wave run as long as
    // its largest star needs, so the brick's vertices are then dealt to the
    // lanes in order of lower-star size (counting sort in LDS): a wave gets
    // stars of similar size.  Any lane can expand any vertex of the brick,
    // since all keys are in the tile.
    {
        const int ox = x0 + tx, oy = y0 + ty, oz = z0 + tz;
        uint32_t own = 0;
        if (ox < X && oy < Y && oz < Z) {
            const int obase = (tx + 1) + kTX * ((ty + 1) + kTY * (tz + 1));
            own = lower_star_mask(FixedKeys64<kTX, kTX * kTY>{tile, obase}, ox, oy, oz, X, Y, Z);
        }
        s_mask[flat] = own;
        const uint32_t w = (uint32_t)__builtin_popcount(own);
        const uint32_t rank = atomicAdd(&s_bucket[w], 1u);
        __syncthreads();
        if (flat < 64) {  // exclusive prefix over the 28 sizes (one wave)
            const uint32_t c = flat < 32 ? s_bucket[flat] : 0u;
            uint32_t incl = c;
#pragma unroll
            for (int o = 1; o < 32; o <<= 1) {
                const uint32_t t = __shfl_up(incl, o);
                if ((flat & 63) >= o) incl += t;
            }
            if (flat < 32) s_bucket[flat] = incl - c;
        }
        __syncthreads();
        s_order[s_bucket[w] + rank] = (uint16_t)flat;
        __syncthreads();
    }
    const int me = s_order[flat];
    const int mx = me % kBX, my = (me / kBX) % kBY, mz = me / (kBX * kBY);
    const int x = x0 + mx, y = y0 + my, z = z0 + mz;
    const uint32_t L = s_mask[me];
    const bool active = L != 0u;
    uint32_t crit = 0;
    uint32_t vcell = 0;
    if (active) {
    const int base = (mx + 1) + kTX * ((my + 1) + kTY * (mz + 1));
    const TableKeys64 keys{tile + base, s_off};
    uint32_t bx, by, bz;
    star_boundary(x, y, z, X, Y, Z, bx, by, bz);

    vcell = 2u * x + Mx * (2u * y + My * 2u * z);  // cells < 2^32
    StarResult r;
    if (!kConstrained) {
        steepest_expand(keys, L, bx, by, bz, r);
    } else {
        const int8_t* rv = restriction + vcell;
        const int64_t sy = Mx, sz = (int64_t)Mx * My;
        auto label = [rv, sy, sz](int t) -> int {
            return rv[tid_x(t) + sy * tid_y(t) + sz * tid_z(t)];
        };
        constrained_expand(keys, L, bx, by, bz, label, r);
    }

    s_res[0][me] = r.c0;
    s_res[1][me] = r.c1;
    s_res[2][me] = r.c2;
    s_res[3][me] = r.crit;
    }  // active
    __syncthreads();
    // each lane writes the owned cells of its own vertex: the lanes of a wave
    // hold neighbouring vertices, so their byte stores share cache lines
    {
        const uint32_t Lh = s_mask[flat];
        const int hx = x0 + tx, hy = y0 + ty, hz = z0 + tz;
        vcell = 2u * hx + Mx * (2u * hy + My * 2u * hz);
        if (Lh) {
            StarResult rh;
            rh.c0 = s_res[0][flat];
            rh.c1 = s_res[1][flat];
            rh.c2 = s_res[2][flat];
            rh.crit = s_res[3][flat];
            uint32_t hbx, hby, hbz;
            star_boundary(hx, hy, hz, X, Y, Z, hbx, hby, hbz);
            const uint32_t honb = hbx | hby | hbz;
            for (uint32_t m = Lh; m; m &= m - 1) {
                const int t = __builtin_ctz(m);
                grad[vcell + (uint32_t)s_cofs[t]] = star_cell_byte(rh, honb, t);
            }
            crit = rh.crit;
        }
    }

    // critical cells: statistics and the ldir start lists, reserved once per
    // block in sub-list blockIdx % kSubs
    const uint32_t n1 = __builtin_popcount(crit & kDim1), n2 = __builtin_popcount(crit & kDim2);
    uint32_t o1 = 0, o2 = 0;
    if (crit) {
        o1 = atomicAdd(&s_cnt[0], n1);
        o2 = atomicAdd(&s_cnt[1], n2);
        if (crit & kDim0) atomicAdd(&s_cnt[2], 1u);
        if (crit & kDim3) atomicAdd(&s_cnt[3], (uint32_t)__builtin_popcount(crit & kDim3));
    }
    __syncthreads();
    const int sub = (int)(blockIdx.x % kSubs);
    if (flat == 0) {
        const uint32_t c1 = s_cnt[0], c2 = s_cnt[1];
        s_slot[0] = c1 ? mlist_reserve(list1, sub, c1) : MSlot{0u, 0u, 0u};
        s_slot[1] = c2 ? mlist_reserve(list2, sub, c2) : MSlot{0u, 0u, 0u};
        const uint32_t st[4] = {s_cnt[2], c1, c2, s_cnt[3]};
        for (int k = 0; k < 4; k++)
            if (st[k]) atomicAdd(&stats[sub * 4 + k], st[k]);
    }
    __syncthreads();
    if (crit & (kDim1 | kDim2)) {
        const MSlot r1 = s_slot[0], r2 = s_slot[1];
        uint32_t j1 = o1, j2 = o2;
        bool lost = false;
        for (uint32_t m = crit & (kDim1 | kDim2); m; m &= m - 1) {
            const int t = __builtin_ctz(m);
            const uint32_t cell = vcell + (uint32_t)s_cofs[t];
            if ((kDim1 >> t) & 1u) lost |= !mlist_store(list1, sub, r1, j1++, cell);
            else lost |= !mlist_store(list2, sub, r2, j2++, cell);
        }
        if (lost) atomicOr(overflow, 1u);
    }
}

}  // namespace

namespace {

template <int kBX, int kBY, int kBZ>
void launch_shape(const Dims& d, const LstarBuffers& b, hipStream_t s) {
    const int nbx = (d.X + kBX - 1) / kBX, nby = (d.Y + kBY - 1) / kBY, nbz = (d.Z + kBZ - 1) / kBZ;
    const dim3 grid((uint32_t)nbx * nby * nbz), block(kThreads);
    if (b.restriction)
        hipLaunchKernelGGL((lstar_kernel<true, kBX, kBY, kBZ>), grid, block, 0, s, b.values, b.grad, b.restriction,
                           d.X, d.Y, d.Z, d.Mx, d.My, *b.list1, *b.list2, b.stats, b.overflow, nbx, nby, nbz);
    else
        hipLaunchKernelGGL((lstar_kernel<false, kBX, kBY, kBZ>), grid, block, 0, s, b.values, b.grad, b.restriction,
                           d.X, d.Y, d.Z, d.Mx, d.My, *b.list1, *b.list2, b.stats, b.overflow, nbx, nby, nbz);
}

}  // namespace

const char* lstar_brick_name(int shape) {
    static const char* names[kLstarBricks] = {"64x2x2", "32x4x2", "16x4x4", "8x8x4"};
    return shape >= 0 && shape < kLstarBricks ? names[shape] : "?";
}

hipError_t launch_lstar(const Dims& d, const LstarBuffers& b, hipStream_t s, int shape) {
    switch (shape) {
        case 1: launch_shape<32, 4, 2>(d, b, s); break;
        case 2: launch_shape<16, 4, 4>(d, b, s); break;
        case 3: launch_shape<8, 8, 4>(d, b, s); break;
        default: launch_shape<64, 2, 2>(d, b, s); break;
    }
    return hipGetLastError();
}

}  // namespace msceer

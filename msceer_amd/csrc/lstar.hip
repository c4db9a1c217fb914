// Lower-star discrete gradient on gfx950.
//
// A work-group owns a brick of vertices.  It stages the halo'd brick's
// values in LDS as 64-bit order keys, every lane builds the lower star of one
// vertex as a 27-bit mask over its star (see star.hpp), and the brick's
// vertices are dealt to the lanes by shape: product lower stars (a minimum,
// an edge, a quadrant, an octant -- three in four vertices of a smooth
// field) take a closed form that writes their bytes at once; the others run
// the relaxed Robins expansion with mask algebra, grouped by size so that a
// wave's loops run about as long as each of its lanes needs.  The bytes go to
// an LDS image of the cells around the brick, which the work-group then
// copies to the .grad with contiguous byte stores, collecting the critical
// cells on the way.  Cell ownership is by highest vertex, so each .grad byte
// is written by exactly one work-group and no pre-clear of the output is
// needed.
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
#include "gbatch.hpp"
#include "mlist.hpp"

namespace msceer {

namespace {

constexpr int kThreads = 256;  // vertices per brick, any of the shapes below

// Brick shapes (vertices along x, y, z); the autotuner picks one per volume
// shape: wide bricks coalesce best, narrow ones waste fewer lanes when the
// x extent is small or not a multiple of the brick width.
//   tile:  the brick's vertices with a one-vertex halo (kTX * kTY * kTZ keys)
//   stage: the cells whose highest vertex can be a brick vertex, cell
//          coordinates 2 x0 - 1 .. 2 (x0 + kBX) - 1 along x and likewise for
//          y and z; rows padded to whole dwords.  The outer layers are shared
//          with the neighbouring bricks, the rest is the brick's alone.
template <int kBX, int kBY, int kBZ>
struct Brick {
    static_assert(kBX * kBY * kBZ == kThreads, "256 lanes");
    static constexpr int kTX = kBX + 2, kTY = kBY + 2, kTZ = kBZ + 2;
    static constexpr int kTileN = kTX * kTY * kTZ;
    static constexpr int kCX = 2 * kBX + 1, kCY = 2 * kBY + 1, kCZ = 2 * kBZ + 1;
    static constexpr int kRowW = (kCX + 3) / 4;  // dwords per staged row
    static constexpr int kRows = kCY * kCZ;
    static constexpr int kStageW = kRowW * kRows;
    // cells of dimension 1 and 2 in the stage (odd coordinates: kB + 1 per
    // axis, even: kB), the room their critical-cell lists need
    static constexpr int kN1 = (kBX + 1) * kBY * kBZ + kBX * (kBY + 1) * kBZ + kBX * kBY * (kBZ + 1);
    static constexpr int kN2 = (kBX + 1) * (kBY + 1) * kBZ + (kBX + 1) * kBY * (kBZ + 1) + kBX * (kBY + 1) * (kBZ + 1);
    // the lists reuse the tile's LDS once the expansions are done
    static constexpr int kSmem = kTileN > (kN1 + kN2 + 1) / 2 ? kTileN : (kN1 + kN2 + 1) / 2;
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

// Keys and byte sink of one lane's product star: tile and stage offsets from
// the lane's vertex and its cell.
template <int kTX, int kTXY, int kRow, int kRowPlane>
struct LaneStar {
    const uint64_t* kc;
    uint8_t* sc;
    static constexpr int ts[3] = {1, kTX, kTXY};
    static constexpr int cs[3] = {1, kRow, kRowPlane};
    __device__ __forceinline__ uint64_t key(int o) const { return kc[o]; }
    __device__ __forceinline__ void put(int o, uint32_t b) const { sc[o] = (uint8_t)b; }
};

// Arguments of one launch: a single volume (values .. nbz), or with kBatch
// the volumes of a group (batch; the lists then hold each volume's cells
// from its first frontier item on).
struct LstarArgs {
    const float* values;
    uint8_t* grad;
    const int8_t* restriction;
    int X, Y, Z;
    uint32_t Mx, My;
    MList list1, list2;
    uint32_t* stats;
    uint32_t* overflow;
    int nbx, nby, nbz;
    GBatch batch;
};

// kMode 0: one volume, a work-group per brick; 1: the bricks of a group of
// volumes, a work-group per brick; 2: the same with persistent work-groups
// claiming bricks from the queues.
template <bool kConstrained, int kMode, int kBX, int kBY, int kBZ>
__global__ void __launch_bounds__(kThreads) lstar_kernel(LstarArgs a) {
    using B = Brick<kBX, kBY, kBZ>;
    constexpr int kTX = B::kTX, kTY = B::kTY, kTXY = kTX * kTY, kTileN = B::kTileN;
    constexpr int kCY = B::kCY, kRowW = B::kRowW, kRow = 4 * kRowW, kRows = B::kRows, kStageW = B::kStageW;
    __shared__ uint64_t smem[B::kSmem];       // tile: (order key << 32 | vertex id); later the lists
    __shared__ uint32_t stage[kStageW];       // .grad bytes of the cells around the brick
    __shared__ uint32_t s_row[kRows];         // .grad index of stage row r at x cell 2 x0
    __shared__ int s_off[27];                 // tile offset of star vertex t
    __shared__ int s_soff[27];                // stage offset of star cell t
    __shared__ uint32_t s_mask[kThreads];     // lower star of each vertex of the brick
    __shared__ uint8_t s_shape[kThreads];     // its product shape, or kNotProduct
    __shared__ uint16_t s_order[kThreads];    // brick vertices by shape and size
    __shared__ uint32_t s_bucket[32];
    __shared__ uint32_t s_cnt[8];             // critical cells by dimension; list fills
    __shared__ MSlot s_slot[2];
    uint64_t* const tile = smem;
    uint8_t* const st8 = reinterpret_cast<uint8_t*>(stage);
    const MList list1 = a.list1, list2 = a.list2;
    uint32_t* const stats = a.stats;
    uint32_t* const overflow = a.overflow;
    __shared__ UnitCursor s_cur;
    if (kMode == 2) cursor_init(&s_cur);

    // one brick, or (persistent) bricks until the queues are empty
    for (;;) {
        const float* __restrict__ values = a.values;
        uint8_t* __restrict__ grad = a.grad;
        const int8_t* __restrict__ restriction = a.restriction;
        int X = a.X, Y = a.Y, Z = a.Z, nbx = a.nbx, nby = a.nby;
        uint32_t Mx = a.Mx, My = a.My, item0 = 0, brick;
        if (kMode == 0) {
            brick = xcd_brick(blockIdx.x, (uint32_t)a.nbx * a.nby * a.nbz);
        } else {
            uint32_t unit, vol;
            if (kMode == 1) {
                unit = xcd_brick(blockIdx.x, a.batch.bricks.total);
                vol = unit_volume(a.batch.bricks, unit);
            } else if (!cursor_next(&s_cur, a.batch.queue, a.batch.bricks, a.batch.chunk, unit, vol)) {
                break;
            }
            const GVol& gv = a.batch.vols[vol];
            values = gv.values;
            grad = gv.grad;
            X = gv.X;
            Y = gv.Y;
            Z = gv.Z;
            Mx = gv.Mx;
            My = gv.My;
            nbx = (X + kBX - 1) / kBX;
            nby = (Y + kBY - 1) / kBY;
            item0 = a.batch.items.first[vol];
            brick = unit - a.batch.bricks.first[vol];
            if (a.batch.times && threadIdx.x == 0) atomicMin(&a.batch.times[2 * vol], batch_clock());
        }
        // an opaque lane index in the persistent loop: without it the
        // compiler keeps lane-invariant values of the brick work live across
        // iterations, 40 more VGPRs
        int flat = threadIdx.x;
        if (kMode == 2) asm volatile("" : "+v"(flat));
        const int x0 = (int)(brick % nbx) * kBX, y0 = (int)((brick / nbx) % nby) * kBY;
        const int z0 = (int)(brick / ((uint32_t)nbx * nby)) * kBZ;
        const int tx = flat % kBX, ty = (flat / kBX) % kBY, tz = flat / (kBX * kBY);
        // every vertex of the brick and its whole star inside the domain
        const bool inner = x0 >= 1 && y0 >= 1 && z0 >= 1 && x0 + kBX < X && y0 + kBY < Y && z0 + kBZ < Z;
        if (inner) {
            // tile row (iy, iz) starts at vertex id corner + X iy + X Y iz
            __shared__ uint32_t s_trow[kTY * B::kTZ];
            if (flat < kTY * B::kTZ)
                s_trow[flat] = (uint32_t)X * ((uint32_t)(flat % kTY) + (uint32_t)Y * (uint32_t)(flat / kTY));
            __syncthreads();
            const uint32_t corner = (uint32_t)(x0 - 1) + (uint32_t)X * ((uint32_t)(y0 - 1) + (uint32_t)Y * (uint32_t)(z0 - 1));
            for (int i = flat; i < kTileN; i += kThreads) {
                const int row = i / kTX, ix = i - row * kTX;
                const uint32_t id = corner + (uint32_t)ix + s_trow[row];
                tile[i] = ((uint64_t)float_key(values[id]) << 32) | id;
            }
        } else {
            for (int i = flat; i < kTileN; i += kThreads) {
                const int ix = i % kTX, iy = (i / kTX) % kTY, iz = i / kTXY;
                int gx = x0 - 1 + ix, gy = y0 - 1 + iy, gz = z0 - 1 + iz;
                gx = gx < 0 ? 0 : (gx >= X ? X - 1 : gx);
                gy = gy < 0 ? 0 : (gy >= Y ? Y - 1 : gy);
                gz = gz < 0 ? 0 : (gz >= Z ? Z - 1 : gz);
                const uint32_t id = (uint32_t)gx + (uint32_t)X * ((uint32_t)gy + (uint32_t)Y * (uint32_t)gz);
                tile[i] = ((uint64_t)float_key(values[id]) << 32) | id;
            }
        }
        for (int i = flat; i < kStageW; i += kThreads) stage[i] = 0u;
        for (int r = flat; r < kRows; r += kThreads) {
            const uint32_t gy = (uint32_t)(2 * y0 - 1 + r % kCY), gz = (uint32_t)(2 * z0 - 1 + r / kCY);
            s_row[r] = (uint32_t)(2 * x0) + Mx * (gy + My * gz);  // wraps for rows outside the domain (never stored)
        }
        if (flat < 8) s_cnt[flat] = 0;
        if (flat < 27) {
            s_off[flat] = tid_x(flat) + kTX * (tid_y(flat) + kTY * tid_z(flat));
            s_soff[flat] = tid_x(flat) + kRow * (tid_y(flat) + kCY * tid_z(flat));
        }
        if (flat < 32) s_bucket[flat] = 0;
        __syncthreads();

        // Lower star and shape of the lane's own vertex (0: outside the volume;
        // a lower star always holds its vertex).  A wave runs each expansion path
        // that any of its lanes takes, the general loops as long as its largest
        // star needs, so the vertices are then dealt to the lanes in bucket order
        // (counting sort in LDS): the four product shapes first, then the rest by
        // size.  Any lane can expand any vertex of the brick, since all keys are
        // in the tile.
        {
            const int ox = x0 + tx, oy = y0 + ty, oz = z0 + tz;
            const int obase = (tx + 1) + kTX * ((ty + 1) + kTY * (tz + 1));
            uint32_t own = 0, shape = kNotProduct;
            if (inner) {
                own = lower_star_from(lower_neighbours(FixedKeys64<kTX, kTXY>{tile, obase}), 0u);
                if (!kConstrained) shape = product_shape(own);
            } else if (ox < X && oy < Y && oz < Z) {
                own = lower_star_from(lower_neighbours(FixedKeys64<kTX, kTXY>{tile, obase}),
                                      star_outside(ox, oy, oz, X, Y, Z));
                const bool onb = ox == 0 || ox == X - 1 || oy == 0 || oy == Y - 1 || oz == 0 || oz == Z - 1;
                if (!kConstrained && !onb) shape = product_shape(own);
            }
            s_mask[flat] = own;
            s_shape[flat] = (uint8_t)shape;
            const uint32_t w = shape != kNotProduct ? (uint32_t)__builtin_popcount(shape >> 3)
                                                    : 4u + (uint32_t)__builtin_popcount(own);
            const uint32_t rank = atomicAdd(&s_bucket[w], 1u);
            __syncthreads();
            if (flat < 64) {  // exclusive prefix over the 32 buckets (one wave)
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
        {
            const int me = s_order[flat];
            const uint32_t L = s_mask[me];
            if (L) {
                const int mx = me % kBX, my = (me / kBX) % kBY, mz = me / (kBX * kBY);
                const int base = (mx + 1) + kTX * ((my + 1) + kTY * (mz + 1));
                const int cbase = (2 * mx + 1) + kRow * ((2 * my + 1) + kCY * (2 * mz + 1));
                const uint32_t shape = s_shape[me];
                if (shape != kNotProduct) {
                    expand_product(shape, LaneStar<kTX, kTXY, kRow, kRow * kCY>{tile + base, st8 + cbase});
                } else {
                    const int x = x0 + mx, y = y0 + my, z = z0 + mz;
                    const TableKeys64 keys{tile + base, s_off};
                    uint32_t bx, by, bz;
                    star_boundary(x, y, z, X, Y, Z, bx, by, bz);
                    // the expansion writes each cell's byte as it pairs it
                    StageBytes r{st8 + cbase, s_soff, bx | by | bz};
                    if (!kConstrained) {
                        steepest_expand(keys, L, bx, by, bz, r);
                    } else {
                        const int8_t* rv = restriction + (2u * x + Mx * (2u * y + My * 2u * z));
                        const int64_t sy = Mx, sz = (int64_t)Mx * My;
                        auto label = [rv, sy, sz](int t) -> int {
                            return rv[tid_x(t) + sy * tid_y(t) + sz * tid_z(t)];
                        };
                        constrained_expand(keys, L, bx, by, bz, label, r);
                    }
                }
            }
        }
        __syncthreads();

        // Copy the stage out, four cells a lane at a time along the rows; a zero
        // byte is a cell no brick vertex owns.  Critical cells (pair code 7) are
        // counted by dimension and the 1- and 2-cells listed for the ldir passes.
        // Cell parity: stage x offset 4 j + b is .grad x 2 x0 - 1 + 4 j + b (odd
        // for even b), stage row (j, k) is y 2 y0 - 1 + j, z 2 z0 - 1 + k.
        uint32_t* const lst = reinterpret_cast<uint32_t*>(smem);
        for (int ch = flat; ch < kStageW; ch += kThreads) {
            const uint32_t w = stage[ch];
            if (!w) continue;
            const int row = ch / kRowW, j4 = ch - row * kRowW;
            uint8_t* const p = grad + (size_t)s_row[row] + 4 * j4;  // byte b of w is p[b - 1]
            if (w & 0xffu) p[-1] = (uint8_t)w;
            if (w & 0xff00u) p[0] = (uint8_t)(w >> 8);
            if (w & 0xff0000u) p[1] = (uint8_t)(w >> 16);
            if (w & 0xff000000u) p[2] = (uint8_t)(w >> 24);
            const uint32_t z = ((w >> 2) & 0x07070707u) ^ 0x07070707u;  // zero byte: critical
            if ((z - 0x01010101u) & ~z & 0x80808080u) {
                const int j = row % kCY, k = row / kCY;
                const int dyz = ((j & 1) ^ 1) + ((k & 1) ^ 1);
    #pragma unroll
                for (int b = 0; b < 4; b++) {
                    const uint32_t v = (w >> (8 * b)) & 0xffu;
                    if (v == 0u || ((v >> 2) & 7u) != 7u) continue;
                    const int dim = dyz + ((b & 1) ^ 1);
                    atomicAdd(&s_cnt[dim], 1u);
                    const uint32_t cell = item0 + s_row[row] + (uint32_t)(4 * j4 + b - 1);
                    if (dim == 1) lst[atomicAdd(&s_cnt[4], 1u)] = cell;
                    else if (dim == 2) lst[B::kN1 + atomicAdd(&s_cnt[5], 1u)] = cell;
                }
            }
        }
        __syncthreads();
        const int sub = (int)(blockIdx.x % kSubs);
        const uint32_t c1 = s_cnt[4], c2 = s_cnt[5];
        if (flat == 0) {
            s_slot[0] = c1 ? mlist_reserve(list1, sub, c1) : MSlot{0u, 0u, 0u};
            s_slot[1] = c2 ? mlist_reserve(list2, sub, c2) : MSlot{0u, 0u, 0u};
            for (int k = 0; k < 4; k++)
                if (s_cnt[k]) atomicAdd(&stats[sub * 4 + k], s_cnt[k]);
        }
        __syncthreads();
        if (c1 + c2) {
            const MSlot r1 = s_slot[0], r2 = s_slot[1];
            bool lost = false;
            for (uint32_t i = flat; i < c1; i += kThreads) lost |= !mlist_store(list1, sub, r1, i, lst[i]);
            for (uint32_t i = flat; i < c2; i += kThreads) lost |= !mlist_store(list2, sub, r2, i, lst[B::kN1 + i]);
            if (lost) atomicOr(overflow, 1u);
        }
        if (kMode != 2) break;
        __syncthreads();  // the next brick's tile overwrites the lists in LDS
    }
}

}  // namespace

namespace {

template <int kBX, int kBY, int kBZ>
void launch_shape(const Dims& d, const LstarBuffers& b, hipStream_t s) {
    const int nbx = (d.X + kBX - 1) / kBX, nby = (d.Y + kBY - 1) / kBY, nbz = (d.Z + kBZ - 1) / kBZ;
    const dim3 grid((uint32_t)nbx * nby * nbz), block(kThreads);
    LstarArgs a{};
    a.values = b.values;
    a.grad = b.grad;
    a.restriction = b.restriction;
    a.X = d.X;
    a.Y = d.Y;
    a.Z = d.Z;
    a.Mx = d.Mx;
    a.My = d.My;
    a.list1 = *b.list1;
    a.list2 = *b.list2;
    a.stats = b.stats;
    a.overflow = b.overflow;
    a.nbx = nbx;
    a.nby = nby;
    a.nbz = nbz;
    if (b.restriction)
        hipLaunchKernelGGL((lstar_kernel<true, 0, kBX, kBY, kBZ>), grid, block, 0, s, a);
    else
        hipLaunchKernelGGL((lstar_kernel<false, 0, kBX, kBY, kBZ>), grid, block, 0, s, a);
}

template <int kBX, int kBY, int kBZ>
void launch_batch_shape(const GBatch& g, const LstarBuffers& b, uint32_t blocks, hipStream_t s) {
    LstarArgs a{};
    a.list1 = *b.list1;
    a.list2 = *b.list2;
    a.stats = b.stats;
    a.overflow = b.overflow;
    a.batch = g;
    if (blocks == 0)
        hipLaunchKernelGGL((lstar_kernel<false, 1, kBX, kBY, kBZ>), dim3(g.bricks.total), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL((lstar_kernel<false, 2, kBX, kBY, kBZ>), dim3(blocks), dim3(kThreads), 0, s, a);
}

template <int kBX, int kBY, int kBZ>
int batch_occupancy() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, lstar_kernel<false, 2, kBX, kBY, kBZ>, kThreads, 0) !=
        hipSuccess)
        return 1;
    return n > 0 ? n : 1;
}

}  // namespace

const char* lstar_brick_name(int shape) {
    static const char* names[kLstarBricks] = {"64x2x2", "32x4x2", "16x4x4", "8x8x4"};
    return shape >= 0 && shape < kLstarBricks ? names[shape] : "?";
}

void lstar_brick_dims(int shape, int* bx, int* by, int* bz) {
    static const int dims[kLstarBricks][3] = {{64, 2, 2}, {32, 4, 2}, {16, 4, 4}, {8, 8, 4}};
    const int k = shape >= 0 && shape < kLstarBricks ? shape : 0;
    *bx = dims[k][0];
    *by = dims[k][1];
    *bz = dims[k][2];
}

int lstar_batch_occupancy(int shape) {
    switch (shape) {
        case 1: return batch_occupancy<32, 4, 2>();
        case 2: return batch_occupancy<16, 4, 4>();
        case 3: return batch_occupancy<8, 8, 4>();
        default: return batch_occupancy<64, 2, 2>();
    }
}

hipError_t launch_lstar_batch(const GBatch& g, const LstarBuffers& b, uint32_t blocks, hipStream_t s, int shape) {
    switch (shape) {
        case 1: launch_batch_shape<32, 4, 2>(g, b, blocks, s); break;
        case 2: launch_batch_shape<16, 4, 4>(g, b, blocks, s); break;
        case 3: launch_batch_shape<8, 8, 4>(g, b, blocks, s); break;
        default: launch_batch_shape<64, 2, 2>(g, b, blocks, s); break;
    }
    return hipGetLastError();
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

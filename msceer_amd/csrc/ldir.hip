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
//  1. ldir_records_kernel, one streaming pass over the .grad (tiles of
//     vertices, their cells staged in LDS by aligned dword loads): for every
//     edge the pair directions of its cofacet quads that lead on to another
//     edge (rec1, 4 bits per cofacet slot), for every quad those of its
//     cofacet hexes that lead on to another quad (rec2).  A walk step then
//     reads one record instead of four scattered .grad bytes.
//  2. Frontier rounds (ldir_round_kernel) walk pass 1 over rec1 and pass 2
//     over rec2; a reached item is claimed in a bitmap (visited1 / visited2,
//     one bit per record, atomic OR), so each item is expanded once.  The
//     .grad bytes are not written during the walks.
//  3. ldir_finish_kernel writes every cell's ldir from the closed form
//     above, one streaming pass over the same tiles (the marks of a quad --
//     marked1, and with which hex it is paired -- computed from the pass-1
//     bits where a cell needs them).
//
// Records and bits are indexed by 3 * (vertex id of the cell's lowest corner)
// + axis, axis = the odd axis of an edge or the even axis of a quad.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "internal.hpp"
#include "gbatch.hpp"
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
    uint32_t X, Y;        // vertex grid
    uint32_t V;           // vertices
    uint32_t nbx, nby;    // 4x4x4 vertex blocks of the record layout along x, y
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

// Records and visited bits are laid out by 4x4x4 blocks of vertices (64
// vertices, 3 entries each: 384 record bytes of rec1, 192 bits): a walk
// step moves to a neighbouring cell, whose entries then mostly share the
// cache lines of the current one.  Slot of vertex (i, j, k):
__device__ __forceinline__ uint32_t vslot(const LGeom& g, uint32_t i, uint32_t j, uint32_t k) {
    return ((((k >> 2) * g.nby + (j >> 2)) * g.nbx + (i >> 2)) << 6) | ((k & 3u) << 4) | ((j & 3u) << 2) | (i & 3u);
}

// Record / bit index of the edge or quad at cell coordinates (x, y, z).
__device__ __forceinline__ uint32_t rec_index(const LGeom& g, uint32_t x, uint32_t y, uint32_t z, uint32_t axis) {
    return 3u * vslot(g, x >> 1, y >> 1, z >> 1) + axis;
}

__device__ __forceinline__ bool test_bit(const uint32_t* bits, uint32_t i) { return (bits[i >> 5] >> (i & 31u)) & 1u; }

// Claim bit i; true if this call set it.
__device__ __forceinline__ bool claim_bit(uint32_t* bits, uint32_t i) {
    const uint32_t m = 1u << (i & 31u);
    return !(atomicOr(&bits[i >> 5], m) & m);
}

// Pair direction code d (0..5 = +x,-x,+y,-y,+z,-z) applied to coordinates.
// (selects rather than branches on the axis, so the coordinates stay in
// registers when the axis is known only at run time)
__device__ __forceinline__ void apply_dir(uint32_t d, uint32_t& x, uint32_t& y, uint32_t& z) {
    const uint32_t a = d >> 1;
    const uint32_t s = (d & 1u) ? 0xFFFFFFFFu : 1u;  // -1 or +1 (wrapping add)
    x += a == 0u ? s : 0u;
    y += a == 1u ? s : 0u;
    z += a == 2u ? s : 0u;
}

__device__ __forceinline__ void axis_step(uint32_t a, int sgn, uint32_t& x, uint32_t& y, uint32_t& z) {
    const uint32_t s = sgn > 0 ? 1u : 0xFFFFFFFFu;
    x += a == 0u ? s : 0u;
    y += a == 1u ? s : 0u;
    z += a == 2u ? s : 0u;
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

// ---- LDS staging of .grad boxes and bitmap rows ----------------------------
//
// The streaming passes read a box of .grad cells around a tile of vertices.
// Rows of the box are fetched with aligned dword loads (the .grad rows start
// at any byte: Mx is odd) into LDS rows that keep the global alignment: row r
// holds the dwords from the one containing the box's first cell, which lies
// sh[r] bytes into the LDS row.  Only bytes inside the cell grid are
// fetched; the rest of an LDS row is never read.  A dword that is not wholly
// inside the .grad allocation's [0, C) (at most one at each end) is fetched
// byte by byte.
template <int kRX, int kRY, int kRZ>
struct CellBox {
    static constexpr int kRows = kRY * kRZ;
    static constexpr int kDw = (kRX + 6) / 4;  // dwords one row can span
    static constexpr int kPitch = 4 * kDw;      // LDS bytes per row
};

struct BoxRow {
    int32_t dbase;   // aligned dword holding the box's first cell of the row
    int32_t dfirst;  // first and last aligned dwords with cells of the grid
    int32_t dlast;   // (dlast < dfirst: the row lies outside the grid)
    int32_t off;     // LDS byte offset of the box's first cell: r * kPitch + sh
    uint32_t lo;     // .grad bytes [lo, hi) of the row's cells from the box's first one
    uint32_t hi;     // up to kWX of them (the cells a finish tile writes)
    int32_t wfirst;  // aligned dwords wholly inside [lo, hi)
    int32_t wlast;
};

// Aligned dword index of .grad byte i: (align + i) >> 2, align = address & 3.
// kWX: the cells of a row a finish tile writes, from the box's first one.
template <int kRX, int kRY, int kRZ, int kWX = 0>
__device__ __forceinline__ void box_rows(const LGeom& g, uint32_t align, int cx0, int cy0, int cz0, BoxRow* rows) {
    using Bx = CellBox<kRX, kRY, kRZ>;
    for (int r = threadIdx.x; r < Bx::kRows; r += blockDim.x) {
        const int cy = cy0 + r % kRY, cz = cz0 + r / kRY;
        BoxRow b;
        b.dfirst = 1;
        b.dlast = 0;
        b.dbase = 0;
        b.off = r * Bx::kPitch;
        b.lo = b.hi = 0u;
        b.wfirst = 1;
        b.wlast = 0;
        if (cy >= 0 && cz >= 0 && (uint32_t)cy < g.My && (uint32_t)cz < g.Mz) {
            const int64_t rs = (int64_t)g.Mx * ((int64_t)cy + (int64_t)g.My * cz);
            const int64_t gs = (int64_t)align + rs + cx0;  // aligned position of the box's first cell (maybe -1 + rs)
            const int64_t lo = (int64_t)align + rs + (cx0 > 0 ? cx0 : 0);
            const int64_t hi = (int64_t)align + rs + (cx0 + kRX < (int)g.Mx ? cx0 + kRX : (int)g.Mx);
            b.dbase = (int32_t)(gs >> 2);
            b.dfirst = (int32_t)(lo >> 2);
            b.dlast = (int32_t)((hi - 1) >> 2);
            b.off += (int32_t)(gs - 4 * (int64_t)b.dbase);
            if (kWX > 0 && cx0 >= 0) {
                const int64_t wl = rs + cx0, wh = rs + (cx0 + kWX < (int)g.Mx ? cx0 + kWX : (int)g.Mx);
                b.lo = (uint32_t)wl;
                b.hi = (uint32_t)wh;
                b.wfirst = (int32_t)((wl + align + 3) >> 2);
                b.wlast = (int32_t)((wh + align) >> 2) - 1;
            }
        }
        rows[r] = b;
    }
}

// jf, jl: the aligned dwords wholly inside the .grad allocation's [0, C).
template <int kRX, int kRY, int kRZ>
__device__ __forceinline__ void box_load(const uint8_t* __restrict__ grad, uint64_t cells, uint32_t align,
                                         const BoxRow* rows, uint32_t* stage) {
    using Bx = CellBox<kRX, kRY, kRZ>;
    const uint32_t* g32 = reinterpret_cast<const uint32_t*>(grad - align);
    const int32_t jf = (int32_t)((align + 3u) >> 2), jl = (int32_t)(((uint64_t)align + cells) >> 2) - 1;
    for (int d = threadIdx.x; d < Bx::kRows * Bx::kDw; d += blockDim.x) {
        const int r = d / Bx::kDw, k = d - r * Bx::kDw;
        const int32_t j = rows[r].dfirst + k;
        if (j > rows[r].dlast) continue;
        uint32_t w;
        if (j >= jf && j <= jl) {
            w = g32[j];
        } else {  // the allocation's first or last dword: byte by byte
            w = 0;
            const int64_t first = 4 * (int64_t)j - (int64_t)align;  // .grad index of the dword's first byte
            for (int t = 0; t < 4; t++)
                if (first + t >= 0 && first + t < (int64_t)cells) w |= (uint32_t)grad[first + t] << (8 * t);
        }
        stage[(r * Bx::kPitch) / 4 + (j - rows[r].dbase)] = w;
    }
}

// The visited bits of the 4x4x4 vertex blocks around a tile (kBX x kBY x
// kBZ blocks from block (bx0, by0, bz0)): 6 words a block, as in the bitmap
// (a block's 192 bits are word-aligned there), one more word of padding.
template <int kBX, int kBY, int kBZ>
__device__ __forceinline__ void bits_load(const LGeom& g, const uint32_t* __restrict__ bits, uint32_t nwords,
                                          uint32_t bx0, uint32_t by0, uint32_t bz0, uint32_t* lds) {
    constexpr int kN = kBX * kBY * kBZ * 6;
    const uint32_t nbz = (((g.Mz + 1) >> 1) + 3u) >> 2;
    for (int d = threadIdx.x; d <= kN; d += blockDim.x) {
        uint32_t v = 0;
        if (d < kN) {
            const int blk = d / 6, w = d - blk * 6;
            const uint32_t bx = bx0 + (uint32_t)(blk % kBX), by = by0 + (uint32_t)((blk / kBX) % kBY);
            const uint32_t bz = bz0 + (uint32_t)(blk / (kBX * kBY));
            if (bx < g.nbx && by < g.nby && bz < nbz) {
                const uint32_t i = 6u * ((bz * g.nby + by) * g.nbx + bx) + (uint32_t)w;
                if (i < nwords) v = bits[i];
            }
        }
        lds[d] = v;
    }
}

// 3 bits of vertex (i, j, k) from the staged blocks (its block kBX x kBY x
// kBZ-relative coordinates rb*).
template <int kBX, int kBY>
__device__ __forceinline__ uint32_t bits3_lds(const uint32_t* lds, uint32_t i, uint32_t j, uint32_t k, uint32_t rbx,
                                              uint32_t rby, uint32_t rbz) {
    const uint32_t inner = ((k & 3u) << 4) | ((j & 3u) << 2) | (i & 3u);
    const uint32_t bit = 3u * inner;  // < 192: the 3 bits stay inside the block's 6 words
    const uint32_t* p = lds + 6u * ((rbz * kBY + rby) * kBX + rbx) + (bit >> 5);
    const uint64_t w = ((uint64_t)p[1] << 32) | p[0];
    return (uint32_t)(w >> (bit & 31u)) & 7u;
}

// ---- phase 1: records ------------------------------------------------------


// Record slot of the cofacet (byte b) reached from the item along axis a
// with sign sgn; the item is an edge with odd axis o (kO < 3), or a quad
// (kO == 3) whose pair partners are always quads.  8 | pair direction when
// the walk goes on through it: assigned, not critical, the pair a facet
// (for an edge item: along o or a) and not the item itself.
// The pair directions that qualify form a constant set per slot: for an
// edge item the two directions of axes o and a, for a quad item all six,
// less the one pointing straight back along a (code 7, critical, never; nor
// 6, which no .grad byte holds).
template <int kO, int kA, int kSgn>
__device__ __forceinline__ uint32_t slot_code(uint32_t b) {
    constexpr uint32_t kBack = 2u * (uint32_t)kA + (kSgn > 0 ? 1u : 0u);
    constexpr uint32_t kOk = (kO < 3 ? ((3u << (2 * kO)) | (3u << (2 * kA))) : 0x3Fu) & ~(1u << kBack);
    const uint32_t d = (b >> 2) & 7u;
    return ((b & 1u) && ((kOk >> d) & 1u)) ? (8u | d) : 0u;
}

// Vertex tile of the streaming passes: 64 x 4 x 4 vertices (16 x 1 x 1 of the
// 4x4x4 blocks of the record layout).
constexpr int kTX = 64, kTY = 4, kTZ = 4;
static_assert(kTX * kTY == kThreads, "one lane per tile column");
static_assert(kTX % 4 == 0 && kTY % 4 == 0 && kTZ == 4, "tiles of whole 4x4x4 vertex blocks");
using RecBox = CellBox<2 * kTX + 1, 2 * kTY + 1, 2 * kTZ + 1>;  // cells 2 x0 - 1 .. 2 (x0 + kTX) - 1
using FinBox = CellBox<2 * kTX + 1, 2 * kTY + 1, 2 * kTZ + 1>;  // cells 2 x0 .. 2 (x0 + kTX)

// Records of every vertex (i, j, k) of a tile: its three "+" edges and three
// "+" quads.  rec1[3v + o] (edge with odd axis o): slot 2m + (sgn < 0) is the
// cofacet quad along the edge's m-th even axis; rec2[3v + a] (quad with even
// axis a): slot (sgn < 0) is the cofacet hex along a.  The cofacets lie
// within one cell of the vertex's cell, all inside the tile's staged box.
// Arguments of the streaming passes: one volume, or (kGroup) the tiles of
// every volume of a group, a work-group per tile.
struct TileArgs {
    uint8_t* grad;
    LGeom g;
    uint64_t cells;
    int ntx, nty;
    uint16_t* rec1;
    uint8_t* rec2;
    const uint32_t* v1;
    const uint32_t* v2;
    uint32_t* bits;      // records: bitmaps to clear (both, or every volume's of a group)
    uint32_t bit_words;  // records: words to clear; finish: words of one bitmap
    GBatch batch;
};

// The volume of the work-group's tile: from the arguments, or from the
// group's volume table.
template <bool kGroup>
__device__ __forceinline__ int tile_of(TileArgs& a, uint32_t& vol) {
    vol = 0;
    if (!kGroup) return (int)blockIdx.x;
    const uint32_t unit = blockIdx.x;
    vol = unit_volume(a.batch.tiles, unit);
    const GVol& gv = a.batch.vols[vol];
    a.grad = gv.grad;
    a.g = LGeom{gv.Mx, gv.My, gv.Mz, gv.Mx * gv.My, (uint32_t)gv.X, (uint32_t)gv.Y,
                (uint32_t)gv.X * (uint32_t)gv.Y * (uint32_t)gv.Z, ((uint32_t)gv.X + 3u) / 4u, ((uint32_t)gv.Y + 3u) / 4u};
    a.cells = gv.cells;
    a.ntx = (int)gv.ntx;
    a.nty = (int)gv.nty;
    a.rec1 = gv.rec1;
    a.rec2 = gv.rec2;
    a.v1 = gv.bits1;
    a.v2 = gv.bits2;
    a.bit_words = gv.bit_words;
    return (int)(unit - a.batch.tiles.first[vol]);
}

template <bool kGroup>
__global__ void __launch_bounds__(kThreads) ldir_records_kernel(TileArgs a) {
    constexpr int kRY = 2 * kTY + 1;
    __shared__ uint32_t stage[RecBox::kRows * RecBox::kDw];
    __shared__ BoxRow rows[RecBox::kRows];
    for (uint32_t w = blockIdx.x * kThreads + threadIdx.x; w < a.bit_words; w += gridDim.x * kThreads) a.bits[w] = 0u;
    uint32_t vol;
    const int tile = tile_of<kGroup>(a, vol);
    const uint8_t* __restrict__ grad = a.grad;
    const LGeom g = a.g;
    const uint64_t cells = a.cells;
    const uint32_t align = (uint32_t)((uintptr_t)grad & 3u);
    const int ntx = a.ntx, nty = a.nty;
    uint16_t* __restrict__ rec1 = a.rec1;
    uint8_t* __restrict__ rec2 = a.rec2;
    const int x0 = (tile % ntx) * kTX, y0 = ((tile / ntx) % nty) * kTY, z0 = (tile / (ntx * nty)) * kTZ;
    box_rows<2 * kTX + 1, kRY, 2 * kTZ + 1>(g, align, 2 * x0 - 1, 2 * y0 - 1, 2 * z0 - 1, rows);
    __syncthreads();
    box_load<2 * kTX + 1, kRY, 2 * kTZ + 1>(grad, cells, align, rows, stage);
    __syncthreads();
    const uint8_t* st8 = reinterpret_cast<const uint8_t*>(stage);
    const uint32_t Zv = (g.Mz + 1) >> 1;
    // a wave per 4x4x4 vertex block of the tile (8 x 2 x 1 blocks), so that
    // its records are written as one contiguous run
    for (int it = 0; it < (kTX / 4) * (kTY / 4) * (kTZ / 4) / (kThreads / 64); it++) {
        const int blk = it * (kThreads / 64) + (int)(threadIdx.x >> 6), inner = (int)(threadIdx.x & 63u);
        const int lx = (blk % (kTX / 4)) * 4 + (inner & 3), ly = (blk / (kTX / 4)) * 4 + ((inner >> 2) & 3);
        const int lz = inner >> 4;
        const uint32_t i = (uint32_t)(x0 + lx), j = (uint32_t)(y0 + ly), k = (uint32_t)(z0 + lz);
        if (i >= g.X || j >= g.Y || k >= Zv) continue;
        const uint32_t v = vslot(g, i, j, k);
        // byte of the cell at offset (dx, dy, dz) from the vertex's cell
        auto B = [&](int dx, int dy, int dz) -> uint32_t {
            return st8[rows[(2 * ly + 1 + dy) + kRY * (2 * lz + 1 + dz)].off + (2 * lx + 1 + dx)];
        };
        const bool px = i + 1 < g.X, py = j + 1 < g.Y, pz = k + 1 < Zv;  // "+" neighbours exist
        const bool mx = i > 0, my = j > 0, mz = k > 0;                     // "-" neighbours exist
        if (px) {  // x-edge: cofacets along y, z
            uint32_t r = 0;
            if (py) r |= slot_code<0, 1, 1>(B(1, 1, 0));
            if (my) r |= slot_code<0, 1, -1>(B(1, -1, 0)) << 4;
            if (pz) r |= slot_code<0, 2, 1>(B(1, 0, 1)) << 8;
            if (mz) r |= slot_code<0, 2, -1>(B(1, 0, -1)) << 12;
            rec1[3u * v] = (uint16_t)r;
        }
        if (py) {  // y-edge: cofacets along x, z
            uint32_t r = 0;
            if (px) r |= slot_code<1, 0, 1>(B(1, 1, 0));
            if (mx) r |= slot_code<1, 0, -1>(B(-1, 1, 0)) << 4;
            if (pz) r |= slot_code<1, 2, 1>(B(0, 1, 1)) << 8;
            if (mz) r |= slot_code<1, 2, -1>(B(0, 1, -1)) << 12;
            rec1[3u * v + 1u] = (uint16_t)r;
        }
        if (pz) {  // z-edge: cofacets along x, y
            uint32_t r = 0;
            if (px) r |= slot_code<2, 0, 1>(B(1, 0, 1));
            if (mx) r |= slot_code<2, 0, -1>(B(-1, 0, 1)) << 4;
            if (py) r |= slot_code<2, 1, 1>(B(0, 1, 1)) << 8;
            if (my) r |= slot_code<2, 1, -1>(B(0, -1, 1)) << 12;
            rec1[3u * v + 2u] = (uint16_t)r;
        }
        if (py && pz) {  // quad with even axis x: cofacet hexes along x
            uint32_t r = 0;
            if (px) r |= slot_code<3, 0, 1>(B(1, 1, 1));
            if (mx) r |= slot_code<3, 0, -1>(B(-1, 1, 1)) << 4;
            rec2[3u * v] = (uint8_t)r;
        }
        if (px && pz) {  // even axis y
            uint32_t r = 0;
            if (py) r |= slot_code<3, 1, 1>(B(1, 1, 1));
            if (my) r |= slot_code<3, 1, -1>(B(1, -1, 1)) << 4;
            rec2[3u * v + 1u] = (uint8_t)r;
        }
        if (px && py) {  // even axis z
            uint32_t r = 0;
            if (pz) r |= slot_code<3, 2, 1>(B(1, 1, 1));
            if (mz) r |= slot_code<3, 2, -1>(B(1, 1, -1)) << 4;
            rec2[3u * v + 2u] = (uint8_t)r;
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

constexpr uint32_t kNoItem = 0xFFFFFFFFu;

// Claim the bits of the (up to 4) items reached in one step from `item`
// (cell id): out[slot] is the cell id of the item reached through record
// slot `slot` when this call claimed it, else kNoItem.  `start`: the item is
// a critical cell of the start list, claimed here first (skipped if already
// reached).  The claims of the slots are independent atomics, issued
// together.
// Coordinates, axis (an edge item: its odd axis; a quad item: its even axis)
// and record index of an item.
__device__ __forceinline__ uint32_t item_record(const LGeom& g, bool edges, uint32_t item, uint32_t& x, uint32_t& y,
                                                uint32_t& z, uint32_t& o) {
    cell_coords(g, item, x, y, z);
    o = edges ? ((x & 1u) ? 0u : ((y & 1u) ? 1u : 2u)) : (!(x & 1u) ? 0u : (!(y & 1u) ? 1u : 2u));
    return rec_index(g, x, y, z, o);
}

// Claim an item's bit; true if this call set it.
__device__ __forceinline__ bool claim_item(const LGeom& g, const Walk& w, uint32_t item) {
    uint32_t x, y, z, o;
    return claim_bit(w.bits, item_record(g, w.pass == 1, item, x, y, z, o));
}

__device__ __forceinline__ void expand(const LGeom& g, const Walk& w, uint32_t item, bool start, uint32_t out[4]) {
#pragma unroll
    for (int s = 0; s < 4; s++) out[s] = kNoItem;
    uint32_t x, y, z, o;
    const bool edges = w.pass == 1;
    const uint32_t ri = item_record(g, edges, item, x, y, z, o);
    if (start) {
        const uint32_t m = 1u << (ri & 31u);
        if (atomicOr(&w.bits[ri >> 5], m) & m) return;
    }
    const uint32_t r = edges ? (uint32_t)w.rec1[ri] : (uint32_t)w.rec2[ri];
    if (!r) return;
    uint32_t a1, a2;
    other_axes(o, a1, a2);
    uint32_t cand[4], word[4], mask[4];
#pragma unroll
    for (int s = 0; s < 4; s++) {
        const uint32_t f = (r >> (4 * s)) & 15u;
        cand[s] = kNoItem;
        if (!(f & 8u) || (!edges && s >= 2)) continue;
        uint32_t px = x, py = y, pz = z;
        axis_step(edges ? (s < 2 ? a1 : a2) : o, (s & 1) ? -1 : 1, px, py, pz);
        apply_dir(f & 7u, px, py, pz);
        // the item reached: an edge (its odd axis) or a quad (its even axis)
        const uint32_t pa = edges ? ((px & 1u) ? 0u : ((py & 1u) ? 1u : 2u))
                                  : (!(px & 1u) ? 0u : (!(py & 1u) ? 1u : 2u));
        const uint32_t pi = rec_index(g, px, py, pz, pa);
        cand[s] = cell_id(g, px, py, pz);
        word[s] = pi >> 5;
        mask[s] = 1u << (pi & 31u);
    }
    uint32_t old[4];
#pragma unroll
    for (int s = 0; s < 4; s++) old[s] = cand[s] != kNoItem ? atomicOr(&w.bits[word[s]], mask[s]) : 0u;
#pragma unroll
    for (int s = 0; s < 4; s++)
        if (cand[s] != kNoItem && !(old[s] & mask[s])) out[s] = cand[s];
}

// One round.  A block takes an even share of the incoming frontier into an
// LDS pool and works it breadth-first: each step, every lane pops up to
// kPerLane items, and the items found behind them go back into the pool, so
// a large manifold reached from one start item is spread over the whole
// block.  After kMaxSteps steps whatever is left goes to the outgoing
// frontier, which the next round shares out again.  Counts are read on the
// device, so rounds are enqueued back to back; `clear` (the counts the next
// round appends to) is zeroed here.
// Where the items of a round live: one volume (items are its cell ids), or
// a group of volumes (items are cell ids from each volume's first item on).
struct OneSpace {
    LGeom g;
    Walk w;
    __device__ __forceinline__ void expand_item(uint32_t item, bool start, uint32_t out[4]) const {
        expand(g, w, item, start, out);
    }
    __device__ __forceinline__ bool claim(uint32_t item) const { return claim_item(g, w, item); }
};

struct GroupSpace {
    GBatch b;
    int pass;
    // the item's volume: its geometry, walk and first item
    __device__ __forceinline__ uint32_t locate(uint32_t item, LGeom& g, Walk& w) const {
        const uint32_t vol = unit_volume(b.items, item);
        const GVol& gv = b.vols[vol];
        g = LGeom{gv.Mx, gv.My, gv.Mz, gv.Mx * gv.My, (uint32_t)gv.X, (uint32_t)gv.Y,
                  (uint32_t)gv.X * (uint32_t)gv.Y * (uint32_t)gv.Z, ((uint32_t)gv.X + 3u) / 4u,
                  ((uint32_t)gv.Y + 3u) / 4u};
        w = Walk{gv.rec1, gv.rec2, pass == 1 ? gv.bits1 : gv.bits2, pass};
        return b.items.first[vol];
    }
    __device__ __forceinline__ void expand_item(uint32_t item, bool start, uint32_t out[4]) const {
        LGeom g;
        Walk w;
        const uint32_t base = locate(item, g, w);
        expand(g, w, item - base, start, out);
#pragma unroll
        for (int s = 0; s < 4; s++)
            if (out[s] != kNoItem) out[s] += base;
    }
    __device__ __forceinline__ bool claim(uint32_t item) const {
        LGeom g;
        Walk w;
        const uint32_t base = locate(item, g, w);
        return claim_item(g, w, item - base);
    }
};

template <class Space>
__global__ void __launch_bounds__(kThreads)
ldir_round_kernel(Space sp, MList in, MList out, uint32_t* __restrict__ clear, uint32_t* __restrict__ overflow,
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
        uint32_t nb[kPerLane][4];
#pragma unroll
        for (int k = 0; k < kPerLane; k++) {
            if (items[k] == kNoItem) {
#pragma unroll
                for (int j = 0; j < 4; j++) nb[k][j] = kNoItem;
            } else {
                sp.expand_item(items[k], from_share[k], nb[k]);
            }
        }
        // one pool reservation per wave for everything its lanes reached
        uint32_t cnt = 0;
#pragma unroll
        for (int k = 0; k < kPerLane; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) cnt += nb[k][j] != kNoItem ? 1u : 0u;
        const uint32_t lane = tid & 63u;
        uint32_t incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o);
            if (lane >= (uint32_t)o) incl += t;
        }
        const uint32_t wave_total = __shfl(incl, 63);
        uint32_t base = 0;
        if (lane == 63 && wave_total) base = atomicAdd(&s_n, wave_total);
        base = __shfl(base, 63) + incl - cnt;
#pragma unroll
        for (int k = 0; k < kPerLane; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                if (nb[k][j] == kNoItem) continue;
                if (base < (uint32_t)kPool) {
                    pool[base] = nb[k][j];
                } else {
                    if (!mlist_push(out, sub, nb[k][j])) atomicOr(overflow, 1u);
                }
                base++;
            }
        __syncthreads();
    }
    // hand on what is left: the pool (claimed items), then the unread part of
    // the share -- in a start round claimed here first, as their expansion
    // would have claimed them, and dropped if already reached (the rounds
    // after a start round hold claimed items only)
    __syncthreads();
    const uint32_t n = min(s_n, (uint32_t)kPool);
    mlist_flush_block(out, sub, pool, n, &s_base, overflow);
    for (uint32_t nx = s_next; nx < hi; nx += kPool) {
        const uint32_t take = min(hi - nx, (uint32_t)kPool);
        __syncthreads();
        if (!start) {
            for (uint32_t j = tid; j < take; j += kThreads) pool[j] = mlist_at(in, prefix, nx + j);
            __syncthreads();
            mlist_flush_block(out, sub, pool, take, &s_base, overflow);
        } else {
            if (tid == 0) s_n = 0;
            __syncthreads();
            for (uint32_t j = tid; j < take; j += kThreads) {
                const uint32_t item = mlist_at(in, prefix, nx + j);
                if (sp.claim(item)) pool[atomicAdd(&s_n, 1u)] = item;
            }
            __syncthreads();
            mlist_flush_block(out, sub, pool, s_n, &s_base, overflow);
        }
    }
}

// Frontier from the records and bits alone (after any overflow): the
// unreached critical cells of the pass (claimed here), and every unreached
// cell one step from a reached one (claimed here).
__global__ void __launch_bounds__(kThreads)
ldir_sweep_kernel(const uint8_t* __restrict__ grad, LGeom g, Walk w, MList out, uint32_t* __restrict__ overflow,
                  uint32_t item0) {
    __shared__ uint32_t stage[kStage];
    __shared__ uint32_t s_n;
    __shared__ MSlot s_base;
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
    const int sub = (int)(blockIdx.x % kSubs);
    auto push = [&](uint32_t item) {
        const uint32_t slot = atomicAdd(&s_n, 1u);
        if (slot < (uint32_t)kStage) stage[slot] = item0 + item;
        else if (!mlist_push(out, sub, item0 + item)) atomicOr(overflow, 1u);
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
                expand(g, w, c, false, nb);
#pragma unroll
                for (int t = 0; t < 4; t++)
                    if (nb[t] != kNoItem) push(nb[t]);
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

// Per vertex v, the marks of its three "+" quads (quad a: even axis a) that
// the finish pass needs from its neighbours: bit a = marked1 (assigned, not
// critical, a facet edge other than its pair reached in pass 1); bit 3 + a =
// marked1 and paired with its "+" hex along a; bit 6 + a = marked1 and paired
// with its "-" hex.  The facet edges of quad a lie at v (odd axis u) and at
// v + e_u' (odd axis u) for its two odd axes u, u'.
template <int kA>
__device__ __forceinline__ uint32_t quad_marks(uint32_t b, uint32_t e_v, uint32_t e_n1, uint32_t e_n2) {
    // odd axes u1 < u2 of the quad
    constexpr int u1 = kA == 0 ? 1 : 0, u2 = kA == 2 ? 1 : 2;
    const uint32_t d = (b >> 2) & 7u;
    if (!(b & 1u) || d == 7u) return 0u;
    // facets: -u1: edge of v with odd axis u2; +u1: edge of v + e_u1, odd axis u2;
    //         -u2: edge of v with odd axis u1; +u2: edge of v + e_u2, odd axis u1
    bool m = false;
    if (d != 2u * u1 + 1u && ((e_v >> u2) & 1u)) m = true;
    if (d != 2u * u1 && ((e_n1 >> u2) & 1u)) m = true;
    if (d != 2u * u2 + 1u && ((e_v >> u1) & 1u)) m = true;
    if (d != 2u * u2 && ((e_n2 >> u1) & 1u)) m = true;
    if (!m) return 0u;
    return (1u << kA) | (d == 2u * kA ? 1u << (3 + kA) : 0u) | (d == 2u * kA + 1u ? 1u << (6 + kA) : 0u);
}

__device__ __forceinline__ uint8_t with_ldir(uint32_t b, uint32_t l) { return (uint8_t)((b & 0x1Fu) | (l << 5)); }

// The ldir bits of every cell of the tile's "+" octants (cells 2 x0 .. 2 (x0 +
// kTX) - 1 along x, likewise y and z), written back row by row: aligned
// dwords wholly inside the tile's cells as dwords, the rest of a row as
// bytes.  The box staged holds one more cell layer on each "+" side (the
// facet quads of the tile's last hexes) and the bitmaps one more vertex
// layer; the marks of a quad (marked1, and with which hex it is paired) are
// computed where needed.  The cells of the next tiles that are read here are
// only read for their pair and assigned bits, which no tile changes.
template <bool kGroup>
__global__ void __launch_bounds__(kThreads) ldir_finish_kernel(TileArgs a) {
    constexpr int kRX = 2 * kTX + 1, kRY = 2 * kTY + 1, kRZ = 2 * kTZ + 1;
    // vertex blocks of the tile and its "+" halo layer
    constexpr int kBX = kTX / 4 + 1, kBY = kTY / 4 + 1, kBZ = kTZ / 4 + 1, kBW = kBX * kBY * kBZ * 6 + 1;
    __shared__ uint32_t stage[FinBox::kRows * FinBox::kDw];
    __shared__ BoxRow rows[FinBox::kRows];
    // per vertex of the tile and its halo: visited1 bits | visited2 bits << 3
    constexpr int kVX = kTX + 1, kVY = kTY + 1, kVZ = kTZ + 1;
    __shared__ uint32_t b1[kBW], b2[kBW];
    __shared__ uint8_t vbits[kVX * kVY * kVZ];
    uint32_t vol;
    const int tile = tile_of<kGroup>(a, vol);
    uint8_t* __restrict__ grad = a.grad;
    const LGeom g = a.g;
    const uint64_t cells = a.cells;
    const uint32_t align = (uint32_t)((uintptr_t)grad & 3u);
    const int ntx = a.ntx, nty = a.nty;
    const uint32_t* __restrict__ v1 = a.v1;
    const uint32_t* __restrict__ v2 = a.v2;
    const uint32_t nwords = a.bit_words;
    const int x0 = (tile % ntx) * kTX, y0 = ((tile / ntx) % nty) * kTY, z0 = (tile / (ntx * nty)) * kTZ;
    const uint32_t bx0 = (uint32_t)x0 >> 2, by0 = (uint32_t)y0 >> 2, bz0 = (uint32_t)z0 >> 2;
    box_rows<kRX, kRY, kRZ, 2 * kTX>(g, align, 2 * x0, 2 * y0, 2 * z0, rows);
    bits_load<kBX, kBY, kBZ>(g, v1, nwords, bx0, by0, bz0, b1);
    bits_load<kBX, kBY, kBZ>(g, v2, nwords, bx0, by0, bz0, b2);
    __syncthreads();
    box_load<kRX, kRY, kRZ>(grad, cells, align, rows, stage);
    for (int t = threadIdx.x; t < kVX * kVY * kVZ; t += kThreads) {
        const uint32_t a = (uint32_t)(x0 + t % kVX), b = (uint32_t)(y0 + (t / kVX) % kVY);
        const uint32_t c = (uint32_t)(z0 + t / (kVX * kVY));
        const uint32_t rbx = (a >> 2) - bx0, rby = (b >> 2) - by0, rbz = (c >> 2) - bz0;
        vbits[t] = (uint8_t)(bits3_lds<kBX, kBY>(b1, a, b, c, rbx, rby, rbz) |
                             (bits3_lds<kBX, kBY>(b2, a, b, c, rbx, rby, rbz) << 3));
    }
    __syncthreads();
    uint8_t* st8 = reinterpret_cast<uint8_t*>(stage);
    const uint32_t Zv = (g.Mz + 1) >> 1;
    const int lx = (int)(threadIdx.x % kTX), ly = (int)(threadIdx.x / kTX);
    const uint32_t i = (uint32_t)(x0 + lx), j = (uint32_t)(y0 + ly);
    for (int lz = 0; lz < kTZ; lz++) {
        const uint32_t k = (uint32_t)(z0 + lz);
        if (i >= g.X || j >= g.Y || k >= Zv) break;
        // cell at offset (dx, dy, dz) from the cell of vertex (lx + ex, ly + ey, lz + ez)
        auto at = [&](int dx, int dy, int dz) -> uint8_t& {
            return st8[rows[(2 * ly + dy) + kRY * (2 * lz + dz)].off + (2 * lx + dx)];
        };
        // bits of the vertex (ex, ey, ez) vertices further on
        auto e1 = [&](int ex, int ey, int ez) -> uint32_t {
            return vbits[(lx + ex) + kVX * ((ly + ey) + kVY * (lz + ez))] & 7u;
        };
        auto e2 = [&](int ex, int ey, int ez) -> uint32_t {
            return (uint32_t)vbits[(lx + ex) + kVX * ((ly + ey) + kVY * (lz + ez))] >> 3;
        };
        const bool px = i + 1 < g.X, py = j + 1 < g.Y, pz = k + 1 < Zv;
        const uint32_t e_v = e1(0, 0, 0), q_v = e2(0, 0, 0);
        const uint32_t e_x = px ? e1(1, 0, 0) : 0u, e_y = py ? e1(0, 1, 0) : 0u, e_z = pz ? e1(0, 0, 1) : 0u;
        uint32_t m_v = 0;
        if (py && pz) m_v |= quad_marks<0>(at(0, 1, 1), e_v, e_y, e_z);
        if (px && pz) m_v |= quad_marks<1>(at(1, 0, 1), e_v, e_x, e_z);
        if (px && py) m_v |= quad_marks<2>(at(1, 1, 0), e_v, e_x, e_y);
        if (px) { uint8_t& c = at(1, 0, 0); c = with_ldir(c, (e_v & 1u) ? 2u : 3u); }
        if (py) { uint8_t& c = at(0, 1, 0); c = with_ldir(c, (e_v & 2u) ? 2u : 3u); }
        if (pz) { uint8_t& c = at(0, 0, 1); c = with_ldir(c, (e_v & 4u) ? 2u : 3u); }
        // quads: reached in pass 2 -> 1; marked in pass 1 -> 2
        if (py && pz) { uint8_t& c = at(0, 1, 1); c = with_ldir(c, (q_v & 1u) ? 1u : ((m_v & 1u) ? 2u : 3u)); }
        if (px && pz) { uint8_t& c = at(1, 0, 1); c = with_ldir(c, (q_v & 2u) ? 1u : ((m_v & 2u) ? 2u : 3u)); }
        if (px && py) { uint8_t& c = at(1, 1, 0); c = with_ldir(c, (q_v & 4u) ? 1u : ((m_v & 4u) ? 2u : 3u)); }
        if (px && py && pz) {
            uint8_t& c = at(1, 1, 1);
            const uint32_t b = c;
            const uint32_t d = (b >> 2) & 7u;
            if (d != 7u) {
                uint32_t l = 3u;
                if (b & 1u) {
                    // facet quads: -x, +x, -y, +y, -z, +z; the pair excluded
                    const uint32_t q_x = e2(1, 0, 0), q_y = e2(0, 1, 0), q_z = e2(0, 0, 1);
                    const bool f[6] = {(q_x & 1u) != 0u, (q_v & 1u) != 0u, (q_y & 2u) != 0u,
                                       (q_v & 2u) != 0u, (q_z & 4u) != 0u, (q_v & 4u) != 0u};
                    bool m2 = false;
#pragma unroll
                    for (uint32_t t = 0; t < 6; t++)
                        if (t != d && f[t]) m2 = true;
                    if (m2) {
                        l = 1u;
                    } else if (d & 1u) {
                        // the pair is the vertex's own quad: marked in pass 1 and paired with its "+" hex
                        if ((m_v >> (3u + (d >> 1))) & 1u) l = 2u;
                    } else {
                        // the pair is the next vertex's quad: marked in pass 1 and paired with its "-" hex
                        uint32_t mq = 0;
                        if (d == 0u) mq = quad_marks<0>(at(2, 1, 1), e_x, e1(1, 1, 0), e1(1, 0, 1));
                        else if (d == 2u) mq = quad_marks<1>(at(1, 2, 1), e_y, e1(1, 1, 0), e1(0, 1, 1));
                        else mq = quad_marks<2>(at(1, 1, 2), e_z, e1(1, 0, 1), e1(0, 1, 1));
                        if ((mq >> (6u + (d >> 1))) & 1u) l = 2u;
                    }
                }
                c = with_ldir(b, l);
            }
        }
    }
    __syncthreads();
    // write back the rows of the tile's cells
    constexpr int kWY = 2 * kTY, kWZ = 2 * kTZ;  // rows written: cells 2 y0 .. 2 (y0 + kTY) - 1, likewise z
    uint32_t* g32 = reinterpret_cast<uint32_t*>(grad - align);
    for (int d = threadIdx.x; d < kWY * kWZ * FinBox::kDw; d += kThreads) {
        const int wr = d / FinBox::kDw, kk = d - wr * FinBox::kDw;
        const int r = (wr % kWY) + kRY * (wr / kWY);
        const int32_t jj = rows[r].dfirst + kk;
        if (jj > rows[r].dlast) continue;
        const uint32_t w = stage[(r * FinBox::kPitch) / 4 + (jj - rows[r].dbase)];
        if (jj >= rows[r].wfirst && jj <= rows[r].wlast) {
            g32[jj] = w;
        } else {  // a dword at an end of the tile's cells of the row: its own bytes only
            const uint32_t lo = rows[r].lo, hi = rows[r].hi;
            const int64_t first = 4 * (int64_t)jj - (int64_t)align;
            for (int t = 0; t < 4; t++)
                if (first + t >= (int64_t)lo && first + t < (int64_t)hi) grad[first + t] = (uint8_t)(w >> (8 * t));
        }
    }
    if (kGroup && a.batch.times) {
        __syncthreads();
        if (threadIdx.x == 0) atomicMax(&a.batch.times[2 * vol + 1], batch_clock());
    }
}

__global__ void zero_kernel(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0u;
}

LGeom make_geom(const Dims& d) {
    return LGeom{d.Mx, d.My, d.Mz, d.Mx * d.My, (uint32_t)d.X, (uint32_t)d.Y, (uint32_t)d.verts,
                 (uint32_t)((d.X + 3) / 4), (uint32_t)((d.Y + 3) / 4)};
}

// Entries of the block layout: 3 per vertex slot, 64 slots per 4x4x4 block.
uint64_t layout_entries(const Dims& d) {
    return 3ull * 64ull * (uint64_t)((d.X + 3) / 4) * (uint64_t)((d.Y + 3) / 4) * (uint64_t)((d.Z + 3) / 4);
}

uint32_t grid_of(uint64_t n, uint32_t cap) {
    uint64_t b = (n + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    return (uint32_t)(b > cap ? cap : b);
}

}  // namespace

size_t ldir_scratch_bytes(const Dims& d, LdirScratch* s) {
    const uint64_t n = layout_entries(d);
    const uint64_t words = (n + 31) / 32 + 1;  // one word of padding: bits3 reads a word pair
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

// Tiles of the streaming passes (fewer than 2^31: the volume has fewer than
// 2^32 cells).
static void tiles_of(const Dims& d, int& ntx, int& nty, uint32_t& n) {
    ntx = (d.X + kTX - 1) / kTX;
    nty = (d.Y + kTY - 1) / kTY;
    const int ntz = (d.Z + kTZ - 1) / kTZ;
    n = (uint32_t)ntx * (uint32_t)nty * (uint32_t)ntz;
}

__global__ void __launch_bounds__(256) clear_ldir_words_kernel(uint32_t* __restrict__ w, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) w[i] &= 0x1F1F1F1Fu;
}

__global__ void __launch_bounds__(256) clear_ldir_bytes_kernel(uint8_t* __restrict__ g, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) g[i] &= 0x1Fu;
}

hipError_t launch_clear_ldir_bits(uint8_t* grad, int64_t n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    // 4-byte words where the buffer is aligned, bytes for the rest
    const int64_t lead = (int64_t)((4 - (reinterpret_cast<uintptr_t>(grad) & 3)) & 3);
    const int64_t head = lead < n ? lead : n;
    const int64_t words = (n - head) / 4;
    const int64_t tail = n - head - 4 * words;
    auto blocks = [](int64_t m) { return (unsigned)std::min<int64_t>((m + 255) / 256, 8192); };
    if (head) hipLaunchKernelGGL(clear_ldir_bytes_kernel, dim3(1), dim3(256), 0, st, grad, head);
    if (words)
        hipLaunchKernelGGL(clear_ldir_words_kernel, dim3(blocks(words)), dim3(256), 0, st,
                           reinterpret_cast<uint32_t*>(grad + head), words);
    if (tail) hipLaunchKernelGGL(clear_ldir_bytes_kernel, dim3(1), dim3(256), 0, st, grad + head + 4 * words, tail);
    return hipGetLastError();
}

hipError_t launch_ldir_records(const Dims& d, const uint8_t* grad, void* scratch, const LdirScratch& s,
                               hipStream_t st) {
    uint8_t* base = static_cast<uint8_t*>(scratch);
    int ntx, nty;
    uint32_t n;
    tiles_of(d, ntx, nty, n);
    TileArgs a{};
    a.grad = const_cast<uint8_t*>(grad);  // read only here
    a.g = make_geom(d);
    a.cells = d.cells;
    a.ntx = ntx;
    a.nty = nty;
    a.rec1 = reinterpret_cast<uint16_t*>(base + s.rec1_off);
    a.rec2 = base + s.rec2_off;
    // visited1 and visited2 are adjacent: one clear covers both
    a.bits = reinterpret_cast<uint32_t*>(base + s.bits1_off);
    a.bit_words = (uint32_t)((s.bits2_off - s.bits1_off) / 4) + s.bit_words;
    hipLaunchKernelGGL(ldir_records_kernel<false>, dim3(n), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

static Walk make_walk(void* scratch, const LdirScratch& s, int pass) {
    uint8_t* base = static_cast<uint8_t*>(scratch);
    return Walk{reinterpret_cast<const uint16_t*>(base + s.rec1_off), base + s.rec2_off,
                reinterpret_cast<uint32_t*>(base + (pass == 1 ? s.bits1_off : s.bits2_off)), pass};
}

hipError_t launch_ldir_round(const Dims& d, void* scratch, const LdirScratch& s, int pass, bool start, const MList& in,
                             const MList& out, uint32_t* clear_counts, uint32_t* overflow, hipStream_t st) {
    hipLaunchKernelGGL(ldir_round_kernel<OneSpace>, dim3(kRoundBlocks), dim3(kThreads), 0, st,
                       OneSpace{make_geom(d), make_walk(scratch, s, pass)}, in, out, clear_counts, overflow,
                       start ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_ldir_sweep(const Dims& d, const uint8_t* grad, void* scratch, const LdirScratch& s, int pass,
                             const MList& out, uint32_t* overflow, hipStream_t st, uint32_t item0) {
    hipLaunchKernelGGL(ldir_sweep_kernel, dim3(grid_of(d.verts, 8192)), dim3(kThreads), 0, st, grad, make_geom(d),
                       make_walk(scratch, s, pass), out, overflow, item0);
    return hipGetLastError();
}

hipError_t launch_ldir_finish(const Dims& d, uint8_t* grad, void* scratch, const LdirScratch& s, hipStream_t st) {
    uint8_t* base = static_cast<uint8_t*>(scratch);
    int ntx, nty;
    uint32_t n;
    tiles_of(d, ntx, nty, n);
    TileArgs a{};
    a.grad = grad;
    a.g = make_geom(d);
    a.cells = d.cells;
    a.ntx = ntx;
    a.nty = nty;
    a.v1 = reinterpret_cast<const uint32_t*>(base + s.bits1_off);
    a.v2 = reinterpret_cast<const uint32_t*>(base + s.bits2_off);
    a.bit_words = s.bit_words;
    hipLaunchKernelGGL(ldir_finish_kernel<false>, dim3(n), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

void ldir_tiles(const Dims& d, uint32_t* ntx, uint32_t* nty, uint32_t* tiles) {
    int x, y;
    tiles_of(d, x, y, *tiles);
    *ntx = (uint32_t)x;
    *nty = (uint32_t)y;
}

hipError_t launch_ldir_records_group(const GBatch& b, uint32_t* bits, uint64_t bit_words, hipStream_t st) {
    TileArgs a{};
    a.batch = b;
    a.bits = bits;
    a.bit_words = (uint32_t)bit_words;
    hipLaunchKernelGGL(ldir_records_kernel<true>, dim3(b.tiles.total), dim3(kThreads), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_ldir_round_group(const GBatch& b, int pass, bool start, const MList& in, const MList& out,
                                   uint32_t* clear_counts, uint32_t* overflow, hipStream_t st) {
    hipLaunchKernelGGL(ldir_round_kernel<GroupSpace>, dim3(kRoundBlocks), dim3(kThreads), 0, st, GroupSpace{b, pass},
                       in, out, clear_counts, overflow, start ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_ldir_finish_group(const GBatch& b, hipStream_t st) {
    TileArgs a{};
    a.batch = b;
    hipLaunchKernelGGL(ldir_finish_kernel<true>, dim3(b.tiles.total), dim3(kThreads), 0, st, a);
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

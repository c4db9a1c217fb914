// Internal interfaces between the C ABI layer and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msceer {

// Vertex grid X*Y*Z and its cell complex (2X-1)(2Y-1)(2Z-1).
struct Dims {
    int X, Y, Z;
    uint32_t Mx, My, Mz;
    uint64_t cells;
    uint64_t verts;
};

inline Dims make_dims(int64_t x, int64_t y, int64_t z) {
    Dims d;
    d.X = (int)x; d.Y = (int)y; d.Z = (int)z;
    d.Mx = (uint32_t)(2 * x - 1); d.My = (uint32_t)(2 * y - 1); d.Mz = (uint32_t)(2 * z - 1);
    d.cells = (uint64_t)d.Mx * d.My * d.Mz;
    d.verts = (uint64_t)x * y * z;
    return d;
}

struct MList;

struct LstarBuffers {
    const float* values;        // device, X*Y*Z
    uint8_t* grad;              // device, cells
    const int8_t* restriction;  // device, cells, or null (steepest pass)
    const MList* list1;         // critical 1-cells (ldir pass 1 start)
    const MList* list2;         // critical 2-cells (ldir pass 2 start)
    uint32_t* stats;            // device, kSubs*4: critical cells by sub-list and dimension
    uint32_t* overflow;         // device flag: a critical list ran out of room
};

// Lower-star pairing of every vertex; writes every cell byte exactly once
// with assigned/flag/pair set and ldir preset to 3 (0 for critical 3-cells).
// Lower-star kernel with brick shape `shape` (0: 64x2x2, 1: 32x4x2,
// 2: 16x4x4, 3: 8x8x4 vertices); every shape gives the same bytes.
constexpr int kLstarBricks = 4;
const char* lstar_brick_name(int shape);
hipError_t launch_lstar(const Dims& d, const LstarBuffers& b, hipStream_t s, int shape = 0);
void lstar_brick_dims(int shape, int* bx, int* by, int* bz);

// Batched launches over a group of volumes (gbatch.hpp): the unit tables,
// queues and volume table in g.
struct GBatch;
// Lower-star bricks of every volume of the group (list entries: each volume's
// cells from its first frontier item on; restriction unused): a work-group
// per brick (blocks == 0) or `blocks` persistent work-groups claiming bricks
// (queue cleared beforehand).
hipError_t launch_lstar_batch(const GBatch& g, const LstarBuffers& b, uint32_t blocks, hipStream_t s, int shape);
// Persistent work-groups of the batched lower-star kernel one CU holds at once.
int lstar_batch_occupancy(int shape);

// ldir (ascending-manifold dimension) passes, see ldir.hip.  Scratch: edge
// and quad walk records and the two visited bitmaps, at these byte offsets of
// one device buffer of ldir_scratch_bytes().
struct LdirScratch {
    size_t rec1_off, rec2_off, bits1_off, bits2_off;
    uint32_t bit_words;  // words of one bitmap
};
size_t ldir_scratch_bytes(const Dims& d, LdirScratch* s);
// Clears the ascending-manifold dimension bits (5-7) of n .grad bytes.
hipError_t launch_clear_ldir_bits(uint8_t* grad, int64_t n, hipStream_t st);
// Records of the walks (from the lower-star bytes) and cleared bitmaps.
hipError_t launch_ldir_records(const Dims& d, const uint8_t* grad, void* scratch, const LdirScratch& s,
                               hipStream_t st);
// One frontier round of pass 1 (edges) or 2 (quads): expands the items of
// `in` (critical cells, claimed first, when `start`), appends what is left to
// `out` (start items it did not reach claimed first) and zeroes
// `clear_counts` for the next round.  overflow |= 1 when an item did not fit
// `out`.
hipError_t launch_ldir_round(const Dims& d, void* scratch, const LdirScratch& s, int pass, bool start, const MList& in,
                             const MList& out, uint32_t* clear_counts, uint32_t* overflow, hipStream_t st);
// Rebuild a frontier from the bitmaps and records alone (after an overflow).
// (The frontier entries are the cell ids plus item0.)
hipError_t launch_ldir_sweep(const Dims& d, const uint8_t* grad, void* scratch, const LdirScratch& s, int pass,
                             const MList& out, uint32_t* overflow, hipStream_t st, uint32_t item0 = 0);
// Every cell's ldir from what the two passes reached.
hipError_t launch_ldir_finish(const Dims& d, uint8_t* grad, void* scratch, const LdirScratch& s, hipStream_t st);
// Tiles of the ldir streaming passes of a volume (along x, y, in all).
void ldir_tiles(const Dims& d, uint32_t* ntx, uint32_t* nty, uint32_t* tiles);
// The same passes over a group of volumes (one launch each; records and
// bitmaps of every volume in its table entry): records, clearing bit_words
// words of the group's bitmaps from `bits`; one round of pass 1 or 2 over
// frontier items of the group; the ldir bits of every volume.
hipError_t launch_ldir_records_group(const GBatch& b, uint32_t* bits, uint64_t bit_words, hipStream_t st);
hipError_t launch_ldir_round_group(const GBatch& b, int pass, bool start, const MList& in, const MList& out,
                                   uint32_t* clear_counts, uint32_t* overflow, hipStream_t st);
hipError_t launch_ldir_finish_group(const GBatch& b, hipStream_t st);
// Zero n words.
hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t s);

}  // namespace msceer

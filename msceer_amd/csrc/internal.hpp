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

// ldir (ascending-manifold dimension) passes, see ldir.hip.  Scratch: edge
// and quad walk records and the two visited bitmaps, at these byte offsets of
// one device buffer of ldir_scratch_bytes().
struct LdirScratch {
    size_t rec1_off, rec2_off, bits1_off, bits2_off;
    uint32_t bit_words;  // words of one bitmap
};
size_t ldir_scratch_bytes(const Dims& d, LdirScratch* s);
// Records of the walks (from the lower-star bytes) and cleared bitmaps.
hipError_t launch_ldir_records(const Dims& d, const uint8_t* grad, void* scratch, const LdirScratch& s,
                               hipStream_t st);
// One frontier round of pass 1 (edges) or 2 (quads): expands the items of
// `in` (critical cells, claimed first, when `start`), appends what is left to
// `out` and zeroes `clear_counts` for the next round.  overflow |= 1 when an
// item did not fit `out`, |= 4 when a start round left items unread.
hipError_t launch_ldir_round(const Dims& d, void* scratch, const LdirScratch& s, int pass, bool start, const MList& in,
                             const MList& out, uint32_t* clear_counts, uint32_t* overflow, hipStream_t st);
// Rebuild a frontier from the bitmaps and records alone (after an overflow).
hipError_t launch_ldir_sweep(const Dims& d, const uint8_t* grad, void* scratch, const LdirScratch& s, int pass,
                             const MList& out, uint32_t* overflow, hipStream_t st);
// Every cell's ldir from what the two passes reached.
hipError_t launch_ldir_finish(const Dims& d, uint8_t* grad, void* scratch, const LdirScratch& s, hipStream_t st);
// Zero n words.
hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t s);

}  // namespace msceer

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

// ldir passes.  `tail` is one device uint32 shadowing the bytes of the last,
// partial 32-bit word of the grad array: pack it before the passes
// (unpack=false) and write it back after them (unpack=true).
hipError_t launch_ldir_tail(const Dims& d, uint8_t* grad, uint32_t* tail, bool unpack, hipStream_t s);
// One frontier round (see ldir.hip): follows the items of `in`, appends what
// is left to `out` and zeroes `clear_counts` (kSubs words) for the next round.
hipError_t launch_ldir_round(const Dims& d, uint8_t* grad, uint32_t* tail, const MList& in, const MList& out,
                             uint32_t* clear_counts, uint32_t* overflow, uint32_t value, hipStream_t s);
// Rebuild a frontier from the grad bytes alone (after an overflow).
hipError_t launch_ldir_sweep(const Dims& d, uint8_t* grad, uint32_t* tail, uint32_t item_dim,
                             uint32_t value, const MList& out, uint32_t* overflow, hipStream_t s);
// Zero n words.
hipError_t launch_zero(uint32_t* p, uint32_t n, hipStream_t s);

}  // namespace msceer

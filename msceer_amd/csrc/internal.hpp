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

// Device counters block (uint32 slots).
enum Counter : int {
    kCritCount1 = 0,    // critical 1-cells appended to list 1
    kCritCount2 = 1,    // critical 2-cells appended to list 2
    kCritOverflow = 2,  // a critical list ran out of room
    kFrontCount = 3,    // entries written to the outgoing frontier
    kFrontOverflow = 4, // the outgoing frontier ran out of room
    kCritDim0 = 8,      // critical cells by dimension (statistics)
    kCritDim1 = 9,
    kCritDim2 = 10,
    kCritDim3 = 11,
    kNumCounters = 16
};

struct LstarBuffers {
    const float* values;   // device, X*Y*Z
    uint8_t* grad;         // device, cells
    const int8_t* restriction;  // device, cells, or null (steepest pass)
    uint32_t* crit_list1;  // device, crit_cap entries
    uint32_t* crit_list2;
    uint32_t crit_cap;
    uint32_t* counters;    // device, kNumCounters
};

// Lower-star pairing of every vertex; writes every cell byte exactly once
// with assigned/flag/pair set and ldir preset to 3 (0 for critical 3-cells).
hipError_t launch_lstar(const Dims& d, const LstarBuffers& b, hipStream_t s);

// ldir passes.  `tail` is one device uint32 shadowing the bytes of the last,
// partial 32-bit word of the grad array: pack it before the passes
// (unpack=false) and write it back after them (unpack=true).
hipError_t launch_ldir_tail(const Dims& d, uint8_t* grad, uint32_t* tail, bool unpack, hipStream_t s);
// One frontier round (see ldir.hip for the semantics).
hipError_t launch_ldir_expand(const Dims& d, uint8_t* grad, uint32_t* tail, const uint32_t* in,
                              uint32_t n_in, uint32_t* out, uint32_t out_cap, uint32_t* counters,
                              uint32_t value, hipStream_t s);
// Rebuild a frontier from the grad bytes alone (start, and after an overflow).
hipError_t launch_ldir_sweep(const Dims& d, uint8_t* grad, uint32_t* tail, uint32_t item_dim,
                             uint32_t value, uint32_t* out, uint32_t out_cap, uint32_t* counters,
                             hipStream_t s);

}  // namespace msceer

// Batched lower-star .grad: many volumes in one launch per stage.
//
// A group of volumes is laid out as one index space per kind of work unit:
// lower-star bricks, ldir tiles and frontier items (cells) of volume v are
// numbered from first[v] on, first[] an ascending table with first[n] the
// total.  Work-groups are mapped to units so that every XCD takes a
// contiguous eighth of them (bricks that share halos and partly written
// cache lines of the .grad then meet in that XCD's L2): either one
// work-group per unit, dealt out by the hardware, or persistent work-groups
// that claim units from eight queues, one per XCD, and take units from the
// other queues once their own is empty.  Frontier items are the volume's
// cell id plus its first[], found again through a coarse bucket table.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlist.hpp"

namespace msceer {

// One volume of a batch (device-resident table).
struct GVol {
    const float* values;  // X*Y*Z
    uint8_t* grad;        // cells
    uint16_t* rec1;       // ldir records and visited bitmaps of the volume
    uint8_t* rec2;
    uint32_t* bits1;
    uint32_t* bits2;
    uint64_t cells;
    uint32_t bit_words;  // words of one bitmap
    int X, Y, Z;
    uint32_t Mx, My, Mz;
    uint32_t ntx, nty;  // ldir tiles along x and y
};

constexpr int kQueues = 8;  // claim queues, one per XCD

// Unit table of a group (device): volume v holds units first[v] ..
// first[v + 1] - 1.  bucket[u >> shift] is the volume holding unit
// (u >> shift) << shift (nbuckets + 1 entries, the last one for the end), so
// the volume of unit u lies between bucket[u >> shift] and the next entry.
struct GUnits {
    const uint32_t* first;  // n + 1 entries
    const uint32_t* bucket;
    uint32_t shift;
    uint32_t total;
};

struct GBatch {
    const GVol* vols;
    uint32_t nvols;
    GUnits bricks;  // lower-star bricks of the launch's brick shape
    GUnits tiles;   // ldir tiles
    GUnits items;   // frontier items: the volume's cells
    uint32_t* queue;            // kQueues claim counters (persistent launches; zeroed before each)
    unsigned long long* times;  // 2 per volume: first brick started, last tile finished (wall clock); or null
    uint32_t chunk;             // units per claim
};

// Wall clock of the device (constant rate, hipDeviceAttributeWallClockRate).
__device__ __forceinline__ unsigned long long batch_clock() { return wall_clock64(); }

// Volume holding unit u (u < t.total).
__device__ __forceinline__ uint32_t unit_volume(const GUnits& t, uint32_t u) {
    uint32_t lo = t.bucket[u >> t.shift], hi = t.bucket[(u >> t.shift) + 1];
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (t.first[mid] <= u) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Claim the next `chunk` units of a stage's queues for a work-group (called by
// one lane).  The work-group's own XCD queue first, then the others in turn;
// `q` (the queue tried next, relative to the home queue) advances past
// exhausted queues.  false once every queue is exhausted.
__device__ __forceinline__ bool claim_units(uint32_t* queues, uint32_t total, uint32_t chunk, uint32_t home,
                                            uint32_t& q, uint32_t& lo, uint32_t& hi) {
    while (q < (uint32_t)kQueues) {
        const uint32_t x = (home + q) % kQueues;
        const uint32_t qlo = (uint32_t)((uint64_t)total * x / kQueues);
        const uint32_t qhi = (uint32_t)((uint64_t)total * (x + 1) / kQueues);
        if (qlo < qhi) {
            const uint32_t old = atomicAdd(&queues[x], chunk);
            if (old < qhi - qlo) {
                lo = qlo + old;
                hi = lo + chunk < qhi ? lo + chunk : qhi;
                return true;
            }
        }
        q++;
    }
    return false;
}

// Per-work-group unit iterator over a stage's queues (state in LDS; next()
// must be called by every lane of the work-group, as it synchronises).
struct UnitCursor {
    uint32_t lo, hi, q, vol, unit;
    int done;
};

__device__ __forceinline__ void cursor_init(UnitCursor* c) {
    if (threadIdx.x == 0) {
        c->lo = c->hi = 0;
        c->q = 0;
        c->vol = 0;
        c->done = 0;
    }
}

// Next unit and its volume; false when the stage's units are all taken.
__device__ __forceinline__ bool cursor_next(UnitCursor* c, uint32_t* queues, const GUnits& t, uint32_t chunk,
                                            uint32_t& unit, uint32_t& vol) {
    if (threadIdx.x == 0 && !c->done) {
        if (c->lo == c->hi) {
            uint32_t q = c->q, lo = 0, hi = 0;
            if (claim_units(queues, t.total, chunk, blockIdx.x % kQueues, q, lo, hi)) {
                c->lo = lo;
                c->hi = hi;
            } else {
                c->done = 1;
            }
            c->q = q;
        }
        if (!c->done) {
            c->unit = c->lo++;
            c->vol = unit_volume(t, c->unit);
        }
    }
    __syncthreads();
    // uniform over the work-group: kept in scalar registers, the volume's
    // fields then come through scalar loads
    const bool more = __builtin_amdgcn_readfirstlane(c->done) == 0;
    unit = __builtin_amdgcn_readfirstlane(c->unit);
    vol = __builtin_amdgcn_readfirstlane(c->vol);
    __syncthreads();
    return more;
}

}  // namespace msceer

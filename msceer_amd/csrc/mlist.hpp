// Append-only work lists split into kSubs sub-lists.
//
// Millions of work-groups appending to one list would serialise on a single
// counter in L2; a work-group instead appends to sub-list blockIdx % kSubs
// with one atomic reservation for its whole batch.  Readers see the
// concatenation of the sub-lists through a prefix table built in LDS.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msceer {

constexpr int kSubs = 256;
// Words of one count array: kSubs sub-list counters, then the spill counter.
constexpr int kCountWords = kSubs + 8;

// Sub-list s holds sub_cap entries at items + s * sub_cap; entries that do
// not fit their sub-list go to the shared spill list of spill_cap entries at
// items + kSubs * sub_cap (counted in counts[kSubs]).  The capacity that
// matters is therefore the total, however unevenly the work-groups fill
// their sub-lists.
struct MList {
    uint32_t* items;
    uint32_t* counts;  // kSubs + 1 reservation counters (may exceed the capacities)
    uint32_t sub_cap;
    uint32_t spill_cap;
};

// A reservation of n consecutive entries: the first `split` go to the
// sub-list from slot `base`, the rest to the spill list from `spill_base`.
struct MSlot {
    uint32_t base, split, spill_base;
};

__device__ __forceinline__ MSlot mlist_reserve(const MList& l, int sub, uint32_t n) {
    MSlot r;
    r.base = atomicAdd(&l.counts[sub], n);
    r.split = r.base >= l.sub_cap ? 0u : (n < l.sub_cap - r.base ? n : l.sub_cap - r.base);
    r.spill_base = r.split < n ? atomicAdd(&l.counts[kSubs], n - r.split) : 0u;
    return r;
}

// Store entry j of a reservation; false if the spill list is full too.
__device__ __forceinline__ bool mlist_store(const MList& l, int sub, const MSlot& r, uint32_t j, uint32_t item) {
    if (j < r.split) {
        l.items[(size_t)sub * l.sub_cap + r.base + j] = item;
        return true;
    }
    const uint32_t k = r.spill_base + (j - r.split);
    if (k >= l.spill_cap) return false;
    l.items[(size_t)kSubs * l.sub_cap + k] = item;
    return true;
}

// Single-entry append.
__device__ __forceinline__ bool mlist_push(const MList& l, int sub, uint32_t item) {
    return mlist_store(l, sub, mlist_reserve(l, sub, 1u), 0u, item);
}

// Block-level flush of n staged items (LDS) into sub-list `sub`.  Must be
// called by every thread of the block; `s_slot` lives in LDS.
__device__ __forceinline__ void mlist_flush_block(const MList& l, int sub, const uint32_t* stage, uint32_t n,
                                                  MSlot* s_slot, uint32_t* overflow) {
    const uint32_t tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    if (tid == 0) *s_slot = n ? mlist_reserve(l, sub, n) : MSlot{0u, 0u, 0u};
    __syncthreads();
    const MSlot r = *s_slot;
    const uint32_t nthreads = blockDim.x * blockDim.y * blockDim.z;
    bool lost = false;
    for (uint32_t j = tid; j < n; j += nthreads)
        if (!mlist_store(l, sub, r, j, stage[j])) lost = true;
    if (lost) atomicOr(overflow, 1u);
}

// Prefix table of a list's counts, built by one block in LDS (kSubs + 2
// entries: the sub-lists, the spill list, the total).  Every thread of the
// block must call it.
__device__ __forceinline__ uint32_t mlist_prefix(const MList& l, uint32_t* prefix) {
    const uint32_t tid = threadIdx.x;
    if (tid < (uint32_t)kSubs) {
        const uint32_t c = l.counts[tid];
        prefix[tid + 1] = c < l.sub_cap ? c : l.sub_cap;
    }
    if (tid == 0) {
        prefix[0] = 0;
        const uint32_t c = l.counts[kSubs];
        prefix[kSubs + 1] = c < l.spill_cap ? c : l.spill_cap;
    }
    __syncthreads();
    if (tid == 0) {
        for (int s = 1; s <= kSubs + 1; s++) prefix[s] += prefix[s - 1];
    }
    __syncthreads();
    return prefix[kSubs + 1];
}

// Item i of the concatenation (i < total).  The spill list is list kSubs and
// starts right after the sub-lists.
__device__ __forceinline__ uint32_t mlist_at(const MList& l, const uint32_t* prefix, uint32_t i) {
    int lo = 0, hi = kSubs + 1;  // find the list s with prefix[s] <= i < prefix[s+1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= i) lo = mid; else hi = mid;
    }
    return l.items[(size_t)lo * l.sub_cap + (i - prefix[lo])];
}

}  // namespace msceer

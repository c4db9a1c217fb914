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

struct MList {
    uint32_t* items;   // kSubs * sub_cap entries; sub-list s starts at s * sub_cap
    uint32_t* counts;  // kSubs reservation counters (may exceed sub_cap: overflow)
    uint32_t sub_cap;
};

// Reserve n slots in sub-list `sub` for the calling thread's batch; returns
// the first slot (caller checks it against sub_cap).
__device__ __forceinline__ uint32_t mlist_reserve(const MList& l, int sub, uint32_t n) {
    return atomicAdd(&l.counts[sub], n);
}

// Block-level flush of n staged items (LDS) into sub-list `sub`.  Must be
// called by every thread of the block; `s_base` is one word of LDS.
__device__ __forceinline__ void mlist_flush_block(const MList& l, int sub, const uint32_t* stage, uint32_t n,
                                                  uint32_t* s_base, uint32_t* overflow) {
    if (threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z) == 0)
        *s_base = n ? mlist_reserve(l, sub, n) : 0u;
    __syncthreads();
    const uint32_t base = *s_base;
    const uint32_t nthreads = blockDim.x * blockDim.y * blockDim.z;
    const uint32_t tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    bool lost = false;
    for (uint32_t j = tid; j < n; j += nthreads) {
        const uint32_t slot = base + j;
        if (slot < l.sub_cap) l.items[(size_t)sub * l.sub_cap + slot] = stage[j];
        else lost = true;
    }
    if (lost) atomicOr(overflow, 1u);
}

// Prefix table of a list's sub-list counts, built by one block in LDS
// (prefix[kSubs] entries + total).  Every thread of the block must call it.
__device__ __forceinline__ uint32_t mlist_prefix(const MList& l, uint32_t* prefix) {
    const uint32_t tid = threadIdx.x;
    if (tid < (uint32_t)kSubs) {
        const uint32_t c = l.counts[tid];
        prefix[tid + 1] = c < l.sub_cap ? c : l.sub_cap;
    }
    if (tid == 0) prefix[0] = 0;
    __syncthreads();
    if (tid == 0) {
        for (int s = 1; s <= kSubs; s++) prefix[s] += prefix[s - 1];
    }
    __syncthreads();
    return prefix[kSubs];
}

// Item i of the concatenation (i < total).
__device__ __forceinline__ uint32_t mlist_at(const MList& l, const uint32_t* prefix, uint32_t i) {
    int lo = 0, hi = kSubs;  // find the sub-list s with prefix[s] <= i < prefix[s+1]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (prefix[mid] <= i) lo = mid; else hi = mid;
    }
    return l.items[(size_t)lo * l.sub_cap + (i - prefix[lo])];
}

}  // namespace msceer

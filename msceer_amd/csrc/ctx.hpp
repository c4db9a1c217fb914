// Context internals shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "msceer.h"
#include "internal.hpp"
#include "mlist.hpp"

namespace msceer {

// Grow-only device buffer.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (ptr) { (void)hipFree(ptr); ptr = nullptr; bytes = 0; }
        need = (need + 255) & ~(size_t)255;
        hipError_t e = hipMalloc(&ptr, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

constexpr uint32_t kMinListCap = 1u << 16;
constexpr uint32_t kMaxListCap = 1u << 26;   // entries per critical list
constexpr uint32_t kMaxFrontCap = 1u << 27;  // entries per frontier buffer

// Control block (uint32 words, device; mirrored to pinned host memory):
//   sub-list counts of the two critical lists and the three rotating
//   frontier count arrays, critical statistics, flags.
enum Ctl : int {
    kCtlList1 = 0,
    kCtlList2 = kSubs,
    kCtlRound = 2 * kSubs,       // three arrays of kSubs
    kCtlStats = 5 * kSubs,       // kSubs * 4
    kCtlOverflow = 9 * kSubs,    // a list or frontier ran out of room
    kCtlTail = 9 * kSubs + 1,    // shadow of the grad array's last partial word
    kCtlFlag = 9 * kSubs + 2,    // change flag of iterative sweeps
    kCtlCount = 9 * kSubs + 3,   // a kernel-reported count
    kCtlIntStats = 9 * kSubs + 8,  // eight integrator counters
    kCtlWords = 9 * kSubs + 16
};

int hip_fail(hipError_t e);
int check_dims(int64_t nx, int64_t ny, int64_t nz);

}  // namespace msceer

struct msceer_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    msceer::DevBuf values, grad, lists, front0, front1, ctl;
    // accurate labelling
    msceer::DevBuf gradf, label, label2, dest, seedmask, redo, stamp, rep, rep2, restriction, host_io, mask;
    uint32_t* h_ctl = nullptr;  // pinned mirror of the control block
    hipEvent_t ev[12] = {};
    msceer_timing last{};
    msceer_label_stats last_labels{};
};

namespace msceer {

uint32_t* ctl_dev(msceer_ctx* ctx);
int read_ctl(msceer_ctx* ctx, int first, int n);
uint64_t list_total(const uint32_t* counts, uint32_t sub_cap);
int ldir_pass(msceer_ctx* ctx, const Dims& d, uint8_t* grad, uint32_t item_dim, uint32_t value,
              const MList& start, bool start_valid, uint32_t front_sub_cap, int* rounds);
// Critical lists and frontiers sized for a volume (sub-list capacities out).
int ensure_lists(msceer_ctx* ctx, const Dims& d, uint32_t* list_sub, uint32_t* front_sub);
// ldir passes + tail shadow on a .grad whose lstar pass filled list1/list2.
int run_ldir(msceer_ctx* ctx, const Dims& d, uint8_t* grad, const MList& list1, const MList& list2, bool lists_ok,
             uint32_t front_sub, int* rounds);

#define CHECK_HIP(expr)                                        \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return ::msceer::hip_fail(_e);   \
    } while (0)

}  // namespace msceer

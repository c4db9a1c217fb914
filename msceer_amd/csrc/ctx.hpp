// Context internals shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <vector>

#include "msceer.h"
#include "internal.hpp"
#include "mlist.hpp"

namespace msceer {

struct BatchState;
// Releases the batched .grad state of a context (graphs, buffers, streams).
void batch_state_free(msceer_ctx* ctx);

// Grow-only device buffer.  A borrowed buffer is a view of memory the
// buffer does not own (a slice of a static scratch array): it is never freed
// or grown, and asking it for more than it holds fails.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool borrowed = false;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (borrowed) return hipErrorOutOfMemory;
        if (ptr) { (void)hipFree(ptr); ptr = nullptr; bytes = 0; }
        need = (need + 255) & ~(size_t)255;
        hipError_t e = hipMalloc(&ptr, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
    void release() {
        if (ptr && !borrowed) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        borrowed = false;
    }
    void borrow(void* p, size_t n) {
        ptr = p;
        bytes = n;
        borrowed = true;
    }
};

// Pinned host staging memory (hipHostMalloc): copies to and from it run at
// full link speed.  Grown with 25% headroom, because the sizes it holds vary
// a little from call to call and pinning a fresh buffer costs more than the
// copy; never cleared.
struct HostBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (ptr) { (void)hipHostFree(ptr); ptr = nullptr; bytes = 0; }
        need = ((need + need / 4) + 4095) & ~(size_t)4095;
        hipError_t e = hipHostMalloc(&ptr, need, hipHostMallocDefault);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    template <typename T>
    T* as() const { return static_cast<T*>(ptr); }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

constexpr uint32_t kMinListCap = 1u << 16;
constexpr uint32_t kMaxListCap = 1u << 26;   // entries per critical list
constexpr uint32_t kMaxFrontCap = 1u << 27;  // entries per frontier buffer

// Control block (uint32 words, device; mirrored to pinned host memory):
//   count arrays (kCountWords each) of the two critical lists and of the
//   three rotating frontiers, critical statistics, flags.
enum Ctl : int {
    kCtlList1 = 0,
    kCtlList2 = kCountWords,
    kCtlRound = 2 * kCountWords,   // three count arrays
    kCtlStats = 5 * kCountWords,   // kSubs * 4
    kCtlOverflow = 5 * kCountWords + 4 * kSubs,  // a list or frontier ran out of room
    kCtlTail = kCtlOverflow + 1,   // (unused)
    kCtlFlag = kCtlOverflow + 2,   // change flag of iterative sweeps
    kCtlCount = kCtlOverflow + 3,  // a kernel-reported count
    kCtlCursor = kCtlOverflow + 4,  // head, tail, trace-ahead flag, chunk head, first count of the re-integration queue
    kCtlIntStats = kCtlOverflow + 12,  // sixteen integrator counters (8-byte aligned)
    kCtlWords = kCtlOverflow + 28
};

// Capacities of the critical lists and the frontiers of one volume.
struct ListCaps {
    uint32_t list_sub, list_spill;
    uint32_t front_sub, front_spill;
};

// Lower-star brick shapes chosen per x extent (shared by a context and the
// worker contexts of its batches).
struct Tuning {
    std::mutex mu;
    int mode = 1;                       // 0: fixed 64x2x2; 1: measured per x extent; 2+k: shape k
    std::map<int, int> brick_by_x;      // x extent -> brick shape
};

int hip_fail(hipError_t e);
int check_dims(int64_t nx, int64_t ny, int64_t nz);

}  // namespace msceer

struct msceer_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    msceer::DevBuf values, grad, lists, front0, front1, ctl;
    msceer::DevBuf ldir;  // ldir walk records and visited bitmaps
    // accurate labelling
    msceer::DevBuf gradf, gradd, scan, label, label2, dest, seedmask, redo, stamp, rep, rep2, restriction, host_io, mask;
    // accurate 1-manifolds: per-cell line stamps (two words, kept cleared
    // between calls) and the saddle / window scratch
    msceer::DevBuf line_stamps, lines;
    // descending 3-manifolds: their labels and seed mask
    msceer::DevBuf label_dsc, seedmask2;
    bool line_stamps_clear = false;
    uint64_t line_stamps_cells = 0;
    msceer::HostBuf host_stage;  // pinned staging of host-side merges (cleanup)
    uint32_t* h_ctl = nullptr;  // pinned mirror of the control block
    // 8, 9: certain-round groups; 12-15: integration passes; 16, 17: the
    // descending branch (on the auxiliary context); 18: its start
    static constexpr int kEvents = 19;
    hipEvent_t ev[kEvents] = {};
    msceer_timing last{};
    msceer_label_stats last_labels{};
    std::shared_ptr<msceer::Tuning> tuning = std::make_shared<msceer::Tuning>();
    std::vector<msceer_ctx*> workers;  // contexts of msceer_batch's worker threads
    msceer_ctx* aux = nullptr;         // the descending labelling's own stream and buffers
    msceer::BatchState* gbatch = nullptr;  // msceer_batch_grad: buffers, plans and graphs
};

namespace msceer {

uint32_t* ctl_dev(msceer_ctx* ctx);
int read_ctl(msceer_ctx* ctx, int first, int n);
// Entries held by a list whose count array (host mirror) is `counts`.
uint64_t list_total(const uint32_t* counts, const MList& l);
// Critical lists and frontiers sized for a volume.
int ensure_lists(msceer_ctx* ctx, const Dims& d, ListCaps* caps);
// Critical 1-cell (which = 0) or 2-cell (which = 1) list.
MList crit_list(msceer_ctx* ctx, const ListCaps& c, int which);
// Frontier k of a chain of rounds: buffer k % 2, count array k % 3.
MList front_list(msceer_ctx* ctx, const ListCaps& c, int k);
int ldir_pass(msceer_ctx* ctx, const Dims& d, uint8_t* grad, const LdirScratch& ls, int pass, const MList& start,
              bool start_valid, const ListCaps& caps, int* rounds);
// Brick shape for the lower-star kernel on this volume: tuned on first use of
// an x extent (times every shape on the given buffers; call before the
// control block is cleared for the real launch).
int lstar_shape(msceer_ctx* ctx, const Dims& d, const LstarBuffers& b, int* shape);
// The shape lstar_shape would return without timing any: the tuned one, else `fallback`.
int tuned_shape(msceer_ctx* ctx, const Dims& d, int fallback);
// Resolves the device address of the allocation-free entry point's scratch
// array on `device` (current), once per device.
int noalloc_setup(int device);
// ldir passes on a .grad whose lstar pass filled the critical lists.
int run_ldir(msceer_ctx* ctx, const Dims& d, uint8_t* grad, const ListCaps& caps, bool lists_ok, int* rounds);
// Capacities of the critical lists and frontiers of a volume, and the bytes
// ensure_lists gives them (two lists, two frontiers).
ListCaps list_caps(const Dims& d);
size_t list_bytes(const ListCaps& c);
size_t front_bytes(const ListCaps& c);
// Flag of lstar_grad_device (beside the MSCEER_* flags): use the brick shape
// already tuned for this x extent (64x2x2 if none) instead of tuning one.
constexpr uint32_t kLstarNoTuning = 0x80000000u;
// The whole .grad of a volume in device memory (msceer_lstar_grad's work after
// staging): lists, frontiers and ldir scratch from the context; host_out
// (may be null) receives a copy before the final wait.
int lstar_grad_device(msceer_ctx* ctx, const Dims& d, const float* d_values, uint8_t* d_grad, uint8_t* host_out,
                      uint32_t flags);

#define CHECK_HIP(expr)                                        \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return ::msceer::hip_fail(_e);   \
    } while (0)

}  // namespace msceer

// C ABI of libmsceer_hip.so (declared in include/msceer.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/msceer.h"
#include "internal.hpp"

using namespace msceer;

namespace {

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
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
};

constexpr uint32_t kMinListCap = 1u << 16;
constexpr uint32_t kMaxListCap = 1u << 26;   // entries per critical list
constexpr uint32_t kMaxFrontCap = 1u << 26;  // entries per frontier buffer

}  // namespace

struct msceer_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    DevBuf values, grad, lists, front0, front1, counters;
    uint32_t* h_counters = nullptr;  // pinned
    hipEvent_t ev[6] = {};
    msceer_timing last{};
};

namespace {

int hip_fail(hipError_t e) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return MSCEER_ENOMEM;
    return MSCEER_EHIP;
}

#define CHECK_HIP(expr)                       \
    do {                                      \
        hipError_t _e = (expr);               \
        if (_e != hipSuccess) return hip_fail(_e); \
    } while (0)

int check_dims(int64_t nx, int64_t ny, int64_t nz) {
    if (nx < 2 || ny < 2 || nz < 2) return MSCEER_EINVAL;
    if (nx > 65536 || ny > 65536 || nz > 65536) return MSCEER_ETOOBIG;
    const double cells = (2.0 * nx - 1) * (2.0 * ny - 1) * (2.0 * nz - 1);
    if (cells >= 4294967295.0) return MSCEER_ETOOBIG;
    return MSCEER_OK;
}

// Read back counters [0, n) synchronously.
int read_counters(msceer_ctx* ctx, int n) {
    CHECK_HIP(hipMemcpyAsync(ctx->h_counters, ctx->counters.ptr, sizeof(uint32_t) * n,
                             hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return MSCEER_OK;
}

int zero_counter(msceer_ctx* ctx, int first, int n) {
    CHECK_HIP(hipMemsetAsync(static_cast<uint32_t*>(ctx->counters.ptr) + first, 0, sizeof(uint32_t) * n,
                             ctx->stream));
    return MSCEER_OK;
}

// One ldir pass: frontier rounds from `start` (or from a sweep), with a sweep
// after any overflow, until no item is left.
int ldir_pass(msceer_ctx* ctx, const Dims& d, uint8_t* grad, uint32_t* tail, uint32_t item_dim,
              uint32_t value, const uint32_t* start, uint32_t n_start, bool start_valid,
              uint32_t front_cap, int* rounds) {
    uint32_t* counters = static_cast<uint32_t*>(ctx->counters.ptr);
    uint32_t* bufs[2] = {static_cast<uint32_t*>(ctx->front0.ptr), static_cast<uint32_t*>(ctx->front1.ptr)};
    int cur = 0;
    const uint32_t* in = start;
    uint32_t n = n_start;
    bool need_sweep = !start_valid;
    for (int guard = 0; guard < 1000000; guard++) {
        if (need_sweep) {
            int rc = zero_counter(ctx, kFrontCount, 2);
            if (rc) return rc;
            CHECK_HIP(launch_ldir_sweep(d, grad, tail, item_dim, value, bufs[cur], front_cap, counters,
                                        ctx->stream));
            rc = read_counters(ctx, kNumCounters);
            if (rc) return rc;
            n = std::min(ctx->h_counters[kFrontCount], front_cap);
            need_sweep = ctx->h_counters[kFrontOverflow] != 0;
            in = bufs[cur];
            cur ^= 1;
            (*rounds)++;
            if (n == 0 && !need_sweep) return MSCEER_OK;
        }
        while (n > 0) {
            int rc = zero_counter(ctx, kFrontCount, 2);
            if (rc) return rc;
            CHECK_HIP(launch_ldir_expand(d, grad, tail, in, n, bufs[cur], front_cap, counters, value,
                                         ctx->stream));
            rc = read_counters(ctx, kNumCounters);
            if (rc) return rc;
            (*rounds)++;
            if (ctx->h_counters[kFrontOverflow]) need_sweep = true;
            n = std::min(ctx->h_counters[kFrontCount], front_cap);
            in = bufs[cur];
            cur ^= 1;
        }
        if (!need_sweep) return MSCEER_OK;
    }
    return MSCEER_EHIP;
}

}  // namespace

extern "C" {

int msceer_version(void) { return MSCEER_VERSION; }

const char* msceer_strerror(int code) {
    switch (code) {
        case MSCEER_OK: return "ok";
        case MSCEER_EINVAL: return "invalid argument";
        case MSCEER_EHIP: return "HIP runtime error";
        case MSCEER_ENOMEM: return "device out of memory";
        case MSCEER_ETOOBIG: return "volume too large (needs fewer than 2^32 cells)";
        case MSCEER_ENODEV: return "no usable GPU";
    }
    return "unknown error";
}

int msceer_create(msceer_ctx** out, int device) {
    if (!out) return MSCEER_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return MSCEER_ENODEV;
    if (device < 0 || device >= n) return MSCEER_EINVAL;
    msceer_ctx* ctx = new (std::nothrow) msceer_ctx();
    if (!ctx) return MSCEER_ENOMEM;
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->h_counters, sizeof(uint32_t) * kNumCounters, hipHostMallocDefault);
    for (int i = 0; i < 6 && e == hipSuccess; i++) e = hipEventCreate(&ctx->ev[i]);
    if (e == hipSuccess) e = ctx->counters.ensure(sizeof(uint32_t) * kNumCounters);
    if (e != hipSuccess) {
        msceer_destroy(ctx);
        return hip_fail(e);
    }
    ctx->stream = ctx->own;
    *out = ctx;
    return MSCEER_OK;
}

int msceer_destroy(msceer_ctx* ctx) {
    if (!ctx) return MSCEER_EINVAL;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->values.release();
    ctx->grad.release();
    ctx->lists.release();
    ctx->front0.release();
    ctx->front1.release();
    ctx->counters.release();
    if (ctx->h_counters) (void)hipHostFree(ctx->h_counters);
    for (int i = 0; i < 6; i++)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return MSCEER_OK;
}

int msceer_set_stream(msceer_ctx* ctx, void* hip_stream) {
    if (!ctx) return MSCEER_EINVAL;
    ctx->stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : ctx->own;
    return MSCEER_OK;
}

int msceer_last_timing(const msceer_ctx* ctx, msceer_timing* out) {
    if (!ctx || !out) return MSCEER_EINVAL;
    *out = ctx->last;
    return MSCEER_OK;
}

int msceer_lstar_grad(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                      uint8_t* grad_out, uint32_t flags) {
    if (!ctx || !values || !grad_out) return MSCEER_EINVAL;
    if (flags & ~(MSCEER_DEVICE_PTRS | MSCEER_SKIP_LDIR | MSCEER_TIMING)) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const Dims d = make_dims(nx, ny, nz);
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const bool timing = flags & MSCEER_TIMING;
    hipStream_t s = ctx->stream;

    const float* d_values = values;
    uint8_t* d_grad = grad_out;
    if (!dev) {
        CHECK_HIP(ctx->values.ensure(sizeof(float) * d.verts));
        CHECK_HIP(ctx->grad.ensure(d.cells));
        d_values = static_cast<const float*>(ctx->values.ptr);
        d_grad = static_cast<uint8_t*>(ctx->grad.ptr);
        CHECK_HIP(hipMemcpyAsync(ctx->values.ptr, values, sizeof(float) * d.verts, hipMemcpyHostToDevice, s));
    }
    // critical-cell lists: sized to the volume, capped; overflow falls back to a sweep
    const uint32_t list_cap = (uint32_t)std::min<uint64_t>(kMaxListCap, std::max<uint64_t>(kMinListCap, d.verts / 4));
    const uint32_t front_cap = (uint32_t)std::min<uint64_t>(kMaxFrontCap, std::max<uint64_t>(kMinListCap, d.verts / 2));
    CHECK_HIP(ctx->lists.ensure(sizeof(uint32_t) * 2 * (size_t)list_cap + 256));
    CHECK_HIP(ctx->front0.ensure(sizeof(uint32_t) * (size_t)front_cap));
    CHECK_HIP(ctx->front1.ensure(sizeof(uint32_t) * (size_t)front_cap));
    uint32_t* counters = static_cast<uint32_t*>(ctx->counters.ptr);
    uint32_t* list1 = static_cast<uint32_t*>(ctx->lists.ptr);
    uint32_t* list2 = list1 + list_cap;
    uint32_t* tail = list2 + list_cap;  // one word after the lists

    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[0], s));
    rc = zero_counter(ctx, 0, kNumCounters);
    if (rc) return rc;
    LstarBuffers b{d_values, d_grad, nullptr, list1, list2, list_cap, counters};
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[1], s));
    CHECK_HIP(launch_lstar(d, b, s));
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[2], s));
    rc = read_counters(ctx, kNumCounters);
    if (rc) return rc;
    msceer_timing t{};
    for (int i = 0; i < 4; i++) t.critical[i] = ctx->h_counters[kCritDim0 + i];
    const bool lists_ok = ctx->h_counters[kCritOverflow] == 0;
    const uint32_t n1 = std::min(ctx->h_counters[kCritCount1], list_cap);
    const uint32_t n2 = std::min(ctx->h_counters[kCritCount2], list_cap);

    int rounds = 0;
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[3], s));
    if (!(flags & MSCEER_SKIP_LDIR)) {
        CHECK_HIP(launch_ldir_tail(d, d_grad, tail, false, s));
        rc = ldir_pass(ctx, d, d_grad, tail, 1, 2, list1, n1, lists_ok, front_cap, &rounds);
        if (rc) return rc;
        rc = ldir_pass(ctx, d, d_grad, tail, 2, 1, list2, n2, lists_ok, front_cap, &rounds);
        if (rc) return rc;
        CHECK_HIP(launch_ldir_tail(d, d_grad, tail, true, s));
    }
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[4], s));
    if (!dev) CHECK_HIP(hipMemcpyAsync(grad_out, d_grad, d.cells, hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    if (timing) {
        CHECK_HIP(hipEventElapsedTime(&t.lstar_ms, ctx->ev[1], ctx->ev[2]));
        CHECK_HIP(hipEventElapsedTime(&t.ldir_ms, ctx->ev[3], ctx->ev[4]));
        CHECK_HIP(hipEventElapsedTime(&t.total_ms, ctx->ev[0], ctx->ev[4]));
    }
    t.ldir_rounds = rounds;
    ctx->last = t;
    return MSCEER_OK;
}

}  // extern "C"

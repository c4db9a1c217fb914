// C ABI of libmsceer_hip.so (declared in include/msceer.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "ctx.hpp"

using namespace msceer;

namespace msceer {

int hip_fail(hipError_t e) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return MSCEER_ENOMEM;
    return MSCEER_EHIP;
}

int check_dims(int64_t nx, int64_t ny, int64_t nz) {
    if (nx < 2 || ny < 2 || nz < 2) return MSCEER_EINVAL;
    if (nx > 65536 || ny > 65536 || nz > 65536) return MSCEER_ETOOBIG;
    const double cells = (2.0 * nx - 1) * (2.0 * ny - 1) * (2.0 * nz - 1);
    if (cells >= 4294967295.0) return MSCEER_ETOOBIG;
    return MSCEER_OK;
}

uint32_t* ctl_dev(msceer_ctx* ctx) { return static_cast<uint32_t*>(ctx->ctl.ptr); }

// Copy control words [first, first+n) to the host mirror and wait.
int read_ctl(msceer_ctx* ctx, int first, int n) {
    CHECK_HIP(hipMemcpyAsync(ctx->h_ctl + first, ctl_dev(ctx) + first, sizeof(uint32_t) * n,
                             hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return MSCEER_OK;
}

uint64_t list_total(const uint32_t* counts, const MList& l) {
    uint64_t t = std::min(counts[kSubs], l.spill_cap);
    for (int i = 0; i < kSubs; i++) t += std::min(counts[i], l.sub_cap);
    return t;
}

MList crit_list(msceer_ctx* ctx, const ListCaps& c, int which) {
    const size_t per = (size_t)c.list_sub * kSubs + c.list_spill;
    return MList{ctx->lists.as<uint32_t>() + which * per, ctl_dev(ctx) + (which ? kCtlList2 : kCtlList1), c.list_sub,
                 c.list_spill};
}

MList front_list(msceer_ctx* ctx, const ListCaps& c, int k) {
    uint32_t* buf = (k % 2 ? ctx->front1 : ctx->front0).as<uint32_t>();
    return MList{buf, ctl_dev(ctx) + kCtlRound + (k % 3) * kCountWords, c.front_sub, c.front_spill};
}

// One ldir pass (1: from the critical edges, 2: from the critical quads).
// Rounds are enqueued in groups without host synchronisation (each reads its
// counts on the device); after each group one read-back shows whether the
// frontier emptied.  The first group is short: the start round's work-groups
// usually finish their whole manifolds, and the round after it the rest.  A sweep rebuilds the frontier after any overflow of the
// start list or of a frontier.
int ldir_pass(msceer_ctx* ctx, const Dims& d, uint8_t* grad, const LdirScratch& ls, int pass, const MList& start,
              bool start_valid, const ListCaps& caps, int* rounds) {
    uint32_t* ctl = ctl_dev(ctx);
    uint32_t* overflow = ctl + kCtlOverflow;
    void* scratch = ctx->ldir.ptr;
    auto front = [&](int k) { return front_list(ctx, caps, k); };
    auto host_total = [&](int k) { return list_total(ctx->h_ctl + kCtlRound + (k % 3) * kCountWords, front(k)); };
    constexpr int kFirstGroup = 2, kGroup = 8;
    int group = kFirstGroup;
    MList in = start;
    bool is_start = true;
    int k = 0;
    bool need_sweep = !start_valid;
    CHECK_HIP(launch_zero(front(0).counts, kCountWords, ctx->stream));
    CHECK_HIP(launch_zero(overflow, 1, ctx->stream));
    for (int guard = 0; guard < 100000; guard++) {
        if (need_sweep) {
            CHECK_HIP(launch_zero(front(k).counts, kCountWords, ctx->stream));
            CHECK_HIP(launch_zero(overflow, 1, ctx->stream));
            CHECK_HIP(launch_ldir_sweep(d, grad, scratch, ls, pass, front(k), overflow, ctx->stream));
            int rc = read_ctl(ctx, 0, kCtlWords);
            if (rc) return rc;
            need_sweep = ctx->h_ctl[kCtlOverflow] != 0;
            (*rounds)++;
            const uint64_t n = host_total(k);
            if (n == 0 && !need_sweep) return MSCEER_OK;
            CHECK_HIP(launch_zero(overflow, 1, ctx->stream));
            in = front(k);
            is_start = false;
            k++;
            CHECK_HIP(launch_zero(front(k).counts, kCountWords, ctx->stream));
        }
        for (int j = 0; j < group; j++) {
            const MList out = front(k);
            CHECK_HIP(launch_ldir_round(d, scratch, ls, pass, is_start, in, out, front(k + 1).counts, overflow,
                                        ctx->stream));
            (*rounds)++;
            in = out;
            is_start = false;
            k++;
        }
        group = kGroup;
        int rc = read_ctl(ctx, 0, kCtlWords);
        if (rc) return rc;
        if (ctx->h_ctl[kCtlOverflow]) {
            need_sweep = true;
            continue;
        }
        if (host_total(k - 1) == 0) return MSCEER_OK;
    }
    return MSCEER_EHIP;
}

// Critical lists: a quarter of the vertices per sub-list set, spill for as
// many again as there are vertices; frontiers: the vertex count in sub-lists
// and as much again in spill.  Entries are cell ids (< 2^32).
int tuned_shape(msceer_ctx* ctx, const Dims& d, int fallback) {
    std::lock_guard<std::mutex> lk(ctx->tuning->mu);
    if (ctx->tuning->mode == 0) return 0;
    if (ctx->tuning->mode > 1) return ctx->tuning->mode - 2;
    auto it = ctx->tuning->brick_by_x.find(d.X);
    return it == ctx->tuning->brick_by_x.end() ? fallback : it->second;
}

int lstar_shape(msceer_ctx* ctx, const Dims& d, const LstarBuffers& b, int* shape) {
    Tuning& t = *ctx->tuning;
    {
        std::lock_guard<std::mutex> lk(t.mu);
        if (t.mode != 1) {
            *shape = t.mode == 0 ? 0 : t.mode - 2;
            return MSCEER_OK;
        }
        auto it = t.brick_by_x.find(d.X);
        if (it != t.brick_by_x.end()) {
            *shape = it->second;
            return MSCEER_OK;
        }
    }
    // time every shape twice on this volume, keep the fastest second run
    float best = 1e30f;
    int pick = 0;
    for (int sh = 0; sh < kLstarBricks; sh++) {
        float ms = 0.f;
        for (int rep = 0; rep < 2; rep++) {
            CHECK_HIP(launch_zero(ctl_dev(ctx), kCtlWords, ctx->stream));
            CHECK_HIP(hipEventRecord(ctx->ev[10], ctx->stream));
            CHECK_HIP(launch_lstar(d, b, ctx->stream, sh));
            CHECK_HIP(hipEventRecord(ctx->ev[11], ctx->stream));
            CHECK_HIP(hipEventSynchronize(ctx->ev[11]));
            CHECK_HIP(hipEventElapsedTime(&ms, ctx->ev[10], ctx->ev[11]));
        }
        if (ms < best) {
            best = ms;
            pick = sh;
        }
    }
    {
        std::lock_guard<std::mutex> lk(t.mu);
        t.brick_by_x[d.X] = pick;
    }
    *shape = pick;
    return MSCEER_OK;
}

ListCaps list_caps(const Dims& d) {
    ListCaps c{};
    const uint64_t list_cap = std::min<uint64_t>(kMaxListCap, std::max<uint64_t>(kMinListCap, d.verts / 4));
    const uint64_t front_cap = std::min<uint64_t>(kMaxFrontCap, std::max<uint64_t>(kMinListCap, d.verts));
    c.list_sub = (uint32_t)((list_cap + kSubs - 1) / kSubs);
    c.front_sub = (uint32_t)((front_cap + kSubs - 1) / kSubs);
    c.list_spill = (uint32_t)std::min<uint64_t>(0xFFFFFFFFull, std::max<uint64_t>(kMinListCap, d.verts));
    c.front_spill = c.list_spill;
    return c;
}

size_t list_bytes(const ListCaps& c) { return sizeof(uint32_t) * 2 * ((size_t)c.list_sub * kSubs + c.list_spill); }
size_t front_bytes(const ListCaps& c) { return sizeof(uint32_t) * ((size_t)c.front_sub * kSubs + c.front_spill); }

int ensure_lists(msceer_ctx* ctx, const Dims& d, ListCaps* caps) {
    *caps = list_caps(d);
    CHECK_HIP(ctx->lists.ensure(list_bytes(*caps)));
    CHECK_HIP(ctx->front0.ensure(front_bytes(*caps)));
    CHECK_HIP(ctx->front1.ensure(front_bytes(*caps)));
    return MSCEER_OK;
}

int run_ldir(msceer_ctx* ctx, const Dims& d, uint8_t* grad, const ListCaps& caps, bool lists_ok, int* rounds) {
    LdirScratch ls{};
    CHECK_HIP(ctx->ldir.ensure(ldir_scratch_bytes(d, &ls)));
    CHECK_HIP(launch_ldir_records(d, grad, ctx->ldir.ptr, ls, ctx->stream));
    int rc = ldir_pass(ctx, d, grad, ls, 1, crit_list(ctx, caps, 0), lists_ok, caps, rounds);
    if (rc) return rc;
    rc = ldir_pass(ctx, d, grad, ls, 2, crit_list(ctx, caps, 1), lists_ok, caps, rounds);
    if (rc) return rc;
    CHECK_HIP(launch_ldir_finish(d, grad, ctx->ldir.ptr, ls, ctx->stream));
    return MSCEER_OK;
}

int lstar_grad_device(msceer_ctx* ctx, const Dims& d, const float* d_values, uint8_t* d_grad, uint8_t* host_out,
                      uint32_t flags) {
    const bool timing = flags & MSCEER_TIMING;
    hipStream_t s = ctx->stream;
    // critical-cell lists and frontiers: sized to the volume, capped; any
    // overflow falls back to a sweep
    ListCaps caps{};
    int rc = ensure_lists(ctx, d, &caps);
    if (rc) return rc;
    uint32_t* ctl = ctl_dev(ctx);
    const MList list1 = crit_list(ctx, caps, 0), list2 = crit_list(ctx, caps, 1);

    LstarBuffers b{d_values, d_grad, nullptr, &list1, &list2, ctl + kCtlStats, ctl + kCtlOverflow};
    int shape = 0;
    if (flags & kLstarNoTuning) {
        shape = tuned_shape(ctx, d, 0);
    } else {
        rc = lstar_shape(ctx, d, b, &shape);
        if (rc) return rc;
    }
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[0], s));
    CHECK_HIP(launch_zero(ctl, kCtlWords, s));
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[1], s));
    CHECK_HIP(launch_lstar(d, b, s, shape));
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[2], s));
    rc = read_ctl(ctx, 0, kCtlWords);
    if (rc) return rc;
    msceer_timing t{};
    for (int i = 0; i < kSubs; i++)
        for (int k = 0; k < 4; k++) t.critical[k] += ctx->h_ctl[kCtlStats + 4 * i + k];
    const bool lists_ok = ctx->h_ctl[kCtlOverflow] == 0;

    int rounds = 0;
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[3], s));
    if (flags & MSCEER_TILED_LAYOUT) {
        CHECK_HIP(launch_clear_ldir_bits(d_grad, (int64_t)d.cells, s));
    } else if (!(flags & MSCEER_SKIP_LDIR)) {
        rc = run_ldir(ctx, d, d_grad, caps, lists_ok, &rounds);
        if (rc) return rc;
    }
    if (timing) CHECK_HIP(hipEventRecord(ctx->ev[4], s));
    if (host_out) CHECK_HIP(hipMemcpyAsync(host_out, d_grad, d.cells, hipMemcpyDeviceToHost, s));
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

}  // namespace msceer

extern "C" {

int msceer_version(void) { return MSCEER_VERSION; }

int msceer_set_autotune(msceer_ctx* ctx, int32_t mode) {
    if (!ctx || mode < 0 || mode > 1 + kLstarBricks) return MSCEER_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->tuning->mu);
    ctx->tuning->mode = mode;
    return MSCEER_OK;
}

int msceer_lstar_brick(msceer_ctx* ctx, int64_t nx) {
    if (!ctx) return MSCEER_EINVAL;
    std::lock_guard<std::mutex> lk(ctx->tuning->mu);
    if (ctx->tuning->mode != 1) return ctx->tuning->mode == 0 ? 0 : ctx->tuning->mode - 2;
    auto it = ctx->tuning->brick_by_x.find((int)nx);
    return it == ctx->tuning->brick_by_x.end() ? -1 : it->second;
}

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
    if (e == hipSuccess && noalloc_setup(device) != MSCEER_OK) e = hipErrorInvalidSymbol;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipHostMalloc((void**)&ctx->h_ctl, sizeof(uint32_t) * kCtlWords, hipHostMallocDefault);
    for (int i = 0; i < msceer_ctx::kEvents && e == hipSuccess; i++) e = hipEventCreate(&ctx->ev[i]);
    if (e == hipSuccess) e = ctx->ctl.ensure(sizeof(uint32_t) * kCtlWords);
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
    for (msceer_ctx* w : ctx->workers) msceer_destroy(w);
    ctx->workers.clear();
    if (ctx->aux) msceer_destroy(ctx->aux);
    ctx->aux = nullptr;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    batch_state_free(ctx);
    ctx->values.release();
    ctx->grad.release();
    ctx->lists.release();
    ctx->front0.release();
    ctx->front1.release();
    ctx->ctl.release();
    ctx->ldir.release();
    DevBuf* label_bufs[] = {&ctx->gradf, &ctx->gradd, &ctx->scan, &ctx->label, &ctx->label2, &ctx->dest, &ctx->seedmask, &ctx->redo,
                            &ctx->stamp, &ctx->rep, &ctx->rep2, &ctx->restriction, &ctx->host_io, &ctx->mask,
                            &ctx->label_dsc, &ctx->seedmask2, &ctx->line_stamps, &ctx->lines};
    for (DevBuf* b : label_bufs) b->release();
    ctx->host_stage.release();
    if (ctx->h_ctl) (void)hipHostFree(ctx->h_ctl);
    for (int i = 0; i < msceer_ctx::kEvents; i++)
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
    if (flags & ~(MSCEER_DEVICE_PTRS | MSCEER_SKIP_LDIR | MSCEER_TIMING | MSCEER_TILED_LAYOUT)) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const Dims d = make_dims(nx, ny, nz);
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const float* d_values = values;
    uint8_t* d_grad = grad_out;
    if (!dev) {
        CHECK_HIP(ctx->values.ensure(sizeof(float) * d.verts));
        CHECK_HIP(ctx->grad.ensure(d.cells));
        d_values = static_cast<const float*>(ctx->values.ptr);
        d_grad = static_cast<uint8_t*>(ctx->grad.ptr);
        CHECK_HIP(hipMemcpyAsync(ctx->values.ptr, values, sizeof(float) * d.verts, hipMemcpyHostToDevice,
                                 ctx->stream));
    }
    return lstar_grad_device(ctx, d, d_values, d_grad, dev ? nullptr : grad_out, flags);
}

}  // extern "C"

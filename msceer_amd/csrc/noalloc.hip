// msceer_lstar_grad_noalloc: the lower-star .grad of a volume in device memory
// without any device allocation.  Every scratch buffer of msceer_lstar_grad
// (control words, critical-cell lists, frontiers, ldir walk records and
// bitmaps) is carved out of one static device array per GPU, `scratch`, of
// kScratchBytes; inputs and outputs are the caller's.
#include <hip/hip_runtime.h>

#include <mutex>

#include "ctx.hpp"

using namespace msceer;

namespace {

constexpr size_t kScratchBytes = (size_t)48 << 20;  // per device, held until the process ends
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// Device address of `scratch` on each device (each GPU holds its own copy),
// from hipGetSymbolAddress with that device current; null until set up.
std::mutex g_setup_mu;
void* g_scratch[kMaxDevices] = {};
// One call at a time per device: the calls share that device's array.
std::mutex g_call_mu[kMaxDevices];

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

}  // namespace

namespace msceer {

// Called by msceer_create with `device` current: resolves this device's
// scratch address once.
int noalloc_setup(int device) {
    if (device < 0 || device >= kMaxDevices) return MSCEER_OK;  // the entry point then reports ENODEV
    std::lock_guard<std::mutex> lk(g_setup_mu);
    if (g_scratch[device]) return MSCEER_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return MSCEER_EHIP;
    g_scratch[device] = p;
    return MSCEER_OK;
}

}  // namespace msceer

extern "C" int msceer_lstar_grad_noalloc(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                                         uint8_t* grad_out, uint32_t flags) {
    if (!ctx || !values || !grad_out) return MSCEER_EINVAL;
    if (flags & ~(MSCEER_SKIP_LDIR | MSCEER_TIMING)) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return MSCEER_ENODEV;
    uint8_t* base = static_cast<uint8_t*>(g_scratch[ctx->device]);
    if (!base) return MSCEER_EHIP;
    const Dims d = make_dims(nx, ny, nz);
    // the scratch of this volume: control words, lists, two frontiers, ldir
    const ListCaps caps = list_caps(d);
    LdirScratch ls{};
    const size_t n_ctl = align256(sizeof(uint32_t) * kCtlWords), n_lists = align256(list_bytes(caps));
    const size_t n_front = align256(front_bytes(caps)), n_ldir = align256(ldir_scratch_bytes(d, &ls));
    if (n_ctl + n_lists + 2 * n_front + n_ldir > kScratchBytes) return MSCEER_ETOOBIG;
    std::lock_guard<std::mutex> lk(g_call_mu[ctx->device]);
    CHECK_HIP(hipSetDevice(ctx->device));
    // the context's own buffers are set aside for the call and views of the
    // scratch take their place, so the shared pipeline finds every buffer big
    // enough and allocates nothing
    DevBuf* bufs[5] = {&ctx->ctl, &ctx->lists, &ctx->front0, &ctx->front1, &ctx->ldir};
    const size_t sizes[5] = {n_ctl, n_lists, n_front, n_front, n_ldir};
    DevBuf saved[5];
    size_t off = 0;
    for (int i = 0; i < 5; i++) {
        saved[i] = *bufs[i];
        bufs[i]->borrow(base + off, sizes[i]);
        off += sizes[i];
    }
    rc = lstar_grad_device(ctx, d, values, grad_out, nullptr, flags | kLstarNoTuning);
    for (int i = 0; i < 5; i++) *bufs[i] = saved[i];
    return rc;
}

// C ABI of the near-separatrix classifier (msceer_cnn_*, include/msceer.h).
#include <hip/hip_runtime.h>

#include <new>
#include <vector>

#include "cnn.hpp"
#include "ctx.hpp"

using namespace msceer;

struct msceer_cnn {
    int device = 0;
    CnnState s;
};

namespace {

int check_cnn_dims(int64_t nx, int64_t ny, int64_t nz) {
    if (nx < 1 || ny < 1 || nz < 1) return MSCEER_EINVAL;
    if (nx > 65536 || ny > 65536 || nz > 65536) return MSCEER_ETOOBIG;
    if (nx * ny * nz >= ((int64_t)1 << 34)) return MSCEER_ETOOBIG;
    return MSCEER_OK;
}

// Device view of a host or device input of n elements.
template <typename T>
int stage(msceer_ctx* ctx, DevBuf& buf, const T* src, size_t n, bool dev, const T** out) {
    if (dev) {
        *out = src;
        return MSCEER_OK;
    }
    CHECK_HIP(buf.ensure(sizeof(T) * n));
    CHECK_HIP(hipMemcpyAsync(buf.ptr, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream));
    *out = buf.as<T>();
    return MSCEER_OK;
}

int upload_params(msceer_ctx* ctx, msceer_cnn* net, const float* p) {
    CnnState& s = net->s;
    CHECK_HIP(hipMemcpyAsync(s.params.ptr, p, sizeof(float) * kNumParams, hipMemcpyHostToDevice, ctx->stream));
    CHECK_HIP(hipMemsetAsync(s.m.ptr, 0, sizeof(float) * kNumParams, ctx->stream));
    CHECK_HIP(hipMemsetAsync(s.v.ptr, 0, sizeof(float) * kNumParams, ctx->stream));
    s.step = 0;
    CHECK_HIP(cnn_pack(s, ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return MSCEER_OK;
}

}  // namespace

extern "C" {

int msceer_cnn_num_params(void) { return kNumParams; }

int msceer_cnn_create(msceer_ctx* ctx, uint64_t seed, msceer_cnn** out) {
    if (!ctx || !out) return MSCEER_EINVAL;
    *out = nullptr;
    CHECK_HIP(hipSetDevice(ctx->device));
    msceer_cnn* net = new (std::nothrow) msceer_cnn;
    if (!net) return MSCEER_ENOMEM;
    net->device = ctx->device;
    CnnState& s = net->s;
    hipError_t e = hipSuccess;
    if ((e = s.params.ensure(sizeof(float) * kNumParams)) != hipSuccess ||
        (e = s.grads.ensure(sizeof(float) * kNumParams)) != hipSuccess ||
        (e = s.m.ensure(sizeof(float) * kNumParams)) != hipSuccess ||
        (e = s.v.ensure(sizeof(float) * kNumParams)) != hipSuccess ||
        (e = s.pk.ensure(kPackedBytes)) != hipSuccess ||
        (e = s.parts.ensure(sizeof(float) * (size_t)kPartWidth * kGradBlocks)) != hipSuccess ||
        (e = s.loss.ensure(sizeof(float) * 32)) != hipSuccess) {
        s.release_all();
        delete net;
        return hip_fail(e);
    }
    const std::vector<float> p = cnn_initial_params(seed);
    const int rc = upload_params(ctx, net, p.data());
    if (rc) {
        s.release_all();
        delete net;
        return rc;
    }
    *out = net;
    return MSCEER_OK;
}

int msceer_cnn_destroy(msceer_cnn* net) {
    if (!net) return MSCEER_EINVAL;
    (void)hipSetDevice(net->device);
    (void)hipDeviceSynchronize();
    net->s.release_all();
    delete net;
    return MSCEER_OK;
}

int msceer_cnn_get_params(msceer_ctx* ctx, msceer_cnn* net, float* out) {
    if (!ctx || !net || !out) return MSCEER_EINVAL;
    CHECK_HIP(hipSetDevice(ctx->device));
    CHECK_HIP(hipMemcpyAsync(out, net->s.params.ptr, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return MSCEER_OK;
}

int msceer_cnn_set_params(msceer_ctx* ctx, msceer_cnn* net, const float* in) {
    if (!ctx || !net || !in) return MSCEER_EINVAL;
    CHECK_HIP(hipSetDevice(ctx->device));
    return upload_params(ctx, net, in);
}

int msceer_cnn_predict(msceer_ctx* ctx, msceer_cnn* net, const float* values, int64_t nx, int64_t ny, int64_t nz,
                       float* prob_out, uint32_t flags) {
    if (!ctx || !net || !values || !prob_out) return MSCEER_EINVAL;
    if (flags & ~MSCEER_DEVICE_PTRS) return MSCEER_EINVAL;
    int rc = check_cnn_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const int64_t V = nx * ny * nz;
    CnnState& s = net->s;
    CHECK_HIP(s.ensure_volume(V));
    const float* d_values = nullptr;
    rc = stage(ctx, ctx->values, values, (size_t)V, dev, &d_values);
    if (rc) return rc;
    CHECK_HIP(cnn_forward(s, d_values, (int)nx, (int)ny, (int)nz, false, ctx->stream));
    CHECK_HIP(hipMemcpyAsync(prob_out, s.prob.ptr, sizeof(float) * V, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                             ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return MSCEER_OK;
}

int msceer_cnn_gradient(msceer_ctx* ctx, msceer_cnn* net, const float* values, int64_t nx, int64_t ny, int64_t nz,
                        const uint8_t* target, float pos_weight, float* loss_out, float* grad_out, uint32_t flags) {
    if (!ctx || !net || !values || !target) return MSCEER_EINVAL;
    if (flags & ~MSCEER_DEVICE_PTRS) return MSCEER_EINVAL;
    if (!(pos_weight > 0.f)) return MSCEER_EINVAL;
    int rc = check_cnn_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const int64_t V = nx * ny * nz;
    CnnState& s = net->s;
    CHECK_HIP(s.ensure_volume(V));
    CHECK_HIP(s.ensure_train(V));
    const float* d_values = nullptr;
    rc = stage(ctx, ctx->values, values, (size_t)V, dev, &d_values);
    if (rc) return rc;
    const uint8_t* d_target = nullptr;
    rc = stage(ctx, ctx->mask, target, (size_t)V, dev, &d_target);
    if (rc) return rc;
    CHECK_HIP(cnn_forward(s, d_values, (int)nx, (int)ny, (int)nz, true, ctx->stream));
    CHECK_HIP(cnn_backward(s, (int)nx, (int)ny, (int)nz, d_target, pos_weight, ctx->stream));
    if (loss_out) CHECK_HIP(hipMemcpyAsync(loss_out, s.loss.ptr, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    if (grad_out)
        CHECK_HIP(hipMemcpyAsync(grad_out, s.grads.ptr, sizeof(float) * kNumParams, hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (loss_out) *loss_out /= (float)V;
    return MSCEER_OK;
}

int msceer_cnn_adam_step(msceer_ctx* ctx, msceer_cnn* net, float lr) {
    if (!ctx || !net || !(lr > 0.f)) return MSCEER_EINVAL;
    CHECK_HIP(hipSetDevice(ctx->device));
    CHECK_HIP(cnn_adam(net->s, lr, ctx->stream));
    return MSCEER_OK;
}

int msceer_cnn_mask(msceer_ctx* ctx, msceer_cnn* net, const float* values, int64_t nx, int64_t ny, int64_t nz,
                    float threshold, int32_t dilation, uint8_t* mask_out, int64_t* flagged, uint32_t flags) {
    if (!ctx || !net || !values || !mask_out) return MSCEER_EINVAL;
    if (flags & ~MSCEER_DEVICE_PTRS) return MSCEER_EINVAL;
    if (dilation < 0 || dilation > 16) return MSCEER_EINVAL;
    int rc = check_cnn_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const int64_t V = nx * ny * nz;
    CnnState& s = net->s;
    CHECK_HIP(s.ensure_volume(V));
    const float* d_values = nullptr;
    rc = stage(ctx, ctx->values, values, (size_t)V, dev, &d_values);
    if (rc) return rc;
    uint8_t* d_mask = mask_out;
    if (!dev) {
        CHECK_HIP(ctx->mask.ensure(V));
        d_mask = ctx->mask.as<uint8_t>();
    }
    unsigned long long* d_count = reinterpret_cast<unsigned long long*>(s.loss.as<float>() + 24);
    CHECK_HIP(hipMemsetAsync(d_count, 0, sizeof(unsigned long long), ctx->stream));
    CHECK_HIP(cnn_forward(s, d_values, (int)nx, (int)ny, (int)nz, false, ctx->stream));
    CHECK_HIP(cnn_mask(s, (int)nx, (int)ny, (int)nz, threshold, dilation, d_mask, d_count, ctx->stream));
    if (!dev) CHECK_HIP(hipMemcpyAsync(mask_out, d_mask, V, hipMemcpyDeviceToHost, ctx->stream));
    unsigned long long cnt = 0;
    CHECK_HIP(hipMemcpyAsync(&cnt, d_count, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (flagged) *flagged = (int64_t)cnt;
    return MSCEER_OK;
}

}  // extern "C"

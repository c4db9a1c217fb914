// msceer_batch: many independent volumes of different sizes on one GPU.
//
// Worker threads each drive their own context (own HIP stream, own device
// scratch, shared brick tuning) and take volumes largest first from a shared
// counter, so the host synchronisations of one volume's pipeline overlap with
// kernels of the others and the uneven volume sizes balance out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

#include "ctx.hpp"

using namespace msceer;

extern "C" int msceer_batch(msceer_ctx* ctx, msceer_volume* vols, int32_t n, int32_t streams, float err_thr,
                            float grad_thr, int32_t max_iters, float presimp, uint32_t flags) {
    if (!ctx || (!vols && n > 0) || n < 0) return MSCEER_EINVAL;
    if (flags != 0) return MSCEER_EINVAL;  // host buffers only
    for (int32_t i = 0; i < n; i++) {
        const msceer_volume& v = vols[i];
        if (!v.values || (!v.grad_out && !v.labels_out)) return MSCEER_EINVAL;
        const int rc = check_dims(v.nx, v.ny, v.nz);
        if (rc) return rc;
    }
    if (n == 0) return MSCEER_OK;
    if (streams <= 0) streams = 16;
    streams = std::min(std::min(streams, 32), n);
    // worker contexts persist in the parent context between calls
    while ((int)ctx->workers.size() < streams) {
        msceer_ctx* w = nullptr;
        const int rc = msceer_create(&w, ctx->device);
        if (rc) return rc;
        w->tuning = ctx->tuning;
        ctx->workers.push_back(w);
    }
    std::vector<int32_t> order(n);
    for (int32_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
        return vols[a].nx * vols[a].ny * vols[a].nz > vols[b].nx * vols[b].ny * vols[b].nz;
    });
    std::atomic<int32_t> next{0};
    auto work = [&](msceer_ctx* w) {
        while (true) {
            const int32_t k = next.fetch_add(1);
            if (k >= n) return;
            msceer_volume& v = vols[order[k]];
            const auto t0 = std::chrono::steady_clock::now();
            int rc;
            if (v.labels_out)
                rc = msceer_ondemand_asc(w, v.values, v.nx, v.ny, v.nz, err_thr, grad_thr, max_iters, presimp,
                                         nullptr, v.grad_out, v.labels_out, nullptr, 0);
            else
                rc = msceer_lstar_grad(w, v.values, v.nx, v.ny, v.nz, v.grad_out, 0);
            const auto t1 = std::chrono::steady_clock::now();
            v.ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
            v.status = rc;
        }
    };
    std::vector<std::thread> threads;
    threads.reserve(streams);
    for (int32_t k = 0; k < streams; k++) threads.emplace_back(work, ctx->workers[k]);
    for (auto& t : threads) t.join();
    for (int32_t i = 0; i < n; i++)
        if (vols[i].status) return vols[i].status;
    return MSCEER_OK;
}

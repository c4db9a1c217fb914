// msceer_ondemand_asc: accurate ascending labelling and its conforming .grad
// (orchestration of the kernels in lstar.hip, ldir.hip and integrate.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <thread>
#include <vector>

#include "ctx.hpp"
#include "extgraph.hpp"
#include "integrate.hpp"
#include "lines.hpp"

using namespace msceer;

namespace msceer {
namespace {

// Run an iterative kernel in groups of `group` launches until a whole group
// reports no change (flag word in the control block).
template <typename Launch>
int iterate_until_stable(msceer_ctx* ctx, int group, int max_groups, int* count, Launch launch) {
    uint32_t* flag = ctl_dev(ctx) + kCtlFlag;
    for (int gi = 0; gi < max_groups; gi++) {
        CHECK_HIP(launch_zero(flag, 1, ctx->stream));
        for (int j = 0; j < group; j++) {
            CHECK_HIP(launch(flag));
            (*count)++;
        }
        int rc = read_ctl(ctx, kCtlFlag, 1);
        if (rc) return rc;
        if (ctx->h_ctl[kCtlFlag] == 0) return MSCEER_OK;
    }
    return MSCEER_EHIP;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

// IsolatedCCRegionRemoverNEW::ComputeOutput (gi_isolated_region_remover.h:155-192)
// on b.label, in place.  Same-label components (root = earliest vertex) on the
// GPU; the roots then join the region of their earliest lower neighbour with
// another label, in vertex-id order with every earlier merge visible, which
// is a sequential scan over the (few) roots on the host.
int run_cleanup(msceer_ctx* ctx, const VGrid& g, const LabelBuffers& b) {
    hipStream_t s = ctx->stream;
    const int64_t V = g.verts();
    CHECK_HIP(launch_cc_init(g, b, s));
    CHECK_HIP(launch_cc_union(g, b, s));
    uint32_t* count = ctl_dev(ctx) + kCtlCount;
    // candidate roots (all others keep their region) and their slots in
    // vertex order; b.stamp is free here
    uint32_t* cand = b.stamp;
    size_t temp_bytes = 0;
    CHECK_HIP(launch_exclusive_sum(cand, nullptr, V, nullptr, &temp_bytes, s));
    CHECK_HIP(ctx->scan.ensure(sizeof(uint32_t) * V + temp_bytes + 256));
    uint32_t* slot_of = ctx->scan.as<uint32_t>();
    void* temp = reinterpret_cast<uint8_t*>(ctx->scan.ptr) + ((sizeof(uint32_t) * V + 255) & ~(size_t)255);
    CHECK_HIP(launch_zero(count, 1, s));
    CHECK_HIP(launch_cc_candidates(g, b, cand, count, s));
    CHECK_HIP(launch_exclusive_sum(cand, slot_of, V, temp, &temp_bytes, s));
    int rc = read_ctl(ctx, kCtlCount, 1);
    if (rc) return rc;
    const uint32_t R = ctx->h_ctl[kCtlCount];
    const size_t row = sizeof(int32_t) * kCollectStride;
    // rows of the candidates, then the (vertex, new label) pairs (pinned
    // staging, not cleared: every entry read is written first)
    CHECK_HIP(ctx->host_stage.ensure(row * R + sizeof(int32_t) * 2 * (size_t)R + 64));
    const int32_t* h = ctx->host_stage.as<int32_t>();
    int32_t* pairs = ctx->host_stage.as<int32_t>() + (size_t)R * kCollectStride;
    if (R) {
        CHECK_HIP(ctx->host_io.ensure(std::max(row * R, sizeof(int32_t) * 2 * (size_t)R)));
        CHECK_HIP(launch_cc_collect(g, b, cand, slot_of, ctx->host_io.as<int32_t>(), s));
        CHECK_HIP(hipMemcpyAsync(ctx->host_stage.as<void>(), ctx->host_io.ptr, row * R, hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipStreamSynchronize(s));
    }
    // candidates in vertex-id order; a set of candidates merged together has
    // a representative q, and term[q] >= 0 once the set joined a region that
    // no candidate represents (that root, with label termlab[q])
    std::vector<int32_t> parent(R), term(R, -1), termlab(R, -1);
    for (uint32_t k = 0; k < R; k++) parent[k] = (int32_t)k;
    auto find = [&](int32_t x) {
        int32_t r = x;
        while (parent[r] != r) r = parent[r];
        while (parent[x] != r) { const int32_t n = parent[x]; parent[x] = r; x = n; }
        return r;
    };
    auto region_label = [&](int32_t q) { return term[q] >= 0 ? termlab[q] : h[(size_t)q * kCollectStride + 1]; };
    for (uint32_t k = 0; k < R; k++) {
        const int32_t* e = &h[(size_t)k * kCollectStride];
        const int32_t lab = e[1];
        for (int j = 0; j < 6 && e[2 + j] >= 0; j++) {
            if (e[8 + j] >= 0) {
                const int32_t fr = find(e[8 + j]);
                if (region_label(fr) == lab) continue;
                parent[k] = fr;  // k is still its set's representative
            } else {
                if (e[14 + j] == lab) continue;  // a root that never moves: its own label
                term[k] = e[2 + j];
                termlab[k] = e[14 + j];
            }
            break;
        }
    }
    for (uint32_t k = 0; k < R; k++) {
        const int32_t q = find((int32_t)k);
        pairs[2 * (size_t)k] = h[(size_t)k * kCollectStride];
        pairs[2 * (size_t)k + 1] = term[q] >= 0 ? term[q] : h[(size_t)q * kCollectStride];
    }
    if (R) {
        CHECK_HIP(hipMemcpyAsync(ctx->host_io.ptr, pairs, sizeof(int32_t) * 2 * (size_t)R, hipMemcpyHostToDevice, s));
        CHECK_HIP(launch_cc_scatter(ctx->host_io.as<int32_t>(), R, b.rep2, s));
    }
    CHECK_HIP(launch_cc_apply(g, b, s));
    // the staging buffer is reused by the next call: wait for the copy
    CHECK_HIP(hipStreamSynchronize(s));
    return MSCEER_OK;
}

// Passes 1-3 of the integrator (NumericIntegratorPathCompressingToTerminal::
// BeginIntegration, gi_numeric_integrator_path_compressing.h:75-433): labels
// >= 0 in b.label on entry are certain terminals, b.gradf and b.seedmask are
// set, the integrator counters are zero.  With d_mask, vertices outside the
// mask take their combinatorial label from d_grad and count as assigned.
// sign -1 integrates ascending manifolds (downhill), +1 descending (uphill).
int run_integration(msceer_ctx* ctx, const VGrid& g, const LabelBuffers& b, const uint8_t* d_mask,
                    const uint8_t* d_grad, float grad_thr, float err_thr, int32_t max_iters, int sign, bool fp32_first,
                    msceer_label_stats& st) {
    hipStream_t s = ctx->stream;
    uint32_t* ctl = ctl_dev(ctx);
    const int64_t V = g.verts();
    CHECK_HIP(launch_dest_init(g, b, s));
    if (d_mask) {
        // vertices outside the mask: combinatorial label, treated as assigned
        CHECK_HIP(launch_combinatorial_labels(g, d_grad, b.label, b.label2, nullptr, s));
        CHECK_HIP(launch_apply_mask(g, b, d_mask, s));
    }
    // coarse-to-fine strides (the reference visits each thread's range in a
    // bisection order): 2, then 1.  Each stride's launch lasts as long as its
    // longest streamline.
    // Pass 1 is the FP32 first guess: a streamline that exhausts this budget is
    // flagged and re-traced in FP64 by pass 3 with the full max_iters, as the
    // ones exhausting max_iters are.  The few that run that long set each
    // stride's length (256^3: 47.6 -> 43.2 ms for the whole call, FP64-oracle
    // mismatch 0.050 % either way; 537 of 5.0 M streamlines hit it).
    constexpr int kPass1Budget = 1024;
    constexpr int kPass1TopStride = 2;
    const int max_dim = std::max(g.X, std::max(g.Y, g.Z));
    int top = 1;
    while (top * 2 <= kPass1TopStride && top * 2 < max_dim) top *= 2;
    CHECK_HIP(hipEventRecord(ctx->ev[12], s));
    for (int stride = top; stride >= 1; stride /= 2) {
        CHECK_HIP(launch_pass1(g, b, stride, stride == top, grad_thr, err_thr, std::min(max_iters, kPass1Budget),
                               sign, s));
        CHECK_HIP(launch_pass1_finish(g, b, stride, stride == top, stride == 1, s));
    }
    CHECK_HIP(hipEventRecord(ctx->ev[13], s));
    // pass 2 fills the re-integration queue, pass 3 drains it in one launch
    // each vertex is queued at most once as a re-integration and at most once
    // to be traced ahead, so up to 2V items: the queue takes the whole
    // frontier buffer (at least V words; items beyond it report overflow)
    const uint64_t qcap = std::min<uint64_t>(std::min<uint64_t>(ctx->front0.bytes / 4, 2 * (uint64_t)V), 0xFFFFFFFFull);
    const WorkQueue q{ctx->front0.as<uint32_t>(), (uint32_t)qcap, ctl + kCtlCursor, ctl + kCtlCursor + 1,
                      ctl + kCtlOverflow, ctl + kCtlCursor + 2, ctl + kCtlCursor + 3, ctl + kCtlCursor + 4};
    CHECK_HIP(launch_fill_u32(q.items, (int64_t)qcap, 0xFFFFFFFFu, s));
    CHECK_HIP(launch_zero(ctl + kCtlCursor, 5, s));
    CHECK_HIP(launch_zero(ctl + kCtlOverflow, 1, s));
    CHECK_HIP(launch_pass2(g, b, q, s));
    // the labels pass 3 reads for its results
    CHECK_HIP(hipMemcpyAsync(const_cast<int32_t*>(b.snap), b.label, sizeof(int32_t) * V, hipMemcpyDeviceToDevice, s));
    CHECK_HIP(launch_grad_field64(g, b, s));
    CHECK_HIP(launch_fill_i32(b.label2, V, -1, s));  // no vertex traced ahead yet
    CHECK_HIP(hipEventRecord(ctx->ev[14], s));
    CHECK_HIP(launch_pass3(g, b, q, fp32_first, grad_thr, err_thr, max_iters, sign, s));
    CHECK_HIP(hipEventRecord(ctx->ev[15], s));
    int rc = read_ctl(ctx, 0, kCtlWords);
    if (rc) return rc;
    st.ms_pass1 = elapsed(ctx->ev[12], ctx->ev[13]);
    st.ms_pass3 = elapsed(ctx->ev[14], ctx->ev[15]);
    if (ctx->h_ctl[kCtlOverflow] & 2u) return MSCEER_EHIP;  // the queue stalled (never expected)
    if (ctx->h_ctl[kCtlOverflow]) return MSCEER_ETOOBIG;
    st.pass3_rounds = 1;
    st.traced_fp32 = ctx->h_ctl[kCtlIntStats + kStatTraced1];
    st.over_budget_fp32 = ctx->h_ctl[kCtlIntStats + kStatOver1];
    st.reintegrated_pass3 = ctx->h_ctl[kCtlIntStats + kStatTraced3f];
    st.escalated_fp64 = ctx->h_ctl[kCtlIntStats + kStatDoubt];
    st.reintegrated_fp64 = fp32_first ? st.escalated_fp64 : st.reintegrated_pass3;
    st.over_budget_fp64 = ctx->h_ctl[kCtlIntStats + kStatOver3];
    st.integrated_vertices = d_mask ? (int64_t)ctx->h_ctl[kCtlIntStats + kStatMasked] : V;
    auto u64 = [&](int k) {
        return (int64_t)((uint64_t)ctx->h_ctl[kCtlIntStats + k] | ((uint64_t)ctx->h_ctl[kCtlIntStats + k + 1] << 32));
    };
    st.steps_fp32 = u64(kStatSteps1) + u64(kStatSteps3f);
    st.steps_fp64 = u64(kStatSteps3);
    st.ahead_fp64 = ctx->h_ctl[kCtlIntStats + kStatAhead3];
    return MSCEER_OK;
}

}  // namespace
}  // namespace msceer

namespace {

// OnDemandAccurate's accuracy branches: the steepest pairing; the accurate
// ascending labelling and its MIN boundary (needASC2); the digitized
// 1-manifolds (needASC1, needDSC1); the conforming lower stars and ldir.
// Certain regions from the labelled seeds (b.label, b.seedmask): rounds in
// dependency order (certain_init / certain_round), enqueued in groups; a
// read-back per group shows whether the next round's list is empty.  Uses
// the frontier buffers (lists) and b.stamp; same_bdry: SameBdry6 neighbours.
int run_certain_rounds(msceer_ctx* ctx, const VGrid& g, const float* d_values, const LabelBuffers& b, bool same_bdry,
                       int* sweeps_out, bool desc = false) {
    hipStream_t s = ctx->stream;
    uint32_t* ctl = ctl_dev(ctx);
    const int64_t V = g.verts();
    int rc;
    uint64_t* lists[2] = {ctx->front0.as<uint64_t>(), ctx->front1.as<uint64_t>()};
    const uint32_t lcap = (uint32_t)std::min<uint64_t>(std::min(ctx->front0.bytes, ctx->front1.bytes) / 8, V);
    uint32_t* counts = ctl + kCtlRound;  // three rotating count words
    CHECK_HIP(launch_zero(counts, 3, s));
    CHECK_HIP(launch_zero(ctl + kCtlOverflow, 1, s));
    CHECK_HIP(launch_certain_init(g, d_values, b.seedmask, b.stamp, lists[0], counts, lcap, ctl + kCtlOverflow, s,
                                  same_bdry, desc));
    rc = read_ctl(ctx, 0, kCtlWords);
    if (rc) return rc;
    int k = 0;
    // Groups are pipelined: group i+1 is enqueued before the host reads the
    // list length that group i left (copied to pinned memory behind it), so
    // the GPU does not wait for the host between groups; the group after
    // the last non-empty list runs empty rounds.  The launches of a group
    // are sized by the list length last read back.
    uint64_t expected = ctx->h_ctl[kCtlRound];
    CHECK_HIP(ctx->host_stage.ensure(8 * sizeof(uint32_t)));
    const volatile uint32_t* seen = ctx->host_stage.as<uint32_t>();  // two slots of the three counts
    int done = 0, kprev = 0;
    for (int guard = 0; expected > 0; guard++) {
        if (guard > (int)(V / 16 + 16)) return MSCEER_EHIP;  // a round resolves at least one vertex
        constexpr int kGroup = 16;
        for (int j = 0; j < kGroup; j++, k++)
            CHECK_HIP(launch_certain_round(g, b.label, b.stamp, lists[k % 2], counts + k % 3, lists[(k + 1) % 2],
                                           counts + (k + 1) % 3, counts + (k + 2) % 3, lcap, ctl + kCtlOverflow,
                                           expected, s));
        const int slot = guard & 1;
        CHECK_HIP(hipMemcpyAsync(ctx->host_stage.as<uint32_t>() + 4 * slot, counts, 3 * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipEventRecord(ctx->ev[8 + slot], s));
        if (guard > 0) {
            CHECK_HIP(hipEventSynchronize(ctx->ev[8 + (slot ^ 1)]));
            const uint32_t left = seen[4 * (slot ^ 1) + kprev % 3];
            if (left == 0) {
                done = kprev;
                break;
            }
            expected = left;
        }
        kprev = k;
    }
    rc = read_ctl(ctx, 0, kCtlWords);  // also waits for the last group
    if (rc) return rc;
    if (ctx->h_ctl[kCtlOverflow]) return MSCEER_ETOOBIG;
    *sweeps_out = done ? done : k;
    return MSCEER_OK;
}

// Inputs of the descending branch.  The values and the base .grad are read
// only by both branches; list2 holds the base pass's critical quads (n_sad
// of them, counted in the calling context); asc_gradf / asc_seedmask are the
// ascending branch's FP32 field and seed mask after its seeding (null
// without needASC2), ready once `ready` has completed.
struct DscIn {
    VGrid g;
    Dims d;
    const float* values;
    const uint8_t* grad;
    MList list2;
    uint64_t n_sad;
    const float4* asc_gradf;
    const uint8_t* asc_seedmask;
    hipEvent_t ready;
    float presimp, grad_thr, err_thr;
    int32_t max_iters;
    bool fp32_first, stop_r, stop_i;
};

// Descending 3-manifolds (needDSC2, ondemandaccurate.cxx:913-991) on the
// context `cx` (its own stream and buffers, so that it runs beside the
// ascending branch): the max extremum graph of the steepest pairing
// (critical hexes joined by interior critical quads), certain regions in the
// descending order, uphill integration and cleanup in that order.  The
// gradient field is the one the ascending branch left (zeroed at its seeds),
// zeroed at the maxima too, as the reference's shared field is.  Labels in
// cx->label; cx->ev[17] completes with the branch.
int run_dsc(msceer_ctx* cx, const DscIn& in, msceer_label_stats& st) {
    CHECK_HIP(hipSetDevice(cx->device));
    hipStream_t s = cx->stream;
    const VGrid& g = in.g;
    const int64_t V = g.verts();
    CHECK_HIP(hipStreamWaitEvent(s, in.ready, 0));
    // frontiers only: the critical lists stay the calling context's
    const ListCaps caps = list_caps(in.d);
    CHECK_HIP(cx->front0.ensure(front_bytes(caps)));
    CHECK_HIP(cx->front1.ensure(front_bytes(caps)));
    int rc = MSCEER_OK;
    CHECK_HIP(cx->gradf.ensure(sizeof(float4) * V));
    CHECK_HIP(cx->gradd.ensure(sizeof(double4) * V));
    DevBuf* i32[] = {&cx->label, &cx->label2, &cx->stamp, &cx->rep, &cx->rep2};
    for (DevBuf* b : i32) CHECK_HIP(b->ensure(sizeof(int32_t) * V));
    DevBuf* u8[] = {&cx->dest, &cx->seedmask, &cx->redo};
    for (DevBuf* b : u8) CHECK_HIP(b->ensure(V));
    uint32_t* ctl = ctl_dev(cx);
    LabelBuffers bd{};
    bd.values = in.values;
    bd.gradf = cx->gradf.as<float4>();
    bd.gradd = cx->gradd.as<double4>();
    bd.label = cx->label.as<int32_t>();
    bd.label2 = cx->label2.as<int32_t>();
    bd.dest = cx->dest.as<uint8_t>();
    bd.seedmask = cx->seedmask.as<uint8_t>();
    bd.redo = cx->redo.as<uint8_t>();
    bd.stamp = cx->stamp.as<uint32_t>();
    bd.rep = cx->rep.as<int32_t>();
    bd.rep2 = cx->rep2.as<int32_t>();
    bd.snap = bd.rep;  // free until the cleanup
    bd.stats = ctl + kCtlIntStats;
    bd.desc = 1;
    CHECK_HIP(hipEventRecord(cx->ev[16], s));
    CHECK_HIP(launch_zero(ctl, kCtlWords, s));
    const MList maxl = front_list(cx, caps, 0);
    CHECK_HIP(launch_collect_maxima(g, in.values, in.grad, maxl, ctl + kCtlOverflow, s));
    std::vector<MaxArcRecord> rec;
    if (in.n_sad) {
        CHECK_HIP(cx->host_io.ensure(sizeof(MaxArcRecord) * in.n_sad + 64));
        MaxArcRecord* d_arcs = cx->host_io.as<MaxArcRecord>();
        CHECK_HIP(launch_saddle2_arcs(g, in.values, in.grad, in.list2, in.presimp, d_arcs, (uint32_t)in.n_sad,
                                      ctl + kCtlCount, s));
        rc = read_ctl(cx, 0, kCtlWords);
        if (rc) return rc;
        rec.resize(std::min<uint64_t>(cx->h_ctl[kCtlCount], in.n_sad));
        if (!rec.empty()) {
            CHECK_HIP(hipMemcpyAsync(rec.data(), d_arcs, sizeof(MaxArcRecord) * rec.size(), hipMemcpyDeviceToHost, s));
            CHECK_HIP(hipStreamSynchronize(s));
        }
    } else {
        rc = read_ctl(cx, 0, kCtlWords);
        if (rc) return rc;
    }
    if (cx->h_ctl[kCtlOverflow]) return MSCEER_ETOOBIG;
    const uint64_t n_max = list_total(cx->h_ctl + kCtlRound, maxl);
    std::vector<MinNode> nodes;
    nodes.reserve(2 * rec.size());
    std::vector<SaddleArc> arcs(rec.size());
    for (size_t i = 0; i < rec.size(); i++) {
        nodes.push_back(MinNode{rec[i].max1, rec[i].v1, rec[i].h1});
        nodes.push_back(MinNode{rec[i].max2, rec[i].v2, rec[i].h2});
        arcs[i] = SaddleArc{rec[i].saddle, rec[i].max1, rec[i].max2, rec[i].sval};
    }
    std::sort(nodes.begin(), nodes.end(), [](const MinNode& x, const MinNode& y) { return x.cell < y.cell; });
    nodes.erase(std::unique(nodes.begin(), nodes.end(),
                            [](const MinNode& x, const MinNode& y) { return x.cell == y.cell; }),
                nodes.end());
    const std::vector<int> rep = simplify_maxima(nodes, arcs, in.presimp);
    std::vector<int32_t> seed_v, seed_l;
    for (size_t i = 0; i < nodes.size(); i++) {
        if (rep[i] == (int)i) continue;
        seed_v.push_back((int32_t)nodes[i].vertex);
        seed_l.push_back((int32_t)nodes[rep[i]].vertex);
    }
    st.maxima = (int64_t)n_max;
    st.saddles2 = (int64_t)in.n_sad;
    st.regions_dsc = (int64_t)n_max - (int64_t)seed_v.size();
    CHECK_HIP(launch_fill_i32(bd.label, V, -1, s));
    CHECK_HIP(launch_fill_u8(bd.seedmask, V, 0, s));
    if (in.asc_gradf)
        CHECK_HIP(hipMemcpyAsync(bd.gradf, in.asc_gradf, sizeof(float4) * V, hipMemcpyDeviceToDevice, s));
    else
        CHECK_HIP(launch_grad_field(g, in.values, bd.gradf, s));
    CHECK_HIP(launch_seed_minima(maxl, bd, s));  // every maximum's highest vertex: its own label
    if (!seed_v.empty()) {
        CHECK_HIP(cx->host_io.ensure(2 * sizeof(int32_t) * seed_v.size() + 64));
        int32_t* d_sv = cx->host_io.as<int32_t>();
        int32_t* d_sl = d_sv + seed_v.size();
        CHECK_HIP(hipMemcpyAsync(d_sv, seed_v.data(), sizeof(int32_t) * seed_v.size(), hipMemcpyHostToDevice, s));
        CHECK_HIP(hipMemcpyAsync(d_sl, seed_l.data(), sizeof(int32_t) * seed_l.size(), hipMemcpyHostToDevice, s));
        CHECK_HIP(launch_seed_scatter(d_sv, d_sl, (int)seed_v.size(), bd, s));
        CHECK_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
    }
    int sweeps = 0;
    rc = run_certain_rounds(cx, g, in.values, bd, false, &sweeps, true);
    if (rc) return rc;
    st.certain_sweeps_dsc = sweeps;
    if (!in.stop_r) {
        // the FP64 field of pass 3 is zeroed at every seed so far
        if (in.asc_seedmask) CHECK_HIP(launch_or_u8(bd.seedmask, in.asc_seedmask, V, s));
        CHECK_HIP(launch_zero(ctl + kCtlIntStats, kStatNum, s));
        msceer_label_stats sd{};
        rc = run_integration(cx, g, bd, nullptr, nullptr, in.grad_thr, in.err_thr, in.max_iters, 1, in.fp32_first, sd);
        if (rc) return rc;
        if (!in.stop_i) {
            rc = run_cleanup(cx, g, bd);
            if (rc) return rc;
        }
    }
    CHECK_HIP(hipEventRecord(cx->ev[17], s));
    CHECK_HIP(hipEventSynchronize(cx->ev[17]));
    float ms = 0.f;
    CHECK_HIP(hipEventElapsedTime(&ms, cx->ev[16], cx->ev[17]));
    st.ms_dsc = ms;
    return MSCEER_OK;
}

int ondemand_impl(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz, float err_thr,
                  float grad_thr, int32_t max_iters, float presimp, uint32_t need, const uint8_t* integrate_mask,
                  uint8_t* grad_out, int32_t* labels_out, int32_t* labels_dsc_out, int8_t* restriction_out,
                  msceer_label_stats* stats, uint32_t flags) {
    if (!ctx || !values) return MSCEER_EINVAL;
    if (flags & ~(MSCEER_DEVICE_PTRS | MSCEER_TIMING | MSCEER_STOP_REGIONS | MSCEER_STOP_INTEGRATED |
                  MSCEER_FP32_FIRST))
        return MSCEER_EINVAL;
    if (!need || (need & ~(MSCEER_NEED_ASC1 | MSCEER_NEED_DSC1 | MSCEER_NEED_ASC2 | MSCEER_NEED_DSC2)))
        return MSCEER_EINVAL;
    const bool asc2 = need & MSCEER_NEED_ASC2, dsc2 = need & MSCEER_NEED_DSC2;
    if (labels_dsc_out && !dsc2) return MSCEER_EINVAL;
    const bool lines = need & (MSCEER_NEED_ASC1 | MSCEER_NEED_DSC1);
    const bool stop_r = flags & MSCEER_STOP_REGIONS, stop_i = flags & MSCEER_STOP_INTEGRATED;
    if (!asc2 && (integrate_mask || labels_out)) return MSCEER_EINVAL;
    if (!asc2 && !dsc2 && (stop_r || stop_i)) return MSCEER_EINVAL;
    if (!(err_thr > 0.f) || !(grad_thr >= 0.f) || max_iters < 1 || !(presimp >= 0.f)) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    if (nx * ny * nz >= (int64_t)1 << 31) return MSCEER_ETOOBIG;
    CHECK_HIP(hipSetDevice(ctx->device));
    const Dims d = make_dims(nx, ny, nz);
    const VGrid g = make_vgrid(nx, ny, nz);
    const int64_t V = (int64_t)d.verts, C = (int64_t)d.cells;
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    hipStream_t s = ctx->stream;
    hipEvent_t* ev = ctx->ev;
    msceer_label_stats st{};

    // ---- buffers ----
    const float* d_values = values;
    if (!dev) {
        CHECK_HIP(ctx->values.ensure(sizeof(float) * V));
        CHECK_HIP(hipMemcpyAsync(ctx->values.ptr, values, sizeof(float) * V, hipMemcpyHostToDevice, s));
        d_values = ctx->values.as<float>();
    }
    uint8_t* d_grad = (dev && grad_out) ? grad_out : nullptr;
    if (!d_grad) {
        CHECK_HIP(ctx->grad.ensure(C));
        d_grad = ctx->grad.as<uint8_t>();
    }
    CHECK_HIP(ctx->gradf.ensure(sizeof(float4) * V));
    CHECK_HIP(ctx->gradd.ensure(sizeof(double4) * V));
    CHECK_HIP(ctx->label.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->label2.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->rep.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->rep2.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->stamp.ensure(sizeof(uint32_t) * V));
    CHECK_HIP(ctx->dest.ensure(V));
    CHECK_HIP(ctx->seedmask.ensure(V));
    CHECK_HIP(ctx->redo.ensure(V));
    CHECK_HIP(ctx->restriction.ensure(C));
    ListCaps caps{};
    rc = ensure_lists(ctx, d, &caps);
    if (rc) return rc;
    uint32_t* ctl = ctl_dev(ctx);
    const MList list1 = crit_list(ctx, caps, 0), list2 = crit_list(ctx, caps, 1);
    LabelBuffers b{};
    b.values = d_values;
    b.gradf = ctx->gradf.as<float4>();
    b.gradd = ctx->gradd.as<double4>();
    b.label = ctx->label.as<int32_t>();
    b.label2 = ctx->label2.as<int32_t>();
    b.dest = ctx->dest.as<uint8_t>();
    b.seedmask = ctx->seedmask.as<uint8_t>();
    b.redo = ctx->redo.as<uint8_t>();
    b.stamp = ctx->stamp.as<uint32_t>();
    b.rep = ctx->rep.as<int32_t>();
    b.rep2 = ctx->rep2.as<int32_t>();
    b.snap = b.rep;  // free until the cleanup
    b.stats = ctl + kCtlIntStats;
    const uint8_t* d_mask = integrate_mask;
    if (integrate_mask && !dev) {
        CHECK_HIP(ctx->mask.ensure(V));
        CHECK_HIP(hipMemcpyAsync(ctx->mask.ptr, integrate_mask, V, hipMemcpyHostToDevice, s));
        d_mask = ctx->mask.as<uint8_t>();
    }

    // ---- base gradient ----
    LstarBuffers lb{d_values, d_grad, nullptr, &list1, &list2, ctl + kCtlStats, ctl + kCtlOverflow};
    int shape = 0;
    rc = lstar_shape(ctx, d, lb, &shape);
    if (rc) return rc;
    CHECK_HIP(hipEventRecord(ev[0], s));
    CHECK_HIP(launch_zero(ctl, kCtlWords, s));
    CHECK_HIP(launch_lstar(d, lb, s, shape));
    CHECK_HIP(hipEventRecord(ev[1], s));

    // The descending branch runs on an auxiliary context (own stream and
    // buffers) on a host thread of its own, beside the ascending branch once
    // that has seeded its regions (its gradient field, zeroed at its seeds,
    // is the one the descending integration starts from); joined before the
    // restriction labels.  Without needASC2 it runs in this thread, where the
    // ascending branch would.
    LabelBuffers bd = b;
    DscIn din{};
    msceer_label_stats st_dsc{};
    int rc_dsc = MSCEER_OK;
    std::thread dsc_thread;
    struct Joiner {
        std::thread& t;
        ~Joiner() {
            if (t.joinable()) t.join();
        }
    } joiner{dsc_thread};
    bool dsc_started = false;
    if (dsc2) {
        if (!ctx->aux) {
            rc = msceer_create(&ctx->aux, ctx->device);
            if (rc) return rc;
            ctx->aux->tuning = ctx->tuning;
        }
        rc = read_ctl(ctx, 0, kCtlWords);  // the critical quads of the base pass
        if (rc) return rc;
        if (ctx->h_ctl[kCtlOverflow]) return MSCEER_ETOOBIG;
        din.g = g;
        din.d = d;
        din.values = d_values;
        din.grad = d_grad;
        din.list2 = list2;
        din.n_sad = list_total(ctx->h_ctl + kCtlList2, list2);
        din.ready = ev[18];
        din.presimp = presimp;
        din.grad_thr = grad_thr;
        din.err_thr = err_thr;
        din.max_iters = max_iters;
        din.fp32_first = (flags & MSCEER_FP32_FIRST) != 0;
        din.stop_r = stop_r;
        din.stop_i = stop_i;
    }
    auto start_dsc = [&](bool threaded) -> int {
        CHECK_HIP(hipEventRecord(ev[18], s));
        dsc_started = true;
        if (!threaded) {
            rc_dsc = run_dsc(ctx->aux, din, st_dsc);
            return MSCEER_OK;
        }
        try {
            dsc_thread = std::thread([&]() { rc_dsc = run_dsc(ctx->aux, din, st_dsc); });
        } catch (...) {
            dsc_started = false;
            return MSCEER_ENOMEM;
        }
        return MSCEER_OK;
    };
    // waits for the descending branch; its labels then read on this stream
    auto join_dsc = [&]() -> int {
        if (!dsc_started) return MSCEER_OK;
        if (dsc_thread.joinable()) dsc_thread.join();
        dsc_started = false;
        if (rc_dsc) return rc_dsc;
        CHECK_HIP(hipStreamWaitEvent(s, ctx->aux->ev[17], 0));
        CHECK_HIP(hipEventRecord(ev[11], s));
        bd.label = ctx->aux->label.as<int32_t>();
        st.maxima = st_dsc.maxima;
        st.saddles2 = st_dsc.saddles2;
        st.regions_dsc = st_dsc.regions_dsc;
        st.certain_sweeps_dsc = st_dsc.certain_sweeps_dsc;
        st.ms_dsc = st_dsc.ms_dsc;
        return MSCEER_OK;
    };

    // The STOP flags end every labelling asked for at that stage and return
    // its labels alone.
    auto finish_early = [&]() -> int {
        if (labels_out)
            CHECK_HIP(hipMemcpyAsync(labels_out, b.label, sizeof(int32_t) * V,
                                     dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
        if (labels_dsc_out)
            CHECK_HIP(hipMemcpyAsync(labels_dsc_out, bd.label, sizeof(int32_t) * V,
                                     dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
        CHECK_HIP(hipStreamSynchronize(s));
        ctx->last_labels = st;
        if (stats) *stats = st;
        return MSCEER_OK;
    };
    if (asc2) {
        // ---- extremum graph (critical vertices and 1-saddles) ----
        // The minima stay on the device; only the arcs that the simplification can
        // cancel (persistence at most presimp) come to the host.
        const MList minl = front_list(ctx, caps, 0);
        CHECK_HIP(launch_zero(minl.counts, kCountWords, s));
        CHECK_HIP(launch_collect_minima(g, d_grad, minl, ctl + kCtlOverflow, s));
        rc = read_ctl(ctx, 0, kCtlWords);
        if (rc) return rc;
        if (ctx->h_ctl[kCtlOverflow]) return MSCEER_ETOOBIG;  // more critical cells than list room
        const uint64_t n_min = list_total(ctx->h_ctl + kCtlRound, minl);
        const uint64_t n_sad = list_total(ctx->h_ctl + kCtlList1, list1);
        std::vector<ArcRecord> rec;
        if (n_sad) {
            CHECK_HIP(ctx->host_io.ensure(sizeof(ArcRecord) * n_sad + 64));
            ArcRecord* d_arcs = ctx->host_io.as<ArcRecord>();
            CHECK_HIP(launch_zero(ctl + kCtlCount, 1, s));
            CHECK_HIP(launch_saddle_arcs(g, d_values, d_grad, list1, presimp, d_arcs, (uint32_t)n_sad, ctl + kCtlCount, s));
            rc = read_ctl(ctx, kCtlCount, 1);
            if (rc) return rc;
            rec.resize(std::min<uint64_t>(ctx->h_ctl[kCtlCount], n_sad));
            if (!rec.empty()) {
                CHECK_HIP(hipMemcpyAsync(rec.data(), d_arcs, sizeof(ArcRecord) * rec.size(), hipMemcpyDeviceToHost, s));
                CHECK_HIP(hipStreamSynchronize(s));
            }
        }
        // nodes: the minima the cancellable arcs touch, in ascending cell order
        std::vector<MinNode> nodes;
        nodes.reserve(2 * rec.size());
        std::vector<SaddleArc> arcs(rec.size());
        const int64_t Mx0 = 2 * nx - 1, My0 = 2 * ny - 1;
        auto vnum0 = [&](int64_t c) {
            return (c % Mx0) / 2 + nx * (((c / Mx0) % My0) / 2 + ny * ((c / (Mx0 * My0)) / 2));
        };
        for (size_t i = 0; i < rec.size(); i++) {
            nodes.push_back(MinNode{rec[i].min1, rec[i].v1, vnum0(rec[i].min1)});
            nodes.push_back(MinNode{rec[i].min2, rec[i].v2, vnum0(rec[i].min2)});
            arcs[i] = SaddleArc{rec[i].saddle, rec[i].min1, rec[i].min2, rec[i].sval};
        }
        std::sort(nodes.begin(), nodes.end(), [](const MinNode& x, const MinNode& y) { return x.cell < y.cell; });
        nodes.erase(std::unique(nodes.begin(), nodes.end(),
                                [](const MinNode& x, const MinNode& y) { return x.cell == y.cell; }),
                    nodes.end());
        const std::vector<int> rep = simplify_minima(nodes, arcs, presimp);
        // seeds: every minimum labelled by itself, then the cancelled ones by
        // their representative's vertex
        const int64_t Mx = 2 * nx - 1, My = 2 * ny - 1;
        auto vnum = [&](int64_t c) { return (c % Mx) / 2 + nx * (((c / Mx) % My) / 2 + ny * ((c / (Mx * My)) / 2)); };
        std::vector<int32_t> seed_v, seed_l;
        for (size_t i = 0; i < nodes.size(); i++) {
            if (rep[i] == (int)i) continue;
            seed_v.push_back((int32_t)vnum(nodes[i].cell));
            seed_l.push_back((int32_t)vnum(nodes[rep[i]].cell));
        }
        st.minima = (int64_t)n_min;
        st.saddles = (int64_t)n_sad;
        st.regions = (int64_t)n_min - (int64_t)seed_v.size();

        // ---- certain regions ----
        CHECK_HIP(launch_fill_i32(b.label, V, -1, s));
        CHECK_HIP(launch_fill_u8(b.seedmask, V, 0, s));
        CHECK_HIP(launch_grad_field(g, d_values, b.gradf, s));
        CHECK_HIP(launch_seed_minima(minl, b, s));
        if (!seed_v.empty()) {
            CHECK_HIP(ctx->host_io.ensure(2 * sizeof(int32_t) * seed_v.size() + 64));
            int32_t* d_sv = ctx->host_io.as<int32_t>();
            int32_t* d_sl = d_sv + seed_v.size();
            CHECK_HIP(hipMemcpyAsync(d_sv, seed_v.data(), sizeof(int32_t) * seed_v.size(), hipMemcpyHostToDevice, s));
            CHECK_HIP(hipMemcpyAsync(d_sl, seed_l.data(), sizeof(int32_t) * seed_l.size(), hipMemcpyHostToDevice, s));
            CHECK_HIP(launch_seed_scatter(d_sv, d_sl, (int)seed_v.size(), b, s));
        }
        if (dsc2) {
            din.asc_gradf = b.gradf;
            din.asc_seedmask = b.seedmask;
            rc = start_dsc(true);
            if (rc) return rc;
        }
        // rounds in dependency order
        int sweeps = 0;
        rc = run_certain_rounds(ctx, g, d_values, b, false, &sweeps);
        if (rc) return rc;
        st.certain_sweeps = sweeps;
        CHECK_HIP(hipEventRecord(ev[2], s));

        // ---- integration ----
        if (!stop_r) {
            rc = run_integration(ctx, g, b, d_mask, d_grad, grad_thr, err_thr, max_iters, -1,
                                 (flags & MSCEER_FP32_FIRST) != 0, st);
            if (rc) return rc;
        }
        CHECK_HIP(hipEventRecord(ev[3], s));

        // ---- cleanup of isolated components ----
        if (!stop_r && !stop_i) {
            rc = run_cleanup(ctx, g, b);
            if (rc) return rc;
        }
        CHECK_HIP(hipEventRecord(ev[4], s));
    } else {
        CHECK_HIP(hipEventRecord(ev[2], s));
        CHECK_HIP(hipEventRecord(ev[3], s));
        CHECK_HIP(hipEventRecord(ev[4], s));
        if (dsc2) {
            rc = start_dsc(false);
            if (rc) return rc;
        }
    }

    // ---- join the descending branch ----
    CHECK_HIP(hipEventRecord(ev[11], s));
    rc = join_dsc();  // records ev[11] again after the wait
    if (rc) return rc;
    if (stop_r || stop_i) return finish_early();

    // ---- restriction labels, 1-manifolds, conforming gradient, ldir ----
    if (grad_out || restriction_out) {
        int8_t* restr = ctx->restriction.as<int8_t>();
        if (asc2) CHECK_HIP(launch_min_boundary(g, b.label, restr, s));
        else CHECK_HIP(launch_fill_u8(reinterpret_cast<uint8_t*>(restr), C, 0, s));
        if (dsc2) CHECK_HIP(launch_max_boundary(g, d_values, bd.label, restr, s));
        CHECK_HIP(hipEventRecord(ev[5], s));
        if (lines) {
            // the streamlines follow the FP64 field, which the labelling has
            // zeroed at its region extrema (as the reference's shared field is)
            if (!asc2 && !dsc2) {
                CHECK_HIP(launch_fill_u8(b.seedmask, V, 0, s));
                CHECK_HIP(launch_grad_field64(g, b, s));
            }
            // after needDSC2 the descending branch's field, zeroed at both families' seeds
            const double4* field = dsc2 ? ctx->aux->gradd.as<double4>() : b.gradd;
            LineStats la, ld;
            if (need & MSCEER_NEED_ASC1) {
                rc = run_lines(ctx, d, d_values, d_grad, field, restr, true, grad_thr, err_thr, max_iters, &la);
                if (rc) return rc;
            }
            if (need & MSCEER_NEED_DSC1) {
                rc = run_lines(ctx, d, d_values, d_grad, field, restr, false, grad_thr, err_thr, max_iters, &ld);
                if (rc) return rc;
            }
            st.lines_asc = (int64_t)la.traced;
            st.lines_dsc = (int64_t)ld.traced;
            st.lines_replayed = (int64_t)(la.replayed + ld.replayed);
            st.line_windows = (int32_t)(la.windows + ld.windows);
            st.line_cuts = (int32_t)(la.cuts + ld.cuts);
            st.line_alone = (int32_t)(la.direct + ld.direct);
            st.line_overflows = (int32_t)(la.overflows + ld.overflows);
        }
        CHECK_HIP(hipEventRecord(ev[10], s));
        if (restriction_out)
            CHECK_HIP(hipMemcpyAsync(restriction_out, restr, C, dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                                     s));
    } else {
        CHECK_HIP(hipEventRecord(ev[5], s));
        CHECK_HIP(hipEventRecord(ev[10], s));
    }
    if (grad_out) {
        int8_t* restr = ctx->restriction.as<int8_t>();
        CHECK_HIP(launch_zero(ctl + kCtlList1, 2 * kCountWords, s));
        CHECK_HIP(launch_zero(ctl + kCtlOverflow, 1, s));
        LstarBuffers cb{d_values, d_grad, restr, &list1, &list2, ctl + kCtlStats, ctl + kCtlOverflow};
        CHECK_HIP(launch_lstar(d, cb, s, shape));
        CHECK_HIP(hipEventRecord(ev[6], s));
        rc = read_ctl(ctx, kCtlOverflow, 1);
        if (rc) return rc;
        int ldir_rounds = 0;
        rc = run_ldir(ctx, d, d_grad, caps, ctx->h_ctl[kCtlOverflow] == 0, &ldir_rounds);
        if (rc) return rc;
        if (!dev) CHECK_HIP(hipMemcpyAsync(grad_out, d_grad, C, hipMemcpyDeviceToHost, s));
    } else {
        CHECK_HIP(hipEventRecord(ev[6], s));
    }
    CHECK_HIP(hipEventRecord(ev[7], s));
    if (labels_out) {
        CHECK_HIP(hipMemcpyAsync(labels_out, b.label, sizeof(int32_t) * V,
                                 dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    }
    if (labels_dsc_out) {
        CHECK_HIP(hipMemcpyAsync(labels_dsc_out, bd.label, sizeof(int32_t) * V,
                                 dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    }
    CHECK_HIP(hipStreamSynchronize(s));
    st.ms_lstar = elapsed(ev[0], ev[1]) + elapsed(ev[10], ev[6]);
    st.ms_regions = elapsed(ev[1], ev[2]);
    st.ms_integrate = elapsed(ev[2], ev[3]);
    st.ms_cleanup = elapsed(ev[3], ev[4]) + elapsed(ev[11], ev[5]);  // and the restriction labels
    st.ms_dsc_wait = elapsed(ev[4], ev[11]);
    st.ms_lines = elapsed(ev[5], ev[10]);
    st.ms_grad = elapsed(ev[10], ev[7]);
    st.ms_total = elapsed(ev[0], ev[7]);
    ctx->last_labels = st;
    if (stats) *stats = st;
    return MSCEER_OK;
}

}  // namespace

extern "C" int msceer_ondemand_asc(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                                   float err_thr, float grad_thr, int32_t max_iters, float presimp,
                                   const uint8_t* integrate_mask, uint8_t* grad_out, int32_t* labels_out,
                                   msceer_label_stats* stats, uint32_t flags) {
    return ondemand_impl(ctx, values, nx, ny, nz, err_thr, grad_thr, max_iters, presimp, MSCEER_NEED_ASC2,
                         integrate_mask, grad_out, labels_out, nullptr, nullptr, stats, flags);
}

extern "C" int msceer_ondemand(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                               float err_thr, float grad_thr, int32_t max_iters, float presimp, uint32_t need,
                               const uint8_t* integrate_mask, uint8_t* grad_out, int32_t* labels_out,
                               int8_t* restriction_out, msceer_label_stats* stats, uint32_t flags) {
    return ondemand_impl(ctx, values, nx, ny, nz, err_thr, grad_thr, max_iters, presimp, need, integrate_mask,
                         grad_out, labels_out, nullptr, restriction_out, stats, flags);
}

extern "C" int msceer_ondemand2(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                                float err_thr, float grad_thr, int32_t max_iters, float presimp, uint32_t need,
                                const uint8_t* integrate_mask, uint8_t* grad_out, int32_t* labels_out,
                                int32_t* labels_dsc_out, int8_t* restriction_out, msceer_label_stats* stats,
                                uint32_t flags) {
    return ondemand_impl(ctx, values, nx, ny, nz, err_thr, grad_thr, max_iters, presimp, need, integrate_mask,
                         grad_out, labels_out, labels_dsc_out, restriction_out, stats, flags);
}

namespace {

// Device copy of a host or device input (returns the device pointer).
template <typename T>
int stage_in(msceer_ctx* ctx, msceer::DevBuf& buf, const T* src, size_t n, bool dev, const T** out) {
    if (dev) {
        *out = src;
        return MSCEER_OK;
    }
    CHECK_HIP(buf.ensure(sizeof(T) * n));
    CHECK_HIP(hipMemcpyAsync(buf.ptr, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream));
    *out = buf.as<T>();
    return MSCEER_OK;
}

}  // namespace

// IsolatedCCRegionRemoverNEW(func, labels).ComputeOutput()
// (gi_isolated_region_remover.h:155-192) on a label array, in place.
extern "C" int msceer_clean_labels(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                                   int32_t* labels_inout, uint32_t flags) {
    if (!ctx || !values || !labels_inout) return MSCEER_EINVAL;
    if (flags & ~MSCEER_DEVICE_PTRS) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    if (nx * ny * nz >= (int64_t)1 << 31) return MSCEER_ETOOBIG;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const VGrid g = make_vgrid(nx, ny, nz);
    const int64_t V = g.verts();
    hipStream_t s = ctx->stream;
    const float* d_values = nullptr;
    rc = stage_in(ctx, ctx->values, values, (size_t)V, dev, &d_values);
    if (rc) return rc;
    CHECK_HIP(ctx->label.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->label2.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->rep.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->rep2.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->stamp.ensure(sizeof(uint32_t) * V));
    CHECK_HIP(hipMemcpyAsync(ctx->label.ptr, labels_inout, sizeof(int32_t) * V,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    LabelBuffers b{};
    b.values = d_values;
    b.label = ctx->label.as<int32_t>();
    b.label2 = ctx->label2.as<int32_t>();
    b.rep = ctx->rep.as<int32_t>();
    b.rep2 = ctx->rep2.as<int32_t>();
    b.stamp = ctx->stamp.as<uint32_t>();
    rc = run_cleanup(ctx, g, b);
    if (rc) return rc;
    CHECK_HIP(hipMemcpyAsync(labels_inout, b.label, sizeof(int32_t) * V,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    return MSCEER_OK;
}

// VertexLabelingToBoundaryLabeling::ComputeMINBoundary (gi_labeling_to_bounary_labeling.h:48-116)
// starting from an all-zero restriction (InitializeFirst).
extern "C" int msceer_restriction_labels(msceer_ctx* ctx, const int32_t* labels, int64_t nx, int64_t ny, int64_t nz,
                                         int8_t* restriction_out, uint32_t flags) {
    if (!ctx || !labels || !restriction_out) return MSCEER_EINVAL;
    if (flags & ~MSCEER_DEVICE_PTRS) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const Dims d = make_dims(nx, ny, nz);
    const VGrid g = make_vgrid(nx, ny, nz);
    hipStream_t s = ctx->stream;
    const int32_t* d_labels = nullptr;
    rc = stage_in(ctx, ctx->label, labels, (size_t)d.verts, dev, &d_labels);
    if (rc) return rc;
    int8_t* d_r = restriction_out;
    if (!dev) {
        CHECK_HIP(ctx->restriction.ensure(d.cells));
        d_r = ctx->restriction.as<int8_t>();
    }
    CHECK_HIP(launch_min_boundary(g, d_labels, d_r, s));
    if (!dev) CHECK_HIP(hipMemcpyAsync(restriction_out, d_r, d.cells, hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    return MSCEER_OK;
}

// VertexLabelingToBoundaryLabeling::ComputeRegionBoundaryKind
// (gi_labeling_to_bounary_labeling.h:262-393), as Accurate3DMorseLabeling
// writes it (accurate3dmorselabeling.cxx:337-343).
extern "C" int msceer_region_boundary_kind(msceer_ctx* ctx, const int32_t* labels, int64_t nx, int64_t ny, int64_t nz,
                                           int8_t* kind_out, uint32_t flags) {
    if (!ctx || !labels || !kind_out) return MSCEER_EINVAL;
    if (flags & ~MSCEER_DEVICE_PTRS) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const Dims d = make_dims(nx, ny, nz);
    const VGrid g = make_vgrid(nx, ny, nz);
    hipStream_t s = ctx->stream;
    const int32_t* d_labels = nullptr;
    rc = stage_in(ctx, ctx->label, labels, (size_t)d.verts, dev, &d_labels);
    if (rc) return rc;
    int8_t* d_k = kind_out;
    if (!dev) {
        CHECK_HIP(ctx->restriction.ensure(d.cells));
        d_k = ctx->restriction.as<int8_t>();
    }
    CHECK_HIP(launch_region_kind(g, d_labels, d_k, s));
    if (!dev) CHECK_HIP(hipMemcpyAsync(kind_out, d_k, d.cells, hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    return MSCEER_OK;
}

// VertexLabelingToBoundaryLabeling::ComputeMAXBoundary (gi_labeling_to_bounary_labeling.h:118-195)
// added onto an existing restriction labelling (in/out), as ondemandaccurate
// does after the MIN boundary of the ascending labels (ondemandaccurate.cxx:988).
extern "C" int msceer_restriction_add_max(msceer_ctx* ctx, const float* values, const int32_t* labels, int64_t nx,
                                          int64_t ny, int64_t nz, int8_t* restriction_inout, uint32_t flags) {
    if (!ctx || !values || !labels || !restriction_inout) return MSCEER_EINVAL;
    if (flags & ~MSCEER_DEVICE_PTRS) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const Dims d = make_dims(nx, ny, nz);
    const VGrid g = make_vgrid(nx, ny, nz);
    hipStream_t s = ctx->stream;
    const float* d_values = nullptr;
    rc = stage_in(ctx, ctx->values, values, (size_t)d.verts, dev, &d_values);
    if (rc) return rc;
    const int32_t* d_labels = nullptr;
    rc = stage_in(ctx, ctx->label, labels, (size_t)d.verts, dev, &d_labels);
    if (rc) return rc;
    int8_t* d_r = restriction_inout;
    if (!dev) {
        CHECK_HIP(ctx->restriction.ensure(d.cells));
        d_r = ctx->restriction.as<int8_t>();
        CHECK_HIP(hipMemcpyAsync(d_r, restriction_inout, d.cells, hipMemcpyHostToDevice, s));
    }
    CHECK_HIP(launch_max_boundary(g, d_values, d_labels, d_r, s));
    if (!dev) CHECK_HIP(hipMemcpyAsync(restriction_inout, d_r, d.cells, hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    return MSCEER_OK;
}

// Lower stars recomputed under a restriction labelling (MyRobinsNoalloc<...,4,6>,
// gi_modified_robins.h:1281-1343; ondemandaccurate.cxx:1014-1052) followed by
// RecordGrad (ldir + flags, ondemandaccurate.cxx:693-699) unless MSCEER_SKIP_LDIR.
// Every lower star is recomputed: a star without restricted cells splits by
// boundary value alone, exactly as in the steepest pass.
extern "C" int msceer_constrained_lstar(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                                        const int8_t* restriction, uint8_t* grad_inout, uint32_t flags) {
    if (!ctx || !values || !restriction || !grad_inout) return MSCEER_EINVAL;
    if (flags & ~(MSCEER_DEVICE_PTRS | MSCEER_SKIP_LDIR)) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const Dims d = make_dims(nx, ny, nz);
    hipStream_t s = ctx->stream;
    const float* d_values = nullptr;
    rc = stage_in(ctx, ctx->values, values, (size_t)d.verts, dev, &d_values);
    if (rc) return rc;
    const int8_t* d_r = nullptr;
    rc = stage_in(ctx, ctx->restriction, restriction, (size_t)d.cells, dev, &d_r);
    if (rc) return rc;
    uint8_t* d_grad = grad_inout;
    if (!dev) {
        CHECK_HIP(ctx->grad.ensure(d.cells));
        d_grad = ctx->grad.as<uint8_t>();
    }
    ListCaps caps{};
    rc = ensure_lists(ctx, d, &caps);
    if (rc) return rc;
    uint32_t* ctl = ctl_dev(ctx);
    const MList list1 = crit_list(ctx, caps, 0), list2 = crit_list(ctx, caps, 1);
    LstarBuffers cb{d_values, d_grad, d_r, &list1, &list2, ctl + kCtlStats, ctl + kCtlOverflow};
    int shape = 0;
    {
        std::lock_guard<std::mutex> lk(ctx->tuning->mu);
        auto it = ctx->tuning->brick_by_x.find(d.X);
        if (ctx->tuning->mode == 1 && it != ctx->tuning->brick_by_x.end()) shape = it->second;
        else if (ctx->tuning->mode > 1) shape = ctx->tuning->mode - 2;
    }
    CHECK_HIP(launch_zero(ctl, kCtlWords, s));
    CHECK_HIP(launch_lstar(d, cb, s, shape));
    if (!(flags & MSCEER_SKIP_LDIR)) {
        rc = read_ctl(ctx, kCtlOverflow, 1);
        if (rc) return rc;
        int rounds = 0;
        rc = run_ldir(ctx, d, d_grad, caps, ctx->h_ctl[kCtlOverflow] == 0, &rounds);
        if (rc) return rc;
    }
    if (!dev) CHECK_HIP(hipMemcpyAsync(grad_inout, d_grad, d.cells, hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    return MSCEER_OK;
}

// NumericIntegratorPathCompressingToTerminal::BeginIntegration on given
// terminals (gi_numeric_integrator_path_compressing.h:75-433): labels >= 0 are
// certain terminals, -1 vertices are integrated; the gradient is zero where
// zero_grad is set (the seeds of BeginIntegration's remap).  sign -1 flows
// downhill (ascending manifolds, AdaptiveEulerAdvector3D<F,-1>), +1 uphill.
// Passes 1-3 as in msceer_ondemand_asc; no cleanup.
extern "C" int msceer_integrate_labels(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                                       const int32_t* terminals, const uint8_t* zero_grad, float err_thr,
                                       float grad_thr, int32_t max_iters, int32_t sign, int32_t* labels_out,
                                       msceer_label_stats* stats, uint32_t flags) {
    if (!ctx || !values || !terminals || !labels_out) return MSCEER_EINVAL;
    if (flags & ~(MSCEER_DEVICE_PTRS | MSCEER_TIMING | MSCEER_FP32_FIRST)) return MSCEER_EINVAL;
    if (sign != -1 && sign != 1) return MSCEER_EINVAL;
    if (!(err_thr > 0.f) || !(grad_thr >= 0.f) || max_iters < 1) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    if (nx * ny * nz >= (int64_t)1 << 31) return MSCEER_ETOOBIG;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const Dims d = make_dims(nx, ny, nz);
    const VGrid g = make_vgrid(nx, ny, nz);
    const int64_t V = (int64_t)d.verts;
    hipStream_t s = ctx->stream;
    hipEvent_t* ev = ctx->ev;
    msceer_label_stats st{};
    const float* d_values = nullptr;
    rc = stage_in(ctx, ctx->values, values, (size_t)V, dev, &d_values);
    if (rc) return rc;
    CHECK_HIP(ctx->gradf.ensure(sizeof(float4) * V));
    CHECK_HIP(ctx->gradd.ensure(sizeof(double4) * V));
    CHECK_HIP(ctx->label.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->label2.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->stamp.ensure(sizeof(uint32_t) * V));
    CHECK_HIP(ctx->rep.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->dest.ensure(V));
    CHECK_HIP(ctx->seedmask.ensure(V));
    CHECK_HIP(ctx->redo.ensure(V));
    ListCaps caps{};
    rc = ensure_lists(ctx, d, &caps);  // the re-integration queue lives in the frontier buffer
    if (rc) return rc;
    uint32_t* ctl = ctl_dev(ctx);
    LabelBuffers b{};
    b.values = d_values;
    b.gradf = ctx->gradf.as<float4>();
    b.gradd = ctx->gradd.as<double4>();
    b.label = ctx->label.as<int32_t>();
    b.label2 = ctx->label2.as<int32_t>();
    b.dest = ctx->dest.as<uint8_t>();
    b.seedmask = ctx->seedmask.as<uint8_t>();
    b.redo = ctx->redo.as<uint8_t>();
    b.stamp = ctx->stamp.as<uint32_t>();
    b.snap = ctx->rep.as<int32_t>();
    b.stats = ctl + kCtlIntStats;
    CHECK_HIP(hipEventRecord(ev[0], s));
    CHECK_HIP(hipMemcpyAsync(b.label, terminals, sizeof(int32_t) * V,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    if (zero_grad)
        CHECK_HIP(hipMemcpyAsync(b.seedmask, zero_grad, V, dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
    else
        CHECK_HIP(launch_fill_u8(b.seedmask, V, 0, s));
    CHECK_HIP(launch_grad_field(g, d_values, b.gradf, s));
    CHECK_HIP(launch_zero_seed_gradient(g, b, s));
    CHECK_HIP(launch_zero(ctl, kCtlWords, s));
    rc = run_integration(ctx, g, b, nullptr, nullptr, grad_thr, err_thr, max_iters, sign,
                         (flags & MSCEER_FP32_FIRST) != 0, st);
    if (rc) return rc;
    CHECK_HIP(hipMemcpyAsync(labels_out, b.label, sizeof(int32_t) * V,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipEventRecord(ev[1], s));
    CHECK_HIP(hipStreamSynchronize(s));
    st.ms_integrate = elapsed(ev[0], ev[1]);
    st.ms_total = st.ms_integrate;
    if (stats) *stats = st;
    return MSCEER_OK;
}

// NumericIntegratorExpandingRegionStop<AdaptiveEulerAdvector3D<F, -1>,
// IndexCompareLessThan<F>>::BeginIntegration
// (gi_numeric_integrator_expanding_region_stop.h:204-440): extrema by the
// SameBdry6 rule, certain regions grown from each (the priority flood of
// Expand_Lower_Neighborhood, :153-180, has one solution, computed here in
// dependency order), pass 1 to the certain regions, unambiguous vertices made
// terminals, and a re-integration of the rest with 500 iterations.  The
// streamlines are traced in FP64, each vertex alone in both passes (the
// reference's pass 1 compresses paths in its threads' visiting order, and
// its re-integration reads labels that earlier re-integrations wrote; here
// the re-integration reads pass 1's labels).
extern "C" int msceer_expanding_region_stop(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                                            float err_thr, float grad_thr, int32_t max_iters, int32_t* labels_out,
                                            int64_t* stats_out, uint32_t flags) {
    if (!ctx || !values || !labels_out || max_iters < 0) return MSCEER_EINVAL;
    if (flags & ~MSCEER_DEVICE_PTRS) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    if (nx * ny * nz >= (int64_t)1 << 31) return MSCEER_ETOOBIG;
    CHECK_HIP(hipSetDevice(ctx->device));
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    const VGrid g = make_vgrid(nx, ny, nz);
    const Dims d = make_dims(nx, ny, nz);
    const int64_t V = g.verts();
    hipStream_t s = ctx->stream;
    const float* d_values = nullptr;
    rc = stage_in(ctx, ctx->values, values, (size_t)V, dev, &d_values);
    if (rc) return rc;
    CHECK_HIP(ctx->gradd.ensure(sizeof(double4) * V));
    CHECK_HIP(ctx->label.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->label2.ensure(sizeof(int32_t) * V));
    CHECK_HIP(ctx->stamp.ensure(sizeof(uint32_t) * V));
    CHECK_HIP(ctx->dest.ensure(V));
    CHECK_HIP(ctx->seedmask.ensure(V));
    ListCaps caps{};
    rc = ensure_lists(ctx, d, &caps);
    if (rc) return rc;
    uint32_t* ctl = ctl_dev(ctx);
    uint32_t* counts = ctl + kCtlIntStats;  // extrema, certain, pass-1 streamlines, unambiguous, re-integrated
    CHECK_HIP(launch_zero(counts, 5, s));
    LabelBuffers b{};
    b.values = d_values;
    b.gradd = ctx->gradd.as<double4>();
    b.label = ctx->label.as<int32_t>();
    b.label2 = ctx->label2.as<int32_t>();
    b.dest = ctx->dest.as<uint8_t>();
    b.seedmask = ctx->seedmask.as<uint8_t>();
    b.stamp = ctx->stamp.as<uint32_t>();
    // the gradient of the image, zeroed nowhere
    LabelBuffers nb = b;
    nb.seedmask = nullptr;
    CHECK_HIP(launch_grad_field64(g, nb, s));
    CHECK_HIP(launch_ers_seeds(g, d_values, b.seedmask, b.label, counts, s));
    int sweeps = 0;
    rc = run_certain_rounds(ctx, g, d_values, b, true, &sweeps);
    if (rc) return rc;
    CHECK_HIP(launch_ers_state(g, b.label, b.dest, counts, s));
    // pass 1: label2 = the certain labels, and the label each streamline reaches
    CHECK_HIP(hipMemcpyAsync(b.label2, b.label, sizeof(int32_t) * V, hipMemcpyDeviceToDevice, s));
    CHECK_HIP(launch_ers_trace(g, b.gradd, b.dest, b.label, b.label2, grad_thr, err_thr, max_iters, counts, 2, s));
    CHECK_HIP(launch_ers_mark(g, b.label2, b.dest, counts, s));
    // re-integration of the rest: label = pass 1's labels, then the label each streamline reaches
    CHECK_HIP(hipMemcpyAsync(b.label, b.label2, sizeof(int32_t) * V, hipMemcpyDeviceToDevice, s));
    CHECK_HIP(launch_ers_trace(g, b.gradd, b.dest, b.label2, b.label, grad_thr, err_thr, 500, counts, 4, s));
    CHECK_HIP(hipMemcpyAsync(labels_out, b.label, sizeof(int32_t) * V,
                             dev ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s));
    rc = read_ctl(ctx, 0, kCtlWords);
    if (rc) return rc;
    if (stats_out)
        for (int k = 0; k < 5; k++) stats_out[k] = (int64_t)ctx->h_ctl[kCtlIntStats + k];
    return MSCEER_OK;
}

// msceer_batch_grad: the lower-star .grad of many volumes, one launch per
// stage for a whole group of volumes, each group's launches captured once in a
// hipGraph and replayed.
//
// A group is a run of consecutive volumes with fewer than 2^32 - 1 cells in
// all (frontier items are 32-bit), optionally capped in vertices.  Its
// pipeline: clear the control block; lower-star bricks of every volume (one
// launch: a work-group per brick, or persistent work-groups over the claim
// queues, whichever the tuner measured faster); the ldir records of every
// volume (one launch); pass 1 and pass 2 of the ldir walks as two chains of R
// frontier rounds on two streams (the passes are independent until the
// finish); the finish of every volume (one launch).  A round whose frontier
// is empty returns at once, so R rounds cover every walk that ends within R.
// After a replay one read-back of the control block shows whether a list
// overflowed or a frontier is left; then the pass goes on eagerly (rounds,
// sweeps after an overflow) and the finish runs again (it writes every ldir
// bit from the visited bitmaps alone), and the next capture takes more rounds.
//
// Per group and call: the captured graph, keyed by the group's volume shapes,
// buffer addresses and launch parameters; the tuning (brick shape, persistent
// or not, work-groups, queue chunk), keyed by the shapes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <list>
#include <map>
#include <string>
#include <vector>

#include "ctx.hpp"
#include "gbatch.hpp"

using namespace msceer;

namespace msceer {

namespace {

// Control block of a group (uint32 words).
enum GCtl : int {
    kGList1 = 0,                           // count arrays of the critical lists
    kGList2 = kCountWords,
    kGRound = 2 * kCountWords,             // 3 count arrays per pass (pass p at kGRound + 3 (p - 1) kCountWords)
    kGStats = 8 * kCountWords,             // kSubs * 4 critical counts
    kGOverflow = 8 * kCountWords + 4 * kSubs,  // lists, pass 1, pass 2
    kGQueue = kGOverflow + 4,              // kQueues claim counters
    kGWords = kGQueue + kQueues
};

constexpr uint32_t kMaxItems = 0xFFFFFFFEu;  // cells of a group (kNoItem is 0xFFFFFFFF)
constexpr int kDefaultRounds = 3;
constexpr int kMaxPlans = 4;

__global__ void times_init_kernel(unsigned long long* t, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        t[2 * i] = ~0ull;
        t[2 * i + 1] = 0ull;
    }
}

// Launch parameters of the lower-star stage.
struct LaunchTune {
    int shape = 0;
    uint32_t blocks = 0;  // 0: a work-group per brick; else persistent work-groups
    uint32_t chunk = 1;
    float ms = 0.f;       // measured time of the chosen launch
};

struct GroupLayout {
    int first, count;  // volumes [first, first + count)
    uint64_t verts, cells;
    // host copies of the unit tables
    std::vector<uint32_t> brick_first[kLstarBricks], tile_first, item_first;
    size_t ldir_bytes = 0, bits_offset = 0;  // ldir scratch: records, then every volume's bitmaps
    uint64_t bit_words = 0;
    std::vector<LdirScratch> ls;          // per volume, offsets from the group's ldir buffer
    std::vector<size_t> vals_off, grad_off;  // packed host-staging offsets (bytes)
    size_t vals_bytes = 0, grad_bytes = 0;
};

struct GroupPlan {
    std::string key;  // shapes, buffers and launch parameters of the capture
    DevBuf tables;    // GVol table and unit tables
    GBatch batch{};   // device pointers into tables (bricks: the tuned shape)
    GUnits bricks[kLstarBricks];
    hipGraphExec_t exec = nullptr;
    int rounds = kDefaultRounds;
    LaunchTune tune;
    bool tuned = false;
};

}  // namespace

struct BatchState {
    hipStream_t side = nullptr;  // second stream (pass 2)
    hipEvent_t fork = nullptr, join = nullptr, t0 = nullptr, t1 = nullptr;
    DevBuf vals, grad, ldir, lists, fronts, ctl, times;
    uint64_t generation = 0;  // bumped whenever a shared buffer moves (captured graphs then stale)
    std::list<GroupPlan> plans;  // most recent first
    std::map<std::string, LaunchTune> tunes;
    msceer_batch_stats last{};
    uint32_t* h_ctl = nullptr;
    int cus = 0;
    int clock_khz = 0;
};

namespace {

// Grow a shared buffer; captured graphs hold its address, so a move
// invalidates them.
hipError_t grow(BatchState& s, DevBuf& b, size_t bytes) {
    if (bytes <= b.bytes) return hipSuccess;
    s.generation++;
    return b.ensure(bytes);
}

// Bucket table of a unit table (n volumes, total units T): bucket b holds the
// volume of unit b << shift; one more entry for the end.
std::vector<uint32_t> buckets(const std::vector<uint32_t>& first, uint32_t* shift) {
    const uint32_t n = (uint32_t)first.size() - 1, T = first[n];
    uint32_t sh = 0;
    while (((uint64_t)(T - 1) >> sh) >= 16384u) sh++;
    const uint32_t nb = ((T - 1) >> sh) + 1;
    std::vector<uint32_t> b(nb + 1);
    uint32_t v = 0;
    for (uint32_t k = 0; k < nb; k++) {
        const uint32_t u = k << sh;
        while (first[v + 1] <= u) v++;
        b[k] = v;
    }
    b[nb] = n - 1;
    *shift = sh;
    return b;
}

ListCaps group_caps(uint64_t verts) {
    Dims d{};
    d.verts = verts;
    return list_caps(d);
}

MList glist(BatchState& s, const ListCaps& c, int which) {
    const size_t per = (size_t)c.list_sub * kSubs + c.list_spill;
    return MList{s.lists.as<uint32_t>() + which * per, s.ctl.as<uint32_t>() + (which ? kGList2 : kGList1), c.list_sub,
                 c.list_spill};
}

// Frontier k of pass p: buffer (p - 1) * 2 + k % 2, count array k % 3.
MList gfront(BatchState& s, const ListCaps& c, int pass, int k) {
    const size_t per = (size_t)c.front_sub * kSubs + c.front_spill;
    uint32_t* buf = s.fronts.as<uint32_t>() + ((size_t)(pass - 1) * 2 + (size_t)(k % 2)) * per;
    return MList{buf, s.ctl.as<uint32_t>() + kGRound + (3 * (pass - 1) + k % 3) * kCountWords, c.front_sub,
                 c.front_spill};
}

std::string shapes_key(const msceer_volume* v, const GroupLayout& g) {
    std::string k;
    char buf[64];
    for (int i = g.first; i < g.first + g.count; i++) {
        snprintf(buf, sizeof buf, "%lldx%lldx%lld;", (long long)v[i].nx, (long long)v[i].ny, (long long)v[i].nz);
        k += buf;
    }
    return k;
}

// Layout of a group: unit tables, ldir scratch and staging offsets.
void layout(const msceer_volume* vols, GroupLayout& g) {
    const int n = g.count;
    for (int sh = 0; sh < kLstarBricks; sh++) g.brick_first[sh].assign(n + 1, 0);
    g.tile_first.assign(n + 1, 0);
    g.item_first.assign(n + 1, 0);
    g.ls.resize(n);
    g.vals_off.resize(n);
    g.grad_off.resize(n);
    size_t rec = 0, bits = 0, vo = 0, go = 0;
    std::vector<size_t> bits_at(n);
    for (int i = 0; i < n; i++) {
        const msceer_volume& v = vols[g.first + i];
        const Dims d = make_dims(v.nx, v.ny, v.nz);
        for (int sh = 0; sh < kLstarBricks; sh++) {
            int bx, by, bz;
            lstar_brick_dims(sh, &bx, &by, &bz);
            const uint64_t nb = (uint64_t)((d.X + bx - 1) / bx) * ((d.Y + by - 1) / by) * ((d.Z + bz - 1) / bz);
            g.brick_first[sh][i + 1] = g.brick_first[sh][i] + (uint32_t)nb;
        }
        uint32_t ntx, nty, nt;
        ldir_tiles(d, &ntx, &nty, &nt);
        g.tile_first[i + 1] = g.tile_first[i] + nt;
        g.item_first[i + 1] = g.item_first[i] + (uint32_t)d.cells;
        LdirScratch one{};
        const size_t total = ldir_scratch_bytes(d, &one);
        const size_t bw = total - one.bits2_off;  // bytes of one bitmap (padded)
        // records of the volume at `rec`, its two bitmaps at `bits` (placed after all records)
        g.ls[i].rec1_off = rec + one.rec1_off;
        g.ls[i].rec2_off = rec + one.rec2_off;
        g.ls[i].bit_words = one.bit_words;
        bits_at[i] = bits;
        rec += one.bits1_off;
        bits += 2 * bw;
        g.vals_off[i] = vo;
        g.grad_off[i] = go;
        vo += (sizeof(float) * d.verts + 255) & ~(size_t)255;
        go += (d.cells + 255) & ~(size_t)255;
    }
    g.bits_offset = rec;
    for (int i = 0; i < n; i++) {
        const msceer_volume& v = vols[g.first + i];
        const Dims d = make_dims(v.nx, v.ny, v.nz);
        LdirScratch one{};
        const size_t total = ldir_scratch_bytes(d, &one);
        const size_t bw = total - one.bits2_off;
        g.ls[i].bits1_off = rec + bits_at[i];
        g.ls[i].bits2_off = rec + bits_at[i] + bw;
    }
    g.bit_words = bits / 4;
    g.ldir_bytes = rec + bits;
    g.vals_bytes = vo;
    g.grad_bytes = go;
}

// Volume and unit tables of a group in device memory (uploaded once per plan).
int upload_tables(msceer_ctx* ctx, BatchState& s, const msceer_volume* vols, const GroupLayout& g, bool dev,
                  GroupPlan& p) {
    const int n = g.count;
    std::vector<GVol> gv(n);
    uint8_t* lbase = s.ldir.as<uint8_t>();
    for (int i = 0; i < n; i++) {
        const msceer_volume& v = vols[g.first + i];
        const Dims d = make_dims(v.nx, v.ny, v.nz);
        GVol& o = gv[i];
        o.values = dev ? v.values : reinterpret_cast<const float*>(s.vals.as<uint8_t>() + g.vals_off[i]);
        o.grad = dev ? v.grad_out : s.grad.as<uint8_t>() + g.grad_off[i];
        o.rec1 = reinterpret_cast<uint16_t*>(lbase + g.ls[i].rec1_off);
        o.rec2 = lbase + g.ls[i].rec2_off;
        o.bits1 = reinterpret_cast<uint32_t*>(lbase + g.ls[i].bits1_off);
        o.bits2 = reinterpret_cast<uint32_t*>(lbase + g.ls[i].bits2_off);
        o.cells = d.cells;
        o.bit_words = g.ls[i].bit_words;
        o.X = d.X;
        o.Y = d.Y;
        o.Z = d.Z;
        o.Mx = d.Mx;
        o.My = d.My;
        o.Mz = d.Mz;
        uint32_t nt;
        ldir_tiles(d, &o.ntx, &o.nty, &nt);
    }
    // pack: GVol[n] | per table: first[n + 1], bucket[]
    std::vector<uint8_t> host;
    auto put = [&](const void* p, size_t bytes) -> size_t {
        const size_t at = (host.size() + 15) & ~(size_t)15;
        host.resize(at + bytes);
        memcpy(host.data() + at, p, bytes);
        return at;
    };
    const size_t at_vols = put(gv.data(), sizeof(GVol) * n);
    struct T { size_t first, bucket; uint32_t shift, total; };
    auto table = [&](const std::vector<uint32_t>& first) {
        T t;
        std::vector<uint32_t> b = buckets(first, &t.shift);
        t.first = put(first.data(), sizeof(uint32_t) * first.size());
        t.bucket = put(b.data(), sizeof(uint32_t) * b.size());
        t.total = first.back();
        return t;
    };
    T tb[kLstarBricks];
    for (int sh = 0; sh < kLstarBricks; sh++) tb[sh] = table(g.brick_first[sh]);
    const T tt = table(g.tile_first), ti = table(g.item_first);
    CHECK_HIP(p.tables.ensure(host.size()));
    CHECK_HIP(hipMemcpyAsync(p.tables.ptr, host.data(), host.size(), hipMemcpyHostToDevice, ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));  // `host` goes out of scope
    const uint8_t* base = p.tables.as<uint8_t>();
    auto units = [&](const T& t) {
        return GUnits{reinterpret_cast<const uint32_t*>(base + t.first), reinterpret_cast<const uint32_t*>(base + t.bucket),
                      t.shift, t.total};
    };
    p.batch = GBatch{};
    p.batch.vols = reinterpret_cast<const GVol*>(base + at_vols);
    p.batch.nvols = (uint32_t)n;
    for (int sh = 0; sh < kLstarBricks; sh++) p.bricks[sh] = units(tb[sh]);
    p.batch.tiles = units(tt);
    p.batch.items = units(ti);
    p.batch.queue = s.ctl.as<uint32_t>() + kGQueue;
    return MSCEER_OK;
}

LstarBuffers lstar_bufs(BatchState& s, const ListCaps& caps, MList* l1, MList* l2) {
    *l1 = glist(s, caps, 0);
    *l2 = glist(s, caps, 1);
    uint32_t* ctl = s.ctl.as<uint32_t>();
    return LstarBuffers{nullptr, nullptr, nullptr, l1, l2, ctl + kGStats, ctl + kGOverflow};
}

// Lower-star stage of a group with the given launch parameters.
hipError_t lstar_stage(BatchState& s, GroupPlan& p, const ListCaps& caps, const LaunchTune& t,
                       unsigned long long* times, hipStream_t st) {
    GBatch b = p.batch;
    b.bricks = p.bricks[t.shape];
    b.chunk = t.chunk;
    b.times = times;
    MList l1, l2;
    const LstarBuffers lb = lstar_bufs(s, caps, &l1, &l2);
    if (t.blocks) {
        hipError_t e = launch_zero(s.ctl.as<uint32_t>() + kGQueue, kQueues, st);
        if (e != hipSuccess) return e;
    }
    return launch_lstar_batch(b, lb, t.blocks, st, t.shape);
}

// Measure the lower-star launches of a group on its own data: every brick
// shape a work-group per brick, then the best shape persistent, with claim
// chunks of 1, 4 and 16 bricks, as many work-groups as the CUs hold and half
// as many.  (A brick is always four wavefronts, 256 vertices: the tile, the
// stage and the dealing of vertices to lanes by shape are laid out for it.)
int tune_group(msceer_ctx* ctx, BatchState& s, GroupPlan& p, const ListCaps& caps, LaunchTune* out) {
    // a fixed launch "shape:blocks_per_cu:chunk" (blocks_per_cu 0: a
    // work-group per brick), for tests of each launch form
    if (const char* e = getenv("MSCEER_BATCH_LAUNCH")) {
        int sh = 0, bpc = 0, ch = 1;
        if (sscanf(e, "%d:%d:%d", &sh, &bpc, &ch) == 3 && sh >= 0 && sh < kLstarBricks && bpc >= 0 && ch >= 1) {
            out->shape = sh;
            out->blocks = (uint32_t)(s.cus * std::min(bpc, lstar_batch_occupancy(sh)));
            out->chunk = (uint32_t)ch;
            return MSCEER_OK;
        }
    }
    int mode;
    {
        std::lock_guard<std::mutex> lk(ctx->tuning->mu);
        mode = ctx->tuning->mode;
    }
    if (mode != 1) {
        out->shape = mode == 0 ? 0 : mode - 2;
        out->blocks = 0;
        out->chunk = 1;
        return MSCEER_OK;
    }
    auto time_one = [&](LaunchTune& t) -> int {
        float best = 1e30f;
        for (int rep = 0; rep < 2; rep++) {
            CHECK_HIP(launch_zero(s.ctl.as<uint32_t>(), kGWords, ctx->stream));
            CHECK_HIP(hipEventRecord(s.t0, ctx->stream));
            CHECK_HIP(lstar_stage(s, p, caps, t, nullptr, ctx->stream));
            CHECK_HIP(hipEventRecord(s.t1, ctx->stream));
            CHECK_HIP(hipEventSynchronize(s.t1));
            float ms = 0.f;
            CHECK_HIP(hipEventElapsedTime(&ms, s.t0, s.t1));
            best = std::min(best, ms);
        }
        t.ms = best;
        return MSCEER_OK;
    };
    LaunchTune best;
    best.ms = 1e30f;
    for (int sh = 0; sh < kLstarBricks; sh++) {
        LaunchTune t;
        t.shape = sh;
        int rc = time_one(t);
        if (rc) return rc;
        if (t.ms < best.ms) best = t;
    }
    // persistent: as many work-groups as the CUs hold, then half as many
    const int occ = lstar_batch_occupancy(best.shape);
    const uint32_t chunks[3] = {1, 4, 16};
    const LaunchTune grid = best;
    LaunchTune pbest;
    pbest.ms = 1e30f;
    for (uint32_t c : chunks) {
        LaunchTune t = grid;
        t.blocks = (uint32_t)(s.cus * occ);
        t.chunk = c;
        int rc = time_one(t);
        if (rc) return rc;
        if (t.ms < pbest.ms) pbest = t;
    }
    if (occ > 1) {
        LaunchTune t = pbest;
        t.blocks = (uint32_t)(s.cus * (occ / 2));
        int rc = time_one(t);
        if (rc) return rc;
        if (t.ms < pbest.ms) pbest = t;
    }
    if (pbest.ms < best.ms) best = pbest;
    *out = best;
    return MSCEER_OK;
}

// Enqueue the group's pipeline (captured or not): clear, lower star,
// records, R rounds of each pass on two streams, finish.
int enqueue_group(msceer_ctx* ctx, BatchState& s, GroupPlan& p, const GroupLayout& g, const ListCaps& caps, int R,
                  bool timing) {
    hipStream_t st = ctx->stream;
    uint32_t* ctl = s.ctl.as<uint32_t>();
    CHECK_HIP(launch_zero(ctl, kGWords, st));
    if (timing) {
        hipLaunchKernelGGL(times_init_kernel, dim3((g.count + 255) / 256), dim3(256), 0, st,
                           s.times.as<unsigned long long>(), (uint32_t)g.count);
        CHECK_HIP(hipGetLastError());
    }
    CHECK_HIP(lstar_stage(s, p, caps, p.tune, timing ? s.times.as<unsigned long long>() : nullptr, st));
    GBatch b = p.batch;
    b.bricks = p.bricks[p.tune.shape];
    b.times = timing ? s.times.as<unsigned long long>() : nullptr;
    CHECK_HIP(launch_ldir_records_group(b, reinterpret_cast<uint32_t*>(s.ldir.as<uint8_t>() + g.bits_offset),
                                        g.bit_words, st));
    CHECK_HIP(hipEventRecord(s.fork, st));
    CHECK_HIP(hipStreamWaitEvent(s.side, s.fork, 0));
    for (int pass = 1; pass <= 2; pass++) {
        hipStream_t ps = pass == 1 ? st : s.side;
        MList in = glist(s, caps, pass - 1);
        for (int k = 0; k < R; k++) {
            const MList out = gfront(s, caps, pass, k);
            CHECK_HIP(launch_ldir_round_group(b, pass, k == 0, in, out, gfront(s, caps, pass, k + 1).counts,
                                              ctl + kGOverflow + pass, ps));
            in = out;
        }
    }
    CHECK_HIP(hipEventRecord(s.join, s.side));
    CHECK_HIP(hipStreamWaitEvent(st, s.join, 0));
    CHECK_HIP(launch_ldir_finish_group(b, st));
    return MSCEER_OK;
}

int read_gctl(msceer_ctx* ctx, BatchState& s) {
    CHECK_HIP(hipMemcpyAsync(s.h_ctl, s.ctl.ptr, sizeof(uint32_t) * kGWords, hipMemcpyDeviceToHost, ctx->stream));
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return MSCEER_OK;
}

// Finish a pass eagerly after the captured rounds: from frontier R - 1 (or a
// sweep of every volume after an overflow) until the frontier is empty.
int continue_pass(msceer_ctx* ctx, BatchState& s, GroupPlan& p, const GroupLayout& g, const msceer_volume* vols,
                  bool dev, const ListCaps& caps, int pass, int R, bool need_sweep, int* extra) {
    uint32_t* ctl = s.ctl.as<uint32_t>();
    uint32_t* overflow = ctl + kGOverflow + pass;
    GBatch b = p.batch;
    b.bricks = p.bricks[p.tune.shape];
    int k = R;
    MList in = R > 0 ? gfront(s, caps, pass, R - 1) : glist(s, caps, pass - 1);
    auto total = [&](int kk) {
        return list_total(s.h_ctl + kGRound + (3 * (pass - 1) + kk % 3) * kCountWords, gfront(s, caps, pass, kk));
    };
    for (int guard = 0; guard < 100000; guard++) {
        if (need_sweep) {
            const MList out = gfront(s, caps, pass, k);
            CHECK_HIP(launch_zero(out.counts, kCountWords, ctx->stream));
            CHECK_HIP(launch_zero(overflow, 1, ctx->stream));
            uint8_t* lbase = s.ldir.as<uint8_t>();
            for (int i = 0; i < g.count; i++) {
                const msceer_volume& v = vols[g.first + i];
                const Dims d = make_dims(v.nx, v.ny, v.nz);
                const uint8_t* grad = dev ? v.grad_out : s.grad.as<uint8_t>() + g.grad_off[i];
                CHECK_HIP(launch_ldir_sweep(d, grad, lbase, g.ls[i], pass, out, overflow, ctx->stream,
                                            g.item_first[i]));
            }
            int rc = read_gctl(ctx, s);
            if (rc) return rc;
            need_sweep = s.h_ctl[kGOverflow + pass] != 0;
            (*extra)++;
            if (total(k) == 0 && !need_sweep) return MSCEER_OK;
            CHECK_HIP(launch_zero(overflow, 1, ctx->stream));
            in = out;
            k++;
            CHECK_HIP(launch_zero(gfront(s, caps, pass, k).counts, kCountWords, ctx->stream));
        } else if (total(k - 1) == 0) {
            return MSCEER_OK;
        }
        for (int j = 0; j < 4; j++) {
            const MList out = gfront(s, caps, pass, k);
            CHECK_HIP(launch_ldir_round_group(b, pass, false, in, out, gfront(s, caps, pass, k + 1).counts, overflow,
                                              ctx->stream));
            (*extra)++;
            in = out;
            k++;
        }
        int rc = read_gctl(ctx, s);
        if (rc) return rc;
        if (s.h_ctl[kGOverflow + pass]) need_sweep = true;
    }
    return MSCEER_EHIP;
}

int ensure_state(msceer_ctx* ctx) {
    if (ctx->gbatch) return MSCEER_OK;
    BatchState* s = new (std::nothrow) BatchState();
    if (!s) return MSCEER_ENOMEM;
    ctx->gbatch = s;
    CHECK_HIP(hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking));
    CHECK_HIP(hipEventCreateWithFlags(&s->fork, hipEventDisableTiming));
    CHECK_HIP(hipEventCreateWithFlags(&s->join, hipEventDisableTiming));
    CHECK_HIP(hipEventCreate(&s->t0));
    CHECK_HIP(hipEventCreate(&s->t1));
    CHECK_HIP(hipHostMalloc((void**)&s->h_ctl, sizeof(uint32_t) * kGWords, hipHostMallocDefault));
    CHECK_HIP(hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    CHECK_HIP(hipDeviceGetAttribute(&s->clock_khz, hipDeviceAttributeWallClockRate, ctx->device));
    return MSCEER_OK;
}

// Run one group: plan (tables, tuning, graph) from the cache or new, replay,
// finish what the replay left, copy out.
int run_group(msceer_ctx* ctx, BatchState& s, msceer_volume* vols, GroupLayout& g, bool dev, bool timing,
              msceer_batch_stats& st) {
    const ListCaps caps = group_caps(g.verts);
    const size_t lbytes = sizeof(uint32_t) * 2 * ((size_t)caps.list_sub * kSubs + caps.list_spill);
    const size_t fbytes = sizeof(uint32_t) * 4 * ((size_t)caps.front_sub * kSubs + caps.front_spill);
    CHECK_HIP(grow(s, s.lists, lbytes));
    CHECK_HIP(grow(s, s.fronts, fbytes));
    CHECK_HIP(grow(s, s.ctl, sizeof(uint32_t) * kGWords));
    CHECK_HIP(grow(s, s.ldir, g.ldir_bytes));
    CHECK_HIP(grow(s, s.times, sizeof(unsigned long long) * 2 * (size_t)g.count));
    if (!dev) {
        CHECK_HIP(grow(s, s.vals, g.vals_bytes));
        CHECK_HIP(grow(s, s.grad, g.grad_bytes));
    }
    // tunings by shapes and tuning mode (and a fixed launch form, if set);
    // plans by these and every address the graph holds
    std::string tune_suffix;
    {
        std::lock_guard<std::mutex> lk(ctx->tuning->mu);
        tune_suffix = "|m" + std::to_string(ctx->tuning->mode);
    }
    if (const char* e = getenv("MSCEER_BATCH_LAUNCH")) tune_suffix += std::string("|") + e;
    std::string key = shapes_key(vols, g) + tune_suffix;
    char buf[160];
    snprintf(buf, sizeof buf, "|gen%llu|dev%d|t%d", (unsigned long long)s.generation, dev ? 1 : 0, timing ? 1 : 0);
    key += buf;
    if (dev) {
        for (int i = g.first; i < g.first + g.count; i++) {
            snprintf(buf, sizeof buf, "%p,%p;", (const void*)vols[i].values, (void*)vols[i].grad_out);
            key += buf;
        }
    }
    auto it = std::find_if(s.plans.begin(), s.plans.end(), [&](const GroupPlan& p) { return p.key == key; });
    if (it == s.plans.end()) {
        s.plans.emplace_front();
        it = s.plans.begin();
        it->key = key;
        // frontier rounds per pass in a first capture (later captures take
        // what the calls needed)
        // (0: none, each pass then starts from a sweep after the replay)
        if (const char* e = getenv("MSCEER_BATCH_ROUNDS")) it->rounds = std::max(0, std::min(64, atoi(e)));
        int rc = upload_tables(ctx, s, vols, g, dev, *it);
        if (rc) return rc;
        while ((int)s.plans.size() > kMaxPlans) {
            if (s.plans.back().exec) (void)hipGraphExecDestroy(s.plans.back().exec);
            s.plans.back().tables.release();
            s.plans.pop_back();
        }
    } else if (it != s.plans.begin()) {
        s.plans.splice(s.plans.begin(), s.plans, it);
        it = s.plans.begin();
    }
    GroupPlan& p = *it;
    if (!dev) {
        for (int i = 0; i < g.count; i++) {
            const msceer_volume& v = vols[g.first + i];
            CHECK_HIP(hipMemcpyAsync(s.vals.as<uint8_t>() + g.vals_off[i], v.values,
                                     sizeof(float) * (size_t)(v.nx * v.ny * v.nz), hipMemcpyHostToDevice,
                                     ctx->stream));
        }
    }
    if (!p.tuned) {
        const std::string skey = shapes_key(vols, g) + tune_suffix;
        auto t = s.tunes.find(skey);
        if (t != s.tunes.end()) {
            p.tune = t->second;
        } else {
            int rc = tune_group(ctx, s, p, caps, &p.tune);
            if (rc) return rc;
            s.tunes[skey] = p.tune;
        }
        p.tuned = true;
    }
    if (!p.exec) {
        hipGraph_t graph = nullptr;
        CHECK_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue_group(ctx, s, p, g, caps, p.rounds, timing);
        hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        CHECK_HIP(e);
        e = hipGraphInstantiate(&p.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        CHECK_HIP(e);
        st.captures++;
    }
    CHECK_HIP(hipEventRecord(s.t0, ctx->stream));
    CHECK_HIP(hipGraphLaunch(p.exec, ctx->stream));
    st.replays++;
    int rc = read_gctl(ctx, s);
    if (rc) return rc;
    const bool lists_lost = s.h_ctl[kGOverflow] != 0;
    int extra = 0, grow_by = 0;
    bool redo = false;
    for (int pass = 1; pass <= 2; pass++) {
        const int R = p.rounds;
        const bool left =
            R > 0 && list_total(s.h_ctl + kGRound + (3 * (pass - 1) + (R - 1) % 3) * kCountWords,
                                gfront(s, caps, pass, R - 1)) != 0;
        const bool sweep = R == 0 || lists_lost || s.h_ctl[kGOverflow + pass] != 0;
        if (!left && !sweep) continue;
        int used = 0;
        rc = continue_pass(ctx, s, p, g, vols, dev, caps, pass, R, sweep, &used);
        if (rc) return rc;
        extra += used;
        // more captured rounds help a walk that was still going, not an overflow
        if (!sweep || R == 0) grow_by = std::max(grow_by, used);
        redo = true;
    }
    if (redo) {
        GBatch b = p.batch;
        b.bricks = p.bricks[p.tune.shape];
        b.times = timing ? s.times.as<unsigned long long>() : nullptr;
        CHECK_HIP(launch_ldir_finish_group(b, ctx->stream));
        // the next capture of this plan takes the rounds this call needed
        if (grow_by > 0) {
            if (p.exec) (void)hipGraphExecDestroy(p.exec);
            p.exec = nullptr;
            p.rounds = std::min(p.rounds + grow_by, 64);
        }
        st.eager_rounds += extra;
    }
    CHECK_HIP(hipEventRecord(s.t1, ctx->stream));
    if (!dev) {
        for (int i = 0; i < g.count; i++) {
            msceer_volume& v = vols[g.first + i];
            const Dims d = make_dims(v.nx, v.ny, v.nz);
            CHECK_HIP(hipMemcpyAsync(v.grad_out, s.grad.as<uint8_t>() + g.grad_off[i], d.cells, hipMemcpyDeviceToHost,
                                     ctx->stream));
        }
    }
    std::vector<unsigned long long> t;
    if (timing) {
        t.resize(2 * (size_t)g.count);
        CHECK_HIP(hipMemcpyAsync(t.data(), s.times.ptr, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost,
                                 ctx->stream));
    }
    CHECK_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0.f;
    CHECK_HIP(hipEventElapsedTime(&ms, s.t0, s.t1));
    st.ms_gpu += ms;
    for (int k = 0; k < kSubs; k++)
        for (int d = 0; d < 4; d++) st.critical[d] += s.h_ctl[kGStats + 4 * k + d];
    st.shape = p.tune.shape;
    st.blocks = (int32_t)p.tune.blocks;
    st.chunk = (int32_t)p.tune.chunk;
    st.rounds = p.rounds;
    unsigned long long t_first = ~0ull;
    if (timing)
        for (int i = 0; i < g.count; i++) t_first = std::min(t_first, t[2 * i]);
    for (int i = 0; i < g.count; i++) {
        msceer_volume& v = vols[g.first + i];
        v.status = MSCEER_OK;
        // device times: from the group's first brick to the volume's last finished tile
        v.ms = timing && s.clock_khz > 0 && t[2 * i + 1] >= t_first
                   ? (float)((double)(t[2 * i + 1] - t_first) / (double)s.clock_khz)
                   : ms;
    }
    return MSCEER_OK;
}

}  // namespace

void batch_state_free(msceer_ctx* ctx) {
    BatchState* s = ctx->gbatch;
    if (!s) return;
    for (GroupPlan& p : s->plans) {
        if (p.exec) (void)hipGraphExecDestroy(p.exec);
        p.tables.release();
    }
    DevBuf* bufs[] = {&s->vals, &s->grad, &s->ldir, &s->lists, &s->fronts, &s->ctl, &s->times};
    for (DevBuf* b : bufs) b->release();
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
    hipEvent_t evs[] = {s->fork, s->join, s->t0, s->t1};
    for (hipEvent_t e : evs)
        if (e) (void)hipEventDestroy(e);
    if (s->side) (void)hipStreamDestroy(s->side);
    delete s;
    ctx->gbatch = nullptr;
}

}  // namespace msceer

extern "C" int msceer_batch_grad(msceer_ctx* ctx, msceer_volume* vols, int32_t n, int64_t max_group_verts,
                                 uint32_t flags) {
    if (!ctx || (!vols && n > 0) || n < 0 || max_group_verts < 0) return MSCEER_EINVAL;
    if (flags & ~(MSCEER_DEVICE_PTRS | MSCEER_TIMING)) return MSCEER_EINVAL;
    for (int32_t i = 0; i < n; i++) {
        const msceer_volume& v = vols[i];
        if (!v.values || !v.grad_out || v.labels_out) return MSCEER_EINVAL;
        const int rc = check_dims(v.nx, v.ny, v.nz);
        if (rc) return rc;
    }
    if (n == 0) return MSCEER_OK;
    CHECK_HIP(hipSetDevice(ctx->device));
    int rc = ensure_state(ctx);
    if (rc) return rc;
    BatchState& s = *ctx->gbatch;
    const bool dev = flags & MSCEER_DEVICE_PTRS, timing = flags & MSCEER_TIMING;
    for (int32_t i = 0; i < n; i++) vols[i].status = MSCEER_EINVAL;  // until its group has run
    msceer_batch_stats st{};
    int first = 0;
    while (first < n) {
        GroupLayout g;
        g.first = first;
        g.count = 0;
        g.verts = g.cells = 0;
        while (first + g.count < n) {
            const msceer_volume& v = vols[first + g.count];
            const Dims d = make_dims(v.nx, v.ny, v.nz);
            const bool fits = g.cells + d.cells <= kMaxItems &&
                              (max_group_verts == 0 || g.verts + d.verts <= (uint64_t)max_group_verts);
            if (g.count > 0 && !fits) break;
            g.verts += d.verts;
            g.cells += d.cells;
            g.count++;
        }
        layout(vols, g);
        rc = run_group(ctx, s, vols, g, dev, timing, st);
        if (rc) return rc;
        st.groups++;
        first += g.count;
    }
    s.last = st;
    return MSCEER_OK;
}

extern "C" int msceer_last_batch_stats(const msceer_ctx* ctx, msceer_batch_stats* out) {
    if (!ctx || !out) return MSCEER_EINVAL;
    *out = ctx->gbatch ? ctx->gbatch->last : msceer_batch_stats{};
    return MSCEER_OK;
}

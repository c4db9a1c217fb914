// msceer_msc_base: what extractmsc reads out of a .grad before any
// simplification (extractmsc/extractmsc.cxx:467-556).
//
//   nodes:    one per critical cell, numbered in cell-id order
//             (MorseSmaleComplexBasic::ComputeFromGrad,
//             gi_morse_smale_complex_basic.h:676-690);
//   basins:   per vertex, the node of the minimum its V-path descends to --
//             vertex, its paired edge, the edge's other vertex, ... --
//             (FillAscendingIndexVolumeByVPaths, extractmsc.cxx:371-440,
//             written to TESTBASINS.raw);
//   mounts:   per vertex, the node of the maximum whose descending manifold
//             holds the hex at that vertex (fillGeometry of the maxima,
//             rec_man_trace_down, gi_morse_smale_complex_basic.h:1181-1198:
//             a hex paired with a facet quad belongs to the manifold of the
//             quad's other hex; TESTMOUNTS.raw), -1 where there is none.
//
// On the GPU: (1) node numbering -- every work-group stages 4096 .grad bytes
// in LDS and counts its critical cells; one work-group scans the counts; every
// work-group then numbers its critical cells and writes, for each vertex and
// each hex, the next one on its path (itself at a critical cell, with the
// node id beside it).  (2) Pointer jumping over the two forests until every
// entry points at its root.  (3) The labels from the roots' node ids.
#include <hip/hip_runtime.h>

#include "ctx.hpp"

using namespace msceer;

namespace {

constexpr int kThreads = 256;
constexpr int kPerThread = 16;
constexpr int kChunk = kThreads * kPerThread;  // cells per work-group

struct MscGeom {
    uint32_t X, Y, Z, Mx, My, Mz;
    uint64_t cells;
};

__device__ __forceinline__ bool is_critical(uint32_t b) { return ((b >> 2) & 7u) == 7u; }

// The chunk's bytes in LDS (bytes past the end read as not critical).
__device__ __forceinline__ void stage_chunk(const uint8_t* __restrict__ grad, uint64_t cells, uint64_t base,
                                            uint8_t* lds) {
    for (int i = threadIdx.x; i < kChunk; i += kThreads) {
        const uint64_t c = base + (uint64_t)i;
        lds[i] = c < cells ? grad[c] : (uint8_t)0;
    }
}

// Exclusive prefix over the work-group of one value per lane; total in *total.
__device__ __forceinline__ uint32_t block_exclusive(uint32_t v, uint32_t* s_wave, uint32_t* total) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t t = __shfl_up(incl, o);
        if (lane >= (uint32_t)o) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    uint32_t before = 0, all = 0;
    for (uint32_t w = 0; w < kThreads / 64; w++) {
        if (w < wave) before += s_wave[w];
        all += s_wave[w];
    }
    *total = all;
    return before + incl - v;
}

// Critical cells per chunk, and by dimension.
__global__ void __launch_bounds__(kThreads)
msc_count_kernel(const uint8_t* __restrict__ grad, MscGeom g, uint32_t* __restrict__ chunk_counts,
                 unsigned long long* __restrict__ by_dim) {
    __shared__ uint8_t lds[kChunk];
    __shared__ uint32_t s_dim[4];
    __shared__ uint32_t s_wave[kThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kChunk;
    if (threadIdx.x < 4) s_dim[threadIdx.x] = 0;
    stage_chunk(grad, g.cells, base, lds);
    __syncthreads();
    uint32_t n = 0, dims[4] = {0, 0, 0, 0};
    const uint64_t first = base + (uint64_t)threadIdx.x * kPerThread;
    // coordinates of the lane's first cell, then stepped along x with carries
    uint32_t x = (uint32_t)(first % g.Mx), y = (uint32_t)((first / g.Mx) % g.My), z = (uint32_t)(first / ((uint64_t)g.Mx * g.My));
    for (int k = 0; k < kPerThread; k++) {
        if (first + k >= g.cells) break;
        if (is_critical(lds[threadIdx.x * kPerThread + k])) {
            n++;
            dims[(x & 1u) + (y & 1u) + (z & 1u)]++;
        }
        if (++x == g.Mx) {
            x = 0;
            if (++y == g.My) {
                y = 0;
                z++;
            }
        }
    }
    for (int d = 0; d < 4; d++)
        if (dims[d]) atomicAdd(&s_dim[d], dims[d]);
    uint32_t total;
    (void)block_exclusive(n, s_wave, &total);
    __syncthreads();
    if (threadIdx.x == 0) chunk_counts[blockIdx.x] = total;
    if (threadIdx.x < 4 && s_dim[threadIdx.x]) atomicAdd(&by_dim[threadIdx.x], (unsigned long long)s_dim[threadIdx.x]);
}

// Exclusive scan of n chunk counts in place, by one work-group.
__global__ void __launch_bounds__(1024) msc_scan_kernel(uint32_t* __restrict__ counts, uint32_t n) {
    __shared__ uint32_t s_sum[1024];
    const uint32_t per = (n + 1023u) / 1024u;
    const uint32_t lo = threadIdx.x * per, hi = min(n, lo + per);
    uint32_t sum = 0;
    for (uint32_t i = lo; i < hi; i++) sum += counts[i];
    s_sum[threadIdx.x] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int t = 0; t < 1024; t++) {
            const uint32_t v = s_sum[t];
            s_sum[t] = run;
            run += v;
        }
    }
    __syncthreads();
    uint32_t run = s_sum[threadIdx.x];
    for (uint32_t i = lo; i < hi; i++) {
        const uint32_t v = counts[i];
        counts[i] = run;
        run += v;
    }
}

// Node ids of the chunk's critical cells; the next vertex / hex of every
// vertex and hex of the chunk (itself at a root, its node id or -1 in rank).
__global__ void __launch_bounds__(kThreads)
msc_link_kernel(const uint8_t* __restrict__ grad, MscGeom g, const uint32_t* __restrict__ chunk_first,
                int32_t* __restrict__ vnext, int32_t* __restrict__ vrank, int32_t* __restrict__ hnext,
                int32_t* __restrict__ hrank) {
    __shared__ uint8_t lds[kChunk];
    __shared__ uint32_t s_wave[kThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kChunk;
    stage_chunk(grad, g.cells, base, lds);
    __syncthreads();
    const uint64_t first = base + (uint64_t)threadIdx.x * kPerThread;
    uint32_t n = 0;
    for (int k = 0; k < kPerThread; k++)
        if (first + k < g.cells && is_critical(lds[threadIdx.x * kPerThread + k])) n++;
    uint32_t total;
    uint32_t rank = chunk_first[blockIdx.x] + block_exclusive(n, s_wave, &total);
    const uint32_t Xh = g.X - 1, Yh = g.Y - 1, Zh = g.Z - 1;
    uint32_t x = (uint32_t)(first % g.Mx), y = (uint32_t)((first / g.Mx) % g.My), z = (uint32_t)(first / ((uint64_t)g.Mx * g.My));
    for (int k = 0; k < kPerThread; k++) {
        if (first + k >= g.cells) break;
        const uint32_t b = lds[threadIdx.x * kPerThread + k];
        const bool crit = is_critical(b);
        const uint32_t odd = (x & 1u) + (y & 1u) + (z & 1u);
        if (odd == 0u || odd == 3u) {
            // a vertex (i, j, k) or a hex at vertex (i, j, k)
            const uint32_t i = x >> 1, j = y >> 1, kk = z >> 1;
            const bool vert = odd == 0u;
            const uint32_t nx = vert ? g.X : Xh, ny = vert ? g.Y : Yh, nz = vert ? g.Z : Zh;
            const int32_t self = (int32_t)(i + nx * (j + ny * kk));
            int32_t next = self, r = crit ? (int32_t)rank : -1;
            if (!crit) {
                // one cell along the pair direction is the paired edge / quad, two the next vertex / hex
                const uint32_t d = (b >> 2) & 7u, a = d >> 1;
                const int s = (d & 1u) ? -1 : 1;
                const int64_t ti = (int64_t)i + (a == 0 ? s : 0), tj = (int64_t)j + (a == 1 ? s : 0);
                const int64_t tk = (int64_t)kk + (a >= 2 ? s : 0);
                if (ti >= 0 && tj >= 0 && tk >= 0 && ti < nx && tj < ny && tk < nz) {
                    next = (int32_t)(ti + (int64_t)nx * (tj + (int64_t)ny * tk));
                    r = -1;
                }
                // else: off the domain (a hex paired with a boundary quad): a root of no node
            }
            (vert ? vnext : hnext)[self] = next;
            if (next == self) (vert ? vrank : hrank)[self] = r;
        }
        if (crit) rank++;
        if (++x == g.Mx) {
            x = 0;
            if (++y == g.My) {
                y = 0;
                z++;
            }
        }
    }
}

// Pointer jumping: every entry moves up to 8 links towards its root; *changed
// set when an entry moved.
__global__ void __launch_bounds__(kThreads)
msc_jump_kernel(int32_t* __restrict__ next, uint32_t n, uint32_t* __restrict__ changed) {
    bool moved = false;
    for (uint32_t i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads) {
        const int32_t p0 = next[i];
        int32_t p = p0;
        for (int k = 0; k < 8; k++) {
            const int32_t q = next[p];
            if (q == p) break;
            p = q;
        }
        if (p != p0) {
            next[i] = p;
            moved = true;
        }
    }
    if (__any(moved) && (threadIdx.x & 63u) == 0u) atomicOr(changed, 1u);
}

__global__ void __launch_bounds__(kThreads)
msc_label_kernel(MscGeom g, const int32_t* __restrict__ vnext, const int32_t* __restrict__ vrank,
                 const int32_t* __restrict__ hnext, const int32_t* __restrict__ hrank, int32_t* __restrict__ basins,
                 int32_t* __restrict__ mounts) {
    const uint64_t V = (uint64_t)g.X * g.Y * g.Z;
    for (uint64_t v = (uint64_t)blockIdx.x * kThreads + threadIdx.x; v < V; v += (uint64_t)gridDim.x * kThreads) {
        if (basins) basins[v] = vrank[vnext[v]];
        if (mounts) {
            const uint32_t i = (uint32_t)(v % g.X), j = (uint32_t)((v / g.X) % g.Y), k = (uint32_t)(v / ((uint64_t)g.X * g.Y));
            int32_t m = -1;
            if (i + 1 < g.X && j + 1 < g.Y && k + 1 < g.Z) {
                const uint32_t h = i + (g.X - 1) * (j + (g.Y - 1) * k);
                m = hrank[hnext[h]];
            }
            mounts[v] = m;
        }
    }
}

uint32_t grid_for(uint64_t n) {
    uint64_t b = (n + kThreads - 1) / kThreads;
    return (uint32_t)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

// Jump until no entry moves (one read-back per launch).
int jump_all(msceer_ctx* ctx, int32_t* next, uint32_t n, uint32_t* flag, int* rounds) {
    if (n == 0) return MSCEER_OK;
    for (int it = 0; it < 64; it++) {
        CHECK_HIP(launch_zero(flag, 1, ctx->stream));
        hipLaunchKernelGGL(msc_jump_kernel, dim3(grid_for(n)), dim3(kThreads), 0, ctx->stream, next, n, flag);
        CHECK_HIP(hipGetLastError());
        (*rounds)++;
        uint32_t moved = 0;
        CHECK_HIP(hipMemcpyAsync(&moved, flag, sizeof moved, hipMemcpyDeviceToHost, ctx->stream));
        CHECK_HIP(hipStreamSynchronize(ctx->stream));
        if (!moved) return MSCEER_OK;
    }
    return MSCEER_EHIP;  // 8^64 links: a cycle, not a gradient
}

}  // namespace

extern "C" int msceer_msc_base(msceer_ctx* ctx, const uint8_t* grad, int64_t nx, int64_t ny, int64_t nz,
                               int32_t* basins_out, int32_t* mounts_out, int64_t* counts_out, uint32_t flags) {
    if (!ctx || !grad) return MSCEER_EINVAL;
    if (flags & ~(MSCEER_DEVICE_PTRS)) return MSCEER_EINVAL;
    int rc = check_dims(nx, ny, nz);
    if (rc) return rc;
    if (nx * ny * nz >= (int64_t)1 << 31) return MSCEER_ETOOBIG;  // int32 vertex ids and node ids
    CHECK_HIP(hipSetDevice(ctx->device));
    const Dims d = make_dims(nx, ny, nz);
    const bool dev = flags & MSCEER_DEVICE_PTRS;
    hipStream_t s = ctx->stream;
    const uint64_t V = d.verts, H = (uint64_t)(nx - 1) * (ny - 1) * (nz - 1);
    const uint32_t chunks = (uint32_t)((d.cells + kChunk - 1) / kChunk);
    // scratch: next and rank of vertices and hexes, chunk counts, dimension counts, flag
    const size_t vb = (4 * V + 255) & ~(size_t)255, hb = (4 * (H ? H : 1) + 255) & ~(size_t)255;
    const size_t cb = (4 * (size_t)chunks + 255) & ~(size_t)255;
    CHECK_HIP(ctx->scan.ensure(2 * vb + 2 * hb + cb + 256));
    uint8_t* sc = ctx->scan.as<uint8_t>();
    int32_t* vnext = reinterpret_cast<int32_t*>(sc);
    int32_t* vrank = reinterpret_cast<int32_t*>(sc + vb);
    int32_t* hnext = reinterpret_cast<int32_t*>(sc + 2 * vb);
    int32_t* hrank = reinterpret_cast<int32_t*>(sc + 2 * vb + hb);
    uint32_t* ccount = reinterpret_cast<uint32_t*>(sc + 2 * vb + 2 * hb);
    unsigned long long* by_dim = reinterpret_cast<unsigned long long*>(sc + 2 * vb + 2 * hb + cb);
    uint32_t* flag = reinterpret_cast<uint32_t*>(sc + 2 * vb + 2 * hb + cb + 64);
    const uint8_t* dgrad = grad;
    if (!dev) {
        CHECK_HIP(ctx->grad.ensure(d.cells));
        CHECK_HIP(hipMemcpyAsync(ctx->grad.ptr, grad, d.cells, hipMemcpyHostToDevice, s));
        dgrad = ctx->grad.as<uint8_t>();
    }
    int32_t* dbasins = basins_out;
    int32_t* dmounts = mounts_out;
    if (!dev && (basins_out || mounts_out)) {
        CHECK_HIP(ctx->label.ensure(8 * V));
        dbasins = basins_out ? ctx->label.as<int32_t>() : nullptr;
        dmounts = mounts_out ? ctx->label.as<int32_t>() + V : nullptr;
    }
    const MscGeom g{(uint32_t)nx, (uint32_t)ny, (uint32_t)nz, d.Mx, d.My, d.Mz, d.cells};
    CHECK_HIP(launch_zero(reinterpret_cast<uint32_t*>(by_dim), 8, s));
    hipLaunchKernelGGL(msc_count_kernel, dim3(chunks), dim3(kThreads), 0, s, dgrad, g, ccount, by_dim);
    CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(msc_scan_kernel, dim3(1), dim3(1024), 0, s, ccount, chunks);
    CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(msc_link_kernel, dim3(chunks), dim3(kThreads), 0, s, dgrad, g, ccount, vnext, vrank, hnext,
                       hrank);
    CHECK_HIP(hipGetLastError());
    int rounds = 0;
    if (basins_out) {
        rc = jump_all(ctx, vnext, (uint32_t)V, flag, &rounds);
        if (rc) return rc;
    }
    if (mounts_out) {
        rc = jump_all(ctx, hnext, (uint32_t)H, flag, &rounds);
        if (rc) return rc;
    }
    if (basins_out || mounts_out) {
        hipLaunchKernelGGL(msc_label_kernel, dim3(grid_for(V)), dim3(kThreads), 0, s, g, vnext, vrank, hnext, hrank,
                           dbasins, dmounts);
        CHECK_HIP(hipGetLastError());
    }
    if (!dev) {
        if (basins_out) CHECK_HIP(hipMemcpyAsync(basins_out, dbasins, 4 * V, hipMemcpyDeviceToHost, s));
        if (mounts_out) CHECK_HIP(hipMemcpyAsync(mounts_out, dmounts, 4 * V, hipMemcpyDeviceToHost, s));
    }
    unsigned long long dims[4] = {0, 0, 0, 0};
    CHECK_HIP(hipMemcpyAsync(dims, by_dim, sizeof dims, hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    if (counts_out)
        for (int k = 0; k < 4; k++) counts_out[k] = (int64_t)dims[k];
    return MSCEER_OK;
}

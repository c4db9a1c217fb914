// Region boundary kinds of a vertex labelling (the `_edges_` output of
// Accurate3DMorseLabeling): per cell of the complex, how many regions meet
// there.
//
// Restates VertexLabelingToBoundaryLabeling::ComputeRegionBoundaryKind
// (gi_labeling_to_bounary_labeling.h:262-393), used by
// accurate3dmorselabeling.cxx:337-343:
//   edge:  1 if its two vertices carry different valid labels, 0 if equal;
//   quad:  0 with no such edge, 1 with one or two, else 2 if its four
//          vertices carry more than two labels, else 1;
//   hex:   with more than two quads of kind 2, 3 if its eight vertices carry
//          more than three labels, else 2; otherwise the largest kind of its
//          quads.
// The reference leaves vertices, and edges with an unlabelled (-1) vertex,
// unwritten (its array is not initialised); they are 0 here.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "ctx.hpp"
#include "integrate.hpp"

namespace msceer {
namespace {

constexpr int kThreads = 256;

// Distinct values among n labels, counting up to `stop` (the reference stops
// counting once the answer is decided).
__device__ __forceinline__ int distinct(const int32_t* l, int n, int stop) {
    int32_t seen[8];
    int u = 0;
    for (int i = 0; i < n && u <= stop; i++) {
        bool fresh = true;
        for (int j = 0; j < u; j++) fresh = fresh && seen[j] != l[i];
        if (fresh) seen[u++] = l[i];
    }
    return u;
}

template <int kDim>
__global__ void __launch_bounds__(kThreads)
region_kind_kernel(VGrid g, const int32_t* __restrict__ label, int8_t* __restrict__ out) {
    const int64_t Mx = 2 * (int64_t)g.X - 1, My = 2 * (int64_t)g.Y - 1, Mz = 2 * (int64_t)g.Z - 1;
    const int64_t C = Mx * My * Mz;
    const int64_t off[3] = {1, Mx, Mx * My};
    for (int64_t c = (int64_t)blockIdx.x * kThreads + threadIdx.x; c < C; c += (int64_t)gridDim.x * kThreads) {
        const int64_t p[3] = {c % Mx, (c / Mx) % My, c / (Mx * My)};
        const int odd = (int)((p[0] & 1) + (p[1] & 1) + (p[2] & 1));
        if (odd != kDim) continue;
        if (kDim == 0) {
            out[c] = 0;
            continue;
        }
        // the cell's vertices: corners (p - odd axes + 2 * bits) / 2
        int32_t l[8];
        int n = 0;
        for (int k = 0; k < 8; k++) {
            if (((k & 1) && !(p[0] & 1)) || ((k & 2) && !(p[1] & 1)) || ((k & 4) && !(p[2] & 1))) continue;
            const int64_t vx = (p[0] >> 1) + (k & 1), vy = (p[1] >> 1) + ((k >> 1) & 1), vz = (p[2] >> 1) + (k >> 2);
            l[n++] = label[vx + (int64_t)g.X * (vy + (int64_t)g.Y * vz)];
        }
        if (kDim == 1) {
            out[c] = (l[0] == -1 || l[1] == -1) ? 0 : (l[0] != l[1] ? 1 : 0);
            continue;
        }
        // facets along the odd axes
        int sum = 0, maxv = 0;
        for (int a = 0; a < 3; a++) {
            if (!(p[a] & 1)) continue;
            for (int sg = -1; sg <= 1; sg += 2) {
                const int8_t v = out[c + sg * off[a]];
                if (kDim == 2) sum += v == 1;
                else {
                    maxv = v > maxv ? v : maxv;
                    sum += v == 2;
                }
            }
        }
        if (kDim == 2) out[c] = sum == 0 ? 0 : (sum <= 2 ? 1 : (distinct(l, 4, 2) > 2 ? 2 : 1));
        else out[c] = sum > 2 ? (distinct(l, 8, 3) > 3 ? 3 : 2) : (int8_t)maxv;
    }
}

}  // namespace

hipError_t launch_region_kind(const VGrid& g, const int32_t* label, int8_t* out, hipStream_t s) {
    const int64_t C = (2 * (int64_t)g.X - 1) * (2 * (int64_t)g.Y - 1) * (2 * (int64_t)g.Z - 1);
    const unsigned blocks = (unsigned)std::min<int64_t>((C + kThreads - 1) / kThreads, 8192);
    hipLaunchKernelGGL(region_kind_kernel<0>, dim3(blocks), dim3(kThreads), 0, s, g, label, out);
    hipLaunchKernelGGL(region_kind_kernel<1>, dim3(blocks), dim3(kThreads), 0, s, g, label, out);
    hipLaunchKernelGGL(region_kind_kernel<2>, dim3(blocks), dim3(kThreads), 0, s, g, label, out);
    hipLaunchKernelGGL(region_kind_kernel<3>, dim3(blocks), dim3(kThreads), 0, s, g, label, out);
    return hipGetLastError();
}

}  // namespace msceer

// Host emulation of the accurate 1-manifold kernels (test infrastructure):
// the same lines_core.hpp the gfx950 kernels run, driven window by window
// on the CPU -- every speculative trace, the union-find over footprints, the
// in-order replay of each component, the commit of the lines before a cut --
// with the lanes of a launch taken one after another in a chosen order (the
// order only decides which line's mark a shared cell keeps, which the
// algorithm must not depend on).  Its restriction labels are compared with
// the oracle's one-thread saddle loops.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

// what the HIP build provides to the device headers
#define __host__
#define __device__
#define __forceinline__ inline
#define __restrict__
#define __HIP_MEMORY_SCOPE_AGENT 0
using std::min;
struct float4 { float x, y, z, w; };
struct double4 { double x, y, z, w; };
inline uint32_t atomicMin(uint32_t* p, uint32_t v) { const uint32_t o = *p; if (v < o) *p = v; return o; }
inline uint32_t atomicAdd(uint32_t* p, uint32_t v) { const uint32_t o = *p; *p = o + v; return o; }
inline uint32_t __hip_atomic_load(uint32_t* p, int, int) { return *p; }
inline void __hip_atomic_store(uint32_t* p, uint32_t v, int, int) { *p = v; }
inline bool __hip_atomic_compare_exchange_strong(uint32_t* p, uint32_t* expected, uint32_t desired, int, int, int) {
    if (*p == *expected) { *p = desired; return true; }
    *expected = *p;
    return false;
}
namespace msceer {
struct VGrid {
    int X, Y, Z;
    int64_t sxy;
    int64_t verts() const { return sxy * Z; }
};
}  // namespace msceer

#include "lines_core.hpp"

using namespace msceer;

namespace {

struct Window {
    std::vector<uint32_t> fhead, whead, mhead, parent;
    std::vector<uint8_t> dep;
};

}  // namespace

// stats (8): saddles, windows, cuts, traced alone, speculative traces,
//            replayed traces, replay stops on a foreign test, on a foreign label
extern "C" int emul_lines(const float* f, long long X, long long Y, long long Z, const uint8_t* grad,
                          const double* field3, int8_t* restr, int asc, float ethr, float gthr, int iters,
                          long long window, int order, long long* stats) {
    const CellGeom m{2 * X - 1, 2 * Y - 1, 2 * Z - 1, (2 * X - 1) * (2 * Y - 1)};
    const uint64_t C = (uint64_t)m.sxy * (uint64_t)m.Mz;
    const VGrid g{(int)X, (int)Y, (int)Z, X * Y};
    std::vector<double4> gd((size_t)(X * Y * Z));
    for (size_t v = 0; v < gd.size(); v++) gd[v] = double4{field3[3 * v], field3[3 * v + 1], field3[3 * v + 2], 0.0};
    const GradSrc<double> src{gd.data()};
    // saddles in processing order
    const int sdim = asc ? 2 : 1;
    std::vector<uint64_t> keys;
    for (uint64_t c = 0; c < C; c++) {
        if (!critical(grad, (uint32_t)c)) continue;
        int64_t x, y, z;
        m.xyz((uint32_t)c, x, y, z);
        if ((int)((x & 1) + (y & 1) + (z & 1)) != sdim) continue;
        int64_t best = -1;
        for (int k = 0; k < 8; k++) {
            if (((k & 1) && !(x & 1)) || ((k & 2) && !(y & 1)) || ((k & 4) && !(z & 1))) continue;
            const int64_t v = ((x >> 1) + (k & 1)) + X * (((y >> 1) + ((k >> 1) & 1)) + Y * ((z >> 1) + (k >> 2)));
            if (best < 0 || f[v] > f[best] || (!(f[best] > f[v]) && v > best)) best = v;
        }
        const float val = f[best];
        keys.push_back(((uint64_t)value_key(asc ? val : -1.0f * val) << 32) | c);
    }
    std::sort(keys.rbegin(), keys.rend());
    const uint64_t S = keys.size(), nslots = 7 * S;
    std::vector<uint32_t> slots(nslots);
    for (uint64_t k = 0; k < S; k++) saddle_seeds(grad, m, (uint32_t)keys[k], asc != 0, slots.data() + 7 * k);

    std::vector<uint32_t> wstamp(C, kNoLine), pstamp(C, kNoLine), lstamp(C, kNoLine), own(C, kNoLine);
    const uint32_t pool_cap = 1u << 26;
    std::vector<uint32_t> pool(pool_cap), touched(pool_cap);
    const int8_t target = asc ? 4 : 5;
    long long st[8] = {(long long)S, 0, 0, 0, 0, 0, 0, 0};
    std::mt19937 rng(12345);
    const uint32_t W = (uint32_t)std::max<long long>(1, window);
    uint64_t w0 = 0;
    while (w0 < nslots) {
        const uint32_t a0 = (uint32_t)w0, w1 = (uint32_t)std::min<uint64_t>(nslots, w0 + W), n = w1 - a0;
        st[1]++;
        Window w;
        w.fhead.assign(n, kNoLine);
        w.whead.assign(n, kNoLine);
        w.mhead.assign(n, kNoLine);
        uint32_t top = 0, ntouched = 0, overflow = kNoLine;
        // 1. speculation, lanes in the chosen order
        std::vector<uint32_t> lanes(n);
        for (uint32_t r = 0; r < n; r++) lanes[r] = r;
        if (order == 1) std::reverse(lanes.begin(), lanes.end());
        if (order == 2) std::shuffle(lanes.begin(), lanes.end(), rng);
        for (uint32_t r : lanes) {
            const uint32_t i = a0 + r;
            SpecLabels lab{restr, target, wstamp.data(), pstamp.data(), own.data(), i,
                           Pool{pool.data(), pool_cap, &top}, touched.data(), &ntouched};
            const uint32_t c = slots[i];
            if (c != kNoLine) {
                if (i % 7 == 6) lab.write(c);
                else {
                    uint32_t vox = 0;
                    trace_line(src, g, m, c, asc != 0, (double)gthr, (double)ethr, iters, lab, &vox);
                    st[4]++;
                }
            }
            lab.finish();
            w.fhead[r] = lab.over ? kNoLine : lab.all.head;
            w.whead[r] = lab.over ? kNoLine : lab.wrote.head;
            w.mhead[r] = lab.over ? kNoLine : lab.maybe.head;
            if (lab.over) overflow = std::min(overflow, i);
        }
        if (overflow == a0) return -1;  // the emulation's pool is sized never to run out
        const uint32_t wc = overflow < w1 ? overflow : w1, nc = wc - a0;
        // 2. components and dependence
        w.parent.resize(nc);
        w.dep.assign(nc, 0);
        for (uint32_t r = 0; r < nc; r++) w.parent[r] = r;
        for (uint32_t r = 0; r < nc; r++) {
            const uint32_t i = a0 + r;
            for (uint32_t ch = w.fhead[r]; ch != kNoLine; ch = pool[ch])
                for (uint32_t j = 1; j < kChunk; j++) {
                    const uint32_t c = pool[ch + j];
                    if (c == kNoLine) break;
                    for (uint32_t ws : {wstamp[c], pstamp[c]})
                        if (ws != i && ws >= a0 && ws < wc) {
                            lf_union(w.parent.data(), r, ws - a0);
                            if (ws < i) w.dep[r] = 1;
                        }
                }
        }
        std::vector<uint64_t> ck(nc);
        for (uint32_t r = 0; r < nc; r++) ck[r] = ((uint64_t)lf_find(w.parent.data(), r) << 32) | r;
        std::sort(ck.begin(), ck.end());
        // 3. replay, one component after another
        uint32_t fail = kNoLine, why[2] = {0, 0};
        for (uint32_t q0 = 0; q0 < nc;) {
            const uint32_t root = (uint32_t)(ck[q0] >> 32);
            uint32_t q1 = q0;
            while (q1 < nc && (uint32_t)(ck[q1] >> 32) == root) q1++;
            for (uint32_t q = q0; q < q1; q++) {
                const uint32_t r = (uint32_t)ck[q], i = a0 + r;
                if (!w.dep[r]) {
                    for (uint32_t ch = w.fhead[r]; ch != kNoLine; ch = pool[ch])
                        for (uint32_t j = 1; j < kChunk; j++) {
                            const uint32_t c = pool[ch + j];
                            if (c == kNoLine) break;
                            if (wstamp[c] == i) lstamp[c] = i;
                        }
                    continue;
                }
                ReplayLabels lab{restr,       target,     wstamp.data(),   pstamp.data(), lstamp.data(), own.data(),
                                 pool.data(), w.whead[r], w.mhead[r], w.parent.data(), a0, wc, root, i, why, false};
                const uint32_t c = slots[i];
                if (c != kNoLine) {
                    if (i % 7 == 6) lab.write(c);
                    else {
                        uint32_t vox = 0;
                        trace_line(src, g, m, c, asc != 0, (double)gthr, (double)ethr, iters, lab, &vox);
                        st[5]++;
                    }
                }
                if (lab.bad) {
                    fail = std::min(fail, i);
                    break;
                }
            }
            q0 = q1;
        }
        st[6] += why[0];
        st[7] += why[1];
        uint32_t cut = fail < wc ? fail : wc;
        const uint32_t nt = std::min(ntouched, pool_cap);
        if (cut == a0) {
            // traced alone, in order
            for (uint32_t t = 0; t < nt; t++) {
                const uint32_t c = touched[t];
                wstamp[c] = pstamp[c] = lstamp[c] = own[c] = kNoLine;
            }
            DirectLabels lab{restr, target};
            const uint32_t c = slots[a0];
            if (c != kNoLine) {
                if (a0 % 7 == 6) lab.write(c);
                else {
                    uint32_t vox = 0;
                    trace_line(src, g, m, c, asc != 0, (double)gthr, (double)ethr, iters, lab, &vox);
                }
            }
            st[3]++;
            w0 = a0 + 1;
            continue;
        }
        // 4. commit, clear
        for (uint32_t t = 0; t < nt; t++) {
            const uint32_t c = touched[t];
            if (lstamp[c] < cut) restr[c] = target;
        }
        for (uint32_t t = 0; t < nt; t++) {
            const uint32_t c = touched[t];
            wstamp[c] = pstamp[c] = lstamp[c] = own[c] = kNoLine;
        }
        if (cut < wc) st[2]++;
        w0 = cut;
    }
    if (stats) std::copy(st, st + 8, stats);
    return 0;
}

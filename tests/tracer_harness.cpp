// Host build of the GPU streamline tracer (msceer_amd/csrc/tracer.hpp, the
// code pass 3 runs in FP64) against the oracle's FP64 advector
// (oracle/integ_oracle.cpp, included so that its internal functions are
// visible): from every start vertex, flowing down (sign -1) and up (+1), both
// must visit the same vertices and end at the same vertex with the same event.
// Test infrastructure only.
//
//   usage: tracer_harness  -> prints "streamlines N differing D", exit 0 iff D == 0
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <algorithm>
#include <vector>

// what the HIP build provides to the tracer header
#define __host__
#define __device__
#define __forceinline__ inline
#define __restrict__
using std::min;
struct float4 { float x, y, z, w; };
struct double4 { double x, y, z, w; };
namespace msceer {
struct VGrid {
    int X, Y, Z;
    int64_t sxy;
    int64_t verts() const { return sxy * Z; }
};
}  // namespace msceer

#include "../msceer_amd/csrc/tracer.hpp"
#include "../oracle/integ_oracle.cpp"

namespace {

uint64_t lcg(uint64_t& s) {
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return s >> 11;
}
double unit(uint64_t& s) { return (double)lcg(s) * (1.0 / 9007199254740992.0); }

// tracer streamline: visited vertices (after each voxel and at the end), event
int gpu_trace(const msceer::GradSrc<double>& src, const msceer::VGrid& g, const uint8_t* dest, bool stop, double dir,
              double gthr, double ethr, int x, int y, int z, int max_iters, std::vector<long long>& path) {
    using namespace msceer;
    Tracer<double> t;
    tracer_start(src, g, t, x, y, z, max_iters);
    while (true) {
        int64_t next = 0;
        bool moved = false;
        int ev = tracer_iterate(src, g, dest, stop, dir, gthr, ethr, t, next, moved);
        if (ev == kOutOfVoxel) {
            next = tracer_reenter(src, g, t);
            moved = true;
            ev = kNone;
        }
        if (moved) path.push_back(next);
        if (ev == kNone) continue;
        if (ev != kHitPreassigned) next = nearest(g, t.cur, t.near);
        path.push_back(next);
        return ev;
    }
}

int oracle_event(int ev) {
    switch (ev) {
        case msceer::kHitPreassigned: return HIT_PREASSIGNED;
        case msceer::kLowGradient: return LOW_GRADIENT;
        case msceer::kOverMax: return OVER_MAX_ITERATIONS;
    }
    return -1;
}

}  // namespace

int main() {
    const int X = 40, Y = 36, Z = 32;
    const long long V = (long long)X * Y * Z;
    // a sum of 30 Gaussians (sigma 2-7) plus noise of 1e-4
    uint64_t s = 12345;
    std::vector<double> cx, cy, cz, sg, am;
    for (int k = 0; k < 30; k++) {
        cx.push_back(unit(s) * (X - 1)); cy.push_back(unit(s) * (Y - 1)); cz.push_back(unit(s) * (Z - 1));
        sg.push_back(2.0 + 5.0 * unit(s)); am.push_back(2.0 * unit(s) - 1.0);
    }
    std::vector<float> f((size_t)V);
    for (int z = 0; z < Z; z++)
        for (int y = 0; y < Y; y++)
            for (int x = 0; x < X; x++) {
                double v = 1e-4 * (2.0 * unit(s) - 1.0);
                for (size_t k = 0; k < cx.size(); k++) {
                    const double d2 = (x - cx[k]) * (x - cx[k]) + (y - cy[k]) * (y - cy[k]) + (z - cz[k]) * (z - cz[k]);
                    v += am[k] * std::exp(-d2 / (2.0 * sg[k] * sg[k]));
                }
                f[(size_t)(x + X * (y + Y * (long long)z))] = (float)v;
            }
    const Grid og(X, Y, Z);
    std::vector<V3d> grad;
    gradient_field(og, f.data(), grad);
    std::vector<double4> gd((size_t)V);
    for (long long i = 0; i < V; i++) gd[(size_t)i] = double4{grad[(size_t)i].v[0], grad[(size_t)i].v[1], grad[(size_t)i].v[2], 0.0};
    const msceer::GradSrc<double> src{gd.data()};
    const msceer::VGrid g{X, Y, Z, (int64_t)X * Y};
    std::vector<uint8_t> dest((size_t)V);
    const float gthr = 1e-6f, ethr = 1e-3f;
    long long total = 0, differing = 0, steps_path = 0;
    for (int trial = 0; trial < 3; trial++) {
        // certain terminals, assigned and unassigned vertices at random
        for (long long i = 0; i < V; i++) {
            const double r = unit(s);
            const double term = trial == 0 ? 0.002 : (trial == 1 ? 0.02 : 0.1);
            dest[(size_t)i] = r < term ? CERTAIN_TERMINAL : (r < 0.3 ? ASSIGNED : UNASSIGNED);
        }
        for (int stop = 0; stop < 2; stop++) {
            const Checker chk{reinterpret_cast<const Dest*>(dest.data()), stop != 0, &og};
            for (int mi : {4096, 60, 7})
            for (int sign : {-1, 1}) {
                for (long long v = trial; v < V; v += 5) {
                    const int x = (int)(v % X), y = (int)((v / X) % Y), z = (int)(v / ((long long)X * Y));
                    std::vector<long long> po, pg;
                    Event eo;
                    const long long end = trace<true>(og, grad, chk, gthr, ethr, mi, v, &po, &eo, sign);
                    const int eg = gpu_trace(src, g, dest.data(), stop != 0, (double)sign, (double)gthr, (double)ethr,
                                             x, y, z, mi, pg);
                    total++;
                    steps_path += (long long)po.size();
                    if (oracle_event(eg) != (int)eo || pg != po || pg.back() != end) {
                        if (differing < 5)
                            printf("differ: start %lld stop %d budget %d sign %d: event %d/%d, path %zu/%zu vertices\n",
                                   v, stop, mi, sign, (int)eo, eg, po.size(), pg.size());
                        differing++;
                    }
                }
            }
        }
    }
    printf("streamlines %lld differing %lld path vertices %lld\n", total, differing, steps_path);
    return differing == 0 ? 0 : 1;
}

// Streamline tracer of the accurate labelling: adaptive Euler with step
// doubling, one lane's state advanced one iteration at a time
// (gi_adaptive_euler_advector.h:51-305, gi_advection_checker.h:212-262,
// gi_regular_grid_trilinear_function.h:211-229).
//
// Device code of integrate.hip, kept in this header so that a host build of
// the same source can be checked against the CPU oracle.  The includer
// provides uint8_t/int64_t, the HIP qualifiers, float4/double4, sqrt, min and
// VGrid (integrate.hpp).  Products and sums are not fused into FMAs, so each
// rounds as in the reference's scalar code.
#pragma once

#pragma clang fp contract(off)

namespace msceer {
namespace {

enum Dest : uint8_t { kBackground = 0, kUnassigned = 1, kAssigned = 2, kCertainTerminal = 3, kCertainNonterminal = 4 };
enum Event : int { kNone = 0, kOutOfVoxel = 1, kOverMax = 2, kLowGradient = 3, kHitPreassigned = 5 };

template <typename T>
struct V3 { T x, y, z; };

// Gradient at a vertex: the stored FP32 field, or (FP64) the central
// difference recomputed from the values in double, as the reference does.
template <typename T>
struct GradSrc;
template <>
struct GradSrc<float> {
    const float4* gradf;
    const float* f;
    const uint8_t* seedmask;
    VGrid g;
    __device__ __forceinline__ V3<float> at(int64_t v) const {
        const float4 r = gradf[v];
        return {r.x, r.y, r.z};
    }
};
template <>
struct GradSrc<double> {
    const double4* gradd;
    __device__ __forceinline__ V3<double> at(int64_t v) const {
        const double4 r = gradd[v];
        return {r.x, r.y, r.z};
    }
};

template <typename T>
__device__ __forceinline__ V3<T> lerp3(const V3<T>& a, const V3<T>& b, T t) {
    const T u = (T)1 - t;
    return {a.x * u + b.x * t, a.y * u + b.y * t, a.z * u + b.z * t};
}

template <typename T>
__device__ __forceinline__ V3<T> trilin(const V3<T> c[8], T fx, T fy, T fz) {
    const V3<T> x0 = lerp3(c[0], c[1], fx), x1 = lerp3(c[2], c[3], fx);
    const V3<T> x2 = lerp3(c[4], c[5], fx), x3 = lerp3(c[6], c[7], fx);
    const V3<T> y0 = lerp3(x0, x1, fy), y1 = lerp3(x2, x3, fy);
    return lerp3(y0, y1, fz);
}

template <typename T>
__device__ __forceinline__ T clampc(T p, int n) {
    if (p < (T)0) p = (T)0;
    if (p >= (T)n) p = (T)(n - 1);
    return p;
}

// Nearest vertex of a position (truncation of p + 0.5, clamped near the
// boundary as Inbounds does).
__device__ __forceinline__ int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// (Away from the boundary the position lies inside the grid and the final
// index clamp changes nothing; it only keeps a corrupt position in bounds.)
template <typename T>
__device__ __forceinline__ int64_t nearest_xyz(const VGrid& g, const V3<T>& p, bool near, int& x, int& y, int& z) {
    T qx = p.x + (T)0.5, qy = p.y + (T)0.5, qz = p.z + (T)0.5;
    if (near) { qx = clampc(qx, g.X); qy = clampc(qy, g.Y); qz = clampc(qz, g.Z); }
    x = clampi((int)qx, g.X);
    y = clampi((int)qy, g.Y);
    z = clampi((int)qz, g.Z);
    return (int64_t)x + (int64_t)g.X * ((int64_t)y + (int64_t)g.Y * (int64_t)z);
}

template <typename T>
__device__ __forceinline__ int64_t nearest(const VGrid& g, const V3<T>& p, bool near) {
    int x, y, z;
    return nearest_xyz(g, p, near, x, y, z);
}

__device__ __forceinline__ int dist_to_boundary(const VGrid& g, int x, int y, int z) {
    int m = min(x, g.X - 1 - x);
    m = min(m, min(y, g.Y - 1 - y));
    return min(m, min(z, g.Z - 1 - z));
}

// Adaptive Euler with step doubling (gi_adaptive_euler_advector.h:51-305) in
// precision T, as one lane's state advanced one iteration of the step loop at
// a time.  The step loop of the reference runs voxel by voxel: on entering a
// voxel it gathers the 8 corner gradients and restarts the step length; an
// iteration recomputes the gradient when the position moved, then takes or
// rejects one step.  Here a step that leaves the voxel enters the next one in
// the same iteration (its corner loads are issued beside the destination
// check of that step), so the lanes of a wave stay in step instead of each
// voxel's loop running until its slowest lane leaves.  Per lane, the
// arithmetic and the iteration count are the reference's: leaving a voxel
// costs one iteration without a step.
constexpr double kMaxStep = 0.5;

template <typename T>
struct Tracer {
    V3<T> cur, old, gr;
    V3<T> c[8];
    T gm, step;
    int bx, by, bz;
    bool near;
    int iters;
};

// Corner gradients of the voxel of t.cur; (cx, cy, cz) is the vertex nearest
// to the position on entry, which decides whether the boundary clamps apply.
template <typename T>
__device__ __forceinline__ void tracer_load(const GradSrc<T>& src, const VGrid& g, Tracer<T>& t, int cx, int cy,
                                            int cz) {
    t.near = dist_to_boundary(g, cx, cy, cz) <= 1;
    V3<T> b0 = t.cur;
    if (t.near) { b0.x = clampc(b0.x, g.X); b0.y = clampc(b0.y, g.Y); b0.z = clampc(b0.z, g.Z); }
    t.bx = (int)b0.x;
    t.by = (int)b0.y;
    t.bz = (int)b0.z;
    int xs[2] = {t.bx, t.bx + 1}, ys[2] = {t.by, t.by + 1}, zs[2] = {t.bz, t.bz + 1};
    if (t.near) {
        if (t.bx >= g.X - 1) xs[0] = xs[1] = g.X - 1;
        if (t.by >= g.Y - 1) ys[0] = ys[1] = g.Y - 1;
        if (t.bz >= g.Z - 1) zs[0] = zs[1] = g.Z - 1;
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        xs[k] = clampi(xs[k], g.X);
        ys[k] = clampi(ys[k], g.Y);
        zs[k] = clampi(zs[k], g.Z);
    }
#pragma unroll
    for (int k = 0; k < 8; k++)
        t.c[k] = src.at((int64_t)xs[k & 1] + (int64_t)g.X * ((int64_t)ys[(k >> 1) & 1] + (int64_t)g.Y * zs[k >> 2]));
}

// Voxel entry: gradient at the position and a fresh step length.
template <typename T>
__device__ __forceinline__ void tracer_begin(Tracer<T>& t) {
    t.old = t.cur;
    t.gr = trilin(t.c, t.cur.x - (T)t.bx, t.cur.y - (T)t.by, t.cur.z - (T)t.bz);
    t.gm = sqrt(t.gr.x * t.gr.x + t.gr.y * t.gr.y + t.gr.z * t.gr.z);
    t.step = (T)kMaxStep / t.gm;
}

// One iteration.  kNone: the streamline goes on (moved: it entered the voxel
// of vertex `next`); kHitPreassigned (next: the vertex reached), kLowGradient
// and kOverMax end it; kOutOfVoxel asks the caller to enter the voxel of the
// position (only a position changed outside a step could lead there).
template <typename T>
__device__ __forceinline__ int tracer_iterate(const GradSrc<T>& src, const VGrid& g,
                                              const uint8_t* __restrict__ dest, bool stop_on_assigned, T dir,
                                              T gthr, T ethr, Tracer<T>& t, int64_t& next, bool& moved) {
    moved = false;
    if (!(t.iters-- > 0)) return kOverMax;
    if (!(t.old.x == t.cur.x && t.old.y == t.cur.y && t.old.z == t.cur.z)) {
        if ((int)t.cur.x != (int)t.old.x || (int)t.cur.y != (int)t.old.y || (int)t.cur.z != (int)t.old.z)
            return kOutOfVoxel;
        t.gr = trilin(t.c, t.cur.x - (T)t.bx, t.cur.y - (T)t.by, t.cur.z - (T)t.bz);
        t.gm = sqrt(t.gr.x * t.gr.x + t.gr.y * t.gr.y + t.gr.z * t.gr.z);
        if (t.gm * t.step > (T)kMaxStep) t.step = (T)kMaxStep / t.gm;
        t.old = t.cur;
    }
    if (t.gm < gthr) return kLowGradient;
    // dir = -1: ascending manifolds, flow downhill; +1: descending, uphill
    const T s = dir * t.step, hs = (dir * (T)0.5) * t.step;
    const V3<T> full{t.gr.x * s, t.gr.y * s, t.gr.z * s};
    const V3<T> half{t.gr.x * hs, t.gr.y * hs, t.gr.z * hs};
    const V3<T> hg = trilin(t.c, t.cur.x + half.x - (T)t.bx, t.cur.y + half.y - (T)t.by, t.cur.z + half.z - (T)t.bz);
    const V3<T> two{hg.x * hs + half.x, hg.y * hs + half.y, hg.z * hs + half.z};
    const T dx = two.x - full.x, dy = two.y - full.y, dz = two.z - full.z;
    const T err = dx * dx + dy * dy + dz * dz;
    if (!(err < ethr)) {
        t.step = t.step * (T)0.5;
        return kNone;
    }
    t.cur = {t.cur.x + two.x, t.cur.y + two.y, t.cur.z + two.z};
    if (t.near) { t.cur.x = clampc(t.cur.x, g.X); t.cur.y = clampc(t.cur.y, g.Y); t.cur.z = clampc(t.cur.z, g.Z); }
    const T grow = t.step * (T)1.8, cap = (T)kMaxStep / t.gm;
    t.step = grow < cap ? grow : cap;
    int nx, ny, nz;
    next = nearest_xyz(g, t.cur, t.near, nx, ny, nz);
    const uint8_t d = dest[next];
    // the next iteration would leave the voxel: gather the next voxel's
    // corners while the destination byte is in flight
    const bool leave = t.iters > 0 &&
                       ((int)t.cur.x != (int)t.old.x || (int)t.cur.y != (int)t.old.y || (int)t.cur.z != (int)t.old.z);
    if (leave) tracer_load(src, g, t, nx, ny, nz);
    if (d == kCertainTerminal || (stop_on_assigned && d == kAssigned)) return kHitPreassigned;
    if (leave) {
        t.iters--;  // the iteration that leaves the voxel
        tracer_begin(t);
        moved = true;
    }
    return kNone;
}

// Enter the voxel of the position after kOutOfVoxel; returns the vertex nearest to it.
template <typename T>
__device__ __forceinline__ int64_t tracer_reenter(const GradSrc<T>& src, const VGrid& g, Tracer<T>& t) {
    int x, y, z;
    const int64_t next = nearest_xyz(g, t.cur, t.near, x, y, z);
    tracer_load(src, g, t, x, y, z);
    tracer_begin(t);
    return next;
}

// Start a streamline at vertex (x, y, z).
template <typename T>
__device__ __forceinline__ void tracer_start(const GradSrc<T>& src, const VGrid& g, Tracer<T>& t, int x, int y, int z,
                                             int max_iters) {
    t.cur = {(T)x, (T)y, (T)z};
    t.iters = max_iters;
    tracer_load(src, g, t, x, y, z);
    tracer_begin(t);
}

}  // namespace
}  // namespace msceer

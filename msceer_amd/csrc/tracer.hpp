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

// The 8 corner gradients of a tracer's voxel: in registers, or in LDS
// (component-major, kStride lanes apart) where the registers are needed for
// another tracer state.
template <typename T>
struct RegCorners {
    V3<T> c[8];
    __device__ __forceinline__ V3<T> get(int k) const { return c[k]; }
    __device__ __forceinline__ void set(int k, const V3<T>& v) { c[k] = v; }
};
template <typename T, int kStride>
struct LdsCorners {
    T* p;  // this lane's slot of component 0 of corner 0
    __device__ __forceinline__ V3<T> get(int k) const {
        return {p[(3 * k) * kStride], p[(3 * k + 1) * kStride], p[(3 * k + 2) * kStride]};
    }
    __device__ __forceinline__ void set(int k, const V3<T>& v) {
        p[(3 * k) * kStride] = v.x;
        p[(3 * k + 1) * kStride] = v.y;
        p[(3 * k + 2) * kStride] = v.z;
    }
};

template <typename T, class C>
__device__ __forceinline__ V3<T> trilin(const C& c, T fx, T fy, T fz) {
    const V3<T> x0 = lerp3(c.get(0), c.get(1), fx), x1 = lerp3(c.get(2), c.get(3), fx);
    const V3<T> x2 = lerp3(c.get(4), c.get(5), fx), x3 = lerp3(c.get(6), c.get(7), fx);
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

template <typename T, class C = RegCorners<T>>
struct Tracer {
    V3<T> cur, old, gr;
    C c;
    T gm, step;
    int bx, by, bz;
    bool near;
    int iters;
};

// Corner gradients of the voxel of t.cur; (cx, cy, cz) is the vertex nearest
// to the position on entry, which decides whether the boundary clamps apply.
template <typename T, class C>
__device__ __forceinline__ void tracer_load(const GradSrc<T>& src, const VGrid& g, Tracer<T, C>& t, int cx, int cy,
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
        t.c.set(k, src.at((int64_t)xs[k & 1] + (int64_t)g.X * ((int64_t)ys[(k >> 1) & 1] + (int64_t)g.Y * zs[k >> 2])));
}

// Voxel entry: gradient at the position and a fresh step length.
template <typename T, class C>
__device__ __forceinline__ void tracer_begin(Tracer<T, C>& t) {
    t.old = t.cur;
    t.gr = trilin(t.c, t.cur.x - (T)t.bx, t.cur.y - (T)t.by, t.cur.z - (T)t.bz);
    t.gm = sqrt(t.gr.x * t.gr.x + t.gr.y * t.gr.y + t.gr.z * t.gr.z);
    t.step = (T)kMaxStep / t.gm;
}

// One iteration.  kNone: the streamline goes on (moved: it entered the voxel
// of vertex `next`); kHitPreassigned (next: the vertex reached), kLowGradient
// and kOverMax end it; kOutOfVoxel asks the caller to enter the voxel of the
// position (only a position changed outside a step could lead there).
template <typename T, class C>
__device__ __forceinline__ int tracer_iterate(const GradSrc<T>& src, const VGrid& g,
                                              const uint8_t* __restrict__ dest, bool stop_on_assigned, T dir,
                                              T gthr, T ethr, Tracer<T, C>& t, int64_t& next, bool& moved) {
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
template <typename T, class C>
__device__ __forceinline__ int64_t tracer_reenter(const GradSrc<T>& src, const VGrid& g, Tracer<T, C>& t) {
    int x, y, z;
    const int64_t next = nearest_xyz(g, t.cur, t.near, x, y, z);
    tracer_load(src, g, t, x, y, z);
    tracer_begin(t);
    return next;
}

// Start a streamline at vertex (x, y, z).
template <typename T, class C>
__device__ __forceinline__ void tracer_start(const GradSrc<T>& src, const VGrid& g, Tracer<T, C>& t, int x, int y,
                                             int z, int max_iters) {
    t.cur = {(T)x, (T)y, (T)z};
    t.iters = max_iters;
    tracer_load(src, g, t, x, y, z);
    tracer_begin(t);
}

// ---- FP32 tracer in voxel-local coordinates, with decision margins ---------
//
// The position is kept as the corner (bx, by, bz) of its voxel plus an
// offset.  One step moves it by about half a voxel at most, so the offset
// stays within (-1, 2) and rounds to about 2^-24 voxel however large the
// volume (a global FP32 coordinate near 256 rounds to 2^-16).  An FP64 trace
// of the same streamline in global coordinates makes the same discrete
// decisions unless one of them lies within a margin of its threshold.  Every
// such near-tie, and a trace that uses up its step budget, sets `doubt`; the
// caller then traces the streamline again in FP64.  The near-ties:
//   a coordinate moved by a step to within kTieAbs of a voxel face (which
//   voxel holds it) or of a half-integer (which vertex is nearest) -- one
//   that a step leaves in place (a zero gradient component, as on a boundary
//   face) or clamps onto the boundary is exact in both precisions;
//   the step error within kTieRel of its threshold (step taken or halved);
//   the gradient magnitude within kTieRel of its threshold (streamline ends).
// Step-length caps are minima of continuous quantities, not decisions.  Per
// step the FP32 and FP64 displacements differ by ~2^-23 voxel (the FP32
// gradient field and its interpolation round at ~2^-22 relative; the offset
// at 2^-25), so over ~100 steps the positions drift apart by ~2^-20 voxel
// (random rounding) and no more than ~2^-16 (all in one direction); the
// step error is a difference of interpolated gradients ~1/8 of their size,
// which an FP32 evaluation gets to ~2^-19 relative.
constexpr float kTieAbs = 1.0f / 65536.0f;  // 2^-16 voxel
constexpr float kTieRel = 1.0f / 16384.0f;  // 2^-14

template <class C = RegCorners<float>>
struct TracerL {
    V3<float> cur, old, gr;  // cur, old: offsets from the voxel corner (bx, by, bz)
    C c;
    float gm, step;
    int bx, by, bz;
    bool near, doubt;
    int iters;
};

// Within kTieAbs of a multiple of one half?
__device__ __forceinline__ bool half_tie(float l) {
    const float t = l + l;
    return fabsf(t - rintf(t)) < 2.0f * kTieAbs;
}

// Nearest vertex of the position (global p + 0.5 truncated and clamped, as
// nearest_xyz).
template <class C>
__device__ __forceinline__ int64_t nearest_l(const VGrid& g, const TracerL<C>& t, int& x, int& y, int& z) {
    x = clampi(t.bx + (int)floorf(t.cur.x + 0.5f), g.X);
    y = clampi(t.by + (int)floorf(t.cur.y + 0.5f), g.Y);
    z = clampi(t.bz + (int)floorf(t.cur.z + 0.5f), g.Z);
    return (int64_t)x + (int64_t)g.X * ((int64_t)y + (int64_t)g.Y * (int64_t)z);
}
template <class C>
__device__ __forceinline__ int64_t tracer_nearest(const VGrid& g, const TracerL<C>& t) {
    int x, y, z;
    return nearest_l(g, t, x, y, z);
}
template <typename T, class C>
__device__ __forceinline__ int64_t tracer_nearest(const VGrid& g, const Tracer<T, C>& t) {
    return nearest(g, t.cur, t.near);
}

// Clamp of the global coordinate b + l to [0, n - 1], in offsets.
__device__ __forceinline__ float clamp_off(float l, int b, int n) {
    const float lo = (float)(-b), hi = (float)(n - 1 - b);
    return l < lo ? lo : (l > hi ? hi : l);
}

// Enter the voxel of the position: new corner (clamped as tracer_load
// clamps), offsets re-based exactly, corner gradients gathered.
template <class C>
__device__ __forceinline__ void tracer_load(const GradSrc<float>& src, const VGrid& g, TracerL<C>& t, int cx, int cy,
                                            int cz) {
    t.near = dist_to_boundary(g, cx, cy, cz) <= 1;
    int nbx = t.bx + (int)floorf(t.cur.x), nby = t.by + (int)floorf(t.cur.y), nbz = t.bz + (int)floorf(t.cur.z);
    if (t.near) {
        nbx = nbx < 0 ? 0 : (nbx > g.X - 1 ? g.X - 1 : nbx);
        nby = nby < 0 ? 0 : (nby > g.Y - 1 ? g.Y - 1 : nby);
        nbz = nbz < 0 ? 0 : (nbz > g.Z - 1 ? g.Z - 1 : nbz);
    }
    t.cur.x -= (float)(nbx - t.bx);
    t.cur.y -= (float)(nby - t.by);
    t.cur.z -= (float)(nbz - t.bz);
    t.bx = nbx;
    t.by = nby;
    t.bz = nbz;
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
        t.c.set(k, src.at((int64_t)xs[k & 1] + (int64_t)g.X * ((int64_t)ys[(k >> 1) & 1] + (int64_t)g.Y * zs[k >> 2])));
}

template <class C>
__device__ __forceinline__ void tracer_begin(TracerL<C>& t) {
    t.old = t.cur;
    t.gr = trilin(t.c, t.cur.x, t.cur.y, t.cur.z);
    t.gm = sqrtf(t.gr.x * t.gr.x + t.gr.y * t.gr.y + t.gr.z * t.gr.z);
    t.step = (float)kMaxStep / t.gm;
}

// One iteration, as tracer_iterate<float> in voxel-local coordinates.
template <class C>
__device__ __forceinline__ int tracer_iterate(const GradSrc<float>& src, const VGrid& g,
                                              const uint8_t* __restrict__ dest, bool stop_on_assigned, float dir,
                                              float gthr, float ethr, TracerL<C>& t, int64_t& next, bool& moved) {
    moved = false;
    if (!(t.iters-- > 0)) {
        t.doubt = true;
        return kOverMax;
    }
    if (!(t.old.x == t.cur.x && t.old.y == t.cur.y && t.old.z == t.cur.z)) {
        if (floorf(t.cur.x) != floorf(t.old.x) || floorf(t.cur.y) != floorf(t.old.y) ||
            floorf(t.cur.z) != floorf(t.old.z))
            return kOutOfVoxel;
        t.gr = trilin(t.c, t.cur.x, t.cur.y, t.cur.z);
        t.gm = sqrtf(t.gr.x * t.gr.x + t.gr.y * t.gr.y + t.gr.z * t.gr.z);
        if (t.gm * t.step > (float)kMaxStep) t.step = (float)kMaxStep / t.gm;
        t.old = t.cur;
    }
    if (fabsf(t.gm - gthr) < kTieRel * gthr) t.doubt = true;
    if (t.gm < gthr) return kLowGradient;
    const float s = dir * t.step, hs = (dir * 0.5f) * t.step;
    const V3<float> full{t.gr.x * s, t.gr.y * s, t.gr.z * s};
    const V3<float> half{t.gr.x * hs, t.gr.y * hs, t.gr.z * hs};
    const V3<float> hg = trilin(t.c, t.cur.x + half.x, t.cur.y + half.y, t.cur.z + half.z);
    const V3<float> two{hg.x * hs + half.x, hg.y * hs + half.y, hg.z * hs + half.z};
    const float dx = two.x - full.x, dy = two.y - full.y, dz = two.z - full.z;
    const float err = dx * dx + dy * dy + dz * dz;
    if (fabsf(err - ethr) < kTieRel * ethr) t.doubt = true;
    if (!(err < ethr)) {
        t.step = t.step * 0.5f;
        return kNone;
    }
    V3<float> nc{t.cur.x + two.x, t.cur.y + two.y, t.cur.z + two.z};
    if (t.near) {
        nc.x = clamp_off(nc.x, t.bx, g.X);
        nc.y = clamp_off(nc.y, t.by, g.Y);
        nc.z = clamp_off(nc.z, t.bz, g.Z);
    }
    // a coordinate that did not move (a gradient component exactly zero, as
    // on a boundary face) or was clamped onto the boundary is exact in both
    // precisions: only coordinates this step changed freely can tie
    auto tie = [](float step, float after, float unclamped) {
        return step != 0.f && after == unclamped && half_tie(after);
    };
    if (tie(two.x, nc.x, t.cur.x + two.x) || tie(two.y, nc.y, t.cur.y + two.y) || tie(two.z, nc.z, t.cur.z + two.z))
        t.doubt = true;
    t.cur = nc;
    const float grow = t.step * 1.8f, cap = (float)kMaxStep / t.gm;
    t.step = grow < cap ? grow : cap;
    int nx, ny, nz;
    next = nearest_l(g, t, nx, ny, nz);
    const uint8_t d = dest[next];
    const bool leave = t.iters > 0 && (floorf(t.cur.x) != floorf(t.old.x) || floorf(t.cur.y) != floorf(t.old.y) ||
                                       floorf(t.cur.z) != floorf(t.old.z));
    if (leave) tracer_load(src, g, t, nx, ny, nz);
    if (d == kCertainTerminal || (stop_on_assigned && d == kAssigned)) return kHitPreassigned;
    if (leave) {
        t.iters--;  // the iteration that leaves the voxel
        tracer_begin(t);
        moved = true;
    }
    return kNone;
}

template <class C>
__device__ __forceinline__ int64_t tracer_reenter(const GradSrc<float>& src, const VGrid& g, TracerL<C>& t) {
    int x, y, z;
    const int64_t next = nearest_l(g, t, x, y, z);
    tracer_load(src, g, t, x, y, z);
    tracer_begin(t);
    return next;
}

template <class C>
__device__ __forceinline__ void tracer_start(const GradSrc<float>& src, const VGrid& g, TracerL<C>& t, int x, int y,
                                             int z, int max_iters) {
    t.bx = x;
    t.by = y;
    t.bz = z;
    t.cur = {0.f, 0.f, 0.f};
    t.iters = max_iters;
    t.doubt = false;
    tracer_load(src, g, t, x, y, z);
    tracer_begin(t);
}

}  // namespace
}  // namespace msceer

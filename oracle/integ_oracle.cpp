// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference's accurate ascending-manifold labelling
// (the needASC2 branch of ondemandaccurate/ondemandaccurate.cxx:728-1061).
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
// this library; the product path never calls it.
//
// Two arithmetic modes of the numeric integrator:
//   mode 0 "reference": the reference's mixed precision, operation for
//          operation (FP32 step size, full-step displacement, magnitudes and
//          error; FP64 positions, gradients and half steps);
//   mode 1 "fp64": the same algorithm with every quantity in FP64 (the
//          FP64 oracle the GPU labels are measured against);
//   mode 2 "exact fp64": every vertex's own destination in FP64 -- each
//          non-certain vertex integrated alone until it reaches a certain
//          terminal, as pass 3 re-integrates (no path compression, which is
//          the reference's approximation of exactly this), so the labels do
//          not depend on any visiting order.
// Run single-threaded it reproduces the reference run with one OpenMP
// thread, whose first integration pass depends on the visiting order.
//
// Labels are canonical: a region is named by the vertex id of its lowest
// vertex (its representative minimum), -1 = unassigned.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <functional>
#include <type_traits>
#include <map>
#include <queue>
#include <set>
#include <unordered_map>
#include <vector>

extern "C" int oracle_steepest_pairs_only(const float* f, long long X, long long Y, long long Z, unsigned char* grad);
extern "C" int oracle_lstar_pairs(const float* f, long long X, long long Y, long long Z, const signed char* restriction,
                                  unsigned char* grad);
extern "C" int oracle_maxmin(const float* f, long long X, long long Y, long long Z, unsigned char* maxv,
                             unsigned char* minv);
extern "C" int oracle_lstar_vertices(const float* f, long long X, long long Y, long long Z,
                                     const signed char* restriction, const long long* vcells, long long n,
                                     unsigned char* grad);
extern "C" int oracle_ldir(unsigned char* grad, long long X, long long Y, long long Z);
extern "C" int oracle_flags(unsigned char* grad, long long X, long long Y, long long Z);

namespace {

typedef long long idx_t;

struct V3d { double v[3]; };
struct V3l { idx_t v[3]; };

// DestType (gi_basic_types.h:43-49)
enum Dest : unsigned char { BACKGROUND, UNASSIGNED, ASSIGNED, CERTAIN_TERMINAL, CERTAIN_NONTERMINAL };
// ADVECTION_EVENT (gi_advection_events.h:5-13)
enum Event { NONE, OUT_OF_VOXEL, OVER_MAX_ITERATIONS, LOW_GRADIENT, HIT_EXTREMUM, HIT_PREASSIGNED, OTHER };

struct Grid {
    idx_t X, Y, Z;
    idx_t n[3];
    Grid(idx_t x, idx_t y, idx_t z) : X(x), Y(y), Z(z) { n[0] = x; n[1] = y; n[2] = z; }
    idx_t index(const V3l& p) const { return p.v[0] + X * (p.v[1] + Y * p.v[2]); }
    V3l coords(idx_t id) const { return V3l{{id % X, (id / X) % Y, id / (X * Y)}}; }
    // gi_regular_grid.h:319-333
    idx_t dist_to_boundary(const V3l& a) const {
        idx_t m = a.v[0];
        for (int i = 0; i < 3; i++) {
            m = std::min(m, a.v[i]);
            m = std::min(m, (n[i] - 1) - a.v[i]);
        }
        return m;
    }
    // Inbounds(Vec3d), gi_regular_grid.h:102-120 (non-periodic)
    V3d inbounds(V3d p) const {
        for (int i = 0; i < 3; i++) {
            if (p.v[i] < 0) p.v[i] = 0;
            if (p.v[i] >= (double)n[i]) p.v[i] = (double)n[i] - 1;
        }
        return p;
    }
    // All6 neighbours in +x,-x,+y,-y,+z,-z order (gi_regular_grid.h:179-210)
    int all6(idx_t id, idx_t out[6]) const {
        V3l c = coords(id);
        int k = 0;
        const idx_t st[3] = {1, X, X * Y};
        for (int a = 0; a < 3; a++) {
            if (c.v[a] < n[a] - 1) out[k++] = id + st[a];
            if (c.v[a] > 0) out[k++] = id - st[a];
        }
        return k;
    }
};

inline V3l trunc_l(const V3d& p) { return V3l{{(idx_t)p.v[0], (idx_t)p.v[1], (idx_t)p.v[2]}}; }

// gi_regular_grid_trilinear_function.h:309-314
inline bool is_greater(const float* f, idx_t a, idx_t b) {
    if (f[a] > f[b]) return true;
    if (f[b] > f[a]) return false;
    return a > b;
}
inline bool before(const float* f, idx_t a, idx_t b) { return is_greater(f, b, a); }

// Central differences, zero where the vertex lies on the boundary of that
// axis (GradientFromImage with rklevel 1, gi_regular_grid_trilinear_function.h:285-341;
// coefficients src/gi_regular_grid_trilinear_function.cxx:5-11).
void gradient_field(const Grid& g, const float* f, std::vector<V3d>& out) {
    out.resize((size_t)(g.X * g.Y * g.Z));
    const idx_t st[3] = {1, g.X, g.X * g.Y};
    const float c[3] = {-1.0f / 2.0f, 0.0f, 1.0f / 2.0f};
    for (idx_t z = 0; z < g.Z; z++)
        for (idx_t y = 0; y < g.Y; y++)
            for (idx_t x = 0; x < g.X; x++) {
                const idx_t id = x + g.X * (y + g.Y * z);
                const idx_t p[3] = {x, y, z};
                V3d r;
                for (int a = 0; a < 3; a++) {
                    double res = 0.0;
                    const idx_t d = std::min(p[a], g.n[a] - 1 - p[a]);
                    if (d >= 1) {
                        res += c[0] * f[id - st[a]];
                        res += c[1] * f[id];
                        res += c[2] * f[id + st[a]];
                    } else {
                        res += 0.0f * f[id];
                    }
                    r.v[a] = res;
                }
                out[(size_t)id] = r;
            }
}

// 8 corner gradients of the voxel at base (GatherSurrounding[NoBoundaryCheck],
// gi_regular_grid.h:335-420): near the boundary, a corner index at or past the
// last vertex collapses onto the last vertex.
void gather8(const Grid& g, const std::vector<V3d>& grad, const V3l& b, bool check, V3d vals[8]) {
    idx_t xs[3][2];
    for (int a = 0; a < 3; a++) {
        xs[a][0] = b.v[a];
        xs[a][1] = b.v[a] + 1;
        if (check) {
            if (b.v[a] < 0) xs[a][0] = 0;
            if (b.v[a] >= g.n[a] - 1) { xs[a][0] = xs[a][1] = g.n[a] - 1; }
        }
    }
    int k = 0;
    for (int kz = 0; kz < 2; kz++)
        for (int jy = 0; jy < 2; jy++)
            for (int ix = 0; ix < 2; ix++)
                vals[k++] = grad[(size_t)(xs[0][ix] + g.X * (xs[1][jy] + g.Y * xs[2][kz]))];
}

inline V3d lerp(const V3d& a, const V3d& b, double t) {
    V3d r;
    for (int i = 0; i < 3; i++) r.v[i] = a.v[i] * (1.0 - t) + b.v[i] * t;
    return r;
}

// TriLinInterpGrad(s, base, vals) (gi_regular_grid_trilinear_function.h:211-229)
V3d trilin(const V3d& s, const V3l& base, const V3d vals[8]) {
    double fx = s.v[0] - (double)base.v[0], fy = s.v[1] - (double)base.v[1], fz = s.v[2] - (double)base.v[2];
    V3d x0 = lerp(vals[0], vals[1], fx), x1 = lerp(vals[2], vals[3], fx);
    V3d x2 = lerp(vals[4], vals[5], fx), x3 = lerp(vals[6], vals[7], fx);
    V3d y0 = lerp(x0, x1, fy), y1 = lerp(x2, x3, fy);
    return lerp(y0, y1, fz);
}

struct Checker {
    const Dest* dest;
    bool stop_on_assigned;  // pass 1: ASSIGNED or CERTAIN_TERMINAL; pass 3: CERTAIN_TERMINAL only
    const Grid* g;
    // TerminateNearPathCompressedRegion / TerminateNearOriginalCertain
    // (gi_advection_checker.h:212-262)
    Event check(const V3d& p, bool near_boundary) const {
        V3d q{{p.v[0] + 0.5, p.v[1] + 0.5, p.v[2] + 0.5}};
        if (near_boundary) q = g->inbounds(q);
        const idx_t id = g->index(trunc_l(q));
        const Dest d = dest[id];
        if (d == CERTAIN_TERMINAL || (stop_on_assigned && d == ASSIGNED)) return HIT_PREASSIGNED;
        return NONE;
    }
};

inline double mag(const V3d& a) { return std::sqrt(a.v[0] * a.v[0] + a.v[1] * a.v[1] + a.v[2] * a.v[2]); }

// AdaptiveEulerAdvector3D<F,-1>::AdvectThroughVoxel{NoCheck,NearBoundary}
// (gi_adaptive_euler_advector.h:51-305).  kFp64 selects all-FP64 arithmetic.
template <bool kFp64, class Chk>
Event advect_voxel(const Grid& g, const std::vector<V3d>& grad, const Chk& chk, float gthr, float ethr,
                   bool near_boundary, V3d& cur, int& iters, int kMul) {
    typedef typename std::conditional<kFp64, double, float>::type F;
    const double kMaxStep = 0.5;  // MAX_STEP_IN_GRID_CELLS
    // kMul: -1 for ascending manifolds (flow downhill), +1 for descending
    // (and for the ascending 1-manifolds traced up from 2-saddles)
    V3d vals[8];
    V3d old = cur;
    const V3l base = near_boundary ? trunc_l(g.inbounds(cur)) : trunc_l(cur);
    gather8(g, grad, base, near_boundary, vals);
    V3d gr = trilin(cur, base, vals);
    F gm = (F)mag(gr);
    // MAX_STEP_IN_GRID_CELLS / gradient_magnitude: a double literal over a
    // float, rounded to the float step size in the reference arithmetic
    F step = (F)(kMaxStep / (double)gm);
    while (iters-- > 0) {
        if (!(old.v[0] == cur.v[0] && old.v[1] == cur.v[1] && old.v[2] == cur.v[2])) {
            if ((int)cur.v[0] != (int)old.v[0] || (int)cur.v[1] != (int)old.v[1] || (int)cur.v[2] != (int)old.v[2])
                return OUT_OF_VOXEL;
            gr = trilin(cur, base, vals);
            gm = (F)mag(gr);
            if (kFp64) {
                if (gm * step > kMaxStep) step = kMaxStep / gm;
            } else {
                if ((double)(float)(gm * step) > kMaxStep) step = (F)(kMaxStep / (double)gm);
            }
            old = cur;
        }
        if (gm < (F)gthr) return LOW_GRADIENT;
        V3d full, half, two;
        if (kFp64) {
            const double s = (double)kMul * step;
            for (int i = 0; i < 3; i++) full.v[i] = gr.v[i] * s;
        } else {
            // Vec3d * float multiplies in FP32 (gi_vectors.h:250-254)
            const float s = (float)kMul * (float)step;
            for (int i = 0; i < 3; i++) full.v[i] = (double)((float)gr.v[i] * s);
        }
        const double hs = (double)step * ((double)kMul * 0.5);
        for (int i = 0; i < 3; i++) half.v[i] = gr.v[i] * hs;
        V3d hp{{cur.v[0] + half.v[0], cur.v[1] + half.v[1], cur.v[2] + half.v[2]}};
        V3d hg = trilin(hp, base, vals);
        for (int i = 0; i < 3; i++) two.v[i] = hg.v[i] * hs + half.v[i];
        V3d d{{two.v[0] - full.v[0], two.v[1] - full.v[1], two.v[2] - full.v[2]}};
        const F err = (F)(d.v[0] * d.v[0] + d.v[1] * d.v[1] + d.v[2] * d.v[2]);
        if (err < (F)ethr) {
            V3d np{{cur.v[0] + two.v[0], cur.v[1] + two.v[1], cur.v[2] + two.v[2]}};
            cur = near_boundary ? g.inbounds(np) : np;
            if (kFp64) {
                step = std::min(step * 1.8, kMaxStep / gm);
            } else {
                const float a = (float)((double)step * 1.8), b = (float)(kMaxStep / (double)gm);
                step = a < b ? a : b;
            }
            const Event e = chk.check(cur, near_boundary);
            if (e != NONE) return e;
        } else {
            step = (F)((double)step * 0.5);
        }
    }
    return OVER_MAX_ITERATIONS;
}

// Integrate one streamline from vertex `start`; returns the nearest vertex of
// the end point and the terminating event; visits the nearest vertex after
// every voxel into `path` when given.
template <bool kFp64>
idx_t trace(const Grid& g, const std::vector<V3d>& grad, const Checker& chk, float gthr, float ethr, int max_iters,
            idx_t start, std::vector<idx_t>* path, Event* ev, int mul = -1) {
    V3l c = g.coords(start);
    V3d cur{{(double)c.v[0], (double)c.v[1], (double)c.v[2]}};
    int iters = max_iters;
    while (true) {
        Event e;
        if (g.dist_to_boundary(c) <= 1) {
            e = advect_voxel<kFp64>(g, grad, chk, gthr, ethr, true, cur, iters, mul);
            c = trunc_l(g.inbounds(V3d{{cur.v[0] + 0.5, cur.v[1] + 0.5, cur.v[2] + 0.5}}));
        } else {
            e = advect_voxel<kFp64>(g, grad, chk, gthr, ethr, false, cur, iters, mul);
            c = trunc_l(V3d{{cur.v[0] + 0.5, cur.v[1] + 0.5, cur.v[2] + 0.5}});
        }
        const idx_t next = g.index(c);
        if (path) path->push_back(next);
        if (e == OUT_OF_VOXEL) continue;
        if (e == LOW_GRADIENT || e == HIT_EXTREMUM || e == HIT_PREASSIGNED || e == OVER_MAX_ITERATIONS) {
            *ev = e;
            return next;
        }
    }
}

// Every non-certain vertex integrated alone, in FP64, to a certain terminal
// (the stop rule of pass 3, TerminateNearOriginalCertain,
// gi_advection_checker.h:245-262); independent vertices, so OpenMP threads
// share them in any order.
void integrate_exact(const Grid& g, std::vector<V3d> grad, const std::vector<std::vector<idx_t>>& seeds, float ethr,
                     float gthr, int max_iters, int* label, long long* stats, int mul) {
    const idx_t n = g.X * g.Y * g.Z;
    std::vector<Dest> dest((size_t)n);
    for (idx_t i = 0; i < n; i++) dest[(size_t)i] = label[i] == -1 ? UNASSIGNED : CERTAIN_TERMINAL;
    for (const auto& s : seeds)
        for (idx_t id : s) grad[(size_t)id] = V3d{{0, 0, 0}};
    const Checker chk{dest.data(), false, &g};
    std::vector<int> out(label, label + n);
    long long traced = 0, over = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : traced, over)
    for (idx_t v = 0; v < n; v++) {
        if (dest[(size_t)v] != UNASSIGNED) continue;
        Event ev;
        const idx_t end = trace<true>(g, grad, chk, gthr, ethr, max_iters, v, nullptr, &ev, mul);
        traced++;
        if (ev == OVER_MAX_ITERATIONS) over++;
        // a terminal's label, or the label of a vertex the streamline stopped
        // at otherwise (low gradient, step budget): a certain one or -1
        out[(size_t)v] = label[end];
    }
    std::copy(out.begin(), out.end(), label);
    if (stats) { stats[0] = traced; stats[1] = over; stats[2] = traced; }
}

// NumericIntegratorPathCompressingToTerminal::BeginIntegration with one thread
// (gi_numeric_integrator_path_compressing.h:75-433).
template <bool kFp64>
void integrate(const Grid& g, const float* f, std::vector<V3d> grad, const std::vector<std::vector<idx_t>>& seeds,
               float ethr, float gthr, int max_iters, int* label, long long* stats, int mul = -1) {
    (void)f;
    const idx_t n = g.X * g.Y * g.Z;
    std::vector<Dest> dest((size_t)n);
    for (idx_t i = 0; i < n; i++) dest[(size_t)i] = label[i] == -1 ? UNASSIGNED : CERTAIN_TERMINAL;
    for (const auto& s : seeds)
        for (idx_t id : s) grad[(size_t)id] = V3d{{0, 0, 0}};
    Checker chk1{dest.data(), true, &g};
    std::vector<idx_t> path;
    long long traced = 0, steps_over = 0;
    // pass 1: coarse-to-fine visiting order of one thread's range [0, n)
    const idx_t num = n;
    for (idx_t kk = 1; kk <= num * 4; kk *= 2) {
        const idx_t startsize = num / kk, stepsize = (num * 2) / kk;
        if (stepsize == 0) continue;
        for (idx_t i = startsize; i < n; i += stepsize) {
            if (dest[(size_t)i] != UNASSIGNED) continue;
            path.clear();
            path.push_back(i);
            Event ev;
            const idx_t end = trace<kFp64>(g, grad, chk1, gthr, ethr, max_iters, i, &path, &ev, mul);
            traced++;
            if (ev == OVER_MAX_ITERATIONS) steps_over++;
            const int lab = label[end];
            for (idx_t jj : path)
                if (dest[(size_t)jj] == UNASSIGNED) { label[jj] = lab; dest[(size_t)jj] = ASSIGNED; }
        }
    }
    // pass 2: assigned vertices next to a different label
    std::set<idx_t> todo;
    for (idx_t i = 0; i < n; i++) {
        if (dest[(size_t)i] == BACKGROUND) continue;
        idx_t nb[6];
        const int k = g.all6(i, nb);
        for (int j = 0; j < k; j++) {
            const idx_t v2 = nb[j];
            if (v2 > i) continue;
            if (label[i] != label[v2]) {
                if (dest[(size_t)i] == ASSIGNED) todo.insert(i);
                if (dest[(size_t)v2] == ASSIGNED) todo.insert(v2);
            }
        }
    }
    // pass 3: re-integrate to original certain regions only
    Checker chk3{dest.data(), false, &g};
    long long fixed = 0;
    while (!todo.empty()) {
        std::vector<idx_t> work(todo.begin(), todo.end());
        todo.clear();
        fixed += (long long)work.size();
        for (idx_t v : work) {
            const int init = label[v];
            Event ev;
            const idx_t end = trace<kFp64>(g, grad, chk3, gthr, ethr, max_iters, v, nullptr, &ev, mul);
            traced++;
            if (ev == OVER_MAX_ITERATIONS) steps_over++;
            const int nl = label[end];
            label[v] = nl;
            dest[(size_t)v] = CERTAIN_NONTERMINAL;
            if (nl != init) {
                idx_t nb[6];
                const int k = g.all6(v, nb);
                for (int j = 0; j < k; j++)
                    if (dest[(size_t)nb[j]] == ASSIGNED && label[nb[j]] != nl) todo.insert(nb[j]);
            }
        }
    }
    if (stats) { stats[0] = traced; stats[1] = steps_over; stats[2] = fixed; }
}

// ---- extremum graph and certain regions -----------------------------------

struct CellMesh {
    idx_t X, Y, Z, M[3];
    CellMesh(idx_t x, idx_t y, idx_t z) : X(x), Y(y), Z(z) { M[0] = 2 * x - 1; M[1] = 2 * y - 1; M[2] = 2 * z - 1; }
    void coords(idx_t c, idx_t p[3]) const { p[0] = c % M[0]; p[1] = (c / M[0]) % M[1]; p[2] = c / (M[0] * M[1]); }
    int dim(idx_t c) const { idx_t p[3]; coords(c, p); return (int)((p[0] & 1) + (p[1] & 1) + (p[2] & 1)); }
    int bv(idx_t c) const {
        idx_t p[3]; coords(c, p);
        return (p[0] == 0 || p[0] == M[0] - 1) + (p[1] == 0 || p[1] == M[1] - 1) + (p[2] == 0 || p[2] == M[2] - 1);
    }
    idx_t vertex_number(idx_t c) const { idx_t p[3]; coords(c, p); return p[0] / 2 + X * (p[1] / 2 + Y * (p[2] / 2)); }
    idx_t off(int a) const { return a == 0 ? 1 : (a == 1 ? M[0] : M[0] * M[1]); }
    int facets(idx_t c, idx_t out[6]) const {
        idx_t p[3]; coords(c, p);
        int n = 0;
        for (int a = 0; a < 3; a++) if (p[a] & 1) { out[n++] = c + off(a); out[n++] = c - off(a); }
        return n;
    }
    idx_t pair(const unsigned char* g, idx_t c) const {
        const unsigned d = (g[c] >> 2) & 7;
        switch (d) {
            case 0: return c + 1;
            case 1: return c - 1;
            case 2: return c + M[0];
            case 3: return c - M[0];
            case 4: return c + M[0] * M[1];
            case 5: return c - M[0] * M[1];
        }
        return c;
    }
    bool critical(const unsigned char* g, idx_t c) const { return ((g[c] >> 2) & 7) == 7; }
};

// rec_td (gi_extrema_region_builder.h:286-301): follow the V-path down from a
// 1-cell or vertex to a critical vertex.
idx_t trace_down(const CellMesh& m, const unsigned char* g, idx_t c) {
    while (true) {
        if (m.critical(g, c)) return c;
        const idx_t p = m.pair(g, c);
        if (m.critical(g, p)) return p;
        idx_t fc[6];
        m.facets(p, fc);
        c = (fc[0] == c) ? fc[1] : fc[0];
    }
}

// rec_tu (gi_extrema_region_builder.h:330-352): follow the V-path up from a
// hex (its paired quad's other hex) to a critical hex; -1 where a hex is
// paired with a quad on the domain boundary (the reference reads past its one
// cofacet there).
idx_t trace_up(const CellMesh& m, const unsigned char* g, idx_t c) {
    while (true) {
        if (m.critical(g, c)) return c;
        const idx_t p = m.pair(g, c);
        if (m.critical(g, p)) return p;
        idx_t q[3];
        m.coords(p, q);
        int a = 0;
        while (a < 3 && (q[a] & 1)) a++;  // the quad's even axis: its cofacets lie along it
        const idx_t nxt = (c == p + m.off(a)) ? p - m.off(a) : p + m.off(a);
        const idx_t along = (c == p + m.off(a)) ? q[a] - 1 : q[a] + 1;
        if (along < 0 || along >= m.M[a]) return -1;
        c = nxt;
    }
}

// cellValue / lessThan (gi_topological_max_vertex_mesh_function.h:45-62):
// value of the highest vertex; ties by highest-vertex id then cell id.
struct CellFunc {
    const CellMesh* m;
    const float* f;
    const unsigned char* maxv;
    idx_t highest(idx_t c) const {
        const int t = maxv[c];
        const idx_t d = (t % 3 - 1) + m->M[0] * (((t / 3) % 3 - 1) + m->M[1] * (t / 9 - 1));
        return m->vertex_number(c + d);
    }
    float value(idx_t c) const { return f[highest(c)]; }
    bool less(idx_t a, idx_t b) const {
        const idx_t ha = highest(a), hb = highest(b);
        if (f[ha] < f[hb]) return true;
        if (f[hb] < f[ha]) return false;
        if (ha != hb) return ha < hb;
        return a < b;
    }
};

// UFMergeGraph (MINIMAL) + SimplifyToThreshold (gi_extrema_region_builder.h:28-268).
struct MinGraph {
    struct Ext { idx_t cell; int rep; float val; };
    struct Arc {
        idx_t saddle; int e1, e2; float pers, sval;
    };
    struct ArcOrder {  // priority_queue comparator of MergeArc: lowest persistence first, then lowest saddle id
        bool operator()(const Arc& l, const Arc& r) const {
            if (l.pers < r.pers) return false;
            if (l.pers > r.pers) return true;
            return l.saddle > r.saddle;
        }
    };
    const CellFunc* fn;
    bool min_hier = true;  // MINIMAL (true) or MAXIMAL: which extremum of a merge stays representative
    std::vector<Ext> ext;
    std::unordered_map<idx_t, int> id;
    std::priority_queue<Arc, std::vector<Arc>, ArcOrder> arcs;
    float thr = 0;
    void add_node(idx_t c) { int l = (int)ext.size(); ext.push_back(Ext{c, l, fn->value(c)}); id[c] = l; }
    int find(int l) {
        Ext& e = ext[l];
        if (e.rep == l) return l;
        int r = find(e.rep);
        ext[l].rep = r;
        return r;
    }
    void add_arc(idx_t a, idx_t b, idx_t s) {
        if (a == b) return;
        Arc m;
        m.saddle = s; m.e1 = id[a]; m.e2 = id[b];
        m.sval = fn->value(s);
        float d1 = ext[m.e1].val - m.sval, d2 = ext[m.e2].val - m.sval;
        if (d1 < 0) d1 *= -1;
        if (d2 < 0) d2 *= -1;
        m.pers = d1 < d2 ? d1 : d2;
        arcs.push(m);
    }
    void merge(int l1, int l2) {
        Ext& e1 = ext[l1];
        Ext& e2 = ext[l2];
        if (fn->less(e1.cell, e2.cell) ^ min_hier) e1.rep = e2.rep;
        else e2.rep = e1.rep;
    }
    void cancel(Arc m) {
        if (m.pers > thr) return;
        int r1 = find(m.e1), r2 = find(m.e2);
        if (r1 == r2) return;
        // note the crossed references of PerformCancel (:98-99)
        Ext& c1 = ext[r2];
        Ext& c2 = ext[r1];
        if (m.e1 != r1 || m.e2 != r2) {
            float d1 = std::fabs(m.sval - c1.val), d2 = std::fabs(m.sval - c2.val);
            if (d1 < d2) {
                if (d1 > thr) return;
                m.pers = d1; m.e1 = r2; m.e2 = r1;
                arcs.push(m);
            } else {
                if (d2 > thr) return;
                m.pers = d2; m.e1 = r1; m.e2 = r2;
                arcs.push(m);
            }
            return;
        }
        merge(r1, r2);
    }
    void simplify(float t) {
        thr = t;
        while (!arcs.empty()) { Arc a = arcs.top(); arcs.pop(); cancel(a); }
        for (int i = 0; i < (int)ext.size(); i++) find(i);
    }
};

// Certain regions: seeds are the vertices of the minima sharing one
// representative; a vertex joins a region when all of its earlier All6
// neighbours carry that region's label (GridSimplifiedExtremalRegionBuilder,
// gi_extrema_region_builder.h:806-1042).  The fixed point does not depend on
// the order regions are grown in.
// desc: the descending order of ComparerDSC (IndexCompareGreaterThan,
// gi_numeric_integrator_expanding_region_stop.h:41-59): a vertex comes
// before another when it is greater (value, then id); flooded from the top.
void certain_regions(const Grid& g, const float* f, const std::vector<std::vector<idx_t>>& seeds,
                     const std::vector<int>& seed_label, int* lab, bool desc = false) {
    const idx_t n = g.X * g.Y * g.Z;
    for (idx_t i = 0; i < n; i++) lab[i] = -1;
    auto before = [desc](const float* ff, idx_t a, idx_t b) { return desc ? is_greater(ff, a, b) : ::before(ff, a, b); };
    auto cmp = [f, desc](idx_t a, idx_t b) { return desc ? is_greater(f, b, a) : is_greater(f, a, b); };  // earliest on top
    for (size_t r = 0; r < seeds.size(); r++) {
        std::set<idx_t> enq;
        std::priority_queue<idx_t, std::vector<idx_t>, decltype(cmp)> pq(cmp);
        for (idx_t s : seeds[r]) { enq.insert(s); lab[s] = seed_label[r]; }
        auto enqueue_later = [&](idx_t v) {
            idx_t nb[6];
            const int k = g.all6(v, nb);
            for (int j = 0; j < k; j++)
                if (lab[nb[j]] == -1 && before(f, v, nb[j]) && !enq.count(nb[j])) { enq.insert(nb[j]); pq.push(nb[j]); }
        };
        for (idx_t s : seeds[r]) enqueue_later(s);
        while (!pq.empty()) {
            const idx_t v = pq.top();
            pq.pop();
            int val = lab[v];
            bool has = false;
            bool ok = true;
            idx_t nb[6];
            const int k = g.all6(v, nb);
            for (int j = 0; j < k && ok; j++) {
                if (before(f, nb[j], v)) {
                    if (lab[nb[j]] == -1) { ok = false; break; }
                    if (!has) { val = lab[nb[j]]; has = true; }
                    else if (val != lab[nb[j]]) ok = false;
                }
            }
            if (ok && val >= 0) { lab[v] = val; enqueue_later(v); }
        }
    }
}

// IsolatedCCRegionRemoverNEW::ComputeOutput (gi_isolated_region_remover.h:155-192).
void clean_regions(const Grid& g, const float* f, int* lab, bool desc = false) {
    auto before = [desc](const float* ff, idx_t a, idx_t b) { return desc ? is_greater(ff, a, b) : ::before(ff, a, b); };
    const idx_t n = g.X * g.Y * g.Z;
    std::vector<idx_t> dst((size_t)n);
    for (idx_t i = 0; i < n; i++) dst[(size_t)i] = i;
    std::function<idx_t(idx_t)> findc = [&](idx_t i) -> idx_t {
        idx_t r = i;
        while (dst[(size_t)r] != r) r = dst[(size_t)r];
        while (dst[(size_t)i] != r) { idx_t nx = dst[(size_t)i]; dst[(size_t)i] = r; i = nx; }
        return r;
    };
    for (idx_t i = 0; i < n; i++) {
        idx_t nb[6];
        const int k = g.all6(i, nb);
        for (int j = 0; j < k; j++) {
            if (lab[i] != lab[nb[j]]) continue;
            const idx_t a = findc(i), b = findc(nb[j]);
            if (before(f, a, b)) dst[(size_t)b] = a; else dst[(size_t)a] = b;
        }
    }
    std::vector<idx_t> roots;
    for (idx_t i = 0; i < n; i++) if (findc(i) == i) roots.push_back(i);
    auto region = [&](idx_t i) { return lab[findc(i)]; };
    for (idx_t id : roots) {
        idx_t nb[6];
        const int k = g.all6(id, nb);
        const int idreg = region(id);
        idx_t earliest = id;
        for (int j = 0; j < k; j++) {
            if (before(f, nb[j], id)) {
                if (idreg == region(nb[j])) continue;
                if (before(f, nb[j], earliest)) earliest = nb[j];
            }
        }
        if (k == 0) continue;
        dst[(size_t)id] = findc(earliest);
    }
    std::vector<int> out(lab, lab + n);
    for (idx_t i = 0; i < n; i++) if (lab[i] != -1) out[(size_t)i] = lab[findc(i)];
    std::copy(out.begin(), out.end(), lab);
}

// ComputeMINBoundary (gi_labeling_to_bounary_labeling.h:48-116).
void min_boundary(const CellMesh& m, const int* lab, signed char* r) {
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    for (idx_t c = 0; c < C; c++) {
        if (m.dim(c) != 1) continue;
        idx_t fc[6];
        m.facets(c, fc);
        const int a = lab[m.vertex_number(fc[0])], b = lab[m.vertex_number(fc[1])];
        if (a == -1 || b == -1) continue;
        if (a != b) r[c] = 1;
    }
    for (int d = 2; d <= 3; d++)
        for (idx_t c = 0; c < C; c++) {
            if (m.dim(c) != d) continue;
            idx_t fc[6];
            const int k = m.facets(c, fc);
            for (int j = 0; j < k; j++) if (r[fc[j]] == 1) { r[c] = 1; break; }
        }
}


// ---- accurate 1-manifolds: digitized streamlines from the saddles ---------
//
// DigitizingNumericStreamlineIntegrator3dASC / 3dDSC
// (gi_numeric_streamline_integrator.h:526-990) driven by
// ReIntegrateUpFrom2Saddles / ReIntegrateDownFrom1Saddles
// (ondemandaccurate.cxx:254-500), as one OpenMP thread runs them: saddles in
// decreasing (value, id) order (ASC) or increasing value, decreasing id
// (DSC: keys (-value, id) taken from the largest), each saddle's seed cells
// gathered by a short V-path search, one streamline per seed that labels the
// cells it crosses with the target value (4 ASC, 5 DSC) in the restriction
// array and stops near a cell already carrying that value; the saddle takes
// the target after its streamlines.
//
// ASC: cells are hexes (voxels) and the quads between them; the flow is
// uphill (advector sign +1).  DSC: the same on the dual grid (vertices and
// the edges between them, positions shifted by half a voxel); downhill.

struct LineChecker {
    // TerminateNearAssignedHex / TerminateNearAssignedVert
    // (gi_advection_checker.h:59-100)
    const signed char* lab;
    const CellMesh* m;
    int target;
    bool asc;
    Event check(const V3d& p, bool near_boundary) const {
        if (near_boundary) return HIT_PREASSIGNED;
        idx_t c[3];
        for (int a = 0; a < 3; a++)
            c[a] = asc ? 2 * (idx_t)std::floor(p.v[a]) + 1 : 2 * (idx_t)std::floor(p.v[a] + 0.5);
        return lab[c[0] + m->M[0] * (c[1] + m->M[1] * c[2])] == target ? HIT_PREASSIGNED : NONE;
    }
};

// DigitizeSegmentInternal (gi_numeric_streamline_integrator.h:566-700 ASC,
// 824-985 DSC): the cell faces the segment a -> b crosses, in the order of
// their crossing parameter; for each crossing whose face bounds the current
// cell, label the current cell and the face (stop at the first that already
// carries the target) and move to the cell on the other side.
void digitize(const CellMesh& m, V3d a, V3d b, signed char* lab, int target, bool asc, long long* nlab) {
    if (!asc)
        for (int d = 0; d < 3; d++) { a.v[d] += 0.5; b.v[d] += 0.5; }
    const V3d ab{{b.v[0] - a.v[0], b.v[1] - a.v[1], b.v[2] - a.v[2]}};
    idx_t face[3];
    double tc[3];
    int n = 0;
    for (int d = 0; d < 3; d++) {
        const int va = (int)std::floor(a.v[d]), vb = (int)std::floor(b.v[d]);
        if (va == vb) continue;
        const int plane = va > vb ? va : vb;
        const double t = ((double)plane - a.v[d]) / (b.v[d] - a.v[d]);
        tc[n] = t;
        if (t < 0 || t > 1) return;
        idx_t fc[3];
        for (int k = 0; k < 3; k++) fc[k] = 2 * (idx_t)std::floor(ab.v[k] * t + a.v[k]) + (k == d ? 0 : 1);
        if (!asc)
            for (int k = 0; k < 3; k++) fc[k] -= 1;
        face[n++] = fc[0] + m.M[0] * (fc[1] + m.M[1] * fc[2]);
    }
    if (n == 0) return;
    for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++)
            if (tc[i] > tc[j]) {
                std::swap(face[i], face[j]);
                std::swap(tc[i], tc[j]);
            }
    idx_t sc[3];
    for (int k = 0; k < 3; k++) sc[k] = 2 * (idx_t)std::floor(a.v[k]) + (asc ? 1 : 0);
    idx_t cur = sc[0] + m.M[0] * (sc[1] + m.M[1] * sc[2]);
    for (int i = 0; i < n; i++) {
        // the two cells the face separates: cofacets (hexes) of a quad for
        // ASC, facets (vertices) of an edge for DSC, along the face's
        // crossing axis, the domain-clipped ones only
        idx_t p[3];
        m.coords(face[i], p);
        bool has = false;
        idx_t next = cur;
        for (int d = 0; d < 3; d++) {
            if (asc ? (p[d] & 1) != 0 : (p[d] & 1) == 0) continue;
            if (p[d] < m.M[d] - 1) { const idx_t c = face[i] + m.off(d); if (c == cur) has = true; else next = c; }
            if (p[d] > 0) { const idx_t c = face[i] - m.off(d); if (c == cur) has = true; else next = c; }
        }
        if (!has) continue;
        if (lab[cur] == target) return;
        lab[cur] = (signed char)target;
        if (nlab) (*nlab)++;
        if (lab[face[i]] == target) return;
        lab[face[i]] = (signed char)target;
        if (nlab) (*nlab)++;
        cur = next;
    }
}

// IntegrateStreamline (gi_numeric_streamline_integrator.h:734-788, 1018-1072).
template <bool kFp64>
Event trace_line(const Grid& g, const CellMesh& m, const std::vector<V3d>& grad, const LineChecker& chk, float gthr,
                 float ethr, int max_iters, V3d seed, signed char* lab, long long* nvox, long long* nlab) {
    V3l c = trunc_l(V3d{{seed.v[0] + 0.5, seed.v[1] + 0.5, seed.v[2] + 0.5}});
    V3d cur = seed;
    int iters = max_iters;
    const int mul = chk.asc ? 1 : -1;
    while (true) {
        const V3d old = cur;
        Event e;
        if (g.dist_to_boundary(c) <= 1) {
            e = advect_voxel<kFp64>(g, grad, chk, gthr, ethr, true, cur, iters, mul);
            c = trunc_l(g.inbounds(V3d{{cur.v[0] + 0.5, cur.v[1] + 0.5, cur.v[2] + 0.5}}));
        } else {
            e = advect_voxel<kFp64>(g, grad, chk, gthr, ethr, false, cur, iters, mul);
            c = trunc_l(V3d{{cur.v[0] + 0.5, cur.v[1] + 0.5, cur.v[2] + 0.5}});
            digitize(m, old, cur, lab, chk.target, chk.asc, nlab);
        }
        if (nvox) (*nvox)++;
        if (e == LOW_GRADIENT || e == HIT_EXTREMUM || e == HIT_PREASSIGNED || e == OVER_MAX_ITERATIONS) return e;
    }
}

// ReIntegrateUpFrom2Saddles / ReIntegrateDownFrom1Saddles with one thread.
// stats (8): saddles, streamlines, voxel traversals, cells labelled,
//            stops on labelled cells or the boundary, low-gradient stops, budget stops,
//            microseconds of the saddle gathering, sorting and streamlines
template <bool kFp64>
void saddle_lines(const Grid& g, const CellMesh& m, const float* f, const unsigned char* grad,
                  const std::vector<V3d>& field, float ethr, float gthr, int max_iters, bool asc, signed char* lab,
                  long long* stats) {
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    std::vector<unsigned char> maxv((size_t)C), minv((size_t)C);
    oracle_maxmin(f, g.X, g.Y, g.Z, maxv.data(), minv.data());
    const CellFunc cf{&m, f, maxv.data()};
    const auto t0 = std::chrono::steady_clock::now();
    const int sdim = asc ? 2 : 1, target = asc ? 4 : 5;
    std::vector<std::pair<float, idx_t>> sad;
    for (idx_t c = 0; c < C; c++)
        if (m.dim(c) == sdim && m.critical(grad, c)) sad.push_back({asc ? cf.value(c) : -1 * cf.value(c), c});
    std::sort(sad.begin(), sad.end());
    const LineChecker chk{lab, &m, target, asc};
    long long st[8] = {(long long)sad.size(), 0, 0, 0, 0, 0, 0, 0};
    for (idx_t k = (idx_t)sad.size() - 1; k >= 0; k--) {
        const idx_t s = sad[(size_t)k].second;
        // a few V-path steps on either side of the saddle
        // (ondemandaccurate.cxx:319-351 ASC, 448-480 DSC)
        std::vector<idx_t> result, queue{s};
        std::set<idx_t> seen;
        size_t qh = 0;
        for (int counter = 4; qh < queue.size() && counter >= 0; counter--) {
            const idx_t cur = queue[qh++];
            seen.insert(cur);
            idx_t p[3];
            m.coords(cur, p);
            for (int d = 0; d < 3; d++) {
                if (asc ? (p[d] & 1) != 0 : (p[d] & 1) == 0) continue;  // cofacets (ASC) / facets (DSC) along d
                for (int sg = 0; sg < 2; sg++) {
                    if (asc && (sg == 0 ? p[d] >= m.M[d] - 1 : p[d] <= 0)) continue;
                    const idx_t t = cur + (sg == 0 ? m.off(d) : -m.off(d));
                    if (seen.count(t)) continue;
                    result.push_back(t);
                    seen.insert(t);
                    if (!m.critical(grad, t)) {
                        const idx_t pr = m.pair(grad, t);
                        seen.insert(pr);
                        queue.push_back(pr);
                    }
                }
            }
        }
        for (idx_t t : result) {
            if (m.dim(t) != (asc ? 3 : 0)) continue;
            idx_t p[3];
            m.coords(t, p);
            const V3d seed{{(double)p[0] * 0.5, (double)p[1] * 0.5, (double)p[2] * 0.5}};
            const Event e = trace_line<kFp64>(g, m, field, chk, gthr, ethr, max_iters, seed, lab, &st[2], &st[3]);
            st[1]++;
            if (e == HIT_PREASSIGNED) st[4]++;
            else if (e == LOW_GRADIENT) st[5]++;
            else if (e == OVER_MAX_ITERATIONS) st[6]++;
        }
        lab[s] = (signed char)target;
    }
    st[7] = (long long)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0)
                .count();
    if (stats) std::copy(st, st + 8, stats);
}


// ---- NumericIntegratorExpandingRegionStop ---------------------------------
//
// gi_numeric_integrator_expanding_region_stop.h:63-440 with
// AdaptiveEulerAdvector3D<F, -1> and IndexCompareLessThan (minima, flow
// downhill), run with one thread.  SameBdry6 neighbours: both neighbours
// along an axis, and only where the vertex is interior along it
// (gi_regular_grid.h:232-265).
int same_bdry6(const Grid& g, idx_t id, idx_t out[6]) {
    const V3l c = g.coords(id);
    const idx_t st[3] = {1, g.X, g.X * g.Y};
    int k = 0;
    for (int a = 0; a < 3; a++)
        if (c.v[a] > 0 && c.v[a] < g.n[a] - 1) {
            out[k++] = id + st[a];
            out[k++] = id - st[a];
        }
    return k;
}

// mode 0: the reference's mixed arithmetic, one thread's order (pass 1 with
// path compression: every path vertex takes the label the path ends on,
// later paths overwrite earlier ones; the re-integration reads the labels as
// it leaves them).  mode 1: the same in FP64.  mode 2 (FP64, no visiting
// order): every non-certain vertex traced alone in pass 1, and the
// re-integration reads pass 1's labels.  labels: the minimum's vertex id or
// -1.  stats: extrema, certain vertices, pass-1 streamlines, unambiguous
// vertices, re-integrated vertices.
void expanding_region_stop(const Grid& g, const float* f, float ethr, float gthr, int max_iters, int mode,
                           int* label, long long* stats) {
    const idx_t n = g.X * g.Y * g.Z;
    std::vector<V3d> grad;
    gradient_field(g, f, grad);  // ComputeGradFromImage(1); no gradient is zeroed
    std::vector<int> certain((size_t)n, -1);
    std::vector<idx_t> dest((size_t)n, -1);
    // extrema: no SameBdry6 neighbour before the vertex (:94-107, :218-233)
    std::vector<idx_t> extrema;
    for (idx_t i = 0; i < n; i++) {
        idx_t nb[6];
        const int k = same_bdry6(g, i, nb);
        bool ext = true;
        for (int j = 0; j < k && ext; j++)
            if (before(f, nb[j], i)) ext = false;
        if (ext) {
            dest[(size_t)i] = i;
            extrema.push_back(i);
        }
    }
    // certain regions, one priority flood per extremum in order (:153-180, :235-251)
    struct After {
        const float* f;
        bool operator()(idx_t a, idx_t b) const { return is_greater(f, a, b); }  // lowest on top
    };
    long long ncertain = 0;
    for (size_t m = 0; m < extrema.size(); m++) {
        const idx_t start = extrema[m];
        certain[(size_t)start] = (int)m;
        std::set<idx_t> seen;
        std::priority_queue<idx_t, std::vector<idx_t>, After> front(After{f});
        seen.insert(start);
        auto enqueue_later = [&](idx_t t) {
            idx_t nb[6];
            const int k = same_bdry6(g, t, nb);
            for (int j = 0; j < k; j++) {
                const idx_t u = nb[j];
                if (certain[(size_t)u] == -1 && before(f, t, u) && !seen.count(u)) {
                    seen.insert(u);
                    front.push(u);
                }
            }
        };
        enqueue_later(start);
        while (!front.empty()) {
            const idx_t cur = front.top();
            front.pop();
            // Inspect_Higher_Certains (:113-151): the earlier neighbours all certain, with one label
            idx_t nb[6];
            const int k = same_bdry6(g, cur, nb);
            int lab = -1;
            bool has = false, ok = true;
            for (int j = 0; j < k && ok; j++) {
                if (!before(f, nb[j], cur)) continue;
                const int c = certain[(size_t)nb[j]];
                if (c < 0) ok = false;
                else if (!has) { lab = c; has = true; }
                else if (lab != c) ok = false;
            }
            if (ok && has) {
                certain[(size_t)cur] = lab;
                dest[(size_t)cur] = start;
                enqueue_later(cur);
            }
        }
    }
    for (idx_t i = 0; i < n; i++) ncertain += certain[(size_t)i] != -1;
    auto run = [&](auto fp64) {
        constexpr bool kFp64 = decltype(fp64)::value;
        std::vector<Dest> stop((size_t)n);
        for (idx_t i = 0; i < n; i++) stop[(size_t)i] = certain[(size_t)i] != -1 ? CERTAIN_TERMINAL : UNASSIGNED;
        const Checker chk{stop.data(), false, &g};  // TerminateNearCertain (gi_advection_checker.h:183-208)
        std::vector<idx_t> path;
        long long traced = 0;
        // pass 1 (:252-337)
        const std::vector<idx_t> before_pass1 = dest;
        for (idx_t i = 0; i < n; i++) {
            if (certain[(size_t)i] != -1) continue;
            if (mode != 2 && dest[(size_t)i] != -1) continue;
            path.clear();
            path.push_back(i);
            Event ev;
            const idx_t end = trace<kFp64>(g, grad, chk, gthr, ethr, max_iters, i, &path, &ev);
            traced++;
            if (mode == 2) {
                dest[(size_t)i] = before_pass1[(size_t)end];
            } else {
                const idx_t lab = dest[(size_t)end];
                for (idx_t p : path) dest[(size_t)p] = lab;
            }
        }
        // unambiguous vertices become certain (:348-373)
        long long unamb = 0;
        for (idx_t i = 0; i < n; i++) {
            if (certain[(size_t)i] != -1) continue;
            idx_t nb[6];
            const int k = same_bdry6(g, i, nb);
            bool same = true;
            for (int j = 0; j < k && same; j++)
                if (dest[(size_t)nb[j]] != dest[(size_t)i]) same = false;
            if (same) {
                certain[(size_t)i] = 1;
                stop[(size_t)i] = CERTAIN_TERMINAL;
                unamb++;
            }
        }
        // re-integration with a budget of 500 iterations (:380-420)
        const std::vector<idx_t> after_pass1 = dest;
        long long re = 0;
        for (idx_t i = 0; i < n; i++) {
            if (certain[(size_t)i] != -1) continue;
            Event ev;
            const idx_t end = trace<kFp64>(g, grad, chk, gthr, ethr, 500, i, nullptr, &ev);
            dest[(size_t)i] = mode == 2 ? after_pass1[(size_t)end] : dest[(size_t)end];
            re++;
        }
        if (stats) {
            stats[2] = traced;
            stats[3] = unamb;
            stats[4] = re;
        }
    };
    if (mode == 0) run(std::false_type{});
    else run(std::true_type{});
    for (idx_t i = 0; i < n; i++) label[i] = (int)dest[(size_t)i];
    if (stats) {
        stats[0] = (long long)extrema.size();
        stats[1] = ncertain;
    }
}

}  // namespace

extern "C" {

int oracle_gradient_field(const float* f, long long X, long long Y, long long Z, double* out3) {
    Grid g(X, Y, Z);
    std::vector<V3d> gr;
    gradient_field(g, f, gr);
    for (size_t i = 0; i < gr.size(); i++) for (int a = 0; a < 3; a++) out3[3 * i + a] = gr[i].v[a];
    return 0;
}

// Certain regions from a base (steepest) gradient's critical cells: labels
// named by the representative minimum's vertex id.  Also returns, per
// vertex, the seed flag (1 for vertices of minima) in seeds_out if non-null.
int oracle_certain_regions(const float* f, long long X, long long Y, long long Z, const unsigned char* grad,
                           float presimp, int* labels, signed char* seeds_out) {
    CellMesh m(X, Y, Z);
    Grid g(X, Y, Z);
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    std::vector<unsigned char> maxv((size_t)C), minv((size_t)C);
    oracle_maxmin(f, X, Y, Z, maxv.data(), minv.data());
    CellFunc fn{&m, f, maxv.data()};
    MinGraph mg;
    mg.fn = &fn;
    std::vector<idx_t> sad1;
    for (idx_t c = 0; c < C; c++) {
        if (!m.critical(grad, c)) continue;
        const int d = m.dim(c);
        if (d == 0) mg.add_node(c);
        else if (d == 1) sad1.push_back(c);
    }
    for (idx_t s : sad1) {
        idx_t fc[6];
        m.facets(s, fc);
        const idx_t a = trace_down(m, grad, fc[0]), b = trace_down(m, grad, fc[1]);
        if (a != b) mg.add_arc(a, b, s);
    }
    mg.simplify(presimp);
    // group minima by representative; name a region by its representative's vertex
    std::map<int, std::vector<idx_t>> groups;
    for (int l = 0; l < (int)mg.ext.size(); l++) groups[mg.ext[l].rep].push_back(m.vertex_number(mg.ext[l].cell));
    std::vector<std::vector<idx_t>> seeds;
    std::vector<int> seed_label;
    for (auto& kv : groups) {
        seeds.push_back(kv.second);
        seed_label.push_back((int)m.vertex_number(mg.ext[kv.first].cell));
    }
    certain_regions(g, f, seeds, seed_label, labels);
    if (seeds_out) {
        for (idx_t i = 0; i < X * Y * Z; i++) seeds_out[i] = 0;
        for (auto& s : seeds) for (idx_t v : s) seeds_out[v] = 1;
    }
    return (int)seeds.size();
}

// The descending side (ondemandaccurate.cxx:913-940): maxima (critical hexes)
// joined by interior 2-saddles (critical quads off the boundary, traced up
// through both cofacet hexes, + side first: trace_up_2saddle,
// gi_extrema_region_builder.h:354-372), UFMergeGraph MAXIMAL +
// SimplifyToThreshold, seeds at each maximum's highest vertex grouped by
// representative (a region is named by its representative's highest vertex),
// then the certain regions in the descending order.
int oracle_certain_regions_dsc(const float* f, long long X, long long Y, long long Z, const unsigned char* grad,
                               float presimp, int* labels, signed char* seeds_out) {
    CellMesh m(X, Y, Z);
    Grid g(X, Y, Z);
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    std::vector<unsigned char> maxv((size_t)C), minv((size_t)C);
    oracle_maxmin(f, X, Y, Z, maxv.data(), minv.data());
    CellFunc fn{&m, f, maxv.data()};
    MinGraph mg;
    mg.fn = &fn;
    mg.min_hier = false;
    std::vector<idx_t> sad2;
    for (idx_t c = 0; c < C; c++) {
        if (!m.critical(grad, c)) continue;
        const int d = m.dim(c);
        if (d == 3) mg.add_node(c);
        else if (d == 2 && m.bv(c) == 0) sad2.push_back(c);
    }
    for (idx_t s : sad2) {
        idx_t q[3];
        m.coords(s, q);
        int a = 0;
        while (a < 3 && (q[a] & 1)) a++;
        const idx_t h1 = trace_up(m, grad, s + m.off(a)), h2 = trace_up(m, grad, s - m.off(a));
        if (h1 != h2 && h1 >= 0 && h2 >= 0) mg.add_arc(h1, h2, s);
    }
    mg.simplify(presimp);
    std::map<int, std::vector<idx_t>> groups;
    for (int l = 0; l < (int)mg.ext.size(); l++) groups[mg.ext[l].rep].push_back(fn.highest(mg.ext[l].cell));
    std::vector<std::vector<idx_t>> seeds;
    std::vector<int> seed_label;
    for (auto& kv : groups) {
        seeds.push_back(kv.second);
        seed_label.push_back((int)fn.highest(mg.ext[kv.first].cell));
    }
    certain_regions(g, f, seeds, seed_label, labels, true);
    if (seeds_out) {
        for (idx_t i = 0; i < X * Y * Z; i++) seeds_out[i] = 0;
        for (auto& sd : seeds) for (idx_t v : sd) seeds_out[v] = 1;
    }
    return (int)seeds.size();
}

int oracle_clean_dsc(const float* f, long long X, long long Y, long long Z, int* labels) {
    Grid g(X, Y, Z);
    clean_regions(g, f, labels, true);
    return 0;
}

// Path-compressing integration from certain regions (labels in/out).
// seeds: per-vertex flag of region seeds (their gradient is zeroed).
// sign -1 integrates ascending manifolds (downhill), +1 descending (uphill).
int oracle_integrate_signed(const float* f, long long X, long long Y, long long Z, const signed char* seeds,
                            float ethr, float gthr, int max_iters, int sign, int mode, int* labels, long long* stats) {
    if (sign != -1 && sign != 1) return -1;
    Grid g(X, Y, Z);
    std::vector<V3d> gr;
    gradient_field(g, f, gr);
    std::vector<std::vector<idx_t>> sd(1);
    for (idx_t i = 0; i < X * Y * Z; i++) if (seeds && seeds[i]) sd[0].push_back(i);
    if (mode == 0) integrate<false>(g, f, gr, sd, ethr, gthr, max_iters, labels, stats, sign);
    else if (mode == 1) integrate<true>(g, f, gr, sd, ethr, gthr, max_iters, labels, stats, sign);
    else integrate_exact(g, gr, sd, ethr, gthr, max_iters, labels, stats, sign);
    return 0;
}

int oracle_integrate(const float* f, long long X, long long Y, long long Z, const signed char* seeds, float ethr,
                     float gthr, int max_iters, int mode, int* labels, long long* stats) {
    return oracle_integrate_signed(f, X, Y, Z, seeds, ethr, gthr, max_iters, -1, mode, labels, stats);
}

int oracle_clean(const float* f, long long X, long long Y, long long Z, int* labels) {
    Grid g(X, Y, Z);
    clean_regions(g, f, labels);
    return 0;
}

// VertexLabelingToBoundaryLabeling::ComputeMAXBoundary (gi_labeling_to_bounary_labeling.h:118-195),
// added onto restriction (in/out): the highest vertex of each hex comes from
// the max/min vertex labelling (oracle_maxmin, gi_max_vertex_labeling.h).
int oracle_max_boundary(const float* f, const int* labels, long long X, long long Y, long long Z,
                        signed char* restriction) {
    CellMesh m(X, Y, Z);
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    std::vector<unsigned char> maxv((size_t)C), minv((size_t)C);
    oracle_maxmin(f, X, Y, Z, maxv.data(), minv.data());
    auto highest = [&](idx_t c) {
        const int t = maxv[(size_t)c];
        const idx_t d = (idx_t)(t % 3 - 1) + m.M[0] * ((idx_t)((t / 3) % 3 - 1) + m.M[1] * (idx_t)(t / 9 - 1));
        return c + d;  // vertex cell
    };
    auto interior = [&](idx_t c) { return m.bv(c) == 0; };
    // quads between hexes whose highest vertices carry different valid labels
    for (idx_t c = 0; c < C; c++) {
        if (m.dim(c) != 2 || !interior(c)) continue;
        idx_t p[3];
        m.coords(c, p);
        int a = 0;
        while (p[a] & 1) a++;  // the even axis
        const idx_t v1 = highest(c + m.off(a)), v2 = highest(c - m.off(a));  // cofacets +axis, -axis
        if (v1 == v2) continue;
        const int l1 = labels[m.vertex_number(v1)], l2 = labels[m.vertex_number(v2)];
        if (l1 == -1 || l2 == -1) continue;
        if (l1 != l2) restriction[c] = (signed char)(restriction[c] + 2);
    }
    // then edges under such a quad, then vertices under such an edge
    for (int d = 1; d >= 0; d--)
        for (idx_t c = 0; c < C; c++) {
            if (m.dim(c) != d || !interior(c)) continue;
            idx_t p[3];
            m.coords(c, p);
            for (int a = 0; a < 3; a++) {
                if (p[a] & 1) continue;  // cofacets lie along the even axes
                if (restriction[c + m.off(a)] > 1 || restriction[c - m.off(a)] > 1) {
                    restriction[c] = (signed char)(restriction[c] + 2);
                    break;
                }
            }
        }
    return 0;
}

// VertexLabelingToBoundaryLabeling::ComputeRegionBoundaryKind
// (gi_labeling_to_bounary_labeling.h:262-393): edges, then quads, then hexes,
// each from the previous dimension's kinds; cells the reference leaves
// unwritten (vertices, edges with an unlabelled vertex) are 0.
int oracle_region_boundary_kind(const int* labels, long long X, long long Y, long long Z, signed char* out) {
    CellMesh m(X, Y, Z);
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    std::fill(out, out + C, (signed char)0);
    auto verts = [&](idx_t c, std::vector<int>& l) {
        l.clear();
        idx_t p[3];
        m.coords(c, p);
        for (int k = 0; k < 8; k++) {
            if (((k & 1) && !(p[0] & 1)) || ((k & 2) && !(p[1] & 1)) || ((k & 4) && !(p[2] & 1))) continue;
            const idx_t v = (p[0] / 2 + (k & 1)) + X * ((p[1] / 2 + ((k >> 1) & 1)) + Y * (p[2] / 2 + (k >> 2)));
            l.push_back(labels[v]);
        }
    };
    auto unique_beyond = [](const std::vector<int>& l, int limit) {
        std::vector<int> u;
        for (int x : l) {
            if (std::find(u.begin(), u.end(), x) == u.end()) u.push_back(x);
            if ((int)u.size() > limit) break;
        }
        return (int)u.size() > limit;
    };
    std::vector<int> l;
    for (int d = 1; d <= 3; d++)
        for (idx_t c = 0; c < C; c++) {
            if (m.dim(c) != d) continue;
            verts(c, l);
            if (d == 1) {
                if (l[0] == -1 || l[1] == -1) continue;
                out[c] = l[0] != l[1] ? 1 : 0;
                continue;
            }
            idx_t fc[6];
            const int k = m.facets(c, fc);
            int sum = 0, maxv = 0;
            for (int j = 0; j < k; j++) {
                const int v = out[fc[j]];
                if (d == 2) sum += v == 1;
                else { maxv = std::max(maxv, v); sum += v == 2; }
            }
            if (d == 2) out[c] = (signed char)(sum == 0 ? 0 : (sum <= 2 ? 1 : (unique_beyond(l, 2) ? 2 : 1)));
            else out[c] = (signed char)(sum > 2 ? (unique_beyond(l, 3) ? 3 : 2) : maxv);
        }
    return 0;
}

int oracle_min_boundary(const int* labels, long long X, long long Y, long long Z, signed char* restriction) {
    CellMesh m(X, Y, Z);
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    for (idx_t c = 0; c < C; c++) restriction[c] = 0;
    min_boundary(m, labels, restriction);
    return 0;
}

// Vertex cells whose lower star must be recomputed under the restriction
// (ondemandaccurate.cxx:1014-1052); returns their count.
long long oracle_conforming_vertices(const float* f, long long X, long long Y, long long Z,
                                     const signed char* restriction, long long* out) {
    CellMesh m(X, Y, Z);
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    std::vector<unsigned char> maxv((size_t)C), minv((size_t)C);
    oracle_maxmin(f, X, Y, Z, maxv.data(), minv.data());
    long long n = 0;
    for (idx_t vz = 0; vz < Z; vz++)
        for (idx_t vy = 0; vy < Y; vy++)
            for (idx_t vx = 0; vx < X; vx++) {
                const idx_t p[3] = {2 * vx, 2 * vy, 2 * vz};
                const idx_t v = p[0] + m.M[0] * (p[1] + m.M[1] * p[2]);
                bool hit = restriction[v] > 0;
                for (int t = 0; t < 27 && !hit; t++) {
                    const idx_t q[3] = {p[0] + t % 3 - 1, p[1] + (t / 3) % 3 - 1, p[2] + t / 9 - 1};
                    if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= m.M[0] || q[1] >= m.M[1] || q[2] >= m.M[2]) continue;
                    const idx_t c = q[0] + m.M[0] * (q[1] + m.M[1] * q[2]);
                    if (restriction[c] > 0 && maxv[c] == 26 - t) hit = true;
                }
                if (hit) { if (out) out[n] = v; n++; }
            }
    return n;
}

// The whole accurate ascending pipeline; returns the .grad and the cleaned
// labels.  mode: 0 reference arithmetic, 1 FP64, 2 exact FP64 destinations.
// NumericIntegratorExpandingRegionStop (see expanding_region_stop above).
// labels: X*Y*Z int32; stats: 5 entries (may be null).
int oracle_expanding_region_stop(const float* f, long long X, long long Y, long long Z, float ethr, float gthr,
                                 int max_iters, int mode, int* labels, long long* stats) {
    const Grid g(X, Y, Z);
    expanding_region_stop(g, f, ethr, gthr, max_iters, mode, labels, stats);
    return 0;
}

int oracle_ondemand_asc(const float* f, long long X, long long Y, long long Z, float ethr, float gthr, int max_iters,
                        float presimp, int mode, unsigned char* grad, int* labels, signed char* restriction_out,
                        long long* stats) {
    CellMesh m(X, Y, Z);
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    const idx_t V = X * Y * Z;
    std::vector<unsigned char> base((size_t)C);
    oracle_lstar_pairs(f, X, Y, Z, nullptr, base.data());
    std::vector<signed char> seeds((size_t)V);
    oracle_certain_regions(f, X, Y, Z, base.data(), presimp, labels, seeds.data());
    oracle_integrate(f, X, Y, Z, seeds.data(), ethr, gthr, max_iters, mode, labels, stats);
    oracle_clean(f, X, Y, Z, labels);
    std::vector<signed char> r((size_t)C, 0);
    oracle_min_boundary(labels, X, Y, Z, r.data());
    const long long nv = oracle_conforming_vertices(f, X, Y, Z, r.data(), nullptr);
    std::vector<long long> vs((size_t)nv);
    oracle_conforming_vertices(f, X, Y, Z, r.data(), vs.data());
    oracle_lstar_vertices(f, X, Y, Z, r.data(), vs.data(), nv, base.data());
    oracle_ldir(base.data(), X, Y, Z);
    oracle_flags(base.data(), X, Y, Z);
    std::copy(base.begin(), base.end(), grad);
    if (restriction_out) std::copy(r.begin(), r.end(), restriction_out);
    if (stats) stats[3] = nv;
    return 0;
}

// The digitized 1-manifolds of one family added to restriction (in/out):
// asc 1 = from the 2-saddles up (value 4), asc 0 = from the 1-saddles down
// (value 5).  grad: a .grad whose pairs are the steepest lower-star pairing;
// seeds (may be null): vertices whose gradient is zeroed first, as the
// accurate 3-manifold labelling leaves the shared gradient field
// (gi_numeric_integrator_path_compressing.h:118-123).  mode 0: the
// reference's mixed arithmetic, else FP64.  stats: 8 entries (saddle_lines).
int oracle_saddle_lines(const float* f, long long X, long long Y, long long Z, const unsigned char* grad,
                        const signed char* seeds, float ethr, float gthr, int max_iters, int asc, int mode,
                        signed char* restriction, long long* stats) {
    Grid g(X, Y, Z);
    CellMesh m(X, Y, Z);
    std::vector<V3d> field;
    gradient_field(g, f, field);
    if (seeds)
        for (idx_t i = 0; i < X * Y * Z; i++) if (seeds[i]) field[(size_t)i] = V3d{{0, 0, 0}};
    if (mode == 0) saddle_lines<false>(g, m, f, grad, field, ethr, gthr, max_iters, asc != 0, restriction, stats);
    else saddle_lines<true>(g, m, f, grad, field, ethr, gthr, max_iters, asc != 0, restriction, stats);
    return 0;
}

// OnDemandAccurate with any of needASC1, needDSC1, needASC2 (needDSC2 is not
// restated), ondemandaccurate.cxx:728-1157: the steepest pairing; the
// accurate ascending labelling and its MIN boundary when need_asc2; the
// digitized 1-manifolds (ASC then DSC); the conforming lower stars; ldir and
// flags.  labels (V) are written only when need_asc2 (else left as they are);
// stats: [0..3] the labelling's, [4..11] ASC lines, [12..19] DSC lines.
// OnDemandAccurate with any of needASC1, needDSC1, needASC2, needDSC2
// (ondemandaccurate.cxx:728-1157).  labels_dsc (may be null when dsc2 is 0):
// the descending labels.  The gradient field is shared: the descending
// integration and the 1-manifolds see it zeroed at every seed so far.
int oracle_ondemand2(const float* f, long long X, long long Y, long long Z, float ethr, float gthr, int max_iters,
                     float presimp, int need_asc1, int need_dsc1, int need_asc2, int need_dsc2, int mode,
                     unsigned char* grad, int* labels, int* labels_dsc, signed char* restriction_out,
                     long long* stats) {
    CellMesh m(X, Y, Z);
    const idx_t C = m.M[0] * m.M[1] * m.M[2];
    const idx_t V = X * Y * Z;
    std::vector<unsigned char> base((size_t)C);
    oracle_lstar_pairs(f, X, Y, Z, nullptr, base.data());
    std::vector<signed char> r((size_t)C, 0), seeds((size_t)V, 0);
    long long st[20] = {0};
    if (need_asc2) {
        oracle_certain_regions(f, X, Y, Z, base.data(), presimp, labels, seeds.data());
        oracle_integrate(f, X, Y, Z, seeds.data(), ethr, gthr, max_iters, mode, labels, st);
        oracle_clean(f, X, Y, Z, labels);
        oracle_min_boundary(labels, X, Y, Z, r.data());
    }
    if (need_dsc2) {
        std::vector<signed char> ds((size_t)V, 0);
        oracle_certain_regions_dsc(f, X, Y, Z, base.data(), presimp, labels_dsc, ds.data());
        for (idx_t i = 0; i < V; i++) seeds[(size_t)i] |= ds[(size_t)i];
        oracle_integrate_signed(f, X, Y, Z, seeds.data(), ethr, gthr, max_iters, 1, mode, labels_dsc, st + 16);
        oracle_clean_dsc(f, X, Y, Z, labels_dsc);
        oracle_max_boundary(f, labels_dsc, X, Y, Z, r.data());
    }
    const int lmode = mode == 0 ? 0 : 1;
    const signed char* zs = (need_asc2 || need_dsc2) ? seeds.data() : nullptr;
    if (need_asc1) oracle_saddle_lines(f, X, Y, Z, base.data(), zs, ethr, gthr, max_iters, 1, lmode, r.data(), st + 4);
    if (need_dsc1) oracle_saddle_lines(f, X, Y, Z, base.data(), zs, ethr, gthr, max_iters, 0, lmode, r.data(), st + 12);
    if (need_asc2 || need_dsc2 || need_asc1 || need_dsc1) {
        const long long nv = oracle_conforming_vertices(f, X, Y, Z, r.data(), nullptr);
        std::vector<long long> vs((size_t)nv);
        oracle_conforming_vertices(f, X, Y, Z, r.data(), vs.data());
        oracle_lstar_vertices(f, X, Y, Z, r.data(), vs.data(), nv, base.data());
        st[3] = nv;
    }
    oracle_ldir(base.data(), X, Y, Z);
    oracle_flags(base.data(), X, Y, Z);
    std::copy(base.begin(), base.end(), grad);
    if (restriction_out) std::copy(r.begin(), r.end(), restriction_out);
    if (stats) std::copy(st, st + 20, stats);
    return 0;
}

int oracle_ondemand(const float* f, long long X, long long Y, long long Z, float ethr, float gthr, int max_iters,
                    float presimp, int need_asc1, int need_dsc1, int need_asc2, int mode, unsigned char* grad,
                    int* labels, signed char* restriction_out, long long* stats) {
    return oracle_ondemand2(f, X, Y, Z, ethr, gthr, max_iters, presimp, need_asc1, need_dsc1, need_asc2, 0, mode,
                            grad, labels, nullptr, restriction_out, stats);
}

}  // extern "C"

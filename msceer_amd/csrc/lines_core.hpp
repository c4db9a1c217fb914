// Device code of the accurate 1-manifolds (lines.hip): cell geometry, the
// saddles' seed search, one digitizing streamline and the label policies of
// its speculative, replayed and in-order traces.  Kept in a header so that a
// host build of the same source can be checked against the CPU oracle; the
// includer provides what tracer.hpp needs plus the HIP atomics used here
// (atomicMin, atomicAdd, __hip_atomic_load / _store /
// _compare_exchange_strong).
#pragma once

#include "tracer.hpp"

#pragma clang fp contract(off)

namespace msceer {
namespace {

constexpr uint32_t kNoLine = 0xFFFFFFFFu;
constexpr int kAbort = 100;  // trace event: the footprint rule stopped it

struct CellGeom {
    int64_t Mx, My, Mz, sxy;
    __device__ uint32_t id(int64_t x, int64_t y, int64_t z) const { return (uint32_t)(x + Mx * (y + My * z)); }
    __device__ void xyz(uint32_t c, int64_t& x, int64_t& y, int64_t& z) const {
        x = (int64_t)c % Mx;
        y = ((int64_t)c / Mx) % My;
        z = (int64_t)c / sxy;
    }
    __device__ int64_t off(int a) const { return a == 0 ? 1 : (a == 1 ? Mx : sxy); }
    __device__ int64_t ext(int a) const { return a == 0 ? Mx : (a == 1 ? My : Mz); }
    // Cell coordinates inside the complex.  A streamline away from the
    // boundary stays inside; this only keeps a step that a huge error
    // threshold let run far from addressing outside the arrays.
    __device__ bool inside(int64_t x, int64_t y, int64_t z) const {
        return x >= 0 && y >= 0 && z >= 0 && x < Mx && y < My && z < Mz;
    }
};

// Order key of a float: monotone, -0 and +0 equal (as float comparison has them).
__device__ __forceinline__ uint32_t value_key(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) == 0u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ bool critical(const uint8_t* grad, uint32_t c) { return ((grad[c] >> 2) & 7u) == 7u; }

__device__ __forceinline__ uint32_t pair_of(const CellGeom& m, const uint8_t* grad, uint32_t c) {
    const uint32_t code = (grad[c] >> 2) & 7u;
    const int64_t o = m.off((int)(code >> 1));
    return (uint32_t)((int64_t)c + ((code & 1u) ? -o : o));
}

__device__ __forceinline__ double comp(const V3<double>& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// DigitizeSegmentInternal: the faces the segment a -> b crosses, by their
// crossing parameter; for each face that bounds the current cell, the
// current cell and the face, then the cell beyond becomes current.  ASC:
// voxel faces, cells = hexes; DSC: the dual grid (positions shifted by half a
// voxel), faces = edges, cells = vertices.  digitize_cells lists those cells
// in order (cell, face, cell, face, ...: at most six) without looking at any
// label; digitize labels them in that order and stops at the first one that
// already carries the target.
__device__ inline int digitize_cells(const CellGeom& m, V3<double> a, V3<double> b, bool asc, uint32_t* out) {
    if (!asc) {
        a = {a.x + 0.5, a.y + 0.5, a.z + 0.5};
        b = {b.x + 0.5, b.y + 0.5, b.z + 0.5};
    }
    const V3<double> ab{b.x - a.x, b.y - a.y, b.z - a.z};
    uint32_t face[3];
    double tc[3];
    int n = 0;
    for (int d = 0; d < 3; d++) {
        const double av = comp(a, d), bv = comp(b, d);
        const int va = (int)floor(av), vb = (int)floor(bv);
        if (va == vb) continue;
        const int plane = va > vb ? va : vb;
        const double t = ((double)plane - av) / (bv - av);
        tc[n] = t;
        if (t < 0 || t > 1) return 0;
        int64_t fc[3];
        for (int k = 0; k < 3; k++)
            fc[k] = 2 * (int64_t)floor(comp(ab, k) * t + comp(a, k)) + (k == d ? 0 : 1) - (asc ? 0 : 1);
        if (!m.inside(fc[0], fc[1], fc[2])) return 0;
        face[n++] = m.id(fc[0], fc[1], fc[2]);
    }
    if (n == 0) return 0;
    for (int i = 0; i < n; i++)
        for (int j = i; j < n; j++)
            if (tc[i] > tc[j]) {
                const uint32_t tf = face[i]; face[i] = face[j]; face[j] = tf;
                const double tt = tc[i]; tc[i] = tc[j]; tc[j] = tt;
            }
    const int64_t s0 = 2 * (int64_t)floor(a.x) + (asc ? 1 : 0), s1 = 2 * (int64_t)floor(a.y) + (asc ? 1 : 0),
                  s2 = 2 * (int64_t)floor(a.z) + (asc ? 1 : 0);
    if (!m.inside(s0, s1, s2)) return 0;
    uint32_t cur = m.id(s0, s1, s2);
    int k = 0;
    for (int i = 0; i < n; i++) {
        int64_t p[3];
        m.xyz(face[i], p[0], p[1], p[2]);
        bool has = false;
        uint32_t next = cur;
        for (int d = 0; d < 3; d++) {
            if (asc ? (p[d] & 1) != 0 : (p[d] & 1) == 0) continue;
            if (p[d] < m.ext(d) - 1) {
                const uint32_t c = (uint32_t)((int64_t)face[i] + m.off(d));
                if (c == cur) has = true; else next = c;
            }
            if (p[d] > 0) {
                const uint32_t c = (uint32_t)((int64_t)face[i] - m.off(d));
                if (c == cur) has = true; else next = c;
            }
        }
        if (!has) continue;
        out[k++] = cur;
        out[k++] = face[i];
        cur = next;
    }
    return k;
}

// The seed cells of saddle s from the short V-path search
// (ondemandaccurate.cxx:319-351 ASC, 448-480 DSC), at most six, in out[0..5]
// (kNoLine where absent), and the saddle in out[6].  The search pops at most
// five cells: the saddle, whose two cofacets (ASC) / facets (DSC) are new,
// then cells that each add one new neighbour.
__device__ inline void saddle_seeds(const uint8_t* grad, const CellGeom& m, uint32_t s, bool asc, uint32_t* out) {
    uint32_t result[6], queue[6], seen[16];
    int nres = 0, qh = 0, qt = 0, ns = 0;
    auto is_seen = [&](uint32_t c) {
        for (int i = 0; i < ns; i++)
            if (seen[i] == c) return true;
        return false;
    };
    queue[qt++] = s;
    for (int counter = 4; qh < qt && counter >= 0; counter--) {
        const uint32_t cur = queue[qh++];
        if (!is_seen(cur) && ns < 16) seen[ns++] = cur;
        int64_t p[3];
        m.xyz(cur, p[0], p[1], p[2]);
        for (int a = 0; a < 3; a++) {
            if (asc ? (p[a] & 1) != 0 : (p[a] & 1) == 0) continue;
            for (int sg = 0; sg < 2; sg++) {
                if (sg == 0 ? p[a] >= m.ext(a) - 1 : p[a] <= 0) continue;
                const uint32_t t = (uint32_t)((int64_t)cur + (sg == 0 ? m.off(a) : -m.off(a)));
                if (is_seen(t)) continue;
                if (nres < 6) result[nres++] = t;
                if (ns < 16) seen[ns++] = t;
                if (!critical(grad, t)) {
                    const uint32_t pr = pair_of(m, grad, t);
                    if (ns < 16) seen[ns++] = pr;
                    if (qt < 6) queue[qt++] = pr;
                }
            }
        }
    }
    for (int j = 0; j < 6; j++) out[j] = j < nres ? result[j] : kNoLine;
    out[6] = s;
}

// ---- one streamline ----------------------------------------------------------

// Label policies of a trace: test(c) -- does cell c carry the target (a
// checker or digitizing test); write(c) -- give it the target; ok() -- false
// once the trace must stop for a footprint rule.

// Alone and in order: the restriction labels themselves.
struct DirectLabels {
    int8_t* restr;
    int8_t target;
    __device__ bool test(uint32_t c) { return restr[c] == target; }
    __device__ void write(uint32_t c) { restr[c] = target; }
    __device__ bool ok() const { return true; }
    __device__ void stop_here(const CellGeom&, const V3<double>&, const V3<double>&, bool) {}
};

// Footprints: every cell a speculative trace tests or labels, appended (with
// repeats) to chunks of a shared pool, one chain per line (and a second chain
// of the cells it labelled): chunk = the index of the next chunk (or kNoLine)
// and kChunk - 1 cells, the last chunk of a footprint ended by kNoLine.
constexpr uint32_t kChunk = 32;

struct Pool {
    uint32_t* cells;     // capacity words
    uint32_t capacity;   // a multiple of kChunk
    uint32_t* top;       // next free word
};

// Speculation: committed labels plus the line's own labels, known exactly
// (own[c]: the last line of the window that labelled c; its second chain
// where another line overwrote that mark).  Labelled cells claim wstamp (the
// lowest line of the window labelling them); the first claim of a cell
// appends it to the window's list of labelled cells.
struct SpecLabels {
    const int8_t* restr;
    int8_t target;
    uint32_t* wstamp;
    uint32_t* pstamp;
    uint32_t* own;
    uint32_t line;
    Pool pool;
    uint32_t* touched;     // cells labelled (or possibly labelled) in the window
    uint32_t* ntouched;
    struct Chain {
        uint32_t head = kNoLine, chunk = kNoLine, fill = 0;
    } all, wrote, maybe;   // the footprint; the cells labelled; the cells a stop could label
    uint32_t last[6];      // the cells of the last stop_here, not recorded again
    int nlast = -1;
    bool over = false;
    __device__ void append(Chain& ch, uint32_t c) {
        if (over) return;
        if (ch.chunk == kNoLine || ch.fill == kChunk - 1) {
            const uint32_t at = atomicAdd(pool.top, kChunk);
            if (at + kChunk > pool.capacity) {
                over = true;
                return;
            }
            pool.cells[at] = kNoLine;
            if (ch.chunk == kNoLine) ch.head = at;
            else pool.cells[ch.chunk] = at;
            ch.chunk = at;
            ch.fill = 0;
        }
        pool.cells[ch.chunk + 1 + ch.fill] = c;
        ch.fill++;
    }
    __device__ void claim(uint32_t* stamp, uint32_t c) {
        if (atomicMin(&stamp[c], line) == kNoLine) {
            const uint32_t at = atomicAdd(ntouched, 1u);
            if (at < pool.capacity) touched[at] = c;  // (never more cells than pool words)
        }
    }
    // Did this line label c?  own[c] says yes unless another line of the
    // window labelled c since; wstamp[c] > line says no (a line's own claim
    // keeps wstamp at or below it); otherwise the line's labels are searched.
    __device__ bool mine(uint32_t c) const {
        if (own[c] == line) return true;
        if (wstamp[c] > line) return false;
        for (uint32_t ch = wrote.head; ch != kNoLine; ch = pool.cells[ch]) {
            const uint32_t n = ch == wrote.chunk ? wrote.fill : kChunk - 1;
            for (uint32_t j = 0; j < n; j++)
                if (pool.cells[ch + 1 + j] == c) return true;
        }
        return false;
    }
    __device__ bool test(uint32_t c) {
        append(all, c);
        return restr[c] == target || mine(c);
    }
    __device__ void write(uint32_t c) {
        append(all, c);
        append(wrote, c);
        if (over) return;  // the window ends before this line
        own[c] = line;
        claim(wstamp, c);
    }
    // Cells a stop at this point would label (its digitization start ->
    // here): footprint, and claimed in pstamp, so that a replay stopping here
    // stays within what the speculation saw.
    __device__ void stop_here(const CellGeom& m, const V3<double>& a, const V3<double>& b, bool asc) {
        uint32_t cells[6];
        const int n = digitize_cells(m, a, b, asc, cells);
        bool same = n == nlast;
        for (int k = 0; same && k < n; k++) same = cells[k] == last[k];
        if (same) return;
        nlast = n;
        for (int k = 0; k < n; k++) {
            last[k] = cells[k];
            append(all, cells[k]);
            append(maybe, cells[k]);
            if (over) return;
            claim(pstamp, cells[k]);
        }
    }
    // End marks of the last chunks.
    __device__ void finish() {
        if (over) return;
        if (all.chunk != kNoLine && all.fill < kChunk - 1) pool.cells[all.chunk + 1 + all.fill] = kNoLine;
        if (wrote.chunk != kNoLine && wrote.fill < kChunk - 1) pool.cells[wrote.chunk + 1 + wrote.fill] = kNoLine;
        if (maybe.chunk != kNoLine && maybe.fill < kChunk - 1) pool.cells[maybe.chunk + 1 + maybe.fill] = kNoLine;
    }
    __device__ bool ok() const { return !over; }
};

// Union-find over the lines of a window (relative indices).
__device__ __forceinline__ uint32_t lf_parent(uint32_t* p, uint32_t x) {
    return __hip_atomic_load(&p[x], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t lf_find(uint32_t* p, uint32_t x) {
    uint32_t q = lf_parent(p, x);
    while (q != x) {
        const uint32_t gq = lf_parent(p, q);
        if (gq != q) __hip_atomic_store(&p[x], gq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        x = q;
        q = gq;
    }
    return x;
}
// Hook the later root under the earlier one, so a component's root is its
// first line.
__device__ void lf_union(uint32_t* p, uint32_t a, uint32_t b) {
    while (true) {
        a = lf_find(p, a);
        b = lf_find(p, b);
        if (a == b) return;
        if (a > b) { const uint32_t t = a; a = b; b = t; }
        uint32_t expected = b;
        if (__hip_atomic_compare_exchange_strong(&p[b], &expected, a, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT))
            return;
    }
}


// Replay of one component, its lines in order: committed labels plus those
// its earlier lines gave (lstamp).  A cell that lines of the window labelled,
// or could label by stopping, in speculation belongs to their component; a
// replayed line may test it only in its own component, and may label only
// cells it labelled, or could label, itself in speculation.  Anything else
// means the line went beyond what the speculation saw, and the window is cut
// before it.
struct ReplayLabels {
    const int8_t* restr;
    int8_t target;
    const uint32_t* wstamp;
    const uint32_t* pstamp;
    uint32_t* lstamp;
    const uint32_t* own;
    const uint32_t* pool;
    uint32_t whead, mhead;  // chains of the cells this line labelled / a stop could label, in speculation
    uint32_t* parent;
    uint32_t w0, wc, root, line;
    uint32_t* why;          // failure counters: [0] tests, [1] labels
    bool bad;
    __device__ static bool in_chain(const uint32_t* pool, uint32_t head, uint32_t c) {
        for (uint32_t ch = head; ch != kNoLine; ch = pool[ch])
            for (uint32_t j = 1; j < kChunk; j++) {
                const uint32_t x = pool[ch + j];
                if (x == kNoLine) break;
                if (x == c) return true;
            }
        return false;
    }
    __device__ bool foreign(uint32_t w) const { return w >= w0 && w < wc && lf_find(parent, w - w0) != root; }
    __device__ bool test(uint32_t c) {
        if (foreign(wstamp[c]) || foreign(pstamp[c])) {
            if (!bad) atomicAdd(&why[0], 1u);
            bad = true;
            return false;
        }
        return restr[c] == target || lstamp[c] != kNoLine;
    }
    __device__ void write(uint32_t c) {
        const bool seen = own[c] == line || (wstamp[c] <= line && in_chain(pool, whead, c)) ||
                          (pstamp[c] <= line && in_chain(pool, mhead, c));
        if (!seen) {
            if (!bad) atomicAdd(&why[1], 1u);
            bad = true;
            return;
        }
        lstamp[c] = line;
    }
    __device__ bool ok() const { return !bad; }
    __device__ void stop_here(const CellGeom&, const V3<double>&, const V3<double>&, bool) {}
};

// Label the cells of digitize_cells in order; stop at the first one that
// already carries the target.
template <class L>
__device__ void digitize(const CellGeom& m, const V3<double>& a, const V3<double>& b, bool asc, L& lab) {
    uint32_t cells[6];
    const int n = digitize_cells(m, a, b, asc, cells);
    for (int k = 0; k < n; k++) {
        if (lab.test(cells[k])) return;
        lab.write(cells[k]);
        if (!lab.ok()) return;
    }
}

template <class L>
__device__ int advect_voxel(const VGrid& g, const CellGeom& m, Tracer<double>& t, const V3<double>& start, double dir,
                            double gthr, double ethr, bool asc, L& lab) {
    while (true) {
        if (!(t.iters-- > 0)) return kOverMax;
        if (!(t.old.x == t.cur.x && t.old.y == t.cur.y && t.old.z == t.cur.z)) {
            if ((int)t.cur.x != (int)t.old.x || (int)t.cur.y != (int)t.old.y || (int)t.cur.z != (int)t.old.z)
                return kOutOfVoxel;
            t.gr = trilin(t.c, t.cur.x - (double)t.bx, t.cur.y - (double)t.by, t.cur.z - (double)t.bz);
            t.gm = sqrt(t.gr.x * t.gr.x + t.gr.y * t.gr.y + t.gr.z * t.gr.z);
            if (t.gm * t.step > kMaxStep) t.step = kMaxStep / t.gm;
            t.old = t.cur;
        }
        if (t.gm < gthr) return kLowGradient;
        const double s = dir * t.step, hs = (dir * 0.5) * t.step;
        const V3<double> full{t.gr.x * s, t.gr.y * s, t.gr.z * s};
        const V3<double> half{t.gr.x * hs, t.gr.y * hs, t.gr.z * hs};
        const V3<double> hg = trilin(t.c, t.cur.x + half.x - (double)t.bx, t.cur.y + half.y - (double)t.by,
                                     t.cur.z + half.z - (double)t.bz);
        const V3<double> two{hg.x * hs + half.x, hg.y * hs + half.y, hg.z * hs + half.z};
        const double dx = two.x - full.x, dy = two.y - full.y, dz = two.z - full.z;
        const double err = dx * dx + dy * dy + dz * dz;
        if (!(err < ethr)) {
            t.step = t.step * 0.5;
            continue;
        }
        t.cur = {t.cur.x + two.x, t.cur.y + two.y, t.cur.z + two.z};
        if (t.near) { t.cur.x = clampc(t.cur.x, g.X); t.cur.y = clampc(t.cur.y, g.Y); t.cur.z = clampc(t.cur.z, g.Z); }
        const double grow = t.step * 1.8, cap = kMaxStep / t.gm;
        t.step = grow < cap ? grow : cap;
        if (t.near) return kHitPreassigned;
        // a stop at a step that stays in the voxel digitizes start -> here
        // instead of the whole voxel segment
        if ((int)t.cur.x == (int)t.old.x && (int)t.cur.y == (int)t.old.y && (int)t.cur.z == (int)t.old.z)
            lab.stop_here(m, start, t.cur, asc);
        const int64_t cx = asc ? 2 * (int64_t)floor(t.cur.x) + 1 : 2 * (int64_t)floor(t.cur.x + 0.5);
        const int64_t cy = asc ? 2 * (int64_t)floor(t.cur.y) + 1 : 2 * (int64_t)floor(t.cur.y + 0.5);
        const int64_t cz = asc ? 2 * (int64_t)floor(t.cur.z) + 1 : 2 * (int64_t)floor(t.cur.z + 0.5);
        if (!m.inside(cx, cy, cz)) continue;
        const bool hit = lab.test(m.id(cx, cy, cz));
        if (!lab.ok()) return kAbort;
        if (hit) return kHitPreassigned;
    }
}

// IntegrateStreamline from the centre of seed cell `seed` (cell coordinates / 2).
template <class L>
__device__ int trace_line(const GradSrc<double>& src, const VGrid& g, const CellGeom& m, uint32_t seed, bool asc,
                          double gthr, double ethr, int max_iters, L& lab, uint32_t* voxels) {
    int64_t sx, sy, sz;
    m.xyz(seed, sx, sy, sz);
    Tracer<double> t;
    t.cur = {(double)sx * 0.5, (double)sy * 0.5, (double)sz * 0.5};
    t.iters = max_iters;
    tracer_load(src, g, t, (int)(t.cur.x + 0.5), (int)(t.cur.y + 0.5), (int)(t.cur.z + 0.5));
    tracer_begin(t);
    const double dir = asc ? 1.0 : -1.0;
    while (true) {
        const V3<double> start = t.cur;
        const int ev = advect_voxel(g, m, t, start, dir, gthr, ethr, asc, lab);
        (*voxels)++;
        if (ev == kAbort) return kAbort;
        if (!t.near) {
            digitize(m, start, t.cur, asc, lab);
            if (!lab.ok()) return kAbort;
        }
        if (ev != kOutOfVoxel) return ev;
        int x, y, z;
        nearest_xyz(g, t.cur, t.near, x, y, z);
        tracer_load(src, g, t, x, y, z);
        tracer_begin(t);
    }
}

}  // namespace
}  // namespace msceer

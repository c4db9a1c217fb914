// Per-vertex lower-star logic shared by the gfx950 kernels and by their
// host-side emulation in the tests.  Bit conventions are those of star.hpp.
#pragma once
#include <stdint.h>

#include "star.hpp"

namespace msceer {

// Order-preserving key of a float: +0 and -0 map to the same key.
MSC_HD uint32_t float_key(float f) {
    uint32_t u = __builtin_bit_cast(uint32_t, f);
    if ((u & 0x7fffffffu) == 0u) u = 0u;
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

MSC_HD int tid_x(int t) { return t - 3 * ((t * 11) >> 5) - 1; }
MSC_HD int tid_y(int t) { return ((t * 11) >> 5) - 3 * ((t * 57) >> 9) - 1; }
MSC_HD int tid_z(int t) { return ((t * 57) >> 9) - 1; }

// Components of bit t from the constant axis masks (cheaper than the
// divisions above when t varies per lane).
MSC_HD int comp_x(int t) { return (int)((kXP >> t) & 1u) - (int)((kXM >> t) & 1u); }
MSC_HD int comp_y(int t) { return (int)((kYP >> t) & 1u) - (int)((kYM >> t) & 1u); }
MSC_HD int comp_z(int t) { return (int)((kZP >> t) & 1u) - (int)((kZM >> t) & 1u); }

// Keys of the 27 vertices around one vertex, read through fixed strides (an
// LDS brick) or run-time strides (a padded host volume).
// Every key source gives at64(t): the order key of star vertex t as a 64-bit
// integer, comparing as the reference order restricted to one neighbourhood.
// 32-bit sources append the star bit (star-bit order is global-id order for
// the vertices of one 3x3x3 block); 64-bit sources hold (key << 32 | id).
template <int SX, int SXY>
struct FixedKeys {
    const uint32_t* p;
    int base;
    MSC_HD uint32_t at(int t) const { return p[base + tid_x(t) + SX * tid_y(t) + SXY * tid_z(t)]; }
    MSC_HD uint64_t at64(int t) const { return ((uint64_t)at(t) << 5) | (uint64_t)t; }
};
// 64-bit (key << 32 | vertex id) entries through fixed strides or a table.
template <int SX, int SXY>
struct FixedKeys64 {
    const uint64_t* p;
    int base;
    MSC_HD uint64_t at64(int t) const { return p[base + tid_x(t) + SX * tid_y(t) + SXY * tid_z(t)]; }
};
struct TableKeys64 {
    const uint64_t* p;   // centre of the lane's star in the tile
    const int* off;
    MSC_HD uint64_t at64(int t) const { return p[off[t]]; }
};
// One lane's 27 keys laid out t-major with a fixed lane stride (an LDS slice):
// key t of the lane is p[t * kStride].
template <int kStride>
struct SliceKeys {
    const uint32_t* p;
    MSC_HD uint32_t at(int t) const { return p[t * kStride]; }
    MSC_HD uint64_t at64(int t) const { return ((uint64_t)at(t) << 5) | (uint64_t)t; }
};
// Keys read through a per-star offset table (tile offsets of the 27 star
// vertices, shared by the work-group) for star indices known only at run time.
struct TableKeys {
    const uint32_t* p;   // centre of the lane's star in the tile
    const int* off;
    MSC_HD uint32_t at(int t) const { return p[off[t]]; }
    MSC_HD uint64_t at64(int t) const { return ((uint64_t)at(t) << 5) | (uint64_t)t; }
};
struct StridedKeys {
    const uint32_t* p;
    int64_t base, sx, sxy;
    MSC_HD uint32_t at(int t) const { return p[base + tid_x(t) + sx * tid_y(t) + sxy * tid_z(t)]; }
    MSC_HD uint64_t at64(int t) const { return ((uint64_t)at(t) << 5) | (uint64_t)t; }
};

// Order key of star vertex t (see at64 above).
template <class Keys>
MSC_HD uint64_t vkey(const Keys& k, int t) {
    return k.at64(t);
}

// Bits of the cofacets of the cell at bit a (one zero component moved to +-1).
MSC_HD uint32_t cofacet_mask(int a) {
    const uint32_t ba = 1u << a;
    uint32_t m = 0;
    if (kX0 & ba) m |= (ba << 1) | (ba >> 1);
    if (kY0 & ba) m |= (ba << 3) | (ba >> 3);
    if (kZ0 & ba) m |= (ba << 9) | (ba >> 9);
    return m;
}

// Lowest vertex of a 2-cell (quad) at bit c of the lower star: its vertices
// other than the centre are its two edge ends and its far corner.
template <class Keys>
MSC_HD uint64_t quad_lowest(const Keys& k, int c) {
    const int dx = comp_x(c), dy = comp_y(c), dz = comp_z(c);
    // the two edges: zero one of the two nonzero components
    const int e1 = dx ? c - dx : c - 3 * dy;           // first nonzero axis zeroed
    const int e2 = dz ? c - 9 * dz : c - 3 * dy;       // last nonzero axis zeroed
    uint64_t a = vkey(k, c), b = vkey(k, e1), d = vkey(k, e2);
    a = b < a ? b : a;
    return d < a ? d : a;
}

// Lowest vertex of a 3-cell (hex): all 7 of its vertices below the centre.
template <class Keys>
MSC_HD uint64_t hex_lowest(const Keys& k, int c) {
    const int dx = comp_x(c), dy = comp_y(c), dz = comp_z(c);
    uint64_t best = vkey(k, c);
#pragma unroll
    for (int s = 1; s < 7; s++) {
        const int q = c - ((s & 1) ? dx : 0) - ((s & 2) ? 3 * dy : 0) - ((s & 4) ? 9 * dz : 0);
        const uint64_t kq = vkey(k, q);
        best = kq < best ? kq : best;
    }
    return best;
}

struct StarResult {
    uint32_t c0 = 0, c1 = 0, c2 = 0;  // bits of each cell's 3-bit pair code
    uint32_t crit = 0;                // critical cells
    MSC_HD void set_pair(int a, int b) {
        const uint32_t ca = star_dir_code(a, b), cb = ca ^ 1u;
        c0 |= ((ca & 1u) << a) | ((cb & 1u) << b);
        c1 |= (((ca >> 1) & 1u) << a) | (((cb >> 1) & 1u) << b);
        c2 |= ((ca >> 2) << a) | ((cb >> 2) << b);
    }
    MSC_HD void set_critical(uint32_t m) {
        c0 |= m; c1 |= m; c2 |= m; crit |= m;
    }
};

// The same results written straight as .grad bytes (star_cell_byte's) into a
// byte image of the cells: the stage byte of star cell t is sc[soff[t]];
// onb holds the cells on a domain face (their flag bit).
struct StageBytes {
    uint8_t* sc;
    const int* soff;
    uint32_t onb;
    MSC_HD void set_pair(int a, int b) const {
        const uint32_t ca = star_dir_code(a, b);
        sc[soff[a]] = (uint8_t)(0x61u | (((onb >> a) & 1u) << 1) | (ca << 2));
        sc[soff[b]] = (uint8_t)(0x61u | (((onb >> b) & 1u) << 1) | ((ca ^ 1u) << 2));
    }
    MSC_HD void set_critical(uint32_t m) const {
        for (; m; m &= m - 1) {
            const int t = __builtin_ctz(m);
            const uint32_t ldir = ((kDim3 >> t) & 1u) ? 0u : 3u;
            sc[soff[t]] = (uint8_t)(0x1du | (((onb >> t) & 1u) << 1) | (ldir << 5));
        }
    }
};

// First candidate, in cofacet order +x,-x,+y,-y,+z,-z, whose lowest vertex is
// strictly smallest (gi_fast_robins_noalloc.h:124-137 over the cofacet
// iteration order of :256-270).
template <int kCandDim, class Keys>
MSC_HD int pick_candidate(const Keys& kk, int a, uint32_t cset) {
    if (!(cset & (cset - 1))) return __builtin_ctz(cset);
    // only the candidates are visited; comparing (lowest key, position in the
    // cofacet order) makes the earliest of equal keys win, as the ordered scan
    int best = -1, brank = 6;
    uint64_t bk = ~0ull;
    for (uint32_t m = cset; m; m &= m - 1) {
        const int c = __builtin_ctz(m);
        const int d = c - a;
        const int rank = d == 1 ? 0 : (d == -1 ? 1 : (d == 3 ? 2 : (d == -3 ? 3 : (d == 9 ? 4 : 5))));
        const uint64_t kc = kCandDim == 2 ? quad_lowest(kk, c) : hex_lowest(kk, c);
        if (kc < bk || (kc == bk && rank < brank)) { bk = kc; best = c; brank = rank; }
    }
    return best;
}

// Relaxed Robins expansion of one subset S of the lower star
// (gi_fast_robins_noalloc.h:147-321, gi_modified_robins.h:999-1199):
//   vertex: pair with the edge whose other vertex comes first, else critical;
//   then for d = 1, 2: repeatedly take the first unprocessed d-cell (star
//   order) having a cofacet in S with exactly one unprocessed facet, and pair
//   it with the first lowest such cofacet; when none has one, the first
//   unprocessed d-cell with the lowest lowest vertex becomes critical;
//   leftover 3-cells are critical.
template <class Keys, class Res>
MSC_HD void expand_subset(uint32_t S, const Keys& k, Res& r) {
    uint32_t P = 0;  // processed (paired or critical)
    if (S & kCenter) {
        const uint32_t E = S & kDim1;
        if (!E) {
            r.set_critical(kCenter);
            P |= kCenter;
        } else {
            int best = __builtin_ctz(E);
            if (E & (E - 1)) {
                uint64_t bk = ~0ull;
                for (uint32_t m = E; m; m &= m - 1) {
                    const int t = __builtin_ctz(m);
                    const uint64_t kt = vkey(k, t);
                    if (kt < bk) { bk = kt; best = t; }
                }
            }
            r.set_pair(13, best);
            P |= kCenter | (1u << best);
        }
    }
    // edges
    while (true) {
        const uint32_t U = S & ~P;
        const uint32_t Ud = U & kDim1;
        if (!Ud) break;
        // quads have two edge facets: exactly one unprocessed <=> xor
        const uint32_t cand = S & kDim2 & (star_facet_x(U) ^ star_facet_y(U) ^ star_facet_z(U));
        const uint32_t ready = Ud & star_has_cofacet_in(cand);
        if (ready) {
            const int a = __builtin_ctz(ready);
            const int c = pick_candidate<2>(k, a, cand & cofacet_mask(a));
            r.set_pair(a, c);
            P |= (1u << a) | (1u << c);
        } else {
            int best = __builtin_ctz(Ud);
            if (Ud & (Ud - 1)) {
                uint64_t bk = ~0ull;
                for (uint32_t m = Ud; m; m &= m - 1) {
                    const int t = __builtin_ctz(m);
                    const uint64_t kt = vkey(k, t);
                    if (kt < bk) { bk = kt; best = t; }
                }
            }
            r.set_critical(1u << best);
            P |= 1u << best;
        }
    }
    // quads
    while (true) {
        const uint32_t U = S & ~P;
        const uint32_t Ud = U & kDim2;
        if (!Ud) break;
        const uint32_t fx = star_facet_x(U), fy = star_facet_y(U), fz = star_facet_z(U);
        const uint32_t cand = S & kDim3 & (fx ^ fy ^ fz) & ~(fx & fy & fz);
        const uint32_t ready = Ud & star_has_cofacet_in(cand);
        if (ready) {
            const int a = __builtin_ctz(ready);
            const int c = pick_candidate<3>(k, a, cand & cofacet_mask(a));
            r.set_pair(a, c);
            P |= (1u << a) | (1u << c);
        } else {
            int best = __builtin_ctz(Ud);
            if (Ud & (Ud - 1)) {
                uint64_t bk = ~0ull;
                for (uint32_t m = Ud; m; m &= m - 1) {
                    const int t = __builtin_ctz(m);
                    const uint64_t kt = quad_lowest(k, t);
                    if (kt < bk) { bk = kt; best = t; }
                }
            }
            r.set_critical(1u << best);
            P |= 1u << best;
        }
    }
    // 3-cells left over have no cofacets: all critical
    r.set_critical(S & kDim3 & ~P);
}


// Product lower stars: one edge on each of n = 0..3 axes and every cell
// those edges span -- the vertex alone (a minimum), an edge, a quadrant, an
// octant; on smoothed fields about three in four interior vertices.
// product_shape returns has << 3 | neg (bit a of has: axis a carries an
// edge; bit a of neg: that edge points to -1) for those, kNotProduct for
// any other lower star.
constexpr uint32_t kNotProduct = 0xffu;
constexpr uint32_t kEdgeXM = 1u << 12, kEdgeXP = 1u << 14, kEdgeYM = 1u << 10, kEdgeYP = 1u << 16;
constexpr uint32_t kEdgeZM = 1u << 4, kEdgeZP = 1u << 22;

MSC_HD uint32_t product_shape(uint32_t L) {
    const uint32_t hx = (L & (kEdgeXM | kEdgeXP)) != 0u, hy = (L & (kEdgeYM | kEdgeYP)) != 0u;
    const uint32_t hz = (L & (kEdgeZM | kEdgeZP)) != 0u;
    const uint32_t nx = (L & kEdgeXP) == 0u, ny = (L & kEdgeYP) == 0u, nz = (L & kEdgeZP) == 0u;
    const uint32_t p = (hx ? (kX0 | (nx ? kXM : kXP)) : kX0) & (hy ? (kY0 | (ny ? kYM : kYP)) : kY0) &
                       (hz ? (kZ0 | (nz ? kZM : kZP)) : kZ0);
    return L == p ? ((hx | hy << 1 | hz << 2) << 3) | ((hx & nx) | (hy & ny) << 1 | (hz & nz) << 2) : kNotProduct;
}

// .grad byte of a paired cell of an interior star (assigned, pair code d,
// ldir preset 3), and of a critical vertex.
MSC_HD uint32_t paired_byte(uint32_t d) { return 0x61u | (d << 2); }
constexpr uint32_t kCriticalVertexByte = 0x7du;

MSC_HD int sel3(int a, int v0, int v1, int v2) { return a == 0 ? v0 : (a == 1 ? v1 : v2); }
MSC_HD uint64_t sel3(int a, uint64_t v0, uint64_t v1, uint64_t v2) { return a == 0 ? v0 : (a == 1 ? v1 : v2); }

// The relaxed expansion (expand_subset) of a product lower star in closed
// form, for a vertex whose star lies inside the domain:
//   the vertex pairs with the steepest edge e*;
//   quadrant: the other edge pairs with the quad;
//   octant: of the other two edges, the first in star order, a, has one
//     candidate, q(e*, a); the second, b, then has two, q(e*, b) and
//     q(a, b), and takes the one whose lowest vertex is lower (they cannot
//     tie: their only common vertices are the centre and b, and e* is below
//     b); the quad left pairs with the hex.
// Edge a comes before edge b in star order when a is on the lower axis and
// the edge on the higher axis points to +1, or else the other way round.
// g.key(o) is the order key of the vertex at tile offset o from the centre,
// g.put(o, byte) stores the byte of the star cell at staging offset o, and
// g.ts[a], g.cs[a] are the tile and staging strides of axis a.
template <class G>
MSC_HD void expand_product(uint32_t shape, const G& g) {
    const uint32_t has = shape >> 3, neg = shape & 7u;
    const int n = __builtin_popcount(has);
    const int t0 = (neg & 1u) ? -g.ts[0] : g.ts[0], t1 = (neg & 2u) ? -g.ts[1] : g.ts[1];
    const int t2 = (neg & 4u) ? -g.ts[2] : g.ts[2];
    const int c0 = (neg & 1u) ? -g.cs[0] : g.cs[0], c1 = (neg & 2u) ? -g.cs[1] : g.cs[1];
    const int c2 = (neg & 4u) ? -g.cs[2] : g.cs[2];
    auto code = [neg](int a) { return 2u * (uint32_t)a + ((neg >> a) & 1u); };
    if (n == 0) {
        g.put(0, kCriticalVertexByte);
    } else if (n == 1) {
        const int a = __builtin_ctz(has);
        const uint32_t d = code(a);
        g.put(0, paired_byte(d));
        g.put(sel3(a, c0, c1, c2), paired_byte(d ^ 1u));
    } else if (n == 2) {
        const int p = __builtin_ctz(has), q = 31 - __builtin_clz(has);
        const int tp = sel3(p, t0, t1, t2), tq = sel3(q, t0, t1, t2);
        const int cp = sel3(p, c0, c1, c2), cq = sel3(q, c0, c1, c2);
        const bool low = g.key(tp) < g.key(tq);  // p's edge is the steepest
        const uint32_t d = code(low ? p : q);
        g.put(0, paired_byte(d));
        g.put(low ? cp : cq, paired_byte(d ^ 1u));
        g.put(low ? cq : cp, paired_byte(d));
        g.put(cp + cq, paired_byte(d ^ 1u));
    } else {
        const uint64_t kx = g.key(t0), ky = g.key(t1), kz = g.key(t2);
        const int es = (kx < ky && kx < kz) ? 0 : (ky < kz ? 1 : 2);
        // the other two axes p < q; a is first in star order
        const int p = es == 0 ? 1 : 0, q = es == 2 ? 1 : 2;
        const bool pfirst = !((neg >> q) & 1u);
        const int a = pfirst ? p : q, b = pfirst ? q : p;
        const int te = sel3(es, t0, t1, t2), ta = sel3(a, t0, t1, t2), tb = sel3(b, t0, t1, t2);
        const uint64_t ke = sel3(es, kx, ky, kz), ka = sel3(a, kx, ky, kz), kb = sel3(b, kx, ky, kz);
        const uint64_t keb = g.key(te + tb), kab = g.key(ta + tb);
        const uint64_t l1 = ke < keb ? ke : keb;
        uint64_t l2 = ka < kb ? ka : kb;
        l2 = kab < l2 ? kab : l2;
        const bool take1 = l1 < l2;  // b pairs with q(e*, b)
        const int ce = sel3(es, c0, c1, c2), ca = sel3(a, c0, c1, c2), cb = sel3(b, c0, c1, c2);
        const uint32_t de = code(es), da = code(a);
        g.put(0, paired_byte(de));
        g.put(ce, paired_byte(de ^ 1u));
        g.put(ca, paired_byte(de));
        g.put(ce + ca, paired_byte(de ^ 1u));
        g.put(cb, paired_byte(take1 ? de : da));
        g.put(ce + cb, paired_byte(take1 ? de ^ 1u : da));
        g.put(ca + cb, paired_byte(take1 ? de : da ^ 1u));
        g.put(c0 + c1 + c2, paired_byte(take1 ? de ^ 1u : da ^ 1u));
    }
}

// Neighbours of the centre that come before it, as a star mask (bit 13
// clear).
// 2 acc + (a < b).  On the GPU the comparison lands in VCC and is shifted in
// as the carry of acc + acc: two instructions a neighbour.
MSC_HD uint32_t shift_in_less(uint32_t acc, uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    asm("v_cmp_lt_u64 vcc, %1, %2\n\tv_addc_co_u32 %0, vcc, %0, %0, vcc" : "+v"(acc) : "v"(a), "v"(b) : "vcc");
    return acc;
#else
    return 2u * acc + (a < b ? 1u : 0u);
#endif
}

template <class Keys>
MSC_HD uint32_t lower_neighbours(const Keys& k) {
    const uint64_t kc = k.at64(13);
    uint32_t lower = 0;  // built from bit 26 down
#pragma unroll
    for (int t = 26; t >= 0; t--) lower = t == 13 ? 2u * lower : shift_in_less(lower, k.at64(t), kc);
    return lower;
}

// Lower star from the lower neighbours: the cells whose every vertex other
// than the centre comes before it, less those touching `invalid` (star
// vertices outside the domain).
MSC_HD uint32_t lower_star_from(uint32_t lower, uint32_t invalid) {
    return ~star_dilate(((~lower) & kAll & ~kCenter) | invalid) & kAll;
}

// Star vertices of vertex (x,y,z) that fall outside an X*Y*Z grid.
MSC_HD uint32_t star_outside(int x, int y, int z, int X, int Y, int Z) {
    return (x == 0 ? kXM : 0u) | (x == X - 1 ? kXP : 0u) | (y == 0 ? kYM : 0u) | (y == Y - 1 ? kYP : 0u) |
           (z == 0 ? kZM : 0u) | (z == Z - 1 ? kZP : 0u);
}

// Lower star of vertex (x,y,z) of an X*Y*Z grid as a star mask: the cells
// whose every vertex other than the centre comes before it, restricted to
// the domain (gi_fast_robins_noalloc.h:359-366, gi_modified_robins.h:1293-1297).
template <class Keys>
MSC_HD uint32_t lower_star_mask(const Keys& k, int x, int y, int z, int X, int Y, int Z) {
    return lower_star_from(lower_neighbours(k), star_outside(x, y, z, X, Y, Z));
}

// Boundary strata of the star cells of vertex (x,y,z): bit t of b[a] is set
// when the cell lies on a domain face normal to axis a
// (gi_topological_regular_grid.h:1291-1303).
MSC_HD void star_boundary(int x, int y, int z, int X, int Y, int Z, uint32_t& bx, uint32_t& by, uint32_t& bz) {
    bx = (x == 0 || x == X - 1) ? kX0 : 0u;
    by = (y == 0 || y == Y - 1) ? kY0 : 0u;
    bz = (z == 0 || z == Z - 1) ? kZ0 : 0u;
}

// Steepest pass: subsets keyed by boundaryValue 0..3 (slots of <5,4>).
template <class Keys, class Res>
MSC_HD void steepest_expand(const Keys& k, uint32_t L, uint32_t bx, uint32_t by, uint32_t bz, Res& r) {
    const uint32_t onb = bx | by | bz;
    if (!onb) {
        expand_subset(L, k, r);
        return;
    }
    const uint32_t odd = bx ^ by ^ bz, two = (bx & by) | (bx & bz) | (by & bz);
    const uint32_t sub[4] = {~onb, odd & ~two, two & ~odd, odd & two};
    for (int i = 0; i < 4; i++) {
        const uint32_t S = L & sub[i];
        if (S) expand_subset(S, k, r);
    }
}

// Constrained pass <4,6>: subsets keyed restriction(c)*3 + boundaryValue(c),
// clamped to 23 (gi_modified_robins.h:1299-1313); label(t) returns the
// restriction label of star cell t.  Subsets are independent, so they are
// expanded in any order.
template <class Keys, class Labels, class Res>
MSC_HD void constrained_expand(const Keys& k, uint32_t L, uint32_t bx, uint32_t by, uint32_t bz,
                               const Labels& label, Res& r) {
    uint32_t pending = L;
    while (pending) {
        int kmin = 1 << 30;
        uint32_t S = 0;
        for (uint32_t m = pending; m; m &= m - 1) {
            const int t = __builtin_ctz(m);
            const int bvt = (int)((bx >> t) & 1u) + (int)((by >> t) & 1u) + (int)((bz >> t) & 1u);
            int key = (int)label(t) * 3 + bvt;
            key = key > 23 ? 23 : key;
            if (key < kmin) {
                kmin = key;
                S = 1u << t;
            } else if (key == kmin) {
                S |= 1u << t;
            }
        }
        pending &= ~S;
        expand_subset(S, k, r);
    }
}

// .grad byte of star cell t after the expansion: assigned, flag, pair code,
// ldir preset (3, or 0 for a critical 3-cell).
MSC_HD uint8_t star_cell_byte(const StarResult& r, uint32_t onb, int t) {
    const uint32_t code = ((r.c0 >> t) & 1u) | (((r.c1 >> t) & 1u) << 1) | (((r.c2 >> t) & 1u) << 2);
    const uint32_t flag = (onb >> t) & 1u;
    const uint32_t ldir = ((kDim3 >> t) & (r.crit >> t) & 1u) ? 0u : 3u;
    return (uint8_t)(1u | (flag << 1) | (code << 2) | (ldir << 5));
}

}  // namespace msceer

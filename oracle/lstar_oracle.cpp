// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference's lower-star discrete gradient
// (steepest_lstar path).  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load this library, and only as the checker or as the
// timed CPU baseline; the product path never calls it.
//
// Every function cites the reference file:line it restates.  Parity of this
// restatement is pinned against golden vectors produced by the reference's own
// classes (oracle/refdrv/ref_steepest.cpp) and against closed forms and
// invariants in tests/test_oracle_lstar.py.
//
// Layout conventions (gi_topological_regular_grid.h:113-126,1268-1278):
//   vertex grid X*Y*Z, vertex id x + X*y + X*Y*z;
//   cell complex Mx*My*Mz with M = 2N-1, cell id cx + Mx*cy + Mx*My*cz;
//   a cell's dimension is the number of odd coordinates.
#include <cstdint>
#include <cstring>
#include <queue>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef long long idx_t;

struct Mesh {
    idx_t X, Y, Z;      // vertices per axis
    idx_t M[3];         // cells per axis
    idx_t num_cells;
    Mesh(idx_t x, idx_t y, idx_t z) : X(x), Y(y), Z(z) {
        M[0] = 2 * x - 1; M[1] = 2 * y - 1; M[2] = 2 * z - 1;
        num_cells = M[0] * M[1] * M[2];
    }
    inline void coords(idx_t c, idx_t p[3]) const {
        p[0] = c % M[0]; p[1] = (c / M[0]) % M[1]; p[2] = c / (M[0] * M[1]);
    }
    inline idx_t id(const idx_t p[3]) const { return p[0] + M[0] * (p[1] + M[1] * p[2]); }
    inline int dim(const idx_t p[3]) const { return (int)((p[0] & 1) + (p[1] & 1) + (p[2] & 1)); }
    inline int dim(idx_t c) const { idx_t p[3]; coords(c, p); return dim(p); }
    // gi_topological_regular_grid.h:1291-1303 (non-periodic mesh)
    inline int bv(const idx_t p[3]) const {
        return (p[0] == 0 || p[0] == M[0] - 1) + (p[1] == 0 || p[1] == M[1] - 1) +
               (p[2] == 0 || p[2] == M[2] - 1);
    }
    inline int bv(idx_t c) const { idx_t p[3]; coords(c, p); return bv(p); }
    // gi_topological_regular_grid.h:1337-1341
    inline idx_t vertex_number(idx_t c) const {
        idx_t p[3]; coords(c, p);
        return p[0] / 2 + X * (p[1] / 2 + Y * (p[2] / 2));
    }
    inline idx_t off(int axis) const { return axis == 0 ? 1 : (axis == 1 ? M[0] : M[0] * M[1]); }
    // facets: odd axes, +axis then -axis, x,y,z order; never clipped
    // (gi_topological_regular_grid.h:281-306, FacetsIterator :413-500)
    inline int facets(idx_t c, idx_t out[6]) const {
        idx_t p[3]; coords(c, p);
        int n = 0;
        for (int a = 0; a < 3; a++)
            if (p[a] & 1) { out[n++] = c + off(a); out[n++] = c - off(a); }
        return n;
    }
    // cofacets: even axes, +axis (unless at M-1) then -axis (unless at 0)
    // (gi_topological_regular_grid.h:308-340, CofacetsIterator :502-591)
    inline int cofacets(idx_t c, idx_t out[6]) const {
        idx_t p[3]; coords(c, p);
        int n = 0;
        for (int a = 0; a < 3; a++)
            if (!(p[a] & 1)) {
                if (p[a] != M[a] - 1) out[n++] = c + off(a);
                if (p[a] != 0) out[n++] = c - off(a);
            }
        return n;
    }
    // gi_topological_regular_grid.h:1182-1192 (pair direction byte)
    inline unsigned char dir_byte(idx_t base, idx_t other) const {
        idx_t d = other - base;
        if (d == 1) return 0;
        if (d == -1) return 1;
        if (d == M[0]) return 2;
        if (d == -M[0]) return 3;
        if (d == M[0] * M[1]) return 4;
        if (d == -M[0] * M[1]) return 5;
        return 7;
    }
    inline idx_t dir_target(idx_t base, unsigned dir) const {
        switch (dir) {
            case 0: return base + 1;
            case 1: return base - 1;
            case 2: return base + M[0];
            case 3: return base - M[0];
            case 4: return base + M[0] * M[1];
            case 5: return base - M[0] * M[1];
        }
        return base;
    }
};

// Total order on vertices: a before b iff f[a] < f[b], ties by smaller id
// (gi_regular_grid_trilinear_function.h:309-314, gi_max_vertex_labeling.h:758-760).
inline bool before(const float* f, idx_t a, idx_t b) {
    if (f[b] > f[a]) return true;
    if (f[a] > f[b]) return false;
    return b > a;
}

inline int tid_of(int dx, int dy, int dz) { return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1); }

// Max / min vertex of every cell as a 27-tid byte (gi_max_vertex_labeling.h:1119-1339;
// the 3-slab sweep there computes exactly the argmax / argmin under the order above).
void maxmin_labels(const Mesh& m, const float* f, unsigned char* maxv, unsigned char* minv) {
#pragma omp parallel for schedule(static)
    for (idx_t cz = 0; cz < m.M[2]; cz++) {
        for (idx_t cy = 0; cy < m.M[1]; cy++) {
            for (idx_t cx = 0; cx < m.M[0]; cx++) {
                idx_t p[3] = {cx, cy, cz};
                idx_t c = m.id(p);
                idx_t best_hi = -1, best_lo = -1;
                int hi_tid = 13, lo_tid = 13;
                int nx = (cx & 1) ? 2 : 1, ny = (cy & 1) ? 2 : 1, nz = (cz & 1) ? 2 : 1;
                for (int kz = 0; kz < nz; kz++)
                    for (int ky = 0; ky < ny; ky++)
                        for (int kx = 0; kx < nx; kx++) {
                            int dx = (cx & 1) ? (kx ? 1 : -1) : 0;
                            int dy = (cy & 1) ? (ky ? 1 : -1) : 0;
                            int dz = (cz & 1) ? (kz ? 1 : -1) : 0;
                            idx_t vx = (cx + dx) / 2, vy = (cy + dy) / 2, vz = (cz + dz) / 2;
                            idx_t v = vx + m.X * (vy + m.Y * vz);
                            int t = tid_of(dx, dy, dz);
                            if (best_hi < 0 || before(f, best_hi, v)) { best_hi = v; hi_tid = t; }
                            if (best_lo < 0 || before(f, v, best_lo)) { best_lo = v; lo_tid = t; }
                        }
                maxv[c] = (unsigned char)hi_tid;
                minv[c] = (unsigned char)lo_tid;
            }
        }
    }
}

inline idx_t tid_offset(const Mesh& m, int t) {
    int dx = t % 3 - 1, dy = (t / 3) % 3 - 1, dz = t / 9 - 1;
    return dx + m.M[0] * (dy + m.M[1] * dz);
}

struct Expander {
    const Mesh& m;
    const float* f;
    const unsigned char* minv;
    unsigned char* grad;  // bit0 assigned, bits2-4 pair

    // one subset of a lower star; ids in star order
    idx_t ids[27];
    int nm[27];
    idx_t pair[27];
    bool paired[27];
    int n;

    Expander(const Mesh& mm, const float* ff, const unsigned char* mn, unsigned char* g)
        : m(mm), f(ff), minv(mn), grad(g), n(0) {}

    int find(idx_t c) const {
        for (int i = 0; i < n; i++) if (ids[i] == c) return i;
        return -1;
    }
    idx_t lowest_vertex(idx_t c) const { return m.vertex_number(c + tid_offset(m, minv[c])); }
    void dec_cofacets(idx_t c) {
        idx_t cf[6]; int k = m.cofacets(c, cf);
        for (int i = 0; i < k; i++) { int j = find(cf[i]); if (j >= 0) nm[j]--; }
    }
    // first candidate with the strictly smallest lowest vertex
    // (gi_modified_robins.h:774-790, gi_fast_robins_noalloc.h:124-137)
    int pick_lowest(const int* cand, int nc) const {
        int best = cand[0];
        idx_t lv = lowest_vertex(ids[best]);
        for (int i = 1; i < nc; i++) {
            idx_t o = lowest_vertex(ids[cand[i]]);
            if (before(f, o, lv)) { lv = o; best = cand[i]; }
        }
        return best;
    }

    // Relaxed Robins homotopy expansion of one subset
    // (gi_modified_robins.h:999-1199; same rules as gi_fast_robins_noalloc.h:147-321).
    void expand(const idx_t* cells, int count) {
        n = count;
        int lists[4][27], ln[4] = {0, 0, 0, 0};
        for (int i = 0; i < n; i++) { ids[i] = cells[i]; pair[i] = -1; paired[i] = false; nm[i] = 0; }
        for (int i = 0; i < n; i++) {
            int d = m.dim(ids[i]);
            lists[d][ln[d]++] = i;
            idx_t fc[6]; int k = m.facets(ids[i], fc);
            for (int j = 0; j < k; j++) if (find(fc[j]) >= 0) nm[i]++;
        }
        if (ln[0] > 0) {
            int vi = lists[0][0];
            idx_t vid = ids[vi];
            if (ln[1] == 0) {
                pair[vi] = vid; paired[vi] = true; dec_cofacets(vid);
            } else {
                // steepest edge: the edge whose other vertex comes first
                // (gi_modified_robins.h:1062-1104)
                int best = lists[1][0];
                idx_t best_v = lowest_vertex(ids[best]);
                for (int i = 1; i < ln[1]; i++) {
                    idx_t o = lowest_vertex(ids[lists[1][i]]);
                    if (before(f, o, best_v)) { best_v = o; best = lists[1][i]; }
                }
                pair[vi] = ids[best]; paired[vi] = true;
                pair[best] = vid; paired[best] = true;
                dec_cofacets(vid); dec_cofacets(ids[best]);
            }
        }
        for (int d = 0; d < 4; d++) {
            int todo = 0, done = 0;
            for (int j = 0; j < ln[d]; j++) if (!paired[lists[d][j]]) todo++;
            while (done < todo) {
                bool progressed = false;
                for (int j = 0; j < ln[d] && !progressed; j++) {
                    int a = lists[d][j];
                    if (paired[a]) continue;
                    int cand[6], nc = 0;
                    idx_t cf[6]; int k = m.cofacets(ids[a], cf);
                    for (int i = 0; i < k; i++) {
                        int ci = find(cf[i]);
                        if (ci >= 0 && nm[ci] == 1) cand[nc++] = ci;
                    }
                    if (nc > 0) {
                        int c = pick_lowest(cand, nc);
                        pair[a] = ids[c]; paired[a] = true;
                        pair[c] = ids[a]; paired[c] = true;
                        dec_cofacets(ids[a]); dec_cofacets(ids[c]);
                        done++;
                        progressed = true;
                    }
                }
                if (!progressed) {
                    int cand[27], nc = 0;
                    for (int j = 0; j < ln[d]; j++) if (!paired[lists[d][j]]) cand[nc++] = lists[d][j];
                    int c = pick_lowest(cand, nc);
                    pair[c] = ids[c]; paired[c] = true;
                    dec_cofacets(ids[c]);
                    done++;
                }
            }
        }
        // write back (gi_modified_robins.h:1326-1340)
        for (int i = 0; i < n; i++) {
            idx_t c = ids[i];
            unsigned char dir = (pair[i] == c) ? 7 : m.dir_byte(c, pair[i]);
            grad[c] = (unsigned char)((grad[c] & ~0x1D) | 1 | (dir << 2));
            if (pair[i] != c) {
                idx_t q = pair[i];
                unsigned char back = m.dir_byte(q, c);
                grad[q] = (unsigned char)((grad[q] & ~0x1D) | 1 | (back << 2));
            }
        }
    }
};

// Lower star of one vertex split into subsets keyed label*3 + boundaryValue
// (gi_modified_robins.h:1281-1343).  slots = LMAXDIM*LMAX_LABEL (20 for the
// steepest pass <5,4>, 24 for the constrained pass <4,6>).
void lower_star(const Mesh& m, const float* f, const unsigned char* maxv, const unsigned char* minv,
                const signed char* restriction, int slots, idx_t vcell, unsigned char* grad) {
    idx_t subsets[24][27];
    int sn[24];
    for (int i = 0; i < slots; i++) sn[i] = 0;
    idx_t vp[3]; m.coords(vcell, vp);
    for (int t = 0; t < 27; t++) {
        int d[3] = {t % 3 - 1, (t / 3) % 3 - 1, t / 9 - 1};
        idx_t p[3];
        bool ok = true;
        for (int a = 0; a < 3; a++) { p[a] = vp[a] + d[a]; if (p[a] < 0 || p[a] >= m.M[a]) ok = false; }
        if (!ok) continue;
        idx_t c = m.id(p);
        if (maxv[c] != 26 - t) continue;  // highest vertex of c is not this vertex
        int key = (restriction ? (int)restriction[c] : 0) * 3 + m.bv(p);
        if (key > slots - 1) key = slots - 1;
        subsets[key][sn[key]++] = c;
    }
    Expander ex(m, f, minv, grad);
    for (int i = 0; i < slots; i++)
        if (sn[i] > 0) ex.expand(subsets[i], sn[i]);
}

// ldir pass (gi_topological_gradient_using_algorithms.h:395-478).
void ldir_pass(const Mesh& m, unsigned char* g) {
    const idx_t C = m.num_cells;
    auto critical = [&](idx_t c) { return ((g[c] >> 2) & 7) == 7; };
    auto assigned = [&](idx_t c) { return (g[c] & 1) != 0; };
    auto getl = [&](idx_t c) { return (int)(g[c] >> 5); };
    auto setl = [&](idx_t c, int v) { g[c] = (unsigned char)((g[c] & 0x1F) | (v << 5)); };
    std::vector<idx_t> crit[4];
    for (idx_t c = 0; c < C; c++) {
        setl(c, 3);
        if (critical(c)) crit[m.dim(c)].push_back(c);
    }
    auto trace = [&](idx_t start, int value) {
        std::queue<idx_t> q;
        q.push(start);
        while (!q.empty()) {
            idx_t item = q.front(); q.pop();
            if (getl(item) == value) continue;
            setl(item, value);
            idx_t cf[6]; int k = m.cofacets(item, cf);
            for (int i = 0; i < k; i++) {
                idx_t t = cf[i];
                if (critical(t) || !assigned(t)) continue;
                idx_t p = m.dir_target(t, (g[t] >> 2) & 7);
                if (p == item) continue;
                if (getl(t) == value) continue;
                setl(t, value);
                q.push(p);
            }
        }
    };
    for (idx_t c : crit[1]) trace(c, 2);
    for (idx_t c : crit[2]) trace(c, 1);
    for (idx_t c : crit[3]) setl(c, 0);
}

}  // namespace

extern "C" {

int oracle_version(void) { return 1; }

// 27-tid bytes of the highest / lowest vertex of every cell.
int oracle_maxmin(const float* f, long long X, long long Y, long long Z,
                  unsigned char* maxv, unsigned char* minv) {
    if (X < 2 || Y < 2 || Z < 2) return -1;
    Mesh m(X, Y, Z);
    maxmin_labels(m, f, maxv, minv);
    return 0;
}

// Lower-star pairing of every vertex (assigned + pair bits only; flag/ldir 0).
// restriction may be NULL (steepest pass, slots 20) or a per-cell label
// (constrained pass, slots 24).  Vertex traversal order does not matter:
// each vertex writes only the cells of its own lower star.
int oracle_lstar_pairs(const float* f, long long X, long long Y, long long Z,
                       const signed char* restriction, unsigned char* grad) {
    if (X < 2 || Y < 2 || Z < 2) return -1;
    Mesh m(X, Y, Z);
    std::vector<unsigned char> maxv(m.num_cells), minv(m.num_cells);
    maxmin_labels(m, f, maxv.data(), minv.data());
    memset(grad, 0, (size_t)m.num_cells);
    int slots = restriction ? 24 : 20;
#pragma omp parallel for schedule(dynamic, 1)
    for (idx_t vz = 0; vz < Z; vz++)
        for (idx_t vy = 0; vy < Y; vy++)
            for (idx_t vx = 0; vx < X; vx++) {
                idx_t p[3] = {2 * vx, 2 * vy, 2 * vz};
                lower_star(m, f, maxv.data(), minv.data(), restriction, slots, m.id(p), grad);
            }
    return 0;
}

// Recompute the lower star of the listed vertex cells only (constrained pass).
int oracle_lstar_vertices(const float* f, long long X, long long Y, long long Z,
                          const signed char* restriction, const long long* vcells, long long n,
                          unsigned char* grad) {
    Mesh m(X, Y, Z);
    std::vector<unsigned char> maxv(m.num_cells), minv(m.num_cells);
    maxmin_labels(m, f, maxv.data(), minv.data());
    int slots = restriction ? 24 : 20;
    for (long long i = 0; i < n; i++)
        lower_star(m, f, maxv.data(), minv.data(), restriction, slots, vcells[i], grad);
    return 0;
}

int oracle_ldir(unsigned char* grad, long long X, long long Y, long long Z) {
    Mesh m(X, Y, Z);
    ldir_pass(m, grad);
    return 0;
}

// flag bit = boundaryValue != 0 (gi_discrete_gradient_labeling.h:200-226)
int oracle_flags(unsigned char* grad, long long X, long long Y, long long Z) {
    Mesh m(X, Y, Z);
    for (idx_t c = 0; c < m.num_cells; c++) {
        int b = m.bv(c) != 0;
        grad[c] = (unsigned char)((grad[c] & ~2) | (b << 1));
    }
    return 0;
}

// Whole steepest_lstar .grad (steepest_lstar.cxx:129-258).
int oracle_steepest_grad(const float* f, long long X, long long Y, long long Z, unsigned char* grad) {
    int rc = oracle_lstar_pairs(f, X, Y, Z, nullptr, grad);
    if (rc) return rc;
    oracle_ldir(grad, X, Y, Z);
    oracle_flags(grad, X, Y, Z);
    return 0;
}

int oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // extern "C"

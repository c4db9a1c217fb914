// ORACLE -- test infrastructure only.
//
// CPU restatement of what extractmsc reads out of a .grad before any
// simplification (extractmsc/extractmsc.cxx:467-556): the Morse-Smale
// complex's nodes, one per critical cell in cell-id order
// (MorseSmaleComplexBasic::ComputeFromGrad, gi_morse_smale_complex_basic.h:676-690),
// the base basins (FillAscendingIndexVolumeByVPaths, extractmsc.cxx:371-440:
// every vertex follows its V-path down -- vertex, its paired edge, the edge's
// other vertex -- to a critical vertex, and takes that minimum's node id) and
// the base mountains (fillGeometry of every maximum, descending, painted on
// the lower corner vertex of each hex: rec_man_trace_down,
// gi_morse_smale_complex_basic.h:1181-1198, reaches a hex h' from a hex h when
// h' is paired with a facet quad of h).  Only tests/ load this library.
//
// .grad byte: bit 0 assigned, bits 2-4 pair direction (0..5 = +x,-x,+y,-y,
// +z,-z; 7 = critical).
#include <cstdint>
#include <vector>

namespace {

struct Cells {
    long long X, Y, Z, Mx, My, Mz;
    long long id(long long x, long long y, long long z) const { return x + Mx * (y + My * z); }
};

bool critical(unsigned char b) { return ((b >> 2) & 7u) == 7u; }

// Cell reached from (x, y, z) by pair direction d, `steps` cells on.
void step(unsigned d, long long steps, long long& x, long long& y, long long& z) {
    const long long s = (d & 1u) ? -steps : steps;
    if (d >> 1 == 0) x += s;
    else if (d >> 1 == 1) y += s;
    else z += s;
}

}  // namespace

extern "C" {

// basins, mounts: X*Y*Z int32; counts: critical cells by dimension (4).
// Returns the number of nodes.
long long oracle_msc_base(const unsigned char* grad, long long X, long long Y, long long Z, int* basins, int* mounts,
                          long long* counts) {
    const Cells c{X, Y, Z, 2 * X - 1, 2 * Y - 1, 2 * Z - 1};
    const long long cells = c.Mx * c.My * c.Mz;
    // node ids: rank among the critical cells in cell-id order
    std::vector<int> node(cells, -1);
    long long n = 0;
    for (int k = 0; k < 4; k++) counts[k] = 0;
    for (long long i = 0; i < cells; i++) {
        if (!critical(grad[i])) continue;
        node[i] = (int)n++;
        const long long x = i % c.Mx, y = (i / c.Mx) % c.My, z = i / (c.Mx * c.My);
        counts[(x & 1) + (y & 1) + (z & 1)]++;
    }
    // basins: V-paths down from every vertex, painted back along the path
    const long long V = X * Y * Z;
    for (long long v = 0; v < V; v++) basins[v] = -1;
    std::vector<long long> path;
    for (long long v = 0; v < V; v++) {
        if (basins[v] != -1) continue;
        long long x = 2 * (v % X), y = 2 * ((v / X) % Y), z = 2 * (v / (X * Y));
        path.clear();
        int dest = -1;
        for (;;) {
            const long long vid = x / 2 + X * (y / 2 + Y * (z / 2));
            path.push_back(vid);
            if (basins[vid] != -1) {
                dest = basins[vid];
                break;
            }
            const long long cid = c.id(x, y, z);
            if (critical(grad[cid])) {
                dest = node[cid];
                break;
            }
            step((grad[cid] >> 2) & 7u, 2, x, y, z);
            if (x < 0 || y < 0 || z < 0 || x >= c.Mx || y >= c.My || z >= c.Mz) break;  // not a gradient: -1
        }
        for (long long p : path) basins[p] = dest;
    }
    // mountains: every hex follows its pairing up to a critical hex (its
    // paired facet quad's other hex), or to a quad on the domain boundary
    for (long long v = 0; v < V; v++) mounts[v] = -1;
    const long long H = (X - 1) * (Y - 1) * (Z - 1);
    std::vector<int> hex_dest(H > 0 ? H : 0, -2);  // -2: not yet known
    auto hex_index = [&](long long x, long long y, long long z) { return (x / 2) + (X - 1) * ((y / 2) + (Y - 1) * (z / 2)); };
    std::vector<long long> hpath;
    for (long long h = 0; h < H; h++) {
        if (hex_dest[h] != -2) continue;
        long long x = 2 * (h % (X - 1)) + 1, y = 2 * ((h / (X - 1)) % (Y - 1)) + 1, z = 2 * (h / ((X - 1) * (Y - 1))) + 1;
        hpath.clear();
        int dest = -1;
        for (;;) {
            const long long hi = hex_index(x, y, z);
            if (hex_dest[hi] != -2) {
                dest = hex_dest[hi];
                break;
            }
            hpath.push_back(hi);
            const long long cid = c.id(x, y, z);
            if (critical(grad[cid])) {
                dest = node[cid];
                break;
            }
            step((grad[cid] >> 2) & 7u, 2, x, y, z);
            if (x < 0 || y < 0 || z < 0 || x >= c.Mx || y >= c.My || z >= c.Mz) {
                dest = -1;  // paired with a boundary quad: no maximum reaches it
                break;
            }
        }
        for (long long p : hpath) hex_dest[p] = dest;
    }
    for (long long h = 0; h < H; h++) {
        const long long hx = h % (X - 1), hy = (h / (X - 1)) % (Y - 1), hz = h / ((X - 1) * (Y - 1));
        mounts[hx + X * (hy + Y * hz)] = hex_dest[h];
    }
    return n;
}

}  // extern "C"

// Host emulation of the lower-star kernel's per-lane logic (test
// infrastructure): the same lstar_core.hpp the gfx950 kernel runs, driven
// over every vertex on the CPU, so kernel-logic changes are checked against
// the oracle before they reach a GPU.
#include <cstdint>
#include <cstring>
#include <vector>

#include "lstar_core.hpp"

using namespace msceer;

extern "C" int emul_lstar_prepass(const float* f, long long X, long long Y, long long Z,
                                  const int8_t* restriction, uint8_t* grad) {
    const int64_t PX = X + 2, PY = Y + 2, PZ = Z + 2;
    std::vector<uint32_t> keys((size_t)(PX * PY * PZ));
    for (int64_t z = 0; z < PZ; z++)
        for (int64_t y = 0; y < PY; y++)
            for (int64_t x = 0; x < PX; x++) {
                int64_t gx = x - 1, gy = y - 1, gz = z - 1;
                gx = gx < 0 ? 0 : (gx >= X ? X - 1 : gx);
                gy = gy < 0 ? 0 : (gy >= Y ? Y - 1 : gy);
                gz = gz < 0 ? 0 : (gz >= Z ? Z - 1 : gz);
                keys[(size_t)(x + PX * (y + PY * z))] = float_key(f[gx + X * (gy + Y * gz)]);
            }
    const int64_t Mx = 2 * X - 1, My = 2 * Y - 1;
    int soff[27];  // .grad offset of star cell t (the kernel's stage offsets, here in the .grad itself)
    for (int t = 0; t < 27; t++) soff[t] = (int)(tid_x(t) + Mx * (tid_y(t) + My * tid_z(t)));
#pragma omp parallel for schedule(dynamic, 1)
    for (long long z = 0; z < Z; z++)
        for (long long y = 0; y < Y; y++)
            for (long long x = 0; x < X; x++) {
                StridedKeys k{keys.data(), (x + 1) + PX * ((y + 1) + PY * (z + 1)), PX, PX * PY};
                const uint32_t L = lower_star_mask(k, (int)x, (int)y, (int)z, (int)X, (int)Y, (int)Z);
                uint32_t bx, by, bz;
                star_boundary((int)x, (int)y, (int)z, (int)X, (int)Y, (int)Z, bx, by, bz);
                const int64_t vcell = 2 * x + Mx * (2 * y + My * 2 * z);
                // the kernel's byte sink: every cell of L written as it is paired
                StageBytes r{grad + vcell, soff, bx | by | bz};
                if (!restriction) {
                    steepest_expand(k, L, bx, by, bz, r);
                } else {
                    const int8_t* rv = restriction + vcell;
                    auto label = [rv, Mx, My](int t) -> int {
                        return rv[tid_x(t) + Mx * tid_y(t) + Mx * My * tid_z(t)];
                    };
                    constrained_expand(k, L, bx, by, bz, label, r);
                }
            }
    return 0;
}

// Every interior vertex: the closed-form expansion of product lower stars
// (expand_product, through star-bit offsets) against the general expansion
// (expand_subset) and its bytes.  counts[0]: vertices the closed form took,
// counts[1]: of those, stars whose bytes differ or whose written cells are
// not exactly the lower star.
namespace {
struct StarBytes {
    const StridedKeys* k;
    uint8_t* out;     // 27 bytes by star bit
    uint32_t* wrote;  // star bits written
    int ts[3] = {1, 3, 9}, cs[3] = {1, 3, 9};
    uint64_t key(int o) const { return k->at64(13 + o); }
    void put(int o, uint32_t b) const {
        out[13 + o] = (uint8_t)b;
        *wrote |= 1u << (13 + o);
    }
};
}  // namespace

extern "C" int emul_check_product(const float* f, long long X, long long Y, long long Z, long long* counts) {
    const int64_t PX = X + 2, PY = Y + 2, PZ = Z + 2;
    std::vector<uint32_t> keys((size_t)(PX * PY * PZ));
    for (int64_t z = 0; z < PZ; z++)
        for (int64_t y = 0; y < PY; y++)
            for (int64_t x = 0; x < PX; x++) {
                int64_t gx = x - 1, gy = y - 1, gz = z - 1;
                gx = gx < 0 ? 0 : (gx >= X ? X - 1 : gx);
                gy = gy < 0 ? 0 : (gy >= Y ? Y - 1 : gy);
                gz = gz < 0 ? 0 : (gz >= Z ? Z - 1 : gz);
                keys[(size_t)(x + PX * (y + PY * z))] = float_key(f[gx + X * (gy + Y * gz)]);
            }
    long long taken = 0, differ = 0;
#pragma omp parallel for reduction(+ : taken, differ) schedule(dynamic, 1)
    for (long long z = 1; z < Z - 1; z++)
        for (long long y = 1; y < Y - 1; y++)
            for (long long x = 1; x < X - 1; x++) {
                StridedKeys k{keys.data(), (x + 1) + PX * ((y + 1) + PY * (z + 1)), PX, PX * PY};
                const uint32_t L = lower_star_mask(k, (int)x, (int)y, (int)z, (int)X, (int)Y, (int)Z);
                const uint32_t shape = product_shape(L);
                if (shape == kNotProduct) continue;
                taken++;
                uint8_t out[27] = {0};
                uint32_t wrote = 0;
                expand_product(shape, StarBytes{&k, out, &wrote});
                StarResult r;
                expand_subset(L, k, r);
                bool bad = wrote != L;
                for (int t = 0; t < 27; t++)
                    if ((L >> t) & 1u) bad |= out[t] != star_cell_byte(r, 0u, t);
                differ += bad;
            }
    counts[0] = taken;
    counts[1] = differ;
    return 0;
}

// Lower star of every vertex (star mask) and its product shape
// (kNotProduct for vertices on the domain boundary), for shape statistics.
extern "C" int emul_lower_stars(const float* f, long long X, long long Y, long long Z, uint32_t* L, uint8_t* shape) {
    const int64_t PX = X + 2, PY = Y + 2, PZ = Z + 2;
    std::vector<uint32_t> keys((size_t)(PX * PY * PZ));
    for (int64_t z = 0; z < PZ; z++)
        for (int64_t y = 0; y < PY; y++)
            for (int64_t x = 0; x < PX; x++) {
                int64_t gx = x - 1, gy = y - 1, gz = z - 1;
                gx = gx < 0 ? 0 : (gx >= X ? X - 1 : gx);
                gy = gy < 0 ? 0 : (gy >= Y ? Y - 1 : gy);
                gz = gz < 0 ? 0 : (gz >= Z ? Z - 1 : gz);
                keys[(size_t)(x + PX * (y + PY * z))] = float_key(f[gx + X * (gy + Y * gz)]);
            }
#pragma omp parallel for schedule(dynamic, 1)
    for (long long z = 0; z < Z; z++)
        for (long long y = 0; y < Y; y++)
            for (long long x = 0; x < X; x++) {
                StridedKeys k{keys.data(), (x + 1) + PX * ((y + 1) + PY * (z + 1)), PX, PX * PY};
                const uint32_t l = lower_star_mask(k, (int)x, (int)y, (int)z, (int)X, (int)Y, (int)Z);
                const bool onb = x == 0 || y == 0 || z == 0 || x == X - 1 || y == Y - 1 || z == Z - 1;
                const size_t i = (size_t)(x + X * (y + Y * z));
                L[i] = l;
                shape[i] = (uint8_t)(onb ? kNotProduct : product_shape(l));
            }
    return 0;
}

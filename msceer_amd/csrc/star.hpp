// Bit algebra over the star of a vertex in the cubical cell complex.
//
// The 27 cells that contain vertex v sit at cell offsets d in {-1,0,1}^3 from
// v's cell; bit t = (dx+1) + 3(dy+1) + 9(dz+1) of a 27-bit mask stands for the
// cell at offset d.  This is the 27-adjacency order of the reference
// (gi_topological_regular_grid.h:196-205), so "first in star order" is
// "lowest set bit".  A cell's vertices are v + s for the sub-offsets s of d
// (s_a in {0, d_a}); the cell at offset d and the neighbour vertex at offset d
// share the same bit, which is what the dilations below exploit.
#pragma once
#include <stdint.h>

// Functions usable from kernels and from host-side emulation of the kernels.
#if defined(__HIPCC__)
#define MSC_HD __host__ __device__ __forceinline__
#else
#define MSC_HD inline
#endif

namespace msceer {

constexpr uint32_t star_mask_where(int axis, int val) {
    uint32_t m = 0;
    for (int t = 0; t < 27; t++) {
        int c = axis == 0 ? t % 3 - 1 : (axis == 1 ? (t / 3) % 3 - 1 : t / 9 - 1);
        if (c == val) m |= 1u << t;
    }
    return m;
}
constexpr uint32_t star_dim_mask(int dim) {
    uint32_t m = 0;
    for (int t = 0; t < 27; t++) {
        int d = (t % 3 != 1) + ((t / 3) % 3 != 1) + (t / 9 != 1);
        if (d == dim) m |= 1u << t;
    }
    return m;
}

constexpr uint32_t kAll = (1u << 27) - 1;
constexpr uint32_t kCenter = 1u << 13;
constexpr uint32_t kXP = star_mask_where(0, 1), kXM = star_mask_where(0, -1), kX0 = star_mask_where(0, 0);
constexpr uint32_t kYP = star_mask_where(1, 1), kYM = star_mask_where(1, -1), kY0 = star_mask_where(1, 0);
constexpr uint32_t kZP = star_mask_where(2, 1), kZM = star_mask_where(2, -1), kZ0 = star_mask_where(2, 0);
constexpr uint32_t kDim0 = star_dim_mask(0), kDim1 = star_dim_mask(1);
constexpr uint32_t kDim2 = star_dim_mask(2), kDim3 = star_dim_mask(3);

static_assert(kDim0 == kCenter, "center");
static_assert(__builtin_popcount(kDim1) == 6 && __builtin_popcount(kDim2) == 12 &&
              __builtin_popcount(kDim3) == 8, "star cell counts");

// Cells whose vertex set meets NL (a mask of "bad" neighbour vertices):
// dilate from zero components to +-1 along each axis.
MSC_HD uint32_t star_dilate(uint32_t nl) {
    nl |= ((nl & kX0) << 1) | ((nl & kX0) >> 1);
    nl |= ((nl & kY0) << 3) | ((nl & kY0) >> 3);
    nl |= ((nl & kZ0) << 9) | ((nl & kZ0) >> 9);
    return nl;
}

// For every cell at bit t: is its facet along x (the cell with the x
// component zeroed) set in u?  Only cells with a nonzero x component answer.
MSC_HD uint32_t star_facet_x(uint32_t u) { return ((u << 1) & kXP) | ((u >> 1) & kXM); }
MSC_HD uint32_t star_facet_y(uint32_t u) { return ((u << 3) & kYP) | ((u >> 3) & kYM); }
MSC_HD uint32_t star_facet_z(uint32_t u) { return ((u << 9) & kZP) | ((u >> 9) & kZM); }

// For every cell at bit t: does any cofacet of it (one zero component moved
// to +-1) lie in c?
MSC_HD uint32_t star_has_cofacet_in(uint32_t c) {
    return (((c >> 1) | (c << 1)) & kX0) | (((c >> 3) | (c << 3)) & kY0) |
           (((c >> 9) | (c << 9)) & kZ0);
}

// Pair-direction byte of the step from bit a to bit b (b - a in {+-1,+-3,+-9}):
// +x 0, -x 1, +y 2, -y 3, +z 4, -z 5 (gi_topological_regular_grid.h:1182-1192).
MSC_HD uint32_t star_dir_code(int a, int b) {
    int d = b - a;
    int ad = d < 0 ? -d : d;
    uint32_t base = ad == 1 ? 0u : (ad == 3 ? 2u : 4u);
    return base + (d < 0 ? 1u : 0u);
}

}  // namespace msceer

// Accurate 1-manifolds (lines.hip): digitized streamlines from the saddles.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "internal.hpp"

struct msceer_ctx;

namespace msceer {

struct LineStats {
    uint64_t saddles = 0;     // critical 2-cells (ASC) or 1-cells (DSC)
    uint64_t traced = 0;      // streamlines traced, speculation and replays included
    uint64_t replayed = 0;    // streamlines re-traced in order inside a component
    uint64_t voxels = 0;      // voxel traversals of all those traces
    uint64_t windows = 0;     // speculation windows
    uint64_t components = 0;  // components replayed (lone lines included)
    uint64_t cuts = 0;        // windows cut short by a replay that left its footprint
    uint64_t direct = 0;      // lines traced alone (first of a window and beyond the footprint pool)
    uint64_t overflows = 0;   // windows cut short where the footprint pool ran out
    uint64_t fail_tests = 0;  // replays stopped by a test of another component's cell
    uint64_t fail_labels = 0; // replays stopped by a label their speculation did not give
};

// The digitized 1-manifolds of one family added onto `restr` (device, one
// int8 per cell): asc = from the 2-saddles up (value 4), else from the
// 1-saddles down (value 5).  grad: the steepest .grad (pairs of the lower-star
// pass); gradd: the FP64 gradient field the streamlines follow.
int run_lines(msceer_ctx* ctx, const Dims& d, const float* values, const uint8_t* grad, const double4* gradd,
              int8_t* restr, bool asc, float gthr, float ethr, int max_iters, LineStats* out);

}  // namespace msceer

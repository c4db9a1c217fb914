/*
 * msceer.h -- C ABI of libmsceer_hip.so, the MI355X implementation of MSCEER's
 * lower-star discrete gradient (steepest_lstar) and accurate labelling
 * (ondemandaccurate) hot path.
 *
 * The reference exposes this path only as C++ templates instantiated in the
 * CLI programs; each entry point below names the reference interface it
 * replaces.  Conventions:
 *   - plain pointers and sizes; the caller owns every buffer passed in;
 *   - return 0 on success or a negative MSCEER_E* code (never throws/exits);
 *   - one context per host thread; calls on one context are serialised on
 *     its HIP stream;
 *   - volumes are float32, x-fastest (vertex id x + X*y + X*Y*z); the cell
 *     complex has (2X-1)(2Y-1)(2Z-1) cells, cell id cx + Mx*cy + Mx*My*cz
 *     (gi_topological_regular_grid.h:113-126,1268-1278);
 *   - input values must be free of NaN (the reference's order is not a total
 *     order in their presence).
 */
#ifndef MSCEER_H
#define MSCEER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSCEER_VERSION 1

/* error codes */
#define MSCEER_OK 0
#define MSCEER_EINVAL (-1)      /* bad argument (sizes, null pointers, flags) */
#define MSCEER_EHIP (-2)        /* a HIP runtime call failed */
#define MSCEER_ENOMEM (-3)      /* device allocation failed */
#define MSCEER_ETOOBIG (-4)     /* volume exceeds the supported cell count (< 2^32 cells) */
#define MSCEER_ENODEV (-5)      /* no usable GPU */

/* flags */
#define MSCEER_DEVICE_PTRS 0x1u /* buffers are device pointers (else host memory) */
#define MSCEER_SKIP_LDIR 0x2u   /* leave ldir bits at their pre-pass values */
#define MSCEER_TIMING 0x4u      /* record per-kernel HIP event timings */

typedef struct msceer_ctx msceer_ctx;

/* Per-call timings (milliseconds, HIP events on the context stream); filled
 * when MSCEER_TIMING is set.  counts[] hold critical cells by dimension. */
typedef struct msceer_timing {
    float lstar_ms;        /* lower-star pairing kernel */
    float ldir_ms;         /* ascending-manifold-dimension (ldir) passes */
    float total_ms;        /* whole device-side call */
    int32_t ldir_rounds;   /* frontier rounds used by the ldir passes */
    int64_t critical[4];   /* critical cells by dimension */
} msceer_timing;

int msceer_version(void);
const char* msceer_strerror(int code);

/* Context: owns a HIP stream (default priority) and grow-only device scratch. */
int msceer_create(msceer_ctx** ctx, int device);
int msceer_destroy(msceer_ctx* ctx);
/* Run subsequent work on an external stream (hipStream_t); NULL restores the
 * context's own stream. */
int msceer_set_stream(msceer_ctx* ctx, void* hip_stream);
int msceer_last_timing(const msceer_ctx* ctx, msceer_timing* out);

/*
 * Full steepest_lstar .grad of a volume: one byte per cell, LSB first
 * assigned(1) flag(1) pair(3) ldir(3) (gi_discrete_gradient_labeling.h:11-19).
 * Replaces the sequence in steepest_lstar/steepest_lstar.cxx:129-258:
 *   RegularGridMaxMinVertexLabeling3D::ComputeOutput (gi_max_vertex_labeling.h:1119),
 *   SlidingWindowRobinsNoalloc::ComputePairing_sliding (gi_fast_robins_noalloc.h:1254),
 *   TopologicalGradientUsingAlgorithms::setAscendingManifoldDimensions
 *     (gi_topological_gradient_using_algorithms.h:430),
 *   DiscreteGradientLabeling::outputToFile flag pass (gi_discrete_gradient_labeling.h:200).
 * grad_out holds (2nx-1)(2ny-1)(2nz-1) bytes.
 */
int msceer_lstar_grad(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                      uint8_t* grad_out, uint32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* MSCEER_H */

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
#define MSCEER_STOP_REGIONS 0x8u     /* msceer_ondemand_asc: return the certain-region labels */
#define MSCEER_STOP_INTEGRATED 0x10u /* msceer_ondemand_asc: return labels before the cleanup */
/* msceer_ondemand_asc, msceer_integrate_labels: re-integrate (pass 3) in FP32
 * first and in FP64 only the streamlines whose FP32 trace meets a near-tie
 * decision (within 2^-16 voxel of a voxel face or half-integer, or an error or
 * gradient test within 2^-14 relative of its threshold) or uses up its step
 * budget; default: pass 3 wholly in FP64.  No counterpart in the reference,
 * whose advector has one fixed mixed precision. */
#define MSCEER_FP32_FIRST 0x20u
/* msceer_lstar_grad: the .grad file of the tiled program
 * (save_discrete_gradient_to_file1, steepest_lstar_tiled/
 * localized_morse_complex.h:1063-1133): pairs and the boundary flag, the
 * ascending-manifold dimension (bits 5-7) left 0; ldir is not run */
#define MSCEER_TILED_LAYOUT 0x40u

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

/* Statistics of one accurate-labelling call (msceer_ondemand_asc). */
typedef struct msceer_label_stats {
    int64_t minima;             /* critical vertices of the base gradient */
    int64_t saddles;            /* critical 1-cells of the base gradient */
    int64_t regions;            /* regions after persistence simplification */
    int64_t traced_fp32;        /* pass-1 streamlines (FP32) */
    int64_t over_budget_fp32;   /* of those, stopped by the step budget */
    int64_t reintegrated_fp64;  /* pass-3 re-integrations traced in FP64 (with MSCEER_FP32_FIRST: those escalated) */
    int64_t over_budget_fp64;   /* of those, stopped by the step budget */
    int64_t integrated_vertices;/* vertices the learned mask sent to integration (all if none) */
    int32_t certain_sweeps;     /* rounds of the certain-region resolution (dependency depth) */
    int32_t pass3_rounds;       /* launches of pass 3 (one persistent launch) */
    float ms_lstar;             /* base + constrained lower-star kernels */
    float ms_regions;           /* extremum graph + certain regions */
    float ms_integrate;         /* passes 1-3 */
    float ms_cleanup;           /* isolated-region cleanup + restriction labels */
    float ms_grad;              /* constrained pass + ldir */
    float ms_total;
    int64_t steps_fp32;         /* FP32 advection steps (accepted and rejected), passes 1 and 3 */
    int64_t steps_fp64;         /* FP64 advection steps (pass 3) */
    int64_t ahead_fp64;         /* pass-3 streamlines traced ahead of a chain of long re-integrations */
    float ms_pass1;             /* integration pass 1 (FP32, all levels) */
    float ms_pass3;             /* integration pass 3 */
    int64_t reintegrated_pass3;  /* pass-3 re-integrations (all of them, whatever their precision) */
    int64_t escalated_fp64;     /* with MSCEER_FP32_FIRST: of those, left in doubt by the FP32 trace and traced again in FP64 */
    int64_t lines_asc;          /* msceer_ondemand, needASC1: streamlines traced from the 2-saddles (speculation included) */
    int64_t lines_dsc;          /* needDSC1: streamlines traced from the 1-saddles */
    int64_t lines_replayed;     /* of the 1-manifold streamlines, re-traced in order inside a component */
    int32_t line_windows;       /* speculation windows of the 1-manifold stage */
    int32_t line_cuts;          /* windows cut short by a replay that left its component */
    float ms_lines;             /* the 1-manifold stage */
    int32_t line_alone;         /* 1-manifold streamlines traced alone, in order */
    int32_t line_overflows;     /* windows cut short where the footprint pool ran out */
    int64_t maxima;             /* needDSC2: critical hexes of the steepest pairing */
    int64_t saddles2;           /* needDSC2: critical quads */
    int64_t regions_dsc;        /* needDSC2: descending certain regions after simplification */
    int32_t certain_sweeps_dsc; /* needDSC2: certain-region rounds */
    float ms_dsc;               /* needDSC2: the whole descending branch (beside the ascending one with needASC2) */
    float ms_dsc_wait;          /* needDSC2: time its end lies past the ascending branch's */
} msceer_label_stats;

int msceer_version(void);
const char* msceer_strerror(int code);

/* Context: owns a HIP stream (default priority) and grow-only device scratch. */
int msceer_create(msceer_ctx** ctx, int device);
int msceer_destroy(msceer_ctx* ctx);
/* Run subsequent work on an external stream (hipStream_t); NULL restores the
 * context's own stream. */
int msceer_set_stream(msceer_ctx* ctx, void* hip_stream);
int msceer_last_timing(const msceer_ctx* ctx, msceer_timing* out);
/* Lower-star brick autotuning (no counterpart in the reference, which has no
 * GPU launch shape): mode 1 (default) times the four brick shapes 64x2x2,
 * 32x4x2, 16x4x4, 8x8x4 on the first volume of each x extent and keeps the
 * fastest; mode 0 always uses 64x2x2, mode 2+k always shape k.  All shapes
 * give identical bytes. */
int msceer_set_autotune(msceer_ctx* ctx, int32_t mode);
/* Brick shape (0..3 as above) used for x extent nx; -1 while not yet tuned. */
int msceer_lstar_brick(msceer_ctx* ctx, int64_t nx);

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

/*
 * msceer_lstar_grad without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  Device pointers only (values in,
 * grad_out out); every scratch buffer (control words, critical-cell lists,
 * frontiers, ldir records) comes from one static 48 MiB array per GPU, so a
 * volume whose scratch would exceed it returns MSCEER_ETOOBIG (up to about
 * 1.3 million vertices fit, e.g. 128x128x64).  The lower-star brick shape is
 * the one an earlier msceer_lstar_grad tuned for this x extent, else 64x2x2;
 * no timing runs.  Flags: MSCEER_SKIP_LDIR, MSCEER_TIMING.  All calls on one
 * GPU share its array, so the library runs them one at a time; each call
 * clears what it needs before use.
 */
int msceer_lstar_grad_noalloc(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                              uint8_t* grad_out, uint32_t flags);

/*
 * Accurate ascending 3-manifold labelling and its conforming .grad: the
 * needASC2 branch of ondemandaccurate (ondemandaccurate/ondemandaccurate.cxx:728-1061):
 *   base gradient (as msceer_lstar_grad, without ldir),
 *   SimplifiedExtremumGraph::ComputeMinMapFromGradient (gi_extrema_region_builder.h:406-551),
 *   GridSimplifiedExtremalRegionBuilder::BeginIntegration (gi_extrema_region_builder.h:982-1042),
 *   NumericIntegratorPathCompressingToTerminal::BeginIntegration
 *     (gi_numeric_integrator_path_compressing.h:75-433) with
 *     AdaptiveEulerAdvector3D<F,-1> (gi_adaptive_euler_advector.h:51-305),
 *   IsolatedCCRegionRemoverNEW::ComputeOutput (gi_isolated_region_remover.h:155-192),
 *   VertexLabelingToBoundaryLabeling::ComputeMINBoundary (gi_labeling_to_bounary_labeling.h:48-116),
 *   the constrained MyRobinsNoalloc<...,4,6> pass (gi_modified_robins.h:1281-1343)
 *   and RecordGrad (ondemandaccurate.cxx:693-699).
 * err_thr bounds the squared step-doubling error, grad_thr the gradient
 * magnitude, max_iters the steps (accepted and rejected) per streamline;
 * presimp is the persistence threshold of the minimum graph.  Pass 1 traces
 * in FP32; pass 3 re-traces in FP64 every vertex next to a label change and
 * every pass-1 streamline that exhausted its budget.
 * labels_out (X*Y*Z int32, may be NULL): region of every vertex, named by the
 * vertex id of the region's lowest minimum.  grad_out (cells bytes, may be
 * NULL): the conforming .grad.  integrate_mask (X*Y*Z bytes, may be NULL):
 * when given, only vertices with a nonzero mask are integrated numerically
 * and every other vertex takes the label of the minimum its discrete-gradient
 * V-path descends to; those vertices are trusted, so streamlines of the
 * integrated vertices (passes 1 and 3) also stop on reaching one.
 */
int msceer_ondemand_asc(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz, float err_thr,
                        float grad_thr, int32_t max_iters, float presimp, const uint8_t* integrate_mask,
                        uint8_t* grad_out, int32_t* labels_out, msceer_label_stats* stats, uint32_t flags);

/* Accuracy branches of OnDemandAccurate (argv needASC1 needDSC1 needASC2,
 * ondemandaccurate.cxx:158-185): bits of `need` for msceer_ondemand.
 * needDSC2 (descending 3-manifolds) is not provided. */
#define MSCEER_NEED_ASC1 0x1u /* ascending 1-manifolds: ReIntegrateUpFrom2Saddles */
#define MSCEER_NEED_DSC1 0x2u /* descending 1-manifolds: ReIntegrateDownFrom1Saddles */
#define MSCEER_NEED_ASC2 0x4u /* ascending 3-manifolds: msceer_ondemand_asc's labelling */
#define MSCEER_NEED_DSC2 0x8u /* descending 3-manifolds: the max-side labelling and its MAX boundary */

/*
 * OnDemandAccurate with any nonzero combination of needASC1, needDSC1 and
 * needASC2 (ondemandaccurate.cxx:704-1157).  Replaces, beside what
 * msceer_ondemand_asc replaces:
 *   DigitizingNumericStreamlineIntegrator3dASC / 3dDSC
 *   (gi_numeric_streamline_integrator.h:526-1072) driven by
 *   ReIntegrateUpFrom2Saddles / ReIntegrateDownFrom1Saddles
 *   (ondemandaccurate.cxx:254-500, 1075-1091): from every critical 2-cell
 *   (critical 1-cell), in the order one reference thread takes them,
 *   streamlines uphill (downhill) from the seed cells of a short V-path
 *   search, digitized into the restriction labels as 4 (5), each stopping on
 *   a cell an earlier one labelled.  The GPU keeps that order exactly.
 * The streamlines run in FP64 on the gradient field, zeroed at the region
 * extrema when needASC2 ran first (as the reference's shared field is).
 * restriction_out (cells int8, may be NULL): the restriction labels the
 * conforming gradient was computed under (MIN boundary 1 when needASC2,
 * 4 / 5 on the digitized 1-manifolds).  labels_out and integrate_mask need
 * needASC2; the STOP flags need needASC2 or needDSC2 (msceer_ondemand2), end
 * every labelling asked for at that stage and return its labels alone.
 * Other arguments as msceer_ondemand_asc.
 */
int msceer_ondemand(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz, float err_thr,
                    float grad_thr, int32_t max_iters, float presimp, uint32_t need, const uint8_t* integrate_mask,
                    uint8_t* grad_out, int32_t* labels_out, int8_t* restriction_out, msceer_label_stats* stats,
                    uint32_t flags);
/*
 * msceer_ondemand with the descending labels too: labels_dsc_out (X*Y*Z
 * int32, may be NULL; needs MSCEER_NEED_DSC2) receives the needDSC2 labels
 * (ondemandaccurate.cxx:913-991: the max extremum graph of the steepest
 * pairing, GridSimplifiedExtremalRegionBuilder<ComparerDSC>,
 * NumericIntegratorPathCompressingToTerminal<AdaptiveEulerAdvector3D<F, 1>>,
 * IsolatedCCRegionRemoverNEW<ComparerDSC>), named by the highest vertex of
 * each region; their MAX boundary (ComputeMAXBoundary) enters the
 * restriction.  The gradient field is shared as in the reference: after
 * needASC2 the descending integration sees it zeroed at the ascending seeds
 * too.  The descending labels follow the ascending ones' arithmetic: certain
 * regions bit-exact, the streamlines FP32 first and FP64 where in doubt.
 */
int msceer_ondemand2(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz, float err_thr,
                     float grad_thr, int32_t max_iters, float presimp, uint32_t need, const uint8_t* integrate_mask,
                     uint8_t* grad_out, int32_t* labels_out, int32_t* labels_dsc_out, int8_t* restriction_out,
                     msceer_label_stats* stats, uint32_t flags);

/*
 * The numeric integrator alone on given terminals: replaces
 * NumericIntegratorPathCompressingToTerminal::BeginIntegration(remap, terminals)
 * (gi_numeric_integrator_path_compressing.h:75-433) with
 * AdaptiveEulerAdvector3D<F,sign> (gi_adaptive_euler_advector.h:51-305).
 * terminals (X*Y*Z int32): -1 = integrate this vertex, any other value = a
 * certain terminal carrying that label.  zero_grad (X*Y*Z bytes, may be NULL):
 * vertices whose gradient is zero (the seeds of the remap argument).
 * sign -1 flows downhill (ascending manifolds), +1 uphill (descending).
 * labels_out (X*Y*Z int32): the label every vertex reaches; pass 1 in FP32,
 * pass 3 in FP64 as in msceer_ondemand_asc; no component cleanup.  Flags:
 * MSCEER_DEVICE_PTRS (all arrays in device memory), MSCEER_TIMING.
 * stats (may be NULL): the integrator fields and ms_integrate.
 */
int msceer_integrate_labels(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                            const int32_t* terminals, const uint8_t* zero_grad, float err_thr, float grad_thr,
                            int32_t max_iters, int32_t sign, int32_t* labels_out, msceer_label_stats* stats,
                            uint32_t flags);

/*
 * Batch of independent volumes (no counterpart in the reference, which
 * processes one volume per program run): for each volume, the lower-star
 * .grad (msceer_lstar_grad) when labels_out is NULL, else the accurate
 * labelling and conforming .grad (msceer_ondemand_asc with the shared
 * parameters; grad_out may then be NULL).  `streams` worker threads (0: 16)
 * each drive a context of their own on the same GPU and take the volumes
 * largest first; the lower-star brick tuning is shared.  HIP maps streams
 * onto GPU_MAX_HW_QUEUES hardware queues (4 unless set before the first HIP
 * call); with fewer queues than workers, a long kernel of one worker delays
 * the queued work of another (16 and 16 measured best).  Host buffers only.
 * Per volume, ms receives the wall time of its pipeline and status its code.
 */
typedef struct msceer_volume {
    const float* values;
    int64_t nx, ny, nz;
    uint8_t* grad_out;
    int32_t* labels_out;
    float ms;
    int32_t status;
} msceer_volume;
int msceer_batch(msceer_ctx* ctx, msceer_volume* vols, int32_t n, int32_t streams, float err_thr, float grad_thr,
                 int32_t max_iters, float presimp, uint32_t flags);

/*
 * Batched lower-star .grad (no counterpart in the reference, which processes
 * one volume per program run): for every volume the .grad of
 * msceer_lstar_grad (labels_out must be NULL), bit for bit.  Consecutive
 * volumes form groups of fewer than 2^32 - 1 cells in all and, if
 * max_group_verts > 0, at most that many vertices (a single larger volume is
 * a group of its own).  Each group runs as one launch per stage for all its
 * volumes (lower star, ldir records, the two ldir passes' frontier rounds on
 * two streams, finish), captured once in a hipGraph and replayed on later
 * calls with the same shapes (and, with device pointers, the same
 * addresses); the lower-star launch (brick shape; a work-group per brick, or
 * persistent work-groups claiming bricks from per-XCD queues in chunks) is
 * tuned on a group's first call unless msceer_set_autotune fixed the shape.
 * One host synchronisation per group.  Flags: MSCEER_DEVICE_PTRS (values and
 * grad_out of every volume in device memory; else host memory, staged
 * through the context's device buffers), MSCEER_TIMING (per-volume device
 * clock: ms = from the first brick of the volume's group to the volume's
 * last finished tile; without it ms is the group's time).  status: 0 once
 * the volume's group has run.
 */
int msceer_batch_grad(msceer_ctx* ctx, msceer_volume* vols, int32_t n, int64_t max_group_verts, uint32_t flags);
typedef struct msceer_batch_stats {
    int32_t groups;        /* groups the last call ran */
    int32_t replays;       /* graph launches */
    int32_t captures;      /* graphs captured (first calls, or after a plan changed) */
    int32_t eager_rounds;  /* frontier rounds and sweeps run after a replay left work */
    int32_t shape;         /* lower-star brick shape (last group), as msceer_lstar_brick */
    int32_t blocks;        /* persistent work-groups, 0: a work-group per brick */
    int32_t chunk;         /* bricks per queue claim */
    int32_t rounds;        /* frontier rounds per pass in the graph */
    int64_t critical[4];   /* critical cells by dimension, all volumes */
    float ms_gpu;          /* device time of the groups (events around each replay and what followed) */
} msceer_batch_stats;
int msceer_last_batch_stats(const msceer_ctx* ctx, msceer_batch_stats* out);

/*
 * The .grad consumer of extractmsc before any simplification
 * (extractmsc/extractmsc.cxx:467-556): Morse-Smale complex nodes = critical
 * cells in cell-id order (MorseSmaleComplexBasic::ComputeFromGrad,
 * gi_morse_smale_complex_basic.h:676-690); basins_out (X*Y*Z int32, may be
 * NULL): per vertex the node id of the minimum its V-path descends to
 * (FillAscendingIndexVolumeByVPaths, extractmsc.cxx:371-440; TESTBASINS.raw);
 * mounts_out (X*Y*Z int32, may be NULL): per vertex the node id of the
 * maximum whose descending manifold holds the hex at that vertex
 * (fillGeometry of the maxima, gi_morse_smale_complex_basic.h:1181-1198;
 * TESTMOUNTS.raw), -1 where there is none (the last vertex layers, hexes
 * paired with a boundary quad).  counts_out (4 int64, may be NULL): critical
 * cells by dimension.  Needs fewer than 2^31 vertices.  Flags:
 * MSCEER_DEVICE_PTRS.  Arcs and the persistence hierarchy
 * (ComputeHierarchy) are not computed.
 */
int msceer_msc_base(msceer_ctx* ctx, const uint8_t* grad, int64_t nx, int64_t ny, int64_t nz, int32_t* basins_out,
                    int32_t* mounts_out, int64_t* counts_out, uint32_t flags);

/*
 * NumericIntegratorExpandingRegionStop<AdaptiveEulerAdvector3D<F, -1>,
 * IndexCompareLessThan<F>>::BeginIntegration
 * (gi_numeric_integrator_expanding_region_stop.h:204-440; no reference
 * program instantiates it): the ascending-manifold labels of the minima of
 * the SameBdry6 rule (no neighbour before the vertex along the axes where it
 * is interior).  Certain regions grow from each minimum over the vertices
 * whose earlier neighbours all lie in it; every other vertex is integrated
 * (budget max_iters) until it nears a certain vertex; vertices whose
 * neighbours all share their label become certain, and the rest are
 * integrated again with 500 iterations.  labels_out (X*Y*Z int32): the vertex
 * id of the minimum, or -1.  The streamlines are traced in FP64, every vertex
 * alone (the reference compresses pass-1 paths in its threads' visiting
 * order, which makes its own result schedule-dependent).  stats_out (5 int64,
 * may be NULL): minima, certain vertices, pass-1 streamlines, unambiguous
 * vertices, re-integrated vertices.  Flags: MSCEER_DEVICE_PTRS.
 */
int msceer_expanding_region_stop(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                                 float err_thr, float grad_thr, int32_t max_iters, int32_t* labels_out,
                                 int64_t* stats_out, uint32_t flags);

/* Stages of the accurate pipeline as separate calls (host or device pointers):
 * IsolatedCCRegionRemoverNEW::ComputeOutput (gi_isolated_region_remover.h:155-192),
 * in place on X*Y*Z int32 labels. */
int msceer_clean_labels(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                        int32_t* labels_inout, uint32_t flags);
/* VertexLabelingToBoundaryLabeling::ComputeMINBoundary (gi_labeling_to_bounary_labeling.h:48-116):
 * int8 restriction label per cell (0 or 1) from X*Y*Z int32 labels. */
int msceer_restriction_labels(msceer_ctx* ctx, const int32_t* labels, int64_t nx, int64_t ny, int64_t nz,
                              int8_t* restriction_out, uint32_t flags);
/* VertexLabelingToBoundaryLabeling::ComputeRegionBoundaryKind
 * (gi_labeling_to_bounary_labeling.h:262-393), the `_edges_` file of
 * Accurate3DMorseLabeling (accurate3dmorselabeling.cxx:337-343): int8 per
 * cell from X*Y*Z int32 labels -- edges 1 where their vertices' valid labels
 * differ; quads 0, 1, or 2 where more than two regions meet; hexes up to 3.
 * Vertices, and edges with an unlabelled vertex, which the reference leaves
 * unwritten, are 0. */
int msceer_region_boundary_kind(msceer_ctx* ctx, const int32_t* labels, int64_t nx, int64_t ny, int64_t nz,
                                int8_t* kind_out, uint32_t flags);
/*
 * VertexLabelingToBoundaryLabeling::ComputeMAXBoundary
 * (gi_labeling_to_bounary_labeling.h:118-195) added onto restriction_inout
 * (cells bytes; e.g. the output of msceer_restriction_labels): +2 on every
 * interior quad whose two hexes have different highest vertices with
 * different valid labels, then on interior edges with such a quad among
 * their cofacets, then on interior vertices with such an edge.  labels: the
 * descending-manifold label of every vertex (-1 = none).
 */
int msceer_restriction_add_max(msceer_ctx* ctx, const float* values, const int32_t* labels, int64_t nx, int64_t ny,
                               int64_t nz, int8_t* restriction_inout, uint32_t flags);

/* Constrained lower-star pass MyRobinsNoalloc<...,4,6>::ComputeLowerStar over the
 * vertices of ondemandaccurate.cxx:1014-1052 (recomputing every lower star is
 * equivalent), then RecordGrad (ldir + flags) unless MSCEER_SKIP_LDIR. */
int msceer_constrained_lstar(msceer_ctx* ctx, const float* values, int64_t nx, int64_t ny, int64_t nz,
                             const int8_t* restriction, uint8_t* grad_inout, uint32_t flags);

/*
 * Near-separatrix classifier (learned integration mask).  No counterpart in
 * the reference: a small 3-D CNN flags the vertices whose accurate label may
 * differ from the combinatorial one, and msceer_ondemand_asc integrates only
 * those (integrate_mask).  Network: the values normalised over each vertex's
 * 9x9x9 window, three 3x3x3 convolutions (1->16->32->16, ReLU; BF16 weights
 * and activations, FP32 accumulation), a 16->1 sigmoid head.  Parameters are
 * a flat FP32 vector of msceer_cnn_num_params() entries.
 */
typedef struct msceer_cnn msceer_cnn;
int msceer_cnn_num_params(void);
/* He-initialised network from a seed (SplitMix64 + Box-Muller). */
int msceer_cnn_create(msceer_ctx* ctx, uint64_t seed, msceer_cnn** net);
int msceer_cnn_destroy(msceer_cnn* net);
int msceer_cnn_get_params(msceer_ctx* ctx, msceer_cnn* net, float* params_out);
/* Replaces the parameters and resets the optimiser state. */
int msceer_cnn_set_params(msceer_ctx* ctx, msceer_cnn* net, const float* params);
/* Per-vertex probability of lying near a separatrix (X*Y*Z floats). */
int msceer_cnn_predict(msceer_ctx* ctx, msceer_cnn* net, const float* values, int64_t nx, int64_t ny, int64_t nz,
                       float* prob_out, uint32_t flags);
/* Mean weighted binary cross-entropy against target (X*Y*Z bytes, nonzero =
 * near a separatrix; positives weighted pos_weight) and its gradient with
 * respect to the parameters (grad_out may be NULL; the gradient is kept for
 * msceer_cnn_adam_step). */
int msceer_cnn_gradient(msceer_ctx* ctx, msceer_cnn* net, const float* values, int64_t nx, int64_t ny, int64_t nz,
                        const uint8_t* target, float pos_weight, float* loss_out, float* grad_out, uint32_t flags);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) with the last gradient. */
int msceer_cnn_adam_step(msceer_ctx* ctx, msceer_cnn* net, float lr);
/* Integration mask: probability > threshold, dilated by `dilation` vertices
 * (Chebyshev distance); flagged (may be NULL) receives the mask's count. */
int msceer_cnn_mask(msceer_ctx* ctx, msceer_cnn* net, const float* values, int64_t nx, int64_t ny, int64_t nz,
                    float threshold, int32_t dilation, uint8_t* mask_out, int64_t* flagged, uint32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* MSCEER_H */

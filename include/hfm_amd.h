/*
 * hfm_amd.h -- C ABI of the MI355X (gfx950) Fast-Marching library libhfm_amd.so.
 *
 * Drop-in boundary for HamiltonFastMarching's CPU solve (Mirebeau/HamiltonFastMarching,
 * conda hfm 0.2.6).  Two layers:
 *
 *  1. The key/value IO object and hfm_run(), which replace the reference's
 *        void Run(IO&)                         Headers/DispatchAndRun.h:224-353
 *     as called by FileHFM (Interfaces/FileHFM/FileHFM.cpp:12-26) and by the pybind module
 *     HFMIO (Interfaces/PythonHFM/src/main.cpp:11-47).  Keys, shapes, array orderings and
 *     outputs follow the reference (JMM_CPPLibs/JMM_CPPLibs/Output/IO.h:77-118,
 *     Headers/Base/Implementation/HFMInterface.hxx:296-579).  The model is chosen at run time
 *     from the "model" key (Riemann2, Riemann3, ReedsSheppForward2, Elastica2) instead of at
 *     compile time (DispatchAndRun.h:12,227-244).
 *
 *  2. A device-resident plan (hfm_plan_*) for callers that keep fields in HBM: stencils,
 *     values and scratch stay on the GPU between calls.  It has no single reference
 *     counterpart; it splits HFMInterface::Run into its phases (Run_SetupStencil :343-352,
 *     Run_RunSolver :422-450, Run_ExportData :526-579, FirstVariation::Finally
 *     Headers/ExtraAlgorithms/FirstVariation.h:60-102) over device pointers.
 *
 * Conventions: every function returns 0 on success and a nonzero code on failure; no C++
 * exception crosses this boundary.  The message of the last failure is returned by
 * hfm_io_last_error() for IO calls and by hfm_last_error() (per thread) for plan calls;
 * it carries the reference's wording where the reference raises the same error
 * (JMM_CPPLibs/JMM_CPPLibs/Macros/Exception.h:9-19).  Arrays crossing the IO layer are FP64
 * like the reference's RawElement (BaseIO.hxx:10-33) and are copied in and out.  Pointers
 * documented as "device" are HIP device pointers; "stream" is a hipStream_t (NULL = default).
 */
#ifndef HFM_AMD_H
#define HFM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------------------------------------- */
/* IO dictionary (replaces IO_<Base>: IO.h:77-118; pybind HFMIO main.cpp:35-47)              */
/* ----------------------------------------------------------------------------------------- */
typedef struct hfm_io hfm_io;

/* new IO object (IO_<FileIO>/IO_<PythonIO> construction) */
hfm_io *hfm_io_create(void);
void hfm_io_destroy(hfm_io *io);
/* set_string (IO.h:98 SetString; main.cpp:37) */
int hfm_io_set_string(hfm_io *io, const char *key, const char *value);
/* set_scalar (IO.h:103 Set<ScalarType>; main.cpp:39) */
int hfm_io_set_scalar(hfm_io *io, const char *key, double value);
/* set_array (IO.h:105 SetArray; main.cpp:41): row-major data of the given shape as stored */
int hfm_io_set_array(hfm_io *io, const char *key, const double *data, int ndim, const int64_t *shape);
/* set_array from FP32 data (the same key store: values are widened to FP64, the reference's
 * only numerical type, RawElement in BaseIO.hxx:10-33) */
int hfm_io_set_array_f32(hfm_io *io, const char *key, const float *data, int ndim, const int64_t *shape);
/* has_field (IO.h:92 HasField; main.cpp:43): 1 if present, 0 otherwise */
int hfm_io_has_field(const hfm_io *io, const char *key);
/* erase_field (main.cpp:45) */
int hfm_io_erase_field(hfm_io *io, const char *key);
/* 0: absent, 1: string, 2: scalar or array */
int hfm_io_field_kind(const hfm_io *io, const char *key);
/* get_string (IO.h:97 GetString; main.cpp:36): copies into buf (NUL-terminated, truncated to
 * bufsize); returns the full length, or -1 if missing / not a string */
long long hfm_io_get_string(hfm_io *io, const char *key, char *buf, size_t bufsize);
/* get_scalar (IO.h:88 Get<ScalarType>; main.cpp:38) */
int hfm_io_get_scalar(hfm_io *io, const char *key, double *value);
/* shape of an array field (ndim = 0 for a scalar); shape must hold max_ndim entries */
int hfm_io_get_array_shape(hfm_io *io, const char *key, int *ndim, int64_t *shape, int max_ndim);
/* get_array (IO.h:91 GetArray; main.cpp:40): copies count values (row-major as stored) */
int hfm_io_get_array(hfm_io *io, const char *key, double *out, size_t count);
/* computed_keys (main.cpp:42): space-separated keys written by hfm_run; returns full length */
long long hfm_io_computed_keys(hfm_io *io, char *buf, size_t bufsize);
const char *hfm_io_last_error(const hfm_io *io);

/* The FileHFM wire format (JMM_CPPLibs/JMM_CPPLibs/Output/FileIO.hxx:7-132): read the input
 * pair <prefix>_Format.txt / <prefix>_Data.dat into the IO object (replaces FileIO::FileIO), and
 * write its computed keys as an output pair (replaces FileIO::~FileIO).  0 or 1
 * (hfm_io_last_error: the reference's messages). */
int hfm_io_read_files(hfm_io *io, const char *prefix);
int hfm_io_write_files(hfm_io *io, const char *prefix);

/* Messages of the last hfm_run at the requested verbosity (the reference prints them with
 * Msg(), HFMInterface.hxx:442-447 and elsewhere). */
const char *hfm_io_messages(const hfm_io *io);

/* Run the solve described by io (Run(IO&), Headers/DispatchAndRun.h:224-353).  Synchronous. */
int hfm_run(hfm_io *io);

/* ----------------------------------------------------------------------------------------- */
/* Device-resident plan (phases of HFMInterface::Run over device pointers)                     */
/* ----------------------------------------------------------------------------------------- */
typedef struct hfm_plan hfm_plan;

typedef struct hfm_plan_stats {
    int64_t iterations;     /* block-FIM rounds of the last solve */
    int64_t tile_visits;    /* tiles processed (64 points each) over the last solve */
    int64_t sweeps;         /* local sweeps summed over tile visits */
    int64_t max_stencil_width; /* MaxStencilWidth (HFMInterface.hxx:357) */
    int64_t n_points;       /* grid points */
    int64_t n_tiles;        /* tiles of 64 points */
    int64_t n_deps;         /* entries of the tile dependency graph */
    double stencil_ms;      /* device time of the last stencil phase (hipEvent) */
    double solve_ms;        /* device time of the last solve (hipEvent) */
    double fim_ms;          /* summed duration of the update-kernel launches of the last solve
                               (hipEvent pairs around each launch on the plan's stream) */
    int64_t fim_launches;   /* update-kernel launches of the last solve */
    int64_t tiles_carried;  /* listed tiles deferred to a later band (not processed) */
    double delta;           /* band width used by the last solve (0: unordered) */
    int64_t adjoint_rounds; /* rounds of the last backward sweep */
    int64_t forward_rounds; /* rounds of the last forward-variation sweep */
    int64_t n_accepted;     /* serial acceptance loop (order 2 / 3, stopping criteria): nAccepted as
                               the reference counts it (the pop that triggered a stop included);
                               -1 after a block fast-iterative solve (every reached point) */
} hfm_plan_stats;

/* model: "Riemann2" | "Riemann3" | "ReedsSheppForward2" | "Elastica2".
 * dims: grid extents (2 or 3).  gridScales: one per axis for Riemann (Riemannian.h:46-61,
 * HFM_ParamDefault.hxx:68-102); gridScales[0] for the curvature models (the theta step is
 * 2 pi / dims[2], Curvature2.h:104-108).  origin: NULL or one per axis (only the two physical
 * components for the curvature models).  precision: 32 or 64 (update arithmetic).
 * xi, eps: curvature-model parameters (Curvature2.h:104-108; ignored for Riemann). */
int hfm_plan_create(hfm_plan **out, const char *model, const int64_t *dims, const double *gridScales,
                    const double *origin, int precision, double xi, double eps);
void hfm_plan_destroy(hfm_plan *plan);

/* Riemann: metric (device, FP64, row-major (dims..., d(d+1)/2) or one tensor if !is_field),
 * primal ("metric") or dual ("dualMetric").  Builds the per-point Selling stencils
 * (StencilRiemann::SetStencil, Riemannian.h:41-44) and the tile dependency graph.
 * Fails with the reference's message on a non positive-definite tensor. */
int hfm_plan_set_metric(hfm_plan *plan, const double *d_metric, int is_field, int dual, void *stream);

/* Curvature models: cost (device, FP64, row-major dims or one value if !is_field);
 * speed = 1/cost (HFM_StencilShare.hxx:61-71). */
int hfm_plan_set_cost(hfm_plan *plan, const double *d_cost, int is_field, void *stream);

/* Walls (WallObstructionTest.h): d_walls = device bytes in row-major point order, nonzero = wall,
 * or NULL to remove them; refine = refineStencilAtWallBoundary.  A difference from an accepted
 * point to the updated one is used only if the discrete segment between them meets no wall
 * (Visible, WallObstructionTest.h:96-150; the distance to the walls is capped at
 * MaxStencilWidth).  Applies to the next solves and to everything computed from them (flow,
 * activeNeighs, geodesics, adjoint, forward variation); kept across set_metric / set_cost. */
int hfm_plan_set_walls(hfm_plan *plan, const uint8_t *d_walls, int refine, void *stream);

/* Solve from seeds (host, physical coordinates, nseeds x dim; seedValues NULL -> 0), the
 * seed placement of StaticFactoring::SetSeeds (StaticFactoring.hxx:242-393).  Runs the
 * block fast-iterative method to its fixed point on the plan's stream. */
int hfm_plan_solve(hfm_plan *plan, int nseeds, const double *seeds, const double *seedValues, void *stream);

/* Solver state from cached results instead of a solve (the bypass of Run_RunSolver,
 * HFMInterface.hxx:423-426): d_values (arrival times) and d_active (activeNeighs bitsets as
 * doubles), both device, FP64, row-major.  A reached point without active differences is taken
 * as a seed.  Geodesics, flow, activeOffsets, adjoint and forward variation then run on these
 * values; their active sets are recomputed from the values (they equal d_active when both come
 * from a solve of the same problem). */
int hfm_plan_load_values(hfm_plan *plan, const double *d_values, const double *d_active, void *stream);

/* Arrival times of the last solve, row-major FP64 (device) -- the 'values' output. */
int hfm_plan_values(hfm_plan *plan, double *d_values, void *stream);

/* Allocation-free Riemann solve.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- tight serving loops, code around HIP graph capture,
 * harnesses that forbid hipMalloc.  It computes what hfm_plan_create (precision 32) +
 * hfm_plan_set_metric + hfm_plan_solve + hfm_plan_values compute, for Riemann2 / Riemann3,
 * from the device metric (row-major FP64, as hfm_plan_set_metric) into d_values (row-major FP64
 * on the device), and calls no hipMalloc, hipFree, hipHostMalloc, event or graph creation:
 * every working buffer (stencils, tile graph, values, keys, lists, counters) lives in one static
 * device workspace of HFM_NOALLOC_WORKSPACE_BYTES, reserved on each GPU the library is loaded on
 * until the process ends.  Limits (an error, never an allocation, beyond them): at most
 * HFM_NOALLOC_MAX_TILES tiles of 64 points (1024 x 1024 in 2-D, 100^3 in 3-D), 4096 seeds, and a
 * tile dependency graph that fits the rest of the workspace.  Calls on one GPU share its
 * workspace: the library serializes them (one call at a time per device, each returning after
 * its work on `stream` has completed), so they never overlap inside one process. */
#define HFM_NOALLOC_MAX_TILES 16384
#define HFM_NOALLOC_WORKSPACE_BYTES (72ull << 20)
int hfm_riemann_solve_noalloc(const char *model, const int64_t *dims, const double *gridScales, const double *origin,
                              const double *d_metric, int metric_is_field, int dual, int nseeds, const double *seeds,
                              const double *seedValues, double *d_values, void *stream);

/* geodesicFlow (N x dim, FP64, device, may be NULL) and activeNeighs (N, FP64 holding the
 * bitset like the reference's export, device, may be NULL) (HFMInterface.hxx:536-553). */
int hfm_plan_flow(hfm_plan *plan, double *d_flow, double *d_active, void *stream);

/* activeOffsets (HFMInterface.hxx:554-577, exportActiveOffsets): N x nslots FP64 (device),
 * nslots = hfm_plan_n_active_slots(plan).  Per point, the offsets of the active differences in
 * stencil slot order (a symmetric difference with the sign of its active side), each packed
 * as sum_q uint8(offset_q) << 8 (dim - 1 - q) and stored as a double; remaining entries 0. */
int hfm_plan_active_offsets(hfm_plan *plan, double *d_offsets, void *stream);
int hfm_plan_n_active_slots(const hfm_plan *plan);

/* Final residual check of the last solve (or of values loaded with hfm_plan_load_values), in
 * FP64: at every reached non-seed point, the root of the local scheme
 * sum_k w_k max(0, u - v_k)^2 = 1 (EulerianStencil.hxx:30-72,104-153; v_k the smaller
 * neighbour value of difference k, over visible sides) is recomputed in the reference's FP64
 * operation order from the stored values and weights.  *max_rel receives the largest
 * |root - u| / max(|u|, |root|), *n_mismatch the number of points where exactly one of root and
 * u is finite.  A converged FP64 solve gives about 1e-15, an FP32 one the FP32 rounding of the
 * local solve (about 1e-7); a point left unconverged shows as a large deviation.  Synchronous
 * on `stream`.  The reference checks nothing of the kind (its heap order makes every accepted
 * value final); this is the check of the iterative solver's convergence. */
int hfm_plan_residual(hfm_plan *plan, double *max_rel, int64_t *n_mismatch, void *stream);

/* Geodesic backtracking from tips (GeodesicDiscreteSolver.hxx:17-231, driver
 * HFMInterface.hxx:452-487): the discrete solver, one wavefront per tip, on the last solve.
 *   tips: ntips x dim host physical points; step / threshold / volumeBound: geodesicStep
 *   (0.25), geodesicWeightThreshold (0.001), geodesicVolumeBound (<= 0: 5 * 1.3^dim).
 *   points: host, receives up to max_points ReDim'd points (x dim), the geodesics concatenated;
 *   lengths: host, ntips point counts (geodesicLengths); *npoints = total;
 *   *nfailed (may be NULL) = tips whose extraction gave up after 50 restarts (their points are
 *   still returned, as the reference does). */
int hfm_plan_geodesics(hfm_plan *plan, int ntips, const double *tips, double step, double threshold,
                       double volumeBound, double *points, int64_t max_points, int64_t *lengths, int64_t *npoints,
                       int *nfailed, void *stream);

/* The ODE geodesic solver (GeodesicODESolver::Run, Headers/Base/Implementation/GeodesicODESolver.hxx
 * :141-217, flow GeodesicODESolver.hxx:42-140; HFMInterface.hxx:464-475 selects it for
 * geodesicSolver = "ODE"): midpoint (RK2) steps of length step (grid units) along the flow
 * interpolated from the 2^d cell corners around the point, each corner's flow taken if its
 * value is within causality_tolerance widths of the smallest.  Stops at a seed of the last
 * solve, when the L-infinity grid distance to the seeds (capped at 6) has not decreased over
 * sqrt(d) / step steps, or (a failure) when a corner's flow has been read more than
 * 6 sqrt(d) / step times.  tips, points, lengths, npoints as hfm_plan_geodesics;
 * tip_failed (may be NULL, ntips entries) = 1 for each tip that failed.  Not available with
 * walls.  Errors: "geodesic ODE solver: ..." when a tip reads more than 32768 distinct grid
 * points or geodesicStep leaves more than 1024 steps of distance history. */
int hfm_plan_geodesics_ode(hfm_plan *plan, int ntips, const double *tips, double step, double causality_tolerance,
                           double *points, int64_t max_points, int64_t *lengths, int64_t *npoints, int *tip_failed,
                           void *stream);

/* Backward differentiation (FirstVariation::BackwardVariation, FirstVariation.h:341-385):
 * gradient of sum_t weights[t] * u(targets[t]) w.r.t. the cost field (row-major FP64 device
 * array costSensitivity; zero outside the targets' dependency cone).  targets are host
 * physical points (ntargets x dim).  Seeds reached are returned in seed_index (row-major
 * linear index) / seed_weight, up to max_seeds entries; *nseed_out = number reached. */
int hfm_plan_adjoint(hfm_plan *plan, int ntargets, const double *targets, const double *weights,
                     double *d_sensitivity, int64_t *seed_index, double *seed_weight, int max_seeds,
                     int *nseed_out, void *stream);

/* Forward differentiation (FirstVariation::ForwardVariation, FirstVariation.h:389-429, driver
 * FirstVariation.h:102-121): first-order variation of the arrival times of the last solve
 * along nvar directions of the cost (costVariation) and of the seed values
 * (seedValueVariation).
 *   d_cost_variation: device, FP64, row-major (dims..., nvar), or NULL (zero);
 *   seed_value_variation: host, FP64, nvar x nseeds row-major (the reference reads it RowMajor
 *   whatever the array ordering), nseeds = the seeds passed to the last hfm_plan_solve, in that
 *   order (a later seed overwrites an earlier one at the same grid point), or NULL (zero);
 *   d_value_variation: device, FP64, row-major (dims..., nvar) -- the 'valueVariation' output.
 * Errors: "Error : Inconsistent size of seedValueVariation" when nseeds differs from the
 * solve's seed count. */
int hfm_plan_forward_variation(hfm_plan *plan, int nvar, const double *d_cost_variation, int nseeds,
                               const double *seed_value_variation, double *d_value_variation, void *stream);

int hfm_plan_get_stats(const hfm_plan *plan, hfm_plan_stats *stats);

/* Per-point stencils of a Riemann plan, copied to host buffers in row-major point order:
 * offsets N x K x dim int8 and weights N x K (FP64 copy of the stored weights), in the
 * reference's slot order (what getStencils prints, HFMInterface.hxx:360-375). */
int hfm_plan_get_stencils(hfm_plan *plan, int8_t *h_offsets, double *h_weights);

/* Width of the value bands that order the block fast-iterative rounds, in units of the mean
 * one-step time 1/sqrt(sum_k w_k) of the stencils (default 8; <= 0 processes every active
 * tile in every round).  It changes the amount of work, never the fixed point. */
int hfm_plan_set_delta(hfm_plan *plan, double factor);

/* Tiles per round that the band width is steered to (x1.5 / /1.5 each round; 0: a default of
 * 64 for 3-D grids of 4096 tiles or more, else 1024).  Changes the work, never the fixed point. */
int hfm_plan_set_round_target(hfm_plan *plan, int tiles);

/* Blocks (of 4 wavefronts, one tile per wavefront at a time) of each update-kernel launch
 * (0: 960; 1280 for 2-D grids scheduled by the listing kernel).  A tuning knob; results do not
 * depend on it. */
int hfm_plan_set_grid(hfm_plan *plan, int blocks);

/* Jacobi sweeps a tile visit may run before an unconverged tile is handed back to the
 * schedule: 4, 8, 16 or 32, or 0 for the default (4 for Riemann3 and Elastica2, 8 for Riemann2
 * and ReedsSheppForward2); any other value is an error.  A tuning knob: it changes the work per visit and the number of visits, never the
 * fixed point. */
int hfm_plan_set_max_sweeps(hfm_plan *plan, int sweeps);

/* Per-launch timing of the update kernel (start/stop events taken from each dispatch's own
 * timestamps; off by default).  Fills fim_ms / fim_launches.  Timed solves run the rounds as
 * individual launches, not as the captured graph. */
int hfm_plan_set_timing(hfm_plan *plan, int enable);

/* Run the update loop as a captured HIP graph of one batch of rounds, replayed until
 * convergence (default on), or as individual launches.  Both compute the same fixed point.
 * No counterpart in the reference (its loop is the serial heap of HamiltonFastMarching.hxx). */
int hfm_plan_set_graph(hfm_plan *plan, int enable);

/* Let every round of the update loop list its own tiles (enable = 1: one kernel a round, each
 * block taking the pending tiles it owns below the round's band bound, then visiting them), or
 * schedule each round with a separate listing kernel (0); -1 (the default): the first on 2-D
 * grids of at most 4096 tiles, the second elsewhere (rounds of a thousand tiles and more
 * balance better over a global list).  Both compute the same fixed point.  No counterpart in
 * the reference. */
int hfm_plan_set_fused_schedule(hfm_plan *plan, int enable);

/* Order of the scheme: 1 (default) or 2 / 3, the recompute of every point at its acceptance with
 * second / third order differences where the next points along an offset are accepted and no
 * larger (replaces the `order` key, HFMInterface.hxx:358, HamiltonFastMarching.hxx:86-123,
 * 187-272).  Values above 3 act as 3, below 2 as 1.  Order 2 / 3 solves run the reference's serial
 * acceptance loop on the device (one wave popping a binary heap with libstdc++'s steps, so that
 * equal keys leave in the reference's order): FP64 plans without walls.  After such a solve,
 * geodesicFlow, activeOffsets and geodesics come from the loop's recompute over its accepted set
 * (order-aware, with the factor's corrections), differentiation reads its activeNeighs and values
 * as the reference's FirstVariation does, and the residual check (first order) refuses it. */
int hfm_plan_set_order(hfm_plan *plan, int order);

/* Stopping criteria (replace nMaxAccepted, stopAtDistance, stopWhenAnyAccepted,
 * stopWhenAllAccepted and their _Unoriented forms, ExtraAlgorithms/CommonStoppingCriteria.h:58-165), checked as
 * each point leaves the heap, before its recompute: stop at the max_accepted-th pop (< 0: no
 * limit), at a point whose value exceeds stop_distance (+inf: none), at any point of any_idx, or
 * once every point of all_idx has left the heap; the _u lists hold (x, y) indices x * n1 + y of
 * the curvature models, matched at any theta.  Indices are row-major grid indices.  A solve with
 * any criterion runs the serial acceptance loop (see hfm_plan_set_order), whose partial result --
 * accepted values, tentative values beyond them, their activeNeighs -- is the reference's; FP64
 * plans without walls.  All lists empty, max_accepted < 0 and stop_distance = +inf remove them. */
int hfm_plan_set_stopping(hfm_plan *plan, int64_t max_accepted, double stop_distance, int n_any,
                          const int64_t *any_idx, int n_all, const int64_t *all_idx, int n_any_u,
                          const int64_t *any_u_idx, int n_all_u, const int64_t *all_u_idx);

/* Static factoring (replaces the imported factor factoringValues / factoringGradients /
 * factoringMask / factoringIndexShift, StaticFactoring.hxx:77-113, and holds the computed one of
 * factoringRadius, :116-239, which hfm_run builds on the host): the factor T over a box of
 * box_dims grid points whose first grid index is shift[] (shift NULL: the whole grid, box_dims
 * = the grid's extents).  Host arrays, row-major over the box: values (T), gradients (grad T in
 * grid-index units, dim per point: a physical gradient times the grid scale of each axis) and
 * mask (NULL: everywhere; nonzero = the factor applies).  They are copied.  A point where the
 * factor applies is recomputed at acceptance, each difference corrected by
 * grad T . e + (one-sided difference of T along e of the order used, with the opposite sign);
 * the solve runs the serial acceptance loop (see hfm_plan_set_order): FP64 plans without walls.
 * values NULL removes the factor. */
int hfm_plan_set_factoring(hfm_plan *plan, const int64_t *box_dims, const int64_t *shift, const double *values,
                           const double *gradients, const uint8_t *mask);

/* ----------------------------------------------------------------------------------------- */
/* Continuous-batching solve server (no reference counterpart: the reference solves one      */
/* problem per Run(IO&) call; this serves many Riemann2 / Riemann3 solves concurrently)      */
/* ----------------------------------------------------------------------------------------- */
typedef struct hfm_server hfm_server;

typedef struct hfm_request {
    const char *model;          /* "Riemann2" | "Riemann3" */
    int64_t dims[3];            /* grid extents (dims[2] ignored in 2-D) */
    double gridScale;           /* h, the same on every axis */
    const double *metric;       /* FP64 row-major (dims..., d(d+1)/2), or one tensor if !metric_is_field */
    int metric_is_field;
    int metric_on_device;       /* 1: device pointer the caller keeps alive until completion;
                                   0: host pointer, copied by hfm_server_submit */
    int dual;                   /* 1: metric is the dual metric (dualMetric key) */
    int nseeds;
    const double *seeds;        /* host, nseeds x dim physical points (copied) */
    const double *seed_values;  /* host, nseeds values or NULL (copied) */
    double *values_out;         /* host buffer of prod(dims) doubles written on completion, or NULL */
} hfm_request;

typedef struct hfm_request_result {
    int status;                 /* 0: solved; nonzero: failed (message in error) */
    double latency_ms;          /* submit -> values available on the host */
    double queue_ms;            /* submit -> admitted (HBM budget and resident-slot limit) */
    double setup_ms;            /* admitted -> stencils and tile graph ready */
    int64_t rounds;             /* scheduling rounds the solve took part in */
    int64_t tile_visits;        /* tiles processed for this solve */
    char error[256];
} hfm_request_result;

typedef struct hfm_server_stats {
    int64_t submitted, completed, failed;
    int64_t resident;           /* solves currently on the device */
    int64_t max_resident_seen;
    int64_t batches, rounds;    /* kernel rounds over the union of resident solves */
    double hbm_in_use;          /* bytes reserved by admitted solves */
    double hbm_budget;
} hfm_server_stats;

/* Create a server.  hbm_budget_bytes: device bytes that resident solves may reserve (requests
 * wait until they fit; a request larger than the budget fails).  max_resident: cap on solves
 * sharing the device.  The update runs in FP32 with FP64 Selling stencils. */
int hfm_server_create(hfm_server **out, double hbm_budget_bytes, int max_resident);
/* Enqueue a solve; returns at once with a ticket.  Thread-safe. */
int hfm_server_submit(hfm_server *srv, const hfm_request *req, int64_t *ticket);
/* *done = 1 once the ticket's result is available. */
int hfm_server_poll(hfm_server *srv, int64_t ticket, int *done);
/* Block until the ticket completes (timeout_ms < 0: no limit); fills res.  Returns nonzero on
 * timeout or unknown ticket. */
int hfm_server_wait(hfm_server *srv, int64_t ticket, double timeout_ms, hfm_request_result *res);
int hfm_server_get_stats(hfm_server *srv, hfm_server_stats *st);
/* Round execution: on = 0 (default) launches a scheduling kernel and an update kernel per
 * round; on = 1 runs each batch of 8 rounds in one persistent launch with grid-wide barriers
 * (one block per CU).  Takes effect from the next batch; returns nonzero if the device cannot
 * hold the persistent grid.  No counterpart in the reference (its solver is a serial heap loop,
 * HamiltonFastMarching.hxx:59-112): this selects between two executions of the same rounds. */
int hfm_server_set_persistent(hfm_server *srv, int on);
/* Stop the worker threads (waits for in-flight solves) and free every device buffer. */
void hfm_server_destroy(hfm_server *srv);

/* Message of the last failing plan call on this thread ("" if none). */
const char *hfm_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
